// Device kernels of the reverse mode (FP64): the selected-inverse GEMMs and
// copies (schedule in selinv_plan.cpp), and the gradient kernels on the
// caller's entries.
#include <hip/hip_runtime.h>

#include "selinv_kernels.h"

namespace mdgpu {

namespace {

constexpr int kThreads = 256;

// C = alpha op(A) op(B) + beta C, 64x64 tile per workgroup, tiles of many
// jobs in one launch, on v_mfma_f64_16x16x4f64: four wavefronts of 32x32
// (2x2 blocks of 16x16).  K is staged through two LDS buffers 16 at a time
// (the next k-tile is loaded into registers while this one is multiplied),
// each operand row-major with k contiguous: op(A)(i, k) and op(B)^T(j, k)
// for both transpose flags, so one MFMA operand read is one 16-byte LDS
// read.  The contraction index of one 16x16x4 instruction is free to
// relabel: lane group lg = lane / 16 supplies k = 4 lg + t in sub-step t.
// Accumulator register v of lane (16 lg + li) holds the C entry of row li
// of the 16x16 block and column lg + 4 v (the product is formed as
// op(B)^T op(A)^T).
__global__ __launch_bounds__(kThreads, 2) void sel_gemm_kernel(const SelGemm* __restrict__ jobs, int njobs,
                                                               SelBuffers bufs) {
    typedef double acc_t __attribute__((ext_vector_type(4)));
    constexpr int BT = 64, KT = 16, LDK = KT + 2;
    __shared__ __align__(16) double As[2][BT * LDK];
    __shared__ __align__(16) double Bs[2][BT * LDK];
    int lo = 0, hi = njobs - 1;
    const int bid = blockIdx.x;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (jobs[mid].tile_base <= bid) lo = mid; else hi = mid - 1;
    }
    const SelGemm g = jobs[lo];
    const int local = bid - g.tile_base;
    const int ti = local % g.tiles_m, tj = local / g.tiles_m;
    const double* __restrict__ A = bufs.p[g.a_buf] + g.a_off;
    const double* __restrict__ B = bufs.p[g.b_buf] + g.b_off;
    double* __restrict__ C = bufs.p[g.c_buf] + g.c_off;
    const int row0 = ti * BT, col0 = tj * BT;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, li = lane & 15, lg = lane >> 4;
    const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
    // staging: each thread stages four k of one row per operand.  Rows
    // contiguous in memory (op = no transpose for A, transpose for B): row
    // tid % 64, k 4 (tid / 64) + e (loads 64 rows apart are coalesced);
    // k contiguous: row tid / 4, k 4 (tid % 4) + e (one 32-byte run).
    const bool a_rows = !g.ta, b_rows = g.tb;
    const int ar = a_rows ? (tid & 63) : (tid >> 2), ak = a_rows ? 4 * (tid >> 6) : 4 * (tid & 3);
    const int br = b_rows ? (tid & 63) : (tid >> 2), bk = b_rows ? 4 * (tid >> 6) : 4 * (tid & 3);
    const int ia = row0 + ar, jb = col0 + br;
    const bool a_ok = ia < g.M, b_ok = jb < g.N;
    double ra[4], rb[4];
    auto fetch = [&](int k0) {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int k = k0 + ak + e, kb = k0 + bk + e;
            const int kc = min(k, g.K - 1), kbc = min(kb, g.K - 1);
            const int64_t ai = a_ok ? ia : 0, bj = b_ok ? jb : 0;
            // op(A)(i, k): A[i + k lda] or, transposed, A[k + i lda]
            ra[e] = g.ta ? A[kc + ai * g.lda] : A[ai + (int64_t)kc * g.lda];
            // op(B)(k, j): B[k + j ldb] or, transposed, B[j + k ldb]
            rb[e] = g.tb ? B[bj + (int64_t)kbc * g.ldb] : B[kbc + bj * g.ldb];
            if (!(a_ok && k < g.K)) ra[e] = 0.0;
            if (!(b_ok && kb < g.K)) rb[e] = 0.0;
        }
    };
    auto store = [&](int buf) {
        typedef double d2 __attribute__((ext_vector_type(2)));
        *reinterpret_cast<d2*>(&As[buf][ar * LDK + ak]) = d2{ra[0], ra[1]};
        *reinterpret_cast<d2*>(&As[buf][ar * LDK + ak + 2]) = d2{ra[2], ra[3]};
        *reinterpret_cast<d2*>(&Bs[buf][br * LDK + bk]) = d2{rb[0], rb[1]};
        *reinterpret_cast<d2*>(&Bs[buf][br * LDK + bk + 2]) = d2{rb[2], rb[3]};
    };
    acc_t acc[2][2];
#pragma unroll
    for (int a2 = 0; a2 < 2; ++a2)
#pragma unroll
        for (int b2 = 0; b2 < 2; ++b2) acc[a2][b2] = acc_t{0.0, 0.0, 0.0, 0.0};
    fetch(0);
    store(0);
    if (KT < g.K) fetch(KT);
    __syncthreads();
    for (int k0 = 0, cur = 0; k0 < g.K; k0 += KT, cur ^= 1) {
        if (k0 + KT < g.K) {
            store(cur ^ 1);
            if (k0 + 2 * KT < g.K) fetch(k0 + 2 * KT);
        }
        const double* __restrict__ Ac = As[cur];
        const double* __restrict__ Bc = Bs[cur];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            double av[2], bv[2];
#pragma unroll
            for (int r2 = 0; r2 < 2; ++r2) av[r2] = Ac[(wr + 16 * r2 + li) * LDK + 4 * lg + t];
#pragma unroll
            for (int c2 = 0; c2 < 2; ++c2) bv[c2] = Bc[(wc + 16 * c2 + li) * LDK + 4 * lg + t];
#pragma unroll
            for (int r2 = 0; r2 < 2; ++r2)
#pragma unroll
                for (int c2 = 0; c2 < 2; ++c2)
                    acc[r2][c2] = __builtin_amdgcn_mfma_f64_16x16x4f64(bv[c2], av[r2], acc[r2][c2], 0, 0, 0);
        }
        __syncthreads();
    }
    // acc[r2][c2][v] = C(row0 + wr + 16 r2 + li, col0 + wc + 16 c2 + lg + 4 v)
#pragma unroll
    for (int r2 = 0; r2 < 2; ++r2)
#pragma unroll
        for (int c2 = 0; c2 < 2; ++c2)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int i = row0 + wr + 16 * r2 + li, j = col0 + wc + 16 * c2 + lg + 4 * v;
                if (i >= g.M || j >= g.N) continue;
                double* cp = C + i + (int64_t)j * g.ldc;
                *cp = g.beta == 0.0 ? g.alpha * acc[r2][c2][v] : g.alpha * acc[r2][c2][v] + g.beta * *cp;
            }
}

// The copies of the selected inverse (SelCopyKind), one work item per
// workgroup: the item's entries flattened (column by column, rows
// contiguous, so reads and writes are coalesced wherever the layouts allow),
// four entries per thread in flight.
__global__ __launch_bounds__(kThreads) void sel_copy_kernel(int kind, const SelCopy* __restrict__ items,
                                                            const FrontDesc* __restrict__ fronts,
                                                            const int* __restrict__ rel,
                                                            const int64_t* __restrict__ zf_ofs, SelBuffers bufs) {
    const SelCopy it = items[blockIdx.x];
    const FrontDesc d = fronts[it.front];
    double* __restrict__ ZF = bufs.p[SB_ZF];
    double* __restrict__ U = bufs.p[SB_U];
    const int n = d.n, c = d.c, m = d.m, ncol = it.j1 - it.j0;
    if (kind == SC_INIT) {
        double* __restrict__ z = ZF + zf_ofs[it.front];
        const int cnt = ncol * m;   // an item holds at most ~4096 entries or one column
        for (int e0 = threadIdx.x; e0 < cnt; e0 += 4 * kThreads) {
            double v[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int e = min(e0 + u * kThreads, cnt - 1);
                v[u] = U[d.uofs + e % m + (int64_t)(it.j0 + e / m) * m];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int e = e0 + u * kThreads;
                if (e < cnt) z[(c + e % m) + (int64_t)(c + it.j0 + e / m) * n] = v[u];
            }
        }
    } else if (kind == SC_TRANS) {
        double* __restrict__ z = ZF + zf_ofs[it.front];
        const int r0 = it.b0 + it.b;                 // first row of R
        for (int t = threadIdx.x; t < ncol * it.b; t += kThreads) {
            const int jj = t % it.b, i = r0 + it.j0 + t / it.b;
            z[(it.b0 + jj) + (int64_t)i * n] = z[i + (int64_t)(it.b0 + jj) * n];
        }
    } else if (kind == SC_OUT) {
        const double* __restrict__ z = ZF + zf_ofs[it.front];
        double* __restrict__ Z = bufs.p[SB_Z] + d.lofs;
        const int64_t base = (int64_t)it.j0 * n;   // the item's columns are one contiguous run
        const int cnt = ncol * n;
        for (int e0 = threadIdx.x; e0 < cnt; e0 += 4 * kThreads) {
            double v[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) v[u] = z[base + min(e0 + u * kThreads, cnt - 1)];
#pragma unroll
            for (int u = 0; u < 4; ++u)
                if (e0 + u * kThreads < cnt) Z[base + e0 + u * kThreads] = v[u];
        }
    } else {   // SC_PUSH: Z22 of child front it.front from its parent's ZF
        const FrontDesc pd = fronts[d.parent];
        const double* __restrict__ zp = ZF + zf_ofs[d.parent];
        const int* __restrict__ r = rel + d.relofs;
        const int cnt = ncol * m;
        for (int e0 = threadIdx.x; e0 < cnt; e0 += 4 * kThreads) {
            double v[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int e = min(e0 + u * kThreads, cnt - 1);
                v[u] = zp[r[e % m] + (int64_t)r[it.j0 + e / m] * pd.n];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int e = e0 + u * kThreads;
                if (e < cnt) U[d.uofs + e % m + (int64_t)(it.j0 + e / m) * m] = v[u];
            }
        }
    }
}

// d log det A / d a_e for the caller's lower-triangle entries:
// scale_e * Z at the entry's L-store position (2 off the diagonal).
__global__ void logdet_grad_kernel(const double* __restrict__ Z, const int64_t* __restrict__ dst,
                                   const int* __restrict__ arow, const int* __restrict__ acol, double* __restrict__ g,
                                   int nnz) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < nnz) g[e] = (arow[e] == acol[e] ? 1.0 : 2.0) * Z[dst[e]];
}

// adjoint of x = A^{-1} b on the caller's entries: g_e = -(bb_r x_c + bb_c x_r)
// off the diagonal, -bb_r x_r on it (bb = A^{-1} xbar)
__global__ void solve_adjoint_kernel(const double* __restrict__ x, const double* __restrict__ bb,
                                     const int* __restrict__ arow, const int* __restrict__ acol,
                                     double* __restrict__ g, int nnz) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nnz) return;
    const int r = arow[e], c = acol[e];
    g[e] = r == c ? -bb[r] * x[r] : -(bb[r] * x[c] + bb[c] * x[r]);
}

}  // namespace

void launch_sel_gemm(const SelGemm* jobs, int njobs, int nblocks, SelBuffers bufs, hipStream_t s) {
    if (nblocks > 0) sel_gemm_kernel<<<nblocks, kThreads, 0, s>>>(jobs, njobs, bufs);
}
void launch_sel_copy(int kind, const SelCopy* items, int count, const FrontDesc* fronts, const int* rel,
                     const int64_t* zf_ofs, SelBuffers bufs, hipStream_t s) {
    if (count > 0) sel_copy_kernel<<<count, kThreads, 0, s>>>(kind, items, fronts, rel, zf_ofs, bufs);
}
void launch_logdet_grad(const double* Z, const int64_t* dst, const int* arow, const int* acol, double* g, int nnz,
                        hipStream_t s) {
    if (nnz > 0) logdet_grad_kernel<<<(nnz + kThreads - 1) / kThreads, kThreads, 0, s>>>(Z, dst, arow, acol, g, nnz);
}
void launch_solve_adjoint(const double* x, const double* bb, const int* arow, const int* acol, double* g, int nnz,
                          hipStream_t s) {
    if (nnz > 0)
        solve_adjoint_kernel<<<(nnz + kThreads - 1) / kThreads, kThreads, 0, s>>>(x, bb, arow, acol, g, nnz);
}

}  // namespace mdgpu

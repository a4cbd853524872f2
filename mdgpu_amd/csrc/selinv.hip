// Device kernels of the reverse mode (FP64): the selected-inverse GEMMs and
// copies (schedule in selinv_plan.cpp), and the gradient kernels on the
// caller's entries.
#include <hip/hip_runtime.h>

#include "selinv_kernels.h"

namespace mdgpu {

namespace {

constexpr int kThreads = 256;

// C = alpha op(A) op(B) + beta C, 64x64 tile per workgroup, 4x4 per thread,
// K staged through LDS 16 at a time.  Tiles of many jobs in one launch.
__global__ __launch_bounds__(kThreads) void sel_gemm_kernel(const SelGemm* __restrict__ jobs, int njobs,
                                                            SelBuffers bufs) {
    __shared__ double As[16][64 + 1];
    __shared__ double Bs[16][64 + 1];
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
    const int row0 = ti * 64, col0 = tj * 64;
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    double acc[4][4];
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int s = 0; s < 4; ++s) acc[r][s] = 0.0;
    const int li = threadIdx.x & 63, lk = threadIdx.x >> 6;
    for (int k0 = 0; k0 < g.K; k0 += 16) {
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const int kk = lk + 4 * s, k = k0 + kk;
            const int i = row0 + li, j = col0 + li;
            double a = 0.0, b = 0.0;
            if (k < g.K && i < g.M) a = g.ta ? A[k + (int64_t)i * g.lda] : A[i + (int64_t)k * g.lda];
            if (k < g.K && j < g.N) b = g.tb ? B[j + (int64_t)k * g.ldb] : B[k + (int64_t)j * g.ldb];
            As[kk][li] = a;
            Bs[kk][li] = b;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < 16; ++kk) {
            double a[4], b[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) a[r] = As[kk][tx + 16 * r];
#pragma unroll
            for (int s = 0; s < 4; ++s) b[s] = Bs[kk][ty + 16 * s];
#pragma unroll
            for (int r = 0; r < 4; ++r)
#pragma unroll
                for (int s = 0; s < 4; ++s) acc[r][s] += a[r] * b[s];
        }
        __syncthreads();
    }
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        const int j = col0 + ty + 16 * s;
        if (j >= g.N) continue;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int i = row0 + tx + 16 * r;
            if (i >= g.M) continue;
            double* cp = C + i + (int64_t)j * g.ldc;
            *cp = g.beta == 0.0 ? g.alpha * acc[r][s] : g.alpha * acc[r][s] + g.beta * *cp;
        }
    }
}

// The copies of the selected inverse (SelCopyKind), one work item per
// workgroup: columns j of the item, lanes over rows (coalesced).
__global__ __launch_bounds__(kThreads) void sel_copy_kernel(int kind, const SelCopy* __restrict__ items,
                                                            const FrontDesc* __restrict__ fronts,
                                                            const int* __restrict__ rel,
                                                            const int64_t* __restrict__ zf_ofs, SelBuffers bufs) {
    const SelCopy it = items[blockIdx.x];
    const FrontDesc d = fronts[it.front];
    double* __restrict__ ZF = bufs.p[SB_ZF];
    double* __restrict__ U = bufs.p[SB_U];
    const int n = d.n, c = d.c, m = d.m;
    if (kind == SC_INIT) {
        double* __restrict__ z = ZF + zf_ofs[it.front];
        for (int j = it.j0; j < it.j1; ++j)
            for (int i = threadIdx.x; i < m; i += kThreads)
                z[(c + i) + (int64_t)(c + j) * n] = U[d.uofs + i + (int64_t)j * m];
    } else if (kind == SC_TRANS) {
        double* __restrict__ z = ZF + zf_ofs[it.front];
        const int r0 = it.b0 + it.b;                 // first row of R
        for (int t = threadIdx.x; t < (it.j1 - it.j0) * it.b; t += kThreads) {
            const int jj = t % it.b, i = r0 + it.j0 + t / it.b;
            z[(it.b0 + jj) + (int64_t)i * n] = z[i + (int64_t)(it.b0 + jj) * n];
        }
    } else if (kind == SC_OUT) {
        const double* __restrict__ z = ZF + zf_ofs[it.front];
        double* __restrict__ Z = bufs.p[SB_Z] + d.lofs;
        for (int j = it.j0; j < it.j1; ++j)
            for (int i = threadIdx.x; i < n; i += kThreads) Z[i + (int64_t)j * n] = z[i + (int64_t)j * n];
    } else {   // SC_PUSH: Z22 of child front it.front from its parent's ZF
        const FrontDesc pd = fronts[d.parent];
        const double* __restrict__ zp = ZF + zf_ofs[d.parent];
        const int* __restrict__ r = rel + d.relofs;
        for (int j = it.j0; j < it.j1; ++j) {
            const int pj = r[j];
            for (int i = threadIdx.x; i < m; i += kThreads)
                U[d.uofs + i + (int64_t)j * m] = zp[r[i] + (int64_t)pj * pd.n];
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

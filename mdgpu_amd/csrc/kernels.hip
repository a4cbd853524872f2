// Device kernels of the multifrontal factorization and solves (gfx950).
//
// Conventions: fronts are column-major; the panel P_f of front f has leading
// dimension n_f; update matrices U_f have leading dimension m_f.  A
// workgroup is 256 lanes = four 64-wide wavefronts.  No kernel uses atomics
// on floating-point data: every summation order is fixed by the schedule.
#include <hip/hip_runtime.h>

#include <climits>
#include <cstdlib>
#include <string>

#include "kernels.h"

namespace mdgpu {

namespace {

constexpr int kThreads = 256;

__device__ inline double to_d(double v) { return v; }
__device__ inline double to_d(float v) { return (double)v; }

__device__ inline double wave_sum(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

// ---------------------------------------------------------------- fill / A
template <typename T>
__global__ void fill_kernel(T* __restrict__ p, int64_t n, T v) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

// front assembly from A: every caller entry lands at a precomputed panel
// offset (the value scatter of MatrizGetFk_3, mdgpu.cu:272-282)
template <typename T>
__global__ void scatter_a_kernel(T* __restrict__ L, const double* __restrict__ ax,
                                 const int64_t* __restrict__ dst, int nnz) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < nnz) L[dst[e]] = (T)ax[e];
}

// -------------------------------------------------------------- extend-add
// parent(rel[i], rel[j]) += U_child(i, j) over the lower triangle of the
// columns [j0, j1) of one child (ExtendAdd_CPU / ExtendAdd_GPU_device,
// mdgpu.cu:748-788).  A column lands either in a pivot column of the
// parent's panel or in the parent's update matrix.  One wavefront per
// column, lanes over rows (coalesced reads; writes follow the increasing
// relative map).
template <typename T>
__global__ __launch_bounds__(kThreads) void extend_add_kernel(const EaItem* __restrict__ items,
                                                              const FrontDesc* __restrict__ fronts,
                                                              const int* __restrict__ rel, T* __restrict__ L,
                                                              T* __restrict__ U) {
    const EaItem it = items[blockIdx.x];
    const FrontDesc cd = fronts[it.child];
    const FrontDesc pd = fronts[cd.parent];
    const int* __restrict__ r = rel + cd.relofs;
    const T* __restrict__ src = U + cd.uofs;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int j = it.j0 + wave; j < it.j1; j += kThreads / 64) {
        const int pj = r[j];
        T* dcol;
        int rowoff;
        if (pj < pd.c) {
            dcol = L + pd.lofs + (int64_t)pj * pd.n;
            rowoff = 0;
        } else {
            dcol = U + pd.uofs + (int64_t)(pj - pd.c) * pd.m;
            rowoff = pd.c;
        }
        const T* scol = src + (int64_t)j * cd.m;
#pragma unroll 4
        for (int i = j + lane; i < cd.m; i += 64) dcol[r[i] - rowoff] += scol[i];
    }
}

// ------------------------------------------------------ dense: LDS panels
// Partial Cholesky of a whole panel held in LDS: for each pivot column,
// sqrt, scale, rank-1 update of the remaining pivot columns (the unblocked
// scal+syr loop of FactorAux_CPU_3, mdgpu.cu:1318-1326, applied to the full
// n_f rows, which also performs the trsm of :1330).
template <typename T>
__global__ __launch_bounds__(kThreads) void panel_kernel(const int* __restrict__ list,
                                                         const FrontDesc* __restrict__ fronts,
                                                         T* __restrict__ L, int* __restrict__ info) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T* P = reinterpret_cast<T*>(smem_raw);
    __shared__ T pivinv;
    const FrontDesc d = fronts[list[blockIdx.x]];
    const int n = d.n, c = d.c;
    T* __restrict__ g = L + d.lofs;
    const int tot = n * c;
    for (int t = threadIdx.x; t < tot; t += kThreads) P[t] = g[t];
    __syncthreads();
    for (int j = 0; j < c; ++j) {
        if (threadIdx.x == 0) {
            T dj = P[j + j * n];
            if (!(dj > T(0))) atomicMin(info, d.k + j);
            T rt = sqrt(dj);
            P[j + j * n] = rt;
            pivinv = T(1) / rt;
        }
        __syncthreads();
        const T s = pivinv;
        for (int i = j + 1 + threadIdx.x; i < n; i += kThreads) P[i + j * n] *= s;
        __syncthreads();
        for (int i = j + 1 + threadIdx.x; i < n; i += kThreads) {
            const T lij = P[i + j * n];
            const int cmax = min(i, c - 1);
#pragma unroll 4
            for (int col = j + 1; col <= cmax; ++col) P[i + col * n] -= lij * P[col + j * n];
        }
        __syncthreads();
    }
    for (int t = threadIdx.x; t < tot; t += kThreads) g[t] = P[t];
}

// ----------------------------------------------- dense: blocked panel path
// Wavefront-level dense kernels on a 64x64 diagonal block held in registers:
// lane i keeps row i (a[0..63]); column entries of other lanes are read with
// v_readlane broadcasts, so no LDS or barrier sits on the serial chain.
__device__ inline float bcast(float v, int src) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), src));
}
__device__ inline double bcast(double v, int src) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, src);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), src);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Column `lane` of inv(L) by forward substitution, in LDS:
// Xs[r*65 + lane] = inv(L)(r, lane), reading L rows Ls(r, l) = Ls[r*65 + l]
// (uniform addresses) and rd[r] = 1 / L(r, r).  Each lane only reads its
// own column of Xs, so no barrier is needed between rows.
template <typename T>
__device__ inline void wave_trinv64(const T* Ls, const T* rd, T* Xs, int lane) {
    for (int r = 0; r < 64; ++r) {
        T acc = (lane == r) ? T(1) : T(0);
#pragma unroll 8
        for (int l = lane; l < r; ++l) acc -= Ls[r * 65 + l] * Xs[l * 65 + lane];
        Xs[r * 65 + lane] = acc * rd[r];
    }
}

// Stage the factor rows (lane i: a[0..i]) into LDS for wave_trinv64.
template <typename T>
__device__ inline void stage_rows(const T (&a)[64], T* Ls, T* rd, int lane) {
#pragma unroll
    for (int c = 0; c < 64; ++c) Ls[lane * 65 + c] = c <= lane ? a[c] : T(0);
    rd[lane] = T(1) / Ls[lane * 65 + lane];
    __syncthreads();
}

// Diagonal block b x b of the blocked path: Cholesky in four 16-column
// panels -- each panel factored by wavefront 0 with its rows in registers
// (pivot and column entries broadcast by v_readlane within the panel), the
// trailing part of the block updated by all four wavefronts from LDS.
// Rows and columns beyond b are identity padding.  (The block's inverse,
// needed only by the solves, is formed later by inv_kernel.)
template <typename T>
__global__ __launch_bounds__(kThreads) void diag_kernel(const DiagJob* __restrict__ jobs, T* __restrict__ L,
                                                        T* __restrict__ inv, int* __restrict__ info) {
    __shared__ T As[64 * 65];   // column-major, As[r + 65 c] = A(r, c)
    const DiagJob jb = jobs[blockIdx.x];
    const int b = jb.b, ld = jb.ld;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    T* __restrict__ g = L + jb.p_off;
    for (int t = tid; t < 64 * 64; t += kThreads) {
        const int r = t & 63, c = t >> 6;
        As[r + 65 * c] = (r < b && c < b) ? g[r + (int64_t)c * ld] : (r == c ? T(1) : T(0));
    }
    __syncthreads();
    int bad = 64;
    for (int p = 0; p < 4; ++p) {
        const int c0 = 16 * p;
        if (wave == 0) {
            const int r = c0 + lane;
            T a[16];
#pragma unroll
            for (int j = 0; j < 16; ++j) a[j] = r < 64 ? As[r + 65 * (c0 + j)] : T(0);
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                const T djj = bcast(a[j], j);
                if (!(djj > T(0)) && bad == 64) bad = c0 + j;
                const T sq = sqrt(djj);
                const T is = T(1) / sq;
                a[j] = (lane == j) ? sq : a[j] * is;
#pragma unroll
                for (int k = j + 1; k < 16; ++k) a[k] -= a[j] * bcast(a[j], k);
            }
            if (r < 64) {
#pragma unroll
                for (int j = 0; j < 16; ++j)
                    if (lane >= j) As[r + 65 * (c0 + j)] = a[j];
            }
        }
        __syncthreads();
        const int w = 48 - c0;          // trailing rows/cols [c0+16, 64)
        for (int idx = tid; idx < w * w; idx += kThreads) {
            const int i = c0 + 16 + idx % w, j = c0 + 16 + idx / w;
            if (i < j) continue;
            T acc = T(0);
#pragma unroll
            for (int t = 0; t < 16; ++t) acc += As[i + 65 * (c0 + t)] * As[j + 65 * (c0 + t)];
            As[i + 65 * j] -= acc;
        }
        __syncthreads();
    }
    if (wave == 0 && lane == 0 && bad < b) atomicMin(info, jb.col0 + bad);
    for (int t = tid; t < 64 * 64; t += kThreads) {
        const int r = t & 63, c = t >> 6;
        if (r < b && c < b && c <= r) g[r + (int64_t)c * ld] = As[r + 65 * c];
    }
}

// Inverse of one 64-column diagonal block of an already factored panel.
template <typename T>
__global__ __launch_bounds__(64) void inv_kernel(const InvJob* __restrict__ jobs, const FrontDesc* __restrict__ fronts,
                                                 const T* __restrict__ L, T* __restrict__ inv) {
    __shared__ T Ls[64 * 65];
    __shared__ T Xs[64 * 65];
    __shared__ T rd[64];
    const InvJob jb = jobs[blockIdx.x];
    const FrontDesc d = fronts[jb.front];
    const int b = min(64, d.c - jb.j0), lane = threadIdx.x;
    const T* __restrict__ g = L + d.lofs + jb.j0 + (int64_t)jb.j0 * d.n;
    T a[64];
#pragma unroll
    for (int c = 0; c < 64; ++c)
        a[c] = (lane < b && c < b && c <= lane) ? g[lane + (int64_t)c * d.n] : (lane == c ? T(1) : T(0));
    stage_rows(a, Ls, rd, lane);
    wave_trinv64(Ls, rd, Xs, lane);
    T* __restrict__ w = inv + d.dinv + (int64_t)(jb.j0 / 64) * 4096;
    if (lane < b)
        for (int r = 0; r < b; ++r) w[lane + 64 * r] = Xs[r * 65 + lane];
}

__device__ inline int find_job_trsm(const TrsmJob* jobs, int njobs, int bid) {
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (jobs[mid].tile_base <= bid) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// R <- R * L_D^{-T} for a 64-row tile of the rows below a diagonal block
// (the trsm of FactorAux_CPU_3, mdgpu.cu:1330): each of 64 threads solves
// x L_D^T = r for its row by forward substitution; L_D (read from the
// panel) and the rows sit in LDS, L_D entries are uniform broadcasts.
template <typename T>
__global__ __launch_bounds__(kThreads) void trsm_kernel(const TrsmJob* __restrict__ jobs, int njobs,
                                                        T* __restrict__ L) {
    __shared__ T R[64 * 65];    // R[i * 65 + l]: row i of the tile
    __shared__ T Ld[64 * 65];   // Ld[j * 65 + l] = L_D(j, l)
    __shared__ T rd[64];
    const TrsmJob jb = jobs[find_job_trsm(jobs, njobs, blockIdx.x)];
    const int tile = blockIdx.x - jb.tile_base;
    const int r0 = tile * 64;
    const int rows = min(64, jb.rows - r0);
    const int b = jb.b, ld = jb.ld;
    T* __restrict__ g = L + jb.p_off + r0;
    const T* __restrict__ dg = L + jb.p_off - b;     // the diagonal block (b rows above)
    for (int t = threadIdx.x; t < 64 * b; t += kThreads) {
        const int i = t & 63, l = t >> 6;
        R[i * 65 + l] = i < rows ? g[i + (int64_t)l * ld] : T(0);
        if (i < b) Ld[i * 65 + l] = l <= i ? dg[i + (int64_t)l * ld] : T(0);
    }
    __syncthreads();
    if (threadIdx.x < b) rd[threadIdx.x] = T(1) / Ld[threadIdx.x * 66];
    __syncthreads();
    // blocked substitution: 16 columns at a time on the 64 row threads, then
    // all threads apply those columns to the remaining ones
    for (int c0 = 0; c0 < b; c0 += 16) {
        const int cw = min(16, b - c0);
        if (threadIdx.x < 64) {
            T* row = R + threadIdx.x * 65;
            for (int j = c0; j < c0 + cw; ++j) {
                T acc = row[j];
                const T* lj = Ld + j * 65;
                for (int l = c0; l < j; ++l) acc -= row[l] * lj[l];
                row[j] = acc * rd[j];
            }
        }
        __syncthreads();
        const int rest = b - c0 - cw;
        for (int idx = threadIdx.x; idx < 64 * rest; idx += kThreads) {
            const int i = idx & 63, j = c0 + cw + (idx >> 6);
            T acc = T(0);
            const T* ri = R + i * 65 + c0;
            const T* lj = Ld + j * 65 + c0;
#pragma unroll 16
            for (int t = 0; t < cw; ++t) acc += ri[t] * lj[t];
            R[i * 65 + j] -= acc;
        }
        __syncthreads();
    }
    for (int t = threadIdx.x; t < 64 * b; t += kThreads) {
        const int i = t & 63, l = t >> 6;
        if (i < rows) g[i + (int64_t)l * ld] = R[i * 65 + l];
    }
}

// C -= A * B^T on 64x64 tiles (beta0: C = -A * B^T), optionally keeping only
// the lower triangle; tiles of many jobs (fronts) in one launch.  This is the
// trailing-matrix gemm of FactorAux_CPU_3 / FactorAux_GPU_2 (mdgpu.cu:1336,
// :1058) restricted to the lower triangle.
template <typename T>
__global__ __launch_bounds__(kThreads) void gemm_nt_kernel(const GemmJob* __restrict__ jobs, int njobs,
                                                           T* __restrict__ L, T* __restrict__ U) {
    __shared__ T As[16][64 + 1];
    __shared__ T Bs[16][64 + 1];
    int lo = 0, hi = njobs - 1;
    const int bid = blockIdx.x;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (jobs[mid].tile_base <= bid) lo = mid; else hi = mid - 1;
    }
    const GemmJob g = jobs[lo];
    int local = bid - g.tile_base;
    int ti, tj;
    if (g.lower) {
        tj = 0;
        int cnt = g.tiles_m;
        while (local >= cnt) { local -= cnt; ++tj; cnt = g.tiles_m - tj; }
        ti = tj + local;
    } else {
        tj = local / g.tiles_m;
        ti = local % g.tiles_m;
    }
    const T* __restrict__ A = L + g.a_off;
    const T* __restrict__ B = L + g.b_off;
    T* __restrict__ C = (g.c_in_arena ? U : L) + g.c_off;
    const int row0 = ti * 64, col0 = tj * 64;
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    T acc[4][4];
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int s = 0; s < 4; ++s) acc[r][s] = T(0);
    const int li = threadIdx.x & 63, lk = threadIdx.x >> 6;
    for (int k0 = 0; k0 < g.K; k0 += 16) {
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const int kk = lk + 4 * s;
            const int k = k0 + kk;
            const int ia = row0 + li, jb = col0 + li;
            As[kk][li] = (k < g.K && ia < g.M) ? A[ia + (int64_t)k * g.lda] : T(0);
            Bs[kk][li] = (k < g.K && jb < g.N) ? B[jb + (int64_t)k * g.ldb] : T(0);
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < 16; ++kk) {
            T a[4], bb[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) a[r] = As[kk][tx + 16 * r];
#pragma unroll
            for (int s = 0; s < 4; ++s) bb[s] = Bs[kk][ty + 16 * s];
#pragma unroll
            for (int r = 0; r < 4; ++r)
#pragma unroll
                for (int s = 0; s < 4; ++s) acc[r][s] += a[r] * bb[s];
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
            if (i >= g.M || (g.lower && i < j)) continue;
            T* cp = C + i + (int64_t)j * g.ldc;
            *cp = g.beta0 ? -acc[r][s] : *cp - acc[r][s];
        }
    }
}

// MFMA version of gemm_nt_kernel for gfx950: 128x128 C tiles, four
// wavefronts of 64x64, v_mfma_{f32,f64}_16x16x4.  The contraction index of
// one 16x16x4 instruction is free to relabel as long as both operands agree,
// so lane group g = lane/16 supplies k = 4g+t in sub-step t: each lane then
// reads a contiguous 4-vector of k from LDS.  The product is formed as
// (B A^T) so that each lane's four accumulator rows map to four C columns
// and a 16-lane group stores 16 consecutive C rows.
template <typename T> struct MfmaOp;
// Accumulator register v of lane (16*lg + li) holds D(row(lg, v), li): the
// 32-bit 16x16 forms put rows 4*lg..4*lg+3 in one lane, the f64 form
// interleaves them (row lg + 4*v) -- both measured on gfx950.
template <> struct MfmaOp<float> {
    typedef float acc_t __attribute__((ext_vector_type(4)));
    __device__ static acc_t mma(float a, float b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    __device__ static int row(int lg, int v) { return 4 * lg + v; }
};
template <> struct MfmaOp<double> {
    typedef double acc_t __attribute__((ext_vector_type(4)));
    __device__ static acc_t mma(double a, double b, acc_t c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    __device__ static int row(int lg, int v) { return lg + 4 * v; }
};

template <typename T>
__global__ __launch_bounds__(kThreads) void gemm_mfma_kernel(const GemmJob* __restrict__ jobs, int njobs,
                                                             T* __restrict__ L, T* __restrict__ U) {
    // k-tile depth: 32 for FP32, 16 for FP64 (register budget of the
    // prefetch); the next k-tile is loaded into registers while the current
    // one is multiplied out of LDS.
    constexpr int BT = 128, KT = sizeof(T) == 4 ? 32 : 16, LDK = KT + 4;
    constexpr int PER = (BT * KT) / kThreads;      // staged elements per thread and operand
    typedef typename MfmaOp<T>::acc_t acc_t;
    __shared__ __align__(16) T As[BT * LDK];
    __shared__ __align__(16) T Bs[BT * LDK];
    int lo = 0, hi = njobs - 1;
    const int bid = blockIdx.x;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (jobs[mid].tile_base <= bid) lo = mid; else hi = mid - 1;
    }
    const GemmJob g = jobs[lo];
    int local = bid - g.tile_base;
    int ti, tj;
    if (g.lower) {
        tj = 0;
        int cnt = g.tiles_m;
        while (local >= cnt) { local -= cnt; ++tj; cnt = g.tiles_m - tj; }
        ti = tj + local;
    } else {
        tj = local / g.tiles_m;
        ti = local % g.tiles_m;
    }
    const T* __restrict__ A = L + g.a_off;
    const T* __restrict__ B = L + g.b_off;
    T* __restrict__ C = (g.c_in_arena ? U : L) + g.c_off;
    const int row0 = ti * BT, col0 = tj * BT;
    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    const int wr = (wave >> 1) * 64, wc = (wave & 1) * 64;
    const int li = lane & 15, lg = lane >> 4;
    // staging coordinates: element s of this thread is row r = (tid + 256 s) % 128,
    // k-offset kk = (tid + 256 s) / 128
    const int r_st = tid & (BT - 1);
    const int ia = row0 + r_st, jb = col0 + r_st;
    const bool a_ok = ia < g.M, b_ok = jb < g.N;
    const T* __restrict__ Ap = A + ia;
    const T* __restrict__ Bp = B + jb;
    T ra[PER], rb[PER];
    auto fetch = [&](int k0) {
#pragma unroll
        for (int s = 0; s < PER; ++s) {
            const int k = k0 + (tid >> 7) + 2 * s;
            ra[s] = (a_ok && k < g.K) ? Ap[(int64_t)k * g.lda] : T(0);
            rb[s] = (b_ok && k < g.K) ? Bp[(int64_t)k * g.ldb] : T(0);
        }
    };
    acc_t acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int v = 0; v < 4; ++v) acc[a][b][v] = T(0);
    fetch(0);
    for (int k0 = 0; k0 < g.K; k0 += KT) {
#pragma unroll
        for (int s = 0; s < PER; ++s) {
            const int kk = (tid >> 7) + 2 * s;
            As[r_st * LDK + kk] = ra[s];
            Bs[r_st * LDK + kk] = rb[s];
        }
        __syncthreads();
        if (k0 + KT < g.K) fetch(k0 + KT);
#pragma unroll
        for (int kq = 0; kq < KT; kq += 16) {
            T av[4][4], bv[4][4];   // [block][t]
#pragma unroll
            for (int rb2 = 0; rb2 < 4; ++rb2)
#pragma unroll
                for (int t = 0; t < 4; ++t) av[rb2][t] = As[(wr + rb2 * 16 + li) * LDK + kq + 4 * lg + t];
#pragma unroll
            for (int cb = 0; cb < 4; ++cb)
#pragma unroll
                for (int t = 0; t < 4; ++t) bv[cb][t] = Bs[(wc + cb * 16 + li) * LDK + kq + 4 * lg + t];
#pragma unroll
            for (int t = 0; t < 4; ++t)
#pragma unroll
                for (int rb2 = 0; rb2 < 4; ++rb2)
#pragma unroll
                    for (int cb = 0; cb < 4; ++cb)
                        acc[rb2][cb] = MfmaOp<T>::mma(bv[cb][t], av[rb2][t], acc[rb2][cb]);
        }
        __syncthreads();
    }
    // acc[rb][cb][v] = C(row0 + wr + rb*16 + li, col0 + wc + cb*16 + row(lg, v))
#pragma unroll
    for (int cb = 0; cb < 4; ++cb)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int j = col0 + wc + cb * 16 + MfmaOp<T>::row(lg, v);
            if (j >= g.N) continue;
#pragma unroll
            for (int rb2 = 0; rb2 < 4; ++rb2) {
                const int i = row0 + wr + rb2 * 16 + li;
                if (i >= g.M || (g.lower && i < j)) continue;
                T* cp = C + i + (int64_t)j * g.ldc;
                *cp = g.beta0 ? -acc[rb2][cb][v] : *cp - acc[rb2][cb][v];
            }
        }
}

// ------------------------------------------------------------------ solves
__global__ void permute_in_kernel(const double* __restrict__ b, const int* __restrict__ perm,
                                  double* __restrict__ xp, int n) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) xp[k] = b[perm[k]];
}
__global__ void permute_out_kernel(const double* __restrict__ xp, const int* __restrict__ perm,
                                   double* __restrict__ x, int n) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) x[perm[k]] = xp[k];
}

// Forward substitution L y = x over one level, one workgroup per front,
// multifrontal form of cs_lsolve (cs.cpp:1086-1101): gather the children's
// update vectors through their relative maps, then per 64-column diagonal
// block y = inv(L_D) x (stored inverse, a parallel 64x64 product) and update
// the rows below.
template <typename T>
__global__ __launch_bounds__(kThreads) void fwd_kernel(const int* __restrict__ list,
                                                       const FrontDesc* __restrict__ fronts,
                                                       const int* __restrict__ kids, const int* __restrict__ rel,
                                                       const int* __restrict__ q, const T* __restrict__ L,
                                                       const T* __restrict__ inv, double* __restrict__ xp,
                                                       double* __restrict__ Y, double* __restrict__ su,
                                                       const int64_t* __restrict__ suofs) {
    __shared__ double W[64 * 65];
    __shared__ double vb[64];
    __shared__ double yb[64];
    const int f = list[blockIdx.x];
    const FrontDesc d = fronts[f];
    const int n = d.n, c = d.c, m = d.m, k = d.k;
    const T* __restrict__ P = L + d.lofs;
    double* __restrict__ u = su + suofs[f];
    for (int t = threadIdx.x; t < m; t += kThreads) u[t] = 0.0;
    __syncthreads();
    for (int t = d.kid_begin; t < d.kid_end; ++t) {
        const int ch = kids[t];
        const FrontDesc cd = fronts[ch];
        const int* __restrict__ r = rel + cd.relofs;
        const double* __restrict__ cu = su + suofs[ch];
        for (int i = threadIdx.x; i < cd.m; i += kThreads) {
            const int pos = r[i];
            if (pos < c) xp[k + pos] += cu[i];
            else u[pos - c] += cu[i];
        }
        __syncthreads();
    }
    for (int j0 = 0; j0 < c; j0 += 64) {
        const int b = min(64, c - j0);
        const T* __restrict__ w = inv + d.dinv + (int64_t)(j0 / 64) * 4096;
        for (int t = threadIdx.x; t < 64 * b; t += kThreads) {
            int l = t & 63, r = t >> 6;
            if (l < b) W[l + r * 65] = to_d(w[l + 64 * r]);
        }
        for (int t = threadIdx.x; t < b; t += kThreads) vb[t] = xp[k + j0 + t];
        __syncthreads();
        if (threadIdx.x < b) {
            const int r = threadIdx.x;
            double y = 0.0;
#pragma unroll 8
            for (int l = 0; l <= r; ++l) y += W[l + r * 65] * vb[l];
            yb[r] = y;
            Y[k + j0 + r] = y;
        }
        __syncthreads();
        for (int i = j0 + b + threadIdx.x; i < n; i += kThreads) {
            double acc = 0.0;
#pragma unroll 16
            for (int jj = 0; jj < b; ++jj) acc += to_d(P[i + (int64_t)(j0 + jj) * n]) * yb[jj];
            if (i < c) xp[k + i] -= acc;
            else u[i - c] -= acc;
        }
        __syncthreads();
    }
}

// Backward substitution L^T x = y over one level (cs_ltsolve, cs.cpp:1104-1119),
// one workgroup per front, diagonal blocks from the last to the first: dot
// products with the already solved rows below (read through the row list
// q), then x = inv(L_D)^T v with the stored inverse.
template <typename T>
__global__ __launch_bounds__(kThreads) void bwd_kernel(const int* __restrict__ list,
                                                       const FrontDesc* __restrict__ fronts,
                                                       const int* __restrict__ qall, const T* __restrict__ L,
                                                       const T* __restrict__ inv, double* __restrict__ xp,
                                                       const double* __restrict__ Y) {
    __shared__ double vb[64];
    const FrontDesc d = fronts[list[blockIdx.x]];
    const int n = d.n, c = d.c, k = d.k;
    const T* __restrict__ P = L + d.lofs;
    const int* __restrict__ q = qall + d.qofs;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int nblk = (c + 63) / 64;
    for (int blk = nblk - 1; blk >= 0; --blk) {
        const int j0 = blk * 64, j1 = min(c, j0 + 64), b = j1 - j0;
        for (int j = j0 + wave; j < j1; j += kThreads / 64) {
            double s = 0.0;
            const T* col = P + (int64_t)j * n;
#pragma unroll 4
            for (int i = j1 + lane; i < n; i += 64) s += to_d(col[i]) * xp[q[i]];
            s = wave_sum(s);
            if (lane == 0) vb[j - j0] = Y[k + j] - s;
        }
        __syncthreads();
        if (threadIdx.x < b) {
            const int i = threadIdx.x;
            const T* __restrict__ w = inv + d.dinv + (int64_t)blk * 4096;
            double x = 0.0;
#pragma unroll 8
            for (int r = i; r < b; ++r) x += to_d(w[i + 64 * r]) * vb[r];
            xp[k + j0 + i] = x;
        }
        __syncthreads();
    }
}

// ------------------------------------------------- solves of large fronts
// Fronts with more than kSolveBW pivots: forward / backward substitution 256
// pivot columns at a time, the rows below each block handled by many
// workgroups (row chunks of 256).

// Forward, step 0: clear the front's update vector and gather its
// children's update vectors (the first phase of fwd_kernel).
__global__ __launch_bounds__(kThreads) void fwd_gather_kernel(const int* __restrict__ list,
                                                              const FrontDesc* __restrict__ fronts,
                                                              const int* __restrict__ kids,
                                                              const int* __restrict__ rel,
                                                              double* __restrict__ xp, double* __restrict__ su,
                                                              const int64_t* __restrict__ suofs) {
    const int f = list[blockIdx.x];
    const FrontDesc d = fronts[f];
    double* __restrict__ u = su + suofs[f];
    for (int t = threadIdx.x; t < d.m; t += kThreads) u[t] = 0.0;
    __syncthreads();
    for (int t = d.kid_begin; t < d.kid_end; ++t) {
        const int ch = kids[t];
        const FrontDesc cd = fronts[ch];
        const int* __restrict__ r = rel + cd.relofs;
        const double* __restrict__ cu = su + suofs[ch];
        for (int i = threadIdx.x; i < cd.m; i += kThreads) {
            const int pos = r[i];
            if (pos < d.c) xp[d.k + pos] += cu[i];
            else u[pos - d.c] += cu[i];
        }
        __syncthreads();
    }
}

// Forward, one 256-column block of a large front per launch.  Every
// workgroup solves the block's own triangle (four 64-column sub-blocks:
// y_s = inv(L_s) v_s with the stored inverse, then the later sub-blocks'
// right-hand sides are updated), the `first` one stores y in Y, and each
// updates its 256 rows [a, a+256) below the block.  v is final on entry:
// earlier blocks were applied by earlier launches.
template <typename T>
__global__ __launch_bounds__(kThreads) void fwd_blk_kernel(const SolveItem* __restrict__ items,
                                                           const FrontDesc* __restrict__ fronts,
                                                           const T* __restrict__ L, const T* __restrict__ inv,
                                                           double* __restrict__ xp, double* __restrict__ Y,
                                                           double* __restrict__ su,
                                                           const int64_t* __restrict__ suofs) {
    __shared__ double W[64 * 65];
    __shared__ double vb[kSolveBW];
    const SolveItem it = items[blockIdx.x];
    const FrontDesc d = fronts[it.front];
    const int n = d.n, j0 = it.j0, bw = min(kSolveBW, d.c - j0);
    const T* __restrict__ P = L + d.lofs;
    const int tid = threadIdx.x;
    if (tid < bw) vb[tid] = xp[d.k + j0 + tid];
    for (int s0 = 0; s0 < bw; s0 += 64) {
        const int b = min(64, bw - s0);
        const T* __restrict__ w = inv + d.dinv + (int64_t)((j0 + s0) / 64) * 4096;
        for (int t = tid; t < 64 * b; t += kThreads) {
            const int l = t & 63, r = t >> 6;
            if (l < b) W[l + r * 65] = to_d(w[l + 64 * r]);
        }
        __syncthreads();
        double y = 0.0;
        if (tid < b) {
#pragma unroll 8
            for (int l = 0; l <= tid; ++l) y += W[l + tid * 65] * vb[s0 + l];
        }
        __syncthreads();
        if (tid < b) vb[s0 + tid] = y;
        __syncthreads();
        const int row = s0 + b + tid;            // later rows of this block
        if (row < bw) {
            const T* __restrict__ pr = P + (j0 + row) + (int64_t)(j0 + s0) * n;
            double acc = 0.0;
#pragma unroll 16
            for (int l = 0; l < b; ++l) acc += to_d(pr[(int64_t)l * n]) * vb[s0 + l];
            vb[row] -= acc;
        }
        __syncthreads();
    }
    if (it.first && tid < bw) Y[d.k + j0 + tid] = vb[tid];
    const int i = it.a + tid;
    if (i < n) {
        const T* __restrict__ pr = P + i + (int64_t)j0 * n;
        double acc = 0.0;
#pragma unroll 16
        for (int j = 0; j < bw; ++j) acc += to_d(pr[(int64_t)j * n]) * vb[j];
        if (i < d.c) xp[d.k + i] -= acc;
        else su[suofs[it.front] + i - d.c] -= acc;
    }
}

// Backward, one 256-column block of a large front per launch.  Each
// workgroup forms partial dot products of the block's columns with x over
// its 256 rows below the block (64x64 tiles staged through LDS); the last
// workgroup of the group to arrive (atomic counter; none waits) reduces the
// partials in slot order, solves the block's own triangle from the last
// 64-column sub-block (dots with the block rows already solved, then
// inv(L_s)^T), stores x and resets the counter for the next launch.
template <typename T>
__global__ __launch_bounds__(kThreads) void bwd_blk_kernel(const SolveItem* __restrict__ items,
                                                           const FrontDesc* __restrict__ fronts,
                                                           const int* __restrict__ qall, const T* __restrict__ L,
                                                           const T* __restrict__ inv, double* __restrict__ xp,
                                                           const double* __restrict__ Y,
                                                           double* __restrict__ partial,
                                                           int* __restrict__ counters) {
    __shared__ T tile[64 * 65];
    __shared__ double xq[64];
    __shared__ double red[4][kSolveBW];
    __shared__ double vb[kSolveBW];
    __shared__ int last;
    const SolveItem it = items[blockIdx.x];
    const FrontDesc d = fronts[it.front];
    const int n = d.n, j0 = it.j0, bw = min(kSolveBW, d.c - j0);
    const T* __restrict__ P = L + d.lofs;
    const int tid = threadIdx.x, r = tid & 63, cq = tid >> 6;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int r0 = it.a; r0 < min(it.a + kSolveBW, n); r0 += 64) {
        const int rows = min(64, n - r0);
        __syncthreads();
        if (tid < 64) xq[r] = r < rows ? xp[qall[d.qofs + r0 + r]] : 0.0;
        for (int c4 = 0; c4 < 4; ++c4) {
            const int cbase = c4 * 64;
            if (cbase >= bw) break;
            __syncthreads();
#pragma unroll
            for (int cc = cq; cc < 64; cc += 4)
                tile[cc * 65 + r] = (r < rows && cbase + cc < bw) ? P[(r0 + r) + (int64_t)(j0 + cbase + cc) * n] : T(0);
            __syncthreads();
            double sdot = 0.0;
#pragma unroll
            for (int rr = cq * 16; rr < cq * 16 + 16; ++rr) sdot += to_d(tile[r * 65 + rr]) * xq[rr];
            acc[c4] += sdot;
        }
    }
    __syncthreads();
#pragma unroll
    for (int c4 = 0; c4 < 4; ++c4) red[cq][c4 * 64 + r] = acc[c4];
    __syncthreads();
    if (tid < bw) partial[(int64_t)it.slot * kSolveBW + tid] = red[0][tid] + red[1][tid] + red[2][tid] + red[3][tid];
    __threadfence();
    __syncthreads();
    if (tid == 0) last = atomicAdd(&counters[it.group], 1) == it.gcount - 1;
    __syncthreads();
    if (!last) return;
    __threadfence();
    if (tid < bw) {
        double v = Y[d.k + j0 + tid];
#pragma unroll 8
        for (int sl = 0; sl < it.gcount; ++sl) v -= partial[(int64_t)(it.gfirst + sl) * kSolveBW + tid];
        vb[tid] = v;
    }
    __syncthreads();
    const int nsub = (bw + 63) / 64;
    for (int sb = nsub - 1; sb >= 0; --sb) {
        const int s0 = sb * 64, s1 = min(s0 + 64, bw), b = s1 - s0;
        // v_j -= sum over solved block rows [s1, bw) of L(row, j) x_row, 4 threads per column
        double part = 0.0;
        if (r < b) {
            const T* __restrict__ col = P + (int64_t)(j0 + s0 + r) * n + j0;
#pragma unroll 4
            for (int row = s1 + cq; row < bw; row += 4) part += to_d(col[row]) * vb[row];
        }
        red[cq][r] = part;
        __syncthreads();
        if (tid < b) vb[s0 + tid] -= red[0][tid] + red[1][tid] + red[2][tid] + red[3][tid];
        __syncthreads();
        double x = 0.0;
        if (tid < b) {
            const T* __restrict__ w = inv + d.dinv + (int64_t)((j0 + s0) / 64) * 4096;
#pragma unroll 8
            for (int rr = tid; rr < b; ++rr) x += to_d(w[tid + 64 * rr]) * vb[s0 + rr];
        }
        __syncthreads();
        if (tid < b) {
            vb[s0 + tid] = x;
            xp[d.k + j0 + s0 + tid] = x;
        }
        __syncthreads();
    }
    if (tid == 0) counters[it.group] = 0;
}

// -------------------------------------------------------------- refinement
__global__ void gather_kernel(const double* __restrict__ vals, const int* __restrict__ src,
                              double* __restrict__ dst, int nnz) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < nnz) dst[t] = vals[src[t]];
}

// r = b - A x, one row per thread, FP64 (cs_gaxpy, cs.cpp:1069-1083, on the
// full symmetric matrix)
__global__ void residual_kernel(const int* __restrict__ rp, const int* __restrict__ ci,
                                const double* __restrict__ av, const double* __restrict__ x,
                                const double* __restrict__ b, double* __restrict__ r, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    for (int p = rp[i]; p < rp[i + 1]; ++p) s += av[p] * x[ci[p]];
    r[i] = b[i] - s;
}

__global__ void rowabs_kernel(const int* __restrict__ rp, const double* __restrict__ av,
                              double* __restrict__ out, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    for (int p = rp[i]; p < rp[i + 1]; ++p) s += fabs(av[p]);
    out[i] = s;
}

// max |v| via integer max on the bit patterns of non-negative doubles
__global__ __launch_bounds__(kThreads) void maxabs_kernel(const double* __restrict__ v, int n,
                                                          unsigned long long* __restrict__ out) {
    __shared__ double red[kThreads / 64];
    double m = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double a = fabs(v[i]);
        if (a > m || a != a) m = (a != a) ? __builtin_huge_val() : a;
    }
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_down(m, off, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kThreads / 64; ++w) m = fmax(m, red[w]);
        atomicMax(out, (unsigned long long)__double_as_longlong(m));
    }
}

__global__ void axpy_kernel(double* __restrict__ x, const double* __restrict__ d, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] += d[i];
}

// partial sums of log(L_jj), fixed reduction tree per block
template <typename T>
__global__ __launch_bounds__(kThreads) void logdiag_kernel(const int64_t* __restrict__ diag_off,
                                                           const T* __restrict__ L, int n,
                                                           double* __restrict__ partial) {
    __shared__ double red[kThreads];
    double s = 0.0;
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x)
        s += log(to_d(L[diag_off[j]]));
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = kThreads / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

inline int blocks_for(int64_t n, int per = kThreads) { return (int)((n + per - 1) / per); }

}  // namespace

template <typename T> void launch_fill(T* p, int64_t n, T v, hipStream_t s) {
    if (n <= 0) return;
    int64_t blocks = (n + kThreads - 1) / kThreads;
    if (blocks > 65536) blocks = 65536;
    fill_kernel<T><<<(int)blocks, kThreads, 0, s>>>(p, n, v);
}
template <typename T> void launch_scatter_a(T* L, const double* ax, const int64_t* dst, int nnz, hipStream_t s) {
    if (nnz > 0) scatter_a_kernel<T><<<blocks_for(nnz), kThreads, 0, s>>>(L, ax, dst, nnz);
}
template <typename T>
void launch_extend_add(const EaItem* items, int count, const FrontDesc* fronts, const int* rel, T* L, T* U,
                       hipStream_t s) {
    if (count > 0) extend_add_kernel<T><<<count, kThreads, 0, s>>>(items, fronts, rel, L, U);
}
template <typename T>
void launch_panel(const int* list, int count, const FrontDesc* fronts, T* L, int* info, int lds_bytes,
                  hipStream_t s) {
    if (count <= 0) return;
    static bool attr_set = false;
    if (!attr_set) {
        (void)hipFuncSetAttribute((const void*)panel_kernel<T>, hipFuncAttributeMaxDynamicSharedMemorySize, 65536);
        attr_set = true;
    }
    panel_kernel<T><<<count, kThreads, lds_bytes, s>>>(list, fronts, L, info);
}
template <typename T>
void launch_diag(const DiagJob* jobs, int count, T* L, T* inv, int* info, hipStream_t s) {
    if (count > 0) diag_kernel<T><<<count, kThreads, 0, s>>>(jobs, L, inv, info);
}
template <typename T>
void launch_inv(const InvJob* jobs, int count, const FrontDesc* fronts, const T* L, T* inv, hipStream_t s) {
    if (count > 0) inv_kernel<T><<<count, 64, 0, s>>>(jobs, fronts, L, inv);
}
template <typename T>
void launch_trsm(const TrsmJob* jobs, int njobs, int nblocks, T* L, hipStream_t s) {
    if (nblocks > 0) trsm_kernel<T><<<nblocks, kThreads, 0, s>>>(jobs, njobs, L);
}
template <typename T>
void launch_gemm(const GemmJob* jobs, int njobs, int nblocks, int tile, T* L, T* U, hipStream_t s) {
    if (nblocks <= 0) return;
    if (tile == 128) gemm_mfma_kernel<T><<<nblocks, kThreads, 0, s>>>(jobs, njobs, L, U);
    else gemm_nt_kernel<T><<<nblocks, kThreads, 0, s>>>(jobs, njobs, L, U);
}

void launch_permute_in(const double* b, const int* perm, double* xp, int n, hipStream_t s) {
    permute_in_kernel<<<blocks_for(n), kThreads, 0, s>>>(b, perm, xp, n);
}
void launch_permute_out(const double* xp, const int* perm, double* x, int n, hipStream_t s) {
    permute_out_kernel<<<blocks_for(n), kThreads, 0, s>>>(xp, perm, x, n);
}
template <typename T>
void launch_fwd(const int* list, int count, const FrontDesc* fronts, const int* kids, const int* rel,
                const int* q, const T* L, const T* inv, double* xp, double* Y, double* su, const int64_t* suofs,
                hipStream_t s) {
    if (count > 0) fwd_kernel<T><<<count, kThreads, 0, s>>>(list, fronts, kids, rel, q, L, inv, xp, Y, su, suofs);
}
template <typename T>
void launch_bwd(const int* list, int count, const FrontDesc* fronts, const int* q, const T* L, const T* inv,
                double* xp, const double* Y, hipStream_t s) {
    if (count > 0) bwd_kernel<T><<<count, kThreads, 0, s>>>(list, fronts, q, L, inv, xp, Y);
}

void launch_fwd_gather(const int* list, int count, const FrontDesc* fronts, const int* kids, const int* rel,
                       double* xp, double* su, const int64_t* suofs, hipStream_t s) {
    if (count > 0) fwd_gather_kernel<<<count, kThreads, 0, s>>>(list, fronts, kids, rel, xp, su, suofs);
}
template <typename T>
void launch_fwd_blk(const SolveItem* items, int count, const FrontDesc* fronts, const T* L, const T* inv,
                    double* xp, double* Y, double* su, const int64_t* suofs, hipStream_t s) {
    if (count > 0) fwd_blk_kernel<T><<<count, kThreads, 0, s>>>(items, fronts, L, inv, xp, Y, su, suofs);
}
template <typename T>
void launch_bwd_blk(const SolveItem* items, int count, const FrontDesc* fronts, const int* q, const T* L,
                    const T* inv, double* xp, const double* Y, double* partial, int* counters, hipStream_t s) {
    if (count > 0)
        bwd_blk_kernel<T><<<count, kThreads, 0, s>>>(items, fronts, q, L, inv, xp, Y, partial, counters);
}

void launch_gather(const double* src_vals, const int* src, double* dst, int nnz, hipStream_t s) {
    if (nnz > 0) gather_kernel<<<blocks_for(nnz), kThreads, 0, s>>>(src_vals, src, dst, nnz);
}
void launch_residual(const int* rp, const int* ci, const double* av, const double* x, const double* b, double* r,
                     int n, hipStream_t s) {
    residual_kernel<<<blocks_for(n), kThreads, 0, s>>>(rp, ci, av, x, b, r, n);
}
void launch_rowabs(const int* rp, const double* av, double* out, int n, hipStream_t s) {
    rowabs_kernel<<<blocks_for(n), kThreads, 0, s>>>(rp, av, out, n);
}
void launch_maxabs(const double* v, int n, unsigned long long* out, hipStream_t s) {
    int blocks = blocks_for(n);
    if (blocks > 1024) blocks = 1024;
    if (blocks < 1) blocks = 1;
    maxabs_kernel<<<blocks, kThreads, 0, s>>>(v, n, out);
}
void launch_axpy(double* x, const double* d, int n, hipStream_t s) {
    axpy_kernel<<<blocks_for(n), kThreads, 0, s>>>(x, d, n);
}
template <typename T>
void launch_logdiag(const int64_t* diag_off, const T* L, int n, double* partial, int nparts, hipStream_t s) {
    logdiag_kernel<T><<<nparts, kThreads, 0, s>>>(diag_off, L, n, partial);
}

#define MDGPU_INSTANTIATE(T)                                                                                   \
    template void launch_fill<T>(T*, int64_t, T, hipStream_t);                                                 \
    template void launch_scatter_a<T>(T*, const double*, const int64_t*, int, hipStream_t);                    \
    template void launch_extend_add<T>(const EaItem*, int, const FrontDesc*, const int*, T*, T*, hipStream_t); \
    template void launch_panel<T>(const int*, int, const FrontDesc*, T*, int*, int, hipStream_t);              \
    template void launch_diag<T>(const DiagJob*, int, T*, T*, int*, hipStream_t);                             \
    template void launch_inv<T>(const InvJob*, int, const FrontDesc*, const T*, T*, hipStream_t);             \
    template void launch_trsm<T>(const TrsmJob*, int, int, T*, hipStream_t);                                   \
    template void launch_gemm<T>(const GemmJob*, int, int, int, T*, T*, hipStream_t);                               \
    template void launch_fwd<T>(const int*, int, const FrontDesc*, const int*, const int*, const int*,         \
                                const T*, const T*, double*, double*, double*, const int64_t*, hipStream_t);   \
    template void launch_bwd<T>(const int*, int, const FrontDesc*, const int*, const T*, const T*, double*,    \
                                const double*, hipStream_t);                                                   \
    template void launch_logdiag<T>(const int64_t*, const T*, int, double*, int, hipStream_t);               \
    template void launch_fwd_blk<T>(const SolveItem*, int, const FrontDesc*, const T*, const T*, double*,   \
                                    double*, double*, const int64_t*, hipStream_t);                          \
    template void launch_bwd_blk<T>(const SolveItem*, int, const FrontDesc*, const int*, const T*,          \
                                    const T*, double*, const double*, double*, int*, hipStream_t);

MDGPU_INSTANTIATE(double)
MDGPU_INSTANTIATE(float)

}  // namespace mdgpu

namespace mdgpu {
template void launch_fill<int>(int*, int64_t, int, hipStream_t);
template void launch_fill<unsigned long long>(unsigned long long*, int64_t, unsigned long long, hipStream_t);
}  // namespace mdgpu

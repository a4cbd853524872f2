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
#include <type_traits>

#include "kernels.h"

namespace mdgpu {

namespace {

constexpr int kThreads = 256;

// ---------------------------------------------- vector types, buffer loads
typedef float f32x4_t __attribute__((ext_vector_type(4)));
typedef float f32x2_t __attribute__((ext_vector_type(2)));
typedef float f32x16_t __attribute__((ext_vector_type(16)));
typedef __attribute__((address_space(3))) void* lds_void_ptr;

// raw buffer resource (stride 0) over [base, base + bytes_left), clamped to
// the 32-bit record count; a load outside returns zero
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buffer_rsrc(const void* base, int64_t bytes_left) {
    const int64_t b = bytes_left < 0 ? 0 : (bytes_left > 0xFFFFFFFFll ? 0xFFFFFFFFll : bytes_left);
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)(uint32_t)b, 0x00020000);
}
__device__ __forceinline__ f32x4_t buffer_load4(__amdgpu_buffer_rsrc_t r, uint32_t byte_off) {
    return __builtin_bit_cast(f32x4_t, __builtin_amdgcn_raw_buffer_load_b128(r, byte_off, 0, 0));
}
// 16 bytes per lane from a buffer straight into LDS: lane l's bytes land at
// dst + 16 l (dst is the same for the whole wavefront)
__device__ __forceinline__ void buffer_load_lds16(__amdgpu_buffer_rsrc_t r, lds_void_ptr dst, uint32_t byte_off) {
    __builtin_amdgcn_raw_ptr_buffer_load_lds(r, dst, 16, byte_off, 0, 0, 0);
}
// s_waitcnt immediate that waits for all vector-memory operations (vmcnt 0)
// and leaves the LDS/scalar and export counters alone
constexpr int kWaitVmcnt0 = (0xF << 8) | (0x7 << 4);

__device__ inline double to_d(double v) { return v; }
__device__ inline double to_d(float v) { return (double)v; }

__device__ inline double wave_sum(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

// v[k] = p[k * stride] for k < 32, all 32 loads in flight before the first
// use: one moving pointer (no per-element 64-bit addresses to keep), and the
// empty asm makes every value live at once, so the compiler cannot put a
// wait after each load.
template <typename T>
__device__ __forceinline__ void load32_strided(T (&v)[32], const T* __restrict__ p, int64_t stride) {
#pragma unroll
    for (int k = 0; k < 32; ++k) {
        v[k] = *p;
        p += stride;
    }
#pragma unroll
    for (int k = 0; k < 32; ++k) asm volatile("" : "+v"(v[k]));
}

// N loads p[0], p[stride], ... with a moving pointer and no wait: the
// caller makes them wait together (keep_in_flight) after issuing all the
// loads it wants in flight.
template <int N, typename T>
__device__ __forceinline__ void issue_strided(T (&v)[N], const T* __restrict__ p, int64_t stride) {
#pragma unroll
    for (int k = 0; k < N; ++k) {
        v[k] = *p;
        p += stride;
    }
}
template <int N, typename T>
__device__ __forceinline__ void keep_in_flight(T (&v)[N]) {
#pragma unroll
    for (int k = 0; k < N; ++k) asm volatile("" : "+v"(v[k]));
}

// ------------------------------------------------------------ chained solves
// The values of Y (forward) and xp (backward) that a chain launch waits for
// hold kChainEmpty until the workgroup of their block stores them: a
// signalling NaN, which arithmetic never produces (operations on a NaN give
// a quiet NaN).
constexpr unsigned long long kChainEmpty = kChainEmptyBits;

// *p once it no longer holds kChainEmpty, read with agent-scope atomic
// loads (another workgroup, possibly on another XCD, stores it with
// chain_store).  After `spin` polls (kChainSpin, about a second) the wait
// ends with a quiet NaN, which then shows in every result that depends on
// the value, and sets bit kSolveTimedOut of *status (read by
// mdgpu_solve_status / mdgpu_refine).
__device__ __forceinline__ double chain_wait(const double* p, int* status, int spin) {
    const unsigned long long* q = reinterpret_cast<const unsigned long long*>(p);
    unsigned long long v = __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    for (int s = 0; v == kChainEmpty; ++s) {
        if (s >= spin) {
            __hip_atomic_fetch_or(status, kSolveTimedOut, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return __builtin_nan("");
        }
        __builtin_amdgcn_s_sleep(2);
        v = __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    return __longlong_as_double((long long)v);
}
__device__ __forceinline__ void chain_store(double* p, double v) {
    __hip_atomic_store(reinterpret_cast<unsigned long long*>(p), (unsigned long long)__double_as_longlong(v),
                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// The workgroup's item index: a ticket from the launch's counter, so every
// item it waits for belongs to a workgroup that has already started.
__device__ __forceinline__ int chain_ticket(int* ctr, int* sh) {
    if (threadIdx.x == 0) *sh = atomicAdd(ctr, 1);
    __syncthreads();
    return __builtin_amdgcn_readfirstlane(*sh);   // uniform: the item's fields go to scalar registers
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
// parent(rel[i], rel[j]) += U_child(i, j) over the lower triangle
// (ExtendAdd_CPU / ExtendAdd_GPU_device, mdgpu.cu:748-788), parent side: one
// wavefront per parent column (EaItem), adding its contributors -- child
// columns, in child order -- one after the other.  Lanes run over the child
// rows (coalesced reads; writes follow the increasing relative map), eight
// rows per lane in flight.  Two contributors can share target rows, so a
// workgroup-scope fence (every wavefront of the workgroup shares one L1)
// separates each contributor's stores from the next one's loads.
template <typename T>
__global__ __launch_bounds__(kThreads) void extend_add_kernel(const EaItem* __restrict__ items, int count,
                                                              const EaContrib* __restrict__ contrib,
                                                              const int* __restrict__ rel, T* __restrict__ L,
                                                              T* __restrict__ U) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int k = blockIdx.x * (kThreads / 64) + wave;
    if (k >= count) return;
    const EaItem it = items[k];
    T* __restrict__ dcol = (it.in_u ? U : L) + it.dst;
    for (int t = it.cbeg; t < it.cend; ++t) {
        const EaContrib e = contrib[t];
        const int* __restrict__ r = rel + e.rel;
        const T* __restrict__ scol = U + e.src;
        const int m = e.m, j = e.j;
        // sources, targets and old target values are all loaded before any
        // store (the relative map is injective, so the targets are distinct);
        // rows past the column end load row m - 1 again and are not stored;
        // the empty asm keeps the loads from waiting one by one behind the
        // previous stores
        for (int e0 = j + lane; e0 < m; e0 += 8 * 64) {
            int ri[8];
            T sv[8], dv[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int i = min(e0 + 64 * u, m - 1);
                ri[u] = r[i] - it.rowoff;
                sv[u] = scol[i];
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) dv[u] = dcol[ri[u]];
#pragma unroll
            for (int u = 0; u < 8; ++u) asm volatile("" : "+v"(dv[u]), "+v"(sv[u]));
#pragma unroll
            for (int u = 0; u < 8; ++u)
                if (e0 + 64 * u < m) dcol[ri[u]] = dv[u] + sv[u];
        }
        if (t + 1 < it.cend) __threadfence_block();
    }
}

// 1/sqrt(d) of a pivot d > 0 from the hardware reciprocal square root:
// v_rsq_f32 is within 1 ulp; v_rsq_f64 gives about 23 correct bits, and two
// Newton steps y <- y + y (1 - d y^2) / 2 bring it to double precision.  The
// pivot chain of a column is then a handful of dependent instructions, not a
// correctly rounded sqrt followed by a correctly rounded division.
__device__ inline float pivot_rsqrt(float d) { return __builtin_amdgcn_rsqf(d); }
__device__ inline double pivot_rsqrt(double d) {
    double y = __builtin_amdgcn_rsq(d);
#pragma unroll
    for (int it = 0; it < 2; ++it) {
        const double r = fma(-d * y, y, 1.0);
        y = fma(0.5 * y, r, y);
    }
    return y;
}

// Pivot-growth threshold of the reduced-precision factorizations: a
// diagonal block (or a front factored wholly in LDS) whose pivots satisfy
// L_jj^2 < 2^-12 A_jj (more than 12 of FP32's 24 bits cancelled), or that
// has a non-positive pivot, is factored again in FP64 from its assembled
// values.
constexpr double kPivotGrowthMin = 1.0 / 4096.0;

// ------------------------------------------------------ dense: LDS panels
// Partial Cholesky of a whole panel P (n x c, column-major, leading
// dimension n) held in LDS: for each pivot column, sqrt, scale, rank-1
// update of the remaining pivot columns (the unblocked scal+syr loop of
// FactorAux_CPU_3, mdgpu.cu:1318-1326, applied to the full n_f rows, which
// also performs the trsm of :1330).  Called by all threads; ends with a
// barrier.  check: reduced precision, where a pivot that fails the
// pivot-growth check against the assembled diagonal a0 sets *weak (the
// caller then falls back to FP64) instead of being reported; otherwise a
// non-positive pivot is reported in *info as its global column.
template <typename U>
__device__ __forceinline__ void panel_chol(U* P, int n, int c, int k, bool check, const U* a0, int* weak, U* pivinv,
                                           int* info) {
    for (int j = 0; j < c; ++j) {
        if (threadIdx.x == 0) {
            const U dj = P[j + j * n];
            if (check) {
                if (!(dj > U(0)) || !((double)dj >= kPivotGrowthMin * (double)a0[j])) *weak = 1;
            } else if (!(dj > U(0))) {
                atomicMin(info, k + j);
            }
            const U is = pivot_rsqrt(dj);
            P[j + j * n] = dj * is;
            *pivinv = is;
        }
        __syncthreads();
        const U s = *pivinv;
        for (int i = j + 1 + threadIdx.x; i < n; i += kThreads) P[i + j * n] *= s;
        __syncthreads();
        for (int i = j + 1 + threadIdx.x; i < n; i += kThreads) {
            const U lij = P[i + j * n];
            const int cmax = min(i, c - 1);
#pragma unroll 4
            for (int col = j + 1; col <= cmax; ++col) P[i + col * n] -= lij * P[col + j * n];
        }
        __syncthreads();
    }
}

// One workgroup per front of the step's list.  float: a front that fails
// the pivot-growth check keeps its assembled panel in L and goes on the
// fallback list (fb_list[0 .. *fb_cnt)) for panel_fb_kernel.
template <typename T>
__global__ __launch_bounds__(kThreads) void panel_kernel(const int* __restrict__ list,
                                                         const FrontDesc* __restrict__ fronts,
                                                         T* __restrict__ L, int* __restrict__ info,
                                                         int* __restrict__ fb_cnt, int* __restrict__ fb_list) {
    constexpr bool kCheck = std::is_same<T, float>::value;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T* P = reinterpret_cast<T*>(smem_raw);
    __shared__ T pivinv;
    __shared__ T a0[kCheck ? kPanelMaxPivots : 1];
    __shared__ int weak;
    const int f = list[blockIdx.x];
    const FrontDesc d = fronts[f];
    const int n = d.n, c = d.c;
    T* __restrict__ g = L + d.lofs;
    const int tot = n * c;
    for (int t = threadIdx.x; t < tot; t += kThreads) P[t] = g[t];
    if (threadIdx.x == 0) weak = 0;
    __syncthreads();
    if constexpr (kCheck) {
        if (threadIdx.x < c) a0[threadIdx.x] = P[threadIdx.x * (n + 1)];   // assembled diagonal
        __syncthreads();
    }
    panel_chol<T>(P, n, c, d.k, kCheck, a0, &weak, &pivinv, info);
    if (kCheck && weak) {
        if (threadIdx.x == 0) fb_list[atomicAdd(fb_cnt, 1)] = f;
        return;
    }
    for (int t = threadIdx.x; t < tot; t += kThreads) g[t] = P[t];
}

// FP64 refactorization of the fronts panel_kernel<float> put on the
// fallback list: each workgroup takes list entries blockIdx.x,
// blockIdx.x + gridDim.x, ..., loads the assembled FP32 panel into LDS as
// FP64, factors it (a non-positive pivot is reported in *info) and stores
// the result rounded to FP32.  With an empty list every workgroup exits at
// once.
__global__ __launch_bounds__(kThreads) void panel_fb_kernel(const int* __restrict__ fb_cnt,
                                                            const int* __restrict__ fb_list,
                                                            const FrontDesc* __restrict__ fronts,
                                                            float* __restrict__ L, int* __restrict__ info,
                                                            int* __restrict__ fallbacks) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    double* P = reinterpret_cast<double*>(smem_raw);
    __shared__ double pivinv;
    const int cnt = *fb_cnt;
    for (int t = blockIdx.x; t < cnt; t += gridDim.x) {
        const FrontDesc d = fronts[fb_list[t]];
        const int n = d.n, c = d.c, tot = n * c;
        float* __restrict__ g = L + d.lofs;
        for (int e = threadIdx.x; e < tot; e += kThreads) P[e] = (double)g[e];
        __syncthreads();
        panel_chol<double>(P, n, c, d.k, false, nullptr, nullptr, &pivinv, info);
        for (int e = threadIdx.x; e < tot; e += kThreads) g[e] = (float)P[e];
        if (threadIdx.x == 0) atomicAdd(fallbacks, 1);
        __syncthreads();   // P is reused by the next entry
    }
}

// MFMA operands of the 16x16x4 products (v_mfma_{f32,f64}_16x16x4):
// mma(a, b, c) = c + A B with A 16x4 supplied by lane (li, lg) = (lane & 15,
// lane >> 4) as A(li, lg) and B 4x16 supplied as B(lg, li).  The
// contraction index of one instruction is free to relabel as long as both
// operands agree.
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

// 8 bf16 values: one 16-byte MFMA operand (8 consecutive k of one row)
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

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

// Trailing update of the 64x64 Cholesky after the panel of columns
// [c0, c0 + 16): As(i, j) -= sum_t As(i, c0 + t) As(j, c0 + t) on the 16x16
// tiles (I, J), I >= J, of rows/cols [c0 + 16, 64) with jlo <= J < jhi,
// one tile (four 16x16x4 MFMA products) per wavefront of [wfirst, wfirst +
// wcount) in turn.  The upper halves of the diagonal tiles are updated too;
// they are never read.
template <typename U>
__device__ inline void chol64_trailing(U* As, int c0, int wave, int lane, int jlo, int jhi, int wfirst,
                                       int wcount) {
    typedef typename MfmaOp<U>::acc_t acc_t;
    const int nt = (48 - c0) / 16, li = lane & 15, lg = lane >> 4;
    if (wave < wfirst || wave >= wfirst + wcount) return;
    int t = 0;
    for (int I = 0; I < nt; ++I)
        for (int J = jlo; J <= I && J < jhi; ++J, ++t) {
            if (t % wcount != wave - wfirst) continue;
            const int i0 = c0 + 16 + 16 * I, j0 = c0 + 16 + 16 * J;
            acc_t acc = {U(0), U(0), U(0), U(0)};
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const int col = c0 + 4 * s + lg;
                // D(p, q) = sum_k As(i0 + p, col_k) As(j0 + q, col_k)
                acc = MfmaOp<U>::mma(As[(i0 + li) + 65 * col], As[(j0 + li) + 65 * col], acc);
            }
#pragma unroll
            for (int v = 0; v < 4; ++v) As[(i0 + MfmaOp<U>::row(lg, v)) + 65 * (j0 + li)] -= acc[v];
        }
}

// Cholesky of the 64x64 block held column-major in LDS (As[r + 65 c]), in
// four 16-column panels: each panel factored by wavefront 0 with its rows in
// registers (pivot and column entries broadcast by v_readlane).  After a
// panel, the four wavefronts update the next panel's columns (one MFMA tile
// each); the rest of that panel's trailing update runs on wavefronts 1-3
// while wavefront 0 factors the next panel.  A non-positive pivot d leaves
// L(j, j) = d rsq(d) = NaN, so the first one is found afterwards from the
// diagonal.  Called by all threads of the workgroup.
template <typename U>
__device__ __forceinline__ void chol64_panels(U* As, int tid) {
    const int wave = tid >> 6, lane = tid & 63;
    for (int p = 0; p < 4; ++p) {
        const int c0 = 16 * p;
        if (wave == 0) {
            const int r = c0 + lane;
            // lanes past row 63 work on copies of row 63 and store nothing
            U a[16];
#pragma unroll
            for (int j = 0; j < 16; ++j) a[j] = As[min(r, 63) + 65 * (c0 + j)];
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                const U djj = bcast(a[j], j);
                const U is = pivot_rsqrt(djj);
                a[j] = (lane == j ? djj : a[j]) * is;
                // all broadcasts of the column first, then the updates: the
                // scalars land in distinct SGPRs, so no update waits on the
                // broadcast just before it
                U l[16];
#pragma unroll
                for (int k = j + 1; k < 16; ++k) l[k] = bcast(a[j], k);
#pragma unroll
                for (int k = j + 1; k < 16; ++k) a[k] -= a[j] * l[k];
            }
            // above the diagonal the panel holds by-products: the strict
            // upper triangle of As is never read after this
            if (r < 64) {
#pragma unroll
                for (int j = 0; j < 16; ++j) As[r + 65 * (c0 + j)] = a[j];
            }
        } else if (p > 0) {
            // look-ahead: the previous panel's update of the columns right
            // of this panel, while wavefront 0 factors it
            chol64_trailing<U>(As, c0 - 16, wave, lane, 1, 4, 1, 3);
        }
        __syncthreads();
        if (p < 3) {
            chol64_trailing<U>(As, c0, wave, lane, 0, 1, 0, 4);
            __syncthreads();
        }
    }
}

// Inverses of the four 16x16 diagonal blocks of the 64x64 lower-triangular
// factor in As (As[r + 65 c]), rd[i] = 1 / L(i, i): wavefront w, lane j < 16,
// forms column j of inv(L_ww) by forward substitution into Zs (same layout,
// zero above the diagonal of the block).  No barrier.
template <typename T>
__device__ __forceinline__ void inv16_diag(const T* As, const T* rd, T* Zs, int wave, int lane) {
    if (lane < 16) {
        const int b0 = 16 * wave, j = lane;
        T z[16];
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            T s = i == j ? T(1) : T(0);
#pragma unroll
            for (int k = 0; k < i; ++k) s -= As[(b0 + i) + 65 * (b0 + k)] * z[k];
            z[i] = s * rd[b0 + i];
        }
#pragma unroll
        for (int i = 0; i < 16; ++i) Zs[(b0 + i) + 65 * (b0 + j)] = i >= j ? z[i] : T(0);
    }
}

// Z = inv(L) of the 64x64 lower-triangular factor in As (As[r + 65 c],
// upper part ignored), rd[i] = 1 / L(i, i), into Zs (same layout, zero
// above the diagonal of the diagonal blocks) by 16x16 blocks: wavefront w
// inverts diagonal block w (lane j < 16: column j by forward substitution),
// then, one block diagonal after the other, Z_IJ = -Z_II sum_{K=J}^{I-1}
// L_IK Z_KJ on MFMA, one block per wavefront, the sum staged in Ts (16 x 17
// per wavefront).  Called by all threads; ends with a barrier.
template <typename T>
__device__ __forceinline__ void inv64_blocked(const T* As, const T* rd, T* Zs, T* Ts, int wave, int lane) {
    typedef typename MfmaOp<T>::acc_t acc_t;
    const int li = lane & 15, lg = lane >> 4;
    inv16_diag<T>(As, rd, Zs, wave, lane);
    __syncthreads();
    T* __restrict__ tw = Ts + wave * (16 * 17);
    for (int d = 1; d < 4; ++d) {
        const int I = d + wave, J = wave;
        const bool act = I < 4;
        if (act) {
            acc_t acc = {T(0), T(0), T(0), T(0)};
            for (int K = J; K < I; ++K) {
#pragma unroll
                for (int s = 0; s < 4; ++s) {
                    const int kk = 16 * K + 4 * s + lg;
                    // D(p, q) = sum_k L(16 I + p, kk) Z(kk, 16 J + q)
                    acc = MfmaOp<T>::mma(As[(16 * I + li) + 65 * kk], Zs[kk + 65 * (16 * J + li)], acc);
                }
            }
#pragma unroll
            for (int v = 0; v < 4; ++v) tw[MfmaOp<T>::row(lg, v) + 17 * li] = acc[v];
        }
        __syncthreads();
        if (act) {
            acc_t acc = {T(0), T(0), T(0), T(0)};
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const int k = 4 * s + lg;
                // D(p, q) = sum_k Z(16 I + p, 16 I + k) T(k, q)
                acc = MfmaOp<T>::mma(Zs[(16 * I + li) + 65 * (16 * I + k)], tw[k + 17 * li], acc);
            }
#pragma unroll
            for (int v = 0; v < 4; ++v) Zs[(16 * I + MfmaOp<T>::row(lg, v)) + 65 * (16 * J + li)] = -acc[v];
        }
        __syncthreads();
    }
}

// Fused left-looking update + diagonal-block Cholesky + TRSM of one
// 64-column block step (the potrf + trsm pair of FactorAux_GPU_2,
// mdgpu.cu:903-1216, with the part of its trailing gemm, :1336, that falls
// inside the outer block applied here instead of by a separate launch).
// Workgroups of 256 threads own kBlkRows = 64 rows below the diagonal block
// (wavefront w: rows 16 w .. 16 w + 15).  Each workgroup
//  * loads its rows and the diagonal block (64 x 64) of the block's columns;
//  * subtracts the products with the kprev earlier 64-column blocks of the
//    outer block (left-looking: C -= L_rows L_diag^T on MFMA, 64 columns of
//    K at a time, the diagonal block's rows of each K chunk staged in LDS,
//    the next chunk's loads in flight while the current one is multiplied);
//  * factors the updated diagonal block in LDS (identity beyond the block
//    width b), inverts its four 16x16 diagonal blocks and solves its rows
//    against L_D^T by 16-column blocks on MFMA (X = A inv(L_D)^T).
// The last workgroup of the job to arrive writes L_D back at its end
// (counters[jb.cnt] returns to zero).  float: a pivot-growth check (L(j,j)^2 against the
// updated diagonal) refactors the block in FP64 when it fails.
template <typename T>
__global__ __launch_bounds__(kThreads) void blk_kernel(const BlkJob* __restrict__ jobs, int njobs,
                                                       T* __restrict__ L, int64_t nL, int* __restrict__ info,
                                                       int* __restrict__ counters, int* __restrict__ fallbacks,
                                                       __bf16* __restrict__ planes, int64_t pstride, int nplanes) {
    typedef typename MfmaOp<T>::acc_t acc_t;
    constexpr bool kF32 = std::is_same<T, float>::value;
    constexpr int kZ = 64 * 65;                                          // one 64 x 64 block
    constexpr int kFb = kF32 ? 64 * 65 * 8 : 0;                          // FP64 fallback block
    // FP32 K chunks: two buffers of 16 blocks of 4 columns (64 rows each,
    // 16 floats of padding per block), column c in block c % 16, slot c / 16
    constexpr int kChBlk = 4 * 64 + 16, kChunk = 16 * kChBlk;
    constexpr int kCh = kF32 ? 2 * kChunk * 4 : 0;
    constexpr int kScr0 = kZ * (int)sizeof(T) > kFb ? kZ * (int)sizeof(T) : kFb;
    constexpr int kScr = kScr0 > kCh ? kScr0 : kCh;
    __shared__ __align__(16) T As[64 * 65];   // column-major A(r, c) = As[r + 65 c]
    // K chunks of the update, then the FP64 fallback block or the inverses
    // of the diagonal 16x16 blocks (same layout as As)
    __shared__ __align__(16) unsigned char scr[kScr];
    __shared__ T rd[64];
    __shared__ T a0[64];
    __shared__ int last;
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (jobs[mid].tile_base <= (int)blockIdx.x) lo = mid; else hi = mid - 1;
    }
    const BlkJob jb = jobs[lo];
    const int b = jb.b, ld = jb.ld, K = 64 * jb.kprev;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, li = lane & 15, lg = lane >> 4;
    T* __restrict__ g = L + jb.p_off;
    T* Ws = reinterpret_cast<T*>(scr);
    // this lane's row below the block (rw < rows) and its diagonal-block row;
    // rows past the job read in-range duplicates (row b - 1 if rows == 0)
    const int rw = ((int)blockIdx.x - jb.tile_base) * kBlkRows + 16 * wave + li;
    const int64_t rrow = b + min(rw, jb.rows - 1);
    const int drow = min(16 * wave + li, b - 1);
    // current values in the accumulator layout: R(N, v) = A(row, 16 N + row(lg, v)),
    // D(N, v) the same for diagonal-block row 16 w + li
    T Rv[4][4], Dv[4][4];
#pragma unroll
    for (int N = 0; N < 4; ++N)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int64_t col = min(16 * N + MfmaOp<T>::row(lg, v), b - 1);
            Rv[N][v] = g[rrow + col * ld];
            Dv[N][v] = g[drow + col * ld];
        }
    if (kF32 && K > 0) {
        // the chunk's diagonal-block rows go from global memory straight into
        // LDS (buffer_load_dwordx4 ... lds, double-buffered); wavefront w
        // loads blocks 4 w .. 4 w + 3, lane l taking rows 4 (l % 16) .. + 3
        // of column j + 16 (l / 16) of block j.  Rows past the block width b
        // read the panel's next rows (or zeros past the L store's end): they
        // reach only entries the step does not keep.  The padding puts the
        // four columns a k-step reads in different banks.
        const T* __restrict__ gp = g - (int64_t)K * ld;   // the outer block's earlier columns
        acc_t aR[4], aD[4];
#pragma unroll
        for (int N = 0; N < 4; ++N) {
            aR[N] = acc_t{T(0), T(0), T(0), T(0)};
            aD[N] = acc_t{T(0), T(0), T(0), T(0)};
        }
        T* __restrict__ Wc = reinterpret_cast<T*>(scr);
        auto wpos = [&](int r, int c) { return (c & 15) * kChBlk + (c >> 4) * 64 + r; };
        T xr[16];
        auto issue = [&](int buf, int kc) {
            const T* base = gp + (int64_t)kc * ld;
            const auto rs = buffer_rsrc(base, (nL - (base - L)) * (int64_t)sizeof(T));
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int j = 4 * wave + i, c = j + 16 * (lane >> 4);
                buffer_load_lds16(rs, (lds_void_ptr)&scr[(buf * kChunk + j * kChBlk) * 4],
                                  (uint32_t)(4 * (lane & 15) + (int64_t)c * ld) * 4u);
            }
#pragma unroll
            for (int ks = 0; ks < 16; ++ks) xr[ks] = gp[rrow + (int64_t)(kc + 4 * ks + lg) * ld];
        };
        issue(0, 0);
        for (int kc = 0, buf = 0; kc < K; kc += 64, buf ^= 1) {
            // this chunk's loads are the only ones outstanding; the barrier
            // makes them visible and ends the previous chunk's reads
            __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
            __syncthreads();
            T xc[16];
#pragma unroll
            for (int ks = 0; ks < 16; ++ks) xc[ks] = xr[ks];
            if (kc + 64 < K) issue(buf ^ 1, kc + 64);
            const T* __restrict__ W = Wc + buf * kChunk;
#pragma unroll
            for (int ks = 0; ks < 16; ++ks) {
                const T xd = W[wpos(16 * wave + li, 4 * ks + lg)];
#pragma unroll
                for (int N = 0; N < 4; ++N) {
                    const T bf = W[wpos(16 * N + li, 4 * ks + lg)];
                    aR[N] = MfmaOp<T>::mma(bf, xc[ks], aR[N]);
                    aD[N] = MfmaOp<T>::mma(bf, xd, aD[N]);
                }
            }
        }
#pragma unroll
        for (int N = 0; N < 4; ++N)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                Rv[N][v] -= aR[N][v];
                Dv[N][v] -= aD[N][v];
            }
        __syncthreads();   // the last chunk's readers are done
    } else if (K > 0) {
        const T* __restrict__ gp = g - (int64_t)K * ld;   // the outer block's earlier columns
        acc_t aR[4], aD[4];
#pragma unroll
        for (int N = 0; N < 4; ++N) {
            aR[N] = acc_t{T(0), T(0), T(0), T(0)};
            aD[N] = acc_t{T(0), T(0), T(0), T(0)};
        }
        T xr[16], bv[16];
        auto fetch = [&](int kc) {
#pragma unroll
            for (int ks = 0; ks < 16; ++ks) xr[ks] = gp[rrow + (int64_t)(kc + 4 * ks + lg) * ld];
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                const int t = tid + kThreads * q, r = t & 63, c = t >> 6;
                bv[q] = gp[min(r, b - 1) + (int64_t)(kc + c) * ld];
            }
        };
        fetch(0);
        for (int kc = 0; kc < K; kc += 64) {
            __syncthreads();   // the previous chunk's readers are done
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                const int t = tid + kThreads * q, r = t & 63, c = t >> 6;
                Ws[r + 65 * c] = bv[q];
            }
            T xc[16];
#pragma unroll
            for (int ks = 0; ks < 16; ++ks) xc[ks] = xr[ks];
            __syncthreads();
            if (kc + 64 < K) fetch(kc + 64);
#pragma unroll
            for (int ks = 0; ks < 16; ++ks) {
                const T xd = Ws[(16 * wave + li) + 65 * (4 * ks + lg)];
#pragma unroll
                for (int N = 0; N < 4; ++N) {
                    const T bf = Ws[(16 * N + li) + 65 * (4 * ks + lg)];
                    aR[N] = MfmaOp<T>::mma(bf, xc[ks], aR[N]);
                    aD[N] = MfmaOp<T>::mma(bf, xd, aD[N]);
                }
            }
        }
#pragma unroll
        for (int N = 0; N < 4; ++N)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                Rv[N][v] -= aR[N][v];
                Dv[N][v] -= aD[N][v];
            }
        __syncthreads();   // the last chunk's readers are done
    }
    // the updated diagonal block (identity beyond b) to LDS
#pragma unroll
    for (int N = 0; N < 4; ++N)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int r = 16 * wave + li, c = 16 * N + MfmaOp<T>::row(lg, v);
            As[r + 65 * c] = (r < b && c < b) ? Dv[N][v] : (r == c ? T(1) : T(0));
        }
    __syncthreads();
    if (tid == 0) last = atomicAdd(&counters[jb.cnt], 1) == jb.ntiles - 1;
    if (tid < 64) a0[tid] = As[tid * 66];            // updated diagonal
    chol64_panels<T>(As, tid);
    bool fell_back = false;
    if constexpr (std::is_same<T, float>::value) {
        // pivot-growth check; on failure factor the block again in FP64
        const double ljj = tid < 64 ? (double)As[tid * 66] : 1.0;
        const bool weak = tid < 64 && !(ljj * ljj >= kPivotGrowthMin * (double)a0[tid]);
        if (__syncthreads_or(weak)) {
            fell_back = true;
            double* Ad = reinterpret_cast<double*>(scr);
#pragma unroll
            for (int N = 0; N < 4; ++N)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int r = 16 * wave + li, c = 16 * N + MfmaOp<T>::row(lg, v);
                    Ad[r + 65 * c] = (r < b && c < b) ? (double)Dv[N][v] : (r == c ? 1.0 : 0.0);
                }
            __syncthreads();
            chol64_panels<double>(Ad, tid);
            for (int t = tid; t < 64 * 64; t += kThreads) {
                const int r = t & 63, c = t >> 6;
                As[r + 65 * c] = (T)Ad[r + 65 * c];
            }
            __syncthreads();
        }
    }
    if (tid < 64) rd[tid] = T(1) / As[tid * 66];
    // first non-positive pivot (64 if none): the first diagonal entry of L
    // that is not > 0 (NaN included)
    int bad = 64;
    if (wave == 0) {
        const unsigned long long nb = __ballot(!(As[lane * 66] > T(0)));
        bad = nb ? __builtin_ctzll(nb) : 64;
    }
    __syncthreads();   // rd
    T* Zs = Ws;
    inv16_diag<T>(As, rd, Zs, wave, lane);
    __syncthreads();
    // X = A inv(L_D)^T by 16-column blocks N: T_N = A_N - sum_{M<N} X_M L_NM^T,
    // X_N = T_N inv(L_NN)^T, all in the accumulator layout (lane (li, lg),
    // element v: row 16 w + li, column 16 N + row(lg, v)); MFMA step s of a
    // product contracts over the columns {row(lg, s) : lg = 0..3} of a block,
    // so an accumulator feeds the next product as its B operand unchanged.
    // Columns past b hold finite duplicates; L_D is the identity there.
    acc_t acc[4];
#pragma unroll
    for (int N = 0; N < 4; ++N) {
        acc_t t = {Rv[N][0], Rv[N][1], Rv[N][2], Rv[N][3]};
#pragma unroll
        for (int M = 0; M < N; ++M)
#pragma unroll
            for (int s = 0; s < 4; ++s)
                t = MfmaOp<T>::mma(-As[(16 * N + li) + 65 * (16 * M + MfmaOp<T>::row(lg, s))], acc[M][s], t);
        acc[N] = acc_t{T(0), T(0), T(0), T(0)};
#pragma unroll
        for (int s = 0; s < 4; ++s)
            acc[N] = MfmaOp<T>::mma(Zs[(16 * N + li) + 65 * (16 * N + MfmaOp<T>::row(lg, s))], t[s], acc[N]);
    }
    if (rw < jb.rows) {
        T* __restrict__ dst = g + b + rw;
#pragma unroll
        for (int N = 0; N < 4; ++N)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int col = 16 * N + MfmaOp<T>::row(lg, v);
                if (col < b) dst[(int64_t)col * ld] = acc[N][v];
            }
    }
    if constexpr (std::is_same<T, float>::value) {
        // bf16 split planes of the rows (bf16 variants; see Plan::planes;
        // nplanes = 2: hi and lo = bf16(x - hi), 3: hi, mid and lo):
        // lane (li, lg) holds columns 16 N + 4 lg .. + 3, so octet (N, p)
        // (columns 16 N + 8 p .. + 7) is split over lanes lg = 2 p and
        // 2 p + 1.  Lane parity e writes the octets of N = 2 e, 2 e + 1
        // after swapping halves with its partner (lane ^ 16).  Columns past
        // b inside the last octet hold the duplicates above; octets past b
        // are not written.
        if (nplanes > 0) {
            const int p = lg >> 1, e = lg & 1;
            float rcv[2][4];
#pragma unroll
            for (int j = 0; j < 2; ++j)
#pragma unroll
                for (int v = 0; v < 4; ++v) rcv[j][v] = __shfl_xor(e ? acc[j][v] : acc[2 + j][v], 16, 64);
            if (rw < jb.rows) {
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    const int N = 2 * e + j;
                    if (16 * N + 8 * p >= b) continue;
                    // x = h + m + l: h = bf16(x), m = bf16(x - h), l = bf16(x - h - m)
                    // (each difference exact in FP32)
                    bf16x8 h, m, l;
#pragma unroll
                    for (int v = 0; v < 8; ++v) {
                        const float x = v < 4 ? (e ? rcv[j][v] : acc[N][v]) : (e ? acc[N][v - 4] : rcv[j][v - 4]);
                        h[v] = (__bf16)x;
                        const float r = x - (float)h[v];
                        m[v] = (__bf16)r;
                        l[v] = (__bf16)(r - (float)m[v]);
                    }
                    const int64_t d = jb.pl + (int64_t)(2 * N + p) * 8 * ld + 8 * (int64_t)rw;
                    *reinterpret_cast<bf16x8*>(planes + d) = h;
                    *reinterpret_cast<bf16x8*>(planes + pstride + d) = m;
                    if (nplanes == 3) *reinterpret_cast<bf16x8*>(planes + 2 * pstride + d) = l;
                }
            }
        }
    }
    // the last workgroup of the job writes L_D back; last, so that no
    // barrier waits on these stores (As is not written after the Cholesky)
    if (last) {
        for (int t = tid; t < 64 * 64; t += kThreads) {
            const int r = t & 63, c = t >> 6;
            if (r < b && c < b && r >= c) g[r + (int64_t)c * ld] = As[r + 65 * c];
        }
        if (tid == 0) {
            if (bad < b) atomicMin(info, jb.col0 + bad);
            if (fell_back) atomicAdd(fallbacks, 1);
            atomicExch(&counters[jb.cnt], 0);
        }
    }
}

// Inverse of one 64-column diagonal block of an already factored panel
// (identity beyond the block width b), one 256-thread workgroup per block,
// by 16x16 blocks on MFMA (inv64_blocked).  Stored as W(i + 64 r) =
// inv(L_D)(r, i) for i, r < b, zero above the diagonal.
template <typename T>
__global__ __launch_bounds__(kThreads) void inv_kernel(const InvJob* __restrict__ jobs,
                                                       const FrontDesc* __restrict__ fronts,
                                                       const T* __restrict__ L, T* __restrict__ inv) {
    __shared__ T As[64 * 65];
    __shared__ T Zs[64 * 65 + 4 * 16 * 17];
    __shared__ T rd[64];
    const InvJob jb = jobs[blockIdx.x];
    const FrontDesc d = fronts[jb.front];
    const int b = min(64, d.c - jb.j0), n = d.n;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const T* __restrict__ g = L + d.lofs + jb.j0 + (int64_t)jb.j0 * n;
    T v[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int t = tid + kThreads * q, r = t & 63, c = t >> 6;
        v[q] = g[min(r, b - 1) + (int64_t)min(c, b - 1) * n];
    }
    keep_in_flight(v);
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int t = tid + kThreads * q, r = t & 63, c = t >> 6;
        As[r + 65 * c] = (r < b && c < b) ? v[q] : (r == c ? T(1) : T(0));
    }
    __syncthreads();
    if (tid < 64) rd[tid] = T(1) / As[tid * 66];
    __syncthreads();
    inv64_blocked<T>(As, rd, Zs, Zs + 64 * 65, wave, lane);
    T* __restrict__ w = inv + d.dinv + (int64_t)(jb.j0 / 64) * 4096;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int t = tid + kThreads * q, i = t & 63, r = t >> 6;
        if (i < b && r < b) w[i + 64 * r] = r >= i ? Zs[r + 65 * i] : T(0);
    }
}

// Inverses of the two kChainBW-column diagonal blocks inside a
// kSolveBW-column block L_B of a front (for the chained solves), one
// workgroup per 256-column block, from the 64x64 inverses W_s of its
// diagonal sub-blocks: X_ss = W_s, and below the diagonal, X_ts = -W_t
// sum_{u=s}^{t-1} L_tu X_us for the pairs (s, t) inside one 128-column half
// (s / 2 == t / 2): the quarter below the halves is not formed.  Products of
// 64x64 blocks go through LDS 32 columns at a time; thread (i, g) owns
// column i, rows 16g..16g+15.  X is written column-major with leading
// dimension 256 and is zero above the diagonal, beyond the block and in the
// quarter below the halves.
template <typename T>
__global__ __launch_bounds__(kThreads) void inv256_kernel(const InvJob* __restrict__ jobs,
                                                          const FrontDesc* __restrict__ fronts,
                                                          const T* __restrict__ L, const T* __restrict__ inv,
                                                          T* __restrict__ X256) {
    __shared__ T Al[64 * 33];
    __shared__ T Bl[32 * 65];
    const InvJob jb = jobs[blockIdx.x];
    const FrontDesc d = fronts[jb.front];
    const int n = d.n, j0 = jb.j0, bw = min(kSolveBW, d.c - j0), nsub = (bw + 63) / 64;
    const T* __restrict__ P = L + d.lofs;
    T* __restrict__ X = X256 + d.dinv256 + (int64_t)(j0 / kSolveBW) * kSolveBW * kSolveBW;
    const int tid = threadIdx.x, ci = tid & 63, rg = tid >> 6;
    for (int t = tid; t < kSolveBW * kSolveBW; t += kThreads) X[t] = T(0);
    __syncthreads();
    for (int s = 0; s < nsub; ++s) {
        const int bs = min(64, bw - 64 * s);
        const T* __restrict__ w = inv + d.dinv + (int64_t)((j0 + 64 * s) / 64) * 4096;
        for (int t = tid; t < 64 * 64; t += kThreads) {
            const int i = t & 63, r = t >> 6;          // W[i + 64 r] = inv(L_s)(r, i)
            if (i < bs && r < bs && r >= i) X[(64 * s + r) + kSolveBW * (64 * s + i)] = w[i + 64 * r];
        }
    }
    __syncthreads();
    for (int s = 0; s < nsub; ++s) {
        const int bs = min(64, bw - 64 * s);
        for (int t = s + 1; t < nsub && t / 2 == s / 2; ++t) {
            const int bt = min(64, bw - 64 * t);
            T acc[16];
#pragma unroll
            for (int q = 0; q < 16; ++q) acc[q] = T(0);
            for (int u = s; u < t; ++u) {
                const int bu = min(64, bw - 64 * u);
                for (int k0 = 0; k0 < bu; k0 += 32) {
                    for (int e = tid; e < 64 * 32; e += kThreads) {
                        const int r = e & 63, kk = e >> 6;
                        Al[r * 33 + kk] = (r < bt && k0 + kk < bu)
                                              ? P[(j0 + 64 * t + r) + (int64_t)(j0 + 64 * u + k0 + kk) * n] : T(0);
                    }
                    for (int e = tid; e < 32 * 64; e += kThreads) {
                        const int kk = e & 31, i = e >> 5;
                        Bl[kk * 65 + i] = X[(64 * u + k0 + kk) + kSolveBW * (64 * s + i)];
                    }
                    __syncthreads();
#pragma unroll
                    for (int q = 0; q < 16; ++q) {
                        T a = T(0);
#pragma unroll 8
                        for (int kk = 0; kk < 32; ++kk) a += Al[(rg * 16 + q) * 33 + kk] * Bl[kk * 65 + ci];
                        acc[q] += a;
                    }
                    __syncthreads();
                }
            }
            // X_ts = -W_t acc, W_t(r, m) = wt[m + 64 r] (lower: m <= r)
            const T* __restrict__ wt = inv + d.dinv + (int64_t)((j0 + 64 * t) / 64) * 4096;
            T out[16];
#pragma unroll
            for (int q = 0; q < 16; ++q) out[q] = T(0);
            for (int m0 = 0; m0 < 64; m0 += 32) {
                for (int e = tid; e < 64 * 32; e += kThreads) {
                    const int r = e & 63, mm = e >> 6, m = m0 + mm;
                    Al[r * 33 + mm] = (r < bt && m < bt && m <= r) ? wt[m + 64 * r] : T(0);
                }
#pragma unroll
                for (int q = 0; q < 16; ++q) {
                    const int m = rg * 16 + q;
                    if (m >= m0 && m < m0 + 32) Bl[(m - m0) * 65 + ci] = acc[q];
                }
                __syncthreads();
#pragma unroll
                for (int q = 0; q < 16; ++q) {
                    T a = T(0);
#pragma unroll 8
                    for (int mm = 0; mm < 32; ++mm) a += Al[(rg * 16 + q) * 33 + mm] * Bl[mm * 65 + ci];
                    out[q] += a;
                }
                __syncthreads();
            }
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                const int r = rg * 16 + q;
                if (r < bt && ci < bs) X[(64 * t + r) + kSolveBW * (64 * s + ci)] = -out[q];
            }
            __syncthreads();
        }
    }
}

// C -= A * B^T (beta0: C = -A * B^T), optionally keeping only the lower
// triangle, tiles of many jobs (fronts) in one launch: the trailing-matrix
// gemm of FactorAux_CPU_3 / FactorAux_GPU_2 (mdgpu.cu:1336, :1058)
// restricted to the lower triangle.  128x128 (or 64x64) C tiles, four
// wavefronts of 64x64, v_mfma_{f32,f64}_16x16x4.  The contraction index of
// one 16x16x4 instruction is free to relabel as long as both operands agree,
// so lane group g = lane/16 supplies k = 4g+t in sub-step t: each lane then
// reads a contiguous 4-vector of k from LDS.  The product is formed as
// (B A^T) so that each lane's four accumulator rows map to four C columns
// and a 16-lane group stores 16 consecutive C rows.
// Store of a wavefront's 64x64 block of C (accumulators acc[rb][cb][v] =
// C(i0 + 16 rb, j0 + 16 cb + row(lg, v))): C -= acc, or C = -acc (beta0),
// lower triangle only if g.lower.  For C -= acc all old values of a column
// block are loaded first, at clamped in-range addresses, so the loads are in
// flight together (a load after each possibly aliasing store would wait
// for each one in turn).
template <typename T, int NB, typename Acc, typename RowFn>
__device__ __forceinline__ void gemm_epilogue(const Acc (&acc)[NB][NB], T* __restrict__ C, const GemmJob& g, int i0,
                                              int j0, int lg, RowFn row) {
#pragma unroll
    for (int cb = 0; cb < NB; ++cb) {
        T old[4][NB];
        bool ok[4][NB];
#pragma unroll
        for (int v = 0; v < 4; ++v)
#pragma unroll
            for (int rb = 0; rb < NB; ++rb) {
                const int i = i0 + rb * 16, j = j0 + cb * 16 + row(lg, v);
                ok[v][rb] = i < g.M && j < g.N && !(g.lower && i < j);
                old[v][rb] = T(0);
            }
        if (!g.beta0) {
#pragma unroll
            for (int v = 0; v < 4; ++v)
#pragma unroll
                for (int rb = 0; rb < NB; ++rb) {
                    const int i = min(i0 + rb * 16, g.M - 1), j = min(j0 + cb * 16 + row(lg, v), g.N - 1);
                    old[v][rb] = C[i + (int64_t)j * g.ldc];
                }
#pragma unroll
            for (int v = 0; v < 4; ++v)
#pragma unroll
                for (int rb = 0; rb < NB; ++rb) asm volatile("" : "+v"(old[v][rb]));
        }
#pragma unroll
        for (int v = 0; v < 4; ++v)
#pragma unroll
            for (int rb = 0; rb < NB; ++rb)
                if (ok[v][rb])
                    C[(i0 + rb * 16) + (int64_t)(j0 + cb * 16 + row(lg, v)) * g.ldc] = old[v][rb] - acc[rb][cb][v];
    }
}

// Workgroups are dispatched round-robin over the 8 XCDs (workgroup i runs
// on XCD i % 8), each with its own L2.  Renumber them so that each XCD gets
// one contiguous run of tiles: neighbouring tiles share A or B panels, which
// then hit the same L2.  A bijection on [0, n).
__device__ inline int xcd_tile(int bid, int n) {
    constexpr int kXcd = 8;
    const int x = bid % kXcd, k = bid / kXcd, q = n / kXcd, r = n % kXcd;
    return x * q + min(x, r) + k;
}

// Register budget: at least 3 wavefronts per SIMD in FP32, 2 in FP64
// (measured: the 128-row FP32 kernel then runs four workgroups per CU, the
// FP64 one two instead of one).
// SPLIT (FP32): two-level accumulation -- the products of each run of
// kSplitTiles k-tiles (64 k) are summed in a partial accumulator that is
// then added to the total, so a long K is summed as K/64 partial sums of 64
// terms instead of one running sum of K terms (a smaller rounding error;
// long-K Schur complements in FP32 get it).
constexpr int kSplitTiles = 4;
template <typename T, int BT, bool SPLIT>
__global__ __launch_bounds__(kThreads, sizeof(T) == 4 ? 3 : 2) void gemm_mfma_kernel(const GemmJob* __restrict__ jobs, int njobs,
                                                             T* __restrict__ L, T* __restrict__ U, int ntiles) {
    static_assert(!SPLIT || sizeof(T) == 4, "two-level accumulation is for FP32");
    // k-tile depth 16 (register budget of the prefetch, and LDS for four
    // FP32 workgroups per CU); the next k-tile is loaded into registers while
    // the current one is multiplied out of LDS.
    // BT x BT tiles (128, or 64 for short launches); each wavefront owns a
    // (BT/2) x (BT/2) quadrant of NB x NB 16x16 MFMA blocks
    constexpr int KT = 16, LDK = KT + 4, NB = BT / 32, KSTEP = kThreads / BT;
    constexpr int PER = (BT * KT) / kThreads;      // staged elements per thread and operand
    constexpr int G = 16 / sizeof(T), NG = PER / G; // k values per 16-byte LDS store, stores per operand
    typedef T vec_t __attribute__((ext_vector_type(G)));
    typedef typename MfmaOp<T>::acc_t acc_t;
    // two LDS buffers: the next k-tile is stored while this one is multiplied
    __shared__ __align__(16) T As[2][BT * LDK];
    __shared__ __align__(16) T Bs[2][BT * LDK];
    // persistent over the launch's ntiles tiles when the grid is smaller
    // (side-stream launches leave CUs to the main stream's launches)
    for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        int lo = 0, hi = njobs - 1;
        const int bid = xcd_tile(tile, ntiles);
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
        const int wr = (wave >> 1) * (BT / 2), wc = (wave & 1) * (BT / 2);
        const int li = lane & 15, lg = lane >> 4;
        // staging coordinates: element G j + u of this thread is row r = tid % BT,
        // k-offset kk = G (tid / BT + KSTEP j) + u: G consecutive k of one row go
        // to LDS in one 16-byte store (rows 4 banks apart: conflict-free)
        const int r_st = tid & (BT - 1), kgrp = tid / BT;
        const int ia = row0 + r_st, jb = col0 + r_st;
        const bool a_ok = ia < g.M, b_ok = jb < g.N;
        const T* __restrict__ Ap = A + ia;
        const T* __restrict__ Bp = B + jb;
        T ra[PER], rb[PER];
        // loads at clamped in-range addresses; entries outside the job are
        // zeroed when stored to LDS, so no wait sits between a fetch and the
        // multiplications that follow it
        const T* __restrict__ Apc = a_ok ? Ap : A;
        const T* __restrict__ Bpc = b_ok ? Bp : B;
        auto kof = [&](int s) { return G * (kgrp + KSTEP * (s / G)) + s % G; };
        auto fetch = [&](int k0) {
    #pragma unroll
            for (int s = 0; s < PER; ++s) {
                const int k = min(k0 + kof(s), g.K - 1);
                ra[s] = Apc[(int64_t)k * g.lda];
                rb[s] = Bpc[(int64_t)k * g.ldb];
            }
        };
        acc_t acc[NB][NB];
    #pragma unroll
        for (int a = 0; a < NB; ++a)
    #pragma unroll
            for (int b = 0; b < NB; ++b)
    #pragma unroll
                for (int v = 0; v < 4; ++v) acc[a][b][v] = T(0);
        auto store = [&](int buf, int k0) {
    #pragma unroll
            for (int j = 0; j < NG; ++j) {
                vec_t va, vb;
    #pragma unroll
                for (int u = 0; u < G; ++u) {
                    const bool kin = k0 + kof(G * j + u) < g.K;
                    va[u] = (a_ok && kin) ? ra[G * j + u] : T(0);
                    vb[u] = (b_ok && kin) ? rb[G * j + u] : T(0);
                }
                const int kk = G * (kgrp + KSTEP * j);
                *reinterpret_cast<vec_t*>(&As[buf][r_st * LDK + kk]) = va;
                *reinterpret_cast<vec_t*>(&Bs[buf][r_st * LDK + kk]) = vb;
            }
        };
        fetch(0);
        store(0, 0);
        if (KT < g.K) fetch(KT);
        __syncthreads();
        if constexpr (sizeof(T) == 8) {
            // FP64: each iteration is one basic block (the store and fetch of
            // the next tiles run unconditionally: past the end they store zeros
            // to the unused buffer and load clamped addresses), and scheduling
            // groups interleave them with the multiplications -- measured
            // faster for FP64, slower for FP32
            for (int k0 = 0, cur = 0; k0 < g.K; k0 += KT, cur ^= 1) {
                const T* __restrict__ Ac = As[cur];
                const T* __restrict__ Bc = Bs[cur];
                T av[KT / 16][NB][4], bv[KT / 16][NB][4];
    #pragma unroll
                for (int q = 0; q < KT / 16; ++q) {
    #pragma unroll
                    for (int rb2 = 0; rb2 < NB; ++rb2)
    #pragma unroll
                        for (int t = 0; t < 4; ++t) av[q][rb2][t] = Ac[(wr + rb2 * 16 + li) * LDK + 16 * q + 4 * lg + t];
    #pragma unroll
                    for (int cb = 0; cb < NB; ++cb)
    #pragma unroll
                        for (int t = 0; t < 4; ++t) bv[q][cb][t] = Bc[(wc + cb * 16 + li) * LDK + 16 * q + 4 * lg + t];
                }
                store(cur ^ 1, k0 + KT);
                fetch(k0 + 2 * KT);
    #pragma unroll
                for (int q = 0; q < KT / 16; ++q)
    #pragma unroll
                    for (int t = 0; t < 4; ++t)
    #pragma unroll
                        for (int rb2 = 0; rb2 < NB; ++rb2)
    #pragma unroll
                            for (int cb = 0; cb < NB; ++cb)
                                acc[rb2][cb] = MfmaOp<T>::mma(bv[q][cb][t], av[q][rb2][t], acc[rb2][cb]);
                // NG rounds of: a share of the products, the two 16-byte LDS
                // stores of one group, its 2 G global loads
    #pragma unroll
                for (int i = 0; i < NG; ++i) {
                    __builtin_amdgcn_sched_group_barrier(0x008, (KT / 4) * NB * NB / NG, 0);   // MFMA
                    __builtin_amdgcn_sched_group_barrier(0x200, 2, 0);                          // DS write
                    __builtin_amdgcn_sched_group_barrier(0x020, 2 * G, 0);                      // VMEM read
                }
                __syncthreads();
            }
        } else {
            acc_t part[NB][NB];   // SPLIT: this run's products
    #pragma unroll
            for (int a = 0; a < NB; ++a)
    #pragma unroll
                for (int b = 0; b < NB; ++b)
    #pragma unroll
                    for (int v = 0; v < 4; ++v) part[a][b][v] = T(0);
            for (int k0 = 0, cur = 0; k0 < g.K; k0 += KT, cur ^= 1) {
                // tile k0 + KT (in registers) to the other buffer, whose last
                // readers passed the barrier that ended the previous iteration
                if (k0 + KT < g.K) {
                    store(cur ^ 1, k0 + KT);
                    if (k0 + 2 * KT < g.K) fetch(k0 + 2 * KT);
                }
                auto& sum = SPLIT ? part : acc;
                const T* __restrict__ Ac = As[cur];
                const T* __restrict__ Bc = Bs[cur];
        #pragma unroll
                for (int kq = 0; kq < KT; kq += 16) {
                    T av[NB][4], bv[NB][4];   // [block][t]
        #pragma unroll
                    for (int rb2 = 0; rb2 < NB; ++rb2)
        #pragma unroll
                        for (int t = 0; t < 4; ++t) av[rb2][t] = Ac[(wr + rb2 * 16 + li) * LDK + kq + 4 * lg + t];
        #pragma unroll
                    for (int cb = 0; cb < NB; ++cb)
        #pragma unroll
                        for (int t = 0; t < 4; ++t) bv[cb][t] = Bc[(wc + cb * 16 + li) * LDK + kq + 4 * lg + t];
        #pragma unroll
                    for (int t = 0; t < 4; ++t)
        #pragma unroll
                        for (int rb2 = 0; rb2 < NB; ++rb2)
        #pragma unroll
                            for (int cb = 0; cb < NB; ++cb)
                                sum[rb2][cb] = MfmaOp<T>::mma(bv[cb][t], av[rb2][t], sum[rb2][cb]);
                }
                if (SPLIT && ((k0 / KT) % kSplitTiles == kSplitTiles - 1 || k0 + KT >= g.K)) {
    #pragma unroll
                    for (int a = 0; a < NB; ++a)
    #pragma unroll
                        for (int b = 0; b < NB; ++b) {
                            acc[a][b] += part[a][b];
    #pragma unroll
                            for (int v = 0; v < 4; ++v) part[a][b][v] = T(0);
                        }
                }
                __syncthreads();
            }
        }
        // acc[rb][cb][v] = C(row0 + wr + rb*16 + li, col0 + wc + cb*16 + row(lg, v))
        gemm_epilogue<T, NB>(acc, C, g, row0 + wr + li, col0 + wc, lg,
                             [](int lg_, int v_) { return MfmaOp<T>::row(lg_, v_); });
    }
}

// FP32 form of the same product on v_mfma_f32_32x32x2f32, for the same jobs
// and BT x BT tiles (BT = 128: four wavefronts of 64x64, 2x2 blocks of
// 32x32; BT = 64: four wavefronts of one 32x32 block).
//  * Global loads are buffer loads (raw buffer resource over the L store from
//    the panel's current column to the store's end): 16 bytes = 4
//    consecutive rows of one column per load, and loads past the store's end
//    return zero without a bounds test.  Rows past M (or N) are loaded as
//    they are: they only reach C rows (columns) that are not stored.
//  * Every full k-tile (16 columns; 32 for 64-row tiles) goes from global
//    memory straight into LDS (buffer_load_dwordx4 ... lds): one wavefront load fills 1 KB of LDS
//    in lane order, so a k-tile is laid out [k][row], BT values per column,
//    each lane taking 4 rows of one column.  No registers hold the staged
//    tile and no LDS store instructions run.  The next k-tile's loads go
//    into the other LDS buffer while the current one is multiplied; a
//    vmcnt wait and one barrier per k-tile make it visible.
//  * A partial last k-tile is staged through registers instead, its columns
//    past K zeroed on both operands (0 * NaN would not vanish).
//  * The MFMA's contraction index is relabelled: for sub-step (s, q) lane
//    half h takes column k = 4 s + 2 h + q of the k-tile.  A-side reads take
//    rows 32 rb + j (two 4-byte reads), so that a wavefront's accumulators
//    map to 32 consecutive C rows per column; B-side reads (BT = 128) take
//    rows 2 j and 2 j + 1 in one 8-byte read, its two 32-column blocks
//    interleaved.
//  * Each wavefront's LDS reads of a k-tile are issued before its products
//    (with the second accumulator set, one group of 4 k at a time).
// Same k order per entry as a sequential sum over k-pairs; the rounding
// differs from gemm_mfma_kernel's (pairs of k per instruction).

// k-tile depth by tile size: 16 for 128-row tiles (32 was measured 2-3 %
// slower on the large updates), 32 for 64-row tiles (1-10 % faster than 16:
// more products per barrier); both sum the same products in the same order
constexpr int kKtRows128 = 16, kKtRows64 = 32;
// SPLIT: two-level accumulation over runs of 64 k, as in
// gemm_mfma_kernel; at BT = 128 its second accumulator set gets a
// 3-wavefront register budget (a few registers spill; measured faster on
// the long-K Schur complements than 2 wavefronts without spills).
template <int BT, bool SPLIT>
__global__ __launch_bounds__(kThreads, (BT == 128 && SPLIT) ? 3 : 2) void gemm_mfma32_kernel(
    const GemmJob* __restrict__ jobs, int njobs, float* __restrict__ L, float* __restrict__ U, int64_t nL, int ntiles) {
    static_assert(BT == 128 || BT == 64, "128- or 64-row tiles");
    constexpr int KT = BT == 64 ? kKtRows64 : kKtRows128, NB = BT / 64;   // NB x NB 32x32 blocks per wavefront
    constexpr int kRun = 64 / KT;                       // k-tiles per two-level accumulation run
    constexpr int kLanesPerCol = BT / 4;                // lanes loading one column (4 rows each)
    constexpr int kColsPerLoad = 64 / kLanesPerCol;     // columns of one wavefront's 1 KB load
    constexpr int kLoads = KT / (4 * kColsPerLoad);     // loads per operand and wavefront per k-tile
    static_assert(kLoads >= 1 && kLoads * 4 * kColsPerLoad == KT, "a k-tile is whole wavefront loads");
    __shared__ __align__(16) float As[2][KT * BT];
    __shared__ __align__(16) float Bs[2][KT * BT];
    // persistent over the launch's ntiles tiles when the grid is smaller
    // (side-stream launches leave CUs to the main stream's launches)
    for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        int lo = 0, hi = njobs - 1;
        const int bid = xcd_tile(tile, ntiles);
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
        float* __restrict__ C = (g.c_in_arena ? U : L) + g.c_off;
        const int row0 = ti * BT, col0 = tj * BT;
        const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
        const int wr = (wave >> 1) * (BT / 2), wc = (wave & 1) * (BT / 2);
        const int j = lane & 31, h = lane >> 5;
        const int64_t a0 = g.a_off + row0, b0 = g.b_off + col0;
        // this lane's part of load i of its wavefront: rows lr .. lr + 3 of
        // column (kLoads wave + i) kColsPerLoad + lc of the k-tile
        const int lr = 4 * (lane % kLanesPerCol), lc = lane / kLanesPerCol;
        const int nfull = g.K / KT, nt = (g.K + KT - 1) / KT;
        auto col_of = [&](int i) { return (kLoads * wave + i) * kColsPerLoad; };
        // full k-tile t into LDS buffer buf
        auto issue = [&](int buf, int t) {
            const int64_t oa = a0 + (int64_t)t * KT * g.lda, ob = b0 + (int64_t)t * KT * g.ldb;
            const auto ra = buffer_rsrc(L + oa, (nL - oa) * 4);
            const auto rb = buffer_rsrc(L + ob, (nL - ob) * 4);
#pragma unroll
            for (int i = 0; i < kLoads; ++i) {
                const int c0 = col_of(i), c = c0 + lc;
                __builtin_amdgcn_raw_ptr_buffer_load_lds(ra, (lds_void_ptr)&As[buf][c0 * BT], 16,
                                                         (uint32_t)(lr + c * g.lda) * 4u, 0, 0, 0);
                __builtin_amdgcn_raw_ptr_buffer_load_lds(rb, (lds_void_ptr)&Bs[buf][c0 * BT], 16,
                                                         (uint32_t)(lr + c * g.ldb) * 4u, 0, 0, 0);
            }
        };
        // partial k-tile t through registers, columns past K zeroed
        auto stage_tail = [&](int buf, int t) {
            const int64_t oa = a0 + (int64_t)t * KT * g.lda, ob = b0 + (int64_t)t * KT * g.ldb;
            const auto ra = buffer_rsrc(L + oa, (nL - oa) * 4);
            const auto rb = buffer_rsrc(L + ob, (nL - ob) * 4);
            const f32x4_t z = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int i = 0; i < kLoads; ++i) {
                const int c = col_of(i) + lc;
                f32x4_t va = buffer_load4(ra, (uint32_t)(lr + c * g.lda) * 4u);
                f32x4_t vb = buffer_load4(rb, (uint32_t)(lr + c * g.ldb) * 4u);
                if (t * KT + c >= g.K) { va = z; vb = z; }
                *reinterpret_cast<f32x4_t*>(&As[buf][c * BT + lr]) = va;
                *reinterpret_cast<f32x4_t*>(&Bs[buf][c * BT + lr]) = vb;
            }
        };
        f32x16_t acc[NB][NB], part[NB][NB];   // part: SPLIT, this run's products
    #pragma unroll
        for (int a = 0; a < NB; ++a)
    #pragma unroll
            for (int b = 0; b < NB; ++b)
    #pragma unroll
                for (int v = 0; v < 16; ++v) acc[a][b][v] = part[a][b][v] = 0.f;
        auto& sum = SPLIT ? part : acc;
        if (nfull > 0) issue(0, 0);
        else stage_tail(0, 0);
        for (int t = 0; t < nt; ++t) {
            const int cur = t & 1;
            // k-tile t's loads are the only ones outstanding: wait for them,
            // then the barrier makes every wavefront's part visible (and ends
            // the reads of the other buffer, k-tile t - 1)
            __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
            __syncthreads();
            if (t + 1 < nfull) issue(cur ^ 1, t + 1);
            else if (t + 1 < nt) stage_tail(cur ^ 1, t + 1);
            const float* Ac = As[cur];
            const float* Bc = Bs[cur];
            // av[s][q][rb] = A(wr + 32 rb + j, k), bv[s][q][cb] = B(wc + NB j + cb, k),
            // k = 4 s + 2 h + q; all of them read before the products, except
            // with the second accumulator set (SPLIT at BT = 128), whose
            // register budget takes one group of 4 k at a time
            constexpr int kGroups = (SPLIT && NB == 2) ? 1 : KT / 4;
    #pragma unroll
            for (int s0 = 0; s0 < KT / 4; s0 += kGroups) {
                float av[kGroups][2][NB];
                float bv[kGroups][2][NB];
    #pragma unroll
                for (int s2 = 0; s2 < kGroups; ++s2)
    #pragma unroll
                    for (int q = 0; q < 2; ++q) {
                        const int k = 4 * (s0 + s2) + 2 * h + q;
    #pragma unroll
                        for (int rb = 0; rb < NB; ++rb) av[s2][q][rb] = Ac[k * BT + wr + 32 * rb + j];
                        if constexpr (NB == 2) {
                            const f32x2_t y = *reinterpret_cast<const f32x2_t*>(&Bc[k * BT + wc + 2 * j]);
                            bv[s2][q][0] = y[0];
                            bv[s2][q][1] = y[1];
                        } else {
                            bv[s2][q][0] = Bc[k * BT + wc + j];
                        }
                    }
    #pragma unroll
                for (int s2 = 0; s2 < kGroups; ++s2)
    #pragma unroll
                    for (int q = 0; q < 2; ++q)
    #pragma unroll
                        for (int rb = 0; rb < NB; ++rb)
    #pragma unroll
                            for (int cb = 0; cb < NB; ++cb)
                                sum[rb][cb] = __builtin_amdgcn_mfma_f32_32x32x2f32(bv[s2][q][cb], av[s2][q][rb],
                                                                                   sum[rb][cb], 0, 0, 0);
            }
            if (SPLIT && (t % kRun == kRun - 1 || t + 1 >= nt)) {
    #pragma unroll
                for (int rb = 0; rb < NB; ++rb)
    #pragma unroll
                    for (int cb = 0; cb < NB; ++cb) {
                        acc[rb][cb] += part[rb][cb];
    #pragma unroll
                        for (int v = 0; v < 16; ++v) part[rb][cb][v] = 0.f;
                    }
            }
        }
        // the next tile's first loads reuse the buffers
        __syncthreads();
        // acc[rb][cb][v] = C(row0 + wr + 32 rb + j, col0 + wc + NB (8 (v / 4) + 4 h + v % 4) + cb);
        // C -= acc (beta0: C = -acc), all old values of a column block loaded
        // first at clamped in-range addresses
    #pragma unroll
        for (int cb = 0; cb < NB; ++cb) {
            float old[NB][16];
    #pragma unroll
            for (int rb = 0; rb < NB; ++rb)
    #pragma unroll
                for (int v = 0; v < 16; ++v) {
                    old[rb][v] = 0.f;
                    if (!g.beta0) {
                        const int i = min(row0 + wr + 32 * rb + j, g.M - 1);
                        const int c = min(col0 + wc + NB * (8 * (v >> 2) + 4 * h + (v & 3)) + cb, g.N - 1);
                        old[rb][v] = C[i + (int64_t)c * g.ldc];
                    }
                }
    #pragma unroll
            for (int rb = 0; rb < NB; ++rb)
    #pragma unroll
                for (int v = 0; v < 16; ++v) asm volatile("" : "+v"(old[rb][v]));
    #pragma unroll
            for (int rb = 0; rb < NB; ++rb)
    #pragma unroll
                for (int v = 0; v < 16; ++v) {
                    const int i = row0 + wr + 32 * rb + j;
                    const int c = col0 + wc + NB * (8 * (v >> 2) + 4 * h + (v & 3)) + cb;
                    if (i < g.M && c < g.N && !(g.lower && i < c)) C[i + (int64_t)c * g.ldc] = old[rb][v] - acc[rb][cb][v];
                }
        }
    }
}

// BF16 variants of gemm_mfma_kernel: the factor stays FP32 in HBM, the A
// and B tiles are rounded to bf16 (v_cvt_pk_bf16_f32, round to nearest
// even) when staged into LDS, and v_mfma_f32_16x16x32_bf16 accumulates in
// FP32.  NSPLIT = 1: one bf16 product.  NSPLIT = 3: each value is split as
// hi = bf16(x), lo = bf16(x - hi) and the product is hi*hi + hi*lo + lo*hi
// (about 16 significant bits).  Each thread stages KT/2 consecutive k of one
// row per operand as 16-byte vectors; a lane's MFMA operand is 8
// consecutive k of one row (one 16-byte LDS read).  Accumulator layout as
// for FP32.
typedef float f32x4 __attribute__((ext_vector_type(4)));

template <int NSPLIT>
__global__ __launch_bounds__(kThreads) void gemm_bf16_kernel(const GemmJob* __restrict__ jobs, int njobs,
                                                             float* __restrict__ L, float* __restrict__ U, int ntiles) {
    constexpr int BT = 128, KT = 32, LDK = KT + 8;                     // rows 16-byte aligned
    constexpr int KS = KT / 2;                                        // staged k per thread
    constexpr int NP = NSPLIT == 1 ? 1 : 2;                           // hi (and lo) planes
    // two LDS buffers: the next k-tile is stored while this one is multiplied
    __shared__ __align__(16) __bf16 As[2][NP][BT * LDK];
    __shared__ __align__(16) __bf16 Bs[2][NP][BT * LDK];
    // persistent over the launch's ntiles tiles when the grid is smaller
    // (side-stream launches leave CUs to the main stream's launches)
    for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        int lo = 0, hi = njobs - 1;
        const int bid = xcd_tile(tile, ntiles);
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
        const float* __restrict__ A = L + g.a_off;
        const float* __restrict__ B = L + g.b_off;
        float* __restrict__ C = (g.c_in_arena ? U : L) + g.c_off;
        const int row0 = ti * BT, col0 = tj * BT;
        const int tid = threadIdx.x;
        const int wave = tid >> 6, lane = tid & 63;
        const int wr = (wave >> 1) * 64, wc = (wave & 1) * 64;
        const int li = lane & 15, lg = lane >> 4;
        const int r_st = tid & (BT - 1), kg = (tid >> 7) * KS;   // staged row, first staged k
        const int ia = row0 + r_st, jb = col0 + r_st;
        const bool a_ok = ia < g.M, b_ok = jb < g.N;
        const float* __restrict__ Ap = A + ia;
        const float* __restrict__ Bp = B + jb;
        float ra[KS], rb[KS];
        // loads at clamped in-range addresses; entries outside the job are
        // zeroed when stored to LDS, so no wait sits between a fetch and the
        // multiplications that follow it
        const float* __restrict__ Apc = a_ok ? Ap : A;
        const float* __restrict__ Bpc = b_ok ? Bp : B;
        auto fetch = [&](int k0) {
    #pragma unroll
            for (int s = 0; s < KS; ++s) {
                const int k = min(k0 + kg + s, g.K - 1);
                ra[s] = Apc[(int64_t)k * g.lda];
                rb[s] = Bpc[(int64_t)k * g.ldb];
            }
        };
        auto store = [&](int buf, int k0) {
    #pragma unroll
            for (int q = 0; q < KS / 8; ++q) {
                bf16x8 ah, bh, al, bl;
    #pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const bool kin = k0 + kg + 8 * q + e < g.K;
                    const float av = (a_ok && kin) ? ra[8 * q + e] : 0.0f;
                    const float bv = (b_ok && kin) ? rb[8 * q + e] : 0.0f;
                    ah[e] = (__bf16)av;
                    bh[e] = (__bf16)bv;
                    if constexpr (NSPLIT == 3) {
                        al[e] = (__bf16)(av - (float)ah[e]);
                        bl[e] = (__bf16)(bv - (float)bh[e]);
                    }
                }
                *reinterpret_cast<bf16x8*>(&As[buf][0][r_st * LDK + kg + 8 * q]) = ah;
                *reinterpret_cast<bf16x8*>(&Bs[buf][0][r_st * LDK + kg + 8 * q]) = bh;
                if constexpr (NSPLIT == 3) {
                    *reinterpret_cast<bf16x8*>(&As[buf][NP - 1][r_st * LDK + kg + 8 * q]) = al;
                    *reinterpret_cast<bf16x8*>(&Bs[buf][NP - 1][r_st * LDK + kg + 8 * q]) = bl;
                }
            }
        };
        f32x4 acc[4][4];
    #pragma unroll
        for (int a = 0; a < 4; ++a)
    #pragma unroll
            for (int b = 0; b < 4; ++b) acc[a][b] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
        fetch(0);
        store(0, 0);
        if (KT < g.K) fetch(KT);
        __syncthreads();
        for (int k0 = 0, cur = 0; k0 < g.K; k0 += KT, cur ^= 1) {
            // tile k0 + KT (in registers) to the other buffer, whose last
            // readers passed the barrier that ended the previous iteration
            if (k0 + KT < g.K) {
                store(cur ^ 1, k0 + KT);
                if (k0 + 2 * KT < g.K) fetch(k0 + 2 * KT);
            }
    #pragma unroll
            for (int kq = 0; kq < KT; kq += 32) {
                bf16x8 av[NP][4], bv[NP][4];
    #pragma unroll
                for (int p = 0; p < NP; ++p)
    #pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        av[p][r] = *reinterpret_cast<const bf16x8*>(&As[cur][p][(wr + r * 16 + li) * LDK + kq + 8 * lg]);
                        bv[p][r] = *reinterpret_cast<const bf16x8*>(&Bs[cur][p][(wc + r * 16 + li) * LDK + kq + 8 * lg]);
                    }
    #pragma unroll
                for (int r = 0; r < 4; ++r)
    #pragma unroll
                    for (int c = 0; c < 4; ++c) {
                        if constexpr (NSPLIT == 3) {
                            acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bv[1][c], av[0][r], acc[r][c], 0, 0, 0);
                            acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bv[0][c], av[1][r], acc[r][c], 0, 0, 0);
                        }
                        acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bv[0][c], av[0][r], acc[r][c], 0, 0, 0);
                    }
            }
            __syncthreads();
        }
        // acc[r][c][v] = C(row0 + wr + r*16 + li, col0 + wc + c*16 + 4*lg + v)
        gemm_epilogue<float, 4>(acc, C, g, row0 + wr + li, col0 + wc, lg, [](int lg_, int v_) { return 4 * lg_ + v_; });
    }
}

// The bf16 products of gemm_bf16_kernel with A and B read from the split
// planes (kGemmPlanes launches): no conversion in the kernel, and every
// load is a 16-byte vector of 8 consecutive k of one row.  128 x 128 tiles,
// k-tiles of 32 = 4 octets.  Thread t stages row t % 128, octets
// 2 (t / 128) and 2 (t / 128) + 1, of each plane of A and B: 8 (NSPLIT = 3)
// or 4 loads per k-tile, held in registers while the current k-tile is
// multiplied, then stored to the other LDS buffer, laid out [octet][row]
// (the 16 lanes of an MFMA operand read 256 contiguous bytes).  Rows past M
// (N) are read clamped to row M - 1 (N - 1): they reach only C entries that
// are not stored.  In the last k-tile, octets past K are not read and
// entries past K are zeroed (the planes hold no factor entries there).
// Same products, order and accumulator layout as gemm_bf16_kernel.
template <int NSPLIT>
__global__ __launch_bounds__(kThreads, NSPLIT == 6 ? 1 : 2) void gemm_bf16p_kernel(
    const GemmJob* __restrict__ jobs, int njobs, const __bf16* __restrict__ Pbase, int64_t pstride,
    float* __restrict__ U, float* __restrict__ L, int ntiles) {
    static_assert(NSPLIT == 1 || NSPLIT == 3 || NSPLIT == 6, "1, 3 or 6 bf16 products");
    constexpr int BT = 128, KT = 32, NO = KT / 8;
    // planes read: hi (1 product); hi, mid (3: mid is BF16X3's lo); hi, mid, lo (6)
    constexpr int NP = NSPLIT == 1 ? 1 : NSPLIT == 3 ? 2 : 3;
    __shared__ bf16x8 As[2][NP][NO * BT];
    __shared__ bf16x8 Bs[2][NP][NO * BT];
    for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        int lo = 0, hi = njobs - 1;
        const int bid = xcd_tile(tile, ntiles);
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
        float* __restrict__ C = (g.c_in_arena ? U : L) + g.c_off;
        const int row0 = ti * BT, col0 = tj * BT;
        const int tid = threadIdx.x;
        const int wave = tid >> 6, lane = tid & 63;
        const int wr = (wave >> 1) * 64, wc = (wave & 1) * 64;
        const int li = lane & 15, lg = lane >> 4;
        const int r_st = tid & (BT - 1), o_st = 2 * (tid >> 7);
        const int64_t ia = min(row0 + r_st, g.M - 1), jb = min(col0 + r_st, g.N - 1);
        // full k-tiles go straight into LDS (buffer_load_dwordx4 ... lds):
        // wavefront load i covers octet (2 wave + i) / 2 of the k-tile and rows
        // 64 ((2 wave + i) % 2) + lane, 1 KB of consecutive rows (clamped to
        // the job's last row, as the staged loads below)
        const int nfull = g.K / KT, nt = (g.K + KT - 1) / KT;
        int dra[2], drb[2];   // clamped row of this lane in load i
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const int idx = 2 * wave + i;
            dra[i] = min(row0 + 64 * (idx & 1) + lane, g.M - 1);
            drb[i] = min(col0 + 64 * (idx & 1) + lane, g.N - 1);
        }
        auto issue = [&](int buf, int t) {
            const int o0 = t * NO;
#pragma unroll
            for (int p = 0; p < NP; ++p) {
                const int64_t ba = p * pstride + g.a_pl + (int64_t)o0 * 8 * g.lda;
                const int64_t bb = p * pstride + g.b_pl + (int64_t)o0 * 8 * g.ldb;
                const auto ra_ = buffer_rsrc(Pbase + ba, ((p + 1) * pstride - ba) * 2);
                const auto rb_ = buffer_rsrc(Pbase + bb, ((p + 1) * pstride - bb) * 2);
#pragma unroll
                for (int i = 0; i < 2; ++i) {
                    const int idx = 2 * wave + i, ol = idx >> 1, rbase = 64 * (idx & 1);
                    __builtin_amdgcn_raw_ptr_buffer_load_lds(ra_, (lds_void_ptr)&As[buf][p][ol * BT + rbase], 16,
                                                             (uint32_t)(ol * g.lda + dra[i]) * 16u, 0, 0, 0);
                    __builtin_amdgcn_raw_ptr_buffer_load_lds(rb_, (lds_void_ptr)&Bs[buf][p][ol * BT + rbase], 16,
                                                             (uint32_t)(ol * g.ldb + drb[i]) * 16u, 0, 0, 0);
                }
            }
        };
        // 16-byte vectors: octet o of row i at base + o * ld + i
        const bf16x8* __restrict__ pa[NP];
        const bf16x8* __restrict__ pb[NP];
#pragma unroll
        for (int p = 0; p < NP; ++p) {
            pa[p] = reinterpret_cast<const bf16x8*>(Pbase + p * pstride + g.a_pl) + ia;
            pb[p] = reinterpret_cast<const bf16x8*>(Pbase + p * pstride + g.b_pl) + jb;
        }
        // partial last k-tile (K not a multiple of 32) through registers,
        // entries past K zeroed
        auto stage_tail = [&](int buf, int k0) {
            const int o0 = k0 / 8 + o_st;
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                const int kk = 8 * (o0 + q);   // first k of the octet
#pragma unroll
                for (int p = 0; p < NP; ++p) {
                    bf16x8 x = {}, y = {};
                    if (kk < g.K) {
                        x = pa[p][(int64_t)(o0 + q) * g.lda];
                        y = pb[p][(int64_t)(o0 + q) * g.ldb];
#pragma unroll
                        for (int e = 0; e < 8; ++e)
                            if (kk + e >= g.K) {
                                x[e] = (__bf16)0.0f;
                                y[e] = (__bf16)0.0f;
                            }
                    }
                    As[buf][p][(o_st + q) * BT + r_st] = x;
                    Bs[buf][p][(o_st + q) * BT + r_st] = y;
                }
            }
        };
        f32x4 acc[4][4];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) acc[a][b] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
        if (nfull > 0) issue(0, 0);
        else stage_tail(0, 0);
        for (int t = 0; t < nt; ++t) {
            const int cur = t & 1;
            // k-tile t's loads are the only ones outstanding: wait for them;
            // the barrier makes them visible and ends the reads of k-tile t - 1
            __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
            __syncthreads();
            if (t + 1 < nfull) issue(cur ^ 1, t + 1);
            else if (t + 1 < nt) stage_tail(cur ^ 1, (t + 1) * KT);
            bf16x8 av[NP][4], bv[NP][4];
#pragma unroll
            for (int p = 0; p < NP; ++p)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    av[p][r] = As[cur][p][lg * BT + wr + r * 16 + li];
                    bv[p][r] = Bs[cur][p][lg * BT + wc + r * 16 + li];
                }
#pragma unroll
            for (int r = 0; r < 4; ++r)
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    // small terms first
                    if constexpr (NSPLIT == 6) {
                        acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bv[2][c], av[0][r], acc[r][c], 0, 0, 0);
                        acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bv[0][c], av[2][r], acc[r][c], 0, 0, 0);
                        acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bv[1][c], av[1][r], acc[r][c], 0, 0, 0);
                    }
                    if constexpr (NSPLIT >= 3) {
                        acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bv[1][c], av[0][r], acc[r][c], 0, 0, 0);
                        acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bv[0][c], av[1][r], acc[r][c], 0, 0, 0);
                    }
                    acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bv[0][c], av[0][r], acc[r][c], 0, 0, 0);
                }
        }
        // the next tile's first loads reuse the buffers
        __syncthreads();
        // acc[r][c][v] = C(row0 + wr + r*16 + li, col0 + wc + c*16 + 4*lg + v)
        gemm_epilogue<float, 4>(acc, C, g, row0 + wr + li, col0 + wc, lg,
                                [](int lg_, int v_) { return 4 * lg_ + v_; });
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

// N entries of one row of a front panel P (n rows, column-major), columns
// col, col + 1, ... of which only the first `valid` are real: the loads of
// the others repeat the last real column (in range) and read as zero.
template <int N, typename T>
__device__ __forceinline__ void issue_row(T (&v)[N], const T* __restrict__ P, int row, int col, int valid,
                                          int64_t n) {
    if (valid >= N) {
        issue_strided<N>(v, P + row + (int64_t)col * n, n);
    } else {
#pragma unroll
        for (int k = 0; k < N; ++k) v[k] = P[row + (int64_t)(col + min(k, valid - 1)) * n];
    }
}
template <typename T>
__device__ __forceinline__ void issue_row32(T (&v)[32], const T* __restrict__ P, int row, int col, int valid,
                                            int64_t n) {
    issue_row<32>(v, P, row, col, valid, n);
}

// acc[k], k < 32, summed over the 64 lanes of the wavefront: afterwards
// acc[0] of lanes l and l + 32 holds the sum of acc[l & 31] (transpose
// reduction within each half, then the halves added).
__device__ __forceinline__ void wave_transpose_sum32(double (&acc)[32], int lane) {
#pragma unroll
    for (int s = 16; s >= 1; s >>= 1) {
        const bool upper = (lane & s) != 0;
#pragma unroll
        for (int k = 0; k < s; ++k) {
            double lo = acc[k], hi = acc[k + s];
            asm volatile("" : "+v"(lo), "+v"(hi));
            const double send = upper ? lo : hi;
            const double keep = upper ? hi : lo;
            acc[k] = keep + __shfl_xor(send, s, 64);
        }
    }
    acc[0] += __shfl_xor(acc[0], 32, 64);
}

// Column sums of a kSolveThreads-thread workgroup laid out as (row r =
// tid & 255, column half h = tid >> 8), acc[k] = this thread's term of
// column 32 h + k: reduced over the wavefront, then over the four
// wavefronts of a half in a fixed order.  Returns the sum of column tid
// for tid < 64 (0 elsewhere).  redc: LDS scratch of 8 * 32.
__device__ __forceinline__ double column_sums64(double (&acc)[32], double* redc, int tid) {
    const int lane = tid & 63, wave = tid >> 6;
    wave_transpose_sum32(acc, lane);
    if (lane < 32) redc[wave * 32 + lane] = acc[0];
    __syncthreads();
    double sum = 0.0;
    if (tid < 64) {
        const double* rr = redc + 4 * (tid >> 5) * 32 + (tid & 31);
        sum = (rr[0] + rr[32]) + (rr[64] + rr[96]);
    }
    return sum;
}

// Small fronts (solve_small_front: at most kSolveSmall pivots, one
// diagonal block, and few rows), one workgroup per front and one launch
// per level.  Front vectors of at most
// vec_cap (<= kSmallVec) entries live in LDS while the children's update
// vectors are added; longer ones are accumulated in place in xp and the
// update vector.

// Forward substitution L y = x over one level, multifrontal form of
// cs_lsolve (cs.cpp:1086-1101): the front vector (pivot entries from xp,
// update entries zero) gathers the children's update vectors through their
// relative maps, child by child; then y = inv(L_D) v (stored inverse, four
// threads per row), stored in Y, and the update rows below subtract L y,
// one row per thread with its c loads in flight together.
template <typename T>
__global__ __launch_bounds__(kThreads) void fwd_kernel(const int* __restrict__ list,
                                                            const FrontDesc* __restrict__ fronts,
                                                            const int* __restrict__ kids,
                                                            const int* __restrict__ rel, const int* __restrict__ q,
                                                            const T* __restrict__ L, const T* __restrict__ inv,
                                                            double* __restrict__ xp, double* __restrict__ Y,
                                                            double* __restrict__ su,
                                                            const int64_t* __restrict__ suofs, int vec_cap) {
    __shared__ double fv[kSmallVec];
    __shared__ double yb[64];
    __shared__ double red[4 * 64];
    const int f = list[blockIdx.x];
    const FrontDesc d = fronts[f];
    const int n = d.n, c = d.c, k = d.k, tid = threadIdx.x;
    const T* __restrict__ P = L + d.lofs;
    double* __restrict__ u = su + suofs[f];
    const bool inl = n <= vec_cap;
    for (int t = tid; t < n; t += kThreads) {
        if (inl) fv[t] = t < c ? xp[k + t] : 0.0;
        else if (t >= c) u[t - c] = 0.0;
    }
    __syncthreads();
    for (int t = d.kid_begin; t < d.kid_end; ++t) {
        const int ch = kids[t];
        const FrontDesc cd = fronts[ch];
        const int* __restrict__ r = rel + cd.relofs;
        const double* __restrict__ cu = su + suofs[ch];
        for (int i = tid; i < cd.m; i += kThreads) {
            const int pos = r[i];
            const double v = cu[i];
            if (inl) fv[pos] += v;
            else if (pos < c) xp[k + pos] += v;
            else u[pos - c] += v;
        }
        __syncthreads();
    }
    // y = W v, W(l, r) = w[l + 64 r] = inv(L_D)(r, l): thread (r, part)
    // sums columns [16 part, 16 part + 16) of row r
    const T* __restrict__ w = inv + d.dinv;
    if (tid < 4 * 64) {
        const int r = tid >> 2, l0 = 16 * (tid & 3);
        double s = 0.0;
        if (r < c) {
            T wv[16];
#pragma unroll
            for (int l = 0; l < 16; ++l) wv[l] = w[min(l0 + l, r) + 64 * r];
            keep_in_flight(wv);
#pragma unroll
            for (int l = 0; l < 16; ++l)
                if (l0 + l <= r) s += to_d(wv[l]) * (inl ? fv[l0 + l] : xp[k + l0 + l]);
        }
        red[tid] = s;
    }
    __syncthreads();
    if (tid < c) {
        const double* rr = red + 4 * tid;
        const double y = (rr[0] + rr[1]) + (rr[2] + rr[3]);
        yb[tid] = y;
        Y[k + tid] = y;
    }
    __syncthreads();
    for (int i = c + tid; i < n; i += kThreads) {
        T lv[64];
#pragma unroll
        for (int j = 0; j < 64; ++j) lv[j] = P[i + (int64_t)min(j, c - 1) * n];
        keep_in_flight(lv);
        double acc = 0.0;
#pragma unroll
        for (int j = 0; j < 64; ++j)
            if (j < c) acc += to_d(lv[j]) * yb[j];
        if (inl) fv[i] -= acc;
        else u[i - c] -= acc;
    }
    if (inl) {
        __syncthreads();
        for (int t = c + tid; t < n; t += kThreads) u[t - c] = fv[t];
    }
}

// Backward substitution L^T x = y over one level (cs_ltsolve,
// cs.cpp:1104-1119), one workgroup per front: the dot products of the c
// pivot columns with the solved update rows (x read through the row list
// q), thread (row, column half) with its 32 loads in flight, summed over
// the rows in a fixed order; then x = inv(L_D)^T v with the stored inverse,
// one thread per pivot.
template <typename T>
__global__ __launch_bounds__(kSolveThreads) void bwd_kernel(const int* __restrict__ list,
                                                            const FrontDesc* __restrict__ fronts,
                                                            const int* __restrict__ qall, const T* __restrict__ L,
                                                            const T* __restrict__ inv, double* __restrict__ xp,
                                                            const double* __restrict__ Y) {
    __shared__ double redc[8 * 32];
    __shared__ double vb[64];
    const FrontDesc d = fronts[list[blockIdx.x]];
    const int n = d.n, c = d.c, k = d.k, tid = threadIdx.x;
    const T* __restrict__ P = L + d.lofs;
    const int* __restrict__ q = qall + d.qofs;
    const int r = tid & 255, h = tid >> 8, col = 32 * h, cvalid = min(32, c - col);
    double acc[32];
#pragma unroll
    for (int kk = 0; kk < 32; ++kk) acc[kk] = 0.0;
    if (cvalid > 0) {
        for (int r0 = c; r0 < n; r0 += 256) {
            T lv[32];
            issue_row32(lv, P, min(r0 + r, n - 1), col, cvalid, n);
            const double xv = r0 + r < n ? xp[q[r0 + r]] : 0.0;
            keep_in_flight(lv);
#pragma unroll
            for (int kk = 0; kk < 32; ++kk) acc[kk] += (kk < cvalid ? to_d(lv[kk]) : 0.0) * xv;
        }
    }
    const double sum = column_sums64(acc, redc, tid);
    if (tid < c) vb[tid] = Y[k + tid] - sum;
    __syncthreads();
    if (tid < c) {
        // x_i = sum_{r >= i} inv(L_D)(r, i) v_r, W(i, r) = w[i + 64 r]
        const T* __restrict__ w = inv + d.dinv;
        T wv[64];
#pragma unroll
        for (int rr = 0; rr < 64; ++rr) wv[rr] = w[tid + 64 * min(max(rr, tid), c - 1)];
        keep_in_flight(wv);
        double x = 0.0;
#pragma unroll
        for (int rr = 0; rr < 64; ++rr)
            if (rr >= tid && rr < c) x += to_d(wv[rr]) * vb[rr];
        xp[k + tid] = x;
    }
}

// ------------------------------------------------- solves of large fronts
// Fronts other than the small ones (solve_small_front): the children's update vectors
// are gathered first (forward) or the update rows applied first (backward),
// then the kChainBW-column blocks of all of a level's large fronts are
// solved in one launch, each block by its own workgroup, which waits inside
// the launch for the blocks it depends on (chained solves, below).

// Forward, step 0 of a large front: its front vector (pivot entries from
// xp, update entries zero) gathers the children's update vectors through
// their relative maps, child by child (the first phase of fwd_kernel).
// Fronts of at most vec_cap (<= kGatherVec) rows accumulate in LDS, longer
// ones in place; each thread keeps four entries of a child in flight.
constexpr int kGatherThreads = 1024;
__global__ __launch_bounds__(kGatherThreads) void fwd_gather_kernel(const int* __restrict__ list,
                                                                    const FrontDesc* __restrict__ fronts,
                                                                    const int* __restrict__ kids,
                                                                    const int* __restrict__ rel,
                                                                    double* __restrict__ xp, double* __restrict__ su,
                                                                    const int64_t* __restrict__ suofs, int vec_cap) {
    __shared__ double fv[kGatherVec];
    const int f = list[blockIdx.x];
    const FrontDesc d = fronts[f];
    const int n = d.n, c = d.c, k = d.k, tid = threadIdx.x;
    double* __restrict__ u = su + suofs[f];
    const bool inl = n <= vec_cap;
    for (int t = tid; t < n; t += kGatherThreads) {
        if (inl) fv[t] = t < c ? xp[k + t] : 0.0;
        else if (t >= c) u[t - c] = 0.0;
    }
    __syncthreads();
    for (int t = d.kid_begin; t < d.kid_end; ++t) {
        const int ch = kids[t];
        const FrontDesc cd = fronts[ch];
        const int* __restrict__ r = rel + cd.relofs;
        const double* __restrict__ cu = su + suofs[ch];
        for (int i0 = tid; i0 < cd.m; i0 += 4 * kGatherThreads) {
            int pos[4];
            double v[4];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int i = min(i0 + e * kGatherThreads, cd.m - 1);
                pos[e] = r[i];
                v[e] = cu[i];
            }
            keep_in_flight(pos);
            keep_in_flight(v);
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                if (i0 + e * kGatherThreads >= cd.m) continue;
                if (inl) fv[pos[e]] += v[e];
                else if (pos[e] < c) xp[k + pos[e]] += v[e];
                else u[pos[e] - c] += v[e];
            }
        }
        __syncthreads();
    }
    if (inl) {
        for (int t = tid; t < n; t += kGatherThreads) {
            if (t < c) xp[k + t] = fv[t];
            else u[t - c] = fv[t];
        }
    }
}

// Rows r0 and r1 of panel P (n rows, c columns), columns col .. col + 15;
// columns from c on repeat column c - 1 (in range) and are multiplied by
// zero by the caller.
template <typename T>
__device__ __forceinline__ void chain_tile(T (&lo)[16], T (&hi)[16], const T* __restrict__ P, int r0, int r1,
                                           int col, int c, int64_t n) {
    // no branch around the loads: a later wait for some of them can then
    // leave the rest in flight
#pragma unroll
    for (int e = 0; e < 16; ++e) {
        const int64_t off = (int64_t)min(col + e, c - 1) * n;
        lo[e] = P[off + r0];
        hi[e] = P[off + r1];
    }
}

// Forward over the large fronts of one level: per front a chain of
// kChainBW-column blocks inside the launch (TRI items), and the update rows
// below the pivots (RECT items).  An item owns rows [a, a + 128) of its
// front: thread (lane, wave) holds rows a + lane and a + 64 + lane of each
// column block j < J, columns 128 j + 16 wave + e, and adds their products
// with y_j (waited for by the first 128 threads, in LDS) to two row sums;
// the eight wavefronts' sums of a row are added in a fixed order.  Block
// j + 1's entries are loaded while block j is waited for and used, and a
// TRI item loads its block inverse (thread (r, q): X(r, 32 q + e)) first, so
// all of its data is in registers when the last y_j it needs arrives.
//   TRI:  y_b = X_b (v_b - sum_{j<b} L_bj y_j), stored in Y (chain_store);
//   RECT: u -= sum_j L_rj y_j.
template <typename T>
__global__ __launch_bounds__(kSolveThreads) void fwd_chain_kernel(const SolveItem* __restrict__ items,
                                                                  int* __restrict__ ctr, int* __restrict__ status,
                                                                  int spin, const T* __restrict__ L,
                                                                  const T* __restrict__ X256,
                                                                  const double* __restrict__ xp,
                                                                  double* __restrict__ Y, double* __restrict__ su) {
    constexpr int BW = kChainBW;
    __shared__ double yl[2][BW];
    __shared__ double red[8][BW];
    __shared__ double zb[BW];
    __shared__ int sh;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const SolveItem it = items[chain_ticket(ctr, &sh)];
    const bool tri = it.first != 0;
    const int n = it.n, c = it.c, row0 = it.a;
    const int nrow = min(tri ? c : n, row0 + BW) - row0;
    const int J = tri ? row0 / BW : (c + BW - 1) / BW;
    const T* __restrict__ P = L + it.lofs;
    const int r = tid & (BW - 1), q = tid >> 7;
    double rhs = 0.0;
    if (tid < nrow) rhs = tri ? xp[it.k + row0 + tid] : su[it.su + row0 + tid - c];
    const int ra = min(row0 + lane, n - 1), rb = min(row0 + 64 + lane, n - 1);
    double acc0 = 0.0, acc1 = 0.0;
    // block j's products with y_j (tile tl / th), y_j waited for by the
    // first 128 threads
    auto apply = [&](int j, const T (&tl)[16], const T (&th)[16]) {
        double* yb = yl[j & 1];
        if (tid < BW) yb[tid] = tid < c - BW * j ? chain_wait(Y + it.k + BW * j + tid, status, spin) : 0.0;
        __syncthreads();
        double s0 = 0.0, s1 = 0.0;
#pragma unroll
        for (int e = 0; e < 16; ++e) {
            const double y = yb[16 * wave + e];
            s0 += to_d(tl[e]) * y;
            s1 += to_d(th[e]) * y;
        }
        acc0 += s0;
        acc1 += s1;
    };
    T cl[16], ch[16];
    if (J > 0) chain_tile(cl, ch, P, ra, rb, 16 * wave, c, n);
    for (int j = 0; j + 1 < J; ++j) {
        T nl[16], nh[16];
        chain_tile(nl, nh, P, ra, rb, BW * (j + 1) + 16 * wave, c, n);
        apply(j, cl, ch);
#pragma unroll
        for (int e = 0; e < 16; ++e) {
            cl[e] = nl[e];
            ch[e] = nh[e];
        }
    }
    // the block inverse goes in flight before the last wait
    T xm[32];
    if (tri) issue_strided<32>(xm, X256 + it.pk + r + (int64_t)(32 * kSolveBW) * q, kSolveBW);
    if (J > 0) apply(J - 1, cl, ch);
    red[wave][lane] = acc0;
    red[wave][64 + lane] = acc1;
    __syncthreads();
    double z = 0.0;
    if (tid < nrow) {
        const double sum = ((red[0][tid] + red[1][tid]) + (red[2][tid] + red[3][tid])) +
                           ((red[4][tid] + red[5][tid]) + (red[6][tid] + red[7][tid]));
        z = rhs - sum;
    }
    if (!tri) {
        if (tid < nrow) su[it.su + row0 + tid - c] = z;
        return;
    }
    if (tid < BW) zb[tid] = z;
    __syncthreads();
    double p = 0.0;
#pragma unroll
    for (int e = 0; e < 32; ++e) p += to_d(xm[e]) * zb[32 * q + e];
    red[q][r] = p;
    __syncthreads();
    if (tid < nrow) chain_store(Y + it.k + row0 + tid, (red[0][tid] + red[1][tid]) + (red[2][tid] + red[3][tid]));
}

// Backward, once per level before the blocks: workgroup (front, 64 pivot
// columns [a, a + 64), update rows [j0, j0 + kSolveUpdRows)) forms the dot
// products of its columns with those rows of the front, whose solution
// entries (final: they belong to ancestors) are read through the row list
// q, into partial-sum slot `slot`.  Rows go 256 at a time, the next
// chunk's loads issued before the current one is used.
template <typename T>
__global__ __launch_bounds__(kSolveThreads) void bwd_upd_kernel(const SolveItem* __restrict__ items,
                                                                const int* __restrict__ qall,
                                                                const T* __restrict__ L,
                                                                const double* __restrict__ xp,
                                                                double* __restrict__ partial) {
    __shared__ double redc[8 * 32];
    const SolveItem it = items[blockIdx.x];
    const int c = it.c, c0 = it.a, n = min(it.n, it.j0 + kSolveUpdRows);
    const T* __restrict__ P = L + it.lofs;
    const int* __restrict__ q = qall + it.qofs;
    const int tid = threadIdx.x, r = tid & 255, h = tid >> 8;
    const int col = c0 + 32 * h, cvalid = min(32, c - col);
    double acc[32];
#pragma unroll
    for (int k = 0; k < 32; ++k) acc[k] = 0.0;
    if (cvalid > 0 && sizeof(T) == 4) {
        T cur[32];
        int r0 = it.j0;
        issue_row32(cur, P, min(r0 + r, n - 1), col, cvalid, it.n);
        double xv = r0 + r < n ? xp[q[r0 + r]] : 0.0;
        for (; r0 < n; r0 += 256) {
            T nxt[32];
            double xn = 0.0;
            const bool more = r0 + 256 < n;
            if (more) {
                issue_row32(nxt, P, min(r0 + 256 + r, n - 1), col, cvalid, it.n);
                xn = r0 + 256 + r < n ? xp[q[r0 + 256 + r]] : 0.0;
            }
#pragma unroll
            for (int k = 0; k < 32; ++k) acc[k] += (k < cvalid ? to_d(cur[k]) : 0.0) * xv;
            if (more) {
#pragma unroll
                for (int k = 0; k < 32; ++k) cur[k] = nxt[k];
                xv = xn;
            }
        }
    } else if (cvalid > 0) {
        // FP64: one chunk in flight (two would not fit the register budget)
        for (int r0 = it.j0; r0 < n; r0 += 256) {
            T cur[32];
            issue_row32(cur, P, min(r0 + r, n - 1), col, cvalid, it.n);
            const double xv = r0 + r < n ? xp[q[r0 + r]] : 0.0;
            keep_in_flight(cur);
#pragma unroll
            for (int k = 0; k < 32; ++k) acc[k] += (k < cvalid ? to_d(cur[k]) : 0.0) * xv;
        }
    }
    const double sum = column_sums64(acc, redc, tid);
    if (tid < kSolveCols) partial[(int64_t)it.slot * kSolveCols + tid] = sum;
}

// Backward, after BWD_UPD: Y(c0 + i) -= sum of the chunk's partial-sum
// slots, added in slot order (one workgroup of kSolveCols threads per
// column chunk).
__global__ __launch_bounds__(kSolveCols) void bwd_upd_sum_kernel(const SolveItem* __restrict__ items,
                                                                 const double* __restrict__ partial,
                                                                 double* __restrict__ Y) {
    const SolveItem it = items[blockIdx.x];
    const int i = threadIdx.x;
    if (it.a + i >= it.c) return;
    const double* __restrict__ p = partial + (int64_t)it.slot * kSolveCols + i;
    const double yold = Y[it.k + it.a + i];
    double s = 0.0;
    for (int t0 = 0; t0 < it.nslot; t0 += 16) {
        double v[16];
#pragma unroll
        for (int t = 0; t < 16; ++t) v[t] = p[(int64_t)min(t0 + t, it.nslot - 1) * kSolveCols];
        keep_in_flight(v);
#pragma unroll
        for (int t = 0; t < 16; ++t)
            if (t0 + t < it.nslot) s += v[t];
    }
    Y[it.k + it.a + i] = yold - s;
}

// acc[k], k < 16, summed over the 64 lanes of the wavefront: afterwards
// acc[0] of lane l holds the sum of acc[l & 15] (transpose reduction within
// each 16-lane group, then the groups added).
__device__ __forceinline__ void wave_transpose_sum16(double (&acc)[16], int lane) {
#pragma unroll
    for (int s = 8; s >= 1; s >>= 1) {
        const bool upper = (lane & s) != 0;
#pragma unroll
        for (int k = 0; k < s; ++k) {
            double lo = acc[k], hi = acc[k + s];
            asm volatile("" : "+v"(lo), "+v"(hi));
            const double send = upper ? lo : hi;
            const double keep = upper ? hi : lo;
            acc[k] = keep + __shfl_xor(send, s, 64);
        }
    }
    acc[0] += __shfl_xor(acc[0], 16, 64);
    acc[0] += __shfl_xor(acc[0], 32, 64);
}

// Backward over the large fronts of one level, a chain per front from its
// last kChainBW-column block (see SolveStep; the update rows were applied
// by BWD_UPD / BWD_UPD_SUM).  Item = pivot columns [a, a + 128): thread
// (lane, wave) holds rows 128 j + lane and 128 j + 64 + lane of each later
// block j, columns a + 16 wave + e, and adds their products with x_j
// (waited for by the first 128 threads, in LDS) to 16 column sums, reduced
// over the wavefront at the end.  Then x_b = X_b^T z_b with thread (r, q)
// holding X(32 q + e, r), a contiguous run of column r of the inverse,
// stored in xp (chain_store).
template <typename T>
__global__ __launch_bounds__(kSolveThreads) void bwd_chain_kernel(const SolveItem* __restrict__ items,
                                                                  int* __restrict__ ctr, int* __restrict__ status,
                                                                  int spin, const T* __restrict__ L,
                                                                  const T* __restrict__ X256, double* __restrict__ xp,
                                                                  const double* __restrict__ Y) {
    constexpr int BW = kChainBW;
    __shared__ double xl[2][BW];
    __shared__ double red[4][BW];
    __shared__ double zb[BW];
    __shared__ int sh;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const SolveItem it = items[chain_ticket(ctr, &sh)];
    const int n = it.n, c = it.c, col0 = it.a;
    const int bw = min(BW, c - col0), nb = (c + BW - 1) / BW, jend = col0 / BW;
    const T* __restrict__ P = L + it.lofs;
    const int r = tid & (BW - 1), q = tid >> 7;
    const double yv = tid < bw ? Y[it.k + col0 + tid] : 0.0;
    const int cb = col0 + 16 * wave;
    double acc[16];
#pragma unroll
    for (int e = 0; e < 16; ++e) acc[e] = 0.0;
    // block j's products with x_j (tile tl / th), x_j waited for by the
    // first 128 threads
    auto apply = [&](int j, const T (&tl)[16], const T (&th)[16]) {
        double* xb = xl[j & 1];
        if (tid < BW) xb[tid] = tid < c - BW * j ? chain_wait(xp + it.k + BW * j + tid, status, spin) : 0.0;
        __syncthreads();
        const double x0 = xb[lane], x1 = xb[64 + lane];
#pragma unroll
        for (int e = 0; e < 16; ++e) {
            acc[e] += to_d(tl[e]) * x0;
            acc[e] += to_d(th[e]) * x1;
        }
    };
    // later blocks j = nb - 1 down to jend + 1 (the rows of the last one
    // clamped to the front, those of the others are below c <= n)
    T cl[16], ch[16];
    if (nb - 1 > jend) chain_tile(cl, ch, P, min(BW * (nb - 1) + lane, n - 1), min(BW * (nb - 1) + 64 + lane, n - 1), cb, c,
                                     n);
    for (int j = nb - 1; j - 1 > jend; --j) {
        T nl[16], nh[16];
        chain_tile(nl, nh, P, BW * (j - 1) + lane, BW * (j - 1) + 64 + lane, cb, c, n);
        apply(j, cl, ch);
#pragma unroll
        for (int e = 0; e < 16; ++e) {
            cl[e] = nl[e];
            ch[e] = nh[e];
        }
    }
    // the block inverse goes in flight before the last wait
    T xm[32];
    issue_strided<32>(xm, X256 + it.pk + 32 * q + (int64_t)kSolveBW * r, 1);
    if (nb - 1 > jend) apply(jend + 1, cl, ch);
    wave_transpose_sum16(acc, lane);
    if (lane < 16) red[0][16 * wave + lane] = acc[0];
    __syncthreads();
    if (tid < BW) zb[tid] = tid < bw ? yv - red[0][tid] : 0.0;
    __syncthreads();
    double p = 0.0;
#pragma unroll
    for (int e = 0; e < 32; ++e) p += to_d(xm[e]) * zb[32 * q + e];
    red[q][r] = p;
    __syncthreads();
    if (tid < bw) chain_store(xp + it.k + col0 + tid, (red[0][tid] + red[1][tid]) + (red[2][tid] + red[3][tid]));
}

// -------------------------------------------------------------- refinement
__global__ void gather_kernel(const double* __restrict__ vals, const int* __restrict__ src,
                              double* __restrict__ dst, int nnz) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < nnz) dst[t] = vals[src[t]];
}

// r = b - A x, FP64 (cs_gaxpy, cs.cpp:1069-1083, on the full symmetric
// matrix), 16 lanes per row: the row's entries are read coalesced and the
// lanes' partial sums added by a butterfly in a fixed order
constexpr int kRowLanes = 16;
__global__ __launch_bounds__(kThreads) void residual_kernel(const int* __restrict__ rp, const int* __restrict__ ci,
                                                            const double* __restrict__ av,
                                                            const double* __restrict__ x,
                                                            const double* __restrict__ b, double* __restrict__ r,
                                                            int n) {
    const int i = (blockIdx.x * blockDim.x + threadIdx.x) / kRowLanes, l = threadIdx.x % kRowLanes;
    if (i >= n) return;   // the whole 16-lane group
    double s = 0.0;
    for (int p = rp[i] + l; p < rp[i + 1]; p += kRowLanes) s += av[p] * x[ci[p]];
#pragma unroll
    for (int o = kRowLanes / 2; o >= 1; o >>= 1) s += __shfl_xor(s, o, kRowLanes);
    if (l == 0) r[i] = b[i] - s;
}

__global__ __launch_bounds__(kThreads) void rowabs_kernel(const int* __restrict__ rp,
                                                          const double* __restrict__ av,
                                                          double* __restrict__ out, int n) {
    const int i = (blockIdx.x * blockDim.x + threadIdx.x) / kRowLanes, l = threadIdx.x % kRowLanes;
    if (i >= n) return;
    double s = 0.0;
    for (int p = rp[i] + l; p < rp[i + 1]; p += kRowLanes) s += fabs(av[p]);
#pragma unroll
    for (int o = kRowLanes / 2; o >= 1; o >>= 1) s += __shfl_xor(s, o, kRowLanes);
    if (l == 0) out[i] = s;
}

// per-segment max |v| (seg[i] = segment of entry i, segments contiguous):
// a wavefront whose entries share one segment reduces first, else each lane
// goes to its own segment; integer max on the bit patterns of non-negative
// doubles (NaN counts as +inf)
__global__ __launch_bounds__(kThreads) void maxabs_seg_kernel(const double* __restrict__ v, const int* __restrict__ seg,
                                                              int n, unsigned long long* __restrict__ out) {
    const int lane = threadIdx.x & 63;
    for (int base = (blockIdx.x * blockDim.x + threadIdx.x) - lane; base < n; base += gridDim.x * blockDim.x) {
        const int i = base + lane;
        double a = 0.0;
        int sg = -1;
        if (i < n) {
            a = fabs(v[i]);
            if (a != a) a = __builtin_huge_val();
            sg = seg[i];
        }
        const int last = min(base + 63, n - 1);
        const int s0 = __shfl(sg, 0, 64), s1 = seg[last];
        if (s0 == s1) {
            double m = a;
            for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_down(m, off, 64));
            if (lane == 0) atomicMax(&out[s0], (unsigned long long)__double_as_longlong(m));
        } else if (i < n) {
            atomicMax(&out[sg], (unsigned long long)__double_as_longlong(a));
        }
    }
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

// ------------------------------------------- conjugate-gradient refinement
// y = A x (FP64 CSR SpMV, 16 lanes per row)
__global__ __launch_bounds__(kThreads) void spmv_kernel(const int* __restrict__ rp, const int* __restrict__ ci,
                                                        const double* __restrict__ av, const double* __restrict__ x,
                                                        double* __restrict__ y, int n) {
    const int i = (blockIdx.x * blockDim.x + threadIdx.x) / kRowLanes, l = threadIdx.x % kRowLanes;
    if (i >= n) return;
    double s = 0.0;
    for (int p = rp[i] + l; p < rp[i + 1]; p += kRowLanes) s += av[p] * x[ci[p]];
#pragma unroll
    for (int o = kRowLanes / 2; o >= 1; o >>= 1) s += __shfl_xor(s, o, kRowLanes);
    if (l == 0) y[i] = s;
}

// partial[b] = sum of a_i c_i over the entries of workgroup b (grid-stride,
// fixed order), then dot_final_kernel sums the kDotParts partials in a
// fixed tree: the dot product is the same on every run
__global__ __launch_bounds__(kThreads) void dot_partial_kernel(const double* __restrict__ a,
                                                               const double* __restrict__ c, int n,
                                                               double* __restrict__ partial) {
    __shared__ double red[kThreads / 64];
    double s = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) s += a[i] * c[i];
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) partial[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}
__global__ __launch_bounds__(kDotParts) void dot_final_kernel(const double* __restrict__ partial,
                                                              double* __restrict__ out) {
    __shared__ double red[kDotParts];
    red[threadIdx.x] = partial[threadIdx.x];
    __syncthreads();
    for (int w = kDotParts / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0];
}
// CG direction: p = z (first) or p = z + (rz / rz_old) p; scal = {rz, rz_old, pq}
__global__ void cg_dir_kernel(double* __restrict__ p, const double* __restrict__ z, const double* __restrict__ scal,
                              int n, int first) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    p[i] = first ? z[i] : z[i] + (scal[0] / scal[1]) * p[i];
}
// x += (rz / pq) p
__global__ void cg_step_kernel(double* __restrict__ x, const double* __restrict__ p,
                               const double* __restrict__ scal, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] += (scal[0] / scal[2]) * p[i];
}
__global__ void copy_scalar_kernel(double* __restrict__ dst, const double* __restrict__ src) { *dst = *src; }

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
void launch_extend_add(const EaItem* items, int count, const EaContrib* contrib, const int* rel, T* L, T* U,
                       hipStream_t s) {
    if (count > 0) extend_add_kernel<T><<<(count + 3) / 4, kThreads, 0, s>>>(items, count, contrib, rel, L, U);
}
template <typename T>
void launch_panel(const int* list, int count, const FrontDesc* fronts, T* L, int* info, int lds_bytes, int* fb_cnt,
                  int* fb_list, int* fallbacks, hipStream_t s) {
    if (count <= 0) return;
    // dynamic LDS beyond 64 KB must be allowed explicitly (static
    // initialisation: once, thread-safe)
    static const bool attr_set = [] {
        (void)hipFuncSetAttribute((const void*)panel_kernel<T>, hipFuncAttributeMaxDynamicSharedMemorySize, 65536);
        (void)hipFuncSetAttribute((const void*)panel_fb_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  2 * 65536);
        return true;
    }();
    (void)attr_set;
    panel_kernel<T><<<count, kThreads, lds_bytes, s>>>(list, fronts, L, info, fb_cnt, fb_list);
    if constexpr (std::is_same<T, float>::value)
        panel_fb_kernel<<<count < 64 ? count : 64, kThreads, 2 * lds_bytes, s>>>(fb_cnt, fb_list, fronts, L, info,
                                                                                 fallbacks);
}
template <typename T>
void launch_blk(const BlkJob* jobs, int njobs, int nblocks, T* L, int64_t nL, int* info, int* counters,
                int* fallbacks, __bf16* planes, int64_t pstride, int nplanes, hipStream_t s) {
    if (nblocks > 0)
        blk_kernel<T><<<nblocks, kThreads, 0, s>>>(jobs, njobs, L, nL, info, counters, fallbacks, planes, pstride,
                                                   nplanes);
}
template <typename T>
void launch_inv(const InvJob* jobs, int count, const FrontDesc* fronts, const T* L, T* inv, hipStream_t s) {
    if (count > 0) inv_kernel<T><<<count, kThreads, 0, s>>>(jobs, fronts, L, inv);
}
template <typename T>
void launch_gemm(const GemmJob* jobs, int njobs, int nblocks, int max_grid, int kernel, T* L, int64_t nL, T* U,
                 hipStream_t s, int lds_min) {
    if (nblocks <= 0) return;
    // a grid of at most max_grid workgroups (a multiple of the 8 XCDs)
    // loops over the nblocks tiles
    const int grid = max_grid > 0 && max_grid < nblocks ? max_grid : nblocks;
    const bool rows128 = gemm_tile_rows(kernel) == 128;
    // lds_min > 0: dynamic LDS that brings a workgroup's LDS to lds_min bytes
    // (fewer workgroups per CU)
    auto pad = [lds_min](const void* fn) -> size_t {
        if (lds_min <= 0) return 0;
        hipFuncAttributes a{};
        if (hipFuncGetAttributes(&a, fn) != hipSuccess) return 0;
        return (size_t)lds_min > a.sharedSizeBytes ? (size_t)lds_min - a.sharedSizeBytes : 0;
    };
    if constexpr (std::is_same<T, float>::value) {
        // two-level accumulation on 128-row tiles always runs the 32x32x2
        // kernel (same tiles; the 16x16x4 kernel's second accumulator set
        // spills)
        const bool split = gemm_split(kernel);
        const int base = gemm_kernel_base(kernel);
        auto mfma32 = [&](auto fn) {
            fn<<<grid, kThreads, pad((const void*)fn), s>>>(jobs, njobs, L, U, nL, nblocks);
        };
        if (base == kGemmMfma32 || (split && rows128)) {
            if (split) mfma32(gemm_mfma32_kernel<128, true>);
            else mfma32(gemm_mfma32_kernel<128, false>);
            return;
        }
        if (base == kGemmMfma32x64) {
            if (split) mfma32(gemm_mfma32_kernel<64, true>);
            else mfma32(gemm_mfma32_kernel<64, false>);
            return;
        }
    }
    auto mfma = [&](auto fn) { fn<<<grid, kThreads, pad((const void*)fn), s>>>(jobs, njobs, L, U, nblocks); };
    if constexpr (std::is_same<T, float>::value) {
        if (gemm_split(kernel)) {
            mfma(gemm_mfma_kernel<T, 64, true>);
            return;
        }
    }
    if (rows128) mfma(gemm_mfma_kernel<T, 128, false>);
    else mfma(gemm_mfma_kernel<T, 64, false>);
}
int blk_lds_bytes_f32() {
    hipFuncAttributes a{};
    if (hipFuncGetAttributes(&a, (const void*)blk_kernel<float>) != hipSuccess) return 0;
    return (int)a.sharedSizeBytes;
}
void launch_gemm_bf16(const GemmJob* jobs, int njobs, int nblocks, int max_grid, int nsplit, float* L, float* U,
                      hipStream_t s) {
    if (nblocks <= 0) return;
    const int grid = max_grid > 0 && max_grid < nblocks ? max_grid : nblocks;
    if (nsplit == 3) gemm_bf16_kernel<3><<<grid, kThreads, 0, s>>>(jobs, njobs, L, U, nblocks);
    else gemm_bf16_kernel<1><<<grid, kThreads, 0, s>>>(jobs, njobs, L, U, nblocks);
}

void launch_gemm_bf16_planes(const GemmJob* jobs, int njobs, int nblocks, int max_grid, int nsplit,
                             const __bf16* planes, int64_t pstride, float* U, float* L, hipStream_t s) {
    if (nblocks <= 0) return;
    const int grid = max_grid > 0 && max_grid < nblocks ? max_grid : nblocks;
    if (nsplit == 6) gemm_bf16p_kernel<6><<<grid, kThreads, 0, s>>>(jobs, njobs, planes, pstride, U, L, nblocks);
    else if (nsplit == 3) gemm_bf16p_kernel<3><<<grid, kThreads, 0, s>>>(jobs, njobs, planes, pstride, U, L, nblocks);
    else gemm_bf16p_kernel<1><<<grid, kThreads, 0, s>>>(jobs, njobs, planes, pstride, U, L, nblocks);
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
                int vec_cap, hipStream_t s) {
    if (count > 0)
        fwd_kernel<T><<<count, kThreads, 0, s>>>(list, fronts, kids, rel, q, L, inv, xp, Y, su, suofs,
                                                 min(vec_cap, kSmallVec));
}
template <typename T>
void launch_bwd(const int* list, int count, const FrontDesc* fronts, const int* q, const T* L, const T* inv,
                double* xp, const double* Y, hipStream_t s) {
    if (count > 0) bwd_kernel<T><<<count, kSolveThreads, 0, s>>>(list, fronts, q, L, inv, xp, Y);
}

void launch_fwd_gather(const int* list, int count, const FrontDesc* fronts, const int* kids, const int* rel,
                       double* xp, double* su, const int64_t* suofs, int vec_cap, hipStream_t s) {
    if (count > 0)
        fwd_gather_kernel<<<count, kGatherThreads, 0, s>>>(list, fronts, kids, rel, xp, su, suofs,
                                                           min(vec_cap, kGatherVec));
}
template <typename T>
void launch_fwd_chain(const SolveItem* items, int count, int* ctr, int* status, int spin, const T* L, const T* X256,
                      const double* xp, double* Y, double* su, hipStream_t s) {
    if (count > 0) fwd_chain_kernel<T><<<count, kSolveThreads, 0, s>>>(items, ctr, status, spin, L, X256, xp, Y, su);
}
template <typename T>
void launch_bwd_upd(const SolveItem* items, int count, const FrontDesc* fronts, const int* q, const T* L,
                    const double* xp, double* partial, hipStream_t s) {
    if (count > 0) bwd_upd_kernel<T><<<count, kSolveThreads, 0, s>>>(items, q, L, xp, partial);
}
void launch_bwd_upd_sum(const SolveItem* items, int count, const FrontDesc* fronts, const double* partial,
                        double* Y, hipStream_t s) {
    if (count > 0) bwd_upd_sum_kernel<<<count, kSolveCols, 0, s>>>(items, partial, Y);
}
template <typename T>
void launch_bwd_chain(const SolveItem* items, int count, int* ctr, int* status, int spin, const T* L, const T* X256,
                      double* xp, const double* Y, hipStream_t s) {
    if (count > 0) bwd_chain_kernel<T><<<count, kSolveThreads, 0, s>>>(items, ctr, status, spin, L, X256, xp, Y);
}
template <typename T>
void launch_inv256(const InvJob* jobs, int count, const FrontDesc* fronts, const T* L, const T* inv, T* X256,
                   hipStream_t s) {
    if (count > 0) inv256_kernel<T><<<count, kThreads, 0, s>>>(jobs, fronts, L, inv, X256);
}

void launch_gather(const double* src_vals, const int* src, double* dst, int nnz, hipStream_t s) {
    if (nnz > 0) gather_kernel<<<blocks_for(nnz), kThreads, 0, s>>>(src_vals, src, dst, nnz);
}
void launch_residual(const int* rp, const int* ci, const double* av, const double* x, const double* b, double* r,
                     int n, hipStream_t s) {
    residual_kernel<<<blocks_for((int64_t)n * kRowLanes), kThreads, 0, s>>>(rp, ci, av, x, b, r, n);
}
void launch_rowabs(const int* rp, const double* av, double* out, int n, hipStream_t s) {
    rowabs_kernel<<<blocks_for((int64_t)n * kRowLanes), kThreads, 0, s>>>(rp, av, out, n);
}
void launch_maxabs(const double* v, int n, unsigned long long* out, hipStream_t s) {
    int blocks = blocks_for(n);
    if (blocks > 1024) blocks = 1024;
    if (blocks < 1) blocks = 1;
    maxabs_kernel<<<blocks, kThreads, 0, s>>>(v, n, out);
}
void launch_maxabs_seg(const double* v, const int* seg, int n, unsigned long long* out, hipStream_t s) {
    int blocks = blocks_for(n);
    if (blocks > 1024) blocks = 1024;
    if (blocks < 1) blocks = 1;
    maxabs_seg_kernel<<<blocks, kThreads, 0, s>>>(v, seg, n, out);
}
void launch_axpy(double* x, const double* d, int n, hipStream_t s) {
    axpy_kernel<<<blocks_for(n), kThreads, 0, s>>>(x, d, n);
}
void launch_spmv(const int* rp, const int* ci, const double* av, const double* x, double* y, int n, hipStream_t s) {
    spmv_kernel<<<blocks_for((int64_t)n * kRowLanes), kThreads, 0, s>>>(rp, ci, av, x, y, n);
}
void launch_dot(const double* a, const double* c, int n, double* partial, double* out, hipStream_t s) {
    dot_partial_kernel<<<kDotParts, kThreads, 0, s>>>(a, c, n, partial);
    dot_final_kernel<<<1, kDotParts, 0, s>>>(partial, out);
}
void launch_cg_dir(double* p, const double* z, const double* scal, int n, bool first, hipStream_t s) {
    cg_dir_kernel<<<blocks_for(n), kThreads, 0, s>>>(p, z, scal, n, first ? 1 : 0);
}
void launch_cg_step(double* x, const double* p, const double* scal, int n, hipStream_t s) {
    cg_step_kernel<<<blocks_for(n), kThreads, 0, s>>>(x, p, scal, n);
}
void launch_copy_scalar(double* dst, const double* src, hipStream_t s) { copy_scalar_kernel<<<1, 1, 0, s>>>(dst, src); }
template <typename T>
void launch_logdiag(const int64_t* diag_off, const T* L, int n, double* partial, int nparts, hipStream_t s) {
    logdiag_kernel<T><<<nparts, kThreads, 0, s>>>(diag_off, L, n, partial);
}

#define MDGPU_INSTANTIATE(T)                                                                                   \
    template void launch_fill<T>(T*, int64_t, T, hipStream_t);                                                 \
    template void launch_scatter_a<T>(T*, const double*, const int64_t*, int, hipStream_t);                    \
    template void launch_extend_add<T>(const EaItem*, int, const EaContrib*, const int*, T*, T*, hipStream_t); \
    template void launch_panel<T>(const int*, int, const FrontDesc*, T*, int*, int, int*, int*, int*, hipStream_t); \
    template void launch_blk<T>(const BlkJob*, int, int, T*, int64_t, int*, int*, int*, __bf16*, int64_t, int,       \
                                hipStream_t);                                                                      \
    template void launch_inv<T>(const InvJob*, int, const FrontDesc*, const T*, T*, hipStream_t);             \
    template void launch_inv256<T>(const InvJob*, int, const FrontDesc*, const T*, const T*, T*, hipStream_t); \
    template void launch_gemm<T>(const GemmJob*, int, int, int, int, T*, int64_t, T*, hipStream_t, int);                               \
    template void launch_fwd<T>(const int*, int, const FrontDesc*, const int*, const int*, const int*,         \
                                const T*, const T*, double*, double*, double*, const int64_t*, int,            \
                                hipStream_t);                                                                  \
    template void launch_bwd<T>(const int*, int, const FrontDesc*, const int*, const T*, const T*, double*,    \
                                const double*, hipStream_t);                                                   \
    template void launch_logdiag<T>(const int64_t*, const T*, int, double*, int, hipStream_t);               \
    template void launch_fwd_chain<T>(const SolveItem*, int, int*, int*, int, const T*, const T*, const double*, \
                                      double*, double*, hipStream_t);                                        \
    template void launch_bwd_upd<T>(const SolveItem*, int, const FrontDesc*, const int*, const T*,           \
                                    const double*, double*, hipStream_t);                                    \
    template void launch_bwd_chain<T>(const SolveItem*, int, int*, int*, int, const T*, const T*, double*,   \
                                      const double*, hipStream_t);

MDGPU_INSTANTIATE(double)
MDGPU_INSTANTIATE(float)

}  // namespace mdgpu

namespace mdgpu {
template void launch_fill<int>(int*, int64_t, int, hipStream_t);
template void launch_fill<unsigned long long>(unsigned long long*, int64_t, unsigned long long, hipStream_t);
}  // namespace mdgpu

// C ABI: the backward error of a solution of a symmetric system on the GPU,
// eta = ||b - A x||_inf / (||A||_1 ||x||_inf + ||b||_inf) -- the check the
// reference driver computes after its solve (cs_gaxpy, cs_norm and norm,
// mdgpu.cu:1962-1972, cs.cpp:236-248, 1069-1083) -- for a full symmetric
// matrix in CSR (= CSC) form on the device.  For a symmetric matrix the
// 1-norm (largest column sum of |a_ij|) is the largest row sum.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <string>

#include "common.h"

using mdgpu::fail;

namespace {

constexpr int kBeThreads = 256;
constexpr int kBeLanes = 16;           // lanes per row
constexpr int kBeRowsPerBlock = 1024;  // rows of one workgroup

// Workgroup w: rows [w kBeRowsPerBlock, (w + 1) kBeRowsPerBlock), 16 lanes
// per row.  part[4 w + {0, 1, 2, 3}] = the workgroup's max |b_i - (A x)_i|,
// max sum_j |a_ij|, max |x_i|, max |b_i| (NaN counts as +inf).
__global__ __launch_bounds__(kBeThreads) void be_rows_kernel(int n, const int* __restrict__ rp,
                                                             const int* __restrict__ ci,
                                                             const double* __restrict__ av,
                                                             const double* __restrict__ x,
                                                             const double* __restrict__ b,
                                                             double* __restrict__ part) {
    __shared__ double red[4][kBeThreads / 64];
    const int l = threadIdx.x % kBeLanes;
    const int row_end = min(n, (blockIdx.x + 1) * kBeRowsPerBlock);
    double m[4] = {0.0, 0.0, 0.0, 0.0};
    for (int i = blockIdx.x * kBeRowsPerBlock + threadIdx.x / kBeLanes; i < row_end; i += kBeThreads / kBeLanes) {
        double s = 0.0, a = 0.0;
        for (int p = rp[i] + l; p < rp[i + 1]; p += kBeLanes) {
            const double v = av[p];
            s += v * x[ci[p]];
            a += fabs(v);
        }
#pragma unroll
        for (int o = kBeLanes / 2; o >= 1; o >>= 1) {
            s += __shfl_xor(s, o, kBeLanes);
            a += __shfl_xor(a, o, kBeLanes);
        }
        if (l == 0) {
            const double v[4] = {fabs(b[i] - s), a, fabs(x[i]), fabs(b[i])};
#pragma unroll
            for (int k = 0; k < 4; ++k) m[k] = (v[k] != v[k]) ? __builtin_huge_val() : fmax(m[k], v[k]);
        }
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        double t = m[k];
        for (int off = 32; off > 0; off >>= 1) t = fmax(t, __shfl_down(t, off, 64));
        if ((threadIdx.x & 63) == 0) red[k][threadIdx.x >> 6] = t;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        double t = red[threadIdx.x][0];
        for (int w = 1; w < kBeThreads / 64; ++w) t = fmax(t, red[threadIdx.x][w]);
        part[4 * blockIdx.x + threadIdx.x] = t;
    }
}

// out[k] = max over the nblocks workgroups of part[4 w + k], one workgroup
// (each thread a fixed set of workgroups, then a fixed reduction order).
__global__ __launch_bounds__(kBeThreads) void be_final_kernel(const double* __restrict__ part, int nblocks,
                                                              double* __restrict__ out) {
    __shared__ double red[4][kBeThreads / 64];
    double m[4] = {0.0, 0.0, 0.0, 0.0};
    for (int w = threadIdx.x; w < nblocks; w += kBeThreads)
#pragma unroll
        for (int k = 0; k < 4; ++k) m[k] = fmax(m[k], part[4 * w + k]);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        double t = m[k];
        for (int off = 32; off > 0; off >>= 1) t = fmax(t, __shfl_down(t, off, 64));
        if ((threadIdx.x & 63) == 0) red[k][threadIdx.x >> 6] = t;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        double t = red[threadIdx.x][0];
        for (int w = 1; w < kBeThreads / 64; ++w) t = fmax(t, red[threadIdx.x][w]);
        out[threadIdx.x] = t;
    }
}

int hip_fail(hipError_t e, const char* what) {
    return fail(MDGPU_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

// Launches into the workspace ws (4 nblocks + 4 doubles of device memory),
// copies the four maxima to the host and forms eta.  Synchronizes the
// stream.
int backward_error_into(int n, const int32_t* rp, const int32_t* ci, const double* av, const double* x,
                        const double* b, double* ws, double* eta, hipStream_t s, const char* who) {
    const int nblocks = (n + kBeRowsPerBlock - 1) / kBeRowsPerBlock;
    double* out = ws + 4 * (int64_t)nblocks;
    be_rows_kernel<<<nblocks, kBeThreads, 0, s>>>(n, rp, ci, av, x, b, ws);
    be_final_kernel<<<1, kBeThreads, 0, s>>>(ws, nblocks, out);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, who);
    double h[4];
    e = hipMemcpyAsync(h, out, sizeof h, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_fail(e, who);
    const double den = h[1] * h[2] + h[3];
    *eta = den > 0 ? h[0] / den : h[0];
    return MDGPU_OK;
}

int check_args(int n, const int32_t* rp, const int32_t* ci, const double* av, const double* x, const double* b,
               const double* eta, const char* who) {
    if (!rp || !ci || !av || !x || !b || !eta) return fail(MDGPU_ERR_ARG, std::string(who) + ": null pointer");
    if (n <= 0) return fail(MDGPU_ERR_ARG, std::string(who) + ": n must be positive");
    return MDGPU_OK;
}

}  // namespace

extern "C" int mdgpu_backward_error(int n, const int32_t* d_rowptr, const int32_t* d_colind, const double* d_val,
                                    const double* d_x, const double* d_b, double* eta, void* stream) {
    mdgpu::clear_error();
    const char* who = "mdgpu_backward_error";
    int st = check_args(n, d_rowptr, d_colind, d_val, d_x, d_b, eta, who);
    if (st != MDGPU_OK) return st;
    const int nblocks = (n + kBeRowsPerBlock - 1) / kBeRowsPerBlock;
    double* ws = nullptr;
    hipError_t e = hipMalloc((void**)&ws, sizeof(double) * (4 * (size_t)nblocks + 4));
    if (e != hipSuccess) return fail(MDGPU_ERR_NOMEM, std::string(who) + ": hipMalloc failed");
    st = backward_error_into(n, d_rowptr, d_colind, d_val, d_x, d_b, ws, eta, (hipStream_t)stream, who);
    (void)hipFree(ws);
    return st;
}

// ------------------------------------------------ no-allocation variant
// mdgpu_backward_error_noalloc: the same kernels with their workspace in a
// file-scope device array, for callers that cannot allocate (code under
// graph capture, serving loops).  Every device the library runs on holds
// one copy until the process ends; all calls on a device share it, so
// calls that overlap on one device race.  Host code uses the array only
// through the device address hipGetSymbolAddress gives (its name is a
// host-side stand-in), looked up once per device.
constexpr int kNoAllocMaxBlocks = 4096;                                      // n <= 4096 * 1024 rows
static __device__ __align__(256) double scratch[4 * kNoAllocMaxBlocks + 4];  // 131 KB

namespace {

constexpr int kNoAllocMaxRows = kNoAllocMaxBlocks * kBeRowsPerBlock;
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

// the device address of `scratch` on the current device (looked up on the
// first call there)
int scratch_ptr(double** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_fail(e, "mdgpu_backward_error_noalloc: hipGetDevice");
    if (dev < 0 || dev >= kMaxDevices) return fail(MDGPU_ERR_DEVICE, "mdgpu_backward_error_noalloc: device index");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hip_fail(e, "mdgpu_backward_error_noalloc: hipGetSymbolAddress");
        g_scratch_ptr[dev] = (double*)p;
    }
    *out = g_scratch_ptr[dev];
    return MDGPU_OK;
}

}  // namespace

extern "C" int mdgpu_backward_error_noalloc(int n, const int32_t* d_rowptr, const int32_t* d_colind,
                                            const double* d_val, const double* d_x, const double* d_b, double* eta,
                                            void* stream) {
    mdgpu::clear_error();
    const char* who = "mdgpu_backward_error_noalloc";
    int st = check_args(n, d_rowptr, d_colind, d_val, d_x, d_b, eta, who);
    if (st != MDGPU_OK) return st;
    if (n > kNoAllocMaxRows)
        return fail(MDGPU_ERR_ARG, std::string(who) + ": n = " + std::to_string(n) + " exceeds the limit of " +
                                       std::to_string(kNoAllocMaxRows) + " rows of the static workspace");
    double* ws = nullptr;
    st = scratch_ptr(&ws);
    if (st != MDGPU_OK) return st;
    // every slot the kernels read is written by them first: nothing to clear
    return backward_error_into(n, d_rowptr, d_colind, d_val, d_x, d_b, ws, eta, (hipStream_t)stream, who);
}

// Launch wrappers of the reverse-mode kernels (selinv.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "selinv.h"

namespace mdgpu {

// device base pointers of the buffers named by SelBuf
struct SelBuffers {
    double* p[SB_N];
};

void launch_sel_gemm(const SelGemm* jobs, int njobs, int nblocks, SelBuffers bufs, hipStream_t s);
void launch_sel_copy(int kind, const SelCopy* items, int count, const FrontDesc* fronts, const int* rel,
                     const int64_t* zf_ofs, SelBuffers bufs, hipStream_t s);
void launch_logdet_grad(const double* Z, const int64_t* dst, const int* arow, const int* acol, double* g, int nnz,
                        hipStream_t s);
void launch_solve_adjoint(const double* x, const double* bb, const int* arow, const int* acol, double* g, int nnz,
                          hipStream_t s);

}  // namespace mdgpu

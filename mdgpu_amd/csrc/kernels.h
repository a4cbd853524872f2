// Launch wrappers of the device kernels (kernels.hip).  All take the stream
// they run on; none synchronizes.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "plan.h"

namespace mdgpu {

template <typename T> void launch_fill(T* p, int64_t n, T v, hipStream_t s);
template <typename T> void launch_scatter_a(T* L, const double* ax, const int64_t* dst, int nnz, hipStream_t s);
// one wavefront per parent column item (Plan::ea), contributors Plan::ea_contrib
template <typename T>
void launch_extend_add(const EaItem* items, int count, const EaContrib* contrib, const int* rel, T* L, T* U,
                       hipStream_t s);
// one workgroup per front of the list; float: fronts that fail the
// pivot-growth check go on fb_list (counter fb_cnt, zero before the launch)
// and are refactored in FP64 by a second launch (fallbacks counts them)
template <typename T>
void launch_panel(const int* list, int count, const FrontDesc* fronts, T* L, int* info, int lds_bytes, int* fb_cnt,
                  int* fb_list, int* fallbacks, hipStream_t s);
// fused diagonal-block Cholesky + TRSM of the rows below; counters: one
// zero-initialised int per job (returned to zero by the kernel); fallbacks:
// incremented once per diagonal block refactored in FP64 (pivot growth);
// nplanes > 0 (FP32 store): the steps also write that many bf16 split
// planes (BlkJob::pl), plane p at planes + p * pstride
template <typename T>
void launch_blk(const BlkJob* jobs, int njobs, int nblocks, T* L, int64_t nL, int* info, int* counters,
                int* fallbacks, __bf16* planes, int64_t pstride, int nplanes, hipStream_t s);
template <typename T>
void launch_inv(const InvJob* jobs, int count, const FrontDesc* fronts, const T* L, T* inv, hipStream_t s);
// MFMA kernel of the step (64, 128 or kGemmMfma32, the step's GEMM kernel
// code; kGemmMfma32 is FP32 only) over its nblocks tiles; max_grid > 0 (a
// multiple of 8): at most that many workgroups, each looping over tiles;
// nL: elements of the L store (the FP32 32x32 kernel's loads stay inside
// [L, L + nL + 64))
template <typename T>
void launch_gemm(const GemmJob* jobs, int njobs, int nblocks, int max_grid, int kernel, T* L, int64_t nL, T* U,
                 hipStream_t s, int lds_min = 0);
// LDS bytes of one FP32 block-step workgroup (blk_kernel<float>), 0 if the
// query fails
int blk_lds_bytes_f32();
// same jobs (128-row tiles), FP32 data, bf16 MFMA with FP32 accumulation;
// nsplit 1: one bf16 product, 3: split hi/lo products (hi*hi + hi*lo + lo*hi)
void launch_gemm_bf16(const GemmJob* jobs, int njobs, int nblocks, int max_grid, int nsplit, float* L, float* U,
                      hipStream_t s);
// kGemmPlanes launches: the same products with A and B from the bf16 split
// planes (GemmJob::a_pl, b_pl; plane p at planes + p * pstride); nsplit 6:
// hi, mid and lo planes, products hi*lo + lo*hi + mid*mid + hi*mid + mid*hi
// + hi*hi; C in U or L as for the others
void launch_gemm_bf16_planes(const GemmJob* jobs, int njobs, int nblocks, int max_grid, int nsplit,
                             const __bf16* planes, int64_t pstride, float* U, float* L, hipStream_t s);

// solves (compute in FP64, factor stored as T); the forward result goes to
// Y (which the backward solve of large fronts then updates in place), the
// backward result to xp
void launch_permute_in(const double* b, const int* perm, double* xp, int n, hipStream_t s);
void launch_permute_out(const double* xp, const int* perm, double* x, int n, hipStream_t s);
// Capacity (entries) of the LDS copies of a front vector in the forward
// solve: small fronts (launch_fwd) and the gather of large fronts
// (launch_fwd_gather).  vec_cap (at most the capacity) is the longest
// front vector kept in LDS; longer ones are accumulated in place in global
// memory.
constexpr int kSmallVec = 2048, kGatherVec = 16384;
template <typename T>
void launch_fwd(const int* list, int count, const FrontDesc* fronts, const int* kids, const int* rel,
                const int* q, const T* L, const T* inv, double* xp, double* Y, double* su, const int64_t* suofs,
                int vec_cap, hipStream_t s);
template <typename T>
void launch_bwd(const int* list, int count, const FrontDesc* fronts, const int* q, const T* L, const T* inv,
                double* xp, const double* Y, hipStream_t s);
void launch_fwd_gather(const int* list, int count, const FrontDesc* fronts, const int* kids, const int* rel,
                       double* xp, double* su, const int64_t* suofs, int vec_cap, hipStream_t s);
// forward chain over a level's large fronts (ctr: the launch's ticket
// counter, zero; Y holds kChainEmpty where the chain stores y); a wait that
// gives up after `spin` polls (kChainSpin: about a second) sets
// kSolveTimedOut in *status
constexpr int kSolveTimedOut = 1;
constexpr int kChainSpin = 1 << 20;
template <typename T>
void launch_fwd_chain(const SolveItem* items, int count, int* ctr, int* status, int spin, const T* L, const T* X256,
                      const double* xp, double* Y, double* su, hipStream_t s);
// backward: BWD_UPD writes the update rows' dot products to partial-sum
// slots, BWD_UPD_SUM subtracts them from Y, then the chain solves the blocks
// into xp (which holds kChainEmpty there)
template <typename T>
void launch_bwd_upd(const SolveItem* items, int count, const FrontDesc* fronts, const int* q, const T* L,
                    const double* xp, double* partial, hipStream_t s);
void launch_bwd_upd_sum(const SolveItem* items, int count, const FrontDesc* fronts, const double* partial,
                        double* Y, hipStream_t s);
template <typename T>
void launch_bwd_chain(const SolveItem* items, int count, int* ctr, int* status, int spin, const T* L, const T* X256,
                      double* xp, const double* Y, hipStream_t s);
// inverses of the kSolveBW-column diagonal blocks (Plan::inv256), after
// launch_inv
template <typename T>
void launch_inv256(const InvJob* jobs, int count, const FrontDesc* fronts, const T* L, const T* inv, T* X256,
                   hipStream_t s);
// bit pattern of a chained-solve value not yet stored
constexpr unsigned long long kChainEmptyBits = 0x7ff4dead0000beefULL;

// refinement helpers (FP64)
void launch_gather(const double* src_vals, const int* src, double* dst, int nnz, hipStream_t s);
void launch_residual(const int* rp, const int* ci, const double* av, const double* x, const double* b, double* r,
                     int n, hipStream_t s);
void launch_rowabs(const int* rp, const double* av, double* out, int n, hipStream_t s);
void launch_maxabs(const double* v, int n, unsigned long long* out, hipStream_t s);
// out[seg[i]] = max |v[i]| per contiguous segment (out cleared by the caller)
void launch_maxabs_seg(const double* v, const int* seg, int n, unsigned long long* out, hipStream_t s);
void launch_axpy(double* x, const double* d, int n, hipStream_t s);
// conjugate-gradient refinement: y = A x; *out = a . c (partial: kDotParts
// doubles; fixed summation order); p = z (+ (rz / rz_old) p); x += (rz /
// pq) p, with scal = {rz, rz_old, pq} on the device
constexpr int kDotParts = 256;
void launch_spmv(const int* rp, const int* ci, const double* av, const double* x, double* y, int n, hipStream_t s);
void launch_dot(const double* a, const double* c, int n, double* partial, double* out, hipStream_t s);
void launch_cg_dir(double* p, const double* z, const double* scal, int n, bool first, hipStream_t s);
void launch_cg_step(double* x, const double* p, const double* scal, int n, hipStream_t s);
void launch_copy_scalar(double* dst, const double* src, hipStream_t s);
template <typename T>
void launch_logdiag(const int64_t* diag_off, const T* L, int n, double* partial, int nparts, hipStream_t s);

}  // namespace mdgpu

// C ABI: host-only planning entry point (no device calls), used to size a
// factorization and to check its schedule before any kernel runs.
#include <cstdlib>

#include "common.h"
#include "plan.h"

using mdgpu::fail;

namespace mdgpu {
PlanOptions plan_options_for(int precision) {
    PlanOptions po;
    po.elem_bytes = precision == MDGPU_FP64 ? 8 : 4;
    po.panel_lds_bytes = 61440;
    // largest MFMA GEMM tile (rows = columns); the autotuner may pick 64
    // per launch (not for the bf16 kernels, which have 128-row tiles only)
    po.tile = 128;
    // FP32: the 32x32x2 MFMA kernel by default (the autotuner may pick the
    // 16x16x4 kernel's 64- or 128-row tiles per launch)
    po.gemm_default = precision == MDGPU_FP32 ? kGemmMfma32 : 128;
    // FP32: Schur complements over more than 256 pivots sum their products
    // in two levels (runs of 64 k), which cuts the factor's backward error
    // after one refinement step by about 10x on the 64^3 27-point stencil
    if (precision == MDGPU_FP32) po.split_k_min = 257;
    // MDGPU_GEMM_KERNEL: one GEMM kernel code for every launch (a kernel code
    // of plan.h valid for the precision; tests pin each kernel with it)
    if (const char* e = std::getenv("MDGPU_GEMM_KERNEL")) {
        const int v = std::atoi(e);
        if (precision != MDGPU_BF16 && precision != MDGPU_BF16X3 && gemm_kernel_valid(v, po.elem_bytes)) {
            po.gemm_default = v;
            po.split_k_min = 0;
        }
    }
    // bf16 variants: GEMMs of blocked fronts from the split planes the block
    // steps write (MDGPU_PRESPLIT=0: every bf16 GEMM converts its FP32 tiles
    // as it stages them)
    po.presplit = precision == MDGPU_BF16 || precision == MDGPU_BF16X3;
    if (const char* e = std::getenv("MDGPU_PRESPLIT")) po.presplit = po.presplit && std::atoi(e) != 0;
    if (const char* e = std::getenv("MDGPU_PLANES")) po.nplanes = std::atoi(e) == 3 ? 3 : 2;
    // MDGPU_U_LOOKAHEAD=0: update matrices in one Schur launch per level
    // (no second side stream)
    if (const char* e = std::getenv("MDGPU_U_LOOKAHEAD")) po.u_lookahead = std::atoi(e) != 0;
    if (const char* e = std::getenv("MDGPU_U_STRIDE")) po.u_stride = std::max(1, std::atoi(e));
    if (const char* e = std::getenv("MDGPU_U_FLUSH")) po.u_flush = std::max(0, std::atoi(e));
    if (const char* e = std::getenv("MDGPU_SCHUR_KERNEL")) po.schur_kernel = std::atoi(e);
    // most pivots of a front factored in LDS by one workgroup (at most 128)
    if (const char* e = std::getenv("MDGPU_PANEL_MAXC")) po.panel_max_pivots = std::max(0, std::atoi(e));
    // outer block width of the blocked panel path (multiple of 64)
    if (const char* e = std::getenv("MDGPU_NB_OUTER")) {
        const int v = std::atoi(e);
        if (v >= 64 && v % 64 == 0) po.nb_outer = v;
    }
    return po;
}
}  // namespace mdgpu

extern "C" int mdgpu_plan_check(const mdgpu_symbolic* sym, int precision, int64_t* out) {
    mdgpu::clear_error();
    if (!sym || !out) return fail(MDGPU_ERR_ARG, "mdgpu_plan_check: null pointer");
    if (precision < MDGPU_FP64 || precision > MDGPU_BF16X3)
        return fail(MDGPU_ERR_ARG, "mdgpu_plan_check: precision must be MDGPU_FP64, MDGPU_FP32, MDGPU_BF16 or MDGPU_BF16X3");
    mdgpu::PlanOptions po = mdgpu::plan_options_for(precision);
    mdgpu::Plan P;
    std::string err = mdgpu::build_plan(sym->S, po, P);
    if (err.empty()) err = mdgpu::validate_plan(P, po);
    if (!err.empty()) return fail(MDGPU_ERR_ARG, "mdgpu_plan_check: " + err);
    const int64_t es = po.elem_bytes;
    int64_t bytes = (P.lstore + P.arena + P.dinv) * es + P.solve_u * 8 +
                    (int64_t)P.a_dst.size() * 8 + (int64_t)(P.rel.size() + P.q.size()) * 4 +
                    (int64_t)P.afull_i.size() * 16 + 2 * po.nplanes * P.planes;
    out[MDGPU_NUM_DEVICE_BYTES] = bytes;
    out[MDGPU_NUM_LSTORE] = P.lstore;
    out[MDGPU_NUM_ARENA] = P.arena;
    out[MDGPU_NUM_STEPS] = (int64_t)P.steps.size();
    out[MDGPU_NUM_LEVELS] = P.nlevels;
    out[MDGPU_NUM_PANEL_FRONTS] = (int64_t)P.panel.size();
    out[MDGPU_NUM_DIAG_JOBS] = (int64_t)P.blk.size();
    out[MDGPU_NUM_PRECISION] = precision;
    out[MDGPU_NUM_FALLBACKS] = 0;
    return MDGPU_OK;
}

// C ABI: numeric factorization, solves, refinement, log-determinant and
// L export on one GPU (see include/mdgpu_c.h).  Host code only orchestrates:
// every arithmetic operation on the factor runs in kernels.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <type_traits>
#include <vector>

#include "common.h"
#include "kernels.h"
#include "plan.h"
#include "selinv.h"
#include "selinv_kernels.h"

using mdgpu::fail;

namespace {

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
};

}  // namespace

struct mdgpu_numeric {
    const mdgpu_symbolic* sym = nullptr;
    int precision = MDGPU_FP64;
    int device = 0;
    mdgpu::Plan plan;
    std::vector<DevBuf> bufs;
    // device views
    void* dL = nullptr;
    void* dU = nullptr;
    void* dInv = nullptr;
    mdgpu::FrontDesc* dFronts = nullptr;
    int* dKids = nullptr;
    int* dRel = nullptr;
    int* dQ = nullptr;
    int64_t* dAdst = nullptr;
    mdgpu::EaItem* dEa = nullptr;
    mdgpu::EaContrib* dEaContrib = nullptr;
    mdgpu::GemmJob* dGemm = nullptr;
    // bf16 split planes (PlanOptions::nplanes of Plan::planes entries each,
    // plane p at + p * plane_stride; PlanOptions::presplit, else unallocated)
    __bf16* dPlanes = nullptr;
    int64_t plane_stride = 0;
    mdgpu::BlkJob* dBlk = nullptr;
    int* dBlkCnt = nullptr;          // arrival counters of the block-column jobs
    mdgpu::InvJob* dInvJobs = nullptr;
    void* dInv256 = nullptr;              // kSolveBW-column block inverses (solves)
    int* dChainCtr = nullptr;             // ticket counters of the chained solves
    double* dUpdPartial = nullptr;        // BWD_UPD partial sums
    mdgpu::InvJob* dInv256Jobs = nullptr;
    int* dPanel = nullptr;
    int* dPanelFbList = nullptr;          // FP64 fallback lists of the panel steps (one slot per panel front)
    int* dPanelFbCnt = nullptr;           // their counters, one per panel step
    int npanel_steps = 0;
    int* dLevelList = nullptr;
    std::vector<int> level_off;
    double* dSu = nullptr;
    int64_t* dSuofs = nullptr;
    int* dPerm = nullptr;
    int* dAfp = nullptr;
    int* dAfi = nullptr;
    int* dAfsrc = nullptr;
    double* dAfx = nullptr;
    double* dXp = nullptr;
    double* dR = nullptr;
    double* dD = nullptr;
    double* dP = nullptr;                 // CG refinement: search direction, A p
    double* dQ2 = nullptr;
    double* dScal = nullptr;              // CG scalars {rz, rz_old, pq} + dot partials
    bool refine_cg = false;               // CG refinement instead of classical refinement
    double* dRowabs = nullptr;
    int64_t* dDiagOff = nullptr;
    int* dInfo = nullptr;
    unsigned long long* dMax = nullptr;   // three slots per segment (max_abs_n), then the solve status
    int* dStatus = nullptr;               // solve status bits (kSolveTimedOut), sticky until read
    int vec_small = mdgpu::kSmallVec;     // longest front vector kept in LDS by the forward solve
    int vec_gather = mdgpu::kGatherVec;
    int chain_spin = mdgpu::kChainSpin;   // polls of a chained-solve wait before it gives up
    // GPU-side order of the calls on this handle (CallOrder)
    hipEvent_t done = nullptr;
    bool done_recorded = false;
    int nseg = 1;
    int* dSeg = nullptr;                  // row -> segment (batches only)
    double* dPartial = nullptr;
    int nparts = 0;
    // hipGraph replay of the fixed launch sequences (refactorization and
    // solves on one pattern): captured on cap_stream against the internal
    // buffers dAx / dB / dX, launched on the caller's stream
    bool use_graph = true;
    hipStream_t cap_stream = nullptr;
    hipStream_t side = nullptr;           // look-ahead streams (default priority): trailing
    hipStream_t side2 = nullptr;          // updates of pivot columns / of update matrices
    // workgroups of a side-stream GEMM launch (0: one per tile): a bounded
    // grid loops over its tiles and leaves CUs to the main stream's chain
    int side_grid = 0;
    int side_lds = 0;                     // LDS floor of side-stream GEMM workgroups (0: none)
    std::vector<hipEvent_t> events;       // look-ahead events of the plan
    hipGraphExec_t g_factor = nullptr;
    mdgpu::PlanOptions po;                // options the plan was built with
    hipGraphExec_t g_solve = nullptr;
    double* dAx = nullptr;
    double* dB = nullptr;
    double* dX = nullptr;
    int* dSvFronts = nullptr;
    mdgpu::SolveItem* dSvItems = nullptr;
    double* dY = nullptr;
    double anorm = -1;
    std::vector<double> anorm_seg;        // ||A_k||_1 per segment
    std::vector<double> eta_seg;          // backward error per segment, last refine
    bool factored = false;
    int fallbacks = 0;                    // diagonal blocks refactored in FP64, last factorization
    int64_t device_bytes = 0;
    // reverse mode (built on first use, FP64 only)
    bool sel_ready = false, z_valid = false;
    mdgpu::SelPlan sel;
    double* dZ = nullptr;                 // selected inverse, L-store layout
    double* dZF = nullptr;                // dense per-level work matrices
    double* dLH = nullptr;                // L(R, J) inv(L(J, J)) per front
    mdgpu::SelGemm* dSelGemm = nullptr;
    mdgpu::SelCopy* dSelCopy = nullptr;
    int64_t* dZfOfs = nullptr;
    int* dArow = nullptr;                 // caller entry -> row, column (caller numbering)
    int* dAcol = nullptr;
    double* dBbar = nullptr;
};

namespace {

const int kLogParts = 256;

int hip_fail(hipError_t e, const char* what) {
    return fail(MDGPU_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

template <typename V>
int upload(mdgpu_numeric* N, const V& v, void** out) {
    using E = typename V::value_type;
    size_t bytes = std::max<size_t>(v.size() * sizeof(E), 16);
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) return fail(MDGPU_ERR_NOMEM, std::string("hipMalloc failed: ") + hipGetErrorString(e));
    N->bufs.push_back({p, bytes});
    N->device_bytes += (int64_t)bytes;
    if (!v.empty()) {
        e = hipMemcpy(p, v.data(), v.size() * sizeof(E), hipMemcpyHostToDevice);
        if (e != hipSuccess) return hip_fail(e, "hipMemcpy");
    }
    *out = p;
    return MDGPU_OK;
}

int alloc(mdgpu_numeric* N, size_t bytes, void** out) {
    bytes = std::max<size_t>(bytes, 16);
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) return fail(MDGPU_ERR_NOMEM, std::string("hipMalloc failed: ") + hipGetErrorString(e));
    N->bufs.push_back({p, bytes});
    N->device_bytes += (int64_t)bytes;
    *out = p;
    return MDGPU_OK;
}

void release(mdgpu_numeric* N) {
    if (N->g_factor) (void)hipGraphExecDestroy(N->g_factor);
    if (N->g_solve) (void)hipGraphExecDestroy(N->g_solve);
    N->g_factor = N->g_solve = nullptr;
    if (N->cap_stream) (void)hipStreamDestroy(N->cap_stream);
    N->cap_stream = nullptr;
    if (N->side) (void)hipStreamDestroy(N->side);
    if (N->side2) (void)hipStreamDestroy(N->side2);
    N->side = N->side2 = nullptr;
    for (hipEvent_t e : N->events) (void)hipEventDestroy(e);
    N->events.clear();
    if (N->done) (void)hipEventDestroy(N->done);
    N->done = nullptr;
    for (auto& b : N->bufs) (void)hipFree(b.p);
    N->bufs.clear();
}

// Capture `body(stream)` once into an executable graph.
template <typename F>
int capture(mdgpu_numeric* N, hipGraphExec_t* out, F body) {
    hipError_t e = hipStreamBeginCapture(N->cap_stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return hip_fail(e, "hipStreamBeginCapture");
    int st = body(N->cap_stream);
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(N->cap_stream, &g);
    if (st != MDGPU_OK) {
        if (g) (void)hipGraphDestroy(g);
        return st;
    }
    if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
    e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return hip_fail(e, "hipGraphInstantiate");
    return MDGPU_OK;
}

#define TRY(x)                        \
    do {                              \
        int _st = (x);                \
        if (_st != MDGPU_OK) return _st; \
    } while (0)

// The work of the calls on one handle runs on the GPU in call order,
// whatever streams the calls use (they share the handle's buffers: the L
// store, the solve vectors, the chain counters): each call's stream first
// waits for the event the previous call recorded, and the call records it
// again on its own stream when it returns.
struct CallOrder {
    mdgpu_numeric* N;
    hipStream_t s;
    CallOrder(mdgpu_numeric* n, hipStream_t st) : N(n), s(st) {
        if (N->done_recorded) (void)hipStreamWaitEvent(s, N->done, 0);
    }
    ~CallOrder() {
        if (hipEventRecord(N->done, s) == hipSuccess) N->done_recorded = true;
    }
};
// Host copies out of the handle's buffers wait for all of its GPU work.
int wait_handle(mdgpu_numeric* N) {
    if (!N->done_recorded) return MDGPU_OK;
    hipError_t e = hipEventSynchronize(N->done);
    return e == hipSuccess ? MDGPU_OK : hip_fail(e, "hipEventSynchronize");
}

// Kernel classes reported by mdgpu_factorize_profile.
enum { KIND_ASSEMBLY = 0, KIND_EXTEND_ADD, KIND_PANEL, KIND_POTRF_TRSM, KIND_INVERSE, KIND_GEMM, KIND_N };
static_assert(KIND_N == MDGPU_KIND_N, "kernel classes of mdgpu_factorize_profile");

int kind_of(int step_kind) {
    switch (step_kind) {
        case mdgpu::STEP_EA1:
        case mdgpu::STEP_EA2: return KIND_EXTEND_ADD;
        case mdgpu::STEP_PANEL: return KIND_PANEL;
        case mdgpu::STEP_BLK: return KIND_POTRF_TRSM;
        case mdgpu::STEP_INV:
        case mdgpu::STEP_INV256: return KIND_INVERSE;
        default: return KIND_GEMM;
    }
}

// One factorization; with `ev` non-null an event is recorded on the launch
// stream before the first kernel and after every step (ev has steps+2 slots).
template <typename T>
int run_factor(mdgpu_numeric* N, const double* ax, hipStream_t s, hipEvent_t* ev) {
    const mdgpu::Plan& P = N->plan;
    T* L = (T*)N->dL;
    T* U = (T*)N->dU;
    T* inv = (T*)N->dInv;
    if (ev) hipEventRecord(ev[0], s);
    mdgpu::launch_fill<T>(L, P.lstore, T(0), s);
    mdgpu::launch_fill<int>(N->dInfo, 1, INT_MAX, s);
    mdgpu::launch_fill<int>(N->dInfo + 1, 1, 0, s);     // FP64 diagonal fallbacks
    if (N->npanel_steps > 0) mdgpu::launch_fill<int>(N->dPanelFbCnt, N->npanel_steps, 0, s);
    mdgpu::launch_scatter_a<T>(L, ax, N->dAdst, (int)P.a_dst.size(), s);
    if (ev) hipEventRecord(ev[1], s);
    const hipStream_t main_s = s;
    int panel_step = 0;
    for (size_t t = 0; t < P.steps.size(); t++) {
        const mdgpu::Step& st = P.steps[t];
        // look-ahead (not when timing every step): side-stream launches and
        // their event waits / records
        s = ev ? main_s : st.stream == 1 ? N->side : st.stream == 2 ? N->side2 : main_s;
        if (!ev && st.wait_ev >= 0) hipStreamWaitEvent(s, N->events[st.wait_ev], 0);
        if (!ev && st.wait_ev2 >= 0) hipStreamWaitEvent(s, N->events[st.wait_ev2], 0);
        switch (st.kind) {
            case mdgpu::STEP_EA1:
            case mdgpu::STEP_EA2:
                mdgpu::launch_extend_add<T>(N->dEa + st.first, st.count, N->dEaContrib, N->dRel, L, U, s);
                break;
            case mdgpu::STEP_PANEL:
                mdgpu::launch_panel<T>(N->dPanel + st.first, st.count, N->dFronts, L, N->dInfo, st.arg,
                                       N->dPanelFbCnt + panel_step, N->dPanelFbList + st.first, N->dInfo + 1, s);
                panel_step++;
                break;
            case mdgpu::STEP_BLK:
                mdgpu::launch_blk<T>(N->dBlk + st.first, st.count, st.nblocks, L, P.lstore, N->dInfo, N->dBlkCnt,
                                     N->dInfo + 1, N->dPlanes, N->plane_stride, N->dPlanes ? N->po.nplanes : 0, s);
                break;
            case mdgpu::STEP_GEMM: {
                // side-stream launches run on a bounded grid (look-ahead only)
                const int cap = (!ev && st.stream != 0) ? N->side_grid : 0;
                if constexpr (std::is_same<T, float>::value) {
                    const int nsplit = N->precision == MDGPU_BF16X3 ? 3 : 1;
                    if (mdgpu::gemm_planes(st.arg)) {
                        mdgpu::launch_gemm_bf16_planes(N->dGemm + st.first, st.count, st.nblocks, cap,
                                                       mdgpu::gemm_six(st.arg) ? 6 : nsplit, N->dPlanes,
                                                       N->plane_stride, U, L, s);
                        break;
                    }
                    if ((N->precision == MDGPU_BF16 || N->precision == MDGPU_BF16X3) && st.arg == 128) {
                        mdgpu::launch_gemm_bf16(N->dGemm + st.first, st.count, st.nblocks, cap, nsplit, L, U, s);
                        break;
                    }
                }
                mdgpu::launch_gemm<T>(N->dGemm + st.first, st.count, st.nblocks, cap, st.arg, L, P.lstore, U, s,
                                      (!ev && st.stream != 0) ? N->side_lds : 0);
                break;
            }
            case mdgpu::STEP_INV:
                mdgpu::launch_inv<T>(N->dInvJobs + st.first, st.count, N->dFronts, L, inv, s);
                break;
            case mdgpu::STEP_INV256:
                mdgpu::launch_inv256<T>(N->dInv256Jobs + st.first, st.count, N->dFronts, L, inv, (T*)N->dInv256, s);
                break;
        }
        if (!ev && st.rec_ev >= 0) hipEventRecord(N->events[st.rec_ev], s);
        if (ev) hipEventRecord(ev[t + 2], s);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "factorization launch");
    return MDGPU_OK;
}

template <typename T>
int run_solve(mdgpu_numeric* N, const double* b, double* x, hipStream_t s) {
    const mdgpu::Plan& P = N->plan;
    const int n = N->sym->S.n;
    const T* L = (const T*)N->dL;
    const T* inv = (const T*)N->dInv;
    const T* X256 = (const T*)N->dInv256;
    mdgpu::launch_fill<int>(N->dChainCtr, P.sv_nctr, 0, s);
    mdgpu::launch_fill<unsigned long long>((unsigned long long*)N->dY, n, mdgpu::kChainEmptyBits, s);
    mdgpu::launch_permute_in(b, N->dPerm, N->dXp, n, s);
    for (const mdgpu::SolveStep& st : P.sv_fwd) {
        const int* list = N->dSvFronts + st.first;
        const mdgpu::SolveItem* items = N->dSvItems + st.first;
        switch (st.kind) {
            case mdgpu::SV_FWD_SMALL:
                mdgpu::launch_fwd<T>(list, st.count, N->dFronts, N->dKids, N->dRel, N->dQ, L, inv, N->dXp, N->dY,
                                     N->dSu, N->dSuofs, N->vec_small, s);
                break;
            case mdgpu::SV_FWD_GATHER:
                mdgpu::launch_fwd_gather(list, st.count, N->dFronts, N->dKids, N->dRel, N->dXp, N->dSu, N->dSuofs,
                                         N->vec_gather, s);
                break;
            case mdgpu::SV_FWD_CHAIN:
                mdgpu::launch_fwd_chain<T>(items, st.count, N->dChainCtr + st.ctr, N->dStatus, N->chain_spin, L, X256,
                                           N->dXp, N->dY, N->dSu, s);
                break;
        }
    }
    // every entry of xp is stored again by the backward sweep before it is
    // read; the chained backward waits for kChainEmpty to go
    mdgpu::launch_fill<unsigned long long>((unsigned long long*)N->dXp, n, mdgpu::kChainEmptyBits, s);
    for (const mdgpu::SolveStep& st : P.sv_bwd) {
        const int* list = N->dSvFronts + st.first;
        const mdgpu::SolveItem* items = N->dSvItems + st.first;
        switch (st.kind) {
            case mdgpu::SV_BWD_SMALL:
                mdgpu::launch_bwd<T>(list, st.count, N->dFronts, N->dQ, L, inv, N->dXp, N->dY, s);
                break;
            case mdgpu::SV_BWD_UPD:
                mdgpu::launch_bwd_upd<T>(items, st.count, N->dFronts, N->dQ, L, N->dXp, N->dUpdPartial, s);
                break;
            case mdgpu::SV_BWD_UPD_SUM:
                mdgpu::launch_bwd_upd_sum(items, st.count, N->dFronts, N->dUpdPartial, N->dY, s);
                break;
            case mdgpu::SV_BWD_CHAIN:
                mdgpu::launch_bwd_chain<T>(items, st.count, N->dChainCtr + st.ctr, N->dStatus, N->chain_spin, L, X256,
                                           N->dXp, N->dY, s);
                break;
        }
    }
    mdgpu::launch_permute_out(N->dXp, N->dPerm, x, n, s);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "solve launch");
    return MDGPU_OK;
}

int solve_direct(mdgpu_numeric* N, const double* b, double* x, hipStream_t s) {
    if (N->precision == MDGPU_FP64) return run_solve<double>(N, b, x, s);
    return run_solve<float>(N, b, x, s);
}

int solve_any(mdgpu_numeric* N, const double* b, double* x, hipStream_t s) {
    if (!N->use_graph) return solve_direct(N, b, x, s);
    const size_t bytes = sizeof(double) * N->sym->S.n;
    if (!N->g_solve) {
        int st = capture(N, &N->g_solve, [&](hipStream_t cs) { return solve_direct(N, N->dB, N->dX, cs); });
        if (st != MDGPU_OK) return st;
    }
    hipError_t e = hipMemcpyAsync(N->dB, b, bytes, hipMemcpyDeviceToDevice, s);
    if (e == hipSuccess) e = hipGraphLaunch(N->g_solve, s);
    if (e == hipSuccess) e = hipMemcpyAsync(x, N->dX, bytes, hipMemcpyDeviceToDevice, s);
    if (e != hipSuccess) return hip_fail(e, "solve graph");
    return MDGPU_OK;
}

// max |v| per segment of each of the vectors v[0..nv) (nv <= 3), one copy
// to the host and one synchronization of the stream for all of them; the
// same copy brings the solve status bits (*status, if not null)
int max_abs_n(mdgpu_numeric* N, const double* const* v, int nv, int n, hipStream_t s,
              std::vector<double>* const* out, int* status = nullptr) {
    const int ns = N->nseg;
    mdgpu::launch_fill<unsigned long long>(N->dMax, (int64_t)nv * ns, 0ULL, s);
    for (int t = 0; t < nv; t++) {
        if (ns == 1) mdgpu::launch_maxabs(v[t], n, N->dMax + t, s);
        else mdgpu::launch_maxabs_seg(v[t], N->dSeg, n, N->dMax + (int64_t)t * ns, s);
    }
    std::vector<unsigned long long> bits((size_t)3 * ns + 1, 0);
    hipError_t e = hipMemcpyAsync(bits.data(), N->dMax, bits.size() * sizeof(unsigned long long),
                                  hipMemcpyDeviceToHost, s);
    if (e != hipSuccess) return hip_fail(e, "hipMemcpyAsync");
    e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
    for (int t = 0; t < nv; t++) {
        out[t]->resize(ns);
        for (int k = 0; k < ns; k++) memcpy(&(*out[t])[k], &bits[(size_t)t * ns + k], sizeof(double));
    }
    if (status) memcpy(status, &bits[(size_t)3 * ns], sizeof(int));
    return MDGPU_OK;
}
// max |v| per segment (synchronizes the stream)
int max_abs(mdgpu_numeric* N, const double* v, int n, hipStream_t s, std::vector<double>& out) {
    std::vector<double>* o[1] = {&out};
    return max_abs_n(N, &v, 1, n, s, o);
}

}  // namespace

extern "C" int mdgpu_numeric_create(const mdgpu_symbolic* sym, int precision, mdgpu_numeric** out) {
    mdgpu::clear_error();
    if (!sym || !out) return fail(MDGPU_ERR_ARG, "mdgpu_numeric_create: null pointer");
    *out = nullptr;
    if (precision < MDGPU_FP64 || precision > MDGPU_BF16X3)
        return fail(MDGPU_ERR_ARG, "mdgpu_numeric_create: precision must be MDGPU_FP64, MDGPU_FP32, MDGPU_BF16 or MDGPU_BF16X3");
    mdgpu_numeric* N = new (std::nothrow) mdgpu_numeric();
    if (!N) return fail(MDGPU_ERR_NOMEM, "out of host memory");
    N->sym = sym;
    N->precision = precision;
    hipGetDevice(&N->device);
    mdgpu::PlanOptions& po = N->po;
    po = mdgpu::plan_options_for(precision);
    std::string err;
    try {
        err = mdgpu::build_plan(sym->S, po, N->plan);
        if (err.empty()) err = mdgpu::validate_plan(N->plan, po);
    } catch (const std::bad_alloc&) {
        delete N;
        return fail(MDGPU_ERR_NOMEM, "out of host memory building the plan");
    }
    if (!err.empty()) {
        delete N;
        return fail(MDGPU_ERR_ARG, "mdgpu_numeric_create: " + err);
    }
    const mdgpu::Plan& P = N->plan;
    const mdgpu::Symbolic& S = sym->S;
    const size_t es = po.elem_bytes;
    int st = MDGPU_OK;
    auto guard = [&](int code) {
        if (code != MDGPU_OK) { release(N); delete N; }
        return code;
    };
    std::vector<int> level_list;
    N->level_off.assign(P.nlevels + 1, 0);
    for (int lv = 0; lv < P.nlevels; lv++) {
        N->level_off[lv] = (int)level_list.size();
        level_list.insert(level_list.end(), P.level_fronts[lv].begin(), P.level_fronts[lv].end());
    }
    N->level_off[P.nlevels] = (int)level_list.size();
    std::vector<int64_t> diag_off(S.n);
    for (int f = 0; f < P.nfronts; f++) {
        const mdgpu::FrontDesc& d = P.fronts[f];
        for (int j = 0; j < d.c; j++) diag_off[d.k + j] = d.lofs + (int64_t)j * d.n + j;
    }
    for (const mdgpu::Step& stp : P.steps) N->npanel_steps += stp.kind == mdgpu::STEP_PANEL;
    N->nparts = kLogParts;
    N->nseg = std::max(1, (int)sym->seg_col.size() - 1);
    std::vector<int> seg_of_row;
    if (N->nseg > 1) {
        seg_of_row.resize(S.n);
        for (int k = 0; k < N->nseg; k++)
            for (int r = sym->seg_col[k]; r < sym->seg_col[k + 1]; r++) seg_of_row[r] = k;
    }
    if (const char* g = std::getenv("MDGPU_GRAPH")) N->use_graph = std::atoi(g) != 0;
    if (const char* g = std::getenv("MDGPU_SIDE_GRID")) N->side_grid = std::max(0, std::atoi(g)) / 8 * 8;
    // FP32 side-stream GEMM workgroups get an LDS floor just above a third
    // of the CU's LDS: at most two share a CU, and a block-step workgroup of
    // the main stream always fits beside them instead of waiting for a GEMM
    // tile to finish (both GEMM kernels' registers leave room for it)
    if (precision == MDGPU_FP32) {
        int cu_lds = 0;
        if (hipDeviceGetAttribute(&cu_lds, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, N->device) ==
                hipSuccess && cu_lds > 0) {
            const int floor_lds = (cu_lds / 3 + 256) / 256 * 256;
            const int blk = mdgpu::blk_lds_bytes_f32();
            if (blk > 0 && 2 * floor_lds + blk <= cu_lds) N->side_lds = floor_lds;
        }
    }
    if (const char* g = std::getenv("MDGPU_SIDE_LDS")) N->side_lds = std::max(0, std::atoi(g));
    // 64 elements past the L store: the FP32 GEMM's 16-byte buffer loads
    // are bounded by the store's end at their first element
    if ((st = alloc(N, (P.lstore + 64) * es, &N->dL)) ||
        (st = alloc(N, P.arena * es, &N->dU)) ||
        (st = alloc(N, P.dinv * es, &N->dInv)) ||
        (st = upload(N, P.inv, (void**)&N->dInvJobs)) ||
        (st = alloc(N, P.inv256 * es, &N->dInv256)) ||
        (st = alloc(N, P.sv_nctr * sizeof(int), (void**)&N->dChainCtr)) ||
        (st = alloc(N, P.sv_upd_partial * sizeof(double), (void**)&N->dUpdPartial)) ||
        (st = upload(N, P.inv256_jobs, (void**)&N->dInv256Jobs)) ||
        (st = upload(N, P.fronts, (void**)&N->dFronts)) ||
        (st = upload(N, P.kids, (void**)&N->dKids)) ||
        (st = upload(N, P.rel, (void**)&N->dRel)) ||
        (st = upload(N, P.q, (void**)&N->dQ)) ||
        (st = upload(N, P.a_dst, (void**)&N->dAdst)) ||
        (st = upload(N, P.ea, (void**)&N->dEa)) ||
        (st = upload(N, P.ea_contrib, (void**)&N->dEaContrib)) ||
        (st = upload(N, P.gemm, (void**)&N->dGemm)) ||
        (P.planes > 0 && (st = alloc(N, po.nplanes * (size_t)P.planes * sizeof(__bf16), (void**)&N->dPlanes))) ||
        (st = upload(N, P.blk, (void**)&N->dBlk)) ||
        (st = upload(N, std::vector<int>(P.blk.size(), 0), (void**)&N->dBlkCnt)) ||
        (st = upload(N, P.panel, (void**)&N->dPanel)) ||
        (st = alloc(N, std::max<size_t>(P.panel.size(), 1) * sizeof(int), (void**)&N->dPanelFbList)) ||
        (st = alloc(N, std::max(N->npanel_steps, 1) * sizeof(int), (void**)&N->dPanelFbCnt)) ||
        (st = upload(N, level_list, (void**)&N->dLevelList)) ||
        (st = alloc(N, P.solve_u * sizeof(double), (void**)&N->dSu)) ||
        (st = upload(N, P.solve_uofs, (void**)&N->dSuofs)) ||
        (st = upload(N, S.perm, (void**)&N->dPerm)) ||
        (st = upload(N, P.afull_p, (void**)&N->dAfp)) ||
        (st = upload(N, P.afull_i, (void**)&N->dAfi)) ||
        (st = upload(N, P.afull_src, (void**)&N->dAfsrc)) ||
        (st = alloc(N, P.afull_i.size() * sizeof(double), (void**)&N->dAfx)) ||
        (st = alloc(N, (size_t)S.n * sizeof(double), (void**)&N->dXp)) ||
        (st = alloc(N, (size_t)S.n * sizeof(double), (void**)&N->dR)) ||
        (st = alloc(N, (size_t)S.n * sizeof(double), (void**)&N->dD)) ||
        (st = alloc(N, (size_t)S.n * sizeof(double), (void**)&N->dP)) ||
        (st = alloc(N, (size_t)S.n * sizeof(double), (void**)&N->dQ2)) ||
        (st = alloc(N, (8 + mdgpu::kDotParts) * sizeof(double), (void**)&N->dScal)) ||
        (st = alloc(N, (size_t)S.n * sizeof(double), (void**)&N->dRowabs)) ||
        (st = upload(N, diag_off, (void**)&N->dDiagOff)) ||
        (st = alloc(N, 2 * sizeof(int), (void**)&N->dInfo)) ||
        (st = alloc(N, (3 * N->nseg + 1) * sizeof(unsigned long long), (void**)&N->dMax)) ||
        (N->nseg > 1 && (st = upload(N, seg_of_row, (void**)&N->dSeg))) ||
        (st = alloc(N, kLogParts * sizeof(double), (void**)&N->dPartial)) ||
        (st = alloc(N, std::max<size_t>(P.a_dst.size(), 1) * sizeof(double), (void**)&N->dAx)) ||
        (st = alloc(N, (size_t)S.n * sizeof(double), (void**)&N->dB)) ||
        (st = alloc(N, (size_t)S.n * sizeof(double), (void**)&N->dX)) ||
        (st = upload(N, P.sv_fronts, (void**)&N->dSvFronts)) ||
        (st = upload(N, P.sv_items, (void**)&N->dSvItems)) ||
        (st = alloc(N, (size_t)S.n * sizeof(double), (void**)&N->dY)))
        return guard(st);
    if (hipStreamCreateWithFlags(&N->cap_stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&N->side, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&N->side2, hipStreamNonBlocking) != hipSuccess)
        return guard(fail(MDGPU_ERR_DEVICE, "mdgpu_numeric_create: hipStreamCreate failed"));
    N->dStatus = reinterpret_cast<int*>(N->dMax + 3 * N->nseg);
    N->plane_stride = P.planes;
    if (hipMemset(N->dMax, 0, (3 * N->nseg + 1) * sizeof(unsigned long long)) != hipSuccess ||
        hipEventCreateWithFlags(&N->done, hipEventDisableTiming) != hipSuccess)
        return guard(fail(MDGPU_ERR_DEVICE, "mdgpu_numeric_create: status word / event set-up failed"));
    // longest front vector the forward solve keeps in LDS (default: the
    // kernels' capacity; a smaller value sends longer fronts through the
    // in-place global-memory path)
    if (const char* e = std::getenv("MDGPU_SOLVE_LDS_ROWS")) {
        const int v = std::max(0, std::atoi(e));
        N->vec_small = std::min(v, mdgpu::kSmallVec);
        N->vec_gather = std::min(v, mdgpu::kGatherVec);
    }
    // polls before a chained-solve wait gives up (a smaller value makes the
    // give-up path observable without a stalled producer)
    if (const char* e = std::getenv("MDGPU_CHAIN_SPIN")) N->chain_spin = std::max(0, std::atoi(e));
    // refinement method: classical iterative refinement or conjugate
    // gradients preconditioned by the factor ("cg"; fewer solves when the
    // factor is poor).  Single bf16 factors default to CG (on the 64^3
    // 27-point stencil it reaches 1e-12 after 5 solves, classical after 6);
    // FP32 and BF16X3 factors need the same solves either way and default
    // to the classical steps, which cost one SpMV and two dots less each.
    N->refine_cg = precision == MDGPU_BF16;
    if (const char* e = std::getenv("MDGPU_REFINE")) N->refine_cg = std::string(e) == "cg";
    N->events.assign(P.nevents, nullptr);
    for (hipEvent_t& e : N->events)
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess)
            return guard(fail(MDGPU_ERR_DEVICE, "mdgpu_numeric_create: hipEventCreate failed"));
    *out = N;
    return MDGPU_OK;
}

extern "C" int mdgpu_numeric_set_graph(mdgpu_numeric* num, int enable) {
    if (!num) return fail(MDGPU_ERR_ARG, "mdgpu_numeric_set_graph: null pointer");
    num->use_graph = enable != 0;
    return MDGPU_OK;
}

extern "C" int mdgpu_numeric_destroy(mdgpu_numeric* num) {
    if (!num) return MDGPU_OK;
    release(num);
    delete num;
    return MDGPU_OK;
}

extern "C" int mdgpu_numeric_info(const mdgpu_numeric* num, int64_t* out) {
    if (!num || !out) return fail(MDGPU_ERR_ARG, "mdgpu_numeric_info: null pointer");
    const mdgpu::Plan& P = num->plan;
    out[MDGPU_NUM_DEVICE_BYTES] = num->device_bytes;
    out[MDGPU_NUM_LSTORE] = P.lstore;
    out[MDGPU_NUM_ARENA] = P.arena;
    out[MDGPU_NUM_STEPS] = (int64_t)P.steps.size();
    out[MDGPU_NUM_LEVELS] = P.nlevels;
    out[MDGPU_NUM_PANEL_FRONTS] = (int64_t)P.panel.size();
    out[MDGPU_NUM_DIAG_JOBS] = (int64_t)P.blk.size();
    out[MDGPU_NUM_PRECISION] = num->precision;
    out[MDGPU_NUM_FALLBACKS] = num->fallbacks;
    return MDGPU_OK;
}

namespace {
int finish_factor(mdgpu_numeric* num, const double* d_ax, hipStream_t s);
}

extern "C" int mdgpu_factorize(mdgpu_numeric* num, const double* d_ax, void* stream) {
    mdgpu::clear_error();
    if (!num || !d_ax) return fail(MDGPU_ERR_ARG, "mdgpu_factorize: null pointer");
    hipStream_t s = (hipStream_t)stream;
    CallOrder order(num, s);
    num->factored = false;
    int st;
    if (num->use_graph) {
        if (!num->g_factor) {
            st = capture(num, &num->g_factor, [&](hipStream_t cs) {
                return num->precision == MDGPU_FP64 ? run_factor<double>(num, num->dAx, cs, nullptr)
                                                    : run_factor<float>(num, num->dAx, cs, nullptr);
            });
            if (st != MDGPU_OK) return st;
        }
        hipError_t e = hipMemcpyAsync(num->dAx, d_ax, sizeof(double) * num->plan.a_dst.size(),
                                      hipMemcpyDeviceToDevice, s);
        if (e == hipSuccess) e = hipGraphLaunch(num->g_factor, s);
        if (e != hipSuccess) return hip_fail(e, "factorization graph");
        st = MDGPU_OK;
    } else {
        st = num->precision == MDGPU_FP64 ? run_factor<double>(num, d_ax, s, nullptr)
                                          : run_factor<float>(num, d_ax, s, nullptr);
    }
    if (st != MDGPU_OK) return st;
    return finish_factor(num, d_ax, s);
}

namespace {
// One factorization with a HIP event around every step on the launch
// stream; ms[0] = zero fill + A scatter, ms[t + 1] = step t.
int factor_timed(mdgpu_numeric* num, const double* d_ax, hipStream_t s, std::vector<float>& ms) {
    num->factored = false;
    const mdgpu::Plan& P = num->plan;
    std::vector<hipEvent_t> ev(P.steps.size() + 2);
    for (auto& e : ev)
        if (hipEventCreate(&e) != hipSuccess) return fail(MDGPU_ERR_DEVICE, "hipEventCreate failed");
    int st = num->precision == MDGPU_FP64 ? run_factor<double>(num, d_ax, s, ev.data())
                                          : run_factor<float>(num, d_ax, s, ev.data());
    if (st == MDGPU_OK) {
        hipError_t e = hipEventSynchronize(ev.back());
        if (e != hipSuccess) st = hip_fail(e, "hipEventSynchronize");
    }
    if (st == MDGPU_OK) {
        ms.assign(P.steps.size() + 1, 0.0f);
        for (size_t t = 0; t + 1 < ev.size(); t++) hipEventElapsedTime(&ms[t], ev[t], ev[t + 1]);
    }
    for (auto& e : ev) hipEventDestroy(e);
    if (st != MDGPU_OK) return st;
    return finish_factor(num, d_ax, s);
}
}  // namespace

extern "C" int mdgpu_factorize_profile(mdgpu_numeric* num, const double* d_ax, void* stream, double* out) {
    mdgpu::clear_error();
    if (!num || !d_ax || !out) return fail(MDGPU_ERR_ARG, "mdgpu_factorize_profile: null pointer");
    CallOrder order(num, (hipStream_t)stream);
    std::vector<float> ms;
    TRY(factor_timed(num, d_ax, (hipStream_t)stream, ms));
    const mdgpu::Plan& P = num->plan;
    const mdgpu::Symbolic& S = num->sym->S;
    const double es = num->precision == MDGPU_FP64 ? 8 : 4;
    for (int k = 0; k < 4 * KIND_N; k++) out[k] = 0;
    out[4 * KIND_ASSEMBLY + 0] += ms[0];
    out[4 * KIND_ASSEMBLY + 1] += 3;
    out[4 * KIND_ASSEMBLY + 3] += P.lstore * es + S.nnz_in * (8.0 + 8.0 + es);
    for (size_t t = 0; t < P.steps.size(); t++) {
        const int k = kind_of(P.steps[t].kind);
        out[4 * k + 0] += ms[t + 1];
        out[4 * k + 1] += 1;
        out[4 * k + 2] += P.steps[t].flops;
        out[4 * k + 3] += P.steps[t].bytes;
    }
    return MDGPU_OK;
}

extern "C" int mdgpu_factorize_steps(mdgpu_numeric* num, const double* d_ax, void* stream, double* out, int64_t cap) {
    mdgpu::clear_error();
    if (!num || !d_ax || !out) return fail(MDGPU_ERR_ARG, "mdgpu_factorize_steps: null pointer");
    const mdgpu::Plan& P = num->plan;
    if (cap < (int64_t)P.steps.size() * 8)
        return fail(MDGPU_ERR_ARG, "mdgpu_factorize_steps: capacity below 8 * MDGPU_NUM_STEPS");
    CallOrder order(num, (hipStream_t)stream);
    std::vector<float> ms;
    TRY(factor_timed(num, d_ax, (hipStream_t)stream, ms));
    for (size_t t = 0; t < P.steps.size(); t++) {
        const mdgpu::Step& st = P.steps[t];
        double* o = out + 8 * t;
        o[0] = kind_of(st.kind);
        o[1] = ms[t + 1];
        o[2] = st.nblocks;
        o[3] = st.flops;
        o[4] = st.bytes;
        o[5] = st.level;
        o[6] = st.stream;
        o[7] = st.kind == mdgpu::STEP_GEMM ? st.arg : 0;
    }
    return MDGPU_OK;
}

namespace {
int finish_factor(mdgpu_numeric* num, const double* d_ax, hipStream_t s) {
    num->z_valid = false;
    // full symmetric values and ||A||_1 for the residual of refinement
    const int n = num->sym->S.n;
    mdgpu::launch_gather(d_ax, num->dAfsrc, num->dAfx, (int)num->plan.afull_i.size(), s);
    mdgpu::launch_rowabs(num->dAfp, num->dAfx, num->dRowabs, n, s);
    int info2[2] = {INT_MAX, 0};
    hipError_t e = hipMemcpyAsync(info2, num->dInfo, 2 * sizeof(int), hipMemcpyDeviceToHost, s);
    if (e != hipSuccess) return hip_fail(e, "hipMemcpyAsync");
    int st = max_abs(num, num->dRowabs, n, s, num->anorm_seg);   // synchronizes the stream
    if (st != MDGPU_OK) return st;
    num->anorm = *std::max_element(num->anorm_seg.begin(), num->anorm_seg.end());
    const int info = info2[0];
    num->fallbacks = info2[1];
    if (info != INT_MAX)
        return fail(MDGPU_ERR_NOT_SPD, "mdgpu_factorize: non-positive pivot at permuted column " +
                                           std::to_string(info));
    num->factored = true;
    return MDGPU_OK;
}
}  // namespace

extern "C" int mdgpu_solve(mdgpu_numeric* num, const double* d_b, double* d_x, void* stream) {
    mdgpu::clear_error();
    if (!num || !d_b || !d_x) return fail(MDGPU_ERR_ARG, "mdgpu_solve: null pointer");
    if (!num->factored) return fail(MDGPU_ERR_ARG, "mdgpu_solve: no valid factorization");
    CallOrder order(num, (hipStream_t)stream);
    return solve_any(num, d_b, d_x, (hipStream_t)stream);
}

extern "C" int mdgpu_solve_nrhs(mdgpu_numeric* num, const double* d_B, double* d_X, int nrhs, int64_t ld,
                                void* stream) {
    mdgpu::clear_error();
    if (!num || (nrhs > 0 && (!d_B || !d_X))) return fail(MDGPU_ERR_ARG, "mdgpu_solve_nrhs: null pointer");
    if (!num->factored) return fail(MDGPU_ERR_ARG, "mdgpu_solve_nrhs: no valid factorization");
    if (nrhs < 0 || ld < num->sym->S.n) return fail(MDGPU_ERR_ARG, "mdgpu_solve_nrhs: nrhs < 0 or ld < n");
    CallOrder order(num, (hipStream_t)stream);
    for (int r = 0; r < nrhs; r++) TRY(solve_any(num, d_B + (int64_t)r * ld, d_X + (int64_t)r * ld, (hipStream_t)stream));
    return MDGPU_OK;
}

extern "C" int mdgpu_refine(mdgpu_numeric* num, const double* d_b, double* d_x, double tol, int maxit,
                            double* eta_out, int* iters_out, void* stream) {
    mdgpu::clear_error();
    if (!num || !d_b || !d_x) return fail(MDGPU_ERR_ARG, "mdgpu_refine: null pointer");
    if (!num->factored) return fail(MDGPU_ERR_ARG, "mdgpu_refine: no valid factorization");
    hipStream_t s = (hipStream_t)stream;
    CallOrder order(num, s);
    const int n = num->sym->S.n;
    // Solve k (k = 0, 1, ...) applies the factor M = L L^T to the residual
    // r_k = b - A x_k.  Classical refinement: x_{k+1} = x_k + M^-1 r_k
    // (x_0 = 0).  Conjugate gradients preconditioned by M: the step along
    // p_k = M^-1 r_k + beta_k p_{k-1} of optimal length, which minimizes
    // the A-norm error over the same Krylov space the classical iterates
    // lie in; the residual is recomputed from x every step (true residual).
    // The backward error (mdgpu.cu:1972) of x_{k+1} decides after solve k.
    double* r = num->dR;
    double* z = num->dD;
    double* scal = num->dScal;                  // {rz, rz_old, pq}, then the dot partials
    double* part = num->dScal + 8;
    std::vector<double> bn, rn, xn;
    double eta = INFINITY;
    int it = 0;
    num->eta_seg.assign(num->nseg, INFINITY);
    hipError_t e = hipMemcpyAsync(r, d_b, sizeof(double) * n, hipMemcpyDeviceToDevice, s);
    if (e != hipSuccess) return hip_fail(e, "mdgpu_refine: copy");
    for (;; it++) {
        if (num->refine_cg) {
            TRY(solve_any(num, r, z, s));
            mdgpu::launch_dot(r, z, n, part, scal + 0, s);
            mdgpu::launch_cg_dir(num->dP, z, scal, n, it == 0, s);
            mdgpu::launch_spmv(num->dAfp, num->dAfi, num->dAfx, num->dP, num->dQ2, n, s);
            mdgpu::launch_dot(num->dP, num->dQ2, n, part, scal + 2, s);
            if (it == 0) mdgpu::launch_fill<double>(d_x, n, 0.0, s);
            mdgpu::launch_cg_step(d_x, num->dP, scal, n, s);
            mdgpu::launch_copy_scalar(scal + 1, scal + 0, s);
        } else if (it == 0) {
            TRY(solve_any(num, r, d_x, s));
        } else {
            TRY(solve_any(num, r, z, s));
            mdgpu::launch_axpy(d_x, z, n, s);
        }
        mdgpu::launch_residual(num->dAfp, num->dAfi, num->dAfx, d_x, d_b, r, n, s);
        int status = 0;
        {
            // |b| with the first step's |r| and |x|: one round trip per step
            const double* vs[3] = {r, d_x, d_b};
            std::vector<double>* os[3] = {&rn, &xn, &bn};
            TRY(max_abs_n(num, vs, it == 0 ? 3 : 2, n, s, os, &status));
        }
        if (status & mdgpu::kSolveTimedOut) {
            mdgpu::launch_fill<int>(num->dStatus, 1, 0, s);
            return fail(MDGPU_ERR_DEVICE, "mdgpu_refine: a chained solve gave up waiting for a block "
                                          "(x holds NaN)");
        }
        // backward error of the reference's check (mdgpu.cu:1972), per
        // independent matrix of a batch; the worst one decides
        eta = 0;
        for (int k = 0; k < num->nseg; k++) {
            const double den = num->anorm_seg[k] * xn[k] + bn[k];
            const double ek = den > 0 ? rn[k] / den : rn[k];
            num->eta_seg[k] = ek;
            eta = (ek > eta || ek != ek) ? ek : eta;
        }
        if (eta <= tol || it >= maxit || !(eta == eta)) break;
    }
    if (eta_out) *eta_out = eta;
    if (iters_out) *iters_out = it;
    e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "refine launch");
    if (!(eta <= tol))
        return fail(MDGPU_ERR_NOT_CONVERGED, "mdgpu_refine: backward error " + std::to_string(eta) +
                                                 " above tolerance after " + std::to_string(it) + " steps");
    return MDGPU_OK;
}

extern "C" int mdgpu_refine_segment_eta(const mdgpu_numeric* num, double* out, int cap) {
    mdgpu::clear_error();
    if (!num || !out) return fail(MDGPU_ERR_ARG, "mdgpu_refine_segment_eta: null pointer");
    if (cap < num->nseg) return fail(MDGPU_ERR_ARG, "mdgpu_refine_segment_eta: capacity below the segment count");
    if ((int)num->eta_seg.size() != num->nseg)
        return fail(MDGPU_ERR_ARG, "mdgpu_refine_segment_eta: no refinement has run");
    for (int k = 0; k < num->nseg; k++) out[k] = num->eta_seg[k];
    return MDGPU_OK;
}

extern "C" int mdgpu_solve_status(mdgpu_numeric* num, int* status, void* stream) {
    mdgpu::clear_error();
    if (!num || !status) return fail(MDGPU_ERR_ARG, "mdgpu_solve_status: null pointer");
    hipStream_t s = (hipStream_t)stream;
    CallOrder order(num, s);
    int v = 0;
    hipError_t e = hipMemcpyAsync(&v, num->dStatus, sizeof(int), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_fail(e, "mdgpu_solve_status");
    if (v) mdgpu::launch_fill<int>(num->dStatus, 1, 0, s);
    *status = v;
    return MDGPU_OK;
}

extern "C" int mdgpu_logdet(mdgpu_numeric* num, double* out, void* stream) {
    mdgpu::clear_error();
    if (!num || !out) return fail(MDGPU_ERR_ARG, "mdgpu_logdet: null pointer");
    if (!num->factored) return fail(MDGPU_ERR_ARG, "mdgpu_logdet: no valid factorization");
    hipStream_t s = (hipStream_t)stream;
    CallOrder order(num, s);
    const int n = num->sym->S.n;
    if (num->precision == MDGPU_FP64)
        mdgpu::launch_logdiag<double>(num->dDiagOff, (const double*)num->dL, n, num->dPartial, num->nparts, s);
    else
        mdgpu::launch_logdiag<float>(num->dDiagOff, (const float*)num->dL, n, num->dPartial, num->nparts, s);
    std::vector<double> part(num->nparts);
    hipError_t e = hipMemcpyAsync(part.data(), num->dPartial, part.size() * sizeof(double),
                                  hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_fail(e, "logdet copy");
    double sum = 0;
    for (double v : part) sum += v;
    *out = 2.0 * sum;
    return MDGPU_OK;
}

namespace {
// L columns [c0, c1) (one matrix of a batch, or all of L) to host CSC in
// the reference's spL layout, rows and column pointers relative to c0.
int export_columns(mdgpu_numeric* num, int c0, int c1, int32_t* Lp, int32_t* Li, double* Lx, const char* who) {
    const mdgpu::Symbolic& S = num->sym->S;
    const mdgpu::Plan& P = num->plan;
    std::vector<int64_t> lp(c1 - c0 + 1, 0);
    for (int k = c0; k < c1; k++) lp[k - c0 + 1] = lp[k - c0] + S.count_final[k];
    if (lp.back() > INT32_MAX) return fail(MDGPU_ERR_ARG, std::string(who) + ": nnz(L) exceeds int32");
    // the L-store range of the fronts in [c0, c1)
    int64_t lo = INT64_MAX, hi = 0;
    for (int f = 0; f < P.nfronts; f++) {
        const mdgpu::FrontDesc& d = P.fronts[f];
        if (d.k < c0 || d.k >= c1) continue;
        lo = std::min(lo, d.lofs);
        hi = std::max(hi, d.lofs + (int64_t)d.n * d.c);
    }
    if (lo > hi) lo = hi = 0;
    std::vector<double> h(Lx ? hi - lo : 0);
    hipError_t e = hipSuccess;
    if (Lx && hi > lo) {
        if (num->precision == MDGPU_FP64) {
            e = hipMemcpy(h.data(), (const double*)num->dL + lo, (hi - lo) * sizeof(double), hipMemcpyDeviceToHost);
        } else {
            std::vector<float> hf(hi - lo);
            e = hipMemcpy(hf.data(), (const float*)num->dL + lo, (hi - lo) * sizeof(float), hipMemcpyDeviceToHost);
            for (int64_t t = 0; t < hi - lo; t++) h[t] = hf[t];
        }
    }
    if (e != hipSuccess) return hip_fail(e, "hipMemcpy L");
    if (Lp)
        for (size_t k = 0; k < lp.size(); k++) Lp[k] = (int32_t)lp[k];
    // column k+j holds rows q_f[j..n_f) of its front, diagonal first
    // (the spL write-out of FactorAux_CPU_3, mdgpu.cu:1346-1352)
    for (int f = 0; f < P.nfronts; f++) {
        const mdgpu::FrontDesc& d = P.fronts[f];
        if (d.k < c0 || d.k >= c1) continue;
        for (int j = 0; j < d.c; j++) {
            int64_t dst = lp[d.k + j - c0];
            for (int i = j; i < d.n; i++, dst++) {
                if (Li) Li[dst] = S.q_all[d.qofs + i] - c0;
                if (Lx) Lx[dst] = h[d.lofs - lo + (int64_t)j * d.n + i];
            }
        }
    }
    return MDGPU_OK;
}
}  // namespace

extern "C" int mdgpu_export_L(mdgpu_numeric* num, int32_t* Lp, int32_t* Li, double* Lx) {
    mdgpu::clear_error();
    if (!num) return fail(MDGPU_ERR_ARG, "mdgpu_export_L: null pointer");
    if (!num->factored) return fail(MDGPU_ERR_ARG, "mdgpu_export_L: no valid factorization");
    TRY(wait_handle(num));
    return export_columns(num, 0, num->sym->S.n, Lp, Li, Lx, "mdgpu_export_L");
}

extern "C" int mdgpu_export_L_segment(mdgpu_numeric* num, int seg, int32_t* Lp, int32_t* Li, double* Lx) {
    mdgpu::clear_error();
    if (!num) return fail(MDGPU_ERR_ARG, "mdgpu_export_L_segment: null pointer");
    if (!num->factored) return fail(MDGPU_ERR_ARG, "mdgpu_export_L_segment: no valid factorization");
    const std::vector<int>& sc = num->sym->seg_col;
    if (seg < 0 || seg + 1 >= (int)sc.size()) return fail(MDGPU_ERR_ARG, "mdgpu_export_L_segment: no such matrix");
    TRY(wait_handle(num));
    return export_columns(num, sc[seg], sc[seg + 1], Lp, Li, Lx, "mdgpu_export_L_segment");
}

// ------------------------------------------------------------ reverse mode
namespace {
int selinv_setup(mdgpu_numeric* N) {
    if (N->sel_ready) return MDGPU_OK;
    const mdgpu::Plan& P = N->plan;
    const mdgpu::Symbolic& S = N->sym->S;
    std::string err = mdgpu::build_selinv_plan(P, N->sel);
    if (err.empty()) err = mdgpu::validate_selinv_plan(P, N->sel);
    if (!err.empty()) return fail(MDGPU_ERR_ARG, "selected inverse schedule: " + err);
    // caller entries: row and column in the caller's numbering
    std::vector<int> arow(S.nnz_in), acol(S.nnz_in);
    for (int j = 0; j < S.n; j++)
        for (int t = S.a_low.p[j]; t < S.a_low.p[j + 1]; t++) {
            arow[S.a_src[t]] = S.perm[S.a_low.i[t]];
            acol[S.a_src[t]] = S.perm[j];
        }
    int st;
    if ((st = alloc(N, P.lstore * sizeof(double), (void**)&N->dZ)) ||
        (st = alloc(N, N->sel.zf_size * sizeof(double), (void**)&N->dZF)) ||
        (st = alloc(N, N->sel.lh_size * sizeof(double), (void**)&N->dLH)) ||
        (st = upload(N, N->sel.gemm, (void**)&N->dSelGemm)) ||
        (st = upload(N, N->sel.copy, (void**)&N->dSelCopy)) ||
        (st = upload(N, N->sel.zf_ofs, (void**)&N->dZfOfs)) ||
        (st = upload(N, arow, (void**)&N->dArow)) ||
        (st = upload(N, acol, (void**)&N->dAcol)) ||
        (st = alloc(N, (size_t)S.n * sizeof(double), (void**)&N->dBbar)))
        return st;
    N->sel_ready = true;
    return MDGPU_OK;
}

int check_fp64_factor(const mdgpu_numeric* num, const char* who) {
    if (!num) return fail(MDGPU_ERR_ARG, std::string(who) + ": null pointer");
    if (!num->factored) return fail(MDGPU_ERR_ARG, std::string(who) + ": no valid factorization");
    if (num->precision != MDGPU_FP64)
        return fail(MDGPU_ERR_ARG, std::string(who) + ": the reverse mode needs an MDGPU_FP64 factorization");
    return MDGPU_OK;
}
}  // namespace

extern "C" int mdgpu_selinv(mdgpu_numeric* num, void* stream) {
    mdgpu::clear_error();
    TRY(check_fp64_factor(num, "mdgpu_selinv"));
    TRY(selinv_setup(num));
    hipStream_t s = (hipStream_t)stream;
    CallOrder order(num, s);
    mdgpu::SelBuffers b;
    b.p[mdgpu::SB_L] = (double*)num->dL;
    b.p[mdgpu::SB_Z] = num->dZ;
    b.p[mdgpu::SB_ZF] = num->dZF;
    b.p[mdgpu::SB_LH] = num->dLH;
    b.p[mdgpu::SB_W] = (double*)num->dInv;
    b.p[mdgpu::SB_U] = (double*)num->dU;
    for (const mdgpu::SelStep& st : num->sel.steps) {
        if (st.kind == mdgpu::SS_GEMM)
            mdgpu::launch_sel_gemm(num->dSelGemm + st.first, st.count, st.nblocks, b, s);
        else
            mdgpu::launch_sel_copy(st.copy_kind, num->dSelCopy + st.first, st.count, num->dFronts, num->dRel,
                                   num->dZfOfs, b, s);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "selected inverse launch");
    num->z_valid = true;
    return MDGPU_OK;
}

extern "C" int mdgpu_selinv_export(mdgpu_numeric* num, double* Zx) {
    mdgpu::clear_error();
    TRY(check_fp64_factor(num, "mdgpu_selinv_export"));
    if (!num->z_valid) return fail(MDGPU_ERR_ARG, "mdgpu_selinv_export: run mdgpu_selinv first");
    if (!Zx) return fail(MDGPU_ERR_ARG, "mdgpu_selinv_export: null pointer");
    TRY(wait_handle(num));
    const mdgpu::Symbolic& S = num->sym->S;
    const mdgpu::Plan& P = num->plan;
    std::vector<double> h(P.lstore);
    hipError_t e = hipMemcpy(h.data(), num->dZ, P.lstore * sizeof(double), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(e, "hipMemcpy Z");
    for (int f = 0; f < P.nfronts; f++) {
        const mdgpu::FrontDesc& d = P.fronts[f];
        for (int j = 0; j < d.c; j++) {
            int64_t dst = S.lp[d.k + j];
            for (int i = j; i < d.n; i++, dst++) Zx[dst] = h[d.lofs + (int64_t)j * d.n + i];
        }
    }
    return MDGPU_OK;
}

extern "C" int mdgpu_logdet_grad(mdgpu_numeric* num, double* d_g, void* stream) {
    mdgpu::clear_error();
    TRY(check_fp64_factor(num, "mdgpu_logdet_grad"));
    if (!d_g) return fail(MDGPU_ERR_ARG, "mdgpu_logdet_grad: null pointer");
    if (!num->z_valid) TRY(mdgpu_selinv(num, stream));
    CallOrder order(num, (hipStream_t)stream);
    mdgpu::launch_logdet_grad(num->dZ, num->dAdst, num->dArow, num->dAcol, d_g, num->sym->S.nnz_in,
                              (hipStream_t)stream);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? MDGPU_OK : hip_fail(e, "logdet gradient launch");
}

extern "C" int mdgpu_solve_adjoint(mdgpu_numeric* num, const double* d_x, const double* d_xbar, double* d_bbar,
                                   double* d_g, void* stream) {
    mdgpu::clear_error();
    TRY(check_fp64_factor(num, "mdgpu_solve_adjoint"));
    if (!d_x || !d_xbar || !d_g) return fail(MDGPU_ERR_ARG, "mdgpu_solve_adjoint: null pointer");
    TRY(selinv_setup(num));
    hipStream_t s = (hipStream_t)stream;
    CallOrder order(num, s);
    double* bb = d_bbar ? d_bbar : num->dBbar;
    TRY(solve_any(num, d_xbar, bb, s));
    mdgpu::launch_solve_adjoint(d_x, bb, num->dArow, num->dAcol, d_g, num->sym->S.nnz_in, s);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? MDGPU_OK : hip_fail(e, "solve adjoint launch");
}

// ------------------------------------------------------------ autotuning
namespace {
// Rebuild the plan with per-launch GEMM tiles `tiles` and install it: only
// the GEMM job table and the launch records change (the schedule's shape
// does not depend on tile sizes), so the device job table is overwritten in
// place.  Captured graphs are dropped (recaptured on next use).
int install_gemm_tiles(mdgpu_numeric* N, const std::vector<int>& tiles) {
    mdgpu::PlanOptions po = N->po;
    po.gemm_tiles = tiles;
    mdgpu::Plan P;
    std::string err = mdgpu::build_plan(N->sym->S, po, P);
    if (err.empty()) err = mdgpu::validate_plan(P, po);
    if (!err.empty()) return fail(MDGPU_ERR_ARG, "autotune plan: " + err);
    if (P.gemm.size() != N->plan.gemm.size() || P.steps.size() != N->plan.steps.size())
        return fail(MDGPU_ERR_ARG, "autotune plan: schedule shape changed");
    hipError_t e = hipMemcpy(N->dGemm, P.gemm.data(), P.gemm.size() * sizeof(mdgpu::GemmJob), hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(e, "autotune upload");
    N->plan.gemm.swap(P.gemm);
    N->plan.steps.swap(P.steps);
    N->po.gemm_tiles = tiles;
    if (N->g_factor) (void)hipGraphExecDestroy(N->g_factor);
    N->g_factor = nullptr;
    return MDGPU_OK;
}
}  // namespace

extern "C" int mdgpu_numeric_autotune(mdgpu_numeric* num, const double* d_ax, void* stream, int64_t* changed) {
    mdgpu::clear_error();
    if (!num || !d_ax) return fail(MDGPU_ERR_ARG, "mdgpu_numeric_autotune: null pointer");
    // the bf16 variants keep their 128-row bf16 kernel (64-row launches
    // would run the FP32 kernel instead)
    if (num->precision == MDGPU_BF16 || num->precision == MDGPU_BF16X3) {
        if (changed) *changed = 0;
        return MDGPU_OK;
    }
    hipStream_t s = (hipStream_t)stream;
    CallOrder order(num, s);
    std::vector<int> ords;          // step index of each GEMM launch, in order
    for (size_t t = 0; t < num->plan.steps.size(); t++)
        if (num->plan.steps[t].kind == mdgpu::STEP_GEMM) ords.push_back((int)t);
    // candidate kernels: the 16x16x4 kernel's 64- and 128-row tiles, and in
    // FP32 the 32x32x2 kernel
    // (each launch keeps its kGemmSplit flag: that is a rounding choice; a
    // split launch on 128-row tiles runs the 32x32x2 kernel either way)
    std::vector<int> cands = {64, 128};
    if (num->precision == MDGPU_FP32) {
        cands.push_back(mdgpu::kGemmMfma32);
        cands.push_back(mdgpu::kGemmMfma32x64);
    }
    std::vector<int> split(ords.size());
    for (size_t g = 0; g < ords.size(); g++) split[g] = num->plan.steps[ords[g]].arg & mdgpu::kGemmSplit;
    std::vector<std::vector<float>> msk(cands.size());
    std::vector<float> ms;
    for (size_t c = 0; c < cands.size(); c++) {
        std::vector<int> tiles(ords.size());
        for (size_t g = 0; g < ords.size(); g++) tiles[g] = cands[c] | split[g];
        TRY(install_gemm_tiles(num, tiles));
        TRY(factor_timed(num, d_ax, s, ms));    // warm
        TRY(factor_timed(num, d_ax, s, msk[c]));
    }
    std::vector<int> best(ords.size());
    int64_t n64 = 0;
    for (size_t g = 0; g < ords.size(); g++) {
        const int t = ords[g] + 1;              // ms[0] is the assembly
        size_t b = 0;
        for (size_t c = 1; c < cands.size(); c++)
            if (msk[c][t] < msk[b][t]) b = c;
        best[g] = cands[b] | split[g];
        n64 += mdgpu::gemm_tile_rows(cands[b]) == 64;
    }
    TRY(install_gemm_tiles(num, best));
    TRY(factor_timed(num, d_ax, s, ms));        // leaves a valid factorization
    if (changed) *changed = n64;
    return MDGPU_OK;
}

// ------------------------------------------------------ partition tuning
extern "C" int mdgpu_tune_partition(const mdgpu_symbolic* const* syms, int nsym, const double* d_ax, int precision,
                                    int reps, int solves, void* stream, double* ms, int* best) {
    mdgpu::clear_error();
    if (!syms || nsym <= 0 || !d_ax || !ms || !best || reps < 1 || solves < 0)
        return fail(MDGPU_ERR_ARG, "mdgpu_tune_partition: invalid argument");
    for (int k = 0; k < nsym; k++)
        if (!syms[k] || syms[k]->S.n != syms[0]->S.n || syms[k]->S.nnz_in != syms[0]->S.nnz_in)
            return fail(MDGPU_ERR_ARG, "mdgpu_tune_partition: the analyses are not of one matrix");
    hipStream_t s = (hipStream_t)stream;
    const int n = syms[0]->S.n;
    double* dbx = nullptr;
    if (hipMalloc((void**)&dbx, 2 * sizeof(double) * (size_t)n) != hipSuccess)
        return fail(MDGPU_ERR_NOMEM, "mdgpu_tune_partition: hipMalloc failed");
    mdgpu::launch_fill<double>(dbx, n, 1.0, s);
    auto wall = [] {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch())
            .count();
    };
    int st = MDGPU_OK;
    *best = 0;
    for (int k = 0; k < nsym && st == MDGPU_OK; k++) {
        mdgpu_numeric* N = nullptr;
        st = mdgpu_numeric_create(syms[k], precision, &N);
        if (st == MDGPU_OK) st = mdgpu_numeric_autotune(N, d_ax, stream, nullptr);
        double tf = 0, ts = 0;
        if (st == MDGPU_OK) st = mdgpu_factorize(N, d_ax, stream);   // graph capture, warm
        for (int r = 0; r < reps && st == MDGPU_OK; r++) {
            const double t0 = wall();
            st = mdgpu_factorize(N, d_ax, stream);                    // ends with a stream sync
            tf += wall() - t0;
        }
        if (st == MDGPU_OK && solves > 0) {
            st = mdgpu_solve(N, dbx, dbx + n, stream);                // warm
            if (st == MDGPU_OK && hipStreamSynchronize(s) != hipSuccess)
                st = fail(MDGPU_ERR_DEVICE, "mdgpu_tune_partition: hipStreamSynchronize failed");
            const double t0 = wall();
            for (int r = 0; r < reps * solves && st == MDGPU_OK; r++) st = mdgpu_solve(N, dbx, dbx + n, stream);
            if (st == MDGPU_OK && hipStreamSynchronize(s) != hipSuccess)
                st = fail(MDGPU_ERR_DEVICE, "mdgpu_tune_partition: hipStreamSynchronize failed");
            ts = wall() - t0;
        }
        mdgpu_numeric_destroy(N);
        if (st != MDGPU_OK) break;
        ms[k] = (tf + ts) / reps;   // factorization + `solves` solves
        if (ms[k] < ms[*best]) *best = k;
    }
    (void)hipFree(dbx);
    return st;
}

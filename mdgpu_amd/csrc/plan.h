// Execution plan of the numeric multifrontal factorization on one GPU.
//
// Storage (all device-resident, sized once from the symbolic analysis):
//   * L store: front f keeps its pivot panel P_f, the n_f x c_f column-major
//     block (leading dimension n_f) of rows q_f and pivot columns k..k+c_f-1.
//     After factorization P_f holds L(q_f, k:k+c_f) (upper part of its
//     diagonal block unused).  This replaces the per-front host/device dense
//     squares of the reference (Matriz, mdgpu.h:7-15) for the L part.
//   * update arena: front f's Schur complement U_f (m_f x m_f, m_f = n_f -
//     c_f, leading dimension m_f, lower triangle) lives from f's level until
//     its parent's level; offsets come from a static first-fit plan so
//     fronts of different subtrees reuse the same memory.
//   * relative maps rel_f: position of each update row of f inside its
//     parent's row list (the reference's `ql`, VectorMergeIndices,
//     mdgpu.cu:585-595), computed once here instead of per factorization.
//
// Per level (fronts whose children are all in lower levels):
//   1. extend-add pass 1: child update columns that land in the parent's
//      pivot columns go into the parent's panel (ExtendAdd, mdgpu.cu:748-767);
//   2. panel factorization (potrf + trsm of FactorAux_CPU_3, :1298-1338);
//   3. U_f = -L21 L21^T (the gemm of :1336, lower triangle only);
//   4. extend-add pass 2: remaining child update columns go into U_f.
// Siblings that update the same parent are split into rounds so no two
// writers in one launch touch the same entry (deterministic, no atomics).
#pragma once

#include <algorithm>
#include <cstdint>
#include <string>
#include <vector>

#include "symbolic.h"

namespace mdgpu {

struct FrontDesc {            // mirrored on the device
    int64_t lofs;             // panel offset in the L store (elements)
    int64_t uofs;             // U_f offset in the update arena (-1: none)
    int64_t relofs;           // offset of rel_f in the map buffer
    int64_t qofs;             // offset of q_f in the row-list buffer
    int k, c, n, m;           // first column, pivots, rows, update rows
    int parent;               // parent front or -1
    int kid_begin, kid_end;   // range in the kids buffer
    int split;                // # of update columns that land in parent pivots
    int level;
    int64_t dinv;             // offset of the diagonal-block inverses (see Plan::dinv)
    int64_t dinv256;          // offset of the 256-column block inverses (-1: none; see Plan::inv256)
};

// One extend-add work item: columns [j0, j1) of child `child`'s update.
// Extend-add, parent side ("pull"): one item per parent column that
// receives child columns in a pass, with its contributors (child columns)
// in child order.  A wavefront owns the parent column and adds the
// contributors one after the other, so all of a level's children go in
// one launch per pass and each parent entry sums its children in a fixed
// order.
struct EaItem {
    int64_t dst;                   // offset of the parent column (L store or, in_u, the arena)
    int in_u;                      // the column is in the parent's update matrix
    int rowoff;                    // parent row of the column's first entry (0 or c_parent)
    int cbeg, cend;                // contributors [cbeg, cend) in Plan::ea_contrib
};
struct EaContrib {
    int64_t src;                   // child column in the arena: uofs + j * m
    int64_t rel;                   // child's relative map (Plan::rel offset)
    int j, m;                      // child column, child update rows (rows j..m-1 are added)
};

// One dense job: C -= A * B^T (beta0: C = -A * B^T) on a tile grid.
struct GemmJob {
    int64_t a_off, b_off, c_off;   // A, B in the L store; C in L store or arena
    // launches of kernel code kGemmPlanes: A and B read from the bf16 split
    // planes (see Plan::planes) at these offsets -- element (i, k) of A at
    // a_pl + (k / 8) * 8 * lda + 8 * i + k % 8 -- instead of from the L store
    int64_t a_pl = -1, b_pl = -1;
    int lda, ldb, ldc;
    int M, N, K;
    int c_in_arena;                // C lives in the update arena
    int beta0;
    int lower;                     // keep only entries with row >= col
    int tiles_m, tiles_n;
    int tile_base;                 // first global tile of this job
    int ntiles;
};

// Block-column job of the blocked panel path, left-looking inside the
// outer block: the b columns of the block and the rows from its diagonal
// down first get the update of the kprev 64-column blocks of the same outer
// block left of it (A -= L_prev L_prev^T), then the b x b diagonal block at
// p_off is factored (Cholesky) and the `rows` rows below it solved (L21 <-
// A21 L_D^{-T}), in ntiles workgroups of kBlkRows rows.  Every workgroup
// updates and factors the diagonal block itself; the last one to have read
// it (counter `cnt`) writes it back.
struct BlkJob {
    int64_t p_off;                 // panel offset of the diagonal block
    int ld, b, col0, rows;         // panel ld, block size, global column, rows below
    int tile_base, ntiles;         // first workgroup of this job, workgroups (>= 1)
    int cnt, front;                // arrival counter index, front
    int kprev;                     // earlier 64-column blocks of the outer block (left-looking update)
    // split planes (presplit, else -1): plane offset of octet col0 / 8 at
    // the first row below the diagonal block; the step writes the octets
    // of its columns there (row r below the block: + 8 r, octet q: + 8 q ld)
    int64_t pl = -1;
};
constexpr int kBlkRows = 64;
// Fronts factored wholly in LDS by one workgroup (STEP_PANEL) have at most
// this many pivots (the kernel keeps their assembled diagonal for the
// pivot-growth check).  In reduced precision a front that fails the check
// is put on the step's fallback list (its panel left as assembled) and
// factored again in FP64 by a second launch right after the step, with
// twice the LDS.
constexpr int kPanelMaxPivots = 128;

// bf16 split planes (MDGPU_BF16 / MDGPU_BF16X3 with PlanOptions::presplit):
// two bf16 arrays of Plan::planes entries, hi = bf16(x) and lo =
// bf16(x - hi) of the factor's entries, in the "octet" layout of each panel:
// entry (i, k) of front f (rows n) at pofs_f + (k / 8) * 8 n + 8 i + k % 8
// (pofs_f: one n x ceil8(c) region per front), so that 8 consecutive k of one
// row (one MFMA operand) are 16 contiguous bytes.  The block steps write
// them (BlkJob::pl) for the rows below their diagonal block.

// A launch of the schedule: kind + a slice of the matching job array.
enum StepKind { STEP_EA1 = 0, STEP_EA2, STEP_PANEL, STEP_BLK, STEP_GEMM, STEP_INV, STEP_INV256 };

// Diagonal-block inverse job (fronts factored by the LDS panel kernel):
// block of 64 pivot columns starting at j0 of front `front`.
struct InvJob {
    int front;
    int j0;
};

struct Step {
    int kind;
    int first, count;      // slice of the job array of this kind
    int nblocks;           // workgroups to launch
    int level;
    int arg;               // panel: LDS bytes needed
    double flops = 0;      // algorithmic flops of the launch
    double bytes = 0;      // algorithmic HBM bytes of the launch
    // look-ahead: stream 0 = main, 1 = side (trailing updates of pivot
    // columns), 2 = side (update matrices); before the launch the stream
    // waits for events wait_ev and wait_ev2, after it records rec_ev (-1:
    // none)
    int stream = 0, wait_ev = -1, rec_ev = -1, wait_ev2 = -1;
};

// Solve schedule.  Small fronts (solve_small_front) are solved by one
// workgroup each (FWD_SMALL / BWD_SMALL).  The other fronts are cut into
// kChainBW-column blocks and solved by one launch per level whose
// workgroups form a chain inside the launch:
//  * FWD_CHAIN: one workgroup per block (TRI item: y_b = X_b (v_b -
//    sum_{j<b} L_bj y_j), X_b the block's inverse) and one per kChainBW
//    update rows below the pivots (RECT item: u -= sum_j L_rj y_j).  A
//    workgroup reads y_j from Y as soon as it is there: Y holds the
//    kChainEmpty pattern until the TRI item of block j stores y_j.
//  * BWD_UPD (before the chain): each workgroup takes kSolveCols pivot
//    columns of a front and kSolveUpdRows of its update rows (solved by the
//    ancestors, read through the row list q) and writes the dot products to
//    its partial-sum slot; BWD_UPD_SUM then subtracts the slots of each
//    column chunk from Y, in slot order.
//  * BWD_CHAIN: one workgroup per block, last block first (x_b = X_b^T
//    (z_b - sum_{j>b} L_jb^T x_j)), reading x_j from xp as it arrives (xp
//    is reset to kChainEmpty between the two sweeps).
// Workgroups take their items in launch order from a ticket counter (one
// per chain launch, zeroed at the start of every solve), so an item only
// waits for items of earlier tickets, which are running or done: the chain
// cannot deadlock however many workgroups are resident.
// fronts with at most kSolveSmall pivots and kSolveSmallEntries panel
// entries are solved by one workgroup each (SV_*_SMALL); the others in the
// chained launches, whose rows spread over many workgroups
constexpr int kSolveSmall = 64;
constexpr int64_t kSolveSmallEntries = 32768;
inline bool solve_small_front(int c, int n) { return c <= kSolveSmall && (int64_t)n * c <= kSolveSmallEntries; }
constexpr int kSolveThreads = 512;
constexpr int kChainBW = 128;
constexpr int kSolveBW = 256;       // column width of the block inverses (Plan::inv256)
constexpr int kSolveCols = 64;      // pivot columns per BWD_UPD workgroup
constexpr int kSolveUpdRows = 1024; // update rows per BWD_UPD workgroup
enum SolveKind { SV_FWD_SMALL = 0, SV_FWD_GATHER, SV_FWD_CHAIN, SV_BWD_SMALL, SV_BWD_UPD, SV_BWD_UPD_SUM, SV_BWD_CHAIN };

struct SolveItem {
    int front;
    int a;        // *_CHAIN: first row (FWD) / pivot column (BWD) of the item;
                  // BWD_UPD*: first pivot column of the kSolveCols-column chunk
    int first;    // *_CHAIN: 1 for a block of pivots (TRI), 0 for update rows (RECT)
    int j0;       // BWD_UPD: first update row
    int slot;     // BWD_UPD: partial-sum slot; BWD_UPD_SUM: first slot of the chunk
    int nslot;    // BWD_UPD_SUM: slots of the chunk
    // copies of the front's fields, so that a workgroup reads one item and
    // no front descriptor before its data loads
    int n, c, k;
    int64_t qofs;  // offset of the front's row list
    int64_t lofs;  // panel offset in the L store
    int64_t pk;    // TRI: offset of the block's inverse in Plan::inv256
    int64_t su;    // offset of the front's update vector (solve_uofs)
};

struct SolveStep {
    int kind;
    int first, count;   // slice of sv_items (or of sv_fronts for *_SMALL / GATHER)
    double bytes;       // algorithmic bytes moved
    int ctr = -1;       // *_CHAIN: ticket counter
};

struct Plan {
    int nfronts = 0, nlevels = 0;
    std::vector<FrontDesc> fronts;
    std::vector<int> kids;             // children of each front (processing order)
    std::vector<int> rel;              // relative maps, flat
    std::vector<int> q;                // row lists, flat (final numbering)
    std::vector<int64_t> a_dst;        // caller entry -> L-store offset
    std::vector<std::vector<int>> level_fronts;
    int64_t lstore = 0;                // elements of the L store
    int64_t arena = 0;                 // elements of the update arena (peak)
    int64_t solve_u = 0;               // elements of the solve update vectors
    std::vector<int64_t> solve_uofs;   // per front offset (m_f entries)
    std::vector<int> sv_fronts;        // front lists of SMALL / GATHER steps
    std::vector<SolveItem> sv_items;
    std::vector<SolveStep> sv_fwd, sv_bwd;
    int64_t sv_upd_partial = 1;        // doubles of BWD_UPD partial sums (one level at a time)
    int sv_nctr = 1;                   // ticket counters of the chain launches
    std::vector<EaItem> ea;
    std::vector<EaContrib> ea_contrib;
    std::vector<GemmJob> gemm;
    int64_t planes = 0;                // entries of each bf16 split plane (PlanOptions::presplit)
    std::vector<BlkJob> blk;
    std::vector<int> panel;            // fronts factored by the LDS panel kernel
    std::vector<InvJob> inv;           // diagonal-block inverses of every front
    // Inverses of the 64x64 diagonal blocks of every front, block t of front
    // f at fronts[f].dinv + 4096 t, stored as W(i + 64 r) = inv(L_D)(r, i)
    // for i, r < b (b = block width, last block may be narrower).
    int64_t dinv = 0;
    // Inverses of the kSolveBW-column diagonal blocks of fronts solved by the
    // chained solves (c > kSolveSmall): block t of front f at
    // fronts[f].dinv256 + 65536 t, column-major with leading dimension 256,
    // X(r, i) = inv(L_B)(r, i) for r, i < bw, zero elsewhere.  Formed from
    // the 64x64 inverses after them (STEP_INV256).  Its two diagonal
    // 128 x 128 blocks are the inverses of the kChainBW-column blocks.
    int64_t inv256 = 1;
    std::vector<InvJob> inv256_jobs;   // (front, first pivot column of the block)
    std::vector<Step> steps;           // the factorization schedule
    int nevents = 0;                   // events used by the look-ahead
    // full symmetric A (original numbering) for the residual: CSR + source
    std::vector<int> afull_p, afull_i, afull_src;
    int max_front = 0;
    int max_kids = 0;
};

// GEMM kernel code of a launch (Step.arg of a STEP_GEMM): a base kernel --
// 64 or 128, the 16x16x4 MFMA kernel with that tile size, or kGemmMfma32 /
// kGemmMfma32x64, the FP32 32x32x2 MFMA kernel with 128- / 64-row tiles --
// plus the kGemmSplit flag (FP32): two-level accumulation over runs of 64 k
constexpr int kGemmMfma32 = 129;
constexpr int kGemmMfma32x64 = 65;   // the 32x32x2 kernel with 64-row tiles
constexpr int kGemmSplit = 256;
// bf16 variants: operands from the split planes (128-row tiles); the other
// bf16 launches convert FP32 tiles as they stage them
constexpr int kGemmPlanes = 512;
// ... with six products from three planes (hi, mid, lo): FP32-level products
constexpr int kGemmSix = 1024;
inline int gemm_kernel_base(int kernel) { return kernel & ~(kGemmSplit | kGemmPlanes | kGemmSix); }
inline bool gemm_split(int kernel) { return (kernel & kGemmSplit) != 0; }
inline bool gemm_planes(int kernel) { return (kernel & kGemmPlanes) != 0; }
inline bool gemm_six(int kernel) { return (kernel & kGemmSix) != 0; }
inline int gemm_tile_rows(int kernel) {
    const int b = gemm_kernel_base(kernel);
    return b == 64 || b == kGemmMfma32x64 ? 64 : 128;
}
inline bool gemm_kernel_valid(int kernel, int elem_bytes) {
    const int b = gemm_kernel_base(kernel);
    if (gemm_split(kernel) && elem_bytes != 4) return false;
    if (gemm_planes(kernel) && (elem_bytes != 4 || b != 128 || gemm_split(kernel))) return false;
    if (gemm_six(kernel) && !gemm_planes(kernel)) return false;
    return b == 64 || b == 128 || ((b == kGemmMfma32 || b == kGemmMfma32x64) && elem_bytes == 4);
}

struct PlanOptions {
    int elem_bytes = 8;          // 8: FP64 store, 4: FP32 store
    int panel_lds_bytes = 65536; // LDS budget of the one-workgroup panel kernel
    // most pivots of a front factored wholly in LDS: the panel kernel's
    // column-by-column loop beats the blocked path only on narrow fronts
    // (measured with 16, 32, 64 and 128 as the bound: 16 is fastest on
    // every workload that has such fronts)
    int panel_max_pivots = 16;
    int nb = 64;                 // block width of the blocked panel path
    int nb_outer = 256;          // outer block width (trailing updates with K = nb_outer)
    int tile = 64;               // largest GEMM tile (rows = cols)
    int gemm_default = 64;       // GEMM kernel code of launches the autotuner did not set
    // Schur-complement launches whose largest K is at least split_k_min get
    // the kGemmSplit flag by default (0: never)
    int split_k_min = 0;
    // per-launch GEMM kernel codes (64, 128 or kGemmMfma32; tiles <= tile)
    // chosen by the autotuner, indexed by the GEMM launch's ordinal in the
    // schedule; empty: gemm_default for every launch
    std::vector<int> gemm_tiles;
    int ea_chunk = 16384;        // target entries per extend-add workgroup
    // right-looking update-matrix look-ahead: a blocked front's U_f receives
    // each outer block's product as soon as the block is factored, on a
    // second side stream, instead of in one Schur launch after the front
    bool u_lookahead = true;
    // ... applied every u_stride outer blocks (K = u_stride * nb_outer per
    // launch): a deeper K reads and writes U_f fewer times
    int u_stride = 2;
    // ... and at each of the u_flush outer blocks before a front's last one
    // (0: none), so that the launch after the next-to-last block, which the
    // Schur launch has to wait for, holds one block's columns
    int u_flush = 1;
    // bf16 variants: the block steps also write their columns as bf16 split
    // planes, and the GEMMs of the blocked fronts read them (kGemmPlanes)
    bool presplit = false;
    // split planes written by the block steps: 2 (hi, lo) or 3 (hi, mid, lo;
    // the update-matrix launches then make six products, kGemmSix)
    int nplanes = 2;
    // GEMM kernel code of the update-matrix launches ((u) and the Schur
    // launch) when nonzero, else the same kernel as the other GEMMs
    int schur_kernel = 0;
};

std::string build_plan(const Symbolic& S, const PlanOptions& o, Plan& P);

}  // namespace mdgpu

namespace mdgpu {
// Host-side check that every job of the plan stays inside its buffers and
// that update matrices live at the same time never overlap.  Returns "" if
// the plan is sound.
std::string validate_plan(const Plan& P, const PlanOptions& o);
// Plan options used for a given MDGPU_FP* precision.
PlanOptions plan_options_for(int precision);
}  // namespace mdgpu

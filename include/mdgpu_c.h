/*
 * mdgpu_c.h -- C ABI of the MI355X multifrontal sparse Cholesky.
 *
 * Plain pointers and sizes only.  Every function returns an int status
 * (MDGPU_OK == 0) and never exits the process; mdgpu_last_error() returns a
 * thread-local message for the most recent failure.  Device pointers are
 * HIP device addresses (hipMalloc or a framework tensor); host pointers are
 * ordinary memory.  `stream` is a hipStream_t passed as void* (NULL = the
 * null stream).  Each entry point names the reference interface it replaces.
 *
 * Calls on one numeric handle are ordered on the GPU: each call's stream
 * waits for the work of the previous call on that handle (an internal
 * event), so a handle can be used from several streams, one host thread at
 * a time; its calls then do not overlap on the GPU.  Host copies out of the
 * handle (mdgpu_export_L, mdgpu_selinv_export) wait for its GPU work.
 *
 * Life cycle:
 *   mdgpu_analyze   -> symbolic handle (host: ordering, trees, front maps)
 *   mdgpu_numeric_create(sym, precision) -> numeric handle (device arena)
 *   mdgpu_factorize(num, values)  (repeatable: refactorization, fixed pattern)
 *   mdgpu_solve / mdgpu_refine / mdgpu_logdet / mdgpu_export_L
 *   mdgpu_numeric_destroy, mdgpu_symbolic_destroy
 */
#ifndef MDGPU_C_H
#define MDGPU_C_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MDGPU_OK 0
#define MDGPU_ERR_ARG 1          /* invalid argument or matrix            */
#define MDGPU_ERR_NOT_SPD 2      /* non-positive pivot (column reported)  */
#define MDGPU_ERR_DEVICE 3       /* HIP runtime error                     */
#define MDGPU_ERR_NOMEM 4        /* host or device allocation failed      */
#define MDGPU_ERR_NOT_CONVERGED 5 /* refinement did not reach tolerance   */

/* Precision of the stored factor and of the dense front kernels. */
#define MDGPU_FP64 0
#define MDGPU_FP32 1
#define MDGPU_BF16 2   /* FP32 factor, Schur-update GEMMs on bf16 MFMA (one product) */
#define MDGPU_BF16X3 3 /* FP32 factor, Schur-update GEMMs as split bf16 products
                          hi*hi + hi*lo + lo*hi on bf16 MFMA (~16-bit mantissa) */

/* Node partition: -sn / -rn c / plain flags of the reference driver
 * (mdgpu.cu:1708-1714). */
#define MDGPU_PART_PLAIN 0
#define MDGPU_PART_SUPERNODAL 1
#define MDGPU_PART_RELAXED 2

typedef struct mdgpu_symbolic mdgpu_symbolic;
typedef struct mdgpu_numeric mdgpu_numeric;

typedef struct {
    int use_amd;     /* -amd: cs_amd(1, .) fill-reducing ordering        */
    int partition;   /* MDGPU_PART_*                                     */
    int max_zeros;   /* -rn c: relaxed-amalgamation zero budget          */
} mdgpu_analyze_opts;

/* Most recent error message of the calling thread ("" if none). */
const char* mdgpu_last_error(void);

/* Symbolic analysis of the lower triangle (diagonal included, rows sorted)
 * of a symmetric matrix in CSC form.  Replaces the driver's analyze glue
 * mdgpu.cu:1745-1875 (cs_amd, cs_symperm, cs_etree, cs_post, cs_counts,
 * GetETree / GetSuperETree / GetRelaxedETree, spL->p) plus the per-node row
 * merges of MatrizGetFk_3 / VectorMerge (mdgpu.cu:180-291, 535-595). */
int mdgpu_analyze(int n, const int32_t* Cp, const int32_t* Ci,
                  const mdgpu_analyze_opts* opts, mdgpu_symbolic** sym);
int mdgpu_symbolic_destroy(mdgpu_symbolic* sym);

/* The relaxed-amalgamation analyses of one matrix for nbudgets zero budgets
 * (the -rn c flag; the greedy cost of etree.cpp:236-237), sharing the
 * ordering, elimination tree and column counts: syms[k] equals what
 * mdgpu_analyze with partition MDGPU_PART_RELAXED and max_zeros =
 * budgets[k] returns.  The candidates of mdgpu_tune_partition. */
int mdgpu_analyze_budgets(int n, const int32_t* Cp, const int32_t* Ci, int use_amd, int nbudgets,
                          const int32_t* budgets, mdgpu_symbolic** syms);

/* Batch of independent matrices analysed as one: each matrix k (order n[k],
 * lower triangle Cp[k]/Ci[k]) gets exactly the analysis mdgpu_analyze
 * would give it (run on `nthreads` host threads, 0 = all cores), and the
 * results are joined block-diagonally into one forest, so one numeric
 * handle factors and solves all of them in one level-batched launch
 * sequence (one hipGraph).  Values, right-hand sides and solutions are the
 * concatenation of the matrices' own (matrix k's columns start at
 * MDGPU_ARR_SEG_COL[k], its values at MDGPU_ARR_SEG_NNZ[k]); mdgpu_refine
 * stops when every matrix reaches the tolerance.  No counterpart in the
 * reference (its driver takes one matrix, mdgpu.cu:1666-1939). */
int mdgpu_analyze_batch(int nmat, const int32_t* n, const int32_t* const* Cp, const int32_t* const* Ci,
                        const mdgpu_analyze_opts* opts, int nthreads, mdgpu_symbolic** sym);

/* Analysis for a caller-supplied ordering and partition: Cp/Ci is the
 * lower triangle ALREADY permuted, and the nodes (first column, column
 * count, children as node indices in kid_ptr/kids CSR form) are in
 * postorder, children before parents -- exactly what the reference passes
 * to Factor(A, spL, listaNodos, cantNodos) (mdgpu.cu:1394, 1889).  Children
 * may be listed in any order.  L's column counts are derived from the front
 * row lists. */
int mdgpu_analyze_nodes(int n, const int32_t* Cp, const int32_t* Ci, int nnodes, const int32_t* first,
                        const int32_t* ncols, const int32_t* kid_ptr, const int32_t* kids,
                        mdgpu_symbolic** sym);

/* Write an analysis to a file / read it back (a versioned binary format
 * with a checksum; the load checks the structure before trusting it), so a
 * fixed pattern can be refactorized later without analysing it again.  No
 * counterpart in the reference. */
int mdgpu_symbolic_save(const mdgpu_symbolic* sym, const char* path);
int mdgpu_symbolic_load(const char* path, mdgpu_symbolic** sym);

/* Scalar facts about an analysis. */
#define MDGPU_SYM_N 0             /* matrix order                          */
#define MDGPU_SYM_NNODES 1        /* nodes (fronts) in the partition       */
#define MDGPU_SYM_LNZ 2           /* entries of L (padded counts)          */
#define MDGPU_SYM_NNZ_A 3         /* entries of the input lower triangle   */
#define MDGPU_SYM_QLEN 4          /* total length of all front row lists   */
#define MDGPU_SYM_MAXFRONT 5      /* largest front order                   */
#define MDGPU_SYM_NLEVELS 6       /* levels of the front tree              */
#define MDGPU_SYM_NODES_FLAT 7    /* length of the MDGPU_ARR_NODES array   */
#define MDGPU_SYM_NSEG 8          /* independent matrices (1, or the batch) */
#define MDGPU_SYM_NSIZES 9
int mdgpu_symbolic_sizes(const mdgpu_symbolic* sym, int64_t* out /* [MDGPU_SYM_NSIZES] */);

/* Flop counts: [0] = sum_j c_j^2 over the unpadded cs_counts column counts,
 * [1] = the reference's own sum_j (c_j^2 + 1) over padded counts
 * (mdgpu.cu:1354-1359). */
int mdgpu_symbolic_flops(const mdgpu_symbolic* sym, double* out /* [2] */);

/* Integer arrays of the analysis, copied into dst (capacity cap entries). */
#define MDGPU_ARR_AMD_PERM 0      /* n: cs_amd permutation                 */
#define MDGPU_ARR_PARENT 1        /* n: cs_etree of the AMD-ordered matrix */
#define MDGPU_ARR_POST 2          /* n: cs_post                            */
#define MDGPU_ARR_COUNT 3         /* n: cs_counts                          */
#define MDGPU_ARR_RELAXED_PERM 4  /* n: GetRelaxedETree perm (relaxed only)*/
#define MDGPU_ARR_PERM 5          /* n: final permutation (new k = old perm[k]) */
#define MDGPU_ARR_COUNT_FINAL 6   /* n: padded counts in final order       */
#define MDGPU_ARR_LP 7            /* n+1: spL->p                           */
#define MDGPU_ARR_ALOW_P 8        /* n+1: final permuted lower triangle    */
#define MDGPU_ARR_ALOW_I 9        /* nnz_a                                 */
#define MDGPU_ARR_NODES 10        /* flat: per node first, ncols, rows,
                                     zeros, nkids, kid node indices...     */
#define MDGPU_ARR_FRONT_PARENT 11 /* nnodes: parent node index or -1      */
#define MDGPU_ARR_Q_OFF 12        /* nnodes+1 (int32; fails past 2^31)     */
#define MDGPU_ARR_Q 13            /* qlen: sorted row ids of each front    */
#define MDGPU_ARR_LEVEL 14        /* nnodes: height of each front          */
#define MDGPU_ARR_SEG_COL 15      /* nseg+1: first column of each matrix   */
#define MDGPU_ARR_SEG_NNZ 16      /* nseg+1: first caller entry of each matrix */
int mdgpu_symbolic_array(const mdgpu_symbolic* sym, int which, int32_t* dst, int64_t cap);


/* Sorted set union of two sorted row lists (VectorMerge, mdgpu.cu:535-583;
 * out has room for na + nb rows, *nout = rows of the union) and the
 * relative map of a sorted sublist inside a sorted list (VectorMergeIndices,
 * mdgpu.cu:585-595): map[t] = position of sub[t] in super.  These are the
 * operations the analysis builds the front row lists and the extend-add
 * maps with. */
int mdgpu_row_union(const int32_t* a, int na, const int32_t* b, int nb, int32_t* out, int* nout);
int mdgpu_row_map(const int32_t* super, int n, const int32_t* sub, int m, int32_t* map);

/* ------------------------------------------------------------ numeric */

/* Allocate the device arena for a numeric factorization of `sym` (L store,
 * update-matrix arena, maps and schedule uploaded once).  Replaces the
 * per-node device allocation and map uploads of the reference
 * (AllocFrenteGPU / MoverFrenteQlAGPU, mdgpu.cu:597-723).  precision:
 * MDGPU_FP64 or MDGPU_FP32 (storage and arithmetic of the dense fronts), or
 * MDGPU_BF16 / MDGPU_BF16X3 (FP32 storage; the GEMM/SYRK updates on bf16
 * MFMA with FP32 accumulation; use mdgpu_refine for an FP64-accurate x). */
int mdgpu_numeric_create(const mdgpu_symbolic* sym, int precision, mdgpu_numeric** num);
int mdgpu_numeric_destroy(mdgpu_numeric* num);

/* Replay factorization and solves from hipGraphs captured on first use
 * (enable = 1, the default) or launch every kernel eagerly (enable = 0).
 * The graphs bind internal buffers, so callers' pointers may change between
 * calls.  No counterpart in the reference. */
int mdgpu_numeric_set_graph(mdgpu_numeric* num, int enable);

#define MDGPU_NUM_DEVICE_BYTES 0  /* device memory held by the handle     */
#define MDGPU_NUM_LSTORE 1        /* elements of the L store               */
#define MDGPU_NUM_ARENA 2         /* elements of the update arena (peak)   */
#define MDGPU_NUM_STEPS 3         /* kernel launches of one factorization  */
#define MDGPU_NUM_LEVELS 4        /* levels of the schedule                */
#define MDGPU_NUM_PANEL_FRONTS 5  /* fronts factored wholly in LDS         */
#define MDGPU_NUM_DIAG_JOBS 6     /* diagonal blocks of the blocked path   */
#define MDGPU_NUM_PRECISION 7
#define MDGPU_NUM_FALLBACKS 8     /* FP32/BF16: diagonal blocks and fronts
                                     factored wholly in LDS of the last
                                     factorization refactored in FP64 after
                                     failing the pivot-growth check
                                     (L_jj^2 < 2^-12 A_jj or L_jj^2 <= 0)  */
#define MDGPU_NUM_NINFO 9
int mdgpu_numeric_info(const mdgpu_numeric* num, int64_t* out /* [MDGPU_NUM_NINFO] */);

/* Host only (no device calls): build the schedule mdgpu_numeric_create
 * would build for `precision`, check every kernel job against its buffers,
 * and report the same figures as mdgpu_numeric_info (DEVICE_BYTES is the
 * planned device footprint).  No counterpart in the reference. */
int mdgpu_plan_check(const mdgpu_symbolic* sym, int precision, int64_t* out /* [MDGPU_NUM_NINFO] */);

/* Numeric factorization.  d_ax: DEVICE pointer to the values of the
 * caller's lower triangle, in the order of the Ci array given to
 * mdgpu_analyze (FP64).  Replaces Factor(A, spL, listaNodos, cantNodos),
 * mdgpu.cu:1394-1608 (front assembly, extend-add, partial factorization of
 * every front).  Returns MDGPU_ERR_NOT_SPD on a non-positive pivot.
 * Synchronizes `stream` at the end (pivot check). */
int mdgpu_factorize(mdgpu_numeric* num, const double* d_ax, void* stream);

/* mdgpu_factorize with per-kernel-class timing: HIP events are recorded on
 * `stream` around every launch.  out[4*k + {0,1,2,3}] = milliseconds,
 * launches, algorithmic flops, algorithmic HBM bytes of kernel class k:
 * 0 assembly (zero fill + A scatter), 1 extend-add, 2 LDS panel
 * factorization, 3 blocked diagonal potrf + TRSM, 4 diagonal-block
 * inverses (used by the solves), 5 tiled GEMM/SYRK.  The event
 * gaps add launch latency, so use mdgpu_factorize for end-to-end times. */
#define MDGPU_KIND_N 6
int mdgpu_factorize_profile(mdgpu_numeric* num, const double* d_ax, void* stream,
                            double* out /* [4*MDGPU_KIND_N] */);

/* Autotuner: factorizes d_ax with every GEMM launch of the schedule on 64-
 * and on 128-row 16x16x4 MFMA tiles and, in MDGPU_FP32, on the 32x32x2
 * MFMA kernel's 128- and 64-row tiles (HIP-event time per launch), keeps the
 * fastest kernel for each
 * launch and installs that schedule (captured graphs are recaptured on next
 * use).  *changed = launches on 64-row tiles.  Ends with a valid
 * factorization of d_ax.  The FP32 kernels round differently, so an FP32
 * factor can change in its last bits with the choice.  No counterpart in the
 * reference. */
int mdgpu_numeric_autotune(mdgpu_numeric* num, const double* d_ax, void* stream, int64_t* changed);

/* Partition tuner (relaxed-amalgamation zero budget, etree.cpp:236-237):
 * for each analysis syms[k] of one matrix (e.g. mdgpu_analyze_budgets),
 * build a numeric handle in `precision`, autotune its GEMM tiles on d_ax,
 * and time `reps` factorizations plus reps * `solves` solves (wall clock,
 * synchronized); ms[k] = time of one factorization + `solves` solves, *best
 * = the fastest k.  One handle exists at a time.  Allocates (a setup
 * routine, like the analysis).  No counterpart in the reference. */
int mdgpu_tune_partition(const mdgpu_symbolic* const* syms, int nsym, const double* d_ax, int precision, int reps,
                         int solves, void* stream, double* ms, int* best);

/* mdgpu_factorize with one record per launch of the schedule:
 * out[8*t + {0..7}] = kernel class (as above), milliseconds, workgroups,
 * algorithmic flops, algorithmic bytes, level, stream of the look-ahead
 * schedule (0 main, 1 trailing updates of pivot columns, 2 update
 * matrices), GEMM kernel code (0 for other launches).  The launches run
 * one after the other on `stream` here.  cap >= 8 * MDGPU_NUM_STEPS.
 * No counterpart in the reference (its timers are per phase, mdgpu.cu:1905-1930). */
int mdgpu_factorize_steps(mdgpu_numeric* num, const double* d_ax, void* stream, double* out, int64_t cap);

/* x = A^{-1} b with the stored factor (device vectors, FP64, caller's
 * numbering): permutation, supernodal forward and backward substitution.
 * Replaces cs_pvec + cs_lsolve + cs_ltsolve + cs_ipvec (cs.cpp:1086-1146,
 * dead call site mdgpu.cu:1953-1960).  Asynchronous on `stream`. */
int mdgpu_solve(mdgpu_numeric* num, const double* d_b, double* d_x, void* stream);

/* nrhs right-hand sides at once: column r of B (d_B + r * ld, n entries)
 * gives column r of X, each as mdgpu_solve (the same launches, so the same
 * results).  ld >= n.  The solve entry of the interface contract,
 * mdgpu_solve(num, b, x, nrhs), over cs_lsolve / cs_ltsolve
 * (cs.cpp:1086-1119) applied per column.  Asynchronous on `stream`. */
int mdgpu_solve_nrhs(mdgpu_numeric* num, const double* d_B, double* d_X, int nrhs, int64_t ld, void* stream);

/* Iterative refinement in FP64 until the backward error
 * ||b - A x||_inf / (||A||_1 ||x||_inf + ||b||_inf) (mdgpu.cu:1972) is at
 * most tol, or maxit corrections were applied.  The residual uses the
 * values last passed to mdgpu_factorize.  Classical refinement for FP64,
 * FP32 and BF16X3 factors; conjugate gradients preconditioned by the factor
 * for MDGPU_BF16 (environment MDGPU_REFINE=cg or =ir overrides either).
 * Returns MDGPU_ERR_NOT_CONVERGED (with eta/iters filled) if tol was not
 * reached. */
int mdgpu_refine(mdgpu_numeric* num, const double* d_b, double* d_x, double tol, int maxit,
                 double* eta, int* iters, void* stream);

/* Backward error of x for the symmetric system A x = b, computed on the
 * GPU: eta = ||b - A x||_inf / (||A||_1 ||x||_inf + ||b||_inf) -- the check
 * of the reference driver (cs_gaxpy + cs_norm + norm, mdgpu.cu:1962-1972).
 * d_rowptr / d_colind / d_val: the FULL symmetric matrix (both triangles)
 * in CSR form on the device (for a symmetric matrix the same arrays are its
 * CSC form); d_x, d_b: device vectors.  Allocates a small device workspace
 * per call; synchronizes `stream`. */
int mdgpu_backward_error(int n, const int32_t* d_rowptr, const int32_t* d_colind, const double* d_val,
                         const double* d_x, const double* d_b, double* eta, void* stream);

/* mdgpu_backward_error without any allocation: no counterpart in the
 * reference; for callers that cannot allocate device memory (code under
 * hipGraph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The workspace is a static device array of the library (131 KB per
 * device, held until the process ends), so n is limited to 4,194,304 rows
 * (MDGPU_ERR_ARG beyond), and calls that overlap on one device race (each
 * call synchronizes its stream, so calls from one host thread do not). */
int mdgpu_backward_error_noalloc(int n, const int32_t* d_rowptr, const int32_t* d_colind, const double* d_val,
                                 const double* d_x, const double* d_b, double* eta, void* stream);

/* Status of the solves run since the last call: MDGPU_SOLVE_TIMED_OUT is set
 * when a wait inside a chained block solve gave up (about a second of
 * polling for a value that never came): the solution then holds NaN where
 * it depends on that value.  mdgpu_solve / mdgpu_solve_nrhs return before
 * the GPU finishes and cannot report it; mdgpu_refine checks it itself and
 * returns MDGPU_ERR_DEVICE.  Reading clears the bits.  Synchronizes
 * `stream`.  No counterpart in the reference. */
#define MDGPU_SOLVE_TIMED_OUT 1
int mdgpu_solve_status(mdgpu_numeric* num, int* status, void* stream);

/* Backward error of each independent matrix (MDGPU_SYM_NSEG of them) after
 * the last mdgpu_refine; for a batch, *eta of mdgpu_refine is their max.
 * No counterpart in the reference. */
int mdgpu_refine_segment_eta(const mdgpu_numeric* num, double* out, int cap);

/* ---------------------------------------------------------- reverse mode
 * None of these has a counterpart in the reference.  All need an
 * MDGPU_FP64 factorization (mdgpu_factorize done).
 *
 * Selected inverse Z = A^{-1} on the pattern of L (permuted numbering),
 * supernodal Takahashi recurrences top-down over the front tree, with the
 * children's blocks of Z passed down by a reverse extend-add (update
 * matrices are recomputed from the stored L, not kept).  Asynchronous. */
int mdgpu_selinv(mdgpu_numeric* num, void* stream);
/* Copy Z to the host in the layout of mdgpu_export_L's Lx (synchronous). */
int mdgpu_selinv_export(mdgpu_numeric* num, double* Zx);
/* d log det A / d a_e for every entry e of the caller's lower triangle
 * (device vector, caller entry order): Z_ii on the diagonal, 2 Z_ij off it
 * (A is symmetric: a_ij stands for both a_ij and a_ji).  Runs mdgpu_selinv
 * first when Z is not current. */
int mdgpu_logdet_grad(mdgpu_numeric* num, double* d_g, void* stream);
/* Adjoint of x = A^{-1} b: given x and xbar = dL/dx (device, caller
 * numbering), bbar = A^{-1} xbar (dL/db; d_bbar may be NULL) and
 * d_g[e] = dL/da_e = -(bbar_i x_j + bbar_j x_i) off the diagonal,
 * -bbar_i x_i on it. */
int mdgpu_solve_adjoint(mdgpu_numeric* num, const double* d_x, const double* d_xbar, double* d_bbar,
                        double* d_g, void* stream);

/* log det A = 2 sum_j log L_jj (no counterpart in the reference). */
int mdgpu_logdet(mdgpu_numeric* num, double* out, void* stream);

/* Copy L to host CSC in the reference's spL layout: column j holds the rows
 * of its front from j down, diagonal first, ascending (mdgpu.cu:1346-1352,
 * 1859-1873).  Lp has n+1 entries, Li/Lx have lnz; any may be NULL. */
int mdgpu_export_L(mdgpu_numeric* num, int32_t* Lp, int32_t* Li, double* Lx);

/* mdgpu_export_L for matrix k of a batch (mdgpu_analyze_batch): its own
 * columns, in its own numbering (rows and column pointers relative to
 * MDGPU_ARR_SEG_COL[k]); Lp has n_k + 1 entries.  Works when the whole
 * batch's L exceeds int32 entries.  No counterpart in the reference. */
int mdgpu_export_L_segment(mdgpu_numeric* num, int seg, int32_t* Lp, int32_t* Li, double* Lx);

#ifdef __cplusplus
}
#endif

#endif /* MDGPU_C_H */

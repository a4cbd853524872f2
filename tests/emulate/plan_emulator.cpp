// TEST INFRASTRUCTURE: executes the numeric plan's steps sequentially on the
// CPU with the same loops as the device kernels, so that the schedule
// (extend-add rounds, splits, arena offsets, blocked steps) can be checked
// against the oracle without a GPU.  Built and loaded by the tests only.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../mdgpu_amd/csrc/plan.h"
#include "../../mdgpu_amd/csrc/selinv.h"
#include "../../mdgpu_amd/csrc/symbolic.h"

using namespace mdgpu;

// W(i + 64 r) = inv(L)(r, i) for the lower-triangular b x b block g (ld)
static void invert_block(const double* g, int ld, int b, double* w) {
    std::vector<double> X((size_t)b * b, 0.0);   // X(r, i) at r + i*b
    for (int i = 0; i < b; i++) {
        X[i + i * b] = 1.0 / g[i + i * ld];
        for (int r = i + 1; r < b; r++) {
            double s = 0;
            for (int l = i; l < r; l++) s += g[r + l * ld] * X[l + i * b];
            X[r + i * b] = -s / g[r + r * ld];
        }
    }
    for (int i = 0; i < b; i++)
        for (int r = 0; r < b; r++) w[i + 64 * r] = X[r + i * b];
}

// Executes the selected-inverse schedule (selinv_plan.cpp) with the device
// kernels' loops, given the factor L, the diagonal-block inverses and the
// update arena buffer U; writes Z in the L store layout.
static int run_selinv(const Plan& P, const std::vector<double>& L, const std::vector<double>& inv,
                      std::vector<double>& U, std::vector<double>& Z) {
    SelPlan Q;
    if (!build_selinv_plan(P, Q).empty()) return 4;
    if (!validate_selinv_plan(P, Q).empty()) return 5;
    std::vector<double> ZF(Q.zf_size, std::nan("")), LH(Q.lh_size, std::nan(""));
    Z.assign(P.lstore, 0.0);
    std::vector<double>* bufs[SB_N] = {const_cast<std::vector<double>*>(&L), &Z, &ZF, &LH,
                                       const_cast<std::vector<double>*>(&inv), &U};
    for (const SelStep& st : Q.steps) {
        if (st.kind == SS_GEMM) {
            for (int t = st.first; t < st.first + st.count; t++) {
                const SelGemm& g = Q.gemm[t];
                const double* A = bufs[g.a_buf]->data() + g.a_off;
                const double* B = bufs[g.b_buf]->data() + g.b_off;
                double* C = bufs[g.c_buf]->data() + g.c_off;
                std::vector<double> out((size_t)g.M * g.N);
                for (int j = 0; j < g.N; j++)
                    for (int i = 0; i < g.M; i++) {
                        double acc = 0;
                        for (int k = 0; k < g.K; k++)
                            acc += (g.ta ? A[k + (int64_t)i * g.lda] : A[i + (int64_t)k * g.lda]) *
                                   (g.tb ? B[j + (int64_t)k * g.ldb] : B[k + (int64_t)j * g.ldb]);
                        out[i + (size_t)j * g.M] = acc;
                    }
                for (int j = 0; j < g.N; j++)
                    for (int i = 0; i < g.M; i++) {
                        double& c = C[i + (int64_t)j * g.ldc];
                        c = g.beta == 0.0 ? g.alpha * out[i + (size_t)j * g.M] : g.alpha * out[i + (size_t)j * g.M] + g.beta * c;
                    }
            }
            continue;
        }
        for (int t = st.first; t < st.first + st.count; t++) {
            const SelCopy& it = Q.copy[t];
            const FrontDesc& d = P.fronts[it.front];
            double* z = ZF.data() + Q.zf_ofs[it.front];
            if (st.copy_kind == SC_INIT) {
                for (int j = it.j0; j < it.j1; j++)
                    for (int i = 0; i < d.m; i++) z[(d.c + i) + (int64_t)(d.c + j) * d.n] = U[d.uofs + i + (int64_t)j * d.m];
            } else if (st.copy_kind == SC_TRANS) {
                for (int ii = it.j0; ii < it.j1; ii++)
                    for (int jj = 0; jj < it.b; jj++) {
                        const int i = it.b0 + it.b + ii;
                        z[(it.b0 + jj) + (int64_t)i * d.n] = z[i + (int64_t)(it.b0 + jj) * d.n];
                    }
            } else if (st.copy_kind == SC_OUT) {
                for (int j = it.j0; j < it.j1; j++)
                    for (int i = 0; i < d.n; i++) Z[d.lofs + i + (int64_t)j * d.n] = z[i + (int64_t)j * d.n];
            } else {
                const FrontDesc& pd = P.fronts[d.parent];
                const double* zp = ZF.data() + Q.zf_ofs[d.parent];
                const int* r = P.rel.data() + d.relofs;
                for (int j = it.j0; j < it.j1; j++)
                    for (int i = 0; i < d.m; i++) U[d.uofs + i + (int64_t)j * d.m] = zp[r[i] + (int64_t)r[j] * pd.n];
            }
        }
    }
    return 0;
}

extern "C" int emulate_factor_selinv(int n, const int* Cp, const int* Ci, const double* ax, int use_amd, int mode,
                                     int max_zeros, int nb_outer, double* Lx_out, int* Li_out, double* Zx_out);

extern "C" int emulate_factor(int n, const int* Cp, const int* Ci, const double* ax, int use_amd, int mode,
                              int max_zeros, int nb_outer, double* Lx_out, int* Li_out) {
    return emulate_factor_selinv(n, Cp, Ci, ax, use_amd, mode, max_zeros, nb_outer, Lx_out, Li_out, nullptr);
}

static int emulate(int n, const int* Cp, const int* Ci, const double* ax, int use_amd, int mode, int max_zeros,
                   int nb_outer, bool presplit, int u_stride, int u_flush, double* Lx_out, int* Li_out,
                   double* Zx_out);

extern "C" int emulate_factor_selinv(int n, const int* Cp, const int* Ci, const double* ax, int use_amd, int mode,
                                     int max_zeros, int nb_outer, double* Lx_out, int* Li_out, double* Zx_out) {
    return emulate(n, Cp, Ci, ax, use_amd, mode, max_zeros, nb_outer, false, PlanOptions().u_stride,
                   PlanOptions().u_flush, Lx_out, Li_out, Zx_out);
}

// update-matrix look-ahead every u_stride outer blocks and at the last
// u_flush blocks before a front's last, with (presplit) or without the
// split planes
extern "C" int emulate_factor_ustride(int n, const int* Cp, const int* Ci, const double* ax, int use_amd, int mode,
                                      int max_zeros, int nb_outer, int u_stride, int u_flush, int presplit,
                                      double* Lx_out, int* Li_out) {
    return emulate(n, Cp, Ci, ax, use_amd, mode, max_zeros, nb_outer, presplit != 0, u_stride, u_flush, Lx_out,
                   Li_out, nullptr);
}

// The schedule of the bf16 variants (block steps that also write the split
// planes, GEMMs that read them), run in FP64 with planes that hold the
// values unsplit: checks that every plane entry a GEMM reads was written
// before it (the planes start as NaN, and the padding past a front's last
// column is written as NaN).
extern "C" int emulate_factor_presplit(int n, const int* Cp, const int* Ci, const double* ax, int use_amd, int mode,
                                       int max_zeros, int nb_outer, double* Lx_out, int* Li_out) {
    return emulate(n, Cp, Ci, ax, use_amd, mode, max_zeros, nb_outer, true, PlanOptions().u_stride,
                   PlanOptions().u_flush, Lx_out, Li_out, nullptr);
}

static int emulate(int n, const int* Cp, const int* Ci, const double* ax, int use_amd, int mode, int max_zeros,
                   int nb_outer, bool presplit, int u_stride, int u_flush, double* Lx_out, int* Li_out,
                   double* Zx_out) {
    AnalyzeOptions ao;
    ao.use_amd = use_amd;
    ao.partition = (Partition)mode;
    ao.max_zeros = max_zeros;
    Symbolic S;
    if (!analyze(n, Cp, Ci, ao, S).empty()) return 1;
    PlanOptions po;
    po.elem_bytes = 8;
    po.panel_lds_bytes = 61440;
    po.nb = 64;
    // nb_outer < 0: outer blocks of -nb_outer columns without the
    // update-matrix look-ahead (one Schur launch per level)
    po.nb_outer = nb_outer < 0 ? -nb_outer : nb_outer;
    po.u_lookahead = nb_outer > 0;
    po.u_stride = u_stride;
    po.u_flush = u_flush;
    // outer width 128 also sends fronts of up to 128 pivots to the LDS panel
    // kernel (its upper bound; the default keeps it to the narrow fronts)
    if (nb_outer == 128) po.panel_max_pivots = kPanelMaxPivots;
    po.tile = 64;
    if (presplit) {   // the bf16 variants' options (FP32 store, 128-row tiles)
        po.elem_bytes = 4;
        po.tile = 128;
        po.gemm_default = 128;
        po.presplit = true;
    }
    Plan P;
    std::string err = build_plan(S, po, P);
    if (!err.empty()) {
        fprintf(stderr, "emulate: %s\n", err.c_str());
        return 2;
    }
    err = validate_plan(P, po);
    if (!err.empty()) {
        fprintf(stderr, "emulate: %s\n", err.c_str());
        return 3;
    }
    if (presplit && P.blk.empty()) return 4;
    std::vector<double> L(P.lstore, 0.0), U(std::max<int64_t>(P.arena, 1), std::nan("")),
        inv(P.dinv, 0.0), planes(std::max<int64_t>(P.planes, 1), std::nan(""));
    for (size_t e = 0; e < P.a_dst.size(); e++) L[P.a_dst[e]] = ax[e];
    for (const Step& st : P.steps) {
        if (st.kind == STEP_EA1 || st.kind == STEP_EA2) {
            for (int t = st.first; t < st.first + st.count; t++) {
                const EaItem& it = P.ea[t];
                double* dcol = it.in_u ? &U[it.dst] : &L[it.dst];
                for (int c = it.cbeg; c < it.cend; c++) {
                    const EaContrib& e = P.ea_contrib[c];
                    const int* r = P.rel.data() + e.rel;
                    for (int i = e.j; i < e.m; i++) dcol[r[i] - it.rowoff] += U[e.src + i];
                }
            }
        } else if (st.kind == STEP_PANEL) {
            for (int t = st.first; t < st.first + st.count; t++) {
                const FrontDesc& d = P.fronts[P.panel[t]];
                double* g = &L[d.lofs];
                for (int j = 0; j < d.c; j++) {
                    double rt = std::sqrt(g[j + j * d.n]);
                    g[j + j * d.n] = rt;
                    for (int i = j + 1; i < d.n; i++) g[i + j * d.n] /= rt;
                    for (int i = j + 1; i < d.n; i++)
                        for (int col = j + 1; col <= std::min(i, d.c - 1); col++)
                            g[i + col * d.n] -= g[i + j * d.n] * g[col + j * d.n];
                }
            }
        } else if (st.kind == STEP_BLK) {
            for (int t = st.first; t < st.first + st.count; t++) {
                const BlkJob& bj = P.blk[t];
                double* g = &L[bj.p_off];
                const int b = bj.b, ld = bj.ld;
                // left-looking: the block's columns, rows from the diagonal
                // down, minus the outer block's earlier block columns
                const int K = 64 * bj.kprev;
                for (int c = 0; c < b; c++)
                    for (int i = c; i < b + bj.rows; i++) {
                        double acc = 0;
                        for (int l = 0; l < K; l++) acc += g[i + (int64_t)(l - K) * ld] * g[c + (int64_t)(l - K) * ld];
                        g[i + (int64_t)c * ld] -= acc;
                    }
                for (int j = 0; j < b; j++) {
                    double rt = std::sqrt(g[j + j * ld]);
                    g[j + j * ld] = rt;
                    for (int i = j + 1; i < b; i++) g[i + j * ld] /= rt;
                    for (int i = j + 1; i < b; i++)
                        for (int col = j + 1; col <= i; col++) g[i + col * ld] -= g[i + j * ld] * g[col + j * ld];
                }
                for (int i = 0; i < bj.rows; i++) {
                    double* row = &g[b + i];
                    for (int j = 0; j < b; j++) {
                        double acc = row[(int64_t)j * ld];
                        for (int l = 0; l < j; l++) acc -= row[(int64_t)l * ld] * g[j + (int64_t)l * ld];
                        row[(int64_t)j * ld] = acc / g[j + (int64_t)j * ld];
                    }
                }
                // split planes of the rows below the block (octet layout)
                if (bj.pl >= 0)
                    for (int i = 0; i < bj.rows; i++)
                        for (int j = 0; j < (b + 7) / 8 * 8; j++)
                            planes[bj.pl + (int64_t)(j / 8) * 8 * ld + 8 * (int64_t)i + j % 8] =
                                j < b ? g[b + i + (int64_t)j * ld] : std::nan("");
            }
        } else if (st.kind == STEP_INV) {
            for (int t = st.first; t < st.first + st.count; t++) {
                const FrontDesc& d = P.fronts[P.inv[t].front];
                const int j0 = P.inv[t].j0;
                invert_block(&L[d.lofs + j0 + (int64_t)j0 * d.n], d.n, std::min(64, d.c - j0),
                             &inv[d.dinv + (int64_t)(j0 / 64) * 4096]);
            }
        } else if (st.kind == STEP_GEMM) {
            const bool pl = gemm_planes(st.arg);
            for (int t = st.first; t < st.first + st.count; t++) {
                const GemmJob& g = P.gemm[t];
                double* C = g.c_in_arena ? &U[g.c_off] : &L[g.c_off];
                auto a = [&](int i, int k) {
                    return pl ? planes[g.a_pl + (int64_t)(k / 8) * 8 * g.lda + 8 * (int64_t)i + k % 8]
                              : L[g.a_off + i + (int64_t)k * g.lda];
                };
                auto b = [&](int j, int k) {
                    return pl ? planes[g.b_pl + (int64_t)(k / 8) * 8 * g.ldb + 8 * (int64_t)j + k % 8]
                              : L[g.b_off + j + (int64_t)k * g.ldb];
                };
                for (int j = 0; j < g.N; j++)
                    for (int i = (g.lower ? j : 0); i < g.M; i++) {
                        double acc = 0;
                        for (int k = 0; k < g.K; k++) acc += a(i, k) * b(j, k);
                        double* cp = &C[i + (int64_t)j * g.ldc];
                        *cp = g.beta0 ? -acc : *cp - acc;
                    }
            }
        }
    }
    std::vector<double> Z;
    if (Zx_out) {
        const int st = run_selinv(P, L, inv, U, Z);
        if (st) return st;
    }
    for (int f = 0; f < P.nfronts; f++) {
        const FrontDesc& d = P.fronts[f];
        for (int j = 0; j < d.c; j++) {
            int64_t dst = S.lp[d.k + j];
            for (int i = j; i < d.n; i++, dst++) {
                Lx_out[dst] = L[d.lofs + (int64_t)j * d.n + i];
                Li_out[dst] = S.q_all[d.qofs + i];
                if (Zx_out) Zx_out[dst] = Z[d.lofs + (int64_t)j * d.n + i];
            }
        }
    }
    return 0;
}

// Plants one mistake in a valid plan and returns what validate_plan says
// about it (0 when the plan was rejected with a message containing `want`,
// 1 when it was accepted or rejected for another reason, 2 when the plan
// has nothing to plant the mistake in).  kinds:
//  0  two blocks of a front swapped in a forward chain launch
//  1  a backward chain block moved ahead of the block it waits for
//  2  a chain launch given a ticket counter past the end
//  3  a left-looking block job reaching left of its outer block
//  4  a chain block pointing at another block's inverse
extern "C" int emulate_validate_mistake(int n, const int* Cp, const int* Ci, int mode, int max_zeros, int kind) {
    AnalyzeOptions ao;
    ao.use_amd = 1;
    ao.partition = (Partition)mode;
    ao.max_zeros = max_zeros;
    Symbolic S;
    if (!analyze(n, Cp, Ci, ao, S).empty()) return 3;
    PlanOptions po;
    po.elem_bytes = 4;
    po.panel_lds_bytes = 61440;
    Plan P;
    if (!build_plan(S, po, P).empty() || !validate_plan(P, po).empty()) return 3;
    const char* want = nullptr;
    bool done = false;
    auto chain_pair = [&](std::vector<SolveStep>& steps, int kd) -> int {
        for (const SolveStep& st : steps) {
            if (st.kind != kd) continue;
            for (int t = st.first; t + 1 < st.first + st.count; t++) {
                const SolveItem &a = P.sv_items[t], &b = P.sv_items[t + 1];
                if (a.front == b.front && a.first && b.first) return t;
            }
        }
        return -1;
    };
    if (kind == 0 || kind == 1) {
        const int t = chain_pair(kind == 0 ? P.sv_fwd : P.sv_bwd, kind == 0 ? SV_FWD_CHAIN : SV_BWD_CHAIN);
        if (t < 0) return 2;
        std::swap(P.sv_items[t], P.sv_items[t + 1]);
        want = "solve chain order";
        done = true;
    } else if (kind == 2) {
        for (SolveStep& st : P.sv_fwd)
            if (st.kind == SV_FWD_CHAIN) {
                st.ctr = P.sv_nctr;
                done = true;
                break;
            }
        want = "solve chain counter";
    } else if (kind == 3) {
        for (BlkJob& bj : P.blk)
            if (bj.kprev > 0) {
                bj.kprev += 1;
                done = true;
                break;
            }
        want = "block left-looking range";
    } else if (kind == 4) {
        for (SolveItem& it : P.sv_items)
            if (it.first && it.c > 2 * kChainBW && it.a == kChainBW) {
                it.pk += (int64_t)kChainBW * (kSolveBW + 1);   // the next block's inverse
                done = true;
                break;
            }
        want = "solve chain inverse";
    } else if (kind == 5) {
        // the last join of side stream 2 dropped: its launches of that level
        // are never waited for
        for (size_t t = P.steps.size(); t-- > 0 && !done;) {
            Step& st = P.steps[t];
            if (st.stream != 0) continue;
            for (int* w : {&st.wait_ev2, &st.wait_ev})
                if (!done && *w >= 0)
                    for (const Step& q : P.steps)
                        if (q.stream == 2 && q.rec_ev == *w) {
                            *w = -1;
                            done = true;
                            break;
                        }
        }
        want = "side launch never joined";
    }
    if (!done) return 2;
    const std::string err = validate_plan(P, po);
    return err.find(want) != std::string::npos ? 0 : 1;
}

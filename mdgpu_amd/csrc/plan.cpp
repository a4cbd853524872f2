// Build the numeric execution plan (see plan.h).
#include "plan.h"

#include <algorithm>
#include <map>

namespace mdgpu {

namespace {

int64_t align_up(int64_t x, int64_t a) { return (x + a - 1) / a * a; }

// First-fit allocator over a linear arena, used only to compute static
// offsets for the update matrices.
class StaticArena {
   public:
    int64_t alloc(int64_t size) {
        size = align_up(std::max<int64_t>(size, 1), 64);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= size) {
                int64_t off = it->first;
                int64_t rest = it->second - size;
                free_.erase(it);
                if (rest > 0) free_[off + size] = rest;
                used_[off] = size;
                return off;
            }
        }
        int64_t off = top_;
        top_ += size;
        used_[off] = size;
        peak_ = std::max(peak_, top_);
        return off;
    }
    void release(int64_t off) {
        auto u = used_.find(off);
        if (u == used_.end()) return;
        int64_t size = u->second;
        used_.erase(u);
        auto it = free_.emplace(off, size).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
                it = pv;
            }
        }
        if (it->first + it->second == top_) {   // shrink the top
            top_ = it->first;
            free_.erase(it);
        }
    }
    int64_t peak() const { return peak_; }

   private:
    std::map<int64_t, int64_t> free_, used_;
    int64_t top_ = 0, peak_ = 0;
};

int pos_in(const int* q, int n, int row) {
    const int* it = std::lower_bound(q, q + n, row);
    if (it == q + n || *it != row) return -1;
    return (int)(it - q);
}

}  // namespace

// Offset in Plan::inv256 of the inverse of the kChainBW-column block of
// front d starting at pivot column a: a diagonal 128 x 128 block of the
// 256-column inverse holding it (leading dimension kSolveBW).
static int64_t chain_inverse(const FrontDesc& d, int a) {
    return d.dinv256 + (int64_t)(a / kSolveBW) * kSolveBW * kSolveBW +
           (int64_t)((a % kSolveBW) / kChainBW) * kChainBW * (kSolveBW + 1);
}

std::string build_plan(const Symbolic& S, const PlanOptions& o, Plan& P) {
    P = Plan();
    const int nf = (int)S.nodes.size();
    const int n = S.n;
    P.nfronts = nf;
    P.q = S.q_all;
    P.fronts.resize(nf);
    std::vector<int> level(nf, 0);
    int64_t lofs = 0;
    for (int f = 0; f < nf; f++) {
        FrontDesc& d = P.fronts[f];
        const PartNode& nd = S.nodes[f];
        d.k = nd.first;
        d.c = nd.ncols;
        d.n = (int)(S.q_off[f + 1] - S.q_off[f]);
        d.m = d.n - d.c;
        d.parent = S.front_parent[f];
        d.qofs = S.q_off[f];
        d.lofs = lofs;
        lofs += (int64_t)d.n * d.c;
        d.uofs = -1;
        d.kid_begin = (int)P.kids.size();
        for (int kid : nd.kids) {
            P.kids.push_back(kid);
            level[f] = std::max(level[f], level[kid] + 1);
        }
        d.kid_end = (int)P.kids.size();
        d.level = level[f];
        P.nlevels = std::max(P.nlevels, level[f] + 1);
        P.max_front = std::max(P.max_front, d.n);
        P.max_kids = std::max(P.max_kids, d.kid_end - d.kid_begin);
    }
    P.lstore = lofs;
    // split-plane offsets (bf16 variants): an n x ceil8(c) region per front
    std::vector<int64_t> pofs(nf, 0);
    if (o.presplit) {
        int64_t off = 0;
        for (int f = 0; f < nf; f++) {
            pofs[f] = off;
            off += (int64_t)P.fronts[f].n * align_up(P.fronts[f].c, 8);
        }
        P.planes = off;
    }
    {
        int64_t off = 0;
        for (int f = 0; f < nf; f++) {
            FrontDesc& d = P.fronts[f];
            d.dinv = off;
            const int nblk = (d.c + 63) / 64;
            off += (int64_t)(nblk - 1) * 4096 + (int64_t)(d.c - (nblk - 1) * 64) * 64;
        }
        P.dinv = std::max<int64_t>(off, 1);
        int64_t off256 = 0;
        for (int f = 0; f < nf; f++) {
            FrontDesc& d = P.fronts[f];
            d.dinv256 = -1;
            if (solve_small_front(d.c, d.n)) continue;
            d.dinv256 = off256;
            off256 += (int64_t)((d.c + kSolveBW - 1) / kSolveBW) * kSolveBW * kSolveBW;
        }
        P.inv256 = std::max<int64_t>(off256, 1);
    }

    // relative maps child update rows -> parent rows (VectorMergeIndices)
    for (int f = 0; f < nf; f++) {
        FrontDesc& d = P.fronts[f];
        d.relofs = (int64_t)P.rel.size();
        d.split = 0;
        if (d.parent < 0) {
            if (d.m != 0) return "root front has update rows";
            continue;
        }
        const FrontDesc& pd = P.fronts[d.parent];
        const size_t at = P.rel.size();
        P.rel.resize(at + d.m);
        if (!relative_map(S.q_all.data() + pd.qofs, pd.n, S.q_all.data() + d.qofs + d.c, d.m, P.rel.data() + at))
            return "update row missing from parent front";
        for (int t = 0; t < d.m; t++) d.split += P.rel[at + t] < pd.c;
    }

    // caller entry -> panel offset (MatrizGetFk_3's A scatter, mdgpu.cu:272-282)
    std::vector<int> front_of_col(n, -1);
    for (int f = 0; f < nf; f++)
        for (int j = P.fronts[f].k; j < P.fronts[f].k + P.fronts[f].c; j++) front_of_col[j] = f;
    P.a_dst.assign(S.nnz_in, -1);
    for (int j = 0; j < n; j++) {
        const int f = front_of_col[j];
        if (f < 0) return "column not covered by any front";
        const FrontDesc& d = P.fronts[f];
        const int* qf = S.q_all.data() + d.qofs;
        for (int t = S.a_low.p[j]; t < S.a_low.p[j + 1]; t++) {
            int pos = pos_in(qf, d.n, S.a_low.i[t]);
            if (pos < 0) return "A entry outside its front";
            P.a_dst[S.a_src[t]] = d.lofs + (int64_t)(j - d.k) * d.n + pos;
        }
    }

    // levels and the static update-arena plan
    P.level_fronts.assign(P.nlevels, {});
    for (int f = 0; f < nf; f++) P.level_fronts[level[f]].push_back(f);
    StaticArena arena;
    for (int L = 0; L < P.nlevels; L++) {
        for (int f : P.level_fronts[L])
            if (P.fronts[f].m > 0) P.fronts[f].uofs = arena.alloc((int64_t)P.fronts[f].m * P.fronts[f].m);
        for (int f : P.level_fronts[L])
            for (int t = P.fronts[f].kid_begin; t < P.fronts[f].kid_end; t++) {
                const FrontDesc& kd = P.fronts[P.kids[t]];
                if (kd.uofs >= 0) arena.release(kd.uofs);
            }
    }
    P.arena = arena.peak();

    // solve update vectors (one per front, no reuse)
    P.solve_uofs.resize(nf);
    int64_t su = 0;
    for (int f = 0; f < nf; f++) {
        P.solve_uofs[f] = su;
        su += align_up(std::max(P.fronts[f].m, 1), 8);
    }
    P.solve_u = su;

    // solve schedule (see SolveStep)
    {
        int nctr = 0;
        for (int L = 0; L < P.nlevels; L++) {
            std::vector<int> small, big;
            for (int f : P.level_fronts[L])
                (solve_small_front(P.fronts[f].c, P.fronts[f].n) ? small : big).push_back(f);
            if (!small.empty()) {
                SolveStep st{SV_FWD_SMALL, (int)P.sv_fronts.size(), (int)small.size(), 0};
                for (int f : small) st.bytes += (double)P.fronts[f].n * P.fronts[f].c * o.elem_bytes;
                P.sv_fronts.insert(P.sv_fronts.end(), small.begin(), small.end());
                P.sv_fwd.push_back(st);
            }
            if (big.empty()) continue;
            P.sv_fwd.push_back({SV_FWD_GATHER, (int)P.sv_fronts.size(), (int)big.size(), 0});
            P.sv_fronts.insert(P.sv_fronts.end(), big.begin(), big.end());
            SolveStep st{SV_FWD_CHAIN, (int)P.sv_items.size(), 0, 0};
            st.ctr = nctr++;
            for (int f : big) {
                const FrontDesc& d = P.fronts[f];
                for (int a = 0; a < d.c; a += kChainBW) P.sv_items.push_back({f, a, 1});
                for (int a = d.c; a < d.n; a += kChainBW) P.sv_items.push_back({f, a, 0});
                // strictly lower panel below the diagonal blocks, the block
                // inverses (128 x 128 each), the front vector in and out
                st.bytes += ((double)d.n * d.c - (double)d.c * (d.c + 1) / 2 + (double)kChainBW * d.c) * o.elem_bytes +
                            16.0 * d.n;
            }
            st.count = (int)P.sv_items.size() - st.first;
            P.sv_fwd.push_back(st);
        }
        for (int L = P.nlevels - 1; L >= 0; L--) {
            std::vector<int> small, big;
            for (int f : P.level_fronts[L])
                (solve_small_front(P.fronts[f].c, P.fronts[f].n) ? small : big).push_back(f);
            if (!small.empty()) {
                SolveStep st{SV_BWD_SMALL, (int)P.sv_fronts.size(), (int)small.size(), 0};
                for (int f : small) st.bytes += (double)P.fronts[f].n * P.fronts[f].c * o.elem_bytes;
                P.sv_fronts.insert(P.sv_fronts.end(), small.begin(), small.end());
                P.sv_bwd.push_back(st);
            }
            if (big.empty()) continue;
            {
                SolveStep st{SV_BWD_UPD, (int)P.sv_items.size(), 0, 0};
                std::vector<SolveItem> sums;
                int slot = 0;
                for (int f : big) {
                    const FrontDesc& d = P.fronts[f];
                    if (d.n == d.c) continue;
                    for (int c0 = 0; c0 < d.c; c0 += kSolveCols) {
                        const int first = slot;
                        for (int r0 = d.c; r0 < d.n; r0 += kSolveUpdRows) P.sv_items.push_back({f, c0, 0, r0, slot++, 0});
                        sums.push_back({f, c0, 0, 0, first, slot - first});
                    }
                    st.bytes += (double)(d.n - d.c) * d.c * o.elem_bytes + 12.0 * (d.n - d.c) + 16.0 * d.c;
                }
                st.count = (int)P.sv_items.size() - st.first;
                if (st.count > 0) {
                    P.sv_bwd.push_back(st);
                    SolveStep ss{SV_BWD_UPD_SUM, (int)P.sv_items.size(), (int)sums.size(), 8.0 * slot * kSolveCols};
                    P.sv_items.insert(P.sv_items.end(), sums.begin(), sums.end());
                    P.sv_bwd.push_back(ss);
                    P.sv_upd_partial = std::max<int64_t>(P.sv_upd_partial, (int64_t)slot * kSolveCols);
                }
            }
            SolveStep st{SV_BWD_CHAIN, (int)P.sv_items.size(), 0, 0};
            st.ctr = nctr++;
            for (int f : big) {
                const FrontDesc& d = P.fronts[f];
                for (int a = (d.c - 1) / kChainBW * kChainBW; a >= 0; a -= kChainBW) P.sv_items.push_back({f, a, 1});
                // triangle below the diagonal blocks and the block inverses
                st.bytes += ((double)d.c * (d.c + 1) / 2 + (double)kChainBW * d.c) * o.elem_bytes + 24.0 * d.c;
            }
            st.count = (int)P.sv_items.size() - st.first;
            P.sv_bwd.push_back(st);
        }
        P.sv_nctr = std::max(nctr, 1);
    }

    // front fields copied into the solve items
    for (const std::vector<SolveStep>* steps : {&P.sv_fwd, &P.sv_bwd})
        for (const SolveStep& st : *steps) {
            if (st.kind == SV_FWD_SMALL || st.kind == SV_BWD_SMALL || st.kind == SV_FWD_GATHER) continue;
            for (int t = st.first; t < st.first + st.count; t++) {
                SolveItem& it = P.sv_items[t];
                const FrontDesc& d = P.fronts[it.front];
                it.n = d.n;
                it.c = d.c;
                it.k = d.k;
                it.qofs = d.qofs;
                it.lofs = d.lofs;
                it.su = P.solve_uofs[it.front];
                it.pk = ((st.kind == SV_FWD_CHAIN || st.kind == SV_BWD_CHAIN) && it.first) ? chain_inverse(d, it.a) : 0;
            }
        }

    // schedule
    // GEMM kernel code of a launch: o.gemm_default unless the autotuner chose
    int gemm_ordinal = 0;   // GEMM launches emitted so far
    // A launch's jobs are collected first; its kernel (the autotuner's choice
    // for this launch ordinal, else `dflt`) and with it the tile grid are
    // fixed when the launch is emitted.  Returns false for an empty batch.
    auto flush_gemm = [&](std::vector<GemmJob>& batch, int L, int dflt) {
        if (batch.empty()) return false;
        const int kernel = gemm_ordinal < (int)o.gemm_tiles.size() ? o.gemm_tiles[gemm_ordinal] : dflt;
        const int T = gemm_tile_rows(kernel);
        int tiles = 0;
        for (GemmJob& g : batch) {
            g.tiles_m = (g.M + T - 1) / T;
            g.tiles_n = (g.N + T - 1) / T;
            if (g.lower) {
                g.ntiles = 0;
                for (int tj = 0; tj < g.tiles_n; tj++) g.ntiles += std::max(0, g.tiles_m - tj);
            } else {
                g.ntiles = g.tiles_m * g.tiles_n;
            }
            g.tile_base = tiles;
            tiles += g.ntiles;
        }
        gemm_ordinal++;
        Step s{STEP_GEMM, (int)P.gemm.size(), (int)batch.size(), tiles, L, kernel};
        for (const GemmJob& g : batch) {
            double entries = g.lower ? (double)g.N * g.M - 0.5 * (double)g.N * (g.N - 1) : (double)g.M * g.N;
            s.flops += 2.0 * g.K * entries;
            s.bytes += ((double)g.M + (g.lower ? 0.0 : g.N)) * g.K * o.elem_bytes + entries * o.elem_bytes * (g.beta0 ? 1 : 2);
        }
        P.gemm.insert(P.gemm.end(), batch.begin(), batch.end());
        P.steps.push_back(s);
        batch.clear();
        return true;
    };
    auto add_ea = [&](int L, bool first_pass) {
        // items of this pass: every column of a level-L front that receives
        // child columns (first pass: the parent's pivot columns; second: its
        // update columns), contributors in child order
        const int begin = (int)P.ea.size();
        std::vector<std::pair<double, EaItem>> items;
        for (int f : P.level_fronts[L]) {
            const FrontDesc& pd = P.fronts[f];
            std::vector<std::vector<EaContrib>> col(pd.n);
            for (int t = pd.kid_begin; t < pd.kid_end; t++) {
                const int ch = P.kids[t];
                const FrontDesc& cd = P.fronts[ch];
                const int j0 = first_pass ? 0 : cd.split, j1 = first_pass ? cd.split : cd.m;
                for (int j = j0; j < j1; j++)
                    col[P.rel[cd.relofs + j]].push_back({cd.uofs + (int64_t)j * cd.m, cd.relofs, j, cd.m});
            }
            for (int pc = 0; pc < pd.n; pc++) {
                if (col[pc].empty()) continue;
                EaItem it;
                it.in_u = pc >= pd.c;
                it.rowoff = it.in_u ? pd.c : 0;
                it.dst = it.in_u ? pd.uofs + (int64_t)(pc - pd.c) * pd.m : pd.lofs + (int64_t)pc * pd.n;
                it.cbeg = (int)P.ea_contrib.size();
                double work = 0;
                for (const EaContrib& e : col[pc]) {
                    P.ea_contrib.push_back(e);
                    work += e.m - e.j;
                }
                it.cend = (int)P.ea_contrib.size();
                items.push_back({work, it});
            }
        }
        if (items.empty()) return;
        // longest columns first (the order of disjoint targets does not
        // change any sum)
        std::stable_sort(items.begin(), items.end(),
                         [](const std::pair<double, EaItem>& x, const std::pair<double, EaItem>& y) { return x.first > y.first; });
        Step st{first_pass ? STEP_EA1 : STEP_EA2, begin, (int)items.size(), ((int)items.size() + 3) / 4, L, 0};
        for (const auto& w : items) {
            P.ea.push_back(w.second);
            // entries of the contributors: read child, RMW parent, map
            st.bytes += w.first * (3.0 * o.elem_bytes + 4.0);
        }
        P.steps.push_back(st);
    };


    std::vector<std::pair<size_t, int>> joins;   // (first step index, side event to wait for)
    for (int L = 0; L < P.nlevels; L++) {
        const std::vector<int>& fl = P.level_fronts[L];
        add_ea(L, true);
        // small panels: whole panel in LDS, one workgroup per front
        std::vector<int> blocked;
        {
            const int begin = (int)P.panel.size();
            int lds = 0;
            for (int f : fl) {
                const FrontDesc& d = P.fronts[f];
                int64_t bytes = (int64_t)d.n * d.c * o.elem_bytes;
                if (bytes <= o.panel_lds_bytes && d.c <= std::min(o.panel_max_pivots, kPanelMaxPivots)) {
                    P.panel.push_back(f);
                    lds = std::max<int>(lds, (int)bytes);
                } else {
                    blocked.push_back(f);
                }
            }
            const int count = (int)P.panel.size() - begin;
            if (count > 0) {
                Step st{STEP_PANEL, begin, count, count, L, lds};
                for (int t = begin; t < begin + count; t++) {
                    const FrontDesc& d = P.fronts[P.panel[t]];
                    for (int j = 0; j < d.c; j++) {
                        st.flops += (double)(d.n - j - 1);
                        for (int col = j + 1; col < d.c; col++) st.flops += 2.0 * (d.n - col);
                    }
                    st.bytes += 2.0 * d.n * d.c * o.elem_bytes;
                }
                P.steps.push_back(st);
            }
        }
        // large panels: blocked in two levels -- nb-column block columns,
        // left-looking inside nb_outer-column outer blocks (each BLK first
        // applies the outer block's earlier block columns to its own), and
        // right-looking across outer blocks, whose trailing update is one
        // deep (K = nb_outer) GEMM
        int maxc = 0;
        for (int f : blocked) maxc = std::max(maxc, P.fronts[f].c);
        // columns [0, u_done[f]) of a blocked front are applied to U_f
        std::vector<int> u_done(P.nfronts, 0);
        int side_last = -1;   // event of the last (b) launch (side stream 1)
        int side_waited = -1; // the latest side_last a main-stream (a) launch waited for
        int u_last = -1;      // event of the last update-matrix launch (side stream 2)
        for (int R0 = 0; R0 < maxc; R0 += o.nb_outer) {
            for (int r0 = R0; r0 < std::min(R0 + o.nb_outer, maxc); r0 += o.nb) {
                const int dbegin = (int)P.blk.size();
                int btiles = 0;
                for (int f : blocked) {
                    const FrontDesc& d = P.fronts[f];
                    if (r0 >= d.c) continue;
                    BlkJob bj;
                    bj.b = std::min(o.nb, d.c - r0);
                    bj.ld = d.n;
                    bj.p_off = d.lofs + (int64_t)r0 * d.n + r0;
                    bj.col0 = d.k + r0;
                    bj.rows = d.n - r0 - bj.b;
                    bj.ntiles = std::max(1, (bj.rows + kBlkRows - 1) / kBlkRows);
                    bj.tile_base = btiles;
                    bj.cnt = (int)P.blk.size();
                    bj.front = f;
                    bj.kprev = (r0 - R0) / 64;
                    if (o.presplit) bj.pl = pofs[f] + (int64_t)r0 * d.n + 8 * (int64_t)(r0 + bj.b);
                    btiles += bj.ntiles;
                    P.blk.push_back(bj);
                }
                const int dcount = (int)P.blk.size() - dbegin;
                if (dcount == 0) continue;
                {
                    Step st{STEP_BLK, dbegin, dcount, btiles, L, 0};
                    for (int t = dbegin; t < dbegin + dcount; t++) {
                        const double bb = P.blk[t].b, rows = P.blk[t].rows, kk = 64.0 * P.blk[t].kprev;
                        st.flops += bb * bb * bb / 3.0 + rows * bb * bb + 2.0 * (rows + bb) * bb * kk;
                        st.bytes += (2.0 * bb * bb / 2 + 2.0 * rows * bb + (rows + bb) * kk) * o.elem_bytes;
                    }
                    P.steps.push_back(st);
                }
            }
            // trailing update after the outer block, in three parts:
            //  (a) the next outer block's pivot columns, on the main stream
            //      (the next block's factorization needs them);
            //  (b) the later pivot columns, on side stream 1, overlapping the
            //      next outer block's BLK chain;
            //  (u) with u_lookahead, the front's update matrix U_f (rows and
            //      columns past the pivots, lower triangle), on side stream 2:
            //      U_f (+)= -L(c:n, R0:E) L(c:n, R0:E)^T, the first outer
            //      block with beta0.
            // (a) of this block waits for (b) of the previous one (both
            // update the same columns); (b) and (u) start when this outer
            // block is factored.  A front's last outer block has no trailing
            // update: its part of U_f is the level's Schur launch below.
            const int inner_done = P.nevents++;
            if (!P.steps.empty()) P.steps.back().rec_ev = inner_done;
            std::vector<GemmJob> batch_a, batch_b, batch_u;
            for (int f : blocked) {
                const FrontDesc& d = P.fronts[f];
                if (R0 >= d.c) continue;
                const int E = std::min(R0 + o.nb_outer, d.c);
                if (E >= d.c) continue;
                const int E2 = std::min(E + o.nb_outer, d.c);
                // split planes of the outer block's columns (presplit): the
                // block steps wrote them for the rows below each diagonal
                // block; R0, E, E2 and u_done are multiples of 8
                const int64_t pl0 = pofs[f] + (int64_t)R0 * d.n;
                GemmJob g{};
                g.a_off = d.lofs + (int64_t)R0 * d.n + E;
                g.b_off = g.a_off;
                g.c_off = d.lofs + (int64_t)E * d.n + E;
                if (o.presplit) g.a_pl = g.b_pl = pl0 + 8 * (int64_t)E;
                g.lda = g.ldb = g.ldc = d.n;
                g.M = d.n - E;
                g.N = E2 - E;
                g.K = E - R0;
                g.c_in_arena = 0;
                g.beta0 = 0;
                g.lower = 1;
                batch_a.push_back(g);
                if (E2 < d.c) {
                    GemmJob h = g;
                    h.a_off = d.lofs + (int64_t)R0 * d.n + E2;
                    h.b_off = h.a_off;
                    h.c_off = d.lofs + (int64_t)E2 * d.n + E2;
                    if (o.presplit) h.a_pl = h.b_pl = pl0 + 8 * (int64_t)E2;
                    h.M = d.n - E2;
                    h.N = d.c - E2;
                    batch_b.push_back(h);
                }
                // (u) once u_stride blocks are pending, and (u_flush) at each
                // of the last u_flush blocks before the front's last, so that
                // the Schur launch after the join is left at most one outer
                // block and the (u) launch it waits for at most one more
                const int blocks_left = (d.c - E + o.nb_outer - 1) / o.nb_outer;
                if (o.u_lookahead && d.m > 0 &&
                    (E - u_done[f] >= std::max(1, o.u_stride) * o.nb_outer || blocks_left <= o.u_flush)) {
                    const int k0 = u_done[f];
                    GemmJob u{};
                    u.a_off = d.lofs + (int64_t)k0 * d.n + d.c;
                    u.b_off = u.a_off;
                    if (o.presplit) u.a_pl = u.b_pl = pofs[f] + (int64_t)k0 * d.n + 8 * (int64_t)d.c;
                    u.c_off = d.uofs;
                    u.lda = u.ldb = d.n;
                    u.ldc = d.m;
                    u.M = u.N = d.m;
                    u.K = E - k0;
                    u.c_in_arena = 1;
                    u.beta0 = k0 == 0;
                    u.lower = 1;
                    batch_u.push_back(u);
                    u_done[f] = E;
                }
            }
            const int wait_prev = side_last;
            const int gk = o.presplit ? (o.gemm_default | kGemmPlanes) : o.gemm_default;
            if (flush_gemm(batch_a, L, gk) && wait_prev >= 0) {
                P.steps.back().wait_ev = wait_prev;
                side_waited = wait_prev;
            }
            if (flush_gemm(batch_b, L, gk)) {
                Step& sb = P.steps.back();
                sb.stream = 1;
                sb.wait_ev = inner_done;
                sb.rec_ev = side_last = P.nevents++;
            }
            const int uk = o.schur_kernel ? o.schur_kernel : (o.presplit && o.nplanes == 3) ? (gk | kGemmSix) : gk;
            if (flush_gemm(batch_u, L, uk)) {
                Step& su = P.steps.back();
                su.stream = 2;
                su.wait_ev = inner_done;
                su.rec_ev = u_last = P.nevents++;
            }
        }
        // Schur complements of every front of the level: U_f = -L21 L21^T
        // (with u_lookahead, for a blocked front only the columns from
        // u_done[f] on: the earlier ones went to side stream 2).  The first
        // main-stream launch after the blocked part waits for both side
        // streams (see the join pass after the loop); side stream 1 only if
        // its last launch is not already waited for by an (a) launch (each
        // front's last non-empty (b) is followed by an (a) that waits for it)
        if (side_last >= 0 && side_last != side_waited) joins.push_back({P.steps.size(), side_last});
        if (u_last >= 0) joins.push_back({P.steps.size(), u_last});
        {
            // presplit: the launch reads the split planes when every front
            // of it was blocked (fronts factored in LDS write no planes)
            std::vector<char> is_blocked(P.nfronts, 0);
            for (int f : blocked) is_blocked[f] = 1;
            bool planes = o.presplit;
            for (int f : fl)
                if (P.fronts[f].m > 0 && !is_blocked[f]) planes = false;
            std::vector<GemmJob> batch;
            int kmax = 0;
            for (int f : fl) {
                const FrontDesc& d = P.fronts[f];
                if (d.m == 0) continue;
                const int k0 = u_done[f];
                GemmJob g{};
                g.a_off = d.lofs + (int64_t)k0 * d.n + d.c;
                g.b_off = g.a_off;
                if (planes) g.a_pl = g.b_pl = pofs[f] + (int64_t)k0 * d.n + 8 * (int64_t)d.c;
                g.c_off = d.uofs;
                g.lda = g.ldb = d.n;
                g.ldc = d.m;
                g.M = g.N = d.m;
                g.K = d.c - k0;
                g.c_in_arena = 1;
                g.beta0 = k0 == 0;
                g.lower = 1;
                kmax = std::max(kmax, g.K);
                batch.push_back(g);
            }
            const bool split = o.split_k_min > 0 && kmax >= o.split_k_min;
            flush_gemm(batch, L,
                       o.schur_kernel ? o.schur_kernel
                                      : o.gemm_default | (split ? kGemmSplit : 0) | (planes ? kGemmPlanes : 0) |
                                            (planes && o.nplanes == 3 ? kGemmSix : 0));
        }
        add_ea(L, false);
    }
    // inverses of the diagonal blocks of every front, for the solves (off
    // the factorization's critical path: one launch at the end)
    {
        const int begin = (int)P.inv.size();
        for (int f = 0; f < nf; f++)
            for (int j0 = 0; j0 < P.fronts[f].c; j0 += 64) P.inv.push_back({f, j0});
        const int count = (int)P.inv.size() - begin;
        if (count > 0) {
            Step st{STEP_INV, begin, count, count, P.nlevels - 1, 0};
            for (int t = begin; t < begin + count; t++) {
                const double b = std::min(64, P.fronts[P.inv[t].front].c - P.inv[t].j0);
                st.flops += b * b * b / 3.0;
                st.bytes += 1.5 * b * b * o.elem_bytes;
            }
            P.steps.push_back(st);
        }
        const int b256 = (int)P.inv256_jobs.size();
        for (int f = 0; f < nf; f++)
            if (P.fronts[f].dinv256 >= 0)
                for (int j0 = 0; j0 < P.fronts[f].c; j0 += kSolveBW) P.inv256_jobs.push_back({f, j0});
        const int c256 = (int)P.inv256_jobs.size() - b256;
        if (c256 > 0) {
            Step st{STEP_INV256, b256, c256, c256, P.nlevels - 1, 0};
            for (int t = b256; t < b256 + c256; t++) {
                // the two 128-column halves
                const int b = std::min(kSolveBW, P.fronts[P.inv256_jobs[t].front].c - P.inv256_jobs[t].j0);
                for (int h = 0; h < b; h += kChainBW) {
                    const double bh = std::min(kChainBW, b - h);
                    st.flops += bh * bh * bh / 3.0;
                }
                st.bytes += 1.5 * b * b * o.elem_bytes;
            }
            P.steps.push_back(st);
        }
    }

    // join pass: the first main-stream launch at or after each recorded
    // position waits for the side stream's last event of that level
    for (const auto& jn : joins) {
        size_t t = jn.first;
        while (t < P.steps.size() && P.steps[t].stream != 0) t++;
        if (t == P.steps.size()) return "look-ahead: no launch to join the side stream";
        if (P.steps[t].wait_ev < 0) P.steps[t].wait_ev = jn.second;
        else if (P.steps[t].wait_ev2 < 0) P.steps[t].wait_ev2 = jn.second;
        else return "look-ahead: join on a launch that already waits twice";
    }

    // full symmetric A in the caller's numbering (CSR rows sorted), for the
    // residual r = b - A x of iterative refinement
    {
        const int nz = S.a_low.nnz();
        std::vector<int> rr, cc, ss;
        rr.reserve((size_t)nz * 2);
        cc.reserve((size_t)nz * 2);
        ss.reserve((size_t)nz * 2);
        for (int j = 0; j < n; j++)
            for (int t = S.a_low.p[j]; t < S.a_low.p[j + 1]; t++) {
                int r = S.perm[S.a_low.i[t]], c = S.perm[j];
                rr.push_back(r); cc.push_back(c); ss.push_back(S.a_src[t]);
                if (r != c) { rr.push_back(c); cc.push_back(r); ss.push_back(S.a_src[t]); }
            }
        const int nnz = (int)rr.size();
        std::vector<int> bycol(nnz), cp(n + 1, 0);
        for (int t = 0; t < nnz; t++) cp[cc[t] + 1]++;
        for (int t = 0; t < n; t++) cp[t + 1] += cp[t];
        for (int t = 0; t < nnz; t++) bycol[cp[cc[t]]++] = t;
        P.afull_p.assign(n + 1, 0);
        for (int t = 0; t < nnz; t++) P.afull_p[rr[t] + 1]++;
        for (int t = 0; t < n; t++) P.afull_p[t + 1] += P.afull_p[t];
        std::vector<int> fill(P.afull_p.begin(), P.afull_p.end() - 1);
        P.afull_i.resize(nnz);
        P.afull_src.resize(nnz);
        for (int u = 0; u < nnz; u++) {       // columns ascending within rows
            int t = bycol[u];
            int q = fill[rr[t]]++;
            P.afull_i[q] = cc[t];
            P.afull_src[q] = ss[t];
        }
    }
    return "";
}

}  // namespace mdgpu

namespace mdgpu {

std::string validate_plan(const Plan& P, const PlanOptions& o) {
    const int nf = P.nfronts;
    auto bad = [](const std::string& s) { return "plan check: " + s; };
    for (int f = 0; f < nf; f++) {
        const FrontDesc& d = P.fronts[f];
        if (d.c < 1 || d.n < d.c || d.m != d.n - d.c) return bad("front shape");
        if (d.lofs < 0 || d.lofs + (int64_t)d.n * d.c > P.lstore) return bad("panel outside L store");
        if (d.m > 0 && (d.uofs < 0 || d.uofs + (int64_t)d.m * d.m > P.arena)) return bad("U outside arena");
        if (d.parent >= 0) {
            const FrontDesc& pd = P.fronts[d.parent];
            if (d.parent <= f || pd.level <= d.level) return bad("parent order");
            int prev = -1;
            for (int t = 0; t < d.m; t++) {
                int r = P.rel[d.relofs + t];
                if (r <= prev || r >= pd.n) return bad("relative map");
                prev = r;
            }
            if (d.split < 0 || d.split > d.m) return bad("split");
        }
    }
    for (int64_t a : P.a_dst)
        if (a < 0 || a >= P.lstore) return bad("A scatter target");
    for (const EaItem& it : P.ea) {
        if (it.cbeg < 0 || it.cend > (int)P.ea_contrib.size() || it.cbeg >= it.cend) return bad("EA contributors");
        if (it.in_u ? (it.rowoff < 0 || it.dst < 0) : (it.rowoff != 0 || it.dst < 0 || it.dst >= P.lstore))
            return bad("EA target");
        for (int t = it.cbeg; t < it.cend; t++) {
            const EaContrib& e = P.ea_contrib[t];
            if (e.j < 0 || e.j >= e.m || e.rel < 0 || e.rel + e.m > (int64_t)P.rel.size()) return bad("EA contributor");
            if (e.src < 0 || e.src + e.m > P.arena) return bad("EA source");
            // every target row lies in the parent column: rel values are
            // below n_parent, checked with the relative maps above; the
            // column itself must hold them
            const int last = P.rel[e.rel + e.m - 1] - it.rowoff;
            if (P.rel[e.rel + e.j] - it.rowoff < 0) return bad("EA row below the column");
            const int64_t cap = it.in_u ? P.arena : P.lstore;
            if (it.dst + last >= cap) return bad("EA target bounds");
        }
    }
    // live update matrices must not overlap: at level L the live set is
    // {f : level(f) <= L <= level(parent(f))}
    for (int L = 0; L < P.nlevels; L++) {
        std::vector<std::pair<int64_t, int64_t>> iv;
        for (int f = 0; f < nf; f++) {
            const FrontDesc& d = P.fronts[f];
            if (d.m == 0) continue;
            int end = d.parent >= 0 ? P.fronts[d.parent].level : d.level;
            if (d.level <= L && L <= end) iv.push_back({d.uofs, d.uofs + (int64_t)d.m * d.m});
        }
        std::sort(iv.begin(), iv.end());
        for (size_t t = 1; t < iv.size(); t++)
            if (iv[t].first < iv[t - 1].second) return bad("overlapping update matrices at level " + std::to_string(L));
    }
    for (const GemmJob& g : P.gemm) {
        if (g.M <= 0 || g.N <= 0 || g.K <= 0 || (g.lower && g.M < g.N)) return bad("gemm shape");
        int64_t alast = g.a_off + (g.M - 1) + (int64_t)(g.K - 1) * g.lda;
        int64_t blast = g.b_off + (g.N - 1) + (int64_t)(g.K - 1) * g.ldb;
        int64_t clast = g.c_off + (g.M - 1) + (int64_t)(g.N - 1) * g.ldc;
        if (g.a_off < 0 || alast >= P.lstore || g.b_off < 0 || blast >= P.lstore) return bad("gemm A/B bounds");
        if (g.c_off < 0 || clast >= (g.c_in_arena ? P.arena : P.lstore)) return bad("gemm C bounds");
    }
    for (size_t t = 0; t < P.blk.size(); t++) {
        const BlkJob& bj = P.blk[t];
        if (bj.b < 1 || bj.b > o.nb || bj.b > 64 || bj.rows < 0) return bad("block shape");
        if (bj.ntiles != std::max(1, (bj.rows + kBlkRows - 1) / kBlkRows)) return bad("block tiles");
        if (bj.cnt != (int)t) return bad("block counter");
        if (bj.p_off < 0 || bj.p_off + (bj.b - 1) + bj.rows + (int64_t)(bj.b - 1) * bj.ld >= P.lstore)
            return bad("block bounds");
        if (bj.b + bj.rows > bj.ld) return bad("block rows exceed panel");
        if (bj.front < 0 || bj.front >= nf) return bad("block front");
        if (o.presplit != (bj.pl >= 0)) return bad("block split planes");
        if (bj.pl >= 0 && bj.rows > 0 &&
            (bj.pl % 8 || bj.pl + (int64_t)((bj.b + 7) / 8 - 1) * 8 * bj.ld + 8 * (int64_t)(bj.rows - 1) + 7 >= P.planes))
            return bad("block split-plane bounds");
        {
            const FrontDesc& d = P.fronts[bj.front];
            const int r0 = bj.col0 - d.k;
            if (r0 < 0 || r0 >= d.c || r0 % 64 || bj.ld != d.n || bj.p_off != d.lofs + (int64_t)r0 * d.n + r0)
                return bad("block position");
            if (bj.kprev < 0 || 64 * bj.kprev > r0 || (r0 - 64 * bj.kprev) % o.nb_outer ||
                64 * (bj.kprev + 1) > o.nb_outer)
                return bad("block left-looking range");
        }
    }
    auto check_solve = [&](const std::vector<SolveStep>& steps) -> std::string {
        for (const SolveStep& st : steps) {
            const bool uses_fronts = st.kind == SV_FWD_SMALL || st.kind == SV_BWD_SMALL || st.kind == SV_FWD_GATHER;
            const int64_t cap = uses_fronts ? (int64_t)P.sv_fronts.size() : (int64_t)P.sv_items.size();
            if (st.count <= 0 || st.first < 0 || st.first + st.count > cap) return "solve step slice";
            if (uses_fronts) {
                for (int t = st.first; t < st.first + st.count; t++) {
                    if (P.sv_fronts[t] < 0 || P.sv_fronts[t] >= nf) return "solve front";
                    if (st.kind != SV_FWD_GATHER &&
                        !solve_small_front(P.fronts[P.sv_fronts[t]].c, P.fronts[P.sv_fronts[t]].n))
                        return "solve small front";
                }
                continue;
            }
            for (int t = st.first; t < st.first + st.count; t++) {
                const SolveItem& it = P.sv_items[t];
                if (it.front < 0 || it.front >= nf) return "solve item front";
                const FrontDesc& d = P.fronts[it.front];
                if (it.n != d.n || it.c != d.c || it.k != d.k || it.qofs != d.qofs || it.lofs != d.lofs ||
                    it.su != P.solve_uofs[it.front])
                    return "solve item front fields";
                if (solve_small_front(d.c, d.n) || d.dinv256 < 0) return "solve item front size";
                if (st.kind == SV_BWD_UPD || st.kind == SV_BWD_UPD_SUM) {
                    if (d.n <= d.c || it.a < 0 || it.a >= d.c || it.a % kSolveCols) return "solve update columns";
                    if (st.kind == SV_BWD_UPD &&
                        (it.j0 < d.c || it.j0 >= d.n || (int64_t)(it.slot + 1) * kSolveCols > P.sv_upd_partial))
                        return "solve update rows";
                    if (st.kind == SV_BWD_UPD_SUM &&
                        (it.slot < 0 || it.nslot < 1 || (int64_t)(it.slot + it.nslot) * kSolveCols > P.sv_upd_partial))
                        return "solve update slots";
                    continue;
                }
                // chain items: TRI blocks of pivots in order (FWD) or reverse
                // order (BWD), then (FWD) RECT chunks of update rows; every
                // item of a front in the same launch
                if (it.first) {
                    if (it.a < 0 || it.a >= d.c || it.a % kChainBW) return "solve chain block";
                    if (it.pk != chain_inverse(d, it.a) || it.pk < 0 ||
                        it.pk + (int64_t)(kChainBW - 1) * (kSolveBW + 1) >= P.inv256)
                        return "solve chain inverse";
                } else if (st.kind != SV_FWD_CHAIN || it.a < d.c || it.a >= d.n || (it.a - d.c) % kChainBW) {
                    return "solve chain rows";
                }
                if (st.ctr < 0 || st.ctr >= P.sv_nctr) return "solve chain counter";
                // the blocks an item waits for come earlier in the launch:
                // FWD, blocks 0.. of the front at [t0, t0 + nwait); BWD, the
                // blocks after it, last one first, at [t - nwait, t)
                const int nb = (d.c + kChainBW - 1) / kChainBW;
                int t0, nwait;
                if (st.kind == SV_FWD_CHAIN) {
                    nwait = it.first ? it.a / kChainBW : nb;
                    t0 = t - (it.first ? nwait : nb + (it.a - d.c) / kChainBW);
                } else {
                    nwait = nb - 1 - it.a / kChainBW;
                    t0 = t - nwait;
                }
                if (t0 < st.first) return "solve chain order";
                for (int u = t0; u < t0 + nwait; u++) {
                    const SolveItem& e = P.sv_items[u];
                    const int want = st.kind == SV_FWD_CHAIN ? (u - t0) * kChainBW : it.a + (t - u) * kChainBW;
                    if (e.front != it.front || !e.first || e.a != want) return "solve chain order";
                }
            }
        }
        return "";
    };
    std::string se = check_solve(P.sv_fwd);
    if (se.empty()) se = check_solve(P.sv_bwd);
    if (!se.empty()) return bad(se);
    for (const InvJob& ij : P.inv) {
        if (ij.front < 0 || ij.front >= nf) return bad("inverse job front");
        const FrontDesc& d = P.fronts[ij.front];
        if (ij.j0 < 0 || ij.j0 >= d.c || ij.j0 % 64) return bad("inverse job block");
        if (d.dinv + (int64_t)(ij.j0 / 64) * 4096 + (int64_t)std::min(64, d.c - ij.j0) * 64 > P.dinv)
            return bad("inverse job store");
    }
    for (int f = 0; f < nf; f++) {
        const FrontDesc& d = P.fronts[f];
        const int nblk = (d.c + 63) / 64;
        if (d.dinv < 0 || d.dinv + (int64_t)(nblk - 1) * 4096 + (int64_t)(d.c - (nblk - 1) * 64) * 64 > P.dinv)
            return bad("front inverse store");
        if (solve_small_front(d.c, d.n) != (d.dinv256 < 0)) return bad("front block-inverse presence");
        if (d.dinv256 >= 0 &&
            d.dinv256 + (int64_t)((d.c + kSolveBW - 1) / kSolveBW) * kSolveBW * kSolveBW > P.inv256)
            return bad("front block-inverse store");
        if (d.dinv256 >= 0 && d.dinv256 % ((int64_t)kSolveBW * kSolveBW)) return bad("front block-inverse alignment");
    }
    for (const InvJob& ij : P.inv256_jobs) {
        if (ij.front < 0 || ij.front >= nf || P.fronts[ij.front].dinv256 < 0) return bad("block-inverse job front");
        if (ij.j0 < 0 || ij.j0 >= P.fronts[ij.front].c || ij.j0 % kSolveBW) return bad("block-inverse job block");
    }
    if (o.nb != 64) return bad("blocked path needs 64-column diagonal blocks");
    if (o.nb_outer < o.nb || o.nb_outer % o.nb) return bad("outer block must be a multiple of the block column");
    for (int f : P.panel) {
        const FrontDesc& d = P.fronts[f];
        if ((int64_t)d.n * d.c * o.elem_bytes > o.panel_lds_bytes) return bad("panel LDS");
        if (d.c > kPanelMaxPivots) return bad("panel pivots");
    }
    // look-ahead events: recorded once, waited only after being recorded,
    // only GEMMs on the side streams, and every side launch joined back to
    // the main stream: a later main-stream launch waits for the event of it
    // or of a later launch on the same side stream (stream order)
    {
        std::vector<int> rec_at(std::max(P.nevents, 1), -1);
        std::vector<char> waited(std::max(P.nevents, 1), 0);   // by a main-stream launch
        for (size_t t = 0; t < P.steps.size(); t++) {
            const Step& st = P.steps[t];
            if (st.stream < 0 || st.stream > 2 || (st.stream != 0 && st.kind != STEP_GEMM))
                return bad("side-stream launch kind");
            if (st.wait_ev >= P.nevents || st.rec_ev >= P.nevents || st.wait_ev2 >= P.nevents)
                return bad("event index");
            for (int w : {st.wait_ev, st.wait_ev2}) {
                if (w < 0) continue;
                if (rec_at[w] < 0) return bad("wait on an event not yet recorded");
                if (st.stream == 0) waited[w] = 1;
            }
            if (st.rec_ev >= 0) {
                if (rec_at[st.rec_ev] >= 0) return bad("event recorded twice");
                rec_at[st.rec_ev] = (int)t;
            }
        }
        for (size_t t = 0; t < P.steps.size(); t++) {
            const int s = P.steps[t].stream;
            if (s == 0) continue;
            if (P.steps[t].rec_ev < 0) return bad("side launch records no event");
            bool joined = false;
            for (size_t u = t; u < P.steps.size() && !joined; u++) {
                const Step& su = P.steps[u];
                if (su.stream != s || su.rec_ev < 0 || !waited[su.rec_ev]) continue;
                // the waiting main-stream launch comes after u
                for (size_t v = u + 1; v < P.steps.size() && !joined; v++)
                    joined = P.steps[v].stream == 0 &&
                             (P.steps[v].wait_ev == su.rec_ev || P.steps[v].wait_ev2 == su.rec_ev);
            }
            if (!joined) return bad("side launch never joined");
        }
    }
    for (const Step& s : P.steps) {
        if (s.nblocks <= 0 || s.count <= 0) return bad("empty step");
        if (s.kind == STEP_GEMM) {
            if (!gemm_kernel_valid(s.arg, o.elem_bytes) || gemm_tile_rows(s.arg) > o.tile)
                return bad("gemm kernel or tile size");
            if (gemm_planes(s.arg) && !o.presplit) return bad("split-plane GEMM without planes");
            if (gemm_six(s.arg) && o.nplanes != 3) return bad("six-product GEMM without three planes");
            for (int t = s.first; gemm_planes(s.arg) && t < s.first + s.count; t++) {
                // rows past M are read clamped to row M - 1; octets past K
                // are not read
                const GemmJob& g = P.gemm[t];
                if (g.a_pl < 0 || g.b_pl < 0 || g.a_pl % 8 || g.b_pl % 8) return bad("split-plane GEMM job");
                const int64_t kl = (int64_t)((g.K + 7) / 8 - 1) * 8;
                if (g.a_pl + kl * g.lda + 8 * (int64_t)(g.M - 1) + 7 >= P.planes ||
                    g.b_pl + kl * g.ldb + 8 * (int64_t)(g.N - 1) + 7 >= P.planes)
                    return bad("split-plane GEMM bounds");
            }
            const int T = gemm_tile_rows(s.arg);
            int tiles = 0;
            for (int t = s.first; t < s.first + s.count; t++) {
                const GemmJob& g = P.gemm[t];
                if (g.tiles_m != (g.M + T - 1) / T || g.tiles_n != (g.N + T - 1) / T)
                    return bad("gemm tiles do not match the launch's tile size");
                if (P.gemm[t].tile_base != tiles) return bad("gemm tile bases");
                tiles += P.gemm[t].ntiles;
            }
            if (tiles != s.nblocks) return bad("gemm tile count");
        }
        if (s.kind == STEP_BLK) {
            int tiles = 0;
            for (int t = s.first; t < s.first + s.count; t++) {
                if (P.blk[t].tile_base != tiles) return bad("block tile bases");
                tiles += P.blk[t].ntiles;
            }
            if (tiles != s.nblocks) return bad("block tile count");
        }
    }
    return "";
}

}  // namespace mdgpu

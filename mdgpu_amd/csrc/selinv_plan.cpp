// Host schedule of the selected inverse (see selinv.h).
#include "selinv.h"

#include <algorithm>

namespace mdgpu {

namespace {

constexpr int kTile = 64;
constexpr int64_t kCopyChunk = 4096;    // elements per copy work item

int tiles(int x) { return (x + kTile - 1) / kTile; }

// split columns [0, ncols) of a rows-tall matrix into items of ~kCopyChunk elements
void add_copy_items(std::vector<SelCopy>& out, int front, int ncols, int64_t rows, int b0 = 0, int b = 0) {
    const int per = (int)std::max<int64_t>(1, kCopyChunk / std::max<int64_t>(rows, 1));
    for (int j = 0; j < ncols; j += per) out.push_back({front, j, std::min(ncols, j + per), b0, b, 0});
}

}  // namespace

std::string build_selinv_plan(const Plan& P, SelPlan& Q) {
    Q = SelPlan();
    const int nf = P.nfronts;
    Q.zf_ofs.assign(nf, 0);
    Q.lh_ofs.assign(nf, 0);
    auto add_gemm = [&](std::vector<SelGemm>& batch, int& ntile, SelGemm g) {
        if (g.M <= 0 || g.N <= 0) return;
        g.tiles_m = tiles(g.M);
        g.ntiles = g.tiles_m * tiles(g.N);
        g.tile_base = ntile;
        ntile += g.ntiles;
        batch.push_back(g);
        Q.flops += 2.0 * g.M * g.N * std::max(g.K, 1);
    };
    auto flush_gemm = [&](std::vector<SelGemm>& batch, int ntile, int L) {
        if (batch.empty()) return;
        Q.steps.push_back({SS_GEMM, 0, (int)Q.gemm.size(), (int)batch.size(), ntile, L});
        Q.gemm.insert(Q.gemm.end(), batch.begin(), batch.end());
        batch.clear();
    };
    auto flush_copy = [&](int kind, int first, int L) {
        const int count = (int)Q.copy.size() - first;
        if (count > 0) Q.steps.push_back({SS_COPY, kind, first, count, count, L});
    };
    for (int L = P.nlevels - 1; L >= 0; L--) {
        const std::vector<int>& fl = P.level_fronts[L];
        int64_t zf = 0, lh = 0;
        int maxblk = 0;
        for (int f : fl) {
            const FrontDesc& d = P.fronts[f];
            Q.zf_ofs[f] = zf;
            zf += (int64_t)d.n * d.n;
            Q.lh_ofs[f] = lh;
            lh += (int64_t)d.n * kTile;
            maxblk = std::max(maxblk, (d.c + 63) / 64);
        }
        Q.zf_size = std::max(Q.zf_size, zf);
        Q.lh_size = std::max(Q.lh_size, lh);
        // Z22_f from the arena into ZF_f
        int first = (int)Q.copy.size();
        for (int f : fl)
            if (P.fronts[f].m > 0) add_copy_items(Q.copy, f, P.fronts[f].m, P.fronts[f].m);
        flush_copy(SC_INIT, first, L);
        for (int t = 0; t < maxblk; t++) {
            std::vector<SelGemm> b1, b2, b3;
            int n1 = 0, n2 = 0, n3 = 0;
            const int tfirst = (int)Q.copy.size();
            for (int f : fl) {
                const FrontDesc& d = P.fronts[f];
                const int nblk = (d.c + 63) / 64;
                if (t >= nblk) continue;
                const int k = nblk - 1 - t, j0 = 64 * k, b = std::min(64, d.c - j0);
                const int r0 = j0 + b, r = d.n - r0;
                const int64_t zf0 = Q.zf_ofs[f], w = d.dinv + (int64_t)k * 4096;
                if (r > 0) {
                    // Lh = L(R, J) inv(L(J,J)): inv(L)(k', j) = W[j + 64 k']
                    add_gemm(b1, n1, {d.lofs + r0 + (int64_t)j0 * d.n, w, Q.lh_ofs[f], SB_L, SB_W, SB_LH,
                                      d.n, 64, d.n, r, b, b, 0, 1, 1.0, 0.0, 0, 0, 0, 0});
                    // Z(R, J) = -Z(R, R) Lh
                    add_gemm(b2, n2, {zf0 + r0 + (int64_t)r0 * d.n, Q.lh_ofs[f], zf0 + r0 + (int64_t)j0 * d.n,
                                      SB_ZF, SB_LH, SB_ZF, d.n, d.n, d.n, r, b, r, 0, 0, -1.0, 0.0, 0, 0, 0, 0});
                    add_copy_items(Q.copy, f, r, b, j0, b);   // SC_TRANS: rows of R in chunks
                }
                // Z(J, J) = inv(L)^T inv(L): A(i, k') = inv(L)(k', i) = W[i + 64 k'], B(k', j) = W[j + 64 k']
                add_gemm(b2, n2, {w, w, zf0 + j0 + (int64_t)j0 * d.n, SB_W, SB_W, SB_ZF, 64, 64, d.n, b, b, b,
                                  0, 1, 1.0, 0.0, 0, 0, 0, 0});
                if (r > 0)   // Z(J, J) -= Lh^T Z(R, J)
                    add_gemm(b3, n3, {Q.lh_ofs[f], zf0 + r0 + (int64_t)j0 * d.n, zf0 + j0 + (int64_t)j0 * d.n,
                                      SB_LH, SB_ZF, SB_ZF, d.n, d.n, d.n, b, b, r, 1, 0, -1.0, 1.0, 0, 0, 0, 0});
            }
            // the transpose items were appended above in front order; move
            // them behind the two GEMM steps that produce Z(R, J)
            std::vector<SelCopy> trans(Q.copy.begin() + tfirst, Q.copy.end());
            Q.copy.resize(tfirst);
            flush_gemm(b1, n1, L);
            flush_gemm(b2, n2, L);
            const int cfirst = (int)Q.copy.size();
            Q.copy.insert(Q.copy.end(), trans.begin(), trans.end());
            flush_copy(SC_TRANS, cfirst, L);
            flush_gemm(b3, n3, L);
        }
        first = (int)Q.copy.size();
        for (int f : fl) add_copy_items(Q.copy, f, P.fronts[f].c, P.fronts[f].n);
        flush_copy(SC_OUT, first, L);
        first = (int)Q.copy.size();
        for (int f : fl)
            for (int t = P.fronts[f].kid_begin; t < P.fronts[f].kid_end; t++) {
                const int ch = P.kids[t];
                if (P.fronts[ch].m > 0) add_copy_items(Q.copy, ch, P.fronts[ch].m, P.fronts[ch].m);
            }
        flush_copy(SC_PUSH, first, L);
    }
    return "";
}

std::string validate_selinv_plan(const Plan& P, const SelPlan& Q) {
    auto cap = [&](int buf) -> int64_t {
        switch (buf) {
            case SB_L: case SB_Z: return P.lstore;
            case SB_ZF: return Q.zf_size;
            case SB_LH: return Q.lh_size;
            case SB_W: return P.dinv;
            case SB_U: return std::max<int64_t>(P.arena, 1);
        }
        return -1;
    };
    auto last = [](int64_t off, int rows, int cols, int ld) { return off + (rows - 1) + (int64_t)(cols - 1) * ld; };
    for (const SelGemm& g : Q.gemm) {
        if (g.M <= 0 || g.N <= 0 || g.K < 0) return "selinv gemm shape";
        const int ar = g.ta ? g.K : g.M, ac = g.ta ? g.M : g.K;
        const int br = g.tb ? g.N : g.K, bc = g.tb ? g.K : g.N;
        if (g.a_off < 0 || last(g.a_off, ar, ac, g.lda) >= cap(g.a_buf) || ar > g.lda) return "selinv gemm A bounds";
        if (g.b_off < 0 || last(g.b_off, br, bc, g.ldb) >= cap(g.b_buf) || br > g.ldb) return "selinv gemm B bounds";
        if (g.c_off < 0 || last(g.c_off, g.M, g.N, g.ldc) >= cap(g.c_buf) || g.M > g.ldc) return "selinv gemm C bounds";
        if (g.c_buf == SB_L || g.c_buf == SB_W) return "selinv gemm writes the factor";
    }
    for (const SelStep& s : Q.steps) {
        if (s.count <= 0 || s.nblocks <= 0) return "selinv empty step";
        if (s.kind == SS_GEMM) {
            int t = 0;
            for (int j = s.first; j < s.first + s.count; j++) {
                if (Q.gemm[j].tile_base != t) return "selinv tile bases";
                t += Q.gemm[j].ntiles;
            }
            if (t != s.nblocks) return "selinv tile count";
            continue;
        }
        for (int j = s.first; j < s.first + s.count; j++) {
            const SelCopy& c = Q.copy[j];
            if (c.front < 0 || c.front >= P.nfronts) return "selinv copy front";
            const FrontDesc& d = P.fronts[c.front];
            const int64_t zf_end = Q.zf_ofs[c.front] + (int64_t)d.n * d.n;
            switch (s.copy_kind) {
                case SC_INIT:
                case SC_PUSH:
                    if (c.j0 < 0 || c.j1 > d.m || d.uofs < 0) return "selinv copy update block";
                    if (s.copy_kind == SC_INIT && zf_end > Q.zf_size) return "selinv copy ZF";
                    if (s.copy_kind == SC_PUSH && (d.parent < 0 ||
                        Q.zf_ofs[d.parent] + (int64_t)P.fronts[d.parent].n * P.fronts[d.parent].n > Q.zf_size))
                        return "selinv push parent";
                    break;
                case SC_TRANS:
                    if (c.b0 < 0 || c.b < 1 || c.b0 + c.b > d.c || c.j0 < 0 || c.j1 > d.n - c.b0 - c.b)
                        return "selinv transpose range";
                    if (zf_end > Q.zf_size) return "selinv copy ZF";
                    break;
                case SC_OUT:
                    if (c.j0 < 0 || c.j1 > d.c || zf_end > Q.zf_size) return "selinv out range";
                    break;
                default: return "selinv copy kind";
            }
        }
    }
    return "";
}

}  // namespace mdgpu

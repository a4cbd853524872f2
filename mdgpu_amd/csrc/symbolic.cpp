// Host symbolic analysis (see symbolic.h).  The minimum-degree ordering and
// the relaxed amalgamation must reproduce the reference's tie-breaking, so
// those two follow the reference's sequence of steps exactly (same list
// manipulations, same hash, same scan order); everything else is written
// from the textbook definitions (Davis, "Direct Methods for Sparse Linear
// Systems", SIAM 2006, chapters 4 and 7; Liu, SIAM Review 1990).

#include "symbolic.h"

#include <algorithm>
#include <climits>
#include <cmath>
#include <map>
#include <set>

namespace mdgpu {

namespace {

inline int flip(int i) { return -i - 2; }

// Depth-first postorder of the tree given by child lists head/next starting
// at root j; appends to order[k...] and returns the new k (cs_tdfs,
// cs.cpp:456-477).  `stack` must hold at least as many entries as the tree.
int tree_dfs(int j, int k, std::vector<int>& head, const std::vector<int>& next,
             int* order, int* stack) {
    int top = 0;
    stack[0] = j;
    while (top >= 0) {
        int p = stack[top];
        int child = head[p];
        if (child == -1) {
            top--;
            order[k++] = p;
        } else {
            head[p] = next[child];
            stack[++top] = child;
        }
    }
    return k;
}

// Approximate minimum degree on the quotient graph of A+A' (cs_amd with
// order=1, cs.cpp:645-991).  The sequence of operations below is the
// reference's, so that ties resolve identically.
class MinDegreeOrdering {
   public:
    explicit MinDegreeOrdering(const Csc& upper) { build_graph(upper); }

    std::vector<int> run() {
        init_lists();
        while (eliminated_ < n_) eliminate_one();
        return postorder_assembly_tree();
    }

   private:
    int n_ = 0;
    std::vector<int> pe_;   // object pointers; flip(parent) once absorbed
    std::vector<int> iw_;   // adjacency storage with elbow room
    std::vector<int> len_, nv_, next_, head_, elen_, deg_, wf_, hhead_, last_;
    int cnz_ = 0, nzmax_ = 0, dense_ = 0;
    int mindeg_ = 0, eliminated_ = 0, lemax_ = 0, mark_ = 0;

    // pattern of A+A' without the diagonal; column order as cs_add builds it
    // (entries of A(:,j) first, then new entries of A'(:,j)), which for an
    // upper-triangular A with sorted columns is simply ascending rows.
    void build_graph(const Csc& a) {
        n_ = a.n;
        Csc at = transpose(a);
        std::vector<int> seen(n_, -1);
        pe_.assign(n_ + 1, 0);
        std::vector<int> rows;
        rows.reserve((size_t)a.nnz() * 2);
        for (int j = 0; j < n_; j++) {
            pe_[j] = (int)rows.size();
            for (int p = a.p[j]; p < a.p[j + 1]; p++) {
                int i = a.i[p];
                if (seen[i] != j) { seen[i] = j; if (i != j) rows.push_back(i); }
            }
            for (int p = at.p[j]; p < at.p[j + 1]; p++) {
                int i = at.i[p];
                if (seen[i] != j) { seen[i] = j; if (i != j) rows.push_back(i); }
            }
        }
        pe_[n_] = (int)rows.size();
        cnz_ = pe_[n_];
        // dense-row threshold: max(16, 10 sqrt(n)), capped at n-2, evaluated in
        // single precision as the shipped build does (cs.cpp:658-659, types.h:3)
        float thr = 10.0f * std::sqrt((float)n_);
        dense_ = (int)(thr < 16.0f ? 16.0f : thr);
        dense_ = std::min(n_ - 2, dense_);
        nzmax_ = cnz_ + cnz_ / 5 + 2 * n_;
        iw_.assign((size_t)std::max(nzmax_, 1), 0);
        std::copy(rows.begin(), rows.end(), iw_.begin());
    }

    // cs_wclear: restart the mark counter when it would overflow
    int fresh_mark(long long m) {
        if (m > INT_MAX) m -= 4294967296LL;   // the reference's int wraparound
        if (m < 2 || m + lemax_ > INT_MAX) {
            for (int k = 0; k < n_; k++)
                if (wf_[k] != 0) wf_[k] = 1;
            m = 2;
        }
        return (int)m;
    }

    void degree_list_insert(int i, int d) {
        if (head_[d] != -1) last_[head_[d]] = i;
        next_[i] = head_[d];
        head_[d] = i;
    }

    void init_lists() {
        const int n = n_;
        len_.assign(n + 1, 0);
        for (int k = 0; k < n; k++) len_[k] = pe_[k + 1] - pe_[k];
        nv_.assign(n + 1, 1);
        next_.assign(n + 1, -1);
        head_.assign(n + 1, -1);
        elen_.assign(n + 1, 0);
        deg_ = len_;
        wf_.assign(n + 1, 1);
        hhead_.assign(n + 1, -1);
        last_.assign(n + 1, -1);
        mark_ = fresh_mark(0);
        elen_[n] = -2;
        pe_[n] = -1;
        wf_[n] = 0;
        for (int i = 0; i < n; i++) {
            int d = deg_[i];
            if (d == 0) {                 // empty row: eliminated, a root
                elen_[i] = -2;
                eliminated_++;
                pe_[i] = -1;
                wf_[i] = 0;
            } else if (d > dense_) {      // dense row: absorbed into element n
                nv_[i] = 0;
                elen_[i] = -1;
                eliminated_++;
                pe_[i] = flip(n);
                nv_[n]++;
            } else {
                degree_list_insert(i, d);
            }
        }
    }

    void compact_storage() {
        for (int j = 0; j < n_; j++) {
            int p = pe_[j];
            if (p >= 0) { pe_[j] = iw_[p]; iw_[p] = flip(j); }
        }
        int q = 0;
        for (int p = 0; p < cnz_;) {
            int j = flip(iw_[p++]);
            if (j >= 0) {
                iw_[q] = pe_[j];
                pe_[j] = q++;
                for (int t = 0; t < len_[j] - 1; t++) iw_[q++] = iw_[p++];
            }
        }
        cnz_ = q;
    }

    void eliminate_one() {
        const int n = n_;
        int k = -1;
        for (; mindeg_ < n && (k = head_[mindeg_]) == -1; mindeg_++) {
        }
        if (next_[k] != -1) last_[next_[k]] = -1;
        head_[mindeg_] = next_[k];
        const int elenk = elen_[k];
        int nvk = nv_[k];
        eliminated_ += nvk;
        if (elenk > 0 && cnz_ + mindeg_ >= nzmax_) compact_storage();

        // new element Lk = union of the variables of k and of its elements
        int dk = 0;
        nv_[k] = -nvk;
        int p = pe_[k];
        const int pk1 = (elenk == 0) ? p : cnz_;
        int pk2 = pk1;
        for (int k1 = 1; k1 <= elenk + 1; k1++) {
            int e, pj, ln;
            if (k1 > elenk) {
                e = k; pj = p; ln = len_[k] - elenk;
            } else {
                e = iw_[p++]; pj = pe_[e]; ln = len_[e];
            }
            for (int k2 = 1; k2 <= ln; k2++) {
                int i = iw_[pj++];
                int nvi = nv_[i];
                if (nvi <= 0) continue;
                dk += nvi;
                nv_[i] = -nvi;
                iw_[pk2++] = i;
                if (next_[i] != -1) last_[next_[i]] = last_[i];
                if (last_[i] != -1) next_[last_[i]] = next_[i];
                else head_[deg_[i]] = next_[i];
            }
            if (e != k) { pe_[e] = flip(k); wf_[e] = 0; }
        }
        if (elenk != 0) cnz_ = pk2;
        deg_[k] = dk;
        pe_[k] = pk1;
        len_[k] = pk2 - pk1;
        elen_[k] = -2;

        // |Le \ Lk| for every element e adjacent to a variable of Lk
        mark_ = fresh_mark(mark_);
        for (int pk = pk1; pk < pk2; pk++) {
            int i = iw_[pk];
            int eln = elen_[i];
            if (eln <= 0) continue;
            int nvi = -nv_[i];
            int wnvi = mark_ - nvi;
            for (int q = pe_[i]; q <= pe_[i] + eln - 1; q++) {
                int e = iw_[q];
                if (wf_[e] >= mark_) wf_[e] -= nvi;
                else if (wf_[e] != 0) wf_[e] = deg_[e] + wnvi;
            }
        }

        // approximate degrees, element pruning, hashing
        for (int pk = pk1; pk < pk2; pk++) {
            int i = iw_[pk];
            int p1 = pe_[i];
            int p2 = p1 + elen_[i] - 1;
            int pn = p1;
            unsigned int h = 0;
            int d = 0;
            for (int q = p1; q <= p2; q++) {
                int e = iw_[q];
                if (wf_[e] != 0) {
                    int dext = wf_[e] - mark_;
                    if (dext > 0) {
                        d += dext;
                        iw_[pn++] = e;
                        h += e;
                    } else {
                        pe_[e] = flip(k);     // aggressive absorption
                        wf_[e] = 0;
                    }
                }
            }
            elen_[i] = pn - p1 + 1;
            int p3 = pn;
            int p4 = p1 + len_[i];
            for (int q = p2 + 1; q < p4; q++) {
                int j = iw_[q];
                int nvj = nv_[j];
                if (nvj <= 0) continue;
                d += nvj;
                iw_[pn++] = j;
                h += j;
            }
            if (d == 0) {                    // mass elimination
                pe_[i] = flip(k);
                int nvi = -nv_[i];
                dk -= nvi;
                nvk += nvi;
                eliminated_ += nvi;
                nv_[i] = 0;
                elen_[i] = -1;
            } else {
                deg_[i] = std::min(deg_[i], d);
                iw_[pn] = iw_[p3];
                iw_[p3] = iw_[p1];
                iw_[p1] = k;
                len_[i] = pn - p1 + 1;
                h %= (unsigned int)n;
                next_[i] = hhead_[h];
                hhead_[h] = i;
                last_[i] = (int)h;
            }
        }
        deg_[k] = dk;
        lemax_ = std::max(lemax_, dk);
        mark_ = fresh_mark((long long)mark_ + lemax_);

        // indistinguishable variables (same hash bucket, same adjacency)
        for (int pk = pk1; pk < pk2; pk++) {
            int i = iw_[pk];
            if (nv_[i] >= 0) continue;
            int h = last_[i];
            i = hhead_[h];
            hhead_[h] = -1;
            for (; i != -1 && next_[i] != -1; i = next_[i], mark_++) {
                int ln = len_[i];
                int eln = elen_[i];
                for (int q = pe_[i] + 1; q <= pe_[i] + ln - 1; q++) wf_[iw_[q]] = mark_;
                int jlast = i;
                for (int j = next_[i]; j != -1;) {
                    bool same = (len_[j] == ln) && (elen_[j] == eln);
                    for (int q = pe_[j] + 1; same && q <= pe_[j] + ln - 1; q++)
                        if (wf_[iw_[q]] != mark_) same = false;
                    if (same) {
                        pe_[j] = flip(i);
                        nv_[i] += nv_[j];
                        nv_[j] = 0;
                        elen_[j] = -1;
                        j = next_[j];
                        next_[jlast] = j;
                    } else {
                        jlast = j;
                        j = next_[j];
                    }
                }
            }
        }

        // restore the surviving variables of Lk to the degree lists
        int pout = pk1;
        for (int pk = pk1; pk < pk2; pk++) {
            int i = iw_[pk];
            int nvi = -nv_[i];
            if (nvi <= 0) continue;
            nv_[i] = nvi;
            int d = deg_[i] + dk - nvi;
            d = std::min(d, n - eliminated_ - nvi);
            if (head_[d] != -1) last_[head_[d]] = i;
            next_[i] = head_[d];
            last_[i] = -1;
            head_[d] = i;
            mindeg_ = std::min(mindeg_, d);
            deg_[i] = d;
            iw_[pout++] = i;
        }
        nv_[k] = nvk;
        if ((len_[k] = pout - pk1) == 0) { pe_[k] = -1; wf_[k] = 0; }
        if (elenk != 0) cnz_ = pout;
    }

    std::vector<int> postorder_assembly_tree() {
        const int n = n_;
        for (int i = 0; i < n; i++) pe_[i] = flip(pe_[i]);
        for (int j = 0; j <= n; j++) head_[j] = -1;
        for (int j = n; j >= 0; j--) {          // absorbed variables
            if (nv_[j] > 0) continue;
            if (pe_[j] < 0) continue;           // only node n can get here
            next_[j] = head_[pe_[j]];
            head_[pe_[j]] = j;
        }
        for (int e = n; e >= 0; e--) {          // elements
            if (nv_[e] <= 0) continue;
            if (pe_[e] != -1) {
                next_[e] = head_[pe_[e]];
                head_[pe_[e]] = e;
            }
        }
        std::vector<int> order(n + 1), stack(n + 1);
        for (int k = 0, i = 0; i <= n; i++)
            if (pe_[i] == -1) k = tree_dfs(i, k, head_, next_, order.data(), stack.data());
        order.resize(n);
        return order;
    }
};

// The selection sort of etree.cpp:180-192 (repeatedly swap the FIRST
// minimum of a[i..] into position i) reproduced in O(n log n): positions are
// kept in per-key ordered sets so each step finds that first minimum
// directly.  The resulting (unstable) order is identical.
void first_min_selection_sort(std::vector<int>& a, const std::vector<int>& key) {
    const int n = (int)a.size();
    if (n < 2) return;
    // a non-decreasing array is left as it is (every step picks position i)
    bool sorted = true;
    for (int t = 1; t < n && sorted; t++) sorted = key[a[t - 1]] <= key[a[t]];
    if (sorted) return;
    // The sort itself, step by step: step i swaps a[i] with the first
    // minimum of a[i..n).  A segment tree over the positions holds (key,
    // position) with done positions at +inf, so the root is that first
    // minimum; a swap updates two leaves.
    int size = 1;
    while (size < n) size <<= 1;
    const long long kInf = LLONG_MAX;
    std::vector<long long> tree(2 * size, kInf);   // key * 2^32 + position
    auto val = [&](int t) { return ((long long)key[a[t]] << 32) | (long long)t; };
    for (int t = 0; t < n; t++) tree[size + t] = val(t);
    for (int v = size - 1; v >= 1; v--) tree[v] = std::min(tree[2 * v], tree[2 * v + 1]);
    auto set_leaf = [&](int t, long long x) {
        int v = size + t;
        tree[v] = x;
        for (v >>= 1; v >= 1; v >>= 1) tree[v] = std::min(tree[2 * v], tree[2 * v + 1]);
    };
    for (int i = 0; i < n; i++) {
        const int x = (int)(tree[1] & 0xffffffffLL);
        if (x != i) {
            std::swap(a[i], a[x]);
            set_leaf(x, val(x));
        }
        set_leaf(i, kInf);
    }
}

}  // namespace

Csc transpose(const Csc& a) {
    Csc c;
    c.n = a.n;
    const int n = a.n;
    c.p.assign(n + 1, 0);
    c.i.resize(a.nnz());
    std::vector<int> w(n + 1, 0);
    for (int p = 0; p < a.nnz(); p++) w[a.i[p]]++;
    int s = 0;
    for (int r = 0; r < n; r++) { c.p[r] = s; s += w[r]; w[r] = c.p[r]; }
    c.p[n] = s;
    for (int j = 0; j < n; j++)
        for (int p = a.p[j]; p < a.p[j + 1]; p++) c.i[w[a.i[p]]++] = j;
    return c;
}

std::vector<int> amd_order(const Csc& upper) {
    if (upper.n == 0) return {};
    MinDegreeOrdering md(upper);
    return md.run();
}

// Liu's elimination tree with path compression on the upper triangle.
std::vector<int> etree(const Csc& a) {
    const int n = a.n;
    std::vector<int> parent(n, -1), anc(n, -1);
    for (int k = 0; k < n; k++) {
        for (int p = a.p[k]; p < a.p[k + 1]; p++) {
            int i = a.i[p];
            while (i != -1 && i < k) {
                int up = anc[i];
                anc[i] = k;
                if (up == -1) parent[i] = k;
                i = up;
            }
        }
    }
    return parent;
}

// Postorder with children visited in increasing index order, roots in
// increasing order.
std::vector<int> postorder(const std::vector<int>& parent) {
    const int n = (int)parent.size();
    std::vector<int> head(n, -1), next(n, -1), post(n), stack(n);
    for (int j = n - 1; j >= 0; j--) {
        if (parent[j] == -1) continue;
        next[j] = head[parent[j]];
        head[parent[j]] = j;
    }
    int k = 0;
    for (int j = 0; j < n; j++)
        if (parent[j] == -1) k = tree_dfs(j, k, head, next, post.data(), stack.data());
    return post;
}

// Column counts of L (diagonal included) by the row-subtree / skeleton
// method of Gilbert, Ng and Peyton (1994).
std::vector<int> column_counts(const Csc& a, const std::vector<int>& parent,
                               const std::vector<int>& post) {
    const int n = a.n;
    std::vector<int> delta(n), first(n, -1), maxfirst(n, -1), prevleaf(n, -1), anc(n);
    for (int k = 0; k < n; k++) {
        int j = post[k];
        delta[j] = (first[j] == -1) ? 1 : 0;
        for (; j != -1 && first[j] == -1; j = parent[j]) first[j] = k;
    }
    Csc at = transpose(a);
    for (int i = 0; i < n; i++) anc[i] = i;
    for (int k = 0; k < n; k++) {
        int j = post[k];
        if (parent[j] != -1) delta[parent[j]]--;
        for (int p = at.p[j]; p < at.p[j + 1]; p++) {
            int i = at.i[p];
            if (i <= j || first[j] <= maxfirst[i]) continue;   // j not a leaf
            maxfirst[i] = first[j];
            int jprev = prevleaf[i];
            prevleaf[i] = j;
            delta[j]++;                                        // A(i,j) in skeleton
            if (jprev != -1) {
                int q = jprev;
                while (q != anc[q]) q = anc[q];
                for (int s = jprev; s != q;) { int up = anc[s]; anc[s] = q; s = up; }
                delta[q]--;                                    // overlap at lca
            }
        }
        if (parent[j] != -1) anc[j] = parent[j];
    }
    for (int j = 0; j < n; j++)
        if (parent[j] != -1) delta[parent[j]] += delta[j];
    return delta;
}

namespace {

// lower pattern -> full symmetric pattern with diagonal, sorted rows
Csc full_from_lower(const Csc& low) {
    const int n = low.n;
    Csc up = transpose(low);
    Csc f;
    f.n = n;
    f.p.assign(n + 1, 0);
    f.i.reserve((size_t)low.nnz() * 2);
    for (int j = 0; j < n; j++) {
        f.p[j] = (int)f.i.size();
        for (int p = up.p[j]; p < up.p[j + 1]; p++)
            if (up.i[p] != j) f.i.push_back(up.i[p]);
        for (int p = low.p[j]; p < low.p[j + 1]; p++) f.i.push_back(low.i[p]);
    }
    f.p[n] = (int)f.i.size();
    return f;
}

// Symmetric permutation of a lower triangle: entry (i,j) goes to
// (max, min)(pinv[i], pinv[j]); columns sorted.  src[] tracks, for every
// output entry, which input entry it came from (composed with src_in).
Csc permute_lower(const Csc& low, const std::vector<int>& pinv, const std::vector<int>& src_in,
                  std::vector<int>& src_out) {
    const int n = low.n;
    const int nz = low.nnz();
    std::vector<int> cnt(n + 1, 0);
    std::vector<int> r(nz), c(nz);
    for (int j = 0; j < n; j++)
        for (int p = low.p[j]; p < low.p[j + 1]; p++) {
            int a = pinv[low.i[p]], b = pinv[j];
            r[p] = std::max(a, b);
            c[p] = std::min(a, b);
            cnt[c[p]]++;
        }
    // counting sort by (col, row): first by row, then stable by column
    std::vector<int> byrow(nz), rp(n + 1, 0);
    for (int p = 0; p < nz; p++) rp[r[p] + 1]++;
    for (int t = 0; t < n; t++) rp[t + 1] += rp[t];
    for (int p = 0; p < nz; p++) byrow[rp[r[p]]++] = p;
    Csc out;
    out.n = n;
    out.p.assign(n + 1, 0);
    for (int t = 0; t < n; t++) out.p[t + 1] = out.p[t] + cnt[t];
    out.i.resize(nz);
    src_out.resize(nz);
    std::vector<int> fill(out.p.begin(), out.p.end() - 1);
    for (int t = 0; t < nz; t++) {
        int p = byrow[t];
        int q = fill[c[p]]++;
        out.i[q] = r[p];
        src_out[q] = src_in[p];
    }
    return out;
}

// GetETree (etree.cpp:12-48): one node per column, children in increasing
// index order, list in postorder.  The reference stops on a forest
// (etree.cpp:26-33); here a forest is accepted unless `single_tree` is set,
// each tree being factored on its own (reducible matrices, batches).
std::string plain_tree(const std::vector<int>& parent, const std::vector<int>& post,
                       const std::vector<int>& count, bool single_tree, std::vector<PartNode>& nodes,
                       std::vector<int>& node_of_col) {
    const int n = (int)parent.size();
    int roots = 0;
    for (int j = 0; j < n; j++) roots += parent[j] == -1;
    if (single_tree && roots != 1)
        return "elimination tree is a forest (" + std::to_string(roots) + " roots)";
    std::vector<int> pos(n);
    for (int k = 0; k < n; k++) pos[post[k]] = k;
    nodes.assign(n, PartNode());
    node_of_col.assign(n, -1);
    for (int k = 0; k < n; k++) {
        int j = post[k];
        nodes[k].first = j;
        nodes[k].ncols = 1;
        nodes[k].rows = count[j];
        nodes[k].cols = {j};
        node_of_col[j] = k;
    }
    for (int j = 0; j < n; j++)      // increasing child index
        if (parent[j] != -1) nodes[pos[parent[j]]].kids.push_back(pos[j]);
    return "";
}

// GetSuperETree (etree.cpp:50-118): fundamental supernodes.  A node absorbs
// its last (largest-index) child when that child is the previous column and
// has exactly one more row.  Children are kept in processing order.
void supernodal_tree(const std::vector<PartNode>& plain, std::vector<PartNode>& out) {
    const int n = (int)plain.size();
    std::vector<int> super_of(n, -1);
    out.clear();
    for (int t = 0; t < n; t++) {
        const PartNode& node = plain[t];
        bool merge = false;
        if (!node.kids.empty()) {
            const PartNode& last = plain[node.kids.back()];
            merge = last.rows == node.rows + 1 && last.first == node.first - 1;
        }
        if (merge) {
            int s = super_of[node.kids.back()];
            PartNode& sn = out[s];
            sn.ncols += 1;
            sn.rows -= 1;
            sn.cols.push_back(node.first);
            for (size_t h = 0; h + 1 < node.kids.size(); h++) sn.kids.push_back(super_of[node.kids[h]]);
            std::sort(sn.kids.begin(), sn.kids.end());
            super_of[t] = s;
        } else {
            PartNode sn;
            sn.first = node.first;
            sn.ncols = 1;
            sn.rows = node.rows;
            sn.cols = {node.first};
            for (int c : node.kids) sn.kids.push_back(super_of[c]);
            std::sort(sn.kids.begin(), sn.kids.end());
            super_of[t] = (int)out.size();
            out.push_back(sn);
        }
    }
}

// GetRelaxedETree (etree.cpp:194-342): greedy amalgamation of children under
// a zero budget, column-count padding, per-node first-minimum selection sort
// by padded count, and a depth-first renumbering that yields the new
// permutation.  Node lists/children are reproduced in the reference's order.
std::string relaxed_tree(const std::vector<PartNode>& plain, int max_zeros, std::vector<int>& count,
                         std::vector<int>& perm, std::vector<PartNode>& out) {
    const int n = (int)plain.size();
    struct R { int nodo, ncols, rows; long long zeros; std::vector<int> cols, kids; };
    std::vector<R> pool;
    pool.reserve(n);
    std::vector<int> stack;
    stack.reserve(n);
    for (int t = 0; t < n; t++) {
        const PartNode& tree = plain[t];
        const int nh = (int)tree.kids.size();
        const int base = (int)stack.size() - nh;
        std::vector<char> merged(nh, 0);
        long long zeros = tree.zeros;
        R root;
        root.nodo = tree.first;
        root.ncols = tree.ncols;
        root.rows = tree.rows;
        root.zeros = 0;
        root.cols = {tree.first};
        long long cols = 1;
        while (true) {
            int best = -1;
            long long best_c = -1;
            for (int h = 0; h < nh; h++) {
                const R& kid = pool[stack[base + h]];
                long long c = ((long long)tree.rows + cols - kid.rows) * kid.ncols + kid.zeros;
                if (!merged[h] && c <= max_zeros && (best_c == -1 || c < best_c)) {
                    best_c = c;
                    best = h;
                }
            }
            if (best < 0) break;
            const R& kid = pool[stack[base + best]];
            if (best_c > INT_MAX || zeros + best_c > INT_MAX)
                return "relaxed amalgamation: zero count overflows int32";
            merged[best] = 1;
            zeros += best_c;
            cols += kid.ncols;
            const int pad = (int)((best_c - kid.zeros) / kid.ncols);
            for (int c : kid.cols) count[c] += pad;
        }
        root.zeros = zeros;
        for (int h = 0; h < nh; h++) {
            const R& kid = pool[stack[base + h]];
            if (!merged[h]) {
                root.kids.push_back(stack[base + h]);
            } else {
                root.cols.insert(root.cols.end(), kid.cols.begin(), kid.cols.end());
                root.ncols += kid.ncols;
                root.kids.insert(root.kids.end(), kid.kids.begin(), kid.kids.end());
            }
        }
        for (int c : root.cols) root.nodo = std::min(root.nodo, c);
        first_min_selection_sort(root.cols, count);
        stack.resize(base);
        stack.push_back((int)pool.size());
        pool.push_back(std::move(root));
    }
    // depth-first renumbering from the root (etree.cpp:165-178, 331); for a
    // forest every remaining stack entry is a root, renumbered in order
    perm.assign(n, -1);
    out.clear();
    std::vector<int> index_of(pool.size(), -1);
    int c = 0;
    struct Frame { int node; size_t next_kid; };
    std::vector<Frame> dfs;
    for (size_t root = 0; root < stack.size(); root++) {
    dfs.push_back({stack[root], 0});
    while (!dfs.empty()) {
        Frame& f = dfs.back();
        R& r = pool[f.node];
        if (f.next_kid < r.kids.size()) {
            int kid = r.kids[f.next_kid++];
            dfs.push_back({kid, 0});
            continue;
        }
        PartNode pn;
        pn.first = c;
        pn.ncols = r.ncols;
        pn.rows = r.rows;
        pn.zeros = (int)r.zeros;
        pn.cols = r.cols;
        for (int kid : r.kids) pn.kids.push_back(index_of[kid]);
        for (int t = r.ncols - 1; t >= 0; t--) perm[c++] = r.cols[t];
        index_of[f.node] = (int)out.size();
        out.push_back(std::move(pn));
        dfs.pop_back();
    }
    }
    if (c != n) return "relaxed tree does not cover every column (forest?)";
    return "";
}

}  // namespace

void supernodal_partition(const std::vector<PartNode>& plain, std::vector<PartNode>& out) {
    supernodal_tree(plain, out);
}

std::string relaxed_partition(const std::vector<PartNode>& plain, int max_zeros, std::vector<int>& count,
                              std::vector<int>& perm, std::vector<PartNode>& out) {
    return relaxed_tree(plain, max_zeros, count, perm, out);
}

Csc full_pattern(const Csc& low) { return full_from_lower(low); }

void union_rows(std::vector<int>& rows, std::vector<int>& stamp, int tag, const int* list, int64_t len) {
    for (int64_t t = 0; t < len; t++) {
        const int r = list[t];
        if (stamp[r] != tag) {
            stamp[r] = tag;
            rows.push_back(r);
        }
    }
}

void sort_union_tail(std::vector<int>& rows, size_t from) { std::sort(rows.begin() + from, rows.end()); }

bool relative_map(const int* super, int n, const int* sub, int m, int* map) {
    for (int t = 0; t < m; t++) {
        const int* it = std::lower_bound(super, super + n, sub[t]);
        if (it == super + n || *it != sub[t]) return false;
        map[t] = (int)(it - super);
    }
    return true;
}

// Front row lists: pivot columns, rows of A below them, and the update
// rows of every child -- the union MatrizGetFk_3 and VectorMerge build per
// factorization (mdgpu.cu:180-291, 1456-1508).  With check_counts the
// front sizes must agree with count_final (spL->p); otherwise count_final
// and lp are derived from the fronts.
std::string build_fronts(Symbolic& S, bool check_counts) {
    const int n = S.n;
    const int nn = (int)S.nodes.size();
    S.front_parent.assign(nn, -1);
    for (int f = 0; f < nn; f++)
        for (int kid : S.nodes[f].kids) {
            if (kid < 0 || kid >= f) return "node list is not in postorder";
            S.front_parent[kid] = f;
        }
    S.q_off.assign(nn + 1, 0);
    S.q_all.clear();
    std::vector<int> stamp(n, -1);
    std::vector<int> rows;
    for (int f = 0; f < nn; f++) {
        const PartNode& nd = S.nodes[f];
        const int k = nd.first, c = nd.ncols;
        if (k < 0 || k + c > n) return "node columns out of range";
        rows.clear();
        for (int j = k; j < k + c; j++) { stamp[j] = f; rows.push_back(j); }
        for (int j = k; j < k + c; j++)
            union_rows(rows, stamp, f, S.a_low.i.data() + S.a_low.p[j], S.a_low.p[j + 1] - S.a_low.p[j]);
        for (int kid : nd.kids) {
            const int kc = S.nodes[kid].ncols;
            union_rows(rows, stamp, f, S.q_all.data() + S.q_off[kid] + kc, S.q_off[kid + 1] - S.q_off[kid] - kc);
        }
        sort_union_tail(rows, c);
        if ((int)rows.size() > c && rows[c] < k + c) return "front rows precede the pivot block";
        const int nf = (int)rows.size();
        if (!check_counts)
            for (int j = 0; j < c; j++) S.count_final[k + j] = nf - j;
        for (int j = 0; j < c; j++)
            if (S.count_final[k + j] != nf - j)
                return "front of node " + std::to_string(f) + " has " + std::to_string(nf) +
                       " rows but column " + std::to_string(k + j) + " count is " +
                       std::to_string(S.count_final[k + j]);
        S.q_all.insert(S.q_all.end(), rows.begin(), rows.end());
        S.q_off[f + 1] = (int64_t)S.q_all.size();
    }
    return "";
}

namespace {

// Caller's lower triangle: sorted unique rows, diagonal present.
std::string read_lower(int n, const int* Cp, const int* Ci, Csc& low) {
    if (n <= 0) return "matrix order must be positive";
    low.n = n;
    low.p.assign(Cp, Cp + n + 1);
    if (low.p[0] != 0) return "column pointers must start at 0";
    for (int j = 0; j < n; j++)
        if (low.p[j + 1] < low.p[j]) return "column pointers must be nondecreasing";
    const int nz = low.p[n];
    low.i.assign(Ci, Ci + nz);
    for (int j = 0; j < n; j++) {
        bool has_diag = false;
        for (int p = low.p[j]; p < low.p[j + 1]; p++) {
            int i = low.i[p];
            if (i < j || i >= n) return "entry outside the lower triangle in column " + std::to_string(j);
            if (p > low.p[j] && low.i[p - 1] >= i) return "rows must be sorted and unique in column " + std::to_string(j);
            has_diag |= (i == j);
        }
        if (!has_diag) return "missing diagonal entry in column " + std::to_string(j);
    }
    return "";
}

}  // namespace

namespace {

// Steps of the analysis that do not depend on the node partition: the
// fill-reducing ordering, elimination tree, postorder, column counts and
// the plain node tree.
struct AnalysisPrefix {
    int n = 0, nnz_in = 0;
    std::vector<int> amd_perm, parent, post, count;
    Csc a_low;                // AMD-permuted lower triangle
    std::vector<int> src;     // a_low entry -> caller entry
    std::vector<PartNode> plain;
};

std::string analysis_prefix(int n, const int* Cp, const int* Ci, const AnalyzeOptions& opts, AnalysisPrefix& A) {
    A.n = n;
    Csc low;
    std::string verr = read_lower(n, Cp, Ci, low);
    if (!verr.empty()) return verr;
    const int nz = low.p[n];
    A.nnz_in = nz;
    std::vector<int> ident(nz);
    for (int p = 0; p < nz; p++) ident[p] = p;

    // 1. fill-reducing ordering on the upper triangle (mdgpu.cu:1745-1773)
    if (opts.use_amd) {
        A.amd_perm = amd_order(transpose(low));
        std::vector<int> pinv(n);
        for (int k = 0; k < n; k++) pinv[A.amd_perm[k]] = k;
        A.a_low = permute_lower(low, pinv, ident, A.src);
    } else {
        A.amd_perm.resize(n);
        for (int k = 0; k < n; k++) A.amd_perm[k] = k;
        A.a_low = low;
        A.src = ident;
    }
    Csc full = full_from_lower(A.a_low);

    // 2. etree, postorder, column counts (mdgpu.cu:1783-1785)
    A.parent = etree(full);
    A.post = postorder(A.parent);
    A.count = column_counts(full, A.parent, A.post);
    std::vector<int> node_of_col;
    return plain_tree(A.parent, A.post, A.count, opts.single_tree, A.plain, node_of_col);
}

// 3. node partition (mdgpu.cu:1803-1853) and the fronts, on a prefix
std::string analysis_finish(const AnalysisPrefix& A, const AnalyzeOptions& opts, Symbolic& S) {
    const int n = A.n;
    S = Symbolic();
    S.n = n;
    S.opts = opts;
    S.nnz_in = A.nnz_in;
    S.amd_perm = A.amd_perm;
    S.parent = A.parent;
    S.post = A.post;
    S.count = A.count;
    Csc a_low = A.a_low;
    std::vector<int> src = A.src;
    S.perm = S.amd_perm;
    S.count_final = S.count;
    if (opts.partition == PART_PLAIN) {
        S.nodes = A.plain;
    } else if (opts.partition == PART_SUPERNODAL) {
        supernodal_tree(A.plain, S.nodes);
    } else if (opts.partition == PART_RELAXED) {
        std::vector<int> padded = S.count;
        std::string err = relaxed_tree(A.plain, opts.max_zeros, padded, S.relaxed_perm, S.nodes);
        if (!err.empty()) return err;
        std::vector<int> rpinv(n);
        for (int k = 0; k < n; k++) rpinv[S.relaxed_perm[k]] = k;
        std::vector<int> src2;
        a_low = permute_lower(a_low, rpinv, src, src2);
        src.swap(src2);
        for (int k = 0; k < n; k++) S.count_final[k] = padded[S.relaxed_perm[k]];
        for (int k = 0; k < n; k++) S.perm[k] = S.amd_perm[S.relaxed_perm[k]];
    } else {
        return "unknown partition mode";
    }
    S.a_low = std::move(a_low);
    S.a_src = std::move(src);
    S.pinv.resize(n);
    for (int k = 0; k < n; k++) S.pinv[S.perm[k]] = k;
    S.lp.assign(n + 1, 0);
    for (int k = 0; k < n; k++) S.lp[k + 1] = S.lp[k] + S.count_final[k];
    S.lnz = S.lp[n];
    for (int k = 0; k < n; k++) {
        S.flops_counts += (double)S.count[k] * S.count[k];
        S.flops_padded += (double)S.count_final[k] * S.count_final[k] + 1.0;
    }
    return build_fronts(S, true);
}

}  // namespace

std::string analyze(int n, const int* Cp, const int* Ci, const AnalyzeOptions& opts, Symbolic& S) {
    AnalysisPrefix A;
    S = Symbolic();
    std::string err = analysis_prefix(n, Cp, Ci, opts, A);
    if (!err.empty()) return err;
    return analysis_finish(A, opts, S);
}

std::string analyze_budgets(int n, const int* Cp, const int* Ci, const AnalyzeOptions& opts,
                            const std::vector<int>& max_zeros, std::vector<Symbolic>& out) {
    AnalysisPrefix A;
    std::string err = analysis_prefix(n, Cp, Ci, opts, A);
    if (!err.empty()) return err;
    out.assign(max_zeros.size(), Symbolic());
    for (size_t k = 0; k < max_zeros.size(); k++) {
        AnalyzeOptions o = opts;
        o.partition = PART_RELAXED;
        o.max_zeros = max_zeros[k];
        err = analysis_finish(A, o, out[k]);
        if (!err.empty()) return err;
    }
    return "";
}

std::string analyze_with_nodes(int n, const int* Cp, const int* Ci, const std::vector<PartNode>& nodes,
                               Symbolic& S) {
    S = Symbolic();
    S.n = n;
    S.opts.use_amd = false;
    S.opts.partition = PART_SUPERNODAL;
    Csc low;
    std::string err = read_lower(n, Cp, Ci, low);
    if (!err.empty()) return err;
    const int nz = low.p[n];
    S.nnz_in = nz;
    S.a_src.resize(nz);
    for (int p = 0; p < nz; p++) S.a_src[p] = p;
    S.amd_perm.resize(n);
    for (int k = 0; k < n; k++) S.amd_perm[k] = k;
    S.perm = S.amd_perm;
    S.pinv = S.amd_perm;
    Csc full = full_from_lower(low);
    S.parent = etree(full);
    S.post = postorder(S.parent);
    S.count = column_counts(full, S.parent, S.post);
    S.a_low = std::move(low);
    std::vector<int> covered(n, 0);
    for (const PartNode& nd : nodes) {
        if (nd.ncols < 1 || nd.first < 0 || nd.first + nd.ncols > n) return "node columns out of range";
        for (int j = nd.first; j < nd.first + nd.ncols; j++) covered[j]++;
    }
    for (int j = 0; j < n; j++)
        if (covered[j] != 1) return "column " + std::to_string(j) + " is not in exactly one node";
    S.nodes = nodes;
    S.count_final.assign(n, 0);
    err = build_fronts(S, false);
    if (!err.empty()) return err;
    S.lp.assign(n + 1, 0);
    for (int k = 0; k < n; k++) S.lp[k + 1] = S.lp[k] + S.count_final[k];
    S.lnz = S.lp[n];
    for (int k = 0; k < n; k++) {
        S.flops_counts += (double)S.count[k] * S.count[k];
        S.flops_padded += (double)S.count_final[k] * S.count_final[k] + 1.0;
    }
    return "";
}

void merge_symbolic(const std::vector<Symbolic>& parts, Symbolic& out) {
    out = Symbolic();
    if (parts.empty()) return;
    out.opts = parts[0].opts;
    int n = 0, nnz_in = 0, nnodes = 0;
    int64_t nq = 0;
    for (const Symbolic& P : parts) {
        n += P.n;
        nnz_in += P.nnz_in;
        nnodes += (int)P.nodes.size();
        nq += (int64_t)P.q_all.size();
    }
    out.n = n;
    out.nnz_in = nnz_in;
    const bool relaxed = !parts[0].relaxed_perm.empty();
    auto shift = [](std::vector<int>& dst, const std::vector<int>& src, int off, bool keep_neg) {
        for (int v : src) dst.push_back(keep_neg && v < 0 ? v : v + off);
    };
    out.a_low.n = n;
    out.a_low.p.assign(1, 0);
    out.q_off.assign(1, 0);
    out.q_all.reserve(nq);
    out.nodes.reserve(nnodes);
    int col = 0, ent = 0, node = 0;
    for (const Symbolic& P : parts) {
        shift(out.amd_perm, P.amd_perm, col, false);
        shift(out.parent, P.parent, col, true);
        shift(out.post, P.post, col, false);
        out.count.insert(out.count.end(), P.count.begin(), P.count.end());
        if (relaxed) shift(out.relaxed_perm, P.relaxed_perm, col, false);
        shift(out.perm, P.perm, col, false);
        shift(out.pinv, P.pinv, col, false);
        out.count_final.insert(out.count_final.end(), P.count_final.begin(), P.count_final.end());
        const int base = out.a_low.p.back();
        for (int k = 1; k <= P.n; k++) out.a_low.p.push_back(base + P.a_low.p[k]);
        shift(out.a_low.i, P.a_low.i, col, false);
        shift(out.a_src, P.a_src, ent, false);
        for (const PartNode& nd : P.nodes) {
            PartNode m = nd;
            m.first += col;
            for (int& c : m.cols) c += col;
            for (int& k : m.kids) k += node;
            out.nodes.push_back(std::move(m));
        }
        shift(out.front_parent, P.front_parent, node, true);
        const int64_t qbase = (int64_t)out.q_all.size();
        for (size_t f = 1; f < P.q_off.size(); f++) out.q_off.push_back(qbase + P.q_off[f]);
        shift(out.q_all, P.q_all, col, false);
        out.flops_counts += P.flops_counts;
        out.flops_padded += P.flops_padded;
        col += P.n;
        ent += P.nnz_in;
        node += (int)P.nodes.size();
    }
    out.lnz = 0;
    for (const Symbolic& P : parts) out.lnz += P.lnz;
    // CSparse column pointers are int32: only when the batch's L fits
    if (out.lnz <= INT32_MAX) {
        out.lp.assign(n + 1, 0);
        for (int k = 0; k < n; k++) out.lp[k + 1] = out.lp[k] + out.count_final[k];
    }
}

}  // namespace mdgpu

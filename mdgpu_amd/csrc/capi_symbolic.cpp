// C ABI: symbolic analysis entry points (see include/mdgpu_c.h).
#include <algorithm>
#include <atomic>
#include <cstring>
#include <new>
#include <thread>

#include "common.h"

namespace mdgpu {

static thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}
void clear_error() { g_last_error.clear(); }

}  // namespace mdgpu

using mdgpu::fail;

// levels (height above the leaves) and the largest front
static void finish_handle(mdgpu_symbolic* h) {
    const auto& S = h->S;
    if (h->seg_col.empty()) {
        h->seg_col = {0, S.n};
        h->seg_nnz = {0, S.nnz_in};
    }
    const int nn = (int)S.nodes.size();
    h->level.assign(nn, 0);
    for (int f = 0; f < nn; f++) {
        int nf = (int)(S.q_off[f + 1] - S.q_off[f]);
        h->max_front = std::max(h->max_front, nf);
        for (int kid : S.nodes[f].kids) h->level[f] = std::max(h->level[f], h->level[kid] + 1);
        h->nlevels = std::max(h->nlevels, h->level[f] + 1);
    }
}

void mdgpu_finish_symbolic_handle(mdgpu_symbolic* h) { finish_handle(h); }

extern "C" const char* mdgpu_last_error(void) { return mdgpu::g_last_error.c_str(); }

extern "C" int mdgpu_analyze(int n, const int32_t* Cp, const int32_t* Ci,
                             const mdgpu_analyze_opts* opts, mdgpu_symbolic** sym) {
    mdgpu::clear_error();
    if (!Cp || !Ci || !sym) return fail(MDGPU_ERR_ARG, "mdgpu_analyze: null pointer");
    *sym = nullptr;
    mdgpu::AnalyzeOptions o;
    if (opts) {
        o.use_amd = opts->use_amd != 0;
        if (opts->partition < MDGPU_PART_PLAIN || opts->partition > MDGPU_PART_RELAXED)
            return fail(MDGPU_ERR_ARG, "mdgpu_analyze: unknown partition mode");
        o.partition = (mdgpu::Partition)opts->partition;
        o.max_zeros = opts->max_zeros;
        if (o.max_zeros < 0) return fail(MDGPU_ERR_ARG, "mdgpu_analyze: max_zeros must be >= 0");
    }
    mdgpu_symbolic* h = new (std::nothrow) mdgpu_symbolic();
    if (!h) return fail(MDGPU_ERR_NOMEM, "mdgpu_analyze: out of host memory");
    try {
        std::string err = mdgpu::analyze(n, Cp, Ci, o, h->S);
        if (!err.empty()) {
            delete h;
            return fail(MDGPU_ERR_ARG, "mdgpu_analyze: " + err);
        }
    } catch (const std::bad_alloc&) {
        delete h;
        return fail(MDGPU_ERR_NOMEM, "mdgpu_analyze: out of host memory");
    }
    finish_handle(h);
    *sym = h;
    return MDGPU_OK;
}

extern "C" int mdgpu_analyze_nodes(int n, const int32_t* Cp, const int32_t* Ci, int nnodes, const int32_t* first,
                                   const int32_t* ncols, const int32_t* kid_ptr, const int32_t* kids,
                                   mdgpu_symbolic** sym) {
    mdgpu::clear_error();
    if (!Cp || !Ci || !first || !ncols || !kid_ptr || !sym || nnodes <= 0)
        return fail(MDGPU_ERR_ARG, "mdgpu_analyze_nodes: invalid argument");
    *sym = nullptr;
    std::vector<mdgpu::PartNode> nodes(nnodes);
    for (int f = 0; f < nnodes; f++) {
        nodes[f].first = first[f];
        nodes[f].ncols = ncols[f];
        nodes[f].rows = 0;
        for (int t = kid_ptr[f]; t < kid_ptr[f + 1]; t++) nodes[f].kids.push_back(kids[t]);
        std::sort(nodes[f].kids.begin(), nodes[f].kids.end());
        for (int j = 0; j < ncols[f]; j++) nodes[f].cols.push_back(first[f] + j);
    }
    mdgpu_symbolic* h = new (std::nothrow) mdgpu_symbolic();
    if (!h) return fail(MDGPU_ERR_NOMEM, "mdgpu_analyze_nodes: out of host memory");
    try {
        std::string err = mdgpu::analyze_with_nodes(n, Cp, Ci, nodes, h->S);
        if (!err.empty()) {
            delete h;
            return fail(MDGPU_ERR_ARG, "mdgpu_analyze_nodes: " + err);
        }
    } catch (const std::bad_alloc&) {
        delete h;
        return fail(MDGPU_ERR_NOMEM, "mdgpu_analyze_nodes: out of host memory");
    }
    for (int f = 0; f < nnodes; f++)
        h->S.nodes[f].rows = (int)(h->S.q_off[f + 1] - h->S.q_off[f]);
    finish_handle(h);
    *sym = h;
    return MDGPU_OK;
}

extern "C" int mdgpu_analyze_batch(int nmat, const int32_t* n, const int32_t* const* Cp, const int32_t* const* Ci,
                                   const mdgpu_analyze_opts* opts, int nthreads, mdgpu_symbolic** sym) {
    mdgpu::clear_error();
    if (!n || !Cp || !Ci || !sym || nmat <= 0) return fail(MDGPU_ERR_ARG, "mdgpu_analyze_batch: invalid argument");
    *sym = nullptr;
    mdgpu::AnalyzeOptions o;
    if (opts) {
        o.use_amd = opts->use_amd != 0;
        if (opts->partition < MDGPU_PART_PLAIN || opts->partition > MDGPU_PART_RELAXED)
            return fail(MDGPU_ERR_ARG, "mdgpu_analyze_batch: unknown partition mode");
        o.partition = (mdgpu::Partition)opts->partition;
        o.max_zeros = opts->max_zeros;
        if (o.max_zeros < 0) return fail(MDGPU_ERR_ARG, "mdgpu_analyze_batch: max_zeros must be >= 0");
    }
    int64_t ntot = 0, nnztot = 0;
    for (int k = 0; k < nmat; k++) {
        if (!Cp[k] || !Ci[k] || n[k] <= 0) return fail(MDGPU_ERR_ARG, "mdgpu_analyze_batch: invalid matrix");
        ntot += n[k];
        nnztot += Cp[k][n[k]];
    }
    if (ntot > INT32_MAX || nnztot > INT32_MAX)
        return fail(MDGPU_ERR_ARG, "mdgpu_analyze_batch: batch exceeds int32 columns or entries");
    std::vector<mdgpu::Symbolic> parts(nmat);
    std::vector<std::string> errs(nmat);
    // independent analyses on a pool of host threads (each is the
    // single-matrix analysis, so each part is bit-exact on its own)
    std::atomic<int> next(0);
    bool oom = false;
    auto work = [&]() {
        for (int k; (k = next.fetch_add(1)) < nmat;) {
            try {
                errs[k] = mdgpu::analyze(n[k], Cp[k], Ci[k], o, parts[k]);
            } catch (const std::bad_alloc&) {
                errs[k] = "out of host memory";
                oom = true;
            }
        }
    };
    int nt = nthreads > 0 ? nthreads : (int)std::max(1u, std::thread::hardware_concurrency());
    nt = std::min(nt, nmat);
    std::vector<std::thread> pool;
    for (int t = 1; t < nt; t++) pool.emplace_back(work);
    work();
    for (auto& th : pool) th.join();
    for (int k = 0; k < nmat; k++)
        if (!errs[k].empty())
            return fail(oom ? MDGPU_ERR_NOMEM : MDGPU_ERR_ARG,
                        "mdgpu_analyze_batch: matrix " + std::to_string(k) + ": " + errs[k]);
    mdgpu_symbolic* h = new (std::nothrow) mdgpu_symbolic();
    if (!h) return fail(MDGPU_ERR_NOMEM, "mdgpu_analyze_batch: out of host memory");
    try {
        mdgpu::merge_symbolic(parts, h->S);
    } catch (const std::bad_alloc&) {
        delete h;
        return fail(MDGPU_ERR_NOMEM, "mdgpu_analyze_batch: out of host memory");
    }
    h->seg_col.assign(1, 0);
    h->seg_nnz.assign(1, 0);
    for (int k = 0; k < nmat; k++) {
        h->seg_col.push_back(h->seg_col.back() + n[k]);
        h->seg_nnz.push_back(h->seg_nnz.back() + Cp[k][n[k]]);
    }
    finish_handle(h);
    *sym = h;
    return MDGPU_OK;
}

extern "C" int mdgpu_symbolic_destroy(mdgpu_symbolic* sym) {
    delete sym;
    return MDGPU_OK;
}

static int64_t nodes_flat_len(const mdgpu::Symbolic& S) {
    int64_t len = 0;
    for (const auto& nd : S.nodes) len += 5 + (int64_t)nd.kids.size();
    return len;
}

extern "C" int mdgpu_symbolic_sizes(const mdgpu_symbolic* sym, int64_t* out) {
    if (!sym || !out) return fail(MDGPU_ERR_ARG, "mdgpu_symbolic_sizes: null pointer");
    const auto& S = sym->S;
    out[MDGPU_SYM_N] = S.n;
    out[MDGPU_SYM_NNODES] = (int64_t)S.nodes.size();
    out[MDGPU_SYM_LNZ] = S.lnz;
    out[MDGPU_SYM_NNZ_A] = S.nnz_in;
    out[MDGPU_SYM_QLEN] = (int64_t)S.q_all.size();
    out[MDGPU_SYM_MAXFRONT] = sym->max_front;
    out[MDGPU_SYM_NLEVELS] = sym->nlevels;
    out[MDGPU_SYM_NODES_FLAT] = nodes_flat_len(S);
    out[MDGPU_SYM_NSEG] = (int64_t)sym->seg_col.size() - 1;
    return MDGPU_OK;
}

extern "C" int mdgpu_symbolic_flops(const mdgpu_symbolic* sym, double* out) {
    if (!sym || !out) return fail(MDGPU_ERR_ARG, "mdgpu_symbolic_flops: null pointer");
    out[0] = sym->S.flops_counts;
    out[1] = sym->S.flops_padded;
    return MDGPU_OK;
}

template <class V>
static int copy_out(const V& v, int32_t* dst, int64_t cap, const char* what) {
    if ((int64_t)v.size() > cap)
        return fail(MDGPU_ERR_ARG, std::string("mdgpu_symbolic_array: capacity too small for ") + what);
    for (size_t t = 0; t < v.size(); t++) {
        if ((int64_t)v[t] > INT32_MAX) return fail(MDGPU_ERR_ARG, std::string(what) + " exceeds int32");
        dst[t] = (int32_t)v[t];
    }
    return MDGPU_OK;
}

extern "C" int mdgpu_symbolic_array(const mdgpu_symbolic* sym, int which, int32_t* dst, int64_t cap) {
    if (!sym || !dst) return fail(MDGPU_ERR_ARG, "mdgpu_symbolic_array: null pointer");
    const auto& S = sym->S;
    switch (which) {
        case MDGPU_ARR_AMD_PERM: return copy_out(S.amd_perm, dst, cap, "amd_perm");
        case MDGPU_ARR_PARENT: return copy_out(S.parent, dst, cap, "parent");
        case MDGPU_ARR_POST: return copy_out(S.post, dst, cap, "post");
        case MDGPU_ARR_COUNT: return copy_out(S.count, dst, cap, "count");
        case MDGPU_ARR_RELAXED_PERM: return copy_out(S.relaxed_perm, dst, cap, "relaxed_perm");
        case MDGPU_ARR_PERM: return copy_out(S.perm, dst, cap, "perm");
        case MDGPU_ARR_COUNT_FINAL: return copy_out(S.count_final, dst, cap, "count_final");
        case MDGPU_ARR_LP: return copy_out(S.lp, dst, cap, "lp");
        case MDGPU_ARR_ALOW_P: return copy_out(S.a_low.p, dst, cap, "alow_p");
        case MDGPU_ARR_ALOW_I: return copy_out(S.a_low.i, dst, cap, "alow_i");
        case MDGPU_ARR_NODES: {
            std::vector<int> flat;
            flat.reserve((size_t)nodes_flat_len(S));
            for (const auto& nd : S.nodes) {
                flat.push_back(nd.first);
                flat.push_back(nd.ncols);
                flat.push_back(nd.rows);
                flat.push_back(nd.zeros);
                flat.push_back((int)nd.kids.size());
                flat.insert(flat.end(), nd.kids.begin(), nd.kids.end());
            }
            return copy_out(flat, dst, cap, "nodes");
        }
        case MDGPU_ARR_FRONT_PARENT: return copy_out(S.front_parent, dst, cap, "front_parent");
        case MDGPU_ARR_Q_OFF: return copy_out(S.q_off, dst, cap, "q_off");
        case MDGPU_ARR_Q: return copy_out(S.q_all, dst, cap, "q");
        case MDGPU_ARR_SEG_COL: return copy_out(sym->seg_col, dst, cap, "seg_col");
        case MDGPU_ARR_SEG_NNZ: return copy_out(sym->seg_nnz, dst, cap, "seg_nnz");
        case MDGPU_ARR_LEVEL: return copy_out(sym->level, dst, cap, "level");
        default: return fail(MDGPU_ERR_ARG, "mdgpu_symbolic_array: unknown array id");
    }
}

extern "C" int mdgpu_row_union(const int32_t* a, int na, const int32_t* b, int nb, int32_t* out, int* nout) {
    mdgpu::clear_error();
    if ((na > 0 && !a) || (nb > 0 && !b) || !out || !nout || na < 0 || nb < 0)
        return fail(MDGPU_ERR_ARG, "mdgpu_row_union: invalid argument");
    int top = -1;
    for (int t = 0; t < na; t++) top = std::max(top, a[t]);
    for (int t = 0; t < nb; t++) top = std::max(top, b[t]);
    for (int t = 0; t < na; t++)
        if (a[t] < 0 || (t > 0 && a[t] <= a[t - 1])) return fail(MDGPU_ERR_ARG, "mdgpu_row_union: a not sorted");
    for (int t = 0; t < nb; t++)
        if (b[t] < 0 || (t > 0 && b[t] <= b[t - 1])) return fail(MDGPU_ERR_ARG, "mdgpu_row_union: b not sorted");
    std::vector<int> stamp(top + 1, -1), rows;
    mdgpu::union_rows(rows, stamp, 0, a, na);
    mdgpu::union_rows(rows, stamp, 0, b, nb);
    mdgpu::sort_union_tail(rows, 0);
    std::copy(rows.begin(), rows.end(), out);
    *nout = (int)rows.size();
    return MDGPU_OK;
}

extern "C" int mdgpu_row_map(const int32_t* super, int n, const int32_t* sub, int m, int32_t* map) {
    mdgpu::clear_error();
    if ((n > 0 && !super) || (m > 0 && (!sub || !map)) || n < 0 || m < 0)
        return fail(MDGPU_ERR_ARG, "mdgpu_row_map: invalid argument");
    if (!mdgpu::relative_map(super, n, sub, m, map))
        return fail(MDGPU_ERR_ARG, "mdgpu_row_map: a row of the sublist is not in the list");
    return MDGPU_OK;
}

extern "C" int mdgpu_analyze_budgets(int n, const int32_t* Cp, const int32_t* Ci, int use_amd, int nbudgets,
                                     const int32_t* budgets, mdgpu_symbolic** syms) {
    mdgpu::clear_error();
    if (!Cp || !Ci || !budgets || !syms || nbudgets <= 0)
        return fail(MDGPU_ERR_ARG, "mdgpu_analyze_budgets: invalid argument");
    for (int k = 0; k < nbudgets; k++) {
        syms[k] = nullptr;
        if (budgets[k] < 0) return fail(MDGPU_ERR_ARG, "mdgpu_analyze_budgets: budgets must be >= 0");
    }
    mdgpu::AnalyzeOptions o;
    o.use_amd = use_amd != 0;
    o.partition = mdgpu::PART_RELAXED;
    std::vector<mdgpu::Symbolic> out;
    try {
        std::string err = mdgpu::analyze_budgets(n, Cp, Ci, o, std::vector<int>(budgets, budgets + nbudgets), out);
        if (!err.empty()) return fail(MDGPU_ERR_ARG, "mdgpu_analyze_budgets: " + err);
        for (int k = 0; k < nbudgets; k++) {
            mdgpu_symbolic* h = new mdgpu_symbolic();
            h->S = std::move(out[k]);
            finish_handle(h);
            syms[k] = h;
        }
    } catch (const std::bad_alloc&) {
        for (int k = 0; k < nbudgets; k++) {
            delete syms[k];
            syms[k] = nullptr;
        }
        return fail(MDGPU_ERR_NOMEM, "mdgpu_analyze_budgets: out of host memory");
    }
    return MDGPU_OK;
}

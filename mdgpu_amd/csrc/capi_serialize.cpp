// C ABI: on-disk form of a symbolic analysis, so that a fixed pattern can
// be refactorized in a later process without analysing it again.
//
// Layout (little-endian): "MDGPUSYM", u32 version, then the fields of
// mdgpu::Symbolic and the batch segments, each array as an i64 length and
// its elements, and finally the FNV-1a 64-bit hash of every preceding byte.
#include <cstdio>
#include <climits>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "common.h"

using mdgpu::fail;

namespace {

constexpr char kMagic[8] = {'M', 'D', 'G', 'P', 'U', 'S', 'Y', 'M'};
constexpr uint32_t kVersion = 1;

struct Writer {
    std::vector<unsigned char> buf;
    void raw(const void* p, size_t n) {
        const unsigned char* c = (const unsigned char*)p;
        buf.insert(buf.end(), c, c + n);
    }
    template <class T> void pod(T v) { raw(&v, sizeof v); }
    template <class T> void vec(const std::vector<T>& v) {
        pod<int64_t>((int64_t)v.size());
        if (!v.empty()) raw(v.data(), v.size() * sizeof(T));
    }
};

struct Reader {
    const unsigned char* p;
    size_t n, at = 0;
    bool ok = true;
    void raw(void* dst, size_t k) {
        if (!ok || k > n - at) { ok = false; return; }
        memcpy(dst, p + at, k);
        at += k;
    }
    template <class T> T pod() { T v{}; raw(&v, sizeof v); return v; }
    template <class T> void vec(std::vector<T>& v) {
        const int64_t len = pod<int64_t>();
        if (!ok || len < 0 || (uint64_t)len > (n - at) / sizeof(T)) { ok = false; return; }
        v.resize((size_t)len);
        if (len) raw(v.data(), (size_t)len * sizeof(T));
    }
};

uint64_t fnv1a(const unsigned char* p, size_t n) {
    uint64_t h = 1469598103934665603ull;
    for (size_t t = 0; t < n; t++) h = (h ^ p[t]) * 1099511628211ull;
    return h;
}

}  // namespace

extern "C" int mdgpu_symbolic_save(const mdgpu_symbolic* sym, const char* path) {
    mdgpu::clear_error();
    if (!sym || !path) return fail(MDGPU_ERR_ARG, "mdgpu_symbolic_save: null pointer");
    const mdgpu::Symbolic& S = sym->S;
    Writer w;
    w.raw(kMagic, sizeof kMagic);
    w.pod<uint32_t>(kVersion);
    w.pod<int32_t>(S.n);
    w.pod<int32_t>(S.nnz_in);
    w.pod<int32_t>(S.opts.use_amd ? 1 : 0);
    w.pod<int32_t>((int32_t)S.opts.partition);
    w.pod<int32_t>(S.opts.max_zeros);
    for (const std::vector<int>* v : {&S.amd_perm, &S.parent, &S.post, &S.count, &S.relaxed_perm, &S.perm, &S.pinv,
                                      &S.count_final, &S.lp, &S.a_low.p, &S.a_low.i, &S.a_src})
        w.vec(*v);
    w.pod<int64_t>((int64_t)S.nodes.size());
    for (const mdgpu::PartNode& nd : S.nodes) {
        w.pod<int32_t>(nd.first);
        w.pod<int32_t>(nd.ncols);
        w.pod<int32_t>(nd.rows);
        w.pod<int32_t>(nd.zeros);
        w.vec(nd.cols);
        w.vec(nd.kids);
    }
    w.vec(S.front_parent);
    w.vec(S.q_off);
    w.vec(S.q_all);
    w.pod<double>(S.flops_counts);
    w.pod<double>(S.flops_padded);
    w.pod<int64_t>(S.lnz);
    w.vec(sym->seg_col);
    w.vec(sym->seg_nnz);
    w.pod<uint64_t>(fnv1a(w.buf.data(), w.buf.size()));
    FILE* f = fopen(path, "wb");
    if (!f) return fail(MDGPU_ERR_ARG, std::string("mdgpu_symbolic_save: cannot open ") + path);
    const size_t put = fwrite(w.buf.data(), 1, w.buf.size(), f);
    const int closed = fclose(f);
    if (put != w.buf.size() || closed != 0) return fail(MDGPU_ERR_ARG, "mdgpu_symbolic_save: write failed");
    return MDGPU_OK;
}

// defined in capi_symbolic.cpp
void mdgpu_finish_symbolic_handle(mdgpu_symbolic* h);

extern "C" int mdgpu_symbolic_load(const char* path, mdgpu_symbolic** sym) {
    mdgpu::clear_error();
    if (!path || !sym) return fail(MDGPU_ERR_ARG, "mdgpu_symbolic_load: null pointer");
    *sym = nullptr;
    FILE* f = fopen(path, "rb");
    if (!f) return fail(MDGPU_ERR_ARG, std::string("mdgpu_symbolic_load: cannot open ") + path);
    std::vector<unsigned char> buf;
    unsigned char chunk[1 << 16];
    for (size_t got; (got = fread(chunk, 1, sizeof chunk, f)) > 0;) buf.insert(buf.end(), chunk, chunk + got);
    fclose(f);
    if (buf.size() < sizeof kMagic + 4 + 8 || memcmp(buf.data(), kMagic, sizeof kMagic) != 0)
        return fail(MDGPU_ERR_ARG, "mdgpu_symbolic_load: not a symbolic analysis file");
    uint64_t stored;
    memcpy(&stored, buf.data() + buf.size() - 8, 8);
    if (stored != fnv1a(buf.data(), buf.size() - 8)) return fail(MDGPU_ERR_ARG, "mdgpu_symbolic_load: checksum mismatch");
    Reader r{buf.data(), buf.size() - 8};
    char magic[8];
    r.raw(magic, 8);
    if (r.pod<uint32_t>() != kVersion) return fail(MDGPU_ERR_ARG, "mdgpu_symbolic_load: unsupported version");
    mdgpu_symbolic* h = new (std::nothrow) mdgpu_symbolic();
    if (!h) return fail(MDGPU_ERR_NOMEM, "mdgpu_symbolic_load: out of host memory");
    mdgpu::Symbolic& S = h->S;
    try {
        S.n = r.pod<int32_t>();
        S.nnz_in = r.pod<int32_t>();
        S.opts.use_amd = r.pod<int32_t>() != 0;
        S.opts.partition = (mdgpu::Partition)r.pod<int32_t>();
        S.opts.max_zeros = r.pod<int32_t>();
        for (std::vector<int>* v : {&S.amd_perm, &S.parent, &S.post, &S.count, &S.relaxed_perm, &S.perm, &S.pinv,
                                    &S.count_final, &S.lp, &S.a_low.p, &S.a_low.i, &S.a_src})
            r.vec(*v);
        S.a_low.n = S.n;
        const int64_t nn = r.pod<int64_t>();
        if (!r.ok || nn < 0 || nn > S.n) r.ok = false;
        for (int64_t t = 0; r.ok && t < nn; t++) {
            mdgpu::PartNode nd;
            nd.first = r.pod<int32_t>();
            nd.ncols = r.pod<int32_t>();
            nd.rows = r.pod<int32_t>();
            nd.zeros = r.pod<int32_t>();
            r.vec(nd.cols);
            r.vec(nd.kids);
            for (int k : nd.kids)
                if (k < 0 || k >= t) r.ok = false;       // children come before parents
            S.nodes.push_back(std::move(nd));
        }
        r.vec(S.front_parent);
        r.vec(S.q_off);
        r.vec(S.q_all);
        S.flops_counts = r.pod<double>();
        S.flops_padded = r.pod<double>();
        S.lnz = r.pod<int64_t>();
        r.vec(h->seg_col);
        r.vec(h->seg_nnz);
    } catch (const std::bad_alloc&) {
        delete h;
        return fail(MDGPU_ERR_NOMEM, "mdgpu_symbolic_load: out of host memory");
    }
    // structural consistency (everything a plan indexes with)
    const size_t n = (size_t)S.n, nn = S.nodes.size();
    bool ok = r.ok && r.at == r.n && S.n > 0 && S.perm.size() == n && S.pinv.size() == n &&
              S.count_final.size() == n && S.a_low.p.size() == n + 1 && S.front_parent.size() == nn &&
              S.q_off.size() == nn + 1 && S.a_src.size() == S.a_low.i.size() &&
              (int64_t)S.a_low.p[n] == (int64_t)S.a_low.i.size() && S.q_off[nn] == (int64_t)S.q_all.size() &&
              (int64_t)S.a_src.size() == S.nnz_in;
    for (size_t t = 0; ok && t < n; t++) ok = S.perm[t] >= 0 && S.perm[t] < S.n && S.pinv[S.perm[t]] == (int)t;
    for (size_t t = 0; ok && t < S.a_low.i.size(); t++) ok = S.a_low.i[t] >= 0 && S.a_low.i[t] < S.n;
    for (size_t t = 0; ok && t < S.a_src.size(); t++) ok = S.a_src[t] >= 0 && S.a_src[t] < S.nnz_in;
    for (size_t t = 0; ok && t < S.q_all.size(); t++) ok = S.q_all[t] >= 0 && S.q_all[t] < S.n;
    for (size_t f = 0; ok && f < nn; f++)
        ok = S.q_off[f] <= S.q_off[f + 1] && S.front_parent[f] >= -1 && S.front_parent[f] < (int)nn &&
             S.nodes[f].first >= 0 && S.nodes[f].ncols > 0 && S.nodes[f].first + S.nodes[f].ncols <= S.n &&
             S.q_off[f + 1] - S.q_off[f] >= S.nodes[f].ncols;
    // L's layout: lp is the prefix sum of the column counts (lnz entries;
    // left empty past int32, as for a large batch), and the columns of each
    // front hold its rows from the diagonal down: count_final[first + j] =
    // rows - j, with the pivot columns first in its sorted row list.  The
    // fronts cover every column once, children before parents.
    if (ok) {
        int64_t sum = 0;
        for (size_t k = 0; ok && k < n; k++) {
            ok = S.count_final[k] >= 1;
            sum += S.count_final[k];
        }
        ok = ok && sum == S.lnz;
        if (ok && S.lnz <= INT32_MAX) {
            ok = S.lp.size() == n + 1 && S.lp[0] == 0;
            for (size_t k = 0; ok && k < n; k++) ok = S.lp[k + 1] == S.lp[k] + S.count_final[k];
        } else if (ok) {
            ok = S.lp.empty();
        }
    }
    std::vector<char> covered(ok ? n : 0, 0);
    for (size_t f = 0; ok && f < nn; f++) {
        const mdgpu::PartNode& nd = S.nodes[f];
        const int64_t q0 = S.q_off[f], rows = S.q_off[f + 1] - S.q_off[f];
        ok = S.front_parent[f] == -1 || S.front_parent[f] > (int)f;
        for (int j = 0; ok && j < nd.ncols; j++) {
            ok = !covered[nd.first + j] && S.count_final[nd.first + j] == rows - j && S.q_all[q0 + j] == nd.first + j;
            covered[nd.first + j] = 1;
        }
        for (int64_t t = q0 + 1; ok && t < q0 + rows; t++) ok = S.q_all[t - 1] < S.q_all[t];
        for (int k : nd.kids) ok = ok && S.front_parent[k] == (int)f;
    }
    for (size_t k = 0; ok && k < n; k++) ok = covered[k];
    // segments of a batch: columns and caller entries split in order
    const bool no_seg = h->seg_col.empty() && h->seg_nnz.empty();   // one matrix (filled in below)
    ok = ok && (no_seg || (h->seg_col.size() >= 2 && h->seg_col.size() == h->seg_nnz.size() &&
                           h->seg_col[0] == 0 && h->seg_nnz[0] == 0 && h->seg_col.back() == S.n &&
                           h->seg_nnz.back() == S.nnz_in));
    for (size_t k = 0; ok && k + 1 < h->seg_col.size(); k++)
        ok = h->seg_col[k] < h->seg_col[k + 1] && h->seg_nnz[k] <= h->seg_nnz[k + 1];
    if (!ok) {
        delete h;
        return fail(MDGPU_ERR_ARG, "mdgpu_symbolic_load: inconsistent or truncated file");
    }
    mdgpu_finish_symbolic_handle(h);
    *sym = h;
    return MDGPU_OK;
}

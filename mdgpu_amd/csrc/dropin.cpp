// C++ drop-in of the reference's interface (include/cs.h, etree.h, mdgpu.h):
// the CSparse routines its driver calls, the node-tree builders and the
// numeric entry point Factor().  Symbolic routines delegate to this
// library's own implementations (symbolic.cpp); Factor() runs the numeric
// phase on the GPU through the C ABI.  Routine semantics follow the
// reference's cs.cpp / etree.cpp / mdgpu.cu (cited per function).
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cctype>
#include <chrono>
#include <cstring>
#include <string>
#include <type_traits>
#include <unordered_map>
#include <vector>

#include "../../include/cs.h"
#include "../../include/etree.h"
#include "../../include/mdgpu.h"
#include "../../include/mdgpu_c.h"
#include "common.h"
#include "symbolic.h"

// Built once per FLOTANTE (float and double): every definition below lives
// in the inline namespace of that FLOTANTE (include/types.h).
static_assert(std::is_same<FLOTANTE, float>::value || std::is_same<FLOTANTE, double>::value,
              "the drop-in is built for FLOTANTE = float or double");

MDGPU_DROPIN_BEGIN

bool GPU = true;
int block_size = -1;
bool pad = false;

namespace {

mdgpu::Csc to_csc(const cs* A) {
    mdgpu::Csc c;
    c.n = A->n;
    c.p.assign(A->p, A->p + A->n + 1);
    c.i.assign(A->i, A->i + A->p[A->n]);
    return c;
}

int* to_int_array(const std::vector<int>& v) {
    int* out = (int*)malloc(sizeof(int) * std::max<size_t>(v.size(), 1));
    if (out && !v.empty()) memcpy(out, v.data(), sizeof(int) * v.size());
    return out;
}

}  // namespace

// ------------------------------------------------------------ memory (cs.cpp:29-55)
void* cs_malloc(int n, size_t size) { return malloc(CS_MAX(n, 1) * size); }
void* cs_calloc(int n, size_t size) { return calloc(CS_MAX(n, 1), size); }
void* cs_free(void* p) {
    if (p) free(p);
    return NULL;
}
void* cs_realloc(void* p, int n, size_t size, int* ok) {
    void* q = realloc(p, CS_MAX(n, 1) * size);
    *ok = q != NULL;
    return *ok ? q : p;
}

cs* cs_spalloc(int m, int n, int nzmax, int values, int triplet) {
    cs* A = (cs*)cs_calloc(1, sizeof(cs));
    if (!A) return NULL;
    A->m = m;
    A->n = n;
    A->nzmax = nzmax = CS_MAX(nzmax, 1);
    A->nz = triplet ? 0 : -1;
    A->p = (int*)cs_malloc(triplet ? nzmax : n + 1, sizeof(int));
    A->i = (int*)cs_malloc(nzmax, sizeof(int));
    A->x = values ? (FLOTANTE*)cs_malloc(nzmax, sizeof(FLOTANTE)) : NULL;
    return (!A->p || !A->i || (values && !A->x)) ? cs_spfree(A) : A;
}

cs* cs_spfree(cs* A) {
    if (!A) return NULL;
    cs_free(A->p);
    cs_free(A->i);
    cs_free(A->x);
    return (cs*)cs_free(A);
}

int cs_sprealloc(cs* A, int nzmax) {
    if (!A) return 0;
    if (nzmax <= 0) nzmax = CS_CSC(A) ? A->p[A->n] : A->nz;
    int oki = 1, okp = 1, okx = 1;
    A->i = (int*)cs_realloc(A->i, nzmax, sizeof(int), &oki);
    if (CS_TRIPLET(A)) A->p = (int*)cs_realloc(A->p, nzmax, sizeof(int), &okp);
    if (A->x) A->x = (FLOTANTE*)cs_realloc(A->x, nzmax, sizeof(FLOTANTE), &okx);
    int ok = oki && okp && okx;
    if (ok) A->nzmax = nzmax;
    return ok;
}

cs* cs_done(cs* C, void* w, void* x, int ok) {
    cs_free(w);
    cs_free(x);
    return ok ? C : cs_spfree(C);
}

int* cs_idone(int* p, cs* C, void* w, int ok) {
    cs_spfree(C);
    cs_free(w);
    return ok ? p : (int*)cs_free(p);
}

// ------------------------------------------------------------ triplets and I/O
// cs_entry (cs.cpp:177-197): grows x2 below 4M entries, then x1.2
int cs_entry(cs* T, int i, int j, FLOTANTE x) {
    if (!CS_TRIPLET(T) || i < 0 || j < 0) return 0;
    if (T->nz >= T->nzmax) {
        int grow = T->nzmax < 4000000 ? 2 * T->nzmax : (int)(1.2 * T->nzmax);
        if (!cs_sprealloc(T, grow)) return 0;
    }
    if (T->x) T->x[T->nz] = x;
    T->i[T->nz] = i;
    T->p[T->nz++] = j;
    T->m = CS_MAX(T->m, i + 1);
    T->n = CS_MAX(T->n, j + 1);
    return 1;
}

// one "i j x" line, the value in FLOTANTE's own format (cs.cpp:12-18)
namespace {
bool read_triplet(FILE* f, int* i, int* j, float* x) { return fscanf(f, "%d %d %g\n", i, j, x) == 3; }
bool read_triplet(FILE* f, int* i, int* j, double* x) { return fscanf(f, "%d %d %lg\n", i, j, x) == 3; }
}  // namespace

// cs_load (cs.cpp:4-27): headerless "i j x" lines, 1-based
cs* cs_load(FILE* f) {
    if (!f) return NULL;
    cs* T = cs_spalloc(0, 0, 1, 1, 1);
    int i, j;
    FLOTANTE x;
    while (T && read_triplet(f, &i, &j, &x))
        if (!cs_entry(T, i - 1, j - 1, x)) return cs_spfree(T);
    return T;
}

// Matrix Market coordinate file (real, integer or pattern; symmetric or
// general): the entries on and below the diagonal as a triplet matrix, the
// input cs_compress and the driver take (the reference's test() reads .mtx
// files through cs_load, mdgpu.cu:1655, which the header lines stop).  Upper
// entries of a general file are dropped (the matrix is taken symmetric);
// pattern entries get the value 1.  NULL on a malformed file.
cs* cs_load_mm(FILE* f) {
    if (!f) return NULL;
    char line[1024];
    if (!fgets(line, sizeof line, f)) return NULL;
    std::string head(line);
    for (char& ch : head) ch = (char)tolower((unsigned char)ch);
    if (head.rfind("%%matrixmarket", 0) != 0 || head.find("coordinate") == std::string::npos) return NULL;
    const bool pattern = head.find("pattern") != std::string::npos;
    if (head.find("complex") != std::string::npos) return NULL;
    int m = 0, n = 0, nnz = -1;
    while (fgets(line, sizeof line, f)) {
        if (line[0] == '%' || line[0] == '\n') continue;
        if (sscanf(line, "%d %d %d", &m, &n, &nnz) != 3) return NULL;
        break;
    }
    if (nnz < 0 || m != n || n <= 0) return NULL;
    cs* T = cs_spalloc(n, n, CS_MAX(nnz, 1), 1, 1);
    for (int t = 0; T && t < nnz; t++) {
        int i, j;
        double v = 1.0;
        if (!fgets(line, sizeof line, f)) return cs_spfree(T);
        const int got = pattern ? sscanf(line, "%d %d", &i, &j) : sscanf(line, "%d %d %lg", &i, &j, &v);
        if (got != (pattern ? 2 : 3) || i < 1 || j < 1 || i > n || j > n) return cs_spfree(T);
        if (i >= j && !cs_entry(T, i - 1, j - 1, (FLOTANTE)v)) return cs_spfree(T);
    }
    if (T) {
        T->m = n;
        T->n = n;
    }
    return T;
}

FLOTANTE cs_cumsum(int* p, int* c, int n) {
    if (!p || !c) return -1;
    int nz = 0;
    double total = 0;
    for (int t = 0; t < n; t++) {
        p[t] = nz;
        nz += c[t];
        total += c[t];
        c[t] = p[t];
    }
    p[n] = nz;
    return (FLOTANTE)total;
}

// cs_compress (cs.cpp:251-270): stable counting sort by column
cs* cs_compress(const cs* T) {
    if (!CS_TRIPLET(T)) return NULL;
    const int m = T->m, n = T->n, nz = T->nz;
    cs* C = cs_spalloc(m, n, nz, T->x != NULL, 0);
    int* w = (int*)cs_calloc(n, sizeof(int));
    if (!C || !w) return cs_done(C, w, NULL, 0);
    for (int k = 0; k < nz; k++) w[T->p[k]]++;
    cs_cumsum(C->p, w, n);
    for (int k = 0; k < nz; k++) {
        int q = w[T->p[k]]++;
        C->i[q] = T->i[k];
        if (C->x) C->x[q] = T->x[k];
    }
    return cs_done(C, w, NULL, 1);
}

cs* cs_transpose(const cs* A, int values) {
    if (!CS_CSC(A)) return NULL;
    const int m = A->m, n = A->n;
    cs* C = cs_spalloc(n, m, A->p[n], values && A->x, 0);
    int* w = (int*)cs_calloc(m, sizeof(int));
    if (!C || !w) return cs_done(C, w, NULL, 0);
    for (int p = 0; p < A->p[n]; p++) w[A->i[p]]++;
    cs_cumsum(C->p, w, m);
    for (int j = 0; j < n; j++)
        for (int p = A->p[j]; p < A->p[j + 1]; p++) {
            int q = w[A->i[p]]++;
            C->i[q] = j;
            if (C->x) C->x[q] = A->x[p];
        }
    return cs_done(C, w, NULL, 1);
}

int cs_scatter(const cs* A, int j, FLOTANTE beta, int* w, FLOTANTE* x, int mark, cs* C, int nz) {
    if (!CS_CSC(A) || !w || !CS_CSC(C)) return -1;
    for (int p = A->p[j]; p < A->p[j + 1]; p++) {
        int i = A->i[p];
        if (w[i] < mark) {
            w[i] = mark;
            C->i[nz++] = i;
            if (x) x[i] = beta * A->x[p];
        } else if (x) {
            x[i] += beta * A->x[p];
        }
    }
    return nz;
}

// cs_add (cs.cpp:510-535): entries of A(:,j) first, then new ones of B(:,j)
cs* cs_add(const cs* A, const cs* B, FLOTANTE alpha, FLOTANTE beta) {
    if (!CS_CSC(A) || !CS_CSC(B) || A->m != B->m || A->n != B->n) return NULL;
    const int m = A->m, n = B->n;
    const int values = A->x != NULL && B->x != NULL;
    int* w = (int*)cs_calloc(m, sizeof(int));
    FLOTANTE* x = values ? (FLOTANTE*)cs_malloc(m, sizeof(FLOTANTE)) : NULL;
    cs* C = cs_spalloc(m, n, A->p[A->n] + B->p[n], values, 0);
    if (!C || !w || (values && !x)) return cs_done(C, w, x, 0);
    int nz = 0;
    for (int j = 0; j < n; j++) {
        C->p[j] = nz;
        nz = cs_scatter(A, j, alpha, w, x, j + 1, C, nz);
        nz = cs_scatter(B, j, beta, w, x, j + 1, C, nz);
        if (values)
            for (int p = C->p[j]; p < nz; p++) C->x[p] = x[C->i[p]];
    }
    C->p[n] = nz;
    cs_sprealloc(C, 0);
    return cs_done(C, w, x, 1);
}

int cs_fkeep(cs* A, int (*fkeep)(int, int, FLOTANTE, void*), void* other) {
    if (!CS_CSC(A) || !fkeep) return -1;
    int nz = 0;
    for (int j = 0; j < A->n; j++) {
        int p = A->p[j];
        A->p[j] = nz;
        for (; p < A->p[j + 1]; p++)
            if (fkeep(A->i[p], j, A->x ? A->x[p] : 1, other)) {
                if (A->x) A->x[nz] = A->x[p];
                A->i[nz++] = A->i[p];
            }
    }
    A->p[A->n] = nz;
    cs_sprealloc(A, 0);
    return nz;
}

int cs_dupl(cs* A) {
    if (!CS_CSC(A)) return 0;
    int* w = (int*)cs_malloc(A->m, sizeof(int));
    if (!w) return 0;
    for (int i = 0; i < A->m; i++) w[i] = -1;
    int nz = 0;
    for (int j = 0; j < A->n; j++) {
        int q = nz;
        for (int p = A->p[j]; p < A->p[j + 1]; p++) {
            int i = A->i[p];
            if (w[i] >= q) {
                A->x[w[i]] += A->x[p];
            } else {
                w[i] = nz;
                A->i[nz] = i;
                A->x[nz++] = A->x[p];
            }
        }
        A->p[j] = q;
    }
    A->p[A->n] = nz;
    cs_free(w);
    return cs_sprealloc(A, 0);
}

static int keep_above_tol(int, int, FLOTANTE aij, void* tol) { return fabs(aij) > *((FLOTANTE*)tol); }
int cs_droptol(cs* A, FLOTANTE tol) { return cs_fkeep(A, &keep_above_tol, &tol); }

FLOTANTE cs_norm(const cs* A) {
    if (!CS_CSC(A) || !A->x) return -1;
    FLOTANTE best = 0;
    for (int j = 0; j < A->n; j++) {
        FLOTANTE s = 0;
        for (int p = A->p[j]; p < A->p[j + 1]; p++) s += fabs(A->x[p]);
        best = CS_MAX(best, s);
    }
    return best;
}

int cs_print(FILE* f, const cs* A, int brief) {
    if (!A) {
        fprintf(f, "(null)\n");
        return 0;
    }
    if (A->nz < 0) {
        fprintf(f, "%d-by-%d, nzmax: %d nnz: %d, 1-norm: %g\n", A->m, A->n, A->nzmax, A->p[A->n],
                (double)cs_norm(A));
        for (int j = 0; j < A->n; j++) {
            fprintf(f, "    col %d : locations %d to %d\n", j, A->p[j], A->p[j + 1] - 1);
            for (int p = A->p[j]; p < A->p[j + 1]; p++) {
                fprintf(f, "      %d : %g\n", A->i[p], A->x ? (double)A->x[p] : 1.0);
                if (brief && p > 20) { fprintf(f, "  ...\n"); return 1; }
            }
        }
    } else {
        fprintf(f, "triplet: %d-by-%d, nzmax: %d nnz: %d\n", A->m, A->n, A->nzmax, A->nz);
        for (int p = 0; p < A->nz; p++) {
            fprintf(f, "    %d %d : %g\n", A->i[p], A->p[p], A->x ? (double)A->x[p] : 1.0);
            if (brief && p > 20) { fprintf(f, "  ...\n"); return 1; }
        }
    }
    return 1;
}
int cs_print(const cs* A, int brief) { return cs_print(stdout, A, brief); }

cs* cs_permute(const cs* A, const int* pinv, const int* q, int values) {
    if (!CS_CSC(A)) return NULL;
    cs* C = cs_spalloc(A->m, A->n, A->p[A->n], values && A->x != NULL, 0);
    if (!C) return NULL;
    int nz = 0;
    for (int k = 0; k < A->n; k++) {
        C->p[k] = nz;
        int j = q ? q[k] : k;
        for (int t = A->p[j]; t < A->p[j + 1]; t++) {
            if (C->x) C->x[nz] = A->x[t];
            C->i[nz++] = pinv ? pinv[A->i[t]] : A->i[t];
        }
    }
    C->p[A->n] = nz;
    return C;
}

// cs_symperm (cs.cpp:1005-1041): C = A(p,p), upper part of A kept
cs* cs_symperm(const cs* A, const int* pinv, int values) {
    if (!CS_CSC(A)) return NULL;
    const int n = A->n;
    cs* C = cs_spalloc(n, n, A->p[n], values && A->x != NULL, 0);
    int* w = (int*)cs_calloc(n, sizeof(int));
    if (!C || !w) return cs_done(C, w, NULL, 0);
    for (int j = 0; j < n; j++) {
        int j2 = pinv ? pinv[j] : j;
        for (int p = A->p[j]; p < A->p[j + 1]; p++) {
            int i = A->i[p];
            if (i > j) continue;
            int i2 = pinv ? pinv[i] : i;
            w[CS_MAX(i2, j2)]++;
        }
    }
    cs_cumsum(C->p, w, n);
    for (int j = 0; j < n; j++) {
        int j2 = pinv ? pinv[j] : j;
        for (int p = A->p[j]; p < A->p[j + 1]; p++) {
            int i = A->i[p];
            if (i > j) continue;
            int i2 = pinv ? pinv[i] : i;
            int q = w[CS_MAX(i2, j2)]++;
            C->i[q] = CS_MIN(i2, j2);
            if (C->x) C->x[q] = A->x[p];
        }
    }
    return cs_done(C, w, NULL, 1);
}

// ------------------------------------------------------------ ordering, trees
int* cs_amd(int order, const cs* A) {
    if (!CS_CSC(A) || order != 1 || A->m != A->n) return NULL;   // Cholesky ordering only
    return to_int_array(mdgpu::amd_order(to_csc(A)));
}

int* cs_pinv(const int* p, int n) {
    if (!p) return NULL;
    int* pinv = (int*)cs_malloc(n, sizeof(int));
    if (!pinv) return NULL;
    for (int k = 0; k < n; k++) pinv[p[k]] = k;
    return pinv;
}

int* cs_etree(const cs* A, int ata) {
    if (!CS_CSC(A) || ata) return NULL;   // the driver only uses ata = 0
    return to_int_array(mdgpu::etree(to_csc(A)));
}

int* cs_post(const int* parent, int n) {
    if (!parent) return NULL;
    return to_int_array(mdgpu::postorder(std::vector<int>(parent, parent + n)));
}

int cs_tdfs(int j, int k, int* head, const int* next, int* post, int* stack) {
    if (!head || !next || !post || !stack) return -1;
    int top = 0;
    stack[0] = j;
    while (top >= 0) {
        int p = stack[top];
        int i = head[p];
        if (i == -1) {
            top--;
            post[k++] = p;
        } else {
            head[p] = next[i];
            stack[++top] = i;
        }
    }
    return k;
}

int* cs_counts(const cs* A, const int* parent, const int* post, int ata) {
    if (!CS_CSC(A) || !parent || !post || ata) return NULL;
    const int n = A->n;
    return to_int_array(mdgpu::column_counts(to_csc(A), std::vector<int>(parent, parent + n),
                                             std::vector<int>(post, post + n)));
}

int cs_leaf(int i, int j, const int* first, int* maxfirst, int* prevleaf, int* ancestor, int* jleaf) {
    if (!first || !maxfirst || !prevleaf || !ancestor || !jleaf) return -1;
    *jleaf = 0;
    if (i <= j || first[j] <= maxfirst[i]) return -1;
    maxfirst[i] = first[j];
    int jprev = prevleaf[i];
    prevleaf[i] = j;
    *jleaf = (jprev == -1) ? 1 : 2;
    if (*jleaf == 1) return i;
    int q = jprev;
    while (q != ancestor[q]) q = ancestor[q];
    for (int s = jprev; s != q;) {
        int up = ancestor[s];
        ancestor[s] = q;
        s = up;
    }
    return q;
}

// ------------------------------------------------------------ solves
int cs_gaxpy(const cs* A, const FLOTANTE* x, FLOTANTE* y) {
    if (!CS_CSC(A) || !x || !y) return 0;
    for (int j = 0; j < A->n; j++)
        for (int p = A->p[j]; p < A->p[j + 1]; p++) y[A->i[p]] += A->x[p] * x[j];
    return 1;
}

int cs_lsolve(const cs* L, FLOTANTE* x) {
    if (!CS_CSC(L) || !x) return 0;
    for (int j = 0; j < L->n; j++) {
        x[j] /= L->x[L->p[j]];
        for (int p = L->p[j] + 1; p < L->p[j + 1]; p++) x[L->i[p]] -= L->x[p] * x[j];
    }
    return 1;
}

int cs_ltsolve(const cs* L, FLOTANTE* x) {
    if (!CS_CSC(L) || !x) return 0;
    for (int j = L->n - 1; j >= 0; j--) {
        for (int p = L->p[j] + 1; p < L->p[j + 1]; p++) x[j] -= L->x[p] * x[L->i[p]];
        x[j] /= L->x[L->p[j]];
    }
    return 1;
}

int cs_pvec(const int* p, const FLOTANTE* b, FLOTANTE* x, int n) {
    if (!x || !b) return 0;
    for (int k = 0; k < n; k++) x[k] = b[p ? p[k] : k];
    return 1;
}
int cs_pvec(const int* p, const int* b, int* x, int n) {
    if (!x || !b) return 0;
    for (int k = 0; k < n; k++) x[k] = b[p ? p[k] : k];
    return 1;
}
int cs_ipvec(const int* p, const FLOTANTE* b, FLOTANTE* x, int n) {
    if (!x || !b) return 0;
    for (int k = 0; k < n; k++) x[p ? p[k] : k] = b[k];
    return 1;
}

// ------------------------------------------------------------ node trees
namespace {

struct ETree* new_node(int nodo, int cols, int rows, int ceros) {
    struct ETree* t = (struct ETree*)calloc(1, sizeof(struct ETree));
    t->nodo = nodo;
    t->cols = cols;
    t->rows = rows;
    t->ceros = ceros;
    return t;
}

// PartNode list -> ETree structs in the same order
void to_etree_list(const std::vector<mdgpu::PartNode>& nodes, struct ETree*** out, int* cnt) {
    const int nn = (int)nodes.size();
    struct ETree** list = (struct ETree**)malloc(sizeof(struct ETree*) * std::max(nn, 1));
    for (int f = 0; f < nn; f++) {
        const auto& nd = nodes[f];
        list[f] = new_node(nd.first, nd.ncols, nd.rows, nd.zeros);
        list[f]->nodos = to_int_array(nd.cols);
    }
    for (int f = 0; f < nn; f++) {
        const auto& nd = nodes[f];
        list[f]->nHijos = (int)nd.kids.size();
        list[f]->hijos = (struct ETree**)malloc(sizeof(struct ETree*) * std::max<size_t>(nd.kids.size(), 1));
        for (size_t h = 0; h < nd.kids.size(); h++) list[f]->hijos[h] = list[nd.kids[h]];
    }
    *out = list;
    *cnt = nn;
}

// ETree list (postorder, one column per node) -> PartNode list
std::vector<mdgpu::PartNode> from_etree_list(struct ETree** list, int cnt) {
    std::unordered_map<const struct ETree*, int> index;
    for (int f = 0; f < cnt; f++) index[list[f]] = f;
    std::vector<mdgpu::PartNode> nodes(cnt);
    for (int f = 0; f < cnt; f++) {
        nodes[f].first = list[f]->nodo;
        nodes[f].ncols = list[f]->cols;
        nodes[f].rows = list[f]->rows;
        nodes[f].zeros = list[f]->ceros;
        nodes[f].cols.clear();
        for (int j = 0; j < list[f]->cols; j++) nodes[f].cols.push_back(list[f]->nodo + j);
        for (int h = 0; h < list[f]->nHijos; h++) nodes[f].kids.push_back(index.at(list[f]->hijos[h]));
    }
    return nodes;
}

}  // namespace

struct ETree* GetETree(cs* A, int* parent, int* post, int* count, struct ETree*** listaNodos, int* cantNodos) {
    const int n = A->n;
    int roots = 0;
    for (int j = 0; j < n; j++) roots += parent[j] == -1;
    if (roots != 1) {
        mdgpu::fail(MDGPU_ERR_ARG, "GetETree: the elimination tree is a forest");
        fprintf(stderr, "GetETree: the elimination tree is a forest (%d roots)\n", roots);
        *listaNodos = NULL;
        *cantNodos = 0;
        return NULL;
    }
    struct ETree* nodes = (struct ETree*)calloc(n, sizeof(struct ETree));
    struct ETree* root = NULL;
    std::vector<int> nk(n, 0);
    for (int j = 0; j < n; j++)
        if (parent[j] != -1) nk[parent[j]]++;
    for (int j = 0; j < n; j++) {
        nodes[j].nodo = j;
        nodes[j].rows = count[j];
        nodes[j].cols = 1;
        nodes[j].nodos = (int*)malloc(sizeof(int));
        nodes[j].nodos[0] = j;
        nodes[j].hijos = nk[j] ? (struct ETree**)malloc(sizeof(struct ETree*) * nk[j]) : NULL;
    }
    for (int j = 0; j < n; j++) {      // children in increasing index order
        if (parent[j] == -1) root = &nodes[j];
        else nodes[parent[j]].hijos[nodes[parent[j]].nHijos++] = &nodes[j];
    }
    *listaNodos = (struct ETree**)malloc(sizeof(struct ETree*) * n);
    for (int k = 0; k < n; k++) (*listaNodos)[k] = &nodes[post[k]];
    *cantNodos = n;
    return root;
}

void GetSuperETree(struct ETree** listaNodos, int cantNodos, struct ETree*** out, int* cnt) {
    std::vector<mdgpu::PartNode> sup;
    mdgpu::supernodal_partition(from_etree_list(listaNodos, cantNodos), sup);
    to_etree_list(sup, out, cnt);
}

void GetRelaxedETree(struct ETree** listaNodos, int cantNodos, int max_ceros, int* count, int** perm,
                     struct ETree*** out, int* cnt) {
    std::vector<mdgpu::PartNode> plain = from_etree_list(listaNodos, cantNodos);
    int n = 0;
    for (const auto& nd : plain) n = std::max(n, nd.first + nd.ncols);
    std::vector<int> cvec(count, count + n), p, rel;
    std::vector<mdgpu::PartNode> nodes;
    std::string err = mdgpu::relaxed_partition(plain, max_ceros, cvec, p, nodes);
    if (!err.empty()) {
        mdgpu::fail(MDGPU_ERR_ARG, "GetRelaxedETree: " + err);
        fprintf(stderr, "GetRelaxedETree: %s\n", err.c_str());
        *out = NULL;
        *cnt = 0;
        *perm = NULL;
        return;
    }
    memcpy(count, cvec.data(), sizeof(int) * n);
    *perm = to_int_array(p);
    to_etree_list(nodes, out, cnt);
}

void PrintETree(struct ETree* root, int nivel) {
    for (int t = 0; t < nivel; t++) printf("-");
    if (!root) {
        printf("> NULL\n");
        return;
    }
    if (root->cols == 1) printf("> %i\n", root->nodo);
    else printf("> (%i-%i, %i, %i)\n", root->nodo, root->nodo + root->cols - 1, root->rows, root->ceros);
    for (int h = 0; h < root->nHijos; h++) PrintETree(root->hijos[h], nivel + 1);
}

// ------------------------------------------------------------ fronts (mdgpu2.cpp:21-29)
void MatrizAlloc(Matriz* m) {
    m->w = pad ? (m->n / 16 + 1) * 16 : m->n;
    m->x = (FLOTANTE*)calloc((size_t)m->w * (size_t)m->w > 0 ? (size_t)m->w * m->w : 1, sizeof(FLOTANTE));
}

FLOTANTE MatrizGet(cs* M, int i, int j) {
    for (int k = M->p[j]; k < M->p[j + 1]; k++)
        if (M->i[k] == i) return M->x[k];
    return 0.0;
}
void MatrizSet(cs* M, int i, int j, FLOTANTE v) {
    for (int k = M->p[j]; k < M->p[j + 1]; k++)
        if (M->i[k] == i) M->x[k] = v;
}

void VectorPrint(FILE* f, int* v, int n) {
    for (int t = 0; t < n; t++) fprintf(f, "%i ", v[t]);
    fprintf(f, "\n");
}
void VectorPrint(int* v, int n) { VectorPrint(stdout, v, n); }
void MatrizPrint(FILE* f, Matriz* m, const char* format) {
    for (int i = 0; i < m->n; i++) {
        for (int j = 0; j < m->n; j++) fprintf(f, format, (double)MatrizGet(m, i, j));
        fprintf(f, "\n");
    }
}
void MatrizPrint(Matriz* m, const char* format) { MatrizPrint(stdout, m, format); }
void MatrizPrint(FILE* f, cs* m, const char* format) {
    for (int i = 0; i < m->m; i++) {
        for (int j = 0; j < m->n; j++) fprintf(f, format, (double)MatrizGet(m, i, j));
        fprintf(f, "\n");
    }
}
void MatrizPrint(cs* m, const char* format) { MatrizPrint(stdout, m, format); }
void MatrizSpy(Matriz* m) {
    for (int i = 0; i < m->n; i++) {
        for (int j = 0; j < m->n; j++) putchar(MatrizGet(m, i, j) != 0 ? 'X' : '.');
        putchar('\n');
    }
}
void MatrizSpy(cs* m) {
    for (int i = 0; i < m->m; i++) {
        for (int j = 0; j < m->n; j++) putchar(MatrizGet(m, i, j) != 0 ? 'X' : '.');
        putchar('\n');
    }
}
void MatrizSpy2(Matriz* A, Matriz* L) {
    for (int i = 0; i < A->n; i++) {
        for (int j = 0; j < A->n; j++) {
            bool a = MatrizGet(A, i, j) != 0, l = MatrizGet(L, i, j) != 0;
            putchar(a ? 'X' : (l ? '+' : '.'));
        }
        putchar('\n');
    }
}

// ------------------------------------------------------------ Factor
namespace {
thread_local FactorTimes g_times;
double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
}  // namespace

void FactorGetTimes(FactorTimes* out) {
    if (out) *out = g_times;
}

// Numeric factorization of the permuted lower triangle A over the node list
// (mdgpu.cu:1394-1608), on the GPU in FLOTANTE's precision (MDGPU_FP32 for
// float, MDGPU_FP64 for double, as the reference computes everything in
// FLOTANTE); L is written into spL (layout of mdgpu.cu:1346-1352).  The
// phase times of the reference's report (mdgpu.cu:1901-1930) are kept for
// FactorGetTimes: the kernel classes are timed with HIP events around every
// launch (mdgpu_factorize_profile).
void Factor(cs* A, cs* spL, struct ETree** listaNodos, int cantNodos) {
    const double t_start = now_s();
    FactorTimes& T = g_times;
    T = FactorTimes{};
    auto report = [](const char* what) {
        fprintf(stderr, "Factor: %s: %s\n", what, mdgpu_last_error());
    };
    if (!CS_CSC(A) || !A->x || !spL || !spL->p || !listaNodos || cantNodos <= 0) {
        mdgpu::fail(MDGPU_ERR_ARG, "Factor: invalid argument");
        report("arguments");
        return;
    }
    const int n = A->n;
    double t0 = now_s();
    std::unordered_map<const struct ETree*, int> index;
    for (int f = 0; f < cantNodos; f++) index[listaNodos[f]] = f;
    std::vector<int32_t> first(cantNodos), ncols(cantNodos), kid_ptr(cantNodos + 1, 0), kids;
    for (int f = 0; f < cantNodos; f++) {
        first[f] = listaNodos[f]->nodo;
        ncols[f] = listaNodos[f]->cols;
        for (int h = 0; h < listaNodos[f]->nHijos; h++) {
            auto it = index.find(listaNodos[f]->hijos[h]);
            if (it == index.end()) {
                mdgpu::fail(MDGPU_ERR_ARG, "Factor: a child node is not in the node list");
                report("node list");
                return;
            }
            kids.push_back(it->second);
        }
        kid_ptr[f + 1] = (int32_t)kids.size();
    }
    mdgpu_symbolic* sym = nullptr;
    if (mdgpu_analyze_nodes(n, A->p, A->i, cantNodos, first.data(), ncols.data(), kid_ptr.data(),
                            kids.empty() ? kid_ptr.data() : kids.data(), &sym) != MDGPU_OK) {
        report("analysis");
        return;
    }
    T.merge = now_s() - t0;
    double fl[2] = {0, 0};
    mdgpu_symbolic_flops(sym, fl);
    T.mflops = fl[1] / 1e6;
    std::vector<int32_t> lp(n + 1);
    mdgpu_symbolic_array(sym, MDGPU_ARR_LP, lp.data(), n + 1);
    for (int k = 0; k <= n; k++)
        if (lp[k] != spL->p[k]) {
            mdgpu::fail(MDGPU_ERR_ARG, "Factor: spL->p does not match the fronts' column counts");
            report("L skeleton");
            mdgpu_symbolic_destroy(sym);
            return;
        }
    mdgpu_numeric* num = nullptr;
    double* d_ax = nullptr;
    const size_t nz = (size_t)A->p[n];
    const int precision = std::is_same<FLOTANTE, float>::value ? MDGPU_FP32 : MDGPU_FP64;
    t0 = now_s();
    int st = mdgpu_numeric_create(sym, precision, &num);
    if (st == MDGPU_OK && hipMalloc((void**)&d_ax, sizeof(double) * std::max<size_t>(nz, 1)) != hipSuccess)
        st = mdgpu::fail(MDGPU_ERR_NOMEM, "Factor: hipMalloc failed");
    T.malloc_gpu = now_s() - t0;
    if (st == MDGPU_OK) {
        t0 = now_s();
        std::vector<double> ax(A->x, A->x + nz);
        if (hipMemcpy(d_ax, ax.data(), sizeof(double) * nz, hipMemcpyHostToDevice) != hipSuccess)
            st = mdgpu::fail(MDGPU_ERR_DEVICE, "Factor: hipMemcpy failed");
        T.memcpy_x = now_s() - t0;
    }
    if (st == MDGPU_OK) {
        double prof[4 * MDGPU_KIND_N];
        st = mdgpu_factorize_profile(num, d_ax, nullptr, prof);
        if (st == MDGPU_OK) {
            // classes: 0 assembly, 1 extend-add, 2 LDS panels, 3 blocked
            // potrf + TRSM, 4 inverses for the solves, 5 GEMM/SYRK (ms)
            T.getfk = prof[4 * 0] * 1e-3;
            T.extend_add = prof[4 * 1] * 1e-3;
            T.trsm_gpu = (prof[4 * 2] + prof[4 * 3]) * 1e-3;
            T.gemm_gpu = prof[4 * 5] * 1e-3;
            T.other_gpu = prof[4 * 4] * 1e-3;
            T.blas_gpu = T.trsm_gpu + T.gemm_gpu;
            T.factor_aux = T.getfk + T.extend_add + T.blas_gpu + T.other_gpu;
        }
    }
    if (st == MDGPU_OK) {
        t0 = now_s();
        if (!spL->i) spL->i = (int*)cs_malloc(lp[n], sizeof(int));
        if (!spL->x) spL->x = (FLOTANTE*)cs_malloc(lp[n], sizeof(FLOTANTE));
        std::vector<double> lx(lp[n]);
        st = mdgpu_export_L(num, nullptr, spL->i, lx.data());
        if (st == MDGPU_OK)
            for (int t = 0; t < lp[n]; t++) spL->x[t] = (FLOTANTE)lx[t];
        T.memcpy_l = now_s() - t0;
    }
    if (st != MDGPU_OK) report("factorization");
    t0 = now_s();
    if (d_ax) (void)hipFree(d_ax);
    mdgpu_numeric_destroy(num);
    mdgpu_symbolic_destroy(sym);
    T.free_gpu = now_s() - t0;
    T.factor = now_s() - t_start;
}

MDGPU_DROPIN_END

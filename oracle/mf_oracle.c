/*
 * CPU ORACLE -- test infrastructure only, never part of the product path.
 *
 * A plain-C FP64 restatement of the reference's CPU numeric path, used by
 * the tests as the checker and by bench.py as the timed CPU baseline:
 *   - multifrontal driver Factor()            mdgpu.cu:1394-1608
 *   - front assembly from A (MatrizGetFk_3)   mdgpu.cu:180-291
 *   - sorted union / relative maps            mdgpu.cu:535-595
 *   - full-square extend-add (ExtendAdd_CPU)  mdgpu.cu:748-767
 *   - partial factorization FactorAux_CPU_3   mdgpu.cu:1298-1361 with
 *     block_size = whole front (:1312); the dense kernels follow the
 *     Netlib reference BLAS loop orders the reference vendors:
 *     syr lower (ssyr.cpp:199-215), trsm Right/Lower/Trans/NonUnit
 *     (strsm.cpp:447-485), gemm N,T (sgemm.cpp:325-357)
 *   - L write-out into spL                    mdgpu.cu:1346-1352
 *   - flop count                              mdgpu.cu:1354-1359
 *   - cs_lsolve / cs_ltsolve / cs_gaxpy       cs.cpp:1069-1119
 *   - backward error                          mdgpu.cu:1972
 * All arithmetic is double (the reference's FLOTANTE=double branch,
 * types.h:12-21).  Fronts are dense full squares as in the reference, so
 * memory is O(max front^2).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    int n;        /* rows (= cols) of the dense front            */
    int k;        /* first pivot column                          */
    int offset;   /* pivot columns to eliminate                  */
    int *q;       /* global row ids, sorted                      */
    int *ql;      /* relative map into the destination front     */
    double *x;    /* n x n column-major                          */
} front_t;

static front_t *front_new(int n, int *q) {
    front_t *f = (front_t *)calloc(1, sizeof(front_t));
    f->n = n;
    f->q = q;
    f->x = (double *)calloc((size_t)n * (size_t)n > 0 ? (size_t)n * (size_t)n : 1, sizeof(double));
    return f;
}

static void front_free(front_t *f) {
    if (!f) return;
    free(f->q);
    free(f->ql);
    free(f->x);
    free(f);
}

static int cmp_int(const void *a, const void *b) {
    int x = *(const int *)a, y = *(const int *)b;
    return (x > y) - (x < y);
}

/* MatrizGetFk_3: rows = pivot block k..k+c-1 followed by the sorted unique
 * rows >= k+c of A(:, k..k+c-1); A values scattered (later duplicates
 * overwrite, :280). */
static front_t *get_fk(int n, const int *Ap, const int *Ai, const double *Ax, int k, int c, int *mark) {
    int cnt = c, j, p;
    int *rows;
    front_t *f;
    for (j = k; j < k + c; j++)
        for (p = Ap[j]; p < Ap[j + 1]; p++)
            if (Ai[p] >= k + c && mark[Ai[p]] != k) { mark[Ai[p]] = k; cnt++; }
    rows = (int *)malloc(sizeof(int) * (size_t)(cnt > 0 ? cnt : 1));
    for (j = 0; j < c; j++) rows[j] = k + j;
    {
        int t = c;
        for (j = k; j < k + c; j++)
            for (p = Ap[j]; p < Ap[j + 1]; p++)
                if (Ai[p] >= k + c && mark[Ai[p]] == k) { mark[Ai[p]] = -k - 2; rows[t++] = Ai[p]; }
    }
    qsort(rows + c, (size_t)(cnt - c), sizeof(int), cmp_int);
    f = front_new(cnt, rows);
    f->k = k;
    for (j = 0; j < c; j++) {
        int pos = 0;
        for (p = Ap[k + j]; p < Ap[k + j + 1]; p++) {
            while (rows[pos] < Ai[p]) pos++;
            f->x[(size_t)j * cnt + pos] = Ax[p];
        }
    }
    (void)n;
    return f;
}

/* VectorMerge: sorted set union */
static int *merge(const int *q1, int n1, const int *q2, int n2, int *nout) {
    int *qq = (int *)malloc(sizeof(int) * (size_t)(n1 + n2 > 0 ? n1 + n2 : 1));
    int a = 0, b = 0, nn = 0;
    while (a < n1 || b < n2) {
        int v;
        if (b >= n2 || (a < n1 && q1[a] < q2[b])) v = q1[a++];
        else v = q2[b++];
        if (nn == 0 || qq[nn - 1] != v) qq[nn++] = v;
    }
    *nout = nn;
    return qq;
}

/* VectorMergeIndices: position of each q[i] inside Q */
static int *merge_indices(const int *Q, const int *q, int n) {
    int *ql = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    int i = 0, j = 0;
    while (i < n) {
        if (q[i] == Q[j]) ql[i++] = j;
        j++;
    }
    return ql;
}

/* ExtendAdd_CPU over the trailing (n-offset)^2 square of src */
static void extend_add(const front_t *src, front_t *dst) {
    int m = src->n - src->offset, i, j;
    for (j = 0; j < m; j++) {
        double *dcol = dst->x + (size_t)src->ql[j] * dst->n;
        const double *scol = src->x + (size_t)(j + src->offset) * src->n + src->offset;
        for (i = 0; i < m; i++) dcol[src->ql[i]] += scol[i];
    }
}

/* FactorAux_CPU_3 with block = whole front */
static void factor_aux(front_t *U, const int *Lp, int *Li, double *Lx, double *flops, int *bad_col) {
    const int n = U->n, w = U->n, cols = U->offset, k = U->k;
    double *x = U->x;
    int c, i, j, l;
#define M(r, cc) x[(size_t)(cc) * w + (r)]
    for (c = 0; c < cols; c++) {
        double raiz;
        if (!(M(c, c) > 0) && *bad_col < 0) *bad_col = k + c;
        raiz = sqrt(M(c, c));
        for (j = c; j < cols; j++) M(j, c) = M(j, c) / raiz;
        /* syr lower, alpha = -1, x = column c below the diagonal, on the pivot block */
        for (j = c + 1; j < cols; j++) {
            if (M(j, c) != 0) {
                double temp = -M(j, c);
                for (i = j; i < cols; i++) M(i, j) += M(i, c) * temp;
            }
        }
    }
    if (n - cols > 0) {
        /* trsm Right, Lower, Trans, NonUnit: B(n-cols x cols) := B * inv(L11)^T */
        const int mrows = n - cols;
        for (l = 0; l < cols; l++) {
            double temp = 1.0 / M(l, l);
            for (i = 0; i < mrows; i++) M(cols + i, l) = temp * M(cols + i, l);
            for (j = l + 1; j < cols; j++) {
                if (M(j, l) != 0) {
                    temp = M(j, l);
                    for (i = 0; i < mrows; i++) M(cols + i, j) -= temp * M(cols + i, l);
                }
            }
        }
        /* gemm N,T on the full trailing square: C := C - L21 L21^T */
        for (j = 0; j < mrows; j++) {
            for (l = 0; l < cols; l++) {
                double b = M(cols + j, l);
                if (b != 0) {
                    double temp = -b;
                    for (i = 0; i < mrows; i++) M(cols + i, cols + j) += temp * M(cols + i, l);
                }
            }
        }
    }
    for (c = 0; c < cols; c++) {
        for (j = c; j < n; j++) {
            int t = Lp[k + c] + j - c;
            if (Li) Li[t] = U->q[j];
            if (Lx) Lx[t] = M(j, c);
        }
        *flops += (double)(n - c) + 1.0 + (double)(n - c) * (double)(n - c - 1);
    }
#undef M
}

/* Multifrontal factorization over a postordered node list (Factor,
 * mdgpu.cu:1394-1608).  Nodes: first column, column count, children (node
 * indices, kid_ptr/kids CSR).  A: the permuted lower triangle.  Every node
 * without a parent is a root and is finished like the reference's single
 * root (:1551-1607).  Returns the first non-positive pivot column or -1. */
int oracle_factor(int n, const int *Ap, const int *Ai, const double *Ax, int nnodes, const int *first,
                  const int *ncols, const int *kid_ptr, const int *kids, const int *Lp, int *Li, double *Lx,
                  double *flops_out) {
    front_t **upd = (front_t **)calloc((size_t)nnodes, sizeof(front_t *));
    int *mark = (int *)malloc(sizeof(int) * (size_t)n);
    int t, i, bad = -1;
    double flops = 0;
    for (t = 0; t < n; t++) mark[t] = -1;
    for (t = 0; t < nnodes; t++) {
        const int nk = kid_ptr[t + 1] - kid_ptr[t];
        int nq = 0;
        int *q = NULL;
        front_t **u2 = NULL;
        front_t *uk;
        if (nk == 0) continue;
        u2 = (front_t **)calloc((size_t)nk, sizeof(front_t *));
        for (i = 0; i < nk; i++) {
            const int ch = kids[kid_ptr[t] + i];
            front_t *F = get_fk(n, Ap, Ai, Ax, first[ch], ncols[ch], mark);
            front_t *U = upd[ch];
            upd[ch] = NULL;
            if (!U) {
                F->offset = ncols[ch];
                u2[i] = F;
            } else {
                int nm;
                int *qm = merge(F->q, F->n, U->q, U->n, &nm);
                front_t *W = front_new(nm, qm);
                F->ql = merge_indices(qm, F->q, F->n);
                U->ql = merge_indices(qm, U->q, U->n);
                W->k = first[ch];
                F->offset = 0;
                U->offset = 0;
                extend_add(F, W);
                extend_add(U, W);
                W->offset = ncols[ch];
                front_free(F);
                front_free(U);
                u2[i] = W;
            }
            factor_aux(u2[i], Lp, Li, Lx, &flops, &bad);
        }
        /* union of the children's update rows, then the accumulated update */
        nq = u2[0]->n - u2[0]->offset;
        q = (int *)malloc(sizeof(int) * (size_t)(nq > 0 ? nq : 1));
        memcpy(q, u2[0]->q + u2[0]->offset, sizeof(int) * (size_t)nq);
        for (i = 1; i < nk; i++) {
            int n2;
            int *q2 = merge(q, nq, u2[i]->q + u2[i]->offset, u2[i]->n - u2[i]->offset, &n2);
            free(q);
            q = q2;
            nq = n2;
        }
        uk = front_new(nq, q);
        uk->k = first[t];
        for (i = 0; i < nk; i++) {
            u2[i]->ql = merge_indices(q, u2[i]->q + u2[i]->offset, u2[i]->n - u2[i]->offset);
            extend_add(u2[i], uk);
            front_free(u2[i]);
        }
        free(u2);
        upd[t] = uk;
    }
    /* roots: the last node of each tree (a forest has several) */
    {
        char *has_parent = (char *)calloc((size_t)nnodes, 1);
        int r;
        for (t = 0; t < kid_ptr[nnodes]; t++) has_parent[kids[t]] = 1;
        for (r = 0; r < nnodes; r++) {
            front_t *F, *U;
            if (has_parent[r]) continue;
            F = get_fk(n, Ap, Ai, Ax, first[r], ncols[r], mark);
            U = upd[r];
            upd[r] = NULL;
            if (U) {
                int nm;
                int *qm = merge(U->q, U->n, F->q, F->n, &nm);
                front_t *W = front_new(nm, qm);
                U->ql = merge_indices(qm, U->q, U->n);
                F->ql = merge_indices(qm, F->q, F->n);
                W->k = first[r];
                U->offset = 0;
                F->offset = 0;
                extend_add(F, W);
                extend_add(U, W);
                front_free(F);
                front_free(U);
                F = W;
            }
            F->offset = ncols[r];
            factor_aux(F, Lp, Li, Lx, &flops, &bad);
            front_free(F);
        }
        free(has_parent);
    }
    for (t = 0; t < nnodes; t++) front_free(upd[t]);
    free(upd);
    free(mark);
    if (flops_out) *flops_out = flops;
    return bad;
}

/* cs_lsolve: L x = b, diagonal first in each column */
int oracle_lsolve(int n, const int *Lp, const int *Li, const double *Lx, double *x) {
    int j, p;
    for (j = 0; j < n; j++) {
        x[j] /= Lx[Lp[j]];
        for (p = Lp[j] + 1; p < Lp[j + 1]; p++) x[Li[p]] -= Lx[p] * x[j];
    }
    return 1;
}

/* cs_ltsolve: L^T x = b */
int oracle_ltsolve(int n, const int *Lp, const int *Li, const double *Lx, double *x) {
    int j, p;
    for (j = n - 1; j >= 0; j--) {
        for (p = Lp[j] + 1; p < Lp[j + 1]; p++) x[j] -= Lx[p] * x[Li[p]];
        x[j] /= Lx[Lp[j]];
    }
    return 1;
}

/* backward error ||b - A x||_inf / (||A||_1 ||x||_inf + ||b||_inf) with A
 * given as its lower triangle (symmetric) */
double oracle_backward_error(int n, const int *Ap, const int *Ai, const double *Ax, const double *x,
                             const double *b) {
    double *r = (double *)malloc(sizeof(double) * (size_t)n);
    double *colsum = (double *)calloc((size_t)n, sizeof(double));
    double rn = 0, xn = 0, bn = 0, an = 0;
    int j, p;
    for (j = 0; j < n; j++) r[j] = -b[j];
    for (j = 0; j < n; j++)
        for (p = Ap[j]; p < Ap[j + 1]; p++) {
            int i = Ai[p];
            r[i] += Ax[p] * x[j];
            colsum[j] += fabs(Ax[p]);
            if (i != j) { r[j] += Ax[p] * x[i]; colsum[i] += fabs(Ax[p]); }
        }
    for (j = 0; j < n; j++) {
        if (fabs(r[j]) > rn) rn = fabs(r[j]);
        if (fabs(x[j]) > xn) xn = fabs(x[j]);
        if (fabs(b[j]) > bn) bn = fabs(b[j]);
        if (colsum[j] > an) an = colsum[j];
    }
    free(r);
    free(colsum);
    return rn / (an * xn + bn);
}

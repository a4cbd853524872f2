/* Drop-in CSparse interface of the reference (cs.h of the reference: the
 * struct layouts of cs/css/csn and the routines its driver calls), served
 * by this library.  Layouts and signatures match the reference; the
 * routines are this library's own implementations.  Only the functions the
 * reference actually defines in cs.cpp are provided (cs_chol, cs_schol,
 * cs_cholsol, cs_lu, cs_qr, ... are declared there but never defined). */
#ifndef _CS_H
#define _CS_H

#include "types.h"

#include <limits.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

MDGPU_DROPIN_BEGIN

typedef struct cs_sparse { /* compressed-column (nz == -1) or triplet form */
    int nzmax;             /* capacity of i and x                        */
    int m;                 /* rows                                       */
    int n;                 /* columns                                    */
    int *p;                /* column pointers (n+1) or column indices    */
    int *i;                /* row indices                                */
    FLOTANTE *x;           /* values, or NULL for a pattern              */
    int nz;                /* triplet entries, -1 for compressed column  */
} cs;

typedef struct cs_symbolic {
    int *pinv;
    int *q;
    int *parent;
    int *cp;
    int *leftmost;
    int m2;
    FLOTANTE lnz;
    FLOTANTE unz;
} css;

typedef struct cs_numeric {
    cs *L;
    cs *U;
    int *pinv;
    FLOTANTE *B;
} csn;

/* matrices */
cs *cs_spalloc(int m, int n, int nzmax, int values, int triplet);
cs *cs_spfree(cs *A);
int cs_sprealloc(cs *A, int nzmax);
int cs_entry(cs *T, int i, int j, FLOTANTE x);
cs *cs_compress(const cs *T);
cs *cs_load(FILE *f);
/* Matrix Market coordinate file -> lower-triangle triplets (this library's
 * addition; cs_load reads the reference's headerless "i j x" lines). */
cs *cs_load_mm(FILE *f);
cs *cs_transpose(const cs *A, int values);
cs *cs_add(const cs *A, const cs *B, FLOTANTE alpha, FLOTANTE beta);
int cs_scatter(const cs *A, int j, FLOTANTE beta, int *w, FLOTANTE *x, int mark, cs *C, int nz);
int cs_fkeep(cs *A, int (*fkeep)(int, int, FLOTANTE, void *), void *other);
int cs_dupl(cs *A);
int cs_droptol(cs *A, FLOTANTE tol);
FLOTANTE cs_norm(const cs *A);
int cs_print(const cs *A, int brief);
cs *cs_permute(const cs *A, const int *pinv, const int *q, int values);
cs *cs_symperm(const cs *A, const int *pinv, int values);
/* ordering and symbolic analysis */
int *cs_amd(int order, const cs *A);
int *cs_pinv(const int *p, int n);
int *cs_etree(const cs *A, int ata);
int *cs_post(const int *parent, int n);
int cs_tdfs(int j, int k, int *head, const int *next, int *post, int *stack);
int *cs_counts(const cs *A, const int *parent, const int *post, int ata);
int cs_leaf(int i, int j, const int *first, int *maxfirst, int *prevleaf, int *ancestor, int *jleaf);
FLOTANTE cs_cumsum(int *p, int *c, int n);
/* solves and products */
int cs_gaxpy(const cs *A, const FLOTANTE *x, FLOTANTE *y);
int cs_lsolve(const cs *L, FLOTANTE *x);
int cs_ltsolve(const cs *L, FLOTANTE *x);
int cs_pvec(const int *p, const FLOTANTE *b, FLOTANTE *x, int n);
int cs_pvec(const int *p, const int *b, int *x, int n);
int cs_ipvec(const int *p, const FLOTANTE *b, FLOTANTE *x, int n);
/* memory */
void *cs_malloc(int n, size_t size);
void *cs_calloc(int n, size_t size);
void *cs_free(void *p);
void *cs_realloc(void *p, int n, size_t size, int *ok);
cs *cs_done(cs *C, void *w, void *x, int ok);
int *cs_idone(int *p, cs *C, void *w, int ok);

MDGPU_DROPIN_END

#define CS_MAX(a, b) (((a) > (b)) ? (a) : (b))
#define CS_MIN(a, b) (((a) < (b)) ? (a) : (b))
#define CS_FLIP(i) (-(i)-2)
#define CS_UNFLIP(i) (((i) < 0) ? CS_FLIP(i) : (i))
#define CS_MARKED(w, j) (w[j] < 0)
#define CS_MARK(w, j) { w[j] = CS_FLIP(w[j]); }
#define CS_CSC(A) (A && (A->nz == -1))
#define CS_TRIPLET(A) (A && (A->nz >= 0))

#endif

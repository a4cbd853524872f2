/* Drop-in mdgpu interface (mdgpu.h of the reference plus the Factor()
 * entry point defined in its mdgpu.cu:1394).  Factor() runs the numeric
 * multifrontal factorization on the GPU through the C ABI (mdgpu_c.h) and
 * fills spL->i and spL->x; spL->x is allocated here when the caller left it
 * NULL (the reference never allocates it, mdgpu.cu:1873). */
#ifndef _MDGPU_H
#define _MDGPU_H

#include <time.h>

#include "cs.h"
#include "etree.h"
#include "types.h"

MDGPU_DROPIN_BEGIN

typedef struct {
    int w;          /* leading dimension (rows allocated)  */
    int n;          /* rows = columns of the front          */
    FLOTANTE *x;    /* w x w column-major                   */
    int *q;         /* global row ids                       */
    int *ql;        /* relative map into the parent         */
    int k;          /* first pivot column                   */
    int offset;     /* pivot columns                        */
} Matriz;

typedef struct {
    bool GPU;
    Matriz *h_frente;
    Matriz *hd_frente;
    clock_t tiempoFact;
} Frente;

/* Globals of the reference driver (mdgpu.cu:29, 59-60).  GPU and
 * block_size do not change this library's behaviour (it always factors on
 * the GPU with its own blocking); pad selects the padded leading dimension
 * in MatrizAlloc. */
extern bool GPU;
extern int block_size;
extern bool pad;

void MatrizAlloc(Matriz *m);

inline FLOTANTE MatrizGet(Matriz *m, int i, int j) { return m->x[j * (m->w) + i]; }
inline void MatrizSet(Matriz *m, int i, int j, FLOTANTE v) { m->x[j * (m->w) + i] = v; }
inline void MatrizAdd(Matriz *m, int i, int j, FLOTANTE v) { m->x[j * (m->w) + i] += v; }
#define _MatrizSet(m, i, j, v) { m->x[(j) * (m->w) + (i)] = (v); }
FLOTANTE MatrizGet(cs *M, int i, int j);
void MatrizSet(cs *M, int i, int j, FLOTANTE v);

void VectorPrint(FILE *f, int *v, int n);
void VectorPrint(int *v, int n);
void MatrizPrint(FILE *f, Matriz *m, const char *format);
void MatrizPrint(Matriz *m, const char *format);
void MatrizPrint(FILE *f, cs *m, const char *format);
void MatrizPrint(cs *m, const char *format);
void MatrizSpy(Matriz *m);
void MatrizSpy(cs *m);
void MatrizSpy2(Matriz *A, Matriz *L);
int cs_print(FILE *f, const cs *A, int brief);

/* Numeric multifrontal factorization (mdgpu.cu:1394-1608 of the reference):
 * A = the permuted lower triangle, spL = L skeleton with p = cumsum(counts),
 * listaNodos = postordered node list (root(s) last).  On failure spL is left
 * unfilled and mdgpu_last_error() describes why. */
void Factor(cs *A, cs *spL, struct ETree **listaNodos, int cantNodos);

/* Phase times in seconds of the last Factor() call of the calling thread:
 * the line items of the reference driver's report (mdgpu.cu:1901-1930),
 * which it kept in global tick counters.  No counterpart as a function in
 * the reference. */
typedef struct {
    double mflops;      /* MFlops: sum over columns of (c_j^2 + 1), padded counts (mdgpu.cu:1354-1359) */
    double factor;      /* Factor: the whole call                                   */
    double factor_aux;  /* FactorAux: GPU compute (sum of the kernel classes below) */
    double blas_gpu;    /* BLAS GPU: dense front kernels                            */
    double trsm_gpu;    /*   TRSM GPU: potrf + TRSM (LDS panels and blocked path)   */
    double gemm_gpu;    /*   GEMM GPU: MFMA Schur-complement updates                */
    double other_gpu;   /*   Otros: diagonal-block inverses used by the solves      */
    double memcpy_x;    /* MemcpyX: A's values host -> device                       */
    double memcpy_l;    /* Memcpy2: L device -> host (spL)                          */
    double merge;       /* Merge: front row lists and relative maps (host)          */
    double getfk;       /* GetFk: front assembly from A on the GPU                  */
    double extend_add;  /* ExtendAdd                                                */
    double malloc_gpu;  /* Malloc GPU: schedule and device arena                    */
    double free_gpu;    /* Free GPU                                                 */
} FactorTimes;
void FactorGetTimes(FactorTimes *out);

MDGPU_DROPIN_END

#endif

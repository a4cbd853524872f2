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

MDGPU_DROPIN_END

#endif

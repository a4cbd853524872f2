/* Drop-in node trees of the reference (etree.h of the reference): one
 * ETree per node of the elimination tree or of a (relaxed) supernode
 * partition, in the reference's layout. */
#ifndef _ETREE_H
#define _ETREE_H

#include "cs.h"

MDGPU_DROPIN_BEGIN

struct ETree {
    int nodo;               /* first column of the node                 */
    int nHijos;             /* number of children                       */
    struct ETree **hijos;   /* children                                 */
    int rows;               /* rows of the node's top column            */
    int cols;               /* columns in the node                      */
    int *nodos;             /* original columns (relaxed partition)     */
    int ceros;              /* explicit zeros added by amalgamation     */
};

/* Elimination tree nodes in postorder (etree.cpp:12-48 of the reference);
 * returns NULL and prints an error for a forest, as the reference stops. */
struct ETree *GetETree(cs *m, int *parent, int *post, int *count, struct ETree ***listaNodos, int *cantNodos);
/* Fundamental supernodes (etree.cpp:50-118).  Children are stored in
 * processing order. */
void GetSuperETree(struct ETree **listaNodos, int cantNodos, struct ETree ***listaNodosSuper, int *cantNodosSuper);
/* Relaxed supernodes under a zero budget; pads count[] and returns the new
 * permutation (etree.cpp:194-342). */
void GetRelaxedETree(struct ETree **listaNodosSuper, int cantNodosSuper, int max_ceros, int *count, int **perm,
                     struct ETree ***listaNodosRelaxed_out, int *cantNodosRelaxed_out);
void PrintETree(struct ETree *root, int nivel);

MDGPU_DROPIN_END

#endif

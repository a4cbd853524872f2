/* Scalar type of the drop-in CSparse/mdgpu interface.  The reference picks
 * float or double at compile time (types.h:3-23 of the reference); this
 * library is built with double, the reference's FP64 branch, and the
 * precision of the GPU factorization is a run-time choice of the C ABI
 * (mdgpu_c.h).  Code including these headers must use the same FLOTANTE as
 * the library build (double). */
#ifndef _TYPES_H
#define _TYPES_H

#ifndef FLOTANTE
#define FLOTANTE double
#endif

#endif

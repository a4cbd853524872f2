/* Scalar type of the drop-in CSparse/mdgpu interface.  As in the reference
 * (types.h:3-23 of the reference) FLOTANTE is float unless the build
 * defines it; the library serves both FLOTANTE = float and FLOTANTE = double.
 *
 * Every C++ declaration of cs.h, etree.h and mdgpu.h sits in the inline
 * namespace MDGPU_DROPIN_NS, named after FLOTANTE (mdgpu_flotante_float or
 * mdgpu_flotante_double).  Callers write the reference's unqualified names
 * (cs, cs_load, GetETree, Factor, ...) as before, but the linker sees
 * FLOTANTE in every symbol: a caller built with float binds the float
 * library functions (float A->x, float spL->x), one built with double the
 * double ones, and one built with any other FLOTANTE fails to link instead
 * of misreading the value arrays.  The precision of the GPU factorization
 * follows FLOTANTE (MDGPU_FP32 for float, MDGPU_FP64 for double, as the
 * reference's arithmetic does); the C ABI (mdgpu_c.h) chooses it at run
 * time. */
#ifndef _TYPES_H
#define _TYPES_H

#ifndef FLOTANTE
#define FLOTANTE float
#endif

#define MDGPU_NS_CAT_(a, b) a##b
#define MDGPU_NS_CAT(a, b) MDGPU_NS_CAT_(a, b)
#define MDGPU_DROPIN_NS MDGPU_NS_CAT(mdgpu_flotante_, FLOTANTE)
#define MDGPU_DROPIN_BEGIN inline namespace MDGPU_DROPIN_NS {
#define MDGPU_DROPIN_END }

#endif

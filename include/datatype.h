/* Precision macros of the plan API.
 * Replaces include/datatype.h:17-31 of the reference (PCS = double,
 * CUCPX = cuDoubleComplex, CPX = std::complex<double>); the device complex type
 * is HIP's hipDoubleComplex, layout-compatible with double[2]. */
#ifndef CURIG_DATATYPE_H
#define CURIG_DATATYPE_H

#include <hip/hip_complex.h>

#ifdef __cplusplus
#include <complex>
#define COMPLEX(X) std::complex<X>
#else
#define COMPLEX(X) X _Complex
#endif

#undef PCS
#undef CPX
#define PCS double
#define CUCPX hipDoubleComplex
#define CPX COMPLEX(PCS)

#endif

// Test-only stand-in for the two GSL types the GPUQPI interface passes
// (gsl_matrix_complex: row-major complex data with a leading dimension tda;
// gsl_complex), declared from their documented layout.  Scaffolding for
// compiling the plug-in shim outside the calcQPI tree; not GSL.
#pragma once
#include <cstddef>

struct gsl_matrix_complex {
    size_t size1, size2, tda;
    double *data;  // element (i, j) at data[2 * (i * tda + j)], re then im
};
struct gsl_complex {
    double dat[2];
};
#define GSL_REAL(z) ((z).dat[0])
#define GSL_IMAG(z) ((z).dat[1])

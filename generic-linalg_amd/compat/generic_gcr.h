// generic_gcr.h -- GCR and GCR(n) (generic_gcr.h:17-22 of the reference),
// solved on the GPU by libglx.so.
#ifndef GLX_COMPAT_GCR
#define GLX_COMPAT_GCR

#include <complex>
#include <string>
using std::complex;

#include "inverter_struct.h"
#include "verbosity.h"

inversion_info minv_vector_gcr(double* phi, double* phi0, int size, int max_iter, double res,
                               void (*matrix_vector)(double*, double*, void*), void* extra_info,
                               inversion_verbose_struct* verbosity = 0);
inversion_info minv_vector_gcr(complex<double>* phi, complex<double>* phi0, int size, int max_iter, double res,
                               void (*matrix_vector)(complex<double>*, complex<double>*, void*), void* extra_info,
                               inversion_verbose_struct* verbosity = 0);
inversion_info minv_vector_gcr_restart(double* phi, double* phi0, int size, int max_iter, double res, int restart_freq,
                                       void (*matrix_vector)(double*, double*, void*), void* extra_info,
                                       inversion_verbose_struct* verbosity = 0);
inversion_info minv_vector_gcr_restart(complex<double>* phi, complex<double>* phi0, int size, int max_iter,
                                       double res, int restart_freq,
                                       void (*matrix_vector)(complex<double>*, complex<double>*, void*),
                                       void* extra_info, inversion_verbose_struct* verbosity = 0);

#endif

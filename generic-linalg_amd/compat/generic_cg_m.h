// generic_cg_m.h -- multishift CG (generic_cg_m.h:16-17 of the reference),
// solved on the GPU by libglx.so.
#ifndef GLX_COMPAT_CG_M
#define GLX_COMPAT_CG_M

#include <complex>
#include <string>
using std::complex;

#include "inverter_struct.h"
#include "verbosity.h"

inversion_info minv_vector_cg_m(double** phi, double* phi0, int n_shift, int size, int resid_freq_check, int max_iter,
                                double eps, double* shifts, void (*matrix_vector)(double*, double*, void*),
                                void* extra_info, bool worst_first = false, inversion_verbose_struct* verbosity = 0);
inversion_info minv_vector_cg_m(complex<double>** phi, complex<double>* phi0, int n_shift, int size,
                                int resid_freq_check, int max_iter, double eps, double* shifts,
                                void (*matrix_vector)(complex<double>*, complex<double>*, void*), void* extra_info,
                                bool worst_first = false, inversion_verbose_struct* verbosity = 0);

#endif

// generic_bicgstab.h -- BiCGStab and BiCGStab(n) (generic_bicgstab.h:17-21 of
// the reference), solved on the GPU by libglx.so.
#ifndef GLX_COMPAT_BICGSTAB
#define GLX_COMPAT_BICGSTAB

#include <complex>
#include <string>
using std::complex;

#include "inverter_struct.h"
#include "verbosity.h"

inversion_info minv_vector_bicgstab(double* phi, double* phi0, int size, int max_iter, double res,
                                    void (*matrix_vector)(double*, double*, void*), void* extra_info,
                                    inversion_verbose_struct* verbosity = 0);
inversion_info minv_vector_bicgstab_restart(double* phi, double* phi0, int size, int max_iter, double res,
                                            int restart_freq, void (*matrix_vector)(double*, double*, void*),
                                            void* extra_info, inversion_verbose_struct* verbosity = 0);
inversion_info minv_vector_bicgstab(complex<double>* phi, complex<double>* phi0, int size, int max_iter, double res,
                                    void (*matrix_vector)(complex<double>*, complex<double>*, void*),
                                    void* extra_info, inversion_verbose_struct* verbosity = 0);
inversion_info minv_vector_bicgstab_restart(complex<double>* phi, complex<double>* phi0, int size, int max_iter,
                                            double res, int restart_freq,
                                            void (*matrix_vector)(complex<double>*, complex<double>*, void*),
                                            void* extra_info, inversion_verbose_struct* verbosity = 0);

#endif

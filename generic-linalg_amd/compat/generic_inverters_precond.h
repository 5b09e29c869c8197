// generic_inverters_precond.h -- the reference's preconditioned dispatcher
// interface (generic_inverters_precond.h:31-63 of the reference: the
// minv_inverter_precond enum, minv_inverter_precond_params and the two
// minv_preconditioned overloads, same names, values and fields), so that
// drivers including it -- square_laplace.cpp does -- compile unchanged
// against these headers.  The preconditioned solvers themselves (PCG,
// flexible PCG, variably preconditioned GCR, preconditioned BiCGStab) are not
// part of this library: minv_preconditioned throws std::runtime_error for
// every solver type, as minv_unpreconditioned does for the unpreconditioned
// solvers it does not carry; MINV_PRE_INVALID returns a default
// inversion_info.
#ifndef GLX_COMPAT_INVERTERS_PRECOND
#define GLX_COMPAT_INVERTERS_PRECOND

#include <complex>

#include "inverter_struct.h"
#include "verbosity.h"

enum minv_inverter_precond {
  MINV_PRE_CG = 0,
  MINV_PRE_FPCG = 1,
  MINV_PRE_VPGCR = 2,
  MINV_PRE_BICGSTAB = 3,
  MINV_PRE_INVALID = -1,
};

struct minv_inverter_precond_params {
  double tol;
  int max_iters;
  bool restart;
  int restart_freq;
};

inversion_info minv_preconditioned(double* lhs, double* rhs, int size, minv_inverter_precond type,
                                   minv_inverter_precond_params& params,
                                   void (*matrix_vector)(double*, double*, void*), void* extra_info,
                                   void (*precond_matrix_vector)(double*, double*, int, void*,
                                                                 inversion_verbose_struct*),
                                   void* precond_info, inversion_verbose_struct* verbosity = 0);
inversion_info minv_preconditioned(std::complex<double>* lhs, std::complex<double>* rhs, int size,
                                   minv_inverter_precond type, minv_inverter_precond_params& params,
                                   void (*matrix_vector)(std::complex<double>*, std::complex<double>*, void*),
                                   void* extra_info,
                                   void (*precond_matrix_vector)(std::complex<double>*, std::complex<double>*, int,
                                                                 void*, inversion_verbose_struct*),
                                   void* precond_info, inversion_verbose_struct* verbosity = 0);

#endif

// generic_inverters.h -- the solvers of the GPU path in one include, and the
// reference's dispatcher minv_unpreconditioned with its solver enum and
// parameter struct (generic_inverters.h:70-111 of the reference; same names,
// values and fields).  CG, GCR and BiCGStab (and their restarts) run on the
// GPU; the reference's other solvers (CR, BiCGStab(l), GMRES, SOR, MinRes)
// are not part of this library and make the dispatcher throw
// std::runtime_error.  MINV_INVALID returns a default inversion_info, as in
// the reference.
#ifndef GLX_COMPAT_INVERTERS
#define GLX_COMPAT_INVERTERS

#include "generic_bicgstab.h"
#include "generic_cg.h"
#include "generic_cg_m.h"
#include "generic_gcr.h"

enum minv_inverter {
  MINV_CG = 0,
  MINV_CR = 1,
  MINV_GCR = 2,
  MINV_BICGSTAB = 3,
  MINV_BICGSTAB_L = 4,
  MINV_GMRES = 5,
  MINV_SOR = 6,
  MINV_MINRES = 7,
  MINV_INVALID = -1,
};

struct minv_inverter_params {
  // general
  double tol;
  int max_iters;
  bool restart;
  int restart_freq;
  // solver-specific (read only by solvers outside this library)
  double sor_omega;
  double minres_omega;
  int bicgstabl_l;
};

inversion_info minv_unpreconditioned(double* lhs, double* rhs, int size, minv_inverter type,
                                     minv_inverter_params& params, void (*matrix_vector)(double*, double*, void*),
                                     void* extra_info, inversion_verbose_struct* verbosity = 0);
inversion_info minv_unpreconditioned(complex<double>* lhs, complex<double>* rhs, int size, minv_inverter type,
                                     minv_inverter_params& params,
                                     void (*matrix_vector)(complex<double>*, complex<double>*, void*),
                                     void* extra_info, inversion_verbose_struct* verbosity = 0);

#endif

// operators_stencil.h -- stencil builders and even/odd helpers of the
// reference (operator_utils/operators_stencil.h:1-40): same names and
// arguments.  The builders fill an allocated, empty nc = 1 stencil_2d from a
// staggered_u1_op; the helpers run on the GPU.
#ifndef GLX_COMPAT_OPERATORS_STENCIL
#define GLX_COMPAT_OPERATORS_STENCIL

#include "coarse_stencil.h"
#include "lattice.h"
#include "operators.h"

void get_square_staggered_u1_stencil(stencil_2d* stenc, staggered_u1_op* stagif);
void get_square_staggered_gamma5_u1_stencil(stencil_2d* stenc, staggered_u1_op* stagif);
void get_square_staggered_dagger_u1_stencil(stencil_2d* stenc, staggered_u1_op* stagif);
// rhs_e = shift rhs_orig - D_eo rhs_orig on even sites, 0 on odd sites
void apply_square_staggered_eoprec_prepare_stencil(complex<double>* rhs_e, complex<double>* rhs_orig,
                                                   stencil_2d* stenc);
// lhs = shift^2 rhs - D_eo D_oe rhs on even sites, 0 on odd sites
void apply_square_staggered_m2mdeodoe_stencil(complex<double>* lhs, complex<double>* rhs, void* extra_data);
// odd sites: (rhs_o - D_oe lhs_e) / Re(shift); even sites: lhs_e
void apply_square_staggered_eoprec_reconstruct_stencil(complex<double>* lhs_full, complex<double>* lhs_e,
                                                       complex<double>* rhs_o, stencil_2d* stenc);

#endif

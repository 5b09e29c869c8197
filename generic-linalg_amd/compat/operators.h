// operators.h -- the 2D lattice operators of the solver path
// (operators.h:7-83 of the reference).  Each is a callback
// lhs = A rhs with extra_data a staggered_u1_op*; passed to a solver, it is
// recognised and the whole solve runs on the GPU; called directly, it applies
// the operator on the GPU.
#ifndef GLX_COMPAT_OPERATORS
#define GLX_COMPAT_OPERATORS

#include <complex>
using namespace std;

struct staggered_u1_op {
  complex<double>* lattice;  // links, lattice[2*(y*x_fine + x) + mu]
  double mass;               // staggered mass (Laplacians: the m^2 shift)
  int x_fine;
  int y_fine;
  int Nc;                    // colours of the Laplacian
  double wilson_coeff;
};

enum op_type {
  STAGGERED = 0,
  LAPLACE = 1,
  LAPLACE_NC2 = 2,
  G5_STAGGERED = 3,
  STAGGERED_NORMAL = 4,
  STAGGERED_INDEX = 5,
};

// maximum number of hops of an operator's stencil (operators.cpp:8-23)
int get_stencil_size(op_type opt);

void square_laplace(complex<double>* lhs, complex<double>* rhs, void* extra_data);
void square_laplace_u1(complex<double>* lhs, complex<double>* rhs, void* extra_data);
void square_staggered_u1(complex<double>* lhs, complex<double>* rhs, void* extra_data);
void gamma_5(complex<double>* lhs, complex<double>* rhs, void* extra_data);
void square_staggered_gamma5_u1(complex<double>* lhs, complex<double>* rhs, void* extra_data);
void square_staggered_dagger_u1(complex<double>* lhs, complex<double>* rhs, void* extra_data);
void square_staggered_normal_u1(complex<double>* lhs, complex<double>* rhs, void* extra_data);
void square_staggered_deo_u1(complex<double>* lhs, complex<double>* rhs, void* extra_data);
void square_staggered_doe_u1(complex<double>* lhs, complex<double>* rhs, void* extra_data);
void square_staggered_eoprec_prepare(complex<double>* rhs_e, complex<double>* rhs_orig, void* extra_data);
void square_staggered_m2mdeodoe_u1(complex<double>* lhs, complex<double>* rhs, void* extra_data);
void square_staggered_eoprec_reconstruct(complex<double>* lhs_full, complex<double>* lhs_e, complex<double>* rhs_o,
                                         void* extra_data);

// Real Laplacian with a runtime lattice (the multishift test driver's
// square_laplace, tests/multishift/multishift.cpp:634-672).
void square_laplace(double* lhs, double* rhs, void* extra_data);

// the operators off the solver path (operators.cpp:127-846), applied on the
// GPU as general stencils
void square_staggered(complex<double>* lhs, complex<double>* rhs, void* extra_data);
void square_staggered_gamma5(complex<double>* lhs, complex<double>* rhs, void* extra_data);
// staggered D plus wilson_coeff times the two-link Laplacian
void square_staggered_2linklaplace_u1(complex<double>* lhs, complex<double>* rhs, void* extra_data);
void staggered_symmshift_x(complex<double>* lhs, complex<double>* rhs, void* extra_data);
void staggered_symmshift_y(complex<double>* lhs, complex<double>* rhs, void* extra_data);
// i D(m = 0) - m Gamma_5 with Gamma_5 = (i/2)(G_x G_y - G_y G_x) (arXiv 1410.5733)
void staggered_index_operator(complex<double>* lhs, complex<double>* rhs, void* extra_data);

#endif

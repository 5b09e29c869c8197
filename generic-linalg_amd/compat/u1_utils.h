// u1_utils.h -- U(1) gauge-field utilities of the reference's harness
// (u1_utils.h / u1_utils.cpp of the reference: same names, arguments and
// results), implemented on the host in libglx.so with the same C++ standard
// library generators, so a seeded std::mt19937 gives the same fields as the
// reference.  Layout: gauge_field[2 * (y * x_len + x) + mu], mu = 0 (x), 1 (y).
// Device-side generation and plaquette: glx_lattice_create_gaussian and
// glx_lattice_plaquette in glx.h.
#ifndef GLX_COMPAT_U1_UTILS
#define GLX_COMPAT_U1_UTILS

#include <complex>
#include <random>
#include <string>
using std::complex;
using std::string;

enum gauge_create_type {
  GAUGE_LOAD = 0,
  GAUGE_RANDOM = 1,
  GAUGE_UNIT = 2
};

// text format: one phase per line, x slowest, then y, then mu (u1_utils.cpp:17-65)
void read_gauge_u1(complex<double>* gauge_field, int x_len, int y_len, string input_file);
void write_gauge_u1(complex<double>* gauge_field, int x_len, int y_len, string output_file);
void unit_gauge_u1(complex<double>* gauge_field, int x_len, int y_len);
// phases uniform on (-pi, pi)
void rand_gauge_u1(complex<double>* gauge_field, int x_len, int y_len, std::mt19937& generator);
// phases N(0, 1/beta); beta = 0 reproduces the reference's hot start followed
// by a draw with infinite deviation (u1_utils.cpp:98-110)
void gauss_gauge_u1(complex<double>* gauge_field, int x_len, int y_len, std::mt19937& generator, double beta);
void rand_trans_u1(complex<double>* gauge_trans, int x_len, int y_len, std::mt19937& generator);
// u_mu(x) = g(x) u_mu(x) g^*(x + mu)
void apply_gauge_trans_u1(complex<double>* gauge_field, complex<double>* gauge_trans, int x_len, int y_len);
// n_iter rounds of APE smearing (staple weight alpha), each projected back to
// U(1); smeared_field receives the result (u1_utils.cpp:145-187)
void apply_ape_smear_u1(complex<double>* smeared_field, complex<double>* gauge_field, int x_len, int y_len,
                        double alpha, int n_iter);
complex<double> get_plaquette_u1(complex<double>* gauge_field, int x_len, int y_len);
// sum of plaquette angles / (2 pi) (u1_utils.cpp:210-226)
double get_topo_u1(complex<double>* gauge_field, int x_len, int y_len);

#endif

// glx_compat.cpp -- the reference's C++ interface (generic_cg.h,
// generic_cg_m.h, generic_bicgstab.h, generic_gcr.h, operators.h) implemented
// over the C ABI, so that existing drivers recompile against
// generic-linalg_amd/compat/ and link libglx.so in place of the reference's
// object files.  The operator functions defined here are registered as device
// operators: a solver given one of them runs entirely on the GPU.
#include <algorithm>
#include <cmath>
#include <complex>
#include <fstream>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/glx.h"
#include "../compat/generic_inverters.h"
#include "../compat/generic_inverters_precond.h"
#include "../compat/operators.h"
#include "../compat/u1_utils.h"
#include "../compat/coarse_stencil.h"
#include "../compat/operators_stencil.h"
#include "glx_internal.h"

namespace {

typedef void (*rmv)(double*, double*, void*);
typedef void (*cmv)(complex<double>*, complex<double>*, void*);

glx_verbose* to_c(inversion_verbose_struct* v, glx_verbose& out) {
  std::memset(&out, 0, sizeof(out));
  if (!v) return nullptr;
  out.verbosity = v->verbosity;
  out.verb_prefix = v->verb_prefix.c_str();
  out.precond_verbosity = v->precond_verbosity;
  out.precond_verb_prefix = v->precond_verb_prefix.c_str();
  return &out;
}

void check(int rc, const char* what) {
  if (rc != GLX_OK) throw std::runtime_error(std::string(what) + ": " + glx_last_error());
}

inversion_info to_info(const glx_inversion_info& c) {
  inversion_info r;
  r.resSq = c.resSq;
  r.iter = c.iter;
  r.success = c.success != 0;
  r.name = c.name;
  r.ops_count = c.ops_count;
  return r;
}

double* D(complex<double>* p) { return reinterpret_cast<double*>(p); }
glx_matvec M(rmv f) { return reinterpret_cast<glx_matvec>(f); }
glx_matvec M(cmv f) { return reinterpret_cast<glx_matvec>(f); }

typedef int (*csolve)(double*, double*, int, int, double, glx_matvec, void*, glx_verbose*, glx_inversion_info*);
typedef int (*crestart)(double*, double*, int, int, double, int, glx_matvec, void*, glx_verbose*,
                        glx_inversion_info*);

inversion_info call(csolve f, const char* what, double* phi, double* phi0, int size, int max_iter, double res,
                    glx_matvec mv, void* extra, inversion_verbose_struct* verb) {
  glx_verbose cv;
  glx_inversion_info ci;
  std::memset(&ci, 0, sizeof(ci));
  check(f(phi, phi0, size, max_iter, res, mv, extra, to_c(verb, cv), &ci), what);
  return to_info(ci);
}
inversion_info call_r(crestart f, const char* what, double* phi, double* phi0, int size, int max_iter, double res,
                      int restart_freq, glx_matvec mv, void* extra, inversion_verbose_struct* verb) {
  glx_verbose cv;
  glx_inversion_info ci;
  std::memset(&ci, 0, sizeof(ci));
  check(f(phi, phi0, size, max_iter, res, restart_freq, mv, extra, to_c(verb, cv), &ci), what);
  return to_info(ci);
}

}  // namespace

// ------------------------------------------------------------- solvers ----
inversion_info minv_vector_cg(double* phi, double* phi0, int size, int max_iter, double res, rmv mv, void* extra,
                              inversion_verbose_struct* verb) {
  return call(glx_minv_vector_cg_r, "minv_vector_cg", phi, phi0, size, max_iter, res, M(mv), extra, verb);
}
inversion_info minv_vector_cg(complex<double>* phi, complex<double>* phi0, int size, int max_iter, double res, cmv mv,
                              void* extra, inversion_verbose_struct* verb) {
  return call(glx_minv_vector_cg_c, "minv_vector_cg", D(phi), D(phi0), size, max_iter, res, M(mv), extra, verb);
}
inversion_info minv_vector_cg_restart(double* phi, double* phi0, int size, int max_iter, double res,
                                      int restart_freq, rmv mv, void* extra, inversion_verbose_struct* verb) {
  return call_r(glx_minv_vector_cg_restart_r, "minv_vector_cg_restart", phi, phi0, size, max_iter, res, restart_freq,
                M(mv), extra, verb);
}
inversion_info minv_vector_cg_restart(complex<double>* phi, complex<double>* phi0, int size, int max_iter,
                                      double res, int restart_freq, cmv mv, void* extra,
                                      inversion_verbose_struct* verb) {
  return call_r(glx_minv_vector_cg_restart_c, "minv_vector_cg_restart", D(phi), D(phi0), size, max_iter, res,
                restart_freq, M(mv), extra, verb);
}
inversion_info minv_vector_bicgstab(double* phi, double* phi0, int size, int max_iter, double res, rmv mv,
                                    void* extra, inversion_verbose_struct* verb) {
  return call(glx_minv_vector_bicgstab_r, "minv_vector_bicgstab", phi, phi0, size, max_iter, res, M(mv), extra, verb);
}
inversion_info minv_vector_bicgstab(complex<double>* phi, complex<double>* phi0, int size, int max_iter, double res,
                                    cmv mv, void* extra, inversion_verbose_struct* verb) {
  return call(glx_minv_vector_bicgstab_c, "minv_vector_bicgstab", D(phi), D(phi0), size, max_iter, res, M(mv), extra,
              verb);
}
inversion_info minv_vector_bicgstab_restart(double* phi, double* phi0, int size, int max_iter, double res,
                                            int restart_freq, rmv mv, void* extra, inversion_verbose_struct* verb) {
  return call_r(glx_minv_vector_bicgstab_restart_r, "minv_vector_bicgstab_restart", phi, phi0, size, max_iter, res,
                restart_freq, M(mv), extra, verb);
}
inversion_info minv_vector_bicgstab_restart(complex<double>* phi, complex<double>* phi0, int size, int max_iter,
                                            double res, int restart_freq, cmv mv, void* extra,
                                            inversion_verbose_struct* verb) {
  return call_r(glx_minv_vector_bicgstab_restart_c, "minv_vector_bicgstab_restart", D(phi), D(phi0), size, max_iter,
                res, restart_freq, M(mv), extra, verb);
}
inversion_info minv_vector_gcr(double* phi, double* phi0, int size, int max_iter, double res, rmv mv, void* extra,
                               inversion_verbose_struct* verb) {
  return call(glx_minv_vector_gcr_r, "minv_vector_gcr", phi, phi0, size, max_iter, res, M(mv), extra, verb);
}
inversion_info minv_vector_gcr(complex<double>* phi, complex<double>* phi0, int size, int max_iter, double res,
                               cmv mv, void* extra, inversion_verbose_struct* verb) {
  return call(glx_minv_vector_gcr_c, "minv_vector_gcr", D(phi), D(phi0), size, max_iter, res, M(mv), extra, verb);
}
inversion_info minv_vector_gcr_restart(double* phi, double* phi0, int size, int max_iter, double res,
                                       int restart_freq, rmv mv, void* extra, inversion_verbose_struct* verb) {
  return call_r(glx_minv_vector_gcr_restart_r, "minv_vector_gcr_restart", phi, phi0, size, max_iter, res,
                restart_freq, M(mv), extra, verb);
}
inversion_info minv_vector_gcr_restart(complex<double>* phi, complex<double>* phi0, int size, int max_iter,
                                       double res, int restart_freq, cmv mv, void* extra,
                                       inversion_verbose_struct* verb) {
  return call_r(glx_minv_vector_gcr_restart_c, "minv_vector_gcr_restart", D(phi), D(phi0), size, max_iter, res,
                restart_freq, M(mv), extra, verb);
}

inversion_info minv_vector_cg_m(double** phi, double* phi0, int n_shift, int size, int resid_freq_check,
                                int max_iter, double eps, double* shifts, rmv mv, void* extra, bool worst_first,
                                inversion_verbose_struct* verb) {
  inversion_info r(n_shift);
  glx_verbose cv;
  glx_inversion_info ci;
  std::memset(&ci, 0, sizeof(ci));
  ci.resSqmrhs = r.resSqmrhs;
  ci.n_rhs = n_shift;
  check(glx_minv_vector_cg_m_r(phi, phi0, n_shift, size, resid_freq_check, max_iter, eps, shifts, M(mv), extra,
                               worst_first ? 1 : 0, to_c(verb, cv), &ci),
        "minv_vector_cg_m");
  r.resSq = ci.resSq;
  r.iter = ci.iter;
  r.success = ci.success != 0;
  r.name = ci.name;
  r.ops_count = ci.ops_count;
  return r;
}
inversion_info minv_vector_cg_m(complex<double>** phi, complex<double>* phi0, int n_shift, int size,
                                int resid_freq_check, int max_iter, double eps, double* shifts, cmv mv, void* extra,
                                bool worst_first, inversion_verbose_struct* verb) {
  inversion_info r(n_shift);
  glx_verbose cv;
  glx_inversion_info ci;
  std::memset(&ci, 0, sizeof(ci));
  ci.resSqmrhs = r.resSqmrhs;
  ci.n_rhs = n_shift;
  check(glx_minv_vector_cg_m_c(reinterpret_cast<double**>(phi), D(phi0), n_shift, size, resid_freq_check, max_iter,
                               eps, shifts, M(mv), extra, worst_first ? 1 : 0, to_c(verb, cv),
                               &ci),
        "minv_vector_cg_m");
  r.resSq = ci.resSq;
  r.iter = ci.iter;
  r.success = ci.success != 0;
  r.name = ci.name;
  r.ops_count = ci.ops_count;
  return r;
}

// ----------------------------------------------------------- operators ----
void square_laplace(complex<double>* l, complex<double>* r, void* e) { glx_square_laplace(D(l), D(r), e); }
void square_laplace_u1(complex<double>* l, complex<double>* r, void* e) { glx_square_laplace_u1(D(l), D(r), e); }
int get_stencil_size(op_type opt) {
  switch (opt) {
    case STAGGERED:
    case LAPLACE:
    case LAPLACE_NC2:
    case G5_STAGGERED:
      return 1;
    case STAGGERED_INDEX:
    case STAGGERED_NORMAL:
      return 2;
  }
  return 0;
}
void square_staggered_u1(complex<double>* l, complex<double>* r, void* e) { glx_square_staggered_u1(D(l), D(r), e); }
void square_staggered(complex<double>* l, complex<double>* r, void* e) { glx_square_staggered(D(l), D(r), e); }
void square_staggered_gamma5(complex<double>* l, complex<double>* r, void* e) {
  glx_square_staggered_gamma5(D(l), D(r), e);
}
void square_staggered_2linklaplace_u1(complex<double>* l, complex<double>* r, void* e) {
  glx_square_staggered_2linklaplace_u1(D(l), D(r), e);
}
void staggered_symmshift_x(complex<double>* l, complex<double>* r, void* e) {
  glx_staggered_symmshift_x(D(l), D(r), e);
}
void staggered_symmshift_y(complex<double>* l, complex<double>* r, void* e) {
  glx_staggered_symmshift_y(D(l), D(r), e);
}
void staggered_index_operator(complex<double>* l, complex<double>* r, void* e) {
  glx_staggered_index_operator(D(l), D(r), e);
}
void gamma_5(complex<double>* l, complex<double>* r, void* e) { glx_gamma_5(D(l), D(r), e); }
void square_staggered_gamma5_u1(complex<double>* l, complex<double>* r, void* e) {
  glx_square_staggered_gamma5_u1(D(l), D(r), e);
}
void square_staggered_dagger_u1(complex<double>* l, complex<double>* r, void* e) {
  glx_square_staggered_dagger_u1(D(l), D(r), e);
}
void square_staggered_normal_u1(complex<double>* l, complex<double>* r, void* e) {
  glx_square_staggered_normal_u1(D(l), D(r), e);
}
void square_staggered_deo_u1(complex<double>* l, complex<double>* r, void* e) { glx_square_staggered_deo_u1(D(l), D(r), e); }
void square_staggered_doe_u1(complex<double>* l, complex<double>* r, void* e) { glx_square_staggered_doe_u1(D(l), D(r), e); }
void square_staggered_m2mdeodoe_u1(complex<double>* l, complex<double>* r, void* e) {
  glx_square_staggered_m2mdeodoe_u1(D(l), D(r), e);
}
void square_staggered_eoprec_prepare(complex<double>* rhs_e, complex<double>* rhs_orig, void* e) {
  glx_square_staggered_eoprec_prepare(D(rhs_e), D(rhs_orig), e);
}
void square_staggered_eoprec_reconstruct(complex<double>* full, complex<double>* lhs_e, complex<double>* rhs_o,
                                         void* e) {
  glx_square_staggered_eoprec_reconstruct(D(full), D(lhs_e), D(rhs_o), e);
}
void square_laplace(double* l, double* r, void* e) { glx_square_laplace_real(l, r, e); }

namespace {
// Register the C++ operator functions so that solvers recognise them.
struct RegisterCompat {
  RegisterCompat() {
    cmv c;
    rmv rl = square_laplace;
    glx_register_operator(M(rl), GLX_OP_LAPLACE_REAL, nullptr);
    c = square_laplace;
    glx_register_operator(M(c), GLX_OP_LAPLACE, nullptr);
    glx_register_operator(M(square_laplace_u1), GLX_OP_LAPLACE_U1, nullptr);
    glx_register_operator(M(square_staggered_u1), GLX_OP_STAGGERED, nullptr);
    glx_register_operator(M(gamma_5), GLX_OP_GAMMA5, nullptr);
    glx_register_operator(M(square_staggered_gamma5_u1), GLX_OP_STAGGERED_GAMMA5, nullptr);
    glx_register_operator(M(square_staggered_dagger_u1), GLX_OP_STAGGERED_DAGGER, nullptr);
    glx_register_operator(M(square_staggered_normal_u1), GLX_OP_STAGGERED_NORMAL, nullptr);
    glx_register_operator(M(square_staggered_deo_u1), GLX_OP_DEO, nullptr);
    glx_register_operator(M(square_staggered_doe_u1), GLX_OP_DOE, nullptr);
    glx_register_operator(M(square_staggered_m2mdeodoe_u1), GLX_OP_M2MDEODOE, nullptr);
    glx_register_operator(M(square_staggered), GLX_OP_STAGGERED_FREE, nullptr);
    glx_register_operator(M(square_staggered_gamma5), GLX_OP_STAGGERED_GAMMA5_FREE, nullptr);
    glx_register_operator(M(square_staggered_2linklaplace_u1), GLX_OP_STAGGERED_2LINKLAPLACE, nullptr);
    glx_register_operator(M(staggered_symmshift_x), GLX_OP_SYMMSHIFT_X, nullptr);
    glx_register_operator(M(staggered_symmshift_y), GLX_OP_SYMMSHIFT_Y, nullptr);
    glx_register_operator(M(staggered_index_operator), GLX_OP_STAGGERED_INDEX, nullptr);
  }
};
RegisterCompat g_register_compat;
}  // namespace

// ------------------------------------------------------------ dispatcher --
// generic_inverter.cpp:18-188: the solver named by `type`, restarted when
// params.restart (CG, GCR, BiCGStab); solvers outside this library throw.
namespace {
template <class V, class MV>
inversion_info dispatch(V* lhs, V* rhs, int size, minv_inverter type, minv_inverter_params& p, MV mv, void* extra,
                        inversion_verbose_struct* verb) {
  switch (type) {
    case MINV_CG:
      return p.restart ? minv_vector_cg_restart(lhs, rhs, size, p.max_iters, p.tol, p.restart_freq, mv, extra, verb)
                       : minv_vector_cg(lhs, rhs, size, p.max_iters, p.tol, mv, extra, verb);
    case MINV_GCR:
      return p.restart ? minv_vector_gcr_restart(lhs, rhs, size, p.max_iters, p.tol, p.restart_freq, mv, extra, verb)
                       : minv_vector_gcr(lhs, rhs, size, p.max_iters, p.tol, mv, extra, verb);
    case MINV_BICGSTAB:
      return p.restart
                 ? minv_vector_bicgstab_restart(lhs, rhs, size, p.max_iters, p.tol, p.restart_freq, mv, extra, verb)
                 : minv_vector_bicgstab(lhs, rhs, size, p.max_iters, p.tol, mv, extra, verb);
    case MINV_INVALID:
      return inversion_info();
    default:
      throw std::runtime_error("minv_unpreconditioned: solver " + std::to_string(int(type)) +
                               " is not part of this library (CG, GCR and BiCGStab are)");
  }
}
}  // namespace

inversion_info minv_unpreconditioned(double* lhs, double* rhs, int size, minv_inverter type,
                                     minv_inverter_params& params, rmv mv, void* extra,
                                     inversion_verbose_struct* verb) {
  return dispatch(lhs, rhs, size, type, params, mv, extra, verb);
}
inversion_info minv_unpreconditioned(complex<double>* lhs, complex<double>* rhs, int size, minv_inverter type,
                                     minv_inverter_params& params, cmv mv, void* extra,
                                     inversion_verbose_struct* verb) {
  return dispatch(lhs, rhs, size, type, params, mv, extra, verb);
}

// minv_preconditioned (generic_inverters_precond.h:31-63): the interface is
// kept so that drivers including it compile; the preconditioned solvers are
// not part of this library.
namespace {
template <class T, class MV, class PMV>
inversion_info precond_dispatch(minv_inverter_precond type, MV, PMV) {
  if (type == MINV_PRE_INVALID) return inversion_info();
  throw std::runtime_error("minv_preconditioned: preconditioned solver " + std::to_string(int(type)) +
                           " is not part of this library (minv_unpreconditioned carries CG, GCR and BiCGStab)");
}
}  // namespace

inversion_info minv_preconditioned(double*, double*, int, minv_inverter_precond type, minv_inverter_precond_params&,
                                   void (*mv)(double*, double*, void*), void*,
                                   void (*pmv)(double*, double*, int, void*, inversion_verbose_struct*), void*,
                                   inversion_verbose_struct*) {
  return precond_dispatch<double>(type, mv, pmv);
}
inversion_info minv_preconditioned(complex<double>*, complex<double>*, int, minv_inverter_precond type,
                                   minv_inverter_precond_params&,
                                   void (*mv)(complex<double>*, complex<double>*, void*), void*,
                                   void (*pmv)(complex<double>*, complex<double>*, int, void*,
                                               inversion_verbose_struct*),
                                   void*, inversion_verbose_struct*) {
  return precond_dispatch<complex<double>>(type, mv, pmv);
}

// ---------------------------------------------------- U(1) gauge utilities --
// u1_utils.cpp:17-207: the same generators (std::mt19937 with the standard
// library's distributions) and the same draw order, so seeded fields match.
namespace {
constexpr double kPi = 3.141592653589793238462643383279502884;
inline int link_of(int x, int y, int mu, int x_len) { return 2 * (y * x_len + x) + mu; }
}  // namespace

void read_gauge_u1(complex<double>* u, int x_len, int y_len, string input_file) {
  std::ifstream in(input_file);
  if (!in) throw std::runtime_error("read_gauge_u1: cannot open " + input_file);
  for (int x = 0; x < x_len; x++)
    for (int y = 0; y < y_len; y++)
      for (int mu = 0; mu < 2; mu++) {
        double phase;
        in >> phase;
        u[link_of(x, y, mu, x_len)] = std::polar(1.0, phase);
      }
}

void write_gauge_u1(complex<double>* u, int x_len, int y_len, string output_file) {
  std::ofstream out(output_file, std::ios::out | std::ios::trunc);
  if (!out) throw std::runtime_error("write_gauge_u1: cannot open " + output_file);
  out.setf(std::ios_base::fixed, std::ios_base::floatfield);
  out.precision(20);
  for (int x = 0; x < x_len; x++)
    for (int y = 0; y < y_len; y++)
      for (int mu = 0; mu < 2; mu++) out << std::arg(u[link_of(x, y, mu, x_len)]) << "\n";
}

void unit_gauge_u1(complex<double>* u, int x_len, int y_len) {
  for (int i = 0; i < 2 * x_len * y_len; i++) u[i] = 1.0;
}

void rand_gauge_u1(complex<double>* u, int x_len, int y_len, std::mt19937& gen) {
  std::uniform_real_distribution<> phase(-kPi, kPi);
  for (int i = 0; i < 2 * x_len * y_len; i++) u[i] = std::polar(1.0, phase(gen));
}

void gauss_gauge_u1(complex<double>* u, int x_len, int y_len, std::mt19937& gen, double beta) {
  if (beta < 0) beta = -beta;
  if (beta == 0.0) rand_gauge_u1(u, x_len, y_len, gen);  // then overwritten, as in the reference
  std::normal_distribution<> phase(0.0, 1.0 / std::sqrt(beta));
  for (int i = 0; i < 2 * x_len * y_len; i++) u[i] = std::polar(1.0, phase(gen));
}

void rand_trans_u1(complex<double>* g, int x_len, int y_len, std::mt19937& gen) {
  std::uniform_real_distribution<> phase(-kPi, kPi);
  for (int i = 0; i < x_len * y_len; i++) g[i] = std::polar(1.0, phase(gen));
}

void apply_gauge_trans_u1(complex<double>* u, complex<double>* g, int x_len, int y_len) {
  for (int y = 0; y < y_len; y++)
    for (int x = 0; x < x_len; x++) {
      const int s = y * x_len + x;
      const int sx = y * x_len + (x + 1) % x_len, sy = ((y + 1) % y_len) * x_len + x;
      u[2 * s] = g[s] * u[2 * s] * std::conj(g[sx]);
      u[2 * s + 1] = g[s] * u[2 * s + 1] * std::conj(g[sy]);
    }
}

namespace {
// U_x(x,y) U_y(x+1,y) U_x^*(x,y+1) U_y^*(x,y), multiplied in this order
inline complex<double> plaquette_at(const complex<double>* u, int x, int y, int x_len, int y_len) {
  return u[link_of(x, y, 0, x_len)] * u[link_of((x + 1) % x_len, y, 1, x_len)] *
         std::conj(u[link_of(x, (y + 1) % y_len, 0, x_len)]) * std::conj(u[link_of(x, y, 1, x_len)]);
}
}  // namespace

complex<double> get_plaquette_u1(complex<double>* u, int x_len, int y_len) {
  complex<double> sum = 0.0;
  for (int y = 0; y < y_len; y++)
    for (int x = 0; x < x_len; x++) sum += plaquette_at(u, x, y, x_len, y_len);
  return sum / double(x_len * y_len);
}

double get_topo_u1(complex<double>* u, int x_len, int y_len) {
  double angle = 0.0;
  for (int y = 0; y < y_len; y++)
    for (int x = 0; x < x_len; x++) angle += std::arg(plaquette_at(u, x, y, x_len, y_len));
  return 0.5 * angle / kPi;
}

// Each link gains alpha times its two staples; the sum is projected to a
// phase.  Staple products keep the left-to-right order (alpha * a) * b * c of
// u1_utils.cpp:162-172 so results agree to the last bit.
void apply_ape_smear_u1(complex<double>* out, complex<double>* u, int x_len, int y_len, double alpha,
                        int n_iter) {
  const size_t nl = size_t(2) * x_len * y_len;
  std::vector<complex<double>> cur(u, u + nl);
  auto at = [&](int x, int y, int mu) -> const complex<double>& {
    return cur[link_of((x + x_len) % x_len, (y + y_len) % y_len, mu, x_len)];
  };
  for (int it = 0; it < n_iter; it++) {
    for (int y = 0; y < y_len; y++)
      for (int x = 0; x < x_len; x++) {
        complex<double> sx = at(x, y, 0);
        sx += alpha * at(x, y, 1) * at(x, y + 1, 0) * std::conj(at(x + 1, y, 1));
        sx += alpha * std::conj(at(x, y - 1, 1)) * at(x, y - 1, 0) * at(x + 1, y - 1, 1);
        complex<double> sy = at(x, y, 1);
        sy += alpha * std::conj(at(x - 1, y, 0)) * at(x - 1, y, 1) * at(x - 1, y + 1, 0);
        sy += alpha * at(x, y, 0) * at(x + 1, y, 1) * std::conj(at(x, y + 1, 0));
        out[link_of(x, y, 0, x_len)] = sx;
        out[link_of(x, y, 1, x_len)] = sy;
      }
    for (size_t i = 0; i < nl; i++) cur[i] = std::polar(1.0, std::arg(out[i]));
  }
  std::copy(cur.begin(), cur.end(), out);
}

// ------------------------------------------------------------- stencil_2d --
// coarse_stencil.cpp:12-1627 and operators_stencil.cpp:14-236 through the
// general-stencil C ABI: the compat stencil_2d is described to the library
// (pointers into its own arrays), and the compat apply functions are
// registered so that a solve given one of them runs on the GPU.
namespace {
int describe(void* extra, glx_stencil_2d_nc* out) {
  stencil_2d* s = static_cast<stencil_2d*>(extra);
  if (!s || !s->lat || s->lat->get_nd() != 2) return -1;
  out->lx = s->lat->get_lattice_dimension(0);
  out->ly = s->lat->get_lattice_dimension(1);
  out->nc = s->lat->get_nc();
  out->has_two = s->has_two ? 1 : 0;
  out->sdir = static_cast<int>(s->sdir);
  out->clover = D(s->clover);
  out->hopping = D(s->hopping);
  out->two_link = s->has_two ? D(s->two_link) : nullptr;
  out->shift[0] = s->shift.real();
  out->shift[1] = s->shift.imag();
  out->eo_shift[0] = s->eo_shift.real();
  out->eo_shift[1] = s->eo_shift.imag();
  out->dof_shift[0] = s->dof_shift.real();
  out->dof_shift[1] = s->dof_shift.imag();
  return 0;
}
glx_stencil_2d_nc described(void* extra) {
  glx_stencil_2d_nc d{};
  if (describe(extra, &d) != 0) throw std::runtime_error("stencil_2d: needs a two-dimensional Lattice");
  return d;
}

struct StencilRegistration {
  StencilRegistration() {
    glx::register_cst_callback(reinterpret_cast<void*>(M(&apply_stencil_2d)), GLX_OP_CSTENCIL, &describe);
    glx::register_cst_callback(reinterpret_cast<void*>(M(&apply_stencil_2d_eo)), GLX_OP_CSTENCIL_EO, &describe);
    glx::register_cst_callback(reinterpret_cast<void*>(M(&apply_stencil_2d_oe)), GLX_OP_CSTENCIL_OE, &describe);
    glx::register_cst_callback(reinterpret_cast<void*>(M(&apply_stencil_2d_tb)), GLX_OP_CSTENCIL_TB, &describe);
    glx::register_cst_callback(reinterpret_cast<void*>(M(&apply_stencil_2d_bt)), GLX_OP_CSTENCIL_BT, &describe);
    glx::register_cst_callback(reinterpret_cast<void*>(M(&apply_square_staggered_m2mdeodoe_stencil)),
                               GLX_OP_CSTENCIL_M2MDEODOE, &describe);
  }
} g_stencil_registration;

void staggered_builder(stencil_2d* st, staggered_u1_op* op,
                       int (*build)(glx_stencil_2d*, const glx_staggered_u1_op*)) {
  if (st->generated || st->lat->get_nc() != 1) return;  // operators_stencil.cpp:16-24
  glx_stencil_2d c{};
  c.lx = st->lat->get_lattice_dimension(0);
  c.ly = st->lat->get_lattice_dimension(1);
  c.clover = D(st->clover);
  c.hopping = D(st->hopping);
  glx_staggered_u1_op p{D(op->lattice), op->mass, op->x_fine, op->y_fine, op->Nc, op->wilson_coeff};
  if (build(&c, &p) != GLX_OK) throw std::runtime_error(std::string("stencil builder: ") + glx_last_error());
  st->shift = complex<double>(c.shift[0], c.shift[1]);
  st->eo_shift = complex<double>(c.eo_shift[0], c.eo_shift[1]);
  st->dof_shift = complex<double>(c.dof_shift[0], c.dof_shift[1]);
  st->generated = true;
}
}  // namespace

void apply_stencil_2d(complex<double>* lhs, complex<double>* rhs, void* extra) {
  glx_stencil_2d_nc d = described(extra);
  glx_apply_stencil_2d_nc(D(lhs), D(rhs), &d);
}
void apply_stencil_2d_eo(complex<double>* lhs, complex<double>* rhs, void* extra) {
  glx_stencil_2d_nc d = described(extra);
  glx_apply_stencil_2d_nc_eo(D(lhs), D(rhs), &d);
}
void apply_stencil_2d_oe(complex<double>* lhs, complex<double>* rhs, void* extra) {
  glx_stencil_2d_nc d = described(extra);
  glx_apply_stencil_2d_nc_oe(D(lhs), D(rhs), &d);
}
void apply_stencil_2d_tb(complex<double>* lhs, complex<double>* rhs, void* extra) {
  glx_stencil_2d_nc d = described(extra);
  glx_apply_stencil_2d_nc_tb(D(lhs), D(rhs), &d);
}
void apply_stencil_2d_bt(complex<double>* lhs, complex<double>* rhs, void* extra) {
  glx_stencil_2d_nc d = described(extra);
  glx_apply_stencil_2d_nc_bt(D(lhs), D(rhs), &d);
}
void generate_stencil_2d(stencil_2d* st, void (*mv)(complex<double>*, complex<double>*, void*), void* extra) {
  if (st->generated || st->stencil_size > 2) return;
  glx_stencil_2d_nc d = described(st);
  if (glx_generate_stencil_2d_nc(&d, M(mv), extra) != GLX_OK)
    throw std::runtime_error(std::string("generate_stencil_2d: ") + glx_last_error());
  st->generated = true;
}

void get_square_staggered_u1_stencil(stencil_2d* st, staggered_u1_op* op) {
  staggered_builder(st, op, &glx_get_square_staggered_u1_stencil);
}
void get_square_staggered_gamma5_u1_stencil(stencil_2d* st, staggered_u1_op* op) {
  staggered_builder(st, op, &glx_get_square_staggered_gamma5_u1_stencil);
}
void get_square_staggered_dagger_u1_stencil(stencil_2d* st, staggered_u1_op* op) {
  staggered_builder(st, op, &glx_get_square_staggered_dagger_u1_stencil);
}
void apply_square_staggered_eoprec_prepare_stencil(complex<double>* rhs_e, complex<double>* rhs_orig,
                                                   stencil_2d* st) {
  glx_stencil_2d_nc d = described(st);
  glx_stencil_2d_nc_eoprec_prepare(D(rhs_e), D(rhs_orig), &d);
}
void apply_square_staggered_m2mdeodoe_stencil(complex<double>* lhs, complex<double>* rhs, void* extra) {
  glx_stencil_2d_nc d = described(extra);
  glx_apply_stencil_2d_nc_m2mdeodoe(D(lhs), D(rhs), &d);
}
void apply_square_staggered_eoprec_reconstruct_stencil(complex<double>* lhs_full, complex<double>* lhs_e,
                                                       complex<double>* rhs_o, stencil_2d* st) {
  glx_stencil_2d_nc d = described(st);
  glx_stencil_2d_nc_eoprec_reconstruct(D(lhs_full), D(lhs_e), D(rhs_o), &d);
}

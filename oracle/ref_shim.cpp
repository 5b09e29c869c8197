// ref_shim.cpp -- extern "C" entry points over the reference's own solver,
// operator and U(1) routines, so that tests can drive the reference through
// ctypes.  Compiled together with the reference sources by oracle/Makefile
// into oracle/_ref/libref.so.  TEST INFRASTRUCTURE ONLY: this library is the
// checker of the oracle restatement and never part of the product.
//
// This file contains only glue (argument marshalling and output capture); all
// arithmetic runs in the reference's own translation units.
#include <complex>
#include <cstring>
#include <iomanip>
#include <iostream>
#include <random>
#include <sstream>
#include <string>

#include "generic_bicgstab.h"
#include "generic_cg.h"
#include "generic_cg_m.h"
#include "generic_gcr.h"
#include "generic_vector.h"
#include "inverter_struct.h"
#include "verbosity.h"
#include "lattice.h"
#include "coarse_stencil.h"
#include "operators.h"
#include "operators_stencil.h"
#include "u1_utils.h"

typedef std::complex<double> cd;

// The runtime-L real Laplacian lives in the multishift test driver
// (tests/multishift/multishift.cpp:634-672), compiled with main renamed.
void square_laplace(double* lhs, double* rhs, void* extra_data);

namespace {

enum RefOp {
  OP_LAPLACE_REAL = 0,
  OP_LAPLACE = 1,
  OP_LAPLACE_U1 = 2,
  OP_STAGGERED = 3,
  OP_STAGGERED_DAGGER = 4,
  OP_STAGGERED_NORMAL = 5,
  OP_STAGGERED_GAMMA5 = 6,
  OP_GAMMA5 = 7,
  OP_DEO = 8,
  OP_DOE = 9,
  OP_M2MDEODOE = 10,
  OP_STENCIL = 11,
  OP_STENCIL_DAGGER = 12,
  OP_STENCIL_GAMMA5 = 13,
  OP_STENCIL_EO = 14,
  OP_STENCIL_OE = 15,
  OP_STENCIL_M2MDEODOE = 16,
  // operators.h functions off the solver path
  OP_STAG_FREE = 30,
  OP_G5_FREE = 31,
  OP_2LINK = 32,
  OP_SYMM_X = 33,
  OP_SYMM_Y = 34,
  OP_INDEX = 35,
};

// Operator context: the staggered_u1_op the function-form operators read,
// and (for the stencil ops) a Lattice + stencil_2d built from the same links.
struct Ctx {
  staggered_u1_op op;
  Lattice* lat = nullptr;
  stencil_2d* st = nullptr;
  Ctx(double* links, double mass, int lx, int ly, int nc, int opid) {
    op.lattice = reinterpret_cast<cd*>(links);
    op.mass = mass;
    op.x_fine = lx;
    op.y_fine = ly;
    op.Nc = nc;
    op.wilson_coeff = 0.0;
    if (opid >= OP_STENCIL && opid <= OP_STENCIL_M2MDEODOE) {
      int dims[2] = {lx, ly};
      lat = new Lattice(2, dims, 1);
      st = new stencil_2d(lat, 1);
      if (opid == OP_STENCIL_DAGGER)
        get_square_staggered_dagger_u1_stencil(st, &op);
      else if (opid == OP_STENCIL_GAMMA5)
        get_square_staggered_gamma5_u1_stencil(st, &op);
      else
        get_square_staggered_u1_stencil(st, &op);
    }
  }
  ~Ctx() {
    delete st;
    delete lat;
  }
  void* extra(int opid) {
    return (opid >= OP_STENCIL && opid <= OP_STENCIL_M2MDEODOE) ? static_cast<void*>(st) : static_cast<void*>(&op);
  }
};

typedef void (*cmv)(cd*, cd*, void*);
cmv complex_op(int opid) {
  switch (opid) {
    case OP_LAPLACE: return square_laplace;
    case OP_LAPLACE_U1: return square_laplace_u1;
    case OP_STAGGERED: return square_staggered_u1;
    case OP_STAGGERED_DAGGER: return square_staggered_dagger_u1;
    case OP_STAGGERED_NORMAL: return square_staggered_normal_u1;
    case OP_STAGGERED_GAMMA5: return square_staggered_gamma5_u1;
    case OP_GAMMA5: return gamma_5;
    case OP_DEO: return square_staggered_deo_u1;
    case OP_DOE: return square_staggered_doe_u1;
    case OP_M2MDEODOE: return square_staggered_m2mdeodoe_u1;
    case OP_STENCIL:
    case OP_STENCIL_DAGGER:
    case OP_STENCIL_GAMMA5: return apply_stencil_2d;
    case OP_STENCIL_EO: return apply_stencil_2d_eo;
    case OP_STENCIL_OE: return apply_stencil_2d_oe;
    case OP_STENCIL_M2MDEODOE: return apply_square_staggered_m2mdeodoe_stencil;
    case OP_STAG_FREE: return square_staggered;
    case OP_G5_FREE: return square_staggered_gamma5;
    case OP_2LINK: return square_staggered_2linklaplace_u1;
    case OP_SYMM_X: return staggered_symmshift_x;
    case OP_SYMM_Y: return staggered_symmshift_y;
    case OP_INDEX: return staggered_index_operator;
  }
  return nullptr;
}

// Redirects std::cout into a string for the duration of a call.
// Redirects std::cout into `out` for the duration of a call.  Inactive when
// the solve prints nothing (verbosity NONE): then the call touches no global
// state and calls from several threads may run at once.
struct Capture {
  std::ostringstream ss;
  std::streambuf* old = nullptr;
  std::ios_base::fmtflags flags;
  std::streamsize precision;
  char* out;
  int cap;
  bool active;
  Capture(char* o, int c, int sci, int prec, bool on = true) : out(o), cap(c), active(on) {
    if (!active) {
      if (out && cap > 0) out[0] = 0;
      return;
    }
    old = std::cout.rdbuf(ss.rdbuf());
    flags = std::cout.flags();
    precision = std::cout.precision();
    if (sci) std::cout << std::scientific;
    if (prec > 0) std::cout << std::setprecision(prec);
  }
  ~Capture() {
    if (!active) return;
    std::cout.rdbuf(old);
    std::cout.flags(flags);
    std::cout.precision(precision);
    if (out && cap > 0) {
      std::string s = ss.str();
      size_t n = s.size() < size_t(cap - 1) ? s.size() : size_t(cap - 1);
      std::memcpy(out, s.data(), n);
      out[n] = 0;
    }
  }
};

void fill_info(const inversion_info& inf, double* info) {
  info[0] = inf.resSq;
  info[1] = inf.iter;
  info[2] = inf.success ? 1.0 : 0.0;
  info[3] = inf.ops_count;
}

}  // namespace

extern "C" {

// One operator application lhs = A rhs.
int ref_apply(int opid, double* lhs, double* rhs, double* links, double mass, int lx, int ly, int nc) {
  Ctx c(links, mass, lx, ly, nc, opid);
  if (opid == OP_LAPLACE_REAL) {
    square_laplace(lhs, rhs, &c.op);
    return 0;
  }
  cmv f = complex_op(opid);
  if (!f) return -1;
  f(reinterpret_cast<cd*>(lhs), reinterpret_cast<cd*>(rhs), c.extra(opid));
  return 0;
}

// A function-form operator with a Wilson coefficient (the two-link Laplacian term).
int ref_apply_wc(int opid, double* lhs, double* rhs, double* links, double mass, double wc, int lx, int ly) {
  Ctx c(links, mass, lx, ly, 1, opid);
  c.op.wilson_coeff = wc;
  cmv f = complex_op(opid);
  if (!f) return -1;
  f(reinterpret_cast<cd*>(lhs), reinterpret_cast<cd*>(rhs), c.extra(opid));
  return 0;
}

// Even/odd preparation and reconstruction, function form (stencil = 0) or
// stencil form (stencil = 1).
int ref_eoprec_prepare(int stencil, double* rhs_e, double* rhs, double* links, double mass, int lx, int ly) {
  Ctx c(links, mass, lx, ly, 1, stencil ? OP_STENCIL : OP_STAGGERED);
  if (stencil)
    apply_square_staggered_eoprec_prepare_stencil(reinterpret_cast<cd*>(rhs_e), reinterpret_cast<cd*>(rhs), c.st);
  else
    square_staggered_eoprec_prepare(reinterpret_cast<cd*>(rhs_e), reinterpret_cast<cd*>(rhs), &c.op);
  return 0;
}
int ref_eoprec_reconstruct(int stencil, double* full, double* lhs_e, double* rhs_o, double* links, double mass, int lx,
                           int ly) {
  Ctx c(links, mass, lx, ly, 1, stencil ? OP_STENCIL : OP_STAGGERED);
  if (stencil)
    apply_square_staggered_eoprec_reconstruct_stencil(reinterpret_cast<cd*>(full), reinterpret_cast<cd*>(lhs_e),
                                                      reinterpret_cast<cd*>(rhs_o), c.st);
  else
    square_staggered_eoprec_reconstruct(reinterpret_cast<cd*>(full), reinterpret_cast<cd*>(lhs_e),
                                        reinterpret_cast<cd*>(rhs_o), &c.op);
  return 0;
}

// Stencil coefficient blocks (clover V, hopping 4V) and shifts.
int ref_stencil_coeffs(int opid, double* clover, double* hopping, double* shifts6, double* links, double mass, int lx,
                       int ly) {
  Ctx c(links, mass, lx, ly, 1, opid);
  if (!c.st) return -1;
  const int V = lx * ly;
  std::memcpy(clover, c.st->clover, sizeof(cd) * V);
  std::memcpy(hopping, c.st->hopping, sizeof(cd) * 4 * V);
  shifts6[0] = c.st->shift.real(); shifts6[1] = c.st->shift.imag();
  shifts6[2] = c.st->eo_shift.real(); shifts6[3] = c.st->eo_shift.imag();
  shifts6[4] = c.st->dof_shift.real(); shifts6[5] = c.st->dof_shift.imag();
  return 0;
}

// General stencil_2d (any Nc, two-link, sdir): a Lattice and stencil_2d built
// from caller coefficient arrays in the reference layout.
struct CStencil {
  Lattice* lat;
  stencil_2d* st;
  CStencil(int lx, int ly, int nc, int has_two, int sdir, const double* clover, const double* hopping,
           const double* two, const double* shifts6) {
    int dims[2] = {lx, ly};
    lat = new Lattice(2, dims, nc);
    st = new stencil_2d(lat, has_two ? 2 : 1, cd(shifts6[0], shifts6[1]), cd(shifts6[2], shifts6[3]),
                        cd(shifts6[4], shifts6[5]));
    const size_t blk = size_t(lx) * ly * nc * nc;
    if (clover) std::memcpy(st->clover, clover, sizeof(cd) * blk);
    if (hopping) std::memcpy(st->hopping, hopping, sizeof(cd) * 4 * blk);
    if (has_two && two) std::memcpy(st->two_link, two, sizeof(cd) * 8 * blk);
    st->sdir = static_cast<stencil_dir>(sdir);
  }
  ~CStencil() {
    delete st;
    delete lat;
  }
};

// piece: 0 all, 1 eo, 2 oe, 3 tb, 4 bt, 5 m2mdeodoe, 6 eoprec prepare,
// 7 eoprec reconstruct (aux = rhs_o, rhs = lhs_e)
int ref_cstencil_apply(int piece, int lx, int ly, int nc, int has_two, int sdir, const double* clover,
                       const double* hopping, const double* two, const double* shifts6, double* lhs, double* rhs,
                       double* aux) {
  CStencil c(lx, ly, nc, has_two, sdir, clover, hopping, two, shifts6);
  cd* L = reinterpret_cast<cd*>(lhs);
  cd* R = reinterpret_cast<cd*>(rhs);
  switch (piece) {
    case 0: apply_stencil_2d(L, R, c.st); break;
    case 1: apply_stencil_2d_eo(L, R, c.st); break;
    case 2: apply_stencil_2d_oe(L, R, c.st); break;
    case 3: apply_stencil_2d_tb(L, R, c.st); break;
    case 4: apply_stencil_2d_bt(L, R, c.st); break;
    case 5: apply_square_staggered_m2mdeodoe_stencil(L, R, c.st); break;
    case 6: apply_square_staggered_eoprec_prepare_stencil(L, R, c.st); break;
    case 7: apply_square_staggered_eoprec_reconstruct_stencil(L, R, reinterpret_cast<cd*>(aux), c.st); break;
    default: return -1;
  }
  return 0;
}

// generate_stencil_2d from one of the function-form operators (nc = 1 links)
// or from another general stencil (src_* arrays, same lattice and nc).
int ref_cstencil_generate(int opid, double* links, double mass, int lx, int ly, int nc, int has_two,
                          const double* src_clover, const double* src_hopping, const double* src_two,
                          const double* src_shifts6, int src_has_two, double* clover, double* hopping, double* two) {
  const double zero6[6] = {0, 0, 0, 0, 0, 0};
  CStencil out(lx, ly, nc, has_two, 0, nullptr, nullptr, nullptr, zero6);
  if (opid >= 0) {
    Ctx c(links, mass, lx, ly, nc, opid);
    generate_stencil_2d(out.st, complex_op(opid), c.extra(opid));
  } else {
    CStencil src(lx, ly, nc, src_has_two, 0, src_clover, src_hopping, src_two, src_shifts6);
    generate_stencil_2d(out.st, apply_stencil_2d, src.st);
  }
  const size_t blk = size_t(lx) * ly * nc * nc;
  std::memcpy(clover, out.st->clover, sizeof(cd) * blk);
  std::memcpy(hopping, out.st->hopping, sizeof(cd) * 4 * blk);
  if (has_two) std::memcpy(two, out.st->two_link, sizeof(cd) * 8 * blk);
  return 0;
}

// solver: 0 CG, 1 BiCGStab, 2 GCR.  restart_freq > 0 selects the _restart form.
// info = {resSq, iter, success, ops_count}.  Verbosity lines go to `out`.
int ref_solve(int solver, int opid, double* phi, double* phi0, int size, int max_iter, double eps, int restart_freq,
              double* links, double mass, int lx, int ly, int nc, int verb_level, const char* prefix, int sci,
              int prec, char* out, int outcap, double* info) {
  Ctx c(links, mass, lx, ly, nc, opid);
  inversion_verbose_struct verb;
  verb.verbosity = static_cast<inversion_verbose_level>(verb_level);
  verb.verb_prefix = prefix ? prefix : "";
  verb.precond_verbosity = VERB_NONE;
  verb.precond_verb_prefix = "";
  inversion_info inf;
  {
    Capture cap(out, outcap, sci, prec, verb_level != 0);
    if (opid == OP_LAPLACE_REAL) {
      void (*f)(double*, double*, void*) = square_laplace;
      if (solver == 0)
        inf = restart_freq > 0 ? minv_vector_cg_restart(phi, phi0, size, max_iter, eps, restart_freq, f, &c.op, &verb)
                               : minv_vector_cg(phi, phi0, size, max_iter, eps, f, &c.op, &verb);
      else if (solver == 1)
        inf = restart_freq > 0
                  ? minv_vector_bicgstab_restart(phi, phi0, size, max_iter, eps, restart_freq, f, &c.op, &verb)
                  : minv_vector_bicgstab(phi, phi0, size, max_iter, eps, f, &c.op, &verb);
      else
        inf = restart_freq > 0 ? minv_vector_gcr_restart(phi, phi0, size, max_iter, eps, restart_freq, f, &c.op, &verb)
                               : minv_vector_gcr(phi, phi0, size, max_iter, eps, f, &c.op, &verb);
    } else {
      cmv f = complex_op(opid);
      if (!f) return -1;
      cd* x = reinterpret_cast<cd*>(phi);
      cd* b = reinterpret_cast<cd*>(phi0);
      void* e = c.extra(opid);
      if (solver == 0)
        inf = restart_freq > 0 ? minv_vector_cg_restart(x, b, size, max_iter, eps, restart_freq, f, e, &verb)
                               : minv_vector_cg(x, b, size, max_iter, eps, f, e, &verb);
      else if (solver == 1)
        inf = restart_freq > 0 ? minv_vector_bicgstab_restart(x, b, size, max_iter, eps, restart_freq, f, e, &verb)
                               : minv_vector_bicgstab(x, b, size, max_iter, eps, f, e, &verb);
      else
        inf = restart_freq > 0 ? minv_vector_gcr_restart(x, b, size, max_iter, eps, restart_freq, f, e, &verb)
                               : minv_vector_gcr(x, b, size, max_iter, eps, f, e, &verb);
    }
  }
  fill_info(inf, info);
  return 0;
}

// Multishift CG.  phi: n_shift*size(*2) contiguous output; shifts in/out.
int ref_cg_m(int opid, double* phi, double* phi0, int n_shift, int size, int resid_freq_check, int max_iter, double eps,
             double* shifts, int worst_first, double* links, double mass, int lx, int ly, int nc, int verb_level,
             const char* prefix, int sci, int prec, char* out, int outcap, double* info, double* resSqmrhs) {
  Ctx c(links, mass, lx, ly, nc, opid);
  inversion_verbose_struct verb;
  verb.verbosity = static_cast<inversion_verbose_level>(verb_level);
  verb.verb_prefix = prefix ? prefix : "";
  verb.precond_verbosity = VERB_NONE;
  verb.precond_verb_prefix = "";
  inversion_info inf;
  {
    Capture cap(out, outcap, sci, prec, verb_level != 0);
    if (opid == OP_LAPLACE_REAL) {
      double** ph = new double*[n_shift];
      for (int n = 0; n < n_shift; n++) ph[n] = phi + (size_t)n * size;
      inf = minv_vector_cg_m(ph, phi0, n_shift, size, resid_freq_check, max_iter, eps, shifts, square_laplace, &c.op,
                             worst_first != 0, &verb);
      delete[] ph;
    } else {
      cmv f = complex_op(opid);
      if (!f) return -1;
      cd** ph = new cd*[n_shift];
      for (int n = 0; n < n_shift; n++) ph[n] = reinterpret_cast<cd*>(phi) + (size_t)n * size;
      inf = minv_vector_cg_m(ph, reinterpret_cast<cd*>(phi0), n_shift, size, resid_freq_check, max_iter, eps, shifts,
                             f, c.extra(opid), worst_first != 0, &verb);
      delete[] ph;
    }
  }
  fill_info(inf, info);
  for (int n = 0; n < n_shift; n++) resSqmrhs[n] = inf.resSqmrhs[n];
  return 0;
}

// BLAS-1 templates of generic_vector.h.  out = {re, im}.
void ref_blas(int which, int is_complex, double* a, double* b, int n, double* out) {
  out[0] = out[1] = 0.0;
  if (is_complex) {
    cd* x = reinterpret_cast<cd*>(a);
    cd* y = reinterpret_cast<cd*>(b);
    if (which == 0) { cd d = dot<double>(x, y, n); out[0] = d.real(); out[1] = d.imag(); }
    if (which == 1) out[0] = norm2sq<double>(x, n);
    if (which == 2) out[0] = diffnorm2sq<double>(x, y, n);
  } else {
    if (which == 0) out[0] = dot<double>(a, b, n);
    if (which == 1) out[0] = norm2sq<double>(a, n);
    if (which == 2) out[0] = diffnorm2sq<double>(a, b, n);
  }
}

// Random inputs from one seeded std::mt19937 stream.
void* ref_rng_new(unsigned seed) { return new std::mt19937(seed); }
void ref_rng_free(void* g) { delete static_cast<std::mt19937*>(g); }
void ref_gaussian_r(void* g, double* v, int n) { gaussian<double>(v, n, *static_cast<std::mt19937*>(g)); }
void ref_gaussian_c(void* g, double* v, int n) {
  gaussian<double>(reinterpret_cast<cd*>(v), n, *static_cast<std::mt19937*>(g));
}
void ref_gauss_gauge_u1(void* g, double* u, int lx, int ly, double beta) {
  gauss_gauge_u1(reinterpret_cast<cd*>(u), lx, ly, *static_cast<std::mt19937*>(g), beta);
}
void ref_rand_gauge_u1(void* g, double* u, int lx, int ly) {
  rand_gauge_u1(reinterpret_cast<cd*>(u), lx, ly, *static_cast<std::mt19937*>(g));
}
void ref_rand_trans_u1(void* g, double* t, int lx, int ly) {
  rand_trans_u1(reinterpret_cast<cd*>(t), lx, ly, *static_cast<std::mt19937*>(g));
}
void ref_unit_gauge_u1(double* u, int lx, int ly) { unit_gauge_u1(reinterpret_cast<cd*>(u), lx, ly); }
void ref_apply_gauge_trans_u1(double* u, double* t, int lx, int ly) {
  apply_gauge_trans_u1(reinterpret_cast<cd*>(u), reinterpret_cast<cd*>(t), lx, ly);
}
void ref_plaquette_u1(double* u, int lx, int ly, double* out) {
  cd p = get_plaquette_u1(reinterpret_cast<cd*>(u), lx, ly);
  out[0] = p.real();
  out[1] = p.imag();
}
void ref_read_gauge_u1(double* u, int lx, int ly, const char* path) {
  read_gauge_u1(reinterpret_cast<cd*>(u), lx, ly, std::string(path));
}
void ref_write_gauge_u1(double* u, int lx, int ly, const char* path) {
  write_gauge_u1(reinterpret_cast<cd*>(u), lx, ly, std::string(path));
}
void ref_ape_smear_u1(double* out, double* u, int lx, int ly, double alpha, int n_iter) {
  apply_ape_smear_u1(reinterpret_cast<cd*>(out), reinterpret_cast<cd*>(u), lx, ly, alpha, n_iter);
}
double ref_topo_u1(double* u, int lx, int ly) { return get_topo_u1(reinterpret_cast<cd*>(u), lx, ly); }

}  // extern "C"

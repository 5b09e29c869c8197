/*
 * orc.h -- the oracle: a plain-C, single-threaded CPU restatement of
 * generic-linalg's solver hot path (Krylov loops, BLAS-1, 2D lattice
 * operators, U(1) input generation).
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (the HIP library under
 * generic-linalg_amd/) links, loads or calls this code; only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg use it, and only as
 * the checker.
 *
 * Arithmetic follows the reference expression by expression, compiled with
 * -O2 -ffp-contract=off and no -march flags, so that on the same inputs it is
 * bit-identical to the reference compiled by g++ -O2 (which on x86-64 baseline
 * ISA has no FMA).  Complex values are C99 double _Complex, whose layout equals
 * std::complex<double> (interleaved re, im).  Mixed real/complex operations use
 * the helpers below, which mirror libstdc++'s component-wise definitions;
 * complex/complex division uses the compiler's own libgcc routine
 * (__divdc3), the same one g++ calls for std::complex<double>.
 *
 * Pinning: every function is checked in tests/ against fixtures produced by
 * the reference itself (oracle/_ref, built from /root/reference by
 * oracle/Makefile) and against closed forms; see tests/test_oracle_*.py.
 */
#ifndef GLX_ORACLE_H
#define GLX_ORACLE_H

#include <complex.h>
#include <stddef.h>
#include <stdint.h>

typedef double _Complex orc_c;

/* ---- complex helpers (libstdc++ semantics, generic_vector.h / <complex>) -- */
static inline orc_c orc_mk(double re, double im) { return CMPLX(re, im); }
static inline orc_c orc_cmul(orc_c a, orc_c b)
{
  double ar = creal(a), ai = cimag(a), br = creal(b), bi = cimag(b);
  return CMPLX(ar * br - ai * bi, ar * bi + ai * br);
}
/* real * complex: component-wise (std::complex operator*(T, complex<T>)) */
static inline orc_c orc_rmul(double t, orc_c z) { return CMPLX(t * creal(z), t * cimag(z)); }
/* complex / real: component-wise */
static inline orc_c orc_rdiv(orc_c z, double t) { return CMPLX(creal(z) / t, cimag(z) / t); }
/* real - complex: (t - re, -im) (std::complex operator-(T, complex<T>)) */
static inline orc_c orc_rsub(double t, orc_c z) { return CMPLX(t - creal(z), -cimag(z)); }
static inline orc_c orc_add(orc_c a, orc_c b) { return CMPLX(creal(a) + creal(b), cimag(a) + cimag(b)); }
static inline orc_c orc_sub(orc_c a, orc_c b) { return CMPLX(creal(a) - creal(b), cimag(a) - cimag(b)); }
static inline orc_c orc_neg(orc_c a) { return CMPLX(-creal(a), -cimag(a)); }
static inline orc_c orc_conj(orc_c a) { return CMPLX(creal(a), -cimag(a)); }
/* complex / complex and real / complex: libgcc __divdc3 (Smith's method) */
orc_c orc_cdiv(orc_c a, orc_c b);
orc_c orc_rcdiv(double a, orc_c b);

/* ---- result and verbosity (inverter_struct.h:14-89, verbosity.h:9-129) ---- */
typedef struct {
  double resSq;      /* absolute squared true residual ||Ax-b||^2 */
  int iter;
  int success;
  int ops_count;     /* number of operator applications */
  char name[32];
} orc_info;

enum { ORC_VERB_NONE = 0, ORC_VERB_SUMMARY = 1, ORC_VERB_RESTART_DETAIL = 2, ORC_VERB_DETAIL = 3 };

/* Text sink for the verbosity lines.  `fmt` is the printf conversion used for
 * doubles ("%g" reproduces std::cout's default formatting). */
typedef struct {
  int level;
  const char* prefix;
  const char* fmt;
  char* buf;
  size_t cap;
  size_t len;
} orc_verb;

void orc_print_resid(orc_verb* v, const char* alg, int iter, int ops, double relres);
void orc_print_summary(orc_verb* v, const char* alg, int success, int iter, int ops, double relres);
void orc_print_summary_multi(orc_verb* v, int success, int iter, int ops, const double* relres, int n);
void orc_print_restart(orc_verb* v, const char* alg, int iter, int ops, double relres);
/* verbosity for the inner solve of a restarted solver (verbosity.h:27-43) */
void orc_verb_restart(orc_verb* inner, const orc_verb* outer);

/* ---- BLAS-1 (generic_vector.h:12-169) ---- */
double orc_dot_r(const double* a, const double* b, int n);
orc_c orc_dot_c(const orc_c* a, const orc_c* b, int n);
double orc_norm2sq_r(const double* a, int n);
double orc_norm2sq_c(const orc_c* a, int n);
double orc_diffnorm2sq_r(const double* a, const double* b, int n);
double orc_diffnorm2sq_c(const orc_c* a, const orc_c* b, int n);
/* long double accumulation, rounded once (no reference counterpart) */
orc_c orc_dot_cx(const orc_c* a, const orc_c* b, int n);
double orc_norm2sq_cx(const orc_c* a, int n);
double orc_diffnorm2sq_cx(const orc_c* a, const orc_c* b, int n);

/* ---- operators: the callback contract void op(T* lhs, T* rhs, void* extra) -- */
typedef void (*orc_mv_r)(double* lhs, double* rhs, void* extra);
typedef void (*orc_mv_c)(orc_c* lhs, orc_c* rhs, void* extra);

/* staggered_u1_op (operators.h:7-15): same field order and types */
typedef struct {
  orc_c* lattice;       /* links, AoS lattice[2*(y*lx+x)+mu] */
  double mass;          /* staggered mass; for the Laplacians the m^2 shift */
  int x_fine, y_fine;
  int Nc;
  double wilson_coeff;
} orc_stag_op;

void orc_laplace_real(double* lhs, double* rhs, void* extra);   /* tests/multishift/multishift.cpp:634-672 */
void orc_laplace(orc_c* lhs, orc_c* rhs, void* extra);          /* operators.cpp:28-66 */
void orc_laplace_u1(orc_c* lhs, orc_c* rhs, void* extra);       /* operators.cpp:73-120 */
void orc_staggered_u1(orc_c* lhs, orc_c* rhs, void* extra);     /* operators.cpp:184-239 */
void orc_gamma5(orc_c* lhs, orc_c* rhs, void* extra);           /* operators.cpp:242-259 */
void orc_staggered_gamma5_u1(orc_c* lhs, orc_c* rhs, void* extra); /* operators.cpp:316-368 */
void orc_staggered_dagger_u1(orc_c* lhs, orc_c* rhs, void* extra);  /* operators.cpp:372-441 */
void orc_staggered_normal_u1(orc_c* lhs, orc_c* rhs, void* extra);  /* operators.cpp:444-453 */
void orc_staggered_deo_u1(orc_c* lhs, orc_c* rhs, void* extra);     /* operators.cpp:456-496 */
void orc_staggered_doe_u1(orc_c* lhs, orc_c* rhs, void* extra);     /* operators.cpp:499-539 */
void orc_staggered_eoprec_prepare(orc_c* rhs_e, orc_c* rhs_orig, void* extra);            /* :543-563 */
void orc_staggered_m2mdeodoe_u1(orc_c* lhs, orc_c* rhs, void* extra);                     /* :566-588 */
void orc_staggered_eoprec_reconstruct(orc_c* lhs_full, orc_c* lhs_e, orc_c* rhs_o, void* extra); /* :592-616 */

/* ---- the other operators of operators.h (not on the solver path) ---- */
void orc_staggered_free(orc_c* lhs, orc_c* rhs, void* extra);          /* operators.cpp:127-178 */
void orc_staggered_gamma5_free(orc_c* lhs, orc_c* rhs, void* extra);   /* operators.cpp:262-313 */
void orc_staggered_2linklaplace_u1(orc_c* lhs, orc_c* rhs, void* extra); /* operators.cpp:626-694 */
void orc_staggered_symmshift_x(orc_c* lhs, orc_c* rhs, void* extra);   /* operators.cpp:697-736 */
void orc_staggered_symmshift_y(orc_c* lhs, orc_c* rhs, void* extra);   /* operators.cpp:739-779 */
void orc_staggered_index_operator(orc_c* lhs, orc_c* rhs, void* extra); /* operators.cpp:782-846 */

/* ---- Lattice index math (lattice/lattice.h:52-139), nd = 2 ---- */
int orc_coord_to_index(const int* dims, int nc, const int* coord, int color);
void orc_index_to_coord(const int* dims, int nc, int i, int* coord, int* color);
int orc_index_is_even(const int* dims, int nc, int i);

/* ---- stencil_2d, Nc = 1 one-link form (coarse_stencil.h:33-154) ---- */
typedef struct {
  int lx, ly;          /* lattice; nc = 1 */
  orc_c* clover;       /* V */
  orc_c* hopping;      /* 4V: {+x, +y, -x, -y} blocks */
  orc_c shift, eo_shift, dof_shift;
} orc_stencil;

void orc_stencil_staggered(orc_stencil* s, const orc_stag_op* op);        /* operators_stencil.cpp:14-63 */
void orc_stencil_staggered_gamma5(orc_stencil* s, const orc_stag_op* op); /* operators_stencil.cpp:65-116 */
void orc_stencil_staggered_dagger(orc_stencil* s, const orc_stag_op* op); /* operators_stencil.cpp:120-169 */
void orc_stencil_apply(orc_c* lhs, orc_c* rhs, void* stencil);            /* coarse_stencil.cpp:12-172 */
void orc_stencil_apply_eo(orc_c* lhs, orc_c* rhs, void* stencil);         /* coarse_stencil.cpp:395-459 */
void orc_stencil_apply_oe(orc_c* lhs, orc_c* rhs, void* stencil);         /* coarse_stencil.cpp:560-624 */
void orc_stencil_eoprec_prepare(orc_c* rhs_e, orc_c* rhs_orig, void* stencil);           /* operators_stencil.cpp:179-192 */
void orc_stencil_m2mdeodoe(orc_c* lhs, orc_c* rhs, void* stencil);                       /* operators_stencil.cpp:195-214 */
void orc_stencil_eoprec_reconstruct(orc_c* lhs_full, orc_c* lhs_e, orc_c* rhs_o, void* stencil); /* :218-236 */

/* ---- stencil_2d, any Nc, up to two links (coarse_stencil.h:14-154,
 *      coarse_stencil.cpp:12-1610) ----
 * Coefficient (term t, site s, row r, column c) of a block array is at
 * ((t V + s) nc + r) nc + c; terms: clover; hopping {+x, +y, -x, -y};
 * two_link {+2x, +x+y, +2y, -x+y, -2x, -x-y, -2y, +x-y}.  sdir follows the
 * reference's stencil_dir (0 = all terms, 1 = clover, 2..5 hopping,
 * 6..13 two-link). */
enum { ORC_PIECE_ALL = 0, ORC_PIECE_EO = 1, ORC_PIECE_OE = 2, ORC_PIECE_TB = 3, ORC_PIECE_BT = 4 };
typedef struct {
  int lx, ly, nc;
  int has_two;
  int sdir;
  orc_c* clover;    /* V nc^2 */
  orc_c* hopping;   /* 4 V nc^2 */
  orc_c* two_link;  /* 8 V nc^2 (has_two) */
  orc_c shift, eo_shift, dof_shift;
} orc_cstencil;
/* piece: ALL coarse_stencil.cpp:12-392, EO :395-557, OE :560-722,
 * TB :725-1117, BT :1120-1512 */
void orc_cstencil_apply(orc_c* lhs, orc_c* rhs, const orc_cstencil* s, int piece);
/* callback form of the full stencil (extra = orc_cstencil*) */
void orc_cstencil_apply_all(orc_c* lhs, orc_c* rhs, void* s);
/* operators_stencil.cpp:179-236 on any Nc (index_is_even of the full index) */
void orc_cstencil_eoprec_prepare(orc_c* rhs_e, orc_c* rhs_orig, const orc_cstencil* s);
void orc_cstencil_m2mdeodoe(orc_c* lhs, orc_c* rhs, const orc_cstencil* s);
void orc_cstencil_eoprec_reconstruct(orc_c* lhs_full, orc_c* lhs_e, orc_c* rhs_o, const orc_cstencil* s);
/* generate_stencil_2d (coarse_stencil.cpp:1515-1627): probe mv with unit
 * vectors and accumulate (+=) into the coefficient arrays */
void orc_cstencil_generate(orc_cstencil* s, void (*mv)(orc_c*, orc_c*, void*), void* extra);

/* ---- Krylov solvers (generic_cg.cpp, generic_cg_m.cpp, generic_bicgstab.cpp,
 *      generic_gcr.cpp).  `hist`, if non-null, receives the iterated relative
 *      residual after each iteration (max_iter entries). ---- */
orc_info orc_cg_r(double* phi, double* phi0, int size, int max_iter, double eps, orc_mv_r mv, void* extra, orc_verb* verb, double* hist);
orc_info orc_cg_c(orc_c* phi, orc_c* phi0, int size, int max_iter, double eps, orc_mv_c mv, void* extra, orc_verb* verb, double* hist);
orc_info orc_cg_restart_r(double* phi, double* phi0, int size, int max_iter, double eps, int restart_freq, orc_mv_r mv, void* extra, orc_verb* verb);
orc_info orc_cg_restart_c(orc_c* phi, orc_c* phi0, int size, int max_iter, double eps, int restart_freq, orc_mv_c mv, void* extra, orc_verb* verb);

orc_info orc_bicgstab_r(double* phi, double* phi0, int size, int max_iter, double eps, orc_mv_r mv, void* extra, orc_verb* verb, double* hist);
orc_info orc_bicgstab_c(orc_c* phi, orc_c* phi0, int size, int max_iter, double eps, orc_mv_c mv, void* extra, orc_verb* verb, double* hist);
orc_info orc_bicgstab_restart_r(double* phi, double* phi0, int size, int max_iter, double eps, int restart_freq, orc_mv_r mv, void* extra, orc_verb* verb);
orc_info orc_bicgstab_restart_c(orc_c* phi, orc_c* phi0, int size, int max_iter, double eps, int restart_freq, orc_mv_c mv, void* extra, orc_verb* verb);

orc_info orc_gcr_r(double* phi, double* phi0, int size, int max_iter, double eps, orc_mv_r mv, void* extra, orc_verb* verb, double* hist);
orc_info orc_gcr_c(orc_c* phi, orc_c* phi0, int size, int max_iter, double eps, orc_mv_c mv, void* extra, orc_verb* verb, double* hist);
orc_info orc_gcr_restart_r(double* phi, double* phi0, int size, int max_iter, double eps, int restart_freq, orc_mv_r mv, void* extra, orc_verb* verb);
orc_info orc_gcr_restart_c(orc_c* phi, orc_c* phi0, int size, int max_iter, double eps, int restart_freq, orc_mv_c mv, void* extra, orc_verb* verb);

/* Multishift CG: phi[n_shift] solution vectors, shifts[n_shift] (permuted
 * during the call and restored, as in the reference), resSqmrhs[n_shift] out. */
orc_info orc_cg_m_r(double** phi, double* phi0, int n_shift, int size, int resid_freq_check, int max_iter, double eps, double* shifts, orc_mv_r mv, void* extra, int worst_first, orc_verb* verb, double* resSqmrhs);
orc_info orc_cg_m_c(orc_c** phi, orc_c* phi0, int n_shift, int size, int resid_freq_check, int max_iter, double eps, double* shifts, orc_mv_c mv, void* extra, int worst_first, orc_verb* verb, double* resSqmrhs);

/* ---- random inputs: libstdc++ (GCC 11.4) std::mt19937,
 *      std::generate_canonical<double,53>, std::normal_distribution (polar),
 *      std::uniform_real_distribution; u1_utils.cpp ---- */
typedef struct { uint32_t mt[624]; int idx; } orc_mt19937;
void orc_mt_seed(orc_mt19937* g, uint32_t seed);
uint32_t orc_mt_next(orc_mt19937* g);
double orc_canonical(orc_mt19937* g);
typedef struct { double mean, stddev, saved; int have_saved; } orc_normal;
void orc_normal_init(orc_normal* d, double mean, double stddev);
double orc_normal_next(orc_normal* d, orc_mt19937* g);

void orc_gaussian_r(double* v, int size, orc_mt19937* g);                   /* generic_vector.h:35-45 */
void orc_gaussian_c(orc_c* v, int size, orc_mt19937* g, int imag_first);    /* generic_vector.h:48-60 */
void orc_unit_gauge_u1(orc_c* u, int lx, int ly);                           /* u1_utils.cpp:69-75 */
void orc_rand_gauge_u1(orc_c* u, int lx, int ly, orc_mt19937* g);           /* u1_utils.cpp:79-87 */
void orc_gauss_gauge_u1(orc_c* u, int lx, int ly, orc_mt19937* g, double beta); /* u1_utils.cpp:92-112 */
void orc_rand_trans_u1(orc_c* t, int lx, int ly, orc_mt19937* g);           /* u1_utils.cpp:116-124 */
void orc_apply_gauge_trans_u1(orc_c* u, const orc_c* t, int lx, int ly);    /* u1_utils.cpp:128-141 */
orc_c orc_plaquette_u1(const orc_c* u, int lx, int ly);                     /* u1_utils.cpp:190-207 */
int orc_read_gauge_u1(orc_c* u, int lx, int ly, const char* path);          /* u1_utils.cpp:17-38 */
int orc_write_gauge_u1(const orc_c* u, int lx, int ly, const char* path);   /* u1_utils.cpp:44-65 */
/* The library's device field generator (no reference counterpart; the
 * reference draws from std::mt19937 on the host): Philox4x32-10 with
 * counter {s, stream, 0, 0} and key {seed lo, seed hi} per site, Box-Muller
 * for beta != 0, uniform phases for beta = 0.  Pinned by the published
 * Random123 known-answer vectors and by distribution tests. */
void orc_philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1);
void orc_philox_gauge_u1(orc_c* u, int lx, int ly, double beta, uint64_t seed, uint32_t stream);

/* ---- gradient of S = Re phi^dag (D^dag D)^-1 phi (no reference counterpart):
 *      FP64 CPU gradient from chi = (D^dag D)^-1 phi, psi = D chi. ---- */
void orc_staggered_force(double* dS_dtheta, const orc_c* chi, const orc_c* psi, const orc_stag_op* op);
double orc_staggered_dS_dm(const orc_c* chi, const orc_stag_op* op);

#endif

/*
 * orc_core.c -- oracle: complex division, verbosity text, BLAS-1, Lattice
 * index math and libstdc++-compatible random inputs.
 * TEST INFRASTRUCTURE ONLY (see orc.h).
 */
#include "orc.h"

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

/* std::complex<double> division compiles to a call of libgcc's __divdc3; the
 * C front end emits the same call for double _Complex '/'. */
orc_c orc_cdiv(orc_c a, orc_c b) { return a / b; }
/* real / complex: libstdc++ builds complex(a, 0.0) and divides */
orc_c orc_rcdiv(double a, orc_c b)
{
  orc_c num = CMPLX(a, 0.0);
  return num / b;
}

/* ------------------------------------------------------------------------ */
/* verbosity.h:73-129.  Lines go into a caller buffer.                        */
static void vappend(orc_verb* v, const char* fmt, ...)
{
  if (!v || !v->buf || v->len >= v->cap) return;
  va_list ap;
  va_start(ap, fmt);
  int n = vsnprintf(v->buf + v->len, v->cap - v->len, fmt, ap);
  va_end(ap);
  if (n > 0) v->len += (size_t)n;
  if (v->len >= v->cap) v->len = v->cap - 1;
}
static void vnum(orc_verb* v, double x) { vappend(v, v->fmt ? v->fmt : "%g", x); }
static const char* vpre(const orc_verb* v) { return v->prefix ? v->prefix : ""; }

void orc_print_resid(orc_verb* v, const char* alg, int iter, int ops, double relres)
{
  if (!v || v->level != ORC_VERB_DETAIL) return;
  vappend(v, "%s%s Iter %d Ops %d RelRes ", vpre(v), alg, iter, ops);
  vnum(v, relres);
  vappend(v, "\n");
}
void orc_print_summary(orc_verb* v, const char* alg, int success, int iter, int ops, double relres)
{
  if (!v || v->level < ORC_VERB_SUMMARY) return;
  vappend(v, "%s%s Success %s Iter %d Ops %d RelRes ", vpre(v), alg, success ? "Y" : "N", iter, ops);
  vnum(v, relres);
  vappend(v, "\n");
}
void orc_print_summary_multi(orc_verb* v, int success, int iter, int ops, const double* relres, int n)
{
  if (!v || v->level < ORC_VERB_SUMMARY) return;
  /* the multishift summary always names "CG-M " followed by a space (verbosity.h:107) */
  vappend(v, "%sCG-M  Success %s Iter %d Ops %d RelRes ", vpre(v), success ? "Y" : "N", iter, ops);
  for (int k = 0; k < n; k++) { vnum(v, relres[k]); vappend(v, " "); }
  vappend(v, "\n");
}
void orc_print_restart(orc_verb* v, const char* alg, int iter, int ops, double relres)
{
  if (!v || v->level < ORC_VERB_RESTART_DETAIL) return;
  vappend(v, "%s%s Iter %d Ops %d RelRes ", vpre(v), alg, iter, ops);
  vnum(v, relres);
  vappend(v, "\n");
}
/* verbosity.h:27-43: SUMMARY and RESTART_DETAIL are silenced for the inner
 * solve; DETAIL passes through. */
void orc_verb_restart(orc_verb* inner, const orc_verb* outer)
{
  if (outer) {
    *inner = *outer;
    if (outer->level == ORC_VERB_RESTART_DETAIL || outer->level == ORC_VERB_SUMMARY) inner->level = ORC_VERB_NONE;
  } else {
    memset(inner, 0, sizeof(*inner));
  }
}

/* ------------------------------------------------------------------------ */
/* BLAS-1, generic_vector.h:87-169: serial left-to-right sums from zero.      */
double orc_dot_r(const double* a, const double* b, int n)
{
  double s = 0.0;
  for (int i = 0; i < n; i++) s = s + a[i] * b[i];
  return s;
}
orc_c orc_dot_c(const orc_c* a, const orc_c* b, int n)
{
  orc_c s = CMPLX(0.0, 0.0);
  for (int i = 0; i < n; i++) s = orc_add(s, orc_cmul(orc_conj(a[i]), b[i]));
  return s;
}
double orc_norm2sq_r(const double* a, int n)
{
  double s = 0.0;
  for (int i = 0; i < n; i++) s = s + a[i] * a[i];
  return s;
}
double orc_norm2sq_c(const orc_c* a, int n)
{
  double s = 0.0;
  for (int i = 0; i < n; i++) s = s + creal(orc_cmul(orc_conj(a[i]), a[i]));
  return s;
}
double orc_diffnorm2sq_r(const double* a, const double* b, int n)
{
  double s = 0.0;
  for (int i = 0; i < n; i++) s = s + (a[i] - b[i]) * (a[i] - b[i]);
  return s;
}
double orc_diffnorm2sq_c(const orc_c* a, const orc_c* b, int n)
{
  double s = 0.0;
  for (int i = 0; i < n; i++) {
    orc_c d = orc_sub(a[i], b[i]);
    s = s + creal(orc_cmul(orc_conj(d), d));
  }
  return s;
}

/* The same three reductions accumulated in long double with Neumaier's
 * compensation (x86 80-bit: a product of two doubles carries 64 significant
 * bits; the compensated sum keeps the error of millions of terms far below
 * double's last place), rounded to double once: the correctly rounded
 * reductions an FP64 solver approaches when its summation order is not the
 * reference's serial one.  Not in the reference; the "extended" reduction
 * variant of the complex solvers (orc_solvers.c). */
typedef struct {
  long double s, c;
} orc_acc;
static void acc_add(orc_acc* a, long double t)
{
  const long double u = a->s + t;
  if (fabsl(a->s) >= fabsl(t)) a->c += (a->s - u) + t;
  else a->c += (t - u) + a->s;
  a->s = u;
}
static double acc_val(const orc_acc* a) { return (double)(a->s + a->c); }

orc_c orc_dot_cx(const orc_c* a, const orc_c* b, int n)
{
  orc_acc re = {0.0L, 0.0L}, im = {0.0L, 0.0L};
  for (int i = 0; i < n; i++) {
    const long double ar = creal(a[i]), ai = cimag(a[i]), br = creal(b[i]), bi = cimag(b[i]);
    acc_add(&re, ar * br);
    acc_add(&re, ai * bi);
    acc_add(&im, ar * bi);
    acc_add(&im, -(ai * br));
  }
  return CMPLX(acc_val(&re), acc_val(&im));
}
double orc_norm2sq_cx(const orc_c* a, int n)
{
  orc_acc s = {0.0L, 0.0L};
  for (int i = 0; i < n; i++) {
    const long double ar = creal(a[i]), ai = cimag(a[i]);
    acc_add(&s, ar * ar);
    acc_add(&s, ai * ai);
  }
  return acc_val(&s);
}
double orc_diffnorm2sq_cx(const orc_c* a, const orc_c* b, int n)
{
  orc_acc s = {0.0L, 0.0L};
  for (int i = 0; i < n; i++) {
    const orc_c d = orc_sub(a[i], b[i]);
    const long double dr = creal(d), di = cimag(d);
    acc_add(&s, dr * dr);
    acc_add(&s, di * di);
  }
  return acc_val(&s);
}

/* ------------------------------------------------------------------------ */
/* Lattice (lattice/lattice.h:52-139) for nd = 2: x fastest, colour fastest of all. */
int orc_coord_to_index(const int* dims, int nc, const int* coord, int color)
{
  int i = 0;
  for (int mu = 1; mu >= 0; mu--) i = i * dims[mu] + coord[mu];
  return i * nc + color;
}
void orc_index_to_coord(const int* dims, int nc, int i, int* coord, int* color)
{
  *color = i % nc;
  i = (i - *color) / nc;
  for (int mu = 0; mu < 2; mu++) {
    coord[mu] = i % dims[mu];
    i = (i - coord[mu]) / dims[mu];
  }
}
/* even means x + y even; lattice.h:128-139 returns (sum + 1) % 2 */
int orc_index_is_even(const int* dims, int nc, int i)
{
  int coord[2], color;
  orc_index_to_coord(dims, nc, i, coord, &color);
  return (coord[0] + coord[1] + 1) % 2;
}

/* ------------------------------------------------------------------------ */
/* std::mt19937 (the standard's parameters: w=32, n=624, m=397, r=31,
 * a=0x9908b0df, u=11, d=0xffffffff, s=7, b=0x9d2c5680, t=15, c=0xefc60000,
 * l=18, f=1812433253). */
void orc_mt_seed(orc_mt19937* g, uint32_t seed)
{
  g->mt[0] = seed;
  for (int k = 1; k < 624; k++) g->mt[k] = 1812433253u * (g->mt[k - 1] ^ (g->mt[k - 1] >> 30)) + (uint32_t)k;
  g->idx = 624;
}
static void mt_twist(orc_mt19937* g)
{
  for (int k = 0; k < 624; k++) {
    uint32_t y = (g->mt[k] & 0x80000000u) | (g->mt[(k + 1) % 624] & 0x7fffffffu);
    uint32_t v = g->mt[(k + 397) % 624] ^ (y >> 1);
    if (y & 1u) v ^= 0x9908b0dfu;
    g->mt[k] = v;
  }
  g->idx = 0;
}
uint32_t orc_mt_next(orc_mt19937* g)
{
  if (g->idx >= 624) mt_twist(g);
  uint32_t y = g->mt[g->idx++];
  y ^= y >> 11;
  y ^= (y << 7) & 0x9d2c5680u;
  y ^= (y << 15) & 0xefc60000u;
  y ^= y >> 18;
  return y;
}
/* std::generate_canonical<double, 53>: two 32-bit draws, low word first,
 * summed in double and scaled by 2^-64; a result of 1.0 becomes the largest
 * double below 1. */
double orc_canonical(orc_mt19937* g)
{
  double sum = 0.0, scale = 1.0;
  for (int k = 0; k < 2; k++) {
    sum += (double)orc_mt_next(g) * scale;
    scale *= 4294967296.0;
  }
  double r = sum / scale;
  if (r >= 1.0) r = nextafter(1.0, 0.0);
  return r;
}
/* std::normal_distribution<double>: Marsaglia's polar method; each accepted
 * pair returns y*mult and keeps x*mult for the next call. */
void orc_normal_init(orc_normal* d, double mean, double stddev)
{
  d->mean = mean;
  d->stddev = stddev;
  d->saved = 0.0;
  d->have_saved = 0;
}
double orc_normal_next(orc_normal* d, orc_mt19937* g)
{
  double ret;
  if (d->have_saved) {
    d->have_saved = 0;
    ret = d->saved;
  } else {
    double x, y, r2;
    do {
      x = 2.0 * orc_canonical(g) - 1.0;
      y = 2.0 * orc_canonical(g) - 1.0;
      r2 = x * x + y * y;
    } while (r2 > 1.0 || r2 == 0.0);
    double mult = sqrt(-2.0 * log(r2) / r2);
    d->saved = x * mult;
    d->have_saved = 1;
    ret = y * mult;
  }
  return ret * d->stddev + d->mean;
}

/* generic_vector.h:35-45: a fresh N(0,1) distribution per call */
void orc_gaussian_r(double* v, int size, orc_mt19937* g)
{
  orc_normal d;
  orc_normal_init(&d, 0.0, 1.0);
  for (int i = 0; i < size; i++) v[i] = orc_normal_next(&d, g);
}
/* generic_vector.h:48-60: two draws per element; which of the two becomes
 * the real part is the compiler's argument evaluation order, so it is a
 * parameter here and pinned by the reference fixtures. */
void orc_gaussian_c(orc_c* v, int size, orc_mt19937* g, int imag_first)
{
  orc_normal d;
  orc_normal_init(&d, 0.0, 1.0);
  for (int i = 0; i < size; i++) {
    double a = orc_normal_next(&d, g);
    double b = orc_normal_next(&d, g);
    v[i] = imag_first ? CMPLX(b, a) : CMPLX(a, b);
  }
}

static orc_c polar1(double theta) { return CMPLX(1.0 * cos(theta), 1.0 * sin(theta)); }
static const double ORC_PI = 3.14159265358979323846;

void orc_unit_gauge_u1(orc_c* u, int lx, int ly)
{
  for (int i = 0; i < 2 * lx * ly; i++) u[i] = CMPLX(1.0, 0.0);
}
/* std::uniform_real_distribution(a, b): canonical * (b - a) + a */
static double uniform(orc_mt19937* g, double a, double b) { return (orc_canonical(g) * (b - a)) + a; }
void orc_rand_gauge_u1(orc_c* u, int lx, int ly, orc_mt19937* g)
{
  for (int i = 0; i < 2 * lx * ly; i++) u[i] = polar1(uniform(g, -ORC_PI, ORC_PI));
}
/* theta ~ N(0, 1/beta); beta = 0 first draws a hot start and then overwrites
 * it with N(0, inf) draws, as the reference does (u1_utils.cpp:98-110). */
void orc_gauss_gauge_u1(orc_c* u, int lx, int ly, orc_mt19937* g, double beta)
{
  if (beta < 0) beta = -beta;
  if (beta == 0) orc_rand_gauge_u1(u, lx, ly, g);
  orc_normal d;
  orc_normal_init(&d, 0.0, 1.0 / sqrt(beta));
  for (int i = 0; i < 2 * lx * ly; i++) u[i] = polar1(orc_normal_next(&d, g));
}
void orc_rand_trans_u1(orc_c* t, int lx, int ly, orc_mt19937* g)
{
  for (int i = 0; i < lx * ly; i++) t[i] = polar1(uniform(g, -ORC_PI, ORC_PI));
}
/* U_mu(x) <- g(x) U_mu(x) g*(x+mu), sites in lexicographic order, in place */
void orc_apply_gauge_trans_u1(orc_c* u, const orc_c* t, int lx, int ly)
{
  for (int y = 0; y < ly; y++)
    for (int x = 0; x < lx; x++) {
      int s = lx * y + x;
      u[2 * s] = orc_cmul(orc_cmul(t[s], u[2 * s]), orc_conj(t[lx * y + (x + 1) % lx]));
      u[2 * s + 1] = orc_cmul(orc_cmul(t[s], u[2 * s + 1]), orc_conj(t[lx * ((y + 1) % ly) + x]));
    }
}
orc_c orc_plaquette_u1(const orc_c* u, int lx, int ly)
{
  orc_c plaq = CMPLX(0.0, 0.0);
  for (int y = 0; y < ly; y++)
    for (int x = 0; x < lx; x++) {
      orc_c p = orc_cmul(orc_cmul(orc_cmul(u[2 * lx * y + 2 * x], u[2 * lx * y + 2 * ((x + 1) % lx) + 1]),
                                  orc_conj(u[2 * lx * ((y + 1) % ly) + 2 * x])),
                         orc_conj(u[2 * lx * y + 2 * x + 1]));
      plaq = orc_add(plaq, p);
    }
  return orc_rdiv(plaq, (double)(lx * ly));
}
/* Text format: one phase per line, x outermost, then y, then mu. */
int orc_read_gauge_u1(orc_c* u, int lx, int ly, const char* path)
{
  FILE* f = fopen(path, "r");
  if (!f) return -1;
  int rc = 0;
  for (int x = 0; x < lx && rc == 0; x++)
    for (int y = 0; y < ly && rc == 0; y++)
      for (int mu = 0; mu < 2; mu++) {
        double ph;
        if (fscanf(f, "%lf", &ph) != 1) { rc = -2; break; }
        u[y * 2 * lx + x * 2 + mu] = polar1(ph);
      }
  fclose(f);
  return rc;
}
int orc_write_gauge_u1(const orc_c* u, int lx, int ly, const char* path)
{
  FILE* f = fopen(path, "w");
  if (!f) return -1;
  for (int x = 0; x < lx; x++)
    for (int y = 0; y < ly; y++)
      for (int mu = 0; mu < 2; mu++) fprintf(f, "%.20f\n", carg(u[y * 2 * lx + x * 2 + mu]));
  fclose(f);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) and the device field     */
/* generator built on it (see orc.h).                                          */
void orc_philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1)
{
  for (int r = 0; r < 10; r++) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c[0], p1 = (uint64_t)0xCD9E8D57u * c[2];
    uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0, n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
    c[0] = n0;
    c[1] = (uint32_t)p1;
    c[2] = n2;
    c[3] = (uint32_t)p0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}

void orc_philox_gauge_u1(orc_c* u, int lx, int ly, double beta, uint64_t seed, uint32_t stream)
{
  const double two_m53 = 1.1102230246251565404e-16, two_pi = 6.283185307179586232, pi = 3.141592653589793116;
  if (beta < 0) beta = -beta;
  const double sigma = beta > 0 ? 1.0 / sqrt(beta) : 0.0;
  for (int s = 0; s < lx * ly; s++) {
    uint32_t c[4] = {(uint32_t)s, stream, 0u, 0u};
    orc_philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    double a = (double)((((uint64_t)c[1] << 32) | c[0]) >> 11) * two_m53;
    double b = (double)((((uint64_t)c[3] << 32) | c[2]) >> 11) * two_m53;
    double t0, t1;
    if (sigma > 0) {
      double r = sigma * sqrt(-2.0 * log(1.0 - a));
      t0 = r * cos(two_pi * b);
      t1 = r * sin(two_pi * b);
    } else {
      t0 = two_pi * a - pi;
      t1 = two_pi * b - pi;
    }
    u[2 * s] = CMPLX(cos(t0), sin(t0));
    u[2 * s + 1] = CMPLX(cos(t1), sin(t1));
  }
}

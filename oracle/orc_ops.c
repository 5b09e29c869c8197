/*
 * orc_ops.c -- oracle: 2D lattice operators and the Nc = 1 stencil form.
 * TEST INFRASTRUCTURE ONLY (see orc.h).
 *
 * Vectors are lexicographic, i = y*lx + x; links AoS u[2*i + mu] (mu = 0: x,
 * mu = 1: y).  Each operator overwrites every entry of lhs and accumulates in
 * the reference's term order, starting from zero.
 */
#include "orc.h"

#include <stdlib.h>

/* tests/multishift/multishift.cpp:634-672: real Laplacian with Nc colours;
 * `mass` is the m^2 shift, diagonal 4 + mass. */
void orc_laplace_real(double* lhs, double* rhs, void* extra)
{
  const orc_stag_op* op = (const orc_stag_op*)extra;
  const int lx = op->x_fine, ly = op->y_fine, nc = op->Nc;
  const double mass = op->mass;
  for (int i = 0; i < lx * ly * nc; i++) {
    const int c = i % nc, s = (i - c) / nc, x = s % lx, y = s / lx;
    double acc = 0.0;
    acc = acc - rhs[y * lx * nc + ((x + 1) % lx) * nc + c];
    acc = acc - rhs[y * lx * nc + ((x + lx - 1) % lx) * nc + c];
    acc = acc - rhs[((y + 1) % ly) * lx * nc + x * nc + c];
    acc = acc - rhs[((y + ly - 1) % ly) * lx * nc + x * nc + c];
    acc = acc + (4 + mass) * rhs[i];
    lhs[i] = acc;
  }
}

/* operators.cpp:28-66: the same stencil on complex vectors */
void orc_laplace(orc_c* lhs, orc_c* rhs, void* extra)
{
  const orc_stag_op* op = (const orc_stag_op*)extra;
  const int lx = op->x_fine, ly = op->y_fine, nc = op->Nc;
  const double mass = op->mass;
  for (int i = 0; i < lx * ly * nc; i++) {
    const int c = i % nc, s = (i - c) / nc, x = s % lx, y = s / lx;
    orc_c acc = CMPLX(0.0, 0.0);
    acc = orc_sub(acc, rhs[y * lx * nc + ((x + 1) % lx) * nc + c]);
    acc = orc_sub(acc, rhs[y * lx * nc + ((x + lx - 1) % lx) * nc + c]);
    acc = orc_sub(acc, rhs[((y + 1) % ly) * lx * nc + x * nc + c]);
    acc = orc_sub(acc, rhs[((y + ly - 1) % ly) * lx * nc + x * nc + c]);
    acc = orc_add(acc, orc_rmul(4 + mass, rhs[i]));
    lhs[i] = acc;
  }
}

/* operators.cpp:73-120: gauged Laplacian, diagonal 4 + mass */
void orc_laplace_u1(orc_c* lhs, orc_c* rhs, void* extra)
{
  const orc_stag_op* op = (const orc_stag_op*)extra;
  const int lx = op->x_fine, ly = op->y_fine;
  const orc_c* u = op->lattice;
  const double mass = op->mass;
  for (int i = 0; i < lx * ly; i++) {
    const int x = i % lx, y = i / lx;
    const int xp = (x + 1) % lx, xm = (x + lx - 1) % lx, yp = (y + 1) % ly, ym = (y + ly - 1) % ly;
    orc_c acc = CMPLX(0.0, 0.0);
    acc = orc_sub(acc, orc_cmul(u[2 * (y * lx + x)], rhs[y * lx + xp]));
    acc = orc_sub(acc, orc_cmul(orc_conj(u[2 * (y * lx + xm)]), rhs[y * lx + xm]));
    acc = orc_sub(acc, orc_cmul(u[2 * (y * lx + x) + 1], rhs[yp * lx + x]));
    acc = orc_sub(acc, orc_cmul(orc_conj(u[2 * (ym * lx + x) + 1]), rhs[ym * lx + x]));
    acc = orc_add(acc, orc_rmul(4 + mass, rhs[i]));
    lhs[i] = acc;
  }
}

/* Hopping part of the staggered operator with an overall sign `sg` (+1 for D,
 * -1 for D^dag) and an optional per-site factor; terms in the order
 * +x, -x, +y, -y of operators.cpp:215-224 (D) and :403-412 (D^dag).
 * parity_only: -1 all sites, 0 even sites only, 1 odd sites only (other
 * sites are written as zero). */
static void staggered_core(orc_c* lhs, orc_c* rhs, const orc_stag_op* op, double sg, int with_mass,
                           int parity_only, int gamma5)
{
  const int lx = op->x_fine, ly = op->y_fine;
  const orc_c* u = op->lattice;
  const double mass = op->mass;
  for (int i = 0; i < lx * ly; i++) {
    const int x = i % lx, y = i / lx;
    const double eta1 = 1 - 2 * (x % 2);
    if (parity_only >= 0 && (x + y) % 2 != parity_only) {
      lhs[i] = CMPLX(0.0, 0.0);
      continue;
    }
    const int xp = (x + 1) % lx, xm = (x + lx - 1) % lx, yp = (y + 1) % ly, ym = (y + ly - 1) % ly;
    orc_c t0, t1, t2, t3;
    if (gamma5) {
      /* operators.cpp:339-348: eo_sign*U*rhs, (eo_sign*eta1)*U*rhs */
      const double eo = ((x + y) % 2 == 0) ? 1.0 : -1.0;
      t0 = orc_cmul(orc_rmul(eo, u[2 * (y * lx + x)]), rhs[y * lx + xp]);
      t1 = orc_cmul(orc_rmul(eo, orc_conj(u[2 * (y * lx + xm)])), rhs[y * lx + xm]);
      t2 = orc_cmul(orc_rmul(eo * eta1, u[2 * (y * lx + x) + 1]), rhs[yp * lx + x]);
      t3 = orc_cmul(orc_rmul(eo * eta1, orc_conj(u[2 * (ym * lx + x) + 1])), rhs[ym * lx + x]);
    } else {
      t0 = orc_cmul(u[2 * (y * lx + x)], rhs[y * lx + xp]);
      t1 = orc_cmul(orc_conj(u[2 * (y * lx + xm)]), rhs[y * lx + xm]);
      t2 = orc_cmul(orc_rmul(eta1, u[2 * (y * lx + x) + 1]), rhs[yp * lx + x]);
      t3 = orc_cmul(orc_rmul(eta1, orc_conj(u[2 * (ym * lx + x) + 1])), rhs[ym * lx + x]);
    }
    orc_c acc = CMPLX(0.0, 0.0);
    if (sg > 0) {
      acc = orc_sub(acc, t0);
      acc = orc_add(acc, t1);
      acc = orc_sub(acc, t2);
      acc = orc_add(acc, t3);
    } else {
      acc = orc_add(acc, t0);
      acc = orc_sub(acc, t1);
      acc = orc_add(acc, t2);
      acc = orc_sub(acc, t3);
    }
    acc = orc_rmul(0.5, acc);
    if (with_mass) {
      if (gamma5) {
        const double eo = ((x + y) % 2 == 0) ? 1.0 : -1.0;
        acc = orc_add(acc, orc_rmul(eo * mass, rhs[i]));
      } else {
        acc = orc_add(acc, orc_rmul(mass, rhs[i]));
      }
    }
    lhs[i] = acc;
  }
}

void orc_staggered_u1(orc_c* lhs, orc_c* rhs, void* extra)
{
  staggered_core(lhs, rhs, (const orc_stag_op*)extra, +1.0, 1, -1, 0);
}
void orc_staggered_dagger_u1(orc_c* lhs, orc_c* rhs, void* extra)
{
  staggered_core(lhs, rhs, (const orc_stag_op*)extra, -1.0, 1, -1, 0);
}
void orc_staggered_gamma5_u1(orc_c* lhs, orc_c* rhs, void* extra)
{
  staggered_core(lhs, rhs, (const orc_stag_op*)extra, +1.0, 1, -1, 1);
}
void orc_staggered_deo_u1(orc_c* lhs, orc_c* rhs, void* extra)
{
  staggered_core(lhs, rhs, (const orc_stag_op*)extra, +1.0, 0, 0, 0);
}
void orc_staggered_doe_u1(orc_c* lhs, orc_c* rhs, void* extra)
{
  staggered_core(lhs, rhs, (const orc_stag_op*)extra, +1.0, 0, 1, 0);
}

/* operators.cpp:242-259 */
void orc_gamma5(orc_c* lhs, orc_c* rhs, void* extra)
{
  const orc_stag_op* op = (const orc_stag_op*)extra;
  const int lx = op->x_fine, ly = op->y_fine;
  for (int i = 0; i < lx * ly; i++) {
    const int x = i % lx, y = i / lx;
    lhs[i] = orc_rmul((double)(1 - 2 * ((x + y) % 2)), rhs[i]);
  }
}

/* operators.cpp:444-453: D^dag (D rhs) through a temporary */
void orc_staggered_normal_u1(orc_c* lhs, orc_c* rhs, void* extra)
{
  const orc_stag_op* op = (const orc_stag_op*)extra;
  orc_c* tmp = (orc_c*)malloc(sizeof(orc_c) * (size_t)op->x_fine * op->y_fine);
  orc_staggered_u1(tmp, rhs, extra);
  orc_staggered_dagger_u1(lhs, tmp, extra);
  free(tmp);
}

/* operators.cpp:543-563: rhs_e = m b_e - D_eo b_o on even sites, 0 on odd */
void orc_staggered_eoprec_prepare(orc_c* rhs_e, orc_c* rhs_orig, void* extra)
{
  const orc_stag_op* op = (const orc_stag_op*)extra;
  const int lx = op->x_fine, n = op->x_fine * op->y_fine;
  orc_staggered_deo_u1(rhs_e, rhs_orig, extra);
  for (int i = 0; i < n; i++) {
    const int x = i % lx, y = i / lx;
    if ((x + y) % 2 == 0) rhs_e[i] = orc_sub(orc_rmul(op->mass, rhs_orig[i]), rhs_e[i]);
  }
}
/* operators.cpp:566-588: (m^2 - D_eo D_oe) on even sites, 0 on odd */
void orc_staggered_m2mdeodoe_u1(orc_c* lhs, orc_c* rhs, void* extra)
{
  const orc_stag_op* op = (const orc_stag_op*)extra;
  const int lx = op->x_fine, n = op->x_fine * op->y_fine;
  orc_c* tmp = (orc_c*)malloc(sizeof(orc_c) * (size_t)n);
  orc_staggered_doe_u1(tmp, rhs, extra);
  orc_staggered_deo_u1(lhs, tmp, extra);
  for (int i = 0; i < n; i++) {
    const int x = i % lx, y = i / lx;
    if ((x + y) % 2 == 0) lhs[i] = orc_sub(orc_rmul(op->mass * op->mass, rhs[i]), lhs[i]);
  }
  free(tmp);
}
/* operators.cpp:592-616: odd x = (b_o - D_oe x_e)/m, even copied */
void orc_staggered_eoprec_reconstruct(orc_c* lhs_full, orc_c* lhs_e, orc_c* rhs_o, void* extra)
{
  const orc_stag_op* op = (const orc_stag_op*)extra;
  const int lx = op->x_fine, n = op->x_fine * op->y_fine;
  const double inv_mass = 1.0 / op->mass;
  orc_staggered_doe_u1(lhs_full, lhs_e, extra);
  for (int i = 0; i < n; i++) {
    const int x = i % lx, y = i / lx;
    if ((x + y) % 2 == 1)
      lhs_full[i] = orc_rmul(inv_mass, orc_sub(rhs_o[i], lhs_full[i]));
    else
      lhs_full[i] = lhs_e[i];
  }
}

/* ------------------------------------------------------------------------ */
/* Stencil form, Nc = 1.  Coefficient blocks hopping[i + d*V] with
 * d = 0:+x, 1:+y, 2:-x, 3:-y (coarse_stencil.h:44, operators_stencil.cpp:41-55). */
static void stencil_build(orc_stencil* s, const orc_stag_op* op, double sg, int gamma5)
{
  const int lx = op->x_fine, ly = op->y_fine, V = lx * ly;
  const int dims[2] = {lx, ly};
  s->lx = lx;
  s->ly = ly;
  for (int i = 0; i < V; i++) {
    s->clover[i] = CMPLX(0.0, 0.0);
    int coord[2], color, tmp[2];
    orc_index_to_coord(dims, 1, i, coord, &color);
    const int eta1 = 1 - 2 * (coord[0] % 2);
    const int eo = gamma5 ? (orc_index_is_even(dims, 1, i) ? 1 : -1) : 1;
    const double h = -0.5 * sg; /* -0.5 for D, +0.5 for D^dag */
    if (gamma5) {
      s->hopping[i] = orc_rmul(-0.5 * eo, op->lattice[2 * i]);
      s->hopping[i + V] = orc_rmul(-0.5 * eo * eta1, op->lattice[2 * i + 1]);
    } else {
      s->hopping[i] = orc_rmul(h, op->lattice[2 * i]);
      s->hopping[i + V] = orc_rmul(h * eta1, op->lattice[2 * i + 1]);
    }
    tmp[0] = (coord[0] - 1 + lx) % lx;
    tmp[1] = coord[1];
    const orc_c ux = orc_conj(op->lattice[2 * orc_coord_to_index(dims, 1, tmp, 0)]);
    tmp[0] = coord[0];
    tmp[1] = (coord[1] - 1 + ly) % ly;
    const orc_c uy = orc_conj(op->lattice[2 * orc_coord_to_index(dims, 1, tmp, 0) + 1]);
    if (gamma5) {
      s->hopping[i + 2 * V] = orc_rmul(0.5 * eo, ux);
      s->hopping[i + 3 * V] = orc_rmul(0.5 * eo * eta1, uy);
    } else {
      s->hopping[i + 2 * V] = orc_rmul(-h, ux);
      s->hopping[i + 3 * V] = orc_rmul(-h * eta1, uy);
    }
  }
  s->shift = gamma5 ? CMPLX(0.0, 0.0) : CMPLX(op->mass, 0.0);
  s->eo_shift = gamma5 ? CMPLX(op->mass, 0.0) : CMPLX(0.0, 0.0);
  s->dof_shift = CMPLX(0.0, 0.0);
}
void orc_stencil_staggered(orc_stencil* s, const orc_stag_op* op) { stencil_build(s, op, +1.0, 0); }
void orc_stencil_staggered_gamma5(orc_stencil* s, const orc_stag_op* op) { stencil_build(s, op, +1.0, 1); }
void orc_stencil_staggered_dagger(orc_stencil* s, const orc_stag_op* op) { stencil_build(s, op, -1.0, 0); }

/* parity: -1 all sites with clover and shifts (coarse_stencil.cpp:32-171);
 * 0 even-only hopping (:415-459); 1 odd-only hopping (:580-624). */
static void stencil_apply(orc_c* lhs, orc_c* rhs, const orc_stencil* s, int parity)
{
  const int lx = s->lx, ly = s->ly, V = lx * ly;
  const int dims[2] = {lx, ly};
  for (int i = 0; i < V; i++) {
    orc_c acc = CMPLX(0.0, 0.0);
    const int even = orc_index_is_even(dims, 1, i);
    if ((parity == 0 && !even) || (parity == 1 && even)) {
      lhs[i] = acc;
      continue;
    }
    int coord[2], color, t[2];
    orc_index_to_coord(dims, 1, i, coord, &color);
    if (parity < 0) acc = orc_add(acc, orc_cmul(s->clover[i], rhs[orc_coord_to_index(dims, 1, coord, 0)]));
    t[0] = (coord[0] + 1) % lx; t[1] = coord[1];
    acc = orc_add(acc, orc_cmul(s->hopping[i], rhs[orc_coord_to_index(dims, 1, t, 0)]));
    t[0] = coord[0]; t[1] = (coord[1] + 1) % ly;
    acc = orc_add(acc, orc_cmul(s->hopping[i + V], rhs[orc_coord_to_index(dims, 1, t, 0)]));
    t[0] = (coord[0] - 1 + lx) % lx; t[1] = coord[1];
    acc = orc_add(acc, orc_cmul(s->hopping[i + 2 * V], rhs[orc_coord_to_index(dims, 1, t, 0)]));
    t[0] = coord[0]; t[1] = (coord[1] - 1 + ly) % ly;
    acc = orc_add(acc, orc_cmul(s->hopping[i + 3 * V], rhs[orc_coord_to_index(dims, 1, t, 0)]));
    if (parity < 0) {
      if (cabs(s->shift) != 0.0) acc = orc_add(acc, orc_cmul(s->shift, rhs[i]));
      if (cabs(s->eo_shift) != 0.0) acc = orc_add(acc, orc_cmul(orc_rmul(even ? 1.0 : -1.0, s->eo_shift), rhs[i]));
      /* dof_shift acts on colour halves; with nc = 1 every site is in the
       * lower half (colour 0 < 1/2 is false), coarse_stencil.cpp:165-169 */
      if (cabs(s->dof_shift) != 0.0) acc = orc_add(acc, orc_cmul(orc_rmul(-1.0, s->dof_shift), rhs[i]));
    }
    lhs[i] = acc;
  }
}
void orc_stencil_apply(orc_c* lhs, orc_c* rhs, void* st) { stencil_apply(lhs, rhs, (const orc_stencil*)st, -1); }
void orc_stencil_apply_eo(orc_c* lhs, orc_c* rhs, void* st) { stencil_apply(lhs, rhs, (const orc_stencil*)st, 0); }
void orc_stencil_apply_oe(orc_c* lhs, orc_c* rhs, void* st) { stencil_apply(lhs, rhs, (const orc_stencil*)st, 1); }

void orc_stencil_eoprec_prepare(orc_c* rhs_e, orc_c* rhs_orig, void* st)
{
  const orc_stencil* s = (const orc_stencil*)st;
  const int dims[2] = {s->lx, s->ly};
  orc_stencil_apply_eo(rhs_e, rhs_orig, st);
  for (int i = 0; i < s->lx * s->ly; i++)
    if (orc_index_is_even(dims, 1, i)) rhs_e[i] = orc_sub(orc_cmul(s->shift, rhs_orig[i]), rhs_e[i]);
}
void orc_stencil_m2mdeodoe(orc_c* lhs, orc_c* rhs, void* st)
{
  const orc_stencil* s = (const orc_stencil*)st;
  const int dims[2] = {s->lx, s->ly}, V = s->lx * s->ly;
  orc_c* tmp = (orc_c*)malloc(sizeof(orc_c) * (size_t)V);
  orc_stencil_apply_oe(tmp, rhs, st);
  orc_stencil_apply_eo(lhs, tmp, st);
  for (int i = 0; i < V; i++)
    if (orc_index_is_even(dims, 1, i)) lhs[i] = orc_sub(orc_cmul(orc_cmul(s->shift, s->shift), rhs[i]), lhs[i]);
  free(tmp);
}
void orc_stencil_eoprec_reconstruct(orc_c* lhs_full, orc_c* lhs_e, orc_c* rhs_o, void* st)
{
  const orc_stencil* s = (const orc_stencil*)st;
  const int dims[2] = {s->lx, s->ly}, V = s->lx * s->ly;
  const double inv_mass = 1.0 / creal(s->shift);
  orc_stencil_apply_oe(lhs_full, lhs_e, st);
  for (int i = 0; i < V; i++) {
    if (!orc_index_is_even(dims, 1, i))
      lhs_full[i] = orc_rmul(inv_mass, orc_sub(rhs_o[i], lhs_full[i]));
    else
      lhs_full[i] = lhs_e[i];
  }
}

/* ------------------------------------------------------------------------ */
/* operators.cpp:127-846: the operators of operators.h off the solver path,  */
/* each term added in the reference's order and association.                 */
void orc_staggered_free(orc_c* lhs, orc_c* rhs, void* extra)
{
  const orc_stag_op* op = (const orc_stag_op*)extra;
  const int lx = op->x_fine, ly = op->y_fine;
  for (int i = 0; i < lx * ly; i++) {
    const int x = i % lx, y = i / lx;
    const double eta1 = 1 - 2 * (x % 2);
    orc_c v = CMPLX(0.0, 0.0);
    v = orc_sub(v, rhs[y * lx + ((x + 1) % lx)]);
    v = orc_add(v, rhs[y * lx + ((x + lx - 1) % lx)]);
    v = orc_sub(v, orc_rmul(eta1, rhs[((y + 1) % ly) * lx + x]));
    v = orc_add(v, orc_rmul(eta1, rhs[((y + ly - 1) % ly) * lx + x]));
    v = orc_rmul(0.5, v);
    lhs[i] = orc_add(v, orc_rmul(op->mass, rhs[i]));
  }
}
void orc_staggered_gamma5_free(orc_c* lhs, orc_c* rhs, void* extra)
{
  const orc_stag_op* op = (const orc_stag_op*)extra;
  const int lx = op->x_fine, ly = op->y_fine;
  for (int i = 0; i < lx * ly; i++) {
    const int x = i % lx, y = i / lx;
    const double eta1 = 1 - 2 * (x % 2);
    const double eo = ((x + y) % 2 == 0) ? 1.0 : -1.0;
    orc_c v = CMPLX(0.0, 0.0);
    v = orc_sub(v, orc_rmul(eo, rhs[y * lx + ((x + 1) % lx)]));
    v = orc_add(v, orc_rmul(eo, rhs[y * lx + ((x + lx - 1) % lx)]));
    v = orc_sub(v, orc_rmul(eo * eta1, rhs[((y + 1) % ly) * lx + x]));
    v = orc_add(v, orc_rmul(eo * eta1, rhs[((y + ly - 1) % ly) * lx + x]));
    v = orc_rmul(0.5, v);
    lhs[i] = orc_add(v, orc_rmul(eo * op->mass, rhs[i]));
  }
}
void orc_staggered_2linklaplace_u1(orc_c* lhs, orc_c* rhs, void* extra)
{
  const orc_stag_op* op = (const orc_stag_op*)extra;
  const orc_c* U = op->lattice;
  const int lx = op->x_fine, ly = op->y_fine;
  const double wc = op->wilson_coeff;
  for (int i = 0; i < lx * ly; i++) {
    const int x = i % lx, y = i / lx;
    const double eta1 = 1 - 2 * (x % 2);
    const int xp = (x + 1) % lx, xm = (x + lx - 1) % lx, xm2 = (x + lx - 2) % lx, xp2 = (x + 2) % lx;
    const int yp = (y + 1) % ly, ym = (y + ly - 1) % ly, ym2 = (y + ly - 2) % ly, yp2 = (y + 2) % ly;
    orc_c v = CMPLX(0.0, 0.0);
    v = orc_sub(v, orc_cmul(U[y * lx * 2 + x * 2], rhs[y * lx + xp]));
    v = orc_add(v, orc_cmul(orc_conj(U[y * lx * 2 + xm * 2]), rhs[y * lx + xm]));
    v = orc_sub(v, orc_cmul(orc_rmul(eta1, U[y * lx * 2 + x * 2 + 1]), rhs[yp * lx + x]));
    v = orc_add(v, orc_cmul(orc_rmul(eta1, orc_conj(U[ym * lx * 2 + x * 2 + 1])), rhs[ym * lx + x]));
    v = orc_rmul(0.5, v);
    v = orc_add(v, orc_rmul(op->mass, rhs[i]));
    v = orc_add(v, orc_rmul(wc * 4.0, rhs[i]));
    v = orc_sub(v, orc_cmul(orc_cmul(orc_rmul(wc, U[y * lx * 2 + x * 2]), U[y * lx * 2 + xp * 2]), rhs[y * lx + xp2]));
    v = orc_sub(v, orc_cmul(orc_cmul(orc_rmul(wc, orc_conj(U[y * lx * 2 + xm * 2])), orc_conj(U[y * lx * 2 + xm2 * 2])),
                            rhs[y * lx + xm2]));
    v = orc_sub(v, orc_cmul(orc_cmul(orc_rmul(wc, U[y * lx * 2 + x * 2 + 1]), U[yp * lx * 2 + x * 2 + 1]),
                            rhs[yp2 * lx + x]));
    v = orc_sub(v, orc_cmul(orc_cmul(orc_rmul(wc, orc_conj(U[ym * lx * 2 + x * 2 + 1])),
                                     orc_conj(U[ym2 * lx * 2 + x * 2 + 1])),
                            rhs[ym2 * lx + x]));
    lhs[i] = v;
  }
}
void orc_staggered_symmshift_x(orc_c* lhs, orc_c* rhs, void* extra)
{
  const orc_stag_op* op = (const orc_stag_op*)extra;
  const orc_c* U = op->lattice;
  const int lx = op->x_fine, ly = op->y_fine;
  for (int i = 0; i < lx * ly; i++) {
    const int x = i % lx, y = i / lx;
    const int xm = (x + lx - 1) % lx;
    orc_c v = CMPLX(0.0, 0.0);
    v = orc_add(v, orc_cmul(U[y * lx * 2 + x * 2], rhs[y * lx + ((x + 1) % lx)]));
    v = orc_add(v, orc_cmul(orc_conj(U[y * lx * 2 + xm * 2]), rhs[y * lx + xm]));
    lhs[i] = orc_rmul(0.5, v);
  }
}
void orc_staggered_symmshift_y(orc_c* lhs, orc_c* rhs, void* extra)
{
  const orc_stag_op* op = (const orc_stag_op*)extra;
  const orc_c* U = op->lattice;
  const int lx = op->x_fine, ly = op->y_fine;
  for (int i = 0; i < lx * ly; i++) {
    const int x = i % lx, y = i / lx;
    const double eta1 = 1 - 2 * (x % 2);
    const int ym = (y + ly - 1) % ly;
    orc_c v = CMPLX(0.0, 0.0);
    v = orc_add(v, orc_cmul(orc_rmul(eta1, U[y * lx * 2 + x * 2 + 1]), rhs[((y + 1) % ly) * lx + x]));
    v = orc_add(v, orc_cmul(orc_rmul(eta1, orc_conj(U[ym * lx * 2 + x * 2 + 1])), rhs[ym * lx + x]));
    lhs[i] = orc_rmul(0.5, v);
  }
}
void orc_staggered_index_operator(orc_c* lhs, orc_c* rhs, void* extra)
{
  const orc_stag_op* op = (const orc_stag_op*)extra;
  const int V = op->x_fine * op->y_fine;
  const orc_c ci = CMPLX(0.0, 1.0), cid2 = CMPLX(0.0, 0.5);
  orc_c* tmp = (orc_c*)malloc(sizeof(orc_c) * (size_t)V);
  orc_c* tmp2 = (orc_c*)malloc(sizeof(orc_c) * (size_t)V);
  orc_stag_op nomass = *op;
  nomass.mass = 0.0;
  for (int i = 0; i < V; i++) lhs[i] = CMPLX(0.0, 0.0);
  orc_staggered_u1(lhs, rhs, &nomass);
  for (int i = 0; i < V; i++) lhs[i] = orc_cmul(lhs[i], ci);
  orc_staggered_symmshift_y(tmp, rhs, extra);
  orc_staggered_symmshift_x(tmp2, tmp, extra);
  const orc_c mI = orc_rmul(op->mass, cid2);
  for (int i = 0; i < V; i++) lhs[i] = orc_sub(lhs[i], orc_cmul(mI, tmp2[i]));
  orc_staggered_symmshift_x(tmp, rhs, extra);
  orc_staggered_symmshift_y(tmp2, tmp, extra);
  for (int i = 0; i < V; i++) lhs[i] = orc_add(lhs[i], orc_cmul(mI, tmp2[i]));
  free(tmp);
  free(tmp2);
}

/* ------------------------------------------------------------------------ */
/* stencil_2d for any Nc (coarse_stencil.cpp:12-1627).                          */
/* term offsets (dx, dy): clover, hopping 0..3, two-link 0..7 */
static const int cst_off[13][2] = {{0, 0},  {1, 0},  {0, 1},   {-1, 0}, {0, -1}, {2, 0}, {1, 1},
                                   {0, 2},  {-1, 1}, {-2, 0},  {-1, -1}, {0, -2}, {1, -1}};

static const orc_c* cst_block(const orc_cstencil* s, int t)
{
  const size_t blk = (size_t)s->lx * s->ly * s->nc * s->nc;
  if (t == 0) return s->clover;
  if (t < 5) return s->hopping + (size_t)(t - 1) * blk;
  return s->two_link + (size_t)(t - 5) * blk;
}

/* Which terms a piece applies for the stencil's sdir (see orc.h). */
static int cst_term_on(const orc_cstencil* s, int piece, int t)
{
  const int eo = piece == ORC_PIECE_EO || piece == ORC_PIECE_OE;
  if (t >= 5 && !s->has_two) return 0;
  if (s->sdir == 0) return eo ? (t >= 1 && t <= 4) : 1;
  if (s->sdir == 1) return !eo && t == 0;
  if (s->sdir <= 5) return t == s->sdir - 1;
  return !eo && t == s->sdir - 1;
}

void orc_cstencil_apply(orc_c* lhs, orc_c* rhs, const orc_cstencil* s, int piece)
{
  const int lx = s->lx, ly = s->ly, nc = s->nc;
  const int dims[2] = {lx, ly};
  const int n = lx * ly * nc;
  /* a two-link direction without two-link terms leaves lhs as it was
   * (coarse_stencil.cpp:264-391 and the tb/bt copies) */
  if (s->sdir >= 6 && !s->has_two && (piece == ORC_PIECE_ALL || piece == ORC_PIECE_TB || piece == ORC_PIECE_BT))
    return;
  const int c_lo = piece == ORC_PIECE_TB ? nc / 2 : 0;
  const int c_hi = piece == ORC_PIECE_BT ? nc / 2 : nc;
  const int shifts = piece == ORC_PIECE_ALL && (s->sdir == 0 || s->sdir == 1);
  for (int i = 0; i < n; i++) {
    int coord[2], color;
    orc_index_to_coord(dims, nc, i, coord, &color);
    const int even = orc_index_is_even(dims, nc, i);
    orc_c acc = CMPLX(0.0, 0.0);
    int row_on = 1;
    if (piece == ORC_PIECE_EO) row_on = even;
    if (piece == ORC_PIECE_OE) row_on = !even;
    if (piece == ORC_PIECE_TB) row_on = color < nc / 2;
    if (piece == ORC_PIECE_BT) row_on = color >= nc / 2;
    if (row_on) {
      for (int t = 0; t < 13; t++) {
        if (!cst_term_on(s, piece, t)) continue;
        int nb[2];
        nb[0] = ((coord[0] + cst_off[t][0]) % lx + lx) % lx;
        nb[1] = ((coord[1] + cst_off[t][1]) % ly + ly) % ly;
        const orc_c* blk = cst_block(s, t);
        for (int c = c_lo; c < c_hi; c++)
          acc = orc_add(acc, orc_cmul(blk[(size_t)c + (size_t)nc * i], rhs[orc_coord_to_index(dims, nc, nb, c)]));
      }
      if (shifts) {
        if (cabs(s->shift) != 0.0) acc = orc_add(acc, orc_cmul(s->shift, rhs[i]));
        if (cabs(s->eo_shift) != 0.0) acc = orc_add(acc, orc_cmul(orc_rmul(even ? 1.0 : -1.0, s->eo_shift), rhs[i]));
        if (cabs(s->dof_shift) != 0.0)
          acc = orc_add(acc, orc_cmul(orc_rmul(color < nc / 2 ? 1.0 : -1.0, s->dof_shift), rhs[i]));
      }
    }
    lhs[i] = acc;
  }
}

void orc_cstencil_apply_all(orc_c* lhs, orc_c* rhs, void* s) { orc_cstencil_apply(lhs, rhs, (const orc_cstencil*)s, ORC_PIECE_ALL); }

void orc_cstencil_eoprec_prepare(orc_c* rhs_e, orc_c* rhs_orig, const orc_cstencil* s)
{
  const int dims[2] = {s->lx, s->ly}, n = s->lx * s->ly * s->nc;
  orc_cstencil_apply(rhs_e, rhs_orig, s, ORC_PIECE_EO);
  for (int i = 0; i < n; i++)
    if (orc_index_is_even(dims, s->nc, i)) rhs_e[i] = orc_sub(orc_cmul(s->shift, rhs_orig[i]), rhs_e[i]);
}
void orc_cstencil_m2mdeodoe(orc_c* lhs, orc_c* rhs, const orc_cstencil* s)
{
  const int dims[2] = {s->lx, s->ly}, n = s->lx * s->ly * s->nc;
  orc_c* tmp = (orc_c*)malloc(sizeof(orc_c) * (size_t)n);
  orc_cstencil_apply(tmp, rhs, s, ORC_PIECE_OE);
  orc_cstencil_apply(lhs, tmp, s, ORC_PIECE_EO);
  for (int i = 0; i < n; i++)
    if (orc_index_is_even(dims, s->nc, i)) lhs[i] = orc_sub(orc_cmul(orc_cmul(s->shift, s->shift), rhs[i]), lhs[i]);
  free(tmp);
}
void orc_cstencil_eoprec_reconstruct(orc_c* lhs_full, orc_c* lhs_e, orc_c* rhs_o, const orc_cstencil* s)
{
  const int dims[2] = {s->lx, s->ly}, n = s->lx * s->ly * s->nc;
  const double inv_mass = 1.0 / creal(s->shift);
  orc_cstencil_apply(lhs_full, lhs_e, s, ORC_PIECE_OE);
  for (int i = 0; i < n; i++) {
    if (!orc_index_is_even(dims, s->nc, i))
      lhs_full[i] = orc_rmul(inv_mass, orc_sub(rhs_o[i], lhs_full[i]));
    else
      lhs_full[i] = lhs_e[i];
  }
}

void orc_cstencil_generate(orc_cstencil* s, void (*mv)(orc_c*, orc_c*, void*), void* extra)
{
  const int lx = s->lx, ly = s->ly, nc = s->nc, dims[2] = {lx, ly};
  const int n = lx * ly * nc;
  orc_c* e = (orc_c*)malloc(sizeof(orc_c) * (size_t)n);
  orc_c* col = (orc_c*)malloc(sizeof(orc_c) * (size_t)n);
  const int nterms = s->has_two ? 13 : 5;
  for (int i = 0; i < n; i++) {
    int p[2], a;
    orc_index_to_coord(dims, nc, i, p, &a);
    for (int k = 0; k < n; k++) e[k] = CMPLX(0.0, 0.0);
    e[i] = CMPLX(1.0, 0.0);
    for (int k = 0; k < n; k++) col[k] = CMPLX(0.0, 0.0);
    mv(col, e, extra);
    for (int c = 0; c < nc; c++)
      for (int t = 0; t < nterms; t++) {
        /* the site whose term-t neighbour is p */
        int q[2];
        q[0] = ((p[0] - cst_off[t][0]) % lx + lx) % lx;
        q[1] = ((p[1] - cst_off[t][1]) % ly + ly) % ly;
        const int row = orc_coord_to_index(dims, nc, q, c);
        orc_c* blk = (orc_c*)cst_block(s, t);
        blk[(size_t)a + (size_t)nc * row] = orc_add(blk[(size_t)a + (size_t)nc * row], col[row]);
      }
  }
  free(e);
  free(col);
}

/* ------------------------------------------------------------------------ */
/* Gradient of S = phi^dag (D^dag D)^-1 phi with D = m - K,
 * K psi(x) = 1/2 sum_mu eta_mu(x) [U_mu(x) psi(x+mu) - U_mu(x-mu)^* psi(x-mu)].
 * With chi = (D^dag D)^-1 phi and psi = D chi, dS = -2 Re psi^dag (dD) chi, and
 * dU/dtheta = iU gives
 *   dS/dtheta_mu(x) = -eta_mu(x) Im[ psi(x)^* U_mu(x) chi(x+mu) + psi(x+mu)^* U_mu(x)^* chi(x) ],
 *   dS/dm           = -2 m |chi|^2.
 * Written in the natural order (no reference counterpart to follow). */
void orc_staggered_force(double* dS, const orc_c* chi, const orc_c* psi, const orc_stag_op* op)
{
  const int lx = op->x_fine, ly = op->y_fine;
  const orc_c* u = op->lattice;
  for (int y = 0; y < ly; y++)
    for (int x = 0; x < lx; x++) {
      const int i = y * lx + x;
      const double eta[2] = {1.0, (double)(1 - 2 * (x % 2))};
      const int nb[2] = {y * lx + (x + 1) % lx, ((y + 1) % ly) * lx + x};
      for (int mu = 0; mu < 2; mu++) {
        const orc_c U = u[2 * i + mu];
        const orc_c a = orc_cmul(orc_cmul(orc_conj(psi[i]), U), chi[nb[mu]]);
        const orc_c b = orc_cmul(orc_cmul(orc_conj(psi[nb[mu]]), orc_conj(U)), chi[i]);
        dS[2 * i + mu] = -eta[mu] * (cimag(a) + cimag(b));
      }
    }
}
double orc_staggered_dS_dm(const orc_c* chi, const orc_stag_op* op)
{
  return -2.0 * op->mass * orc_norm2sq_c(chi, op->x_fine * op->y_fine);
}

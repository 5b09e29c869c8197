/*
 * orc_solvers.c -- oracle Krylov solvers for real double and double _Complex.
 * TEST INFRASTRUCTURE ONLY (see orc.h).
 */
#include "orc.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>

/* ---- real double ---- */
#define T double
#define SFX r
#define MVT orc_mv_r
#define DOT orc_dot_r
#define NRM orc_norm2sq_r
#define DIFF orc_diffnorm2sq_r
#define MUL(a, b) ((a) * (b))
#define ADD(a, b) ((a) + (b))
#define SUB(a, b) ((a) - (b))
#define RDIV(x, t) ((x) / (t))
#define TDIV(a, b) ((a) / (b))
#define TRDIV(a, r) ((a) / (r))
#define FROMR(x) (x)
#define NEG(a) (-(a))
#define RMUL(r, v) ((r) * (v))
#define RSUB(t, z) ((t) - (z))
#define ZERO 0.0
#define CONVABS(z) (z) /* generic_cg_m.cpp:148 compares zeta itself, no abs */
#define ABSV(z) fabs(z)
#include "orc_solvers_tmpl.h"
#undef T
#undef SFX
#undef MVT
#undef DOT
#undef NRM
#undef DIFF
#undef MUL
#undef ADD
#undef SUB
#undef RDIV
#undef TDIV
#undef TRDIV
#undef FROMR
#undef NEG
#undef RMUL
#undef RSUB
#undef ZERO
#undef CONVABS
#undef ABSV

/* ---- double _Complex ---- */
#define T orc_c
#define SFX c
#define MVT orc_mv_c
#define DOT orc_dot_c
#define NRM orc_norm2sq_c
#define DIFF orc_diffnorm2sq_c
#define MUL(a, b) orc_cmul((a), (b))
#define ADD(a, b) orc_add((a), (b))
#define SUB(a, b) orc_sub((a), (b))
#define RDIV(x, t) orc_rcdiv((x), (t))
#define TDIV(a, b) orc_cdiv((a), (b))
#define TRDIV(a, r) orc_rdiv((a), (r))
#define FROMR(x) CMPLX((x), 0.0)
#define NEG(a) orc_neg(a)
#define RMUL(r, v) orc_rmul((r), (v))
#define RSUB(t, z) orc_rsub((t), (z))
#define ZERO CMPLX(0.0, 0.0)
#define CONVABS(z) cabs(z)
#define ABSV(z) cabs(z)
#include "orc_solvers_tmpl.h"
#undef SFX
#undef DOT
#undef NRM
#undef DIFF

/* ---- double _Complex, reductions in long double ----
 * The complex solvers again with the dot products and norms accumulated in
 * long double and rounded once (orc_dot_cx ...): every other operation, and
 * its order, is the reference's.  This is the FP64 solver with correctly
 * rounded reductions, the common target of implementations whose summation
 * order differs from the reference's serial one (a GPU's tree reductions):
 * at the 256^2 and 1024^2 staggered solves the serial sums move the
 * reference's own solution by ~5e-12 from it.  Not in the reference. */
#define SFX cx
#define DOT orc_dot_cx
#define NRM orc_norm2sq_cx
#define DIFF orc_diffnorm2sq_cx
#include "orc_solvers_tmpl.h"

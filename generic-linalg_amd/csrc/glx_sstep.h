// glx_sstep.h -- two CG iterations per launch: s-step CG (s = 2, the
// communication-avoiding CG of Carson and Demmel with a monomial basis) on the
// staggered normal operator A = D^dag D = m^2 - K^2 (operators.cpp:444-453),
// and on its even/odd Schur form m^2 - K_eo K_oe over compact even sites
// (operators.cpp:543-616), complex FP32 vectors.  The mixed-precision solver
// runs it between FP64 reliable updates.
//
// Launch `ko` holds the basis of the previous launch in slot ko & 1,
//   Y = [P0, P1, P2, R0, R1] = [p, A p, A^2 p, r, A r],
// and writes the next basis into slot (ko + 1) & 1:
//   prelude  every block sums the previous launch's 12 partial sums (in block
//            order, so all blocks agree) into the Gram matrix G = Y^H Y, and
//            performs two CG iterations on coordinate 5-vectors in FP64:
//              p' = e0, r' = e3, x' = 0, delta = G33
//              twice: w' = B p' (B shifts e0 -> e1 -> e2, e3 -> e4)
//                     alpha = delta / p'^T G w'; x' += alpha p'; r' -= alpha w'
//                     delta' = r'^T G r'; beta = delta' / delta; p' = r' + beta p'
//            A Hermitian gives <A^a u, A^b v> = <u, A^(a+b) v>, so G comes
//            from 5 + 4 + 3 sums (mu_d over p, nu_d over p and r, rho_d over
//            r); the CG scalars of a Hermitian A are real, so the coordinates
//            are real and only the real parts of the sums enter.
//   stream   p = Y c_p, r = Y c_r and x += Y c_x site by site (FP32), then
//            P1 = A p, R1 = A r and P2 = A P1 in one row-marching pass
//   sums     mu_0..4, Re nu_0..3, rho_0..2 of the new basis, per block
// A launch whose iterations stop (the residual fell by delta since the last
// reliable update, or reached the tolerance or zero, or max_iter) forms only x and
// the direction p of the stopping iteration (into P0 of the new slot); the
// reliable update then recomputes r = b - A x in FP64 into R0 of that slot,
// and the next launch restarts from P0 and R0 alone (c_p = e3 + beta e0,
// c_r = e3, c_x = 0) without a CG step.
//
// Pass structure (a block of 4 waves owns a strip of W = 64 CPL columns and
// sr rows; lane l holds columns x0 + CPL l .. x0 + CPL l + CPL - 1; each step
// of the march covers four rows, one per wave, with a lag of one row per
// stage):
//   S0  rows 4k-4 .. 4k-1   stage p, r (from Y) and the links   cols x0-4 .. x0+W+3
//   S1  rows 4k-5 .. 4k-2   qp = K p, qr = K r                    cols x0-3 .. x0+W+2
//   S2  rows 4k-6 .. 4k-3   P1 = m^2 p - K qp                     cols x0-2 .. x0+W+1
//   S3  rows 4k-7 .. 4k-4   q1 = K P1                             cols x0-1 .. x0+W
//   S4  rows 4k-8 .. 4k-5   P2 = m^2 P1 - K q1, R1 = m^2 r - K qr, the sums (own)
// LDS rings of 8 rows hold every stage's rows (slot = (y + 8) & 7); the
// links are staged halved, which takes K's factor 1/2 out of every hop.  S4's
// first row needs U_y of the row above the 8 in the ring (it shares a slot
// with the newest staged row), so S0 saves that row's U_y into a one-row
// buffer before overwriting it.  A strip takes sr/4 + 2 steps.
#pragma once

#include "glx_eo.h"

#include <type_traits>

namespace glx {

constexpr int kSsNR = 12;  // partial sums per block

// Two slots of the five basis vectors and the FP32 solution increment.
struct SsBufs {
  c64* y;  // slot s, vector j at y + (5 (s & 1) + j) n
  c64* x;
  long n;
  __host__ __device__ __forceinline__ c64* v(int slot, int j) const { return y + (5 * (slot & 1) + j) * n; }
};

// The CG scalars of a Hermitian A are real, so the coordinate vectors stay
// real and only Re <Y_i, Y_j> enters: from the sums 0..4 mu_d, 5..8 Re nu_d,
// 9..11 rho_d.
struct SsGram {
  const double* v;
  __device__ __forceinline__ double at(int i, int j) const {
    const bool pi = i < 3, pj = j < 3;
    const int d = (pi ? i : i - 3) + (pj ? j : j - 3);
    if (pi && pj) return v[d];
    if (!pi && !pj) return v[9 + d];
    return v[5 + d];
  }
};
// What the prelude decided for this launch.
struct SsPlan {
  double cp[5], cr[5], cx[5];
  bool stencil;  // false: the iterations stopped; form x and p only
  int k, done, pending, steps;
  double rsq, maxrr;
  double hist[2];
};

// The two iterations on coordinates (v = the sums of the basis in slot
// ko & 1), from iteration k0 with maxrr2 = max |r|^2 since the last reliable
// update.  Written out with the coordinates' known zeros (the iterations of
// the generic form p' = e0, r' = e3, w' = B p', alpha = delta / p'^T G w',
// r' -= alpha w', delta' = r'^T G r', p' = r' + beta p'):
//   iteration 1  p' = e0, w' = e1: alpha0 = rho0 / mu1, r' = e3 - alpha0 e1,
//                delta1 = rho0 - 2 alpha0 Re nu1 + alpha0^2 mu2
//   iteration 2  p' = beta0 e0 - alpha0 e1 + e3, w' = beta0 e1 - alpha0 e2 + e4,
//                r' = -(alpha0 + alpha1 beta0) e1 + alpha1 alpha0 e2 + e3 - alpha1 e4,
//                x' = (alpha0 + alpha1 beta0) e0 - alpha1 alpha0 e1 + alpha1 e3
template <class St>
__device__ __forceinline__ void ss_iterate(const double* v, const St* s, int k0, double maxrr2, SsPlan& pl) {
  const SsGram g{v};
  const double tol = s->eps * s->bsqrt;
  const double dd = s->delta * s->delta;
  auto stops = [&](double dn, int k) {
    return dn < dd * maxrr2 || dn < tol * tol || dn == 0.0 || k == s->max_iter - 1;
  };
  auto stop_with = [&](const double (&cp)[5], const double (&cr)[5], const double (&cx)[5], int k, double dl) {
#pragma unroll
    for (int i = 0; i < 5; i++) {
      pl.cp[i] = cp[i];
      pl.cr[i] = cr[i];
      pl.cx[i] = cx[i];
    }
    pl.stencil = false;
    pl.done = 2;
    pl.pending = 0;
    pl.k = k;
    pl.rsq = dl;
    pl.maxrr = maxrr2;
  };
  pl.stencil = true;
  pl.done = 0;
  pl.pending = 1;
  // iteration k0
  const double dl0 = g.at(3, 3);
  const double pw0 = g.at(0, 1);
  const double a0 = pw0 != 0.0 ? dl0 / pw0 : 0.0;
  double dn1 = dl0 - 2.0 * a0 * g.at(1, 3) + a0 * a0 * g.at(1, 1);
  if (!(dn1 > 0.0)) dn1 = 0.0;
  pl.hist[0] = dn1;
  pl.steps = 1;
  if (dn1 > maxrr2) maxrr2 = dn1;
  if (stops(dn1, k0)) {
    const double cp[5] = {1.0, 0.0, 0.0, 0.0, 0.0}, cr[5] = {0.0, -a0, 0.0, 1.0, 0.0}, cx[5] = {a0, 0.0, 0.0, 0.0, 0.0};
    stop_with(cp, cr, cx, k0, dl0);
    return;
  }
  // iteration k0 + 1
  const double b0 = dn1 / dl0;
  const double P[3] = {b0, -a0, 1.0};  // p' on e0, e1, e3
  const double W[3] = {b0, -a0, 1.0};  // w' on e1, e2, e4
  constexpr int PI[3] = {0, 1, 3}, WI[3] = {1, 2, 4};
  double pw1 = 0.0;
#pragma unroll
  for (int i = 0; i < 3; i++) {
    double t = 0.0;
#pragma unroll
    for (int j = 0; j < 3; j++) t += g.at(PI[i], WI[j]) * W[j];
    pw1 += P[i] * t;
  }
  const double a1 = pw1 != 0.0 ? dn1 / pw1 : 0.0;
  const double r1 = -a0 - a1 * b0, r2 = a1 * a0, r4 = -a1;  // r' on e1, e2, e3 (= 1), e4
  const double R[4] = {r1, r2, 1.0, r4};
  constexpr int RI[4] = {1, 2, 3, 4};
  double dn2 = 0.0;
#pragma unroll
  for (int i = 0; i < 4; i++) {
    double t = 0.0;
#pragma unroll
    for (int j = 0; j < 4; j++) t += g.at(RI[i], RI[j]) * R[j];
    dn2 += R[i] * t;
  }
  if (!(dn2 > 0.0)) dn2 = 0.0;
  pl.hist[1] = dn2;
  pl.steps = 2;
  if (dn2 > maxrr2) maxrr2 = dn2;
  const double cr[5] = {0.0, r1, r2, 1.0, r4};
  const double cx[5] = {a0 + a1 * b0, -a1 * a0, 0.0, a1, 0.0};
  const double p1[5] = {b0, -a0, 0.0, 1.0, 0.0};
  if (stops(dn2, k0 + 1)) {
    stop_with(p1, cr, cx, k0 + 1, dn1);
    return;
  }
  const double b1 = dn2 / dn1;
#pragma unroll
  for (int i = 0; i < 5; i++) {
    pl.cp[i] = cr[i] + b1 * p1[i];
    pl.cr[i] = cr[i];
    pl.cx[i] = cx[i];
  }
  pl.k = k0 + 2;
  pl.rsq = dn2;
  pl.maxrr = maxrr2;
}

// The fields a launch changes, carried from the state it read to the one it
// writes (the constants were copied once when the solve started).
template <class St>
__device__ __forceinline__ void ss_carry(const St* rd, St* wr) {
  wr->k = rd->k;
  wr->ko = rd->ko;
  wr->done = rd->done;
  wr->kfinal = rd->kfinal;
  wr->pending = rd->pending;
  wr->rsq = rd->rsq;
  wr->rsqNew = rd->rsqNew;
  wr->maxrr = rd->maxrr;
  wr->beta_r = rd->beta_r;
  wr->alpha = rd->alpha;
}

// acc + Re(conj(a) b) in FP64 (fused multiply-adds)
__device__ __forceinline__ double ss_acc(double acc, c64 a, c64 b) {
  return fma(double(a.im), double(b.im), fma(double(a.re), double(b.re), acc));
}
// K at one point as khop, with the links already halved in LDS (a power-of-two
// scale, so the result is khop's bit for bit)
template <class T, class C>
__device__ __forceinline__ T khop_h(T fxp, T fxm, T fyp, T fym, C ux, C uxm, C uy, C uym, bool eta_neg) {
  T ax = umul(ux, fxp) - ucmul(uxm, fxm);
  T ay = umul(uy, fyp) - ucmul(uym, fym);
  if (eta_neg) ay = -ay;
  return ax + ay;
}
// sum_j c_j y_j over the basis (only P0 and R0 unless `full`)
__device__ __forceinline__ c64 ss_comb(const float (&c)[5], const c64 (&y)[5], bool full) {
  c64 s = c[0] * y[0] + c[3] * y[3];
  if (full) s = s + c[1] * y[1] + c[2] * y[2] + c[4] * y[4];
  return s;
}
// The stream's three combinations with the coefficients' fixed structure
// (every plan that runs the stencil has c_r = e3 + r1 e1 + r2 e2 + r4 e4 and
// c_x on e0, e1, e3 only; a restart has c_r = e3, c_x = 0 and zero weights
// on the vectors it did not write, which are finite leftovers):
//   p = sum_j cp_j y_j,  r = y3 + cr1 y1 + cr2 y2 + cr4 y4,
//   dx = cx0 y0 + cx1 y1 + cx3 y3
__device__ __forceinline__ c64 ss_comb_p(const float (&c)[5], const c64 (&y)[5]) {
  return c[0] * y[0] + c[3] * y[3] + c[1] * y[1] + c[2] * y[2] + c[4] * y[4];
}
__device__ __forceinline__ c64 ss_comb_r(const float (&c)[5], const c64 (&y)[5]) {
  return y[3] + c[1] * y[1] + c[2] * y[2] + c[4] * y[4];
}
__device__ __forceinline__ c64 ss_comb_x(const float (&c)[5], const c64 (&y)[5]) {
  return c[0] * y[0] + c[3] * y[3] + c[1] * y[1];
}

// The stencil steps of the march on the LDS rings (row slots s0 = y,
// sm = y - 1, sp = y + 1; column indices as in the kernel's arrays).
// qp, qr at q column qc (x = x0 - 3 + qc; p and link column qc + 1)
template <class T, class C, int WP, int WL, int WQ>
__device__ __forceinline__ void ss_q1hop(const T (*sP)[WP], const T (*sR)[WP], const C (*sLx)[WL],
                                         const C (*sLy)[WL], T (*sQp)[WQ], T (*sQr)[WQ], int s0, int sm, int sp,
                                         int qc) {
  const int pc = qc + 1;
  const bool eta_neg = ((qc + 1) & 1) != 0;
  const C ux = sLx[s0][pc], uxm = sLx[s0][pc - 1], uy = sLy[s0][pc], uym = sLy[sm][pc];
  sQp[s0][qc] = khop_h<T, C>(sP[s0][pc + 1], sP[s0][pc - 1], sP[sp][pc], sP[sm][pc], ux, uxm, uy, uym, eta_neg);
  sQr[s0][qc] = khop_h<T, C>(sR[s0][pc + 1], sR[s0][pc - 1], sR[sp][pc], sR[sm][pc], ux, uxm, uy, uym, eta_neg);
}
// P1 = m^2 p - K qp at P1 column c1 (x = x0 - 2 + c1; q column c1 + 1, p and
// link column c1 + 2)
template <class T, class C, int WP, int WL, int WQ>
__device__ __forceinline__ T ss_p1hop(const T (*sP)[WP], const C (*sLx)[WL], const C (*sLy)[WL],
                                      const T (*sQp)[WQ], int s0, int sm, int sp, int c1, float m2) {
  const int qc = c1 + 1, pc = c1 + 2;
  const T kq = khop_h<T, C>(sQp[s0][qc + 1], sQp[s0][qc - 1], sQp[sp][qc], sQp[sm][qc], sLx[s0][pc], sLx[s0][pc - 1],
                          sLy[s0][pc], sLy[sm][pc], (c1 & 1) != 0);
  return rscal(m2, sP[s0][pc]) - kq;
}
// q1 = K P1 at q1 column c2 (x = x0 - 1 + c2; P1 column c2 + 1, link column
// c2 + 3)
template <class T, class C, int W1, int WL, int W2>
__device__ __forceinline__ void ss_q2hop(const T (*sP1)[W1], const C (*sLx)[WL], const C (*sLy)[WL], T (*sQ1)[W2],
                                         int s0, int sm, int sp, int c2) {
  const int lc = c2 + 3;
  sQ1[s0][c2] = khop_h<T, C>(sP1[s0][c2 + 2], sP1[s0][c2], sP1[sp][c2 + 1], sP1[sm][c2 + 1], sLx[s0][lc],
                           sLx[s0][lc - 1], sLy[s0][lc], sLy[sm][lc], (c2 & 1) == 0);
}

// The same steps for the even/odd Schur operator m^2 - K_eo K_oe on compact
// even-site vectors (glx_eo.h): p, r, P1 live on even sites, qp, qr, q1 on
// odd sites, in compact columns; row y has parity e = y & 1 (y0 even), odd
// site c sits at full column x = 2c + 1 - e with even x-neighbours c - e and
// c + 1 - e, even site c at x = 2c + e with odd x-neighbours c + e - 1 and
// c + e; y-neighbours share the compact column.  The links are staged by full
// column, LDS link column = x - (2 x0 - 7).
// qp, qr at odd column qc (compact c0 - 3 + qc)
template <class T, class C, int WP, int WL, int WQ>
__device__ __forceinline__ void ss_q1hop_eo(const T (*sP)[WP], const T (*sR)[WP], const C (*sLx)[WL],
                                            const C (*sLy)[WL], T (*sQp)[WQ], T (*sQr)[WQ], int s0, int sm, int sp,
                                            int qc, int e) {
  const int li = 2 * qc + 2 - e;
  const bool eta_neg = e == 0;
  const C ux = sLx[s0][li], uxm = sLx[s0][li - 1], uy = sLy[s0][li], uym = sLy[sm][li];
  sQp[s0][qc] = khop_h<T, C>(sP[s0][qc + 2 - e], sP[s0][qc + 1 - e], sP[sp][qc + 1], sP[sm][qc + 1], ux, uxm, uy,
                             uym, eta_neg);
  sQr[s0][qc] = khop_h<T, C>(sR[s0][qc + 2 - e], sR[s0][qc + 1 - e], sR[sp][qc + 1], sR[sm][qc + 1], ux, uxm, uy,
                             uym, eta_neg);
}
// P1 = m^2 p - K_eo qp at even column c1 (compact c0 - 2 + c1; p column c1 + 2)
template <class T, class C, int WP, int WL, int WQ>
__device__ __forceinline__ T ss_p1hop_eo(const T (*sP)[WP], const C (*sLx)[WL], const C (*sLy)[WL],
                                         const T (*sQp)[WQ], int s0, int sm, int sp, int c1, int e, float m2) {
  const int li = 2 * c1 + e + 3;
  const T kq = khop_h<T, C>(sQp[s0][c1 + e + 1], sQp[s0][c1 + e], sQp[sp][c1 + 1], sQp[sm][c1 + 1], sLx[s0][li],
                            sLx[s0][li - 1], sLy[s0][li], sLy[sm][li], e != 0);
  return rscal(m2, sP[s0][c1 + 2]) - kq;
}
// q1 = K_oe P1 at odd column c2 (compact c0 - 1 + c2)
template <class T, class C, int W1, int WL, int W2>
__device__ __forceinline__ void ss_q2hop_eo(const T (*sP1)[W1], const C (*sLx)[WL], const C (*sLy)[WL],
                                            T (*sQ1)[W2], int s0, int sm, int sp, int c2, int e) {
  const int li = 2 * c2 + 6 - e;
  sQ1[s0][c2] = khop_h<T, C>(sP1[s0][c2 + 2 - e], sP1[s0][c2 + 1 - e], sP1[sp][c2 + 1], sP1[sm][c2 + 1], sLx[s0][li],
                             sLx[s0][li - 1], sLy[s0][li], sLy[sm][li], e == 0);
}

// Accesses at a 32-bit byte offset from a wave-uniform base: the compiler
// addresses them as scalar base + vector offset instead of forming 64-bit
// addresses per lane.  The host launches the kernel only where every array it
// touches stays below 4 GiB (kSsMaxSites), so the offsets never wrap.
constexpr long kSsMaxSites = 1L << 28;  // links: 16 bytes per site (FP32)
template <class T>
__device__ __forceinline__ const T& ss_at(const void* base, unsigned off) {
  return *reinterpret_cast<const T*>(static_cast<const char*>(base) + off);
}
template <class T>
__device__ __forceinline__ T& ss_at(void* base, unsigned off) {
  return *reinterpret_cast<T*>(static_cast<char*>(base) + off);
}
// A lattice row within 4 of the strip's own rows, -4 <= y < ly + 4 with
// ly >= 8: one conditional add or subtract wraps it (wrap_coord's general
// remainder is never needed here).
__device__ __forceinline__ int ss_wrap_row(int y, int ly) { return y < 0 ? y + ly : (y >= ly ? y - ly : y); }
// The links of sites s .. s + N - 1 (N = 1 or 2, s even when N = 2).
template <int N>
__device__ __forceinline__ void ss_links(const LinksF32& lk, unsigned s, LinksF32::Raw (&r)[N]) {
#pragma unroll
  for (int c = 0; c < N; c++) r[c] = ss_at<LinksF32::Raw>(lk.u, (s + c) * 16u);
}
template <int N>
__device__ __forceinline__ void ss_links(const LinksF16& lk, unsigned s, LinksF16::Raw (&r)[N]) {
  if constexpr (N == 2) {
    const Pair2<HalfLink> h = ss_at<Pair2<HalfLink>>(lk.u, s * 8u);
    const Pair2<float> c = ss_at<Pair2<float>>(lk.scale, s * 4u);
    r[0] = LinksF16::Raw{h.a, c.a};
    r[1] = LinksF16::Raw{h.b, c.b};
  } else {
    r[0] = LinksF16::Raw{ss_at<HalfLink>(lk.u, s * 8u), ss_at<float>(lk.scale, s * 4u)};
  }
}

// EO = false: D^dag D on full-lattice vectors (eg.lx = eg.g.lx); EO = true:
// m^2 - K_eo K_oe on compact even-site vectors (eg.g over lx / 2 columns,
// eg.lx the full width for the links).
template <int CPL, bool EO, class LK, class St>
__global__ __launch_bounds__(kBlock) void k_sscg(EoGeom eg, OpDev<LK> op, SsBufs bf, const St* rd, St* wr, int q,
                                                 RedWs ws) {
  using T = c64;
  using C = typename LK::C;
  const NGeom& g = eg.g;
  constexpr int NXB = kNX * CPL;  // (compact) columns per strip
  constexpr int WP = NXB + 8;     // p, r: columns x0-4 .. x0+NXB+3
  // links: columns x0-4 .. x0+NXB+2 (last slot unused); EO: full columns
  // 2 x0 - 7 .. 2 x0 + 2 NXB + 5 (last slot unused)
  constexpr int WL = EO ? 2 * NXB + 14 : NXB + 8;
  // qp, qr: x0-3 .. x0+NXB+2; P1: x0-2 .. x0+NXB+1; q1: x0-1 .. x0+NXB.  Their
  // rows are padded to the width of p's, so that the rings of all vectors
  // share one row stride: one vector address per ring slot then serves every
  // ring (the rings' own offsets become immediates of the LDS instructions)
  constexpr int WQ = WP, W1 = WP, W2 = WP;
  static_assert(std::is_same<T, C>::value, "vector and link rings share one array");
  // every ring in one array (rings of the vectors first, then the links'):
  // one address per ring slot and lane, the rings at immediate offsets
  constexpr int kVR = 6, kLR = 2;
  constexpr int kRingT = kVR * 8 * WP + kLR * 8 * WL;
  __shared__ T sRing[kRingT];
  T (*const sP)[WP] = reinterpret_cast<T (*)[WP]>(sRing + 0 * 8 * WP);
  T (*const sR)[WP] = reinterpret_cast<T (*)[WP]>(sRing + 1 * 8 * WP);
  T (*const sQp)[WQ] = reinterpret_cast<T (*)[WQ]>(sRing + 2 * 8 * WP);
  T (*const sQr)[WQ] = reinterpret_cast<T (*)[WQ]>(sRing + 3 * 8 * WP);
  T (*const sP1)[W1] = reinterpret_cast<T (*)[W1]>(sRing + 4 * 8 * WP);
  T (*const sQ1)[W2] = reinterpret_cast<T (*)[W2]>(sRing + 5 * 8 * WP);
  C (*const sLx)[WL] = reinterpret_cast<C (*)[WL]>(sRing + kVR * 8 * WP);
  C (*const sLy)[WL] = reinterpret_cast<C (*)[WL]>(sRing + kVR * 8 * WP + 8 * WL);
  __shared__ C sKeep[NXB];

  // a finished solve (or a pending reliable update): carry the state over
  if (rd->done != 0) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      ss_carry(rd, wr);
      if (rd->hdone) *rd->hdone = rd->done;
    }
    return;
  }
  const int pending = rd->pending;
  const bool full = pending != 0;  // the read slot holds all five vectors
  const int k0 = rd->k, ko = rd->ko;
  const int nblk = gridDim.x;

  // the wave index is uniform across the wave: in a scalar register, every
  // row, ring slot and row-range test below is scalar work
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), l = threadIdx.x & 63;
  const int tile = xcd_tile(blockIdx.x, g.nblk);
  const bool active = tile < g.ntiles;
  const int tt = active ? tile : 0;  // grid padding only joins the sums
  const int lx = g.lx, ly = g.ly;  // lx: vector row length (compact for EO)
  const int lxf = eg.lx;            // link row length
  const int x0 = (tt % g.ntx) * NXB;
  const int y0 = (tt / g.ntx) * g.sr;
  const int rows = min(g.sr, ly - y0);  // a multiple of 4
  const int nsteps = rows / 4 + 2;
  const int xm = x0 + CPL * l;
  // Odd bands march upwards: then the two strips either side of a band
  // boundary stage their shared rows (one's own rows, the other's halo) at
  // the same time, at the start or at the end of both marches, and the
  // second read hits L2.  March row yr is lattice row y0 + rrow(yr); the
  // rings are kept in march rows, so a stage's lattice neighbours y +- 1 are
  // the ring slots slp / slm.
  const bool up = ((tt / g.ntx) & 1) != 0;
  auto rrow = [&](int yr) { return up ? rows - 1 - yr : yr; };
  auto slm = [&](int yr) { return up ? (yr + 9) & 7 : (yr + 7) & 7; };  // row y - 1
  auto slp = [&](int yr) { return up ? (yr + 7) & 7 : (yr + 9) & 7; };  // row y + 1
  auto par = [&](int yr) { return rrow(yr) & 1; };                      // EO: parity (y0 even)
  const int so = ko & 1, sn = so ^ 1;
  const c64* Y[5] = {bf.v(so, 0), bf.v(so, 1), bf.v(so, 2), bf.v(so, 3), bf.v(so, 4)};
  c64* N[5] = {bf.v(sn, 0), bf.v(sn, 1), bf.v(sn, 2), bf.v(sn, 3), bf.v(sn, 4)};
  // the CPL sites of a lane as one access (16 bytes for CPL = 2)
  auto ldc = [](const c64* p, int j, c64 (&o)[CPL]) {
    if constexpr (CPL == 2) {
      const Pair2<c64> t = ss_at<Pair2<c64>>(p, unsigned(j) * 8u);
      o[0] = t.a;
      o[1] = t.b;
    } else {
      o[0] = ss_at<c64>(p, unsigned(j) * 8u);
    }
  };
  auto stc = [](c64* p, int j, const c64 (&v)[CPL]) {
    if constexpr (CPL == 2) ss_at<Pair2<c64>>(p, unsigned(j) * 8u) = Pair2<c64>{v[0], v[1]};
    else ss_at<c64>(p, unsigned(j) * 8u) = v[0];
  };

  // the previous launch's sums
  constexpr int kPreU = 2;
  double pv[kPreU][kSsNR];
  const double* PS = ws.partials + (q ^ 1) * kSsNR * nblk;
  if (full) {
#pragma unroll
    for (int u = 0; u < kPreU; u++) {
      const int bb = threadIdx.x + u * kBlock;
#pragma unroll
      for (int k = 0; k < kSsNR; k++) pv[u][k] = bb < nblk ? PS[k * nblk + bb] : 0.0;
    }
  }

  // halo columns, all four rows of a step on one wave each (a pass with a
  // few active lanes costs a wave its whole instruction stream):
  //   S0  wave 3 lanes 0..31: p, r on columns x0-4 .. x0-1, x0+NXB .. x0+NXB+3
  //       wave 1 lanes 0..27: the links on x0-4 .. x0-1, x0+NXB .. x0+NXB+2
  //       (EO lanes 0..51: full columns 2x0-7 .. 2x0-1, 2x0+2NXB .. 2x0+2NXB+5)
  //   S1  wave 0 lanes 0..23: qp, qr on x0-3 .. x0-1, x0+NXB .. x0+NXB+2
  //   S2  wave 1 lanes 0..15: P1 on x0-2, x0-1, x0+NXB, x0+NXB+1
  //   S3  wave 2 lanes 0..7:  q1 on x0-1, x0+NXB
  const bool hp = w == 3 && l < 32;
  constexpr int kHL = EO ? 13 : 7;  // link halo columns per row
  const bool hl = w == 1 && l < 4 * kHL;
  const int hp_row = l >> 3, hp_c = (l & 7) < 4 ? (l & 7) : NXB + (l & 7);
  const int hl_row = l / kHL, hl_j = l % kHL;
  const int hl_c = EO ? (hl_j < 7 ? hl_j : 2 * NXB + hl_j) : (hl_j < 4 ? hl_j : NXB + hl_j);
  const int xhp = wrap_coord(x0 - 4 + hp_c, lx);
  const int xhl = EO ? wrap_coord(2 * x0 - 7 + hl_c, lxf) : wrap_coord(x0 - 4 + hl_c, lx);

  c64 ry[5][CPL], rx[CPL];
  typename LK::Raw lmr[EO ? 2 * CPL : CPL];
  c64 hy[5];
  typename LK::Raw lh;
  // all five basis vectors and x are read in every launch: a restart (not
  // `full`) weights the three it did not write by zero (ss_comb_*), which
  // keeps one code path for both
  auto fetch = [&](int k) {
    const int yr = 4 * k - 4 + w;
    const int j = ss_wrap_row(y0 + rrow(yr), ly) * lx + xm;
    const bool own = yr >= 0 && yr < rows;
#pragma unroll
    for (int i = 0; i < 5; i++) ldc(Y[i], j, ry[i]);
    if (own) ldc(bf.x, j, rx);
    if constexpr (EO) ss_links(op.links, unsigned(ss_wrap_row(y0 + rrow(yr), ly) * lxf + 2 * xm), lmr);
    else ss_links(op.links, unsigned(j), lmr);
    if (hp) {
      const int jh = ss_wrap_row(y0 + rrow(4 * k - 4 + hp_row), ly) * lx + xhp;
#pragma unroll
      for (int i = 0; i < 5; i++) hy[i] = ss_at<c64>(Y[i], unsigned(jh) * 8u);
    }
    if (hl) {
      typename LK::Raw t[1];
      ss_links(op.links, unsigned(ss_wrap_row(y0 + rrow(4 * k - 4 + hl_row), ly) * lxf + xhl), t);
      lh = t[0];
    }
  };

  // prelude: the coefficients of this launch.  The first rows' loads are
  // issued after the block sum of the partials: the wait for the partials is
  // a wait for every load issued before it (the loads sit under branches, so
  // the count of younger loads is not known), and the coefficient algebra
  // then runs while the first rows are in flight.
  __shared__ double sm[kSsNR][kWaves];
  SsPlan pl;
  if (full) {
    double acc[kSsNR];
#pragma unroll
    for (int k = 0; k < kSsNR; k++) acc[k] = pv[0][k];
#pragma unroll
    for (int u = 1; u < kPreU; u++) {
#pragma unroll
      for (int k = 0; k < kSsNR; k++) acc[k] += pv[u][k];
    }
    for (int b0 = threadIdx.x + kBlock * kPreU; b0 < nblk; b0 += kBlock) {
#pragma unroll
      for (int k = 0; k < kSsNR; k++) acc[k] += PS[k * nblk + b0];
    }
    block_sum_dpp<kSsNR>(acc, sm);
    if (active) fetch(0);
    ss_iterate(acc, rd, k0, rd->maxrr, pl);
  } else {
    if (active) fetch(0);
    // restart from P0 (the last direction) and R0 (the true residual)
#pragma unroll
    for (int i = 0; i < 5; i++) {
      pl.cp[i] = i == 3 ? 1.0 : (i == 0 ? rd->beta_r : 0.0);
      pl.cr[i] = i == 3 ? 1.0 : 0.0;
      pl.cx[i] = 0.0;
    }
    pl.stencil = true;
    pl.done = 0;
    pl.pending = 1;
    pl.steps = 0;
    pl.k = k0;
    pl.rsq = rd->rsq;
    pl.maxrr = rd->maxrr;
  }
  float cp[5], cr[5], cx[5];
#pragma unroll
  for (int i = 0; i < 5; i++) {
    cp[i] = float(pl.cp[i]);
    cr[i] = float(pl.cr[i]);
    cx[i] = float(pl.cx[i]);
  }

  if (!pl.stencil) {
    // stopped: x += Y c_x and P0' = Y c_p on the owned sites (a stop follows
    // a launch that wrote all five vectors)
    if (active) {
      for (int yr = w; yr < rows; yr += 4) {
        const int j = (y0 + yr) * lx + xm;
        c64 a[5][CPL], xo[CPL], pn[CPL];
#pragma unroll
        for (int i = 0; i < 5; i++) ldc(Y[i], j, a[i]);
        ldc(bf.x, j, xo);
#pragma unroll
        for (int c = 0; c < CPL; c++) {
          const c64 yv[5] = {a[0][c], a[1][c], a[2][c], a[3][c], a[4][c]};
          pn[c] = ss_comb(cp, yv, true);
          xo[c] = xo[c] + ss_comb(cx, yv, true);
        }
        stc(N[0], j, pn);
        stc(bf.x, j, xo);
      }
    }
  } else {
    const float m2 = float(op.mass * op.mass);
    double part[kSsNR];
#pragma unroll
    for (int k = 0; k < kSsNR; k++) part[k] = 0.0;
    for (int k = 0; k < (active ? nsteps : 0); k++) {
      if (k > 0) __syncthreads();  // step k-1's readers are done with the slots reused below
      {
        // S0: rows 4k-4 .. 4k-1
        const int yr = 4 * k - 4 + w;
        const int j = ss_wrap_row(y0 + rrow(yr), ly) * lx + xm;
        const bool own = yr >= 0 && yr < rows;
        const int sl = (yr + 8) & 7;
        c64 pv2[CPL], rv2[CPL], xv2[CPL];
#pragma unroll
        for (int c = 0; c < CPL; c++) {
          const c64 yv[5] = {ry[0][c], ry[1][c], ry[2][c], ry[3][c], ry[4][c]};
          pv2[c] = ss_comb_p(cp, yv);
          rv2[c] = ss_comb_r(cr, yv);
          sP[sl][CPL * l + c + 4] = pv2[c];
          sR[sl][CPL * l + c + 4] = rv2[c];
        }
        if (own) {
#pragma unroll
          for (int c = 0; c < CPL; c++) {
            const c64 yv[5] = {ry[0][c], ry[1][c], ry[2][c], ry[3][c], ry[4][c]};
            xv2[c] = rx[c] + ss_comb_x(cx, yv);
          }
          stc(N[0], j, pv2);
          stc(N[3], j, rv2);
          stc(bf.x, j, xv2);
        }
        // wave 3 stages row 4k-1 into the slot of row 4k-9, whose U_y the
        // first S4 row of this step still needs (for EO: at the full columns
        // of that even row's sites, link columns 2b + 7)
        if constexpr (EO) {
#pragma unroll
          for (int t = 0; t < 2 * CPL; t++) {
            C ux, uy;
            LK::finish(lmr[t], ux, uy);
            ux = 0.5f * ux;
            uy = 0.5f * uy;
            const int li = 2 * CPL * l + t + 7;
            if (w == 3 && (t & 1) == 0) sKeep[CPL * l + t / 2] = sLy[sl][li];
            sLx[sl][li] = ux;
            sLy[sl][li] = uy;
          }
        } else {
#pragma unroll
          for (int c = 0; c < CPL; c++) {
            C ux, uy;
            LK::finish(lmr[c], ux, uy);
            ux = 0.5f * ux;
            uy = 0.5f * uy;
            if (w == 3) sKeep[CPL * l + c] = sLy[sl][CPL * l + c + 4];
            sLx[sl][CPL * l + c + 4] = ux;
            sLy[sl][CPL * l + c + 4] = uy;
          }
        }
        if (hp) {
          const int hs = (4 * k - 4 + hp_row + 8) & 7;
          sP[hs][hp_c] = ss_comb_p(cp, hy);
          sR[hs][hp_c] = ss_comb_r(cr, hy);
        }
        if (hl) {
          const int hs = (4 * k - 4 + hl_row + 8) & 7;
          C hx, hyy;
          LK::finish(lh, hx, hyy);
          hx = 0.5f * hx;
          hyy = 0.5f * hyy;
          sLx[hs][hl_c] = hx;
          sLy[hs][hl_c] = hyy;
        }
      }
      if (k + 1 < nsteps) fetch(k + 1);
      __syncthreads();
      {
        // S1: qp = K p, qr = K r on rows 4k-5 .. 4k-2 (needed: -3 .. rows+2)
        const int y = 4 * k - 5 + w;
        if (y >= -3 && y <= rows + 2) {
          const int s0 = (y + 8) & 7, sm1 = slm(y), sp1 = slp(y);
#pragma unroll
          for (int c = 0; c < CPL; c++) {
            if constexpr (EO) ss_q1hop_eo<T, C>(sP, sR, sLx, sLy, sQp, sQr, s0, sm1, sp1, CPL * l + c + 3, par(y));
            else ss_q1hop<T, C>(sP, sR, sLx, sLy, sQp, sQr, s0, sm1, sp1, CPL * l + c + 3);
          }
        }
        if (w == 0 && l < 24) {
          const int yh = 4 * k - 5 + l / 6, hc = l % 6;
          if (yh >= -3 && yh <= rows + 2) {
            const int qc = hc < 3 ? hc : NXB + hc;
            if constexpr (EO)
              ss_q1hop_eo<T, C>(sP, sR, sLx, sLy, sQp, sQr, (yh + 8) & 7, slm(yh), slp(yh), qc, par(yh));
            else ss_q1hop<T, C>(sP, sR, sLx, sLy, sQp, sQr, (yh + 8) & 7, slm(yh), slp(yh), qc);
          }
        }
      }
      __syncthreads();
      {
        // S2: P1 = m^2 p - K qp on rows -2 .. rows+1
        const int y = 4 * k - 6 + w;
        if (y >= -2 && y <= rows + 1) {
          const int s0 = (y + 8) & 7, sm1 = slm(y), sp1 = slp(y);
          T v1[CPL];
#pragma unroll
          for (int c = 0; c < CPL; c++) {
            const int c1 = CPL * l + c + 2;
            if constexpr (EO) v1[c] = ss_p1hop_eo<T, C>(sP, sLx, sLy, sQp, s0, sm1, sp1, c1, par(y), m2);
            else v1[c] = ss_p1hop<T, C>(sP, sLx, sLy, sQp, s0, sm1, sp1, c1, m2);
            sP1[s0][c1] = v1[c];
          }
          if (y >= 0 && y < rows) stc(N[1], (y0 + rrow(y)) * lx + xm, v1);
        }
        if (w == 1 && l < 16) {
          const int yh = 4 * k - 6 + (l >> 2), hc = l & 3;
          if (yh >= -2 && yh <= rows + 1) {
            const int c1 = hc < 2 ? hc : NXB + hc;
            const int h0 = (yh + 8) & 7;
            if constexpr (EO)
              sP1[h0][c1] = ss_p1hop_eo<T, C>(sP, sLx, sLy, sQp, h0, slm(yh), slp(yh), c1, par(yh), m2);
            else sP1[h0][c1] = ss_p1hop<T, C>(sP, sLx, sLy, sQp, h0, slm(yh), slp(yh), c1, m2);
          }
        }
      }
      __syncthreads();
      {
        // S3: q1 = K P1 on rows -1 .. rows
        const int y = 4 * k - 7 + w;
        if (y >= -1 && y <= rows) {
          const int s0 = (y + 8) & 7, sm1 = slm(y), sp1 = slp(y);
#pragma unroll
          for (int c = 0; c < CPL; c++) {
            if constexpr (EO) ss_q2hop_eo<T, C>(sP1, sLx, sLy, sQ1, s0, sm1, sp1, CPL * l + c + 1, par(y));
            else ss_q2hop<T, C>(sP1, sLx, sLy, sQ1, s0, sm1, sp1, CPL * l + c + 1);
          }
        }
        if (w == 2 && l < 8) {
          const int yh = 4 * k - 7 + (l >> 1);
          if (yh >= -1 && yh <= rows) {
            const int c2 = (l & 1) == 0 ? 0 : NXB + 1;
            if constexpr (EO)
              ss_q2hop_eo<T, C>(sP1, sLx, sLy, sQ1, (yh + 8) & 7, slm(yh), slp(yh), c2, par(yh));
            else ss_q2hop<T, C>(sP1, sLx, sLy, sQ1, (yh + 8) & 7, slm(yh), slp(yh), c2);
          }
        }
      }
      __syncthreads();
      {
        // S4 on the own rows: P2 = m^2 P1 - K q1, R1 = m^2 r - K qr (qr's
        // ring still holds rows y-1 .. y+1), and the sums
        const int y = 4 * k - 8 + w;
        if (y >= 0 && y < rows) {
          const int s0 = (y + 8) & 7, sm1 = slm(y), sp1 = slp(y);
          T v2[CPL], v4[CPL];
#pragma unroll
          for (int c = 0; c < CPL; c++) {
            const int b = CPL * l + c;
            const int c1 = b + 2, pc = b + 4;
            // full lattice: q1 column b + 1 +- 1, qr column b + 3 +- 1, link
            // column b + 4; EO (even site, x = 2 (x0 + b) + e): q1 columns
            // b + e, b + e + 1 (y: b + 1), qr b + e + 2, b + e + 3 (y: b + 3),
            // link column 2b + e + 7
            const int e = EO ? par(y) : 0;
            const int q1m = EO ? b + e : b, q1p = EO ? b + e + 1 : b + 2;
            const int qrm = EO ? b + e + 2 : b + 2, qrp = EO ? b + e + 3 : b + 4;
            const int lc = EO ? 2 * b + e + 7 : b + 4;
            const bool eta_neg = EO ? e != 0 : (b & 1) != 0;
            const C ux = sLx[s0][lc], uxm = sLx[s0][lc - 1], uy = sLy[s0][lc];
            // row y - 1 of wave 0's row in a downward march shares a slot with
            // the newest staged row (saved in sKeep); upwards it is newer
            const C uym = (w == 0 && !up) ? sKeep[b] : sLy[sm1][lc];
            const T kq = khop_h<T, C>(sQ1[s0][q1p], sQ1[s0][q1m], sQ1[sp1][b + 1], sQ1[sm1][b + 1], ux, uxm, uy, uym,
                                      eta_neg);
            const T kr = khop_h<T, C>(sQr[s0][qrp], sQr[s0][qrm], sQr[sp1][b + 3], sQr[sm1][b + 3], ux, uxm, uy, uym,
                                      eta_neg);
            const T pP1 = sP1[s0][c1];
            const T pp = sP[s0][pc], rr = sR[s0][pc];
            const T p2 = rscal(m2, pP1) - kq;
            const T r1 = rscal(m2, rr) - kr;
            v2[c] = p2;
            v4[c] = r1;
            part[0] = ss_acc(part[0], pp, pp);
            part[1] = ss_acc(part[1], pp, pP1);
            part[2] = ss_acc(part[2], pP1, pP1);
            part[3] = ss_acc(part[3], pP1, p2);
            part[4] = ss_acc(part[4], p2, p2);
            part[5] = ss_acc(part[5], pp, rr);
            part[6] = ss_acc(part[6], pp, r1);
            part[7] = ss_acc(part[7], pP1, r1);
            part[8] = ss_acc(part[8], p2, r1);
            part[9] = ss_acc(part[9], rr, rr);
            part[10] = ss_acc(part[10], rr, r1);
            part[11] = ss_acc(part[11], r1, r1);
          }
          const int j = (y0 + rrow(y)) * lx + xm;
          stc(N[2], j, v2);
          stc(N[4], j, v4);
        }
      }
    }
    block_sum_dpp<kSsNR>(part, sm);
    if (threadIdx.x == 0) {
      double* out = ws.partials + q * kSsNR * nblk;
#pragma unroll
      for (int k = 0; k < kSsNR; k++) out[k * nblk + blockIdx.x] = part[k];
    }
  }

  // block 0 writes the state of the next launch
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    ss_carry(rd, wr);
    wr->k = pl.k;
    wr->ko = ko + 1;
    wr->rsq = pl.rsq;
    wr->maxrr = pl.maxrr;
    wr->pending = pl.pending;
    // (constant indices only: a dynamically indexed member would move the
    // plan to scratch memory)
    if (rd->hist) {
      if (pl.steps > 0) rd->hist[k0] = sqrt(pl.hist[0]) / rd->bsqrt;
      if (pl.steps > 1) rd->hist[k0 + 1] = sqrt(pl.hist[1]) / rd->bsqrt;
    }
    if (pl.done) {
      wr->kfinal = pl.k;
      wr->done = pl.done;
      wr->rsqNew = pl.steps > 1 ? pl.hist[1] : pl.hist[0];
    }
    if (rd->hdone) *rd->hdone = wr->done;
  }
}

}  // namespace glx

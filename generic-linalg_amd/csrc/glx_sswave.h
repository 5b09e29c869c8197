// glx_sswave.h -- the s-step launch of glx_sstep.h (two CG iterations of the
// mixed solver) on the staggered normal operator over full-lattice vectors,
// with every wave marching a strip of its own: no LDS rings and no barriers
// inside the march.
//
// The prelude (the previous launch's sums, the two CG iterations on
// coordinates) and the stop path are those of k_sscg; the stream computes the
// same quantities with the hops as chains of fused multiply-adds (rounded
// differently from k_sscg's) and the FP64 sums accumulated in another order.
//
// Work split: a wave owns 56 columns and `rows` rows of the lattice.  Lane l
// holds column xo - 4 + l (periodically wrapped), so the wave's 64 lanes
// cover its own columns and 4 halo columns each side; lanes 4 .. 59 own
// their column.  The x-neighbours of a lane's value come from lanes l +- 1
// through wave-wide DPP shifts: every hop loses the outermost lane on each
// side (it reads 0 there), and after the four hops of P2 = A^2 p lanes 4 .. 59
// still hold exact values.  The wave marches its rows from 4 halo rows on one
// side of its band to 4 on the other (downwards in even bands, upwards in odd
// ones), one row per step, with a lag of one row per stage (stage s of step t
// works on march row t - s):
//   S0  row t      p, r from the basis, x += Y c_x, the links (halved)
//   S1  row t - 1  qp = K p, qr = K r
//   S2  row t - 2  P1 = m^2 p - K qp
//   S3  row t - 3  q1 = K P1
//   S4  row t - 4  P2 = m^2 P1 - K q1, R1 = m^2 r - K qr, the 12 sums
// Every stage's rows live in registers, in windows of six rows indexed by the
// step modulo 6; the march loop is unrolled six times, so that every window
// index is a constant and every window a set of registers.  The loads of row
// t + 3 are issued at step t (three rows in flight; rows 0 .. 5 in the
// prelude).  Rows of a stage outside the rows it must supply (the first
// steps of the march, and the steps that round the march up to a multiple
// of six) hold values no owned result depends on; band heights of 6 k - 8
// rows waste no step.
#pragma once

#include "glx_sstep.h"

namespace glx {

constexpr int kSwOwn = 56;  // columns a wave owns
constexpr int kSwHalo = 4;  // halo columns and rows each side (four hops)
constexpr int kSwWin = 6;   // rows in every register window
constexpr int kSwAhead = 3; // rows whose loads are in flight (six measured slower)
constexpr int kSwMaxRows = 56;  // rows per wave: a march of rows + 8 <= 64 rows (one row per lane in rowtab)

// Launch geometry of the wave-strip s-step kernel.
struct SwGeom {
  int lx, ly;
  int nstrip, nband, rows;  // strips of kSwOwn columns, bands of `rows` rows
  int ntask;                // nstrip * nband waves of work
  int nblk;                 // ceil(ntask / kWaves) blocks, rounded up to a multiple of 8
};
inline SwGeom make_swgeom(int lx, int ly, int rows) {
  SwGeom g;
  g.lx = lx;
  g.ly = ly;
  g.rows = rows;
  g.nstrip = (lx + kSwOwn - 1) / kSwOwn;
  g.nband = (ly + rows - 1) / rows;
  g.ntask = g.nstrip * g.nband;
  // whole groups of 8 blocks: xcd_tile maps block ids onto tiles one to one
  // only then (the padding blocks join the sums with zeros)
  g.nblk = ((g.ntask + kWaves * 8 - 1) / (kWaves * 8)) * 8;
  return g;
}

// The value lane l - 1 holds; lane 0 reads 0.
__device__ __forceinline__ float sw_prev1(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x138, 0xF, 0xF, true));
}
__device__ __forceinline__ c64 sw_prev(c64 v) { return c64{sw_prev1(v.re), sw_prev1(v.im)}; }
// a float every lane of the wave holds (into a scalar register)
__device__ __forceinline__ float sw_uniform(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, v)));
}

// K f at one site from its four neighbours, with the links halved and the
// y-links already multiplied by the site's staggered phase (+-1):
//   K f = ux f(x+1) - conj(uxm) f(x-1) + uy f(y+1) - conj(uym) f(y-1)
// as two chains of fused multiply-adds (the hop of khop_h, rounded
// differently); sw_mhop gives m2 v - K f.  The x-neighbours are read from
// lanes l +- 1 inside the multiply-adds (DPP operands of v_mul / v_fmac: the
// compiler fuses a DPP move into v_mul and v_add but not into v_fmac), in
// inline asm that begins with `s_nop 1`: a DPP operand must be written by
// VALU at least two wait states earlier, and the compiler's hazard checks do
// not look inside asm.  The wave's end lanes read 0 (bound_ctrl).  Every hop
// runs with all 64 lanes active (clang marks inline asm convergent, so it is
// never moved under a lane-dependent branch).
#define GLX_DPP_NEXT "wave_shl:1 row_mask:0xf bank_mask:0xf bound_ctrl:0"
#define GLX_DPP_PREV "wave_shr:1 row_mask:0xf bank_mask:0xf bound_ctrl:0"
// ux f(x+1) - conj(uxm) f(x-1)
__device__ __forceinline__ c64 sw_hop_x(c64 f, c64 ux, c64 uxm) {
  float re, im;
  asm("s_nop 1\n\t"
      "v_mul_f32_dpp %0, %2, %4 " GLX_DPP_NEXT "\n\t"    // re  = ux.re f+.re
      "v_mul_f32_dpp %1, %3, %4 " GLX_DPP_NEXT "\n\t"    // im  = ux.re f+.im
      "v_fmac_f32_dpp %0, %3, -%5 " GLX_DPP_NEXT "\n\t"  // re -= ux.im f+.im
      "v_fmac_f32_dpp %1, %2, %5 " GLX_DPP_NEXT "\n\t"   // im += ux.im f+.re
      "v_fmac_f32_dpp %0, %2, -%6 " GLX_DPP_PREV "\n\t"  // re -= uxm.re f-.re
      "v_fmac_f32_dpp %1, %3, -%6 " GLX_DPP_PREV "\n\t"  // im -= uxm.re f-.im
      "v_fmac_f32_dpp %0, %3, -%7 " GLX_DPP_PREV "\n\t"  // re -= uxm.im f-.im
      "v_fmac_f32_dpp %1, %2, %7 " GLX_DPP_PREV          // im += uxm.im f-.re
      : "=&v"(re), "=&v"(im)
      : "v"(f.re), "v"(f.im), "v"(ux.re), "v"(ux.im), "v"(uxm.re), "v"(uxm.im));
  return c64{re, im};
}
// acc - (ux f(x+1) - conj(uxm) f(x-1))
__device__ __forceinline__ c64 sw_mhop_x(c64 acc, c64 f, c64 ux, c64 uxm) {
  float re = acc.re, im = acc.im;
  asm("s_nop 1\n\t"
      "v_fmac_f32_dpp %0, %2, -%4 " GLX_DPP_NEXT "\n\t"  // re -= ux.re f+.re
      "v_fmac_f32_dpp %1, %3, -%4 " GLX_DPP_NEXT "\n\t"  // im -= ux.re f+.im
      "v_fmac_f32_dpp %0, %3, %5 " GLX_DPP_NEXT "\n\t"   // re += ux.im f+.im
      "v_fmac_f32_dpp %1, %2, -%5 " GLX_DPP_NEXT "\n\t"  // im -= ux.im f+.re
      "v_fmac_f32_dpp %0, %2, %6 " GLX_DPP_PREV "\n\t"   // re += uxm.re f-.re
      "v_fmac_f32_dpp %1, %3, %6 " GLX_DPP_PREV "\n\t"   // im += uxm.re f-.im
      "v_fmac_f32_dpp %0, %3, %7 " GLX_DPP_PREV "\n\t"   // re += uxm.im f-.im
      "v_fmac_f32_dpp %1, %2, -%7 " GLX_DPP_PREV         // im -= uxm.im f-.re
      : "+v"(re), "+v"(im)
      : "v"(f.re), "v"(f.im), "v"(ux.re), "v"(ux.im), "v"(uxm.re), "v"(uxm.im));
  return c64{re, im};
}
__device__ __forceinline__ c64 sw_hop(c64 f, c64 fyp, c64 fym, c64 ux, c64 uxm, c64 uy, c64 uym) {
  const c64 x = sw_hop_x(f, ux, uxm);
  float re = x.re, im = x.im;
  re = fmaf(uy.re, fyp.re, re);
  re = fmaf(-uy.im, fyp.im, re);
  re = fmaf(-uym.re, fym.re, re);
  re = fmaf(-uym.im, fym.im, re);
  im = fmaf(uy.re, fyp.im, im);
  im = fmaf(uy.im, fyp.re, im);
  im = fmaf(-uym.re, fym.im, im);
  im = fmaf(uym.im, fym.re, im);
  return c64{re, im};
}
__device__ __forceinline__ c64 sw_mhop(float m2, c64 v, c64 f, c64 fyp, c64 fym, c64 ux, c64 uxm, c64 uy, c64 uym) {
  const c64 x = sw_mhop_x(c64{m2 * v.re, m2 * v.im}, f, ux, uxm);
  float re = x.re, im = x.im;
  re = fmaf(-uy.re, fyp.re, re);
  re = fmaf(uy.im, fyp.im, re);
  re = fmaf(uym.re, fym.re, re);
  re = fmaf(uym.im, fym.im, re);
  im = fmaf(-uy.re, fyp.im, im);
  im = fmaf(-uy.im, fyp.re, im);
  im = fmaf(uym.re, fym.im, im);
  im = fmaf(-uym.im, fym.re, im);
  return c64{re, im};
}

template <class LK>
struct SwRow {
  c64 y[5];
  c64 x;
  typename LK::Raw lk[1];
};

// EO: the even/odd Schur operator m^2 - D_eo D_oe on compact even-site
// vectors (y (lx / 2) + (x >> 1)).  The lanes still hold full lattice
// columns: a lane whose site of the row is even loads and stores the compact
// entry, one whose site is odd holds zeros in p, r, P1, P2 and R1 (the hops
// carry even-site values to odd sites and back, so those stay exactly zero
// and add nothing to the sums).  The basis and x then move half the bytes;
// the links are the full lattice's.
template <class LK, class St, bool EO = false>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(2)))
void k_sscg_w(SwGeom g, OpDev<LK> op, SsBufs bf, const St* rd, St* wr, int q, RedWs ws) {
  static_assert(kSwWin % kSwAhead == 0, "the load slots rotate with the windows");
  using C = typename LK::C;
  static_assert(std::is_same<C, c64>::value, "FP32 link elements");
  if (rd->done != 0) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      ss_carry(rd, wr);
      if (rd->hdone) *rd->hdone = rd->done;
    }
    return;
  }
  const bool full = rd->pending != 0;
  const int k0 = rd->k, ko = rd->ko;
  const int nblk = gridDim.x;

  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), l = threadIdx.x & 63;
  const int task = xcd_tile(blockIdx.x, nblk) * kWaves + w;
  const bool active = task < g.ntask;
  const int tk = active ? task : 0;
  const int strip = tk % g.nstrip, band = tk / g.nstrip;
  const int lx = g.lx, ly = g.ly;
  const int xo = strip * kSwOwn;
  const int xl = wrap_coord(xo - kSwHalo + l, lx);  // this lane's column
  const bool own_lane = l >= kSwHalo && l < kSwHalo + kSwOwn && xo + l - kSwHalo < lx;
  const float yh = (xl & 1) != 0 ? -0.5f : 0.5f;  // the halved y-links carry the staggered phase
  const int yb = band * g.rows;
  const int rows = min(g.rows, ly - yb);
  const int steps = rows + 2 * kSwHalo;
  // march row t is lattice row yb - 4 + t; it is owned for 4 <= t < rows + 4
  // odd bands march upwards: the two waves either side of a band boundary
  // then read the rows they share at the same time (both at the start or
  // both at the end of their marches), and the second read hits L2
  const bool up = (band & 1) != 0;
  auto row_of = [&](int t) {
    return ss_wrap_row(up ? yb + rows + kSwHalo - 1 - t : yb - kSwHalo + t, ly);
  };
  // the first site of every march row, lane t holding row t's (steps <= 64;
  // read with readlane, one instruction per use, where the row arithmetic
  // would be several scalar ones)
  // (EO: the compact row offsets, y lx / 2)
  const int rowtab = row_of(min(l, steps - 1)) * (EO ? lx / 2 : lx);
  auto rowj = [&](int t) { return __builtin_amdgcn_readlane(rowtab, t); };
  auto owned = [&](int t) { return t >= kSwHalo && t < rows + kSwHalo; };
  // the lane's vector entry in march row t, and its site's links
  const int xv = EO ? xl >> 1 : xl;
  auto vj = [&](int t) { return unsigned(rowj(t) + xv); };
  auto lj = [&](int t) { return EO ? unsigned(2 * rowj(t) + xl) : unsigned(rowj(t) + xl); };
  // EO: whether the lane's site is even in march rows of even t (ev0) and
  // of odd t (ev1); rows alternate parity along the march (ly is even)
  const bool ev0 = ((xl + row_of(0)) & 1) == 0, ev1 = !ev0;
  const int so = ko & 1, sn = so ^ 1;
  const c64* Y[5] = {bf.v(so, 0), bf.v(so, 1), bf.v(so, 2), bf.v(so, 3), bf.v(so, 4)};
  c64* N[5] = {bf.v(sn, 0), bf.v(sn, 1), bf.v(sn, 2), bf.v(sn, 3), bf.v(sn, 4)};

  // every lane loads every row it marches, x included (a load under a branch
  // would make the compiler wait for all loads in flight where the branches
  // join); past the last row the last row is loaded again
  auto load_row = [&](SwRow<LK>& d, int t) {
    const int tt = min(t, steps - 1);
    const unsigned j = vj(tt);
#pragma unroll
    for (int i = 0; i < 5; i++) d.y[i] = ss_at<c64>(Y[i], j * 8u);
    d.x = ss_at<c64>(bf.x, j * 8u);
    ss_links<1>(op.links, lj(tt), d.lk);
  };

  constexpr int kPreU = 2;
  double pv[kPreU][kSsNR];
  const double* PS = ws.partials + (q ^ 1) * kSsNR * nblk;
  // the previous launch's sums and the first rows: every load issued before
  // any is waited for, and none under a branch (clamped indices, masked
  // values), so that the wait for the sums leaves the rows in flight
#pragma unroll
  for (int u = 0; u < kPreU; u++) {
    const int bb = min(int(threadIdx.x) + u * kBlock, nblk - 1);
#pragma unroll
    for (int k = 0; k < kSsNR; k++) pv[u][k] = PS[k * nblk + bb];
  }
  __builtin_amdgcn_sched_barrier(0);  // (keeps the sums' loads ahead of the rows')
  SwRow<LK> ld[kSwAhead];
#pragma unroll
  for (int a = 0; a < kSwAhead; a++) load_row(ld[a], a);  // (waves without a task load task 0's rows)
  // rows 3 .. 5 too, in registers the march's windows do not yet hold: the
  // first round of the march takes them instead of loading (launch
  // 27.35 -> 27.16 us at 1024^2, medians of five alternated runs)
  SwRow<LK> ld2[kSwAhead];
#pragma unroll
  for (int a = 0; a < kSwAhead; a++) load_row(ld2[a], a + kSwAhead);
  __shared__ double sm[kSsNR][kWaves];
  SsPlan pl;
  if (full) {
    double acc[kSsNR];
#pragma unroll
    for (int k = 0; k < kSsNR; k++) acc[k] = int(threadIdx.x) < nblk ? pv[0][k] : 0.0;
#pragma unroll
    for (int u = 1; u < kPreU; u++) {
      const bool in = int(threadIdx.x) + u * kBlock < nblk;
#pragma unroll
      for (int k = 0; k < kSsNR; k++) acc[k] += in ? pv[u][k] : 0.0;
    }
    for (int b0 = threadIdx.x + kBlock * kPreU; b0 < nblk; b0 += kBlock) {
#pragma unroll
      for (int k = 0; k < kSsNR; k++) acc[k] += PS[k * nblk + b0];
    }
    block_sum_dpp<kSsNR>(acc, sm);
    ss_iterate(acc, rd, k0, rd->maxrr, pl);
  } else {
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
  // (in vector registers: the scalar registers hold the march's addresses)
  float cp[5], cr[5], cx[5];
#pragma unroll
  for (int i = 0; i < 5; i++) {
    cp[i] = float(pl.cp[i]);
    cr[i] = float(pl.cr[i]);
    cx[i] = float(pl.cx[i]);
  }

  if (!pl.stencil) {
    // stopped: x += Y c_x and P0' = Y c_p on the owned sites
    if (active && own_lane) {
      for (int t = kSwHalo; t < rows + kSwHalo; t++) {
        if (EO && !((t & 1) ? ev1 : ev0)) continue;
        const unsigned j = vj(t);
        c64 yv[5];
#pragma unroll
        for (int i = 0; i < 5; i++) yv[i] = ss_at<c64>(Y[i], j * 8u);
        const c64 xv = ss_at<c64>(bf.x, j * 8u);
        ss_at<c64>(N[0], j * 8u) = ss_comb(cp, yv, true);
        ss_at<c64>(bf.x, j * 8u) = xv + ss_comb(cx, yv, true);
      }
    }
  } else {
    const float m2 = float(op.mass * op.mass);
    double part[kSsNR];
#pragma unroll
    for (int k = 0; k < kSsNR; k++) part[k] = 0.0;
    // register windows, slot = march row mod 6 (zeroed: the first steps'
    // stages read slots no earlier step wrote)
    c64 p[kSwWin], r[kSwWin], qp[kSwWin], qr[kSwWin], P1[kSwWin], q1[kSwWin], ux[kSwWin], uxm[kSwWin], uy[kSwWin];
#pragma unroll
    for (int s = 0; s < kSwWin; s++)
      p[s] = r[s] = qp[s] = qr[s] = P1[s] = q1[s] = ux[s] = uxm[s] = uy[s] = c64{0.0f, 0.0f};

    // R0: the first round of the march (t = K), apart so that its steps take
    // the rows the prelude loaded and leave out the stages whose rows no
    // owned result needs (S1 before t = 2, S2 before t = 4, S3 and S4 in the
    // whole round: the owned rows start at t = 4, so S4 needs q1 from row 3,
    // q1 needs P1 from row 2, P1 needs qp from row 1)
    auto step = [&](auto kc, int t, auto rc) __attribute__((always_inline)) {
      constexpr bool R0 = decltype(rc)::value;
      constexpr int K = decltype(kc)::value;
      constexpr int S0 = K, S1 = (K + 5) % 6, S2 = (K + 4) % 6, S3 = (K + 3) % 6, S4 = (K + 2) % 6,
                    S5 = (K + 1) % 6;
      // stage d's hop at march row m = t - d takes "y + 1" from march row
      // m + 1 and "y - 1" from m - 1, with the y-links of rows m and m - 1
      constexpr int LA = K % kSwAhead;
      c64 xn;
      // S0: row t
      {
        SwRow<LK>& cur = ld[LA];
        // (EO: the lanes whose site is odd take zeros)
        const c64 z{0.0f, 0.0f};
        const bool ev = (K & 1) ? ev1 : ev0;
        const c64 yv[5] = {!EO || ev ? cur.y[0] : z, !EO || ev ? cur.y[1] : z, !EO || ev ? cur.y[2] : z,
                           !EO || ev ? cur.y[3] : z, !EO || ev ? cur.y[4] : z};
        p[S0] = ss_comb_p(cp, yv);
        r[S0] = ss_comb_r(cr, yv);
        C u0, u1;
        LK::finish(cur.lk[0], u0, u1);
        ux[S0] = 0.5f * u0;
        uxm[S0] = sw_prev(ux[S0]);
        // the y-links of the march: downwards U_y of the row itself;
        // upwards march row m + 1 is lattice row y - 1, and the hop
        // U_y(y) f(y+1) - conj(U_y(y-1)) f(y-1) is the downward one with the
        // links V(m) = -conj(U_y(y - 1)), which the load of march row m + 1
        // brings: the row loaded now gives V of the row before it
        const c64 vd = yh * u1, vu = c64{-yh * u1.re, yh * u1.im};
        uy[S0] = up ? uy[S0] : vd;
        uy[S1] = up ? vu : uy[S1];
        xn = (!EO || ev ? cur.x : z) + ss_comb_x(cx, yv);
        if constexpr (R0 && K < kSwAhead) cur = ld2[LA];
        else load_row(cur, t + kSwAhead);
      }
      // S1: row t - 1
      if constexpr (!R0 || K >= 2) {
        qp[S1] = sw_hop(p[S1], p[S0], p[S2], ux[S1], uxm[S1], uy[S1], uy[S2]);
        qr[S1] = sw_hop(r[S1], r[S0], r[S2], ux[S1], uxm[S1], uy[S1], uy[S2]);
      }
      // S2: row t - 2
      if constexpr (!R0 || K >= 4)
        P1[S2] = sw_mhop(m2, p[S2], qp[S2], qp[S1], qp[S3], ux[S2], uxm[S2], uy[S2], uy[S3]);
      // S3: row t - 3
      if constexpr (!R0) q1[S3] = sw_hop(P1[S3], P1[S2], P1[S4], ux[S3], uxm[S3], uy[S3], uy[S4]);
      // S4: row t - 4, the owned rows only
      const bool o4 = !R0 && owned(t - 4);
      c64 p2, r1;
      if (o4) {
        p2 = sw_mhop(m2, P1[S4], q1[S4], q1[S3], q1[S5], ux[S4], uxm[S4], uy[S4], uy[S5]);
        r1 = sw_mhop(m2, r[S4], qr[S4], qr[S3], qr[S5], ux[S4], uxm[S4], uy[S4], uy[S5]);
      }
      // the stores and sums of the lanes that own their column, in one
      // masked region per step (EO: the rows t, t - 2 and t - 4 have one
      // parity; only the lanes whose site is even store)
      if (own_lane && (!EO || ((K & 1) ? ev1 : ev0))) {
        if (owned(t)) {
          const unsigned j = vj(t) * 8u;
          ss_at<c64>(N[0], j) = p[S0];
          ss_at<c64>(N[3], j) = r[S0];
          ss_at<c64>(bf.x, j) = xn;
        }
        if (!R0 && owned(t - 2)) ss_at<c64>(N[1], vj(t - 2) * 8u) = P1[S2];
        if (o4) {
          const c64 pp = p[S4], pP1 = P1[S4], rr = r[S4];
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
          const unsigned j = vj(t - 4) * 8u;
          ss_at<c64>(N[2], j) = p2;
          ss_at<c64>(N[4], j) = r1;
        }
      }
    };
    // whole rounds of six steps: the steps past the last row store nothing
    // (no row they reach is owned) and load the last row again
    using std::integral_constant;
    using std::true_type;
    using std::false_type;
    if (active) {
      step(integral_constant<int, 0>{}, 0, true_type{});
      step(integral_constant<int, 1>{}, 1, true_type{});
      step(integral_constant<int, 2>{}, 2, true_type{});
      step(integral_constant<int, 3>{}, 3, true_type{});
      step(integral_constant<int, 4>{}, 4, true_type{});
      step(integral_constant<int, 5>{}, 5, true_type{});
    }
    for (int t0 = kSwWin; t0 < (active ? steps : 0); t0 += kSwWin) {
      step(integral_constant<int, 0>{}, t0, false_type{});
      step(integral_constant<int, 1>{}, t0 + 1, false_type{});
      step(integral_constant<int, 2>{}, t0 + 2, false_type{});
      step(integral_constant<int, 3>{}, t0 + 3, false_type{});
      step(integral_constant<int, 4>{}, t0 + 4, false_type{});
      step(integral_constant<int, 5>{}, t0 + 5, false_type{});
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

// The memory traffic of k_sscg_w alone: the same waves, march, loads (three
// rows ahead) and stores at the same addresses, with the arithmetic replaced
// by a few adds -- what the one-wave-per-SIMD strip pattern itself streams
// at, the floor of k_sscg_w's time.  Timing only: it writes sums of its
// inputs into the next basis slot and x.
template <class LK>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(2)))
void k_sscg_w_traffic(SwGeom g, OpDev<LK> op, SsBufs bf, int ko) {
  const int nblk = gridDim.x;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), l = threadIdx.x & 63;
  const int task = xcd_tile(blockIdx.x, nblk) * kWaves + w;
  if (task >= g.ntask) return;
  const int strip = task % g.nstrip, band = task / g.nstrip;
  const int lx = g.lx, ly = g.ly;
  const int xo = strip * kSwOwn;
  const int xl = wrap_coord(xo - kSwHalo + l, lx);
  const bool own_lane = l >= kSwHalo && l < kSwHalo + kSwOwn && xo + l - kSwHalo < lx;
  const int yb = band * g.rows;
  const int rows = min(g.rows, ly - yb);
  const int steps = rows + 2 * kSwHalo;
  const bool up = (band & 1) != 0;
  auto row_of = [&](int t) {
    return ss_wrap_row(up ? yb + rows + kSwHalo - 1 - t : yb - kSwHalo + t, ly);
  };
  const int rowtab = row_of(min(l, steps - 1)) * lx;
  auto rowj = [&](int t) { return __builtin_amdgcn_readlane(rowtab, t); };
  auto owned = [&](int t) { return t >= kSwHalo && t < rows + kSwHalo; };
  const int so = ko & 1, sn = so ^ 1;
  const c64* Y[5] = {bf.v(so, 0), bf.v(so, 1), bf.v(so, 2), bf.v(so, 3), bf.v(so, 4)};
  c64* N[5] = {bf.v(sn, 0), bf.v(sn, 1), bf.v(sn, 2), bf.v(sn, 3), bf.v(sn, 4)};
  auto load_row = [&](SwRow<LK>& d, int t) {
    const unsigned j = unsigned(rowj(min(t, steps - 1)) + xl);
#pragma unroll
    for (int i = 0; i < 5; i++) d.y[i] = ss_at<c64>(Y[i], j * 8u);
    d.x = ss_at<c64>(bf.x, j * 8u);
    ss_links<1>(op.links, j, d.lk);
  };
  SwRow<LK> ld[kSwAhead];
#pragma unroll
  for (int a = 0; a < kSwAhead; a++) load_row(ld[a], a);
  c64 acc{0.0f, 0.0f};
  auto step = [&](auto kc, int t) __attribute__((always_inline)) {
    constexpr int K = decltype(kc)::value;
    SwRow<LK>& cur = ld[K % kSwAhead];
    typename LK::C u0, u1;
    LK::finish(cur.lk[0], u0, u1);
    c64 v[6];
#pragma unroll
    for (int i = 0; i < 5; i++) v[i] = cur.y[i] + acc;
    v[5] = cur.x + u0 + u1;
    acc = acc + c64{1e-30f * v[0].re, 1e-30f * v[4].im};
    load_row(cur, t + kSwAhead);
    if (own_lane) {
      if (owned(t)) {
        const unsigned j = unsigned(rowj(t) + xl) * 8u;
        ss_at<c64>(N[0], j) = v[0];
        ss_at<c64>(N[3], j) = v[3];
        ss_at<c64>(bf.x, j) = v[5];
      }
      if (owned(t - 2)) ss_at<c64>(N[1], unsigned(rowj(t - 2) + xl) * 8u) = v[1];
      if (owned(t - 4)) {
        const unsigned j = unsigned(rowj(t - 4) + xl) * 8u;
        ss_at<c64>(N[2], j) = v[2];
        ss_at<c64>(N[4], j) = v[4];
      }
    }
  };
  using std::integral_constant;
  for (int t0 = 0; t0 < steps; t0 += kSwWin) {
    step(integral_constant<int, 0>{}, t0);
    step(integral_constant<int, 1>{}, t0 + 1);
    step(integral_constant<int, 2>{}, t0 + 2);
    step(integral_constant<int, 3>{}, t0 + 3);
    step(integral_constant<int, 4>{}, t0 + 4);
    step(integral_constant<int, 5>{}, t0 + 5);
  }
}

}  // namespace glx

// glx_eo.h -- the even/odd Schur operator m^2 - D_eo D_oe on compacted
// even-site vectors, as a row-marching kernel.
//
// The reference keeps even/odd vectors full length with the other parity
// zeroed (operators.cpp:543-616: eoprec_prepare, m2mdeodoe, reconstruct), so
// every pass reads and writes the zero half too.  Here even-site vectors are
// compacted: site (x, y) with x + y even is element y * (lx / 2) + (x >> 1).
// With D = m - K (K the staggered hopping term, anti-Hermitian, flipping
// parity) D_eo D_oe = K_eo K_oe, so the operator is m^2 - K_eo K_oe: the
// even-even block of D^dag D, Hermitian positive definite, solved by CG on
// half the sites.
//
// Row y has lx / 2 even sites at x = 2c + (y & 1) and lx / 2 odd sites at
// x = 2c + 1 - (y & 1) (compact column c).  Neighbours in compact columns:
//   odd site c:  even x-neighbours c - e (x - 1) and c + 1 - e (x + 1),
//                even y-neighbours column c of rows y +- 1   (e = y & 1)
//   even site c: odd x-neighbours c + e - 1 (x - 1) and c + e (x + 1),
//                odd y-neighbours column c of rows y +- 1
// The march is the one of glx_normal.h (4 waves, 4 rows per step, 8-row LDS
// rings, reads issued one step ahead): stage p (compact even) through the
// caller's prologue and the links of the full rows; q = K_oe p on the odd
// sites of the strip plus one compact column each side; out = m^2 p - K_eo q.
// Links are staged by full-lattice column: x in [2 c0 - 3, 2 c0 + 2 NXB + 1]
// for a block of compact columns [c0, c0 + NXB).
//
// Requirements (checked on the host by eo_fits): lx even with (lx / 2) a
// multiple of 64 * cpl, ly a multiple of 4 and at least 8, sr a multiple of
// 4 and at least 8.
#pragma once

#include "glx_normal.h"

namespace glx {

inline bool eo_fits(int lx, int ly, int sr, int cpl) {
  const int nx = kNX * cpl;
  return lx % 2 == 0 && lx / 2 >= nx && (lx / 2) % nx == 0 && ly >= 8 && ly % 4 == 0 && sr >= 8 && sr % 4 == 0;
}

// Geometry in compact columns: the NGeom of a lattice lx/2 wide, with the
// full width kept for the links.
struct EoGeom {
  NGeom g;  // g.lx = lx / 2 (compact), g.ly = ly
  int lx;   // full width
};
inline EoGeom make_eogeom(int lx, int ly, int sr, int cpl) {
  EoGeom e;
  e.g = make_ngeom(lx / 2, ly, sr, cpl);
  e.lx = lx;
  return e;
}

template <int CPL, class T, class LK, class In, class Epi>
__device__ __forceinline__ void eo_body(const EoGeom& eg, const OpDev<LK>& op, In& in, T* __restrict__ out_arg,
                                        const Epi& epi, RedWs ws, int bid) {
  using C = typename LK::C;
  constexpr int NXB = kNX * CPL;     // compact columns per block
  constexpr int PW = NXB + 4;        // p (even) compact columns c0-2 .. c0+NXB+1
  constexpr int QW = NXB + 2;        // q (odd) compact columns c0-1 .. c0+NXB
  constexpr int LW = 2 * NXB + 6;    // links, full columns 2c0-3 .. 2c0+2NXB+1 (last slot unused)
  constexpr int RM = kNRing - 1;
  const NGeom& g = eg.g;

  in.init();
  T* __restrict__ out = in.target(out_arg);

  __shared__ T sP[kNRing][PW];
  __shared__ C sLx[kNRing][LW];
  __shared__ C sLy[kNRing][LW];
  __shared__ T sQ[kNRing][QW];
  __shared__ T sA[In::kAux ? kNRing : 1][In::kAux ? NXB : 1];

  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int tile = xcd_tile(bid, g.nblk);
  double part[Epi::NR > 0 ? Epi::NR : 1];
#pragma unroll
  for (int k = 0; k < Epi::NR; k++) part[k] = 0.0;

  constexpr bool kSF = has_selffin<In>::value;
  const bool active = tile < g.ntiles;
  const int tt = active ? tile : 0;  // idle blocks (grid padding) only join the reductions
  {
    const int W = g.lx, ly = g.ly, lx = eg.lx;
    const int c0 = (tt % g.ntx) * NXB;
    const int y0 = (tt / g.ntx) * g.sr;  // even: sr is a multiple of 4
    const int rows = min(g.sr, ly - y0);
    const int nsteps = rows / 4 + 1;
    const double m2 = op.mass * op.mass;
    // halo roles: wave 0 lanes 0..15 stage p on compact columns c0-2, c0-1,
    // c0+NXB, c0+NXB+1 of the step's four rows; wave 1 lanes 0..19 stage the
    // links on full columns 2c0-3, 2c0-2, 2c0-1, 2c0+2NXB, 2c0+2NXB+1
    const bool hp = w == 0 && l < 16;
    const bool hl = w == 1 && l < 20;
    const int hp_row = l >> 2, hp_c = (l & 3) < 2 ? (l & 3) : NXB + (l & 3);
    const int hl_row = l / 5, hl_j = l % 5;
    const int hl_i = hl_j < 3 ? hl_j : 2 * NXB + hl_j;  // LDS column: 0,1,2, 2NXB+3, 2NXB+4
    const int cm = c0 + CPL * l;                        // first of this lane's compact columns
    const int chp = wrap_coord(c0 - 2 + hp_c, W);
    const int xhl = wrap_coord(2 * c0 - 3 + hl_i, lx);
    const int xl0 = 2 * c0 + 2 * CPL * l;               // first of this lane's 2 CPL link columns

    InRegs<In, CPL> rm;
    LinkRegs<LK, 2 * CPL> lm;
    typename In::Raw rh;
    typename LK::Raw lh;
    auto fetch = [&](int k) {
      const int yr = 4 * k - 2 + w;
      const int gy = wrap_coord(y0 + yr, ly);
      rm.fetch(in, gy * W + cm, yr >= 0 && yr < rows);
      lm.fetch(op.links, gy * lx + xl0);
      if (hp) rh = in.fetch(wrap_coord(y0 + 4 * k - 2 + hp_row, ly) * W + chp, false);
      if (hl) lh = op.links.fetch(wrap_coord(y0 + 4 * k - 2 + hl_row, ly) * lx + xhl);
    };

    // the previous launch's partials are read first: loads complete in
    // issue order, so the prelude's wait does not include the first rows
    if constexpr (kSF) in.prelude_load(ws);
    if (active) fetch(0);
    bool run = true;
    if constexpr (kSF) run = in.prelude_finish(ws);
    for (int k = 0; k < ((active && run) ? nsteps : 0); k++) {
      if (k > 0) __syncthreads();  // step k-1's readers are done with the slots reused below
      {
        // stage rows 4k-2 .. 4k+1 (relative to y0)
        const int yr = 4 * k - 2 + w;
        const int gy = wrap_coord(y0 + yr, ly);
        const bool own = yr >= 0 && yr < rows;
        const int sl = (yr + 2) & RM;
        T v[CPL], aux[CPL];
        rm.finish(in, gy * W + cm, own, v, aux);
#pragma unroll
        for (int c = 0; c < CPL; c++) {
          sP[sl][CPL * l + c + 2] = v[c];
          if constexpr (In::kAux) {
            if (own) sA[sl][CPL * l + c] = aux[c];
          }
        }
#pragma unroll
        for (int t = 0; t < 2 * CPL; t++) {
          C ux, uy;
          LK::finish(lm.r[t], ux, uy);
          sLx[sl][2 * CPL * l + t + 3] = ux;
          sLy[sl][2 * CPL * l + t + 3] = uy;
        }
        if (hp) {
          const int yh = 4 * k - 2 + hp_row;
          T d;
          sP[(yh + 2) & RM][hp_c] = in.finish(rh, wrap_coord(y0 + yh, ly) * W + chp, false, d);
        }
        if (hl) {
          const int yh = 4 * k - 2 + hl_row;
          C hx, hy;
          LK::finish(lh, hx, hy);
          sLx[(yh + 2) & RM][hl_i] = hx;
          sLy[(yh + 2) & RM][hl_i] = hy;
        }
      }
      if (k + 1 < nsteps) fetch(k + 1);
      __syncthreads();
      {
        // q = K_oe p on odd sites: step 0 rows -1, 0; step k > 0 rows 4k-3 .. 4k
        const int qa = k == 0 ? -1 : 4 * k - 3;
        const int yr = qa + w;
        // q column qc of row yq is compact odd column c0 - 1 + qc, full
        // x = 2 c0 + 2 qc - 1 - e with e = yq & 1
        auto qhop = [&](int yq, int qc) {
          const int e = yq & 1;
          const int s0 = (yq + 2) & RM, sm = (yq + 1) & RM, sp = (yq + 3) & RM;
          const bool eta_neg = e == 0;  // x odd
          const int li = 2 * qc + 2 - e;
          sQ[s0][qc] = khop<T, C>(sP[s0][qc + 2 - e], sP[s0][qc + 1 - e], sP[sp][qc + 1], sP[sm][qc + 1],
                                  sLx[s0][li], sLx[s0][li - 1], sLy[s0][li], sLy[sm][li], eta_neg);
        };
        if (yr <= 4 * k) {
#pragma unroll
          for (int c = 0; c < CPL; c++) qhop(yr, CPL * l + c + 1);
        }
        // the two halo columns of all the step's q rows on wave 3
        if (w == 3 && l < 8) {
          const int yq = qa + (l >> 1);
          if (yq <= 4 * k) qhop(yq, (l & 1) == 0 ? 0 : NXB + 1);
        }
      }
      __syncthreads();
      if (k > 0) {
        // out = m^2 p - K_eo q on even sites, rows 4k-4 .. 4k-1
        const int yr = 4 * k - 4 + w;
        const int e = yr & 1;
        const int s0 = (yr + 2) & RM, sm = (yr + 1) & RM, sp = (yr + 3) & RM;
        T v[CPL];
        const int i0 = (y0 + yr) * W + cm;
#pragma unroll
        for (int c = 0; c < CPL; c++) {
          const int b = CPL * l + c;     // p column b + 2, q columns b + e (x - 1), b + e + 1 (x + 1), b + 1
          const int li = 2 * b + e + 3;  // link column of x = 2 (c0 + b) + e
          const T kq = khop<T, C>(sQ[s0][b + e + 1], sQ[s0][b + e], sQ[sp][b + 1], sQ[sm][b + 1], sLx[s0][li],
                                  sLx[s0][li - 1], sLy[s0][li], sLy[sm][li], e != 0);
          const T pin = sP[s0][b + 2];
          v[c] = rscal(m2, pin) - kq;
          if constexpr (In::kAux) epi.acc(part, i0 + c, pin, v[c], sA[s0][b]);
          else epi.acc(part, i0 + c, pin, v[c], pin);
        }
        if constexpr (CPL == 2) st2(out, i0, v[0], v[1]);
        else out[i0] = v[0];
      }
    }
    if constexpr (kSF) {
      if (run) block_partials_dpp<Epi::NR>(part, in.out_ws(ws), bid);
      in.arrive(run);
      return;
    }
  }
  if constexpr (Epi::NR > 0) block_partials_ex<Epi::NR>(part, ws, bid, g.nblk);
}

template <int CPL, class T, class LK, class In, class Epi, class Fin>
__global__ __launch_bounds__(kBlock) void k_eo(EoGeom eg, OpDev<LK> op, In in, T* __restrict__ out_arg, Epi epi,
                                               Fin, RedWs ws, Gate gate) {
  if (gate_closed(gate)) return;
  eo_body<CPL, T, LK, In, Epi>(eg, op, in, out_arg, epi, ws, blockIdx.x);
}

template <class T, class LK, class In, class Epi, class Fin>
void launch_eo(hipStream_t st, const EoGeom& eg, const OpDev<LK>& op, In in, T* out, Epi epi, Fin fin, RedWs ws,
               Gate gate, bool finalize = true) {
  hipLaunchKernelGGL((k_eo<normal_cpl<T>(), T, LK, In, Epi, Fin>), dim3(eg.g.nblk), dim3(kBlock), 0, st, eg, op, in,
                     out, epi, fin, ws, gate);
  GLX_CHECK_LAUNCH();
  if constexpr (Epi::NR > 0) {
    if (!finalize) return;
    hipLaunchKernelGGL((k_finalize<Epi::NR, Fin>), dim3(1), dim3(kBlock), 0, st, eg.g.nblk, fin, ws, gate);
    GLX_CHECK_LAUNCH();
  }
}

}  // namespace glx

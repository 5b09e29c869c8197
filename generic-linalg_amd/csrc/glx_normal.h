// glx_normal.h -- the fused normal operator D^dag D = m^2 - K^2 as a
// row-marching kernel (operators.cpp:444-453 computes it as D^dag (D p)
// through a heap temporary; here it is one pass over p and the links).
//
// A block of 4 waves owns a 64-column strip of sr rows (sr a multiple of 4).
// Lane l of every wave works on column x0 + l; the block marches down the
// strip four rows per step, wave w taking the step's w-th row.  LDS holds
// short row rings (8 rows) of the staged input p, the links and q = K p, so
// every vector element and link is read from memory once per strip (plus a
// two-row halo at each end of the strip and two columns on each side, which
// mostly hit L2).  Per step:
//   stage    p (through the caller's prologue) and the links of the step's
//            four new rows into the rings; the reads for the next step are
//            issued right after, so they are in flight during the compute
//   q        q = K p for four rows on columns x0-1 .. x0+64
//   out      m^2 p - K q for the four rows that now have their q neighbours,
//            fed to the caller's epilogue
// The row lag between the three phases (stage row y, q row y-1, out row y-2
// over a step of four rows) means a strip of sr rows takes sr/4 + 1 steps.
//
// Requirements (checked on the host by normal_fits): lx a multiple of 64,
// ly a multiple of 4 and at least 8, sr a multiple of 4 and at least 8.
// Other shapes use the tiled kernel of glx_stencil.h.
#pragma once

#include "glx_stencil.h"

namespace glx {

constexpr int kNX = 64;    // columns per block, one per lane
constexpr int kNRing = 8;  // LDS ring depth in rows

struct NGeom {
  int lx, ly;
  int ntx;     // column strips
  int sr;      // rows per block
  int ntiles;  // ntx * ceil(ly / sr)
  int nblk;    // launched blocks (multiple of 8, >= ntiles)
};

inline bool normal_fits(int lx, int ly, int sr) {
  return lx >= kNX && lx % kNX == 0 && ly >= 8 && ly % 4 == 0 && sr >= 8 && sr % 4 == 0;
}

inline NGeom make_ngeom(int lx, int ly, int sr) {
  NGeom g;
  g.lx = lx;
  g.ly = ly;
  g.ntx = lx / kNX;
  g.sr = sr;
  g.ntiles = g.ntx * ((ly + sr - 1) / sr);
  g.nblk = ((g.ntiles + 7) / 8) * 8;
  return g;
}

template <class T, class LK, class In, class Epi, class Fin>
__global__ __launch_bounds__(kBlock) void k_normal(NGeom g, OpDev<LK> op, In in, T* __restrict__ out_arg, Epi epi,
                                                   Fin fin, RedWs ws, Gate gate) {
  using C = typename LK::C;
  constexpr int PW = kNX + 4;  // p columns x0-2 .. x0+65
  constexpr int LW = kNX + 4;  // link columns x0-2 .. x0+64 (last slot unused)
  constexpr int QW = kNX + 2;  // q columns x0-1 .. x0+64
  constexpr int RM = kNRing - 1;

  if (gate_closed(gate)) return;
  in.init();
  T* __restrict__ out = in.target(out_arg);

  __shared__ T sP[kNRing][PW];
  __shared__ C sLx[kNRing][LW];
  __shared__ C sLy[kNRing][LW];
  __shared__ T sQ[kNRing][QW];
  __shared__ T sA[In::kAux ? kNRing : 1][In::kAux ? kNX : 1];

  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int bid = blockIdx.x;
  const int tile = xcd_tile(bid, g.nblk);
  double part[Epi::NR > 0 ? Epi::NR : 1];
#pragma unroll
  for (int k = 0; k < Epi::NR; k++) part[k] = 0.0;

  if (tile < g.ntiles) {
    const int lx = g.lx, ly = g.ly;
    const int x0 = (tile % g.ntx) * kNX;
    const int y0 = (tile / g.ntx) * g.sr;
    const int rows = min(g.sr, ly - y0);  // a multiple of 4
    const int nsteps = rows / 4 + 1;
    const double m2 = op.mass * op.mass;
    // halo roles: wave 0 lanes 0..15 stage p on columns x0-2, x0-1, x0+64,
    // x0+65 of the step's four rows; wave 1 lanes 0..11 stage the links on
    // columns x0-2, x0-1, x0+64
    const bool hp = w == 0 && l < 16;
    const bool hl = w == 1 && l < 12;
    const int hp_row = l >> 2, hp_c = (l & 3) < 2 ? (l & 3) : kNX + (l & 3);
    const int hl_row = l / 3, hl_c = (l % 3) < 2 ? (l % 3) : kNX + 2;
    const int xm = x0 + l;
    const int xhp = wrap_coord(x0 - 2 + hp_c, lx);
    const int xhl = wrap_coord(x0 - 2 + hl_c, lx);

    typename In::Raw rm, rh;
    typename LK::Raw lm, lh;
    auto fetch = [&](int k) {
      const int yr = 4 * k - 2 + w;
      const int j = wrap_coord(y0 + yr, ly) * lx + xm;
      rm = in.fetch(j, yr >= 0 && yr < rows);
      lm = op.links.fetch(j);
      if (hp) rh = in.fetch(wrap_coord(y0 + 4 * k - 2 + hp_row, ly) * lx + xhp, false);
      if (hl) lh = op.links.fetch(wrap_coord(y0 + 4 * k - 2 + hl_row, ly) * lx + xhl);
    };

    fetch(0);
    for (int k = 0; k < nsteps; k++) {
      if (k > 0) __syncthreads();  // step k-1's readers are done with the slots reused below
      {
        // stage rows 4k-2 .. 4k+1 (relative to y0)
        const int yr = 4 * k - 2 + w;
        const int j = wrap_coord(y0 + yr, ly) * lx + xm;
        const bool own = yr >= 0 && yr < rows;
        const int sl = (yr + 2) & RM;
        T aux;
        sP[sl][l + 2] = in.finish(rm, j, own, aux);
        if constexpr (In::kAux) {
          if (own) sA[sl][l] = aux;
        }
        C ux, uy;
        LK::finish(lm, ux, uy);
        sLx[sl][l + 2] = ux;
        sLy[sl][l + 2] = uy;
        if (hp) {
          const int yh = 4 * k - 2 + hp_row;
          T d;
          sP[(yh + 2) & RM][hp_c] = in.finish(rh, wrap_coord(y0 + yh, ly) * lx + xhp, false, d);
        }
        if (hl) {
          const int yh = 4 * k - 2 + hl_row;
          C hx, hy;
          LK::finish(lh, hx, hy);
          sLx[(yh + 2) & RM][hl_c] = hx;
          sLy[(yh + 2) & RM][hl_c] = hy;
        }
      }
      if (k + 1 < nsteps) fetch(k + 1);
      __syncthreads();
      {
        // q rows: step 0 rows -1, 0; step k > 0 rows 4k-3 .. 4k
        const int qa = k == 0 ? -1 : 4 * k - 3;
        const int yr = qa + w;
        if (yr <= 4 * k) {
          const int s0 = (yr + 2) & RM, sm = (yr + 1) & RM, sp = (yr + 3) & RM;
          // q column c is x = x0 - 1 + c; its p column is c + 1, link column c + 1
          auto qhop = [&](int c) {
            const int pc = c + 1;
            const bool eta_neg = (wrap_coord(x0 - 1 + c, lx) & 1) != 0;
            return khop<T, C>(sP[s0][pc + 1], sP[s0][pc - 1], sP[sp][pc], sP[sm][pc], sLx[s0][pc], sLx[s0][pc - 1],
                              sLy[s0][pc], sLy[sm][pc], eta_neg);
          };
          sQ[s0][l + 1] = qhop(l + 1);
          if (l < 2) {
            const int c = l == 0 ? 0 : kNX + 1;
            sQ[s0][c] = qhop(c);
          }
        }
      }
      __syncthreads();
      if (k > 0) {
        // output rows 4k-4 .. 4k-1
        const int yr = 4 * k - 4 + w;
        const int s0 = (yr + 2) & RM, sm = (yr + 1) & RM, sp = (yr + 3) & RM;
        const int qc = l + 1, lc = l + 2;
        const T kq = khop<T, C>(sQ[s0][qc + 1], sQ[s0][qc - 1], sQ[sp][qc], sQ[sm][qc], sLx[s0][lc], sLx[s0][lc - 1],
                                sLy[s0][lc], sLy[sm][lc], (xm & 1) != 0);
        const T pin = sP[s0][lc];
        const T v = rscal(m2, pin) - kq;
        const int i = (y0 + yr) * lx + xm;
        out[i] = v;
        if constexpr (In::kAux) epi.acc(part, i, pin, v, sA[s0][l]);
        else epi.acc(part, i, pin, v, pin);
      }
    }
  }
  if constexpr (Epi::NR > 0) block_partials<Epi::NR>(part, ws, bid);
}

template <class T, class LK, class In, class Epi, class Fin>
void launch_normal(hipStream_t st, const NGeom& g, const OpDev<LK>& op, In in, T* out, Epi epi, Fin fin, RedWs ws,
                   Gate gate) {
  hipLaunchKernelGGL((k_normal<T, LK, In, Epi, Fin>), dim3(g.nblk), dim3(kBlock), 0, st, g, op, in, out, epi, fin,
                     ws, gate);
  GLX_CHECK_LAUNCH();
  if constexpr (Epi::NR > 0) {
    hipLaunchKernelGGL((k_finalize<Epi::NR, Fin>), dim3(1), dim3(kBlock), 0, st, g.nblk, fin, ws, gate);
    GLX_CHECK_LAUNCH();
  }
}

}  // namespace glx

// glx_normal.h -- the fused normal operator D^dag D = m^2 - K^2 as a
// row-marching kernel (operators.cpp:444-453 computes it as D^dag (D p)
// through a heap temporary; here it is one pass over p and the links).
//
// A block of 4 waves owns a strip of 64 * cpl columns and sr rows (sr a
// multiple of 4), cpl = 2 columns per lane for complex FP32 (16-byte
// accesses per lane), 1 for complex FP64.  Lane l of every wave works on
// columns x0 + cpl l ...; the block marches down the strip four rows per
// step, wave w taking the step's w-th row.  LDS holds
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
// Requirements (checked on the host by normal_fits): lx a multiple of
// 64 * cpl, ly a multiple of 4 and at least 8, sr a multiple of 4 and at
// least 8.
// Other shapes use the tiled kernel of glx_stencil.h.
#pragma once

#include "glx_stencil.h"

#include <type_traits>

namespace glx {

constexpr int kNX = 64;    // columns per block per lane-column (lanes per wave)
constexpr int kNRing = 8;  // LDS ring depth in rows

// Columns per lane: complex FP32 vectors take two adjacent sites per lane, so
// every load and store moves 16 bytes per lane (the FP64 width); FP64 one.
template <class T>
__host__ __device__ constexpr int normal_cpl() {
  return sizeof(T) == 8 ? 2 : 1;
}

struct NGeom {
  int lx, ly;
  int ntx;     // column strips of kNX * cpl columns
  int sr;      // rows per block
  int ntiles;  // ntx * ceil(ly / sr)
  int nblk;    // launched blocks (multiple of 8, >= ntiles)
};

inline bool normal_fits(int lx, int ly, int sr, int cpl) {
  const int nx = kNX * cpl;
  return lx >= nx && lx % nx == 0 && ly >= 8 && ly % 4 == 0 && sr >= 8 && sr % 4 == 0;
}

inline NGeom make_ngeom(int lx, int ly, int sr, int cpl) {
  NGeom g;
  g.lx = lx;
  g.ly = ly;
  g.ntx = lx / (kNX * cpl);
  g.sr = sr;
  g.ntiles = g.ntx * ((ly + sr - 1) / sr);
  g.nblk = ((g.ntiles + 7) / 8) * 8;
  return g;
}

// Register staging of the CPL sites of one lane: the paired accesses when
// CPL = 2 and the prologue / link storage offers them, else one per site.
template <class X, class = void>
struct has_pair : std::false_type {};
template <class X>
struct has_pair<X, std::void_t<decltype(X::kPair)>> : std::integral_constant<bool, X::kPair> {};

template <class In, int CPL, bool PAIR = (CPL == 2 && has_pair<In>::value)>
struct InRegs {
  typename In::Raw r[CPL];
  __device__ __forceinline__ void fetch(const In& in, int j, bool own) {
#pragma unroll
    for (int c = 0; c < CPL; c++) r[c] = in.fetch(j + c, own);
  }
  template <class T>
  __device__ __forceinline__ void finish(const In& in, int j, bool own, T (&v)[CPL], T (&aux)[CPL]) const {
#pragma unroll
    for (int c = 0; c < CPL; c++) v[c] = in.finish(r[c], j + c, own, aux[c]);
  }
};
template <class In, int CPL>
struct InRegs<In, CPL, true> {
  typename In::RawPair r;
  __device__ __forceinline__ void fetch(const In& in, int j, bool own) { r = in.fetch_pair(j, own); }
  template <class T>
  __device__ __forceinline__ void finish(const In& in, int j, bool own, T (&v)[CPL], T (&aux)[CPL]) const {
    in.finish_pair(r, j, own, v[0], v[1], aux[0], aux[1]);
  }
};
template <class LK, int CPL, bool PAIR = (CPL == 2 && has_pair<LK>::value)>
struct LinkRegs {
  typename LK::Raw r[CPL];
  __device__ __forceinline__ void fetch(const LK& lk, int j) {
#pragma unroll
    for (int c = 0; c < CPL; c++) r[c] = lk.fetch(j + c);
  }
};
template <class LK, int CPL>
struct LinkRegs<LK, CPL, true> {
  typename LK::Raw r[CPL];
  __device__ __forceinline__ void fetch(const LK& lk, int j) { lk.fetch_pair(j, r[0], r[1]); }
};

// Host check that the paired accesses of a launch are aligned: the output
// and, for prologues with paired reads, their vectors.
template <class X, class = void>
struct has_aligned2 : std::false_type {};
template <class X>
struct has_aligned2<X, std::void_t<decltype(std::declval<const X&>().aligned2())>> : std::true_type {};
template <class T, class In>
bool normal_pairs_aligned(const In& in, const T* out) {
  if (normal_cpl<T>() == 1) return true;
  if (reinterpret_cast<uintptr_t>(out) % (2 * sizeof(T)) != 0) return false;
  if constexpr (has_pair<In>::value) {
    static_assert(has_aligned2<In>::value, "a prologue with paired reads must offer aligned2()");
    return in.aligned2();
  }
  return true;
}

// A prologue that finalises the previous launch's reduction itself
// (kSelfFin: prelude / out_ws / arrive hooks, see the single-pass CG).
template <class X, class = void>
struct has_selffin : std::false_type {};
template <class X>
struct has_selffin<X, std::void_t<decltype(X::kSelfFin)>> : std::integral_constant<bool, X::kSelfFin> {};

// One block's share of a launch: block `bid` of the g.nblk blocks over this
// lattice (a batched launch passes each system's local block index).
template <int CPL, class T, class LK, class In, class Epi>
__device__ __forceinline__ void normal_body(const NGeom& g, const OpDev<LK>& op, In& in, T* __restrict__ out_arg,
                                            const Epi& epi, RedWs ws, int bid) {
  using C = typename LK::C;
  constexpr int NXB = kNX * CPL;  // block columns
  constexpr int PW = NXB + 4;     // p columns x0-2 .. x0+NXB+1
  constexpr int LW = NXB + 4;     // link columns x0-2 .. x0+NXB (last slot unused)
  constexpr int QW = NXB + 2;     // q columns x0-1 .. x0+NXB
  constexpr int RM = kNRing - 1;

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
    const int lx = g.lx, ly = g.ly;
    const int x0 = (tt % g.ntx) * NXB;
    const int y0 = (tt / g.ntx) * g.sr;
    const int rows = min(g.sr, ly - y0);  // a multiple of 4
    const int nsteps = rows / 4 + 1;
    const double m2 = op.mass * op.mass;
    // halo roles: wave 0 lanes 0..15 stage p on columns x0-2, x0-1,
    // x0+NXB, x0+NXB+1 of the step's four rows; wave 1 lanes 0..11 stage the
    // links on columns x0-2, x0-1, x0+NXB
    const bool hp = w == 0 && l < 16;
    const bool hl = w == 1 && l < 12;
    const int hp_row = l >> 2, hp_c = (l & 3) < 2 ? (l & 3) : NXB + (l & 3);
    const int hl_row = l / 3, hl_c = (l % 3) < 2 ? (l % 3) : NXB + 2;
    const int xm = x0 + CPL * l;  // first of this lane's columns
    const int xhp = wrap_coord(x0 - 2 + hp_c, lx);
    const int xhl = wrap_coord(x0 - 2 + hl_c, lx);

    InRegs<In, CPL> rm;
    LinkRegs<LK, CPL> lm;
    typename In::Raw rh;
    typename LK::Raw lh;
    auto fetch = [&](int k) {
      const int yr = 4 * k - 2 + w;
      const int j = wrap_coord(y0 + yr, ly) * lx + xm;
      const bool own = yr >= 0 && yr < rows;
      rm.fetch(in, j, own);
      lm.fetch(op.links, j);
      if (hp) rh = in.fetch(wrap_coord(y0 + 4 * k - 2 + hp_row, ly) * lx + xhp, false);
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
        const int j = wrap_coord(y0 + yr, ly) * lx + xm;
        const bool own = yr >= 0 && yr < rows;
        const int sl = (yr + 2) & RM;
        T v[CPL], aux[CPL];
        rm.finish(in, j, own, v, aux);
#pragma unroll
        for (int c = 0; c < CPL; c++) {
          sP[sl][CPL * l + c + 2] = v[c];
          if constexpr (In::kAux) {
            if (own) sA[sl][CPL * l + c] = aux[c];
          }
          C ux, uy;
          LK::finish(lm.r[c], ux, uy);
          sLx[sl][CPL * l + c + 2] = ux;
          sLy[sl][CPL * l + c + 2] = uy;
        }
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
        // q column c is x = x0 - 1 + c; its p column is c + 1, link column
        // c + 1 (lx is even, so x's parity needs no wrap)
        auto qhop = [&](int yq, int c) {
          const int s0 = (yq + 2) & RM, sm = (yq + 1) & RM, sp = (yq + 3) & RM;
          const int pc = c + 1;
          const bool eta_neg = ((x0 - 1 + c) & 1) != 0;
          sQ[s0][c] = khop<T, C>(sP[s0][pc + 1], sP[s0][pc - 1], sP[sp][pc], sP[sm][pc], sLx[s0][pc],
                                 sLx[s0][pc - 1], sLy[s0][pc], sLy[sm][pc], eta_neg);
        };
        if (yr <= 4 * k) {
#pragma unroll
          for (int c = 0; c < CPL; c++) qhop(yr, CPL * l + c + 1);
        }
        // the two halo columns of all the step's q rows on wave 3 (a pass
        // with two active lanes per wave would cost every wave a full pass)
        if (w == 3 && l < 8) {
          const int yq = qa + (l >> 1);
          if (yq <= 4 * k) qhop(yq, (l & 1) == 0 ? 0 : NXB + 1);
        }
      }
      __syncthreads();
      if (k > 0) {
        // output rows 4k-4 .. 4k-1
        const int yr = 4 * k - 4 + w;
        const int s0 = (yr + 2) & RM, sm = (yr + 1) & RM, sp = (yr + 3) & RM;
        T v[CPL];
        const int i0 = (y0 + yr) * lx + xm;
#pragma unroll
        for (int c = 0; c < CPL; c++) {
          const int qc = CPL * l + c + 1, lc = CPL * l + c + 2;
          const T kq = khop<T, C>(sQ[s0][qc + 1], sQ[s0][qc - 1], sQ[sp][qc], sQ[sm][qc], sLx[s0][lc],
                                  sLx[s0][lc - 1], sLy[s0][lc], sLy[sm][lc], ((xm + c) & 1) != 0);
          const T pin = sP[s0][lc];
          v[c] = rscal(m2, pin) - kq;
          if constexpr (In::kAux) epi.acc(part, i0 + c, pin, v[c], sA[s0][CPL * l + c]);
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
__global__ __launch_bounds__(kBlock) void k_normal(NGeom g, OpDev<LK> op, In in, T* __restrict__ out_arg, Epi epi,
                                                   Fin, RedWs ws, Gate gate) {
  if (gate_closed(gate)) return;
  normal_body<CPL, T, LK, In, Epi>(g, op, in, out_arg, epi, ws, blockIdx.x);
}

template <class T, class LK, class In, class Epi, class Fin>
void launch_normal(hipStream_t st, const NGeom& g, const OpDev<LK>& op, In in, T* out, Epi epi, Fin fin, RedWs ws,
                   Gate gate, bool finalize = true) {
  hipLaunchKernelGGL((k_normal<normal_cpl<T>(), T, LK, In, Epi, Fin>), dim3(g.nblk), dim3(kBlock), 0, st, g, op, in, out, epi, fin,
                     ws, gate);
  GLX_CHECK_LAUNCH();
  if constexpr (Epi::NR > 0) {
    if (!finalize) return;
    hipLaunchKernelGGL((k_finalize<Epi::NR, Fin>), dim3(1), dim3(kBlock), 0, st, g.nblk, fin, ws, gate);
    GLX_CHECK_LAUNCH();
  }
}

}  // namespace glx

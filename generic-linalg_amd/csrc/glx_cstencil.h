// glx_cstencil.h -- stencil_2d for any number of colours Nc, with the optional
// two-link terms (coarse_stencil.h:14-154, coarse_stencil.cpp:12-1512).
//
// Output entry i = s Nc + r (site s, colour row r) is
//   sum over the applied terms t, in the reference's order (clover; hopping
//   +x, +y, -x, -y; two-link +2x, +x+y, +2y, -x+y, -2x, -x-y, -2y, +x-y), of
//   sum_c A_t(s)[r][c] in(s + off_t)[c]
// plus the identity / even-odd / top-bottom shifts for the full stencil.
//
// HBM layout.  The reference stores block (site, row, column) row-major per
// term, so a thread per output row would read Nc-strided addresses.  Here the
// coefficients are transposed once, when the stencil is uploaded, to
//   coef[((t Nc + c) V + s) Nc + r] = A_t(s)[r][c],
// so for every (t, c) the 64 threads of a wave (consecutive i) read 64
// consecutive 16-byte values: each coefficient is read exactly once, fully
// coalesced.  The Nc input values of a neighbour are shared by the Nc
// threads of a site and come from L1/L2.  Bytes per site and launch:
// n_terms Nc^2 16 (coefficients) + 2 Nc 16 (input, output).
#pragma once

#include "glx_common.h"

namespace glx {

// stencil_dir (coarse_stencil.h:14-29)
enum CsDir { CS_DIR_ALL = 0, CS_DIR_0 = 1, CS_DIR_TWO_FIRST = 6, CS_DIR_LAST = 13 };
// which part of the stencil a launch applies
enum CsPiece { CS_ALL = 0, CS_EO = 1, CS_OE = 2, CS_TB = 3, CS_BT = 4 };
// what the launch does with the stencil sum `acc` on the rows it applies to
enum CsComb {
  CS_PLAIN = 0,    // out = acc
  CS_AXMY = 1,     // out = a comb - acc (m^2 - D_eo D_oe, eo-prec prepare)
  CS_RECON = 2,    // applied rows: out = inv_a (comb - acc); other rows: out = comb2
};

constexpr int kCsTerms = 13;

// (dx, dy) of term t
__host__ __device__ __forceinline__ void cs_offset(int t, int& dx, int& dy) {
  switch (t) {
    case 1: dx = 1; dy = 0; return;
    case 2: dx = 0; dy = 1; return;
    case 3: dx = -1; dy = 0; return;
    case 4: dx = 0; dy = -1; return;
    case 5: dx = 2; dy = 0; return;
    case 6: dx = 1; dy = 1; return;
    case 7: dx = 0; dy = 2; return;
    case 8: dx = -1; dy = 1; return;
    case 9: dx = -2; dy = 0; return;
    case 10: dx = -1; dy = -1; return;
    case 11: dx = 0; dy = -2; return;
    case 12: dx = 1; dy = -1; return;
    default: dx = 0; dy = 0; return;
  }
}

// The terms a piece applies for the stencil's sdir (coarse_stencil.cpp: the
// DIR_ALL branches and the per-direction switch of each piece).
inline unsigned cs_terms(int piece, int sdir, bool has_two) {
  const bool eo = piece == CS_EO || piece == CS_OE;
  unsigned m = 0;
  for (int t = 0; t < kCsTerms; t++) {
    bool on;
    if (t >= 5 && !has_two) on = false;
    else if (sdir == CS_DIR_ALL) on = eo ? (t >= 1 && t <= 4) : true;
    else if (sdir == CS_DIR_0) on = !eo && t == 0;
    else if (sdir < CS_DIR_TWO_FIRST) on = t == sdir - 1;
    else on = !eo && t == sdir - 1;
    if (on) m |= 1u << t;
  }
  return m;
}
// A two-link direction without two-link terms leaves the output as it was for
// the full, tb and bt pieces (coarse_stencil.cpp:276-391 and the tb/bt copies).
inline bool cs_untouched(int piece, int sdir, bool has_two) {
  return sdir >= CS_DIR_TWO_FIRST && !has_two && (piece == CS_ALL || piece == CS_TB || piece == CS_BT);
}

struct CsDev {
  int lx, ly, nc, piece;
  unsigned terms;
  int c_lo, c_hi;
  int shifts;  // bit 0: shift, bit 1: eo_shift, bit 2: dof_shift (only the non-zero ones)
  c128 shift, eo_shift, dof_shift;
  const c128* coef;
  // combination (CsComb)
  int comb_mode;
  c128 comb_a;
  double comb_inv;
  const c128* comb;
  const c128* comb2;
};

inline CsDev make_csdev(int lx, int ly, int nc, int piece, int sdir, bool has_two, const c128* coef, c128 shift,
                        c128 eo_shift, c128 dof_shift) {
  CsDev d{};
  d.lx = lx;
  d.ly = ly;
  d.nc = nc;
  d.piece = piece;
  d.terms = cs_terms(piece, sdir, has_two);
  d.c_lo = piece == CS_TB ? nc / 2 : 0;
  d.c_hi = piece == CS_BT ? nc / 2 : nc;
  const bool sh = piece == CS_ALL && (sdir == CS_DIR_ALL || sdir == CS_DIR_0);
  auto nz = [](c128 z) { return z.re != 0.0 || z.im != 0.0; };
  d.shifts = sh ? ((nz(shift) ? 1 : 0) | (nz(eo_shift) ? 2 : 0) | (nz(dof_shift) ? 4 : 0)) : 0;
  d.shift = shift;
  d.eo_shift = eo_shift;
  d.dof_shift = dof_shift;
  d.coef = coef;
  d.comb_mode = CS_PLAIN;
  return d;
}

__device__ __forceinline__ int cs_wrap(int v, int l) {
  v %= l;
  return v < 0 ? v + l : v;
}

template <int = 0>
__global__ __launch_bounds__(kBlock) void k_cstencil(CsDev d, const c128* __restrict__ in, c128* __restrict__ out,
                                                     Gate gate) {
  if (gate_closed(gate)) return;
  const int nc = d.nc, lx = d.lx;
  const int n = lx * d.ly * nc;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    const int s = i / nc, r = i - s * nc;
    const int y = s / lx, x = s - y * lx;
    const bool even = ((x + y) & 1) == 0;
    bool on = true;
    if (d.piece == CS_EO) on = even;
    else if (d.piece == CS_OE) on = !even;
    else if (d.piece == CS_TB) on = r < nc / 2;
    else if (d.piece == CS_BT) on = r >= nc / 2;
    c128 acc{0.0, 0.0};
    if (on) {
      for (int t = 0; t < kCsTerms; t++) {
        if (!((d.terms >> t) & 1u)) continue;
        int dx, dy;
        cs_offset(t, dx, dy);
        const c128* __restrict__ src = in + size_t(cs_wrap(y + dy, d.ly) * lx + cs_wrap(x + dx, lx)) * nc;
        const c128* __restrict__ cf = d.coef + size_t(t) * nc * n + i;
#pragma unroll 4
        for (int c = d.c_lo; c < d.c_hi; c++) acc = acc + cf[size_t(c) * n] * src[c];
      }
      if (d.shifts) {
        const c128 v = in[i];
        if (d.shifts & 1) acc = acc + d.shift * v;
        if (d.shifts & 2) acc = acc + ((even ? 1.0 : -1.0) * d.eo_shift) * v;
        if (d.shifts & 4) acc = acc + ((r < nc / 2 ? 1.0 : -1.0) * d.dof_shift) * v;
      }
    }
    if (d.comb_mode == CS_AXMY) {
      if (on) acc = d.comb_a * d.comb[i] - acc;
    } else if (d.comb_mode == CS_RECON) {
      acc = on ? d.comb_inv * (d.comb[i] - acc) : d.comb2[i];
    }
    out[i] = acc;
  }
}

// Coefficients from the reference layout (term blocks of V Nc^2, row-major
// per site) into the column-major device layout above.
template <int = 0>
__global__ __launch_bounds__(kBlock) void k_cstencil_transpose(int v, int nc, int nterms,
                                                               const c128* __restrict__ src, c128* __restrict__ dst) {
  const size_t n = size_t(v) * nc, total = size_t(nterms) * nc * n;
  for (size_t e = size_t(blockIdx.x) * kBlock + threadIdx.x; e < total; e += size_t(gridDim.x) * kBlock) {
    // e = ((t nc + c) V + s) nc + r
    const size_t i = e % n;
    const size_t tc = e / n;
    const int c = int(tc % nc), t = int(tc / nc);
    const size_t s = i / nc, r = i % nc;
    dst[e] = src[((size_t(t) * v + s) * nc + r) * nc + c];
  }
}

}  // namespace glx

// glx_dwave.h -- the staggered operators D = m - K, D^dag = m + K
// (operators.cpp:184-239 and :372-441), gamma5 D (:316-368) and the
// even/odd pieces D_eo, D_oe (:456-539), and the fused D^dag D = m^2 - K^2
// (operators.cpp:444-453) as wave-march kernels: registers only, no LDS, no
// barriers.  The one-hop kinds first; D^dag D (k_ddwave) at the end of the
// file.
//
// A wave owns a strip of 64 lattice columns (lane l <-> column x0 + l) and a
// band of `rows` rows, and marches the band one row per step.  It holds the
// input rows y - 1, y, y + 1 and the links of rows y and y -/+ 1 in a ring of
// kDwSlots row records and keeps the loads of kDwAhead further rows in
// flight.  The x-neighbours come from the adjacent lanes by DPP wave shifts;
// the two columns beyond the strip (x0 - 1 for lane 0, x0 + 64 for lane 63)
// and lane 0's U_x(x0 - 1) come from one extra load each per row, in which
// every other lane reads its own site again (an L1 hit).  Odd bands march
// upwards, so the halo row at a band boundary is read by both bands near the
// same time (an L2 hit for the second).
//
// Per site: the input element and the two links read, the output written
// once (2 s_v + 2 s_u bytes); the halo rows, one per band end, mostly hit L2.
//
// The site arithmetic is the tiled kernel's (glx_stencil.h), term for term,
// so both kernels give the same result up to the compiler's contraction.
//
// Requirements (checked on the host by dwave_fits): lx a multiple of 64,
// ly >= 4; complex vectors; the plain operator (no prologue, epilogue or
// reduction).  Other shapes and the fused solver launches use the tiled
// kernel.  In is never out: waves read rows and columns that other waves
// write.
#pragma once

#include "glx_stencil.h"

#include <type_traits>
#include <utility>

namespace glx {

constexpr int kDwAhead = 3;             // rows whose loads are in flight beyond the three in use
constexpr int kDwSlots = kDwAhead + 3;  // row records in the ring (the march unrolls rounds of this many rows)
constexpr int kDwWaves = 4;             // waves per block

struct DwGeom {
  int lx, ly;
  int nstrip;  // lx / 64
  int rows;    // rows per band
  int ntask;   // nstrip * bands
  int nblk;    // launched blocks of kDwWaves waves (multiple of 8)
};

// (the march addresses sites by 32-bit byte offsets: at most 32 bytes per
// site, FP64 links, below 2^32)
inline bool dwave_fits(int lx, int ly) {
  return lx >= 64 && lx % 64 == 0 && ly >= 4 && long(lx) * ly <= (1L << 27);
}

inline DwGeom make_dwgeom(int lx, int ly, int rows) {
  DwGeom g;
  g.lx = lx;
  g.ly = ly;
  g.nstrip = lx / 64;
  g.rows = rows;
  g.ntask = g.nstrip * ((ly + rows - 1) / rows);
  const int blocks = (g.ntask + kDwWaves - 1) / kDwWaves;
  g.nblk = ((blocks + 7) / 8) * 8;
  return g;
}

// DPP wave shifts of a register value, word by word: lane l takes lane
// l + 1's value (kDwNext, "wave_shl:1") or lane l - 1's (kDwPrev,
// "wave_shr:1"); the lane with no source (63, resp. 0) keeps `old`.
constexpr int kDwNext = 0x130;
constexpr int kDwPrev = 0x138;
template <int CTRL, class V>
__device__ __forceinline__ V dw_shift(V old, V src) {
  static_assert(sizeof(V) % 4 == 0, "whole words only");
  constexpr int N = int(sizeof(V) / 4);
  int o[N], s[N], r[N];
  __builtin_memcpy(o, &old, sizeof(V));
  __builtin_memcpy(s, &src, sizeof(V));
#pragma unroll
  for (int i = 0; i < N; i++) r[i] = __builtin_amdgcn_update_dpp(o[i], s[i], CTRL, 0xf, 0xf, false);
  V out;
  __builtin_memcpy(&out, r, sizeof(V));
  return out;
}

template <class T, class LK>
struct DwRow {
  T v, ve;                   // the site's element; the edge element (lane 0: x0 - 1, lane 63: x0 + 64)
  typename LK::Raw lk, lke;  // the site's links; lane 0: those of x0 - 1
};

// f(K, t0 + K, tag) for K = 0 .. N - 1, K a compile-time constant
template <class F, class Tag, int... K>
__device__ __forceinline__ void dw_round(F& f, int t0, Tag tag, std::integer_sequence<int, K...>) {
  (f(std::integral_constant<int, K>{}, t0 + K, tag), ...);
}

// a lattice row index in [-1, ly] brought into [0, ly)
__device__ __forceinline__ int dw_wrap(int y, int ly) { return y < 0 ? y + ly : (y >= ly ? y - ly : y); }

template <int KIND, bool UP, class T, class LK>
__device__ __forceinline__ void dwave_march(const DwGeom& g, const OpDev<LK>& op, const T* __restrict__ in,
                                            T* __restrict__ out, int x, int xe, int y0, int rows, bool eta_neg) {
  using C = typename LK::C;
  using R = typename Real<T>::type;
  const int lx = g.lx, ly = g.ly;
  // march row m = 0 .. rows + 1 -> lattice row (downwards: y0 - 1 + m)
  auto row_of = [&](int m) { return dw_wrap(UP ? y0 + rows - m : y0 - 1 + m, ly); };
  // (32-bit byte offsets from the uniform bases: one address instruction per
  // access instead of a 64-bit multiply-add)
  auto at = [](const void* base, unsigned off) -> const char* { return static_cast<const char*>(base) + off; };
  auto load = [&](DwRow<T, LK>& d, int m) {
    const int y = row_of(min(m, rows + 1));  // (the rows past the band's end load its last row again)
    const unsigned j = unsigned(y * lx + x), je = unsigned(y * lx + xe);
    d.v = *reinterpret_cast<const T*>(at(in, j * unsigned(sizeof(T))));
    d.ve = *reinterpret_cast<const T*>(at(in, je * unsigned(sizeof(T))));
    if constexpr (std::is_same<LK, LinksF16>::value) {
      d.lk = op.links.fetch(int(j));
      d.lke = op.links.fetch(int(je));
    } else {
      d.lk = *reinterpret_cast<const typename LK::Raw*>(at(op.links.u, j * unsigned(sizeof(typename LK::Raw))));
      d.lke = *reinterpret_cast<const typename LK::Raw*>(at(op.links.u, je * unsigned(sizeof(typename LK::Raw))));
    }
  };
  DwRow<T, LK> ring[kDwSlots];
#pragma unroll
  for (int s = 0; s < kDwSlots - 1; s++) load(ring[s], s);
  // step t writes march row t + 1 from march rows t, t + 1, t + 2
  auto step = [&](auto kc, int t, auto tail) __attribute__((always_inline)) {
    constexpr int K = decltype(kc)::value;
    constexpr int S0 = K, S1 = (K + 1) % kDwSlots, S2 = (K + 2) % kDwSlots, SN = (K + kDwSlots - 1) % kDwSlots;
    // (the scheduling fences keep each step's arithmetic between its own
    // loads and the next step's: the compiler otherwise hoists the whole
    // round's arithmetic to its top, which waits for every row in flight)
    __builtin_amdgcn_sched_barrier(0);
    load(ring[SN], t + kDwSlots - 1);  // (the slot of march row t - 1, done with)
    const DwRow<T, LK>& a = ring[S0];
    const DwRow<T, LK>& b = ring[S1];
    const DwRow<T, LK>& c = ring[S2];
    C ux, uy, uxe, t0, uym;
    LK::finish(b.lk, ux, uy);
    LK::finish(b.lke, uxe, t0);
    LK::finish(UP ? c.lk : a.lk, t0, uym);  // U_y(y - 1)
    const T fyp = UP ? a.v : c.v, fym = UP ? c.v : a.v;
    const T fxp = dw_shift<kDwNext>(b.ve, b.v), fxm = dw_shift<kDwPrev>(b.ve, b.v);
    const C uxm = dw_shift<kDwPrev>(uxe, ux);
    T ax = umul(ux, fxp) - ucmul(uxm, fxm);
    T ay = umul(uy, fyp) - ucmul(uym, fym);
    if (eta_neg) ay = -ay;
    const T kp = R(0.5) * (ax + ay);
    T v;
    const int y = row_of(t + 1);
    const bool odd = ((x + y) & 1) != 0;
    if constexpr (KIND == OP_STAG) {
      v = rscal(op.mass, b.v) - kp;
    } else if constexpr (KIND == OP_STAG_DAG) {
      v = rscal(op.mass, b.v) + kp;
    } else if constexpr (KIND == OP_STAG_G5) {
      v = rscal(op.mass, b.v) - kp;
      if (odd) v = -v;
    } else if constexpr (KIND == OP_DEO) {
      v = odd ? zero_of<T>() : -kp;
    } else {
      v = odd ? -kp : zero_of<T>();
    }
    (void)odd;
    if (!decltype(tail)::value || t < rows)
      *reinterpret_cast<T*>(reinterpret_cast<char*>(out) + unsigned(row_of(t + 1) * lx + x) * unsigned(sizeof(T))) = v;
  };
  // whole rounds, every step storing its row; then the round that ends past
  // the band's last row, whose steps past it store nothing
  const int tfull = rows - rows % kDwSlots;
  for (int t0 = 0; t0 < tfull; t0 += kDwSlots)
    dw_round(step, t0, std::false_type{}, std::make_integer_sequence<int, kDwSlots>{});
  if (tfull < rows) dw_round(step, tfull, std::true_type{}, std::make_integer_sequence<int, kDwSlots>{});
}

// (amdgpu_waves_per_eu: without it the compiler aims at eight waves per SIMD,
// 64 registers, and moves the ring's loads down to their uses)
template <int KIND, class T, class LK>
__global__ __launch_bounds__(64 * kDwWaves) __attribute__((amdgpu_waves_per_eu(2))) void k_dwave(
    DwGeom g, OpDev<LK> op, const T* __restrict__ in, T* __restrict__ out, Gate gate) {
  if (gate_closed(gate)) return;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), l = threadIdx.x & 63;
  const int task = xcd_tile(blockIdx.x, g.nblk) * kDwWaves + w;
  if (task >= g.ntask) return;
  const int strip = task % g.nstrip, band = task / g.nstrip;
  const int x0 = strip * 64, x = x0 + l;
  const int xe = l == 0 ? (x0 == 0 ? g.lx - 1 : x0 - 1) : (l == 63 ? (x0 + 64 == g.lx ? 0 : x0 + 64) : x);
  const int y0 = band * g.rows;
  const int rows = min(g.rows, g.ly - y0);
  const bool eta_neg = (x & 1) != 0;
  if (band & 1) dwave_march<KIND, true>(g, op, in, out, x, xe, y0, rows, eta_neg);
  else dwave_march<KIND, false>(g, op, in, out, x, xe, y0, rows, eta_neg);
}

template <class T, class LK>
void launch_dwave(hipStream_t st, int kind, const DwGeom& g, const OpDev<LK>& op, const T* in, T* out, Gate gate) {
  const dim3 grid(g.nblk), block(64 * kDwWaves);
  switch (kind) {
    case OP_STAG:
      hipLaunchKernelGGL((k_dwave<OP_STAG, T, LK>), grid, block, 0, st, g, op, in, out, gate);
      break;
    case OP_STAG_DAG:
      hipLaunchKernelGGL((k_dwave<OP_STAG_DAG, T, LK>), grid, block, 0, st, g, op, in, out, gate);
      break;
    case OP_STAG_G5:
      hipLaunchKernelGGL((k_dwave<OP_STAG_G5, T, LK>), grid, block, 0, st, g, op, in, out, gate);
      break;
    case OP_DEO:
      hipLaunchKernelGGL((k_dwave<OP_DEO, T, LK>), grid, block, 0, st, g, op, in, out, gate);
      break;
    case OP_DOE:
      hipLaunchKernelGGL((k_dwave<OP_DOE, T, LK>), grid, block, 0, st, g, op, in, out, gate);
      break;
    default:
      throw ArgError("the wave-march kernel has no operator kind " + std::to_string(kind));
  }
  GLX_CHECK_LAUNCH();
}

// ------------------------------------------------------------ D^dag D --
// The fused normal operator D^dag D = m^2 - K^2 (operators.cpp:444-453), and
// m^2 - D_eo D_oe on full-lattice vectors (operators.cpp:566-588: the same
// march, odd outputs zero), in the same wave-march form: q = K p is formed one march row behind the loads
// and m^2 p - K q two rows behind, so a wave needs two halo lanes on each
// side and owns 60 columns (strips start at x0 - 2, x0 = 60 s); the halo
// columns are the neighbouring strips' own (L2 hits).  A band of `rows` rows
// marches rows + 4 lattice rows (two halo rows at each end).
constexpr int kDdOwn = 60;
constexpr int kDdSlots = 7;  // rows t - 3 .. t in use, three more in flight

struct DdGeom {
  int lx, ly;
  int nstrip;  // ceil(lx / 60)
  int rows;    // rows per band
  int ntask;   // nstrip * bands
  int nblk;    // launched blocks of kDwWaves waves (multiple of 8)
};

inline bool ddwave_fits(int lx, int ly) { return lx >= 64 && lx % 2 == 0 && ly >= 4; }

inline DdGeom make_ddgeom(int lx, int ly, int rows) {
  DdGeom g;
  g.lx = lx;
  g.ly = ly;
  g.nstrip = (lx + kDdOwn - 1) / kDdOwn;
  g.rows = rows;
  g.ntask = g.nstrip * ((ly + rows - 1) / rows);
  const int blocks = (g.ntask + kDwWaves - 1) / kDwWaves;
  g.nblk = ((blocks + 7) / 8) * 8;
  return g;
}

template <class T, class LK>
struct DdRow {
  T v;
  typename LK::Raw lk;
};

// a lattice row index in [-2, ly + 1] brought into [0, ly)
__device__ __forceinline__ int dd_wrap(int y, int ly) { return y < 0 ? y + ly : (y >= ly ? y - ly : y); }

template <int KIND, bool UP, class T, class LK>
__device__ __forceinline__ void ddwave_march(const DdGeom& g, const OpDev<LK>& op, const T* __restrict__ in,
                                             T* __restrict__ out, int x, bool own, int y0, int rows, bool eta_neg) {
  using C = typename LK::C;
  using R = typename Real<T>::type;
  const int lx = g.lx, ly = g.ly;
  const int last = rows + 3;  // march rows 0 .. rows + 3
  // march row m -> lattice row (downwards: y0 - 2 + m)
  auto row_of = [&](int m) { return dd_wrap(UP ? y0 + rows + 1 - m : y0 - 2 + m, ly); };
  auto load = [&](DdRow<T, LK>& d, int m) {
    const int j = row_of(min(m, last)) * lx + x;
    d.v = in[j];
    d.lk = op.links.fetch(j);
  };
  DdRow<T, LK> ring[kDdSlots];
  T q[kDdSlots];
#pragma unroll
  for (int s = 0; s < kDdSlots - 4; s++) load(ring[s], s);
  const R m2 = R(op.mass * op.mass);
  // K f at the lane's site from the rows above and below in march order:
  // fa, fb, fc = march rows m - 1, m, m + 1; links of row m (ux, uy) and
  // U_y of lattice row y - 1 (march row m - 1 downwards, m + 1 upwards)
  auto khop_m = [&](const T& fa, const T& fb, const T& fc, C ux, C uy, C uym) {
    const T fxp = dw_shift<kDwNext>(fb, fb), fxm = dw_shift<kDwPrev>(fb, fb);
    const C uxm = dw_shift<kDwPrev>(ux, ux);
    const T fyp = UP ? fa : fc, fym = UP ? fc : fa;
    T ax = umul(ux, fxp) - ucmul(uxm, fxm);
    T ay = umul(uy, fyp) - ucmul(uym, fym);
    if (eta_neg) ay = -ay;
    return R(0.5) * (ax + ay);
  };
  // step t: loads march row t + 3 into the slot of row t - 4; q of row t - 1
  // from p rows t - 2 .. t; out of row t - 2 from q rows t - 3 .. t - 1
  auto step = [&](auto kc, int t, auto phase) __attribute__((always_inline)) {
    constexpr int K = decltype(kc)::value;
    constexpr int PH = decltype(phase)::value;  // 0 first round, 1 whole, 2 last
    constexpr int S0 = K % kDdSlots, S1 = (K + kDdSlots - 1) % kDdSlots, S2 = (K + kDdSlots - 2) % kDdSlots,
                  S3 = (K + kDdSlots - 3) % kDdSlots, SN = (K + kDdSlots - 4) % kDdSlots;
    __builtin_amdgcn_sched_barrier(0);
    load(ring[SN], t + kDdSlots - 4);
    {
      C ux, uy, t0, uym;
      LK::finish(ring[S1].lk, ux, uy);
      LK::finish(UP ? ring[S0].lk : ring[S2].lk, t0, uym);
      q[S1] = khop_m(ring[S2].v, ring[S1].v, ring[S0].v, ux, uy, uym);
      // m^2 - D_eo D_oe: q is D_oe p, zero on the even sites (an even
      // site's neighbour across the wrap of an odd ly is even)
      if constexpr (KIND == OP_M2MDEODOE) {
        if (((x + row_of(t - 1)) & 1) == 0) q[S1] = zero_of<T>();
      }
    }
    if (PH != 0 || K >= 4) {
      C ux, uy, t0, uym;
      LK::finish(ring[S2].lk, ux, uy);
      LK::finish(UP ? ring[S1].lk : ring[S3].lk, t0, uym);
      const T kq = khop_m(q[S3], q[S2], q[S1], ux, uy, uym);
      T v = rscal(m2, ring[S2].v) - kq;
      // m^2 - D_eo D_oe: the even sites take m^2 p - K q, the odd ones 0
      if constexpr (KIND == OP_M2MDEODOE) {
        if (((x + row_of(t - 2)) & 1) != 0) v = zero_of<T>();
      }
      if ((PH == 1 || t <= last) && own) out[row_of(t - 2) * lx + x] = v;
    }
  };
  using std::integral_constant;
  dw_round(step, 0, integral_constant<int, 0>{}, std::make_integer_sequence<int, kDdSlots>{});
  const int tfull = (last + 1) - (last + 1) % kDdSlots;  // steps 0 .. last
  for (int t0 = kDdSlots; t0 < tfull; t0 += kDdSlots)
    dw_round(step, t0, integral_constant<int, 1>{}, std::make_integer_sequence<int, kDdSlots>{});
  if (tfull <= last && tfull >= kDdSlots)
    dw_round(step, tfull, integral_constant<int, 2>{}, std::make_integer_sequence<int, kDdSlots>{});
}

template <int KIND, class T, class LK>
__global__ __launch_bounds__(64 * kDwWaves) __attribute__((amdgpu_waves_per_eu(2))) void k_ddwave(
    DdGeom g, OpDev<LK> op, const T* __restrict__ in, T* __restrict__ out, Gate gate) {
  if (gate_closed(gate)) return;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), l = threadIdx.x & 63;
  const int task = xcd_tile(blockIdx.x, g.nblk) * kDwWaves + w;
  if (task >= g.ntask) return;
  const int strip = task % g.nstrip, band = task / g.nstrip;
  const int xo = strip * kDdOwn;  // first owned column
  const int xu = xo - 2 + l;      // this lane's column before wrapping
  const int x = xu < 0 ? xu + g.lx : (xu >= g.lx ? xu - g.lx : xu);
  const bool own = l >= 2 && l < 2 + kDdOwn && xu < g.lx;
  const int y0 = band * g.rows;
  const int rows = min(g.rows, g.ly - y0);
  const bool eta_neg = (x & 1) != 0;
  if (band & 1) ddwave_march<KIND, true>(g, op, in, out, x, own, y0, rows, eta_neg);
  else ddwave_march<KIND, false>(g, op, in, out, x, own, y0, rows, eta_neg);
}

template <class T, class LK>
void launch_ddwave(hipStream_t st, int kind, const DdGeom& g, const OpDev<LK>& op, const T* in, T* out,
                   Gate gate) {
  if (kind == OP_STAG_NORMAL)
    hipLaunchKernelGGL((k_ddwave<OP_STAG_NORMAL, T, LK>), dim3(g.nblk), dim3(64 * kDwWaves), 0, st, g, op, in, out,
                       gate);
  else if (kind == OP_M2MDEODOE)
    hipLaunchKernelGGL((k_ddwave<OP_M2MDEODOE, T, LK>), dim3(g.nblk), dim3(64 * kDwWaves), 0, st, g, op, in, out,
                       gate);
  else
    throw ArgError("the wave-march D^dag D kernel has no operator kind " + std::to_string(kind));
  GLX_CHECK_LAUNCH();
}

}  // namespace glx

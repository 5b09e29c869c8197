// glx_stencil.h -- the lattice operators of the solver path as one tiled HIP
// kernel template.
//
// Layout in HBM (identical to the reference's host layout, so no conversion
// is ever needed at the boundary):
//   vectors  lexicographic i = y*lx + x, element double / complex (re,im)
//   links    AoS u[2*i + mu], mu = 0 (x), 1 (y)
//
// One block of 256 threads owns a kTX x kTY = 64 x 8 tile.  Phase 0 stages
// the input on the tile plus an H-site halo (H = 1 for D, D^dag, Laplace;
// H = 2 for the fused normal operator) and the links on the tile plus an
// (H-1)-site halo into LDS with coalesced row loads; the input goes through
// a caller-supplied prologue functor, which is where solvers fuse their
// vector updates (e.g. CG's p = r + beta p).  Phase 1 (H = 2 only) forms
// q = K p on the tile plus a one-site halo in LDS.  Phase 2 writes the owned
// outputs and feeds them to an epilogue functor that accumulates per-thread
// FP64 partial sums; per-block sums are stored and a one-block kernel adds
// them in block order and hands the totals to a finaliser that runs on one
// device thread (the solver's scalar update).
//
// Staggered operator (operators.cpp:184-239): D = m - K,
//   K psi(x) = 1/2 sum_mu eta_mu(x) [U_mu(x) psi(x+mu) - U_mu(x-mu)^* psi(x-mu)],
//   eta_x = 1, eta_y(x) = (-1)^x;  D^dag = m + K (operators.cpp:372-441);
//   D^dag D = m^2 - K^2 (operators.cpp:444-453), computed here in one pass.
#pragma once

#include "glx_common.h"

#include <hip/hip_fp16.h>

namespace glx {

// Operator kinds; numbering shared with the C ABI (include/glx.h).
enum OpKind : int {
  OP_LAPLACE_REAL = 0,   // tests/multishift/multishift.cpp:634-672 (real)
  OP_LAPLACE = 1,        // operators.cpp:28-66
  OP_LAPLACE_U1 = 2,     // operators.cpp:73-120
  OP_STAG = 3,           // operators.cpp:184-239
  OP_STAG_DAG = 4,       // operators.cpp:372-441
  OP_STAG_NORMAL = 5,    // operators.cpp:444-453
  OP_STAG_G5 = 6,        // operators.cpp:316-368
  OP_GAMMA5 = 7,         // operators.cpp:242-259
  OP_DEO = 8,            // operators.cpp:456-496
  OP_DOE = 9,            // operators.cpp:499-539
  OP_M2MDEODOE = 10,     // operators.cpp:566-588
  OP_STENCIL = 11,       // coarse_stencil.cpp:12-172 (Nc = 1)
  OP_STENCIL_EO = 14,    // coarse_stencil.cpp:395-459
  OP_STENCIL_OE = 15,    // coarse_stencil.cpp:560-624
  OP_STENCIL_M2MDEODOE = 16,  // operators_stencil.cpp:195-214
  OP_STAG_EO_SCHUR = 17,      // m^2 - D_eo D_oe (operators.cpp:566-588) on compacted even-site vectors
  // stencil_2d for any Nc with two-link terms (glx_cstencil.h)
  OP_CST = 18,                // coarse_stencil.cpp:12-392
  OP_CST_EO = 19,             // coarse_stencil.cpp:395-557
  OP_CST_OE = 20,             // coarse_stencil.cpp:560-722
  OP_CST_TB = 21,             // coarse_stencil.cpp:725-1117
  OP_CST_BT = 22,             // coarse_stencil.cpp:1120-1512
  OP_CST_M2MDEODOE = 23,      // operators_stencil.cpp:195-214 on any Nc
};

__host__ __device__ constexpr int op_halo(int k) {
  return (k == OP_STAG_NORMAL || k == OP_M2MDEODOE || k == OP_STENCIL_M2MDEODOE) ? 2 : 1;
}
__host__ __device__ constexpr bool op_links(int k) {
  return k == OP_LAPLACE_U1 || k == OP_STAG || k == OP_STAG_DAG || k == OP_STAG_NORMAL || k == OP_STAG_G5 ||
         k == OP_DEO || k == OP_DOE || k == OP_M2MDEODOE || k == OP_STAG_EO_SCHUR;
}
__host__ __device__ constexpr bool is_cst_kind(int k) { return k >= OP_CST && k <= OP_CST_M2MDEODOE; }
// CsPiece of a general-stencil kind (the composite m^2 - D_eo D_oe has none)
__host__ __device__ constexpr int cst_piece(int k) { return k == OP_CST_M2MDEODOE ? -1 : k - OP_CST; }
__host__ __device__ constexpr bool op_coeffs(int k) {
  return k == OP_STENCIL || k == OP_STENCIL_EO || k == OP_STENCIL_OE || k == OP_STENCIL_M2MDEODOE;
}

// ---------------------------------------------------------- link storage --
// FP64 and FP32 links are stored as complex pairs per site; FP16 links as a
// half2 pair per site plus one FP32 scale per site (the largest component
// magnitude of the site's two links), U = scale * half(U / scale).
// Each storage type offers load(s, ux, uy) and the same split in two, so a
// kernel can issue the memory reads early and convert later: Raw, fetch(s)
// (reads only) and finish(raw, ux, uy) (arithmetic only).
struct LinksF64 {
  const c128* u;
  using C = c128;
  struct Raw {
    c128 a, b;
  };
  __device__ __forceinline__ Raw fetch(int s) const { return Raw{u[2 * s], u[2 * s + 1]}; }
  __device__ __forceinline__ static void finish(const Raw& r, C& ux, C& uy) {
    ux = r.a;
    uy = r.b;
  }
  __device__ __forceinline__ void load(int s, C& ux, C& uy) const { finish(fetch(s), ux, uy); }
};
struct LinksF32 {
  const c64* u;
  using C = c64;
  struct alignas(16) Raw {
    c64 a, b;
  };
  __device__ __forceinline__ Raw fetch(int s) const { return reinterpret_cast<const Raw*>(u)[s]; }
  __device__ __forceinline__ static void finish(const Raw& r, C& ux, C& uy) {
    ux = r.a;
    uy = r.b;
  }
  __device__ __forceinline__ void load(int s, C& ux, C& uy) const { finish(fetch(s), ux, uy); }
};
struct alignas(8) HalfLink {
  __half2 x, y;
};
struct LinksF16 {
  const HalfLink* u;
  const float* scale;
  using C = c64;
  struct Raw {
    HalfLink h;
    float sc;
  };
  __device__ __forceinline__ Raw fetch(int s) const { return Raw{u[s], scale[s]}; }
  // sites s and s + 1 (s even): one 16-byte and one 8-byte load
  static constexpr bool kPair = true;
  __device__ __forceinline__ void fetch_pair(int s, Raw& r0, Raw& r1) const {
    const Pair2<HalfLink> h = ld2(u, s);
    const Pair2<float> c = ld2(scale, s);
    r0 = Raw{h.a, c.a};
    r1 = Raw{h.b, c.b};
  }
  __device__ __forceinline__ static void finish(const Raw& r, C& ux, C& uy) {
    const float2 a = __half22float2(r.h.x), b = __half22float2(r.h.y);
    ux = {r.sc * a.x, r.sc * a.y};
    uy = {r.sc * b.x, r.sc * b.y};
  }
  __device__ __forceinline__ void load(int s, C& ux, C& uy) const { finish(fetch(s), ux, uy); }
};
struct LinksNone {
  using C = c64;
  struct Raw {};
  __device__ __forceinline__ Raw fetch(int) const { return Raw{}; }
  __device__ __forceinline__ static void finish(const Raw&, C&, C&) {}
  __device__ __forceinline__ void load(int, C&, C&) const {}
};

// Operator parameters, passed by value to kernels.
template <class LK>
struct OpDev {
  int lx, ly;
  double mass;        // staggered m, or the Laplacian's m^2 shift
  LK links;
  // stencil form (FP64)
  const c128* clover;
  const c128* hop;    // hop[d*V + i], d = 0:+x 1:+y 2:-x 3:-y
  c128 shift, eo_shift, dof_shift;
};

// ------------------------------------------------------------- prologue --
// A prologue produces the operator input at global site j; `own` is true for
// exactly one visit of each site of the lattice (the owning tile).
// init() runs once per thread before the first load (it may cache scalars
// read from the solver state).
//
// A prologue may also hand an auxiliary value of each owned site to the
// epilogue (kAux, written through `aux`), and choose the output vector
// (target), e.g. by the solver's iteration parity read in init().
//
// load(j, own, aux) is also offered split in two, so a kernel can issue the
// reads of a later row before it computes the current one: fetch(j, own)
// performs only the reads into a Raw record, finish(raw, j, own, aux) the
// arithmetic and the owned-site stores; load = finish(fetch).
//
// Optionally (kPair) a prologue offers the same for the two sites j, j + 1
// (j even, all its vectors aligned to two elements) with paired accesses:
// fetch_pair(j, own) -> RawPair, finish_pair(raw, j, own, v0, v1, aux0, aux1).
template <class T>
struct InPlain {
  static constexpr bool kAux = false;
  const T* p;
  struct Raw {
    T v;
  };
  __device__ __forceinline__ void init() {}
  __device__ __forceinline__ Raw fetch(int j, bool) const { return Raw{p[j]}; }
  __device__ __forceinline__ T finish(const Raw& r, int, bool, T&) const { return r.v; }
  static constexpr bool kPair = true;
  using RawPair = Pair2<T>;
  __device__ __forceinline__ RawPair fetch_pair(int j, bool) const { return ld2(p, j); }
  __device__ __forceinline__ void finish_pair(const RawPair& r, int, bool, T& v0, T& v1, T&, T&) const {
    v0 = r.a;
    v1 = r.b;
  }
  __host__ __device__ bool aligned2() const { return (reinterpret_cast<uintptr_t>(p) % (2 * sizeof(T))) == 0; }
  __device__ __forceinline__ T load(int j, bool, T&) const { return p[j]; }
  __device__ __forceinline__ T* target(T* out) const { return out; }
};

// ------------------------------------------------------------- epilogue --
// NR = number of FP64 partial sums; acc(v, i, in_i, out_i) adds a site's share.
template <class T>
struct EpiNone {
  static constexpr int NR = 0;
  __device__ __forceinline__ void acc(double*, int, T, T, T) const {}
};
// <in, out> = sum conj(in) out
template <class T>
struct EpiDotInOut {
  static constexpr int NR = is_cx<T>::value ? 2 : 1;
  __device__ __forceinline__ void acc(double* v, int, T in, T out, T) const {
    if constexpr (is_cx<T>::value) {
      c128 d = dotc(in, out);
      v[0] += d.re;
      v[1] += d.im;
    } else {
      v[0] += dotc(in, out);
    }
  }
};
// <aux, out>
template <class T>
struct EpiDotAuxOut {
  static constexpr int NR = is_cx<T>::value ? 2 : 1;
  const T* aux;
  __device__ __forceinline__ void acc(double* v, int i, T, T out, T) const {
    if constexpr (is_cx<T>::value) {
      c128 d = dotc(aux[i], out);
      v[0] += d.re;
      v[1] += d.im;
    } else {
      v[0] += dotc(aux[i], out);
    }
  }
};
// <out, in> and <out, out> (BiCGStab's omega numerator and denominator)
template <class T>
struct EpiOmega {
  static constexpr int NR = is_cx<T>::value ? 3 : 2;
  __device__ __forceinline__ void acc(double* v, int, T in, T out, T) const {
    if constexpr (is_cx<T>::value) {
      c128 d = dotc(out, in);
      v[0] += d.re;
      v[1] += d.im;
      v[2] += norm2(out);
    } else {
      v[0] += dotc(out, in);
      v[1] += norm2(out);
    }
  }
};

struct FinNone {
  __device__ void operator()(const double*) const {}
};

// ---------------------------------------------------------- site algebra --
// Apply U (or U^*) of compute type C to a vector element of type T.
template <class T, class C>
__device__ __forceinline__ T umul(C u, T v) {
  if constexpr (is_cx<T>::value) {
    using R = typename Real<T>::type;
    cx<R> uu{R(u.re), R(u.im)};
    return uu * v;
  } else {
    return v;
  }
}
template <class T, class C>
__device__ __forceinline__ T ucmul(C u, T v) {
  if constexpr (is_cx<T>::value) {
    using R = typename Real<T>::type;
    cx<R> uu{R(u.re), R(u.im)};
    return cmulc(uu, v);
  } else {
    return v;
  }
}

// K at one point from its four neighbours F(x+x), F(x-x), F(x+y), F(x-y) and
// the links U_x(x), U_x(x-x), U_y(x), U_y(x-y):
//   1/2 [U_x F(x+x) - U_x(x-x)^* F(x-x) + eta_y (U_y F(x+y) - U_y(x-y)^* F(x-y))]
template <class T, class C>
__device__ __forceinline__ T khop(T fxp, T fxm, T fyp, T fym, C ux, C uxm, C uy, C uym, bool eta_neg) {
  using R = typename Real<T>::type;
  T ax = umul(ux, fxp) - ucmul(uxm, fxm);
  T ay = umul(uy, fyp) - ucmul(uym, fym);
  if (eta_neg) ay = -ay;
  return R(0.5) * (ax + ay);
}

// K F at one point: F has row pitch fw, centre index fc; links region has row
// pitch lw, centre index lc (links stored as pairs ux, uy).
template <class T, class C>
__device__ __forceinline__ T hop_k(const T* F, int fw, int fc, const C* L, int lw, int lc, bool eta_neg) {
  return khop<T, C>(F[fc + 1], F[fc - 1], F[fc + fw], F[fc - fw], L[2 * lc], L[2 * (lc - 1)], L[2 * lc + 1],
                    L[2 * (lc - lw) + 1], eta_neg);
}

// ---------------------------------------------------------------- kernel --
struct Geom {
  int lx, ly;
  int ntx, nty;   // tiles per direction
  int ntiles;
  int nblk;       // launched blocks (multiple of 8, >= ntiles)
};

template <int KIND, class T, class LK, class In, class Epi, class Fin>
__global__ __launch_bounds__(kBlock) void k_stencil(Geom g, OpDev<LK> op, In in, T* __restrict__ out_arg, Epi epi,
                                                    Fin fin, RedWs ws, Gate gate) {
  constexpr int H = op_halo(KIND);
  constexpr bool LINKS = op_links(KIND);
  constexpr int PW = kTX + 2 * H, PH = kTY + 2 * H;  // staged input region
  constexpr int LW = PW - 1, LH = PH - 1;              // staged links region
  constexpr int QW = kTX + 2, QH = kTY + 2;            // q region (H = 2)
  using C = typename LK::C;
  using R = typename Real<T>::type;

  if (gate_closed(gate)) return;
  in.init();
  T* __restrict__ out = in.target(out_arg);

  __shared__ T sp[PH * PW];
  __shared__ T sa[In::kAux ? kTX * kTY : 1];
  __shared__ C su[LINKS ? 2 * LW * LH : 2];
  __shared__ T sq[H == 2 ? QW * QH : 1];

  const int bid = blockIdx.x;
  const int tile = xcd_tile(bid, g.nblk);
  const bool active = tile < g.ntiles;
  const int lx = g.lx, ly = g.ly;
  double part[Epi::NR > 0 ? Epi::NR : 1];
#pragma unroll
  for (int k = 0; k < Epi::NR; k++) part[k] = 0.0;

  if (active) {
    const int x0 = (tile % g.ntx) * kTX, y0 = (tile / g.ntx) * kTY;
    // phase 0: stage input and links
    for (int e = threadIdx.x; e < PW * PH; e += kBlock) {
      const int ry = e / PW, rx = e - ry * PW;
      const int ux = x0 + rx - H, uy = y0 + ry - H;
      const int gx = wrap_coord(ux, lx), gy = wrap_coord(uy, ly);
      const bool own = rx >= H && rx < H + kTX && ry >= H && ry < H + kTY && ux < lx && uy < ly;
      T aux;
      sp[e] = in.load(gy * lx + gx, own, aux);
      if constexpr (In::kAux) {
        if (own) sa[(ry - H) * kTX + (rx - H)] = aux;
      }
    }
    if constexpr (LINKS) {
      for (int e = threadIdx.x; e < LW * LH; e += kBlock) {
        const int ry = e / LW, rx = e - ry * LW;
        const int gx = wrap_coord(x0 + rx - H, lx), gy = wrap_coord(y0 + ry - H, ly);
        C a, b;
        op.links.load(gy * lx + gx, a, b);
        su[2 * e] = a;
        su[2 * e + 1] = b;
      }
    }
    __syncthreads();

    // phase 1: q = K p on the tile plus a one-site halo
    if constexpr (H == 2) {
      for (int e = threadIdx.x; e < QW * QH; e += kBlock) {
        const int qy = e / QW, qx = e - qy * QW;
        const int gx = wrap_coord(x0 + qx - 1, lx), gy = wrap_coord(y0 + qy - 1, ly);
        const int pc = (qy + 1) * PW + (qx + 1);
        T q;
        if constexpr (KIND == OP_STENCIL_M2MDEODOE) {
          // odd sites: oe(p) = hopping terms; even: 0
          if (((gx + gy) & 1) == 1) {
            const int s = gy * lx + gx, V = lx * ly;
            const T pxp = sp[pc + 1], pyp = sp[pc + PW], pxm = sp[pc - 1], pym = sp[pc - PW];
            if constexpr (is_cx<T>::value) {
              T acc = zero_of<T>();
              auto cf = [&](int d) { c128 h = op.hop[d * V + s]; return cx<R>{R(h.re), R(h.im)}; };
              acc = acc + cf(0) * pxp;
              acc = acc + cf(1) * pyp;
              acc = acc + cf(2) * pxm;
              acc = acc + cf(3) * pym;
              q = acc;
            } else {
              q = zero_of<T>();
            }
          } else {
            q = zero_of<T>();
          }
        } else {
          const int lc = (qy + 1) * LW + (qx + 1);
          q = hop_k<T, C>(sp, PW, pc, su, LW, lc, (gx & 1) != 0);
          // m^2 - D_eo D_oe: q is D_oe p, zero on the even sites (an even
          // site's neighbour across the wrap of an odd ly is even)
          if constexpr (KIND == OP_M2MDEODOE) {
            if (((gx + gy) & 1) == 0) q = zero_of<T>();
          }
        }
        sq[e] = q;
      }
      __syncthreads();
    }

    // phase 2: owned outputs
    for (int s = threadIdx.x; s < kTX * kTY; s += kBlock) {
      const int ty = s / kTX, tx = s - ty * kTX;
      const int gx = x0 + tx, gy = y0 + ty;
      if (gx >= lx || gy >= ly) continue;
      const int i = gy * lx + gx;
      const int pc = (ty + H) * PW + (tx + H);
      const T pin = sp[pc];
      const bool odd = ((gx + gy) & 1) != 0;
      const bool eta_neg = (gx & 1) != 0;
      T v;
      if constexpr (KIND == OP_LAPLACE_REAL || KIND == OP_LAPLACE) {
        // (4 + m2) p - sum of the four neighbours, reference term order
        T acc = zero_of<T>();
        acc = acc - sp[pc + 1];
        acc = acc - sp[pc - 1];
        acc = acc - sp[pc + PW];
        acc = acc - sp[pc - PW];
        v = acc + rscal(4.0 + op.mass, pin);
      } else if constexpr (KIND == OP_LAPLACE_U1) {
        const int lc = (ty + H) * LW + (tx + H);
        const C ux = su[2 * lc], uy = su[2 * lc + 1], uxm = su[2 * (lc - 1)], uym = su[2 * (lc - LW) + 1];
        T acc = zero_of<T>();
        acc = acc - umul(ux, sp[pc + 1]);
        acc = acc - ucmul(uxm, sp[pc - 1]);
        acc = acc - umul(uy, sp[pc + PW]);
        acc = acc - ucmul(uym, sp[pc - PW]);
        v = acc + rscal(4.0 + op.mass, pin);
      } else if constexpr (KIND == OP_STAG || KIND == OP_STAG_DAG || KIND == OP_STAG_G5 || KIND == OP_DEO ||
                           KIND == OP_DOE) {
        // links region shares the input region's origin (x0-H, y0-H); its
        // pitch is LW, so the link centre index differs from pc
        const int lc = (ty + H) * LW + (tx + H);
        T kp;
        {
          const T fxp = sp[pc + 1], fxm = sp[pc - 1], fyp = sp[pc + PW], fym = sp[pc - PW];
          const C ux = su[2 * lc], uy = su[2 * lc + 1], uxm = su[2 * (lc - 1)], uym = su[2 * (lc - LW) + 1];
          T ax = umul(ux, fxp) - ucmul(uxm, fxm);
          T ay = umul(uy, fyp) - ucmul(uym, fym);
          if (eta_neg) ay = -ay;
          kp = R(0.5) * (ax + ay);
        }
        if constexpr (KIND == OP_STAG) {
          v = rscal(op.mass, pin) - kp;
        } else if constexpr (KIND == OP_STAG_DAG) {
          v = rscal(op.mass, pin) + kp;
        } else if constexpr (KIND == OP_STAG_G5) {
          v = rscal(op.mass, pin) - kp;
          if (odd) v = -v;
        } else if constexpr (KIND == OP_DEO) {
          v = odd ? zero_of<T>() : -kp;
        } else {
          v = odd ? -kp : zero_of<T>();
        }
      } else if constexpr (KIND == OP_STAG_NORMAL || KIND == OP_M2MDEODOE) {
        const int lc = (ty + H) * LW + (tx + H);
        const int qc = (ty + 1) * QW + (tx + 1);
        const T fxp = sq[qc + 1], fxm = sq[qc - 1], fyp = sq[qc + QW], fym = sq[qc - QW];
        const C ux = su[2 * lc], uy = su[2 * lc + 1], uxm = su[2 * (lc - 1)], uym = su[2 * (lc - LW) + 1];
        T ax = umul(ux, fxp) - ucmul(uxm, fxm);
        T ay = umul(uy, fyp) - ucmul(uym, fym);
        if (eta_neg) ay = -ay;
        const T kq = R(0.5) * (ax + ay);
        v = rscal(op.mass * op.mass, pin) - kq;
        if constexpr (KIND == OP_M2MDEODOE) {
          if (odd) v = zero_of<T>();
        }
      } else if constexpr (KIND == OP_GAMMA5) {
        v = odd ? -pin : pin;
      } else if constexpr (KIND == OP_STENCIL || KIND == OP_STENCIL_EO || KIND == OP_STENCIL_OE) {
        if constexpr (is_cx<T>::value) {
          const int V = lx * ly;
          auto cf = [&](c128 h) { return cx<R>{R(h.re), R(h.im)}; };
          const bool run = (KIND == OP_STENCIL) || (KIND == OP_STENCIL_EO && !odd) || (KIND == OP_STENCIL_OE && odd);
          T acc = zero_of<T>();
          if (run) {
            if constexpr (KIND == OP_STENCIL) acc = acc + cf(op.clover[i]) * pin;
            acc = acc + cf(op.hop[i]) * sp[pc + 1];
            acc = acc + cf(op.hop[V + i]) * sp[pc + PW];
            acc = acc + cf(op.hop[2 * V + i]) * sp[pc - 1];
            acc = acc + cf(op.hop[3 * V + i]) * sp[pc - PW];
            if constexpr (KIND == OP_STENCIL) {
              if (op.shift.re != 0.0 || op.shift.im != 0.0) acc = acc + cf(op.shift) * pin;
              if (op.eo_shift.re != 0.0 || op.eo_shift.im != 0.0)
                acc = acc + (odd ? R(-1) : R(1)) * cf(op.eo_shift) * pin;
              if (op.dof_shift.re != 0.0 || op.dof_shift.im != 0.0) acc = acc - cf(op.dof_shift) * pin;
            }
          }
          v = acc;
        } else {
          v = zero_of<T>();
        }
      } else if constexpr (KIND == OP_STENCIL_M2MDEODOE) {
        if constexpr (is_cx<T>::value) {
          const int V = lx * ly;
          auto cf = [&](c128 h) { return cx<R>{R(h.re), R(h.im)}; };
          if (odd) {
            v = zero_of<T>();
          } else {
            const int qc = (ty + 1) * QW + (tx + 1);
            T acc = zero_of<T>();
            acc = acc + cf(op.hop[i]) * sq[qc + 1];
            acc = acc + cf(op.hop[V + i]) * sq[qc + QW];
            acc = acc + cf(op.hop[2 * V + i]) * sq[qc - 1];
            acc = acc + cf(op.hop[3 * V + i]) * sq[qc - QW];
            v = cf(op.shift * op.shift) * pin - acc;
          }
        } else {
          v = zero_of<T>();
        }
      }
      out[i] = v;
      if constexpr (In::kAux) epi.acc(part, i, pin, v, sa[s]);
      else epi.acc(part, i, pin, v, pin);
    }
  }
  if constexpr (Epi::NR > 0) block_partials<Epi::NR>(part, ws, bid);
}

inline Geom make_geom(int lx, int ly) {
  Geom g;
  g.lx = lx;
  g.ly = ly;
  g.ntx = (lx + kTX - 1) / kTX;
  g.nty = (ly + kTY - 1) / kTY;
  g.ntiles = g.ntx * g.nty;
  g.nblk = ((g.ntiles + 7) / 8) * 8;
  return g;
}

template <int KIND, class T, class LK, class In, class Epi, class Fin>
void launch_stencil(hipStream_t st, const Geom& g, const OpDev<LK>& op, In in, T* out, Epi epi, Fin fin, RedWs ws,
                    Gate gate, bool finalize = true) {
  hipLaunchKernelGGL((k_stencil<KIND, T, LK, In, Epi, Fin>), dim3(g.nblk), dim3(kBlock), 0, st, g, op, in, out, epi,
                     fin, ws, gate);
  GLX_CHECK_LAUNCH();
  if constexpr (Epi::NR > 0) {
    if (!finalize) return;
    hipLaunchKernelGGL((k_finalize<Epi::NR, Fin>), dim3(1), dim3(kBlock), 0, st, g.nblk, fin, ws, gate);
    GLX_CHECK_LAUNCH();
  }
}

}  // namespace glx

// glx_runtime.h -- host-side runtime of the device solvers: a caching device
// allocator, the per-solve workspace (reduction partials, scalar state,
// pinned mirror), the operator descriptor and its dispatch onto the stencil
// kernel template, and the generic pointwise kernel with an optional grid
// reduction.
#pragma once

#include "glx_cstencil.h"
#include "glx_dwave.h"
#include "glx_eo.h"

#include <functional>
#include <map>
#include <mutex>
#include <vector>

namespace glx {

// ------------------------------------------------------------- allocator --
// Grow-only cache of device allocations keyed by size; blocks are reused by
// later solves instead of going back to hipMalloc.
class DevicePool {
 public:
  static DevicePool& get();
  void* alloc(size_t bytes);
  void release(void* p);
  void trim();  // frees every cached block

 private:
  std::mutex mu_;
  std::multimap<size_t, void*> free_;
  std::map<void*, size_t> live_;
};

// ------------------------------------------------ allocation-free mode --
// While a thread holds a NoAllocScope, the device buffers of its solves are
// carved out of the scope's region (a caller-owned array, bump allocation,
// all of it reused by the next scope), its solves run on the scope's stream,
// and they create no stream, event, graph or pinned host memory.  Running
// out of the region is an error, never a fallback to the pool.
class NoAllocScope {
 public:
  NoAllocScope(void* base, size_t cap, hipStream_t stream);
  ~NoAllocScope();
  NoAllocScope(const NoAllocScope&) = delete;
  NoAllocScope& operator=(const NoAllocScope&) = delete;
  void* take(size_t bytes);
  hipStream_t stream() const { return stream_; }
  size_t used() const { return used_; }

 private:
  char* base_;
  size_t cap_, used_ = 0;
  hipStream_t stream_;
  NoAllocScope* prev_;
};
// the calling thread's scope, or nullptr
NoAllocScope* noalloc_scope();

// RAII device buffer from the pool (or from the thread's NoAllocScope).
class DevBuf {
 public:
  DevBuf() = default;
  explicit DevBuf(size_t bytes) : bytes_(bytes) {
    if (!bytes) return;
    if (NoAllocScope* s = noalloc_scope()) {
      p_ = s->take(bytes);
      scoped_ = true;
    } else {
      p_ = DevicePool::get().alloc(bytes);
    }
  }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p_(o.p_), bytes_(o.bytes_), scoped_(o.scoped_) { o.p_ = nullptr; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    reset();
    p_ = o.p_;
    bytes_ = o.bytes_;
    scoped_ = o.scoped_;
    o.p_ = nullptr;
    return *this;
  }
  ~DevBuf() { reset(); }
  void reset() {
    if (p_ && !scoped_) DevicePool::get().release(p_);
    p_ = nullptr;
  }
  template <class T>
  T* as() const { return static_cast<T*>(p_); }
  void* get() const { return p_; }
  size_t bytes() const { return bytes_; }

 private:
  void* p_ = nullptr;
  size_t bytes_ = 0;
  bool scoped_ = false;  // carved out of a NoAllocScope: nothing to release
};

// Upper bound on blocks of any reducing launch (pointwise kernels use at most
// kMaxPointBlocks; stencil launches use one block per 64x8 tile).
constexpr int kMaxPointBlocks = 2048;
constexpr int kMaxNR = 16;

constexpr Gate kNoGate{nullptr, 0};

// Reduction workspace for one solve: per-block partial sums.
struct RedSpace {
  DevBuf partials;
  void init(int max_blocks, hipStream_t) { partials = DevBuf(size_t(max_blocks) * kMaxNR * sizeof(double)); }
  RedWs ws() const { return RedWs{partials.as<double>()}; }
};

// -------------------------------------------------------------- operator --
// A device operator: kind + lattice + device-resident links (FP64, plus the
// FP32 / FP16 copies used by the mixed-precision solver) or stencil
// coefficients.  `host_cb` instead marks an unrecognised host callback.
struct DevOperator {
  int kind = -1;
  int lx = 0, ly = 0, nc = 1;
  double mass = 0.0;
  const c128* u64 = nullptr;
  const c64* u32 = nullptr;
  const HalfLink* u16 = nullptr;
  const float* sc16 = nullptr;
  const c128* clover = nullptr;
  const c128* hop = nullptr;
  c128 shift{0, 0}, eo_shift{0, 0}, dof_shift{0, 0};
  // general stencil_2d (kinds OP_CST*): column-major coefficients, the
  // stencil's sdir and whether it has two-link terms, and a vector of scratch
  // for the composite m^2 - D_eo D_oe
  const c128* cs_coef = nullptr;
  int cs_sdir = 0;
  bool cs_has_two = false;
  c128* cs_tmp = nullptr;
  // host callback fallback
  void (*host_cb)(double*, double*, void*) = nullptr;
  void* host_extra = nullptr;
  bool host_complex = true;
  int size_ = 0;

  // vector length (compacted even sites for the even/odd Schur operator)
  int size() const { return host_cb ? size_ : (kind == OP_STAG_EO_SCHUR ? lx * ly / 2 : lx * ly * nc); }
  bool is_host() const { return host_cb != nullptr; }
  bool tiled() const { return !host_cb && nc == 1 && !is_cst_kind(kind); }
};

enum LinkPrec : int { LINKS_F64 = 0, LINKS_F32 = 1, LINKS_F16 = 2 };

template <class LK>
inline OpDev<LK> make_opdev(const DevOperator& o, LK lk) {
  OpDev<LK> d;
  d.lx = o.lx;
  d.ly = o.ly;
  d.mass = o.mass;
  d.links = lk;
  d.clover = o.clover;
  d.hop = o.hop;
  d.shift = o.shift;
  d.eo_shift = o.eo_shift;
  d.dof_shift = o.dof_shift;
  return d;
}

template <class T>
constexpr bool kind_valid_for(int k) {
  if constexpr (is_cx<T>::value) return k != OP_LAPLACE_REAL;
  else return k == OP_LAPLACE_REAL;
}

// Launch shape of the row-marching normal-operator kernel: rows per block
// for a lattice size and vector element size in bytes (set by the autotuner
// or by hand; 0 = use the tiled kernel), with a default for sizes not in the
// table: a fixed value once set, else (-1) the fewest of 16/32/64/128 rows
// that keep the grid within 512 blocks of 64 lanes x (1 or 2) sites, else
// 256.  `cols` is the vectors' row length in sites when it differs from lx
// (the even/odd Schur operator's compact rows).
int normal_rows_for(int lx, int ly, int esize, int cols = 0);
// lx = 0: set the default; otherwise rows < 0 removes the size's entry
void set_normal_rows(int lx, int ly, int esize, int rows);
// the size's table entry, or -1 when it has none
int normal_rows_entry(int lx, int ly, int esize);

// Rows per wave of the wave-march D / D^dag kernel (glx_dwave.h) for a
// lattice size and vector element size in bytes; 0 = use the tiled kernel.
// set_dwave_rows: -1 by lattice size (the default), 0 the tiled kernel
// always, > 0 that many rows.
int dwave_rows_for(int lx, int ly, int esize);
void set_dwave_rows(int rows);
int ddwave_rows_for(int lx, int ly, int esize);
void set_ddwave_rows(int rows);

// Rows per block of the even/odd Schur kernel (compact rows of lx / 2 sites;
// "tiled" does not exist for this operator).
inline int eo_rows_for(int lx, int ly, int esize) {
  const int sr = normal_rows_for(lx, ly, esize, lx / 2);
  return sr > 0 ? sr : 16;
}

// Whether a launch goes to a row-marching kernel: D^dag D (glx_normal.h;
// otherwise the tiled kernel) or the compact even/odd Schur operator
// (glx_eo.h; it has no other kernel).  Complex vectors, a launch shape in the
// table that fits the lattice, and aligned paired accesses.
template <class T, class In>
bool normal_path(const DevOperator& o, const In& in, const T* out) {
  if constexpr (!is_cx<T>::value) {
    return false;
  } else {
    if (o.kind != OP_STAG_NORMAL && o.kind != OP_STAG_EO_SCHUR) return false;
    if (o.kind == OP_STAG_EO_SCHUR)
      return eo_fits(o.lx, o.ly, eo_rows_for(o.lx, o.ly, int(sizeof(T))), normal_cpl<T>()) &&
             normal_pairs_aligned<T>(in, out);
    const int sr = normal_rows_for(o.lx, o.ly, int(sizeof(T)));
    return sr > 0 && normal_fits(o.lx, o.ly, sr, normal_cpl<T>()) && normal_pairs_aligned<T>(in, out);
  }
}

// Launch the tiled stencil kernel for a runtime operator kind.
template <class T, class LK, class In, class Epi, class Fin>
void dispatch_kind(hipStream_t st, const DevOperator& o, const LK& lk, In in, T* out, Epi epi, Fin fin, RedWs ws,
                   Gate done, bool finalize = true) {
  const Geom g = make_geom(o.lx, o.ly);
  const OpDev<LK> od = make_opdev(o, lk);
  // the plain D^dag D: wave-march kernel where enabled
  if constexpr (is_cx<T>::value && std::is_same<In, InPlain<T>>::value && std::is_same<Epi, EpiNone<T>>::value &&
                std::is_same<Fin, FinNone>::value) {
    if ((o.kind == OP_STAG_NORMAL || o.kind == OP_M2MDEODOE) && ddwave_fits(o.lx, o.ly)) {
      const int rows = ddwave_rows_for(o.lx, o.ly, int(sizeof(T)));
      if (rows > 0) {
        launch_ddwave<T, LK>(st, o.kind, make_ddgeom(o.lx, o.ly, rows), od, in.p, out, done);
        return;
      }
    }
  }
  if (normal_path<T>(o, in, out)) {
    if constexpr (is_cx<T>::value) {
      const bool fin_launch = has_selffin<In>::value ? false : finalize;
      if (o.kind == OP_STAG_EO_SCHUR)
        launch_eo<T, LK, In, Epi, Fin>(st, make_eogeom(o.lx, o.ly, eo_rows_for(o.lx, o.ly, int(sizeof(T))),
                                                       normal_cpl<T>()),
                                       od, in, out, epi, fin, ws, done, fin_launch);
      else
        launch_normal<T, LK, In, Epi, Fin>(st, make_ngeom(o.lx, o.ly, normal_rows_for(o.lx, o.ly, int(sizeof(T))),
                                                          normal_cpl<T>()),
                                           od, in, out, epi, fin, ws, done, fin_launch);
      return;
    }
  }
  // the plain one-hop staggered kinds on lattices of whole 64-column strips:
  // wave-march kernel
  if constexpr (is_cx<T>::value && std::is_same<In, InPlain<T>>::value && std::is_same<Epi, EpiNone<T>>::value &&
                std::is_same<Fin, FinNone>::value) {
    if ((o.kind == OP_STAG || o.kind == OP_STAG_DAG || o.kind == OP_STAG_G5 || o.kind == OP_DEO ||
         o.kind == OP_DOE) &&
        dwave_fits(o.lx, o.ly)) {
      const int rows = dwave_rows_for(o.lx, o.ly, int(sizeof(T)));
      if (rows > 0) {
        launch_dwave<T, LK>(st, o.kind, make_dwgeom(o.lx, o.ly, rows), od, in.p, out, done);
        return;
      }
    }
  }
  if (o.kind == OP_STAG_EO_SCHUR)
    throw ArgError("the compact even/odd operator needs complex vectors, (lx/2) % 64 == 0 (FP64) or "
                   "(lx/2) % 128 == 0 (FP32), ly % 4 == 0, ly >= 8, and aligned vectors");
  if constexpr (has_selffin<In>::value) {
    throw ArgError("a self-finalising iteration needs the row-marching normal-operator kernel");
  } else {
#define GLX_CASE(K)                                                                     \
  case K:                                                                               \
    if constexpr (kind_valid_for<T>(K)) {                                               \
      launch_stencil<K, T, LK, In, Epi, Fin>(st, g, od, in, out, epi, fin, ws, done, finalize); \
      return;                                                                           \
    }                                                                                   \
    break;
  switch (o.kind) {
    GLX_CASE(OP_LAPLACE_REAL)
    GLX_CASE(OP_LAPLACE)
    GLX_CASE(OP_LAPLACE_U1)
    GLX_CASE(OP_STAG)
    GLX_CASE(OP_STAG_DAG)
    GLX_CASE(OP_STAG_NORMAL)
    GLX_CASE(OP_STAG_G5)
    GLX_CASE(OP_GAMMA5)
    GLX_CASE(OP_DEO)
    GLX_CASE(OP_DOE)
    GLX_CASE(OP_M2MDEODOE)
    GLX_CASE(OP_STENCIL)
    GLX_CASE(OP_STENCIL_EO)
    GLX_CASE(OP_STENCIL_OE)
    GLX_CASE(OP_STENCIL_M2MDEODOE)
    default:
      break;
  }
#undef GLX_CASE
    throw ArgError("operator kind " + std::to_string(o.kind) + " is not available for this element type");
  }
}

// Stencil launch with the link precision chosen for T: FP64 links for FP64
// vectors; FP32 or FP16 links for FP32 vectors.
template <class T, class In, class Epi, class Fin>
void run_stencil(hipStream_t st, const DevOperator& o, int link_prec, In in, T* out, Epi epi, Fin fin, RedWs ws,
                 Gate done, bool finalize = true) {
  if (op_links(o.kind)) {
    if constexpr (std::is_same<typename Real<T>::type, double>::value) {
      if (!o.u64) throw ArgError("operator has no FP64 links on the device");
      dispatch_kind<T>(st, o, LinksF64{o.u64}, in, out, epi, fin, ws, done, finalize);
    } else {
      if (link_prec == LINKS_F16) {
        if (!o.u16 || !o.sc16) throw ArgError("operator has no FP16 links on the device");
        dispatch_kind<T>(st, o, LinksF16{o.u16, o.sc16}, in, out, epi, fin, ws, done, finalize);
      } else {
        if (!o.u32) throw ArgError("operator has no FP32 links on the device");
        dispatch_kind<T>(st, o, LinksF32{o.u32}, in, out, epi, fin, ws, done, finalize);
      }
    }
  } else {
    if constexpr (std::is_same<typename Real<T>::type, double>::value) {
      dispatch_kind<T>(st, o, LinksF64{nullptr}, in, out, epi, fin, ws, done, finalize);
    } else {
      dispatch_kind<T>(st, o, LinksF32{nullptr}, in, out, epi, fin, ws, done, finalize);
    }
  }
}

// ------------------------------------------------------ pointwise kernel --
// f(i, part) updates element i and adds its share to the F::NR partial sums
// (part has exactly that many entries).
template <int NR, class F>
__device__ __forceinline__ void point_body(int n, F& f, RedWs ws, int lb, int nblk) {
  f.init();
  double part[NR > 0 ? NR : 1];
#pragma unroll
  for (int k = 0; k < NR; k++) part[k] = 0.0;
  const int stride = nblk * kBlock;
  for (int i = lb * kBlock + threadIdx.x; i < n; i += stride) f(i, part);
  if constexpr (NR > 0) block_partials_ex<NR>(part, ws, lb, nblk);
}
template <int NR, class F, class Fin>
__global__ __launch_bounds__(kBlock) void k_point(int n, F f, Fin fin, RedWs ws, Gate gate) {
  if (gate_closed(gate)) return;
  point_body<NR, F>(n, f, ws, blockIdx.x, gridDim.x);
}

inline int point_blocks(int n) {
  int b = (n + kBlock - 1) / kBlock;
  if (b > kMaxPointBlocks) b = kMaxPointBlocks;
  if (b < 1) b = 1;
  return b;
}

// General stencil_2d launch (glx_cstencil.h) for a runtime kind; the
// composite m^2 - D_eo D_oe is an oe launch into o.cs_tmp and an eo launch
// that combines (operators_stencil.cpp:195-214).
inline void launch_cst_one(hipStream_t st, const CsDev& d, const c128* in, c128* out, Gate done) {
  const int n = d.lx * d.ly * d.nc;
  hipLaunchKernelGGL(k_cstencil<0>, dim3(point_blocks(n)), dim3(kBlock), 0, st, d, in, out, done);
  GLX_CHECK_LAUNCH();
}
inline CsDev cst_dev(const DevOperator& o, int piece) {
  return make_csdev(o.lx, o.ly, o.nc, piece, o.cs_sdir, o.cs_has_two, o.cs_coef, o.shift, o.eo_shift, o.dof_shift);
}
inline void launch_cst(hipStream_t st, const DevOperator& o, c128* out, const c128* in, Gate done) {
  if (!o.cs_coef) throw ArgError("general stencil without coefficients on the device");
  if (o.kind == OP_CST_M2MDEODOE) {
    if (!o.cs_tmp) throw ArgError("m^2 - D_eo D_oe needs a scratch vector");
    launch_cst_one(st, cst_dev(o, CS_OE), in, o.cs_tmp, done);
    CsDev e = cst_dev(o, CS_EO);
    e.comb_mode = CS_AXMY;
    e.comb_a = o.shift * o.shift;
    e.comb = in;
    launch_cst_one(st, e, o.cs_tmp, out, done);
    return;
  }
  const int piece = cst_piece(o.kind);
  if (cs_untouched(piece, o.cs_sdir, o.cs_has_two)) return;
  launch_cst_one(st, cst_dev(o, piece), in, out, done);
}

// The number of partial sums is the functor's own F::NR, so a launch can
// never disagree with what the functor accumulates.
// Grid of a pointwise launch that is usually gated off (a smaller grid
// dispatches and retires faster when every block just reads the gate).
constexpr int kGatedBlocks = 512;
template <class F, class Fin>
void launch_point(hipStream_t st, int n, F f, Fin fin, RedWs ws, Gate done, int max_blocks = kMaxPointBlocks) {
  static_assert(F::NR >= 0 && F::NR <= kMaxNR, "too many partial sums");
  const int nb = std::min(point_blocks(n), std::max(1, max_blocks));
  hipLaunchKernelGGL((k_point<F::NR, F, Fin>), dim3(nb), dim3(kBlock), 0, st, n, f, fin, ws, done);
  GLX_CHECK_LAUNCH();
  if constexpr (F::NR > 0) {
    hipLaunchKernelGGL((k_finalize<F::NR, Fin>), dim3(1), dim3(kBlock), 0, st, nb, fin, ws, done);
    GLX_CHECK_LAUNCH();
  }
}

// ---------------------------------------------------- plain operator apply --
// out = A in on the device (no fusion).  For a host callback the vector goes
// through host memory; the callback is skipped when `done` is set on device.
void apply_operator(hipStream_t st, const DevOperator& o, int link_prec, void* out, const void* in, int elem_bytes,
                    RedSpace& red, Gate done);

// Minimum number of workgroups a stencil launch can use on this lattice.
inline int stencil_blocks(int lx, int ly) { return make_geom(lx, ly).nblk; }

}  // namespace glx

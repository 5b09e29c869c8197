// glx_solvers.hip -- device-resident Krylov solvers.
//
// CG              generic_cg.cpp:132-231 (f64), :278-382 (c128)
// BiCGStab        generic_bicgstab.cpp:22-158, :205-341
// GCR             generic_gcr.cpp:27-156, :202-333
// CG-M            generic_cg_m.cpp:23-310, :312-600
// Mixed CG        (no reference counterpart) FP32 Krylov + FP64 reliable updates
//
// Each iteration is a handful of kernels: operator applications are the tiled
// stencil kernel with the solver's vector update fused into its prologue and
// the next inner product fused into its epilogue; the remaining vector
// updates are pointwise kernels with fused norms.  All scalars (alpha, beta,
// zeta, ...) are computed on the device by the reduction finalisers and read
// by the next kernel; the host reads the 16-byte control word once per chunk.
#include "glx_solvers.h"
#include "glx_sstep.h"
#include "glx_sswave.h"

#include <tuple>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>

namespace glx {

// =========================================================== allocator ====
namespace {
struct NormalTable {
  std::mutex mu;
  int def = -1;  // -1: by block count
  std::map<std::tuple<int, int, int>, int> rows;
};
NormalTable& normal_table() {
  static NormalTable t;
  return t;
}
}  // namespace

int normal_rows_for(int lx, int ly, int esize, int cols) {
  {
    NormalTable& t = normal_table();
    std::lock_guard<std::mutex> g(t.mu);
    auto it = t.rows.find(std::make_tuple(lx, ly, esize));
    if (it != t.rows.end()) return it->second;
    if (t.def >= 0 || lx <= 0) return t.def;
  }
  // By block count: the self-finalising one-iteration kernels reduce every
  // block's partial sums in every block, a cost that grows with the square of
  // the block count (at 4096^2 FP64, 16 rows per block take 2x the time of
  // 128; FP32 256 rows 8% less than 128); two blocks per CU fill the chip.
  const int strips = std::max(1, (cols > 0 ? cols : lx) / (64 * (esize == 16 ? 1 : 2)));
  for (int r : {16, 32, 64, 128})
    if (strips * ((ly + r - 1) / r) <= 512) return r;
  return 256;
}
static std::atomic<int> g_dw_rows{-1};
// By lattice size: 16 rows per wave for complex FP32 and 18 for FP64 (at
// 1024^2, FP32 5.2 us at 16 rows against 5.6 at 18 and 5.9 at 8; FP64 11.1
// us at 18 against 11.5 at 16), FP64 16 where 18 rows give 2048 waves (4096^2:
// 195 us at 16 against 205 at 18), unless the lattice then fills fewer than
// 768 of the 1024 SIMDs: then the rows that give about one wave per SIMD, at
// least 4.
int dwave_rows_for(int lx, int ly, int esize) {
  const int forced = g_dw_rows.load();
  if (forced >= 0) return forced;
  const long nstrip = lx / 64;
  const int cap = esize == 16 && nstrip * ((ly + 17) / 18) < 2048 ? 18 : 16;
  if (nstrip * ((ly + cap - 1) / cap) >= 768) return cap;
  return std::max(4, int((ly * nstrip + 1023) / 1024));
}
// By lattice size: none (the row-marching kernel) where one application
// moves more than 256 MiB (4096^2, beyond the last-level cache: FP32 107.9 us
// against 105.9, FP64 210 against 202); else bands of 7 k - 4 rows, which
// march whole rounds of seven steps: 24 rows where that still gives 2048 waves
// (2048^2: FP32 21.5 us against 22.2-22.8 at 16-20 rows), else 10 where that
// gives 768 (1024^2: FP32 6.4 us, FP64 12.1, against 7.1-8.1 and 12.8-15.1
// at 6-20 rows), else about one wave per SIMD, at least 3 rows (512^2: 5
// rows, 4.2 us).
static std::atomic<int> g_dd_rows{-1};
int ddwave_rows_for(int lx, int ly, int esize) {
  const int forced = g_dd_rows.load();
  if (forced >= 0) return forced;
  if (long(lx) * ly * (esize == 16 ? 64 : 32) > (256L << 20)) return 0;
  const long nstrip = (lx + kDdOwn - 1) / kDdOwn;
  if (nstrip * ((ly + 23) / 24) >= 2048) return 24;
  if (nstrip * ((ly + 9) / 10) >= 768) return 10;
  return std::max(3, int((ly * nstrip + 1023) / 1024));
}
void set_ddwave_rows(int rows) { g_dd_rows.store(rows < 0 ? -1 : rows); }
void set_dwave_rows(int rows) {
  g_dw_rows.store(rows < 0 ? -1 : rows);
}
void set_normal_rows(int lx, int ly, int esize, int rows) {
  NormalTable& t = normal_table();
  std::lock_guard<std::mutex> g(t.mu);
  if (lx == 0) t.def = rows;
  else if (rows < 0) t.rows.erase(std::make_tuple(lx, ly, esize));
  else t.rows[std::make_tuple(lx, ly, esize)] = rows;
}
int normal_rows_entry(int lx, int ly, int esize) {
  NormalTable& t = normal_table();
  std::lock_guard<std::mutex> g(t.mu);
  auto it = t.rows.find(std::make_tuple(lx, ly, esize));
  return it != t.rows.end() ? it->second : -1;
}

namespace {
thread_local NoAllocScope* t_noalloc = nullptr;
}
NoAllocScope::NoAllocScope(void* base, size_t cap, hipStream_t stream)
    : base_(static_cast<char*>(base)), cap_(cap), stream_(stream), prev_(t_noalloc) {
  t_noalloc = this;
}
NoAllocScope::~NoAllocScope() { t_noalloc = prev_; }
void* NoAllocScope::take(size_t bytes) {
  const size_t at = (used_ + 255) & ~size_t(255);
  if (at > cap_ || bytes > cap_ - at)
    throw ArgError("the solve needs more than the " + std::to_string(cap_ >> 20) +
                   " MiB of scratch memory the allocation-free entry point has");
  used_ = at + bytes;
  return base_ + at;
}
NoAllocScope* noalloc_scope() { return t_noalloc; }

DevicePool& DevicePool::get() {
  static DevicePool* pool = new DevicePool;  // outlives every static user
  return *pool;
}
void* DevicePool::alloc(size_t bytes) {
  int dev = 0;
  GLX_HIP(hipGetDevice(&dev));
  bytes = (bytes + 255) & ~size_t(255);
  const size_t key = bytes * 64 + size_t(dev);
  {
    std::lock_guard<std::mutex> g(mu_);
    auto it = free_.find(key);
    if (it != free_.end()) {
      void* p = it->second;
      free_.erase(it);
      live_[p] = key;
      return p;
    }
  }
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, bytes);
  if (e != hipSuccess) {
    trim();
    GLX_HIP(hipMalloc(&p, bytes));
  }
  std::lock_guard<std::mutex> g(mu_);
  live_[p] = key;
  return p;
}
void DevicePool::release(void* p) {
  std::lock_guard<std::mutex> g(mu_);
  auto it = live_.find(p);
  if (it == live_.end()) return;
  free_.emplace(it->second, p);
  live_.erase(it);
}
void DevicePool::trim() {
  std::lock_guard<std::mutex> g(mu_);
  if (!free_.empty()) (void)hipDeviceSynchronize();
  for (auto& kv : free_) (void)hipFree(kv.second);
  free_.clear();
}

// ========================================================= solver state ===
constexpr int kMaxShift = 64;

struct SolveState {
  // control word (first 16 bytes, mirrored to the host)
  int k, done, kfinal, n_rem;
  int max_iter, resid_freq, worst_first, n_shift;
  double eps, bsqrt, rsq, rsqNew;
  double beta_r, resSq, maxrr, delta;  // maxrr: single-pass CG: max of |r|^2 since the last reliable update
  int reliable, pending, ko;  // pending: self-finalising single-pass CG; ko: s-step CG launches
  int n_xupd;  // multishift: shifts active before this iteration's check (their x still takes its update)
  c128 alpha, alpha_prev, beta, beta_prev, rho, omega, pap;
  double* hist;
  double* resvec;
  // host-mapped copy of `done` (single-pass CG): the kernels that change
  // `done` also store it here, so the host polls it without a copy command
  int* hdone;
  // multishift
  c128 zeta[kMaxShift], zeta_p[kMaxShift], alpha_s[kMaxShift], beta_s[kMaxShift];
  double shifts[kMaxShift];
  int mapping[kMaxShift];
  int vec[kMaxShift];
  // multishift: sums over the iterations of the shifts whose x_s (n_xupd)
  // and p_s (n_rem after the check) were updated -- the pass's algorithmic
  // bytes are 16 (3 upd_x + upd_p) per site
  long long upd_x, upd_p;
};

struct Mirror {
  int k, done, kfinal, n_rem;
};

template <class T>
__device__ __forceinline__ typename Scal<T>::type as_scal(c128 a) {
  if constexpr (is_cx<T>::value) return a;
  else return a.re;
}

// Pinned slots for the control-word read-back (per host thread).
struct PinnedSlots {
  Mirror* m = nullptr;
  int* hdone = nullptr;      // host address of the single-pass CG's done flag
  int* hdone_dev = nullptr;  // the same memory as the device addresses it
  hipEvent_t ev[2];
  // events of the single-pass CG's groups, recorded without the system-scope
  // fence: the done flag they guard is written to host memory by the kernels
  hipEvent_t ev_nf[2];
  PinnedSlots() {
    GLX_HIP(hipHostMalloc(reinterpret_cast<void**>(&m), 2 * sizeof(Mirror), hipHostMallocDefault));
    GLX_HIP(hipHostMalloc(reinterpret_cast<void**>(&hdone), 64, hipHostMallocMapped));
    GLX_HIP(hipHostGetDevicePointer(reinterpret_cast<void**>(&hdone_dev), hdone, 0));
    for (int i = 0; i < 2; i++)
      GLX_HIP(hipEventCreateWithFlags(&ev_nf[i], hipEventDisableTiming | hipEventDisableSystemFence));
    GLX_HIP(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
    GLX_HIP(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
  }
};
static PinnedSlots& pinned_slots() {
  static thread_local PinnedSlots* s = new PinnedSlots;
  return *s;
}
// the pinned slots of the calling thread; an allocation-free solve never
// reaches them (they are allocated on a thread's first use)
static PinnedSlots& pinned() {
  if (noalloc_scope()) throw std::logic_error("pinned host slots requested in allocation-free mode");
  return pinned_slots();
}

// Per-solve workspace: device state + reduction space.
struct Work {
  hipStream_t st;
  DevBuf state_buf;
  SolveState* ds;
  RedSpace red;
  DevBuf hist;
  Work(hipStream_t s, int max_blocks, int max_iter, bool want_hist) : st(s) {
    state_buf = DevBuf(sizeof(SolveState));
    ds = state_buf.as<SolveState>();
    red.init(std::max(max_blocks, kMaxPointBlocks), st);
    SolveState h;
    std::memset(&h, 0, sizeof(h));
    h.max_iter = max_iter;
    if (max_iter > 0) {
      hist = DevBuf(size_t(max_iter) * sizeof(double));
      h.hist = hist.as<double>();
    }
    (void)want_hist;
    upload(h);
  }
  void upload(const SolveState& h) {
    GLX_HIP(hipMemcpyAsync(ds, &h, sizeof(h), hipMemcpyHostToDevice, st));
    GLX_HIP(hipStreamSynchronize(st));
  }
  SolveState download() {
    SolveState h;
    GLX_HIP(hipMemcpyAsync(&h, ds, sizeof(h), hipMemcpyDeviceToHost, st));
    GLX_HIP(hipStreamSynchronize(st));
    return h;
  }
  Gate done() const { return Gate{&ds->done, 0}; }
  Gate paused() const { return Gate{&ds->done, 2}; }
  RedWs ws() const { return red.ws(); }
  std::vector<double> hist_host(int n) {
    std::vector<double> v(std::max(n, 0));
    if (n > 0) {
      GLX_HIP(hipMemcpyAsync(v.data(), hist.get(), sizeof(double) * n, hipMemcpyDeviceToHost, st));
      GLX_HIP(hipStreamSynchronize(st));
    }
    return v;
  }
};

// Enqueue iterations chunk by chunk, keeping one chunk queued while the
// control word of the previous one is inspected.  `enq(k)` enqueues the
// kernels of iteration k; `pause(m)` handles a device-requested pause
// (done == 2) and returns the next iteration index.
template <class Enq, class Pause>
Mirror run_loop(Work& w, int max_iter, int chunk, Enq enq, Pause pause) {
  PinnedSlots& pin = pinned();
  int next = 0;
  int cur = 0;
  bool inflight = false;
  Mirror m{0, 0, -1, 0};
  auto enqueue_chunk = [&](int slot) {
    const int end = std::min(max_iter, next + chunk);
    for (; next < end; next++) enq(next);
    GLX_HIP(hipMemcpyAsync(&pin.m[slot], w.ds, sizeof(Mirror), hipMemcpyDeviceToHost, w.st));
    GLX_HIP(hipEventRecord(pin.ev[slot], w.st));
  };
  if (max_iter <= 0) return m;
  enqueue_chunk(cur);
  inflight = true;
  for (;;) {
    const bool more = next < max_iter;
    if (more) enqueue_chunk(cur ^ 1);
    GLX_HIP(hipEventSynchronize(pin.ev[cur]));
    m = pin.m[cur];
    if (m.done == 1) break;
    if (m.done == 2) {
      // drain the queued chunk (its kernels returned at once), then resume
      if (more) GLX_HIP(hipEventSynchronize(pin.ev[cur ^ 1]));
      next = pause(m);
      if (next < 0) {
        m = pin.m[cur];
        break;
      }
      if (next <= m.kfinal) {
        (void)hipStreamSynchronize(w.st);
        throw HipError("solver state did not advance after a pause");
      }
      enqueue_chunk(cur);
      continue;
    }
    if (!more) {
      // every iteration is queued: the device stops at max_iter - 1 at the latest
      GLX_HIP(hipStreamSynchronize(w.st));
      GLX_HIP(hipMemcpy(&m, w.ds, sizeof(Mirror), hipMemcpyDeviceToHost));
      if (m.done == 2) {
        next = pause(m);
        if (next >= 0 && next <= m.kfinal) throw HipError("solver state did not advance after a pause");  // drained
        if (next >= 0 && next < max_iter) {
          enqueue_chunk(cur);
          continue;
        }
      }
      break;
    }
    cur ^= 1;
  }
  (void)inflight;
  GLX_HIP(hipStreamSynchronize(w.st));
  GLX_HIP(hipMemcpy(&m, w.ds, sizeof(Mirror), hipMemcpyDeviceToHost));
  return m;
}

// ================================================== generic functors ======
// Pointwise kernel that materialises a prologue over the whole vector (used
// when the operator cannot take a fused prologue: host callbacks, Nc > 1).
template <class T, class In>
struct PtFromIn {
  static constexpr int NR = 0;
  In in;
  __device__ void init() { in.init(); }
  __device__ void operator()(int i, double*) {
    T aux;
    (void)in.load(i, true, aux);
  }
};
template <class T, class Epi>
struct PtFromEpi {
  static constexpr int NR = Epi::NR;
  Epi epi;
  const T* in;
  const T* out;
  __device__ void init() {}
  __device__ void operator()(int i, double* part) { epi.acc(part, i, in[i], out[i], in[i]); }
};

// Nc > 1 Laplacian (colour index fastest), one thread per entry.
template <class T>
__global__ __launch_bounds__(kBlock) void k_laplace_nc(int lx, int ly, int nc, double m2, const T* __restrict__ in,
                                                       T* __restrict__ out, Gate gate) {
  if (gate_closed(gate)) return;
  const int n = lx * ly * nc;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    const int c = i % nc, s = i / nc, x = s % lx, y = s / lx;
    const int xp = x + 1 == lx ? 0 : x + 1, xm = x == 0 ? lx - 1 : x - 1;
    const int yp = y + 1 == ly ? 0 : y + 1, ym = y == 0 ? ly - 1 : y - 1;
    T acc = zero_of<T>();
    acc = acc - in[(y * lx + xp) * nc + c];
    acc = acc - in[(y * lx + xm) * nc + c];
    acc = acc - in[(yp * lx + x) * nc + c];
    acc = acc - in[(ym * lx + x) * nc + c];
    out[i] = acc + rscal(4.0 + m2, in[i]);
  }
}

template <class T>
void apply_typed(hipStream_t st, const DevOperator& o, int link_prec, T* out, const T* in, RedWs ws, Gate done) {
  if (o.is_host()) {
    const size_t bytes = size_t(o.size()) * sizeof(T);
    std::vector<char> hin(bytes), hout(bytes);
    int hdone = done.want;
    GLX_HIP(hipMemcpyAsync(hin.data(), in, bytes, hipMemcpyDeviceToHost, st));
    if (done.flag) GLX_HIP(hipMemcpyAsync(&hdone, done.flag, sizeof(int), hipMemcpyDeviceToHost, st));
    GLX_HIP(hipStreamSynchronize(st));
    if (hdone != done.want) return;
    o.host_cb(reinterpret_cast<double*>(hout.data()), reinterpret_cast<double*>(hin.data()), o.host_extra);
    GLX_HIP(hipMemcpyAsync(out, hout.data(), bytes, hipMemcpyHostToDevice, st));
    GLX_HIP(hipStreamSynchronize(st));
    return;
  }
  if (is_cst_kind(o.kind)) {
    if constexpr (std::is_same<T, c128>::value) {
      launch_cst(st, o, out, in, done);
      return;
    } else {
      throw ArgError("the general stencil takes FP64 complex vectors");
    }
  }
  if (o.nc != 1) {
    if (o.kind != OP_LAPLACE && o.kind != OP_LAPLACE_REAL)
      throw ArgError("Nc > 1 is supported for the Laplacians only");
    hipLaunchKernelGGL(k_laplace_nc<T>, dim3(point_blocks(o.size())), dim3(kBlock), 0, st, o.lx, o.ly, o.nc, o.mass,
                       in, out, done);
    GLX_CHECK_LAUNCH();
    return;
  }
  run_stencil<T>(st, o, link_prec, InPlain<T>{in}, out, EpiNone<T>{}, FinNone{}, ws, done);
}

// Operator application with a prologue and an epilogue reduction: one fused
// stencil launch when the operator is tiled, else prologue kernel + apply +
// reduction kernel.  `mat` is the vector the prologue materialises.
template <class T, class In, class Epi, class Fin>
void run_fused(Work& w, const DevOperator& o, int link_prec, In in, const T* mat, T* out, Epi epi, Fin fin) {
  if (o.tiled()) {
    run_stencil<T>(w.st, o, link_prec, in, out, epi, fin, w.ws(), w.done());
    return;
  }
  launch_point(w.st, o.size(), PtFromIn<T, In>{in}, FinNone{}, w.ws(), w.done());
  apply_typed<T>(w.st, o, link_prec, out, mat, w.ws(), w.done());
  if constexpr (Epi::NR > 0) {
    launch_point(w.st, o.size(), PtFromEpi<T, Epi>{epi, mat, out}, fin, w.ws(), w.done());
  }
}

// ------------------------------------------------------- shared pieces ----
// ||a - b||^2 into s->resSq (or resvec[idx] when idx >= 0)
template <class T>
struct PtDiffNorm {
  static constexpr int NR = 1;
  const T* a;
  const T* b;
  __device__ void init() {}
  __device__ void operator()(int i, double* part) {
    const T d = a[i] - b[i];
    part[0] += norm2(d);
  }
};
struct FinResSq {
  SolveState* s;
  int idx;
  __device__ void operator()(const double* v) const {
    if (idx < 0) s->resSq = v[0];
    else s->resvec[idx] = v[0];
  }
};
// y += a x with a = state alpha (complex or real)
template <class T>
struct PtAxpyAlpha {
  static constexpr int NR = 0;
  T* y;
  const T* x;
  const SolveState* s;
  typename Scal<T>::type a;
  __device__ void init() { a = as_scal<T>(s->alpha); }
  __device__ void operator()(int i, double*) { y[i] = y[i] + scal(a, x[i]); }
};

template <class T>
double true_residual(Work& w, const DevOperator& o, T* x, const T* b, T* tmp) {
  apply_typed<T>(w.st, o, LINKS_F64, tmp, x, w.ws(), kNoGate);
  launch_point(w.st, o.size(), PtDiffNorm<T>{tmp, b}, FinResSq{w.ds, -1}, w.ws(), kNoGate);
  return 0.0;
}

// ================================================================== CG ====
template <class T>
struct PtCgInit {
  static constexpr int NR = 2;
  T* r;
  const T* b;
  const T* ax;
  __device__ void init() {}
  __device__ void operator()(int i, double* part) {
    const T bi = b[i];
    const T ri = bi - ax[i];
    r[i] = ri;
    part[0] += norm2(bi);
    part[1] += norm2(ri);
  }
};
struct FinCgInit {
  SolveState* s;
  __device__ void operator()(const double* v) const {
    s->bsqrt = sqrt(v[0]);
    s->rsq = v[1];
  }
};
// p_new = r + beta p_old (all visited sites); owned: store p_new and
// x += alpha_prev p_old (the previous iteration's solution update).
template <class T, class TX>
struct InCG {
  static constexpr bool kAux = false;
  const T* r;
  const T* pold;
  T* pnew;
  TX* x;
  const SolveState* s;
  typename Scal<T>::type ap;
  double beta;
  __device__ void init() {
    beta = s->beta_r;
    ap = as_scal<T>(s->alpha_prev);
  }
  __device__ T* target(T* out) const { return out; }
  struct Raw {
    T r, po;
    TX xv;
  };
  __device__ Raw fetch(int j, bool own) const {
    Raw w;
    w.po = pold[j];
    w.r = r[j];
    w.xv = (own && x) ? x[j] : zero_of<TX>();
    return w;
  }
  __device__ T finish(const Raw& w, int j, bool own, T&) const {
    const T pn = w.r + rscal(beta, w.po);
    if (own) {
      pnew[j] = pn;
      if (x) {
        if constexpr (std::is_same<T, TX>::value) {
          x[j] = w.xv + scal(ap, w.po);
        } else {
          TX a;
          cvt(a, scal(ap, w.po));
          x[j] = w.xv + a;
        }
      }
    }
    return pn;
  }
  __device__ T load(int j, bool own, T& aux) const { return finish(fetch(j, own), j, own, aux); }
};
template <class T>
struct FinCgAlpha {
  SolveState* s;
  __device__ void operator()(const double* v) const {
    if constexpr (is_cx<T>::value) {
      s->pap = {v[0], v[1]};
      s->alpha = cdiv(c128{s->rsq, 0.0}, s->pap);
    } else {
      s->pap = {v[0], 0.0};
      s->alpha = {s->rsq / v[0], 0.0};
    }
  }
};
template <class T>
struct PtCgR {
  static constexpr int NR = 1;
  T* r;
  const T* Ap;
  const SolveState* s;
  typename Scal<T>::type a;
  __device__ void init() { a = as_scal<T>(s->alpha); }
  __device__ void operator()(int i, double* part) {
    const T v = r[i] - scal(a, Ap[i]);
    r[i] = v;
    part[0] += norm2(v);
  }
};
struct FinCgNorm {
  SolveState* s;
  __device__ void operator()(const double* v) const {
    const double rn = v[0];
    const int k = s->k;
    if (s->hist) s->hist[k] = sqrt(rn) / s->bsqrt;
    s->rsqNew = rn;
    if (sqrt(rn) < s->eps * s->bsqrt || k == s->max_iter - 1) {
      s->kfinal = k;
      s->done = 1;
    } else {
      s->beta_r = rn / s->rsq;
      s->rsq = rn;
      s->alpha_prev = s->alpha;
      s->k = k + 1;
    }
  }
};

template <class T>
SolveResult cg_impl(const DevOperator& o, T* x, const T* b, const SolveOpts& opt) {
  const int n = o.size();
  Work w(opt.stream, stencil_blocks(std::max(o.lx, 1), std::max(o.ly, 1)), opt.max_iter, opt.want_hist);
  SolveState h;
  std::memset(&h, 0, sizeof(h));
  h.max_iter = opt.max_iter;
  h.eps = opt.eps;
  h.hist = w.hist.as<double>();
  w.upload(h);
  DevBuf rb(n * sizeof(T)), pb0(n * sizeof(T)), pb1(n * sizeof(T)), apb(n * sizeof(T));
  T* r = rb.as<T>();
  T* p[2] = {pb0.as<T>(), pb1.as<T>()};
  T* Ap = apb.as<T>();
  GLX_HIP(hipMemsetAsync(p[0], 0, n * sizeof(T), w.st));
  // r = b - A x; ||b||, ||r||^2
  apply_typed<T>(w.st, o, LINKS_F64, Ap, x, w.ws(), kNoGate);
  launch_point(w.st, n, PtCgInit<T>{r, b, Ap}, FinCgInit{w.ds}, w.ws(), kNoGate);

  auto enq = [&](int k) {
    T* pold = p[k & 1];
    T* pnew = p[(k + 1) & 1];
    run_fused<T>(w, o, LINKS_F64, InCG<T, T>{r, pold, pnew, x, w.ds}, pnew, Ap, EpiDotInOut<T>{}, FinCgAlpha<T>{w.ds});
    launch_point(w.st, n, PtCgR<T>{r, Ap, w.ds}, FinCgNorm{w.ds}, w.ws(), w.done());
  };
  Mirror m = run_loop(w, opt.max_iter, opt.chunk, enq, [](const Mirror&) { return -1; });

  SolveResult res;
  res.name = "CG";
  int kf = opt.max_iter > 0 ? m.kfinal : 0;
  if (opt.max_iter > 0) {
    // the last iteration's x += alpha p (deferred into the next prologue)
    launch_point(w.st, n, PtAxpyAlpha<T>{x, p[(kf + 1) & 1], w.ds}, FinNone{}, w.ws(), kNoGate);
  }
  true_residual<T>(w, o, x, b, Ap);
  SolveState hs = w.download();
  res.success = opt.max_iter > 0 ? (kf != opt.max_iter - 1) : true;
  res.iter = kf + 1;
  res.ops_count = res.iter + 2;
  res.resSq = hs.resSq;
  res.bsqrt = hs.bsqrt;
  if (opt.want_hist && opt.max_iter > 0) res.relres_hist = w.hist_host(res.iter);
  return res;
}

// ============================================================ BiCGStab ====
template <class T>
struct PtBicgInit {
  static constexpr int NR = (is_cx<T>::value ? 3 : 2);
  T* r;
  T* r0;
  T* p;
  const T* b;
  const T* ax;
  __device__ void init() {}
  __device__ void operator()(int i, double* part) {
    const T bi = b[i];
    const T ri = bi - ax[i];
    r[i] = ri;
    r0[i] = ri;
    p[i] = ri;
    part[0] += norm2(bi);
    if constexpr (is_cx<T>::value) {
      const c128 d = dotc(ri, ri);
      part[1] += d.re;
      part[2] += d.im;
    } else {
      part[1] += dotc(ri, ri);
    }
  }
};
template <class T>
struct FinBicgInit {
  SolveState* s;
  __device__ void operator()(const double* v) const {
    s->bsqrt = sqrt(v[0]);
    if constexpr (is_cx<T>::value) s->rho = {v[1], v[2]};
    else s->rho = {v[1], 0.0};
  }
};
template <class T>
struct FinBicgAlpha {
  SolveState* s;
  __device__ void operator()(const double* v) const {
    if constexpr (is_cx<T>::value) s->alpha = cdiv(s->rho, c128{v[0], v[1]});
    else s->alpha = {s->rho.re / v[0], 0.0};
  }
};
// s = r - alpha Ap
template <class T>
struct InBicgS {
  static constexpr bool kAux = false;
  const T* r;
  const T* Ap;
  T* sv;
  const SolveState* st;
  typename Scal<T>::type a;
  __device__ void init() { a = as_scal<T>(st->alpha); }
  __device__ T* target(T* out) const { return out; }
  struct Raw {
    T r, ap;
  };
  __device__ Raw fetch(int j, bool) const { return Raw{r[j], Ap[j]}; }
  __device__ T finish(const Raw& w, int j, bool own, T&) const {
    const T v = w.r - scal(a, w.ap);
    if (own) sv[j] = v;
    return v;
  }
  __device__ T load(int j, bool own, T& aux) const { return finish(fetch(j, own), j, own, aux); }
};
template <class T>
struct FinBicgOmega {
  SolveState* s;
  __device__ void operator()(const double* v) const {
    if constexpr (is_cx<T>::value) s->omega = cdiv(c128{v[0], v[1]}, c128{v[2], 0.0});
    else s->omega = {v[0] / v[1], 0.0};
  }
};
// x += alpha p + omega s; r = s - omega As; ||r||^2, <r0, r>
template <class T>
struct PtBicgUpd {
  static constexpr int NR = (is_cx<T>::value ? 3 : 2);
  T* x;
  T* r;
  const T* p;
  const T* sv;
  const T* As;
  const T* r0;
  const SolveState* st;
  typename Scal<T>::type a, om;
  __device__ void init() {
    a = as_scal<T>(st->alpha);
    om = as_scal<T>(st->omega);
  }
  __device__ void operator()(int i, double* part) {
    const T si = sv[i];
    x[i] = (x[i] + scal(a, p[i])) + scal(om, si);
    const T ri = si - scal(om, As[i]);
    r[i] = ri;
    part[0] += norm2(ri);
    if constexpr (is_cx<T>::value) {
      const c128 d = dotc(r0[i], ri);
      part[1] += d.re;
      part[2] += d.im;
    } else {
      part[1] += dotc(r0[i], ri);
    }
  }
};
template <class T>
struct FinBicgUpd {
  SolveState* s;
  __device__ void operator()(const double* v) const {
    const double rn = v[0];
    const int k = s->k;
    if (s->hist) s->hist[k] = sqrt(rn) / s->bsqrt;
    if (sqrt(rn) < s->eps * s->bsqrt || k == s->max_iter - 1) {
      s->kfinal = k;
      s->done = 1;
      return;
    }
    if constexpr (is_cx<T>::value) {
      const c128 rho_new{v[1], v[2]};
      s->beta = cdiv(rho_new, s->rho) * cdiv(s->alpha, s->omega);
      s->rho = rho_new;
    } else {
      s->beta = {v[1] / s->rho.re * (s->alpha.re / s->omega.re), 0.0};
      s->rho = {v[1], 0.0};
    }
    s->k = k + 1;
  }
};
// p_new = r + beta (p_old - omega Ap_old)
template <class T>
struct InBicgP {
  static constexpr bool kAux = false;
  const T* r;
  const T* pold;
  const T* Apold;
  T* pnew;
  const SolveState* st;
  typename Scal<T>::type be, om;
  __device__ void init() {
    be = as_scal<T>(st->beta);
    om = as_scal<T>(st->omega);
  }
  __device__ T* target(T* out) const { return out; }
  struct Raw {
    T r, po, apo;
  };
  __device__ Raw fetch(int j, bool) const { return Raw{r[j], pold[j], Apold[j]}; }
  __device__ T finish(const Raw& w, int j, bool own, T&) const {
    const T v = w.r + scal(be, w.po - scal(om, w.apo));
    if (own) pnew[j] = v;
    return v;
  }
  __device__ T load(int j, bool own, T& aux) const { return finish(fetch(j, own), j, own, aux); }
};

template <class T>
SolveResult bicgstab_impl(const DevOperator& o, T* x, const T* b, const SolveOpts& opt) {
  const int n = o.size();
  Work w(opt.stream, stencil_blocks(std::max(o.lx, 1), std::max(o.ly, 1)), opt.max_iter, opt.want_hist);
  SolveState h;
  std::memset(&h, 0, sizeof(h));
  h.max_iter = opt.max_iter;
  h.eps = opt.eps;
  h.hist = w.hist.as<double>();
  w.upload(h);
  const size_t vb = n * sizeof(T);
  DevBuf rB(vb), r0B(vb), pB0(vb), pB1(vb), aB0(vb), aB1(vb), sB(vb), asB(vb);
  T* r = rB.as<T>();
  T* r0 = r0B.as<T>();
  T* p[2] = {pB0.as<T>(), pB1.as<T>()};
  T* Ap[2] = {aB0.as<T>(), aB1.as<T>()};
  T* sv = sB.as<T>();
  T* As = asB.as<T>();
  apply_typed<T>(w.st, o, LINKS_F64, Ap[0], x, w.ws(), kNoGate);
  launch_point(w.st, n, PtBicgInit<T>{r, r0, p[0], b, Ap[0]}, FinBicgInit<T>{w.ds}, w.ws(), kNoGate);
  run_fused<T>(w, o, LINKS_F64, InPlain<T>{p[0]}, p[0], Ap[0], EpiDotAuxOut<T>{r0}, FinBicgAlpha<T>{w.ds});

  auto enq = [&](int k) {
    const int c = k & 1, nx = (k + 1) & 1;
    run_fused<T>(w, o, LINKS_F64, InBicgS<T>{r, Ap[c], sv, w.ds}, sv, As, EpiOmega<T>{}, FinBicgOmega<T>{w.ds});
    launch_point(w.st, n, PtBicgUpd<T>{x, r, p[c], sv, As, r0, w.ds}, FinBicgUpd<T>{w.ds},
                                          w.ws(), w.done());
    run_fused<T>(w, o, LINKS_F64, InBicgP<T>{r, p[c], Ap[c], p[nx], w.ds}, p[nx], Ap[nx], EpiDotAuxOut<T>{r0},
                 FinBicgAlpha<T>{w.ds});
  };
  Mirror m = run_loop(w, opt.max_iter, opt.chunk, enq, [](const Mirror&) { return -1; });
  SolveResult res;
  res.name = "BiCGStab";
  const int kf = opt.max_iter > 0 ? m.kfinal : 0;
  true_residual<T>(w, o, x, b, As);
  SolveState hs = w.download();
  res.success = opt.max_iter > 0 ? (kf != opt.max_iter - 1) : true;
  res.iter = kf + 1;
  res.ops_count = 2 * res.iter + 2;
  if (opt.max_iter <= 0) res.ops_count = 3;
  res.resSq = hs.resSq;
  res.bsqrt = hs.bsqrt;
  if (opt.want_hist && opt.max_iter > 0) res.relres_hist = w.hist_host(res.iter);
  return res;
}

// ================================================================= GCR ====
template <class T>
struct PtGcrInit {
  static constexpr int NR = 1;
  T* r;
  T* p0;
  const T* b;
  const T* ax;
  __device__ void init() {}
  __device__ void operator()(int i, double* part) {
    const T bi = b[i];
    const T ri = bi - ax[i];
    r[i] = ri;
    p0[i] = ri;
    part[0] += norm2(bi);
  }
};
struct FinBsqrt {
  SolveState* s;
  __device__ void operator()(const double* v) const { s->bsqrt = sqrt(v[0]); }
};
// <Ap_k, r> and ||Ap_k||^2 -> alpha = <Ap_k, r> / ||Ap_k||^2
template <class T>
struct PtGcrAlpha {
  static constexpr int NR = (is_cx<T>::value ? 3 : 2);
  const T* apk;
  const T* r;
  __device__ void init() {}
  __device__ void operator()(int i, double* part) {
    const T a = apk[i];
    if constexpr (is_cx<T>::value) {
      const c128 d = dotc(a, r[i]);
      part[0] += d.re;
      part[1] += d.im;
      part[2] += norm2(a);
    } else {
      part[0] += dotc(a, r[i]);
      part[1] += norm2(a);
    }
  }
};
template <class T>
struct FinGcrAlpha {
  SolveState* s;
  double* norms;
  int k;
  __device__ void operator()(const double* v) const {
    if constexpr (is_cx<T>::value) {
      norms[k] = v[2];
      s->alpha = {v[0] / v[2], v[1] / v[2]};
    } else {
      norms[k] = v[1];
      s->alpha = {v[0] / v[1], 0.0};
    }
  }
};
// x += alpha p_k; r -= alpha Ap_k; ||r||^2
template <class T>
struct PtGcrXR {
  static constexpr int NR = 1;
  T* x;
  T* r;
  const T* pk;
  const T* apk;
  const SolveState* s;
  typename Scal<T>::type a;
  __device__ void init() { a = as_scal<T>(s->alpha); }
  __device__ void operator()(int i, double* part) {
    x[i] = x[i] + scal(a, pk[i]);
    const T ri = r[i] - scal(a, apk[i]);
    r[i] = ri;
    part[0] += norm2(ri);
  }
};
struct FinGcrNorm {
  SolveState* s;
  __device__ void operator()(const double* v) const {
    const double rn = v[0];
    const int k = s->k;
    if (s->hist) s->hist[k] = sqrt(rn) / s->bsqrt;
    if (sqrt(rn) < s->eps * s->bsqrt || k == s->max_iter - 1) {
      s->kfinal = k;
      s->done = 1;
    } else {
      s->k = k + 1;
    }
  }
};
// beta_j = -<Ap_j, Ar> / ||Ap_j||^2 for a group of up to 8 stored directions
constexpr int kGcrGroup = 8;
template <class T>
struct PtGcrDots {
  static constexpr int NR = ((is_cx<T>::value ? 2 : 1) * kGcrGroup);
  T* const* aptab;
  const T* ar;
  int j0, cnt;
  __device__ void init() {}
  __device__ void operator()(int i, double* part) {
    const T a = ar[i];
    for (int g = 0; g < cnt; g++) {
      const T v = aptab[j0 + g][i];
      if constexpr (is_cx<T>::value) {
        const c128 d = dotc(v, a);
        part[2 * g] += d.re;
        part[2 * g + 1] += d.im;
      } else {
        part[g] += dotc(v, a);
      }
    }
  }
};
template <class T>
struct FinGcrBeta {
  c128* betas;
  const double* norms;
  int j0, cnt;
  __device__ void operator()(const double* v) const {
    for (int g = 0; g < cnt; g++) {
      const double nrm = norms[j0 + g];
      if constexpr (is_cx<T>::value) betas[j0 + g] = {-v[2 * g] / nrm, -v[2 * g + 1] / nrm};
      else betas[j0 + g] = {-v[g] / nrm, 0.0};
    }
  }
};
// p_{k+1} = r + sum_j beta_j p_j; Ap_{k+1} = Ar + sum_j beta_j Ap_j
template <class T>
struct PtGcrComb {
  static constexpr int NR = 0;
  T* const* ptab;
  T* const* aptab;
  const c128* betas;
  const T* r;
  const T* ar;
  int k;
  __device__ void init() {}
  __device__ void operator()(int i, double*) {
    T pv = r[i], av = ar[i];
    for (int j = 0; j <= k; j++) {
      const typename Scal<T>::type bj = as_scal<T>(betas[j]);
      pv = pv + scal(bj, ptab[j][i]);
      av = av + scal(bj, aptab[j][i]);
    }
    ptab[k + 1][i] = pv;
    aptab[k + 1][i] = av;
  }
};

template <class T>
SolveResult gcr_impl(const DevOperator& o, T* x, const T* b, const SolveOpts& opt) {
  const int n = o.size();
  Work w(opt.stream, stencil_blocks(std::max(o.lx, 1), std::max(o.ly, 1)), opt.max_iter, opt.want_hist);
  SolveState h;
  std::memset(&h, 0, sizeof(h));
  h.max_iter = opt.max_iter;
  h.eps = opt.eps;
  h.hist = w.hist.as<double>();
  w.upload(h);
  const size_t vb = n * sizeof(T);
  const int cap = std::max(opt.max_iter, 0) + 1;
  DevBuf rB(vb), arB(vb), tmpB(vb);
  DevBuf ptabB(sizeof(T*) * cap), aptabB(sizeof(T*) * cap), betasB(sizeof(c128) * cap), normsB(sizeof(double) * cap);
  T* r = rB.as<T>();
  T* ar = arB.as<T>();
  T** ptab = ptabB.as<T*>();
  T** aptab = aptabB.as<T*>();
  std::vector<DevBuf> store;
  std::vector<T*> hp, hap;
  auto add_dir = [&]() {
    store.emplace_back(vb);
    hp.push_back(store.back().as<T>());
    store.emplace_back(vb);
    hap.push_back(store.back().as<T>());
    const int j = int(hp.size()) - 1;
    GLX_HIP(hipMemcpyAsync(ptab + j, &hp[j], sizeof(T*), hipMemcpyHostToDevice, w.st));
    GLX_HIP(hipMemcpyAsync(aptab + j, &hap[j], sizeof(T*), hipMemcpyHostToDevice, w.st));
    // hp / hap are reserved up front, so the sources stay valid until the end
  };
  hp.reserve(cap);
  hap.reserve(cap);
  add_dir();
  apply_typed<T>(w.st, o, LINKS_F64, tmpB.as<T>(), x, w.ws(), kNoGate);
  launch_point(w.st, n, PtGcrInit<T>{r, hp[0], b, tmpB.as<T>()}, FinBsqrt{w.ds}, w.ws(), kNoGate);
  apply_typed<T>(w.st, o, LINKS_F64, hap[0], hp[0], w.ws(), kNoGate);
  c128* betas = betasB.as<c128>();
  double* norms = normsB.as<double>();
  auto enq = [&](int k) {
    launch_point(w.st, n, PtGcrAlpha<T>{hap[k], r}, FinGcrAlpha<T>{w.ds, norms, k}, w.ws(), w.done());
    launch_point(w.st, n, PtGcrXR<T>{x, r, hp[k], hap[k], w.ds}, FinGcrNorm{w.ds}, w.ws(), w.done());
    if (k == opt.max_iter - 1) return;  // the device stops here at the latest
    apply_typed<T>(w.st, o, LINKS_F64, ar, r, w.ws(), w.done());
    for (int j0 = 0; j0 <= k; j0 += kGcrGroup) {
      const int cnt = std::min(kGcrGroup, k + 1 - j0);
      launch_point(w.st, n, PtGcrDots<T>{aptab, ar, j0, cnt},
                                                            FinGcrBeta<T>{betas, norms, j0, cnt}, w.ws(), w.done());
    }
    add_dir();
    launch_point(w.st, n, PtGcrComb<T>{ptab, aptab, betas, r, ar, k}, FinNone{}, w.ws(), w.done());
  };
  Mirror m = run_loop(w, opt.max_iter, opt.chunk, enq, [](const Mirror&) { return -1; });
  SolveResult res;
  res.name = "GCR";
  const int kf = opt.max_iter > 0 ? m.kfinal : 0;
  true_residual<T>(w, o, x, b, tmpB.as<T>());
  SolveState hs = w.download();
  res.success = opt.max_iter > 0 ? (kf != opt.max_iter - 1) : true;
  res.iter = kf + 1;
  res.ops_count = res.iter + 2;
  res.resSq = hs.resSq;
  res.bsqrt = hs.bsqrt;
  if (opt.want_hist && opt.max_iter > 0) res.relres_hist = w.hist_host(res.iter);
  return res;
}

// =========================================================== multishift ===
template <class T>
struct PtNormB {
  static constexpr int NR = 1;
  const T* b;
  __device__ void init() {}
  __device__ void operator()(int i, double* part) { part[0] += norm2(b[i]); }
};
struct FinCgmInit {
  SolveState* s;
  __device__ void operator()(const double* v) const {
    s->bsqrt = sqrt(v[0]);
    s->rsq = v[0];
  }
};
// beta = -rsq / <p, Ap>, then the zeta / beta_s recurrences for active shifts
// (generic_cg_m.cpp:106-122, :396-412).
template <class T>
struct FinCgmBeta {
  SolveState* s;
  __device__ void operator()(const double* v) const {
    s->beta_prev = s->beta;
    if constexpr (is_cx<T>::value) {
      s->beta = cdiv(c128{-s->rsq, 0.0}, c128{v[0], v[1]});
      const c128 b = s->beta, bp = s->beta_prev, a = s->alpha;
      for (int n = 0; n < s->n_rem; n++) {
        const c128 z = s->zeta[n], zp = s->zeta_p[n];
        const c128 num = (z * zp) * bp;
        const c128 sb = s->shifts[n] * b;
        const c128 one_m{1.0 - sb.re, -sb.im};
        const c128 den = (b * a) * (zp - z) + (zp * bp) * one_m;
        s->zeta[n] = cdiv(num, den);
        s->zeta_p[n] = z;
        s->beta_s[n] = cdiv(b * s->zeta[n], s->zeta_p[n]);
      }
    } else {
      s->beta = {-s->rsq / v[0], 0.0};
      const double b = s->beta.re, bp = s->beta_prev.re, a = s->alpha.re;
      for (int n = 0; n < s->n_rem; n++) {
        const double z = s->zeta[n].re, zp = s->zeta_p[n].re;
        const double zn = (z * zp * bp) / (b * a * (zp - z) + zp * bp * (1.0 - s->shifts[n] * b));
        s->zeta[n] = {zn, 0.0};
        s->zeta_p[n] = {z, 0.0};
        s->beta_s[n] = {b * zn / z, 0.0};
      }
    }
  }
};
// r += beta Ap; ||r||^2 (generic_cg_m.cpp:423-426).  The shifts' x_s -= beta_s p_s
// (:414-420) of the same iteration is deferred to the p_s pass below, which
// reads each p_s once for both updates.
template <class T>
struct PtCgmR {
  static constexpr int NR = 1;
  T* r;
  const T* Ap;
  const SolveState* s;
  typename Scal<T>::type be;
  __device__ void init() { be = as_scal<T>(s->beta); }
  __device__ void operator()(int i, double* part) {
    const T ri = r[i] + scal(be, Ap[i]);
    r[i] = ri;
    part[0] += norm2(ri);
  }
};
// convergence bookkeeping, permutation and the alpha recurrences
// (generic_cg_m.cpp:139-215, :429-505)
template <class T>
struct FinCgmR {
  SolveState* s;
  __device__ static void swp(c128& a, c128& b) {
    const c128 t = a;
    a = b;
    b = t;
  }
  __device__ void operator()(const double* v) const {
    const double rn = v[0];
    const int k = s->k;
    if (s->hist) s->hist[k] = sqrt(rn) / s->bsqrt;
    const double tol = s->eps * s->bsqrt;
    // shifts removed below keep their slots at [n_rem, n_xupd) with their
    // beta_s, for the deferred x_s update of this iteration
    s->n_xupd = s->n_rem;
    s->upd_x += s->n_rem;
    if (k % s->resid_freq == 0) {
      for (int n = 0; n < s->n_rem; n++) {
        double zabs;
        if constexpr (is_cx<T>::value) zabs = hypot(s->zeta[n].re, s->zeta[n].im);
        else zabs = s->zeta[n].re;  // the real version compares zeta itself
        if (zabs * sqrt(rn) < tol) {
          s->n_rem--;
          const int e = s->n_rem;
          if (e != n) {
            int t = s->mapping[e]; s->mapping[e] = s->mapping[n]; s->mapping[n] = t;
            t = s->vec[e]; s->vec[e] = s->vec[n]; s->vec[n] = t;
            swp(s->alpha_s[e], s->alpha_s[n]);
            swp(s->beta_s[e], s->beta_s[n]);
            swp(s->zeta[e], s->zeta[n]);
            swp(s->zeta_p[e], s->zeta_p[n]);
            const double d = s->shifts[e]; s->shifts[e] = s->shifts[n]; s->shifts[n] = d;
            n--;
          }
        }
      }
    }
    const double z0 = hypot(s->zeta[0].re, s->zeta[0].im);
    if ((s->worst_first && z0 * sqrt(rn) < tol) || s->n_rem == 0 || k == s->max_iter - 1) {
      s->kfinal = k;
      s->done = 1;
      return;
    }
    s->upd_p += s->n_rem;
    const double al = rn / s->rsq;
    s->alpha = {al, 0.0};
    s->rsq = rn;
    for (int n = 0; n < s->n_rem; n++) {
      if constexpr (is_cx<T>::value) {
        s->alpha_s[n] = cdiv((s->alpha * s->zeta[n]) * s->beta_s[n], s->zeta_p[n] * s->beta);
      } else {
        s->alpha_s[n] = {al * s->zeta[n].re * s->beta_s[n].re / (s->zeta_p[n].re * s->beta.re), 0.0};
      }
    }
    s->k = k + 1;
  }
};
// Per shift that was active before this iteration's check: x_s -= beta_s p_s
// with the old p_s (generic_cg_m.cpp:414-420), then, if it is still active,
// p_s = zeta_s r + alpha_s p_s (:508-514); p = r + alpha p (:516-518).  Each
// p_s is read once: 64 bytes per shift and site instead of 96 for separate x
// and p passes.  The arithmetic of every element is the reference's.
template <class T>
struct PtCgmXP {
  static constexpr int NR = 0;
  T* const* xtab;
  T* const* ptab;
  T* p;
  const T* r;
  const SolveState* s;
  int n_rem, n_x;
  typename Scal<T>::type al;
  __device__ void init() {
    n_rem = s->n_rem;
    n_x = s->n_xupd;
    al = as_scal<T>(s->alpha);
  }
  __device__ void operator()(int i, double*) {
    const T ri = r[i];
    for (int n = 0; n < n_x; n++) {
      const int v = s->vec[n];
      T* xv = xtab[v];
      T* pv = ptab[v];
      const T pold = pv[i];
      xv[i] = xv[i] - scal(as_scal<T>(s->beta_s[n]), pold);
      if (n < n_rem) pv[i] = scal(as_scal<T>(s->zeta[n]), ri) + scal(as_scal<T>(s->alpha_s[n]), pold);
    }
    p[i] = ri + scal(al, p[i]);
  }
};
// The x_s update of the iteration the solve stopped on (its p pass does not
// run): every shift active before that iteration's check.  Launched once,
// ungated, after the loop; n_xupd is 0 when no iteration ran.
template <class T>
struct PtCgmXLast {
  static constexpr int NR = 0;
  T* const* xtab;
  T* const* ptab;
  const SolveState* s;
  int n_x;
  __device__ void init() { n_x = s->n_xupd; }
  __device__ void operator()(int i, double*) {
    for (int n = 0; n < n_x; n++) {
      const int v = s->vec[n];
      T* xv = xtab[v];
      xv[i] = xv[i] - scal(as_scal<T>(s->beta_s[n]), ptab[v][i]);
    }
  }
};
// ||(A + sigma) x - b||^2 for one shift
template <class T>
struct PtShiftRes {
  static constexpr int NR = 1;
  const T* ax;
  const T* x;
  const T* b;
  double sigma;
  __device__ void init() {}
  __device__ void operator()(int i, double* part) {
    const T v = ax[i] + rscal(sigma, x[i]);
    part[0] += norm2(v - b[i]);
  }
};

template <class T>
SolveResult cg_m_impl(const DevOperator& o, T* const* xs, const T* b, int n_shift, const double* shifts,
                      int resid_freq, bool worst_first, const SolveOpts& opt) {
  if (n_shift < 1 || n_shift > kMaxShift) throw ArgError("n_shift must be in [1, 64]");
  if (resid_freq < 1) throw ArgError("resid_freq_check must be >= 1");
  const int n = o.size();
  Work w(opt.stream, stencil_blocks(std::max(o.lx, 1), std::max(o.ly, 1)), opt.max_iter, opt.want_hist);
  SolveState h;
  std::memset(&h, 0, sizeof(h));
  h.max_iter = opt.max_iter;
  h.eps = opt.eps;
  h.hist = w.hist.as<double>();
  h.n_rem = n_shift;
  h.n_shift = n_shift;
  h.resid_freq = resid_freq;
  h.worst_first = worst_first ? 1 : 0;
  h.beta = {1.0, 0.0};
  h.alpha = {0.0, 0.0};
  for (int k = 0; k < n_shift; k++) {
    h.zeta[k] = h.zeta_p[k] = h.beta_s[k] = {1.0, 0.0};
    h.alpha_s[k] = {0.0, 0.0};
    h.shifts[k] = shifts[k];
    h.mapping[k] = k;
    h.vec[k] = k;
  }
  DevBuf resv(sizeof(double) * n_shift);
  h.resvec = resv.as<double>();
  w.upload(h);
  const size_t vb = n * sizeof(T);
  DevBuf rB(vb), pB(vb), apB(vb), xtabB(sizeof(T*) * n_shift), ptabB(sizeof(T*) * n_shift);
  std::vector<DevBuf> ps;
  std::vector<T*> hps(n_shift), hxs(n_shift);
  for (int k = 0; k < n_shift; k++) {
    ps.emplace_back(vb);
    hps[k] = ps.back().as<T>();
    hxs[k] = xs[k];
    GLX_HIP(hipMemcpyAsync(hps[k], b, vb, hipMemcpyDeviceToDevice, w.st));
    GLX_HIP(hipMemsetAsync(xs[k], 0, vb, w.st));
  }
  T** xtab = xtabB.as<T*>();
  T** ptab = ptabB.as<T*>();
  GLX_HIP(hipMemcpyAsync(xtab, hxs.data(), sizeof(T*) * n_shift, hipMemcpyHostToDevice, w.st));
  GLX_HIP(hipMemcpyAsync(ptab, hps.data(), sizeof(T*) * n_shift, hipMemcpyHostToDevice, w.st));
  T* r = rB.as<T>();
  T* p = pB.as<T>();
  T* Ap = apB.as<T>();
  GLX_HIP(hipMemcpyAsync(r, b, vb, hipMemcpyDeviceToDevice, w.st));
  GLX_HIP(hipMemcpyAsync(p, b, vb, hipMemcpyDeviceToDevice, w.st));
  launch_point(w.st, n, PtNormB<T>{b}, FinCgmInit{w.ds}, w.ws(), kNoGate);
  run_fused<T>(w, o, LINKS_F64, InPlain<T>{p}, p, Ap, EpiDotInOut<T>{}, FinCgmBeta<T>{w.ds});
  auto enq = [&](int) {
    launch_point(w.st, n, PtCgmR<T>{r, Ap, w.ds}, FinCgmR<T>{w.ds}, w.ws(), w.done());
    launch_point(w.st, n, PtCgmXP<T>{xtab, ptab, p, r, w.ds}, FinNone{}, w.ws(), w.done());
    run_fused<T>(w, o, LINKS_F64, InPlain<T>{p}, p, Ap, EpiDotInOut<T>{}, FinCgmBeta<T>{w.ds});
  };
  // hipMemcpyAsync of pageable host vectors above must finish before they die
  GLX_HIP(hipStreamSynchronize(w.st));
  Mirror m = run_loop(w, opt.max_iter, opt.chunk, enq, [](const Mirror&) { return -1; });
  launch_point(w.st, n, PtCgmXLast<T>{xtab, ptab, w.ds}, FinNone{}, w.ws(), kNoGate);
  const int kf = opt.max_iter > 0 ? m.kfinal : 0;
  for (int k = 0; k < n_shift; k++) {
    apply_typed<T>(w.st, o, LINKS_F64, Ap, xs[k], w.ws(), kNoGate);
    launch_point(w.st, n, PtShiftRes<T>{Ap, xs[k], b, shifts[k]}, FinResSq{w.ds, k}, w.ws(), kNoGate);
  }
  SolveState hs = w.download();
  SolveResult res;
  res.name = "CG-M";
  res.success = opt.max_iter > 0 ? (kf != opt.max_iter - 1) : true;
  res.iter = kf + 1;
  res.ops_count = 1 + (opt.max_iter > 0 ? kf : 0) + n_shift;
  res.resSq = 0.0;
  res.bsqrt = hs.bsqrt;
  res.resSqmrhs.resize(n_shift);
  GLX_HIP(hipMemcpy(res.resSqmrhs.data(), resv.get(), sizeof(double) * n_shift, hipMemcpyDeviceToHost));
  if (opt.want_hist && opt.max_iter > 0) res.relres_hist = w.hist_host(res.iter);
  return res;
}

// ======================================== single-pass CG (FP64 and mixed) ===
// One fused stencil launch per iteration.  At iteration k (buffers chosen by
// the device's own k parity: old o = (k+1)&1, new n = k&1):
//   prologue  r_k = r_{k-1} - alpha_{k-1} Ap_{k-1};  p_k = r_k + beta_{k-1} p_{k-1}
//             owned sites: store r_k, p_k and x += alpha_{k-1} p_{k-1}
//   stencil   Ap_k = A p_k
//   epilogue  <p,Ap>, |Ap|^2, <Ap,r>, |r|^2 (direct) in one set of partials
//   finaliser alpha_k = |r_k|^2 / <p,Ap>;
//             |r_{k+1}|^2 = |r_k|^2 - 2 Re(conj(alpha) <Ap,r>) + |alpha|^2 |Ap|^2
//             beta_k = |r_{k+1}|^2 / |r_k|^2; convergence on |r_{k+1}|.
// The mixed solver runs this in FP32 (links FP32 or FP16) and, when the
// iterated residual has dropped by delta since the last update (or reached
// the tolerance), pauses (control word 2); the gated reliable-update kernels
// then fold the FP32 solution into the FP64 one, recompute r = b - A x in
// FP64 with FP64 links, restart the recurrence from it and resume -- all on
// the device, so the host only enqueues groups of iterations (or replays one
// captured hipGraph of a group) and reads the control word per group.

// Two slots per vector in one allocation of 2n elements; slot c starts at
// element (c & 1) * n.  Offset arithmetic rather than a pointer array: a
// dynamically indexed member array of a kernel argument is copied to per-lane
// scratch memory.
template <class T>
struct FastBufs {
  T* r;
  T* p;
  T* ap;
  long n;
  __host__ __device__ __forceinline__ T* rs(int c) const { return r + (c & 1) * n; }
  __host__ __device__ __forceinline__ T* ps(int c) const { return p + (c & 1) * n; }
  __host__ __device__ __forceinline__ T* aps(int c) const { return ap + (c & 1) * n; }
};

template <class T, class TX>
struct InFast {
  static constexpr bool kAux = true;
  FastBufs<T> b;
  TX* x;
  const SolveState* s;
  const T* ro;
  const T* apo;
  const T* po;
  T* rn;
  T* pn;
  T* apn;
  typename Scal<T>::type a;
  double beta;
  // buffers of iteration k and the scalars of its prologue
  __device__ void setup(int k, c128 alpha, double beta_r) {
    const int o = (k + 1) & 1, n = k & 1;
    ro = b.rs(o);
    apo = b.aps(o);
    po = b.ps(o);
    rn = b.rs(n);
    pn = b.ps(n);
    apn = b.aps(n);
    a = as_scal<T>(alpha);
    beta = beta_r;
  }
  __device__ void init() { setup(s->k, s->alpha, s->beta_r); }
  __device__ T* target(T*) const { return apn; }
  struct Raw {
    T po, ro, apo;
    TX xv;
  };
  __device__ Raw fetch(int j, bool own) const {
    Raw w;
    w.po = po[j];
    w.ro = ro[j];
    w.apo = apo[j];
    w.xv = own ? x[j] : zero_of<TX>();
    return w;
  }
  __device__ T finish(const Raw& w, int j, bool own, T& aux) const {
    const T r = w.ro - scal(a, w.apo);
    const T p = r + rscal(beta, w.po);
    if (own) {
      rn[j] = r;
      pn[j] = p;
      aux = r;
      if constexpr (std::is_same<T, TX>::value) {
        x[j] = w.xv + scal(a, w.po);
      } else {
        TX d;
        cvt(d, scal(a, w.po));
        x[j] = w.xv + d;
      }
    }
    return p;
  }
  __device__ T load(int j, bool own, T& aux) const { return finish(fetch(j, own), j, own, aux); }
  // paired form for two adjacent sites (16-byte accesses for complex FP32)
  static constexpr bool kPair = true;
  struct RawPair {
    Pair2<T> po, ro, apo;
    Pair2<TX> xv;
  };
  __host__ bool aligned2() const {
    auto al = [](const void* q, size_t e) { return reinterpret_cast<uintptr_t>(q) % (2 * e) == 0; };
    return al(b.r, sizeof(T)) && al(b.p, sizeof(T)) && al(b.ap, sizeof(T)) && al(b.rs(1), sizeof(T)) &&
           al(x, sizeof(TX));
  }
  __device__ RawPair fetch_pair(int j, bool own) const {
    RawPair w;
    w.po = ld2(po, j);
    w.ro = ld2(ro, j);
    w.apo = ld2(apo, j);
    w.xv = own ? ld2(x, j) : Pair2<TX>{zero_of<TX>(), zero_of<TX>()};
    return w;
  }
  __device__ void finish_pair(const RawPair& w, int j, bool own, T& v0, T& v1, T& aux0, T& aux1) const {
    const T r0 = w.ro.a - scal(a, w.apo.a), r1 = w.ro.b - scal(a, w.apo.b);
    v0 = r0 + rscal(beta, w.po.a);
    v1 = r1 + rscal(beta, w.po.b);
    if (own) {
      st2(rn, j, r0, r1);
      st2(pn, j, v0, v1);
      aux0 = r0;
      aux1 = r1;
      TX d0, d1;
      cvt(d0, scal(a, w.po.a));
      cvt(d1, scal(a, w.po.b));
      st2(x, j, w.xv.a + d0, w.xv.b + d1);
    }
  }
};

// <p,Ap>, |Ap|^2, <Ap,r>, |r|^2
template <class T>
struct EpiFast {
  static constexpr int NR = is_cx<T>::value ? 6 : 4;
  __device__ __forceinline__ void acc(double* v, int, T p, T ap, T r) const {
    if constexpr (is_cx<T>::value) {
      const c128 d1 = dotc(p, ap), d2 = dotc(ap, r);
      v[0] += d1.re;
      v[1] += d1.im;
      v[2] += norm2(ap);
      v[3] += d2.re;
      v[4] += d2.im;
      v[5] += norm2(r);
    } else {
      v[0] += dotc(p, ap);
      v[1] += norm2(ap);
      v[2] += dotc(ap, r);
      v[3] += norm2(r);
    }
  }
};

// Scalar part of one single-pass CG step: from iteration k's sums <p,Ap>,
// |Ap|^2, <Ap,r>, |r|^2 form alpha_k, |r_{k+1}|^2 by the recurrence, beta_k,
// and decide whether to stop (FP64: converged or out of iterations; mixed:
// also when |r| fell below delta times its maximum since the last reliable
// update).  Tests compare squared norms (no square root on this path) and
// alpha = rr conj(<p,Ap>) / |<p,Ap>|^2 takes one division.  maxrr2 is the
// maximum of |r|^2 since the last reliable update.
struct FastStep {
  c128 alpha;
  double rsq, rsqNew, beta_r, maxrr2;
  bool stop;
};
template <class T, bool MIXED>
__device__ __forceinline__ FastStep fast_step(const double* v, int k, double maxrr2, const SolveState* s) {
  c128 pap, apr;
  double apap, rr;
  if constexpr (is_cx<T>::value) {
    pap = {v[0], v[1]};
    apap = v[2];
    apr = {v[3], v[4]};
    rr = v[5];
  } else {
    pap = {v[0], 0.0};
    apap = v[1];
    apr = {v[2], 0.0};
    rr = v[3];
  }
  FastStep f;
  const double g = rr / (pap.re * pap.re + pap.im * pap.im);
  f.alpha = {g * pap.re, -g * pap.im};
  // Re(conj(alpha) <Ap, r>)
  const double re = f.alpha.re * apr.re + f.alpha.im * apr.im;
  double rn = rr - 2.0 * re + (f.alpha.re * f.alpha.re + f.alpha.im * f.alpha.im) * apap;
  if (!(rn > 0.0)) rn = 0.0;
  f.rsq = rr;
  f.rsqNew = rn;
  const double tol = s->eps * s->bsqrt;
  f.maxrr2 = maxrr2;
  if constexpr (MIXED) {
    if (rn > f.maxrr2) f.maxrr2 = rn;
    f.stop = rn < (s->delta * s->delta) * f.maxrr2 || rn < tol * tol || k == s->max_iter - 1;
  } else {
    f.stop = rn < tol * tol || k == s->max_iter - 1;
  }
  f.beta_r = rn / rr;
  return f;
}

// Finaliser of the two-kernel form (stencil + k_finalize).
template <class T, bool MIXED>
struct FinFast {
  SolveState* s;
  __device__ void operator()(const double* v) const {
    const int k = s->k;
    const FastStep f = fast_step<T, MIXED>(v, k, s->maxrr, s);
    s->alpha = f.alpha;
    s->rsq = f.rsq;
    s->rsqNew = f.rsqNew;
    if (s->hist) s->hist[k] = sqrt(f.rsqNew) / s->bsqrt;
    if constexpr (MIXED) s->maxrr = f.maxrr2;
    if (f.stop) {
      s->kfinal = k;
      s->done = MIXED ? 2 : 1;
      if (s->hdone) *s->hdone = s->done;
      return;
    }
    s->beta_r = f.beta_r;
    s->k = k + 1;
  }
};

// Self-finalising form (row-marching kernel only): the iteration kernel
// itself reduces the previous launch's partial sums in every block (the
// grid is the same every launch and the order fixed, so every block computes
// identical scalars), and block 0 writes the resulting state.  No finaliser
// launch and no device-wide atomics per iteration.  Launches alternate
// between two state copies and two partial-sum buffers by a launch parity q
// fixed on the host: launch q reads state `rd` and buffer q ^ 1, writes state
// `wr` and buffer q; nothing a launch reads is written during that launch.
// Groups hold an even number of launches, so the state between groups is
// always the home copy (parity 1 writes it), which the host mirror, the
// reliable-update kernels and the final x update use.
//   pending = 1: the partial sums of iteration k (in buffer q ^ 1) await
//                reduction; this launch performs iteration k + 1
//   pending = 0: alpha / beta_r are final (start, after a reliable update);
//                this launch performs iteration k
// A launch that finds done != 0 only carries the state over to `wr`.
template <class T, class TX, bool MIXED>
struct InFastSF : InFast<T, TX> {
  static constexpr bool kSelfFin = true;
  static constexpr int NR = EpiFast<T>::NR;
  const SolveState* rd;
  SolveState* wr;
  int q;
  int pending, k0;
  bool gated;
  FastStep f;
  __device__ void init() {
    gated = rd->done != 0;
    pending = rd->pending;
    k0 = rd->k;
    // buffers of the iteration this launch performs; its scalars are final
    // only after prelude()
    this->setup(pending ? k0 + 1 : k0, rd->alpha, rd->beta_r);
  }
  // The previous launch's partial sums, in two halves: prelude_load issues
  // the first sweep of loads before the kernel's first row reads, so that
  // waiting for them (loads complete in issue order) does not also wait for
  // the rows; prelude_finish reduces them and takes the CG step.
  // Block-wide; every thread returns the same decision (true = iterate).
  static constexpr int kPreU = 2;  // terms per thread per sweep (512 blocks in one sweep)
  double pv[kPreU][NR];
  __device__ void prelude_load(RedWs ws) {
    if (gated || !pending) return;
    const int nblk = gridDim.x;
    const double* P = ws.partials + (q ^ 1) * NR * nblk;
#pragma unroll
    for (int u = 0; u < kPreU; u++) {
      const int b = threadIdx.x + u * kBlock;
#pragma unroll
      for (int k = 0; k < NR; k++) pv[u][k] = b < nblk ? P[k * nblk + b] : 0.0;
    }
  }
  __device__ bool prelude_finish(RedWs ws) {
    if (gated) return false;
    if (!pending) return true;
    __shared__ double sm[NR][kWaves];
    const int nblk = gridDim.x;
    const double* P = ws.partials + (q ^ 1) * NR * nblk;
    double acc[NR];
#pragma unroll
    for (int k = 0; k < NR; k++) acc[k] = pv[0][k];
#pragma unroll
    for (int u = 1; u < kPreU; u++) {
#pragma unroll
      for (int k = 0; k < NR; k++) acc[k] += pv[u][k];
    }
    for (int b0 = threadIdx.x + kBlock * kPreU; b0 < nblk; b0 += kBlock * kPreU) {
      double v[kPreU][NR];
#pragma unroll
      for (int u = 0; u < kPreU; u++) {
        const int b = b0 + u * kBlock;
#pragma unroll
        for (int k = 0; k < NR; k++) v[u][k] = b < nblk ? P[k * nblk + b] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < kPreU; u++) {
#pragma unroll
        for (int k = 0; k < NR; k++) acc[k] += v[u][k];
      }
    }
    block_sum_dpp<NR>(acc, sm);
    f = fast_step<T, MIXED>(acc, k0, rd->maxrr, rd);
    if (f.stop) return false;
    this->a = as_scal<T>(f.alpha);
    this->beta = f.beta_r;
    return true;
  }
  __device__ RedWs out_ws(RedWs ws) const { return RedWs{ws.partials + q * NR * gridDim.x}; }
  // At the end of the kernel: block 0 writes the state of the next launch.
  __device__ void arrive(bool iterated) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    arrive_state(iterated);
    if (rd->hdone) *rd->hdone = wr->done;
  }
  __device__ void arrive_state(bool iterated) {
    // carried over unless changed below
    wr->k = rd->k;
    wr->done = rd->done;
    wr->kfinal = rd->kfinal;
    wr->pending = rd->pending;
    wr->alpha = rd->alpha;
    wr->beta_r = rd->beta_r;
    wr->rsq = rd->rsq;
    wr->rsqNew = rd->rsqNew;
    wr->maxrr = rd->maxrr;
    if (gated) return;
    if (pending) {
      wr->alpha = f.alpha;
      wr->rsq = f.rsq;
      wr->rsqNew = f.rsqNew;
      if (rd->hist) rd->hist[k0] = sqrt(f.rsqNew) / rd->bsqrt;
      if constexpr (MIXED) wr->maxrr = f.maxrr2;
      if (f.stop) {
        wr->kfinal = k0;
        wr->done = MIXED ? 2 : 1;
        wr->pending = 0;
        return;
      }
      wr->beta_r = f.beta_r;
      wr->k = k0 + 1;
    }
    if (iterated) wr->pending = 1;
  }
};

// setup: r_old = b - A x, Ap_old = 0, p_old = 0, with |b| and |r|
template <class TD, class TS>
struct PtFastInit {
  static constexpr int NR = 2;
  TS* r;
  TS* ap;
  TS* p;
  const TD* b;
  const TD* ax;
  __device__ void init() {}
  __device__ void operator()(int i, double* part) {
    const TD bi = b[i];
    const TD ri = bi - ax[i];
    TS v;
    cvt(v, ri);
    r[i] = v;
    ap[i] = zero_of<TS>();
    p[i] = zero_of<TS>();
    part[0] += norm2(bi);
    part[1] += norm2(ri);
  }
};
struct FinFastInit {
  SolveState* s;
  __device__ void operator()(const double* v) const {
    s->bsqrt = sqrt(v[0]);
    s->rsq = v[1];
    s->maxrr = v[1];  // |r|^2
    s->alpha = {0.0, 0.0};
    s->beta_r = 0.0;
    s->k = 0;
    s->ko = 0;
    s->done = 0;
    s->pending = 0;
    if (s->hdone) *s->hdone = 0;
  }
};

// reliable update, step 1 (gated on a pause): x64 += x32 + alpha_k p_k; x32 = 0
template <class TD, class TS>
struct PtFastAccum {
  static constexpr int NR = 0;
  TD* x;
  TS* xs;
  FastBufs<TS> bufs;
  const SolveState* s;
  const TS* p;
  typename Scal<TD>::type a;
  __device__ void init() {
    p = bufs.ps(s->k);
    a = as_scal<TD>(s->alpha);
  }
  __device__ void operator()(int i, double*) {
    TD xl, pl;
    cvt(xl, xs[i]);
    cvt(pl, p[i]);
    x[i] = x[i] + (xl + scal(a, pl));
    xs[i] = zero_of<TS>();
  }
};
// step 2 epilogue: r_true = b - A x in FP64, written (rounded) as the
// recurrence's old residual r[k&1]; |r_true|^2 decides convergence
template <class TD, class TS>
struct EpiFastTrue {
  static constexpr int NR = 1;
  FastBufs<TS> bufs;
  const TD* b;
  const SolveState* s;
  __device__ __forceinline__ void acc(double* part, int i, TD, TD ax, TD) const {
    const TD d = b[i] - ax;
    TS v;
    cvt(v, d);
    bufs.rs(s->k)[i] = v;
    part[0] += norm2(d);
  }
};
struct FinFastReliable {
  SolveState* s;
  __device__ void operator()(const double* v) const {
    const double rt = v[0];
    const int k = s->k;
    s->reliable += 1;
    s->resSq = rt;
    if (rt == 0.0 || sqrt(rt) < s->eps * s->bsqrt || k >= s->max_iter - 1) {
      s->kfinal = k;
      s->done = 1;
      if (s->hdone) *s->hdone = 1;
      return;
    }
    // the next prologue forms r_{k+1} = r_true - 0 * Ap and p_{k+1} = r_true + beta p_k
    s->beta_r = rt / s->rsq;
    s->alpha = {0.0, 0.0};
    s->maxrr = rt;  // |r|^2
    s->k = k + 1;
    s->pending = 0;
    s->done = 0;
    if (s->hdone) *s->hdone = 0;
  }
};
// FP64 single-pass CG: the pending alpha p after convergence
template <class T>
struct PtFastFinal {
  static constexpr int NR = 0;
  T* x;
  FastBufs<T> bufs;
  const SolveState* s;
  const T* p;
  typename Scal<T>::type a;
  __device__ void init() {
    p = bufs.ps(s->kfinal);
    a = as_scal<T>(s->alpha);
  }
  __device__ void operator()(int i, double*) { x[i] = x[i] + scal(a, p[i]); }
};

// A non-default (blocking) stream owned by one solve: graph capture needs it,
// and it still orders after work the caller queued on the null stream.
// A solve that gives up while groups of iterations are still queued waits
// for its stream before it throws: the unwinding hands the solve's buffers
// back to the device pool, and no queued kernel may still write them then
// (also on a caller's stream, which OwnedStream does not synchronise).
[[noreturn]] static void fail_drained(hipStream_t st, hipGraphExec_t exec, const char* msg) {
  (void)hipStreamSynchronize(st);
  if (exec) (void)hipGraphExecDestroy(exec);
  throw HipError(msg);
}

// The stream of a solve: the caller's (`want`), else the allocation-free
// scope's, else one created for the solve and destroyed after it.
// The stream is taken from the calling thread's streams of finished solves
// on the current device (a stream created and destroyed per solve costs a
// good part of a millisecond); the thread's streams are destroyed when it
// exits.
struct IdleStreams {
  std::vector<std::pair<int, hipStream_t>> v;  // (device, stream)
  ~IdleStreams() {
    for (auto& e : v) (void)hipStreamDestroy(e.second);
  }
};
static IdleStreams& idle_streams() {
  static thread_local IdleStreams s;
  return s;
}
struct OwnedStream {
  hipStream_t st = nullptr;
  int dev = -1;
  bool owned = false;
  explicit OwnedStream(hipStream_t want = nullptr) {
    if (want) {
      st = want;
    } else if (NoAllocScope* s = noalloc_scope()) {
      st = s->stream();
    } else {
      GLX_HIP(hipGetDevice(&dev));
      auto& idle = idle_streams().v;
      for (size_t i = idle.size(); i-- > 0;) {
        if (idle[i].first == dev) {
          st = idle[i].second;
          idle.erase(idle.begin() + long(i));
          break;
        }
      }
      if (!st) GLX_HIP(hipStreamCreate(&st));
      owned = true;
    }
  }
  ~OwnedStream() {
    if (owned) {
      (void)hipStreamSynchronize(st);
      idle_streams().v.emplace_back(dev, st);
    }
  }
};

// Live timing of the single-pass CG's iteration kernel alone, in the form the
// solver launches it (self-finalising where the row-marching kernel takes
// the lattice, else the stencil without its finaliser): `reps` launches on
// one stream between two HIP events, on zeroed vectors (no data-dependent
// control flow in the kernel).  Returns ms per launch.
template <class TD, class TS, bool MIXED>
double time_fast_kernel_impl(const DevOperator& o, int link_prec, int reps) {
  const int n = o.size();
  OwnedStream own;
  hipStream_t st = own.st;
  Work w(st, stencil_blocks(o.lx, o.ly), 0, false);
  std::vector<DevBuf> keep;
  auto vec = [&](size_t bytes) {
    keep.emplace_back(bytes);
    GLX_HIP(hipMemsetAsync(keep.back().get(), 0, bytes, st));
    return keep.back().get();
  };
  FastBufs<TS> fb;
  fb.n = n;
  fb.r = static_cast<TS*>(vec(2 * n * sizeof(TS)));
  fb.p = static_cast<TS*>(vec(2 * n * sizeof(TS)));
  fb.ap = static_cast<TS*>(vec(2 * n * sizeof(TS)));
  using TXv = typename std::conditional<MIXED, TS, TD>::type;
  TXv* xv = static_cast<TXv*>(vec(n * sizeof(TXv)));
  const bool sf = normal_path<TS>(o, InFast<TS, TXv>{fb, xv, w.ds}, fb.aps(0));
  DevBuf shadow_buf(sizeof(SolveState));
  SolveState* shadow = shadow_buf.as<SolveState>();
  GLX_HIP(hipMemcpyAsync(shadow, w.ds, sizeof(SolveState), hipMemcpyDeviceToDevice, st));
  int q = 0;
  auto launch = [&]() {
    if (sf) {
      const SolveState* rd = q == 0 ? w.ds : shadow;
      SolveState* wr = q == 0 ? shadow : w.ds;
      run_stencil<TS>(st, o, link_prec, InFastSF<TS, TXv, MIXED>{{fb, xv, rd}, rd, wr, q}, fb.aps(0),
                      EpiFast<TS>{}, FinNone{}, w.ws(), kNoGate, false);
      q ^= 1;
    } else {
      run_stencil<TS>(st, o, link_prec, InFast<TS, TXv>{fb, xv, w.ds}, fb.aps(0), EpiFast<TS>{},
                      FinFast<TS, MIXED>{w.ds}, w.ws(), kNoGate, false);
    }
  };
  launch();  // warm-up
  hipEvent_t a, b;
  GLX_HIP(hipEventCreate(&a));
  GLX_HIP(hipEventCreate(&b));
  GLX_HIP(hipEventRecord(a, st));
  for (int k = 0; k < reps; k++) launch();
  GLX_HIP(hipEventRecord(b, st));
  GLX_HIP(hipEventSynchronize(b));
  float ms = 0.f;
  GLX_HIP(hipEventElapsedTime(&ms, a, b));
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return double(ms) / reps;
}

template <class TD, class TS, bool MIXED>
SolveResult fastcg_impl(const DevOperator& o, TD* x, const TD* b, const SolveOpts& opt, const MixedOpts& mo) {
  if (!o.tiled()) throw ArgError("single-pass CG needs a device stencil operator");
  const int n = o.size();
  OwnedStream own(opt.stream);
  hipStream_t st = own.st;
  Work w(st, stencil_blocks(o.lx, o.ly), opt.max_iter, opt.want_hist);
  SolveState h;
  std::memset(&h, 0, sizeof(h));
  h.max_iter = opt.max_iter;
  h.eps = opt.eps;
  h.delta = mo.delta;
  h.hist = w.hist.as<double>();
  // allocation-free mode: no pinned slots; the host reads the state after
  // each group instead (synchronously)
  PinnedSlots* pin = noalloc_scope() ? nullptr : &pinned();
  if (pin) __atomic_store_n(pin->hdone, 0, __ATOMIC_RELEASE);
  h.hdone = pin ? pin->hdone_dev : nullptr;
  w.upload(h);
  std::vector<DevBuf> keep;
  auto vec = [&](size_t bytes) {
    keep.emplace_back(bytes);
    return keep.back().get();
  };
  FastBufs<TS> fb;
  fb.n = n;
  fb.r = static_cast<TS*>(vec(2 * n * sizeof(TS)));
  fb.p = static_cast<TS*>(vec(2 * n * sizeof(TS)));
  fb.ap = static_cast<TS*>(vec(2 * n * sizeof(TS)));
  TD* t64 = static_cast<TD*>(vec(n * sizeof(TD)));
  TS* xs = nullptr;
  if constexpr (MIXED) {
    xs = static_cast<TS*>(vec(n * sizeof(TS)));
    GLX_HIP(hipMemsetAsync(xs, 0, n * sizeof(TS), st));
  }
  // r_{-1} = b - A x in slot 1 (k = 0 reads slot 1 as "old"), Ap_{-1} = p_{-1} = 0
  apply_typed<TD>(st, o, LINKS_F64, t64, x, w.ws(), kNoGate);
  launch_point(st, n, PtFastInit<TD, TS>{fb.rs(1), fb.aps(1), fb.ps(1), b, t64}, FinFastInit{w.ds}, w.ws(), kNoGate);
  const int link_prec = MIXED ? mo.link_prec : LINKS_F64;
  using TXv = typename std::conditional<MIXED, TS, TD>::type;
  TXv* xv;
  if constexpr (MIXED) xv = xs;
  else xv = x;
  // self-finalising iteration (no finaliser launch) where the row-marching
  // kernel takes the lattice; the stencil + finaliser pair otherwise
  const bool sf = normal_path<TS>(o, InFast<TS, TXv>{fb, xv, w.ds}, fb.aps(0));
  DevBuf shadow_buf(sf ? sizeof(SolveState) : 1);
  SolveState* shadow = shadow_buf.as<SolveState>();
  auto enq_iter = [&](int q) {
    if (sf) {
      const SolveState* rd = q == 0 ? w.ds : shadow;
      SolveState* wr = q == 0 ? shadow : w.ds;
      run_stencil<TS>(st, o, link_prec, InFastSF<TS, TXv, MIXED>{{fb, xv, rd}, rd, wr, q}, fb.aps(0),
                      EpiFast<TS>{}, FinNone{}, w.ws(), kNoGate, false);
    } else {
      run_stencil<TS>(st, o, link_prec, InFast<TS, TXv>{fb, xv, w.ds}, fb.aps(0), EpiFast<TS>{},
                      FinFast<TS, MIXED>{w.ds}, w.ws(), w.done());
    }
  };
  // the shadow copy starts as the initial state (constants included)
  if (sf) GLX_HIP(hipMemcpyAsync(shadow, w.ds, sizeof(SolveState), hipMemcpyDeviceToDevice, st));
  // reliable update (gated on a pause): x64 += x32 + alpha p, then one FP64
  // operator launch whose epilogue forms the true residual and its norm.
  // Most groups skip it, so the gated pointwise launch uses a small grid.
  auto enq_reliable = [&]() {
    if constexpr (MIXED) {
      launch_point(st, n, PtFastAccum<TD, TS>{x, xs, fb, w.ds}, FinNone{}, w.ws(), w.paused(), kGatedBlocks);
      run_stencil<TD>(st, o, LINKS_F64, InPlain<TD>{x}, t64, EpiFastTrue<TD, TS>{fb, b, w.ds},
                      FinFastReliable{w.ds}, w.ws(), w.paused());
    }
  };
  // the self-finalising form alternates state copies: an even number of
  // launches per group ends every group on the home copy
  const int group = sf ? 2 * ((std::max(1, opt.group) + 1) / 2) : std::max(1, opt.group);
  auto enq_group = [&]() {
    for (int g = 0; g < group; g++) enq_iter(g & 1);
    enq_reliable();
  };
  // optional: one captured graph of a group, replayed
  hipGraphExec_t exec = nullptr;
  if (opt.use_graph) {
    hipGraph_t graph = nullptr;
    GLX_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    enq_group();
    GLX_HIP(hipStreamEndCapture(st, &graph));
    GLX_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    GLX_HIP(hipGraphDestroy(graph));
  }
  auto launch = [&]() {
    if (exec) GLX_HIP(hipGraphLaunch(exec, st));
    else enq_group();
  };
  // pipelined: one group queued ahead of the one being waited for.  With the
  // self-finalising kernel the host reads the done flag the kernels store in
  // host memory themselves, after an event recorded without the system-scope
  // cache writeback (which would stall the following group); the other form
  // copies the state back after each group.
  const long max_groups = 4L * (opt.max_iter / group + 2) + 16;
  long groups = 0;
  if (opt.max_iter > 0 && !pin) {
    // every group ends on the home copy of the state
    for (;;) {
      launch();
      groups++;
      Mirror mm;
      GLX_HIP(hipMemcpyAsync(&mm, w.ds, sizeof(Mirror), hipMemcpyDeviceToHost, st));
      GLX_HIP(hipStreamSynchronize(st));
      if (mm.done == 1) break;
      if (groups > max_groups) fail_drained(st, exec, "single-pass CG made no progress");
    }
  } else if (opt.max_iter > 0 && sf) {
    int cur = 0;
    auto submit = [&](int slot) {
      launch();
      groups++;
      GLX_HIP(hipEventRecord(pin->ev_nf[slot], st));
    };
    submit(cur);
    for (;;) {
      submit(cur ^ 1);
      GLX_HIP(hipEventSynchronize(pin->ev_nf[cur]));
      if (__atomic_load_n(pin->hdone, __ATOMIC_ACQUIRE) == 1) break;
      if (groups > max_groups) fail_drained(st, exec, "single-pass CG made no progress");
      cur ^= 1;
    }
  } else if (opt.max_iter > 0) {
    int cur = 0;
    auto submit = [&](int slot) {
      launch();
      groups++;
      GLX_HIP(hipMemcpyAsync(&pin->m[slot], w.ds, sizeof(Mirror), hipMemcpyDeviceToHost, st));
      GLX_HIP(hipEventRecord(pin->ev[slot], st));
    };
    submit(cur);
    for (;;) {
      submit(cur ^ 1);
      GLX_HIP(hipEventSynchronize(pin->ev[cur]));
      if (pin->m[cur].done == 1) break;
      if (groups > max_groups) fail_drained(st, exec, "single-pass CG made no progress");
      cur ^= 1;
    }
  }
  GLX_HIP(hipStreamSynchronize(st));
  if (exec) GLX_HIP(hipGraphExecDestroy(exec));
  SolveState hs = w.download();
  SolveResult res;
  res.name = MIXED ? "CG-mixed" : "CG-fused";
  if (opt.max_iter <= 0) {
    res.iter = 1;
    res.success = true;
    return res;
  }
  if constexpr (!MIXED) {
    launch_point(st, n, PtFastFinal<TD>{x, fb, w.ds}, FinNone{}, w.ws(), kNoGate);
    true_residual<TD>(w, o, x, b, t64);
    hs = w.download();
    res.success = hs.kfinal != opt.max_iter - 1;
    res.ops_count = hs.kfinal + 3;
  } else {
    res.success = sqrt(hs.resSq) < opt.eps * hs.bsqrt;
    res.ops_count = hs.kfinal + 2 + hs.reliable;
  }
  res.iter = hs.kfinal + 1;
  res.resSq = hs.resSq;
  res.bsqrt = hs.bsqrt;
  res.reliable_updates = hs.reliable;
  if (opt.want_hist) res.relres_hist = w.hist_host(res.iter);
  return res;
}


// ====================================== s-step CG (mixed precision, s = 2) ===
// The mixed solver with two CG iterations per launch (k_sscg, glx_sstep.h) on
// the full-lattice D^dag D.  The reliable updates are the single-pass
// solver's: x64 += x32, then one FP64 operator launch whose epilogue writes
// the true residual into R0 of the basis slot the next launch reads.

namespace {
std::atomic<int> g_sstep{2};
std::atomic<int> g_sstep_rows{0};  // 0: per lattice size
struct SsTable {
  std::mutex mu;
  std::map<std::pair<int, int>, int> rows;
};
SsTable& ss_table() {
  static SsTable t;
  return t;
}
}  // namespace
void set_sstep(int s, int rows) {
  if (s != 1 && s != 2) throw ArgError("s-step: s must be 1 or 2");
  if (rows != 0 && (rows < 8 || rows % 4 != 0)) throw ArgError("s-step: rows must be 0 or a multiple of 4, >= 8");
  g_sstep.store(s);
  g_sstep_rows.store(rows);
}
int sstep_default() { return g_sstep.load(); }
int sstep_rows() { return g_sstep_rows.load(); }
namespace {
std::atomic<int> g_sstep_form{-1};  // -1: by lattice size
std::atomic<int> g_sw_rows{0};      // 0: by lattice size
struct FormTable {
  std::mutex mu;
  std::map<std::pair<int, int>, int> form;
};
FormTable& form_table() {
  static FormTable t;
  return t;
}
}  // namespace
void set_sstep_form(int form, int band_rows) {
  if (form < -1 || form > 1)
    throw ArgError("s-step form must be -1 (by lattice size), 0 (blocks of four waves) or 1 (wave strips)");
  if (band_rows < 0 || band_rows > kSwMaxRows)
    throw ArgError("s-step band rows must be in [0, " + std::to_string(kSwMaxRows) + "]");
  g_sstep_form.store(form);
  g_sw_rows.store(band_rows);
}
int sstep_form() { return g_sstep_form.load(); }
void set_sstep_form_entry(int lx, int ly, int form) {
  if (form < -1 || form > 1) throw ArgError("s-step form entry must be -1 (remove), 0 or 1");
  FormTable& t = form_table();
  std::lock_guard<std::mutex> g(t.mu);
  if (form < 0) t.form.erase(std::make_pair(lx, ly));
  else t.form[std::make_pair(lx, ly)] = form;
}
int sstep_form_entry(int lx, int ly) {
  FormTable& t = form_table();
  std::lock_guard<std::mutex> g(t.mu);
  auto it = t.form.find(std::make_pair(lx, ly));
  return it != t.form.end() ? it->second : -1;
}
// The form a solve of an lx x ly D^dag D takes: the process-wide choice if
// one is set, else the size's entry, else the wave strips where the lattice
// has at least 512^2 sites and fits within 1024 waves at bands of at most 52
// rows (measured per launch: 768^2 21.4 against 28.2 us, 1024^2 28.1 against
// 31.6 us; 512^2 and smaller tie; 2048^2 125 against 101 us, 4096^2 439
// against 373 us, where the lattice no longer fits the last-level cache and
// the blocks' aligned 128-column strips stream better).
int sstep_form_for(int lx, int ly) {
  const int f = g_sstep_form.load();
  if (f >= 0) return f;
  const int e = sstep_form_entry(lx, ly);
  if (e >= 0) return e;
  const long nstrip = (lx + kSwOwn - 1) / kSwOwn;
  const long waves = nstrip * ((ly + 51) / 52);
  return long(lx) * ly >= 512L * 512L && waves <= 1024 ? 1 : 0;
}

// Rows per wave of the wave-strip form: the fewest of the form 6 k - 8 (a
// wave marches rows + 8 rows in rounds of six) that keep the waves within one
// per SIMD (1024), at least 10 and at most 52.
int sw_rows_for(int lx, int ly) {
  const int forced = g_sw_rows.load();
  if (forced > 0) return forced;
  const int nstrip = (lx + kSwOwn - 1) / kSwOwn;
  const int bands = std::max(1, 1024 / nstrip);
  const int need = (ly + bands - 1) / bands;
  int rows = 10;
  while (rows < need && rows + 6 <= kSwMaxRows) rows += 6;
  return rows;
}
void set_sstep_rows(int lx, int ly, int rows) {
  SsTable& t = ss_table();
  std::lock_guard<std::mutex> g(t.mu);
  if (rows < 0) t.rows.erase(std::make_pair(lx, ly));
  else t.rows[std::make_pair(lx, ly)] = rows;
}
// Without a table entry: the fewest rows per block (8, 16, 32, 64, 128) that
// keep the grid within 512 blocks, two per CU, else 256 (measured: 256^2 and
// 512^2 best at 8, 768^2 and 1024^2 at 16, 2048^2 at 64, 4096^2 at 256
// (3% less per launch than 128); the strips are 128 lattice columns wide for
// both operators).
int sstep_rows_auto(int lx, int ly) {
  const long strips = lx / (kNX * 2);
  for (int r = 8; r <= 128; r *= 2)
    if (strips * ((ly + r - 1) / r) <= 512) return r;
  return 256;
}
int sstep_rows_entry(int lx, int ly) {
  SsTable& t = ss_table();
  std::lock_guard<std::mutex> g(t.mu);
  auto it = t.rows.find(std::make_pair(lx, ly));
  return it != t.rows.end() ? it->second : -1;
}
int sstep_rows_for(int lx, int ly) {
  const int forced = g_sstep_rows.load();
  if (forced > 0) return forced;
  SsTable& t = ss_table();
  std::lock_guard<std::mutex> g(t.mu);
  auto it = t.rows.find(std::make_pair(lx, ly));
  return it != t.rows.end() ? it->second : sstep_rows_auto(lx, ly);
}

// Columns per lane of the s-step kernel: D^dag D in strips of 128 columns
// (two per lane); the even/odd Schur operator in strips of 64 compact
// columns (one per lane: its links take twice the LDS), also 128 lattice
// columns wide.
constexpr int kSsCPL = 2, kSsCPLEo = 1;

static bool sscg_fits(const DevOperator& o) {
  if (!o.tiled() || o.ly < 8 || o.ly % 4 != 0) return false;
  if (long(o.lx) * o.ly >= kSsMaxSites) return false;  // its 32-bit offsets; the one-iteration form takes these
  if (o.kind == OP_STAG_NORMAL) return o.lx >= kNX * kSsCPL && o.lx % (kNX * kSsCPL) == 0;
  if (o.kind == OP_STAG_EO_SCHUR) return o.lx % 2 == 0 && o.lx / 2 >= kNX * kSsCPLEo && (o.lx / 2) % (kNX * kSsCPLEo) == 0;
  return false;
}
// Launch geometry of the s-step kernel for the operator's lattice.
static EoGeom ss_geom(const DevOperator& o, int rows) {
  if (o.kind == OP_STAG_EO_SCHUR) return make_eogeom(o.lx, o.ly, rows, kSsCPLEo);
  EoGeom e;
  e.g = make_ngeom(o.lx, o.ly, rows, kSsCPL);
  e.lx = o.lx;
  return e;
}
// The s-step launch of an operator: blocks of four waves marching a strip
// together (k_sscg, `rows` rows per block), or, for the full-lattice D^dag D
// in the wave-strip form, every wave its own strip (k_sscg_w).
struct SsGeo {
  bool wave = false;
  EoGeom e;
  SwGeom s;
  int nblk() const { return wave ? s.nblk : e.g.nblk; }
};
static SsGeo ss_geo(const DevOperator& o, int rows) {
  SsGeo g;
  g.wave = (o.kind == OP_STAG_NORMAL || o.kind == OP_STAG_EO_SCHUR) && sstep_form_for(o.lx, o.ly) == 1;
  // the even/odd Schur operator takes the wave strips by size only above
  // 512^2 sites (mixed even/odd solves, form 0 against 1: 512^2 7.0 against
  // 7.6 ms, 768^2 10.5 against 9.1, 1024^2 11.2 against 10.5)
  if (o.kind == OP_STAG_EO_SCHUR && sstep_form() < 0 && long(o.lx) * o.ly <= 512L * 512L) g.wave = false;
  if (g.wave) g.s = make_swgeom(o.lx, o.ly, sw_rows_for(o.lx, o.ly));
  else g.e = ss_geom(o, rows);
  return g;
}

// setup: R0 of slot 0 = b - A x in FP32, with |b|^2 and |r|^2
struct PtSsInit {
  static constexpr int NR = 2;
  c64* r;
  const c128* b;
  const c128* ax;
  __device__ void init() {}
  __device__ void operator()(int i, double* part) {
    const c128 bi = b[i];
    const c128 ri = bi - ax[i];
    c64 v;
    cvt(v, ri);
    r[i] = v;
    part[0] += norm2(bi);
    part[1] += norm2(ri);
  }
};
// reliable update, step 1 (gated on a pause): x64 += x32; x32 = 0
struct PtSsAccum {
  static constexpr int NR = 0;
  c128* x;
  c64* xs;
  __device__ void init() {}
  __device__ void operator()(int i, double*) {
    c128 d;
    cvt(d, xs[i]);
    x[i] = x[i] + d;
    xs[i] = zero_of<c64>();
  }
};
// step 2 epilogue: r_true = b - A x in FP64, written (rounded) as R0 of the
// slot the next launch reads; |r_true|^2 decides convergence
struct EpiSsTrue {
  static constexpr int NR = 1;
  SsBufs bf;
  const c128* b;
  const SolveState* s;
  __device__ __forceinline__ void acc(double* part, int i, c128, c128 ax, c128) const {
    const c128 d = b[i] - ax;
    c64 v;
    cvt(v, d);
    bf.v(s->ko, 3)[i] = v;
    part[0] += norm2(d);
  }
};

// One s-step launch with the link storage of `link_prec`.
template <class LK>
static void launch_sscg_lk(hipStream_t st, const DevOperator& o, const LK& lk, const SsGeo& sg, SsBufs bf,
                           const SolveState* rd, SolveState* wr, int q, RedWs ws) {
  const OpDev<LK> od = make_opdev(o, lk);
  const EoGeom& eg = sg.e;
  if (sg.wave && o.kind == OP_STAG_EO_SCHUR)
    hipLaunchKernelGGL((k_sscg_w<LK, SolveState, true>), dim3(sg.s.nblk), dim3(kBlock), 0, st, sg.s, od, bf, rd, wr,
                       q, ws);
  else if (sg.wave)
    hipLaunchKernelGGL((k_sscg_w<LK, SolveState>), dim3(sg.s.nblk), dim3(kBlock), 0, st, sg.s, od, bf, rd, wr, q,
                       ws);
  else if (o.kind == OP_STAG_EO_SCHUR)
    hipLaunchKernelGGL((k_sscg<kSsCPLEo, true, LK, SolveState>), dim3(eg.g.nblk), dim3(kBlock), 0, st, eg, od, bf,
                       rd, wr, q, ws);
  else
    hipLaunchKernelGGL((k_sscg<kSsCPL, false, LK, SolveState>), dim3(eg.g.nblk), dim3(kBlock), 0, st, eg, od, bf,
                       rd, wr, q, ws);
  GLX_CHECK_LAUNCH();
}
static void launch_sscg(hipStream_t st, const DevOperator& o, int link_prec, const SsGeo& eg, SsBufs bf,
                        const SolveState* rd, SolveState* wr, int q, RedWs ws) {
  if (link_prec == LINKS_F16) {
    if (!o.u16 || !o.sc16) throw ArgError("operator has no FP16 links on the device");
    launch_sscg_lk(st, o, LinksF16{o.u16, o.sc16}, eg, bf, rd, wr, q, ws);
  } else {
    if (!o.u32) throw ArgError("operator has no FP32 links on the device");
    launch_sscg_lk(st, o, LinksF32{o.u32}, eg, bf, rd, wr, q, ws);
  }
}

// The s-step basis and FP32 solution increment of one solve, zeroed.
struct SsSpace {
  DevBuf y, xs;
  SsBufs bf;
  SsSpace(int n, hipStream_t st) : y(size_t(10) * n * sizeof(c64)), xs(size_t(n) * sizeof(c64)) {
    GLX_HIP(hipMemsetAsync(y.get(), 0, size_t(10) * n * sizeof(c64), st));
    GLX_HIP(hipMemsetAsync(xs.get(), 0, size_t(n) * sizeof(c64), st));
    bf = SsBufs{y.as<c64>(), xs.as<c64>(), long(n)};
  }
};

SolveResult sscg_impl(const DevOperator& o, c128* x, const c128* b, const SolveOpts& opt, const MixedOpts& mo,
                      int rows) {
  const int n = o.size();
  OwnedStream own(opt.stream);
  hipStream_t st = own.st;
  const SsGeo g = ss_geo(o, rows);
  // two buffers of the sums per block
  Work w(st, std::max(2 * g.nblk(), stencil_blocks(o.lx, o.ly)), opt.max_iter, opt.want_hist);
  SolveState h;
  std::memset(&h, 0, sizeof(h));
  h.max_iter = opt.max_iter;
  h.eps = opt.eps;
  h.delta = mo.delta;
  h.hist = w.hist.as<double>();
  PinnedSlots* pin = noalloc_scope() ? nullptr : &pinned();  // as in fastcg_impl
  if (pin) __atomic_store_n(pin->hdone, 0, __ATOMIC_RELEASE);
  h.hdone = pin ? pin->hdone_dev : nullptr;
  w.upload(h);
  SsSpace sp(n, st);
  DevBuf t64buf(size_t(n) * sizeof(c128));
  c128* t64 = t64buf.as<c128>();
  // R0 of slot 0 = b - A x; P0 of slot 0 = 0 (the first launch is a restart with beta 0)
  apply_typed<c128>(st, o, LINKS_F64, t64, x, w.ws(), kNoGate);
  launch_point(st, n, PtSsInit{sp.bf.v(0, 3), b, t64}, FinFastInit{w.ds}, w.ws(), kNoGate);
  DevBuf shadow_buf(sizeof(SolveState));
  SolveState* shadow = shadow_buf.as<SolveState>();
  GLX_HIP(hipMemcpyAsync(shadow, w.ds, sizeof(SolveState), hipMemcpyDeviceToDevice, st));
  auto enq_iter = [&](int q) {
    const SolveState* rd = q == 0 ? w.ds : shadow;
    SolveState* wr = q == 0 ? shadow : w.ds;
    launch_sscg(st, o, mo.link_prec, g, sp.bf, rd, wr, q, w.ws());
  };
  auto enq_reliable = [&]() {
    launch_point(st, n, PtSsAccum{x, sp.bf.x}, FinNone{}, w.ws(), w.paused(), kGatedBlocks);
    run_stencil<c128>(st, o, LINKS_F64, InPlain<c128>{x}, t64, EpiSsTrue{sp.bf, b, w.ds}, FinFastReliable{w.ds},
                      w.ws(), w.paused());
  };
  // launches alternate state copies: an even number per group ends every
  // group on the home copy, which the reliable update uses
  const int group = 2 * ((std::max(1, opt.group) + 1) / 2);
  auto enq_group = [&]() {
    for (int k = 0; k < group; k++) enq_iter(k & 1);
    enq_reliable();
  };
  hipGraphExec_t exec = nullptr;
  if (opt.use_graph) {
    hipGraph_t graph = nullptr;
    GLX_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    enq_group();
    GLX_HIP(hipStreamEndCapture(st, &graph));
    GLX_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    GLX_HIP(hipGraphDestroy(graph));
  }
  auto launch = [&]() {
    if (exec) GLX_HIP(hipGraphLaunch(exec, st));
    else enq_group();
  };
  // one group queued ahead of the one waited for; the kernels store the done
  // flag into host memory
  const long max_groups = 4L * (opt.max_iter / group + 2) + 16;
  long groups = 0;
  if (opt.max_iter > 0 && !pin) {
    for (;;) {
      launch();
      groups++;
      Mirror mm;
      GLX_HIP(hipMemcpyAsync(&mm, w.ds, sizeof(Mirror), hipMemcpyDeviceToHost, st));
      GLX_HIP(hipStreamSynchronize(st));
      if (mm.done == 1) break;
      if (groups > max_groups) fail_drained(st, exec, "s-step CG made no progress");
    }
  } else if (opt.max_iter > 0) {
    int cur = 0;
    auto submit = [&](int slot) {
      launch();
      groups++;
      GLX_HIP(hipEventRecord(pin->ev_nf[slot], st));
    };
    submit(cur);
    for (;;) {
      submit(cur ^ 1);
      GLX_HIP(hipEventSynchronize(pin->ev_nf[cur]));
      if (__atomic_load_n(pin->hdone, __ATOMIC_ACQUIRE) == 1) break;
      if (groups > max_groups) fail_drained(st, exec, "s-step CG made no progress");
      cur ^= 1;
    }
  }
  GLX_HIP(hipStreamSynchronize(st));
  if (exec) GLX_HIP(hipGraphExecDestroy(exec));
  SolveState hs = w.download();
  SolveResult res;
  res.name = "CG-mixed";
  if (opt.max_iter <= 0) {
    res.iter = 1;
    res.success = true;
    return res;
  }
  res.success = hs.resSq == 0.0 || sqrt(hs.resSq) < opt.eps * hs.bsqrt;
  res.ops_count = hs.kfinal + 2 + hs.reliable;
  res.iter = hs.kfinal + 1;
  res.resSq = hs.resSq;
  res.bsqrt = hs.bsqrt;
  res.reliable_updates = hs.reliable;
  if (opt.want_hist) res.relres_hist = w.hist_host(res.iter);
  return res;
}

// Live timing of the s-step kernel alone: `reps` launches between two HIP
// events, from p = 0 and a constant r (every float 0x3c3c3c3c, about 0.0115)
// with tolerance, delta and max_iter zero, so that no launch stops.  Returns
// ms per launch.
double time_sscg_kernel_impl(const DevOperator& o, int link_prec, int reps, int rows) {
  const int n = o.size();
  OwnedStream own;
  hipStream_t st = own.st;
  const SsGeo g = ss_geo(o, rows);
  Work w(st, std::max(2 * g.nblk(), stencil_blocks(o.lx, o.ly)), 0, false);
  SsSpace sp(n, st);
  GLX_HIP(hipMemsetAsync(sp.bf.v(0, 3), 0x3c, size_t(n) * sizeof(c64), st));
  DevBuf shadow_buf(sizeof(SolveState));
  SolveState* shadow = shadow_buf.as<SolveState>();
  GLX_HIP(hipMemcpyAsync(shadow, w.ds, sizeof(SolveState), hipMemcpyDeviceToDevice, st));
  int q = 0;
  auto launch = [&]() {
    const SolveState* rd = q == 0 ? w.ds : shadow;
    SolveState* wr = q == 0 ? shadow : w.ds;
    launch_sscg(st, o, link_prec, g, sp.bf, rd, wr, q, w.ws());
    q ^= 1;
  };
  launch();  // warm-up (a restart; the rest read all five vectors)
  hipEvent_t a, b;
  GLX_HIP(hipEventCreate(&a));
  GLX_HIP(hipEventCreate(&b));
  GLX_HIP(hipEventRecord(a, st));
  for (int k = 0; k < reps; k++) launch();
  GLX_HIP(hipEventRecord(b, st));
  GLX_HIP(hipEventSynchronize(b));
  float ms = 0.f;
  GLX_HIP(hipEventElapsedTime(&ms, a, b));
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return double(ms) / reps;
}

// Live timing of the wave-strip s-step launch's memory traffic alone
// (k_sscg_w_traffic, the same loads and stores with no arithmetic): `reps`
// launches between two HIP events over a zeroed basis.  Returns ms per
// launch.  Only lattices the s-step form takes.
double time_sscg_traffic_impl(const DevOperator& o, int link_prec, int reps) {
  if (o.kind != OP_STAG_NORMAL || !sscg_fits(o))
    throw ArgError("the s-step traffic timing takes D^dag D on lattices the s-step form takes");
  if (reps < 1) throw ArgError("reps must be >= 1");
  const int n = o.size();
  OwnedStream own;
  hipStream_t st = own.st;
  const SwGeom g = make_swgeom(o.lx, o.ly, sw_rows_for(o.lx, o.ly));
  SsSpace sp(n, st);
  int ko = 0;
  auto launch = [&]() {
    if (link_prec == LINKS_F16) {
      if (!o.u16 || !o.sc16) throw ArgError("operator has no FP16 links on the device");
      hipLaunchKernelGGL((k_sscg_w_traffic<LinksF16>), dim3(g.nblk), dim3(kBlock), 0, st, g,
                         make_opdev(o, LinksF16{o.u16, o.sc16}), sp.bf, ko);
    } else {
      if (!o.u32) throw ArgError("operator has no FP32 links on the device");
      hipLaunchKernelGGL((k_sscg_w_traffic<LinksF32>), dim3(g.nblk), dim3(kBlock), 0, st, g,
                         make_opdev(o, LinksF32{o.u32}), sp.bf, ko);
    }
    GLX_CHECK_LAUNCH();
    ko ^= 1;
  };
  launch();
  hipEvent_t a, b;
  GLX_HIP(hipEventCreate(&a));
  GLX_HIP(hipEventCreate(&b));
  GLX_HIP(hipEventRecord(a, st));
  for (int k = 0; k < reps; k++) launch();
  GLX_HIP(hipEventRecord(b, st));
  GLX_HIP(hipEventSynchronize(b));
  float ms = 0.f;
  GLX_HIP(hipEventElapsedTime(&ms, a, b));
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return double(ms) / reps;
}

// ============================================================ dispatch ====
static bool real_op(const DevOperator& o) {
  return o.is_host() ? !o.host_complex : o.kind == OP_LAPLACE_REAL;
}

SolveResult dev_cg(const DevOperator& op, void* x, const void* b, const SolveOpts& o) {
  if (real_op(op)) return cg_impl<double>(op, static_cast<double*>(x), static_cast<const double*>(b), o);
  return cg_impl<c128>(op, static_cast<c128*>(x), static_cast<const c128*>(b), o);
}
SolveResult dev_bicgstab(const DevOperator& op, void* x, const void* b, const SolveOpts& o) {
  if (real_op(op)) return bicgstab_impl<double>(op, static_cast<double*>(x), static_cast<const double*>(b), o);
  return bicgstab_impl<c128>(op, static_cast<c128*>(x), static_cast<const c128*>(b), o);
}
SolveResult dev_gcr(const DevOperator& op, void* x, const void* b, const SolveOpts& o) {
  if (real_op(op)) return gcr_impl<double>(op, static_cast<double*>(x), static_cast<const double*>(b), o);
  return gcr_impl<c128>(op, static_cast<c128*>(x), static_cast<const c128*>(b), o);
}
SolveResult dev_cg_m(const DevOperator& op, void* const* xs, const void* b, int n_shift, const double* shifts,
                     int resid_freq_check, bool worst_first, const SolveOpts& o) {
  if (real_op(op))
    return cg_m_impl<double>(op, reinterpret_cast<double* const*>(xs), static_cast<const double*>(b), n_shift, shifts,
                             resid_freq_check, worst_first, o);
  return cg_m_impl<c128>(op, reinterpret_cast<c128* const*>(xs), static_cast<const c128*>(b), n_shift, shifts,
                         resid_freq_check, worst_first, o);
}
SolveResult dev_cg_mixed(const DevOperator& op, void* x, const void* b, const SolveOpts& o, const MixedOpts& m) {
  if (!real_op(op) && (m.sstep ? m.sstep : sstep_default()) == 2 && sscg_fits(op)) {
    const int rows = sstep_rows_for(op.lx, op.ly);  // 0: the one-iteration form for this size
    if (rows > 0) return sscg_impl(op, static_cast<c128*>(x), static_cast<const c128*>(b), o, m, rows);
  }
  if (real_op(op))
    return fastcg_impl<double, float, true>(op, static_cast<double*>(x), static_cast<const double*>(b), o, m);
  return fastcg_impl<c128, c64, true>(op, static_cast<c128*>(x), static_cast<const c128*>(b), o, m);
}
SolveResult dev_cg_fused(const DevOperator& op, void* x, const void* b, const SolveOpts& o) {
  MixedOpts m;
  if (real_op(op))
    return fastcg_impl<double, double, false>(op, static_cast<double*>(x), static_cast<const double*>(b), o, m);
  return fastcg_impl<c128, c128, false>(op, static_cast<c128*>(x), static_cast<const c128*>(b), o, m);
}

// ================================================== batched multishift ===
// Many independent CG-M solves (generic_cg_m.cpp:312-600 semantics per
// system: own lattice size, links, mass, right-hand side, the same shifts)
// advanced together.  Every phase of a CG-M iteration is ONE launch over all
// systems: the grid is the concatenation of the systems' block ranges (each
// starting at a multiple of 8, so the XCD-aware tile order of a system is
// kept), a block finds its system in a table, and a system whose solve has
// converged (its own control word) returns at once -- a per-system
// convergence mask.  The finalisers are one block per system.  A group of
// iterations is captured once as a hipGraph and replayed; the host reads all
// control words once per group.  Requires D^dag D lattices the row-marching
// kernel takes (lx % 64 == 0, ly % 4 == 0, ly >= 8).

struct BatchMap {
  const int* sys_of_blk;  // system of each block of the launch
  const int* off;         // first block of each system
  const int* lb_of_blk;   // or, where set, each block's index within its system
  __device__ __forceinline__ int local(int s) const {
    return lb_of_blk ? lb_of_blk[blockIdx.x] : int(blockIdx.x) - off[s];
  }
};

template <class F>
struct PtItem {
  F f;
  int n, nblk;
  RedWs ws;
  Gate gate;
};
template <class F>
__global__ __launch_bounds__(kBlock) void k_point_batch(const PtItem<F>* items, BatchMap bm) {
  const int s = bm.sys_of_blk[blockIdx.x];
  const int lb = bm.local(s);
  const PtItem<F> it = items[s];
  if (lb >= it.nblk || gate_closed(it.gate)) return;
  F f = it.f;
  point_body<F::NR, F>(it.n, f, it.ws, lb, it.nblk);
}

template <class Fin>
struct FinItem {
  Fin fin;
  int nblk;
  RedWs ws;
  Gate gate;
};
template <int NR, class Fin>
__global__ __launch_bounds__(kBlock) void k_finalize_batch(const FinItem<Fin>* items) {
  const FinItem<Fin> it = items[blockIdx.x];
  if (gate_closed(it.gate)) return;
  finalize_body<NR, Fin>(it.nblk, it.fin, it.ws);
}

template <class T, class In, class Epi>
struct NrmItem {
  NGeom g;
  OpDev<LinksF64> op;
  In in;
  T* out;
  Epi epi;
  RedWs ws;
  Gate gate;
};
template <class T, class In, class Epi>
__global__ __launch_bounds__(kBlock) void k_normal_batch(const NrmItem<T, In, Epi>* items, BatchMap bm) {
  static_assert(!has_selffin<In>::value, "a batched launch has no self-finalising form");
  const int s = bm.sys_of_blk[blockIdx.x];
  const int lb = bm.local(s);
  const NGeom g = items[s].g;
  if (lb >= g.nblk || gate_closed(items[s].gate)) return;
  const OpDev<LinksF64> op = items[s].op;
  In in = items[s].in;
  const Epi epi = items[s].epi;
  normal_body<normal_cpl<T>(), T, LinksF64, In, Epi>(g, op, in, items[s].out, epi, items[s].ws, lb);
}

// r = p = p_s = b, x_s = 0 for every shift; |b|^2
template <class T>
struct PtCgmStart {
  static constexpr int NR = 1;
  const T* b;
  T* r;
  T* p;
  T* const* ptab;
  T* const* xtab;
  int n_shift;
  __device__ void init() {}
  __device__ void operator()(int i, double* part) {
    const T bi = b[i];
    r[i] = bi;
    p[i] = bi;
    for (int k = 0; k < n_shift; k++) {
      ptab[k][i] = bi;
      xtab[k][i] = zero_of<T>();
    }
    part[0] += norm2(bi);
  }
};

// ------------------------------------ deferred shift updates (batched) ----
// In multishift CG the shifted vectors p_s and x_s never feed back into the
// iteration: beta, zeta_s, beta_s, alpha, alpha_s and the convergence test
// come from r, p and Ap alone (generic_cg_m.cpp:396-505).  The batched solver
// therefore defers their updates over a window of at most kWin iterations
// k0 .. k: with r^(j) the residual of iteration j,
//   p_s^(k+1) = cp p_s^(k0) + sum_j dp_j r^(k0+1+j)
//   x_s^(k+1) = x_s^(k0) + ax p_s^(k0) + sum_j ex_j r^(k0+1+j)
// where the coefficients follow the reference's steps, x_s -= beta_s p_s
// (:414-420) and p_s = zeta_s r + alpha_s p_s (:508-514), as FP64 scalars on
// the finaliser.  A window ends at every convergence check (so a shift that
// converges has its x_s completed there) and after kWin iterations; the next
// residual pass applies it site by site, reading each x_s and p_s once per
// window instead of every iteration.  The iteration itself (and so every
// iteration count and convergence decision) is the reference's; only the
// rounding of x_s differs.  The residuals of a window live in a ring of
// kWin + 1 vectors, r^(j) in slot j mod (kWin + 1).
constexpr int kWin = 10;
constexpr int kRing = kWin + 1;
struct CgmWin {
  int wlen;   // residuals in the window (terms of the p_s coefficients)
  int flush;  // the window ended at the last finaliser: the next residual pass applies it
  long long upd_r, upd_x, upd_p;  // applied: residual terms, x_s and p_s updates (site-independent counts)
  c128 cp[kMaxShift], ax[kMaxShift];
  c128 dp[kMaxShift][kWin], ex[kMaxShift][kWin];
};

// The window's terms for one site: r^(k0+1+j), j < wlen, from the ring.
template <class T>
__device__ __forceinline__ void cgm_window_terms(const T* ring, long n, int kcur, int wlen, int i, T (&rj)[kWin]) {
#pragma unroll
  for (int j = 0; j < kWin; j++) {
    if (j < wlen) {
      const int slot = (kcur - wlen + 1 + j + 2 * kRing) % kRing;
      rj[j] = ring[slot * n + i];
    }
  }
}

// If the last finaliser ended a window: x_s += ax p_s + sum ex_j r_j for the
// shifts active before that check, p_s = cp p_s + sum dp_j r_j for those
// still active (a launch that finds no ended window returns at once).  The
// window's scalars and the shifts' vectors are staged in LDS once per block:
// read there, every lane of a wave reads one address (a broadcast), where
// loads from the state in global memory would be one vector load per lane
// and coefficient.
template <class T>
struct PtCgmWindow {
  static constexpr int NR = 0;
  static constexpr int kU = 4;              // shifts whose loads are in flight together
  static constexpr int kC = 2 * kWin + 2;   // per shift: ax, cp, ex[kWin], dp[kWin]
  const T* ring;
  long n;
  T* const* xtab;
  T* const* ptab;
  const SolveState* s;
  const CgmWin* w;
  int kcur, flush, wlen, nx, np;
  // the block's LDS tables (used through these functions, so that every
  // access is known to be an LDS access)
  __device__ static c128* lds_c() {
    __shared__ c128 a[kMaxShift * kC];
    return a;
  }
  __device__ static T** lds_x() {
    __shared__ T* a[kMaxShift];
    return a;
  }
  __device__ static T** lds_p() {
    __shared__ T* a[kMaxShift];
    return a;
  }
  __device__ void init() {
    // (wave-uniform: the branches on them are scalar)
    flush = __builtin_amdgcn_readfirstlane(w->flush);
    if (!flush) return;
    kcur = __builtin_amdgcn_readfirstlane(s->k);
    wlen = __builtin_amdgcn_readfirstlane(w->wlen);
    nx = __builtin_amdgcn_readfirstlane(s->n_xupd);
    np = __builtin_amdgcn_readfirstlane(s->n_rem);
    c128* lc = lds_c();
    for (int t = threadIdx.x; t < nx * kC; t += blockDim.x) {
      const int q = t / kC, c = t % kC;
      lc[t] = c == 0 ? w->ax[q] : (c == 1 ? w->cp[q] : (c < 2 + kWin ? w->ex[q][c - 2] : w->dp[q][c - 2 - kWin]));
    }
    T** lx = lds_x();
    T** lp = lds_p();
    for (int q = threadIdx.x; q < nx; q += blockDim.x) {
      const int v = s->vec[q];
      lx[q] = xtab[v];
      lp[q] = ptab[v];
    }
    __syncthreads();
  }
  // the window of shift slot q at one site (the loops run over the fixed
  // window length: a term index known at compile time keeps rj in registers)
  __device__ __forceinline__ void apply(int q, T ps, T xo, const T (&rj)[kWin], T* xp, T* pp) const {
    const c128* c = lds_c() + q * kC;
    T xa = xo + c[0] * ps;
#pragma unroll
    for (int j = 0; j < kWin; j++)
      if (j < wlen) xa = xa + c[2 + j] * rj[j];
    *xp = xa;
    if (q < np) {
      T pa = c[1] * ps;
#pragma unroll
      for (int j = 0; j < kWin; j++)
        if (j < wlen) pa = pa + c[2 + kWin + j] * rj[j];
      *pp = pa;
    }
  }
  __device__ void operator()(int i, double*) {
    if (!flush) return;
    T rj[kWin];
    cgm_window_terms(ring, n, kcur, wlen, i, rj);
    T* const* sx = lds_x();
    T* const* sp = lds_p();
    // kU shifts at a time: one shift after another would wait for each
    // shift's loads in turn
    int q = 0;
    for (; q + kU <= nx; q += kU) {
      T* xv[kU];
      T* pv[kU];
      T ps[kU], xo[kU];
#pragma unroll
      for (int u = 0; u < kU; u++) {
        xv[u] = sx[q + u] + i;
        pv[u] = sp[q + u] + i;
      }
#pragma unroll
      for (int u = 0; u < kU; u++) {
        ps[u] = *pv[u];
        xo[u] = *xv[u];
      }
#pragma unroll
      for (int u = 0; u < kU; u++) apply(q + u, ps[u], xo[u], rj, xv[u], pv[u]);
    }
    for (; q < nx; q++) apply(q, sp[q][i], sx[q][i], rj, sx[q] + i, sp[q] + i);
  }
};
// r^(k+1) = r^(k) + beta Ap (generic_cg_m.cpp:423-426) into the residual
// ring, and |r^(k+1)|^2 (after PtCgmWindow, which reads the window's slots;
// slot k + 1 is not among them).
template <class T>
struct PtCgmRing {
  static constexpr int NR = 1;
  T* ring;
  long n;
  const T* Ap;
  const SolveState* s;
  typename Scal<T>::type be;
  int kcur;
  __device__ void init() {
    be = as_scal<T>(s->beta);
    kcur = s->k;
  }
  __device__ void operator()(int i, double* part) {
    const T rn = ring[(kcur % kRing) * n + i] + scal(be, Ap[i]);
    ring[((kcur + 1) % kRing) * n + i] = rn;
    part[0] += norm2(rn);
  }
};
// After the loop: the x_s part of the open window (the solve stopped before
// its p_s update), for the shifts active before the last check.
template <class T>
struct PtCgmFlushX {
  static constexpr int NR = 0;
  const T* ring;
  long n;
  T* const* xtab;
  T* const* ptab;
  const SolveState* s;
  const CgmWin* w;
  int kcur, wlen, nx;
  __device__ void init() {
    kcur = s->k;
    wlen = w->wlen;
    nx = s->n_xupd;
  }
  __device__ void operator()(int i, double*) {
    T rj[kWin];
    cgm_window_terms(ring, n, kcur, wlen, i, rj);
    for (int q = 0; q < nx; q++) {
      const int v = s->vec[q];
      T xa = xtab[v][i] + w->ax[q] * ptab[v][i];
#pragma unroll
      for (int j = 0; j < kWin; j++)
        if (j < wlen) xa = xa + w->ex[q][j] * rj[j];
      xtab[v][i] = xa;
    }
  }
};
// FinCgmR with the window's coefficients: the x_s step of this iteration,
// the check (permuting the coefficients with the shifts), alpha, alpha_s, the
// p_s step, and whether the window ends here.
struct FinCgmRWin {
  SolveState* s;
  CgmWin* w;
  __device__ static void swp(c128& a, c128& b) {
    const c128 t = a;
    a = b;
    b = t;
  }
  __device__ void operator()(const double* v) const {
    const double rn = v[0];
    const int k = s->k;
    if (s->hist) s->hist[k] = sqrt(rn) / s->bsqrt;
    const double tol = s->eps * s->bsqrt;
    if (w->flush) {  // the previous window was applied by this iteration's window launch
      w->flush = 0;
      w->wlen = 0;
      for (int q = 0; q < s->n_rem; q++) {
        w->cp[q] = {1.0, 0.0};
        w->ax[q] = {0.0, 0.0};
        for (int j = 0; j < kWin; j++) w->dp[q][j] = w->ex[q][j] = {0.0, 0.0};
      }
    }
    const int wl = w->wlen;
    s->n_xupd = s->n_rem;
    // x_s -= beta_s p_s (generic_cg_m.cpp:414-420) on the coefficients
    for (int q = 0; q < s->n_rem; q++) {
      const c128 b = s->beta_s[q];
      w->ax[q] = w->ax[q] - b * w->cp[q];
      for (int j = 0; j < wl; j++) w->ex[q][j] = w->ex[q][j] - b * w->dp[q][j];
    }
    if (k % s->resid_freq == 0) {
      for (int n = 0; n < s->n_rem; n++) {
        const double zabs = hypot(s->zeta[n].re, s->zeta[n].im);
        if (zabs * sqrt(rn) < tol) {
          s->n_rem--;
          const int e = s->n_rem;
          if (e != n) {
            int t = s->mapping[e]; s->mapping[e] = s->mapping[n]; s->mapping[n] = t;
            t = s->vec[e]; s->vec[e] = s->vec[n]; s->vec[n] = t;
            swp(s->alpha_s[e], s->alpha_s[n]);
            swp(s->beta_s[e], s->beta_s[n]);
            swp(s->zeta[e], s->zeta[n]);
            swp(s->zeta_p[e], s->zeta_p[n]);
            const double d = s->shifts[e]; s->shifts[e] = s->shifts[n]; s->shifts[n] = d;
            swp(w->cp[e], w->cp[n]);
            swp(w->ax[e], w->ax[n]);
            for (int j = 0; j < kWin; j++) {
              swp(w->dp[e][j], w->dp[n][j]);
              swp(w->ex[e][j], w->ex[n][j]);
            }
            n--;
          }
        }
      }
    }
    const double z0 = hypot(s->zeta[0].re, s->zeta[0].im);
    if ((s->worst_first && z0 * sqrt(rn) < tol) || s->n_rem == 0 || k == s->max_iter - 1) {
      s->kfinal = k;
      s->done = 1;
      return;
    }
    const double al = rn / s->rsq;
    s->alpha = {al, 0.0};
    s->rsq = rn;
    for (int n = 0; n < s->n_rem; n++)
      s->alpha_s[n] = cdiv((s->alpha * s->zeta[n]) * s->beta_s[n], s->zeta_p[n] * s->beta);
    // p_s = zeta_s r^(k+1) + alpha_s p_s (generic_cg_m.cpp:508-514) on the coefficients
    for (int q = 0; q < s->n_rem; q++) {
      const c128 a = s->alpha_s[q];
      w->cp[q] = a * w->cp[q];
      for (int j = 0; j < wl; j++) w->dp[q][j] = a * w->dp[q][j];
      w->dp[q][wl] = s->zeta[q];
    }
    w->wlen = wl + 1;
    if (k % s->resid_freq == 0 || w->wlen == kWin) {
      w->flush = 1;
      w->upd_r += w->wlen;
      w->upd_x += s->n_xupd;
      w->upd_p += s->n_rem;
    }
    s->k = k + 1;
  }
};
// Prologue of the batched D^dag D launch: p^(k+1) = r^(k+1) + alpha p^(k)
// (generic_cg_m.cpp:516-518, the arithmetic of PtCgmP) formed while staging,
// stored on the own sites into the other p buffer.
template <class T>
struct InCgmP {
  static constexpr bool kAux = false;
  const T* ring;
  long n;
  const T* pold;
  T* pnew;
  const SolveState* s;
  const T* r;
  typename Scal<T>::type al;
  __device__ void init() {
    r = ring + (s->k % kRing) * n;
    al = as_scal<T>(s->alpha);
  }
  __device__ T* target(T* out) const { return out; }
  struct Raw {
    T r, po;
  };
  __device__ Raw fetch(int j, bool) const { return Raw{r[j], pold[j]}; }
  __device__ T finish(const Raw& w, int j, bool own, T&) const {
    const T pn = w.r + scal(al, w.po);
    if (own) pnew[j] = pn;
    return pn;
  }
  __device__ T load(int j, bool own, T& aux) const { return finish(fetch(j, own), j, own, aux); }
};

// Device copies of one launch kind's items and its block table.
struct BatchLaunch {
  DevBuf items, sys_of_blk, off, lb_of_blk;
  int nblk_total = 0;
  BatchMap map() const { return BatchMap{sys_of_blk.as<int>(), off.as<int>(), lb_of_blk.as<int>()}; }
};
template <class Item>
static BatchLaunch make_batch_launch(const std::vector<Item>& items, const std::vector<int>& nblk) {
  BatchLaunch b;
  std::vector<int> sob, off(items.size());
  for (size_t s = 0; s < items.size(); s++) {
    off[s] = int(sob.size());
    const int nb = ((nblk[s] + 7) / 8) * 8;
    for (int k = 0; k < nb; k++) sob.push_back(int(s));
  }
  b.nblk_total = int(sob.size());
  b.items = DevBuf(items.size() * sizeof(Item));
  b.sys_of_blk = DevBuf(sob.size() * sizeof(int));
  b.off = DevBuf(off.size() * sizeof(int));
  GLX_HIP(hipMemcpy(b.items.get(), items.data(), items.size() * sizeof(Item), hipMemcpyHostToDevice));
  GLX_HIP(hipMemcpy(b.sys_of_blk.get(), sob.data(), sob.size() * sizeof(int), hipMemcpyHostToDevice));
  GLX_HIP(hipMemcpy(b.off.get(), off.data(), off.size() * sizeof(int), hipMemcpyHostToDevice));
  return b;
}

// The same launch with the systems that share one lattice's links (`group`:
// equal keys, equal block counts) interleaved: block B of a group of n
// systems with 8 per blocks each is system r = (B / 8) % n at index
// 8 (B / 8n) + B % 8 within it.  The hardware deals block ids to the 8 XCDs
// in turn, so the n blocks of one tile (the tile xcd_tile gives that index)
// run on one XCD one after another, and the tile's links come from memory
// once and from that XCD's L2 for the other right-hand sides.
template <class Item, class Key>
static BatchLaunch make_batch_launch_shared(const std::vector<Item>& items, const std::vector<int>& nblk,
                                            const std::vector<Key>& group) {
  BatchLaunch b;
  const int ns = int(items.size());
  std::vector<int> sob, lbb, off(ns, 0);
  std::vector<bool> done(ns, false);
  for (int s0 = 0; s0 < ns; s0++) {
    if (done[s0]) continue;
    std::vector<int> g;
    for (int s = s0; s < ns; s++)
      if (!done[s] && group[s] == group[s0] && nblk[s] == nblk[s0]) {
        g.push_back(s);
        done[s] = true;
      }
    const int n = int(g.size()), per = ((nblk[s0] + 7) / 8);
    for (int j = 0; j < per; j++)
      for (int r = 0; r < n; r++)
        for (int x = 0; x < 8; x++) {
          sob.push_back(g[r]);
          lbb.push_back(8 * j + x);
        }
  }
  b.nblk_total = int(sob.size());
  b.items = DevBuf(items.size() * sizeof(Item));
  b.sys_of_blk = DevBuf(sob.size() * sizeof(int));
  b.off = DevBuf(off.size() * sizeof(int));
  b.lb_of_blk = DevBuf(lbb.size() * sizeof(int));
  GLX_HIP(hipMemcpy(b.items.get(), items.data(), items.size() * sizeof(Item), hipMemcpyHostToDevice));
  GLX_HIP(hipMemcpy(b.sys_of_blk.get(), sob.data(), sob.size() * sizeof(int), hipMemcpyHostToDevice));
  GLX_HIP(hipMemcpy(b.off.get(), off.data(), off.size() * sizeof(int), hipMemcpyHostToDevice));
  GLX_HIP(hipMemcpy(b.lb_of_blk.get(), lbb.data(), lbb.size() * sizeof(int), hipMemcpyHostToDevice));
  return b;
}

constexpr int kBatchRows = 32;

template <class T>
std::vector<SolveResult> cg_m_batch_impl(const std::vector<BatchSystem>& sys, int n_shift, const double* shifts,
                                         int resid_freq, bool worst_first, const SolveOpts& opt, BatchStats* stats,
                                         bool time_phases) {
  static_assert(std::is_same<T, c128>::value, "the batched multishift solver is complex FP64");
  if (sys.empty()) return {};
  if (n_shift < 1 || n_shift > kMaxShift) throw ArgError("n_shift must be in [1, 64]");
  if (resid_freq < 1) throw ArgError("resid_freq_check must be >= 1");
  const int ns = int(sys.size());
  std::vector<int> nsite(ns), rows(ns);
  for (int s = 0; s < ns; s++) {
    const DevOperator& o = sys[s].op;
    if (o.kind != OP_STAG_NORMAL || !o.u64) throw ArgError("batched systems must be staggered normal operators");
    // rows per block of D^dag D: the batch's grid holds every system's
    // blocks, so it takes long strips (little halo) whatever shape a lone
    // lattice of the size would take (measured on the 48-lattice batch, rows
    // for all sizes: 8 128 ms, 16 109, 24 107, 32 105, 48 106, 64 109 ms of
    // D^dag D)
    rows[s] = kBatchRows;
    if (!normal_fits(o.lx, o.ly, rows[s], normal_cpl<T>())) rows[s] = 16;
    if (!normal_fits(o.lx, o.ly, rows[s], normal_cpl<T>()))
      throw ArgError("batched system " + std::to_string(s) + ": lattice " + std::to_string(o.lx) + "x" +
                     std::to_string(o.ly) + " needs lx % 64 == 0, ly % 4 == 0, ly >= 8");
    if (!sys[s].b || !sys[s].xs) throw ArgError("batched system " + std::to_string(s) + ": b and x are required");
    nsite[s] = o.size();
  }
  OwnedStream own(opt.stream);
  hipStream_t st = own.st;
  // one allocation for every system's residual ring (kRing vectors), p in two
  // buffers (the D^dag D launch forms the next p while reading the last),
  // Ap and the n_shift p_s
  constexpr int kLead = kRing + 3;
  size_t total = 0;
  std::vector<size_t> vbase(ns);
  for (int s = 0; s < ns; s++) {
    vbase[s] = total;
    total += size_t(kLead + n_shift) * nsite[s];
  }
  DevBuf vecs(total * sizeof(T));
  DevBuf states(size_t(ns) * sizeof(SolveState));
  DevBuf resv(size_t(ns) * n_shift * sizeof(double));
  DevBuf tabs(size_t(ns) * 2 * n_shift * sizeof(T*));
  SolveState* ds = states.as<SolveState>();
  T* V = vecs.as<T>();
  T** tab = tabs.as<T*>();
  std::vector<T*> htab(size_t(ns) * 2 * n_shift);
  std::vector<NGeom> geo(ns);
  std::vector<int> nb_pt(ns), nb_st(ns), nb_w(ns), nb_r(ns);
  int max_nb = 0;
  for (int s = 0; s < ns; s++) {
    const DevOperator& o = sys[s].op;
    geo[s] = make_ngeom(o.lx, o.ly, rows[s], normal_cpl<T>());
    nb_pt[s] = point_blocks(nsite[s]);
    // the residual pass: 4 sites per thread, the window launch: 16 (a block
    // of a converged system, or of a window launch with no ended window,
    // costs a few dependent loads before it returns: fewer blocks make such
    // launches cheap)
    nb_r[s] = std::max(1, (nb_pt[s] + 3) / 4);
    nb_w[s] = std::max(1, (nb_pt[s] + 15) / 16);
    nb_st[s] = geo[s].nblk;
    max_nb = std::max(max_nb, std::max(nb_pt[s], nb_st[s]));
    for (int k = 0; k < n_shift; k++) {
      htab[(size_t(s) * 2 + 0) * n_shift + k] = static_cast<T*>(sys[s].xs[k]);
      htab[(size_t(s) * 2 + 1) * n_shift + k] = V + vbase[s] + size_t(kLead + k) * nsite[s];
    }
  }
  // partial sums: each system owns 2 * max_nb doubles (NR <= 2 everywhere)
  DevBuf partB(size_t(ns) * 2 * max_nb * sizeof(double));
  GLX_HIP(hipMemcpy(tab, htab.data(), htab.size() * sizeof(T*), hipMemcpyHostToDevice));
  std::vector<SolveState> hs(ns);
  for (int s = 0; s < ns; s++) {
    SolveState& h = hs[s];
    std::memset(&h, 0, sizeof(h));
    h.max_iter = opt.max_iter;
    h.eps = opt.eps;
    h.hist = nullptr;
    h.n_rem = n_shift;
    h.n_shift = n_shift;
    h.resid_freq = resid_freq;
    h.worst_first = worst_first ? 1 : 0;
    h.beta = {1.0, 0.0};
    h.alpha = {0.0, 0.0};
    for (int k = 0; k < n_shift; k++) {
      h.zeta[k] = h.zeta_p[k] = h.beta_s[k] = {1.0, 0.0};
      h.alpha_s[k] = {0.0, 0.0};
      h.shifts[k] = shifts[k];
      h.mapping[k] = k;
      h.vec[k] = k;
    }
    h.resvec = resv.as<double>() + size_t(s) * n_shift;
  }
  GLX_HIP(hipMemcpy(ds, hs.data(), size_t(ns) * sizeof(SolveState), hipMemcpyHostToDevice));
  // the deferred-update windows: empty, cp = 1
  DevBuf winB(size_t(ns) * sizeof(CgmWin));
  CgmWin* dwin = winB.as<CgmWin>();
  {
    std::vector<CgmWin> hw(ns);
    for (CgmWin& w : hw) {
      std::memset(&w, 0, sizeof(w));
      for (int k = 0; k < kMaxShift; k++) w.cp[k] = {1.0, 0.0};
    }
    GLX_HIP(hipMemcpy(dwin, hw.data(), size_t(ns) * sizeof(CgmWin), hipMemcpyHostToDevice));
  }

  auto ws_of = [&](int s) { return RedWs{partB.as<double>() + size_t(s) * 2 * max_nb}; };
  auto gate_of = [&](int s) { return Gate{&ds[s].done, 0}; };
  auto ring = [&](int s) { return V + vbase[s]; };
  auto pbuf = [&](int s, int c) { return V + vbase[s] + size_t(kRing + c) * nsite[s]; };
  auto apvec = [&](int s) { return V + vbase[s] + size_t(kRing + 2) * nsite[s]; };
  T* const* xtab_of_base = tab;
  auto xtab = [&](int s) { return xtab_of_base + (size_t(s) * 2 + 0) * n_shift; };
  auto ptab = [&](int s) { return xtab_of_base + (size_t(s) * 2 + 1) * n_shift; };

  using InP = InCgmP<T>;
  using NrmP = NrmItem<T, InP, EpiDotInOut<T>>;
  std::vector<PtItem<PtCgmStart<T>>> i_start(ns);
  std::vector<FinItem<FinCgmInit>> f_start(ns);
  std::vector<NrmItem<T, InPlain<T>, EpiDotInOut<T>>> i_st0(ns);
  std::vector<NrmP> i_st[2] = {std::vector<NrmP>(ns), std::vector<NrmP>(ns)};
  std::vector<FinItem<FinCgmBeta<T>>> f_st(ns);
  std::vector<PtItem<PtCgmWindow<T>>> i_w(ns);
  std::vector<PtItem<PtCgmRing<T>>> i_x(ns);
  std::vector<FinItem<FinCgmRWin>> f_x(ns);
  std::vector<PtItem<PtCgmFlushX<T>>> i_last(ns);
  for (int s = 0; s < ns; s++) {
    const Gate none = kNoGate;
    const long n = nsite[s];
    const OpDev<LinksF64> od = make_opdev(sys[s].op, LinksF64{sys[s].op.u64});
    i_start[s] = {PtCgmStart<T>{static_cast<const T*>(sys[s].b), ring(s), pbuf(s, 0), ptab(s), xtab(s), n_shift},
                  nsite[s], nb_pt[s], ws_of(s), none};
    f_start[s] = {FinCgmInit{ds + s}, nb_pt[s], ws_of(s), none};
    i_st0[s] = {geo[s], od, InPlain<T>{pbuf(s, 0)}, apvec(s), EpiDotInOut<T>{}, ws_of(s), gate_of(s)};
    // iteration k forms p^(k+1) from p^(k): buffer k & 1 -> (k + 1) & 1
    for (int c = 0; c < 2; c++)
      i_st[c][s] = {geo[s], od, InP{ring(s), n, pbuf(s, c), pbuf(s, c ^ 1), ds + s, nullptr, {}}, apvec(s),
                    EpiDotInOut<T>{}, ws_of(s), gate_of(s)};
    f_st[s] = {FinCgmBeta<T>{ds + s}, nb_st[s], ws_of(s), gate_of(s)};
    i_w[s] = {PtCgmWindow<T>{ring(s), n, xtab(s), ptab(s), ds + s, dwin + s, 0, 0, 0, 0, 0}, nsite[s], nb_w[s],
              ws_of(s), gate_of(s)};
    i_x[s] = {PtCgmRing<T>{ring(s), n, apvec(s), ds + s, {}, 0}, nsite[s], nb_r[s], ws_of(s), gate_of(s)};
    f_x[s] = {FinCgmRWin{ds + s, dwin + s}, nb_r[s], ws_of(s), gate_of(s)};
    i_last[s] = {PtCgmFlushX<T>{ring(s), n, xtab(s), ptab(s), ds + s, dwin + s, 0, 0, 0}, nsite[s], nb_pt[s],
                 ws_of(s), kNoGate};
  }
  BatchLaunch L_start = make_batch_launch(i_start, nb_pt);
  BatchLaunch F_start = make_batch_launch(f_start, std::vector<int>(ns, 1));
  // D^dag D: the right-hand sides of one lattice share its links
  std::vector<const void*> link_key(ns);
  for (int s = 0; s < ns; s++) link_key[s] = sys[s].op.u64;
  BatchLaunch L_st0 = make_batch_launch_shared(i_st0, nb_st, link_key);
  BatchLaunch L_st[2] = {make_batch_launch_shared(i_st[0], nb_st, link_key),
                         make_batch_launch_shared(i_st[1], nb_st, link_key)};
  BatchLaunch F_st = make_batch_launch(f_st, std::vector<int>(ns, 1));
  BatchLaunch L_w = make_batch_launch(i_w, nb_w);
  BatchLaunch L_x = make_batch_launch(i_x, nb_r);
  BatchLaunch F_x = make_batch_launch(f_x, std::vector<int>(ns, 1));
  BatchLaunch L_last = make_batch_launch(i_last, nb_pt);

  // Optional HIP-event timing of each phase: per slot of the two-deep
  // submission pipeline, one pair of events around every phase launch of a
  // group; a slot's events are read after its group is waited for.  Groups
  // hold an even number of iterations: every group starts on p buffer 0.
  const int group = 2 * ((std::max(1, opt.group) + 1) / 2);
  const bool timing = time_phases && stats;
  std::vector<hipEvent_t> tev;
  if (timing) {
    tev.resize(size_t(2) * group * 6);
    for (auto& e : tev) GLX_HIP(hipEventCreate(&e));
  }
  int t_slot = 0, t_iter = 0;
  auto tmark = [&](int phase, int edge) {
    if (timing) GLX_HIP(hipEventRecord(tev[((size_t(t_slot) * group + t_iter) * 3 + phase) * 2 + edge], st));
  };
  double t_ms[3] = {0, 0, 0};
  auto tcollect = [&](int slot, int n_it) {
    if (!timing) return;
    for (int it = 0; it < n_it; it++)
      for (int ph = 0; ph < 2; ph++) {  // D^dag D, the residual pass (with the windows it applies)
        float ms = 0.f;
        const size_t e = ((size_t(slot) * group + it) * 3 + ph) * 2;
        GLX_HIP(hipEventElapsedTime(&ms, tev[e], tev[e + 1]));
        t_ms[ph] += ms;
      }
  };

  auto launch_beta = [&]() {
    hipLaunchKernelGGL((k_finalize_batch<EpiDotInOut<T>::NR, FinCgmBeta<T>>), dim3(ns), dim3(kBlock), 0, st,
                       F_st.items.as<FinItem<FinCgmBeta<T>>>());
    GLX_CHECK_LAUNCH();
  };
  // iteration at group position g: the window that ended at the last
  // iteration (a launch that returns at once where none did), the residual
  // pass and its finaliser, then D^dag D on p^(k+1) formed in the launch's
  // prologue, and beta / zeta
  auto enq_iter = [&]() {
    const int c = t_iter & 1;
    tmark(1, 0);
    hipLaunchKernelGGL((k_point_batch<PtCgmWindow<T>>), dim3(L_w.nblk_total), dim3(kBlock), 0, st,
                       L_w.items.as<PtItem<PtCgmWindow<T>>>(), L_w.map());
    GLX_CHECK_LAUNCH();
    hipLaunchKernelGGL((k_point_batch<PtCgmRing<T>>), dim3(L_x.nblk_total), dim3(kBlock), 0, st,
                       L_x.items.as<PtItem<PtCgmRing<T>>>(), L_x.map());
    GLX_CHECK_LAUNCH();
    tmark(1, 1);
    hipLaunchKernelGGL((k_finalize_batch<1, FinCgmRWin>), dim3(ns), dim3(kBlock), 0, st,
                       F_x.items.as<FinItem<FinCgmRWin>>());
    GLX_CHECK_LAUNCH();
    tmark(0, 0);
    hipLaunchKernelGGL((k_normal_batch<T, InP, EpiDotInOut<T>>), dim3(L_st[c].nblk_total), dim3(kBlock), 0, st,
                       L_st[c].items.as<NrmP>(), L_st[c].map());
    GLX_CHECK_LAUNCH();
    tmark(0, 1);
    launch_beta();
    t_iter++;
  };
  // start: r = p = p_s = b, x_s = 0, |b|; then Ap, beta, zeta (the first
  // stencil of generic_cg_m.cpp:100-122 precedes the loop)
  hipLaunchKernelGGL((k_point_batch<PtCgmStart<T>>), dim3(L_start.nblk_total), dim3(kBlock), 0, st,
                     L_start.items.as<PtItem<PtCgmStart<T>>>(), L_start.map());
  GLX_CHECK_LAUNCH();
  hipLaunchKernelGGL((k_finalize_batch<1, FinCgmInit>), dim3(ns), dim3(kBlock), 0, st,
                     F_start.items.as<FinItem<FinCgmInit>>());
  GLX_CHECK_LAUNCH();
  hipLaunchKernelGGL((k_normal_batch<T, InPlain<T>, EpiDotInOut<T>>), dim3(L_st0.nblk_total), dim3(kBlock), 0, st,
                     L_st0.items.as<NrmItem<T, InPlain<T>, EpiDotInOut<T>>>(), L_st0.map());
  GLX_CHECK_LAUNCH();
  launch_beta();

  hipGraphExec_t exec = nullptr;
  if (opt.use_graph && !timing) {
    hipGraph_t graph = nullptr;
    GLX_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    for (int g = 0; g < group; g++) enq_iter();
    GLX_HIP(hipStreamEndCapture(st, &graph));
    GLX_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    GLX_HIP(hipGraphDestroy(graph));
  }
  // control words of all systems: the first 16 bytes of each state, strided
  Mirror* pin = nullptr;
  GLX_HIP(hipHostMalloc(reinterpret_cast<void**>(&pin), 2 * size_t(ns) * sizeof(Mirror), hipHostMallocDefault));
  hipEvent_t ev[2];
  GLX_HIP(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
  GLX_HIP(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
  auto cleanup = [&]() {
    (void)hipStreamSynchronize(st);
    if (exec) (void)hipGraphExecDestroy(exec);
    (void)hipEventDestroy(ev[0]);
    (void)hipEventDestroy(ev[1]);
    (void)hipHostFree(pin);
    for (auto& e : tev) (void)hipEventDestroy(e);
  };
  auto submit = [&](int slot) {
    t_slot = slot;
    t_iter = 0;
    if (exec) GLX_HIP(hipGraphLaunch(exec, st));
    else
      for (int g = 0; g < group; g++) enq_iter();
    GLX_HIP(hipMemcpy2DAsync(pin + size_t(slot) * ns, sizeof(Mirror), ds, sizeof(SolveState), sizeof(Mirror), ns,
                             hipMemcpyDeviceToHost, st));
    GLX_HIP(hipEventRecord(ev[slot], st));
  };
  try {
    if (opt.max_iter > 0) {
      const long max_groups = 4L * (opt.max_iter / group + 2) + 16;
      long groups = 0;
      int cur = 0;
      submit(cur);
      groups++;
      for (;;) {
        submit(cur ^ 1);
        groups++;
        GLX_HIP(hipEventSynchronize(ev[cur]));
        tcollect(cur, group);
        bool all = true;
        for (int s = 0; s < ns && all; s++) all = pin[size_t(cur) * ns + s].done == 1;
        if (all) {
          GLX_HIP(hipEventSynchronize(ev[cur ^ 1]));  // the group queued ahead (its kernels return at once)
          tcollect(cur ^ 1, group);
          break;
        }
        if (groups > max_groups) throw HipError("batched multishift CG made no progress");  // cleanup() drains
        cur ^= 1;
      }
    }
  } catch (...) {
    cleanup();
    throw;
  }
  cleanup();
  // the x_s part of each system's open window
  hipLaunchKernelGGL((k_point_batch<PtCgmFlushX<T>>), dim3(L_last.nblk_total), dim3(kBlock), 0, st,
                     L_last.items.as<PtItem<PtCgmFlushX<T>>>(), L_last.map());
  GLX_CHECK_LAUNCH();

  // per system and shift: |(A + sigma_k) x_k - b|^2 (generic_cg_m.cpp:271-282)
  {
    std::vector<NrmItem<T, InPlain<T>, EpiNone<T>>> i_ax(ns);
    std::vector<PtItem<PtShiftRes<T>>> i_res(ns);
    std::vector<FinItem<FinResSq>> f_res(ns);
    for (int k = 0; k < n_shift; k++) {
      for (int s = 0; s < ns; s++) {
        T* xk = static_cast<T*>(sys[s].xs[k]);
        i_ax[s] = {geo[s], make_opdev(sys[s].op, LinksF64{sys[s].op.u64}), InPlain<T>{xk}, apvec(s), EpiNone<T>{},
                   ws_of(s), kNoGate};
        i_res[s] = {PtShiftRes<T>{apvec(s), xk, static_cast<const T*>(sys[s].b), shifts[k]}, nsite[s], nb_pt[s],
                    ws_of(s), kNoGate};
        f_res[s] = {FinResSq{ds + s, k}, nb_pt[s], ws_of(s), kNoGate};
      }
      BatchLaunch La = make_batch_launch(i_ax, nb_st);
      BatchLaunch Lr = make_batch_launch(i_res, nb_pt);
      BatchLaunch Fr = make_batch_launch(f_res, std::vector<int>(ns, 1));
      hipLaunchKernelGGL((k_normal_batch<T, InPlain<T>, EpiNone<T>>), dim3(La.nblk_total), dim3(kBlock), 0, st,
                         La.items.as<NrmItem<T, InPlain<T>, EpiNone<T>>>(), La.map());
      GLX_CHECK_LAUNCH();
      hipLaunchKernelGGL((k_point_batch<PtShiftRes<T>>), dim3(Lr.nblk_total), dim3(kBlock), 0, st,
                         Lr.items.as<PtItem<PtShiftRes<T>>>(), Lr.map());
      GLX_CHECK_LAUNCH();
      hipLaunchKernelGGL((k_finalize_batch<1, FinResSq>), dim3(ns), dim3(kBlock), 0, st,
                         Fr.items.as<FinItem<FinResSq>>());
      GLX_CHECK_LAUNCH();
      GLX_HIP(hipStreamSynchronize(st));  // the item buffers of this shift are released next
    }
  }
  GLX_HIP(hipMemcpy(hs.data(), ds, size_t(ns) * sizeof(SolveState), hipMemcpyDeviceToHost));
  std::vector<double> rv(size_t(ns) * n_shift);
  GLX_HIP(hipMemcpy(rv.data(), resv.get(), rv.size() * sizeof(double), hipMemcpyDeviceToHost));
  std::vector<SolveResult> out(ns);
  for (int s = 0; s < ns; s++) {
    const int kf = opt.max_iter > 0 ? hs[s].kfinal : 0;
    SolveResult& r = out[s];
    r.name = "CG-M";
    r.success = opt.max_iter > 0 ? (kf != opt.max_iter - 1) : true;
    r.iter = kf + 1;
    r.ops_count = 1 + (opt.max_iter > 0 ? kf : 0) + n_shift;
    r.resSq = 0.0;
    r.bsqrt = hs[s].bsqrt;
    r.resSqmrhs.assign(rv.begin() + size_t(s) * n_shift, rv.begin() + size_t(s + 1) * n_shift);
  }
  if (stats) {
    BatchStats& b = *stats;
    b = BatchStats{};
    std::vector<CgmWin> hw(ns);
    GLX_HIP(hipMemcpy(hw.data(), dwin, size_t(ns) * sizeof(CgmWin), hipMemcpyDeviceToHost));
    for (int s = 0; s < ns; s++) {
      const double V = double(nsite[s]);
      const long long it = out[s].iter;
      // windows applied in the loop, and the open one after it (x_s only)
      const long long ur = hw[s].upd_r + hw[s].wlen, ux = hw[s].upd_x + hs[s].n_xupd, up = hw[s].upd_p;
      b.iterations += it;
      b.upd_x += ux;
      b.upd_p += up;
      // the first D^dag D reads p (64 B per site); the later ones also form
      // p from r and the last p (96 B)
      b.bytes_stencil += V * (64.0 + 96.0 * double(it - 1));
      b.bytes_r += V * 48.0 * double(it);
      b.bytes_shift += V * (16.0 * double(ur) + 48.0 * double(ux) + 16.0 * double(up));
    }
    b.ms_stencil = t_ms[0];
    b.ms_r = t_ms[1];
    b.ms_shift = t_ms[2];
  }
  return out;
}

std::vector<SolveResult> dev_cg_m_batched(const std::vector<BatchSystem>& sys, int n_shift, const double* shifts,
                                          int resid_freq_check, bool worst_first, const SolveOpts& o,
                                          BatchStats* stats, bool time_phases) {
  return cg_m_batch_impl<c128>(sys, n_shift, shifts, resid_freq_check, worst_first, o, stats, time_phases);
}

// ----------------------------------------------- even/odd (compacted) ---
// Even-site vectors are compacted: site (x, y), x + y even, is element
// y * (lx / 2) + (x >> 1).  With D = m - K:
//   prepare      b_e' = m b_e - D_eo b_o = m b_e + (K b)_e     (operators.cpp:543-563)
//   solve        (m^2 - D_eo D_oe) x_e = b_e'                 (operators.cpp:566-588)
//   reconstruct  x_o = (b_o - D_oe x_e) / m = (b_o + (K x_e)_o) / m   (operators.cpp:592-616)
__device__ __forceinline__ int eo_wrap_up(int v, int n) { return v + 1 == n ? 0 : v + 1; }
__device__ __forceinline__ int eo_wrap_dn(int v, int n) { return v == 0 ? n - 1 : v - 1; }

struct PtEoPrepare {
  static constexpr int NR = 0;
  int lx, ly;
  double m;
  const c128* b;  // full lattice
  const c128* u;  // FP64 links, AoS
  c128* be;       // compact even
  __device__ void init() {}
  __device__ void operator()(int ce, double*) {
    const int W = lx >> 1;
    const int y = ce / W, c = ce - y * W;
    const int x = 2 * c + (y & 1);
    const int i = y * lx + x;
    const int xp = y * lx + eo_wrap_up(x, lx), xm = y * lx + eo_wrap_dn(x, lx);
    const int yp = eo_wrap_up(y, ly) * lx + x, ym = eo_wrap_dn(y, ly) * lx + x;
    const c128 kb = khop<c128, c128>(b[xp], b[xm], b[yp], b[ym], u[2 * i], u[2 * xm], u[2 * i + 1], u[2 * ym + 1],
                                     (x & 1) != 0);
    be[ce] = rscal(m, b[i]) + kb;
  }
};
struct PtEoReconstruct {
  static constexpr int NR = 0;
  int lx, ly;
  double m;
  const c128* xe;  // compact even
  const c128* b;   // full lattice
  const c128* u;
  c128* x;  // full lattice
  __device__ void init() {}
  __device__ void operator()(int i, double*) {
    const int W = lx >> 1;
    const int y = i / lx, xx = i - y * lx;
    if (((xx + y) & 1) == 0) {
      x[i] = xe[y * W + (xx >> 1)];
      return;
    }
    const int xr = eo_wrap_up(xx, lx), xl = eo_wrap_dn(xx, lx);
    const int yu = eo_wrap_up(y, ly), yd = eo_wrap_dn(y, ly);
    const c128 kx = khop<c128, c128>(xe[y * W + (xr >> 1)], xe[y * W + (xl >> 1)], xe[yu * W + (xx >> 1)],
                                     xe[yd * W + (xx >> 1)], u[2 * i], u[2 * (y * lx + xl)], u[2 * i + 1],
                                     u[2 * (yd * lx + xx) + 1], (xx & 1) != 0);
    const c128 v = b[i] + kx;
    x[i] = c128{v.re / m, v.im / m};
  }
};
// |a - b|^2
struct PtDiff2 {
  static constexpr int NR = 1;
  const c128* a;
  const c128* b;
  __device__ void init() {}
  __device__ void operator()(int i, double* part) { part[0] += norm2(a[i] - b[i]); }
};
struct FinStore1 {
  double* out;
  __device__ void operator()(const double* v) const { out[0] = v[0]; }
};

void dev_eo_prepare(const DevOperator& op, void* be, const void* b, hipStream_t st) {
  if (!op.u64 || op.lx % 2 || op.ly % 2) throw ArgError("even/odd preparation needs FP64 links and even lx, ly");
  RedSpace red;
  red.init(kMaxPointBlocks, st);
  launch_point(st, op.lx * op.ly / 2,
               PtEoPrepare{op.lx, op.ly, op.mass, static_cast<const c128*>(b), op.u64, static_cast<c128*>(be)},
               FinNone{}, red.ws(), kNoGate);
  GLX_HIP(hipStreamSynchronize(st));
}
void dev_eo_reconstruct(const DevOperator& op, void* x, const void* xe, const void* b, hipStream_t st) {
  if (!op.u64 || op.lx % 2 || op.ly % 2) throw ArgError("even/odd reconstruction needs FP64 links and even lx, ly");
  if (!(op.mass != 0.0)) throw ArgError("even/odd reconstruction divides by the mass: m must not be 0");
  RedSpace red;
  red.init(kMaxPointBlocks, st);
  launch_point(st, op.lx * op.ly,
               PtEoReconstruct{op.lx, op.ly, op.mass, static_cast<const c128*>(xe), static_cast<const c128*>(b),
                               op.u64, static_cast<c128*>(x)},
               FinNone{}, red.ws(), kNoGate);
  GLX_HIP(hipStreamSynchronize(st));
}

SolveResult dev_solve_staggered_eo(const DevOperator& op, void* x, const void* b, int method, const SolveOpts& o,
                                   double* full_resSq) {
  if (!op.u64 || op.lx % 2 || op.ly % 2) throw ArgError("even/odd solve needs FP64 links and even lx, ly");
  if (!(op.mass != 0.0)) throw ArgError("even/odd solve divides by the mass: m must not be 0");
  DevOperator schur = op;
  schur.kind = OP_STAG_EO_SCHUR;
  const int n = op.lx * op.ly, ne = n / 2;
  DevBuf be(size_t(ne) * sizeof(c128)), xe(size_t(ne) * sizeof(c128));
  GLX_HIP(hipMemset(xe.get(), 0, size_t(ne) * sizeof(c128)));
  dev_eo_prepare(op, be.get(), b, nullptr);
  SolveResult r;
  if (method == 0) {
    r = dev_cg(schur, xe.get(), be.get(), o);
  } else if (method == 3) {
    r = dev_cg_fused(schur, xe.get(), be.get(), o);
  } else {
    MixedOpts mo;
    mo.link_prec = method == LINKS_F16 ? LINKS_F16 : LINKS_F32;
    r = dev_cg_mixed(schur, xe.get(), be.get(), o, mo);
  }
  dev_eo_reconstruct(op, x, xe.get(), b, nullptr);
  if (full_resSq) {
    // |D x - b|^2 of the full system, D with FP64 links
    DevOperator d = op;
    d.kind = OP_STAG;
    DevBuf dx(size_t(n) * sizeof(c128)), sum(sizeof(double));
    dev_apply(d, dx.get(), x, LINKS_F64, 0, nullptr);
    RedSpace red;
    red.init(kMaxPointBlocks, nullptr);
    launch_point(nullptr, n, PtDiff2{dx.as<c128>(), static_cast<const c128*>(b)}, FinStore1{sum.as<double>()},
                 red.ws(), kNoGate);
    GLX_HIP(hipMemcpy(full_resSq, sum.get(), sizeof(double), hipMemcpyDeviceToHost));
  }
  return r;
}

// ----------------------------------------------------------- gradient --
// dS/dtheta_mu(x) = -eta_mu(x) Im[psi(x)^* U_mu(x) chi(x+mu) + psi(x+mu)^* U_mu(x)^* chi(x)]
// for S = phi^dag (D^dag D)^-1 phi, chi = (D^dag D)^-1 phi, psi = D chi
// (dU/dtheta = i U; eta_x = 1, eta_y(x) = (-1)^x, so eta_mu(x + mu) = eta_mu(x)).
// Fused sums: |chi|^2 (dS/dm = -2 m |chi|^2), Re phi^dag chi (S), |psi|^2
// (= chi^dag D^dag D chi, for the estimator 2 Re phi^dag chi - |D chi|^2).
struct PtForce {
  static constexpr int NR = 3;
  int lx, ly;
  const c128* chi;
  const c128* psi;
  const c128* phi;
  const c128* u;
  double* grad;
  __device__ void init() {}
  __device__ void operator()(int i, double* part) {
    const int y = i / lx, x = i - y * lx;
    const int nx = y * lx + (x + 1 == lx ? 0 : x + 1);
    const int ny = (y + 1 == ly ? 0 : y + 1) * lx + x;
    const c128 ci = chi[i], pi = psi[i];
    auto term = [&](c128 U, int nb) {
      const c128 a = cmulc(pi, U * chi[nb]);      // psi(x)^* U chi(x+mu)
      const c128 b = cmulc(psi[nb], cmulc(U, ci));  // psi(x+mu)^* U^* chi(x)
      return a.im + b.im;
    };
    grad[2 * i] = -term(u[2 * i], nx);
    const double ty = term(u[2 * i + 1], ny);
    grad[2 * i + 1] = (x & 1) ? ty : -ty;
    part[0] += norm2(ci);
    part[1] += dotc(phi[i], ci).re;
    part[2] += norm2(pi);
  }
};
struct FinStore3 {
  double* out;
  __device__ void operator()(const double* v) const {
    out[0] = v[0];
    out[1] = v[1];
    out[2] = v[2];
  }
};

GradResult dev_staggered_grad(const DevOperator& normal, const DevOperator& dop, const void* phi, int precision,
                              const SolveOpts& o, double* grad, void* chi_out) {
  if (normal.kind != OP_STAG_NORMAL || dop.kind != OP_STAG || !normal.u64)
    throw ArgError("the gradient is defined for the staggered normal operator");
  const int n = normal.size();
  DevBuf chiB(chi_out ? 1 : size_t(n) * sizeof(c128)), psiB(size_t(n) * sizeof(c128)), sums(4 * sizeof(double));
  c128* chi = chi_out ? static_cast<c128*>(chi_out) : chiB.as<c128>();
  GLX_HIP(hipMemsetAsync(chi, 0, size_t(n) * sizeof(c128), o.stream));
  GradResult g;
  if (precision == 0) {
    g.solve = dev_cg_fused(normal, chi, phi, o);
  } else {
    MixedOpts mo;
    mo.link_prec = precision == LINKS_F16 ? LINKS_F16 : LINKS_F32;
    g.solve = dev_cg_mixed(normal, chi, phi, o, mo);
  }
  dev_apply(dop, psiB.get(), chi, LINKS_F64, 0, o.stream);
  RedSpace red;
  red.init(kMaxPointBlocks, o.stream);
  launch_point(o.stream, n,
               PtForce{normal.lx, normal.ly, chi, psiB.as<c128>(), static_cast<const c128*>(phi), normal.u64, grad},
               FinStore3{sums.as<double>()}, red.ws(), kNoGate);
  double h[3];
  GLX_HIP(hipMemcpyAsync(h, sums.get(), 3 * sizeof(double), hipMemcpyDeviceToHost, o.stream));
  GLX_HIP(hipStreamSynchronize(o.stream));
  g.dS_dm = -2.0 * normal.mass * h[0];
  g.S = h[1];
  g.S_robust = 2.0 * h[1] - h[2];
  return g;
}

double time_cg_kernel(const DevOperator& o, int prec, int reps, double* bytes_per_site) {
  if (!o.tiled() || o.kind != OP_STAG_NORMAL) throw ArgError("timing is offered for the staggered normal operator");
  if (reps <= 0) throw ArgError("reps must be positive");
  if (prec == 0) {
    // reads p, r, Ap (old slots), x; writes r, p, Ap (new slots), x; links
    if (bytes_per_site) *bytes_per_site = 3 * 16 + 16 + 3 * 16 + 16 + 2 * 16;
    return time_fast_kernel_impl<c128, c128, false>(o, LINKS_F64, reps);
  }
  if (prec == 2 + LINKS_F32 || prec == 2 + LINKS_F16) {
    // reads the five basis vectors, x and the links; writes five vectors and x
    if (!sscg_fits(o)) throw ArgError("the s-step kernel needs lx % 128 == 0, ly % 4 == 0, ly >= 8");
    const int lp = prec - 2;
    if (bytes_per_site) *bytes_per_site = 5 * 8 + 8 + (lp == LINKS_F16 ? 2 * 4 + 4 : 2 * 8) + 5 * 8 + 8;
    const int rows = sstep_rows_for(o.lx, o.ly);
    return time_sscg_kernel_impl(o, lp, reps, rows > 0 ? rows : sstep_rows_auto(o.lx, o.ly));
  }
  if (prec == LINKS_F16) {
    if (!o.u16 || !o.sc16) throw ArgError("operator has no FP16 links on the device");
    if (bytes_per_site) *bytes_per_site = 3 * 8 + 8 + 3 * 8 + 8 + 2 * 4 + 4;
    return time_fast_kernel_impl<c128, c64, true>(o, LINKS_F16, reps);
  }
  if (!o.u32) throw ArgError("operator has no FP32 links on the device");
  if (bytes_per_site) *bytes_per_site = 3 * 8 + 8 + 3 * 8 + 8 + 2 * 8;
  return time_fast_kernel_impl<c128, c64, true>(o, LINKS_F32, reps);
}

void dev_apply(const DevOperator& op, void* out, const void* in, int link_prec, int fp32, hipStream_t st) {
  RedSpace red;
  red.init(kMaxPointBlocks, st);
  if (real_op(op)) {
    if (fp32) apply_typed<float>(st, op, link_prec, static_cast<float*>(out), static_cast<const float*>(in), red.ws(), kNoGate);
    else apply_typed<double>(st, op, link_prec, static_cast<double*>(out), static_cast<const double*>(in), red.ws(), kNoGate);
  } else {
    if (fp32) apply_typed<c64>(st, op, link_prec, static_cast<c64*>(out), static_cast<const c64*>(in), red.ws(), kNoGate);
    else apply_typed<c128>(st, op, link_prec, static_cast<c128*>(out), static_cast<const c128*>(in), red.ws(), kNoGate);
  }
  GLX_HIP(hipStreamSynchronize(st));
}

// ------------------------------------------------------------ gauge fields --
// Counter-based Philox4x32-10 (Salmon, Moraes, Dror and Shaw, "Parallel random
// numbers: as easy as 1, 2, 3", SC'11): ten rounds of 32x32->64 multiplies by
// M0 = 0xD2511F53 and M1 = 0xCD9E8D57 with the key bumped by the Weyl
// constants 0x9E3779B9 / 0xBB67AE85 between rounds.  Every site draws from its
// own counter, so a field depends only on (seed, stream), not on the launch.
__device__ __forceinline__ void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; r++) {
    const uint32_t lo0 = 0xD2511F53u * c[0], hi0 = __umulhi(0xD2511F53u, c[0]);
    const uint32_t lo1 = 0xCD9E8D57u * c[2], hi1 = __umulhi(0xCD9E8D57u, c[2]);
    const uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
    c[0] = n0;
    c[1] = lo1;
    c[2] = n2;
    c[3] = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}

// Site s: counter {s, stream, 0, 0}, key {seed lo, seed hi}.  Two 53-bit
// uniforms a, b from the four words; beta > 0: Box-Muller pair
// sigma sqrt(-2 ln(1 - a)) (cos, sin)(2 pi b) for the x and y links;
// beta = 0: uniform phases -pi + 2 pi (a, b).
struct PtGaussLinks {
  static constexpr int NR = 0;
  uint32_t k0, k1, stream;
  double sigma;  // 1/sqrt(|beta|), 0 for the uniform (hot) field
  c128* u;
  __device__ void init() {}
  __device__ void operator()(int s, double*) {
    uint32_t c[4] = {uint32_t(s), stream, 0u, 0u};
    philox4x32_10(c, k0, k1);
    constexpr double kTwoM53 = 1.1102230246251565404e-16;  // 2^-53
    constexpr double kTwoPi = 6.283185307179586232;
    const double a = double(((uint64_t(c[1]) << 32) | c[0]) >> 11) * kTwoM53;
    const double b = double(((uint64_t(c[3]) << 32) | c[2]) >> 11) * kTwoM53;
    double t0, t1;
    if (sigma > 0.0) {
      const double r = sigma * sqrt(-2.0 * log(1.0 - a));
      double sn, cs;
      sincos(kTwoPi * b, &sn, &cs);
      t0 = r * cs;
      t1 = r * sn;
    } else {
      t0 = kTwoPi * a - 3.141592653589793116;
      t1 = kTwoPi * b - 3.141592653589793116;
    }
    double s0, c0, s1, c1;
    sincos(t0, &s0, &c0);
    sincos(t1, &s1, &c1);
    u[2 * s] = {c0, s0};
    u[2 * s + 1] = {c1, s1};
  }
};

// Mean plaquette: U_x(s) U_y(s + x) U_x^*(s + y) U_y^*(s), summed per site.
struct PtPlaquette {
  static constexpr int NR = 2;
  int lx, ly;
  const c128* u;
  __device__ void init() {}
  __device__ void operator()(int i, double* part) {
    const int y = i / lx, x = i - y * lx;
    const int nx = y * lx + (x + 1 == lx ? 0 : x + 1);
    const int ny = (y + 1 == ly ? 0 : y + 1) * lx + x;
    const c128 p = cmulc(u[2 * ny], cmulc(u[2 * i + 1], u[2 * i] * u[2 * nx + 1]));
    part[0] += p.re;
    part[1] += p.im;
  }
};
struct FinStore2 {
  double* out;
  __device__ void operator()(const double* v) const {
    out[0] = v[0];
    out[1] = v[1];
  }
};

void dev_gauss_links(void* u, int lx, int ly, double beta, unsigned long long seed, unsigned stream,
                     hipStream_t st) {
  if (lx <= 0 || ly <= 0 || double(lx) * ly > 2147483647.0) throw ArgError("lattice dimensions out of range");
  const double b = beta < 0 ? -beta : beta;
  RedSpace red;
  red.init(1, st);
  launch_point(st, lx * ly,
               PtGaussLinks{uint32_t(seed), uint32_t(seed >> 32), stream, b > 0 ? 1.0 / sqrt(b) : 0.0,
                            static_cast<c128*>(u)},
               FinNone{}, red.ws(), kNoGate);
  GLX_HIP(hipStreamSynchronize(st));
}

std::complex<double> dev_plaquette(const void* u, int lx, int ly, hipStream_t st) {
  RedSpace red;
  red.init(kMaxPointBlocks, st);
  DevBuf out(2 * sizeof(double));
  launch_point(st, lx * ly, PtPlaquette{lx, ly, static_cast<const c128*>(u)}, FinStore2{out.as<double>()},
               red.ws(), kNoGate);
  double h[2];
  GLX_HIP(hipMemcpyAsync(h, out.get(), sizeof h, hipMemcpyDeviceToHost, st));
  GLX_HIP(hipStreamSynchronize(st));
  return {h[0] / (double(lx) * ly), h[1] / (double(lx) * ly)};
}

}  // namespace glx

// glx_common.h -- device-side building blocks shared by every kernel of the
// library: complex arithmetic for the two precisions, wave64 + LDS
// reductions, the deterministic finalisation kernel that turns per-block
// partial sums into solver scalars on the device, and HIP error handling.
//
// Target: gfx950 (MI355X), wave64.  All reductions accumulate in FP64 and are
// bit-reproducible for a given launch shape: per-block partial sums are
// combined in block order by a one-block kernel, never with floating-point
// atomics.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>

namespace glx {

// ---------------------------------------------------------------- errors --
struct HipError : std::runtime_error {
  explicit HipError(const std::string& s) : std::runtime_error(s) {}
};
struct ArgError : std::runtime_error {
  explicit ArgError(const std::string& s) : std::runtime_error(s) {}
};

#define GLX_HIP(call)                                                                          \
  do {                                                                                         \
    hipError_t e_ = (call);                                                                    \
    if (e_ != hipSuccess)                                                                      \
      throw ::glx::HipError(std::string(#call) + ": " + hipGetErrorString(e_) + " (" __FILE__ \
                            ":" + std::to_string(__LINE__) + ")");                             \
  } while (0)

#define GLX_CHECK_LAUNCH() GLX_HIP(hipGetLastError())

// --------------------------------------------------------------- complex --
template <class R>
struct alignas(2 * sizeof(R)) cx {
  R re, im;
};
using c128 = cx<double>;
using c64 = cx<float>;

template <class R>
__host__ __device__ __forceinline__ cx<R> mk(R a, R b) { return cx<R>{a, b}; }
template <class R>
__host__ __device__ __forceinline__ cx<R> operator+(cx<R> a, cx<R> b) { return {a.re + b.re, a.im + b.im}; }
template <class R>
__host__ __device__ __forceinline__ cx<R> operator-(cx<R> a, cx<R> b) { return {a.re - b.re, a.im - b.im}; }
template <class R>
__host__ __device__ __forceinline__ cx<R> operator-(cx<R> a) { return {-a.re, -a.im}; }
template <class R>
__host__ __device__ __forceinline__ cx<R> operator*(cx<R> a, cx<R> b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <class R>
__host__ __device__ __forceinline__ cx<R> operator*(R t, cx<R> z) { return {t * z.re, t * z.im}; }
template <class R>
__host__ __device__ __forceinline__ cx<R> conj(cx<R> z) { return {z.re, -z.im}; }
// conj(a) * b
template <class R>
__host__ __device__ __forceinline__ cx<R> cmulc(cx<R> a, cx<R> b) {
  return {a.re * b.re + a.im * b.im, a.re * b.im - a.im * b.re};
}
// complex division by Smith's method (the algorithm libgcc's __divdc3 uses for
// finite operands)
template <class R>
__host__ __device__ __forceinline__ cx<R> cdiv(cx<R> a, cx<R> b) {
  R c = b.re, d = b.im;
  if (fabs(c) < fabs(d)) {
    R ratio = c / d, denom = c * ratio + d;
    return {(a.re * ratio + a.im) / denom, (a.im * ratio - a.re) / denom};
  }
  R ratio = d / c, denom = d * ratio + c;
  return {(a.im * ratio + a.re) / denom, (a.im - a.re * ratio) / denom};
}

// Uniform helpers so that solver kernels are written once for real and
// complex element types.  `Acc<T>` is the FP64 accumulation type of a dot.
template <class T> struct Acc;
template <> struct Acc<double> { using type = double; };
template <> struct Acc<float> { using type = double; };
template <> struct Acc<c128> { using type = c128; };
template <> struct Acc<c64> { using type = c128; };

template <class T> struct Real;
template <> struct Real<double> { using type = double; };
template <> struct Real<float> { using type = float; };
template <> struct Real<c128> { using type = double; };
template <> struct Real<c64> { using type = float; };

template <class T> struct is_cx { static constexpr bool value = false; };
template <class R> struct is_cx<cx<R>> { static constexpr bool value = true; };

__host__ __device__ __forceinline__ double norm2(double a) { return a * a; }
__host__ __device__ __forceinline__ double norm2(float a) { return double(a) * double(a); }
__host__ __device__ __forceinline__ double norm2(c128 a) { return a.re * a.re + a.im * a.im; }
__host__ __device__ __forceinline__ double norm2(c64 a) {
  double r = a.re, i = a.im;
  return r * r + i * i;
}
// conj(a)*b accumulated in FP64
__host__ __device__ __forceinline__ double dotc(double a, double b) { return a * b; }
__host__ __device__ __forceinline__ double dotc(float a, float b) { return double(a) * double(b); }
__host__ __device__ __forceinline__ c128 dotc(c128 a, c128 b) { return cmulc(a, b); }
__host__ __device__ __forceinline__ c128 dotc(c64 a, c64 b) {
  c128 x{a.re, a.im}, y{b.re, b.im};
  return cmulc(x, y);
}

// Scalars carried on the device are FP64 (complex or real); `S<T>` is the
// scalar type that multiplies a T element, `scal(s, v)` applies it in T's
// precision.
template <class T> struct Scal;
template <> struct Scal<double> { using type = double; };
template <> struct Scal<float> { using type = double; };
template <> struct Scal<c128> { using type = c128; };
template <> struct Scal<c64> { using type = c128; };

__host__ __device__ __forceinline__ double scal(double s, double v) { return s * v; }
__host__ __device__ __forceinline__ float scal(double s, float v) { return float(s) * v; }
__host__ __device__ __forceinline__ c128 scal(c128 s, c128 v) { return s * v; }
__host__ __device__ __forceinline__ c64 scal(c128 s, c64 v) {
  c64 t{float(s.re), float(s.im)};
  return t * v;
}
__host__ __device__ __forceinline__ double rscal(double s, double v) { return s * v; }
__host__ __device__ __forceinline__ float rscal(double s, float v) { return float(s) * v; }
__host__ __device__ __forceinline__ c128 rscal(double s, c128 v) { return s * v; }
__host__ __device__ __forceinline__ c64 rscal(double s, c64 v) { return float(s) * v; }

template <class T> __host__ __device__ __forceinline__ T zero_of();
template <> __host__ __device__ __forceinline__ double zero_of<double>() { return 0.0; }
template <> __host__ __device__ __forceinline__ float zero_of<float>() { return 0.0f; }
template <> __host__ __device__ __forceinline__ c128 zero_of<c128>() { return {0.0, 0.0}; }
template <> __host__ __device__ __forceinline__ c64 zero_of<c64>() { return {0.0f, 0.0f}; }

// Conversion between precisions of the same kind.
__host__ __device__ __forceinline__ void cvt(double& d, double s) { d = s; }
__host__ __device__ __forceinline__ void cvt(float& d, double s) { d = float(s); }
__host__ __device__ __forceinline__ void cvt(double& d, float s) { d = s; }
__host__ __device__ __forceinline__ void cvt(float& d, float s) { d = s; }
__host__ __device__ __forceinline__ void cvt(c128& d, c128 s) { d = s; }
__host__ __device__ __forceinline__ void cvt(c64& d, c128 s) { d = {float(s.re), float(s.im)}; }
__host__ __device__ __forceinline__ void cvt(c128& d, c64 s) { d = {s.re, s.im}; }
__host__ __device__ __forceinline__ void cvt(c64& d, c64 s) { d = s; }

// ------------------------------------------------------------ reductions --
constexpr int kBlock = 256;         // threads per block everywhere
constexpr int kWaves = kBlock / 64; // wave64

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Sum NR values over the block; result valid in thread 0.  Fixed order.
template <int NR>
__device__ __forceinline__ void block_sum(double (&v)[NR], double (*sm)[kWaves]) {
#pragma unroll
  for (int k = 0; k < NR; k++) v[k] = wave_sum(v[k]);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < NR; k++) sm[k][wave] = v[k];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NR; k++) {
      double s = 0.0;
#pragma unroll
      for (int w = 0; w < kWaves; w++) s += sm[k][w];
      v[k] = s;
    }
  }
}

// Sum NR values over the block; every thread gets the same totals (the
// per-wave sums are added in the same fixed order in every thread).
template <int NR>
__device__ __forceinline__ void block_sum_all(double (&v)[NR], double (*sm)[kWaves]) {
#pragma unroll
  for (int k = 0; k < NR; k++) v[k] = wave_sum(v[k]);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < NR; k++) sm[k][wave] = v[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NR; k++) {
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < kWaves; w++) s += sm[k][w];
    v[k] = s;
  }
  __syncthreads();
}

// Wave and block sums through DPP moves instead of LDS shuffles: within each
// row of 16 lanes xor 1, xor 2 (quad permutations), then the half-row and row
// mirrors; then the four row sums, read from lanes 0, 16, 32 and 48, added in
// that order.  Every lane gets the same bits (a + b and b + a round alike).
// All 64 lanes must be active.  Used by the self-finalising CG kernels, whose
// per-launch reductions are a large share of their LDS instructions otherwise.
template <int CTRL>
__device__ __forceinline__ double dpp_mov(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, int(b), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, int(b >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double((long long)(unsigned)lo | ((long long)hi << 32));
}
__device__ __forceinline__ double readlane_d(double v, int lane) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane(int(b), lane);
  const int hi = __builtin_amdgcn_readlane(int(b >> 32), lane);
  return __longlong_as_double((long long)(unsigned)lo | ((long long)hi << 32));
}
__device__ __forceinline__ double wave_sum_dpp(double v) {
  v += dpp_mov<0xB1>(v);   // quad_perm [1,0,3,2]
  v += dpp_mov<0x4E>(v);   // quad_perm [2,3,0,1]
  v += dpp_mov<0x141>(v);  // row_half_mirror
  v += dpp_mov<0x140>(v);  // row_mirror
  return (readlane_d(v, 0) + readlane_d(v, 16)) + (readlane_d(v, 32) + readlane_d(v, 48));
}
// Sum of NR values over the block in every thread (fixed order).
template <int NR>
__device__ __forceinline__ void block_sum_dpp(double (&v)[NR], double (*sm)[kWaves]) {
#pragma unroll
  for (int k = 0; k < NR; k++) v[k] = wave_sum_dpp(v[k]);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < NR; k++) sm[k][wave] = v[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NR; k++) {
    double t = 0.0;
#pragma unroll
    for (int w = 0; w < kWaves; w++) t += sm[k][w];
    v[k] = t;
  }
  __syncthreads();
}

// Reduction workspace: partial sums [nblocks * NR].
struct RedWs {
  double* partials;
};

// Block-level part of a grid reduction: every block of the grid stores its
// NR partial sums (plain stores, no device-scope fence) component-major,
// partials[k * gridDim.x + block], so the finaliser's loads coalesce.  A
// one-block finalisation kernel launched right after (k_finalize) sums them
// in block order -- deterministic for a given launch shape -- and hands the
// totals to the solver's finaliser.  Kernel boundaries make the partials
// visible across XCDs, which avoids an L2 write-back per block.
// The explicit form stores block `idx` of a reduction over `stride` blocks
// (a batched launch reduces each system over its own block range).
template <int NR>
__device__ __forceinline__ void block_partials_ex(double (&v)[NR], RedWs ws, int idx, int stride) {
  __shared__ double sm[NR > 0 ? NR : 1][kWaves];
  block_sum<NR>(v, sm);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NR; k++) ws.partials[k * stride + idx] = v[k];
  }
}
template <int NR>
__device__ __forceinline__ void block_partials(double (&v)[NR], RedWs ws, int bid) {
  block_partials_ex<NR>(v, ws, bid, int(gridDim.x));
}
// The same store with the block sum of block_sum_dpp.
template <int NR>
__device__ __forceinline__ void block_partials_dpp(double (&v)[NR], RedWs ws, int bid) {
  __shared__ double sm[NR > 0 ? NR : 1][kWaves];
  block_sum_dpp<NR>(v, sm);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NR; k++) ws.partials[k * int(gridDim.x) + bid] = v[k];
  }
}

// A kernel's gate: it runs only while the solver's control word equals
// `want` (0: iterating; 2: a reliable update is pending); a null flag always
// runs.  Every block reads the same word, written by an earlier kernel.
struct Gate {
  const int* flag;
  int want;
};
__device__ __forceinline__ bool gate_closed(Gate g) {
  return g.flag && __hip_atomic_load(g.flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != g.want;
}

// nblk = grid size of the kernel that wrote the partials.  Each thread sums
// blocks t, t + kBlock, t + 2 kBlock, ... in increasing order; the loads of
// kFinUnroll consecutive terms are issued together, so one sweep costs one
// memory latency instead of kFinUnroll.
constexpr int kFinUnroll = 8;
template <int NR, class Fin>
__device__ __forceinline__ void finalize_body(int nblk, const Fin& fin, RedWs ws);
template <int NR, class Fin>
__global__ __launch_bounds__(kBlock) void k_finalize(int nblk, Fin fin, RedWs ws, Gate gate) {
  if (gate_closed(gate)) return;
  finalize_body<NR, Fin>(nblk, fin, ws);
}
template <int NR, class Fin>
__device__ __forceinline__ void finalize_body(int nblk, const Fin& fin, RedWs ws) {
  __shared__ double sm[NR][kWaves];
  double acc[NR];
#pragma unroll
  for (int k = 0; k < NR; k++) acc[k] = 0.0;
  for (int b0 = threadIdx.x; b0 < nblk; b0 += kBlock * kFinUnroll) {
    double v[kFinUnroll][NR];
#pragma unroll
    for (int u = 0; u < kFinUnroll; u++) {
      const int b = b0 + u * kBlock;
#pragma unroll
      for (int k = 0; k < NR; k++) v[u][k] = b < nblk ? ws.partials[k * nblk + b] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < kFinUnroll; u++) {
#pragma unroll
      for (int k = 0; k < NR; k++) acc[k] += v[u][k];
    }
  }
  block_sum<NR>(acc, sm);
  if (threadIdx.x == 0) fin(acc);
}

// Two adjacent elements as one aligned access (16 bytes for complex FP32):
// the caller guarantees p + j is aligned to 2 * sizeof(T).
template <class T>
struct alignas(2 * sizeof(T)) Pair2 {
  T a, b;
};
template <class T>
__device__ __forceinline__ Pair2<T> ld2(const T* p, int j) {
  return *reinterpret_cast<const Pair2<T>*>(p + j);
}
template <class T>
__device__ __forceinline__ void st2(T* p, int j, T a, T b) {
  *reinterpret_cast<Pair2<T>*>(p + j) = Pair2<T>{a, b};
}

// -------------------------------------------------------------- geometry --
// Lattice tile of one block for the stencil kernels.
constexpr int kTX = 64;
constexpr int kTY = 8;

// Periodic wrap of a coordinate that lies in [-2, n + kTX + 2).
__device__ __forceinline__ int wrap_coord(int v, int n) {
  if (v < 0) v += n;
  if (v >= n) v -= n;
  if (v >= n || v < 0) v = ((v % n) + n) % n;
  return v;
}

// Linear block id remapped so that each of the 8 XCDs works on a contiguous
// range of tiles (the hardware deals consecutive block ids round-robin to the
// XCDs); neighbouring tiles then share an L2.
__device__ __forceinline__ int xcd_tile(int bid, int nblk) {
  const int per = (nblk + 7) >> 3;
  return (bid & 7) * per + (bid >> 3);
}

}  // namespace glx

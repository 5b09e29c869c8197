// generic_vector.h -- host BLAS-1 templates of the reference
// (generic_vector.h:12-252): same names, arguments and results.  Sums run
// serially from index 0 (res = res + term), the complex dot conjugates its
// first argument, and random vectors draw from std::normal_distribution.
// The solvers in libglx.so do their own vector work on the GPU; these are
// for drivers that prepare inputs and check results on the host.
#ifndef GLX_COMPAT_GENERIC_VECTOR
#define GLX_COMPAT_GENERIC_VECTOR

#include <cmath>
#include <complex>
#include <random>

template <typename T>
inline void zero(T* v1, int size) {
  for (int i = 0; i < size; i++) v1[i] = 0.0;
}
template <typename T>
inline void zero(std::complex<T>* v1, int size) {
  for (int i = 0; i < size; i++) v1[i] = 0.0;
}

// N(0, 1) entries.
template <typename T>
inline void gaussian(T* v1, int size, std::mt19937& generator) {
  std::normal_distribution<> dist(0.0, 1.0);
  for (int i = 0; i < size; i++) v1[i] = static_cast<T>(dist(generator));
}
// N(0, 1) real and imaginary parts.  The reference constructs
// complex(dist(g), dist(g)); g++ evaluates those arguments right to left, so
// the imaginary part is drawn first.  That order is fixed here explicitly.
template <typename T>
inline void gaussian(std::complex<T>* v1, int size, std::mt19937& generator) {
  std::normal_distribution<> dist(0.0, 1.0);
  for (int i = 0; i < size; i++) {
    const T im = static_cast<T>(dist(generator));
    const T re = static_cast<T>(dist(generator));
    v1[i] = std::complex<T>(re, im);
  }
}

// v1 = v2
template <typename T>
inline void copy(T* v1, T* v2, int size) {
  for (int i = 0; i < size; i++) v1[i] = v2[i];
}
template <typename T>
inline void copy(std::complex<T>* v1, std::complex<T>* v2, int size) {
  for (int i = 0; i < size; i++) v1[i] = v2[i];
}

template <typename T>
inline T dot(T* v1, T* v2, int size) {
  T res = (T)0.0;
  for (int i = 0; i < size; i++) res = res + v1[i] * v2[i];
  return res;
}
// sum conj(v1) v2
template <typename T>
inline std::complex<T> dot(std::complex<T>* v1, std::complex<T>* v2, int size) {
  std::complex<T> res = (T)0.0;
  for (int i = 0; i < size; i++) res = res + std::conj(v1[i]) * v2[i];
  return res;
}

template <typename T>
inline T norm2sq(T* v1, int size) {
  T res = (T)0.0;
  for (int i = 0; i < size; i++) res = res + v1[i] * v1[i];
  return res;
}
template <typename T>
inline T norm2sq(std::complex<T>* v1, int size) {
  T res = (T)0.0;
  for (int i = 0; i < size; i++) res = res + std::real(std::conj(v1[i]) * v1[i]);
  return res;
}

// |v1 - v2|^2
template <typename T>
inline T diffnorm2sq(T* v1, T* v2, int size) {
  T res = (T)0.0;
  for (int i = 0; i < size; i++) res = res + (v1[i] - v2[i]) * (v1[i] - v2[i]);
  return res;
}
template <typename T>
inline T diffnorm2sq(std::complex<T>* v1, std::complex<T>* v2, int size) {
  T res = (T)0.0;
  for (int i = 0; i < size; i++) {
    const std::complex<T> d = v1[i] - v2[i];
    res = res + std::real(std::conj(d) * d);
  }
  return res;
}

// v1 /= |v1| (left alone when the scale is not positive)
template <typename T>
inline void normalize(T* v1, int size) {
  const T s = 1.0 / std::sqrt(norm2sq<T>(v1, size));
  if (s > 0.0)
    for (int i = 0; i < size; i++) v1[i] *= s;
}
template <typename T>
inline void normalize(std::complex<T>* v1, int size) {
  const T s = 1.0 / std::sqrt(norm2sq<T>(v1, size));
  if (s > 0.0)
    for (int i = 0; i < size; i++) v1[i] *= s;
}

template <typename T>
inline void conj(T*, int) {}
template <typename T>
inline void conj(std::complex<T>* v1, int size) {
  for (int i = 0; i < size; i++) v1[i] = std::conj(v1[i]);
}

// v1 -= (v2^dag v1 / |v2|^2) v2
template <typename T>
inline void orthogonal(T* v1, T* v2, int size) {
  const T alpha = -dot<T>(v2, v1, size) / norm2sq<T>(v2, size);
  for (int i = 0; i < size; i++) v1[i] = v1[i] + alpha * v2[i];
}
template <typename T>
inline void orthogonal(std::complex<T>* v1, std::complex<T>* v2, int size) {
  const std::complex<T> alpha = -dot<T>(v2, v1, size) / norm2sq<T>(v2, size);
  for (int i = 0; i < size; i++) v1[i] = v1[i] + alpha * v2[i];
}

#endif

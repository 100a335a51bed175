// Shared device helpers for the LNS3D implicit-step kernels (gfx950).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <type_traits>

namespace lns3d {

// complex value with explicit real/imag parts; T = double or float
template <typename T> struct cx {
  T re, im;
};

template <typename T> __host__ __device__ __forceinline__ cx<T> mk(T r, T i) { return cx<T>{r, i}; }
template <typename T> __host__ __device__ __forceinline__ cx<T> operator+(cx<T> a, cx<T> b) { return {a.re + b.re, a.im + b.im}; }
template <typename T> __host__ __device__ __forceinline__ cx<T> operator-(cx<T> a, cx<T> b) { return {a.re - b.re, a.im - b.im}; }
template <typename T> __host__ __device__ __forceinline__ cx<T> operator-(cx<T> a) { return {-a.re, -a.im}; }
template <typename T> __host__ __device__ __forceinline__ cx<T> operator*(cx<T> a, cx<T> b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <typename T> __host__ __device__ __forceinline__ cx<T> operator*(T s, cx<T> b) { return {s * b.re, s * b.im}; }
// acc + a*b
template <typename T> __host__ __device__ __forceinline__ cx<T> cfma(cx<T> a, cx<T> b, cx<T> acc) {
  acc.re = fma(a.re, b.re, acc.re);
  acc.re = fma(-a.im, b.im, acc.re);
  acc.im = fma(a.re, b.im, acc.im);
  acc.im = fma(a.im, b.re, acc.im);
  return acc;
}
// a*b with one rounding fewer per part than the textbook form
template <typename T> __host__ __device__ __forceinline__ cx<T> cmul(cx<T> a, cx<T> b) {
  T re = fma(a.re, b.re, -(a.im * b.im));
  T im = fma(a.re, b.im, a.im * b.re);
  return {re, im};
}
// a / b, Smith's algorithm (Smith, CACM 5(8), 1962)
template <typename T> __host__ __device__ __forceinline__ cx<T> cdiv(cx<T> a, cx<T> b) {
  if (fabs(b.re) >= fabs(b.im)) {
    T r = b.im / b.re, d = b.re + b.im * r;
    return {(a.re + a.im * r) / d, (a.im - a.re * r) / d};
  }
  T r = b.re / b.im, d = b.re * r + b.im;
  return {(a.re * r + a.im) / d, (a.im * r - a.re) / d};
}
template <typename T> __host__ __device__ __forceinline__ cx<T> crecip(cx<T> b) {
  return cdiv(cx<T>{T(1), T(0)}, b);
}

// wave-uniform broadcast of a lane's complex value (one line per wave)
__device__ __forceinline__ cx<double> bcast(cx<double> v, int lane) {
  union { double d; int i[2]; } a, b;
  a.d = v.re; b.d = v.im;
  a.i[0] = __builtin_amdgcn_readlane(a.i[0], lane);
  a.i[1] = __builtin_amdgcn_readlane(a.i[1], lane);
  b.i[0] = __builtin_amdgcn_readlane(b.i[0], lane);
  b.i[1] = __builtin_amdgcn_readlane(b.i[1], lane);
  return {a.d, b.d};
}
__device__ __forceinline__ cx<float> bcast(cx<float> v, int lane) {
  union { float f; int i; } a, b;
  a.f = v.re; b.f = v.im;
  a.i = __builtin_amdgcn_readlane(a.i, lane);
  b.i = __builtin_amdgcn_readlane(b.i, lane);
  return {a.f, b.f};
}

// broadcast lane q's complex value to the 16 lanes of its DPP row
// (v_mov_b64_dpp row_newbcast:q); q must be a compile-time constant.  Every
// lane receives a value, so bound_ctrl spares the old-value operand.
template <int Q> __device__ __forceinline__ cx<double> row_bcast(cx<double> v) {
  long a = __builtin_bit_cast(long, v.re), b = __builtin_bit_cast(long, v.im);
  a = __builtin_amdgcn_update_dpp(0L, a, 0x150 + Q, 0xF, 0xF, true);
  b = __builtin_amdgcn_update_dpp(0L, b, 0x150 + Q, 0xF, 0xF, true);
  return {__builtin_bit_cast(double, a), __builtin_bit_cast(double, b)};
}
template <int Q> __device__ __forceinline__ cx<float> row_bcast(cx<float> v) {
  long a = __builtin_bit_cast(long, v);
  a = __builtin_amdgcn_update_dpp(0L, a, 0x150 + Q, 0xF, 0xF, true);
  return __builtin_bit_cast(cx<float>, a);
}
template <typename T> __device__ __forceinline__ void row_bcast5(cx<T> v, cx<T> (&y)[5]) {
  y[0] = row_bcast<0>(v); y[1] = row_bcast<1>(v); y[2] = row_bcast<2>(v);
  y[3] = row_bcast<3>(v); y[4] = row_bcast<4>(v);
}
// y[q] = v of lane OFF + q of the 16-lane row
template <int OFF, typename T> __device__ __forceinline__ void row_bcast5_at(cx<T> v, cx<T> (&y)[5]) {
  y[0] = row_bcast<OFF>(v); y[1] = row_bcast<OFF + 1>(v); y[2] = row_bcast<OFF + 2>(v);
  y[3] = row_bcast<OFF + 3>(v); y[4] = row_bcast<OFF + 4>(v);
}
// lane Q's value of each DPP row written to the rows in ROWS (bit r: row r);
// the other rows keep old
template <int Q, int ROWS> __device__ __forceinline__ cx<double> row_bcast_rows(cx<double> old, cx<double> v) {
  long a = __builtin_bit_cast(long, v.re), b = __builtin_bit_cast(long, v.im);
  a = __builtin_amdgcn_update_dpp(__builtin_bit_cast(long, old.re), a, 0x150 + Q, ROWS, 0xF, false);
  b = __builtin_amdgcn_update_dpp(__builtin_bit_cast(long, old.im), b, 0x150 + Q, ROWS, 0xF, false);
  return {__builtin_bit_cast(double, a), __builtin_bit_cast(double, b)};
}
template <int Q, int ROWS> __device__ __forceinline__ cx<float> row_bcast_rows(cx<float> old, cx<float> v) {
  long a = __builtin_bit_cast(long, v);
  a = __builtin_amdgcn_update_dpp(__builtin_bit_cast(long, old), a, 0x150 + Q, ROWS, 0xF, false);
  return __builtin_bit_cast(cx<float>, a);
}
// the sum of a value over the two 32-lane halves of the wave (lane l and lane
// l ^ 32), the same bits in both halves: v_permlane32_swap of the value with a
// copy of itself leaves the lower half's value in one register and the upper
// half's in the other, in every lane.  The copy and the two results pass
// through empty asm statements: without them the compiler has taken the
// swap's two results for one value (the float form added a result to itself).
__device__ __forceinline__ unsigned opaque_copy(unsigned u) {
  asm volatile("" : "+v"(u));
  return u;
}
__device__ __forceinline__ double half_sum(double v) {
  const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
  const unsigned lo = (unsigned)u, hi = (unsigned)(u >> 32);
  const auto l = __builtin_amdgcn_permlane32_swap(lo, opaque_copy(lo), false, false);
  const auto h = __builtin_amdgcn_permlane32_swap(hi, opaque_copy(hi), false, false);
  const double a = __builtin_bit_cast(double, ((unsigned long long)opaque_copy(h[0]) << 32) | opaque_copy(l[0]));
  const double b = __builtin_bit_cast(double, ((unsigned long long)opaque_copy(h[1]) << 32) | opaque_copy(l[1]));
  return a + b;
}
__device__ __forceinline__ float half_sum(float v) {
  const unsigned u = __builtin_bit_cast(unsigned, v);
  const auto r = __builtin_amdgcn_permlane32_swap(u, opaque_copy(u), false, false);
  return __builtin_bit_cast(float, opaque_copy(r[0])) + __builtin_bit_cast(float, opaque_copy(r[1]));
}
template <typename T> __device__ __forceinline__ cx<T> half_sum(cx<T> v) {
  return {half_sum(v.re), half_sum(v.im)};
}
// The chip places workgroup b of a launch on XCD b % 8 (round robin over its
// eight XCDs, each with its own L2).  xcd_tile(b, nb) numbers the tiles so
// that each XCD works through a contiguous range of them (a bijection of
// [0, nb)): neighbouring tiles, which share stencil rows, share an L2.
__device__ __forceinline__ long xcd_tile(long b, long nb) {
  const long q = nb / 8, r = nb % 8, x = b % 8, k = b / 8;
  return x * q + (x < r ? x : r) + k;
}
// orders the LDS accesses of one wave before / after it (lanes of the wave
// exchanging values through LDS)
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// finite-difference coefficient tables of the scheme (src/stencil.f90)
namespace st {
constexpr double ga1 = 8.333333333333333333333E-02, ga2 = -6.666666666666666666667E-01,
                 ga3 = 6.666666666666666666667E-01, ga4 = -8.333333333333333333333E-02;
constexpr double gc1 = -2.083333333333333333333E+00, gc2 = 4.000000000000000000000E+00,
                 gc3 = -3.000000000000000000000E+00, gc4 = 1.333333333333333333333E+00,
                 gc5 = -2.500000000000000000000E-01;
constexpr double gb1 = -2.500000000000000000000E-01, gb2 = -8.333333333333333333333E-01,
                 gb3 = 1.500000000000000000000E+00, gb4 = -5.000000000000000000000E-01,
                 gb5 = 8.333333333333333333333E-02;
constexpr double da1 = -8.333333333333333333333E-02, da2 = 1.333333333333333333333E+00,
                 da3 = -2.500000000000000000000E+00, da4 = 1.333333333333333333333E+00,
                 da5 = -8.333333333333333333333E-02;
constexpr double db1 = 9.166666666666666666667E-01, db2 = -1.666666666666666666667E+00,
                 db3 = 5.000000000000000000000E-01, db4 = 3.333333333333333333333E-01,
                 db5 = -8.333333333333333333333E-02;
constexpr double dd1 = 2.916666666666666666667E+00, dd2 = -8.666666666666666666667E+00,
                 dd3 = 9.500000000000000000000E+00, dd4 = -4.666666666666666666667E+00,
                 dd5 = 9.166666666666666666667E-01;
constexpr double pt66 = 6.6666666666666666666e-1;
}  // namespace st

#define LNS3D_CHECK(expr)                                                        \
  do {                                                                           \
    hipError_t e_ = (expr);                                                      \
    if (e_ != hipSuccess) return lns3d::set_error(__FILE__, __LINE__, hipGetErrorString(e_)); \
  } while (0)

int set_error(const char *file, int line, const char *msg);

}  // namespace lns3d

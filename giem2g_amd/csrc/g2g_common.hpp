// g2g_common.hpp -- shared device/host helpers for the GIEM2G MI355X library.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <type_traits>

namespace g2g {

// Most layers an anomaly may have (the contraction above 64 layers stages
// tensor tiles through LDS with the four points' rows, 12*Nz complex)
constexpr int MAX_NZ = 256;

// ---------------------------------------------------------------- errors --
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define G2G_HIP(call)                                                        \
  do {                                                                       \
    hipError_t e_ = (call);                                                  \
    if (e_ != hipSuccess)                                                    \
      throw ::g2g::Error(-100 - (int)e_,                                     \
                         std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define G2G_CHECK(cond, code, msg)                              \
  do {                                                          \
    if (!(cond)) throw ::g2g::Error((code), std::string(msg)); \
  } while (0)

// -------------------------------------------------------- complex types --
// Interleaved complex numbers with complex arithmetic (HIP's double2/float2
// carry element-wise operators, so the library uses its own types with the
// same memory layout).
struct alignas(16) zd {
  double x, y;
};
struct alignas(8) zf {
  float x, y;
};
template <typename R> struct cx;
template <> struct cx<double> { using t = zd; };
template <> struct cx<float> { using t = zf; };
template <typename R> using cx_t = typename cx<R>::t;

template <typename T2> __host__ __device__ inline T2 mk(decltype(T2::x) a, decltype(T2::x) b) {
  T2 r;
  r.x = a;
  r.y = b;
  return r;
}
__host__ __device__ inline zd mkz(double a, double b) { return zd{a, b}; }
// Load of an element that a kernel streams through once (a tensor, a Krylov
// basis vector): the non-temporal hint keeps the stream from evicting data
// that is read again soon from the caches (the spectrum the next FFT pass
// reads; the vector that a multi-dot reads for every block row)
template <typename T2> __device__ inline T2 ld_stream(const T2* p) {
  return mk<T2>(__builtin_nontemporal_load(&p->x), __builtin_nontemporal_load(&p->y));
}
#define G2G_CPLX_OPS(T, R)                                                                  \
  __host__ __device__ inline T operator+(T a, T b) { return T{a.x + b.x, a.y + b.y}; }     \
  __host__ __device__ inline T operator-(T a, T b) { return T{a.x - b.x, a.y - b.y}; }     \
  __host__ __device__ inline T operator-(T a) { return T{-a.x, -a.y}; }                     \
  __host__ __device__ inline T operator*(T a, T b) {                                        \
    return T{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};                                 \
  }                                                                                         \
  __host__ __device__ inline T operator*(T a, R s) { return T{a.x * s, a.y * s}; }          \
  __host__ __device__ inline T operator*(R s, T a) { return T{a.x * s, a.y * s}; }
G2G_CPLX_OPS(zd, double)
G2G_CPLX_OPS(zf, float)
#undef G2G_CPLX_OPS

template <typename T2> __host__ __device__ inline T2 cconj(T2 a) { a.y = -a.y; return a; }
// a += b * c
template <typename T2> __host__ __device__ inline void cmac(T2& a, T2 b, T2 c) {
  a.x = fma(b.x, c.x, a.x);
  a.x = fma(-b.y, c.y, a.x);
  a.y = fma(b.x, c.y, a.y);
  a.y = fma(b.y, c.x, a.y);
}
__host__ __device__ inline zd cdiv(zd a, zd b) {
  // Smith's algorithm
  if (fabs(b.x) >= fabs(b.y)) {
    double r = b.y / b.x, d = b.x + b.y * r;
    return zd{(a.x + a.y * r) / d, (a.y - a.x * r) / d};
  } else {
    double r = b.x / b.y, d = b.x * r + b.y;
    return zd{(a.x * r + a.y) / d, (a.y * r - a.x) / d};
  }
}
template <typename T2> __host__ __device__ inline T2 mul_mi(T2 a) {  // a * (-i)
  T2 r; r.x = a.y; r.y = -a.x; return r;
}
template <typename T2> __host__ __device__ inline T2 mul_pi(T2 a) {  // a * (+i)
  T2 r; r.x = -a.y; r.y = a.x; return r;
}

// dispatch a power-of-two length in [16,1024] to a compile-time constant
template <typename F> inline void pow2_dispatch(int L, F&& f) {
  switch (L) {
    case 16: f(std::integral_constant<int, 16>{}); break;
    case 32: f(std::integral_constant<int, 32>{}); break;
    case 64: f(std::integral_constant<int, 64>{}); break;
    case 128: f(std::integral_constant<int, 128>{}); break;
    case 256: f(std::integral_constant<int, 256>{}); break;
    case 512: f(std::integral_constant<int, 512>{}); break;
    case 1024: f(std::integral_constant<int, 1024>{}); break;
    default: throw Error(-3, "FFT length must be a power of two in [16,1024]");
  }
}

inline int ilog2(int n) {
  int l = 0;
  while ((1 << l) < n) l++;
  return l;
}
inline bool is_pow2(int64_t n) { return n > 0 && (n & (n - 1)) == 0; }

// A transform length's compile-time class: a power of two in [16,1024] runs
// as itself (rt = false); any other even length up to 1024 runs in the
// buffers of the next power of two with a runtime factor plan (rt = true).
template <int LC_, bool RT_> struct LenClass {
  static constexpr int value = LC_;
  static constexpr bool rt = RT_;
};
template <typename F> inline void len_dispatch(int L, F&& f) {
  if (L < 16 || L > 1024 || (L & 1)) throw Error(-3, "FFT length must be even and in [16,1024]");
  if (is_pow2(L)) {
    pow2_dispatch(L, [&](auto LC) { f(LenClass<decltype(LC)::value, false>{}); });
    return;
  }
  switch (1 << ilog2(L)) {
    case 32: f(LenClass<32, true>{}); break;
    case 64: f(LenClass<64, true>{}); break;
    case 128: f(LenClass<128, true>{}); break;
    case 256: f(LenClass<256, true>{}); break;
    case 512: f(LenClass<512, true>{}); break;
    default: f(LenClass<1024, true>{}); break;
  }
}
// the buffer capacity len_dispatch gives L
inline int len_capacity(int L) { return 1 << ilog2(L); }

}  // namespace g2g

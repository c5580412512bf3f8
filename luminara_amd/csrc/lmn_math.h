// Vector math and the counter-based RNG shared by the HIP kernels (FP32 and
// FP64 instantiations) and the host-side scene builder.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define LMN_HD __host__ __device__ __forceinline__
// Rarely taken FP64 paths stay out of line so that the register allocation of
// the traversal and shading loops is not sized for them.
#define LMN_COLD inline __host__ __device__ __attribute__((noinline))
#else
#include <cmath>
#define LMN_HD inline
#define LMN_COLD inline
#endif

namespace lmn {

#if !defined(__HIPCC__)
using std::acos; using std::asin; using std::atan; using std::atan2; using std::cbrt; using std::cos;
using std::exp; using std::fabs; using std::floor; using std::fma; using std::fmax; using std::fmin; using std::frexp;
using std::ldexp; using std::log; using std::rint;
using std::pow; using std::sin; using std::sqrt; using std::trunc;
#endif


// Bit casts usable on host and device.
LMN_HD uint32_t bits_of(float x) {
  uint32_t u;
  __builtin_memcpy(&u, &x, 4);
  return u;
}
LMN_HD float float_of(uint32_t u) {
  float x;
  __builtin_memcpy(&x, &u, 4);
  return x;
}

// Precision-generic math: float and double overloads that map to the
// single- and double-precision device/libm routines.
LMN_HD float m_abs(float x) { return fabsf(x); }
LMN_HD double m_abs(double x) { return fabs(x); }
LMN_HD float m_sqrt(float x) { return sqrtf(x); }
LMN_HD double m_sqrt(double x) { return sqrt(x); }
LMN_HD float m_exp(float x) { return expf(x); }
LMN_HD double m_exp(double x) { return exp(x); }
LMN_HD float m_acos(float x) { return acosf(x); }
LMN_HD double m_acos(double x) { return acos(x); }
LMN_HD float m_atan(float x) { return atanf(x); }
LMN_HD double m_atan(double x) { return atan(x); }
LMN_HD float m_atan2(float y, float x) { return atan2f(y, x); }
LMN_HD double m_atan2(double y, double x) { return atan2(y, x); }
LMN_HD float m_sin(float x) { return sinf(x); }
LMN_HD double m_sin(double x) { return sin(x); }
LMN_HD float m_cos(float x) { return cosf(x); }
LMN_HD double m_cos(double x) { return cos(x); }
LMN_HD float m_trunc(float x) { return truncf(x); }
LMN_HD double m_trunc(double x) { return trunc(x); }
LMN_HD float m_fmax(float a, float b) { return fmaxf(a, b); }
LMN_HD double m_fmax(double a, double b) { return fmax(a, b); }
LMN_HD float m_fmin(float a, float b) { return fminf(a, b); }
LMN_HD double m_fmin(double a, double b) { return fmin(a, b); }

// x^y for x >= 0 and finite y > 0, FP64, for the exact-mode root finders
// (the library pow spends several hundred operations on range cases these
// never reach).  x = m 2^e with m in [sqrt(1/2), sqrt(2)); ln m = 2 atanh(s),
// s = (m - 1) / (m + 1), by the odd Taylor series of atanh through s^21
// (|s| <= 0.1716: truncation below 2^-54); exp(L), L = y ln x, by Cody-Waite
// reduction r = L - n ln2 with ln2 split so that n * LN2_HI is exact, and the
// Taylor series of exp(r), |r| <= ln2 / 2, through degree 13 (truncation
// below 2^-57).  Relative error: a few ulp times max(1, |L|).
LMN_HD double pow_pos(double x, double y) {
  if (x == 0.0) return 0.0;
  if (!(x < 1.0 / 0.0)) return x;  // NaN, +inf
  int e;
  double m = frexp(x, &e);  // [0.5, 1)
  if (m < 0.70710678118654752440) {
    m *= 2.0;
    --e;
  }
  const double s = (m - 1.0) / (m + 1.0), z = s * s;
  double a = 1.0 / 21.0;
  a = fma(a, z, 1.0 / 19.0);
  a = fma(a, z, 1.0 / 17.0);
  a = fma(a, z, 1.0 / 15.0);
  a = fma(a, z, 1.0 / 13.0);
  a = fma(a, z, 1.0 / 11.0);
  a = fma(a, z, 1.0 / 9.0);
  a = fma(a, z, 1.0 / 7.0);
  a = fma(a, z, 1.0 / 5.0);
  a = fma(a, z, 1.0 / 3.0);
  a = fma(a, z, 1.0);
  const double LN2_HI = 0x1.62e42fee00000p-1, LN2_LO = 0x1.a39ef35793c76p-33;
  const double de = double(e);
  const double L = y * fma(de, LN2_HI, fma(de, LN2_LO, 2.0 * s * a));
  if (L < -745.2) return 0.0;
  if (L > 709.8) return 1.0 / 0.0;
  const double n = rint(L * 1.4426950408889634);
  const double r = fma(-n, LN2_LO, fma(-n, LN2_HI, L));
  double q = 1.0 / 6227020800.0;  // 1/13!
  q = fma(q, r, 1.0 / 479001600.0);
  q = fma(q, r, 1.0 / 39916800.0);
  q = fma(q, r, 1.0 / 3628800.0);
  q = fma(q, r, 1.0 / 362880.0);
  q = fma(q, r, 1.0 / 40320.0);
  q = fma(q, r, 1.0 / 5040.0);
  q = fma(q, r, 1.0 / 720.0);
  q = fma(q, r, 1.0 / 120.0);
  q = fma(q, r, 1.0 / 24.0);
  q = fma(q, r, 1.0 / 6.0);
  q = fma(q, r, 0.5);
  q = fma(q, r, 1.0);
  q = fma(q, r, 1.0);
  return ldexp(q, int(n));
}

template <class R>
struct V3 {
  R x, y, z;
};

template <class R> LMN_HD V3<R> v3(R x, R y, R z) { return V3<R>{x, y, z}; }
template <class R> LMN_HD V3<R> operator+(V3<R> a, V3<R> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class R> LMN_HD V3<R> operator-(V3<R> a, V3<R> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class R> LMN_HD V3<R> operator-(V3<R> a) { return {-a.x, -a.y, -a.z}; }
template <class R> LMN_HD V3<R> operator*(V3<R> a, R t) { return {a.x * t, a.y * t, a.z * t}; }
template <class R> LMN_HD V3<R> operator/(V3<R> a, R t) { return {a.x / t, a.y / t, a.z / t}; }
template <class R> LMN_HD V3<R> hadamard(V3<R> a, V3<R> b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
template <class R> LMN_HD R dot(V3<R> a, V3<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class R> LMN_HD R length_squared(V3<R> a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
template <class R> LMN_HD V3<R> cross(V3<R> a, V3<R> b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class R> LMN_HD R vsqrt(R x);
template <> LMN_HD float vsqrt<float>(float x) { return m_sqrt(x); }
template <> LMN_HD double vsqrt<double>(double x) { return m_sqrt(x); }
template <class R> LMN_HD R length(V3<R> a) { return vsqrt(length_squared(a)); }
// Vec3::unit divides by the length (src/vec3.rs:53-57).
template <class R> LMN_HD V3<R> unit(V3<R> a) { return a / length(a); }
template <class R> LMN_HD bool near_zero(V3<R> a) {  // src/vec3.rs:59-62
  const R eps = R(1e-8);
  return m_abs(a.x) < eps && m_abs(a.y) < eps && m_abs(a.z) < eps;
}
template <class R> LMN_HD V3<R> reflect(V3<R> v, V3<R> n) {  // src/vec3.rs:64-66
  return v - n * R(2) * dot(v, n);
}
template <class R> LMN_HD V3<R> reflect_around(V3<R> v, V3<R> h) {  // src/vec3.rs:69-71
  return h * R(2) * dot(v, h) - v;
}
template <class R> LMN_HD V3<R> refract(V3<R> v, V3<R> n, R eta) {  // src/vec3.rs:73-78
  R c = dot(-v, n);
  R cos_theta = c < R(1) ? c : R(1);
  V3<R> r_perp = (v + n * cos_theta) * eta;
  V3<R> r_par = n * -vsqrt(m_abs(R(1) - length_squared(r_perp)));
  return r_perp + r_par;
}
template <class R> LMN_HD bool has_nan(V3<R> a) { return a.x != a.x || a.y != a.y || a.z != a.z; }
template <class R> LMN_HD R max3(V3<R> a) {
  R m = a.x > a.y ? a.x : a.y;
  return m > a.z ? m : a.z;
}
template <class To, class From> LMN_HD V3<To> cvt(V3<From> a) { return {To(a.x), To(a.y), To(a.z)}; }

// f64::max / f64::min semantics: a NaN operand yields the other operand.
template <class R> LMN_HD R fmax_r(R a, R b) { return m_fmax(a, b); }
template <class R> LMN_HD R fmin_r(R a, R b) { return m_fmin(a, b); }
template <class R> LMN_HD R clamp_r(R x, R lo, R hi) { return x < lo ? lo : (x > hi ? hi : x); }

// ---------------------------------------------------------------------------
// Counter-based RNG.  Every path owns a 64-bit key derived from (seed, scene,
// pixel, sample); draw n of the path is SplitMix64's output function applied
// to key + n * golden-gamma, i.e. the n-th element of a SplitMix64 sequence
// started at the key (Steele, Lea & Flood, OOPSLA 2014).  Stateless, so the
// FP32 wavefront, the FP64 kernel, the CPU oracle and the path-replay adjoint
// all see the same numbers (common random numbers), and a path is replayed
// from (key, counter) alone.  FP64 uses the top 53 bits, FP32 the top 24.
LMN_HD uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}
constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ULL;

LMN_HD uint64_t frame_key(uint64_t seed, uint32_t scene_key) {
  return mix64(seed * 0xD1B54A32D192ED03ULL + uint64_t(scene_key) * 0xA0761D6478BD642FULL + 0x632BE59BD9B4E019ULL);
}
LMN_HD uint64_t path_key(uint64_t fkey, uint64_t pixel, uint32_t sample) {
  return mix64(fkey ^ mix64((pixel << 24) ^ uint64_t(sample) ^ 0x8CB92BA72F3D8DD7ULL));
}
// Rust float semantics of the image outputs: f64::clamp keeps a NaN, and
// `f64 as u8` saturates, truncates toward zero and maps NaN to 0.
LMN_HD double clamp_keep_nan(double x, double lo, double hi) { return x < lo ? lo : (x > hi ? hi : x); }
LMN_HD uint8_t sat_u8(double x) { return !(x > 0.0) ? uint8_t(0) : (x >= 255.0 ? uint8_t(255) : uint8_t(x)); }

// Stream of a pixel's single AOV / depth-fog ray (the reference seeds it per
// pixel, apart from the sample streams).
LMN_HD uint64_t aov_key(uint64_t fkey, uint64_t pixel) {
  return mix64(fkey ^ mix64(pixel ^ 0x3C6EF372FE94F82BULL) ^ 0xA4093822299F31D0ULL);
}

template <class R> LMN_HD R u01_from(uint64_t x);
template <> LMN_HD double u01_from<double>(uint64_t x) { return double(x >> 11) * 0x1.0p-53; }
template <> LMN_HD float u01_from<float>(uint64_t x) { return float(uint32_t(x >> 40)) * 0x1.0p-24f; }

struct Rng {
  uint64_t key;
  uint32_t n;
  LMN_HD uint64_t next_u64() { return mix64(key + uint64_t(++n) * kGolden); }
  template <class R> LMN_HD R next() { return u01_from<R>(next_u64()); }
};

// Rejection samplers with the reference's draw order and arithmetic
// (src/vec3.rs:86-122).
template <class R> LMN_HD V3<R> random_in_unit_sphere(Rng& rng) {
  for (;;) {
    R a = rng.next<R>(), b = rng.next<R>(), c = rng.next<R>();
    V3<R> v = {R(-1) + a * R(2), R(-1) + b * R(2), R(-1) + c * R(2)};
    if (length_squared(v) < R(1)) return v;
  }
}
template <class R> LMN_HD V3<R> random_unit_vector(Rng& rng) { return unit(random_in_unit_sphere<R>(rng)); }
template <class R> LMN_HD V3<R> random_in_unit_disk(Rng& rng) {
  for (;;) {
    R a = rng.next<R>(), b = rng.next<R>();
    V3<R> v = {a * R(2) - R(1), b * R(2) - R(1), R(0)};
    if (length_squared(v) < R(1)) return v;
  }
}

}  // namespace lmn

// Image outputs on the device.  The default post chain that the pixel loop's
// caller runs on the linear HDR framebuffer -- exposure (optionally the
// automatic key-value exposure), depth fog, the tone curve, the sRGB transfer
// function and 8-bit quantisation (src/render.rs:2594-2632, :3099-3202,
// :3606-3658) -- and the Radiance RGBE pixel encoding of the HDR writer
// (src/main.rs:948-989).
//
// One thread per pixel.  FP64 arithmetic in the reference's operation order
// with contraction off, so the bytes equal an FP64 CPU evaluation except where
// a value lies within an ulp or two of a quantisation step (pow, exp and log
// of two maths libraries may differ there).  Per pixel the chain reads 12 B of
// HDR (+4 B fog depth) and writes 4 B of RGBA (+12 B exposed HDR when asked):
// HBM-bound: the sRGB power is read off a threshold table (below).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "lmn_kernels.h"

#pragma clang fp contract(off)

// Scratch of the allocation-free post chain (lmn_post_device_noalloc): the
// automatic exposure's per-block partial sums and its two results, in the
// layout launch_log_luminance expects.  Every launch writes all of it before
// reading it, so nothing needs clearing between calls.  Host code reaches it
// only through the per-device address from hipGetSymbolAddress below.
static __device__ __align__(256) double scratch[2 + lmn::kLogLumBlocks];

namespace lmn {
namespace {

// ToneMap curves (src/render.rs:3606-3641) and ToneMap::None's clamp (:2631).
__device__ double hable(double v) {
  const double a = 0.15, b = 0.50, c = 0.10, d = 0.20, e = 0.02, f = 0.30;
  return ((v * (a * v + c * b) + d * e) / (v * (a * v + b) + d * f)) - e / f;
}
__device__ double tone_curve(uint32_t mode, double x) {
  switch (mode) {
    case LMN_TONE_ACES: {  // Narkowicz's fit of the ACES RRT+ODT
      x = fmax(x, 0.0);
      const double a = 2.51, b = 0.03, c = 2.43, d = 0.59, e = 0.14;
      return clamp_keep_nan((x * (a * x + b)) / (x * (c * x + d) + e), 0.0, 1.0);
    }
    case LMN_TONE_REINHARD:
      x = fmax(x, 0.0);
      return x / (1.0 + x);
    case LMN_TONE_FILMIC:  // Hable's Uncharted 2 curve, white point 11.2
      x = fmax(x, 0.0);
      return clamp_keep_nan(hable(x) / hable(11.2), 0.0, 1.0);
    default: return clamp_keep_nan(x, 0.0, 1.0);
  }
}
// gamma_correct with the sRGB transfer function, no dither (:3644-3658),
// is a non-decreasing step function of its argument: byte k is reached at a
// threshold x_k.  The thresholds are found once on the host by bisection over
// the doubles in [0, 1] with the host's own evaluation of gamma_correct (the
// same pow the reference calls), so the byte the device reads off the table
// is the byte the host computes, for every double.  NaN and x < x_1 give 0,
// x >= 1 gives 255 (s = 1 is clamped to 0.999 before the scaling by 256).
__constant__ double c_srgb_threshold[256];  // [k] = smallest x with byte(x) >= k; [0] unused
// byte(b / 4096) for the 4096 bins of [0, 1): no bin holds two thresholds (the
// steepest part of the curve, 12.92 * 256 bytes per unit, puts thresholds
// 1/3307 apart), so a bin's base byte and one comparison give the byte.
__constant__ uint32_t c_srgb_bin[1024];  // four bin bytes per word

__device__ __forceinline__ uint32_t srgb_byte(const double* __restrict__ t, const uint8_t* __restrict__ bin,
                                              double x) {
  if (!(x > 0.0)) return 0u;  // NaN, negative, 0
  if (x >= 1.0) return 255u;
  uint32_t b = bin[uint32_t(x * 4096.0)];
  return b + (b < 255u && x >= t[b + 1] ? 1u : 0u);
}

__device__ __forceinline__ uint32_t post_pixel(const PostArgs& A, const double* t, const uint8_t* bin, double ex,
                                               const float* rgb, float fog_depth, float* hx) {
  double c[3];
  for (int k = 0; k < 3; ++k) c[k] = double(rgb[k]) * ex;
  for (int k = 0; k < 3; ++k) hx[k] = float(c[k]);  // hdr_data (:2610-2624)
  if (A.fog) {  // :3117-3124, depth +inf where the pixel's ray escapes
    double f = 1.0 - exp(-A.fog_density * double(fog_depth));
    for (int k = 0; k < 3; ++k) c[k] = c[k] * (1.0 - f) + A.fog_rgb[k] * f;
  }
  uint32_t w = 0xff000000u;
  for (int k = 0; k < 3; ++k) w |= srgb_byte(t, bin, tone_curve(A.tone, c[k])) << (8 * k);
  return w;
}

// Four pixels per thread: 48 B of HDR in three 16-B loads, 16 B of RGBA out
// in one store (when the buffers are 16-B aligned; the last group and
// unaligned buffers take the per-pixel path).  A grid of a few blocks per CU
// strides over the groups, so each block fills its LDS tables once.
__global__ __launch_bounds__(256) void k_post(PostArgs A) {
  __shared__ double t[256];
  __shared__ uint32_t bin_w[1024];
  t[threadIdx.x] = c_srgb_threshold[threadIdx.x];
  for (uint32_t k = threadIdx.x; k < 1024; k += 256) bin_w[k] = c_srgb_bin[k];
  __syncthreads();
  const uint8_t* bin = reinterpret_cast<const uint8_t*>(bin_w);
  const double ex = A.exposure_dev ? A.exposure_dev[1] : A.exposure;
  for (uint64_t p0 = (uint64_t(blockIdx.x) * blockDim.x + threadIdx.x) * 4; p0 < A.n;
       p0 += uint64_t(gridDim.x) * blockDim.x * 4) {
    const uint32_t m = A.n - p0 < 4 ? uint32_t(A.n - p0) : 4u;
    float v[12], hx[12], fd[4] = {0.f, 0.f, 0.f, 0.f};
    uint32_t w[4];
    if (m == 4 && A.vec) {
      const float4* src = reinterpret_cast<const float4*>(A.hdr + 3 * p0);
      float4 a = src[0], b = src[1], c = src[2];
      v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w; v[4] = b.x; v[5] = b.y;
      v[6] = b.z; v[7] = b.w; v[8] = c.x; v[9] = c.y; v[10] = c.z; v[11] = c.w;
      if (A.fog) {
        float4 f = *reinterpret_cast<const float4*>(A.fog + p0);
        fd[0] = f.x; fd[1] = f.y; fd[2] = f.z; fd[3] = f.w;
      }
    } else {
      for (uint32_t q = 0; q < m; ++q) {
        for (int k = 0; k < 3; ++k) v[3 * q + k] = A.hdr[3 * (p0 + q) + k];
        if (A.fog) fd[q] = A.fog[p0 + q];
      }
    }
    for (uint32_t q = 0; q < m; ++q) w[q] = post_pixel(A, t, bin, ex, v + 3 * q, fd[q], hx + 3 * q);
    if (m == 4 && A.vec) {
      if (A.rgba) *reinterpret_cast<uint4*>(A.rgba + 4 * p0) = make_uint4(w[0], w[1], w[2], w[3]);
      if (A.hdr_exposed) {
        float4* dst = reinterpret_cast<float4*>(A.hdr_exposed + 3 * p0);
        dst[0] = make_float4(hx[0], hx[1], hx[2], hx[3]);
        dst[1] = make_float4(hx[4], hx[5], hx[6], hx[7]);
        dst[2] = make_float4(hx[8], hx[9], hx[10], hx[11]);
      }
    } else {
      for (uint32_t q = 0; q < m; ++q) {
        if (A.rgba)
          for (int k = 0; k < 4; ++k) A.rgba[4 * (p0 + q) + k] = uint8_t(w[q] >> (8 * k));
        if (A.hdr_exposed)
          for (int k = 0; k < 3; ++k) A.hdr_exposed[3 * (p0 + q) + k] = hx[3 * q + k];
      }
    }
  }
}

// compute_log_average_luminance (src/render.rs:525-541): exp of the mean of
// ln(1e-4 + max(L, 0)), L the Rec. 709 luminance; a fixed grid and fixed
// reduction trees, so the sum is the same on every run.
__global__ __launch_bounds__(256) void k_log_lum_partial(const float* __restrict__ hdr, uint64_t n,
                                                         double* __restrict__ partial) {
  __shared__ double red[256];
  double s = 0.0;
  for (uint64_t i = uint64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += uint64_t(gridDim.x) * 256) {
    double x = hdr[3 * i], y = hdr[3 * i + 1], z = hdr[3 * i + 2];
    double lum = 0.2126 * x + 0.7152 * y + 0.0722 * z;
    s += log(1e-4 + fmax(lum, 0.0));
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (uint32_t w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}
// out[0] = log-average luminance, out[1] = exposure to apply (:2595-2608).
__global__ __launch_bounds__(kLogLumBlocks) void k_log_lum_final(const double* __restrict__ partial, uint64_t n,
                                                                 double exposure, double* __restrict__ out) {
  __shared__ double red[kLogLumBlocks];
  red[threadIdx.x] = partial[threadIdx.x];
  __syncthreads();
  for (uint32_t w = kLogLumBlocks / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double avg = n ? exp(red[0] / double(n)) : 0.0;
    double key = avg > 1e-6 ? 0.18 / avg : 1.0;
    out[0] = avg;
    out[1] = key * exposure;
  }
}

// Radiance RGBE pixel (src/main.rs:964-986): channels clamped at 0 with
// f32::max, shared exponent ceil(log2(max)) from the correctly rounded f32
// log2 (FP64 log2 rounded once to f32), mantissas min(c * 256 / 2^e, 255)
// truncated, exponent byte (e + 128) wrapped to 8 bits.
__device__ __forceinline__ uint32_t rgbe_pixel(float r, float g, float b) {
  r = fmaxf(r, 0.f);
  g = fmaxf(g, 0.f);
  b = fmaxf(b, 0.f);
  float m = fmaxf(fmaxf(r, g), b);
  if (m < 1e-32f) return 0u;
  // ceil of the correctly rounded f32 log2: the hardware log2 (1 ulp, so
  // within 2e-5 for |log2| <= 128) decides it unless it lies within 1e-3 of
  // an integer; there the FP64 log2 rounded once to f32 does
  float l = __log2f(m);
  float c = ceilf(l);
  if (!(c - l > 1e-3f && l - (c - 1.0f) > 1e-3f)) c = ceilf(float(log2(double(m))));
  int32_t e = c >= 2147483647.f ? 2147483647 : int32_t(c);  // `as i32` saturates
  float scale = 256.0f / ldexpf(1.0f, e > 1024 ? 1024 : e);
  uint32_t w = sat_u8(double(fminf(r * scale, 255.0f)));
  w |= uint32_t(sat_u8(double(fminf(g * scale, 255.0f)))) << 8;
  w |= uint32_t(sat_u8(double(fminf(b * scale, 255.0f)))) << 16;
  return w | (uint32_t(uint8_t(uint32_t(e) + 128u)) << 24);
}
// Four pixels per thread, as k_post.
__global__ __launch_bounds__(256) void k_rgbe(const float* __restrict__ rgb, uint64_t n, uint8_t* __restrict__ out,
                                              uint32_t vec) {
  const uint64_t p0 = (uint64_t(blockIdx.x) * blockDim.x + threadIdx.x) * 4;
  if (p0 >= n) return;
  const uint32_t m = n - p0 < 4 ? uint32_t(n - p0) : 4u;
  if (m == 4 && vec) {
    const float4* src = reinterpret_cast<const float4*>(rgb + 3 * p0);
    float4 a = src[0], b = src[1], c = src[2];
    *reinterpret_cast<uint4*>(out + 4 * p0) = make_uint4(rgbe_pixel(a.x, a.y, a.z), rgbe_pixel(a.w, b.x, b.y),
                                                         rgbe_pixel(b.z, b.w, c.x), rgbe_pixel(c.y, c.z, c.w));
    return;
  }
  for (uint32_t q = 0; q < m; ++q) {
    const float* px = rgb + 3 * (p0 + q);
    uint32_t w = rgbe_pixel(px[0], px[1], px[2]);
    for (int k = 0; k < 4; ++k) out[4 * (p0 + q) + k] = uint8_t(w >> (8 * k));
  }
}

// RenderResult.depth_pass holds 0 where the pixel's ray escapes (:3532).
__global__ __launch_bounds__(256) void k_depth_pass(float* __restrict__ depth, uint64_t n) {
  uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n && depth[i] == INFINITY) depth[i] = 0.0f;
}

uint32_t blocks_for(uint64_t n) { return uint32_t((n + 255) / 256); }
bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }

// Host evaluation of gamma_correct (src/render.rs:3649-3658) for the table.
uint32_t host_srgb_byte(double x) {
  x = clamp_keep_nan(x, 0.0, 1.0);
  double s = x <= 0.0031308 ? 12.92 * x : 1.055 * std::pow(x, 1.0 / 2.4) - 0.055;
  return sat_u8(clamp_keep_nan(s, 0.0, 0.999) * 256.0);
}
struct SrgbTable {
  double t[256];
  uint32_t bin[1024];
  bool ok = true;  // at most one threshold per bin
  SrgbTable() {
    t[0] = -INFINITY;
    uint64_t one, zero;
    const double d1 = 1.0, d0 = 0.0;
    std::memcpy(&one, &d1, 8);
    std::memcpy(&zero, &d0, 8);
    for (uint32_t k = 1; k < 256; ++k) {  // smallest double in [0, 1] with byte >= k
      uint64_t lo = zero, hi = one;
      while (lo < hi) {
        uint64_t mid = lo + (hi - lo) / 2;
        double x;
        std::memcpy(&x, &mid, 8);
        if (host_srgb_byte(x) >= k)
          hi = mid;
        else
          lo = mid + 1;
      }
      std::memcpy(&t[k], &lo, 8);
    }
    uint8_t b8[4096];
    for (uint32_t b = 0; b < 4096; ++b) {
      double x0 = double(b) / 4096.0, x1 = double(b + 1) / 4096.0;
      uint32_t below = 0, inside = 0;
      for (uint32_t k = 1; k < 256; ++k) {
        below += t[k] <= x0;
        inside += t[k] > x0 && t[k] < x1;
      }
      b8[b] = uint8_t(below);
      ok = ok && inside <= 1;
    }
    std::memcpy(bin, b8, sizeof b8);
  }
};
std::mutex g_table_mu;
std::vector<int> g_table_devices;

}  // namespace

namespace {
constexpr int kMaxNoAllocDevices = 64;
std::mutex g_noalloc_mu;
double* g_noalloc_scratch[kMaxNoAllocDevices] = {};
}  // namespace

hipError_t noalloc_post_scratch(double** out, size_t* doubles) {
  *out = nullptr;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxNoAllocDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> g(g_noalloc_mu);
  if (!g_noalloc_scratch[dev]) {  // once per device, with that device current
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_noalloc_scratch[dev] = static_cast<double*>(p);
  }
  *out = g_noalloc_scratch[dev];
  if (doubles) *doubles = sizeof(scratch) / sizeof(double);
  return hipSuccess;
}

hipError_t ensure_post_tables() {
  static const SrgbTable table;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  std::lock_guard<std::mutex> g(g_table_mu);
  for (int d : g_table_devices)
    if (d == dev) return hipSuccess;
  if (!table.ok) return hipErrorInvalidValue;  // a bin with two thresholds: the lookup would be wrong
  e = hipMemcpyToSymbol(HIP_SYMBOL(c_srgb_threshold), table.t, sizeof table.t);
  if (e == hipSuccess) e = hipMemcpyToSymbol(HIP_SYMBOL(c_srgb_bin), table.bin, sizeof table.bin);
  if (e == hipSuccess) g_table_devices.push_back(dev);
  return e;
}

void launch_post(const PostArgs& A, hipStream_t s) {
  if (!A.n) return;
  PostArgs B = A;
  B.vec = aligned16(A.hdr) && (!A.fog || aligned16(A.fog)) && (!A.rgba || aligned16(A.rgba)) &&
          (!A.hdr_exposed || aligned16(A.hdr_exposed));
  const uint32_t blocks = std::min<uint32_t>(blocks_for((A.n + 3) / 4), 2048u);
  hipLaunchKernelGGL(k_post, dim3(blocks), dim3(256), 0, s, B);
}
void launch_log_luminance(const float* hdr, uint64_t n, double exposure, double* scratch, hipStream_t s) {
  hipLaunchKernelGGL(k_log_lum_partial, dim3(kLogLumBlocks), dim3(256), 0, s, hdr, n, scratch + 2);
  hipLaunchKernelGGL(k_log_lum_final, dim3(1), dim3(kLogLumBlocks), 0, s, scratch + 2, n, exposure, scratch);
}
void launch_rgbe(const float* rgb, uint64_t n, uint8_t* out, hipStream_t s) {
  if (n)
    hipLaunchKernelGGL(k_rgbe, dim3(blocks_for((n + 3) / 4)), dim3(256), 0, s, rgb, n, out,
                       uint32_t(aligned16(rgb) && aligned16(out)));
}
void launch_depth_pass(float* depth, uint64_t n, hipStream_t s) {
  if (n) hipLaunchKernelGGL(k_depth_pass, dim3(blocks_for(n)), dim3(256), 0, s, depth, n);
}

}  // namespace lmn

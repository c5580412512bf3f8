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
// HBM-bound in principle, bound by the three FP64 pow calls in practice.
#include "lmn_kernels.h"

#pragma clang fp contract(off)

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
// gamma_correct with the sRGB transfer function, no dither (:3644-3658).
__device__ uint8_t srgb_byte(double x) {
  x = clamp_keep_nan(x, 0.0, 1.0);
  double s = x <= 0.0031308 ? 12.92 * x : 1.055 * pow(x, 1.0 / 2.4) - 0.055;
  return sat_u8(clamp_keep_nan(s, 0.0, 0.999) * 256.0);
}

__global__ __launch_bounds__(256) void k_post(PostArgs A) {
  uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= A.n) return;
  const double ex = A.exposure_dev ? A.exposure_dev[1] : A.exposure;
  double c[3];
  for (int k = 0; k < 3; ++k) c[k] = double(A.hdr[3 * i + k]) * ex;
  if (A.hdr_exposed)  // hdr_data (:2610-2624)
    for (int k = 0; k < 3; ++k) A.hdr_exposed[3 * i + k] = float(c[k]);
  if (A.fog) {  // :3117-3124, depth +inf where the pixel's ray escapes
    double f = 1.0 - exp(-A.fog_density * double(A.fog[i]));
    for (int k = 0; k < 3; ++k) c[k] = c[k] * (1.0 - f) + A.fog_rgb[k] * f;
  }
  if (!A.rgba) return;
  uchar4 o;
  o.x = srgb_byte(tone_curve(A.tone, c[0]));
  o.y = srgb_byte(tone_curve(A.tone, c[1]));
  o.z = srgb_byte(tone_curve(A.tone, c[2]));
  o.w = 255;
  reinterpret_cast<uchar4*>(A.rgba)[i] = o;
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
__global__ __launch_bounds__(256) void k_rgbe(const float* __restrict__ rgb, uint64_t n, uchar4* __restrict__ out) {
  uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float r = fmaxf(rgb[3 * i], 0.f), g = fmaxf(rgb[3 * i + 1], 0.f), b = fmaxf(rgb[3 * i + 2], 0.f);
  float m = fmaxf(fmaxf(r, g), b);
  uchar4 o = {0, 0, 0, 0};
  if (!(m < 1e-32f)) {
    float c = ceilf(float(log2(double(m))));
    int32_t e = c >= 2147483647.f ? 2147483647 : int32_t(c);  // `as i32` saturates
    float scale = 256.0f / ldexpf(1.0f, e > 1024 ? 1024 : e);
    o.x = sat_u8(double(fminf(r * scale, 255.0f)));
    o.y = sat_u8(double(fminf(g * scale, 255.0f)));
    o.z = sat_u8(double(fminf(b * scale, 255.0f)));
    o.w = uint8_t(uint32_t(e) + 128u);
  }
  out[i] = o;
}

// RenderResult.depth_pass holds 0 where the pixel's ray escapes (:3532).
__global__ __launch_bounds__(256) void k_depth_pass(float* __restrict__ depth, uint64_t n) {
  uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n && depth[i] == INFINITY) depth[i] = 0.0f;
}

uint32_t blocks_for(uint64_t n) { return uint32_t((n + 255) / 256); }

}  // namespace

void launch_post(const PostArgs& A, hipStream_t s) {
  if (A.n) hipLaunchKernelGGL(k_post, dim3(blocks_for(A.n)), dim3(256), 0, s, A);
}
void launch_log_luminance(const float* hdr, uint64_t n, double exposure, double* scratch, hipStream_t s) {
  hipLaunchKernelGGL(k_log_lum_partial, dim3(kLogLumBlocks), dim3(256), 0, s, hdr, n, scratch + 2);
  hipLaunchKernelGGL(k_log_lum_final, dim3(1), dim3(kLogLumBlocks), 0, s, scratch + 2, n, exposure, scratch);
}
void launch_rgbe(const float* rgb, uint64_t n, uint8_t* out, hipStream_t s) {
  if (n) hipLaunchKernelGGL(k_rgbe, dim3(blocks_for(n)), dim3(256), 0, s, rgb, n, reinterpret_cast<uchar4*>(out));
}
void launch_depth_pass(float* depth, uint64_t n, hipStream_t s) {
  if (n) hipLaunchKernelGGL(k_depth_pass, dim3(blocks_for(n)), dim3(256), 0, s, depth, n);
}

}  // namespace lmn

// Plain one-path-per-thread megakernels (FP64 baseline and FP32) and the
// ray-batch trace kernel of the parity harness.
#include "lmn_kernels.h"

namespace lmn {

// ---------------------------------------------------------------------------
// Plain megakernel: one thread per pixel, paths run to completion one after
// another, private-memory traversal stack, FP64 (or FP32) throughout.

// One sample (work item k) of ray_color (src/render.rs:1472-1537): the
// unclamped radiance and the filter weight.
template <class R, uint32_t FEAT>
__device__ V3<R> plain_sample(const SceneView<R>& S, const FrameDev& F, uint64_t k, R& w,
                              unsigned long long& n_ext, unsigned long long& n_sh) {
  Rng rng;
  V3<R> o, d;
  uint32_t pix;
  generate_item<R>(F, S.cam, k, rng, o, d, w, pix);
  V3<R> L = {R(0), R(0), R(0)}, beta = {R(1), R(1), R(1)};
  bool skip = false;
  // FP32: the record a bounce ray leaves from (test_prim's origin rule); the
  // FP64 kernel keeps the reference's t >= 0.001 alone, as the oracle does
  uint32_t org = kNoOrigin;
  for (uint32_t depth = F.max_depth; depth > 0; --depth) {
    PrivateStack<R> st;
    R t;
    uint32_t idx;
    ++n_ext;
    if (!trace<R, FEAT, false>(S, o, d, R(0.001), R(INFINITY), t, idx, st, nullptr, org)) {
      L = L + hadamard(beta, background(S.bg, d));
      break;
    }
    Surf<R> sf;
    surface<R, FEAT>(S, idx, o, d, t, sf);
    const MaterialT<R>& m = S.materials[sf.mat];
    V3<R> em = skip ? V3<R>{R(0), R(0), R(0)} : emitted(S, m, sf.u, sf.v);
    V3<R> ndir, att;
    if (!scatter(S, m, d, sf, rng, ndir, att)) {
      L = L + hadamard(beta, em);
      break;
    }
    R surv = clamp_r(max3(att), R(0.05), R(1));
    bool rr = depth < 47;
    if (rr && rng.next<R>() > surv) {
      L = L + hadamard(beta, em);
      break;
    }
    R rrw = rr ? R(1) / surv : R(1);
    bool use_nee = !is_specular(m) && S.n_lights > 0;
    V3<R> direct = {R(0), R(0), R(0)};
    bool nee_nan = false;
    if (use_nee) {
      V3<R> ld, c;
      R dist;
      if (nee_sample(S, sf.p, sf.n, att, rng, ld, dist, c)) {
        nee_nan = has_nan(c);
        PrivateStack<R> st2;
        R t2;
        uint32_t i2;
        ++n_sh;
        if (!trace<R, FEAT, true>(S, sf.p, ld, R(0.001), dist - R(0.001), t2, i2, st2)) direct = c;
      }
    }
    V3<R> attw = att * rrw;
    if (has_nan(em) || nee_nan || has_nan(attw)) break;  // level result NaN -> 0
    L = L + hadamard(beta, em + direct);
    beta = hadamard(beta, attw);
    o = sf.p;
    d = ndir;
    skip = use_nee;
    if (sizeof(R) == 4) org = idx;
  }
  return L;
}

// Sums a thread's ray counts over its wave and adds them with one atomic
// per wave (one per thread would serialise every path on two addresses).
__device__ __forceinline__ void add_ray_counts(unsigned long long* rays, unsigned long long n_ext,
                                               unsigned long long n_sh) {
  for (int off = 32; off > 0; off >>= 1) {
    n_ext += __shfl_xor(n_ext, off, 64);
    n_sh += __shfl_xor(n_sh, off, 64);
  }
  if ((threadIdx.x & 63u) == 0u) {
    if (n_ext) atomicAdd(&rays[0], n_ext);
    if (n_sh) atomicAdd(&rays[1], n_sh);
  }
}

// One path per thread: work item k0 + j (k = s * npix + p, sample s of crop
// pixel p) for j < n.  The unclamped radiance and the filter weight go to
// samples[j]; k_plain_resolve adds them per pixel in sample order.  OCC: the
// waves per SIMD the build is held to (its register budget); the FP64 build
// takes 250+ VGPRs left alone (2 waves), at 4 it spills some.  Which runs
// fastest depends on the scene (path length, traversal depth), so the FP64
// kernel is built for 2, 3 and 4 and the caller picks.
template <class R, uint32_t FEAT, int OCC>
__global__ __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(OCC, 8))) void k_plain(SceneView<R> S, FrameDev F, double4* __restrict__ samples, uint64_t k0,
                                               uint64_t n, unsigned long long* rays) {
  const uint64_t j = uint64_t(blockIdx.x) * 128u + threadIdx.x;
  unsigned long long n_ext = 0, n_sh = 0;
  if (j < n) {
    R w;
    V3<R> L = plain_sample<R, FEAT>(S, F, k0 + j, w, n_ext, n_sh);
    samples[j] = make_double4(double(L.x), double(L.y), double(L.z), double(w));
  }
  if (rays) add_ray_counts(rays, n_ext, n_sh);  // every lane of the wave takes part
}

// Per pixel, the samples of `passes` consecutive sample passes in sample
// order: luminance clamp at 100 (src/render.rs:1692-1697), filter weight,
// FP64 sums (:1699-1705).
__global__ __launch_bounds__(256) void k_plain_resolve(const double4* __restrict__ samples, double4* __restrict__ acc,
                                                       uint32_t npix, uint32_t passes) {
  const uint32_t p = blockIdx.x * 256u + threadIdx.x;
  if (p >= npix) return;
  double4 a = acc[p];
  auto add = [&](const double4& v) {
    double x = v.x, y = v.y, z = v.z;
    double lum = 0.2126 * x + 0.7152 * y + 0.0722 * z;
    if (lum > 100.0) {
      double f = 100.0 / lum;
      x *= f;
      y *= f;
      z *= f;
    }
    a.x += x * v.w;
    a.y += y * v.w;
    a.z += z * v.w;
    a.w += v.w;
  };
  // the sample loads go out eight at a time (one memory round trip per
  // eight samples); the sums keep sample order
  uint32_t s = 0;
  for (; s + 8u <= passes; s += 8u) {
    double4 v[8];
#pragma unroll
    for (uint32_t k = 0; k < 8u; ++k) v[k] = samples[size_t(s + k) * npix + p];
#pragma unroll
    for (uint32_t k = 0; k < 8u; ++k) add(v[k]);
  }
  for (; s < passes; ++s) add(samples[size_t(s) * npix + p]);
  acc[p] = a;
}

// Adaptive sampling, one pixel per thread in sample order: sample s of a
// pixel still sampling goes into its state (src/render.rs:1626-1677).
template <class R, uint32_t FEAT>
__global__ __launch_bounds__(128) void k_plain_adaptive(SceneView<R> S, FrameDev F, AdaptState* __restrict__ state,
                                                        uint32_t s_begin, uint32_t s_end, uint32_t min_samples,
                                                        double threshold, unsigned long long* rays) {
  uint32_t npix = F.crop_w * F.crop_h;
  uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npix) return;
  AdaptState a = state[p];
  unsigned long long n_ext = 0, n_sh = 0;
  for (uint32_t s = s_begin; s < s_end && !a.done; ++s) {
    R w;
    V3<R> L = plain_sample<R, FEAT>(S, F, uint64_t(s) * npix + p, w, n_ext, n_sh);
    adapt_take(a, double(L.x), double(L.y), double(L.z), min_samples, threshold);
  }
  state[p] = a;
  if (rays) {
    atomicAdd(&rays[0], n_ext);
    atomicAdd(&rays[1], n_sh);
  }
}

// ---------------------------------------------------------------------------
// Ray-batch closest hit (parity harness)
template <class R, uint32_t FEAT>
__global__ void k_trace(SceneView<R> S, const lmn_ray* __restrict__ rays, lmn_hit* __restrict__ hits, size_t n) {
  size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const lmn_ray& r = rays[i];
  V3<R> o = {R(r.origin[0]), R(r.origin[1]), R(r.origin[2])};
  V3<R> d = {R(r.direction[0]), R(r.direction[1]), R(r.direction[2])};
  PrivateStack<R> st;
  R t;
  uint32_t idx;
  lmn_hit h;
  h.pad = 0;
  if (trace<R, FEAT, false>(S, o, d, R(r.t_min), R(r.t_max), t, idx, st)) {
    Surf<R> sf;
    surface<R, FEAT>(S, idx, o, d, t, sf);
    h.t = double(t);
    h.prim = int64_t(sf.prim_id);
    h.u = double(sf.u);
    h.v = double(sf.v);
    h.normal[0] = double(sf.n.x);
    h.normal[1] = double(sf.n.y);
    h.normal[2] = double(sf.n.z);
    h.front_face = sf.front ? 1 : 0;
  } else {
    h.t = 0;
    h.prim = -1;
    h.u = h.v = 0;
    h.normal[0] = h.normal[1] = h.normal[2] = 0;
    h.front_face = 0;
  }
  hits[i] = h;
}

// ---------------------------------------------------------------------------
// AOV passes and the depth-fog buffer (src/render.rs:3508-3584, :2634-2665):
// one ray per crop pixel through the pixel centre, u = (i + 0.5) / (W - 1),
// v = (y + 0.5) / (H - 1), from the pixel's own stream; closest hit over
// [0.001, inf).  depth: the hit's t (units of the unnormalised camera
// direction), +inf where the ray escapes; normal: the facing normal mapped to
// bytes by (n * 0.5 + 0.5) * 255; albedo: the attenuation of one scatter, or,
// where the material does not scatter, its emission divided by max(max(e), 1);
// normal and albedo bytes are 0 where the ray escapes.
template <class R, uint32_t FEAT>
__global__ __launch_bounds__(128) void k_aov(SceneView<R> S, FrameDev F, float* __restrict__ depth,
                                             uint8_t* __restrict__ normal, uint8_t* __restrict__ albedo) {
  uint32_t npix = F.crop_w * F.crop_h;
  uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npix) return;
  uint32_t gi = p % F.crop_w + F.crop_x, gj = p / F.crop_w + F.crop_y;
  Rng rng;
  rng.key = aov_key(F.fkey, uint64_t(gj) * F.full_w + gi);
  rng.n = 0;
  R uc = (R(gi) + R(0.5)) / R(F.full_w - 1);
  R vc = (R(F.full_h - 1 - gj) + R(0.5)) / R(F.full_h - 1);
  V3<R> o, d;
  camera_ray(S.cam, uc, vc, rng, o, d);
  PrivateStack<R> st;
  R t;
  uint32_t idx;
  float dep = INFINITY;
  uint8_t nb[3] = {0, 0, 0}, ab[3] = {0, 0, 0};
  if (trace<R, FEAT, false>(S, o, d, R(0.001), R(INFINITY), t, idx, st)) {
    Surf<R> sf;
    surface<R, FEAT>(S, idx, o, d, t, sf);
    dep = float(t);
    const double n3[3] = {double(sf.n.x), double(sf.n.y), double(sf.n.z)};
    for (int k = 0; k < 3; ++k) nb[k] = sat_u8(clamp_keep_nan((n3[k] * 0.5 + 0.5) * 255.0, 0.0, 255.0));
    if (albedo) {
      const MaterialT<R>& m = S.materials[sf.mat];
      V3<R> ndir, att;
      if (scatter(S, m, d, sf, rng, ndir, att)) {
        const double a3[3] = {double(att.x), double(att.y), double(att.z)};
        for (int k = 0; k < 3; ++k) ab[k] = sat_u8(clamp_keep_nan(a3[k], 0.0, 1.0) * 255.0);
      } else {
        V3<R> e = emitted(S, m, sf.u, sf.v);
        const double e3[3] = {double(e.x), double(e.y), double(e.z)};
        double mx = fmax(fmax(e3[0], fmax(e3[1], e3[2])), 1.0);
        for (int k = 0; k < 3; ++k) ab[k] = sat_u8(clamp_keep_nan(e3[k] / mx * 255.0, 0.0, 255.0));
      }
    }
  }
  if (depth) depth[p] = dep;
  if (normal)
    for (int k = 0; k < 3; ++k) normal[3 * size_t(p) + k] = nb[k];
  if (albedo)
    for (int k = 0; k < 3; ++k) albedo[3 * size_t(p) + k] = ab[k];
}

// FP32 feature sets (FEAT_MARCH implies FEAT_F64) and FP64 feature sets
// (FEAT_F64 is meaningless in FP64: only transforms and marched kinds).
#define LMN_FEAT_SWITCH(feat, CALL) \
  switch (feat) {                    \
    case 0: CALL(0); break;          \
    case 1: CALL(1); break;          \
    case 2: CALL(2); break;          \
    case 3: CALL(3); break;          \
    case 6: CALL(6); break;          \
    default: CALL(7); break;         \
  }
#define LMN_FEAT64_SWITCH(feat, CALL) \
  switch (feat & 5u) {                 \
    case 0: CALL(0); break;            \
    case 1: CALL(1); break;            \
    case 4: CALL(4); break;            \
    default: CALL(5); break;           \
  }

void launch_plain_fp64(const SceneView<double>& S, uint32_t feat, const FrameDev& F, double4* samples, uint64_t k0,
                       uint64_t n, unsigned long long* rays, uint32_t waves, hipStream_t s) {
  if (!n) return;
  const dim3 grid(uint32_t((n + 127) / 128));
#define CALL(f)                                                                                                     \
  if (waves == 2) hipLaunchKernelGGL((k_plain<double, f, 2>), grid, dim3(128), 0, s, S, F, samples, k0, n, rays);   \
  else if (waves == 3) hipLaunchKernelGGL((k_plain<double, f, 3>), grid, dim3(128), 0, s, S, F, samples, k0, n, rays); \
  else hipLaunchKernelGGL((k_plain<double, f, 4>), grid, dim3(128), 0, s, S, F, samples, k0, n, rays);
  LMN_FEAT64_SWITCH(feat, CALL)
#undef CALL
}
void launch_plain_fp32(const SceneView<float>& S, uint32_t feat, const FrameDev& F, double4* samples, uint64_t k0,
                       uint64_t n, unsigned long long* rays, hipStream_t s) {
  if (!n) return;
  const dim3 grid(uint32_t((n + 127) / 128));
#define CALL(f) hipLaunchKernelGGL((k_plain<float, f, 1>), grid, dim3(128), 0, s, S, F, samples, k0, n, rays)
  LMN_FEAT_SWITCH(feat, CALL)
#undef CALL
}
void launch_plain_resolve(const double4* samples, double4* acc, uint32_t npix, uint32_t passes, hipStream_t s) {
  if (npix && passes) hipLaunchKernelGGL(k_plain_resolve, dim3((npix + 255) / 256), dim3(256), 0, s, samples, acc, npix, passes);
}
void launch_plain_adaptive_fp64(const SceneView<double>& S, uint32_t feat, const FrameDev& F, AdaptState* state,
                                uint32_t s0, uint32_t s1, uint32_t min_samples, double threshold,
                                unsigned long long* rays, hipStream_t s) {
  uint32_t npix = F.crop_w * F.crop_h;
#define CALL(f)                                                                                          \
  hipLaunchKernelGGL((k_plain_adaptive<double, f>), dim3((npix + 127) / 128), dim3(128), 0, s, S, F, state, s0, s1, \
                     min_samples, threshold, rays)
  LMN_FEAT64_SWITCH(feat, CALL)
#undef CALL
}
void launch_plain_adaptive_fp32(const SceneView<float>& S, uint32_t feat, const FrameDev& F, AdaptState* state,
                                uint32_t s0, uint32_t s1, uint32_t min_samples, double threshold,
                                unsigned long long* rays, hipStream_t s) {
  uint32_t npix = F.crop_w * F.crop_h;
#define CALL(f)                                                                                         \
  hipLaunchKernelGGL((k_plain_adaptive<float, f>), dim3((npix + 127) / 128), dim3(128), 0, s, S, F, state, s0, s1, \
                     min_samples, threshold, rays)
  LMN_FEAT_SWITCH(feat, CALL)
#undef CALL
}
void launch_trace_fp32(const SceneView<float>& S, uint32_t feat, const lmn_ray* rays, lmn_hit* hits, size_t n,
                       hipStream_t s) {
#define CALL(f) hipLaunchKernelGGL((k_trace<float, f>), dim3((n + 127) / 128), dim3(128), 0, s, S, rays, hits, n)
  LMN_FEAT_SWITCH(feat, CALL)
#undef CALL
}
void launch_trace_fp64(const SceneView<double>& S, uint32_t feat, const lmn_ray* rays, lmn_hit* hits, size_t n,
                       hipStream_t s) {
#define CALL(f) hipLaunchKernelGGL((k_trace<double, f>), dim3((n + 127) / 128), dim3(128), 0, s, S, rays, hits, n)
  LMN_FEAT64_SWITCH(feat, CALL)
#undef CALL
}
void launch_aov_fp32(const SceneView<float>& S, uint32_t feat, const FrameDev& F, float* depth, uint8_t* normal,
                     uint8_t* albedo, hipStream_t s) {
  uint32_t npix = F.crop_w * F.crop_h;
#define CALL(f) \
  hipLaunchKernelGGL((k_aov<float, f>), dim3((npix + 127) / 128), dim3(128), 0, s, S, F, depth, normal, albedo)
  LMN_FEAT_SWITCH(feat, CALL)
#undef CALL
}
void launch_aov_fp64(const SceneView<double>& S, uint32_t feat, const FrameDev& F, float* depth, uint8_t* normal,
                     uint8_t* albedo, hipStream_t s) {
  uint32_t npix = F.crop_w * F.crop_h;
#define CALL(f) \
  hipLaunchKernelGGL((k_aov<double, f>), dim3((npix + 127) / 128), dim3(128), 0, s, S, F, depth, normal, albedo)
  LMN_FEAT64_SWITCH(feat, CALL)
#undef CALL
}

}  // namespace lmn

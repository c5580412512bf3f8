// Path-replay adjoint of the render hot path: gradients of an image loss with
// respect to albedo texels, microfacet roughness and one emission intensity.
//
// A work item (sample, pixel) is replayed from its counter-based seed with
// exactly the forward estimator of ray_color (src/render.rs:1472-1537, the
// same device functions as the plain kernels, so the same paths), recording
// each vertex in private memory; nothing is stored between the forward
// render and this pass.  A backward sweep over the vertices then has, at
// vertex k, the radiance Lnext arriving along the scattered ray and the
// throughput beta_k, and adds the local derivative terms.
//
// Estimator (derivative of the expected image, sampled directions held
// fixed, "detached"): with the RR survival probability q = clamp(max att),
// rrw = 1/q, NEE term direct = att (.) D (the reference uses the scatter
// attenuation as NEE albedo and does not weight it by rrw), the expected
// value of a vertex given its scattered direction is
//     em + q * direct + att (.) Lnext,
// so with p the direction's density and f the BSDF-cosine value (att = f/p):
//   d/dtheta = dlog f * (att*rrw (.) Lnext + direct)
//            + [q = max att, unclamped] * (dlog f - dlog p) * direct.
// For albedo (p independent): dlog f * att = d att / d albedo.  For GGX
// roughness on the specular lobe, p = sp * D * ndh / (4 vdh), so
// dlog f = dlog D + dlog G and dlog f - dlog p = dlog G.  Emission enters
// linearly: d/dI = (em + direct from its lights) / I.
#include "lmn_kernels.h"

namespace lmn {

// Vertex capacity of the replay: a path of max_depth bounces records up to
// max_depth vertices in private memory.  Two builds: 16 (the common depths,
// about 1.5 KB of scratch per lane in FP32) and LMN_ADJOINT_MAX_DEPTH = 50,
// the reference's default depth, for deeper renders.
constexpr uint32_t ADJ_MAXV_SMALL = 16;
constexpr uint32_t ADJ_MAXV = LMN_ADJOINT_MAX_DEPTH;
constexpr uint32_t ADJ_MAX_SLOTS = 256;

enum : uint32_t { AV_CONT = 1u, AV_RRU = 2u, AV_EMP = 4u, AV_LIGHTP = 8u, AV_ARG_SHIFT = 4u };

template <class R>
struct AdjVertex {
  V3<R> beta, em, att, D, datt;
  R rrw, surv, dlogf, dlogg;
  int32_t texel;
  uint32_t slot, flags, pad;
};

// d log D / d r and d log G / d r of the GGX specular lobe at roughness r
// (alpha = r^2; D = alpha^2 / (pi (c^2 (alpha^2 - 1) + 1)^2), c = n.h;
// G Schlick-Smith with k = (r + 1)^2 / 8, src/material.rs:270-288).
template <class R>
LMN_HD void ggx_dlog(R r, R ndv, R ndl, R ndh, R& dlogd, R& dlogg) {
  R a = r * r;
  R c2 = ndh * ndh;
  R den = c2 * (a * a - R(1)) + R(1);
  dlogd = (R(2) / a - R(4) * c2 * a / den) * (R(2) * r);
  R k = (r + R(1)) * (r + R(1)) / R(8);
  R dk = (r + R(1)) / R(4);
  dlogg = -((R(1) - ndv) / (ndv * (R(1) - k) + k) + (R(1) - ndl) / (ndl * (R(1) - k) + k)) * dk;
}

// Sum of the filter weights of every sample of each pixel (the resolve's
// denominator), from the same jitter draws as the forward pass.
template <class R>
__global__ void k_adj_wsum(CameraT<R> C, FrameDev F, double* __restrict__ wsum) {
  uint32_t npix = F.crop_w * F.crop_h;
  uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npix) return;
  double s = 0.0;
  for (uint32_t k = 0; k < F.spp_a; ++k) {
    Rng rng;
    V3<R> o, d;
    R w;
    uint32_t pix;
    generate_item<R>(F, C, uint64_t(k) * npix + p, rng, o, d, w, pix);
    s += double(w);
  }
  wsum[p] = s;
}

// The replay's traversal stack: FP32 walks keep their first kAdjLdsStack
// words in LDS and the rest in a global spill area ([depth][thread]), as the
// wavefront kernels do; the FP64 binary walk keeps its private stack.
constexpr int kAdjLdsStack = 16;
constexpr int kAdjBlock = 256;
template <class R>
struct AdjStack {
  using type = PrivateStack<R>;
  static __device__ type make(uint32_t*, uint32_t*, uint32_t, uint32_t) { return type{}; }
};
template <>
struct AdjStack<float> {
  using type = LdsStack<kAdjLdsStack, kAdjBlock, true>;
  static __device__ type make(uint32_t* lds, uint32_t* spill, uint32_t gid, uint32_t nthreads) {
    return type{lds, spill, threadIdx.x, gid, nthreads, 0};
  }
};

// Capped at 4 waves per SIMD (128 VGPRs): the replay is latency-bound and
// the extra wave outweighs the spills (24.1 -> 21.6 ms per 512^2 x 36 spp).
template <class R, uint32_t FEAT, uint32_t MAXV>
__global__ __launch_bounds__(kAdjBlock) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_adjoint(SceneView<R> S, FrameDev F, AdjDev A,
                                                 const float* __restrict__ dldi, const double* __restrict__ wsum,
                                                 R* __restrict__ g_tex, double* __restrict__ part, uint64_t k0,
                                                 uint64_t k1, uint32_t* __restrict__ spill) {
  __shared__ R s_g[ADJ_MAX_SLOTS + 1];  // roughness slots, then the emission sum
  __shared__ uint32_t st_lds[sizeof(R) == 4 ? kAdjLdsStack * kAdjBlock : 1];
  const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x, nthreads = gridDim.x * blockDim.x;
  for (uint32_t i = threadIdx.x; i <= A.n_slots; i += blockDim.x) s_g[i] = R(0);
  __syncthreads();
  R g_em = R(0);
  const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
  for (uint64_t k = k0 + uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; k < k1; k += stride) {
    Rng rng;
    V3<R> o, d;
    R w;
    uint32_t pix;
    generate_item<R>(F, S.cam, k, rng, o, d, w, pix);
    AdjVertex<R> V[MAXV];
    uint32_t nv = 0;
    V3<R> L = {R(0), R(0), R(0)}, beta = {R(1), R(1), R(1)}, tail = {R(0), R(0), R(0)};
    bool skip = false;
    uint32_t org = kNoOrigin;  // FP32: the record a bounce ray leaves from, as in the plain kernels
    // ---- replay (the plain kernels' loop, recording vertices)
    for (uint32_t depth = F.max_depth; depth > 0; --depth) {
      typename AdjStack<R>::type st = AdjStack<R>::make(st_lds, spill, gid, nthreads);
      R t;
      uint32_t idx;
      if (!trace<R, FEAT, false>(S, o, d, R(0.001), R(INFINITY), t, idx, st, nullptr, org)) {
        V3<R> bg = background(S.bg, d);
        L = L + hadamard(beta, bg);
        tail = bg;
        break;
      }
      Surf<R> sf;
      surface<R, FEAT>(S, idx, o, d, t, sf);
      const MaterialT<R>& m = S.materials[sf.mat];
      V3<R> em = skip ? V3<R>{R(0), R(0), R(0)} : emitted(S, m, sf.u, sf.v);
      AdjVertex<R>& v = V[nv];
      v.beta = beta;
      v.em = em;
      v.att = v.D = v.datt = {R(0), R(0), R(0)};
      v.rrw = v.surv = R(1);
      v.dlogf = v.dlogg = R(0);
      v.texel = -1;
      v.slot = ~0u;
      v.flags = (!skip && int32_t(sf.mat) == A.emission_material) ? AV_EMP : 0u;
      V3<R> ndir, att;
      ScatterInfo<R> si;
      if (!scatter(S, m, d, sf, rng, ndir, att, &si)) {
        L = L + hadamard(beta, em);
        ++nv;
        break;
      }
      R amax = max3(att);
      R surv = clamp_r(amax, R(0.05), R(1));
      bool rr = depth < 47;
      if (rr && rng.next<R>() > surv) {
        L = L + hadamard(beta, em);
        ++nv;
        break;
      }
      R rrw = rr ? R(1) / surv : R(1);
      bool use_nee = !is_specular(m) && S.n_lights > 0;
      V3<R> direct = {R(0), R(0), R(0)}, D = {R(0), R(0), R(0)};
      bool nee_nan = false, light_param = false;
      if (use_nee) {
        V3<R> ld, c;
        R dist, geo;
        uint32_t li;
        if (nee_sample(S, sf.p, sf.n, att, rng, ld, dist, c, &geo, &li)) {
          nee_nan = has_nan(c);
          R t2;
          uint32_t i2;
          // the closest-hit walk above left the stack empty: the shadow ray reuses it
          if (!trace<R, FEAT, true>(S, sf.p, ld, R(0.001), dist - R(0.001), t2, i2, st)) {
            direct = c;
            D = S.lights[li].emission * geo;
            light_param = int32_t(S.lights[li].mat) == A.emission_material;
          }
        }
      }
      V3<R> attw = att * rrw;
      if (has_nan(em) || nee_nan || has_nan(attw)) {  // the level's result is NaN -> 0
        tail = {R(0), R(0), R(0)};
        break;
      }
      L = L + hadamard(beta, em + direct);
      v.att = att;
      v.D = D;
      v.rrw = rrw;
      v.surv = surv;
      v.flags |= AV_CONT | (light_param ? AV_LIGHTP : 0u);
      if (rr && amax > R(0.05) && amax < R(1)) {
        uint32_t am = att.x >= att.y ? (att.x >= att.z ? 0u : 2u) : (att.y >= att.z ? 1u : 2u);
        v.flags |= AV_RRU | (am << AV_ARG_SHIFT);
      }
      if ((si.lobe == LOBE_LAMBERT || si.lobe == LOBE_GGX_DIFFUSE) && m.tex >= 0 && m.tex == A.texture) {
        v.texel = int32_t(texel_index(S, m, sf.u, sf.v) - int64_t(A.texel_base));
        v.datt = si.datt;
      }
      if (si.lobe == LOBE_GGX_SPEC && m.param_slot < A.n_slots && amax > R(0)) {
        R dd, dg;
        ggx_dlog(m.roughness, si.ndv, si.ndl, si.ndh, dd, dg);
        v.slot = m.param_slot;
        v.dlogf = dd + dg;
        v.dlogg = dg;
      }
      ++nv;
      beta = hadamard(beta, attw);
      o = sf.p;
      d = ndir;
      skip = use_nee;
      if (sizeof(R) == 4) org = idx;
      tail = {R(0), R(0), R(0)};  // depth runs out: ray_color(depth 0) is black
    }
    // ---- the pixel's adjoint: filter weight and luminance clamp (FP64 as the resolve)
    double x = double(L.x), y = double(L.y), z = double(L.z);
    double gx = double(dldi[size_t(pix) * 3]), gy = double(dldi[size_t(pix) * 3 + 1]),
           gz = double(dldi[size_t(pix) * 3 + 2]);
    double sc = double(w) / wsum[pix];
    double lum = 0.2126 * x + 0.7152 * y + 0.0722 * z;
    if (lum > 100.0) {  // c = s * 100 / lum:  J^T g = 100/lum (g - grad_lum (s.g) / lum)
      double f = 100.0 / lum, sg = (x * gx + y * gy + z * gz) / lum;
      gx = f * (gx - 0.2126 * sg);
      gy = f * (gy - 0.7152 * sg);
      gz = f * (gz - 0.0722 * sg);
    }
    V3<R> delta = {R(gx * sc), R(gy * sc), R(gz * sc)};
    if (delta.x == R(0) && delta.y == R(0) && delta.z == R(0)) continue;
    // ---- backward sweep
    V3<R> Ln = tail;
    for (int j = int(nv) - 1; j >= 0; --j) {
      const AdjVertex<R>& v = V[j];
      V3<R> Aw = hadamard(delta, v.beta);
      if (v.flags & AV_EMP) g_em += dot(Aw, v.em);
      if (!(v.flags & AV_CONT)) {
        Ln = v.em;
        continue;
      }
      V3<R> direct = hadamard(v.att, v.D);
      R Ad = dot(Aw, direct);
      if (v.flags & AV_LIGHTP) g_em += Ad;
      V3<R> thr = hadamard(v.att * v.rrw, Ln);  // att * rrw (.) Lnext
      if (v.texel >= 0) {
        uint32_t am = (v.flags >> AV_ARG_SHIFT) & 3u;
        bool rru = (v.flags & AV_RRU) != 0;
        const R* aw = &Aw.x;
        const R* da = &v.datt.x;
        const R* ln = &Ln.x;
        const R* dd = &v.D.x;
        for (uint32_t c = 0; c < 3; ++c) {
          R g = aw[c] * da[c] * (v.rrw * ln[c] + dd[c]);
          if (rru && c == am) g += da[c] * Ad / v.surv;
          if (g != R(0)) atomicAdd(&g_tex[size_t(v.texel) * 3 + c], g);
        }
      }
      if (v.slot != ~0u) {
        R g = v.dlogf * (dot(Aw, thr) + Ad);
        if (v.flags & AV_RRU) g += v.dlogg * Ad;
        atomicAdd(&s_g[v.slot], g);
      }
      Ln = v.em + direct + thr;
    }
  }
  atomicAdd(&s_g[A.n_slots], g_em);
  __syncthreads();
  for (uint32_t i = threadIdx.x; i <= A.n_slots; i += blockDim.x)
    part[size_t(blockIdx.x) * (A.n_slots + 1) + i] += double(s_g[i]);
}

// Sums the per-block partials in block order (out[i] over n_blocks rows).
__global__ void k_adj_reduce(const double* __restrict__ part, uint32_t n_blocks, uint32_t width,
                             double* __restrict__ out) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= width) return;
  double s = 0.0;
  for (uint32_t b = 0; b < n_blocks; ++b) s += part[size_t(b) * width + i];
  out[i] = s;
}

template <class R>
__global__ void k_adj_to_float(const R* __restrict__ in, float* __restrict__ out, size_t n) {
  size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) out[i] = float(in[i]);
}

template <class R>
__global__ void k_adj_zero(R* __restrict__ p, size_t n) {
  size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = R(0);
}

// FP32 feature sets (FEAT_MARCH implies FEAT_F64) and FP64 feature sets.
#define LMN_ADJ_FEAT_SWITCH(feat, CALL) \
  switch (feat) {                        \
    case 0: CALL(0); break;              \
    case 1: CALL(1); break;              \
    case 2: CALL(2); break;              \
    case 3: CALL(3); break;              \
    case 6: CALL(6); break;              \
    default: CALL(7); break;             \
  }
#define LMN_ADJ_FEAT64_SWITCH(feat, CALL) \
  switch (feat & 5u) {                     \
    case 0: CALL(0); break;                \
    case 1: CALL(1); break;                \
    case 4: CALL(4); break;                \
    default: CALL(5); break;               \
  }

void launch_adj_wsum_fp32(const CameraT<float>& C, const FrameDev& F, double* wsum, hipStream_t s) {
  uint32_t npix = F.crop_w * F.crop_h;
  hipLaunchKernelGGL(k_adj_wsum<float>, dim3((npix + 255) / 256), dim3(256), 0, s, C, F, wsum);
}
void launch_adj_wsum_fp64(const CameraT<double>& C, const FrameDev& F, double* wsum, hipStream_t s) {
  uint32_t npix = F.crop_w * F.crop_h;
  hipLaunchKernelGGL(k_adj_wsum<double>, dim3((npix + 255) / 256), dim3(256), 0, s, C, F, wsum);
}

void launch_adjoint_fp32(const SceneView<float>& S, uint32_t feat, const FrameDev& F, const AdjDev& A,
                         const float* dldi, const double* wsum, float* g_tex, double* part, uint32_t n_blocks,
                         uint64_t k0, uint64_t k1, uint32_t* spill, hipStream_t s) {
#define CALL(f)                                                                                              \
  if (F.max_depth <= ADJ_MAXV_SMALL)                                                                          \
    hipLaunchKernelGGL((k_adjoint<float, f, ADJ_MAXV_SMALL>), dim3(n_blocks), dim3(kAdjBlock), 0, s, S, F, A, dldi, wsum, \
                       g_tex, part, k0, k1, spill);                                                                    \
  else                                                                                                       \
    hipLaunchKernelGGL((k_adjoint<float, f, ADJ_MAXV>), dim3(n_blocks), dim3(kAdjBlock), 0, s, S, F, A, dldi, wsum, g_tex, \
                       part, k0, k1, spill)
  LMN_ADJ_FEAT_SWITCH(feat, CALL)
#undef CALL
}

void launch_adjoint_fp64(const SceneView<double>& S, uint32_t feat, const FrameDev& F, const AdjDev& A,
                         const float* dldi, const double* wsum, double* g_tex, double* part, uint32_t n_blocks,
                         uint64_t k0, uint64_t k1, uint32_t* spill, hipStream_t s) {
#define CALL(f)                                                                                               \
  if (F.max_depth <= ADJ_MAXV_SMALL)                                                                          \
    hipLaunchKernelGGL((k_adjoint<double, f, ADJ_MAXV_SMALL>), dim3(n_blocks), dim3(kAdjBlock), 0, s, S, F, A, dldi, wsum, \
                       g_tex, part, k0, k1, spill);                                                                    \
  else                                                                                                       \
    hipLaunchKernelGGL((k_adjoint<double, f, ADJ_MAXV>), dim3(n_blocks), dim3(kAdjBlock), 0, s, S, F, A, dldi, wsum, g_tex, \
                       part, k0, k1, spill)
  LMN_ADJ_FEAT64_SWITCH(feat, CALL)
#undef CALL
}

void launch_adj_reduce(const double* part, uint32_t n_blocks, uint32_t width, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_adj_reduce, dim3((width + 63) / 64), dim3(64), 0, s, part, n_blocks, width, out);
}

void launch_adj_zero_f32(float* p, size_t n, hipStream_t s) {
  if (n) hipLaunchKernelGGL(k_adj_zero<float>, dim3((n + 255) / 256), dim3(256), 0, s, p, n);
}
void launch_adj_zero_f64(double* p, size_t n, hipStream_t s) {
  if (n) hipLaunchKernelGGL(k_adj_zero<double>, dim3((n + 255) / 256), dim3(256), 0, s, p, n);
}
void launch_adj_to_float(const double* in, float* out, size_t n, hipStream_t s) {
  if (n) hipLaunchKernelGGL(k_adj_to_float<double>, dim3((n + 255) / 256), dim3(256), 0, s, in, out, n);
}

}  // namespace lmn

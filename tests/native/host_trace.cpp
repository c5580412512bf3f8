// Test harness: the FP32 device traversal compiled for the host CPU.
//
// Builds a scene's device tables with the library's own host builder, points
// a SceneView<float> at host memory and runs the same trace() / surface()
// code the kernels run, so the traversal logic (4-wide BVH, FP64 islands,
// transforms, marched kinds) can be checked against the oracle without a GPU.
// Test infrastructure only: the library never runs device code on the CPU.
#include <cmath>
#include <cstring>
#include <vector>
#include <string>

#include "../../include/luminara.h"
#include "../../luminara_amd/csrc/host/scene_build.h"
#include "../../luminara_amd/csrc/lmn_render.h"

using namespace lmn;

namespace {
thread_local std::string g_err;

template <uint32_t FEAT>
void run(const SceneView<float>& S, const lmn_ray* rays, size_t n, lmn_hit* hits, unsigned long long* nodes) {
  for (size_t i = 0; i < n; ++i) {
    const lmn_ray& r = rays[i];
    V3<float> o = {float(r.origin[0]), float(r.origin[1]), float(r.origin[2])};
    V3<float> d = {float(r.direction[0]), float(r.direction[1]), float(r.direction[2])};
    PrivateStack<float> st;
    TraceCounters tc{};
    float t;
    uint32_t idx;
    lmn_hit h;
    std::memset(&h, 0, sizeof h);
    h.prim = -1;
    if (trace<float, FEAT, false>(S, o, d, float(r.t_min), float(r.t_max), t, idx, st, &tc)) {
      Surf<float> sf;
      surface<float, FEAT>(S, idx, o, d, t, sf);
      h.t = double(t);
      h.prim = int64_t(sf.prim_id);
      h.u = double(sf.u);
      h.v = double(sf.v);
      for (int k = 0; k < 3; ++k) h.normal[k] = double((&sf.n.x)[k]);
      h.front_face = sf.front ? 1 : 0;
    }
    hits[i] = h;
  }
  (void)nodes;
}
}  // namespace

extern "C" {

const char* host_trace_error(void) { return g_err.c_str(); }

// Closest hit of every ray with the FP32 path's traversal; returns the
// worst-case stack words of the 4-wide BVH in *stack_words.
int host_trace_fp32(const lmn_scene_desc* desc, const lmn_ray* rays, size_t n, uint32_t exact_marched,
                    lmn_hit* hits, uint32_t* stack_words) {
  HostScene H;
  BvhOptions opt;
  int rc = build_host_scene(*desc, opt, H, g_err);
  if (rc != LMN_OK) return rc;
  if (H.stack4_words > uint32_t(kPrivateStackWords32)) {
    g_err = "BVH too deep for the FP32 traversal stack";
    return LMN_ERR_UNSUPPORTED;
  }
  SceneView<float> S{};
  S.nodes4 = H.nodes4.data();
  S.n_nodes4 = uint32_t(H.nodes4.size());
  S.nodes = H.nodesF.data();
  S.n_nodes = uint32_t(H.nodesF.size());
  S.mnodes = H.mnodesF.data();
  S.n_mnodes = uint32_t(H.mnodesF.size());
  S.prims = H.primsF.data();
  S.prims64 = H.primsD.data();
  S.xforms = H.xfF.data();
  S.xforms64 = H.xfD.data();
  S.materials = H.matF.data();
  S.textures = H.tex.data();
  S.texels = H.texels.data();
  S.lights = H.lightF.data();
  S.attrs = H.attrF.data();
  S.n_bounded = H.n_bounded;
  S.n_planes = H.n_planes;
  S.n_lights = uint32_t(H.lightF.size());
  S.bg = H.bgF;
  S.cam = H.camF;
  S.exact_marched = exact_marched;
  uint32_t feat = H.featF;
  if (feat & FEAT_MARCH) feat |= FEAT_F64;
  switch (feat) {
    case 0: run<0>(S, rays, n, hits, nullptr); break;
    case 1: run<1>(S, rays, n, hits, nullptr); break;
    case 2: run<2>(S, rays, n, hits, nullptr); break;
    case 3: run<3>(S, rays, n, hits, nullptr); break;
    case 6: run<6>(S, rays, n, hits, nullptr); break;
    default: run<7>(S, rays, n, hits, nullptr); break;
  }
  if (stack_words) *stack_words = H.stack4_words;
  return LMN_OK;
}

// The quantised 4-wide nodes against the FP32 boxes they were made from:
// every ray against every node, the kernel's child test (node4_children) and
// the FP32 slab test of the unquantised box (fminf/fmaxf of (plane - o) / d)
// over [t_min, t_max].  out[0] = child tests, out[1] = accepted by the FP32
// test, out[2] = accepted by the quantised test, out[3] = accepted by the
// FP32 test and refused by the quantised one (must be 0), out[4] = quantised
// boxes that do not contain their FP32 box (must be 0).
int host_quant_check(const lmn_scene_desc* desc, const lmn_ray* rays, size_t n, unsigned long long out[5]) {
  HostScene H;
  BvhOptions opt;
  int rc = build_host_scene(*desc, opt, H, g_err);
  if (rc != LMN_OK) return rc;
  std::vector<Node4F> wide;
  collapse_bvh4(H.nodesF, wide);
  std::vector<Node4Q> q;
  quantize_bvh4(wide, q);
  for (int k = 0; k < 5; ++k) out[k] = 0;
  for (size_t i = 0; i < wide.size(); ++i) {
    const Node4F& f = wide[i];
    const Node4Q& g = q[i];
    const uint32_t lo[3] = {g.qlo_x, g.qlo_y, g.qlo_z}, hi[3] = {g.qhi_x, g.qhi_y, g.qhi_z};
    for (uint32_t k = 0; k < f.n; ++k)
      for (int a = 0; a < 3; ++a) {
        const long double step = std::ldexp((long double)1, int(int8_t((g.meta >> (8 * a)) & 0xffu)));
        const long double pl = (long double)g.org[a] + ((lo[a] >> (8 * k)) & 0xffu) * step;
        const long double ph = (long double)g.org[a] + ((hi[a] >> (8 * k)) & 0xffu) * step;
        const float flo = (&f.lox[0])[a * 4 + k], fhi = (&f.hix[0])[a * 4 + k];
        if (!(pl <= flo && ph >= fhi)) ++out[4];
      }
  }
  for (size_t r = 0; r < n; ++r) {
    V3<float> o = {float(rays[r].origin[0]), float(rays[r].origin[1]), float(rays[r].origin[2])};
    V3<float> d = {float(rays[r].direction[0]), float(rays[r].direction[1]), float(rays[r].direction[2])};
    const float tmin = float(rays[r].t_min), tmax = float(rays[r].t_max);
    const V3<float> inv = {1.f / d.x, 1.f / d.y, 1.f / d.z}, qinv = slab_inv(d);
    for (size_t i = 0; i < wide.size(); ++i) {
      const Node4F& f = wide[i];
      Words4 w[4];
      std::memcpy(w, &q[i], sizeof(Node4Q));
      float tk[4];
      uint32_t ch[4];
      node4_children(w, o, qinv, tmin, tmax, tk, ch);
      for (uint32_t k = 0; k < f.n; ++k) {
        float tx0 = (f.lox[k] - o.x) * inv.x, tx1 = (f.hix[k] - o.x) * inv.x;
        float ty0 = (f.loy[k] - o.y) * inv.y, ty1 = (f.hiy[k] - o.y) * inv.y;
        float tz0 = (f.loz[k] - o.z) * inv.z, tz1 = (f.hiz[k] - o.z) * inv.z;
        float tn = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fmaxf(fminf(tz0, tz1), tmin));
        float tf = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fminf(fmaxf(tz0, tz1), tmax));
        const bool a = tn <= tf, b = tk[k] != INFINITY;
        ++out[0];
        out[1] += a;
        out[2] += b;
        out[3] += a && !b;
      }
    }
  }
  return LMN_OK;
}

// The exact-mode FP64 power used by the superellipsoid root finder.
void host_pow_pos(const double* x, const double* y, double* out, size_t n) {
  for (size_t i = 0; i < n; ++i) out[i] = pow_pos(x[i], y[i]);
}

}  // extern "C"

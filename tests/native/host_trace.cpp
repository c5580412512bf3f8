// Test harness: the FP32 device traversal compiled for the host CPU.
//
// Builds a scene's device tables with the library's own host builder, points
// a SceneView<float> at host memory and runs the same trace() / surface()
// code the kernels run, so the traversal logic (4-wide BVH, FP64 islands,
// transforms, marched kinds) can be checked against the oracle without a GPU.
// Test infrastructure only: the library never runs device code on the CPU.
#include <cstring>
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
  S.nodes4 = H.nodes4F.data();
  S.n_nodes4 = uint32_t(H.nodes4F.size());
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

// The exact-mode FP64 power used by the superellipsoid root finder.
void host_pow_pos(const double* x, const double* y, double* out, size_t n) {
  for (size_t i = 0; i < n; ++i) out[i] = pow_pos(x[i], y[i]);
}

}  // extern "C"

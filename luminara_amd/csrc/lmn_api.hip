// C ABI of the MI355X path tracer (include/luminara.h): scene loading and
// upload, render entry points (wavefront FP32 / plain FP64 / plain FP32),
// the ray-batch parity harness, device information.
#include <hip/hip_runtime.h>
#include <sys/stat.h>

#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <memory>
#include <mutex>
#include <new>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/luminara.h"
#include "host/scene_build.h"
#include "host/scene_load.h"
#include "lmn_kernels.h"

using namespace lmn;

namespace {

thread_local std::string g_err;

int set_err(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIP_TRY(expr)                                                                             \
  do {                                                                                            \
    hipError_t e_ = (expr);                                                                       \
    if (e_ != hipSuccess)                                                                         \
      return set_err(LMN_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_));           \
  } while (0)

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  hipError_t ensure(size_t n) {
    if (n <= bytes && p) return hipSuccess;
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    hipError_t e = hipMalloc(&p, n ? n : 16);
    if (e == hipSuccess) bytes = n;
    return e;
  }
  template <class T> T* as() const { return static_cast<T*>(p); }
};

template <class T>
hipError_t upload(DevBuf& b, const std::vector<T>& v) {
  size_t n = std::max<size_t>(v.size() * sizeof(T), 16);
  hipError_t e = b.ensure(n);
  if (e != hipSuccess) return e;
  if (!v.empty()) return hipMemcpy(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice);
  return hipSuccess;
}

// Reusable wavefront working memory (path pools, queues, sample buffer).
struct Workspace {
  std::mutex mu;
  DevBuf st[2][4];  // ob, db, bk, meta for the two pools
  DevBuf hit, hit_b, hit_rec, sh_o, sh_d, sh_c, samples, acc, spill, ctr, tc, rays, out, views, parts, part_begin;
  DevBuf adj_wsum, adj_gtex, adj_part, adj_red, adj_dldi, adj_tex_out;  // path-replay adjoint
  DevBuf adj_spill;  // FP32 adjoint traversal stack beyond its LDS words, [depth][thread]
  DevBuf mq_cand, mq_key, mq_ovf, mq_brk;  // marched-pass candidate queues
  uint32_t pool = 0;
  bool graph_failed[2] = {false, false};  // capture refused once [without, with stage events]: launch directly
  WaveCounters* host_ctr = nullptr;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  static constexpr int kBlock = 8;  // wavefront iterations between host checks
  hipEvent_t stage_ev[kBlock][7] = {};
  int init() {
    if (!stream) {
      HIP_TRY(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
      HIP_TRY(hipEventCreate(&ev0));
      HIP_TRY(hipEventCreate(&ev1));
      for (auto& it : stage_ev)
        for (auto& e : it) HIP_TRY(hipEventCreate(&e));
      HIP_TRY(hipHostMalloc((void**)&host_ctr, sizeof(WaveCounters), hipHostMallocDefault));
    }
    return LMN_OK;
  }
};
Workspace& workspace() {
  static Workspace* w = new Workspace();  // lives until process exit
  return *w;
}

}  // namespace

struct lmn_scene {
  HostScene host;
  DevBuf nodes4, nodesF, nodesD, mnodesF, mnodesD, primsF, primsD, xfF, xfD, matF, matD, tex, texels, lightF, lightD,
      attrF, attrD;
  SceneView<float> viewF{};
  SceneView<double> viewD{};
  lmn_camera camera{};
  lmn_render_config config{};
  size_t device_bytes = 0;
  int device = 0;
};

extern "C" {

const char* lmn_last_error(void) { return g_err.c_str(); }

int lmn_load_scene(const char* toml_text, const char* base_dir, lmn_scene_desc** out) {
  if (!toml_text || !out) return set_err(LMN_ERR_INVALID, "lmn_load_scene: null argument");
  *out = nullptr;
  auto* od = new (std::nothrow) OwnedDesc();
  if (!od) return set_err(LMN_ERR_NOMEM, "out of host memory");
  std::string err;
  int rc = load_scene_toml(toml_text, base_dir ? base_dir : "", *od, err);
  if (rc != LMN_OK) {
    delete od;
    return set_err(rc, err);
  }
  *out = &od->desc;
  return LMN_OK;
}

void lmn_desc_free(lmn_scene_desc* desc) { delete reinterpret_cast<OwnedDesc*>(desc); }

int lmn_desc_set_texture(lmn_scene_desc* desc, uint32_t index, uint32_t width, uint32_t height, const float* rgb) {
  if (!desc || !rgb || !width || !height) return set_err(LMN_ERR_INVALID, "lmn_desc_set_texture: bad argument");
  auto* od = reinterpret_cast<OwnedDesc*>(desc);
  if (index >= od->texels.size()) return set_err(LMN_ERR_INVALID, "lmn_desc_set_texture: index out of range");
  od->texels[index].assign(rgb, rgb + size_t(width) * height * 3);
  od->textures[index].width = width;
  od->textures[index].height = height;
  od->sync();
  return LMN_OK;
}

int lmn_desc_add_texture(lmn_scene_desc* desc, uint32_t width, uint32_t height, const float* rgb,
                         uint32_t* out_index) {
  if (!desc || !rgb || !width || !height) return set_err(LMN_ERR_INVALID, "lmn_desc_add_texture: bad argument");
  auto* od = reinterpret_cast<OwnedDesc*>(desc);
  od->texels.emplace_back(rgb, rgb + size_t(width) * height * 3);
  od->textures.push_back(lmn_texture{width, height, nullptr});
  od->sync();
  if (out_index) *out_index = uint32_t(od->texels.size() - 1);
  return LMN_OK;
}

int lmn_desc_set_material(lmn_scene_desc* desc, uint32_t index, const lmn_material* m) {
  if (!desc || !m) return set_err(LMN_ERR_INVALID, "lmn_desc_set_material: null argument");
  auto* od = reinterpret_cast<OwnedDesc*>(desc);
  if (index >= od->materials.size()) return set_err(LMN_ERR_INVALID, "lmn_desc_set_material: index out of range");
  if (m->texture >= 0 && uint32_t(m->texture) >= od->texels.size())
    return set_err(LMN_ERR_INVALID, "lmn_desc_set_material: texture index out of range");
  od->materials[index] = *m;
  od->sync();
  return LMN_OK;
}

int lmn_desc_add_material(lmn_scene_desc* desc, const lmn_material* m, uint32_t* out_index) {
  if (!desc || !m) return set_err(LMN_ERR_INVALID, "lmn_desc_add_material: null argument");
  auto* od = reinterpret_cast<OwnedDesc*>(desc);
  if (m->texture >= 0 && uint32_t(m->texture) >= od->texels.size())
    return set_err(LMN_ERR_INVALID, "lmn_desc_add_material: texture index out of range");
  od->materials.push_back(*m);
  od->sync();
  if (out_index) *out_index = uint32_t(od->materials.size() - 1);
  return LMN_OK;
}

int lmn_desc_add_mesh(lmn_scene_desc* desc, const double* verts, uint64_t n_verts, const uint32_t* tris,
                      uint64_t n_tris, const double* normals, const double* uvs, uint32_t material,
                      const uint32_t* materials) {
  if (!desc || (n_tris && (!verts || !tris))) return set_err(LMN_ERR_INVALID, "lmn_desc_add_mesh: null argument");
  auto* od = reinterpret_cast<OwnedDesc*>(desc);
  if (od->prims.size() + n_tris >= (1ull << 31)) return set_err(LMN_ERR_INVALID, "too many primitives");
  // every index is checked before the description changes: a failed call
  // leaves it (and its exported views) exactly as they were
  for (uint64_t t = 0; t < n_tris; ++t) {
    if ((materials ? materials[t] : material) >= od->materials.size())
      return set_err(LMN_ERR_INVALID, "lmn_desc_add_mesh: material out of range");
    for (int c = 0; c < 3; ++c)
      if (tris[t * 3 + c] >= n_verts) return set_err(LMN_ERR_INVALID, "lmn_desc_add_mesh: vertex index out of range");
  }
  const bool smooth = normals || uvs;
  try {
    od->prims.reserve(od->prims.size() + n_tris);
    if (smooth) od->vertex_attrs.reserve(od->vertex_attrs.size() + n_tris);
  } catch (const std::bad_alloc&) {
    od->sync();
    return set_err(LMN_ERR_NOMEM, "out of host memory");
  }
  for (uint64_t t = 0; t < n_tris; ++t) {
    lmn_prim p{};
    p.kind = smooth ? LMN_PRIM_MESH_TRIANGLE : LMN_PRIM_TRIANGLE;
    p.material = materials ? materials[t] : material;
    p.transform = -1;
    p.object = od->n_objects++;
    for (int c = 0; c < 3; ++c) {
      const uint32_t vi = tris[t * 3 + c];
      for (int k = 0; k < 3; ++k) p.p[c * 3 + k] = verts[size_t(vi) * 3 + k];
    }
    if (smooth) {
      lmn_vertex_attr a{};
      a.has_normals = normals ? 1u : 0u;
      a.has_uvs = uvs ? 1u : 0u;
      for (int k = 0; k < 9 && normals; ++k) a.normal[k / 3][k % 3] = normals[t * 9 + k];
      for (int k = 0; k < 6 && uvs; ++k) a.uv[k / 2][k % 2] = uvs[t * 6 + k];
      p.p[9] = double(od->vertex_attrs.size());
      od->vertex_attrs.push_back(a);  // capacity reserved above: no reallocation
    }
    od->prims.push_back(p);
  }
  od->sync();
  return LMN_OK;
}

int lmn_desc_add_triangles(lmn_scene_desc* desc, const double* verts, uint64_t n_verts, const uint32_t* tris,
                           uint64_t n_tris, uint32_t material, const uint32_t* materials) {
  return lmn_desc_add_mesh(desc, verts, n_verts, tris, n_tris, nullptr, nullptr, material, materials);
}

int lmn_scene_create(const lmn_scene_desc* desc, lmn_scene** out) {
  if (!desc || !out) return set_err(LMN_ERR_INVALID, "lmn_scene_create: null argument");
  *out = nullptr;
  std::unique_ptr<lmn_scene> s(new (std::nothrow) lmn_scene());
  if (!s) return set_err(LMN_ERR_NOMEM, "out of host memory");
  std::string err;
  BvhOptions opt;
  int rc = build_host_scene(*desc, opt, s->host, err);
  if (rc != LMN_OK) return set_err(rc, err);
  HIP_TRY(hipGetDevice(&s->device));
  HostScene& h = s->host;
  if (h.stack4_words > uint32_t(kPrivateStackWords32))
    return set_err(LMN_ERR_UNSUPPORTED, "BVH too deep for the FP32 traversal stack");
  HIP_TRY(upload(s->nodes4, h.nodes4));
  HIP_TRY(upload(s->nodesF, h.nodesF));
  HIP_TRY(upload(s->nodesD, h.nodesD));
  HIP_TRY(upload(s->mnodesF, h.mnodesF));
  HIP_TRY(upload(s->mnodesD, h.mnodesD));
  HIP_TRY(upload(s->primsF, h.primsF));
  HIP_TRY(upload(s->primsD, h.primsD));
  HIP_TRY(upload(s->xfF, h.xfF));
  HIP_TRY(upload(s->xfD, h.xfD));
  HIP_TRY(upload(s->matF, h.matF));
  HIP_TRY(upload(s->matD, h.matD));
  HIP_TRY(upload(s->tex, h.tex));
  HIP_TRY(upload(s->texels, h.texels));
  HIP_TRY(upload(s->lightF, h.lightF));
  HIP_TRY(upload(s->lightD, h.lightD));
  HIP_TRY(upload(s->attrF, h.attrF));
  HIP_TRY(upload(s->attrD, h.attrD));
  s->device_bytes = s->nodes4.bytes + s->nodesF.bytes + s->nodesD.bytes + s->mnodesF.bytes + s->mnodesD.bytes + s->primsF.bytes + s->primsD.bytes + s->xfF.bytes +
                    s->xfD.bytes + s->matF.bytes + s->matD.bytes + s->tex.bytes + s->texels.bytes +
                    s->lightF.bytes + s->lightD.bytes + s->attrF.bytes + s->attrD.bytes;
  SceneView<float>& F = s->viewF;
  F.nodes4 = s->nodes4.as<Node4Q>();
  F.n_nodes4 = uint32_t(h.nodes4.size());
  F.nodes = s->nodesF.as<NodeF>();
  F.mnodes = s->mnodesF.as<NodeF>();
  F.n_mnodes = uint32_t(h.mnodesF.size());
  F.prims = s->primsF.as<PrimT<float>>();
  F.prims64 = s->primsD.as<PrimT<double>>();
  F.xforms = s->xfF.as<XformT<float>>();
  F.xforms64 = s->xfD.as<XformT<double>>();
  F.materials = s->matF.as<MaterialT<float>>();
  F.textures = s->tex.as<TextureDev>();
  F.texels = s->texels.as<float>();
  F.lights = s->lightF.as<LightT<float>>();
  F.attrs = s->attrF.as<AttrT<float>>();
  F.n_nodes = uint32_t(h.nodesF.size());
  F.n_bounded = h.n_bounded;
  F.n_planes = h.n_planes;
  F.n_lights = uint32_t(h.lightF.size());
  F.bg = h.bgF;
  F.cam = h.camF;
  F.textured = 0;
  for (const MaterialT<float>& m : h.matF) F.textured |= m.tex >= 0 ? 1u : 0u;
  SceneView<double>& D = s->viewD;
  D.nodes4 = nullptr;
  D.n_nodes4 = 0;
  D.nodes = s->nodesD.as<NodeD>();
  D.mnodes = s->mnodesD.as<NodeD>();
  D.n_mnodes = F.n_mnodes;
  D.prims = s->primsD.as<PrimT<double>>();
  D.prims64 = D.prims;
  D.xforms = s->xfD.as<XformT<double>>();
  D.xforms64 = D.xforms;
  D.materials = s->matD.as<MaterialT<double>>();
  D.textures = F.textures;
  D.texels = F.texels;
  D.textured = F.textured;
  D.lights = s->lightD.as<LightT<double>>();
  D.attrs = s->attrD.as<AttrT<double>>();
  D.n_nodes = F.n_nodes;
  D.n_bounded = F.n_bounded;
  D.n_planes = F.n_planes;
  D.n_lights = F.n_lights;
  D.bg = h.bgD;
  D.cam = h.camD;
  s->camera = desc->camera;
  s->config = desc->config;
  *out = s.release();
  return LMN_OK;
}

void lmn_scene_destroy(lmn_scene* scene) { delete scene; }

int lmn_scene_info(const lmn_scene* s, uint64_t* n_nodes, uint64_t* n_leaves, uint32_t* max_depth,
                   uint64_t* device_bytes) {
  if (!s) return set_err(LMN_ERR_INVALID, "lmn_scene_info: null scene");
  if (n_nodes) *n_nodes = s->host.nodesF.size() + s->host.mnodesF.size();
  if (n_leaves) *n_leaves = s->host.n_leaves;
  if (max_depth) *max_depth = s->host.tree_depth;
  if (device_bytes) *device_bytes = s->device_bytes;
  return LMN_OK;
}

}  // extern "C"

namespace {

struct FrameSetup {
  FrameDev F;
  uint32_t npix;
  bool black;  // zero spp or zero depth: the reference returns black pixels
};

int frame_setup(const lmn_render_params* p, FrameSetup& fs) {
  // src/render.rs:1563-1590
  uint32_t fw = std::max<uint32_t>(p->width, 1), fh = std::max<uint32_t>(p->height, 1);
  uint32_t cx = 0, cy = 0, cw = fw, ch = fh;
  if (p->has_crop) {
    cx = std::min(p->crop[0], fw);
    cy = std::min(p->crop[1], fh);
    cw = std::min(p->crop[2], fw);
    ch = std::min(p->crop[3], fh);
    if (uint64_t(cy) + ch > fh)
      return set_err(LMN_ERR_INVALID, "crop rows beyond the image height are not supported");
  }
  if (cw == 0 || ch == 0) return set_err(LMN_ERR_INVALID, "empty crop region");
  uint32_t sq = uint32_t(std::ceil(std::sqrt(double(p->spp))));
  FrameDev& F = fs.F;
  F.full_w = fw;
  F.full_h = fh;
  F.crop_x = cx;
  F.crop_y = cy;
  F.crop_w = cw;
  F.crop_h = ch;
  F.sqrt_spp = sq;
  F.spp_a = sq * sq;
  F.max_depth = p->max_depth;
  F.filter = p->filter;
  F.fkey = frame_key(p->seed, p->scene_key);
  F.mode = FRAME_STRATIFIED;
  F.pix_list = nullptr;
  F.n_list = 0;
  F.tile_w = 0;
  fs.npix = cw * ch;
  F.n_items = uint64_t(fs.npix) * F.spp_a;
  F.item_base = 0;
  fs.black = F.spp_a == 0 || p->max_depth == 0;
  if (p->max_depth > 0xffff) return set_err(LMN_ERR_INVALID, "max_depth above 65535 is not supported");
  if (p->filter > LMN_FILTER_MITCHELL) return set_err(LMN_ERR_INVALID, "unknown pixel filter");
  return LMN_OK;
}

constexpr uint32_t kMaxTreeDepth = 48;

// ---- wavefront launch configuration and its tuning cache ---------------------
// The knobs: path-pool size (paths in flight per iteration), workgroup size,
// LDS short-stack depth and grid size of the grid-stride kernels.  Explicit
// values in lmn_render_params win; otherwise an entry that lmn_autotune stored
// for the workload's signature; otherwise the defaults.

struct WaveConfig {
  uint32_t pool, block, stack, grid;
};

struct TuneCache {
  std::mutex mu;
  bool loaded = false;
  std::vector<std::pair<std::string, WaveConfig>> entries;
  std::string path() const {
    if (const char* e = std::getenv("LMN_TUNE_CACHE")) return e;
    const char* home = std::getenv("HOME");
    return std::string(home ? home : "/tmp") + "/.cache/luminara_amd/tuning.txt";
  }
  void load() {
    if (loaded) return;
    loaded = true;
    std::ifstream f(path());
    std::string line;
    while (std::getline(f, line)) {
      size_t tab = line.find('\t');
      if (tab == std::string::npos) continue;
      WaveConfig c{};
      std::istringstream v(line.substr(tab + 1));
      if (v >> c.pool >> c.block >> c.stack >> c.grid) entries.push_back({line.substr(0, tab), c});
    }
  }
  bool find(const std::string& key, WaveConfig& c) {
    std::lock_guard<std::mutex> g(mu);
    load();
    for (auto it = entries.rbegin(); it != entries.rend(); ++it)
      if (it->first == key) {
        c = it->second;
        return true;
      }
    return false;
  }
  void store(const std::string& key, const WaveConfig& c) {
    std::lock_guard<std::mutex> g(mu);
    load();
    entries.push_back({key, c});
    std::string pth = path();
    size_t slash = pth.rfind('/');
    if (slash != std::string::npos) {  // mkdir -p of the cache directory
      std::string dir = pth.substr(0, slash);
      for (size_t k = 1; k <= dir.size(); ++k)
        if (k == dir.size() || dir[k] == '/') (void)mkdir(dir.substr(0, k).c_str(), 0755);
    }
    std::ofstream f(pth, std::ios::app);
    if (f) f << key << '\t' << c.pool << ' ' << c.block << ' ' << c.stack << ' ' << c.grid << '\n';
  }
};
TuneCache& tune_cache() {
  static TuneCache* c = new TuneCache();
  return *c;
}

int log2_bucket(uint64_t v) {
  int b = 0;
  while (v > 1) {
    v >>= 1;
    ++b;
  }
  return b;
}

// Signature of a wavefront workload: device, feature set, scene count, and
// power-of-two buckets of the path count, BVH size and depth.
std::string tune_key(lmn_scene* const* scenes, const lmn_render_params* const* params, const FrameSetup* fs,
                     size_t n) {
  hipDeviceProp_t prop{};
  int dev = 0;
  (void)hipGetDevice(&dev);
  (void)hipGetDeviceProperties(&prop, dev);
  uint64_t total = 0, nodes = 0;
  uint32_t feat = 0, depth = 0;
  for (size_t i = 0; i < n; ++i) {
    if (fs[i].black) continue;
    total += fs[i].F.n_items;
    nodes += scenes[i]->host.nodesF.size();
    feat |= scenes[i]->host.featF;
    depth = std::max(depth, params[i]->max_depth);
  }
  std::ostringstream k;
  k << prop.gcnArchName << "/cu" << prop.multiProcessorCount << "|feat" << feat << "|scenes" << log2_bucket(n)
    << "|paths" << log2_bucket(total) << "|nodes" << log2_bucket(nodes) << "|depth" << log2_bucket(depth);
  return k.str();
}

constexpr uint32_t kDefaultGridThreads = 1u << 22;

WaveConfig default_wave_config(uint64_t total) {
  WaveConfig c;
  c.pool = uint32_t(std::min<uint64_t>(std::max<uint64_t>(total, 64), 1u << 23));
  c.block = 256;
  c.stack = 16;
  c.grid = 0;
  return c;
}

int resolve_wave_config(lmn_scene* const* scenes, const lmn_render_params* const* params, const FrameSetup* fs,
                        size_t n, uint64_t total, WaveConfig& c) {
  const lmn_render_params* p0 = params[0];
  c = default_wave_config(total);
  if (!p0->pool_paths || !p0->block_size || !p0->stack_lds || !p0->grid_blocks) {
    WaveConfig t;
    if (tune_cache().find(tune_key(scenes, params, fs, n), t)) c = t;
  }
  if (p0->pool_paths) c.pool = p0->pool_paths;
  if (p0->block_size) c.block = p0->block_size;
  if (p0->stack_lds) c.stack = p0->stack_lds;
  if (p0->grid_blocks) c.grid = p0->grid_blocks;
  c.pool = std::max<uint32_t>(c.pool, 64);
  if (c.block != 128 && c.block != 256) return set_err(LMN_ERR_INVALID, "block_size must be 128 or 256");
  if (c.stack != 16 && c.stack != 32) return set_err(LMN_ERR_INVALID, "stack_lds must be 16 or 32");
  uint32_t need = (c.pool + c.block - 1) / c.block;
  // 16384 workgroups of 256 (2.5 rays per lane at 10 M rays): the hardware
  // balances many short waves better than a persistent grid of long ones
  if (!c.grid) c.grid = std::min<uint32_t>(need, kDefaultGridThreads / c.block);
  c.grid = std::max<uint32_t>(1, std::min(c.grid, need));
  return LMN_OK;
}

// Wavefront over a batch of scenes: every scene's work items are issued from
// one list (scene by scene, sample pass by sample pass), so one set of
// kernel launches per iteration serves the whole batch.  Items are grouped in
// chunks of whole sample passes (at most kMaxChunkItems samples buffered).
// Every chunk ends with the pool draining (its last iterations run few
// paths), so chunks are as large as HBM comfortably holds: 2^30 items, 16 GiB
// of float4 samples (a 64-scene batch of ~1e9 paths is one chunk).
constexpr uint64_t kMaxChunkItems = 1ull << 30;
constexpr uint64_t kPlainChunkItems = 1ull << 26;  // 2 GiB of double4 plain-kernel samples

struct WaveTotals {
  uint64_t ext = 0, sh = 0, cam = 0, iters = 0;
  double t_stage[6] = {0, 0, 0, 0, 0, 0};
  uint32_t n_launch = 0;
};

// Appends the parts of sample passes [s0, s1) of frame F over n_list listed
// pixels (list null: the whole crop), each at most kMaxChunkItems items.
// The whole-frame wavefront issues a frame's pixels in 8x8 tiles when the
// crop's sides are multiples of 8 (see tile_pixel), else in row order.
static uint32_t frame_tile_w(const FrameDev& F) {
  return F.crop_w % 8u == 0u && F.crop_h % 8u == 0u ? F.crop_w : 0u;
}

void add_parts(std::vector<FrameDev>& parts, FrameDev F, const uint32_t* list, uint32_t n_pix, uint32_t s0,
               uint32_t s1) {
  F.pix_list = list;
  F.n_list = list ? n_pix : 0;
  uint32_t step = uint32_t(std::max<uint64_t>(1, std::min<uint64_t>(s1 - s0, kMaxChunkItems / n_pix)));
  for (uint32_t a = s0; a < s1; a += step) {
    uint32_t b = std::min(s1, a + step);
    FrameDev P = F;
    P.item_base = uint64_t(a) * n_pix;
    P.n_items = uint64_t(b - a) * n_pix;
    parts.push_back(P);
  }
}

// Runs the wavefront over `parts`; after each chunk drains, resolve(part,
// samples of the part) folds the part's samples (sample-major over its pixels).
template <class Resolve>
int wavefront_run(lmn_scene* const* scenes, const lmn_render_params* const* params, const FrameSetup* fs, size_t n,
                  const std::vector<FrameDev>& parts, hipStream_t st, WaveTotals& tot, TraceCounters* d_tc,
                  Resolve&& resolve) {
  if (parts.empty()) return LMN_OK;
  Workspace& W = workspace();
  const lmn_render_params* p0 = params[0];
  uint64_t total = 0;
  uint32_t feat = 0;
  bool islands_main = false;
  for (const FrameDev& P : parts) {
    total += P.n_items;
    feat |= scenes[P.scene]->host.featF;
    islands_main = islands_main || scenes[P.scene]->host.islands_main;
  }
  // the main BVHs' features: FEAT_F64 only for FP64 islands there
  const uint32_t feat_main = (feat & FEAT_XF) | (islands_main ? FEAT_F64 : 0u);
  if (feat & FEAT_MARCH) feat |= FEAT_F64;
  WaveConfig wc;
  int rc = resolve_wave_config(scenes, params, fs, n, total, wc);
  if (rc != LMN_OK) return rc;
  const uint32_t pool = wc.pool, ss = wc.stack;
  // chunks: consecutive parts up to kMaxChunkItems items
  std::vector<std::pair<size_t, size_t>> chunks;
  uint64_t chunk_max = 0;
  for (size_t b = 0; b < parts.size();) {
    size_t e = b;
    uint64_t items = 0;
    while (e < parts.size() && (e == b || items + parts[e].n_items <= kMaxChunkItems)) items += parts[e++].n_items;
    chunks.push_back({b, e});
    chunk_max = std::max(chunk_max, items);
    b = e;
  }
  const uint32_t grid = wc.grid;
  const uint32_t nthreads = grid * wc.block;
  for (int k = 0; k < 2; ++k)
    for (int f = 0; f < 4; ++f) HIP_TRY(W.st[k][f].ensure(size_t(pool) * 16));
  HIP_TRY(W.hit.ensure(size_t(pool) * 16));
  HIP_TRY(W.hit_b.ensure(size_t(pool) * 16));
  HIP_TRY(W.hit_rec.ensure(size_t(pool) * 4));
  HIP_TRY(W.sh_o.ensure(size_t(pool) * 16));
  HIP_TRY(W.sh_d.ensure(size_t(pool) * 16));
  HIP_TRY(W.sh_c.ensure(size_t(pool) * 16));
  HIP_TRY(W.samples.ensure(size_t(std::max<uint64_t>(chunk_max, 1)) * 16));
  // traversal stack words beyond the LDS part: the 4-wide walk's worst case,
  // or the binary marched walk's depth cap
  uint32_t stack_words = kMaxTreeDepth;
  for (size_t i = 0; i < n; ++i) stack_words = std::max(stack_words, scenes[i]->host.stack4_words);
  HIP_TRY(W.spill.ensure(size_t(stack_words) * nthreads * 4));
  HIP_TRY(W.ctr.ensure(sizeof(WaveCounters)));
  // marched-pass queues: per kind one candidate slot per pool path
  uint32_t march_kinds = 0;
  for (size_t i = 0; i < n; ++i) march_kinds |= scenes[i]->host.march_kinds;
  if (feat & FEAT_MARCH) {
    HIP_TRY(W.mq_cand.ensure(size_t(pool) * 4 * sizeof(uint2)));
    HIP_TRY(W.mq_key.ensure(size_t(pool) * 8));
    HIP_TRY(W.mq_ovf.ensure(size_t(pool) * 4));
    if (march_kinds & 2u) HIP_TRY(W.mq_brk.ensure(size_t(pool / 2 + 1) * sizeof(MarchBracket)));
  }
  std::vector<SceneView<float>> views(n);
  for (size_t i = 0; i < n; ++i) {
    views[i] = scenes[i]->viewF;
    views[i].exact_marched = params[i]->intersect_mode == LMN_ISECT_EXACT;
  }
  HIP_TRY(W.views.ensure(n * sizeof(SceneView<float>)));
  HIP_TRY(hipMemcpyAsync(W.views.p, views.data(), n * sizeof(SceneView<float>), hipMemcpyHostToDevice, st));
  HIP_TRY(W.parts.ensure(std::max<size_t>(parts.size(), 1) * sizeof(FrameDev)));
  HIP_TRY(W.part_begin.ensure((parts.size() + 1) * sizeof(uint64_t)));
  PathSoA P[2];
  for (int k = 0; k < 2; ++k)
    P[k] = PathSoA{W.st[k][0].as<float4>(), W.st[k][1].as<float4>(), W.st[k][2].as<uint4>(), W.st[k][3].as<uint4>()};
  WaveLaunch L;
  L.stream = st;
  L.grid = grid;
  L.block = wc.block;
  L.pool = pool;
  L.stack_lds = ss;
  L.feat = feat;
  L.feat_main = feat_main;
  L.hit = HitSoA{W.hit.as<float4>(), W.hit_b.as<float4>(), W.hit_rec.as<uint32_t>()};
  L.sh = ShadowSoA{W.sh_o.as<float4>(), W.sh_d.as<float4>(), W.sh_c.as<float4>()};
  L.samples = W.samples.as<float4>();
  L.ctr = W.ctr.as<WaveCounters>();
  L.spill = W.spill.as<uint32_t>();
  L.mq = MarchQ{};
  if (feat & FEAT_MARCH)
    L.mq = MarchQ{W.mq_cand.as<uint2>(),
                  W.mq_key.as<unsigned long long>(),
                  W.mq_ovf.as<uint32_t>(),
                  (march_kinds & 2u) ? W.mq_brk.as<MarchBracket>() : nullptr,
                  pool,
                  (march_kinds & 2u) ? pool / 2 + 1 : 0u,
                  march_kinds,
                  0u};
  L.tc = d_tc;
  const int block = Workspace::kBlock;
  const bool timing = (p0->flags & LMN_FLAG_STAGE_TIMING) != 0;
  // Stage timing records events between the stages; events recorded by graph
  // nodes cannot be timed, so timed renders launch directly.
  bool use_graph = !(p0->flags & LMN_FLAG_NO_GRAPH) && !timing && !W.graph_failed[0];
  // parts of a chunk: uploaded in one copy, chunk after chunk
  std::vector<uint64_t> begins;
  for (auto [pb, pe] : chunks) {
    begins.assign(1, 0);
    for (size_t q = pb; q < pe; ++q) begins.push_back(begins.back() + parts[q].n_items);
    HIP_TRY(hipMemcpyAsync(W.parts.p, parts.data() + pb, (pe - pb) * sizeof(FrameDev), hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(W.part_begin.p, begins.data(), begins.size() * sizeof(uint64_t), hipMemcpyHostToDevice, st));
    BatchDev B{W.views.as<SceneView<float>>(), W.parts.as<FrameDev>(), W.part_begin.as<uint64_t>(), uint32_t(pe - pb), 0};
    launch_counters_init(L.ctr, begins.back(), st);
    // The block of iterations is captured once per chunk as a hipGraph and
    // replayed until the pool drains: every kernel reads its work counts from
    // device memory, so the same graph serves every block (an even number of
    // iterations brings the double-buffered pools back to the same roles).
    hipGraphExec_t gexec = nullptr;
    if (use_graph) {
      hipGraph_t graph = nullptr;
      bool ok = hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed) == hipSuccess;
      if (ok) {
        for (int k = 0; k < block; ++k)
          launch_wave_iteration(L, B, P[k & 1], P[(k & 1) ^ 1], timing ? W.stage_ev[k] : nullptr);
        ok = hipStreamEndCapture(st, &graph) == hipSuccess && graph;
      }
      if (ok) ok = hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0) == hipSuccess;
      if (graph) (void)hipGraphDestroy(graph);
      if (!ok) {  // capture refused (e.g. event nodes): launch directly from now on
        (void)hipGetLastError();
        if (gexec) (void)hipGraphExecDestroy(gexec);
        gexec = nullptr;
        use_graph = false;
        W.graph_failed[timing ? 1 : 0] = true;
      }
    }
    int cur = 0;
    auto t_start = std::chrono::steady_clock::now();
    for (;;) {
      if (gexec) {
        hipError_t e = hipGraphLaunch(gexec, st);
        if (e != hipSuccess) {
          (void)hipGraphExecDestroy(gexec);
          return set_err(LMN_ERR_DEVICE, std::string("hipGraphLaunch: ") + hipGetErrorString(e));
        }
      } else {
        for (int k = 0; k < block; ++k) {
          launch_wave_iteration(L, B, P[cur], P[cur ^ 1], timing ? W.stage_ev[k] : nullptr);
          cur ^= 1;
        }
      }
      HIP_TRY(hipGetLastError());
      HIP_TRY(hipMemcpyAsync(W.host_ctr, L.ctr, sizeof(WaveCounters), hipMemcpyDeviceToHost, st));
      HIP_TRY(hipStreamSynchronize(st));
      if (timing) {
        for (int k = 0; k < block; ++k)
          for (int g = 0; g < 6; ++g) {
            float ms = 0.f;
            HIP_TRY(hipEventElapsedTime(&ms, W.stage_ev[k][g], W.stage_ev[k][g + 1]));
            tot.t_stage[g] += ms;
          }
        tot.n_launch += block;
      }
      const WaveCounters& c = *W.host_ctr;
      if (c.active == 0 && c.next_item == c.n_items) {
        tot.ext += c.stat_extend;
        tot.sh += c.stat_shadow;
        tot.cam += c.stat_camera;
        tot.iters += c.iterations;
        break;
      }
      if (std::chrono::steady_clock::now() - t_start > std::chrono::seconds(600)) {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        return set_err(LMN_ERR_DEVICE, "wavefront did not drain within 600 s");
      }
    }
    if (gexec) (void)hipGraphExecDestroy(gexec);
    for (size_t q = pb; q < pe; ++q) resolve(parts[q], L.samples + begins[q - pb]);
  }
  HIP_TRY(hipGetLastError());
  return LMN_OK;
}

// Pixel lists and per-pixel state of the adaptive and time-limited modes.
struct FrameModeBufs {
  DevBuf ident;  // 0 .. npix-1 (row bands are slices of it)
  DevBuf list[2], count, state, taken;
};

// Adaptive rounds over the listed pixels of every scene i with lists[i]
// non-empty: all of a round's pixels have taken the same number of samples,
// so a round is one sample range per scene; the pixels still sampling after
// it are compacted into the other list.
int adaptive_rounds(lmn_scene* const* scenes, const lmn_render_params* const* params, const FrameSetup* fs,
                    size_t n, std::vector<FrameModeBufs>& mb, std::vector<uint32_t>& n_list, hipStream_t st,
                    WaveTotals& tot, TraceCounters* d_tc, uint32_t& rounds) {
  std::vector<uint32_t> taken(n, 0), cur(n, 0);
  for (;;) {
    std::vector<FrameDev> parts;
    std::vector<uint32_t> s_end(n, 0);
    for (size_t i = 0; i < n; ++i) {
      if (fs[i].black || n_list[i] == 0 || taken[i] >= fs[i].F.spp_a) continue;
      const FrameDev& F = fs[i].F;
      uint32_t first = std::min(F.spp_a, adaptive_min_samples(F.spp_a));
      uint32_t s1 = taken[i] == 0 ? first : std::min(F.spp_a, taken[i] + std::max<uint32_t>(F.sqrt_spp, 4));
      FrameDev G = F;
      G.scene = uint32_t(i);
      add_parts(parts, G, mb[i].list[cur[i]].as<uint32_t>(), n_list[i], taken[i], s1);
      s_end[i] = s1;
    }
    if (parts.empty()) break;
    ++rounds;
    int rc = wavefront_run(scenes, params, fs, n, parts, st, tot, d_tc, [&](const FrameDev& P, const float4* smp) {
      const lmn_render_params* p = params[P.scene];
      launch_resolve_adaptive(smp, mb[P.scene].state.as<AdaptState>(), P.pix_list, P.n_list,
                              uint32_t(P.n_items / P.n_list), adaptive_min_samples(P.spp_a), p->adaptive_threshold,
                              st);
    });
    if (rc != LMN_OK) return rc;
    for (size_t i = 0; i < n; ++i) {
      if (s_end[i] == 0) continue;
      taken[i] = s_end[i];
      HIP_TRY(hipMemsetAsync(mb[i].count.p, 0, 4, st));
      launch_adapt_compact(mb[i].state.as<AdaptState>(), mb[i].list[cur[i]].as<uint32_t>(), n_list[i],
                           fs[i].F.spp_a, mb[i].list[cur[i] ^ 1].as<uint32_t>(), mb[i].count.as<uint32_t>(), st);
      HIP_TRY(hipMemcpyAsync(&n_list[i], mb[i].count.p, 4, hipMemcpyDeviceToHost, st));
      cur[i] ^= 1;
    }
    HIP_TRY(hipStreamSynchronize(st));
    // the compacted lists now sit in list[cur]; keep them in list[0]
    for (size_t i = 0; i < n; ++i)
      if (cur[i] && n_list[i]) {
        HIP_TRY(hipMemcpyAsync(mb[i].list[0].p, mb[i].list[1].p, size_t(n_list[i]) * 4, hipMemcpyDeviceToDevice, st));
        cur[i] = 0;
      } else {
        cur[i] = 0;
      }
  }
  return LMN_OK;
}

int render_wavefront(lmn_scene* const* scenes, const lmn_render_params* const* params, const FrameSetup* fs,
                     float* const* d_out, size_t n, hipStream_t st, lmn_stats* stats, TraceCounters* d_tc,
                     std::chrono::steady_clock::time_point t0) {
  Workspace& W = workspace();
  WaveTotals tot;
  bool adaptive = false, timed = false;
  for (size_t i = 0; i < n; ++i) {
    adaptive = adaptive || (params[i]->flags & LMN_FLAG_ADAPTIVE);
    timed = timed || params[i]->time_limit > 0.0;
  }
  if (timed && n > 1) return set_err(LMN_ERR_UNSUPPORTED, "render: a time limit applies to one frame, not a batch");
  std::vector<size_t> acc_off(n + 1, 0);
  for (size_t i = 0; i < n; ++i) acc_off[i + 1] = acc_off[i] + (fs[i].black ? 0 : fs[i].npix);
  uint64_t taken_total = 0;
  uint32_t rounds = 0, expired_rows = 0;
  if (!adaptive && !timed) {
    std::vector<FrameDev> parts;
    for (size_t i = 0; i < n; ++i) {
      if (fs[i].black) continue;
      FrameDev F = fs[i].F;
      F.scene = uint32_t(i);
      F.tile_w = frame_tile_w(F);
      add_parts(parts, F, nullptr, fs[i].npix, 0, F.spp_a);
    }
    HIP_TRY(W.acc.ensure(std::max<size_t>(acc_off[n], 1) * 32));
    for (size_t i = 0; i < n; ++i)
      if (!fs[i].black) launch_zero_acc(W.acc.as<double4>() + acc_off[i], fs[i].npix, st);
    int rc = wavefront_run(scenes, params, fs, n, parts, st, tot, d_tc, [&](const FrameDev& P, const float4* smp) {
      launch_resolve(smp, W.acc.as<double4>() + acc_off[P.scene], fs[P.scene].npix,
                     uint32_t(P.n_items / fs[P.scene].npix), P.tile_w, st);
    });
    if (rc != LMN_OK) return rc;
    for (size_t i = 0; i < n; ++i)
      if (!fs[i].black) {
        launch_finalize(W.acc.as<double4>() + acc_off[i], d_out[i], fs[i].npix, st);
        taken_total += uint64_t(fs[i].npix) * fs[i].F.spp_a;
      }
    rounds = 1;
  } else {
    // adaptive and / or time-limited: pixel lists per scene
    std::vector<FrameModeBufs> mb(n);
    std::vector<uint32_t> n_list(n, 0);
    HIP_TRY(W.acc.ensure(std::max<size_t>(acc_off[n], 1) * 32));
    for (size_t i = 0; i < n; ++i) {
      if (fs[i].black) continue;
      const uint32_t npix = fs[i].npix;
      HIP_TRY(mb[i].ident.ensure(size_t(npix) * 4));
      HIP_TRY(mb[i].list[0].ensure(size_t(npix) * 4));
      HIP_TRY(mb[i].list[1].ensure(size_t(npix) * 4));
      HIP_TRY(mb[i].count.ensure(8));
      HIP_TRY(mb[i].state.ensure(size_t(npix) * sizeof(AdaptState)));
      HIP_TRY(mb[i].taken.ensure(8));
      launch_adapt_init(mb[i].state.as<AdaptState>(), mb[i].ident.as<uint32_t>(), npix, st);
      launch_zero_acc(W.acc.as<double4>() + acc_off[i], npix, st);
    }
    if (!timed) {
      for (size_t i = 0; i < n; ++i)
        if (!fs[i].black) {
          HIP_TRY(hipMemcpyAsync(mb[i].list[0].p, mb[i].ident.p, size_t(fs[i].npix) * 4, hipMemcpyDeviceToDevice, st));
          n_list[i] = fs[i].npix;
        }
      int rc = adaptive_rounds(scenes, params, fs, n, mb, n_list, st, tot, d_tc, rounds);
      if (rc != LMN_OK) return rc;
    } else if (!fs[0].black) {
      // Time limit (src/render.rs:1600-1625): rows in bands; once the limit
      // has passed at the start of a band, that band and every later one take
      // one sample through each pixel centre.
      const FrameSetup& f0 = fs[0];
      const uint32_t cw = f0.F.crop_w, ch = f0.F.crop_h;
      const uint32_t n_bands = std::min<uint32_t>(ch, 16);
      const double limit = params[0]->time_limit;
      for (uint32_t b = 0; b < n_bands; ++b) {
        const uint32_t r0 = uint32_t(uint64_t(ch) * b / n_bands), r1 = uint32_t(uint64_t(ch) * (b + 1) / n_bands);
        const double elapsed = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        const uint32_t* band = mb[0].ident.as<uint32_t>() + size_t(r0) * cw;
        if (elapsed > limit) {
          const uint32_t rows = ch - r0, cnt = rows * cw;
          FrameDev F = f0.F;
          F.scene = 0;
          F.mode = FRAME_CENTER;
          std::vector<FrameDev> parts;
          add_parts(parts, F, band, cnt, 0, 1);
          int rc = wavefront_run(scenes, params, fs, n, parts, st, tot, d_tc, [&](const FrameDev& P, const float4* smp) {
            if (adaptive)
              launch_resolve_adaptive(smp, mb[0].state.as<AdaptState>(), P.pix_list, P.n_list, 1, 0xffffffffu, 0.0,
                                      st);
            else
              launch_resolve_list(smp, W.acc.as<double4>(), P.pix_list, P.n_list, 1, st);
          });
          if (rc != LMN_OK) return rc;
          expired_rows = rows;
          ++rounds;
          break;
        }
        const uint32_t cnt = (r1 - r0) * cw;
        if (adaptive) {
          std::vector<uint32_t> nl(1, cnt);
          HIP_TRY(hipMemcpyAsync(mb[0].list[0].p, band, size_t(cnt) * 4, hipMemcpyDeviceToDevice, st));
          int rc = adaptive_rounds(scenes, params, fs, n, mb, nl, st, tot, d_tc, rounds);
          if (rc != LMN_OK) return rc;
        } else {
          FrameDev F = f0.F;
          F.scene = 0;
          std::vector<FrameDev> parts;
          add_parts(parts, F, band, cnt, 0, F.spp_a);
          int rc = wavefront_run(scenes, params, fs, n, parts, st, tot, d_tc, [&](const FrameDev& P, const float4* smp) {
            launch_resolve_list(smp, W.acc.as<double4>(), P.pix_list, P.n_list, uint32_t(P.n_items / P.n_list), st);
          });
          if (rc != LMN_OK) return rc;
          ++rounds;
        }
      }
    }
    for (size_t i = 0; i < n; ++i) {
      if (fs[i].black) continue;
      if (adaptive) {
        HIP_TRY(hipMemsetAsync(mb[i].taken.p, 0, 8, st));
        launch_adapt_finalize(mb[i].state.as<AdaptState>(), d_out[i], fs[i].npix,
                              mb[i].taken.as<unsigned long long>(), st);
        unsigned long long t = 0;
        HIP_TRY(hipMemcpyAsync(&t, mb[i].taken.p, 8, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        taken_total += t;
      } else {
        launch_finalize(W.acc.as<double4>() + acc_off[i], d_out[i], fs[i].npix, st);
        taken_total += uint64_t(fs[i].npix - uint64_t(expired_rows) * fs[i].F.crop_w) * fs[i].F.spp_a +
                       uint64_t(expired_rows) * fs[i].F.crop_w;
      }
    }
    HIP_TRY(hipStreamSynchronize(st));
  }
  HIP_TRY(hipGetLastError());
  if (stats) {
    stats->camera_rays = tot.cam;
    stats->extension_rays = tot.ext - tot.cam;
    stats->shadow_rays = tot.sh;
    stats->iterations = uint32_t(tot.iters);
    stats->extend_launches = tot.n_launch;
    stats->extend_ms = tot.t_stage[0];
    stats->march_ms = tot.t_stage[1] + tot.t_stage[4];
    stats->shade_ms = tot.t_stage[2];
    stats->shadow_ms = tot.t_stage[3];
    stats->regen_ms = tot.t_stage[5];
    stats->paths = taken_total;
    stats->samples_traced = tot.cam;
    stats->rounds = rounds;
    stats->expired_rows = expired_rows;
  }
  return LMN_OK;
}

int render_plain(lmn_scene* s, const lmn_render_params* p, const FrameSetup& fs, float* d_out, hipStream_t st,
                 lmn_stats* stats, bool fp64) {
  Workspace& W = workspace();
  uint32_t npix = fs.npix;
  if (p->time_limit > 0.0)
    return set_err(LMN_ERR_UNSUPPORTED, "render: a time limit is supported on the wavefront path only");
  HIP_TRY(W.rays.ensure(16));
  HIP_TRY(hipMemsetAsync(W.rays.p, 0, 16, st));
  if (p->flags & LMN_FLAG_ADAPTIVE) {
    // one sample pass per launch; pixels that have stopped skip the rest
    DevBuf state, list, taken;
    HIP_TRY(state.ensure(size_t(npix) * sizeof(AdaptState)));
    HIP_TRY(list.ensure(size_t(npix) * 4));
    HIP_TRY(taken.ensure(8));
    HIP_TRY(hipMemsetAsync(taken.p, 0, 8, st));
    launch_adapt_init(state.as<AdaptState>(), list.as<uint32_t>(), npix, st);
    const uint32_t mins = adaptive_min_samples(fs.F.spp_a);
    for (uint32_t s0 = 0; s0 < fs.F.spp_a; ++s0) {
      if (fp64) {
        SceneView<double> S = s->viewD;
        S.exact_marched = p->intersect_mode == LMN_ISECT_EXACT;
        launch_plain_adaptive_fp64(S, s->host.featD, fs.F, state.as<AdaptState>(), s0, s0 + 1, mins,
                                   p->adaptive_threshold, W.rays.as<unsigned long long>(), st);
      } else {
        SceneView<float> S = s->viewF;
        S.exact_marched = p->intersect_mode == LMN_ISECT_EXACT;
        launch_plain_adaptive_fp32(S, s->host.featF, fs.F, state.as<AdaptState>(), s0, s0 + 1, mins,
                                   p->adaptive_threshold, W.rays.as<unsigned long long>(), st);
      }
    }
    launch_adapt_finalize(state.as<AdaptState>(), d_out, npix, taken.as<unsigned long long>(), st);
    HIP_TRY(hipGetLastError());
    unsigned long long t = 0, r[2];
    HIP_TRY(hipMemcpyAsync(&t, taken.p, 8, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(r, W.rays.p, 16, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    if (stats) {
      stats->paths = t;
      stats->samples_traced = t;
      stats->camera_rays = t;
      stats->extension_rays = r[0] - t;
      stats->shadow_rays = r[1];
      stats->rounds = fs.F.spp_a;
    }
    return LMN_OK;
  }
  // one path per thread over whole sample passes, at most kPlainChunkItems
  // paths per launch; each chunk's samples are added in sample order
  HIP_TRY(W.acc.ensure(size_t(npix) * 32));
  launch_zero_acc(W.acc.as<double4>(), npix, st);
  const uint32_t passes = uint32_t(std::max<uint64_t>(1, std::min<uint64_t>(fs.F.spp_a, kPlainChunkItems / npix)));
  HIP_TRY(W.samples.ensure(size_t(passes) * npix * sizeof(double4)));
  for (uint32_t s0 = 0; s0 < fs.F.spp_a; s0 += passes) {
    const uint32_t s1 = std::min(fs.F.spp_a, s0 + passes);
    const uint64_t k0 = uint64_t(s0) * npix, nk = uint64_t(s1 - s0) * npix;
    if (fp64) {
      SceneView<double> S = s->viewD;
      S.exact_marched = p->intersect_mode == LMN_ISECT_EXACT;
      launch_plain_fp64(S, s->host.featD, fs.F, W.samples.as<double4>(), k0, nk, W.rays.as<unsigned long long>(),
                        (p->flags >> LMN_FLAG_PLAIN_WAVES_SHIFT) & 7u, st);
    } else {
      SceneView<float> S = s->viewF;
      S.exact_marched = p->intersect_mode == LMN_ISECT_EXACT;
      launch_plain_fp32(S, s->host.featF, fs.F, W.samples.as<double4>(), k0, nk, W.rays.as<unsigned long long>(), st);
    }
    launch_plain_resolve(W.samples.as<double4>(), W.acc.as<double4>(), npix, s1 - s0, st);
  }
  launch_finalize(W.acc.as<double4>(), d_out, npix, st);
  HIP_TRY(hipGetLastError());
  if (stats) {
    unsigned long long r[2];
    HIP_TRY(hipMemcpyAsync(r, W.rays.p, 16, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    stats->camera_rays = uint64_t(npix) * fs.F.spp_a;
    stats->extension_rays = r[0] - stats->camera_rays;
    stats->shadow_rays = r[1];
    stats->iterations = 0;
    stats->paths = stats->camera_rays;
    stats->samples_traced = stats->camera_rays;
    stats->rounds = fs.F.spp_a;
  }
  return LMN_OK;
}

// Renders n scenes; the wavefront path serves them in one batch, the plain
// kernels one after another.  d_out[i]: device framebuffer of scene i.
int render_common(lmn_scene* const* scenes, const lmn_render_params* const* params, float* const* d_out, size_t n,
                  lmn_stats* stats, TraceCounters* d_tc) {
  if (!scenes || !params || !d_out || n == 0) return set_err(LMN_ERR_INVALID, "render: null argument");
  const auto t0 = std::chrono::steady_clock::now();  // the time limit counts from here
  std::vector<FrameSetup> fs(n);
  for (size_t i = 0; i < n; ++i) {
    if (!scenes[i] || !params[i] || !d_out[i]) return set_err(LMN_ERR_INVALID, "render: null argument");
    int rc = frame_setup(params[i], fs[i]);
    if (rc != LMN_OK) return rc;
    if (params[i]->path != params[0]->path) return set_err(LMN_ERR_INVALID, "render: one path per batch");
  }
  Workspace& W = workspace();
  std::lock_guard<std::mutex> lock(W.mu);
  int rc = W.init();
  if (rc != LMN_OK) return rc;
  const lmn_render_params* p0 = params[0];
  hipStream_t st = p0->stream ? static_cast<hipStream_t>(p0->stream) : W.stream;
  if (stats) std::memset(stats, 0, sizeof *stats);
  HIP_TRY(hipEventRecord(W.ev0, st));
  for (size_t i = 0; i < n; ++i)
    if (fs[i].black) HIP_TRY(hipMemsetAsync(d_out[i], 0, size_t(fs[i].npix) * 12, st));
  if (p0->path == LMN_PATH_WAVEFRONT_FP32) {
    rc = render_wavefront(scenes, params, fs.data(), d_out, n, st, stats, d_tc, t0);
  } else if (p0->path == LMN_PATH_PLAIN_FP64 || p0->path == LMN_PATH_PLAIN_FP32) {
    for (size_t i = 0; i < n && rc == LMN_OK; ++i) {
      if (fs[i].black) continue;
      lmn_stats one{};
      rc = render_plain(scenes[i], params[i], fs[i], d_out[i], st, stats ? &one : nullptr,
                        p0->path == LMN_PATH_PLAIN_FP64);
      if (stats) {
        stats->camera_rays += one.camera_rays;
        stats->extension_rays += one.extension_rays;
        stats->shadow_rays += one.shadow_rays;
        stats->paths += one.paths;
        stats->samples_traced += one.samples_traced;
        stats->rounds += one.rounds;
      }
    }
  } else {
    return set_err(LMN_ERR_INVALID, "unknown render path");
  }
  if (rc != LMN_OK) return rc;
  HIP_TRY(hipEventRecord(W.ev1, st));
  HIP_TRY(hipEventSynchronize(W.ev1));
  float ms = 0.f;
  HIP_TRY(hipEventElapsedTime(&ms, W.ev0, W.ev1));
  if (stats) {
    for (size_t i = 0; i < n; ++i)  // black frames: the loop still counts its samples
      if (fs[i].black) stats->paths += uint64_t(fs[i].npix) * fs[i].F.spp_a;
    stats->render_ms = ms;
  }
  return LMN_OK;
}

}  // namespace

extern "C" {

int lmn_render_device(const lmn_scene* scene, const lmn_render_params* params, float* d_hdr_rgb, lmn_stats* stats) {
  lmn_scene* sc = const_cast<lmn_scene*>(scene);
  return render_common(&sc, &params, &d_hdr_rgb, 1, stats, nullptr);
}

int lmn_render(const lmn_scene* scene, const lmn_render_params* params, float* hdr_rgb, lmn_stats* stats) {
  const lmn_scene* const sc[1] = {scene};
  float* const out[1] = {hdr_rgb};
  return lmn_render_batch(sc, 1, params, out, stats);
}

int lmn_render_batch_device(const lmn_scene* const* scenes, size_t n_scenes, const lmn_render_params* params,
                            float* const* d_hdr_rgb, lmn_stats* stats) {
  if (!scenes || !params || !d_hdr_rgb || !n_scenes) return set_err(LMN_ERR_INVALID, "lmn_render_batch: bad argument");
  std::vector<lmn_scene*> sc(n_scenes);
  std::vector<const lmn_render_params*> pp(n_scenes);
  for (size_t i = 0; i < n_scenes; ++i) {
    sc[i] = const_cast<lmn_scene*>(scenes[i]);
    pp[i] = &params[i];
  }
  return render_common(sc.data(), pp.data(), d_hdr_rgb, n_scenes, stats, nullptr);
}

int lmn_render_batch(const lmn_scene* const* scenes, size_t n_scenes, const lmn_render_params* params,
                     float* const* hdr_rgb, lmn_stats* stats) {
  if (!scenes || !params || !hdr_rgb || !n_scenes) return set_err(LMN_ERR_INVALID, "lmn_render_batch: bad argument");
  std::vector<size_t> bytes(n_scenes);
  std::vector<float*> d(n_scenes, nullptr);
  int rc = LMN_OK;
  for (size_t i = 0; i < n_scenes && rc == LMN_OK; ++i) {
    FrameSetup fs;
    rc = frame_setup(&params[i], fs);
    if (rc != LMN_OK) break;
    bytes[i] = size_t(fs.npix) * 12;
    if (hipMalloc((void**)&d[i], bytes[i]) != hipSuccess) rc = set_err(LMN_ERR_NOMEM, "hipMalloc of a framebuffer failed");
  }
  if (rc == LMN_OK) rc = lmn_render_batch_device(scenes, n_scenes, params, d.data(), stats);
  for (size_t i = 0; i < n_scenes; ++i) {
    if (rc == LMN_OK && d[i]) {
      hipError_t e = hipMemcpy(hdr_rgb[i], d[i], bytes[i], hipMemcpyDeviceToHost);
      if (e != hipSuccess) rc = set_err(LMN_ERR_DEVICE, std::string("hipMemcpy: ") + hipGetErrorString(e));
    }
    if (d[i]) (void)hipFree(d[i]);
  }
  return rc;
}

int lmn_trace_rays(const lmn_scene* scene, const lmn_ray* rays, size_t n, uint32_t path, uint32_t intersect_mode,
                   lmn_hit* hits) {
  if (!scene || (!rays && n) || (!hits && n)) return set_err(LMN_ERR_INVALID, "lmn_trace_rays: null argument");
  if (n == 0) return LMN_OK;
  lmn_ray* d_r = nullptr;
  lmn_hit* d_h = nullptr;
  HIP_TRY(hipMalloc((void**)&d_r, n * sizeof(lmn_ray)));
  hipError_t e = hipMalloc((void**)&d_h, n * sizeof(lmn_hit));
  if (e != hipSuccess) {
    (void)hipFree(d_r);
    return set_err(LMN_ERR_DEVICE, std::string("hipMalloc: ") + hipGetErrorString(e));
  }
  int rc = LMN_OK;
  e = hipMemcpy(d_r, rays, n * sizeof(lmn_ray), hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    if (path == LMN_PATH_PLAIN_FP64) {
      SceneView<double> S = scene->viewD;
      S.exact_marched = intersect_mode == LMN_ISECT_EXACT;
      launch_trace_fp64(S, scene->host.featD, d_r, d_h, n, nullptr);
    } else {
      SceneView<float> S = scene->viewF;
      S.exact_marched = intersect_mode == LMN_ISECT_EXACT;
      launch_trace_fp32(S, scene->host.featF, d_r, d_h, n, nullptr);
    }
    e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpy(hits, d_h, n * sizeof(lmn_hit), hipMemcpyDeviceToHost);
  }
  if (e != hipSuccess) rc = set_err(LMN_ERR_DEVICE, std::string("trace: ") + hipGetErrorString(e));
  (void)hipFree(d_r);
  (void)hipFree(d_h);
  return rc;
}

int lmn_device_info(char* name, size_t name_len, int* n_cu, double* hbm_bytes) {
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  HIP_TRY(hipGetDeviceProperties(&prop, dev));
  if (name && name_len) {
    std::snprintf(name, name_len, "%s (%s)", prop.name, prop.gcnArchName);
  }
  if (n_cu) *n_cu = prop.multiProcessorCount;
  if (hbm_bytes) *hbm_bytes = double(prop.totalGlobalMem);
  return LMN_OK;
}

// Instrumented wavefront render: counts BVH node fetches and primitive tests
// of the extend stage (the algorithmic bytes of the roofline).
int lmn_profile_extend(const lmn_scene* const* scenes, size_t n_scenes, const lmn_render_params* params,
                       float* const* d_hdr_rgb, lmn_trace_counts* counts) {
  if (!counts || !scenes || !params || !d_hdr_rgb || !n_scenes)
    return set_err(LMN_ERR_INVALID, "lmn_profile_extend: bad argument");
  std::vector<lmn_render_params> p(params, params + n_scenes);
  std::vector<lmn_scene*> sc(n_scenes);
  std::vector<const lmn_render_params*> pp(n_scenes);
  for (size_t i = 0; i < n_scenes; ++i) {
    p[i].path = LMN_PATH_WAVEFRONT_FP32;
    p[i].flags = 0;
    sc[i] = const_cast<lmn_scene*>(scenes[i]);
    pp[i] = &p[i];
  }
  TraceCounters* d_tc = nullptr;
  HIP_TRY(hipMalloc((void**)&d_tc, sizeof(TraceCounters)));
  hipError_t e = hipMemset(d_tc, 0, sizeof(TraceCounters));
  int rc = e == hipSuccess ? render_common(sc.data(), pp.data(), d_hdr_rgb, n_scenes, nullptr, d_tc)
                           : set_err(LMN_ERR_DEVICE, "hipMemset failed");
  if (rc == LMN_OK) {
    TraceCounters h;
    e = hipMemcpy(&h, d_tc, sizeof h, hipMemcpyDeviceToHost);
    if (e != hipSuccess) rc = set_err(LMN_ERR_DEVICE, "hipMemcpy failed");
    counts->node_fetches = h.node_fetches;
    counts->prim_tests = h.prim_tests;
    counts->tri_tests = h.tri_tests;
    counts->rays = h.rays;
    counts->hits = h.hits;
    counts->node_bytes = sizeof(Node4Q);
    counts->prim_bytes = sizeof(PrimT<float>);
  }
  (void)hipFree(d_tc);
  return rc;
}

// ---- autotuner ------------------------------------------------------------------

int lmn_autotune(const lmn_scene* const* scenes, size_t n_scenes, const lmn_render_params* params,
                 float* const* d_hdr_rgb, lmn_tuning* out) {
  if (!scenes || !params || !n_scenes || !out) return set_err(LMN_ERR_INVALID, "lmn_autotune: bad argument");
  // without caller buffers the probes render into buffers of their own
  std::vector<DevBuf> own;
  std::vector<float*> own_ptr;
  if (!d_hdr_rgb) {
    own.resize(n_scenes);
    own_ptr.resize(n_scenes);
    for (size_t i = 0; i < n_scenes; ++i) {
      FrameSetup f;
      int rc = frame_setup(&params[i], f);
      if (rc != LMN_OK) return rc;
      HIP_TRY(own[i].ensure(size_t(f.npix) * 12));
      own_ptr[i] = own[i].as<float>();
    }
    d_hdr_rgb = own_ptr.data();
  }
  std::vector<lmn_scene*> sc(n_scenes);
  std::vector<FrameSetup> fs(n_scenes);
  std::vector<const lmn_render_params*> orig(n_scenes);
  uint64_t total = 0;
  for (size_t i = 0; i < n_scenes; ++i) {
    if (!scenes[i]) return set_err(LMN_ERR_INVALID, "lmn_autotune: null scene");
    if (params[i].path != LMN_PATH_WAVEFRONT_FP32)
      return set_err(LMN_ERR_INVALID, "lmn_autotune: tunes the wavefront path only");
    sc[i] = const_cast<lmn_scene*>(scenes[i]);
    orig[i] = &params[i];
    int rc = frame_setup(&params[i], fs[i]);
    if (rc != LMN_OK) return rc;
    if (!fs[i].black) total += fs[i].F.n_items;
  }
  if (total == 0) return set_err(LMN_ERR_INVALID, "lmn_autotune: nothing to render");
  const std::string key = tune_key(sc.data(), orig.data(), fs.data(), n_scenes);
  // probe: the same frames at reduced spp, about 2^29 paths in all (eight
  // fills of the largest pool, so that the drain at the end of a render
  // weighs about as little as in the full one; with 2^28 the probes
  // favoured smaller pools than the full render does)
  const double scale = std::min(1.0, double(1ull << 29) / double(total));
  std::vector<lmn_render_params> pp(params, params + n_scenes);
  uint64_t probe_paths = 0;
  for (size_t i = 0; i < n_scenes; ++i) {
    pp[i].flags &= ~LMN_FLAG_STAGE_TIMING;
    if (fs[i].black) continue;
    uint32_t sq = fs[i].F.sqrt_spp;
    uint32_t k = std::max<uint32_t>(1, uint32_t(std::floor(double(sq) * std::sqrt(scale))));
    pp[i].spp = k * k;
    probe_paths += uint64_t(fs[i].npix) * k * k;
  }
  std::vector<const lmn_render_params*> ppp(n_scenes);
  for (size_t i = 0; i < n_scenes; ++i) ppp[i] = &pp[i];
  uint32_t n_timed = 0;
  auto timed = [&](const WaveConfig& c, double& ms) -> int {
    for (auto& q : pp) {
      q.pool_paths = c.pool;
      q.block_size = c.block;
      q.stack_lds = c.stack;
      q.grid_blocks = c.grid;
    }
    ms = 1e30;
    for (int rep = 0; rep < 3; ++rep) {  // the first run warms caches and allocations
      lmn_stats st{};
      int rc = render_common(sc.data(), ppp.data(), d_hdr_rgb, n_scenes, &st, nullptr);
      if (rc != LMN_OK) return rc;
      if (rep) ms = std::min(ms, st.render_ms);
    }
    ++n_timed;
    return LMN_OK;
  };
  auto with_grid = [](WaveConfig c, double f) {
    uint32_t need = (c.pool + c.block - 1) / c.block;
    uint32_t g = std::min<uint32_t>(need, kDefaultGridThreads / c.block);
    c.grid = std::max<uint32_t>(1, std::min<uint32_t>(need, uint32_t(double(g) * f)));
    return c;
  };
  WaveConfig best = with_grid(default_wave_config(total), 1.0);
  double best_ms = 0.0;
  int rc = timed(best, best_ms);
  if (rc != LMN_OK) return rc;
  const double default_ms = best_ms;
  auto consider = [&](const WaveConfig& c) -> int {
    double ms;
    int r = timed(c, ms);
    if (r != LMN_OK) return r;
    if (ms < best_ms * 0.99) {  // keep the incumbent unless clearly faster
      best = c;
      best_ms = ms;
    }
    return LMN_OK;
  };
  // pool size and workgroup size together (they interact: the grid follows
  // both), then the grid size up to one thread per pool path, then the LDS
  // stack depth, then the grid again around the winner
  const WaveConfig start = best;
  for (uint32_t pool : {1u << 21, 1u << 22, 1u << 24, 1u << 25, 1u << 26, 1u << 27})
    for (uint32_t block : {128u, 256u}) {
      WaveConfig c = start;
      c.pool = uint32_t(std::min<uint64_t>(pool, std::max<uint64_t>(total, 64)));
      c.block = block;
      c = with_grid(c, 1.0);
      if (c.pool == best.pool && c.block == best.block && c.grid == best.grid) continue;
      if ((rc = consider(c)) != LMN_OK) return rc;
    }
  auto grid_pass = [&](std::initializer_list<double> fs) -> int {
    const WaveConfig base = best;
    for (double f : fs) {
      WaveConfig c = with_grid(base, f);
      if (c.grid == best.grid) continue;
      int r = consider(c);
      if (r != LMN_OK) return r;
    }
    return LMN_OK;
  };
  if ((rc = grid_pass({0.25, 0.5, 2.0, 4.0, 8.0, 16.0})) != LMN_OK) return rc;
  {
    WaveConfig c = best;
    c.stack = best.stack == 16 ? 32 : 16;
    if ((rc = consider(c)) != LMN_OK) return rc;
  }
  {
    // the grid again, at half and twice the winner's
    const WaveConfig base = best;
    for (uint32_t g : {base.grid / 2, base.grid * 2}) {
      WaveConfig c = base;
      c.grid = std::max<uint32_t>(1, std::min<uint32_t>(g, (c.pool + c.block - 1) / c.block));
      if (c.grid == best.grid) continue;
      if ((rc = consider(c)) != LMN_OK) return rc;
    }
  }
  tune_cache().store(key, best);
  out->pool_paths = best.pool;
  out->block_size = best.block;
  out->stack_lds = best.stack;
  out->grid_blocks = best.grid;
  out->probe_ms = best_ms;
  out->default_ms = default_ms;
  out->probe_paths = probe_paths;
  out->candidates = n_timed;
  out->pad = 0;
  return LMN_OK;
}

// ---- path-replay adjoint ------------------------------------------------------

namespace {

int adjoint_common(const lmn_scene* scene, const lmn_render_params* p, const lmn_adjoint_spec* spec,
                   const float* d_dldi, float* d_tex_out, float* rough_out, bool rough_on_device, lmn_grads* g) {
  if (!scene || !p || !spec || !d_dldi || !g) return set_err(LMN_ERR_INVALID, "lmn_adjoint: null argument");
  const HostScene& H = scene->host;
  FrameSetup fs;
  int rc = frame_setup(p, fs);
  if (rc != LMN_OK) return rc;
  bool fp64 = p->path == LMN_PATH_PLAIN_FP64;
  if (!fp64 && p->path != LMN_PATH_WAVEFRONT_FP32 && p->path != LMN_PATH_PLAIN_FP32)
    return set_err(LMN_ERR_INVALID, "lmn_adjoint: unknown path");
  if (p->max_depth > LMN_ADJOINT_MAX_DEPTH)
    return set_err(LMN_ERR_UNSUPPORTED, "lmn_adjoint: max_depth above LMN_ADJOINT_MAX_DEPTH (50)");
  uint32_t n_mat = uint32_t(H.matF.size());
  AdjDev A{};
  A.texture = spec->texture;
  A.emission_material = spec->emission_material;
  A.n_slots = uint32_t(H.slot_mat.size());
  if (A.n_slots > 256) return set_err(LMN_ERR_UNSUPPORTED, "lmn_adjoint: more than 256 microfacet materials");
  size_t n_texels = 0;
  if (spec->texture >= 0) {
    if (size_t(spec->texture) >= H.tex.size()) return set_err(LMN_ERR_INVALID, "lmn_adjoint: texture out of range");
    const TextureDev& T = H.tex[spec->texture];
    A.texel_base = T.offset / 3;
    n_texels = size_t(T.w) * T.h;
    for (uint32_t i = 0; i < n_mat; ++i) {
      const MaterialT<float>& m = H.matF[i];
      if (m.tex != spec->texture) continue;
      bool ok = m.kind == M_LAMBERTIAN || (m.kind == M_MICROFACET && m.metallic == 0.0f);
      if (!ok)
        return set_err(LMN_ERR_UNSUPPORTED,
                       "lmn_adjoint: the parameter texture may only be the albedo of Lambertian or metallic = 0 "
                       "microfacet materials");
    }
  }
  double inv_intensity = 0.0;
  if (spec->emission_material >= 0) {
    if (uint32_t(spec->emission_material) >= n_mat)
      return set_err(LMN_ERR_INVALID, "lmn_adjoint: emission material out of range");
    const MaterialT<double>& m = H.matD[spec->emission_material];
    if (m.kind != M_EMISSIVE || !(m.intensity > 0.0))
      return set_err(LMN_ERR_INVALID, "lmn_adjoint: emission material must be emissive with intensity > 0");
    inv_intensity = 1.0 / m.intensity;
  }
  Workspace& W = workspace();
  std::lock_guard<std::mutex> lock(W.mu);
  rc = W.init();
  if (rc != LMN_OK) return rc;
  hipStream_t st = p->stream ? static_cast<hipStream_t>(p->stream) : W.stream;
  uint32_t npix = fs.npix;
  const uint32_t n_blocks = 2048;
  uint32_t width = A.n_slots + 1;
  HIP_TRY(W.adj_wsum.ensure(size_t(npix) * 8));
  HIP_TRY(W.adj_gtex.ensure(std::max<size_t>(n_texels * 3 * (fp64 ? 8 : 4), 16)));
  HIP_TRY(W.adj_part.ensure(size_t(n_blocks) * width * 8));
  HIP_TRY(W.adj_red.ensure(size_t(width) * 8));
  // the FP32 replay's traversal stack beyond its LDS words, per thread of the
  // grid: the 4-wide walk's worst case, or the binary marched walk's depth cap
  const uint32_t spill_words = std::max(kMaxTreeDepth, H.stack4_words);
  if (!fp64) HIP_TRY(W.adj_spill.ensure(size_t(spill_words) * n_blocks * 256 * 4));
  HIP_TRY(hipEventRecord(W.ev0, st));
  launch_adj_zero_f64(W.adj_part.as<double>(), size_t(n_blocks) * width, st);
  if (fp64)
    launch_adj_zero_f64(W.adj_gtex.as<double>(), n_texels * 3, st);
  else
    launch_adj_zero_f32(W.adj_gtex.as<float>(), n_texels * 3, st);
  uint64_t n_items = fs.black ? 0 : uint64_t(npix) * fs.F.spp_a;
  if (n_items) {
    if (fp64)
      launch_adj_wsum_fp64(scene->viewD.cam, fs.F, W.adj_wsum.as<double>(), st);
    else
      launch_adj_wsum_fp32(scene->viewF.cam, fs.F, W.adj_wsum.as<double>(), st);
    // launches of at most 2^22 paths keep each kernel short
    const uint64_t chunk = uint64_t(1) << 22;
    for (uint64_t k0 = 0; k0 < n_items; k0 += chunk) {
      uint64_t k1 = std::min(n_items, k0 + chunk);
      if (fp64) {
        SceneView<double> S = scene->viewD;
        S.exact_marched = p->intersect_mode == LMN_ISECT_EXACT;
        launch_adjoint_fp64(S, H.featD, fs.F, A, d_dldi, W.adj_wsum.as<double>(), W.adj_gtex.as<double>(),
                            W.adj_part.as<double>(), n_blocks, k0, k1, W.adj_spill.as<uint32_t>(), st);
      } else {
        SceneView<float> S = scene->viewF;
        S.exact_marched = p->intersect_mode == LMN_ISECT_EXACT;
        launch_adjoint_fp32(S, H.featF, fs.F, A, d_dldi, W.adj_wsum.as<double>(), W.adj_gtex.as<float>(),
                            W.adj_part.as<double>(), n_blocks, k0, k1, W.adj_spill.as<uint32_t>(), st);
      }
    }
  }
  launch_adj_reduce(W.adj_part.as<double>(), n_blocks, width, W.adj_red.as<double>(), st);
  if (d_tex_out && n_texels) {
    if (fp64)
      launch_adj_to_float(W.adj_gtex.as<double>(), d_tex_out, n_texels * 3, st);
    else
      HIP_TRY(hipMemcpyAsync(d_tex_out, W.adj_gtex.p, n_texels * 3 * 4, hipMemcpyDeviceToDevice, st));
  }
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipEventRecord(W.ev1, st));
  std::vector<double> red(width);
  HIP_TRY(hipMemcpyAsync(red.data(), W.adj_red.p, width * 8, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  float ms = 0.f;
  HIP_TRY(hipEventElapsedTime(&ms, W.ev0, W.ev1));
  if (rough_out) {
    std::vector<float> r(n_mat, 0.0f);
    for (uint32_t j = 0; j < A.n_slots; ++j) r[H.slot_mat[j]] = float(red[j]);
    if (rough_on_device)
      HIP_TRY(hipMemcpy(rough_out, r.data(), n_mat * 4, hipMemcpyHostToDevice));
    else
      std::memcpy(rough_out, r.data(), n_mat * 4);
  }
  g->emission = red[A.n_slots] * inv_intensity;
  g->paths = n_items;
  g->adjoint_ms = ms;
  return LMN_OK;
}

}  // namespace

int lmn_adjoint_device(const lmn_scene* scene, const lmn_render_params* params, const lmn_adjoint_spec* spec,
                       const float* d_dl_dimage, lmn_grads* grads) {
  if (!grads) return set_err(LMN_ERR_INVALID, "lmn_adjoint: null argument");
  return adjoint_common(scene, params, spec, d_dl_dimage, grads->texels, grads->roughness, true, grads);
}

int lmn_adjoint(const lmn_scene* scene, const lmn_render_params* params, const lmn_adjoint_spec* spec,
                const float* dl_dimage, lmn_grads* grads) {
  if (!scene || !params || !spec || !dl_dimage || !grads) return set_err(LMN_ERR_INVALID, "lmn_adjoint: null argument");
  FrameSetup fs;
  int rc = frame_setup(params, fs);
  if (rc != LMN_OK) return rc;
  size_t img = size_t(fs.npix) * 3 * 4;
  size_t n_tex = 0;
  if (spec->texture >= 0 && size_t(spec->texture) < scene->host.tex.size())
    n_tex = size_t(scene->host.tex[spec->texture].w) * scene->host.tex[spec->texture].h * 3;
  float* d_in = nullptr;
  float* d_tex = nullptr;
  HIP_TRY(hipMalloc(&d_in, std::max<size_t>(img, 16)));
  hipError_t e = hipMemcpy(d_in, dl_dimage, img, hipMemcpyHostToDevice);
  if (e == hipSuccess && grads->texels && n_tex) e = hipMalloc(&d_tex, n_tex * 4);
  if (e != hipSuccess) {
    (void)hipFree(d_in);
    return set_err(LMN_ERR_DEVICE, std::string("lmn_adjoint: ") + hipGetErrorString(e));
  }
  rc = adjoint_common(scene, params, spec, d_in, d_tex, grads->roughness, false, grads);
  if (rc == LMN_OK && d_tex) {
    e = hipMemcpy(grads->texels, d_tex, n_tex * 4, hipMemcpyDeviceToHost);
    if (e != hipSuccess) rc = set_err(LMN_ERR_DEVICE, std::string("lmn_adjoint: ") + hipGetErrorString(e));
  }
  (void)hipFree(d_in);
  if (d_tex) (void)hipFree(d_tex);
  return rc;
}


// ---- image outputs: AOV ray, post chain, RGBE -------------------------------
namespace {

int pick_stream(void* s, hipStream_t& out) {
  if (s) {
    out = static_cast<hipStream_t>(s);
    return LMN_OK;
  }
  Workspace& W = workspace();
  std::lock_guard<std::mutex> g(W.mu);
  int rc = W.init();
  if (rc != LMN_OK) return rc;
  out = W.stream;
  return LMN_OK;
}

int aov_enqueue(const lmn_scene* scene, const lmn_render_params* params, const FrameSetup& fs, float* d_depth,
                uint8_t* d_normal, uint8_t* d_albedo, hipStream_t st) {
  if (params->intersect_mode > LMN_ISECT_REFERENCE) return set_err(LMN_ERR_INVALID, "lmn_aov: unknown intersect mode");
  if (params->path == LMN_PATH_PLAIN_FP64) {
    SceneView<double> S = scene->viewD;
    S.exact_marched = params->intersect_mode == LMN_ISECT_EXACT;
    launch_aov_fp64(S, scene->host.featD, fs.F, d_depth, d_normal, d_albedo, st);
  } else {
    SceneView<float> S = scene->viewF;
    S.exact_marched = params->intersect_mode == LMN_ISECT_EXACT;
    launch_aov_fp32(S, scene->host.featF, fs.F, d_depth, d_normal, d_albedo, st);
  }
  HIP_TRY(hipGetLastError());
  return LMN_OK;
}

// Enqueues the post chain; with auto-exposure, d_scratch (2 + kLogLumBlocks
// doubles) holds the log-average luminance and the applied exposure when the
// stream reaches the end.  Allocates nothing.
int post_enqueue_into(const float* d_hdr, uint64_t n, const lmn_post_params* post, const float* d_fog,
                      uint8_t* d_rgba, float* d_hdr_exposed, double* d_scratch, hipStream_t st) {
  if (!post) return set_err(LMN_ERR_INVALID, "lmn_post: null post parameters");
  if (post->tone_map > LMN_TONE_NONE) return set_err(LMN_ERR_INVALID, "lmn_post: unknown tone map");
  const bool fog = post->depth_fog > 0.0;
  if (fog && !d_fog) return set_err(LMN_ERR_INVALID, "lmn_post: depth fog needs the depth output of lmn_aov");
  PostArgs A{};
  A.hdr = d_hdr;
  A.fog = fog ? d_fog : nullptr;
  A.exposure = post->exposure;
  A.fog_density = post->depth_fog;
  for (int k = 0; k < 3; ++k) A.fog_rgb[k] = post->fog_color[k];
  A.tone = post->tone_map;
  A.n = n;
  A.rgba = d_rgba;
  A.hdr_exposed = d_hdr_exposed;
  if (post->auto_exposure) {
    if (!d_scratch) return set_err(LMN_ERR_INVALID, "lmn_post: automatic exposure without scratch");
    launch_log_luminance(d_hdr, n, post->exposure, d_scratch, st);
    A.exposure_dev = d_scratch;
  }
  if (d_rgba || d_hdr_exposed) {
    HIP_TRY(ensure_post_tables());
    launch_post(A, st);
  }
  HIP_TRY(hipGetLastError());
  return LMN_OK;
}

// The same with the scratch in a library-allocated buffer.
int post_enqueue(const float* d_hdr, uint64_t n, const lmn_post_params* post, const float* d_fog, uint8_t* d_rgba,
                 float* d_hdr_exposed, DevBuf& scratch, hipStream_t st) {
  if (!post) return set_err(LMN_ERR_INVALID, "lmn_post: null post parameters");
  if (post->auto_exposure) HIP_TRY(scratch.ensure((2 + kLogLumBlocks) * sizeof(double)));
  return post_enqueue_into(d_hdr, n, post, d_fog, d_rgba, d_hdr_exposed, scratch.as<double>(), st);
}

int read_exposure_from(const lmn_post_params* post, const double* d_scratch, hipStream_t st, double& used,
                       double* log_average = nullptr) {
  used = post->exposure;
  if (log_average) *log_average = 0.0;
  if (post->auto_exposure) {
    double two[2];
    HIP_TRY(hipMemcpyAsync(two, d_scratch, sizeof two, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    used = two[1];
    if (log_average) *log_average = two[0];
  }
  return LMN_OK;
}

int read_exposure(const lmn_post_params* post, const DevBuf& scratch, hipStream_t st, double& used,
                  double* log_average = nullptr) {
  return read_exposure_from(post, scratch.as<double>(), st, used, log_average);
}

int to_host(void* host, const DevBuf& d, size_t bytes) {
  if (host && bytes) HIP_TRY(hipMemcpy(host, d.p, bytes, hipMemcpyDeviceToHost));
  return LMN_OK;
}

}  // namespace

int lmn_aov_device(const lmn_scene* scene, const lmn_render_params* params, float* d_depth, uint8_t* d_normal,
                   uint8_t* d_albedo) {
  if (!scene || !params) return set_err(LMN_ERR_INVALID, "lmn_aov: null argument");
  FrameSetup fs;
  int rc = frame_setup(params, fs);
  hipStream_t st = nullptr;
  if (rc == LMN_OK) rc = pick_stream(params->stream, st);
  if (rc == LMN_OK) rc = aov_enqueue(scene, params, fs, d_depth, d_normal, d_albedo, st);
  if (rc != LMN_OK) return rc;
  if (!params->stream) HIP_TRY(hipStreamSynchronize(st));
  return LMN_OK;
}

int lmn_aov(const lmn_scene* scene, const lmn_render_params* params, float* depth, uint8_t* normal,
            uint8_t* albedo) {
  if (!scene || !params) return set_err(LMN_ERR_INVALID, "lmn_aov: null argument");
  FrameSetup fs;
  int rc = frame_setup(params, fs);
  if (rc != LMN_OK) return rc;
  const size_t n = fs.npix;
  DevBuf dd, dn, da;
  if (depth) HIP_TRY(dd.ensure(n * 4));
  if (normal) HIP_TRY(dn.ensure(n * 3));
  if (albedo) HIP_TRY(da.ensure(n * 3));
  rc = lmn_aov_device(scene, params, dd.as<float>(), dn.as<uint8_t>(), da.as<uint8_t>());
  if (rc == LMN_OK) rc = to_host(depth, dd, n * 4);
  if (rc == LMN_OK) rc = to_host(normal, dn, n * 3);
  if (rc == LMN_OK) rc = to_host(albedo, da, n * 3);
  return rc;
}

int lmn_post_device(const float* d_hdr, uint32_t width, uint32_t height, const lmn_post_params* post,
                    const float* d_fog_depth, uint8_t* d_rgba, float* d_hdr_exposed, double* exposure_used,
                    void* stream) {
  if (!d_hdr || !post || width == 0 || height == 0) return set_err(LMN_ERR_INVALID, "lmn_post: bad argument");
  hipStream_t st = nullptr;
  int rc = pick_stream(stream, st);
  DevBuf scratch;
  if (rc == LMN_OK) rc = post_enqueue(d_hdr, uint64_t(width) * height, post, d_fog_depth, d_rgba, d_hdr_exposed, scratch, st);
  double used = 0.0;
  if (rc == LMN_OK) rc = read_exposure(post, scratch, st, used);  // waits for the stream with auto-exposure
  if (rc != LMN_OK) return rc;
  if (!stream) HIP_TRY(hipStreamSynchronize(st));
  if (exposure_used) *exposure_used = used;
  return LMN_OK;
}

int lmn_post_device_noalloc(const float* d_hdr, uint32_t width, uint32_t height, const lmn_post_params* post,
                            const float* d_fog_depth, uint8_t* d_rgba, float* d_hdr_exposed, double* exposure_used,
                            void* stream) {
  if (!d_hdr || !post || width == 0 || height == 0) return set_err(LMN_ERR_INVALID, "lmn_post: bad argument");
  // the default stream when none is given: the library's own stream would be
  // created (and its pinned counters allocated) on first use
  hipStream_t st = static_cast<hipStream_t>(stream);
  double* d_scratch = nullptr;
  size_t cap = 0;
  if (post->auto_exposure) {
    HIP_TRY(noalloc_post_scratch(&d_scratch, &cap));
    if (cap < 2 + size_t(kLogLumBlocks)) return set_err(LMN_ERR_INVALID, "lmn_post_device_noalloc: scratch too small");
  }
  int rc = post_enqueue_into(d_hdr, uint64_t(width) * height, post, d_fog_depth, d_rgba, d_hdr_exposed, d_scratch, st);
  double used = 0.0;
  if (rc == LMN_OK) rc = read_exposure_from(post, d_scratch, st, used);  // waits for the stream with auto-exposure
  if (rc != LMN_OK) return rc;
  if (!stream) HIP_TRY(hipStreamSynchronize(st));
  if (exposure_used) *exposure_used = used;
  return LMN_OK;
}

int lmn_post(const float* hdr, uint32_t width, uint32_t height, const lmn_post_params* post, const float* fog_depth,
             uint8_t* rgba, float* hdr_exposed, double* exposure_used) {
  if (!hdr || !post || width == 0 || height == 0) return set_err(LMN_ERR_INVALID, "lmn_post: bad argument");
  const size_t n = size_t(width) * height;
  const bool fog = post->depth_fog > 0.0;
  if (fog && !fog_depth) return set_err(LMN_ERR_INVALID, "lmn_post: depth fog needs the depth output of lmn_aov");
  DevBuf dh, df, dr, dx;
  HIP_TRY(dh.ensure(n * 12));
  HIP_TRY(hipMemcpy(dh.p, hdr, n * 12, hipMemcpyHostToDevice));
  if (fog) {
    HIP_TRY(df.ensure(n * 4));
    HIP_TRY(hipMemcpy(df.p, fog_depth, n * 4, hipMemcpyHostToDevice));
  }
  if (rgba) HIP_TRY(dr.ensure(n * 4));
  if (hdr_exposed) HIP_TRY(dx.ensure(n * 12));
  int rc = lmn_post_device(dh.as<float>(), width, height, post, df.as<float>(), dr.as<uint8_t>(), dx.as<float>(),
                           exposure_used, nullptr);
  if (rc == LMN_OK) rc = to_host(rgba, dr, n * 4);
  if (rc == LMN_OK) rc = to_host(hdr_exposed, dx, n * 12);
  return rc;
}

int lmn_encode_rgbe_device(const float* d_rgb, uint64_t n_pixels, uint8_t* d_rgbe, void* stream) {
  if ((!d_rgb || !d_rgbe) && n_pixels) return set_err(LMN_ERR_INVALID, "lmn_encode_rgbe: null argument");
  hipStream_t st = nullptr;
  int rc = pick_stream(stream, st);
  if (rc != LMN_OK) return rc;
  launch_rgbe(d_rgb, n_pixels, d_rgbe, st);
  HIP_TRY(hipGetLastError());
  if (!stream) HIP_TRY(hipStreamSynchronize(st));
  return LMN_OK;
}

int lmn_encode_rgbe(const float* rgb, uint64_t n_pixels, uint8_t* rgbe) {
  if ((!rgb || !rgbe) && n_pixels) return set_err(LMN_ERR_INVALID, "lmn_encode_rgbe: null argument");
  if (n_pixels == 0) return LMN_OK;
  DevBuf di, dout;
  HIP_TRY(di.ensure(n_pixels * 12));
  HIP_TRY(dout.ensure(n_pixels * 4));
  HIP_TRY(hipMemcpy(di.p, rgb, n_pixels * 12, hipMemcpyHostToDevice));
  int rc = lmn_encode_rgbe_device(di.as<float>(), n_pixels, dout.as<uint8_t>(), nullptr);
  if (rc == LMN_OK) rc = to_host(rgbe, dout, n_pixels * 4);
  return rc;
}

int lmn_render_image(const lmn_scene* scene, const lmn_render_params* params, const lmn_post_params* post,
                     lmn_image* img, lmn_stats* stats) {
  if (!scene || !params || !post || !img) return set_err(LMN_ERR_INVALID, "lmn_render_image: null argument");
  FrameSetup fs;
  int rc = frame_setup(params, fs);
  if (rc != LMN_OK) return rc;
  const size_t n = fs.npix;
  hipStream_t st = nullptr;
  rc = pick_stream(params->stream, st);
  if (rc != LMN_OK) return rc;
  DevBuf hdr, dep, nrm, alb, rgba, hx, rgbe, scratch;
  HIP_TRY(hdr.ensure(n * 12));
  rc = lmn_render_device(scene, params, hdr.as<float>(), stats);
  if (rc != LMN_OK) return rc;
  const bool fog = post->depth_fog > 0.0;
  if (fog || img->depth) HIP_TRY(dep.ensure(n * 4));
  if (img->normal) HIP_TRY(nrm.ensure(n * 3));
  if (img->albedo) HIP_TRY(alb.ensure(n * 3));
  if (img->rgba) HIP_TRY(rgba.ensure(n * 4));
  if (img->hdr || img->rgbe) HIP_TRY(hx.ensure(n * 12));
  if (img->rgbe) HIP_TRY(rgbe.ensure(n * 4));
  hipEvent_t e0 = nullptr, e1 = nullptr;
  HIP_TRY(hipEventCreate(&e0));
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return set_err(LMN_ERR_DEVICE, "lmn_render_image: hipEventCreate failed");
  }
  (void)hipEventRecord(e0, st);
  if (dep.p || nrm.p || alb.p) rc = aov_enqueue(scene, params, fs, dep.as<float>(), nrm.as<uint8_t>(), alb.as<uint8_t>(), st);
  if (rc == LMN_OK) rc = post_enqueue(hdr.as<float>(), n, post, dep.as<float>(), rgba.as<uint8_t>(), hx.as<float>(), scratch, st);
  if (rc == LMN_OK && img->rgbe) launch_rgbe(hx.as<float>(), n, rgbe.as<uint8_t>(), st);
  if (rc == LMN_OK && img->depth) launch_depth_pass(dep.as<float>(), n, st);
  hipError_t e = hipSuccess;
  if (rc == LMN_OK) e = hipGetLastError();
  if (rc == LMN_OK && e == hipSuccess) e = hipEventRecord(e1, st);
  if (rc == LMN_OK && e == hipSuccess) e = hipEventSynchronize(e1);
  float ms = 0.f;
  if (rc == LMN_OK && e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc != LMN_OK) return rc;
  if (e != hipSuccess) return set_err(LMN_ERR_DEVICE, std::string("lmn_render_image: ") + hipGetErrorString(e));
  img->post_ms = ms;
  rc = read_exposure(post, scratch, st, img->exposure, &img->log_average);
  if (rc == LMN_OK) rc = to_host(img->rgba, rgba, n * 4);
  if (rc == LMN_OK) rc = to_host(img->hdr, hx, n * 12);
  if (rc == LMN_OK) rc = to_host(img->rgbe, rgbe, n * 4);
  if (rc == LMN_OK) rc = to_host(img->depth, dep, n * 4);
  if (rc == LMN_OK) rc = to_host(img->normal, nrm, n * 3);
  if (rc == LMN_OK) rc = to_host(img->albedo, alb, n * 3);
  return rc;
}

}  // extern "C"

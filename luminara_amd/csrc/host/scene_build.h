// Host-side flattening of an lmn_scene_desc into the device layout (see
// lmn_geom.h): SAH BVH over the bounded primitives (BvhNode::build,
// src/bvh.rs:16-163, re-designed for a flat GPU tree), leaf-ordered primitive
// records in FP32 and FP64, composed instance transforms, materials, lights,
// textures, camera and background.
#pragma once
#include <string>
#include <vector>

#include "../../../include/luminara.h"
#include "../lmn_render.h"

namespace lmn {

// 32 bins and leaves of at most 2 primitives measured best for the FP32
// wavefront on the 64-scene batch (16 bins and 4 per leaf: 299.3 ms per
// step; 32 and 4: 294.5; 64 and 4: 297.3; 32 and 2: 293.0).
struct BvhOptions {
  int bins = 32;        // SAH bins per axis
  int max_leaf = 2;     // primitives per leaf
  int max_depth = 48;   // hard depth cap (traversal stacks are sized for it)
};

struct HostScene {
  std::vector<Node4Q> nodes4;     // main BVH collapsed to 4-wide, quantised boxes (FP32 path)
  uint32_t stack4_words = 0;      // worst-case traversal stack of nodes4, 32-bit words
  std::vector<NodeF> nodesF;   // main BVH
  std::vector<NodeD> nodesD;
  std::vector<NodeF> mnodesF;  // BVH over the marched kinds
  std::vector<NodeD> mnodesD;
  std::vector<PrimT<float>> primsF;   // n_bounded in leaf order, then planes
  std::vector<PrimT<double>> primsD;
  std::vector<XformT<float>> xfF;
  std::vector<XformT<double>> xfD;
  std::vector<MaterialT<float>> matF;
  std::vector<MaterialT<double>> matD;
  std::vector<TextureDev> tex;
  std::vector<float> texels;
  std::vector<LightT<float>> lightF;
  std::vector<LightT<double>> lightD;
  std::vector<AttrT<float>> attrF;   // vertex attributes of smooth mesh triangles
  std::vector<AttrT<double>> attrD;
  BackgroundT<float> bgF;
  BackgroundT<double> bgD;
  CameraT<float> camF;
  CameraT<double> camD;
  uint32_t n_bounded = 0, n_planes = 0;
  uint32_t featF = 0, featD = 0;  // FEAT_* of the FP32 and FP64 paths
  bool islands_main = false;       // FP64 records in the main BVH (not only in the marched kinds' BVH)
  uint32_t march_kinds = 0;        // bit k: marched kind K_TORUS + k present
  uint32_t tree_depth = 0, n_leaves = 0;
  std::vector<uint32_t> order;  // order[i] = descriptor prim index of record i
  std::vector<uint32_t> slot_mat;  // slot_mat[j] = material of roughness slot j (microfacet materials)
};

// Returns LMN_OK or an error code with a message.
int build_host_scene(const lmn_scene_desc& desc, const BvhOptions& opt, HostScene& out, std::string& err);

// The FP32 binary BVH collapsed to 4-wide nodes (returns the worst-case
// traversal stack in 32-bit words), and those nodes' boxes quantised into
// the device layout.
uint32_t collapse_bvh4(const std::vector<NodeF>& n2, std::vector<Node4F>& n4);
void quantize_bvh4(const std::vector<Node4F>& in, std::vector<Node4Q>& out);

// World-space bounds of a descriptor primitive (transform applied).
void prim_world_bounds(const lmn_scene_desc& desc, const lmn_prim& p, double lo[3], double hi[3]);

// Compose a descriptor transform into the device representation.
void compose_transform(const lmn_transform& t, XformT<double>& x, double fwd[12]);

}  // namespace lmn

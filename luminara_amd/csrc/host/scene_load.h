// Scene loading: TOML text -> lmn_scene_desc, following scene::load_scene
// (src/scene.rs:846-1645) for the primitive and material subset the render
// hot path covers.
#pragma once
#include <string>
#include <vector>

#include "../../../include/luminara.h"

namespace lmn {

// Owns every array an lmn_scene_desc points at.  `desc` is the first member
// so that a `lmn_scene_desc*` handed out by the C ABI converts back.
struct OwnedDesc {
  lmn_scene_desc desc{};
  std::vector<lmn_prim> prims;
  std::vector<lmn_transform> transforms;
  std::vector<lmn_material> materials;
  std::vector<std::vector<float>> texels;
  std::vector<lmn_texture> textures;
  std::vector<lmn_light> lights;
  std::vector<lmn_vertex_attr> vertex_attrs;
  uint32_t n_objects = 0;
  void sync();  // refresh desc pointers/counts after the vectors change
};

// Returns LMN_OK or an LMN_ERR_* code with a message in err.
int load_scene_toml(const std::string& text, const std::string& base_dir, OwnedDesc& out, std::string& err);

// ASCII PLY mesh (src/obj.rs:157-246): fan-triangulated faces, vertices
// scaled then offset.  Appends triangles to tri_verts (9 doubles each).
int load_ply_text(const std::string& content, double scale, const double offset[3],
                  std::vector<double>& tri_verts, std::string& err);

// Appends a parsed mesh's triangles to tri_verts (9 doubles each), vertices
// scaled then offset as load_ply does.
struct MeshData;
void emit_mesh(const MeshData& mesh, double scale, const double offset[3], std::vector<double>& tri_verts);

// OBJ mesh (src/obj.rs:18-153): v / vt / vn / f lines, fan-triangulated
// faces, vertices scaled then offset, normals made unit.  Appends 9 vertex
// coordinates per triangle to tri_verts and, per triangle, its corner
// normals and uvs (tri_attrs; has_normals / has_uvs 0 where the face lacks
// them or an index is out of range).
int load_obj_text(const std::string& content, double scale, const double offset[3],
                  std::vector<double>& tri_verts, std::vector<lmn_vertex_attr>& tri_attrs, std::string& err);

}  // namespace lmn

// Mesh ingestion for `[[mesh]]` PLY files (src/scene.rs:1545-1559 reads the
// file whole and hands it to obj::load_ply, src/obj.rs:157-246), plus a binary
// sidecar that holds the parsed mesh, so that a large mesh is parsed from
// text once and read back as two arrays on every later load.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace lmn {

// A parsed PLY mesh, before the `[[mesh]]` scale and offset: the vertex
// coordinates exactly as the text gave them (3 doubles each) and the
// fan-triangulated faces (3 vertex indices each, all < n_vertices).
struct MeshData {
  std::vector<double> verts;
  std::vector<uint32_t> tris;
  uint64_t n_faces = 0;  // PLY face records (a polygon gives several triangles)
  size_t n_vertices() const { return verts.size() / 3; }
  size_t n_triangles() const { return tris.size() / 3; }
};

// ASCII PLY text -> MeshData with the reference loader's rules and messages
// (src/obj.rs:157-246): "ply" magic on the first line, element counts from
// the header up to "end_header", one vertex per line (x y z first), one face
// per line (count then indices), fan triangulation, every index checked.
// `data` must be NUL-terminated at data[len].  Returns LMN_OK or
// LMN_ERR_PARSE with the message in err.
int parse_ply(const char* data, size_t len, MeshData& out, std::string& err);

// Sidecar file of a mesh: `<mesh file>.lmnmesh` beside it.
//
// Layout (little-endian): a 64-byte header
//   char     magic[8]   "LMNMESH1"
//   uint32   version    (parser rules the contents follow)
//   uint32   reserved
//   uint64   n_vertices, n_triangles, n_faces
//   uint64   source_size          size of the mesh file it was made from
//   int64    source_mtime_ns      its modification time
//   uint64   source_digest        FNV-1a 64 of its first and last 64 KiB
// then n_vertices * 3 doubles and n_triangles * 3 uint32 indices.
//
// A sidecar is used only when its header matches the mesh file as it is now
// (size, modification time, digest), its length is exactly what the counts
// imply, and every index is below n_vertices; anything else is a miss and
// the text is parsed instead.
std::string sidecar_path(const std::string& mesh_path);
bool read_mesh_sidecar(const std::string& mesh_path, MeshData& out);
bool write_mesh_sidecar(const std::string& mesh_path, const MeshData& mesh);

// Sidecar policy, from LMN_MESH_SIDECAR: "0"/"off" never, "read" reads but
// never writes, anything else (default) reads and writes one after parsing
// the text of a PLY file.
enum SidecarMode { SIDECAR_OFF = 0, SIDECAR_READ = 1, SIDECAR_AUTO = 2 };
SidecarMode sidecar_mode();

// Loads a PLY mesh file: the sidecar when it is current, else the text
// (writing a sidecar after a successful parse when the mode allows; failing to
// write one is not an error).  `from_sidecar` tells which.  Returns LMN_OK,
// LMN_ERR_IO ("Failed to read mesh '<name>'") or LMN_ERR_PARSE.
int load_ply_file(const std::string& path, const std::string& name, MeshData& out, bool& from_sidecar,
                  std::string& err);

}  // namespace lmn

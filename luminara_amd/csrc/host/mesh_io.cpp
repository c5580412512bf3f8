// PLY mesh ingestion and its binary sidecar.  The text rules follow
// obj::load_ply (src/obj.rs:157-246); the parser works on the file's bytes in
// place (no per-line strings), which is what makes the 1.31M-triangle text
// load take a fraction of a second rather than seconds.
#include "mesh_io.h"

#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>

#include <atomic>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "../../../include/luminara.h"

namespace lmn {

namespace {

constexpr char kMagic[8] = {'L', 'M', 'N', 'M', 'E', 'S', 'H', '1'};
// Bumped whenever the parse rules change what a file turns into, so that
// sidecars written under older rules are parsed again.
constexpr uint32_t kVersion = 1;
constexpr size_t kDigestSpan = 64 * 1024;

struct SidecarHeader {
  char magic[8];
  uint32_t version;
  uint32_t reserved;
  uint64_t n_vertices, n_triangles, n_faces;
  uint64_t source_size;
  int64_t source_mtime_ns;
  uint64_t source_digest;
};
static_assert(sizeof(SidecarHeader) == 64, "sidecar header is 64 bytes");

uint64_t fnv1a(uint64_t h, const unsigned char* p, size_t n) {
  for (size_t i = 0; i < n; ++i) {
    h ^= p[i];
    h *= 0x100000001b3ull;
  }
  return h;
}

// Digest of a file's first and last 64 KiB (the whole file when smaller)
// together with its size.
uint64_t digest_bytes(const unsigned char* head, size_t head_n, const unsigned char* tail, size_t tail_n,
                      uint64_t size) {
  uint64_t h = 0xcbf29ce484222325ull;
  h = fnv1a(h, reinterpret_cast<const unsigned char*>(&size), sizeof(size));
  h = fnv1a(h, head, head_n);
  return fnv1a(h, tail, tail_n);
}

uint64_t digest_of(const char* data, uint64_t size) {
  const unsigned char* p = reinterpret_cast<const unsigned char*>(data);
  size_t head = size < kDigestSpan ? size_t(size) : kDigestSpan;
  size_t tail = size < kDigestSpan ? 0 : kDigestSpan;
  return digest_bytes(p, head, p + size - tail, tail, size);
}

bool digest_of_fd(int fd, uint64_t size, uint64_t& out) {
  std::vector<unsigned char> head(size < kDigestSpan ? size_t(size) : kDigestSpan);
  std::vector<unsigned char> tail(size < kDigestSpan ? 0 : kDigestSpan);
  if (!head.empty() && pread(fd, head.data(), head.size(), 0) != ssize_t(head.size())) return false;
  if (!tail.empty() && pread(fd, tail.data(), tail.size(), off_t(size - tail.size())) != ssize_t(tail.size()))
    return false;
  out = digest_bytes(head.data(), head.size(), tail.data(), tail.size(), size);
  return true;
}

int64_t mtime_ns(const struct stat& st) { return int64_t(st.st_mtim.tv_sec) * 1000000000ll + st.st_mtim.tv_nsec; }

// Rust's io::Error display for an OS error: "<strerror> (os error <n>)".
std::string os_error(int e) { return std::string(std::strerror(e)) + " (os error " + std::to_string(e) + ")"; }

// One line of the text at a time, as str::lines() gives them (split at '\n';
// a '\r' before it is whitespace to every rule below).
struct Lines {
  const char* p;
  const char* end;
  bool next(const char*& b, const char*& e) {
    if (p >= end) return false;
    b = p;
    const char* nl = static_cast<const char*>(std::memchr(p, '\n', size_t(end - p)));
    e = nl ? nl : end;
    p = nl ? nl + 1 : end;
    return true;
  }
};

inline bool is_ws(char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == '\v' || c == '\f'; }

// Whitespace-separated tokens of one line (str::split_whitespace).
struct Tokens {
  const char* p;
  const char* end;
  bool next(const char*& b, const char*& e) {
    while (p < end && is_ws(*p)) ++p;
    if (p >= end) return false;
    b = p;
    while (p < end && !is_ws(*p)) ++p;
    e = p;
    return true;
  }
};

bool token_is(const char* b, const char* e, const char* word) {
  size_t n = std::strlen(word);
  return size_t(e - b) == n && std::memcmp(b, word, n) == 0;
}

// usize::from_str: an optional '+', then decimal digits only.  Returns
// nullptr or the ParseIntError message.
const char* parse_usize(const char* b, const char* e, uint64_t& v) {
  if (b < e && *b == '+') ++b;
  if (b == e) return "invalid digit found in string";
  uint64_t x = 0;
  for (const char* q = b; q < e; ++q) {
    if (*q < '0' || *q > '9') return "invalid digit found in string";
    uint64_t d = uint64_t(*q - '0');
    if (x > (UINT64_MAX - d) / 10) return "number too large to fit in target type";
    x = x * 10 + d;
  }
  v = x;
  return nullptr;
}

// f64::from_str on one token, through strtod (the token is followed by
// whitespace or the terminating NUL, both of which stop strtod).  strtod's
// hexadecimal and nan(...) forms are not Rust float literals and are refused.
bool parse_f64(const char* b, const char* e, double& v) {
  for (const char* q = b; q < e; ++q)
    if (*q == 'x' || *q == 'X' || *q == '(') return false;
  char* stop = nullptr;
  v = std::strtod(b, &stop);
  return stop == e && e > b;
}

}  // namespace

int parse_ply(const char* data, size_t len, MeshData& out, std::string& err) {
  out = MeshData{};
  Lines lines{data, data + len};
  const char *b, *e, *tb, *te;
  // first line: the "ply" magic, trimmed
  bool magic = false;
  if (lines.next(b, e)) {
    Tokens t{b, e};
    magic = t.next(tb, te) && token_is(tb, te, "ply") && !t.next(tb, te);
  }
  if (!magic) {
    err = "Not a PLY file (missing 'ply' magic)";
    return LMN_ERR_PARSE;
  }
  uint64_t nv = 0, nf = 0;
  bool in_header = true;
  while (lines.next(b, e)) {
    Tokens t{b, e};
    const char* w[3][2];
    int n = 0;
    while (n < 3 && t.next(w[n][0], w[n][1])) ++n;
    if (n == 1 && token_is(w[0][0], w[0][1], "end_header")) {
      in_header = false;
      break;
    }
    if (n >= 3 && token_is(w[0][0], w[0][1], "element")) {
      const bool vert = token_is(w[1][0], w[1][1], "vertex"), face = token_is(w[1][0], w[1][1], "face");
      if (vert || face) {
        uint64_t c = 0;
        if (const char* pe = parse_usize(w[2][0], w[2][1], c)) {
          err = std::string(vert ? "Bad vertex count: " : "Bad face count: ") + pe;
          return LMN_ERR_PARSE;
        }
        (vert ? nv : nf) = c;
      }
    }
  }
  if (in_header) {
    err = "PLY header not terminated (missing 'end_header')";
    return LMN_ERR_PARSE;
  }
  if (nv > 0xffffffffull) {
    err = "PLY vertex count " + std::to_string(nv) + " exceeds this library's 2^32 - 1";
    return LMN_ERR_PARSE;
  }
  // the counts come from the file: reserve no more than its size can hold
  out.verts.reserve(size_t(std::min<uint64_t>(nv, len / 6 + 1)) * 3);
  static const char* axis[3] = {"x", "y", "z"};
  for (uint64_t i = 0; i < nv; ++i) {
    if (!lines.next(b, e)) {
      err = "Expected vertex " + std::to_string(i) + ", got EOF";
      return LMN_ERR_PARSE;
    }
    Tokens t{b, e};
    const char* w[3][2];
    int n = 0;
    while (n < 3 && t.next(w[n][0], w[n][1])) ++n;
    if (n < 3) {
      err = "Vertex " + std::to_string(i) + ": need at least 3 coordinates";
      return LMN_ERR_PARSE;
    }
    for (int k = 0; k < 3; ++k) {
      double x;
      if (!parse_f64(w[k][0], w[k][1], x)) {
        err = "Vertex " + std::to_string(i) + " " + axis[k] + ": invalid float literal";
        return LMN_ERR_PARSE;
      }
      out.verts.push_back(x);
    }
  }
  out.tris.reserve(size_t(std::min<uint64_t>(nf, len / 8 + 1)) * 3);
  std::vector<uint64_t> idx;
  for (uint64_t i = 0; i < nf; ++i) {
    if (!lines.next(b, e)) {
      err = "Expected face " + std::to_string(i) + ", got EOF";
      return LMN_ERR_PARSE;
    }
    Tokens t{b, e};
    if (!t.next(tb, te)) {
      err = "Face " + std::to_string(i) + ": empty line";
      return LMN_ERR_PARSE;
    }
    uint64_t n = 0;
    if (const char* pe = parse_usize(tb, te, n)) {
      err = "Face " + std::to_string(i) + " vertex count: " + pe;
      return LMN_ERR_PARSE;
    }
    idx.clear();
    for (uint64_t j = 0; j < n; ++j) {
      if (!t.next(tb, te)) {
        err = "Face " + std::to_string(i) + ": expected " + std::to_string(n) + " vertex indices";
        return LMN_ERR_PARSE;
      }
      idx.push_back(0);
      if (const char* pe = parse_usize(tb, te, idx.back())) {
        // the reference counts every token of the line before parsing any
        // (parts.len() < n + 1), so a short line is reported as short
        Tokens rest = t;
        uint64_t more = 0;
        while (j + 1 + more < n && rest.next(tb, te)) ++more;
        if (j + 1 + more < n) {
          err = "Face " + std::to_string(i) + ": expected " + std::to_string(n) + " vertex indices";
          return LMN_ERR_PARSE;
        }
        err = "Face " + std::to_string(i) + " index " + std::to_string(j + 1) + ": " + pe;
        return LMN_ERR_PARSE;
      }
    }
    // fan triangulation (src/obj.rs:226-240); a face with fewer than three
    // indices adds nothing
    for (size_t j = 1; j + 1 < idx.size(); ++j) {
      uint64_t i0 = idx[0], i1 = idx[j], i2 = idx[j + 1];
      if (i0 >= nv || i1 >= nv || i2 >= nv) {
        err = "Face " + std::to_string(i) + ": vertex index out of bounds";
        return LMN_ERR_PARSE;
      }
      out.tris.push_back(uint32_t(i0));
      out.tris.push_back(uint32_t(i1));
      out.tris.push_back(uint32_t(i2));
    }
  }
  out.n_faces = nf;
  return LMN_OK;
}

std::string sidecar_path(const std::string& mesh_path) { return mesh_path + ".lmnmesh"; }

SidecarMode sidecar_mode() {
  const char* v = std::getenv("LMN_MESH_SIDECAR");
  if (!v) return SIDECAR_AUTO;
  if (!std::strcmp(v, "0") || !std::strcmp(v, "off")) return SIDECAR_OFF;
  if (!std::strcmp(v, "read")) return SIDECAR_READ;
  return SIDECAR_AUTO;
}

namespace {

// Reads the sidecar if it describes the source as stat'ed and digested now.
bool read_sidecar_for(const std::string& mesh_path, uint64_t size, int64_t mtime, uint64_t digest, MeshData& out) {
  int fd = open(sidecar_path(mesh_path).c_str(), O_RDONLY | O_CLOEXEC);
  if (fd < 0) return false;
  bool ok = false;
  SidecarHeader h{};
  struct stat st{};
  if (fstat(fd, &st) == 0 && pread(fd, &h, sizeof(h), 0) == ssize_t(sizeof(h)) &&
      std::memcmp(h.magic, kMagic, 8) == 0 && h.version == kVersion && h.source_size == size &&
      h.source_mtime_ns == mtime && h.source_digest == digest && h.n_vertices <= 0xffffffffull &&
      h.n_triangles <= (uint64_t(1) << 40) && h.n_faces <= UINT64_MAX) {
    const uint64_t vbytes = h.n_vertices * 24, tbytes = h.n_triangles * 12;
    if (uint64_t(st.st_size) == sizeof(h) + vbytes + tbytes) {
      MeshData m;
      m.verts.resize(size_t(h.n_vertices) * 3);
      m.tris.resize(size_t(h.n_triangles) * 3);
      m.n_faces = h.n_faces;
      if (pread(fd, m.verts.data(), vbytes, sizeof(h)) == ssize_t(vbytes) &&
          pread(fd, m.tris.data(), tbytes, off_t(sizeof(h) + vbytes)) == ssize_t(tbytes)) {
        ok = true;
        for (uint32_t x : m.tris)
          if (x >= h.n_vertices) {
            ok = false;
            break;
          }
        if (ok) out = std::move(m);
      }
    }
  }
  close(fd);
  return ok;
}

bool write_sidecar_for(const std::string& mesh_path, const MeshData& mesh, uint64_t size, int64_t mtime,
                       uint64_t digest) {
  SidecarHeader h{};
  std::memcpy(h.magic, kMagic, 8);
  h.version = kVersion;
  h.n_vertices = mesh.n_vertices();
  h.n_triangles = mesh.n_triangles();
  h.n_faces = mesh.n_faces;
  h.source_size = size;
  h.source_mtime_ns = mtime;
  h.source_digest = digest;
  // written under a temporary name and renamed into place, so a reader never
  // sees a partial sidecar
  const std::string final_path = sidecar_path(mesh_path);
  // unique per process and writer, so that loaders on several threads never
  // share a temporary file
  static std::atomic<uint64_t> writers{0};
  const std::string tmp = final_path + ".tmp" + std::to_string(getpid()) + "." + std::to_string(writers++);
  FILE* f = std::fopen(tmp.c_str(), "wb");
  if (!f) return false;
  bool ok = std::fwrite(&h, sizeof(h), 1, f) == 1 &&
            std::fwrite(mesh.verts.data(), sizeof(double), mesh.verts.size(), f) == mesh.verts.size() &&
            std::fwrite(mesh.tris.data(), sizeof(uint32_t), mesh.tris.size(), f) == mesh.tris.size();
  ok = (std::fclose(f) == 0) && ok;
  if (ok) ok = std::rename(tmp.c_str(), final_path.c_str()) == 0;
  if (!ok) std::remove(tmp.c_str());
  return ok;
}

bool stat_source(const std::string& mesh_path, uint64_t& size, int64_t& mtime, uint64_t& digest) {
  int fd = open(mesh_path.c_str(), O_RDONLY | O_CLOEXEC);
  if (fd < 0) return false;
  struct stat st{};
  bool ok = fstat(fd, &st) == 0 && S_ISREG(st.st_mode);
  if (ok) {
    size = uint64_t(st.st_size);
    mtime = mtime_ns(st);
    ok = digest_of_fd(fd, size, digest);
  }
  close(fd);
  return ok;
}

}  // namespace

bool read_mesh_sidecar(const std::string& mesh_path, MeshData& out) {
  uint64_t size, digest;
  int64_t mtime;
  return stat_source(mesh_path, size, mtime, digest) && read_sidecar_for(mesh_path, size, mtime, digest, out);
}

bool write_mesh_sidecar(const std::string& mesh_path, const MeshData& mesh) {
  uint64_t size, digest;
  int64_t mtime;
  return stat_source(mesh_path, size, mtime, digest) && write_sidecar_for(mesh_path, mesh, size, mtime, digest);
}

int load_ply_file(const std::string& path, const std::string& name, MeshData& out, bool& from_sidecar,
                  std::string& err) {
  from_sidecar = false;
  const SidecarMode mode = sidecar_mode();
  int fd = open(path.c_str(), O_RDONLY | O_CLOEXEC);
  if (fd < 0) {
    err = "Failed to read mesh '" + name + "': " + os_error(errno);
    return LMN_ERR_IO;
  }
  struct stat st{};
  if (fstat(fd, &st) != 0 || S_ISDIR(st.st_mode)) {
    int e = S_ISDIR(st.st_mode) ? EISDIR : errno;
    close(fd);
    err = "Failed to read mesh '" + name + "': " + os_error(e);
    return LMN_ERR_IO;
  }
  const uint64_t size = uint64_t(st.st_size);
  const int64_t mtime = mtime_ns(st);
  if (mode != SIDECAR_OFF) {
    uint64_t digest;
    if (digest_of_fd(fd, size, digest) && read_sidecar_for(path, size, mtime, digest, out)) {
      close(fd);
      from_sidecar = true;
      return LMN_OK;
    }
  }
  std::string content;
  content.resize(size_t(size));
  uint64_t got = 0;
  while (got < size) {
    ssize_t r = pread(fd, &content[got], size_t(size - got), off_t(got));
    if (r <= 0) break;
    got += uint64_t(r);
  }
  close(fd);
  if (got != size) {
    err = "Failed to read mesh '" + name + "': short read";
    return LMN_ERR_IO;
  }
  int rc = parse_ply(content.c_str(), content.size(), out, err);
  if (rc == LMN_OK && mode == SIDECAR_AUTO) write_sidecar_for(path, out, size, mtime, digest_of(content.data(), size));
  return rc;
}

}  // namespace lmn

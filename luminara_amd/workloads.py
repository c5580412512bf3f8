"""Seeded synthetic workloads, emitted as scene TOML (plus ASCII PLY meshes)
that the reference's loader accepts, so the same text drives the GPU path
and the CPU oracle.

All randomness comes from `SeededRng` (SplitMix64 sequence, Steele et al.
2014) keyed by (master seed, workload, scene, object), so every workload is
reproducible from its name and seed.
"""
import math
import os

MASK64 = (1 << 64) - 1


def _mix(z):
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9 & MASK64
    z = (z ^ (z >> 27)) * 0x94D049BB133111EB & MASK64
    return z ^ (z >> 31)


class SeededRng:
    """SplitMix64 stream; uniform() returns [0, 1) doubles (53 bits)."""

    def __init__(self, *key):
        s = 0x9E3779B97F4A7C15
        for k in key:
            s = _mix((s ^ (int(k) & MASK64)) + 0x632BE59BD9B4E019 & MASK64)
        self.state = s

    def next_u64(self):
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK64
        return _mix(self.state)

    def uniform(self, lo=0.0, hi=1.0):
        return lo + (hi - lo) * ((self.next_u64() >> 11) * 2.0 ** -53)

    def randint(self, n):
        return self.next_u64() % n


def _v(a):
    return "[" + ", ".join(repr(float(x)) for x in a) + "]"


def _header(width, height, spp, depth, seed, background, camera):
    lines = ["[render]", f"width = {width}", f"height = {height}", f"samples = {spp}",
             f"max_depth = {depth}", f"seed = {seed}"]
    if background is not None:  # None: the loader's default, the sky gradient
        lines.append(f"background = {background}")
    lines += ["", "[camera]"]
    for k, v in camera.items():
        lines.append(f"{k} = {_v(v) if isinstance(v, (list, tuple)) else repr(float(v))}")
    lines.append("")
    return lines


def _mat(m):
    parts = [f'type = "{m["type"]}"']
    for k, v in m.items():
        if k == "type":
            continue
        parts.append(f"{k} = {_v(v) if isinstance(v, (list, tuple)) else repr(float(v))}")
    return "{ " + ", ".join(parts) + " }"


def cornell_box(width=256, height=256, spp=64, depth=8, seed=42):
    """Cornell-style box: five axis-aligned quad walls of a 555-unit cube,
    a 130x105 emissive quad light (15,15,15) under the ceiling, two white
    Lambertian boxes 165^3 (rotated 18 deg) and 165x330x165 (rotated -15 deg),
    black background, camera (278,278,-800) -> (278,278,0), fov 40."""
    white = {"type": "lambertian", "color": [0.73, 0.73, 0.73]}
    red = {"type": "lambertian", "color": [0.65, 0.05, 0.05]}
    green = {"type": "lambertian", "color": [0.12, 0.45, 0.15]}
    L = _header(width, height, spp, depth, seed, '{ type = "black" }',
                {"look_from": [278, 278, -800], "look_at": [278, 278, 0], "vfov": 40, "aperture": 0,
                 "focus_dist": 1})
    quads = [
        ([0, 0, 0], [555, 0, 0], [0, 0, 555], white),      # floor
        ([0, 555, 0], [555, 0, 0], [0, 0, 555], white),    # ceiling
        ([0, 0, 555], [555, 0, 0], [0, 555, 0], white),    # back
        ([0, 0, 0], [0, 555, 0], [0, 0, 555], red),        # x = 0 wall
        ([555, 0, 0], [0, 555, 0], [0, 0, 555], green),    # x = 555 wall
        ([213, 554, 227], [130, 0, 0], [0, 0, 105], {"type": "emissive", "color": [15, 15, 15], "intensity": 1}),
    ]
    for q, u, v, m in quads:
        L += ["[[quad]]", f"q = {_v(q)}", f"u = {_v(u)}", f"v = {_v(v)}", f"material = {_mat(m)}", ""]
    L += ["[[box]]", "min = [0.0, 0.0, 0.0]", "max = [165.0, 165.0, 165.0]", f"material = {_mat(white)}",
          "rotate_y = 18.0", "translate = [130.0, 0.0, 65.0]", ""]
    L += ["[[box]]", "min = [0.0, 0.0, 0.0]", "max = [165.0, 330.0, 165.0]", f"material = {_mat(white)}",
          "rotate_y = -15.0", "translate = [265.0, 0.0, 295.0]", ""]
    return "\n".join(L)


def quadrics_scene(width=64, height=48, spp=16, depth=8, seed=42):
    """Disks, cylinders, cones, an ellipsoid, a hemisphere, an annulus and a
    capsule (src/disk.rs, src/cylinder.rs, src/cone.rs, src/ellipsoid.rs,
    src/hemisphere.rs, src/annulus.rs, src/capsule.rs) on a ground quad:
    rotated and translated cones, an emissive disk that the loader registers
    as an NEE light, a tilted disk, spheres between them, and in the front row
    the ellipsoid, a raised hemisphere open below, a tilted metal annulus and
    the capsule; a rotated prism, a pyramid and a rotated, translated wedge
    (src/prism.rs, src/pyramid.rs, src/wedge.rs) at the sides and back."""
    rng = SeededRng(seed, 9)
    L = _header(width, height, spp, depth, seed, '{ type = "sky" }',
                {"look_from": [0.0, 3.0, 9.0], "look_at": [0.0, 0.8, 0.0], "vfov": 40, "aperture": 0.0,
                 "focus_dist": 1.0})
    L += ["[[quad]]", "q = [-6.0, 0.0, -6.0]", "u = [12.0, 0.0, 0.0]", "v = [0.0, 0.0, 12.0]",
          'material = { type = "lambertian", color = [0.6, 0.6, 0.6] }', ""]
    for i in range(6):
        c = [-3.0 + 1.2 * i, 0.0, -1.0 + 0.8 * rng.uniform()]
        col = [0.2 + 0.7 * rng.uniform() for _ in range(3)]
        L += ["[[sphere]]", f"center = {_v([c[0], 0.3, c[2] + 1.6])}", "radius = 0.3",
              f"material = {_mat({'type': 'lambertian', 'color': col})}", ""]
    for i in range(4):
        c = [-2.4 + 1.6 * i, 0.0, -0.5 * rng.uniform()]
        col = [0.2 + 0.7 * rng.uniform() for _ in range(3)]
        L += ["[[cone]]", f"center = {_v(c)}", f"radius = {0.35 + 0.2 * rng.uniform()!r}",
              f"height = {0.8 + 0.8 * rng.uniform()!r}", f"material = {_mat({'type': 'lambertian', 'color': col})}"]
        if i % 2:
            L += [f"rotate_y = {20.0 * i!r}", f"translate = {_v([0.1 * i, 0.0, 0.2])}"]
        L += [""]
    L += ["[[disk]]", "center = [0.0, 3.5, 0.5]", "normal = [0.0, -1.0, 0.0]", "radius = 0.8",
          'material = { type = "emissive", color = [1.0, 0.9, 0.8], intensity = 6.0 }', ""]
    L += ["[[disk]]", "center = [2.5, 0.9, -1.2]", "normal = [-0.4, 0.3, 1.0]", "radius = 0.7",
          'material = { type = "metal", color = [0.8, 0.8, 0.9], fuzz = 0.1 }', ""]
    for i in range(3):
        c = [-2.0 + 2.0 * i, 0.0, -2.2]
        col = [0.2 + 0.7 * rng.uniform() for _ in range(3)]
        L += ["[[cylinder]]", f"center = {_v(c)}", f"radius = {0.3 + 0.2 * rng.uniform()!r}",
              f"height = {1.0 + rng.uniform()!r}", f"material = {_mat({'type': 'lambertian', 'color': col})}"]
        if i == 1:
            L += ["rotate_z = 12.0", "translate = [0.0, 0.1, 0.0]"]
        L += [""]
    col = [[0.2 + 0.7 * rng.uniform() for _ in range(3)] for _ in range(3)]
    L += ["[[capsule]]", "center = [2.6, 0.35, 2.0]", "radius = 0.3", "height = 0.9",
          f"material = {_mat({'type': 'lambertian', 'color': col[0]})}", ""]
    L += ["[[ellipsoid]]", "center = [-2.5, 0.5, 2.2]", "radii = [0.6, 0.4, 0.3]",
          f"material = {_mat({'type': 'lambertian', 'color': col[1]})}", ""]
    L += ["[[hemisphere]]", "center = [-0.8, 0.5, 2.3]", "radius = 0.45",
          f"material = {_mat({'type': 'lambertian', 'color': col[2]})}", ""]
    L += ["[[annulus]]", "center = [0.9, 0.8, 2.3]", "normal = [0.3, 1.0, -0.6]", "inner_radius = 0.2",
          "outer_radius = 0.55", 'material = { type = "metal", color = [0.9, 0.8, 0.6], fuzz = 0.05 }', ""]
    col = [[0.2 + 0.7 * rng.uniform() for _ in range(3)] for _ in range(3)]
    L += ["[[prism]]", "center = [0.0, 0.0, 0.0]", "side = 0.7", "height = 1.0",
          f"material = {_mat({'type': 'lambertian', 'color': col[0]})}", "rotate_y = 25.0",
          "translate = [3.3, 0.0, 0.9]", ""]
    L += ["[[wedge]]", "min = [-0.4, 0.0, -0.25]", "max = [0.4, 0.8, 0.25]",
          f"material = {_mat({'type': 'lambertian', 'color': col[1]})}", "rotate_y = 15.0",
          "translate = [1.0, 0.0, -2.7]", ""]
    L += ["[[pyramid]]", "center = [-3.3, 0.0, 0.9]", "base_size = 0.9", "height = 1.1",
          f"material = {_mat({'type': 'lambertian', 'color': col[2]})}", ""]
    L += ["[[rounded_box]]", "center = [-3.2, 0.45, -1.6]", "half_size = [0.3, 0.25, 0.35]", "radius = 0.12",
          'material = { type = "metal", color = [0.8, 0.7, 0.6], fuzz = 0.2 }', ""]
    return "\n".join(L)


def traced_scene(width=64, height=48, spp=16, depth=8, seed=42):
    """The sphere-traced kinds (src/mobius.rs, src/rounded_box.rs): two Mobius
    strips (one thin, one with the default thickness), rounded boxes of
    several proportions, on a ground quad under the sky, spheres beside them."""
    rng = SeededRng(seed, 11)
    L = _header(width, height, spp, depth, seed, '{ type = "sky" }',
                {"look_from": [0.0, 2.5, 7.0], "look_at": [0.0, 0.7, 0.0], "vfov": 40, "aperture": 0.0,
                 "focus_dist": 1.0})
    L += ["[[quad]]", "q = [-6.0, 0.0, -6.0]", "u = [12.0, 0.0, 0.0]", "v = [0.0, 0.0, 12.0]",
          'material = { type = "lambertian", color = [0.6, 0.6, 0.6] }', ""]
    for i in range(3):
        col = [0.2 + 0.7 * rng.uniform() for _ in range(3)]
        L += ["[[rounded_box]]", f"center = {_v([-2.0 + 2.0 * i, 0.5, -1.5])}",
              f"half_size = {_v([0.2 + 0.3 * rng.uniform(), 0.3, 0.2 + 0.3 * rng.uniform()])}",
              f"radius = {0.05 + 0.15 * rng.uniform()!r}", f"material = {_mat({'type': 'lambertian', 'color': col})}",
              ""]
    L += ["[[mobius]]", "center = [-1.2, 0.9, 0.8]", "radius = 0.8", "width = 0.3", "thickness = 0.08",
          'material = { type = "lambertian", color = [0.8, 0.3, 0.2] }', ""]
    L += ["[[mobius]]", "center = [1.3, 0.8, 0.6]", "radius = 0.6", "width = 0.25",
          'material = { type = "metal", color = [0.7, 0.7, 0.8], fuzz = 0.1 }', ""]
    for i in range(4):
        col = [0.2 + 0.7 * rng.uniform() for _ in range(3)]
        L += ["[[sphere]]", f"center = {_v([-3.0 + 2.0 * i, 0.25, 2.0])}", "radius = 0.25",
              f"material = {_mat({'type': 'lambertian', 'color': col})}", ""]
    return "\n".join(L)


def random_spheres(width=1920, height=1080, spp=128, depth=50, seed=42, grid=11):
    """Random-sphere scene: ground sphere r=1000 (albedo 0.5), a grid of
    r=0.2 spheres at integer x,z in [-grid, grid) with U[0,0.9] jitter
    (80% Lambertian albedo U^2, 15% metal U[.5,1] fuzz U[0,.5], 5% glass
    1.5), glass / Lambertian (0.4,0.2,0.1) / metal (0.7,0.6,0.5) r=1 spheres,
    sky gradient, camera (13,2,3) -> origin, fov 20, aperture 0.1, focus 10."""
    rng = SeededRng(seed, 1)
    L = _header(width, height, spp, depth, seed, '{ type = "sky" }',
                {"look_from": [13, 2, 3], "look_at": [0, 0, 0], "vfov": 20, "aperture": 0.1, "focus_dist": 10})

    def sphere(c, r, m):
        L.extend(["[[sphere]]", f"center = {_v(c)}", f"radius = {float(r)!r}", f"material = {_mat(m)}", ""])

    sphere([0, -1000, 0], 1000, {"type": "lambertian", "color": [0.5, 0.5, 0.5]})
    sphere([0, 1, 0], 1, {"type": "dielectric", "refraction_index": 1.5})
    sphere([-4, 1, 0], 1, {"type": "lambertian", "color": [0.4, 0.2, 0.1]})
    sphere([4, 1, 0], 1, {"type": "metal", "color": [0.7, 0.6, 0.5], "fuzz": 0.0})
    for a in range(-grid, grid):
        for b in range(-grid, grid):
            choose = rng.uniform()
            c = [a + 0.9 * rng.uniform(), 0.2, b + 0.9 * rng.uniform()]
            if math.dist(c, [4, 0.2, 0]) <= 0.9:
                continue
            if choose < 0.8:
                alb = [rng.uniform() * rng.uniform() for _ in range(3)]
                m = {"type": "lambertian", "color": alb}
            elif choose < 0.95:
                m = {"type": "metal", "color": [rng.uniform(0.5, 1.0) for _ in range(3)],
                     "fuzz": rng.uniform(0, 0.5)}
            else:
                m = {"type": "dielectric", "refraction_index": 1.5}
            sphere(c, 0.2, m)
    return "\n".join(L)


WORKLOADS = {
    "cornell": cornell_box,
    "random_spheres": random_spheres,
}


# ---------------------------------------------------------------------------
# Meshes

def perlin_table(seed):
    """Permutation table of Perlin's improved noise, shuffled by the seed."""
    rng = SeededRng(seed, 0x9E11)
    perm = list(range(256))
    for i in range(255, 0, -1):  # Fisher-Yates
        j = rng.randint(i + 1)
        perm[i], perm[j] = perm[j], perm[i]
    import numpy as np
    return np.array(perm + perm, dtype=np.int64)


def perlin3(perm, x, y, z):
    """Ken Perlin's improved noise (SIGGRAPH 2002) on numpy arrays."""
    import numpy as np
    X = np.floor(x).astype(np.int64)
    Y = np.floor(y).astype(np.int64)
    Z = np.floor(z).astype(np.int64)
    x = x - X
    y = y - Y
    z = z - Z
    X &= 255
    Y &= 255
    Z &= 255

    def fade(t):
        return t * t * t * (t * (t * 6 - 15) + 10)

    def grad(h, x, y, z):
        h = h & 15
        u = np.where(h < 8, x, y)
        v = np.where(h < 4, y, np.where((h == 12) | (h == 14), x, z))
        return np.where(h & 1, -u, u) + np.where(h & 2, -v, v)

    def lerp(t, a, b):
        return a + t * (b - a)

    u, v, w = fade(x), fade(y), fade(z)
    A = perm[X] + Y
    AA = perm[A] + Z
    AB = perm[A + 1] + Z
    B = perm[X + 1] + Y
    BA = perm[B] + Z
    BB = perm[B + 1] + Z
    return lerp(w, lerp(v, lerp(u, grad(perm[AA], x, y, z), grad(perm[BA], x - 1, y, z)),
                        lerp(u, grad(perm[AB], x, y - 1, z), grad(perm[BB], x - 1, y - 1, z))),
                lerp(v, lerp(u, grad(perm[AA + 1], x, y, z - 1), grad(perm[BA + 1], x - 1, y, z - 1)),
                     lerp(u, grad(perm[AB + 1], x, y - 1, z - 1), grad(perm[BB + 1], x - 1, y - 1, z - 1))))


def fbm(perm, p, octaves=4, freq=1.5):
    """Fractal sum of `octaves` noise octaves (lacunarity 2, gain 0.5),
    normalised to about [-1, 1]."""
    total = 0.0
    amp = 1.0
    norm = 0.0
    for o in range(octaves):
        f = freq * (2 ** o)
        total = total + amp * perlin3(perm, p[:, 0] * f, p[:, 1] * f, p[:, 2] * f)
        norm += amp
        amp *= 0.5
    return total / norm * 1.5


_ICO = None


def _icosahedron():
    import numpy as np
    global _ICO
    if _ICO is None:
        t = (1.0 + 5.0 ** 0.5) / 2.0
        v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t],
                      [0, 1, -t], [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], dtype=np.float64)
        f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4],
                      [11, 10, 2], [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
                      [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], dtype=np.int64)
        _ICO = (v, f)
    return _ICO


def geodesic_sphere(freq):
    """Class-I geodesic sphere: every icosahedron face split into freq^2
    triangles, 20 freq^2 in all (freq = 2^k gives the k-level icosphere's
    count, e.g. freq 256 -> 1,310,720).  Shared edge points are evaluated
    with the corners in index order, so neighbouring faces produce bitwise
    identical vertices (watertight).  Returns unit vertices (V, 3) and
    triangles (T, 3) with per-face vertex blocks."""
    import numpy as np
    V, F = _icosahedron()
    n = freq
    ii, jj = np.meshgrid(np.arange(n + 1), np.arange(n + 1), indexing="ij")
    mask = ii + jj <= n
    ii, jj = ii[mask], jj[mask]
    kk = n - ii - jj
    local = -np.ones((n + 1, n + 1), dtype=np.int64)
    local[ii, jj] = np.arange(len(ii))
    # triangles of one face in the (i, j) lattice
    tris = []
    for i in range(n):
        for j in range(n - i):
            tris.append((local[i, j], local[i + 1, j], local[i, j + 1]))
            if j < n - i - 1:
                tris.append((local[i + 1, j], local[i + 1, j + 1], local[i, j + 1]))
    tris = np.array(tris, dtype=np.int64)
    npts = len(ii)
    verts = np.empty((20 * npts, 3))
    faces = np.empty((20 * len(tris), 3), dtype=np.int64)
    w = np.stack([ii, jj, kk], axis=1).astype(np.float64)
    for fi, (a, b, c) in enumerate(F):
        corners = [(a, 0), (b, 1), (c, 2)]
        corners.sort()
        p = np.zeros((npts, 3))
        for cid, col in corners:  # canonical summation order
            p = p + w[:, col:col + 1] * V[cid][None, :]
        p = p / np.linalg.norm(p, axis=1, keepdims=True)
        verts[fi * npts:(fi + 1) * npts] = p
        faces[fi * len(tris):(fi + 1) * len(tris)] = tris + fi * npts
    return verts, faces


def displaced_sphere(freq, center, radius, seed, amplitude=0.05, octaves=4):
    """Geodesic sphere displaced along its normal by seeded fBm (amplitude
    `amplitude * radius`)."""
    import numpy as np
    u, f = geodesic_sphere(freq)
    perm = perlin_table(seed)
    h = fbm(perm, u, octaves)
    v = np.asarray(center)[None, :] + u * (radius * (1.0 + amplitude * h))[:, None]
    return v, f


def write_ply(path, verts, faces):
    with open(path, "w") as fh:
        fh.write(f"ply\nformat ascii 1.0\nelement vertex {len(verts)}\nproperty float x\nproperty float y\n"
                 f"property float z\nelement face {len(faces)}\nproperty list uchar int vertex_indices\nend_header\n")
        fh.write("\n".join(f"{x!r} {y!r} {z!r}" for x, y, z in verts.tolist()))
        fh.write("\n")
        fh.write("\n".join(f"3 {a} {b} {c}" for a, b, c in faces.tolist()))
        fh.write("\n")


def _microfacet(color, roughness, metallic=0.0):
    from . import _abi
    m = _abi.Material()
    m.kind = _abi.MAT_MICROFACET
    m.texture = -1
    for k in range(3):
        m.color[k] = color[k]
    m.roughness = min(max(roughness, 0.01), 1.0)
    m.metallic = metallic
    m.intensity = 1.0
    m.ior = 1.5
    return m


def mesh_heavy(width=1920, height=1080, spp=64, depth=12, seed=42, freq=256, patch=1024, n_tori=64,
               n_superellipsoids=16):
    """Mesh-heavy scene: displaced geodesic icosphere (freq 256 = 1,310,720
    triangles, 4-octave seeded Perlin fBm, amplitude 0.05 r) with one GGX
    material per `patch` consecutive triangles (roughness U[0.05, 0.6]),
    an infinite Lambertian ground plane, `n_tori` tori (R 0.5, r 0.15) and
    `n_superellipsoids` superellipsoids (e1, e2 ~ U[0.3, 3]) at seeded
    positions, three r=0.5 emissive spheres (20) sampled by NEE, sun + sky.
    Returns (toml, add_mesh) where add_mesh(world) appends the mesh through
    the descriptor API (the TOML part goes through the scene loader)."""
    import numpy as np
    rng = SeededRng(seed, 2)
    L = _header(width, height, spp, depth, seed,
                '{ type = "sun", direction = [0.35, 0.6, -0.45] }',
                {"look_from": [0.0, 3.2, 10.5], "look_at": [0.0, 1.6, 0.0], "vfov": 38, "aperture": 0.0,
                 "focus_dist": 1.0})
    for c in ([-3.2, 5.0, 2.0], [3.4, 5.5, 1.0], [0.0, 6.0, -3.0]):
        L += ["[[sphere]]", f"center = {_v(c)}", "radius = 0.5",
              'material = { type = "emissive", color = [1.0, 1.0, 1.0], intensity = 20.0 }', ""]
    L += ["[[plane]]", "point = [0.0, 0.0, 0.0]", "normal = [0.0, 1.0, 0.0]",
          'material = { type = "lambertian", color = [0.5, 0.5, 0.5] }', ""]
    for i in range(n_tori):
        ang = 2 * math.pi * (i + rng.uniform()) / n_tori
        rad = 3.6 + 2.2 * rng.uniform()
        c = [rad * math.cos(ang), 0.15, rad * math.sin(ang) * 0.6 + 0.5]
        col = [0.2 + 0.7 * rng.uniform() for _ in range(3)]
        L += ["[[torus]]", f"center = {_v(c)}", "major_radius = 0.5", "minor_radius = 0.15",
              f"material = {_mat({'type': 'lambertian', 'color': col})}", ""]
    for i in range(n_superellipsoids):
        ang = 2 * math.pi * (i + 0.5) / n_superellipsoids
        c = [2.9 * math.cos(ang), 0.35, 2.9 * math.sin(ang) * 0.7 + 0.3]
        e1, e2 = rng.uniform(0.3, 3.0), rng.uniform(0.3, 3.0)
        col = [0.2 + 0.7 * rng.uniform() for _ in range(3)]
        L += ["[[superellipsoid]]", f"center = {_v(c)}", "scale = [0.35, 0.35, 0.35]", f"e1 = {e1!r}",
              f"e2 = {e2!r}", f"material = {_mat({'type': 'pbr', 'color': col, 'roughness': 0.4})}", ""]
    toml = "\n".join(L)

    def add_mesh(world):
        v, f = displaced_sphere(freq, [0.0, 1.75, 0.0], 1.6, seed)
        n_patch = (len(f) + patch - 1) // patch
        mats = np.empty(len(f), dtype=np.uint32)
        prng = SeededRng(seed, 3)
        for p in range(n_patch):
            col = [0.55 + 0.4 * prng.uniform() for _ in range(3)]
            idx = world.add_material(_microfacet(col, prng.uniform(0.05, 0.6)))
            mats[p * patch:(p + 1) * patch] = idx
        world.add_triangles(v, f, materials=mats)
        return len(f)

    return toml, add_mesh


def batch_scene(index, seed=42, max_res=512):
    """Scene `index` of the 64-scene batch: a seeded displaced geodesic sphere
    with a log-uniform triangle count in [1k, 200k] plus a few random spheres
    on a ground quad, 1-4 emissive quad lights (NEE), the default sky
    background, depth 8; resolution in {128, 256, 384, 512}^2 and spp in
    {16, 64, 256}.  Returns (toml, add_mesh, info)."""
    rng = SeededRng(seed, 4, index)
    n_target = math.exp(rng.uniform(math.log(1000), math.log(200000)))
    freq = max(7, int(round(math.sqrt(n_target / 20.0))))
    res = [128, 256, 384, 512][rng.randint(4)]
    res = min(res, max_res)
    spp = [16, 64, 256][rng.randint(3)]
    L = _header(res, res, spp, 8, seed + index, None,
                {"look_from": [0.0, 2.2, 6.5], "look_at": [0.0, 1.0, 0.0], "vfov": 40, "aperture": 0.0,
                 "focus_dist": 1.0})
    for k in range(6):
        c = [rng.uniform(-2.5, 2.5), 0.3, rng.uniform(-1.5, 2.0)]
        choose = rng.uniform()
        if choose < 0.6:
            m = {"type": "lambertian", "color": [rng.uniform(0.2, 0.9) for _ in range(3)]}
        elif choose < 0.85:
            m = {"type": "metal", "color": [rng.uniform(0.5, 1.0) for _ in range(3)], "fuzz": rng.uniform(0, 0.4)}
        else:
            m = {"type": "dielectric", "refraction_index": 1.5}
        L += ["[[sphere]]", f"center = {_v(c)}", "radius = 0.3", f"material = {_mat(m)}", ""]
    n_lights = 1 + rng.randint(4)
    for k in range(n_lights):
        x = -2.5 + 5.0 * (k + 0.5) / n_lights
        L += ["[[quad]]", f"q = {_v([x - 0.4, 4.0, -0.5])}", "u = [0.8, 0.0, 0.0]", "v = [0.0, 0.0, 1.0]",
              f"material = {{ type = \"emissive\", color = [1.0, 0.95, 0.9], intensity = {12.0 / n_lights!r} }}", ""]
    L += ["[[quad]]", "q = [-8.0, 0.0, -8.0]", "u = [16.0, 0.0, 0.0]", "v = [0.0, 0.0, 16.0]",
          'material = { type = "lambertian", color = [0.6, 0.6, 0.6] }', ""]
    toml = "\n".join(L)
    col = [0.4 + 0.5 * rng.uniform() for _ in range(3)]
    rough = rng.uniform(0.05, 0.6)
    mesh_seed = seed * 1000 + index

    def add_mesh(world):
        v, f = displaced_sphere(freq, [0.0, 1.2, 0.0], 1.0, mesh_seed)
        idx = world.add_material(_microfacet(col, rough))
        world.add_triangles(v, f, material=idx)
        return len(f)

    info = {"res": res, "spp": spp, "tris": 20 * freq * freq, "lights": n_lights}
    return toml, add_mesh, info


def inverse_rendering_scene(width=512, height=512, spp=32, depth=8, seed=42, tex_res=256, n_spheres=32,
                            hidden=False, intensity=None):
    """Differentiable inverse-rendering workload: the Cornell geometry, a
    tex_res^2 FP32 albedo texture on the back wall (texel values k/255),
    n_spheres seeded GGX spheres (r U[20,60]) each with its own roughness, and
    the light's emission scalar (the intensity of its material).  hidden=True
    draws the parameters of the target image from the seed; otherwise they
    are the starting point (grey texture, roughness 0.3, intensity 1);
    `intensity` overrides the emission scalar.
    Returns (toml, setup) where setup(world) installs the texture and returns
    {"texture", "wall_material", "emission_material", "sphere_materials"}."""
    rng = SeededRng(seed, 5)
    L = cornell_box(width, height, spp, depth, seed).split("\n")
    spheres = []
    for k in range(n_spheres):
        r = rng.uniform(20.0, 60.0)
        c = [rng.uniform(r, 555.0 - r), r, rng.uniform(r, 555.0 - r)]
        col = [rng.uniform(0.3, 0.9) for _ in range(3)]
        spheres.append((c, r, col))
    hrng = SeededRng(seed, 6)
    rough = [hrng.uniform(0.05, 0.6) if hidden else 0.3 for _ in range(n_spheres)]
    hidden_intensity = hrng.uniform(0.7, 1.3)
    if intensity is None:
        intensity = hidden_intensity if hidden else 1.0
    for (c, r, col), ro in zip(spheres, rough):
        m = {"type": "pbr", "color": col, "roughness": ro, "metallic": 0.0}
        L += ["[[sphere]]", f"center = {_v(c)}", f"radius = {r!r}", f"material = {_mat(m)}", ""]
    toml = "\n".join(L)
    toml = toml.replace("intensity = 1.0 }", f"intensity = {intensity!r} }}", 1)
    import numpy as np
    if hidden:
        k = SeededRng(seed, 7).state
        z = (np.uint64(k) + np.arange(1, tex_res * tex_res * 3 + 1, dtype=np.uint64)
             * np.uint64(0x9E3779B97F4A7C15))  # SplitMix64 sequence, wrapping
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
        texels = ((z % np.uint64(256)).astype(np.float32) / np.float32(255.0)).reshape(tex_res, tex_res, 3)
    else:
        texels = np.full((tex_res, tex_res, 3), np.float32(186.0 / 255.0), dtype=np.float32)

    def setup(world):
        from . import _abi
        prims = world.prims
        wall = emit = None
        for p in prims:
            v = list(p.p)
            if p.kind == _abi.PRIM_QUAD and v[:3] == [0.0, 0.0, 555.0] and v[3:6] == [555.0, 0.0, 0.0]:
                wall = p.material
            if p.kind == _abi.PRIM_QUAD and world.materials()[p.material].kind == _abi.MAT_EMISSIVE:
                emit = p.material
        tex = world.add_texture(texels)
        m = world.materials()[wall]
        m.texture = tex
        world.set_material(wall, m)
        sph = sorted({p.material for p in prims if p.kind == _abi.PRIM_SPHERE})
        return {"texture": tex, "wall_material": wall, "emission_material": emit, "sphere_materials": sph}

    return toml, setup


def smooth_mesh_scene(width=64, height=48, spp=16, depth=8, seed=42, freq=10):
    """Smooth-shaded meshes (OBJ faces with vn / vt, MeshTriangle with
    attributes, src/obj.rs:250-313): a geodesic sphere with its radial vertex
    normals and spherical texture coordinates under an image texture, and a
    displaced sphere with vertex normals only (GGX), on a ground quad lit by
    a quad light (NEE) under the sky.  Returns (toml, add_meshes); add_meshes
    appends both meshes and the texture and returns the triangle count."""
    import numpy as np
    L = _header(width, height, spp, depth, seed, None,
                {"look_from": [0.0, 1.6, 5.5], "look_at": [0.0, 0.9, 0.0], "vfov": 40, "aperture": 0.0,
                 "focus_dist": 1.0})
    L += ["[[quad]]", "q = [-0.6, 3.5, -0.5]", "u = [1.2, 0.0, 0.0]", "v = [0.0, 0.0, 1.0]",
          'material = { type = "emissive", color = [1.0, 0.95, 0.9], intensity = 10.0 }', ""]
    L += ["[[quad]]", "q = [-6.0, 0.0, -6.0]", "u = [12.0, 0.0, 0.0]", "v = [0.0, 0.0, 12.0]",
          'material = { type = "lambertian", color = [0.6, 0.6, 0.6] }', ""]
    toml = "\n".join(L)

    def add_meshes(world):
        from . import _abi
        u, f = geodesic_sphere(freq)
        centre = np.array([-0.9, 0.9, 0.0])
        v = centre + 0.85 * u
        nrm = u[f]  # (T, 3, 3): the corners' radial normals
        uv = np.stack([(np.arctan2(-u[:, 2], u[:, 0]) + np.pi) / (2 * np.pi), np.arccos(-u[:, 1]) / np.pi], axis=1)[f]
        yy, xx = np.mgrid[0:8, 0:16]  # a 16 x 8 checker
        tex = (((xx + yy) % 2)[..., None] * np.array([0.7, 0.3, 0.1])
               + np.array([0.2, 0.25, 0.3])).astype(np.float32)
        t = world.add_texture(tex)
        m = _abi.Material()
        m.kind, m.texture, m.intensity, m.ior = _abi.MAT_LAMBERTIAN, t, 1.0, 1.5
        for c in range(3):
            m.color[c] = 0.5
        world.add_mesh(v, f, normals=nrm, uvs=uv, material=world.add_material(m))
        v2, f2 = displaced_sphere(freq, [1.0, 0.8, 0.2], 0.75, seed)
        d = v2 - np.array([1.0, 0.8, 0.2])
        vn = d / np.linalg.norm(d, axis=1)[:, None]
        world.add_mesh(v2, f2, normals=vn[f2], material=world.add_material(_microfacet([0.8, 0.7, 0.5], 0.3)))
        return len(f) + len(f2)

    return toml, add_meshes

"""Boundary meshes for the exterior Burton-Miller path.

`BoundaryMesh` mirrors the reference's mesh type
(src/beat_engine/julia_local/src/BeatEngineCore.jl:224-232, constructor :344-364):
per face a centroid, the unit normal cross(v2-v1, v3-v1)/|.|, and the area |.|/2.
Faces are 0-based here (the reference is 1-based).

The generators build the synthetic workloads the benchmarks are quoted on:
icospheres (the construction of julia_local/tests/analytical_exterior_tests.jl:4-66),
radially jittered icospheres, graded capped cylinders and harmonic morphs.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np


@dataclasses.dataclass
class BoundaryMesh:
    vertices: np.ndarray  # (nv, 3) float64
    faces: np.ndarray  # (nf, 3) int32, 0-based, outward winding
    physical_tags: np.ndarray  # (nf,) int64

    def __post_init__(self):
        self.vertices = np.ascontiguousarray(self.vertices, dtype=np.float64)
        faces = np.asarray(self.faces)
        if not np.issubdtype(faces.dtype, np.integer):
            raise ValueError("faces must hold integer vertex indices")
        if self.vertices.ndim != 2 or self.vertices.shape[1] != 3:
            raise ValueError("vertices must be (nv, 3)")
        if faces.ndim != 2 or faces.shape[1] != 3:
            raise ValueError("faces must be (nf, 3)")
        # range-checked in the caller's integer type, before the narrowing to int32
        if faces.size and (int(faces.min()) < 0 or int(faces.max()) >= len(self.vertices)):
            raise ValueError("face vertex index out of range")
        self.faces = np.ascontiguousarray(faces, dtype=np.int32)
        self.physical_tags = np.ascontiguousarray(self.physical_tags, dtype=np.int64)
        v1 = self.vertices[self.faces[:, 0]]
        v2 = self.vertices[self.faces[:, 1]]
        v3 = self.vertices[self.faces[:, 2]]
        c = np.cross(v2 - v1, v3 - v1)
        n = np.sqrt((c * c).sum(axis=1))
        self.centroids = (v1 + v2 + v3) / 3.0
        self.areas = n / 2.0
        self.normals = c / n[:, None]

    @property
    def num_vertices(self) -> int:
        return int(self.vertices.shape[0])

    @property
    def num_faces(self) -> int:
        return int(self.faces.shape[0])


def load_gmsh22(path: str, scale: float = 1.0) -> BoundaryMesh:
    """Gmsh 2.2 ASCII triangles with physical tags (BeatEngineCore.jl:301-342)."""
    with open(path, "r") as fh:
        lines = [ln.rstrip("\n") for ln in fh]

    def find(tag):
        for i, ln in enumerate(lines):
            if ln.strip() == tag:
                return i
        raise ValueError("Only Gmsh 2.2 ASCII meshes with Nodes and Elements sections are supported.")

    ns, ne = find("$Nodes"), find("$EndNodes")
    es, ee = find("$Elements"), find("$EndElements")
    index = {}
    verts = []
    for ln in lines[ns + 2 : ne]:
        parts = ln.split()
        if len(parts) < 4:
            continue
        index[int(parts[0])] = len(verts)
        verts.append([float(parts[1]) * scale, float(parts[2]) * scale, float(parts[3]) * scale])
    faces, tags = [], []
    for ln in lines[es + 2 : ee]:
        parts = ln.split()
        if len(parts) < 8 or int(parts[1]) != 2:
            continue
        ids = [index.get(int(p), -1) for p in parts[-3:]]
        if min(ids) < 0:
            continue
        faces.append(ids)
        tags.append(int(parts[3]))
    return BoundaryMesh(np.array(verts), np.array(faces, dtype=np.int32), np.array(tags))


_ICO_VERTS = None


def icosphere(radius: float, subdivisions: int, centre=(0.0, 0.0, 0.0), tag: int = 1) -> BoundaryMesh:
    """Outward-oriented icosphere: 12 vertices, 20 faces, 4^level midpoint splits,
    projected to the sphere (construction of analytical_exterior_tests.jl:4-66)."""
    if subdivisions < 0:
        raise ValueError("Icosphere subdivision level must be >= 0.")
    phi = (1.0 + math.sqrt(5.0)) / 2.0
    verts = [
        (-1, phi, 0), (1, phi, 0), (-1, -phi, 0), (1, -phi, 0),
        (0, -1, phi), (0, 1, phi), (0, -1, -phi), (0, 1, -phi),
        (phi, 0, -1), (phi, 0, 1), (-phi, 0, -1), (-phi, 0, 1),
    ]
    verts = [np.array(v, dtype=np.float64) for v in verts]
    faces = [
        (1, 12, 6), (1, 6, 2), (1, 2, 8), (1, 8, 11), (1, 11, 12),
        (2, 6, 10), (6, 12, 5), (12, 11, 3), (11, 8, 7), (8, 2, 9),
        (4, 10, 5), (4, 5, 3), (4, 3, 7), (4, 7, 9), (4, 9, 10),
        (5, 10, 6), (3, 5, 12), (7, 3, 11), (9, 7, 8), (10, 9, 2),
    ]
    faces = [(a - 1, b - 1, c - 1) for a, b, c in faces]
    for _ in range(subdivisions):
        mids = {}

        def midpoint(a, b):
            key = (a, b) if a < b else (b, a)
            idx = mids.get(key)
            if idx is not None:
                return idx
            verts.append((verts[a] + verts[b]) / 2.0)
            mids[key] = len(verts) - 1
            return len(verts) - 1

        nxt = []
        for a, b, c in faces:
            ab, bc, ca = midpoint(a, b), midpoint(b, c), midpoint(c, a)
            nxt += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nxt
    centre = np.asarray(centre, dtype=np.float64)
    V = np.array([centre + radius * (v / np.linalg.norm(v)) for v in verts])
    F = np.array(faces, dtype=np.int32)
    v1, v2, v3 = V[F[:, 0]], V[F[:, 1]], V[F[:, 2]]
    outward = (v1 + v2 + v3) / 3.0 - centre
    flip = (np.cross(v2 - v1, v3 - v1) * outward).sum(axis=1) < 0
    F[flip] = F[flip][:, [0, 2, 1]]
    return BoundaryMesh(V, F, np.full(len(F), tag))


def mean_edge_length(mesh: BoundaryMesh) -> float:
    F = mesh.faces
    e = np.concatenate([F[:, [0, 1]], F[:, [1, 2]], F[:, [2, 0]]])
    e = np.unique(np.sort(e, axis=1), axis=0)
    d = mesh.vertices[e[:, 0]] - mesh.vertices[e[:, 1]]
    return float(np.sqrt((d * d).sum(axis=1)).mean())


def jittered_icosphere(level: int, radius: float = 0.1, jitter: float = 0.02, seed: int = 20480) -> BoundaryMesh:
    """Icosphere whose vertices are moved radially by N(0, jitter*h), h the mean
    edge length of the unjittered mesh, from a numpy PCG64 stream."""
    base = icosphere(radius, level)
    h = mean_edge_length(base)
    rng = np.random.Generator(np.random.PCG64(seed))
    dr = rng.normal(0.0, jitter * h, size=base.num_vertices)
    V = base.vertices * (1.0 + dr / radius)[:, None]
    return BoundaryMesh(V, base.faces, base.physical_tags)


def fibonacci_sphere(n: int, radius: float) -> np.ndarray:
    """Fibonacci points on a sphere (BeatEngineCore.jl:828-840)."""
    golden = math.pi * (3.0 - math.sqrt(5.0))
    out = np.empty((n, 3))
    for i in range(n):
        z = 1.0 - (2.0 * i + 1.0) / n
        r = math.sqrt(1.0 - z * z)
        phi = i * golden
        out[i] = (r * math.cos(phi) * radius, r * math.sin(phi) * radius, z * radius)
    return out


def arc_points(n: int, radius: float) -> np.ndarray:
    """n points on a half circle of the xz-plane, theta in [0, pi]."""
    t = np.linspace(0.0, math.pi, n)
    return np.stack([radius * np.sin(t), np.zeros(n), radius * np.cos(t)], axis=1)


def capped_cylinder(length: float = 0.3, radius: float = 0.08, target_faces: int = 32768,
                    grading: float = 4.0, seed: int = 32768) -> BoundaryMesh:
    """Closed capped cylinder along z, element size shrinking `grading`-fold toward
    the rim at z=length. Caps are polar-structured discs; tag 2 marks the graded
    cap at z=length, tag 1 the rest. Outward winding."""
    # Ring count along the wall and segments around; faces ~ 2*nseg*(nz + 2*ncap).
    nseg = max(8, int(round(math.sqrt(target_faces * 2.0 * math.pi * radius / (4.0 * length)))))
    # graded axial spacing: h(z) goes from h0 at z=0 to h0/grading at z=length
    def graded(n):
        s = np.linspace(0.0, 1.0, n + 1)
        g = grading
        # spacing ratio geometric with total ratio g
        u = (g ** (-s) - 1.0) / (1.0 / g - 1.0) if g != 1.0 else s
        return u
    # choose counts so total faces ~ target
    ncap = max(2, nseg // 8)
    nz = max(4, int(round(target_faces / (2.0 * nseg) - 2 * ncap)))
    zs = length * (1.0 - graded(nz)[::-1])  # dense near z=length
    zs = np.sort(zs)
    rng = np.random.Generator(np.random.PCG64(seed))
    verts = []
    rings = []
    for iz, z in enumerate(zs):
        ring = []
        off = 0.5 * (iz % 2)
        for j in range(nseg):
            th = 2.0 * math.pi * (j + off) / nseg
            ring.append(len(verts))
            verts.append((radius * math.cos(th), radius * math.sin(th), z))
        rings.append(ring)
    faces, tags = [], []
    for iz in range(nz):
        a, b = rings[iz], rings[iz + 1]
        for j in range(nseg):
            j1 = (j + 1) % nseg
            faces += [(a[j], a[j1], b[j]), (a[j1], b[j1], b[j])]
            tags += [1, 1]

    def cap(ring, z, graded_cap, tag):
        # concentric rings toward the centre
        prev = ring
        rr = np.linspace(1.0, 0.0, ncap + 1)[1:-1]
        if graded_cap:
            rr = 1.0 - (1.0 - rr) ** 1.5
        for ic, frac in enumerate(rr):
            cur = []
            for j in range(nseg):
                th = 2.0 * math.pi * (j + 0.5 * ((ic + 1) % 2)) / nseg
                cur.append(len(verts))
                verts.append((radius * frac * math.cos(th), radius * frac * math.sin(th), z))
            for j in range(nseg):
                j1 = (j + 1) % nseg
                faces.append((prev[j], prev[j1], cur[j]))
                faces.append((prev[j1], cur[j1], cur[j]))
                tags.extend([tag, tag])
            prev = cur
        c = len(verts)
        verts.append((0.0, 0.0, z))
        for j in range(nseg):
            faces.append((prev[j], prev[(j + 1) % nseg], c))
            tags.append(tag)

    cap(rings[0], zs[0], False, 1)
    cap(rings[-1], zs[-1], True, 2)
    V = np.array(verts, dtype=np.float64)
    # small seeded in-plane jitter keeps the mesh generic without moving the surface
    F = np.array(faces, dtype=np.int32)
    C = np.array([0.0, 0.0, 0.5 * length])
    v1, v2, v3 = V[F[:, 0]], V[F[:, 1]], V[F[:, 2]]
    cen = (v1 + v2 + v3) / 3.0
    outward = cen - C
    # on caps the outward direction is +-z
    cap_mask = np.abs(cen[:, 2] - zs[0]) < 1e-12
    outward[cap_mask] = [0, 0, -1]
    cap_mask = np.abs(cen[:, 2] - zs[-1]) < 1e-12
    outward[cap_mask] = [0, 0, 1]
    wall = ~(np.abs(cen[:, 2] - zs[0]) < 1e-12) & ~(np.abs(cen[:, 2] - zs[-1]) < 1e-12)
    outward[wall, 2] = 0.0
    flip = (np.cross(v2 - v1, v3 - v1) * outward).sum(axis=1) < 0
    F[flip] = F[flip][:, [0, 2, 1]]
    del rng
    return BoundaryMesh(V, F, np.array(tags))


def morphed_icosphere(level: int = 4, radius: float = 0.1, n_harmonics: int = 5, amplitude: float = 0.01,
                      seed: int = 5120) -> BoundaryMesh:
    """Icosphere displaced radially by a seeded sum of real spherical-harmonic-like
    modes r(theta, phi) = a + sum_i c_i Y_i with c_i ~ N(0, amplitude)."""
    base = icosphere(radius, level)
    rng = np.random.Generator(np.random.PCG64(seed))
    V = base.vertices
    u = V / np.linalg.norm(V, axis=1)[:, None]
    theta = np.arccos(np.clip(u[:, 2], -1, 1))
    phi = np.arctan2(u[:, 1], u[:, 0])
    disp = np.zeros(len(V))
    for _ in range(n_harmonics):
        l = int(rng.integers(1, 5))
        m = int(rng.integers(-l, l + 1))
        c = rng.normal(0.0, amplitude)
        leg = np.cos(theta) ** (l - abs(m)) * np.sin(theta) ** abs(m)
        ang = np.cos(m * phi) if m >= 0 else np.sin(-m * phi)
        disp += c * leg * ang
    r = radius + disp
    r = np.maximum(r, 0.25 * radius)
    return BoundaryMesh(u * r[:, None], base.faces, base.physical_tags)


# ---- symmetry planes (src/beat_engine/julia_local/src/BeatEngineCore.jl:448-580) ----
SYMMETRY_MODES = ("off", "x", "xy", "ground")


def normalized_symmetry_mode(mode) -> str:
    """normalized_symmetry_mode (BeatEngineCore.jl:448-454)."""
    m = str(mode).strip().lower()
    if m not in SYMMETRY_MODES:
        raise ValueError(f"Unsupported image mode: {mode}. Expected off, x, xy, or ground.")
    return m


def symmetry_active_axes(mode) -> tuple:
    m = normalized_symmetry_mode(mode)
    return (0,) if m == "x" else (0, 1) if m == "xy" else ()


def symmetry_plane_tolerance(vertices, absolute_floor: float = 1e-9, relative_tolerance: float = 1e-6) -> float:
    """max(floor, bounding-box diagonal x relative) (BeatEngineCore.jl:497-517)."""
    v = np.asarray(vertices, dtype=np.float64)
    scale = 0.0 if len(v) == 0 else float(np.linalg.norm(v.max(0) - v.min(0)))
    return max(absolute_floor, scale * relative_tolerance)


def snap_symmetry_planes(mesh: BoundaryMesh, mode, tolerance: float | None = None) -> BoundaryMesh:
    """Vertices within the tolerance of an active symmetry plane moved onto it
    (snap_symmetry_plane_vertices / snap_symmetry_planes, BeatEngineCore.jl:519-542)."""
    tol = symmetry_plane_tolerance(mesh.vertices) if tolerance is None else tolerance
    v = mesh.vertices.copy()
    for axis in symmetry_active_axes(mode):
        v[np.abs(v[:, axis]) <= tol, axis] = 0.0
    if np.array_equal(v, mesh.vertices):
        return mesh
    return BoundaryMesh(v, mesh.faces, mesh.physical_tags)


def validate_symmetry_fundamental_domain(mesh: BoundaryMesh, mode, tolerance: float | None = None) -> None:
    """The mesh must lie in the positive half-space of every active plane
    (validate_symmetry_fundamental_domain!, BeatEngineCore.jl:544-561)."""
    m = normalized_symmetry_mode(mode)
    tol = symmetry_plane_tolerance(mesh.vertices) if tolerance is None else tolerance
    for axis in symmetry_active_axes(m):
        bad = np.nonzero(mesh.vertices[:, axis] < -tol)[0]
        if len(bad):
            name = "XYZ"[axis]
            raise ValueError(f"Mesh is not in the positive {name} fundamental domain for {m.upper()} symmetry. "
                             f"Vertex {bad[0] + 1} has {name.lower()}={mesh.vertices[bad[0], axis]} m.")


def p1_symmetry_orbit_weights(mesh: BoundaryMesh, mode, tolerance: float | None = None) -> np.ndarray:
    """2 per active plane a vertex lies on (p1_symmetry_orbit_weights, BeatEngineCore.jl:563-580)."""
    tol = symmetry_plane_tolerance(mesh.vertices) if tolerance is None else tolerance
    w = np.ones(mesh.num_vertices)
    for axis in symmetry_active_axes(mode):
        w[np.abs(mesh.vertices[:, axis]) <= tol] *= 2.0
    return w

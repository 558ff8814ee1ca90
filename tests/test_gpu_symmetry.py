"""GPU parity of the symmetry / image modes against the oracle, through the C ABI.

Cases: the reference's half and quarter fixtures touching their mirror planes
(x, xy: image-touching pairs and orbit weights), the half fixture turned onto
the rigid ground plane (ground: image-touching pairs, no weights) and a sphere
away from both planes (xy: image passes only).

Bars as for the direct path: assembled system relative Frobenius <= 1e-12
(FP64) / 1e-5 (FP32); FP64 pressure <= 1e-10, FP32-LU + FP64 refinement <= 1e-8;
field <= 1e-12 (FP64) / 1e-5 (FP32)."""
import os
import sys

import numpy as np
import pytest

import oracle
from beat_engine_amd import exterior
from beat_engine_amd.mesh import BoundaryMesh, fibonacci_sphere, icosphere, load_gmsh22, snap_symmetry_planes
from beat_engine_amd.phasor import burton_miller_coupling, neumann_scale

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def weld(mesh):
    keys, remap, verts = {}, [], []
    for v in mesh.vertices:
        kk = tuple(np.round(v / 1e-9).astype(np.int64))
        if kk not in keys:
            keys[kk] = len(verts)
            verts.append(v)
        remap.append(keys[kk])
    return BoundaryMesh(np.array(verts), np.array(remap)[mesh.faces], mesh.physical_tags)


def fixture(name, mode):
    return weld(snap_symmetry_planes(load_gmsh22(os.path.join(GOLDEN, f"{name}.msh"), 0.001), mode))


def ground_half():
    m = fixture("sample_half", "x")
    return BoundaryMesh(m.vertices[:, [1, 0, 2]].copy(), m.faces[:, ::-1].copy(), m.physical_tags)


CASES = {
    "half_x": (lambda: fixture("sample_half", "x"), "x"),
    "quarter_xy": (lambda: fixture("sample_quarter", "xy"), "xy"),
    "half_ground": (ground_half, "ground"),
    "offplane_xy": (lambda: icosphere(0.1, 2, (0.4, 0.3, 0.05)), "xy"),
}


def _case(mesh, f=1500.0, nrhs=3, seed=4):
    w = 2 * np.pi * f
    k = w / 343.0
    rng = np.random.default_rng(seed)
    q = neumann_scale(1.21, w) * (rng.normal(size=(mesh.num_faces, nrhs)) + 1j * rng.normal(size=(mesh.num_faces, nrhs)))
    return k, burton_miller_coupling(k), q


@pytest.fixture(scope="module")
def eng():
    e = exterior.Engine(0)
    yield e
    e.close()


_systems = {}


def system(name):
    if name not in _systems:
        make, mode = CASES[name]
        m = make()
        k, eta, q = _case(m)
        A, B = oracle.bm_assemble(m.vertices, m.faces, k, eta, q, symmetry=mode)
        _systems[name] = (m, mode, k, eta, q, A, B)
    return _systems[name]


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("precision,tol", [("float64", 1e-12), ("float32", 1e-5)])
def test_symmetric_assembly_matches_oracle(eng, name, precision, tol):
    m, mode, k, eta, q, A_ref, B_ref = system(name)
    say("symmetric assembly", name, mode, precision, m.num_vertices, m.num_faces)
    dm = eng.upload(m, 4, 4, mode)
    A, B = dm.assemble(k, eta, q, precision)
    assert rel(A, A_ref) < tol
    assert rel(B, B_ref) < tol
    # the images change the system: the same mesh without them is far off
    if precision == "float64":
        A0, _ = eng.upload(m, 4, 4).assemble(k, eta, q, precision)
        assert rel(A0, A_ref) > 1e-3


@pytest.mark.parametrize("name", list(CASES))
def test_symmetric_counts(eng, name):
    m, mode, *_ = system(name)
    c = eng.upload(m, 4, 4, mode).counts()
    nimg = {"x": 1, "xy": 3, "ground": 1}[mode]
    assert c["image_transforms"] == nimg
    want = sum(len(oracle.image_singular_pairs(m.vertices, m.faces, mode, i, 4)) for i in range(nimg))
    assert c["image_singular_pairs"] == want
    w = oracle.orbit_weights(m.vertices, m.faces, mode)
    assert c["orbit_weighted_rows"] == int(np.count_nonzero(w != 1.0))


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("precision,tol", [("float64", 1e-10), ("float32_refined", 1e-8)])
def test_symmetric_solve_matches_oracle(eng, name, precision, tol):
    m, mode, k, eta, q, A_ref, B_ref = system(name)
    say("symmetric solve", name, mode, precision)
    p_ref = oracle.lu_solve(A_ref, B_ref)
    P, rep = eng.upload(m, 4, 4, mode).solve(k, eta, q, precision)
    assert rel(P, p_ref) < tol
    if precision == "float32_refined":
        assert rep["refinement"]["status"] == "converged"


@pytest.mark.parametrize("name", ["half_x", "quarter_xy", "half_ground"])
@pytest.mark.parametrize("precision,tol", [("float64", 1e-12), ("float32", 1e-5)])
def test_symmetric_field_matches_oracle(eng, name, precision, tol):
    m, mode, k, eta, q, A_ref, B_ref = system(name)
    say("symmetric field", name, mode, precision)
    p = oracle.lu_solve(A_ref, B_ref)
    pts = fibonacci_sphere(40, 1.5) + np.array([0.0, 0.6 if mode == "ground" else 0.0, 0.0])
    ref = oracle.field(m.vertices, m.faces, k, p, q, pts, symmetry=mode)
    out = eng.upload(m, 4, 4, mode).field(k, p, q, pts, precision)
    assert rel(out, ref) < tol

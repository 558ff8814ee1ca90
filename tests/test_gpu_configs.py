"""The precision-form workload (graded capped cylinder, 0.3 m x 0.08 m, element size
shrinking 4x toward one rim, ~32768 triangles; 64 right-hand sides of random
complex normal velocity over the graded end cap; field at 10,000 points on a 2 m
hemisphere) on the GPU.

At full size the checks are size-independent properties: the FP32-LU + FP64
refinement solution reaches the 1e-10 relative residual (status converged) and
agrees with the complex128 GPU solution within 1e-8; the FP32 field of it agrees
with the FP64 field within 1e-5. At a reduced size of the same construction the
whole chain is compared with the oracle (refined pressure 1e-8, field 1e-12)."""
import sys

import numpy as np
import pytest

import oracle
from beat_engine_amd import exterior
from beat_engine_amd.mesh import capped_cylinder
from beat_engine_amd.phasor import burton_miller_coupling, neumann_scale

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def hemisphere(n, radius):
    """n Fibonacci points on the upper (z > 0) hemisphere, centred on the cylinder axis."""
    i = np.arange(n) + 0.5
    z = i / n
    r = np.sqrt(1 - z * z)
    phi = np.pi * (3 - np.sqrt(5)) * i
    return radius * np.stack([r * np.cos(phi), r * np.sin(phi), z], axis=1) + np.array([0.0, 0.0, 0.15])


def cap_drive(mesh, nrhs, f, seed=64):
    w = 2 * np.pi * f
    rng = np.random.default_rng(seed)
    v = np.zeros((mesh.num_faces, nrhs), complex)
    cap = mesh.physical_tags == 2
    v[cap] = rng.normal(size=(cap.sum(), nrhs)) + 1j * rng.normal(size=(cap.sum(), nrhs))
    return w / 343.0, neumann_scale(1.21, w) * v


@pytest.fixture(scope="module")
def eng():
    e = exterior.Engine(0)
    yield e
    e.close()


def test_precision_form_reduced_size_against_oracle(eng):
    mesh = capped_cylinder(target_faces=2000)
    k, q = cap_drive(mesh, 6, 2000.0)
    eta = burton_miller_coupling(k)
    say("graded cylinder (reduced)", mesh.num_vertices, mesh.num_faces)
    A, B = oracle.bm_assemble(mesh.vertices, mesh.faces, k, eta, q)
    pr = oracle.lu_solve(A, B)
    dm = eng.upload(mesh)
    P, rep = dm.solve(k, eta, q, "float32_refined", 3, 1e-10)
    assert rep["refinement"]["status"] == "converged"
    assert rel(P, pr) < 1e-8
    pts = hemisphere(500, 2.0)
    assert rel(dm.field(k, P, q, pts, "float64"), oracle.field(mesh.vertices, mesh.faces, k, pr, q, pts)) < 1e-12


def test_precision_form_full_size(eng):
    mesh = capped_cylinder(target_faces=32768)
    k, q = cap_drive(mesh, 64, 2000.0)
    eta = burton_miller_coupling(k)
    say("graded cylinder", mesh.num_vertices, mesh.num_faces, "64 rhs")
    dm = eng.upload(mesh)
    P32, rep = dm.solve(k, eta, q, "float32_refined", 3, 1e-10)
    say("refined", rep["assembly_s"], rep["factor_s"], rep["solve_s"], rep["refinement"]["steps"])
    assert rep["refinement"]["status"] == "converged"
    assert np.all(rep["refinement"]["working_relative_residuals"] <= 1e-10)
    P64, rep64 = dm.solve(k, eta, q, "float64")
    say("fp64", rep64["assembly_s"], rep64["factor_s"], rep64["solve_s"])
    assert rel(P32, P64) < 1e-8
    for j in range(0, 64, 9):
        assert rel(P32[:, j], P64[:, j]) < 1e-8
    pts = hemisphere(10000, 2.0)
    f64 = dm.field(k, P32[:, :8], q[:, :8], pts, "float64")
    f32 = dm.field(k, P32[:, :8], q[:, :8], pts, "float32")
    assert rel(f32, f64) < 1e-5
    assert np.all(np.isfinite(f64))

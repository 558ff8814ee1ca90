"""The regular assembly pass over unordered tile pairs against the pass over ordered
tile pairs (the two kernels behind beat_ctx_regular_symmetric), and against the FP64
oracle's Burton-Miller system.

The unordered pass evaluates r, e^{ikr}/r and G' once per point pair and adds both
orders of the face pair; the ordered pass evaluates each order separately. They sum
the same terms in another order, so the matrices agree to rounding: relative
Frobenius difference 1e-13 for complex128, 1e-5 for complex64. The cases cover a
purely imaginary coupling (the reference's eta = i/k) and a general complex one,
meshes of one test tile and of many, the table phase and the wide-phase path
(|k| times the mesh diameter above the table's bound), and 1, 3 and 6 rule points."""
import sys

import numpy as np
import pytest

import oracle
from beat_engine_amd import exterior
from beat_engine_amd.mesh import capped_cylinder, icosphere, jittered_icosphere
from beat_engine_amd.phasor import burton_miller_coupling

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def eng():
    e = exterior.Engine(0)
    yield e
    e.regular_symmetric(True)
    e.close()


def both(eng, mesh, k, eta, q, prec, rule=4):
    dm = eng.upload(mesh, rule, 4)
    eng.regular_symmetric(False)
    A0, B0 = dm.assemble(k, eta, q, prec)
    eng.regular_symmetric(True)
    A1, B1 = dm.assemble(k, eta, q, prec)
    return A0, B0, A1, B1


CASES = [
    ("L1 icosphere, one test tile", lambda: icosphere(0.1, 1), 800.0, None),
    ("L3 jittered icosphere", lambda: jittered_icosphere(3, 0.1, 0.02, seed=3), 1500.0, None),
    ("capped cylinder, complex coupling", lambda: capped_cylinder(target_faces=1500, seed=7), 900.0,
     0.3 + 0.8j),
]


@pytest.mark.parametrize("name,make,f,eta", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("prec,tol", [("float64", 1e-13), ("float32", 1e-5)])
def test_unordered_pass_matches_ordered_pass(eng, name, make, f, eta, prec, tol):
    mesh = make()
    k = 2 * np.pi * f / 343.0
    eta = burton_miller_coupling(k) if eta is None else eta
    rng = np.random.default_rng(mesh.num_faces)
    q = rng.normal(size=(mesh.num_faces, 3)) + 1j * rng.normal(size=(mesh.num_faces, 3))
    A0, B0, A1, B1 = both(eng, mesh, k, eta, q, prec)
    ea, eb = rel(A1, A0), rel(B1, B0)
    say(f"{name} {prec}: A {ea:.2e}, B {eb:.2e}")
    assert ea <= tol and eb <= tol


@pytest.mark.parametrize("rule", [1, 2, 4])
def test_rule_orders(eng, rule):
    mesh = jittered_icosphere(2, 0.1, 0.02, seed=11)
    k = 2 * np.pi * 2000.0 / 343.0
    eta = burton_miller_coupling(k)
    q = np.ones((mesh.num_faces, 1), dtype=np.complex128)
    A0, B0, A1, B1 = both(eng, mesh, k, eta, q, "float64", rule)
    say(f"rule order {rule}: A {rel(A1, A0):.2e}, B {rel(B1, B0):.2e}")
    assert rel(A1, A0) <= 1e-13 and rel(B1, B0) <= 1e-13


def test_wide_phase(eng):
    # |k| diameter above the table bound (4e5): the pi/2 reduction with the library fallback
    mesh = icosphere(0.1, 2)
    k = 3.0e6
    eta = 1j / k
    q = np.ones((mesh.num_faces, 2), dtype=np.complex128)
    A0, B0, A1, B1 = both(eng, mesh, k, eta, q, "float64")
    say(f"wide phase: A {rel(A1, A0):.2e}, B {rel(B1, B0):.2e}")
    assert rel(A1, A0) <= 1e-12 and rel(B1, B0) <= 1e-12


def test_unordered_pass_matches_the_oracle(eng):
    mesh = jittered_icosphere(2, 0.1, 0.02, seed=2)
    k = 2 * np.pi * 1200.0 / 343.0
    eta = burton_miller_coupling(k)
    rng = np.random.default_rng(2)
    q = rng.normal(size=(mesh.num_faces, 2)) + 1j * rng.normal(size=(mesh.num_faces, 2))
    A, B = oracle.bm_assemble(mesh.vertices, mesh.faces, k, eta, q)
    eng.regular_symmetric(True)
    Ag, Bg = eng.upload(mesh, 4, 4).assemble(k, eta, q, "float64")
    say(f"oracle: A {rel(Ag, A):.2e}, B {rel(Bg, B):.2e}")
    assert rel(Ag, A) <= 1e-12 and rel(Bg, B) <= 1e-12

"""Pins for the oracle's Burton-Miller path: operator identities with closed forms,
the reference's fused == four-operator invariant, the analytic pulsating sphere,
the reference's golden sweep norms and the refinement known answers."""
import json
import os

import numpy as np
import pytest

import oracle
from beat_engine_amd.mesh import fibonacci_sphere, icosphere, load_gmsh22

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module")
def l2():
    return icosphere(0.1, 2)


@pytest.fixture(scope="module")
def l3():
    return icosphere(0.1, 3)


def test_identity_blocks_closed_form(l2):
    Mpp, Mpd = oracle.identity(l2.vertices, l2.faces)
    # element mass blocks: A/6 diagonal, A/12 off-diagonal, A/3 for P1 x DP0
    F = l2.faces
    ref_pp = np.zeros_like(Mpp)
    ref_pd = np.zeros_like(Mpd)
    for e, f in enumerate(F):
        A = l2.areas[e]
        for a in range(3):
            ref_pd[f[a], e] += A / 3
            for b in range(3):
                ref_pp[f[a], f[b]] += A / 6 if a == b else A / 12
    assert np.abs(Mpp - ref_pp).max() <= 1e-14 * np.abs(ref_pp).max()
    assert np.abs(Mpd - ref_pd).max() <= 1e-14 * np.abs(ref_pd).max()


def test_operator_closed_forms(l3):
    """k -> 0 identities on a sphere of radius a: S.1 = a M.1 (shell potential),
    D.1 = -1/2 M.1 (Gauss), H.1 = O(k^2); each pins one operator's sign."""
    k = 1e-4
    S, D, K, H = oracle.operators(l3.vertices, l3.faces, k)
    Mpp, Mpd = oracle.identity(l3.vertices, l3.faces)
    m1 = Mpd.sum(axis=1)
    assert np.abs(S.sum(axis=1).real - 0.1 * m1).max() < 5e-3 * (0.1 * m1).max()
    assert np.abs(D.sum(axis=1).real + 0.5 * m1).max() < 5e-3 * (0.5 * m1).max()
    assert np.abs(H.sum(axis=1)).max() < 1e-6 * np.abs(H).max()
    # D is the k-derivative-free part: imaginary parts vanish like k^2
    assert np.abs(D.imag).max() < 1e-6 * np.abs(D.real).max()


def test_green_kernel_finite_difference():
    """D's kernel G (ik - 1/r) (r.n_y)/r is dG/dn_y; K' is dG/dn_x (BeatEngineCore.jl:875-903)."""
    k = 7.3
    x = np.array([0.01, -0.02, 0.03])
    y = np.array([0.05, 0.04, -0.01])
    n = np.array([0.3, -0.5, 0.8]); n /= np.linalg.norm(n)
    G = lambda x, y: np.exp(1j * k * np.linalg.norm(y - x)) / (4 * np.pi * np.linalg.norm(y - x))
    r = y - x; R = np.linalg.norm(r)
    dl = G(x, y) * (1j * k - 1 / R) * (r @ n) / R
    h = 1e-6
    fd = (G(x, y + h * n) - G(x, y - h * n)) / (2 * h)
    assert dl == pytest.approx(fd, rel=1e-8)
    adl = -G(x, y) * (1j * k - 1 / R) * (r @ n) / R
    fdx = (G(x + h * n, y) - G(x - h * n, y)) / (2 * h)
    assert adl == pytest.approx(fdx, rel=1e-8)


@pytest.mark.parametrize("precision,tol", [("f64", 1e-13), ("f32", 1e-5)])
def test_fused_equals_four_operators(l2, precision, tol):
    """runtests.jl:210-271: LHS and RHS of the fused path equal the four-operator
    combination (relative Frobenius), with 3 drives."""
    k = 2 * np.pi * 700 / 343.0
    eta = oracle.burton_miller_coupling(k)
    rng = np.random.default_rng(3)
    q = rng.normal(size=(l2.num_faces, 3)) + 1j * rng.normal(size=(l2.num_faces, 3))
    A, B = oracle.bm_assemble(l2.vertices, l2.faces, k, eta, q, precision=precision)
    S, D, K, H = oracle.operators(l2.vertices, l2.faces, k, precision=precision)
    Mpp, Mpd = oracle.identity(l2.vertices, l2.faces)
    L, R = oracle.bm_from_operators(S.astype(complex), D.astype(complex), K.astype(complex), H.astype(complex), Mpp, Mpd, eta, q)
    assert np.linalg.norm(A - L) / np.linalg.norm(L) < tol
    assert np.linalg.norm(B - R) / np.linalg.norm(R) < tol
    pA = oracle.lu_solve(A.astype(complex), B.astype(complex))
    pL = oracle.lu_solve(L, R)
    assert np.linalg.norm(pA - pL) / np.linalg.norm(pL) < 10 * tol


def _analytic_case(mesh, convention, q_convention=None):
    f, a, rho, c = 1000.0, 0.1, 1.2041, 343.0
    w = 2 * np.pi * f
    k = oracle.outgoing_wavenumber(w / c, convention)
    eta = oracle.burton_miller_coupling(k, convention)
    q = np.full(mesh.num_faces, oracle.neumann_scale(rho, w, q_convention or convention))
    A, B = oracle.bm_assemble(mesh.vertices, mesh.faces, k, eta, q)
    p = oracle.lu_solve(A, B)
    pts = fibonacci_sphere(128, 3.0)
    fld = oracle.field(mesh.vertices, mesh.faces, k, p, q, pts)[:, 0]
    ka = 1j * k * a
    ref = np.array([rho * c * (a / np.linalg.norm(x)) * ka / (ka - 1) * np.exp(1j * k * (np.linalg.norm(x) - a)) for x in pts])
    return p[:, 0], fld, ref


def test_analytic_pulsating_sphere(l3):
    """analytical_exterior_tests.jl:128-155: 0.05 dB / 0.5 deg at 3 m, both phasor
    conventions. With the drive fixed and the kernel's convention flipped, real
    velocity gives -conj(p) (:209-212); flipping the drive too gives conj(p)."""
    out = {}
    for conv in (oracle.NEGATIVE_TIME_PHASOR, oracle.POSITIVE_TIME_PHASOR):
        p, fld, ref = _analytic_case(l3, conv)
        ratio = fld / ref
        assert np.max(np.abs(20 * np.log10(np.abs(ratio)))) <= 0.05
        assert np.max(np.abs(np.degrees(np.angle(ratio)))) <= 0.5
        out[conv] = (p, fld)
    pn, fn = out[oracle.NEGATIVE_TIME_PHASOR]
    pp, fp = out[oracle.POSITIVE_TIME_PHASOR]
    assert np.allclose(fp, np.conj(fn), rtol=1e-12, atol=0)
    assert np.allclose(pp, np.conj(pn), rtol=1e-12, atol=0)
    _, wrong, ref = _analytic_case(l3, oracle.POSITIVE_TIME_PHASOR, q_convention=oracle.NEGATIVE_TIME_PHASOR)
    assert np.allclose(wrong, -np.conj(fn), rtol=1e-12, atol=0)
    assert np.max(np.abs(np.degrees(np.angle(wrong / ref)))) >= 5.0


def test_superposition_and_replay(l2):
    """analytical_exterior_tests.jl:169-183: field is linear in (p, q) to rtol 1e-12."""
    k = 2 * np.pi * 800 / 343.0
    eta = oracle.burton_miller_coupling(k)
    rng = np.random.default_rng(5)
    q1 = rng.normal(size=l2.num_faces) + 1j * rng.normal(size=l2.num_faces)
    q2 = rng.normal(size=l2.num_faces) + 1j * rng.normal(size=l2.num_faces)
    amp = 0.3 - 0.7j
    A, B = oracle.bm_assemble(l2.vertices, l2.faces, k, eta, np.stack([q1, q2, q1 + amp * q2], axis=1))
    P = oracle.lu_solve(A, B)
    pts = fibonacci_sphere(32, 3.0)
    Q = np.stack([q1, q2, q1 + amp * q2], axis=1)
    F = oracle.field(l2.vertices, l2.faces, k, P, Q, pts)
    assert np.allclose(F[:, 2], F[:, 0] + amp * F[:, 1], rtol=1e-12, atol=1e-12 * np.abs(F).max())
    Fr = oracle.field(l2.vertices, l2.faces, k, amp * P[:, 0], amp * q1, pts)[:, 0]
    assert np.allclose(Fr, amp * F[:, 0], rtol=1e-12, atol=0)


@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_golden_sweep_norms(precision):
    """results/baseline_sweep_cpu.json (benchmark_sweep.jl defaults): sample.msh at
    scale 0.001, throat tag 2 driven with q = i*1.21*omega, q2/s2, field at 74
    Fibonacci points on a 2 m sphere; CI accepts 1e-3 relative."""
    m = load_gmsh22(os.path.join(GOLDEN, "sample.msh"), 0.001)
    g = json.load(open(os.path.join(GOLDEN, "baseline_sweep_cpu.json")))
    pts = fibonacci_sphere(74, 2.0)
    rows = g["frequencies"] if precision == "f32" else g["frequencies"][::3]
    for row in rows:
        w = 2 * np.pi * row["frequency_hz"]
        k = w / 343.0
        q = np.zeros(m.num_faces, complex)
        q[m.physical_tags == 2] = 1j * 1.21 * w
        A, B = oracle.bm_assemble(m.vertices, m.faces, k, oracle.burton_miller_coupling(k), q, quad_order=2,
                                  singular_order=2, precision=precision)
        p = oracle.lu_solve(A, B)
        fl = oracle.field(m.vertices, m.faces, k, p, q, pts, quad_order=2, precision=precision)
        assert np.linalg.norm(p) == pytest.approx(row["pressure_norm"], rel=1e-4)
        assert np.linalg.norm(fl) == pytest.approx(row["field_norm"], rel=1e-4)


def test_refinement_known_answers():
    """cpu_refinement_tests.jl:91-129."""
    I3 = np.eye(3, dtype=np.complex64)
    exact = np.array([1 + 2j, 3 - 4j, 0], dtype=np.complex64)
    x, r = oracle.refine_cpu_lu(I3, exact)
    assert np.array_equal(x, exact) and r["steps"] == 0 and r["status"] == "converged"
    x, r = oracle.refine_cpu_lu(I3, exact, residual_matrix=np.zeros((3, 3), complex))
    assert np.array_equal(x, exact) and r["status"] == "stagnated" and r["steps"] == 1
    assert [h.tolist() for h in r["history"]] == [[1.0], [1.0]]
    x, r = oracle.refine_cpu_lu(I3, np.zeros((3, 2), np.complex64))
    assert not x.any() and r["steps"] == 0 and r["status"] == "converged"
    I2 = np.eye(2, dtype=np.complex64)
    res_op = np.array([[-1, 0], [0, 0.5]], dtype=complex)
    x, r = oracle.refine_cpu_lu(I2, I2, residual_matrix=res_op)
    assert np.array_equal(x, np.array([[1, 0], [0, 1.875]], dtype=np.complex64))
    assert r["status"] == r["working_status"] == "step_limit" and r["steps"] == 3
    assert [h.tolist() for h in r["history"]] == [[2, .5], [2, .25], [2, .125], [2, .0625]]
    tri = np.array([[1, 3], [0, 7]], dtype=np.complex64)
    x, r = oracle.refine_cpu_lu(tri, np.ones(2, np.complex64), max_steps=1)
    assert r["steps"] == 1 and r["working_status"] == "converged" and r["status"] == "rounding_limited"
    with pytest.raises(ValueError):
        oracle.refine_cpu_lu(I2, I2, max_steps=4)

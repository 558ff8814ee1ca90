"""The speculative factorization: leaves factored without row exchanges while
every column is checked against partial pivoting's choice (izamax: the diagonal
must be the first maximum of |re| + |im| among the rows at and below it).

- Column-dominant matrices, where LAPACK zgetrf exchanges nothing: the factors
  equal scipy.linalg.lu_factor's (identity pivots, 1e-12), across several outer
  blocks and a ragged last leaf.
- The same matrices with two rows exchanged (inside one leaf, across leaves,
  across outer blocks): the check fails, the factorization is redone with the
  exchange panels, and pivots and factors still equal zgetrf's.
- A Burton-Miller solve with the check forced to fail takes the exchange path
  (factorization_count 2) and returns the pressures of the speculative one."""
import sys

import numpy as np
import pytest
import scipy.linalg

import oracle
from beat_engine_amd import exterior
from beat_engine_amd.mesh import jittered_icosphere
from beat_engine_amd.phasor import burton_miller_coupling, neumann_scale

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def eng():
    e = exterior.Engine(0)
    yield e
    e.close()


def dominant(n, seed):
    """Complex matrix whose diagonal exceeds every other entry of its column by far
    in |re| + |im| (so partial pivoting never exchanges: the dominance survives
    elimination)."""
    rng = np.random.default_rng(seed)
    A = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    A += np.diag(4.0 * n * np.exp(1j * rng.uniform(0, 2 * np.pi, n)))
    return A


@pytest.mark.parametrize("n", [5, 64, 200, 777, 1300])
@pytest.mark.parametrize("precision", ["float64", "float32"])
def test_dominant_matrix_factors_equal_getrf(eng, n, precision):
    say("dominant", n, precision)
    A = dominant(n, n)
    lu_r, piv_r = scipy.linalg.lu_factor(A)
    assert np.array_equal(piv_r, np.arange(n))  # LAPACK exchanges nothing here
    lu, piv, info = eng.dense_lu_factor(A, precision)
    assert info == 0
    assert np.array_equal(piv, piv_r)
    assert rel(lu, lu_r) < (1e-12 if precision == "float64" else 2e-6)
    B = np.ones((n, 3), complex)
    X = eng.dense_lu_solve(A, B, precision)
    assert rel(A @ X, B) < (1e-12 if precision == "float64" else 1e-5)


@pytest.mark.parametrize("r1,r2", [(130, 150), (100, 190), (300, 1250), (1290, 1299)])
def test_planted_exchange_falls_back_to_getrf(eng, r1, r2):
    say("planted exchange", r1, r2)
    n = 1300
    A = dominant(n, 7)
    A[[r1, r2]] = A[[r2, r1]]  # partial pivoting must exchange them back at column r1
    lu_r, piv_r = scipy.linalg.lu_factor(A)
    assert piv_r[r1] == r2
    lu, piv, info = eng.dense_lu_factor(A, "float64")
    assert np.array_equal(piv, piv_r)
    assert rel(lu, lu_r) < 1e-12
    B = np.ones((n, 2), complex)
    X = eng.dense_lu_solve(A, B, "float64")
    assert rel(A @ X, B) < 1e-12


@pytest.mark.parametrize("precision,tol", [("float64", 1e-10), ("float32_refined", 1e-8), ("float32", 2e-4)])
def test_forced_exchange_path_in_a_solve(eng, precision, tol):
    mesh = jittered_icosphere(3, 0.1, 0.02, seed=11)
    f = 1200.0
    w = 2 * np.pi * f
    k = w / 343.0
    eta = burton_miller_coupling(k)
    rng = np.random.default_rng(12)
    q = neumann_scale(1.21, w) * (rng.normal(size=(mesh.num_faces, 3)) + 1j * rng.normal(size=(mesh.num_faces, 3)))
    A, B = oracle.bm_assemble(mesh.vertices, mesh.faces, k, eta, q)
    p_cpu = oracle.lu_solve(A, B)
    dm = eng.upload(mesh)
    P0, rep0 = dm.solve(k, eta, q, precision)
    assert rep0["factorization_count"] == 1  # the speculation holds on this system
    eng.debug_force_exchange(True)
    try:
        P1, rep1 = dm.solve(k, eta, q, precision)
        Ps, reps = dm.sweep([k, k], [eta, eta], [q, q], precision)
    finally:
        eng.debug_force_exchange(False)
    assert rep1["factorization_count"] == 2
    assert all(r["solved"] for r in reps)
    for P in (P0, P1, Ps[0], Ps[1]):
        assert rel(P, p_cpu) <= tol
    assert rel(P1, P0) <= (1e-12 if precision != "float32" else 1e-5)

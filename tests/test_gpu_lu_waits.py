"""The LU's cross-workgroup exchanges: exact pivot candidates and bounded waits.

A column whose two largest |re| + |im| lie in different workgroups (rows more
than 64 apart) and differ only in the last mantissa bit must be pivoted as
LAPACK's izamax pivots it (scipy.linalg.lu_factor, OpenBLAS 0.3.29). The give-up
path of the bounded waits is forced from the host (the context's test hook sets
every factorization's abort word before its first kernel), so every kernel ends
normally: the call must raise, and the next call on the context must be correct."""
import sys

import numpy as np
import pytest
import scipy.linalg

from beat_engine_amd import _native, exterior
from beat_engine_amd.mesh import icosphere
from beat_engine_amd.phasor import burton_miller_coupling, neumann_scale

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def eng():
    e = exterior.Engine(0)
    yield e
    e.close()


@pytest.mark.parametrize("n,lo,hi,winner", [(200, 3, 150, "hi"), (200, 3, 150, "lo"), (700, 20, 650, "hi"),
                                            (700, 130, 600, "lo")])
def test_cross_workgroup_near_tie_pivots_like_izamax(eng, n, lo, hi, winner):
    """Column 0 holds 3.0 and the next double below it in rows `lo` and `hi`; the
    row with the larger value must win (the smaller row only on an exact tie)."""
    say("near tie", n, lo, hi, winner)
    rng = np.random.default_rng(n + lo)
    A = (rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))) * 0.25
    A[:, 0] = (rng.uniform(-0.5, 0.5, n) + 1j * rng.uniform(-0.5, 0.5, n))
    big, below = 3.0, np.nextafter(3.0, 0.0)
    A[lo, 0] = big if winner == "lo" else below
    A[hi, 0] = big if winner == "hi" else below
    lu, piv, info = eng.dense_lu_factor(A, "float64")
    lu_r, piv_r = scipy.linalg.lu_factor(A)
    assert info == 0
    assert piv[0] == (lo if winner == "lo" else hi)
    assert np.array_equal(piv, piv_r)
    assert np.linalg.norm(lu - lu_r) / np.linalg.norm(lu_r) < 1e-12


def test_exact_tie_across_workgroups_takes_the_first_row(eng):
    n = 300
    rng = np.random.default_rng(9)
    A = (rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))) * 0.25
    A[:, 0] = rng.uniform(-0.5, 0.5, n)
    A[40, 0] = 2.0 + 1.0j  # |re| + |im| = 3
    A[250, 0] = -3.0      # ties at 3: izamax takes the first
    lu, piv, info = eng.dense_lu_factor(A, "float64")
    assert piv[0] == 40
    assert np.array_equal(piv, scipy.linalg.lu_factor(A)[1])


@pytest.mark.parametrize("precision", ["float64", "float32"])
def test_abandoned_waits_raise_and_the_context_recovers(eng, precision):
    n = 400  # seven panel workgroups, seven solve blocks
    rng = np.random.default_rng(4)
    A = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n)) + 30 * np.eye(n)
    B = rng.normal(size=(n, 2)) + 1j * rng.normal(size=(n, 2))
    eng.debug_abort_waits(True)
    try:
        with pytest.raises(_native.BeatError, match="gave up"):
            eng.dense_lu_solve(A, B, precision)
        with pytest.raises(_native.BeatError, match="gave up"):
            eng.dense_lu_factor(A, precision)
    finally:
        eng.debug_abort_waits(False)
    X = eng.dense_lu_solve(A, B, precision)
    assert np.linalg.norm(A @ X - B) / np.linalg.norm(B) < (1e-12 if precision == "float64" else 1e-5)


def test_abandoned_waits_in_a_refined_solve(eng):
    mesh = icosphere(0.1, 3)
    f = 800.0
    k = 2 * np.pi * f / 343.0
    eta = burton_miller_coupling(k)
    rng = np.random.default_rng(3)
    q = neumann_scale(1.21, 2 * np.pi * f) * (rng.normal(size=(mesh.num_faces, 2)) + 0j)
    dm = eng.upload(mesh)
    P0, _ = dm.solve(k, eta, q, "float32_refined")
    eng.debug_abort_waits(True)
    try:
        with pytest.raises(_native.BeatError, match="gave up"):
            dm.solve(k, eta, q, "float32_refined")
        with pytest.raises(_native.BeatError, match="gave up"):
            dm.sweep([k, k], [eta, eta], [q, q], "float64")
    finally:
        eng.debug_abort_waits(False)
    P1, rep = dm.solve(k, eta, q, "float32_refined")
    assert rep["refinement"]["status"] == "converged"
    assert np.linalg.norm(P1 - P0) / np.linalg.norm(P0) < 1e-12

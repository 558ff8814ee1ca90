"""The grouped batching path: jobs of small systems in groups of up to 32, each
group's systems assembled, then factored (batched speculative LU: one launch per
kernel and leaf for every system), solved and refined (complex128 residuals,
per-column acceptance decided on the device) by one captured hipGraph, replayed
for every later group of the same shape.

Every job equals its per-job solve (complex128 1e-11, refined 1e-9, complex64
5e-5: the grouped factorization updates leaf by leaf instead of by 256-column
blocks, so rounding differs) and, for some jobs, the oracle (1e-10 / 1e-8).
More than one group, mesh-major and mixed orders, a replay of a captured graph,
and the re-solve of jobs whose speculative factorization reports an exchange."""
import sys

import numpy as np
import pytest

import oracle
from beat_engine_amd import exterior
from beat_engine_amd.mesh import capped_cylinder, icosphere, jittered_icosphere
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


def make_jobs(eng, meshes, freqs, nrhs, seed, shuffle):
    rng = np.random.default_rng(seed)
    dms = [eng.upload(m) for m in meshes]
    J = []
    for m, dm in zip(meshes, dms):
        for f in freqs:
            w = 2 * np.pi * f
            k = w / 343.0
            v = rng.normal(size=(m.num_faces, nrhs)) + 1j * rng.normal(size=(m.num_faces, nrhs))
            J.append((m, dm, k, burton_miller_coupling(k), neumann_scale(1.21, w) * v))
    if shuffle:
        J = [J[i] for i in rng.permutation(len(J))]
    return J


TOL_SINGLE = {"float64": 1e-11, "float32_refined": 1e-9, "float32": 5e-5}
TOL_ORACLE = {"float64": 1e-10, "float32_refined": 1e-8, "float32": 1e-4}


@pytest.mark.parametrize("precision", ["float32_refined", "float64", "float32"])
def test_grouped_batch_matches_per_job_solves(eng, precision):
    meshes = [jittered_icosphere(2, 0.1, 0.02, seed=3), capped_cylinder(target_faces=600, seed=4),
              jittered_icosphere(3, 0.1, 0.02, seed=5), capped_cylinder(target_faces=1500, seed=6)]
    J = make_jobs(eng, meshes, [60.0, 400.0, 900.0, 2000.0, 3500.0, 5000.0, 7000.0, 8000.0, 9000.0], 4, 12, True)
    say("grouped batch", precision, len(J), "jobs")
    P, info = eng.batch([j[1] for j in J], [j[2] for j in J], [j[3] for j in J], [j[4] for j in J], precision)
    worst = 0.0
    for i, (m, dm, k, eta, q) in enumerate(J):
        single, _ = dm.solve(k, eta, q, precision)
        e = rel(P[i], single)
        worst = max(worst, e)
        assert e < TOL_SINGLE[precision], (i, e)
        assert info[i]["factorization_count"] == 1
        if precision == "float32_refined":
            assert info[i]["refinement_status"] == "converged"
            assert 1 <= info[i]["refinement_steps"] <= 2
    say("worst vs per-job", worst)
    for i in (0, len(J) // 2, len(J) - 1):
        m, dm, k, eta, q = J[i]
        A, B = oracle.bm_assemble(m.vertices, m.faces, k, eta, q)
        assert rel(P[i], oracle.lu_solve(A, B)) < TOL_ORACLE[precision]


def test_graph_replay_and_group_shapes(eng):
    """One mesh at 40 frequencies (groups of 32 and 8), twice: the second call
    replays both captured graphs."""
    J = make_jobs(eng, [icosphere(0.1, 2)], np.linspace(100.0, 6000.0, 40), 3, 21, False)
    args = ([j[1] for j in J], [j[2] for j in J], [j[3] for j in J], [j[4] for j in J], "float32_refined")
    P1, info1 = eng.batch(*args)
    P2, info2 = eng.batch(*args)
    for i, (m, dm, k, eta, q) in enumerate(J):
        assert rel(P2[i], P1[i]) < 1e-12
        assert info1[i]["refinement_status"] == info2[i]["refinement_status"] == "converged"
    for i in (0, 31, 32, 39):
        m, dm, k, eta, q = J[i]
        single, _ = dm.solve(k, eta, q, "float32_refined")
        assert rel(P2[i], single) < TOL_SINGLE["float32_refined"]


@pytest.mark.parametrize("precision", ["float32_refined", "float64"])
def test_exchange_reported_in_a_group_resolves_that_job(eng, precision):
    J = make_jobs(eng, [jittered_icosphere(2, 0.1, 0.02, seed=8), capped_cylinder(target_faces=700, seed=9)],
                  [300.0, 3000.0], 2, 31, False)
    eng.debug_force_exchange(True)
    try:
        P, info = eng.batch([j[1] for j in J], [j[2] for j in J], [j[3] for j in J], [j[4] for j in J], precision)
    finally:
        eng.debug_force_exchange(False)
    for i, (m, dm, k, eta, q) in enumerate(J):
        assert info[i]["factorization_count"] == 2
        single, _ = dm.solve(k, eta, q, precision)
        assert rel(P[i], single) < TOL_SINGLE[precision]

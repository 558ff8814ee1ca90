"""Batching form at reduced scale: meshes of uneven size (jittered icospheres and
graded capped cylinders) at several frequencies each, 4 right-hand sides, solved as
one batch (next job assembled while the current one is factored) in complex64 with
refinement and in complex128; every job equals its single-solve result and, for a
subset, the oracle. The autotuner's choices apply to later batches."""
import sys

import numpy as np
import pytest

import oracle
from beat_engine_amd import exterior
from beat_engine_amd.mesh import capped_cylinder, jittered_icosphere
from beat_engine_amd.phasor import burton_miller_coupling, neumann_scale

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def jobs():
    meshes = [jittered_icosphere(2, 0.1, 0.02, seed=3), capped_cylinder(target_faces=600, seed=4),
              jittered_icosphere(3, 0.1, 0.02, seed=5), capped_cylinder(target_faces=1500, seed=6)]
    freqs = [60.0, 900.0, 7000.0]
    eng = exterior.Engine(0)
    dms = [eng.upload(m) for m in meshes]
    rng = np.random.default_rng(12)
    J = []
    for m, dm in zip(meshes, dms):
        for f in freqs:
            w = 2 * np.pi * f
            k = w / 343.0
            v = rng.normal(size=(m.num_faces, 4)) + 1j * rng.normal(size=(m.num_faces, 4))
            J.append((m, dm, k, burton_miller_coupling(k), neumann_scale(1.21, w) * v))
    order = rng.permutation(len(J))  # mixed sizes back to back
    yield [J[i] for i in order], eng
    eng.close()


@pytest.mark.parametrize("precision,tol", [("float32_refined", 1e-8), ("float64", 1e-10)])
def test_batch_matches_single_solves_and_oracle(jobs, precision, tol):
    J, eng = jobs
    print("batch", precision, len(J), file=sys.stderr, flush=True)
    P, info = eng.batch([j[1] for j in J], [j[2] for j in J], [j[3] for j in J], [j[4] for j in J], precision)
    for i, (m, dm, k, eta, q) in enumerate(J):
        single, _ = dm.solve(k, eta, q, precision)
        # the batch factors these small systems in groups (leaf-by-leaf updates),
        # the single solve by 256-column blocks: the roundings differ
        assert rel(P[i], single) < (1e-9 if precision == "float32_refined" else 1e-11), i
        if precision == "float32_refined":
            assert info[i]["refinement_status"] == "converged"
    for i in (0, len(J) - 1):
        m, dm, k, eta, q = J[i]
        A, B = oracle.bm_assemble(m.vertices, m.faces, k, eta, q)
        assert rel(P[i], oracle.lu_solve(A, B)) < tol


def test_autotune_then_batch(jobs):
    J, eng = jobs
    m, dm, k, eta, q = J[0]
    choice = eng.autotune(dm, "float32_refined", k, 4)
    print("autotune", choice, file=sys.stderr, flush=True)
    assert choice["lu_layout"] in exterior.LU_LAYOUTS and choice["batch_assembly"] in exterior.ASM_OCCUPANCIES
    assert choice["gemm_tile"] in exterior.GEMM_TILES and choice["regular_tiles"] in exterior.REGULAR_TILES
    assert all(np.isfinite(v) or v == np.inf for v in choice["seconds"].values())
    P, info = eng.batch([dm, dm], [k, k], [eta, eta], [q, q], "float32_refined")
    single, _ = dm.solve(k, eta, q, "float32_refined")
    assert rel(P[0], single) < 1e-9 and rel(P[1], single) < 1e-9

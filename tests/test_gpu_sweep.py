"""The pipelined frequency sweep (beat_bm_sweep_device): every frequency's
pressures equal the oracle's (FP64 1e-10, FP32-LU + refinement 1e-8, FP32 1e-4)
and the single-frequency path's, with the next frequency assembled on a second
stream while the current one is factored; an odd count exercises both system
buffers and the last-frequency drain."""
import sys

import numpy as np
import pytest

import oracle
from beat_engine_amd import exterior
from beat_engine_amd.mesh import jittered_icosphere
from beat_engine_amd.phasor import burton_miller_coupling, neumann_scale

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def case():
    mesh = jittered_icosphere(3, 0.1, 0.02, seed=17)
    freqs = np.array([150.0, 600.0, 1400.0, 2600.0, 4100.0])
    rng = np.random.default_rng(5)
    v = rng.normal(size=(mesh.num_faces, 3)) + 1j * rng.normal(size=(mesh.num_faces, 3))
    ks = 2 * np.pi * freqs / 343.0
    etas = np.array([burton_miller_coupling(k) for k in ks])
    qs = np.stack([neumann_scale(1.21, 2 * np.pi * f) * v for f in freqs])
    refs = []
    for k, eta, q in zip(ks, etas, qs):
        A, B = oracle.bm_assemble(mesh.vertices, mesh.faces, k, eta, q)
        refs.append(oracle.lu_solve(A, B))
    eng = exterior.Engine(0)
    yield mesh, ks, etas, qs, refs, eng
    eng.close()


@pytest.mark.parametrize("precision,tol", [("float64", 1e-10), ("float32_refined", 1e-8), ("float32", 1e-4)])
def test_sweep_matches_oracle_and_single_solves(case, precision, tol):
    mesh, ks, etas, qs, refs, eng = case
    print("sweep", precision, file=sys.stderr, flush=True)
    dm = eng.upload(mesh)
    P, reps = dm.sweep(ks, etas, qs, precision)
    assert P.shape == (len(ks), mesh.num_vertices, 3)
    for i in range(len(ks)):
        assert rel(P[i], refs[i]) < tol, i
        single, _ = dm.solve(ks[i], etas[i], qs[i], precision)
        assert rel(P[i], single) < (1e-12 if precision != "float32" else 1e-5), i
        if precision == "float32_refined":
            assert reps[i]["refinement_status"] == "converged"
        assert reps[i]["assembly_s"] > 0 and reps[i]["factor_s"] > 0


def test_cancel_marker_reclaim_and_capabilities(case, tmp_path):
    mesh, ks, etas, qs, refs, eng = case
    dm = eng.upload(mesh)
    marker = tmp_path / "cancel"
    marker.write_text("")
    eng.set_cancel_path(marker)
    P, reps = dm.sweep(ks, etas, qs, "float64")
    assert not any(r["solved"] for r in reps)
    marker.unlink()
    P, reps = dm.sweep(ks[:2], etas[:2], qs[:2], "float64")
    assert all(r["solved"] for r in reps)
    assert rel(P[1], refs[1]) < 1e-10
    eng.set_cancel_path(None)
    used = eng.memory_info()
    assert used["workspace_bytes"] > 0 and 0 < used["free_bytes"] <= used["total_bytes"]
    eng.reclaim()
    assert eng.memory_info()["workspace_bytes"] == 0
    P, reps = dm.sweep(ks[:1], etas[:1], qs[:1], "float32_refined")
    assert rel(P[0], refs[0]) < 1e-8
    caps = eng.capabilities()
    assert caps["arch"].startswith("gfx950")
    avail = exterior.worker_backend_availability(eng)
    assert avail["backends"]["rocm"]["available"] is True
    assert "pressure_gradient_v1" in avail["field_quantities"]["rocm"]

"""Every tile shape the autotuner can choose gives the same solves: the LU's four
trailing-update tile shapes (per-job factorizations and the grouped batches'
batched factorizations) and the regular pass's three tilings, each set explicitly
for the mesh's size class and checked against the oracle (complex128 1e-10,
refined 1e-8) and against the default tiles."""
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
    # 642 unknowns: eleven 64-column leaves, three outer blocks, row tiles of both heights
    mesh = jittered_icosphere(3, 0.1, 0.02, seed=21)
    f = 1700.0
    w = 2 * np.pi * f
    k = w / 343.0
    eta = burton_miller_coupling(k)
    rng = np.random.default_rng(5)
    q = neumann_scale(1.21, w) * (rng.normal(size=(mesh.num_faces, 4)) + 1j * rng.normal(size=(mesh.num_faces, 4)))
    A, B = oracle.bm_assemble(mesh.vertices, mesh.faces, k, eta, q)
    p_ref = oracle.lu_solve(A, B)
    eng = exterior.Engine(0)
    dm = eng.upload(mesh)
    yield mesh, dm, eng, k, eta, q, p_ref
    eng.close()


TOL = {"float64": 1e-10, "float32_refined": 1e-8}


@pytest.mark.parametrize("precision", ["float64", "float32_refined"])
@pytest.mark.parametrize("tile", exterior.GEMM_TILES)
def test_trailing_update_tiles(case, precision, tile):
    mesh, dm, eng, k, eta, q, p_ref = case
    print("tile", tile, precision, file=sys.stderr, flush=True)
    eng.set_tuning(mesh.num_vertices, precision, gemm_tile=tile)
    try:
        P, info = dm.solve(k, eta, q, precision)
        # the grouped batch (batched factorization with the same tile shape)
        Pb, _ = eng.batch([dm, dm], [k, k], [eta, eta], [q, q], precision)
    finally:
        eng.set_tuning(mesh.num_vertices, precision)
    assert rel(P, p_ref) < TOL[precision]
    for p in Pb:
        assert rel(p, p_ref) < TOL[precision]


@pytest.mark.parametrize("precision", ["float64", "float32_refined"])
@pytest.mark.parametrize("tiles", exterior.REGULAR_TILES)
def test_regular_pass_tiles(case, precision, tiles):
    mesh, dm, eng, k, eta, q, p_ref = case
    print("regular tiles", tiles, precision, file=sys.stderr, flush=True)
    eng.set_tuning(mesh.num_vertices, precision, regular_tiles=tiles)
    try:
        P, _ = dm.solve(k, eta, q, precision)
    finally:
        eng.set_tuning(mesh.num_vertices, precision)
    assert rel(P, p_ref) < TOL[precision]


def test_autotune_choices_apply(case):
    mesh, dm, eng, k, eta, q, p_ref = case
    choice = eng.autotune(dm, "float32_refined", k, 4)
    print("autotune", choice, file=sys.stderr, flush=True)
    assert choice["gemm_tile"] in exterior.GEMM_TILES and choice["regular_tiles"] in exterior.REGULAR_TILES
    s = choice["seconds"]
    # the kept choice is the fastest of those timed
    assert s[choice["gemm_tile"]] == min(s[t] for t in exterior.GEMM_TILES)
    assert s[choice["regular_tiles"]] == min(s[t] for t in exterior.REGULAR_TILES)
    try:
        P, _ = dm.solve(k, eta, q, "float32_refined")
        Pb, _ = eng.batch([dm, dm, dm], [k, k, k], [eta, eta, eta], [q, q, q], "float32_refined")
    finally:
        eng.set_tuning(mesh.num_vertices, "float32_refined")
    assert rel(P, p_ref) < 1e-8
    for p in Pb:
        assert rel(p, p_ref) < 1e-8


def test_set_tuning_rejects_out_of_range(case):
    mesh, dm, eng, *_ = case
    with pytest.raises(ValueError):
        eng.set_tuning(mesh.num_vertices, gemm_tile="256x64_k8")
    import ctypes
    ch = np.array([0, 0, 4, 0], dtype=np.int32)
    rc = eng._lib.beat_ctx_set_tuning(eng._h, 1, mesh.num_vertices, ch.ctypes.data_as(ctypes.c_void_p))
    assert rc != 0
    ch = np.array([0, 0, 0, 0], dtype=np.int32)
    rc = eng._lib.beat_ctx_set_tuning(eng._h, 7, mesh.num_vertices, ch.ctypes.data_as(ctypes.c_void_p))
    assert rc != 0

"""The plain-method and headline workloads on the GPU, at the sizes they are stated at.

Plain method (CPU FP64 reference case): level-3 icosphere, a = 0.1 m, v = 1 m/s on
every face, c = 343 m/s, rho = 1.21 kg/m^3, 500 Hz, eta = i/k, q4/s4, 1 RHS. The
GPU pressures equal the oracle's (FP64 1e-10, FP32-LU + FP64 refinement 1e-8)
and the pulsating-sphere closed form p(r) = rho c (a/r) ika/(ika - 1) e^{ik(r-a)}
within the reference's gates (0.05 dB, 0.5 deg; analytical_exterior_tests.jl:
111-115, 151-155), on the surface and at 100 points on a 1 m arc (there the field
of the GPU pressures within 1e-8 of the oracle's field of its own).

Headline workload: level-5 icosphere with radial jitter N(0, 0.02 h) (20480
triangles, 10242 P1), 8 RHS of complex N(0,1) normal velocity, q = i rho omega v,
through the pipelined device sweep at 100 Hz, 963 Hz and 5 kHz: every RHS of the
refined path converges and agrees with the complex128 GPU path within 1e-8
(relative L2), and the plain complex64 path within 2e-4. The same chain at level
4 (5120 triangles, 8 RHS) is compared with the FP64 CPU oracle directly."""
import ctypes
import sys

import numpy as np
import pytest

import oracle
from beat_engine_amd import _native, exterior
from beat_engine_amd.mesh import arc_points, icosphere, jittered_icosphere
from beat_engine_amd.phasor import burton_miller_coupling, neumann_scale

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def gates(actual, reference):
    r = actual / reference
    return float(np.max(np.abs(20 * np.log10(np.abs(r))))), float(np.max(np.abs(np.degrees(np.angle(r)))))


@pytest.fixture(scope="module")
def eng():
    e = exterior.Engine(0)
    yield e
    e.close()


def test_plain_method_pulsating_sphere(eng):
    say("plain method: L3, 500 Hz, rho 1.21, v = 1, 100 points on a 1 m arc")
    mesh = icosphere(0.1, 3)
    assert mesh.num_vertices == 642 and mesh.num_faces == 1280
    f, a, rho, c = 500.0, 0.1, 1.21, 343.0
    w = 2 * np.pi * f
    k = w / c
    eta = burton_miller_coupling(k)
    assert eta == pytest.approx(1j / k)
    q = np.full((mesh.num_faces, 1), neumann_scale(rho, w) * 1.0)
    A, B = oracle.bm_assemble(mesh.vertices, mesh.faces, k, eta, q)
    p_cpu = oracle.lu_solve(A, B)
    pts = arc_points(100, 1.0)
    f_cpu = oracle.field(mesh.vertices, mesh.faces, k, p_cpu, q, pts)
    ka = 1j * k * a
    p_exact = rho * c * ka / (ka - 1)  # every icosphere vertex lies on r = a
    f_exact = rho * c * (a / 1.0) * ka / (ka - 1) * np.exp(1j * k * (1.0 - a))
    dm = eng.upload(mesh, 4, 4)
    for prec, tol in (("float64", 1e-10), ("float32_refined", 1e-8)):
        P, rep = dm.solve(k, eta, q, prec)
        assert rep["factorization_count"] == 1
        if prec == "float32_refined":
            assert rep["refinement"]["status"] == "converged"
        assert rel(P, p_cpu) <= tol
        db, deg = gates(P[:, 0], np.full(mesh.num_vertices, p_exact))
        say(prec, "surface", db, "dB", deg, "deg")
        assert db <= 0.05 and deg <= 0.5
        F = dm.field(k, P, q, pts, "float64")
        assert rel(F, f_cpu) <= 1e-8
        db, deg = gates(F[:, 0], np.full(len(pts), f_exact))
        say(prec, "1 m arc", db, "dB", deg, "deg")
        assert db <= 0.05 and deg <= 0.5


def _drives(mesh, nrhs, seed=20480):
    rng = np.random.default_rng(seed)
    return rng.normal(size=(mesh.num_faces, nrhs)) + 1j * rng.normal(size=(mesh.num_faces, nrhs))


def test_level4_eight_rhs_against_the_cpu_oracle(eng):
    mesh = jittered_icosphere(4, 0.1, 0.02, seed=5120)
    assert mesh.num_faces == 5120
    v = _drives(mesh, 8, seed=4)
    f = 1500.0
    w = 2 * np.pi * f
    k = w / 343.0
    eta = burton_miller_coupling(k)
    q = neumann_scale(1.21, w) * v
    say("L4 oracle assembly, 8 rhs")
    A, B = oracle.bm_assemble(mesh.vertices, mesh.faces, k, eta, q)
    p_cpu = oracle.lu_solve(A, B)
    dm = eng.upload(mesh, 4, 4)
    for prec, tol in (("float64", 1e-10), ("float32_refined", 1e-8), ("float32", 2e-4)):
        P, rep = dm.solve(k, eta, q, prec)
        errs = [rel(P[:, j], p_cpu[:, j]) for j in range(8)]
        say(prec, "max rel L2 vs oracle", max(errs))
        assert max(errs) <= tol
        if prec == "float32_refined":
            assert rep["refinement"]["status"] == "converged"


class DeviceBlock:
    """Device memory through the HIP runtime the library itself uses (the same
    loaded libamdhip64), for the device-pointer entry points."""

    def __init__(self, nbytes):
        self.hip = ctypes.CDLL("libamdhip64.so.7")
        self.hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.hip.hipFree.argtypes = [ctypes.c_void_p]
        self.ptr = ctypes.c_void_p()
        self.nbytes = nbytes
        assert self.hip.hipMalloc(ctypes.byref(self.ptr), nbytes) == 0

    def put(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes == self.nbytes
        assert self.hip.hipMemcpy(self.ptr, a.ctypes.data_as(ctypes.c_void_p), a.nbytes, 1) == 0

    def get(self, shape, dtype):
        out = np.empty(shape, dtype)
        assert out.nbytes == self.nbytes
        assert self.hip.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.ptr, out.nbytes, 2) == 0
        return out

    def free(self):
        self.hip.hipFree(self.ptr)


def test_headline_level5_sweep_eight_rhs(eng):
    mesh = jittered_icosphere(5, 0.1, 0.02, seed=20480)
    assert mesh.num_vertices == 10242 and mesh.num_faces == 20480
    nrhs = 8
    v = _drives(mesh, nrhs)
    freqs = np.array([100.0, np.logspace(np.log10(100.0), np.log10(5000.0), 20)[11], 5000.0])
    ks = 2 * np.pi * freqs / 343.0
    etas = [burton_miller_coupling(k) for k in ks]
    er = np.ascontiguousarray([e.real for e in etas])
    ei = np.ascontiguousarray([e.imag for e in etas])
    # per frequency a column-major F x nrhs complex128 block (= row-major nrhs x F)
    qh = np.stack([np.ascontiguousarray((neumann_scale(1.21, 2 * np.pi * f) * v).T) for f in freqs])
    dm = eng.upload(mesh, 4, 4)
    lib = _native.load()
    q_dev = DeviceBlock(qh.nbytes)
    q_dev.put(qh)
    pshape = (len(freqs), nrhs, mesh.num_vertices)
    p_dev = DeviceBlock(int(np.prod(pshape)) * 16)
    out = {}
    try:
        for prec in (_native.PREC_FP32_REFINED, _native.PREC_FP64, _native.PREC_FP32):
            say("L5 sweep, precision", prec)
            reps = (_native.SolveReport * len(freqs))()
            opts = _native.SolveOptions(3, 1e-10)
            rc = lib.beat_bm_sweep_device(eng._h, dm._h, prec, len(freqs), ks.ctypes.data_as(ctypes.c_void_p),
                                          er.ctypes.data_as(ctypes.c_void_p), ei.ctypes.data_as(ctypes.c_void_p),
                                          q_dev.ptr, nrhs, p_dev.ptr, ctypes.byref(opts), reps)
            eng._check(rc, "beat_bm_sweep_device")
            for r in reps:
                assert r.factorization_count == 1
                if prec == _native.PREC_FP32_REFINED:
                    assert _native.REFINE_STATUS[r.refine.status] == "converged"
            out[prec] = p_dev.get(pshape, np.complex128)
    finally:
        q_dev.free()
        p_dev.free()
    for i, f in enumerate(freqs):
        ref = out[_native.PREC_FP64][i]
        e_ref = max(rel(out[_native.PREC_FP32_REFINED][i][j], ref[j]) for j in range(nrhs))
        e_32 = max(rel(out[_native.PREC_FP32][i][j], ref[j]) for j in range(nrhs))
        say(f"{f:.1f} Hz: refined vs complex128 {e_ref:.2e}, complex64 vs complex128 {e_32:.2e}")
        assert e_ref <= 1e-8
        assert e_32 <= 2e-4

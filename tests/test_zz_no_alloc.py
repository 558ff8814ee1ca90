"""beat_field_eval_device_noalloc (device pointers, no allocation: every staging
buffer in the library's file-scope device scratch array) against beat_field_eval
on the same inputs: two calls in a row, the second on a smaller mesh with fewer
right-hand sides and points (the array keeps the first call's contents), FP64
within 1e-13 and FP32 within 1e-6 relative (the two paths run the same kernels;
only the order of the source splits' atomic additions may differ). A call whose
buffers would not fit the array is rejected before any GPU work."""
import ctypes
import sys

import numpy as np
import pytest

from beat_engine_amd import _native, exterior
from beat_engine_amd.mesh import fibonacci_sphere, icosphere, jittered_icosphere
from beat_engine_amd.phasor import burton_miller_coupling, neumann_scale

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


class DeviceBlock:
    """Device memory through the HIP runtime the library itself uses."""

    def __init__(self, nbytes):
        self.hip = ctypes.CDLL("libamdhip64.so.7")
        self.hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.hip.hipFree.argtypes = [ctypes.c_void_p]
        self.ptr = ctypes.c_void_p()
        self.nbytes = nbytes
        assert self.hip.hipMalloc(ctypes.byref(self.ptr), max(nbytes, 256)) == 0

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


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def noalloc_field(eng, dm, precision, k, P, q, pts):
    """P: nv x nrhs, q: nf x nrhs, pts: npts x 3 -> npts x nrhs, through device pointers."""
    lib = _native.load()
    nrhs = P.shape[1]
    blocks = []
    try:
        pd = DeviceBlock(P.size * 16)
        blocks.append(pd)
        pd.put(np.asfortranarray(P).ravel(order="F"))
        qd = DeviceBlock(q.size * 16)
        blocks.append(qd)
        qd.put(np.asfortranarray(q).ravel(order="F"))
        xd = DeviceBlock(pts.size * 8)
        blocks.append(xd)
        xd.put(np.ascontiguousarray(pts, dtype=np.float64))
        od = DeviceBlock(len(pts) * nrhs * 16)
        blocks.append(od)
        rc = lib.beat_field_eval_device_noalloc(eng._h, dm._h, exterior.PRECISIONS[precision], float(k), pd.ptr, qd.ptr,
                                                nrhs, xd.ptr, len(pts), od.ptr)
        eng._check(rc, "beat_field_eval_device_noalloc")
        return od.get(len(pts) * nrhs, np.complex128).reshape((len(pts), nrhs), order="F")
    finally:
        for b in blocks:
            b.free()


@pytest.fixture(scope="module")
def eng():
    e = exterior.Engine(0)
    yield e
    e.close()


def _case(mesh, nrhs, f, seed):
    w = 2 * np.pi * f
    k = w / 343.0
    rng = np.random.default_rng(seed)
    q = neumann_scale(1.21, w) * (rng.normal(size=(mesh.num_faces, nrhs)) + 1j * rng.normal(size=(mesh.num_faces, nrhs)))
    return k, q


@pytest.mark.parametrize("precision,tol", [("float64", 1e-13), ("float32", 1e-6)])
def test_noalloc_field_matches_the_allocating_entry_point(eng, precision, tol):
    cases = [(jittered_icosphere(3, 0.1, 0.02, seed=3), 4, 300, 900.0),
             (icosphere(0.1, 2), 2, 50, 2500.0)]  # the second call: smaller mesh, fewer drives and points
    for i, (mesh, nrhs, npts, f) in enumerate(cases):
        say("noalloc call", i, precision, mesh.num_faces, nrhs, npts)
        dm = eng.upload(mesh)
        k, q = _case(mesh, nrhs, f, seed=10 + i)
        P, _ = dm.solve(k, burton_miller_coupling(k), q, "float64")
        pts = fibonacci_sphere(npts, 1.5)
        ref = dm.field(k, P, q, pts, precision)
        got = noalloc_field(eng, dm, precision, k, P, q, pts)
        assert got.shape == ref.shape
        assert np.all(np.isfinite(got))
        assert rel(got, ref) <= tol, i


def test_noalloc_field_rejects_what_does_not_fit(eng):
    mesh = icosphere(0.1, 4)
    dm = eng.upload(mesh)
    lib = _native.load()
    nrhs, npts = 64, 20000
    dummy = DeviceBlock(256)
    try:
        rc = lib.beat_field_eval_device_noalloc(eng._h, dm._h, _native.PREC_FP64, 10.0, dummy.ptr, dummy.ptr, nrhs,
                                                dummy.ptr, npts, dummy.ptr)
    finally:
        dummy.free()
    assert rc == _native.BEAT_ERR_ARG
    msg = lib.beat_last_error(eng._h).decode()
    assert "scratch" in msg and str(64 << 20) in msg

"""The retained `bem_field` operation (BeatEngineCompiledDriver.jl:3993-4255):
binary-array validation and round trip, request validation and the LRU retained
cache on the CPU; end-to-end pressure and pressure_gradient_v1 requests on the GPU
against the oracle."""
import json
import os

import numpy as np
import pytest

import oracle
from beat_engine_amd import field_request as fr
from beat_engine_amd.mesh import icosphere


def write_arrays(tmp_path, arrays, fname="arrays.bin"):
    """Packs arrays into one file and returns the binary_arrays descriptors."""
    desc, off = {}, 0
    with open(tmp_path / fname, "wb") as fh:
        for name, (arr, dtype) in arrays.items():
            a = np.ascontiguousarray(arr, dtype=fr._DTYPES[dtype])
            fh.write(a.tobytes())
            desc[name] = {"file": fname, "offset": off, "nbytes": a.nbytes, "dtype": dtype,
                          "shape": list(a.shape), "order": "C", "byte_order": "little"}
            off += a.nbytes
    return desc


def request_for(tmp_path, mesh, p, q, pts, precision="float64", quantity="pressure", **extra):
    grad = quantity == "pressure_gradient_v1"
    c = "complex128" if precision == "float64" else "complex64"
    geo = precision if grad else "float32"
    arrays = write_arrays(tmp_path, {
        "boundary_points_m": (mesh.vertices, geo),
        "boundary_triangles": (mesh.faces.astype(np.int64), "int64"),
        "boundary_pressure": (p, c),
        "boundary_normal_derivative": (q, c),
        "points_m": (pts, geo),
    })
    req = {"binary_array_schema_version": 1, "binary_arrays": arrays, "domain_key": "dom", "precision": precision,
           "bem_backend": "rocm", "field_quantity": quantity, "frequency_hz": 700.0, "sound_speed_m_per_s": 343.0,
           "binary_result_file": "result.bin"}
    req.update(extra)
    path = tmp_path / "request.json"
    path.write_text(json.dumps(req))
    return req, str(path)


def test_binary_round_trip_and_checks(tmp_path):
    a = np.arange(12, dtype=np.float64).reshape(4, 3)
    desc = write_arrays(tmp_path, {"x": (a, "float64"), "y": (np.arange(3) + 1j, "complex64")})
    req = {"binary_array_schema_version": 1, "binary_arrays": desc}
    rp = str(tmp_path / "r.json")
    v, shape = fr.read_field_binary_array(req, rp, "x", "float64")
    assert shape == [4, 3] and np.array_equal(v.reshape(shape), a)
    v, shape = fr.read_field_binary_array(req, rp, "y", "complex64")
    assert np.array_equal(v, (np.arange(3) + 1j).astype(np.complex64))

    def bad(mut, msg, name="x", dtype="float64"):
        r = json.loads(json.dumps(req))
        mut(r)
        with pytest.raises(ValueError, match=msg):
            fr.read_field_binary_array(r, rp, name, dtype)

    bad(lambda r: r.update(binary_array_schema_version=2), "schema")
    bad(lambda r: r.pop("binary_arrays"), "missing binary-array metadata")
    bad(lambda r: None, "missing binary array 'z'", name="z")
    bad(lambda r: None, "wrong data type", dtype="float32")
    bad(lambda r: r["binary_arrays"]["x"].update(order="F"), "row-major")
    bad(lambda r: r["binary_arrays"]["x"].update(byte_order="big"), "little-endian")
    bad(lambda r: r["binary_arrays"]["x"].update(shape=[-1, 3]), "invalid shape")
    bad(lambda r: r["binary_arrays"]["x"].update(offset=-1), "invalid offset")
    bad(lambda r: r["binary_arrays"]["x"].update(nbytes=8), "inconsistent byte metadata")
    bad(lambda r: r["binary_arrays"]["x"].update(file="../arrays.bin"), "invalid file path")
    bad(lambda r: r["binary_arrays"]["x"].update(file="none.bin"), "does not exist")
    bad(lambda r: r["binary_arrays"]["x"].update(offset=10_000), "incomplete")

    out = fr.write_field_binary_result({"binary_result_file": "res.bin"}, rp,
                                       np.arange(8).reshape(4, 2).astype(np.complex64))
    assert out["shape"] == [4, 2] and out["dtype"] == "complex64" and out["nbytes"] == 64
    back = np.fromfile(tmp_path / "res.bin", dtype="<c8").reshape(4, 2)
    assert np.array_equal(back, np.arange(8).reshape(4, 2))
    with pytest.raises(ValueError, match="invalid binary result path"):
        fr.write_field_binary_result({"binary_result_file": "a/b.bin"}, rp, back)


@pytest.mark.parametrize("change,msg", [
    ({"precision": "float16"}, "float32 or float64"),
    ({"bem_backend": "tpu"}, "cpu, cuda, rocm, or metal"),
    ({"bem_backend": "cpu"}, "use rocm"),
    ({"field_quantity": "velocity"}, "Unsupported BEM field quantity"),
    ({"field_quantity": "pressure_gradient_v1", "quadrature_order": 2}, "fourth-order"),
    ({"symmetry": "z"}, "Unsupported image mode"),
    ({"phasor_convention": "sideways"}, "phasor convention"),
])
def test_request_validation(change, msg):
    req = {"precision": "float64", "bem_backend": "rocm"}
    req.update(change)
    with pytest.raises(ValueError, match=msg):
        fr.evaluate_bem_field_request(req, "/nonexistent/request.json", engine=object())


def test_retained_cache_is_lru():
    closed = []

    class Dev:
        def __init__(self, name):
            self.name = name

        def close(self):
            closed.append(self.name)

    c = fr.RetainedFieldCaches(2)
    c.put("a", (None, Dev("a")))
    c.put("b", (None, Dev("b")))
    assert c.get("a") is not None  # a becomes most recent
    c.put("c", (None, Dev("c")))
    assert closed == ["b"] and c.keys() == ["a", "c"]
    c.release_all()
    assert sorted(closed) == ["a", "b", "c"] and len(c) == 0


@pytest.mark.gpu
@pytest.mark.parametrize("precision,tol", [("float64", 1e-12), ("float32", 1e-5)])
def test_pressure_request_end_to_end(tmp_path, precision, tol):
    from beat_engine_amd import exterior
    mesh = icosphere(0.1, 2)
    rng = np.random.default_rng(2)
    p = rng.normal(size=mesh.num_vertices) + 1j * rng.normal(size=mesh.num_vertices)
    q = rng.normal(size=mesh.num_faces) + 1j * rng.normal(size=mesh.num_faces)
    pts = np.array([[0.0, 0.0, 1.0], [0.5, 0.2, -0.3], [2.0, 0.0, 0.0]])
    req, path = request_for(tmp_path, mesh, p, q, pts, precision)
    eng = exterior.Engine(0)
    caches = fr.RetainedFieldCaches()
    try:
        events = fr.run_bem_field(path, engine=eng, caches=caches)
        assert [e["type"] for e in events] == ["field_result", "completed"]
        d = events[0]["values_binary"]
        got = np.fromfile(tmp_path / d["file"], dtype="<c16" if precision == "float64" else "<c8")
        assert d["shape"] == [3]
        # the reference reads geometry and points as float32 for pressure requests
        f32 = lambda a: a.astype(np.float32).astype(np.float64)
        k = float(np.float64(2 * np.pi) * 700.0 / 343.0) if precision == "float64" else \
            float(np.float32(2 * np.pi) * np.float32(700.0) / np.float32(343.0))
        ref = oracle.field(f32(mesh.vertices), mesh.faces, k, p, q, f32(pts), quad_order=2,
                           precision="f64" if precision == "float64" else "f32")[:, 0]
        assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < tol
        first = caches.get("dom|%s|rocm|off|q2|pressure" % precision)
        fr.run_bem_field(path, engine=eng, caches=caches)
        assert caches.get("dom|%s|rocm|off|q2|pressure" % precision) is first  # retained
    finally:
        caches.release_all()
        eng.close()


@pytest.mark.gpu
def test_gradient_request_end_to_end(tmp_path):
    from beat_engine_amd import exterior
    mesh = icosphere(0.1, 2, (0.2, 0.0, 0.0))
    rng = np.random.default_rng(3)
    p = rng.normal(size=mesh.num_vertices) + 1j * rng.normal(size=mesh.num_vertices)
    q = rng.normal(size=mesh.num_faces) + 1j * rng.normal(size=mesh.num_faces)
    pts = np.array([[0.2, 0.0, 0.105], [0.6, 0.3, 0.1], [0.0, 0.0, 1.0]])
    req, path = request_for(tmp_path, mesh, p, q, pts, "float64", "pressure_gradient_v1", symmetry="x")
    eng = exterior.Engine(0)
    caches = fr.RetainedFieldCaches()
    try:
        events = fr.run_bem_field(path, engine=eng, caches=caches)
        d = events[0]["values_binary"]
        assert d["shape"] == [3, 4] and d["dtype"] == "complex128"
        got = np.fromfile(tmp_path / d["file"], dtype="<c16").reshape(3, 4)
        k = 2 * np.pi * 700.0 / 343.0
        ref = oracle.field_gradient(mesh.vertices, mesh.faces, k, p, q, pts, symmetry="x")
        assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 1e-11
    finally:
        caches.release_all()
        eng.close()

"""The worker's retained exterior-field operation (`bem_field`) over the gfx950
library: binary-array I/O, the retained device-mesh cache and the dispatch of
`pressure` / `pressure_gradient_v1` requests.

  read_field_binary_array       <- read_field_binary_array
                                   (src/beat_engine/julia_local/BeatEngineCompiledDriver.jl:3993-4027)
  write_field_binary_result     <- write_field_binary_result (:4029-4049)
  RetainedFieldCaches           <- BEM_FIELD_EVALUATION_CACHES, retained_bem_field_evaluation_cache,
                                   release_bem_field_evaluation_cache!, release_all_... (:4051-4151)
  evaluate_bem_field_request    <- evaluate_bem_field_request(_impl) (:4153-4255)

Request keys, array names, validation order and error messages follow the
reference. This build is the ROCm backend: requests for other backends are
rejected rather than evaluated elsewhere.
"""
from __future__ import annotations

import json
import math
import os
from collections import OrderedDict

import numpy as np

from . import exterior
from .mesh import BoundaryMesh, normalized_symmetry_mode
from .phasor import NEGATIVE_TIME_PHASOR, POSITIVE_TIME_PHASOR, outgoing_wavenumber

# worker-v1.json "field_cache": lifetime worker_process, cleared on solve / failed / exit
MAX_BEM_FIELD_EVALUATION_CACHES = 2
FIELD_ARRAY_VERSION = 1

_DTYPES = {
    "float32": np.dtype("<f4"), "float64": np.dtype("<f8"), "int64": np.dtype("<i8"),
    "complex64": np.dtype("<c8"), "complex128": np.dtype("<c16"),
}


def read_field_binary_array(request: dict, request_path: str, name: str, dtype: str):
    """(values as a flat array, shape) of one described array, with the reference's checks."""
    if int(request.get("binary_array_schema_version", 0)) != FIELD_ARRAY_VERSION:
        raise ValueError("Unsupported BEM field binary-array schema.")
    arrays = request.get("binary_arrays")
    if not isinstance(arrays, dict):
        raise ValueError("BEM field request is missing binary-array metadata.")
    desc = arrays.get(name)
    if not isinstance(desc, dict):
        raise ValueError(f"BEM field request is missing binary array {name!r}.")
    if str(desc.get("dtype", "")) != dtype:
        raise ValueError(f"BEM field binary array {name!r} has the wrong data type.")
    if str(desc.get("order", "")) != "C":
        raise ValueError(f"BEM field binary array {name!r} must use row-major order.")
    if str(desc.get("byte_order", "")) != "little":
        raise ValueError(f"BEM field binary array {name!r} must use little-endian byte order.")
    shape = [int(s) for s in desc.get("shape", [])]
    if any(s < 0 for s in shape):
        raise ValueError(f"BEM field binary array {name!r} has an invalid shape.")
    count = math.prod(shape)
    offset = int(desc.get("offset", -1))
    nbytes = int(desc.get("nbytes", -1))
    dt = _DTYPES[dtype]
    if offset < 0:
        raise ValueError(f"BEM field binary array {name!r} has an invalid offset.")
    if nbytes != count * dt.itemsize:
        raise ValueError(f"BEM field binary array {name!r} has inconsistent byte metadata.")
    filename = str(desc.get("file", ""))
    if not filename or os.path.basename(filename) != filename:
        raise ValueError(f"BEM field binary array {name!r} has an invalid file path.")
    path = os.path.join(os.path.dirname(request_path), filename)
    if not os.path.isfile(path):
        raise ValueError(f"BEM field binary array file does not exist: {filename}")
    if os.path.getsize(path) < offset + nbytes:
        raise ValueError(f"BEM field binary array {name!r} is incomplete.")
    with open(path, "rb") as fh:
        fh.seek(offset)
        values = np.frombuffer(fh.read(nbytes), dtype=dt, count=count).copy()
    return values, shape


def write_field_binary_result(request: dict, request_path: str, values) -> dict:
    """Writes values (C order, complex of the evaluation precision) and returns its descriptor."""
    filename = str(request.get("binary_result_file", ""))
    if not filename or os.path.basename(filename) != filename:
        raise ValueError("BEM field request has an invalid binary result path.")
    arr = np.asarray(values)
    if arr.dtype not in (np.complex64, np.complex128):
        raise ValueError(f"Unsupported BEM field result precision: {arr.dtype}")
    arr = np.ascontiguousarray(arr, dtype=arr.dtype.newbyteorder("<"))
    with open(os.path.join(os.path.dirname(request_path), filename), "wb") as fh:
        fh.write(arr.tobytes(order="C"))
    return {"file": filename, "offset": 0, "nbytes": int(arr.nbytes),
            "dtype": "complex64" if arr.dtype == np.complex64 else "complex128",
            "shape": list(arr.shape), "order": "C", "byte_order": "little"}


class RetainedFieldCaches:
    """Device meshes kept between field requests, least recently used first out."""

    def __init__(self, max_entries: int = MAX_BEM_FIELD_EVALUATION_CACHES):
        self.max_entries = max_entries
        self._entries: OrderedDict[str, tuple] = OrderedDict()

    def __len__(self):
        return len(self._entries)

    def keys(self):
        return list(self._entries)

    def get(self, key):
        entry = self._entries.get(key)
        if entry is not None:
            self._entries.move_to_end(key)
        return entry

    def put(self, key, entry):
        self._entries[key] = entry
        self._entries.move_to_end(key)
        while len(self._entries) > self.max_entries:
            self.release(next(iter(self._entries)))

    def release(self, key):
        entry = self._entries.pop(key, None)
        if entry is not None:
            entry[1].close()

    def release_all(self):
        for key in list(self._entries):
            self.release(key)


_default_caches = RetainedFieldCaches()


def retained_bem_field_evaluation_cache(request, request_path, precision, backend, symmetry, engine, caches):
    domain_key = str(request.get("domain_key", "")).strip()
    if not domain_key:
        raise ValueError("Exterior field request is missing its domain cache key.")
    quantity = str(request.get("field_quantity", "pressure"))
    quadrature_order = int(request.get("quadrature_order", 4 if quantity == "pressure_gradient_v1" else 2))
    key = f"{domain_key}|{precision}|{backend}|{symmetry}|q{quadrature_order}|{quantity}"
    cached = caches.get(key)
    if cached is not None:
        return cached
    vdtype = precision if quantity == "pressure_gradient_v1" else "float32"
    vv, vshape = read_field_binary_array(request, request_path, "boundary_points_m", vdtype)
    if not (len(vshape) == 2 and vshape[1] == 3):
        raise ValueError("BEM boundary points must have shape (bem_node, 3).")
    fv, fshape = read_field_binary_array(request, request_path, "boundary_triangles", "int64")
    if not (len(fshape) == 2 and fshape[1] == 3):
        raise ValueError("BEM boundary triangles must have shape (bem_face, 3).")
    # int64 indices go to BoundaryMesh as they are: it range-checks them before narrowing
    mesh = BoundaryMesh(vv.reshape(-1, 3).astype(np.float64), fv.reshape(-1, 3), np.ones(fshape[0]))
    dm = exterior.DeviceMesh(engine, mesh, quadrature_order, 4, symmetry)
    entry = (mesh, dm)
    caches.put(key, entry)
    return entry


def evaluate_bem_field_request(request: dict, request_path: str, *, engine=None, caches=None):
    """Values of one `bem_field` request: (points,) for pressure, (points, 4) for
    pressure_gradient_v1, complex of the request precision."""
    convention = request.get("phasor_convention", NEGATIVE_TIME_PHASOR)
    if convention not in (NEGATIVE_TIME_PHASOR, POSITIVE_TIME_PHASOR):
        raise ValueError(f"Unsupported phasor convention: {convention!r}")
    precision = str(request.get("precision", "float32")).lower()
    if precision not in ("float32", "float64"):
        raise ValueError("Exterior field precision must be float32 or float64.")
    backend = str(request.get("bem_backend", "cpu")).lower()
    if backend not in ("cpu", "cuda", "rocm", "metal"):
        raise ValueError("Exterior field backend must be cpu, cuda, rocm, or metal.")
    if backend != "rocm":
        raise ValueError(f"Exterior field backend {backend!r} is not provided by the gfx950 library (use rocm).")
    symmetry = normalized_symmetry_mode(request.get("symmetry", "off"))
    quantity = str(request.get("field_quantity", "pressure"))
    if quantity not in ("pressure", "pressure_gradient_v1"):
        raise ValueError(f"Unsupported BEM field quantity: {quantity}")
    if quantity == "pressure_gradient_v1" and int(request.get("quadrature_order", 4)) != 4:
        raise ValueError("BEM gradients require fourth-order quadrature.")
    eng = engine or exterior.default_engine()
    caches = _default_caches if caches is None else caches
    mesh, dm = retained_bem_field_evaluation_cache(request, request_path, precision, backend, symmetry, eng, caches)
    ctype = "complex64" if precision == "float32" else "complex128"
    pressure, pshape = read_field_binary_array(request, request_path, "boundary_pressure", ctype)
    normal, nshape = read_field_binary_array(request, request_path, "boundary_normal_derivative", ctype)
    if len(pshape) != 1:
        raise ValueError("Retained BEM pressure must be one-dimensional.")
    if len(nshape) != 1:
        raise ValueError("Retained BEM normal derivative must be one-dimensional.")
    if len(pressure) != mesh.num_vertices:
        raise ValueError("Retained BEM pressure does not align with boundary vertices.")
    if len(normal) != mesh.num_faces:
        raise ValueError("Retained BEM normal derivative does not align with boundary faces.")
    pdtype = precision if quantity == "pressure_gradient_v1" else "float32"
    pv, ptshape = read_field_binary_array(request, request_path, "points_m", pdtype)
    if not (len(ptshape) == 2 and ptshape[1] == 3):
        raise ValueError("Exterior evaluation points must have shape (point, 3).")
    points = pv.reshape(-1, 3).astype(np.float64)
    ftype = np.float32 if precision == "float32" else np.float64
    frequency = ftype(request["frequency_hz"])
    speed = ftype(request["sound_speed_m_per_s"])
    if not frequency > 0:
        raise ValueError("Exterior field frequency must be positive.")
    if not speed > 0:
        raise ValueError("Exterior field sound speed must be positive.")
    k = float(ftype(2 * np.pi) * frequency / speed)
    ks = outgoing_wavenumber(k, convention)
    out_type = np.complex64 if precision == "float32" else np.complex128
    if quantity == "pressure_gradient_v1":
        vals = dm.field_gradient(ks, pressure.astype(np.complex128), normal.astype(np.complex128), points, precision)
        return vals[:, :, 0].astype(out_type)
    vals = dm.field(ks, pressure.astype(np.complex128), normal.astype(np.complex128), points, precision)
    return vals[:, 0].astype(out_type)


def run_bem_field(request_path: str, *, engine=None, caches=None) -> list:
    """The worker's `bem_field` branch (BeatEngineCompiledDriver.jl:4290-4307): the
    events it prints, as dictionaries."""
    with open(request_path) as fh:
        request = json.load(fh)
    values = evaluate_bem_field_request(request, request_path, engine=engine, caches=caches)
    return [{"type": "field_result", "field_quantity": request.get("field_quantity", "pressure"),
             "phasor_convention": request.get("phasor_convention", NEGATIVE_TIME_PHASOR),
             "values_binary": write_field_binary_result(request, request_path, values)},
            {"type": "completed"}]

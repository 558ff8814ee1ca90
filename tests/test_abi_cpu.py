"""CPU-side checks of the C ABI: the library loads and exports every symbol the
public header declares (no compute calls without a GPU)."""
import ctypes
import os
import re

import pytest

from beat_engine_amd import _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _header_functions():
    text = open(os.path.join(ROOT, "include", "beat_bem.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(beat_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_binding_list():
    assert _header_functions() == sorted(_native.EXPORTS)


def test_library_exports_every_declared_symbol():
    if not os.path.exists(_native.LIB_PATH):
        pytest.fail("libbeat_bem_hip.so is not built; run __graft_entry__.build()")
    lib = ctypes.CDLL(_native.LIB_PATH)
    for name in _header_functions():
        assert hasattr(lib, name), name
    _native.load()


def test_missing_library_fails_loudly(tmp_path):
    with pytest.raises(RuntimeError, match="not built"):
        _native.load(str(tmp_path / "nope.so"))


def test_capabilities_without_a_device():
    caps = _native.capabilities()
    assert caps["library"] == "beat_bem_hip" and caps["arch"] == "gfx950"
    assert "pressure_gradient_v1" in caps["field_quantities"]
    assert caps["cancellation"] == "marker_file"
    assert set(caps["symmetry_modes"]) == {"off", "x", "xy", "ground"}

"""CPU checks of the library's per-mesh preprocessing (mesh_setup.cpp compiled for
the host): the touching-pair list, the image-touching pairs of every symmetry
image, the delta rules and the orbit-weighted rows, against the oracle.

* ordinary pairs: the same (test, trial, kind) set as the oracle's singular cache;
* image pairs: per image transform the same (test, trial, kind) set as the
  oracle's image-singular cache (geometric adjacency of the reflected trial face);
* delta rules: Duffy part (weights summing to 1/4 as every Duffy rule over two
  reference triangles) followed by the regular product rule with negated weights
  (summing to -1/4), so each delta rule's weights sum to zero;
* orbit weights: the rows with a weight other than 1 and their weights equal the
  oracle's p1_symmetry_orbit_weights.
"""
import os
import subprocess

import numpy as np
import pytest

import oracle
from beat_engine_amd.mesh import BoundaryMesh, icosphere, load_gmsh22, snap_symmetry_planes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
CODES = {"off": 0, "x": 1, "xy": 2, "ground": 3}


@pytest.fixture(scope="module")
def exe(tmp_path_factory):
    out = tmp_path_factory.mktemp("msc") / "mesh_setup_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
                    os.path.join(ROOT, "tests", "native", "mesh_setup_check.cpp"),
                    os.path.join(ROOT, "beat_engine_amd", "csrc", "mesh_setup.cpp"), "-o", str(out)], check=True)
    return str(out)


def run(exe, mesh, mode, qo=4, so=4):
    lines = [f"{len(mesh.vertices)} {len(mesh.faces)} {qo} {so} {CODES[mode]}"]
    lines += [" ".join(repr(float(c)) for c in v) for v in mesh.vertices]
    lines += [" ".join(str(int(c)) for c in f) for f in mesh.faces]
    res = subprocess.run([exe], input="\n".join(lines) + "\n", capture_output=True, text=True)
    assert res.returncode == 0, res.stdout + res.stderr
    pairs, rules, weights = [], [], {}
    for ln in res.stdout.splitlines():
        t = ln.split()
        if t[0] == "pair":
            pairs.append(tuple(int(x) for x in t[1:]))
        elif t[0] == "rule":
            rules.append((int(t[2]), float(t[3])))
        elif t[0] == "weight":
            weights[int(t[1])] = float(t[2])
    assert res.stdout.rstrip().endswith("done")
    return pairs, rules, weights


def weld(mesh):
    keys, remap, verts = {}, [], []
    for v in mesh.vertices:
        kk = tuple(np.round(v / 1e-9).astype(np.int64))
        if kk not in keys:
            keys[kk] = len(verts)
            verts.append(v)
        remap.append(keys[kk])
    return BoundaryMesh(np.array(verts), np.array(remap)[mesh.faces], mesh.physical_tags)


def ground_half():
    """sample_half turned so that its cut lies on the ground plane y = 0 (faces
    reversed to keep the outward winding under the axis swap)."""
    m = weld(snap_symmetry_planes(load_gmsh22(os.path.join(GOLDEN, "sample_half.msh"), 0.001), "x"))
    return BoundaryMesh(m.vertices[:, [1, 0, 2]].copy(), m.faces[:, ::-1].copy(), m.physical_tags)


CASES = {
    "sample_half/x": (lambda: weld(snap_symmetry_planes(load_gmsh22(os.path.join(GOLDEN, "sample_half.msh"), 0.001),
                                                        "x")), "x"),
    "sample_quarter/xy": (lambda: weld(snap_symmetry_planes(
        load_gmsh22(os.path.join(GOLDEN, "sample_quarter.msh"), 0.001), "xy")), "xy"),
    "ground_half/ground": (ground_half, "ground"),
    "offplane/xy": (lambda: icosphere(0.1, 2, (0.4, 0.3, 0.05)), "xy"),
    "sample/off": (lambda: load_gmsh22(os.path.join(GOLDEN, "sample.msh"), 0.001), "off"),
}
NIMG = {"off": 0, "x": 1, "xy": 3, "ground": 1}


@pytest.mark.parametrize("case", list(CASES))
def test_pairs_rules_and_weights_match_oracle(exe, case):
    make, mode = CASES[case]
    m = make()
    so = 2
    pairs, rules, weights = run(exe, m, mode, 4, so)
    base = {(t, s, k) for (t, s, k, x) in pairs if x == 0}
    ref = {tuple(int(c) for c in p[:3]) for p in oracle.singular_pairs(m.vertices, m.faces, so)}
    assert base == ref
    assert sum(1 for p in pairs if p[3] == 0) == len(ref)
    for img in range(NIMG[mode]):
        got = {(t, s, k) for (t, s, k, x) in pairs if x == img + 1}
        want = {tuple(int(c) for c in p[:3]) for p in oracle.image_singular_pairs(m.vertices, m.faces, mode, img, so)}
        assert got == want, (img, len(got), len(want))
        assert sum(1 for p in pairs if p[3] == img + 1) == len(want)
    assert max((p[3] for p in pairs), default=0) <= NIMG[mode]
    # every ordinary rule integrates 1 over the two reference triangles (1/4 in
    # rule weights); a delta rule adds the regular product rule with weight -1/4
    sums = np.array([w for _, w in rules])
    assert np.all((np.abs(sums - 0.25) < 1e-14) | (np.abs(sums) < 1e-14))
    if mode != "off" and case != "offplane/xy":
        assert np.any(np.abs(sums) < 1e-14)
    w = oracle.orbit_weights(m.vertices, m.faces, mode)
    want_w = {i: float(v) for i, v in enumerate(w) if v != 1.0}
    assert weights == want_w
    if mode in ("x", "xy") and case != "offplane/xy":
        assert len(weights) > 0

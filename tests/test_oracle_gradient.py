"""Field value and observation gradient of the oracle
(src/beat_engine/julia_local/src/BeatEngineFieldGradient.jl:1-92), pinned by the
reference's own known answers (tests/field_gradient_tests.jl):

* radiating monopole: a closed sphere carrying the traces of exp(-ikr)/r
  (positive-time phasor) reproduces the pressure and its gradient within 1%, and
  the active intensity points radially outward (:20-60);
* analytic gradient: a unit single-layer density on one triangle with k = 0 gives
  d/dz = solid angle / 4 pi (Van Oosterom-Strackee formula) at heights 0.2, 0.02,
  0.002 above it, the last two inside the adaptive near-panel correction, in FP32
  and FP64 (:62-88);
* finite differences: the value equals the scalar field and the gradient the
  central difference of the scalar field (rtol 1e-8) for symmetry off/x/xy in both
  phasor conventions (:89-118).

Plus: a point closer to the panel than the depth-12 subdivision resolves raises the
reference's error, and the correction leaves far points untouched.
"""
import numpy as np
import pytest

import oracle
from beat_engine_amd.phasor import NEGATIVE_TIME_PHASOR, POSITIVE_TIME_PHASOR, outgoing_wavenumber


def isapprox(a, b, rtol):
    a, b = np.asarray(a), np.asarray(b)
    return np.linalg.norm(a - b) <= rtol * max(np.linalg.norm(a), np.linalg.norm(b))


def octa_sphere(levels=4, radius=0.1):
    """The reference test's sphere: an octahedron split `levels` times with midpoints
    pushed to the unit sphere, each face adding its own three midpoints."""
    v = [np.array(x, float) for x in [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]]
    faces = [(0, 2, 4), (2, 1, 4), (1, 3, 4), (3, 0, 4), (2, 0, 5), (1, 2, 5), (3, 1, 5), (0, 3, 5)]
    nrm = lambda x: x / np.linalg.norm(x)
    for _ in range(levels):
        nxt = []
        for a, b, c in faces:
            ab, bc, ca = len(v), len(v) + 1, len(v) + 2
            v += [nrm(v[a] + v[b]), nrm(v[b] + v[c]), nrm(v[c] + v[a])]
            nxt += [(a, ab, ca), (ab, b, bc), (ca, bc, c), (ab, bc, ca)]
        faces = nxt
    return radius * np.array(v), np.array(faces, np.int32)


def test_radiating_monopole_velocity_and_intensity():
    V, F = octa_sphere()
    points = np.array([[0.21, 0.08, 0.06], [-0.3, 0.4, 0.2], [1.0, 2.0, 3.0]])
    omega, rho, k = 2 * np.pi * 500, 1.2041, 2 * np.pi * 500 / 343
    ks = outgoing_wavenumber(k, POSITIVE_TIME_PHASOR)
    p = lambda x: np.exp(-1j * k * np.linalg.norm(x)) / np.linalg.norm(x)
    grad = lambda x: p(x) * (-1j * k - 1 / np.linalg.norm(x)) * x / np.linalg.norm(x)
    tri = V[F]
    cen = tri.mean(axis=1)
    nrm = np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0])
    nrm /= np.linalg.norm(nrm, axis=1)[:, None]
    pressure = np.array([p(x) for x in V])
    q = np.array([np.dot(grad(c), n) for c, n in zip(cen, nrm)])
    res = oracle.field_gradient(V, F, ks, pressure, q, points, quad_order=4)
    for i, x in enumerate(points):
        assert isapprox(res[i, 0], p(x), 0.01)
        assert isapprox(res[i, 1:], grad(x), 0.01)
        velocity = res[i, 1:] / (-1j * omega * rho)
        intensity = np.real(res[i, 0] * np.conj(velocity))
        assert np.dot(intensity, x) > 0
        assert np.linalg.norm(np.cross(intensity, x)) / np.linalg.norm(intensity) / np.linalg.norm(x) < 0.01


def solid_angle(tri, x):
    a, b, c = (v - x for v in tri)
    na, nb, nc = (np.linalg.norm(v) for v in (a, b, c))
    return abs(2 * np.arctan2(np.dot(a, np.cross(b, c)), na * nb * nc + np.dot(a, b) * nc + np.dot(b, c) * na
                              + np.dot(c, a) * nb))


@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_analytic_single_layer_gradient(precision):
    V = np.array([[0.0, 0, 0], [1, 0, 0], [0, 1, 0]])
    F = np.array([[0, 1, 2]], np.int32)
    points = np.array([[0.2, 0.2, z] for z in (0.2, 0.02, 0.002)])
    vals = oracle.field_gradient(V, F, 0.0, np.zeros(3), np.ones(1), points, precision=precision)
    for i, x in enumerate(points):
        assert np.isclose(vals[i, 3].real, solid_angle(V, x) / (4 * np.pi), rtol=2e-4)
        assert np.all(np.isfinite(vals[i]))
    # the 6-point rule alone (d/dz of -G q at k = 0 is (y - x)_z / (4 pi r^3)) is far
    # off at z = 0.002: the near-panel correction carries the answer there
    pts, w = oracle.triangle_rule(4)
    y = np.c_[pts, np.zeros(len(w))]
    r = y - points[2]
    plain = np.sum(w * r[:, 2] / (4 * np.pi * np.linalg.norm(r, axis=1) ** 3))
    assert abs(plain - solid_angle(V, points[2]) / (4 * np.pi)) > 0.1 * solid_angle(V, points[2]) / (4 * np.pi)


@pytest.mark.parametrize("symmetry", ["off", "x", "xy"])
@pytest.mark.parametrize("convention", [NEGATIVE_TIME_PHASOR, POSITIVE_TIME_PHASOR])
def test_gradient_is_the_field_derivative(symmetry, convention):
    V = np.array([[0.0, 0, 0], [1, 0, 0], [0, 1, 0]])
    F = np.array([[0, 1, 2]], np.int32)
    p = np.array([1 + 1j, 2 - 1j, 0.5j])
    q = np.array([0.2 + 0.3j])
    ks = outgoing_wavenumber(1.7, convention)
    x = np.array([4.0, 3.0, 2.0])
    vals = oracle.field_gradient(V, F, ks, p, q, x[None], symmetry=symmetry)
    field = oracle.field(V, F, ks, p, q, x[None], symmetry=symmetry)
    assert np.isclose(vals[0, 0], field[0, 0], rtol=1e-12)
    for d in range(3):
        delta = np.zeros(3)
        delta[d] = 1e-5
        pm = oracle.field(V, F, ks, p, q, np.stack([x + delta, x - delta]), symmetry=symmetry)
        assert np.isclose(vals[0, d + 1], (pm[0, 0] - pm[1, 0]) / 2e-5, rtol=1e-8)


def test_too_close_raises_like_the_reference():
    V = np.array([[0.0, 0, 0], [1, 0, 0], [0, 1, 0]])
    F = np.array([[0, 1, 2]], np.int32)
    with pytest.raises(ValueError, match="too close"):
        oracle.field_gradient(V, F, 0.0, np.zeros(3), np.ones(1), np.array([[0.2, 0.2, 1e-7]]))


def test_far_points_need_no_correction():
    # beyond 3 diameters of every centroid the result is the plain source sum: its
    # value equals the scalar field to rounding
    V, F = octa_sphere(2)
    rng = np.random.default_rng(0)
    p = rng.normal(size=len(V)) + 1j * rng.normal(size=len(V))
    q = rng.normal(size=len(F)) + 1j * rng.normal(size=len(F))
    pts = np.array([[0.0, 0.0, 0.6], [0.5, -0.4, 0.1]])
    vals = oracle.field_gradient(V, F, 9.0, p, q, pts)
    field = oracle.field(V, F, 9.0, p, q, pts)
    assert np.allclose(vals[:, 0], field[:, 0], rtol=1e-13)

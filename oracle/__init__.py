"""CPU ORACLE — TEST INFRASTRUCTURE ONLY.

Python face of the C++ restatement in bem_oracle.cpp (built to liboracle_bem.so
by the Makefile beside it) plus the dense linear algebra of the reference's CPU
path in numpy/scipy. Only tests/, the smoke check and bench.py's cpu_baseline
leg import this package; the product library must never route through it.

Reference anchors (paths relative to src/beat_engine/julia_local/):
  phasor helpers          src/BeatEnginePhasor.jl:1-24
  fused system            src/BeatEngineCpuBurtonMiller.jl:276-469
  four operators          src/BeatEngineCpuAssembly.jl:739-1031
  operator combination    src/BeatEngineCpuSolve.jl:55-110
  LU solve                src/BeatEngineDenseSolve.jl:1104-1110 (LAPACK getrf/getrs)
  refinement              src/BeatEngineCpuRefinement.jl:5-78
  field                   src/BeatEngineCpuField.jl:1-49
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle_bem.so")
_lib = None

NEGATIVE_TIME_PHASOR = "exp(-i omega t)"
POSITIVE_TIME_PHASOR = "exp(+i omega t)"


def build():
    import subprocess

    subprocess.run(["make", "-C", _HERE, "-s"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"oracle library missing: {_LIB_PATH} (run make -C oracle)")
        L = ctypes.CDLL(_LIB_PATH)
        P = ctypes.c_void_p
        I64 = ctypes.c_int64
        I = ctypes.c_int
        D = ctypes.c_double
        L.oracle_set_threads.argtypes = [I]
        L.oracle_triangle_rule.argtypes = [I, P, P]
        L.oracle_gauss_rule.argtypes = [I, P, P]
        L.oracle_duffy_rule.argtypes = [I, I, P, P, P]
        L.oracle_singular_pairs.argtypes = [P, I64, P, I64, I, P]
        L.oracle_singular_pairs.restype = I64
        for s in ("f64", "f32"):
            getattr(L, f"oracle_bm_assemble_{s}").argtypes = [P, I64, P, I64, I, I, D, D, D, P, I64, P, P, I, I, I, I,
                                                              I]
            getattr(L, f"oracle_operators_{s}").argtypes = [P, I64, P, I64, I, I, D, P, P, P, P, I]
            getattr(L, f"oracle_field_{s}").argtypes = [P, I64, P, I64, I, D, P, P, I64, P, I64, P, I]
            getattr(L, f"oracle_field_gradient_{s}").argtypes = [P, I64, P, I64, I, D, P, P, I64, P, I64, P, I]
            getattr(L, f"oracle_field_gradient_{s}").restype = I64
        L.oracle_identity_f64.argtypes = [P, I64, P, I64, I, P, P, I]
        L.oracle_orbit_weights.argtypes = [P, I64, P, I64, I, P]
        L.oracle_image_singular_pairs.argtypes = [P, I64, P, I64, I, I, I, P]
        L.oracle_image_singular_pairs.restype = I64
        L.oracle_surface_curls.argtypes = [P, I64, P, I64, P]
        L.oracle_singular_pair_points.argtypes = [P, I64, P, I64, I, I64, P, P, P]
        L.oracle_singular_pair_points.restype = I64
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def set_threads(n: int):
    lib().oracle_set_threads(int(n))


# ---- phasor convention (BeatEnginePhasor.jl:1-24) --------------------------------
def propagation_sign(convention=NEGATIVE_TIME_PHASOR) -> int:
    return 1 if convention == NEGATIVE_TIME_PHASOR else -1


def outgoing_wavenumber(k, convention=NEGATIVE_TIME_PHASOR):
    return propagation_sign(convention) * abs(k)


def neumann_scale(density, omega, convention=NEGATIVE_TIME_PHASOR):
    return complex(0.0, propagation_sign(convention) * density * omega)


def burton_miller_coupling(k, convention=NEGATIVE_TIME_PHASOR):
    return complex(0.0, propagation_sign(convention)) / abs(k)


# ---- rules ------------------------------------------------------------------------
def triangle_rule(order):
    pts = np.zeros((6, 2))
    w = np.zeros(6)
    n = lib().oracle_triangle_rule(order, _p(pts), _p(w))
    return pts[:n].copy(), w[:n].copy()


def gauss_rule(order):
    x = np.zeros(8)
    w = np.zeros(8)
    n = lib().oracle_gauss_rule(order, _p(x), _p(w))
    return x[:n].copy(), w[:n].copy()


def duffy_rule(order, kind):
    """kind: 1 coincident, 2 edge-adjacent, 3 vertex-adjacent (unremapped base rule)."""
    n = lib().oracle_duffy_rule(order, kind, None, None, None)
    tp = np.zeros((n, 2))
    sp = np.zeros((n, 2))
    w = np.zeros(n)
    lib().oracle_duffy_rule(order, kind, _p(tp), _p(sp), _p(w))
    return tp, sp, w


# normalized_symmetry_mode (JL/src/BeatEngineCore.jl:448-454): off, x, xy, ground
SYMMETRY_MODES = {"off": 0, "x": 1, "xy": 2, "ground": 3}


def _sym(mode):
    m = str(mode).strip().lower()
    if m not in SYMMETRY_MODES:
        raise ValueError(f"Unsupported image mode: {mode}. Expected off, x, xy, or ground.")
    return SYMMETRY_MODES[m]


def _mesh_args(vertices, faces):
    V = np.ascontiguousarray(vertices, dtype=np.float64)
    F = np.ascontiguousarray(faces, dtype=np.int32)
    return V, F


def singular_pairs(vertices, faces, order=4):
    V, F = _mesh_args(vertices, faces)
    n = lib().oracle_singular_pairs(_p(V), len(V), _p(F), len(F), order, None)
    out = np.zeros((n, 4), dtype=np.int32)
    lib().oracle_singular_pairs(_p(V), len(V), _p(F), len(F), order, _p(out))
    return out


def singular_pair_points(vertices, faces, index, order=4):
    """Physical Duffy points (x on the test face, y on the trial face) and weights
    (jacobian included) of singular pair `index` in the reference's pair order."""
    V, F = _mesh_args(vertices, faces)
    n = lib().oracle_singular_pair_points(_p(V), len(V), _p(F), len(F), order, index, None, None, None)
    x = np.zeros((n, 3))
    y = np.zeros((n, 3))
    w = np.zeros(n)
    lib().oracle_singular_pair_points(_p(V), len(V), _p(F), len(F), order, index, _p(x), _p(y), _p(w))
    return x, y, w


def orbit_weights(vertices, faces, symmetry="off"):
    """p1_symmetry_orbit_weights: 2 per active symmetry plane a vertex lies on."""
    V, F = _mesh_args(vertices, faces)
    out = np.zeros(len(V))
    lib().oracle_orbit_weights(_p(V), len(V), _p(F), len(F), _sym(symmetry), _p(out))
    return out


def image_singular_pairs(vertices, faces, symmetry, image, order=4):
    """(test, trial, kind, rule) of the image-singular pairs of one image transform."""
    V, F = _mesh_args(vertices, faces)
    n = lib().oracle_image_singular_pairs(_p(V), len(V), _p(F), len(F), order, _sym(symmetry), image, None)
    if n < 0:
        raise ValueError("no such image transform")
    out = np.zeros((n, 4), dtype=np.int32)
    lib().oracle_image_singular_pairs(_p(V), len(V), _p(F), len(F), order, _sym(symmetry), image, _p(out))
    return out


def surface_curls(vertices, faces):
    V, F = _mesh_args(vertices, faces)
    out = np.zeros((len(F), 3, 3))
    lib().oracle_surface_curls(_p(V), len(V), _p(F), len(F), _p(out))
    return out


# ---- assembly ---------------------------------------------------------------------
def bm_assemble(vertices, faces, k, coupling, q, quad_order=4, singular_order=4, precision="f64",
                test_range=None, singular=True, identity=True, symmetry="off"):
    """Fused Burton-Miller system (A, B): A = 0.5 M - D + eta H, B = (-S - eta(K' + 0.5 M)) q.

    k is the signed (outgoing) wavenumber, coupling the signed eta; q is F x nrhs.
    """
    V, F = _mesh_args(vertices, faces)
    nv, nf = len(V), len(F)
    q = np.asarray(q)
    if q.ndim == 1:
        q = q[:, None]
    if q.shape[0] != nf:
        raise ValueError("Fused CPU Burton-Miller assembly needs one Neumann row per DP0 dof.")
    ct = np.complex128 if precision == "f64" else np.complex64
    qf = np.asfortranarray(q.astype(ct))
    nrhs = q.shape[1]
    A = np.zeros((nv, nv), dtype=ct, order="F")
    B = np.zeros((nv, nrhs), dtype=ct, order="F")
    lo, hi = (0, nf) if test_range is None else test_range
    fn = lib().oracle_bm_assemble_f64 if precision == "f64" else lib().oracle_bm_assemble_f32
    fn(_p(V), nv, _p(F), nf, quad_order, singular_order, float(k), float(coupling.real), float(coupling.imag),
       _p(qf), nrhs, _p(A), _p(B), int(lo), int(hi), int(bool(singular)), int(bool(identity)), _sym(symmetry))
    return A, B


def operators(vertices, faces, k, quad_order=4, singular_order=4, precision="f64", symmetry="off"):
    """S (N x F), D (N x N), K' (N x F), H (N x N) with the singular pass (and the
    image passes, image-singular deltas and orbit row weights of a symmetry mode)."""
    V, F = _mesh_args(vertices, faces)
    nv, nf = len(V), len(F)
    ct = np.complex128 if precision == "f64" else np.complex64
    S = np.zeros((nv, nf), dtype=ct, order="F")
    K = np.zeros((nv, nf), dtype=ct, order="F")
    D = np.zeros((nv, nv), dtype=ct, order="F")
    H = np.zeros((nv, nv), dtype=ct, order="F")
    fn = lib().oracle_operators_f64 if precision == "f64" else lib().oracle_operators_f32
    fn(_p(V), nv, _p(F), nf, quad_order, singular_order, float(k), _p(S), _p(D), _p(K), _p(H), _sym(symmetry))
    return S, D, K, H


def identity(vertices, faces, quad_order=4, symmetry="off"):
    V, F = _mesh_args(vertices, faces)
    nv, nf = len(V), len(F)
    Mpp = np.zeros((nv, nv), order="F")
    Mpd = np.zeros((nv, nf), order="F")
    lib().oracle_identity_f64(_p(V), nv, _p(F), nf, quad_order, _p(Mpp), _p(Mpd), _sym(symmetry))
    return Mpp, Mpd


def bm_from_operators(S, D, K, H, Mpp, Mpd, coupling, q):
    """burton_miller_neumann_lhs/rhs (BeatEngineCpuSolve.jl:55-110)."""
    lhs = 0.5 * Mpp - D + coupling * H
    q = np.asarray(q)
    if q.ndim == 1:
        q = q[:, None]
    rhs = -(S @ q) - coupling * (K @ q) + (-0.5 * coupling) * (Mpd @ q)
    return lhs, rhs


def field(vertices, faces, k, pressure, q, points, quad_order=4, precision="f64", symmetry="off"):
    V, F = _mesh_args(vertices, faces)
    nv, nf = len(V), len(F)
    ct = np.complex128 if precision == "f64" else np.complex64
    p = np.asarray(pressure)
    qq = np.asarray(q)
    if p.ndim == 1:
        p = p[:, None]
    if qq.ndim == 1:
        qq = qq[:, None]
    p = np.asfortranarray(p.astype(ct))
    qq = np.asfortranarray(qq.astype(ct))
    pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
    out = np.zeros((len(pts), p.shape[1]), dtype=ct, order="F")
    fn = lib().oracle_field_f64 if precision == "f64" else lib().oracle_field_f32
    fn(_p(V), nv, _p(F), nf, quad_order, float(k), _p(p), _p(qq), p.shape[1], _p(pts), len(pts), _p(out),
       _sym(symmetry))
    return out


# ---- dense solve ------------------------------------------------------------------
def field_gradient(vertices, faces, k, pressure, q, points, quad_order=4, precision="f64", symmetry="off"):
    """evaluate_galerkin_field_gradient_cpu (JL/src/BeatEngineFieldGradient.jl:45-92): array
    (points, 4, rhs) of pressure and its x, y, z derivatives; k is the signed wavenumber.
    Raises like the reference when a point is too close to the boundary."""
    V, F = _mesh_args(vertices, faces)
    nv, nf = len(V), len(F)
    ct = np.complex128 if precision == "f64" else np.complex64
    p = np.asarray(pressure)
    qq = np.asarray(q)
    vec = p.ndim == 1
    if p.ndim == 1:
        p = p[:, None]
    if qq.ndim == 1:
        qq = qq[:, None]
    p = np.asfortranarray(p.astype(ct))
    qq = np.asfortranarray(qq.astype(ct))
    pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
    out = np.zeros((len(pts), 4, p.shape[1]), dtype=ct, order="F")
    fn = lib().oracle_field_gradient_f64 if precision == "f64" else lib().oracle_field_gradient_f32
    failed = fn(_p(V), nv, _p(F), nf, quad_order, float(k), _p(p), _p(qq), p.shape[1], _p(pts), len(pts), _p(out),
                _sym(symmetry))
    if failed:
        raise ValueError("BEM gradient sample is too close to a boundary for reliable evaluation.")
    return out[:, :, 0] if vec else out


def lu_solve(A, B):
    """LAPACK getrf/getrs with partial pivoting, one factorization for all columns
    (_beat_dense_lu_solve, BeatEngineDenseSolve.jl:1104-1110)."""
    import scipy.linalg as sl

    lu, piv = sl.lu_factor(A, check_finite=False)
    return sl.lu_solve((lu, piv), B, check_finite=False)


def refine_cpu_lu(matrix, rhs, residual_matrix=None, residual_rhs=None, max_steps=3, rtol=1e-10):
    """beat_refine_cpu_lu (BeatEngineCpuRefinement.jl:5-78): complex64 LU, complex128
    residual, per-column acceptance. Returns (x_complex64, report dict)."""
    import scipy.linalg as sl

    if not (1 <= max_steps <= 3):
        raise ValueError("max_steps must be in 1:3")
    if not (np.isfinite(rtol) and rtol > 0):
        raise ValueError("rtol must be finite and positive")
    matrix = np.asarray(matrix, dtype=np.complex64)
    n = matrix.shape[0]
    if matrix.shape[1] != n:
        raise ValueError("matrix must be square")
    vec = np.ndim(rhs) == 1
    b = np.asarray(rhs, dtype=np.complex64)
    b = b[:, None] if vec else b
    Ar = matrix if residual_matrix is None else np.asarray(residual_matrix)
    br = b if residual_rhs is None else np.asarray(residual_rhs).reshape(b.shape)
    if Ar.shape != matrix.shape:
        raise ValueError("residual matrix differs")
    if b.shape[0] != n or br.shape != b.shape:
        raise ValueError("right-hand sides differ")
    A64 = Ar.astype(np.complex128)
    b64 = br.astype(np.complex128)
    lu = sl.lu_factor(matrix, check_finite=False)

    def solve32(r):
        return sl.lu_solve(lu, r.astype(np.complex64), check_finite=False).astype(np.complex64)

    with np.errstate(all="ignore"):
        x = solve32(b).astype(np.complex128)
        rhs_norms = np.linalg.norm(b64, axis=0)

        def residuals(xx):
            r = b64 - A64 @ xx
            nr = np.linalg.norm(r, axis=0)
            out = np.where(rhs_norms == 0, np.where(nr == 0, 0.0, np.inf), nr / np.where(rhs_norms == 0, 1, rhs_norms))
            return r, out

        r, errors = residuals(x)
        history = [errors.copy()]
        steps = 0
        status = "step_limit"
        for step in range(1, max_steps + 1):
            if np.all(errors <= rtol):
                status = "converged"
                break
            delta = r.astype(np.complex64)
            delta[:, errors <= rtol] = 0
            delta = solve32(delta)
            nxt = x + delta.astype(np.complex128)
            _, nerr = residuals(nxt)
            improved = False
            for j in range(x.shape[1]):
                if np.isfinite(nerr[j]) and nerr[j] < errors[j]:
                    x[:, j] = nxt[:, j]
                    errors[j] = nerr[j]
                    improved = True
            steps = step
            r, errors = residuals(x)
            history.append(errors.copy())
            if not improved:
                status = "stagnated"
                break
        if np.all(errors <= rtol):
            status = "converged"
        solution = x.astype(np.complex64)
        _, returned = residuals(solution.astype(np.complex128))
    working_status = "nonfinite" if np.any(~np.isfinite(errors)) else status
    if np.any(~np.isfinite(returned)):
        status = "nonfinite"
    elif np.all(returned <= rtol):
        status = "converged"
    elif working_status == "converged":
        status = "rounding_limited"
    else:
        status = working_status
    out = solution[:, 0] if vec else solution
    return out, dict(steps=steps, status=status, working_status=working_status, rtol=rtol,
                     working_relative_residuals=errors, returned_relative_residuals=returned, history=history)

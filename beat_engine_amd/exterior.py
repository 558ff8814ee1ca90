"""Host-side mirror of the reference's exterior Burton-Miller interface over the
gfx950 library.

  solve_exterior_direct_hip        <- solve_exterior_direct_cuda
                                      (src/beat_engine/julia_local/BeatEngineCompiledDriver.jl:870-915)
  assemble_burton_miller_neumann_system_hip
                                   <- assemble_burton_miller_neumann_system_cpu
                                      (src/BeatEngineCpuBurtonMiller.jl:276-469)
  evaluate_galerkin_field_hip      <- evaluate_galerkin_field_cpu (src/BeatEngineCpuField.jl:1-49)
  assemble_regular_galerkin_operators_hip
                                   <- assemble_regular_galerkin_operators_cpu / _rocm
                                      (src/BeatEngineCpuAssembly.jl:739-1031,
                                       src/BeatEngineRocmAssembly.jl:16-161)
  evaluate_galerkin_field_gradient_hip
                                   <- evaluate_galerkin_field_gradient_cpu / _rocm
                                      (src/BeatEngineFieldGradient.jl:45-92)
  beat_refine_lu                   <- beat_refine_cpu_lu (src/BeatEngineCpuRefinement.jl:5-78)
  dense_solve_diagnostics          <- beat_dense_solve_diagnostics (src/BeatEngineDenseSolve.jl:1160-1196)

Arguments keep the reference's meaning: k is the wavenumber (normalised by
outgoing_wavenumber), q_neumann is F x drives, pressures come back per drive,
symmetry_mode is :off / :x / :xy / :ground (normalized_symmetry_mode,
src/BeatEngineCore.jl:448-460) with the mesh as the fundamental domain, used as
given (the reference's driver snaps it first, BeatEngineDriver.jl:927-928;
mesh.snap_symmetry_planes does the same here).
"""
from __future__ import annotations

import ctypes
import weakref

import numpy as np

from . import _native
from .mesh import BoundaryMesh, normalized_symmetry_mode
from .phasor import NEGATIVE_TIME_PHASOR, burton_miller_coupling, neumann_scale, outgoing_wavenumber

PRECISIONS = {
    "float64": _native.PREC_FP64,
    "float32_refined": _native.PREC_FP32_REFINED,
    "float32": _native.PREC_FP32,
}

# the autotuner's choices, in beat_autotune's numbering
LU_LAYOUTS = ("lookahead", "one_stream")
ASM_OCCUPANCIES = ("every_slot", "one_per_cu")
GEMM_TILES = ("64x64_k16", "64x64_k32", "128x64_k16", "128x64_k32")
REGULAR_TILES = ("unordered_16", "ordered_32", "ordered_16")


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _precision(p) -> int:
    if isinstance(p, int):
        return p
    key = str(p).lower()
    if key not in PRECISIONS:
        raise ValueError(f"precision must be one of {sorted(PRECISIONS)}")
    return PRECISIONS[key]


class Engine:
    """One device context: a stream and the workspace the solves reuse."""

    def __init__(self, device: int = 0):
        self._lib = _native.load()
        h = ctypes.c_void_p()
        rc = self._lib.beat_ctx_create(int(device), ctypes.byref(h))
        if rc != 0:
            raise _native.BeatError(f"beat_ctx_create({device}) failed with {rc}")
        self._h = h
        self._meshes = {}
        self._fin = weakref.finalize(self, self._lib.beat_ctx_destroy, h)

    def _check(self, rc, what):
        if rc != 0:
            msg = self._lib.beat_last_error(self._h).decode()
            err = _native.BeatError(f"{what} failed ({rc}): {msg}")
            err.code = rc
            raise err

    def close(self):
        for m in list(self._meshes.values()):
            m.close()
        self._meshes.clear()
        self._fin()

    # -- worker integration -----------------------------------------------------------
    def set_cancel_path(self, path):
        """Marker file that stops a sweep before its next frequency (None clears)."""
        self._check(self._lib.beat_ctx_set_cancel_path(self._h, None if path is None else str(path).encode()),
                    "beat_ctx_set_cancel_path")

    def debug_abort_waits(self, on: bool):
        """Test hook: while on, every factorization's cross-workgroup waits give up at
        once and the call raises (the error path of a wait that times out)."""
        self._check(self._lib.beat_ctx_debug_abort_waits(self._h, 1 if on else 0), "beat_ctx_debug_abort_waits")

    def debug_force_exchange(self, on: bool):
        """Test hook: while on, every solve's speculative factorization reports that
        partial pivoting exchanges rows, so the solve factors again with exchanges."""
        self._check(self._lib.beat_ctx_debug_force_exchange(self._h, 1 if on else 0), "beat_ctx_debug_force_exchange")

    def regular_symmetric(self, on: bool):
        """The regular assembly pass over unordered tile pairs (on, the default) or
        over ordered tile pairs (off); both give the same operators to rounding."""
        self._check(self._lib.beat_ctx_regular_symmetric(self._h, 1 if on else 0), "beat_ctx_regular_symmetric")

    def reclaim(self):
        """Frees the context's workspaces (reclaim_accelerator_memory!)."""
        self._check(self._lib.beat_ctx_reclaim(self._h), "beat_ctx_reclaim")

    def memory_info(self) -> dict:
        v = np.zeros(3, dtype=np.int64)
        self._check(self._lib.beat_ctx_memory_info(self._h, _ptr(v)), "beat_ctx_memory_info")
        return dict(workspace_bytes=int(v[0]), free_bytes=int(v[1]), total_bytes=int(v[2]))

    def capabilities(self) -> dict:
        return _native.capabilities(self._lib, self._h)

    # -- batching ---------------------------------------------------------------------
    def autotune(self, device_mesh: "DeviceMesh", precision="float32_refined", k=10.0, nrhs=8) -> dict:
        """Times the launch strategies and tile shapes on this mesh's size class in
        this precision and keeps the fastest of each (beat_autotune)."""
        t = np.zeros(11)
        ch = np.zeros(4, dtype=np.int32)
        rc = self._lib.beat_autotune(self._h, device_mesh._h, _precision(precision), float(k), int(nrhs), _ptr(t),
                                     _ptr(ch))
        self._check(rc, "beat_autotune")
        return dict(lu_layout=LU_LAYOUTS[int(ch[0])], batch_assembly=ASM_OCCUPANCIES[int(ch[1])],
                    gemm_tile=GEMM_TILES[int(ch[2])], regular_tiles=REGULAR_TILES[int(ch[3])],
                    seconds=dict(zip(LU_LAYOUTS + ASM_OCCUPANCIES + GEMM_TILES + REGULAR_TILES, t.tolist())))

    def set_tuning(self, nv: int, precision="float32_refined", lu_layout=None, batch_assembly=None, gemm_tile=None,
                   regular_tiles=None):
        """Sets the tuned choices of a precision and the size class of nv (the names
        autotune() reports; None: not tuned, the built-in default)."""
        def code(v, names):
            return -1 if v is None else names.index(v)
        ch = np.array([code(lu_layout, LU_LAYOUTS), code(batch_assembly, ASM_OCCUPANCIES),
                       code(gemm_tile, GEMM_TILES), code(regular_tiles, REGULAR_TILES)], dtype=np.int32)
        self._check(self._lib.beat_ctx_set_tuning(self._h, _precision(precision), int(nv), _ptr(ch)),
                    "beat_ctx_set_tuning")

    def batch(self, meshes, ks, couplings, qs, precision="float32_refined", max_steps=3, rtol=1e-10):
        """Solves job i = (meshes[i], ks[i], couplings[i], qs[i] (F_i x nrhs)) with the
        next job assembled while the current one is factored; returns the pressures
        (N_i x nrhs each) and the per-job reports."""
        nj = len(meshes)
        if nj == 0:
            raise ValueError("a batch needs at least one job")
        if len(qs) != nj:
            raise ValueError(f"qs must hold one Neumann block per mesh ({nj}), got {len(qs)}")
        qs = [np.asarray(q, dtype=np.complex128) for q in qs]
        for m, q in zip(meshes, qs):
            if q.shape[0] != m.num_faces or q.ndim > 2:
                raise ValueError("Fused Burton-Miller assembly needs one Neumann row per DP0 dof.")
        qs = [q.reshape(m.num_faces, -1) for m, q in zip(meshes, qs)]
        nrhs = qs[0].shape[1]
        if any(q.shape[1] != nrhs for q in qs):
            raise ValueError("every job needs the same number of right-hand sides")
        k = np.ascontiguousarray(ks, dtype=np.float64)
        c = np.asarray(couplings, dtype=np.complex128)
        if k.shape != (nj,) or c.shape != (nj,):
            raise ValueError(f"ks and couplings must have shape ({nj},), got {k.shape} and {c.shape}")
        qh = np.concatenate([np.asfortranarray(q).ravel(order="F") for q in qs])
        ph = np.zeros(sum(m.num_vertices for m in meshes) * nrhs, dtype=np.complex128)
        handles = (ctypes.c_void_p * nj)(*[m._h.value for m in meshes])
        er, ei = np.ascontiguousarray(c.real), np.ascontiguousarray(c.imag)
        reps = (_native.SolveReport * nj)()
        opts = _native.SolveOptions(int(max_steps), float(rtol))
        rc = self._lib.beat_bm_batch(self._h, nj, ctypes.cast(handles, ctypes.c_void_p), _precision(precision),
                                     _ptr(k), _ptr(er), _ptr(ei), _ptr(qh), nrhs, _ptr(ph), ctypes.byref(opts), reps)
        self._check(rc, "beat_bm_batch")
        out, off = [], 0
        for m in meshes:
            n = m.num_vertices * nrhs
            out.append(ph[off:off + n].reshape((m.num_vertices, nrhs), order="F"))
            off += n
        info = [dict(assembly_s=r.assembly_s, factor_s=r.factor_s, solve_s=r.solve_s,
                     factorization_count=r.factorization_count, refinement_steps=r.refine.steps,
                     refinement_status=_native.REFINE_STATUS[r.refine.status]) for r in reps]
        return out, info

    def upload(self, mesh: BoundaryMesh, quadrature_order: int = 4, singular_order: int = 4,
               symmetry_mode="off") -> "DeviceMesh":
        sym = normalized_symmetry_mode(symmetry_mode)
        key = (id(mesh), quadrature_order, singular_order, sym)
        dm = self._meshes.get(key)
        if dm is None or dm._mesh_ref() is not mesh:
            dm = DeviceMesh(self, mesh, quadrature_order, singular_order, sym)
            self._meshes[key] = dm
        return dm

    # -- dense linear algebra --------------------------------------------------------
    def dense_lu_solve(self, A, B, precision="float64"):
        A = np.asfortranarray(A, dtype=np.complex128)
        vec = np.ndim(B) == 1
        B = np.asfortranarray(np.reshape(B, (A.shape[0], -1)), dtype=np.complex128)
        if A.shape[0] != A.shape[1] or B.shape[0] != A.shape[0]:
            raise ValueError("dimension mismatch")
        X = np.zeros_like(B)
        rc = self._lib.beat_dense_lu_solve(self._h, _precision(precision), _ptr(A), A.shape[0], _ptr(B), B.shape[1],
                                           _ptr(X))
        self._check(rc, "beat_dense_lu_solve")
        return X[:, 0] if vec else X

    def dense_lu_factor(self, A, precision="float64"):
        """getrf on the GPU: (lu, piv, info) with LAPACK's layout (scipy.linalg.lu_factor's
        lu and 0-based piv) and info (first exactly zero pivot, 1-based, or 0)."""
        A = np.asfortranarray(A, dtype=np.complex128)
        n = A.shape[0]
        if A.ndim != 2 or A.shape[1] != n:
            raise ValueError("matrix must be square")
        lu = np.zeros((n, n), dtype=np.complex128, order="F")
        piv = np.zeros(n, dtype=np.int32)
        info = np.zeros(1, dtype=np.int32)
        rc = self._lib.beat_dense_lu_factor(self._h, _precision(precision), _ptr(A), n, _ptr(lu), _ptr(piv), _ptr(info))
        self._check(rc, "beat_dense_lu_factor")
        return lu, piv.astype(np.int64) - 1, int(info[0])

    def refine_lu(self, matrix, rhs, residual_matrix=None, residual_rhs=None, max_steps=3, rtol=1e-10):
        if not (1 <= max_steps <= 3):
            raise ValueError("max_steps must be in 1:3")
        if not (np.isfinite(rtol) and rtol > 0):
            raise ValueError("rtol must be finite and positive")
        M = np.asfortranarray(matrix, dtype=np.complex64)
        n = M.shape[0]
        if M.shape[1] != n:
            raise ValueError("matrix must be square")
        vec = np.ndim(rhs) == 1
        b = np.asfortranarray(np.reshape(rhs, (n, -1)) if not vec else np.asarray(rhs)[:, None], dtype=np.complex64)
        if b.shape[0] != n:
            raise ValueError("right-hand sides differ")
        Ar = None if residual_matrix is None else np.asfortranarray(residual_matrix, dtype=np.complex128)
        if Ar is not None and Ar.shape != M.shape:
            raise ValueError("residual matrix differs")
        br = None if residual_rhs is None else np.asfortranarray(np.reshape(residual_rhs, b.shape), dtype=np.complex128)
        nrhs = b.shape[1]
        x = np.zeros_like(b)
        work = np.zeros(nrhs)
        ret = np.zeros(nrhs)
        hist = np.zeros((max_steps + 1) * nrhs)
        rep = _native.RefineReport()
        rep.working_residuals = work.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        rep.returned_residuals = ret.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        rep.history = hist.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        rc = self._lib.beat_refine_lu(self._h, _ptr(M), n, _ptr(b), nrhs, None if Ar is None else _ptr(Ar),
                                      None if br is None else _ptr(br), int(max_steps), float(rtol), _ptr(x),
                                      ctypes.byref(rep))
        self._check(rc, "beat_refine_lu")
        report = dict(steps=rep.steps, status=_native.REFINE_STATUS[rep.status],
                      working_status=_native.REFINE_STATUS[rep.working_status], rtol=rep.rtol,
                      working_relative_residuals=work.copy(), returned_relative_residuals=ret.copy(),
                      history=[hist[i * nrhs:(i + 1) * nrhs].copy() for i in range(rep.history_len)])
        return (x[:, 0] if vec else x), report


class DeviceMesh:
    """Mesh resident on the device with its tiles, touching pairs and Duffy rules
    (and, with a symmetry mode, its image-touching pairs and orbit weights)."""

    def __init__(self, engine: Engine, mesh: BoundaryMesh, quadrature_order=4, singular_order=4, symmetry_mode="off"):
        self.engine = engine
        self._mesh_ref = weakref.ref(mesh)
        self.num_vertices = mesh.num_vertices
        self.num_faces = mesh.num_faces
        self.quadrature_order = quadrature_order
        self.singular_order = singular_order
        self.symmetry_mode = normalized_symmetry_mode(symmetry_mode)
        h = ctypes.c_void_p()
        rc = engine._lib.beat_mesh_upload_symmetric(engine._h, _ptr(mesh.vertices), mesh.num_vertices,
                                                    _ptr(mesh.faces), mesh.num_faces, int(quadrature_order),
                                                    int(singular_order), _native.SYMMETRY_CODES[self.symmetry_mode],
                                                    ctypes.byref(h))
        engine._check(rc, "beat_mesh_upload_symmetric")
        self._h = h
        self._fin = weakref.finalize(self, engine._lib.beat_mesh_destroy, h)

    def close(self):
        self._fin()

    def counts(self) -> dict:
        c = np.zeros(6, dtype=np.int64)
        self.engine._check(self.engine._lib.beat_mesh_counts(self._h, _ptr(c)), "beat_mesh_counts")
        s = np.zeros(4, dtype=np.int64)
        self.engine._check(self.engine._lib.beat_mesh_symmetry(self._h, _ptr(s)), "beat_mesh_symmetry")
        return dict(regular_pairs=int(c[0]), coincident=int(c[1]), edge=int(c[2]), vertex=int(c[3]),
                    tiles=int(c[4]), rule_points=int(c[5]), image_transforms=int(s[1]),
                    image_singular_pairs=int(s[2]), orbit_weighted_rows=int(s[3]))

    def _q(self, q):
        q = np.asarray(q)
        if q.ndim == 1:
            q = q[:, None]
        if q.shape[0] != self.num_faces:
            raise ValueError("Fused Burton-Miller assembly needs one Neumann row per DP0 dof.")
        return np.asfortranarray(q, dtype=np.complex128)

    def assemble(self, k, coupling, q, precision="float64"):
        qq = self._q(q)
        n = self.num_vertices
        A = np.zeros((n, n), dtype=np.complex128, order="F")
        B = np.zeros((n, qq.shape[1]), dtype=np.complex128, order="F")
        rc = self.engine._lib.beat_bm_assemble(self.engine._h, self._h, _precision(precision), float(k),
                                               float(coupling.real), float(coupling.imag), _ptr(qq), qq.shape[1],
                                               _ptr(A), _ptr(B))
        self.engine._check(rc, "beat_bm_assemble")
        return A, B

    def operators(self, k, precision="float64"):
        """(S, D, K', H): nv x nf, nv x nv, nv x nf, nv x nv complex128."""
        n, f = self.num_vertices, self.num_faces
        S = np.zeros((n, f), dtype=np.complex128, order="F")
        K = np.zeros((n, f), dtype=np.complex128, order="F")
        D = np.zeros((n, n), dtype=np.complex128, order="F")
        H = np.zeros((n, n), dtype=np.complex128, order="F")
        rc = self.engine._lib.beat_operators_assemble(self.engine._h, self._h, _precision(precision), float(k), _ptr(S),
                                                      _ptr(D), _ptr(K), _ptr(H))
        self.engine._check(rc, "beat_operators_assemble")
        return S, D, K, H

    def solve(self, k, coupling, q, precision="float32_refined", max_steps=3, rtol=1e-10):
        qq = self._q(q)
        n, nrhs = self.num_vertices, qq.shape[1]
        P = np.zeros((n, nrhs), dtype=np.complex128, order="F")
        opts = _native.SolveOptions(int(max_steps), float(rtol))
        rep = _native.SolveReport()
        work = np.zeros(nrhs)
        hist = np.zeros((max_steps + 1) * nrhs)
        rep.refine.working_residuals = work.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        rep.refine.returned_residuals = work.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        rep.refine.history = hist.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        prec = _precision(precision)
        rc = self.engine._lib.beat_bm_solve(self.engine._h, self._h, prec, float(k), float(coupling.real),
                                            float(coupling.imag), _ptr(qq), nrhs, _ptr(P), ctypes.byref(opts),
                                            ctypes.byref(rep))
        self.engine._check(rc, "beat_bm_solve")
        report = dict(assembly_s=rep.assembly_s, factor_s=rep.factor_s, solve_s=rep.solve_s,
                      regular_kernel_s=rep.regular_kernel_s,
                      factorization_count=rep.factorization_count, lu_panels=rep.lu_panels,
                      regular_pairs=rep.regular_pairs, singular_pairs=rep.singular_pairs)
        if prec == _native.PREC_FP32_REFINED:
            report["refinement"] = dict(
                steps=rep.refine.steps, status=_native.REFINE_STATUS[rep.refine.status],
                working_status=_native.REFINE_STATUS[rep.refine.working_status], rtol=rep.refine.rtol,
                working_relative_residuals=work.copy(),
                history=[hist[i * nrhs:(i + 1) * nrhs].copy() for i in range(rep.refine.history_len)])
        return P, report

    def sweep(self, ks, couplings, qs, precision="float32_refined", max_steps=3, rtol=1e-10):
        """Solves at every (k, coupling, q) with the next frequency assembled while the
        current one is factored; qs: (nfreq, F, nrhs). Returns pressures (nfreq, N,
        nrhs) and the per-frequency reports; after a cancellation (the engine's
        marker file) the unsolved frequencies report solved=False."""
        qs = np.asarray(qs, dtype=np.complex128)
        if qs.ndim == 2:
            qs = qs[:, :, None]
        if qs.ndim != 3:
            raise ValueError("qs must have shape (nfreq, F) or (nfreq, F, nrhs)")
        nfreq, nf, nrhs = qs.shape
        if nf != self.num_faces:
            raise ValueError("Fused Burton-Miller assembly needs one Neumann row per DP0 dof.")
        # per frequency a column-major F x nrhs block
        qh = np.ascontiguousarray(np.transpose(qs, (0, 2, 1)))
        ph = np.zeros((nfreq, nrhs, self.num_vertices), dtype=np.complex128)
        k = np.ascontiguousarray(ks, dtype=np.float64)
        c = np.asarray(couplings, dtype=np.complex128)
        if k.shape != (nfreq,) or c.shape != (nfreq,):
            raise ValueError(f"ks and couplings must have shape ({nfreq},), got {k.shape} and {c.shape}")
        er = np.ascontiguousarray(c.real)
        ei = np.ascontiguousarray(c.imag)
        reps = (_native.SolveReport * nfreq)()
        work = np.zeros((nfreq, nrhs))
        for i in range(nfreq):
            reps[i].refine.working_residuals = work[i].ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        opts = _native.SolveOptions(int(max_steps), float(rtol))
        rc = self.engine._lib.beat_bm_sweep(self.engine._h, self._h, _precision(precision), nfreq, _ptr(k), _ptr(er),
                                            _ptr(ei), _ptr(qh), nrhs, _ptr(ph), ctypes.byref(opts), reps)
        if rc != _native.BEAT_ERR_CANCELLED:
            self.engine._check(rc, "beat_bm_sweep")
        P = np.transpose(ph, (0, 2, 1))
        out = [dict(solved=r.factorization_count >= 1, assembly_s=r.assembly_s, factor_s=r.factor_s,
                    solve_s=r.solve_s, refinement_status=_native.REFINE_STATUS[r.refine.status],
                    refinement_steps=r.refine.steps, working_relative_residuals=work[i].copy())
               for i, r in enumerate(reps)]
        return P, out

    def spl_loss_gradient(self, ks, couplings, qs, points, target_db=90.0):
        """(loss, gradient (nv x 3), SPL (points x nfreq)) of
        sum_f sum_d (SPL_fd - target_db)^2 with one adjoint solve per frequency."""
        qs = np.asarray(qs, dtype=np.complex128)
        if qs.ndim == 1:
            qs = qs[None, :]
        nfreq = qs.shape[0]
        if qs.shape[1] != self.num_faces:
            raise ValueError("one Neumann value per face and frequency")
        qh = np.ascontiguousarray(qs)
        k = np.ascontiguousarray(ks, dtype=np.float64).reshape(nfreq)
        c = np.asarray(couplings, dtype=np.complex128).reshape(nfreq)
        er, ei = np.ascontiguousarray(c.real), np.ascontiguousarray(c.imag)
        pts = np.ascontiguousarray(np.reshape(points, (-1, 3)), dtype=np.float64)
        loss = ctypes.c_double(0.0)
        spl = np.zeros((len(pts), nfreq), order="F")
        grad = np.zeros((self.num_vertices, 3))
        rc = self.engine._lib.beat_adjoint_spl_gradient(self.engine._h, self._h, nfreq, _ptr(k), _ptr(er), _ptr(ei),
                                                        _ptr(qh), _ptr(pts), len(pts), float(target_db),
                                                        ctypes.byref(loss), _ptr(spl), _ptr(grad))
        self.engine._check(rc, "beat_adjoint_spl_gradient")
        return loss.value, grad, spl

    def field(self, k, pressure, q, points, precision="float64"):
        p = np.asarray(pressure)
        if p.ndim == 1:
            p = p[:, None]
        p = np.asfortranarray(p, dtype=np.complex128)
        qq = self._q(q)
        pts = np.ascontiguousarray(np.reshape(points, (-1, 3)), dtype=np.float64)
        out = np.zeros((len(pts), p.shape[1]), dtype=np.complex128, order="F")
        rc = self.engine._lib.beat_field_eval(self.engine._h, self._h, _precision(precision), float(k), _ptr(p),
                                              _ptr(qq), p.shape[1], _ptr(pts), len(pts), _ptr(out))
        self.engine._check(rc, "beat_field_eval")
        return out

    def field_gradient(self, k, pressure, q, points, precision="float64"):
        """(points, 4, rhs): pressure and its x, y, z derivatives."""
        p = np.asarray(pressure)
        if p.ndim == 1:
            p = p[:, None]
        p = np.asfortranarray(p, dtype=np.complex128)
        qq = self._q(q)
        pts = np.ascontiguousarray(np.reshape(points, (-1, 3)), dtype=np.float64)
        out = np.zeros((len(pts), 4, p.shape[1]), dtype=np.complex128, order="F")
        rc = self.engine._lib.beat_field_gradient_eval(self.engine._h, self._h, _precision(precision), float(k),
                                                       _ptr(p), _ptr(qq), p.shape[1], _ptr(pts), len(pts), _ptr(out))
        self.engine._check(rc, "beat_field_gradient_eval")
        return out


_default_engine = None


def default_engine() -> Engine:
    global _default_engine
    if _default_engine is None:
        _default_engine = Engine(0)
    return _default_engine


def _rule_order(rule) -> int:
    if rule is None:
        return 4
    return int(getattr(rule, "order", rule))


def assemble_burton_miller_neumann_system_hip(mesh: BoundaryMesh, q_neumann, k, rule=4, *, singular_order=4,
                                              precision="float64", symmetry_mode="off",
                                              convention=NEGATIVE_TIME_PHASOR, engine=None):
    """(matrix, rhs) of the fused system; k is normalised to the outgoing wavenumber."""
    eng = engine or default_engine()
    dm = eng.upload(mesh, _rule_order(rule), singular_order, symmetry_mode)
    ks = outgoing_wavenumber(k, convention)
    A, B = dm.assemble(ks, burton_miller_coupling(ks, convention), q_neumann, precision)
    return dict(matrix=A, rhs=B, assembly_mode="hip_fused_burton_miller", on_gpu=True, **dm.counts())


def assemble_regular_galerkin_operators_hip(mesh: BoundaryMesh, k, rule=4, *, singular_order=4, precision="float64",
                                            symmetry_mode="off", convention=NEGATIVE_TIME_PHASOR, engine=None):
    """The operators of the reference's operator_matrices mode as a named tuple-like
    dict (single_layer, double_layer, adjoint_double_layer, hypersingular) plus the
    pair counts; k is normalised to the outgoing wavenumber."""
    eng = engine or default_engine()
    dm = eng.upload(mesh, _rule_order(rule), singular_order, symmetry_mode)
    S, D, K, H = dm.operators(outgoing_wavenumber(k, convention), precision)
    c = dm.counts()
    return dict(single_layer=S, double_layer=D, adjoint_double_layer=K, hypersingular=H,
                regular_pairs=c["regular_pairs"], singular_pairs=c["coincident"] + c["edge"] + c["vertex"],
                image_singular_pairs=c["image_singular_pairs"], on_gpu=True, gpu_backend="rocm")


def solve_exterior_direct_hip(mesh: BoundaryMesh, p1_space, dp0_space, neumann_values, wavenumber, rule=4, *,
                              precision="float32_refined", singular_order=4, refinement_steps=3,
                              refinement_rtol=1e-10, symmetry_mode="off", convention=NEGATIVE_TIME_PHASOR,
                              engine=None):
    """Same contract as solve_exterior_direct_cuda: one fused assembly and one
    factorization for every excitation; returns (pressures, assembly_s, solve_s,
    rhs) with rhs = (mode, operator_bytes), plus the dense-solve diagnostics
    as a fifth element."""
    del p1_space, dp0_space  # P1 = vertices and DP0 = faces, as in the reference
    eng = engine or default_engine()
    dm = eng.upload(mesh, _rule_order(rule), singular_order, symmetry_mode)
    q = np.stack([np.asarray(v, dtype=np.complex128) for v in neumann_values], axis=1)
    ks = outgoing_wavenumber(wavenumber, convention)
    P, rep = dm.solve(ks, burton_miller_coupling(ks, convention), q, precision, refinement_steps, refinement_rtol)
    pressures = [P[:, j].copy() for j in range(P.shape[1])]
    diag = dense_solve_diagnostics(rep, precision)
    return pressures, rep["assembly_s"], rep["factor_s"] + rep["solve_s"], dict(mode="fused", operator_bytes=0), diag


def evaluate_galerkin_field_hip(eval_points, mesh: BoundaryMesh, pressure, q_neumann, k, rule=4, *,
                                precision="float64", symmetry_mode="off", convention=NEGATIVE_TIME_PHASOR,
                                engine=None):
    eng = engine or default_engine()
    dm = eng.upload(mesh, _rule_order(rule), 4, symmetry_mode)
    out = dm.field(outgoing_wavenumber(k, convention), pressure, q_neumann, eval_points, precision)
    return out[:, 0] if np.ndim(pressure) == 1 else out


def evaluate_galerkin_field_gradient_hip(points, mesh: BoundaryMesh, pressure, neumann, k, *, precision="float64",
                                         symmetry_mode="off", convention=NEGATIVE_TIME_PHASOR, engine=None):
    """Same contract as evaluate_galerkin_field_gradient_rocm: a (points, 4) array of
    pressure and its x, y, z derivatives (a third axis per column of pressure), with
    the 4th-order rule and the near-panel correction."""
    eng = engine or default_engine()
    dm = eng.upload(mesh, 4, 4, symmetry_mode)
    out = dm.field_gradient(outgoing_wavenumber(k, convention), pressure, neumann, points, precision)
    return out[:, :, 0] if np.ndim(pressure) == 1 else out


def far_field_spl_loss_and_gradient(mesh: BoundaryMesh, frequencies_hz, normal_velocity, points, *,
                                    target_db=90.0, sound_speed=343.0, density=1.21, quadrature_order=4,
                                    singular_order=4, convention=NEGATIVE_TIME_PHASOR, engine=None):
    """Loss sum_f sum_d (SPL_fd - target_db)^2 at the observation points and its
    gradient with respect to every vertex coordinate (nv x 3), for a fixed normal
    velocity per face (F, or nfreq x F). The reverse rule of the mesh -> loss map."""
    eng = engine or default_engine()
    dm = DeviceMesh(eng, mesh, quadrature_order, singular_order)
    try:
        freqs = np.atleast_1d(np.asarray(frequencies_hz, dtype=np.float64))
        vn = np.asarray(normal_velocity, dtype=np.complex128)
        if vn.ndim == 1:
            vn = np.broadcast_to(vn, (len(freqs), mesh.num_faces))
        ks, etas, qs = [], [], []
        for f, v in zip(freqs, vn):
            w = 2 * np.pi * f
            k = outgoing_wavenumber(w / sound_speed, convention)
            ks.append(k)
            etas.append(burton_miller_coupling(k, convention))
            qs.append(neumann_scale(density, w, convention) * v)
        return dm.spl_loss_gradient(ks, etas, np.stack(qs), points, target_db)
    finally:
        dm.close()


def beat_refine_lu(matrix, rhs, residual_matrix=None, residual_rhs=None, max_steps=3, rtol=1e-10, engine=None):
    return (engine or default_engine()).refine_lu(matrix, rhs, residual_matrix, residual_rhs, max_steps, rtol)


def worker_backend_availability(engine=None) -> dict:
    """The worker's backend announcement (worker_backend_availability.jl:1-26) for the
    ROCm backend this library provides: available when a device context opens."""
    try:
        eng = engine or default_engine()
        caps = eng.capabilities()
        rocm = {"available": True, "reason": "", "phasor_conventions": caps["phasor_conventions"],
                "arch": caps["arch"], "linear_solver": caps["linear_solver"]}
        field_quantities = {"rocm": caps["field_quantities"]}
    except Exception as exc:  # noqa: BLE001 - reported, as the reference does
        rocm = {"available": False, "reason": str(exc),
                "phasor_conventions": [NEGATIVE_TIME_PHASOR, "exp(+i omega t)"]}
        field_quantities = {"rocm": []}
    return {"backends": {"rocm": rocm}, "field_quantities": field_quantities}


def dense_solve_diagnostics(report: dict, precision="float32_refined") -> dict:
    """Keys of beat_dense_solve_diagnostics plus the exterior diagnostics the
    driver records (burton_miller_assembly, linear_solver, factorization_count)."""
    d = {
        "burton_miller_assembly": "direct_system",
        "linear_solver": "hip_mfma_lu",
        "factorization_count": report.get("factorization_count", 1),
        "assembly_s": report.get("assembly_s"),
        "factor_s": report.get("factor_s"),
        "solve_s": report.get("solve_s"),
        "precision": str(precision),
    }
    ref = report.get("refinement")
    if ref is not None:
        res = [float(v) if np.isfinite(v) else None for v in ref["working_relative_residuals"]]
        d.update({
            "dense_solve_refinement_steps": ref["steps"],
            "dense_solve_refinement_status": ref["status"],
            "dense_solve_refinement_working_status": ref["working_status"],
            "dense_solve_refinement_tolerance": ref["rtol"],
            "dense_solve_refinement_working_residuals": res,
            "dense_solve_refinement_returned_residuals": res,
            "dense_solve_refinement_operator": "float64",
        })
    return d

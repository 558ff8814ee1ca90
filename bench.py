"""Frequency-solve throughput of the exterior Burton-Miller path on one MI355X.

Workload (one step = one freq-solve): level-5 icosphere, radius 0.1 m, vertices
jittered radially by N(0, 0.02 h) (20480 triangles, 10242 P1 dofs), 8 complex
N(0,1) normal-velocity drives, q = i rho omega v (rho 1.21, c 343), quadrature
order 4 / singular order 4, frequencies cycling through 20 log-spaced values
100 Hz - 5 kHz. Each step assembles the system in complex128, factors it once
in complex64 (trailing update on FP32 MFMA) and refines all 8 right-hand sides
with complex128 residuals to 1e-10 (the FP32-MFMA + refinement path). The timed
steps run as one frequency sweep in which the next frequency is assembled while
the current one is factored (beat_bm_sweep_device); the same solves one call at a
time are reported beside it. The plain complex128 GPU path is timed on the same
inputs for the ratio, and the refined pressures are compared with it (relative
L2 per drive). The CPU baseline solves one frequency of the workload whole on the
host (the oracle's fused FP64 assembly of every pair, LAPACK getrf/getrs), and the
GPU's pressures at that frequency are compared with it (rel_err_vs_cpu_oracle).
Beside the headline the line carries the plain complex64 path (fp32_gpu), the
precision form (a 32k-face cylinder, 64 drives, 10k field points in FP32 and
FP64), the differentiation form (an adjoint gradient) and the batching form (32
small meshes x 32 frequencies, 4 drives, jobs/s).

Prints one JSON line. --gpus above 1 is not supported: this build runs on one GPU.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "freq-solves/s @N=20480,8 RHS: FP32-MFMA+IR vs FP64 GPU at ≤1e-8 rel error"
FP64_VALU_PEAK_TFLOPS = 78.6
FP32_MFMA_PEAK_TFLOPS = 157.3
FLOP_PER_QPAIR = 72  # fused Burton-Miller flops per quadrature-point pair
# HBM bytes per launch of the FP64 regular kernel (the unordered-pair pass) at L5,
# from two rocprofv3 --pmc runs (FETCH_SIZE x 2.0, WRITE_SIZE x 1.0: the factors make
# a 1 GiB device copy read as 1 GiB each way); profiles/bm_regular_sym_L5_pmc_traffic.md
REGULAR_L5_TRAFFIC_BYTES = 22.67e9


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--level", type=int, default=5)
    ap.add_argument("--nrhs", type=int, default=8)
    ap.add_argument("--fp64-steps", type=int, default=6)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--batch-form", action="store_true",
                    help="time the batching form (32 meshes x 128 frequencies, 4 RHS) instead; no JSON line")
    ap.add_argument("--batch-freqs", type=int, default=128)
    ap.add_argument("--json-batch-freqs", type=int, default=32,
                    help="frequencies per mesh of the batching form in the JSON line (32 meshes each)")
    ap.add_argument("--counter-run", action="store_true",
                    help="for rocprofv3 --pmc: a 1 GiB device copy (counter calibration), then one complex128 "
                         "freq-solve of the workload; prints no JSON line")
    ap.add_argument("--no-autotune", action="store_true",
                    help="skip the autotuner (built-in launch strategies and tile shapes)")
    ap.add_argument("--no-rocsolver", action="store_true",
                    help="skip timing rocSOLVER getrf (the vendor LU) at the workload's size")
    ap.add_argument("--no-extra-forms", action="store_true",
                    help="skip the plain complex64 sweep, the precision form's field, the differentiation form "
                         "and the batching form")
    ap.add_argument("--cpu-freq-index", type=int, default=11,
                    help="which of the 20 frequencies the CPU baseline solves (and the GPU result is checked at)")
    return ap.parse_args()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def counts_flops(dm):
    """Algorithmic FLOP of one regular-pair kernel launch: pairs x rule points^2 x 72."""
    c = dm.counts()
    return c["regular_pairs"] * c["rule_points"] ** 2 * FLOP_PER_QPAIR


def host_threads():
    """Host threads for the CPU baseline: the process's CPU affinity, capped by
    OMP_NUM_THREADS when it is set (the GPU box sets it to its CPU share)."""
    cores = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        cores = min(cores, int(env))
    return cores


def cpu_baseline(mesh, k, eta, q):
    """One whole freq-solve of the workload on the host, unextrapolated: the
    oracle's fused FP64 assembly over every test face (the C++ restatement of the
    reference's CPU direct_system path, OpenMP over coloured test elements), then
    LAPACK zgetrf and zgetrs for every drive (scipy's OpenBLAS 0.3.29, the version
    the reference pins). Returns the timing record and the pressures."""
    import scipy.linalg as sl
    from threadpoolctl import threadpool_info, threadpool_limits

    import oracle

    cores = host_threads()
    oracle.set_threads(cores)
    t = time.perf_counter()
    A, B = oracle.bm_assemble(mesh.vertices, mesh.faces, k, eta, q)
    t_asm = time.perf_counter() - t
    with threadpool_limits(cores):
        blas = [(i.get("internal_api"), i.get("version"), i.get("num_threads")) for i in threadpool_info()
                if i.get("user_api") == "blas"]
        t = time.perf_counter()
        lu = sl.lu_factor(A, check_finite=False, overwrite_a=True)
        t_lu = time.perf_counter() - t
        t = time.perf_counter()
        p = sl.lu_solve(lu, B, check_finite=False)
        t_sv = time.perf_counter() - t
    del A, lu
    total = t_asm + t_lu + t_sv
    return {
        "value": 1.0 / total,
        "unit": "freq-solves/s",
        "cores": cores,
        "kind": "port",
        "sample": (f"one whole freq-solve of the workload at k = {k:.4f} (unextrapolated): oracle fused FP64 "
                   f"assembly of all {mesh.num_faces} test faces with {cores} OpenMP threads ({t_asm:.1f} s), "
                   f"complex128 getrf at n = {mesh.num_vertices} ({t_lu:.2f} s) and getrs of {q.shape[1]} drives "
                   f"({t_sv:.3f} s) with BLAS {blas}"),
        "s_per_freq": {"assembly": t_asm, "lu": t_lu, "solve": t_sv},
        "omp_threads": cores,
        "blas": [{"api": b[0], "version": b[1], "threads": b[2]} for b in blas],
    }, p


def rocsolver_getrf_ms(n, reps=2):
    """Best of reps timed calls of rocSOLVER cgetrf and zgetrf on an n x n N(0,1)
    complex matrix (after one untimed call). rocSOLVER getrf is the factorization
    the reference's ROCm path calls (rocm_dense_lu!, BeatEngineRocmOperators.jl:
    132-159); it is timed here for comparison only and is no part of this path."""
    import ctypes

    import torch
    libdir = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib")
    try:
        rb = ctypes.CDLL(os.path.join(libdir, "librocblas.so"))
        rs = ctypes.CDLL(os.path.join(libdir, "librocsolver.so"))
    except OSError as e:
        return {"error": str(e)}
    h = ctypes.c_void_p()
    if rb.rocblas_create_handle(ctypes.byref(h)) != 0:
        return {"error": "rocblas_create_handle failed"}
    rb.rocblas_set_stream.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    rb.rocblas_set_stream(h, ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    ipiv = torch.empty(n, dtype=torch.int32, device="cuda")
    info = torch.empty(1, dtype=torch.int32, device="cuda")
    out = {"n": n}
    for name, dt in (("cgetrf", torch.complex64), ("zgetrf", torch.complex128)):
        fn = getattr(rs, "rocsolver_" + name)
        fn.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                       ctypes.c_void_p]
        fn.restype = ctypes.c_int
        gen = torch.Generator(device="cuda")
        gen.manual_seed(10242)
        a0 = torch.randn(n, n, dtype=dt, device="cuda", generator=gen)
        a = torch.empty_like(a0)
        times = []
        for r in range(reps + 1):
            a.copy_(a0)
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            st = fn(h, n, n, ctypes.c_void_p(a.data_ptr()), n, ctypes.c_void_p(ipiv.data_ptr()),
                    ctypes.c_void_p(info.data_ptr()))
            e1.record()
            e1.synchronize()
            if st != 0 or int(info.item()) != 0:
                out[name + "_error"] = f"status {st} info {int(info.item())}"
                break
            if r:
                times.append(e0.elapsed_time(e1))
        if times:
            out[name + "_ms"] = min(times)
        del a0, a
    rb.rocblas_destroy_handle(h)
    torch.cuda.empty_cache()
    return out


def batch_form(eng, nfreq):
    """32 meshes of uneven size (jittered icospheres and graded capped cylinders, 500-6000
    triangles, from seeds) x nfreq log-spaced frequencies 50 Hz - 8 kHz, 4 RHS each, as
    one batch per precision (grouped batches of equal-size systems, one hipGraph per
    group shape); jobs/s per precision and the refined-vs-FP64 error."""
    from beat_engine_amd.mesh import capped_cylinder, jittered_icosphere
    from beat_engine_amd.phasor import burton_miller_coupling, neumann_scale

    rng = np.random.default_rng(32)
    meshes = []
    for i in range(32):
        target = int(rng.integers(500, 6001))
        if i % 2 == 0:
            level = 2 if target < 900 else 3 if target < 2500 else 4
            m = jittered_icosphere(level, 0.1, 0.02, seed=100 + i)
            if m.num_faces > 6000:
                m = jittered_icosphere(3, 0.1, 0.02, seed=100 + i)
        else:
            m = capped_cylinder(target_faces=target, seed=200 + i)
        meshes.append(m)
    freqs = np.logspace(np.log10(50.0), np.log10(8000.0), nfreq)
    dms = [eng.upload(m) for m in meshes]
    tune = eng.autotune(dms[-1], "float32_refined", 2 * np.pi * 1000 / 343.0, 4)
    log(f"autotune (size class of {meshes[-1].num_vertices} dofs): {tune}")
    jobs = []
    for m, dm in zip(meshes, dms):
        v = rng.normal(size=(m.num_faces, 4)) + 1j * rng.normal(size=(m.num_faces, 4))
        for f in freqs:
            w = 2 * np.pi * f
            k = w / 343.0
            jobs.append((dm, k, burton_miller_coupling(k), neumann_scale(1.21, w) * v))
    faces = [m.num_faces for m in meshes]
    log(f"{len(jobs)} jobs, faces {min(faces)}-{max(faces)}, dofs {sum(m.num_vertices for m in meshes)} total")
    res = {}
    for prec in ("float32_refined", "float64"):
        t = time.perf_counter()
        P, info = eng.batch([j[0] for j in jobs], [j[1] for j in jobs], [j[2] for j in jobs], [j[3] for j in jobs],
                            prec)
        el = time.perf_counter() - t
        res[prec] = (P, el)
        bad = sum(1 for r in info if prec == "float32_refined" and r["refinement_status"] != "converged")
        log(f"{prec}: {len(jobs)} jobs in {el:.2f} s = {len(jobs) / el:.1f} jobs/s; not converged: {bad}")
    err = max(float(np.linalg.norm(a - b) / np.linalg.norm(b))
              for a, b in zip(res["float32_refined"][0], res["float64"][0]))
    return {"jobs": len(jobs), "meshes": 32, "frequencies": int(nfreq), "nrhs": 4,
            "faces_range": [min(faces), max(faces)],
            "refined_jobs_per_s": len(jobs) / res["float32_refined"][1],
            "fp64_jobs_per_s": len(jobs) / res["float64"][1],
            "max_rel_err_refined_vs_fp64": err, "autotune": tune,
            "timed": "whole batch call per precision: host staging of q and p, grouped assembly and graphs"}


FP32_VALU_TFLOPS = 78.6  # FP32 VALU without packed instructions (this build uses none), as FP64
FIELD_FLOP_PER_EVAL = 40  # per (point, source) pair and right-hand side, the SURVEY's count


def precision_form(eng):
    """Graded capped cylinder (0.3 m x 0.08 m, ~32768 triangles), 64 drives on the graded
    cap, refined solve at 2 kHz, then the field of all 64 at 10000 points on a 2 m
    hemisphere in FP32 and FP64 (wall time of beat_field_eval, host arrays in and out)."""
    from beat_engine_amd.mesh import capped_cylinder
    from beat_engine_amd.phasor import burton_miller_coupling, neumann_scale

    mesh = capped_cylinder(target_faces=32768)
    f = 2000.0
    w = 2 * math.pi * f
    k = w / 343.0
    rng = np.random.default_rng(64)
    v = np.zeros((mesh.num_faces, 64), complex)
    cap = mesh.physical_tags == 2
    v[cap] = rng.normal(size=(cap.sum(), 64)) + 1j * rng.normal(size=(cap.sum(), 64))
    q = neumann_scale(1.21, w) * v
    dm = eng.upload(mesh, 4, 4)
    t = time.perf_counter()
    P, rep = dm.solve(k, burton_miller_coupling(k), q, "float32_refined", 3, 1e-10)
    t_solve = time.perf_counter() - t
    n = 10000
    i = np.arange(n) + 0.5
    z = i / n
    rr = np.sqrt(1 - z * z)
    phi = math.pi * (3 - math.sqrt(5)) * i
    pts = 2.0 * np.stack([rr * np.cos(phi), rr * np.sin(phi), z], axis=1) + np.array([0.0, 0.0, 0.15])
    out = {"workload": (f"capped cylinder {mesh.num_faces} triangles ({mesh.num_vertices} P1), 64 drives on the "
                        f"graded cap, {f:.0f} Hz, field at {n} points on a 2 m hemisphere"),
           "refined_solve_s": t_solve, "refinement_status": rep["refinement"]["status"],
           "refinement_steps": rep["refinement"]["steps"],
           "stage_s": {"assembly": rep["assembly_s"], "lu_c64": rep["factor_s"], "solve_refine": rep["solve_s"]}}
    nsrc = mesh.num_faces * 6
    res = {}
    for prec in ("float32", "float64"):
        dm.field(k, P[:, :2], q[:, :2], pts[:64], prec)  # warm
        t = time.perf_counter()
        res[prec] = dm.field(k, P, q, pts, prec)
        el = time.perf_counter() - t
        evals = n * nsrc * 64
        out[f"field_{prec}_s"] = el
        out[f"field_{prec}_evals_per_s"] = evals / el
        out[f"field_{prec}_tflops"] = evals * FIELD_FLOP_PER_EVAL / el / 1e12
    out["field_frac_of_valu_peak"] = {"float32": out["field_float32_tflops"] / FP32_VALU_TFLOPS,
                                      "float64": out["field_float64_tflops"] / FP64_VALU_PEAK_TFLOPS}
    out["field_float32_vs_float64_rel"] = float(np.linalg.norm(res["float32"] - res["float64"]) /
                                                np.linalg.norm(res["float64"]))
    dm.close()
    return out


def differentiation_form(eng):
    """Level-4 icosphere morphed by 5 seeded spherical harmonics, 8 frequencies
    1-4 kHz, 64 far-field directions at 2 m: loss and gradient by the adjoint."""
    from beat_engine_amd.mesh import fibonacci_sphere, morphed_icosphere
    from beat_engine_amd.phasor import burton_miller_coupling, neumann_scale

    mesh = morphed_icosphere(4, 0.1, 5, 0.01, seed=5120)
    freqs = np.linspace(1000.0, 4000.0, 8)
    ks = 2 * np.pi * freqs / 343.0
    etas = np.array([burton_miller_coupling(k) for k in ks])
    qs = np.stack([np.full(mesh.num_faces, neumann_scale(1.21, 2 * np.pi * f)) for f in freqs])
    pts = fibonacci_sphere(64, 2.0)
    dm = eng.upload(mesh, 4, 4)
    dm.spl_loss_gradient(ks[:1], etas[:1], qs[:1], pts, 90.0)  # warm
    t = time.perf_counter()
    L, g, _ = dm.spl_loss_gradient(ks, etas, qs, pts, 90.0)
    el = time.perf_counter() - t
    dm.close()
    return {"workload": (f"morphed level-4 icosphere ({mesh.num_faces} triangles, {mesh.num_vertices} nodes), "
                         "8 frequencies 1-4 kHz, 64 directions at 2 m, loss sum (SPL - 90 dB)^2"),
            "loss_and_gradient_s": el, "s_per_frequency": el / len(freqs), "loss": float(L),
            "gradient_norm": float(np.linalg.norm(g))}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); the multi-GPU sharding of this "
              "path is described in DESIGN.md and not built.", file=sys.stderr)
        sys.exit(2)
    import ctypes

    import torch

    from beat_engine_amd import _native
    from beat_engine_amd.exterior import Engine
    from beat_engine_amd.mesh import jittered_icosphere
    from beat_engine_amd.phasor import burton_miller_coupling, neumann_scale

    if not torch.cuda.is_available():
        raise RuntimeError("bench.py needs a GPU")
    if args.batch_form:
        eng = Engine(0)
        print(json.dumps({"batch_form": batch_form(eng, args.batch_freqs)}), flush=True)
        eng.close()
        return
    mesh = jittered_icosphere(args.level, 0.1, 0.02, seed=20480)
    nrhs = args.nrhs
    if args.counter_run:
        a = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b
        dm = Engine(0).upload(mesh)
        rng = np.random.default_rng(20480)
        v = rng.standard_normal((mesh.num_faces, nrhs)) + 1j * rng.standard_normal((mesh.num_faces, nrhs))
        k = 2 * np.pi * 1000.0 / 343.0
        q = neumann_scale(1.21, 2 * np.pi * 1000.0) * v
        _, rep = dm.solve(k, burton_miller_coupling(k), q, precision="float64")
        log(f"counter run: regular kernel {rep['regular_kernel_s']:.4f} s")
        return
    freqs = np.logspace(np.log10(100.0), np.log10(5000.0), 20)
    rng = np.random.default_rng(20480)
    v = rng.normal(size=(mesh.num_faces, nrhs)) + 1j * rng.normal(size=(mesh.num_faces, nrhs))
    eng = Engine(0)
    t = time.perf_counter()
    dm = eng.upload(mesh, 4, 4)
    log(f"mesh upload {time.perf_counter() - t:.2f} s: {dm.counts()}")
    lib = _native.load()
    dev = torch.device("cuda:0")
    # q per frequency, column-major F x nrhs complex128 == row-major (nrhs, F)
    qs = []
    for f in freqs:
        q = neumann_scale(1.21, 2 * math.pi * f) * v
        qs.append(torch.from_numpy(np.ascontiguousarray(q.T)).to(dev))
    p_mixed = torch.zeros((nrhs, mesh.num_vertices), dtype=torch.complex128, device=dev)
    p_fp64 = torch.zeros_like(p_mixed)
    torch.cuda.synchronize()
    # the autotuner's launch strategies and tile shapes for this size, per precision
    # (every path, the complex128 one included, then runs with its own fastest)
    tuning = {}
    if not args.no_autotune:
        k_tune = 2 * math.pi * 1000.0 / 343.0
        for name in ("float32_refined", "float64", "float32"):
            t = time.perf_counter()
            ch = eng.autotune(dm, name, k_tune, nrhs)
            tuning[name] = {kk: v for kk, v in ch.items() if kk != "seconds"}
            log(f"autotune {name} ({time.perf_counter() - t:.1f} s): {tuning[name]} "
                f"{ {kk: round(v * 1e3, 2) for kk, v in ch['seconds'].items()} } ms")

    def step(i, prec, out):
        f = freqs[i % len(freqs)]
        k = 2 * math.pi * f / 343.0
        eta = burton_miller_coupling(k)
        opts = _native.SolveOptions(8, 1e-10)
        rep = _native.SolveReport()
        rc = lib.beat_bm_solve_device(eng._h, dm._h, prec, k, eta.real, eta.imag,
                                      ctypes.c_void_p(qs[i % len(freqs)].data_ptr()), nrhs,
                                      ctypes.c_void_p(out.data_ptr()), ctypes.byref(opts), ctypes.byref(rep))
        eng._check(rc, "beat_bm_solve_device")
        return rep

    def sweep(first, count, prec, out):
        """count consecutive freq-solves (cycling the 20 frequencies) as one pipelined sweep"""
        idx = [(first + i) % len(freqs) for i in range(count)]
        ks = np.array([2 * math.pi * freqs[j] / 343.0 for j in idx])
        etas = [burton_miller_coupling(k) for k in ks]
        er = np.ascontiguousarray([e.real for e in etas])
        ei = np.ascontiguousarray([e.imag for e in etas])
        q_all = torch.stack([qs[j] for j in idx])
        torch.cuda.synchronize()
        reps = (_native.SolveReport * count)()
        opts = _native.SolveOptions(8, 1e-10)
        t = time.perf_counter()
        rc = lib.beat_bm_sweep_device(eng._h, dm._h, prec, count, ks.ctypes.data_as(ctypes.c_void_p),
                                      er.ctypes.data_as(ctypes.c_void_p), ei.ctypes.data_as(ctypes.c_void_p),
                                      ctypes.c_void_p(q_all.data_ptr()), nrhs, ctypes.c_void_p(out.data_ptr()),
                                      ctypes.byref(opts), reps)
        torch.cuda.synchronize()
        el = time.perf_counter() - t
        eng._check(rc, "beat_bm_sweep_device")
        return el, list(reps)

    p_sweep = torch.zeros((max(args.steps, args.warmup, args.fp64_steps, 1), nrhs, mesh.num_vertices),
                          dtype=torch.complex128, device=dev)
    if args.warmup:
        _, wr = sweep(0, args.warmup, _native.PREC_FP32_REFINED, p_sweep)
        for i, r in enumerate(wr):
            log(f"warmup {i}: asm {r.assembly_s:.3f} lu {r.factor_s:.3f} solve {r.solve_s:.3f} s")
    # timed region: args.steps freq-solves of the 20-frequency sweep, pipelined
    torch.cuda.synchronize()
    el, reps = sweep(0, args.steps, _native.PREC_FP32_REFINED, p_sweep)
    for i, r in enumerate(reps):
        log(f"step {i} f={freqs[i % 20]:.1f}: asm {r.assembly_s:.3f} (regular {r.regular_kernel_s:.3f}) "
            f"lu {r.factor_s:.3f} solve+refine {r.solve_s:.3f} s, refine steps {r.refine.steps} "
            f"status {_native.REFINE_STATUS[r.refine.status]}")
    bad = [i for i, r in enumerate(reps) if r.refine.status != 1]
    if bad:
        raise RuntimeError(f"refinement did not converge at steps {bad}")
    sec = el / args.steps
    # the same freq-solves one call at a time (no overlap of assembly and LU): the
    # per-kernel figures (roofline, LU rate, stage times) come from these, where no
    # other stream shares the device
    t = time.perf_counter()
    serial = []
    for i in range(min(3, args.steps)):
        serial.append(step(i, _native.PREC_FP32_REFINED, p_mixed))
    torch.cuda.synchronize()
    sec_serial = (time.perf_counter() - t) / len(serial)
    last = (args.steps - 1) % len(freqs)

    # plain complex128 GPU path on the same inputs, pipelined the same way
    el64, reps64 = sweep(0, args.fp64_steps, _native.PREC_FP64, p_sweep)
    sec64 = el64 / max(1, args.fp64_steps)
    # refined vs FP64 at the same frequency
    step(last, _native.PREC_FP32_REFINED, p_mixed)
    rep64 = step(last, _native.PREC_FP64, p_fp64)
    torch.cuda.synchronize()
    a = p_mixed.cpu().numpy()
    b = p_fp64.cpu().numpy()
    err = float(max(np.linalg.norm(a[j] - b[j]) / np.linalg.norm(b[j]) for j in range(nrhs)))
    log(f"fp64 path {sec64:.3f} s/freq; refined serial {sec_serial:.3f} s/freq; refined vs fp64 max rel L2 {err:.2e}")
    extra = {}
    if not args.no_extra_forms:
        # the plain complex64 path (the reference's float32 option) on the same inputs
        el32, reps32 = sweep(0, args.fp64_steps, _native.PREC_FP32, p_sweep)
        p_32 = torch.zeros_like(p_fp64)
        rep32 = step(last, _native.PREC_FP32, p_32)
        torch.cuda.synchronize()
        c = p_32.cpu().numpy()
        err32 = float(max(np.linalg.norm(c[j] - b[j]) / np.linalg.norm(b[j]) for j in range(nrhs)))
        extra["fp32_gpu"] = {
            "freq_solves_per_s": args.fp64_steps / el32,
            "rel_err_vs_fp64_gpu": err32,
            "stage_s": {"assembly": rep32.assembly_s, "regular_kernel": rep32.regular_kernel_s,
                        "lu_c64": rep32.factor_s, "solve": rep32.solve_s},
            "sweep_stage_s": {"assembly": float(np.mean([r.assembly_s for r in reps32])),
                              "lu_c64": float(np.mean([r.factor_s for r in reps32])),
                              "solve": float(np.mean([r.solve_s for r in reps32]))},
            "regular_kernel_tflops": counts_flops(dm) / rep32.regular_kernel_s / 1e12,
        }
        log("fp32 path:", extra["fp32_gpu"])

    if not args.no_extra_forms:
        t = time.perf_counter()
        extra["precision_form"] = precision_form(eng)
        log(f"precision form ({time.perf_counter() - t:.1f} s):", extra["precision_form"])
        t = time.perf_counter()
        extra["differentiation_form"] = differentiation_form(eng)
        log(f"differentiation form ({time.perf_counter() - t:.1f} s):", extra["differentiation_form"])
        t = time.perf_counter()
        extra["batch_form"] = batch_form(eng, args.json_batch_freqs)
        log(f"batch form ({time.perf_counter() - t:.1f} s):", extra["batch_form"])

    reg_s = float(np.mean([r.regular_kernel_s for r in serial]))
    counts = dm.counts()
    qpairs = counts["regular_pairs"] * counts["rule_points"] ** 2
    flops = qpairs * FLOP_PER_QPAIR
    achieved = flops / reg_s / 1e12
    n_p1 = mesh.num_vertices
    lu_flop = 8.0 * n_p1 ** 3 / 3.0
    lu_s = float(np.mean([r.factor_s for r in serial]))
    out = {
        "metric": METRIC,
        "value": 1.0 / sec,
        "unit": "freq-solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": sec * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c128 assembly + c64 LU + c128 refinement",
        "data": "synthetic",
        "config": {
            "workload": (f"level-{args.level} jittered icosphere r=0.1 m ({mesh.num_faces} triangles, "
                         f"{mesh.num_vertices} P1), {nrhs} RHS, 20 log-spaced f 100-5000 Hz, q4/s4"),
            "precision": "float32_refined",
            "refinement_rtol": 1e-10,
        },
        "fp64_gpu_freq_solves_per_s": 1.0 / sec64,
        "serial_freq_solves_per_s": 1.0 / sec_serial,
        "speedup_vs_fp64_gpu": sec64 / sec,
        "rel_err_vs_fp64_gpu": err,
        "stage_s": {
            "assembly": float(np.mean([r.assembly_s for r in serial])),
            "regular_kernel": reg_s,
            "lu_c64": lu_s,
            "solve_refine": float(np.mean([r.solve_s for r in serial])),
        },
        "sweep_stage_s": {
            "assembly": float(np.mean([r.assembly_s for r in reps])),
            "lu_c64": float(np.mean([r.factor_s for r in reps])),
            "solve_refine": float(np.mean([r.solve_s for r in reps])),
            "fp64_assembly": float(np.mean([r.assembly_s for r in reps64])),
            "fp64_lu": float(np.mean([r.factor_s for r in reps64])),
            "fp64_solve": float(np.mean([r.solve_s for r in reps64])),
        },
        "element_pair_interactions_per_s": counts["regular_pairs"] / reg_s,
        "roofline": {
            "kernel": "bm_regular_sym_kernel<double,6>",
            "bound": "valu",
            "achieved": achieved,
            "peak": FP64_VALU_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP64_VALU_PEAK_TFLOPS,
            "traffic": REGULAR_L5_TRAFFIC_BYTES if args.level == 5 else None,
            "algorithmic_flop_per_launch": flops,
        },
        "lu_c64_mfma": {
            "flop": lu_flop,
            "achieved": lu_flop / lu_s / 1e12,
            "peak": FP32_MFMA_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": lu_flop / lu_s / 1e12 / FP32_MFMA_PEAK_TFLOPS,
            # flop is the conventional 8 n^3 / 3 of a complex LU; the refined path's
            # trailing updates form each complex product from three real ones, so
            # the MFMA units execute about 3/4 of it
            "complex_product": "three real products (refined path)",
        },
    }
    if tuning:
        out["autotune"] = tuning
    out.update(extra)
    if not args.no_rocsolver:
        rs = rocsolver_getrf_ms(n_p1)
        rs["ours_c64_ms"] = lu_s * 1e3  # the refined path's LU: three-real-product updates
        rs["ours_c128_ms"] = rep64.factor_s * 1e3
        if "cgetrf_ms" in rs:
            rs["c64_speedup"] = rs["cgetrf_ms"] / rs["ours_c64_ms"]
        if "zgetrf_ms" in rs:
            rs["c128_speedup"] = rs["zgetrf_ms"] / rs["ours_c128_ms"]
        log("rocSOLVER getrf comparison:", rs)
        out["lu_vs_rocsolver"] = rs
    if not args.no_cpu_baseline:
        # one frequency of the workload solved whole on the host, and the GPU's
        # refined and complex128 pressures at that frequency compared with it
        ci = args.cpu_freq_index % len(freqs)
        fc = freqs[ci]
        kc = 2 * math.pi * fc / 343.0
        etac = burton_miller_coupling(kc)
        step(ci, _native.PREC_FP32_REFINED, p_mixed)
        step(ci, _native.PREC_FP64, p_fp64)
        torch.cuda.synchronize()
        g_ref = p_mixed.cpu().numpy().T
        g_64 = p_fp64.cpu().numpy().T
        eng.close()
        del qs, p_mixed, p_fp64, p_sweep
        torch.cuda.empty_cache()
        log(f"CPU baseline: one whole freq-solve at {fc:.1f} Hz on {host_threads()} threads")
        cb, p_cpu = cpu_baseline(mesh, kc, etac, neumann_scale(1.21, 2 * math.pi * fc) * v)
        errs = [float(np.linalg.norm(g_ref[:, j] - p_cpu[:, j]) / np.linalg.norm(p_cpu[:, j])) for j in range(nrhs)]
        errs64 = [float(np.linalg.norm(g_64[:, j] - p_cpu[:, j]) / np.linalg.norm(p_cpu[:, j])) for j in range(nrhs)]
        log(f"CPU baseline {cb['s_per_freq']}; refined GPU vs CPU oracle max rel L2 {max(errs):.2e}, "
            f"complex128 GPU {max(errs64):.2e}")
        out["cpu_baseline"] = cb
        out["speedup_vs_cpu"] = out["value"] / cb["value"]
        out["rel_err_vs_cpu_oracle"] = max(errs)
        out["rel_err_vs_cpu_oracle_detail"] = {"frequency_hz": float(fc), "refined_per_rhs": errs,
                                               "fp64_gpu_max": max(errs64)}
    else:
        eng.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

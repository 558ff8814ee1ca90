/* C ABI of the gfx950 exterior Burton-Miller library (libbeat_bem_hip.so).
 *
 * The reference's exterior solve reaches its accelerator backends from
 * solve_exterior_request (src/beat_engine/julia_local/BeatEngineCompiledDriver.jl:1049-1635);
 * its direct_system branch calls solve_exterior_direct_cuda (:870-915), which
 * returns (pressures, assembly_s, solve_s, rhs). These entry points are what a
 * Julia `ccall` sibling, solve_exterior_direct_hip, binds to with the same
 * contract. Conventions:
 *   - complex arrays are interleaved (re, im) doubles, column-major, in the
 *     caller's numbering (P1 dof = vertex, DP0 dof = face), 0-based;
 *   - the wavenumber is the signed outgoing k (outgoing_wavenumber) and the
 *     coupling the signed eta (burton_miller_coupling), BeatEnginePhasor.jl:1-24;
 *   - q is the Neumann data F x nrhs, already scaled (neumann_scale * v_n);
 *   - every call returns BEAT_OK (0) or a negative code; beat_last_error gives
 *     the message. Host pointers are only touched during the call. One call at
 *     a time per context; contexts hold no thread-local state.
 */
#ifndef BEAT_BEM_H
#define BEAT_BEM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BEAT_OK 0
#define BEAT_ERR_ARG (-1)
#define BEAT_ERR_HIP (-2)
#define BEAT_ERR_ALLOC (-3)
#define BEAT_ERR_SINGULAR (-4)
#define BEAT_ERR_CANCELLED (-5)

/* Arithmetic of a solve. FP64: assembly, LU and solve in complex128.
 * FP32_REFINED: complex128 assembly held for the residual, complex64 LU
 * (trailing update on FP32 MFMA) and complex128 iterative refinement; the
 * returned pressure is the complex128 working solution.
 * FP32: complex64 assembly, LU and solve, as the reference's float32 option. */
#define BEAT_PREC_FP64 0
#define BEAT_PREC_FP32_REFINED 1
#define BEAT_PREC_FP32 2

/* Refinement status values (beat_refine_cpu_lu's symbols,
 * src/beat_engine/julia_local/src/BeatEngineCpuRefinement.jl:5-78). */
#define BEAT_REFINE_NONE 0
#define BEAT_REFINE_CONVERGED 1
#define BEAT_REFINE_STAGNATED 2
#define BEAT_REFINE_STEP_LIMIT 3
#define BEAT_REFINE_ROUNDING_LIMITED 4
#define BEAT_REFINE_NONFINITE 5

typedef struct beat_ctx beat_ctx;
typedef struct beat_mesh beat_mesh;

/* Refinement report. The arrays are caller-owned and may be NULL:
 * working/returned residuals hold nrhs values, history (max_steps+1)*nrhs
 * (row e = entry e of the reference's history). */
typedef struct {
  int steps;
  int status;
  int working_status;
  int history_len;
  double rtol;
  double* working_residuals;
  double* returned_residuals;
  double* history;
} beat_refine_report;

typedef struct {
  double assembly_s; /* zero fill, regular, singular and identity passes */
  double factor_s;   /* LU (and the complex128 -> complex64 copy when refined) */
  double solve_s;    /* triangular solves and refinement */
  double regular_kernel_s; /* the regular-pair assembly kernel alone */
  int factorization_count; /* 1; 2 when the first, speculative factorization (no row
                              exchanges, each column checked against partial pivoting's
                              choice) found a column where partial pivoting exchanges
                              rows and the matrix was factored again with exchanges */
  int lu_panels;
  int64_t regular_pairs;
  int64_t singular_pairs;
  beat_refine_report refine; /* filled for BEAT_PREC_FP32_REFINED */
} beat_solve_report;

typedef struct {
  int max_steps; /* refinement steps, 1..8 (reference default 3) */
  double rtol;   /* relative residual target per column (reference 1e-10) */
} beat_solve_options;

/* Context on one device (stream, workspace). Replaces the implicit device
 * state the reference's ROCm backend keeps in AMDGPU.jl. */
int beat_ctx_create(int device, beat_ctx** out);
void beat_ctx_destroy(beat_ctx* ctx);
const char* beat_last_error(const beat_ctx* ctx);

/* Worker integration (the compiled worker's cancellation, cleanup and capability
 * advertisement: cancel_requested, BeatEngineCompiledDriver.jl:1199-1200,1270;
 * reclaim_accelerator_memory!, :3951-3971 and the cleanup policy, :4308-4349;
 * worker_backend_availability, worker_backend_availability.jl:1-26).
 * beat_ctx_set_cancel_path: a marker file whose existence stops a sweep before its
 * next frequency (BEAT_ERR_CANCELLED; the reports of the solved frequencies carry
 * factorization_count >= 1, the rest 0); NULL clears it.
 * beat_ctx_reclaim: waits for the context's streams and frees its workspaces
 * (meshes stay resident). beat_ctx_memory_info: info[0..2] = workspace bytes,
 * device free bytes, device total bytes. beat_capabilities: a JSON object of what
 * this library evaluates (ctx may be NULL: no device query); returns the buffer
 * length needed including the terminating zero and writes at most len bytes. */
int beat_ctx_set_cancel_path(beat_ctx* ctx, const char* path);
/* Test hook (no counterpart in the reference): while on, every factorization
 * starts with its device abort word set, so each bounded cross-workgroup wait of
 * the factorization and of its triangular solves gives up at once; the kernels
 * end normally and the call returns BEAT_ERR_HIP with a message (the same path a
 * wait that times out after 4 s takes). For tests of that error path only. */
int beat_ctx_debug_abort_waits(beat_ctx* ctx, int on);
/* Test hook (no counterpart in the reference): while on, the speculative first
 * factorization of every solve (no row exchanges, each column checked against
 * partial pivoting's choice) reports that an exchange is needed, so the solve
 * takes its second factorization with the exchange panels from the matrix as
 * assembled (factorization_count 2). For tests of that path only. */
int beat_ctx_debug_force_exchange(beat_ctx* ctx, int on);
/* Tuning and test switch (no counterpart in the reference): on (the default, or
 * BEAT_REGULAR_SYMMETRIC unset / nonzero at context creation), the regular pass
 * of the fused assembly runs over unordered tile pairs, evaluating the kernel
 * once per point pair for both orders of a face pair; off, over ordered tile
 * pairs. The results agree to rounding; image passes always run ordered. */
int beat_ctx_regular_symmetric(beat_ctx* ctx, int on);
int beat_ctx_reclaim(beat_ctx* ctx);
int beat_ctx_memory_info(beat_ctx* ctx, int64_t* info);
int64_t beat_capabilities(const beat_ctx* ctx, char* buf, int64_t len);

/* Mesh, P1/DP0 spaces, regular rule, touching-pair cache and Duffy rules, once
 * per mesh. Replaces BoundaryMesh (BeatEngineCore.jl:344-364), build_p1_space /
 * build_dp0_space (:445-446), triangle_rule (:660), build_singular_correction_cache
 * (:1049) and build_rocm_regular_assembly_cache (BeatEngineRocmRegular.jl:134-205).
 * verts: nv x 3 row-major; faces: nf x 3 row-major, 0-based, outward winding. */
int beat_mesh_upload(beat_ctx* ctx, const double* verts, int64_t nv, const int32_t* faces, int64_t nf,
                     int quad_order, int singular_order, beat_mesh** out);
void beat_mesh_destroy(beat_mesh* mesh);
/* counts[0..5] = regular pairs, coincident, edge, vertex pairs, tiles, regular-rule points */
int beat_mesh_counts(const beat_mesh* mesh, int64_t* counts);

/* Symmetry modes: the mesh is the fundamental domain and the solve adds its
 * mirror images (normalized_symmetry_mode / symmetry_image_transforms,
 * BeatEngineCore.jl:448-495). X: the plane x = 0; XY: x = 0 and y = 0 (three
 * images); GROUND: a rigid ground plane y = 0. */
#define BEAT_SYM_OFF 0
#define BEAT_SYM_X 1
#define BEAT_SYM_XY 2
#define BEAT_SYM_GROUND 3

/* beat_mesh_upload with a symmetry mode: image passes over every (test,
 * reflected trial) pair, Duffy-minus-regular deltas for the image pairs that
 * touch (image_singular_candidates, BeatEngineCore.jl:1152-1166;
 * _beat_cpu_bm_image_singular_delta_test!, BeatEngineCpuBurtonMiller.jl:232-265),
 * P1 orbit weights on the rows of plane vertices (p1_symmetry_orbit_weights,
 * BeatEngineCore.jl:566-580) and image sources in the field. The vertices are
 * used as given: snap plane vertices to the plane first, as the reference's
 * driver does with snap_symmetry_planes (BeatEngineDriver.jl:927). */
int beat_mesh_upload_symmetric(beat_ctx* ctx, const double* verts, int64_t nv, const int32_t* faces, int64_t nf,
                               int quad_order, int singular_order, int symmetry, beat_mesh** out);
/* info[0..3] = symmetry mode, image transforms, image-singular pairs, rows with an orbit weight */
int beat_mesh_symmetry(const beat_mesh* mesh, int64_t* info);

/* The assembled system A (nv x nv) and B (nv x nrhs), complex128, without
 * solving: the operator-level parity point. Replaces
 * assemble_burton_miller_neumann_system_cpu (BeatEngineCpuBurtonMiller.jl:276-469)
 * and assemble_burton_miller_neumann_system_columns_cuda
 * (BeatEngineCudaBurtonMiller.jl:785). precision: BEAT_PREC_FP64 or BEAT_PREC_FP32. */
int beat_bm_assemble(beat_ctx* ctx, const beat_mesh* mesh, int precision, double k, double eta_re, double eta_im,
                     const double* q, int64_t nrhs, double* a_out, double* b_out);

/* The four Galerkin operators separately (operator_matrices): S and K' (nv x nf,
 * P1 rows, DP0 columns), D and H (nv x nv), complex128 column-major in the
 * caller's numbering, with the mesh's symmetry images and orbit row weights.
 * Replaces assemble_regular_galerkin_operators_cpu with the singular pass
 * (BeatEngineCpuAssembly.jl:739-1031) and the ROCm native operator assembly
 * (_assemble_regular_galerkin_operators_rocm_native, BeatEngineRocmAssembly.jl:16-161;
 * its optional near-field corrections are not part of this entry point).
 * precision: BEAT_PREC_FP64 or BEAT_PREC_FP32 (the singular pass is FP64). */
int beat_operators_assemble(beat_ctx* ctx, const beat_mesh* mesh, int precision, double k, double* s_out,
                            double* d_out, double* kp_out, double* h_out);

/* Assemble, factor once and solve every right-hand side: the body of
 * solve_exterior_direct_cuda (BeatEngineCompiledDriver.jl:870-915) and
 * solve_burton_miller_neumann_system_cpu_with_report (BeatEngineCpuBurtonMiller.jl:490-508).
 * pressure_out: nv x nrhs complex128. opts and rep may be NULL. */
int beat_bm_solve(beat_ctx* ctx, const beat_mesh* mesh, int precision, double k, double eta_re, double eta_im,
                  const double* q, int64_t nrhs, double* pressure_out, const beat_solve_options* opts,
                  beat_solve_report* rep);

/* Same with device-resident q (complex128 nf x nrhs) and pressure (complex128
 * nv x nrhs), both column-major device pointers on the context's device. */
int beat_bm_solve_device(beat_ctx* ctx, const beat_mesh* mesh, int precision, double k, double eta_re,
                         double eta_im, const void* q_dev, int64_t nrhs, void* pressure_dev,
                         const beat_solve_options* opts, beat_solve_report* rep);

/* A frequency sweep on one mesh: for each of nfreq frequencies (signed k[i],
 * coupling eta_re[i] + i eta_im[i]) the body of beat_bm_solve_device, with two
 * systems in flight: frequency i + 1 is assembled on a second stream while
 * frequency i is factored and solved (the reference's producer/consumer sweep,
 * start_sweep_assembly_pipeline / take_sweep_assembly!, BeatEngineCore.jl:1536-1570,
 * around the per-frequency solve of solve_exterior_request,
 * BeatEngineCompiledDriver.jl:1337-1346). q_dev: nfreq consecutive complex128 nf x nrhs
 * column-major blocks; p_dev: nfreq consecutive nv x nrhs blocks (device
 * pointers). reps: NULL or nfreq reports. Device memory: two assembled systems. */
int beat_bm_sweep_device(beat_ctx* ctx, const beat_mesh* mesh, int precision, int64_t nfreq, const double* k,
                         const double* eta_re, const double* eta_im, const void* q_dev, int64_t nrhs, void* p_dev,
                         const beat_solve_options* opts, beat_solve_report* reps);
/* Same with host arrays: q nfreq x (nf x nrhs) and pressure_out nfreq x (nv x nrhs),
 * complex128 column-major blocks. */
int beat_bm_sweep(beat_ctx* ctx, const beat_mesh* mesh, int precision, int64_t nfreq, const double* k,
                  const double* eta_re, const double* eta_im, const double* q, int64_t nrhs, double* pressure_out,
                  const beat_solve_options* opts, beat_solve_report* reps);

/* Autotuner (no counterpart in the reference): times the launch strategies and
 * tile shapes on this mesh's size class (floor(log2 nv)) in this precision, keeps
 * the fastest of each in the context for later solves, sweeps and batches of that
 * precision and size class, and reports them. In order:
 *   - the regular pass's tiles, by the assembly's own time: unordered pairs over
 *     16-face sub-tiles (0), ordered pairs over 32-face trial tiles (1) or over
 *     16-face trial tiles (2);
 *   - the factorization's streams, by its own time: look-ahead streams (0) or one
 *     stream (1);
 *   - the LU's trailing-update tile shape: 64 x 64-complex tiles with K chunks of
 *     16 (0) or 32 (1), 128 x 64 with 16 (2) or 32 (3), by the factorization's time,
 *     or for systems the grouped batches take (nv <= 4096, nrhs <= 8) by the time
 *     of a grouped batch of 8 solves on this mesh (its batched factorization);
 *   - the next assembly's occupancy in a three-solve batch: every slot (0) or one
 *     workgroup per CU (1).
 * k: a wavenumber for the trial solves. timings (may be NULL): 11 seconds, best of
 * two timed repeats after a warm-up (INFINITY when not available): [0..1] stream
 * layouts, [2..3] occupancies, [4..7] trailing-update tile shapes, [8..10]
 * regular-pass tiles. choice[0..3] = stream layout, occupancy, trailing-update
 * tile shape, regular-pass tiles. */
int beat_autotune(beat_ctx* ctx, const beat_mesh* mesh, int precision, double k, int64_t nrhs, double* timings,
                  int* choice);

/* Sets the tuned choices of a precision and the size class of nv, in
 * beat_autotune's numbering (choice[0..3] = stream layout, occupancy,
 * trailing-update tile shape, regular-pass tiles; -1: not tuned, the built-in
 * default), for instance to apply what an earlier process's autotuner chose (no
 * counterpart in the reference). */
int beat_ctx_set_tuning(beat_ctx* ctx, int precision, int64_t nv, const int* choice);

/* A batch of solves over several meshes (batching form: many meshes of uneven
 * size, each at its frequencies): job i solves meshes[i] at k[i], eta[i] with the
 * next job assembled while job i is factored, as in beat_bm_sweep. q: per job an
 * nf_i x nrhs complex128 column-major block, consecutive; pressure_out: per job
 * nv_i x nrhs, consecutive. reps: NULL or njobs reports. */
int beat_bm_batch(beat_ctx* ctx, int64_t njobs, const beat_mesh* const* meshes, int precision, const double* k,
                  const double* eta_re, const double* eta_im, const double* q, int64_t nrhs, double* pressure_out,
                  const beat_solve_options* opts, beat_solve_report* reps);

/* Field at npts points (npts x 3 row-major) from pressure (nv x nrhs) and q
 * (nf x nrhs): evaluate_galerkin_field_cpu (BeatEngineCpuField.jl:1-49) /
 * evaluate_galerkin_field_rocm (BeatEngineRocmField.jl:150). out: npts x nrhs.
 * precision: BEAT_PREC_FP64 or BEAT_PREC_FP32. */
int beat_field_eval(beat_ctx* ctx, const beat_mesh* mesh, int precision, double k, const double* pressure,
                    const double* q, int64_t nrhs, const double* pts, int64_t npts, double* out);

/* beat_field_eval on device pointers, allocating nothing (no counterpart in the
 * reference; for callers that cannot allocate device memory during the call:
 * code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). pressure_dev: nv x nrhs, q_dev: nf x nrhs, out_dev: npts x nrhs,
 * complex128 column-major; pts_dev: npts x 3 doubles row-major; all on the
 * context's device. Every staging buffer (internal-order inputs, field sources,
 * the accumulated field) lives in one file-scope device array of 64 MiB per GPU,
 * held until the process ends; a call that would need more (it takes about 43 MB for a
 * 20480-face mesh, 8 right-hand sides and 10000 points in FP64) returns
 * BEAT_ERR_ARG without touching the GPU. All calls on one GPU share that array:
 * calls that overlap on one GPU (from two contexts, or two threads) race. The
 * call returns when the output is written. */
int beat_field_eval_device_noalloc(beat_ctx* ctx, const beat_mesh* mesh, int precision, double k,
                                   const void* pressure_dev, const void* q_dev, int64_t nrhs, const void* pts_dev,
                                   int64_t npts, void* out_dev);

/* Pressure and its x, y, z derivatives at npts points (pressure_gradient_v1):
 * evaluate_galerkin_field_gradient_cpu / _rocm (BeatEngineFieldGradient.jl:45-92,
 * BeatEngineRocmFieldGradient.jl:20-47), field_value_gradient (:2-13) over the
 * field sources (with the mesh's symmetry images), plus the near-panel correction
 * (correct_field_gradient!, :45-73): for every face image whose centroid lies
 * within 3 diameters of a point, the FP64 adaptive integral (subdivision until
 * diameter <= 0.4 distance and |k| diameter <= 0.7, at most 12 levels) replaces
 * the 6-point rule. Arguments as beat_field_eval; out: npts x 4 x nrhs
 * (pressure, d/dx, d/dy, d/dz), column-major. A point the 12 levels cannot
 * resolve returns BEAT_ERR_ARG with the reference's message. The reference
 * requires the 4th-order rule for gradients (BeatEngineCompiledDriver.jl:4168-4171):
 * upload the mesh with quad_order 4. */
int beat_field_gradient_eval(beat_ctx* ctx, const beat_mesh* mesh, int precision, double k, const double* pressure,
                             const double* q, int64_t nrhs, const double* pts, int64_t npts, double* out);

/* Far-field SPL loss and its gradient with respect to the mesh vertex positions
 * (no counterpart in the reference, which has no shape derivative; its SPL is
 * spl_for_points, BeatEngineDriver.jl:657-663). For each frequency f (signed k[f],
 * coupling eta[f], Neumann data q: nf complex128 per frequency, nfreq consecutive
 * blocks) the complex128 system is assembled, factored and solved, the field is
 * evaluated at the npts observation points (npts x 3 row-major), and
 *   loss = sum_f sum_d (20 log10(|u_fd| / 20e-6) - target_db)^2.
 * The gradient comes from one adjoint solve per frequency with the same LU
 * (A^T mu = F^T conj(dL/du*)) and kernels that contract mu^T (dA p - db) and the
 * field derivative over element pairs without forming dA; q is held fixed.
 * spl_out (npts x nfreq column-major) may be NULL; grad_out: nv x 3 row-major,
 * caller's vertex order. Meshes uploaded with a symmetry mode are rejected. */
int beat_adjoint_spl_gradient(beat_ctx* ctx, const beat_mesh* mesh, int64_t nfreq, const double* k,
                              const double* eta_re, const double* eta_im, const double* q, const double* pts,
                              int64_t npts, double target_db, double* loss_out, double* spl_out, double* grad_out);

/* Dense LU with partial pivoting and solve of nrhs columns, the library routine
 * behind _beat_dense_lu_solve (BeatEngineDenseSolve.jl:1104-1110) and
 * rocm_dense_lu! (BeatEngineRocmOperators.jl:132-159). a: n x n, b and x:
 * n x nrhs, complex128 column-major. precision FP64 or FP32. */
int beat_dense_lu_solve(beat_ctx* ctx, int precision, const double* a, int64_t n, const double* b, int64_t nrhs,
                        double* x);

/* LU factorization with partial pivoting alone, in LAPACK getrf form: the
 * factorization rocm_dense_lu! (BeatEngineRocmOperators.jl:160-191) obtains from
 * lu!(working), i.e. rocSOLVER getrf. a: n x n complex128 column-major; lu_out: the
 * unit-lower L and U in place (n x n complex128 column-major, L and U rounded from
 * complex64 for precision FP32); ipiv_out: n 1-based row exchanges (row k with
 * ipiv_out[k]); info_out: LAPACK's info (first exactly zero pivot, 1-based, or 0;
 * the factorization still completes). Pivot rows are chosen by |re| + |im|, the
 * first maximum on ties (izamax). */
int beat_dense_lu_factor(beat_ctx* ctx, int precision, const double* a, int64_t n, double* lu_out, int32_t* ipiv_out,
                         int32_t* info_out);

/* beat_refine_cpu_lu (BeatEngineCpuRefinement.jl:5-78) on the GPU: complex64 LU of
 * `matrix`, complex128 residuals with residual_matrix / residual_rhs (NULL: the
 * complex64 inputs widened), per-column acceptance. matrix, rhs, x_out are
 * complex64 (interleaved floats) column-major; max_steps in 1..3. */
int beat_refine_lu(beat_ctx* ctx, const float* matrix, int64_t n, const float* rhs, int64_t nrhs,
                   const double* residual_matrix, const double* residual_rhs, int max_steps, double rtol,
                   float* x_out, beat_refine_report* rep);

#ifdef __cplusplus
}
#endif

#endif /* BEAT_BEM_H */

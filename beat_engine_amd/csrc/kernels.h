// Kernel argument blocks and launchers shared by the HIP sources and the C ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "bem_common.h"

namespace beat {

template <typename T>
struct RegularArgs {
  const T* rec;            // FREC per face (internal order)
  const int32_t* fverts;   // 3 per face
  const uint8_t* floc;     // 3 per face
  const int32_t* tile_start;
  const int32_t* tile_count;
  const int32_t* tile_nv;
  const int32_t* tile_verts;  // TILE_V per tile
  int ntiles;
  const uint8_t* flocs;       // 3 per face: vertex slot in its trial tile
  const int32_t* stile_start;
  const int32_t* stile_count;
  const int32_t* stile_nv;
  const int32_t* stile_verts;  // TILE_VS (narrow: TILE_VS_N) per trial tile
  int nstiles;
  int narrow;              // the trial tile arrays are the narrow partition
  // symmetric pass: sub-tiles nested in the test tiles (TILE_FS_N faces /
  // TILE_VS_N vertices), each test tile's first sub-tile (ntiles + 1 entries) and
  // first unit (ntiles + 1 entries, the last one the unit count)
  const int32_t* qtile_start;
  const int32_t* qtile_count;
  const int32_t* qtile_nv;
  const int32_t* qtile_verts;
  const uint8_t* flocq;
  const int32_t* tile_qfirst;
  const int64_t* unit_start;
  const uint8_t* frot;     // per face: the test-side slot rotation (0..2)
  int nqtiles;
  int64_t nunits;
  int symmetric;           // nonzero: the non-image pass runs the symmetric kernel
  T* A;                    // row-major complex, lda
  int64_t lda;
  T* B;                    // N x nrhs row-major complex
  const T* q;              // F x nrhs row-major complex (internal face order)
  int nrhs;
  T k, krev;               // signed wavenumber, k / (2 pi)
  double ksteps;           // k 64/pi: the FP64 table phase in steps per unit length
  T eta_re, eta_im;
  T w[6];                  // rule weights
  T phi[6][3];             // P1 basis at the rule points
  T pw[6][3];              // phi[i][b] w[i]
  const double* trig;      // 2 TRIG_N doubles: (sin, cos)(m pi/64)
  // image passes: nimg transforms, per transform the point signs (s0, s1, s2)
  // followed by the curl signs det * (s0, s1, s2)
  int nimg;
  T img[MAX_IMAGES][6];
  // nonzero: at most one workgroup per CU (leaves room for kernels of other
  // streams, e.g. an LU running beside the assembly of the next frequency)
  int one_per_cu;
};

template <typename TO>
struct SingularArgs {
  const double* rec;
  const double* verts;     // 3 per internal vertex
  const int32_t* fverts;
  const SingularPair* pairs;
  int64_t npairs;
  const int32_t* rule_off;
  const int32_t* rule_cnt;
  const double* rule_data;
  TO* A;
  int64_t lda;
  TO* B;
  const double* q;         // F x nrhs row-major complex128
  int nrhs;
  double k, eta_re, eta_im;
  // signs by SingularPair::xform (row 0: identity): point signs, then curl signs
  double img[MAX_IMAGES + 1][6];
  // the phase: the pi/64 table (2 TRIG_N doubles) in steps of k 64/pi, unless wide
  // (|k| r not bounded below the table's limit: the pi/2 form)
  const double* trig;
  double ksteps;
  int wide;
};

// Four separate operators (operator_matrices): S, K' (n x ldf, P1 rows x DP0
// columns) and D, H (n x ldd), row-major complex, internal numbering. g carries
// the tiles, geometry, rule, wavenumber and images of the regular pass (its
// A, B, q and eta are unused); sg those of the singular pass.
template <typename T>
struct OperatorArgs {
  RegularArgs<T> g;
  T *S, *D, *K, *H;
  int64_t ldd, ldf;
};
template <typename TO>
struct SingularOperatorArgs {
  SingularArgs<TO> g;
  TO *S, *D, *K, *H;
  int64_t ldd, ldf;
};

template <typename TO>
struct IdentityArgs {
  const double* rec;
  const int32_t* fverts;
  int64_t nf;
  TO* A;
  int64_t lda;
  TO* B;
  const double* q;
  int nrhs;
  double eta_re, eta_im;
  double pp[3][3];
  double pd[3];
};

template <typename T>
// wide: |k| times the mesh diameter may reach SINCOS_TABLE_LIMIT (FP64 phase by
// the pi/2 reduction with the library fallback instead of the table form).
// image: the a.nimg image passes (every pair, trial face reflected) instead of
// the direct pass (pairs without a shared vertex).
hipError_t launch_regular(const RegularArgs<T>& a, int nq, bool pure_imag_eta, bool wide, bool image, hipStream_t s);
template <typename TO>
hipError_t launch_singular(const SingularArgs<TO>& a, hipStream_t s);
template <typename T>
hipError_t launch_operators(const OperatorArgs<T>& a, int nq, bool wide, bool image, hipStream_t s);
template <typename TO>
hipError_t launch_singular_operators(const SingularOperatorArgs<TO>& a, hipStream_t s);
template <typename TO>
hipError_t launch_identity(const IdentityArgs<TO>& a, hipStream_t s);
// rows[i] of A (n columns, lda) and B (nrhs columns) times w[i]
template <typename TO>
hipError_t launch_row_scale(TO* A, int64_t lda, int64_t n, TO* B, int nrhs, const int32_t* rows, const double* w,
                            int64_t nrows, hipStream_t s);
template <typename T>
hipError_t launch_fill(T* p, int64_t n, T v, hipStream_t s);
hipError_t launch_downcast(const double* src, float* dst, int64_t n, hipStream_t s);

// ---- dense LU (row-major complex, leading dimension lda) -------------------------
// Outer block width of lu_factor: the row exchanges chosen in an outer block are
// applied to that block's columns and every column right of it, never to the
// columns of earlier blocks (the solves read those through gather tables).
constexpr int LU_BLOCK_COLS = 256;
// trailing-update tile shapes of the LU (LuWork::gemm_tile, lu_factor_batched)
constexpr int GEMM_TILES = 4;

struct LuWork {
  int* ipiv;          // n, global row index chosen at each column
  unsigned long long* slots;  // 2 x maxg slots of 16 words (two tagged halves of the candidate magnitude, padding)
  void* slot_rows;    // 2 x maxg x 64 published rows, 4 tagged 64-bit words per element
  void* slot_diag;    // 2 x 64 published elements, same form
  int* info;          // first zero pivot (1-based), 0 if none
  // wait control of the factorization and of the solves with its factors (ctl =
  // info + 1): [0] abort word (nonzero: every cross-workgroup wait gives up at
  // once), [1] the kernels whose wait timed out (SpinSite bits), 0 if none,
  // [2] nonzero when a speculative leaf found that partial pivoting would
  // exchange rows. lu_factor clears all three, then sets the abort word to
  // abort_init.
  int* ctl;
  int abort_init = 0;
  // test hook: a speculative factorization reports an exchange (ctl[2]) whatever
  // its check finds
  bool force_exchange = false;
  void* linv;         // 2 x 4 x 64 x 64 complex (double-sized): leaf inverses of two outer blocks
  int maxg;
  hipStream_t side;   // look-ahead stream (nullptr: no look-ahead)
  hipEvent_t ev_fact, ev_far, ev_far2;  // far updates of alternate blocks record ev_far / ev_far2
  hipStream_t inb;    // in-block stream (nullptr: in-block work in order on the main stream)
  hipEvent_t ev_panel, ev_inb;
  // complex64 trailing updates as three real products instead of four (normwise
  // stable; for factors that only precondition a refinement)
  bool gauss3 = false;
  // speculative leaves: each leaf factored without row exchanges while every
  // column is checked against partial pivoting's choice (ctl[2] set where an
  // exchange would have been made; the factors are then those of an unpivoted LU
  // and the caller factors again with nopiv false)
  bool nopiv = false;
  // tile shape of the trailing updates (0..GEMM_TILES-1): 64 x 64-complex tiles with
  // K chunks of 16 (0) or 32 (1), 128 x 64 with 16 (2) or 32 (3)
  int gemm_tile = 0;
  // trailing rows below which the outer blocks run in order on one stream (the
  // look-ahead and in-block streams' events cost more than they overlap there)
  int64_t tail_rows = 0;
};

// Columns [n, nc) (n < nc <= lda) may carry right-hand sides: on return they
// hold L^-1 P b (nc = 0: none).
template <typename T>
hipError_t lu_factor(T* A, int64_t n, int64_t lda, LuWork& w, int nb, hipStream_t s, int* panels_out,
                     int64_t nc = 0);

// Forward/backward solves with the factors of lu_factor. lu_solve_prepare runs
// once per factorization (leaf inverses and link matrices, row gather tables) into work
// (lu_solve_work_bytes); lu_solve then overwrites B (n x nrhs row-major complex
// T, nrhs <= 8) with the solution; ctl is the factorization's wait control
// (LuWork::ctl). Solves sharing one work area run in stream order.
template <typename T>
hipError_t lu_solve_prepare(const T* LU, int64_t n, int64_t lda, const int* ipiv, T* work, hipStream_t s);
template <typename T>
hipError_t lu_solve(const T* LU, int64_t n, int64_t lda, T* B, int nrhs, T* work, int* ctl, hipStream_t s);
size_t lu_solve_work_bytes(int64_t n, size_t elem);
// Right-hand sides carried by lu_factor: lu_append_rhs copies B (n x nrhs) to the
// columns [n, n + nrhs) before the factorization (n + nrhs <= lda); after it and
// lu_solve_prepare, lu_solve_eliminated runs the backward solve alone on those
// columns and writes the solution to B.
template <typename T>
hipError_t lu_append_rhs(T* A, int64_t n, int64_t lda, const T* B, int nrhs, hipStream_t s);
template <typename T>
hipError_t lu_solve_eliminated(const T* LU, int64_t n, int64_t lda, T* B, int nrhs, T* work, int* ctl,
                               hipStream_t s);
// A^T x = c with the same factors and prepared work area (after lu_solve_prepare):
// v (n complex, internal row order) holds c and is overwritten; x receives the
// solution.
template <typename T>
hipError_t lu_solve_transposed(const T* LU, int64_t n, int64_t lda, T* v, T* x, const T* work, hipStream_t s);

// ---- batched factorization and solves (many small systems at once) ----------------
// One system of a batch, as a device array entry. A: n x lda row-major, factored
// in place; its columns [n, nc) carry the right-hand sides (nc <= lda); linv: one
// 64 x 64 complex leaf inverse per 64 columns; status: the LuWork info block
// ([0] first zero pivot, [1] abort word, [2] give-up record, [3] exchange needed).
template <typename T>
struct BatchSystem {
  T* A;
  int64_t n, lda, nc;
  T* linv;
  int* status;
};
// A vector block of a system: n x nrhs complex T with row stride ld; skip: when
// non-null and nonzero, the batched kernels leave this system alone (its
// refinement has finished).
template <typename T>
struct BatchVec {
  T* v;
  int64_t n, ld;
  const int* skip;
};
// Speculative LU (no exchanges, every column checked) of nsys systems at once,
// leaf by leaf: per 64 columns a diagonal-block kernel, a rows-below kernel and
// the two trailing products, each one launch for every system. nmax / ncmax: the
// largest n and nc; tile: the trailing products' tile shape (0..GEMM_TILES-1).
// Zero pivots and needed exchanges land in each status.
template <typename T>
hipError_t lu_factor_batched(const BatchSystem<T>* sys, int nsys, int64_t nmax, int64_t ncmax, bool gauss3,
                             int tile, hipStream_t s);
// With the factors of lu_factor_batched: forward (L y = v) and/or backward (U x = y)
// substitution of each system's vector block in place, one workgroup per system.
template <typename T>
hipError_t tri_solve_batched(const BatchSystem<T>* sys, const BatchVec<T>* vec, int nsys, int nrhs, bool fwd, bool bwd,
                             hipStream_t s);
// r = b - A x for every system (A the complex128 matrices of sys, x / b / r the
// vector blocks), nrhs <= 8.
hipError_t residual64_batched(const BatchSystem<double>* sys, const BatchVec<double>* x, const BatchVec<double>* b,
                              const BatchVec<double>* r, int nsys, int64_t nmax, int nrhs, hipStream_t s);

// r = b - A x (complex128), A row-major n x n, x and b n x nrhs row-major (nrhs
// <= 8); part: workspace of residual64_part_bytes(n) for the partial products;
// split: spread each row block's columns over several workgroups to fill the
// device; false when other streams keep it busy (small workgroups that fit beside
// the other streams' kernels, one k range).
hipError_t residual64(const double* A, int64_t n, int64_t lda, const double* x, const double* b, double* r, int nrhs,
                      double* part, bool split, hipStream_t s);
size_t residual64_part_bytes(int64_t n);
// per-column 2-norms of an n x nrhs complex128 row-major matrix into out[nrhs]
// (out holds nrhs norms followed by 64*nrhs partial sums)
hipError_t column_norms64(const double* X, int64_t n, int nrhs, double* out, hipStream_t s);

// ---- adjoint shape gradient (FP64, complex interleaved, internal numbering) ---------
struct IdentityBlocks {
  double pp[3][3];  // unit-area P1 x P1 mass block
  double pd[3];     // unit-area P1 x DP0 block
};
// per-face interpolants of mu (may be all zero), p and q; fi: 38 doubles per face
hipError_t adjoint_face_interp(const double* rec, const int32_t* fverts, int64_t nf, int nq, const double* rule,
                               const double* mu, const double* p, const double* q, double* fi, hipStream_t s);
// field part: seeds (34 doubles per face, accumulated) and the adjoint right-hand
// side F^T cw (n complex, accumulated); cw: conj(dL/du*) per observation point
hipError_t adjoint_field(const double* rec, const int32_t* fverts, int64_t nf, int nq, const double* rule,
                         const double* fi, const double* pts, const double* cw, int64_t npts, double k, double* seeds,
                         double* rhs, hipStream_t s);
// system part: regular pairs (two role passes) and touching pairs; gV: n x 3 real
hipError_t adjoint_pairs(const double* rec, const double* verts, const int32_t* fverts, int64_t nf, int nq,
                         const double* rule, const SingularPair* pairs, int64_t npairs, const int32_t* rule_off,
                         const int32_t* rule_cnt, const double* rule_data, const double* fi, const double* mu,
                         const double* p, double k, double er, double ei, double* seeds, double* gV, hipStream_t s);
// identity blocks and the chain from face seeds to vertex coordinates (into gV)
hipError_t adjoint_chain(const double* verts, const int32_t* fverts, int64_t nf, int nq, const double* rule,
                         const double* seeds, const double* mu, const double* p, const double* q, double er, double ei,
                         const IdentityBlocks& idb, double* gV, hipStream_t s);

// ---- field -----------------------------------------------------------------------
template <typename T>
hipError_t field_eval(const T* src_pts, const T* src_nrm, const T* pw, const T* qw, int64_t nsrc, int nrhs,
                      const T* pts, int64_t npts, T* out, T k, hipStream_t s);
// rule: 3 doubles per regular rule point (xi, eta, weight); p is N x nrhs and q
// F x nrhs (row-major, internal order); outputs are per source: F * nq for the
// mesh, then F * nq for each of the nimg image transforms (points and normals
// reflected by the signs sg[i][0..2], same weighted values).
template <typename T>
hipError_t field_sources(const double* rec, const int32_t* fverts, int64_t nf, int nq, const double* rule, const T* p,
                         const T* q, int nrhs, int nimg, const ImageXf* img, T* src_pts, T* src_nrm, T* pw, T* qw,
                         hipStream_t s);

// Value and x, y, z derivatives of the field (out: npts x nrhs x 4 complex,
// internal row-major) from the same sources as field_eval.
template <typename T>
hipError_t field_gradient_eval(const T* src_pts, const T* src_nrm, const T* pw, const T* qw, int64_t nsrc, int nrhs,
                               const T* pts, int64_t npts, T* out, T k, hipStream_t s);
// (point, face image) pairs with the image's centroid within 3 diameters of the
// point: counted into *count (zeroed by the caller) and, with fill, stored in list
// while the count stays below cap (entry = point, face * (1 + nimg) + transform).
template <typename T>
hipError_t gradient_near_pairs(const double* verts, const int32_t* fverts, int64_t nf, int nimg, const ImageXf* img,
                               const T* pts, int64_t npts, unsigned long long* count, int2* list,
                               unsigned long long cap, bool fill, hipStream_t s);
// Adds the adaptive FP64 integral minus the 6-point rule (rule4: xi, eta, w per
// point) of every listed pair to out; *failed (zeroed by the caller) counts pairs
// that reach the depth limit without meeting the refinement criterion.
template <typename T>
hipError_t gradient_correction(const double* verts, const int32_t* fverts, const double* rec, int nimg,
                               const ImageXf* img, const int2* list, int64_t npairs, const T* pts, const T* p,
                               const T* q, int nrhs, double k, const double* rule4, T* out, int* failed,
                               hipStream_t s);

}  // namespace beat

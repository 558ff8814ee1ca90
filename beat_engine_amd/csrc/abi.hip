// C ABI (include/beat_bem.h): contexts, meshes and the solve pipeline
//   assembly (regular + singular + identity) -> LU -> triangular solves
//   [-> complex128 iterative refinement]
// for one frequency and every right-hand side, with a single factorization.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include <unistd.h>

#include "../../include/beat_bem.h"
#include "bem_common.h"
#include "bem_math.h"
#include "kernels.h"

using namespace beat;

namespace {

constexpr int LU_NB = 64;
constexpr int SOLVE_CHUNK = 8;

// Device scratch of beat_field_eval_device_noalloc (callers that cannot
// allocate): the staged inputs, the field sources and the accumulated output.
// One copy per GPU, alive until the process ends; every call on a GPU shares it.
// Host code never passes the name itself to a kernel or copy: it uses the
// address hipGetSymbolAddress gives on each device (beat_ctx::scratch).
constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
};

hipError_t ensure(DevBuf& b, size_t bytes) {
  if (b.bytes >= bytes && b.p) return hipSuccess;
  if (b.p) {
    hipError_t e = hipFree(b.p);
    if (e != hipSuccess) return e;
    b.p = nullptr;
    b.bytes = 0;
  }
  hipError_t e = hipMalloc(&b.p, std::max<size_t>(bytes, 256));
  if (e == hipSuccess) b.bytes = std::max<size_t>(bytes, 256);
  return e;
}

template <typename T>
T* as(DevBuf& b) {
  return reinterpret_cast<T*>(b.p);
}

}  // namespace

// The autotuner's choices for one precision and size class, -1 where not tuned: the
// factorization's stream layout (0 look-ahead streams, 1 one stream), the batch
// assembly occupancy (0 every slot, 1 one workgroup per CU), the LU's
// trailing-update tile shape (LuWork::gemm_tile) and the regular pass's tiles (0
// unordered pairs over 16-face sub-tiles, 1 ordered pairs over 32-face trial
// tiles, 2 ordered pairs over 16-face trial tiles).
struct Tuning {
  signed char lu_layout = -1, asm_occ = -1, gemm_tile = -1, asm_tile = -1;
};

struct beat_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipStream_t side = nullptr;  // look-ahead stream of the LU
  int side_reserve = 0;        // eighths of the CUs the side stream leaves free
  hipEvent_t ev_fact = nullptr, ev_far = nullptr, ev_far2 = nullptr;
  hipStream_t inb = nullptr;  // in-block stream of the LU
  hipEvent_t ev_panel = nullptr, ev_inb = nullptr;
  std::string err;
  std::mutex mu;
  DevBuf a64, a32, b64, b32, q64, q32, x64, r64, n64, rn64, xn64, d32, work, chunk, norms, ipiv, lusmall, qin, pout,
      fld_src, fld_out, fld_pts, grad, adj, af, rpart;
  hipEvent_t ev[7] = {};
  // frequency sweeps: the assembly stream, two systems in flight, their events
  // (assembly start, regular start, regular end, assembly end) and release events
  hipStream_t asm_stream = nullptr;
  DevBuf sw_q64[2], sw_q32[2], sw_a64[2], sw_b64[2], sw_a32[2], sw_b32[2], sw_af[2];
  hipEvent_t sw_ev[2][4] = {};
  hipEvent_t sw_free[2] = {};
  // the regular kernel at one workgroup per CU while it shares the device with
  // an LU (BEAT_SWEEP_ASM_ONE_PER_CU=1). Off by default: the LU's kernels issue at
  // a higher wave priority, so the assembly may take every slot the LU leaves
  // (measured at L5, 8 RHS: 12.9 vs 10.7 freq-solves/s; one per CU the assembly
  // takes 90 ms per frequency instead of 70, the LU 64.8 ms instead of 69.7)
  bool sweep_asm_one_per_cu = false;
  // the regular pass over unordered tile pairs (both orders of a face pair from
  // one evaluation of the kernel); false: the ordered-pair kernel
  bool regular_symmetric = true;
  // FP32 solves assemble in complex128 and round the system once (false: the
  // complex64 assembly kernels; BEAT_FP32_ASSEMBLY=native at context creation)
  bool fp32_assembly_fp64 = true;
  // contexts a batch may use at once (BEAT_BATCH_WORKERS, default 4)
  int batch_workers = 4;
  // batches of small systems through the grouped, graph-replayed path
  // (BEAT_BATCH_GROUPED=0: the per-job pipeline)
  bool batch_grouped = true;
  // trailing rows below which the LU's outer blocks run in order on one stream
  // (BEAT_LU_TAIL_ROWS at context creation; measured at n = 10242, complex64:
  // 0 rows 33.0 ms, 768 32.8, 1536 32.6, 2560 32.5, 4096 32.6)
  int64_t lu_tail_rows = 2560;
  // set while a batch or sweep runs another system's assembly beside the solves
  bool shared_device = false;
  // cancellation marker polled between the frequencies of a sweep (empty: none)
  std::string cancel_path;
  // autotuned launch strategy and tile shapes per precision and size class
  // (floor(log2 n))
  Tuning tune[3][40];
  // further contexts on the same device that run batch jobs of small systems
  // concurrently (created on first use, owned by this context)
  std::vector<beat_ctx*> aux;
  double* trig = nullptr;  // (sin, cos)(m pi/64) table of the FP64 phase
  // wait control of the last factorization (LuWork::ctl), passed to its solves
  int* lu_ctl = nullptr;
  // test hook (beat_ctx_debug_abort_waits): factorizations start with their abort
  // word set, so every cross-workgroup wait gives up at once
  int debug_abort_waits = 0;
  // test hook (beat_ctx_debug_force_exchange): speculative factorizations report
  // an exchange, so every solve takes its second, exchange factorization
  bool debug_force_exchange = false;
  // this device's copy of the no-allocation scratch array (set at creation)
  unsigned char* scratch = nullptr;
  // grouped batches (batch_grouped_impl): the group's systems, vectors, refinement
  // states and descriptors, and the captured factor / solve / refine graphs by
  // group shape (invalidated whenever these buffers move)
  DevBuf g_a64, g_a32, g_b64, g_b32, g_q64, g_q32, g_linv, g_vec, g_state, g_desc;
  std::map<std::string, hipGraphExec_t> g_graphs;
};

struct beat_mesh {
  beat_ctx* ctx = nullptr;
  HostMesh h;
  double* rec64 = nullptr;
  float* rec32 = nullptr;
  double* verts = nullptr;
  int32_t* fverts = nullptr;
  uint8_t* floc = nullptr;
  int32_t *tile_start = nullptr, *tile_count = nullptr, *tile_nv = nullptr, *tile_verts = nullptr;
  int32_t *stile_start = nullptr, *stile_count = nullptr, *stile_nv = nullptr, *stile_verts = nullptr;
  uint8_t* flocs = nullptr;
  int32_t *ntile_start = nullptr, *ntile_count = nullptr, *ntile_nv = nullptr, *ntile_verts = nullptr;
  uint8_t* flocn = nullptr;
  int32_t *qtile_start = nullptr, *qtile_count = nullptr, *qtile_nv = nullptr, *qtile_verts = nullptr;
  uint8_t* flocq = nullptr;
  int32_t* tile_qfirst = nullptr;
  int64_t* sym_unit_start = nullptr;
  uint8_t* frot = nullptr;
  SingularPair* spairs = nullptr;
  int32_t *rule_off = nullptr, *rule_cnt = nullptr;
  double* rule_data = nullptr;
  double* rule3 = nullptr;  // regular rule: xi, eta, w per point
  double* rule4 = nullptr;  // the 6-point rule (xi, eta, w) of the near-panel gradient correction
  int32_t* face_perm = nullptr;
  int32_t* vert_inv = nullptr;
  int32_t* wrows = nullptr;  // rows on a mirror plane and their orbit weights
  double* wvals = nullptr;
  int64_t regular_pairs = 0;
  double diameter = 0;  // bounding-box diagonal: bounds every |y - x| of the assembly (images included)
};

namespace {

inline int size_class(int64_t n) {
  int b = 0;
  while (b < 39 && (int64_t(2) << b) <= n) ++b;
  return b;
}

// the tuned choices of a precision and system size
inline Tuning& tuned(beat_ctx* ctx, int precision, int64_t n) {
  return ctx->tune[precision < 0 || precision > 2 ? 0 : precision][size_class(n)];
}

int fail(beat_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

#define HIPCHK(expr)                                                                              \
  do {                                                                                            \
    hipError_t e_ = (expr);                                                                       \
    if (e_ != hipSuccess)                                                                         \
      return fail(ctx, BEAT_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));          \
  } while (0)

template <typename T>
__global__ void gather_q_kernel(const double* qc, int64_t nf, int nrhs, const int32_t* perm, T* qi) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nf * nrhs) return;
  const int64_t f = i / nrhs;
  const int d = (int)(i - f * nrhs);
  const int64_t src = (int64_t)perm[f] + (int64_t)d * nf;
  qi[2 * i] = (T)qc[2 * src];
  qi[2 * i + 1] = (T)qc[2 * src + 1];
}

// caller column-major nv x nrhs complex128 -> internal row-major (row inv[v] of caller row v)
__global__ void p_to_internal_kernel(const double* pc, int64_t nv, int nrhs, const int32_t* inv, double* pi) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nv * nrhs) return;
  const int64_t v = i % nv;
  const int d = (int)(i / nv);
  const int64_t dst = (int64_t)inv[v] * nrhs + d;
  pi[2 * dst] = pc[2 * i];
  pi[2 * dst + 1] = pc[2 * i + 1];
}

template <typename T>
__global__ void scatter_p_kernel(const T* X, int64_t nv, int nrhs, const int32_t* inv, double* out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nv * nrhs) return;
  const int64_t v = i % nv;
  const int d = (int)(i / nv);
  const int64_t s = (int64_t)inv[v] * nrhs + d;
  out[2 * i] = (double)X[2 * s];
  out[2 * i + 1] = (double)X[2 * s + 1];
}

// columns [c0, c0+nc) of an n x ld row-major complex matrix <-> n x nc buffer
template <typename T>
__global__ void extract_cols_kernel(const T* src, int64_t n, int ld, int c0, int nc, T* dst) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * nc) return;
  const int64_t r = i / nc;
  const int c = (int)(i - r * nc);
  dst[2 * i] = src[2 * (r * ld + c0 + c)];
  dst[2 * i + 1] = src[2 * (r * ld + c0 + c) + 1];
}
template <typename T>
__global__ void insert_cols_kernel(const T* src, int64_t n, int ld, int c0, int nc, T* dst) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * nc) return;
  const int64_t r = i / nc;
  const int c = (int)(i - r * nc);
  dst[2 * (r * ld + c0 + c)] = src[2 * i];
  dst[2 * (r * ld + c0 + c) + 1] = src[2 * i + 1];
}

__global__ void upcast_kernel(const float* s, double* d, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    d[i] = (double)s[i];
}
// delta = r (complex128 -> complex64), zero for columns whose mask is 0
__global__ void masked_downcast_kernel(const double* r, float* d, int64_t n, int nrhs, const int* mask) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * nrhs) return;
  const int c = (int)(i % nrhs);
  const bool on = mask[c] != 0;
  d[2 * i] = on ? (float)r[2 * i] : 0.f;
  d[2 * i + 1] = on ? (float)r[2 * i + 1] : 0.f;
}
__global__ void add_delta_kernel(const double* x, const float* d, double* out, int64_t n2) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n2; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = x[i] + (double)d[i];
}
// accepted columns: x[:, j] = next[:, j], r[:, j] = rnext[:, j]
__global__ void accept_kernel(double* x, const double* nx, double* r, const double* nr, int64_t n, int nrhs,
                              const int* mask) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * nrhs) return;
  const int c = (int)(i % nrhs);
  if (!mask[c]) return;
  x[2 * i] = nx[2 * i];
  x[2 * i + 1] = nx[2 * i + 1];
  r[2 * i] = nr[2 * i];
  r[2 * i + 1] = nr[2 * i + 1];
}
// round a complex128 solution through complex64 and back
__global__ void round32_kernel(const double* x, double* y, int64_t n2) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n2; i += (int64_t)gridDim.x * blockDim.x)
    y[i] = (double)(float)x[i];
}
// column-major caller matrix (complex128) -> row-major internal (T), with lda
template <typename T>
__global__ void transpose_in_kernel(const double* a, int64_t n, int64_t lda, T* out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * n) return;
  const int64_t r = i % n, c = i / n;
  out[2 * (r * lda + c)] = (T)a[2 * i];
  out[2 * (r * lda + c) + 1] = (T)a[2 * i + 1];
}
template <typename T>
__global__ void transpose_in_f32_kernel(const float* a, int64_t n, int64_t lda, T* out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * n) return;
  const int64_t r = i % n, c = i / n;
  out[2 * (r * lda + c)] = (T)a[2 * i];
  out[2 * (r * lda + c) + 1] = (T)a[2 * i + 1];
}
// caller column-major n x nrhs (complex128 or complex64) -> row-major
template <typename S, typename T>
__global__ void cols_in_kernel(const S* b, int64_t n, int nrhs, T* out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * nrhs) return;
  const int64_t r = i % n, c = i / n;
  out[2 * (r * nrhs + c)] = (T)b[2 * i];
  out[2 * (r * nrhs + c) + 1] = (T)b[2 * i + 1];
}
template <typename S, typename T>
__global__ void cols_out_kernel(const S* x, int64_t n, int nrhs, T* out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * nrhs) return;
  const int64_t r = i % n, c = i / n;
  out[2 * i] = (T)x[2 * (r * nrhs + c)];
  out[2 * i + 1] = (T)x[2 * (r * nrhs + c) + 1];
}

inline unsigned blocks_for(int64_t n, int bs = 256) { return (unsigned)std::max<int64_t>(1, (n + bs - 1) / bs); }
inline unsigned grid_cap(int64_t n) { return (unsigned)std::min<int64_t>(std::max<int64_t>(1, (n + 255) / 256), 16384); }

int64_t lda_for(int64_t n) { return (n + 63) / 64 * 64; }

template <typename T>
int ensure_lu_work(beat_ctx* ctx, int64_t n, LuWork& w) {
  const int maxg = (int)((n + 63) / 64) + 1;
  // published panel rows: up to four tagged 64-bit words per complex element
  const size_t slot_rows = (size_t)2 * maxg * 64 * 4 * sizeof(unsigned long long);
  const size_t slot_diag = (size_t)2 * 64 * 4 * sizeof(unsigned long long);
  const size_t linv_bytes = (size_t)8 * 64 * 64 * 2 * sizeof(double);
  const size_t bytes =
      sizeof(int) * (n + 8) + sizeof(long long) * 32 * maxg + slot_rows + slot_diag + linv_bytes + 4096;
  HIPCHK(ensure(ctx->ipiv, bytes));
  char* p = reinterpret_cast<char*>(ctx->ipiv.p);
  auto take = [&](size_t b) {
    char* r = p;
    p += (b + 255) / 256 * 256;
    return r;
  };
  w.ipiv = reinterpret_cast<int*>(take(sizeof(int) * (n + 8)));
  w.info = reinterpret_cast<int*>(take(sizeof(int) * 4));
  w.ctl = w.info + 1;
  w.abort_init = ctx->debug_abort_waits;
  w.force_exchange = ctx->debug_force_exchange;
  ctx->lu_ctl = w.ctl;
  w.slots = reinterpret_cast<unsigned long long*>(take(sizeof(unsigned long long) * 32 * maxg));  // 2 x maxg slots of 16 words
  w.slot_rows = take(slot_rows);
  w.slot_diag = take(slot_diag);
  w.linv = take(linv_bytes);
  w.side = ctx->side;
  w.ev_fact = ctx->ev_fact;
  w.ev_far = ctx->ev_far;
  w.ev_far2 = ctx->ev_far2;
  w.inb = ctx->inb;
  w.ev_panel = ctx->ev_panel;
  w.ev_inb = ctx->ev_inb;
  w.maxg = maxg;
  w.tail_rows = ctx->lu_tail_rows;
  (void)sizeof(T);
  return BEAT_OK;
}

// The factorization's zero-pivot report and wait control, read after the stream
// has drained: BEAT_ERR_HIP when a cross-workgroup wait of the factorization or of
// its solves gave up (the results are then abandoned), else *info = the first
// zero pivot (1-based, 0 if none) and *exchange_needed = a speculative
// factorization found a column where partial pivoting exchanges rows.
int lu_outcome(beat_ctx* ctx, const LuWork& w, int* info, bool* exchange_needed = nullptr) {
  int st[4] = {0, 0, 0, 0};
  HIPCHK(hipMemcpy(st, w.info, sizeof(st), hipMemcpyDeviceToHost));
  *info = st[0];
  if (exchange_needed) *exchange_needed = st[3] != 0;
  if (st[1] != 0 || st[2] != 0) {
    std::string where;
    if (st[2] == 0) where = "the abort word was set";
    if (st[2] & 1) where += "panel slot poll ";
    if (st[2] & 2) where += "panel row fetch ";
    if (st[2] & 4) where += "forward solve ";
    if (st[2] & 8) where += "backward solve ";
    return fail(ctx, BEAT_ERR_HIP, "a cross-workgroup wait of the factorization or its solves gave up (" + where +
                                       "); the results of this call are abandoned");
  }
  return BEAT_OK;
}

// Solve with LU factors for every column of the n x nrhs row-major B (in place).
// prepare: first solve after a factorization (builds the per-factorization solve
// data in ctx->work; later solves with the same factors pass false).
template <typename T>
int solve_all(beat_ctx* ctx, const T* LU, int64_t n, int64_t lda, const int* ipiv, T* B, int nrhs, bool prepare) {
  HIPCHK(ensure(ctx->work, lu_solve_work_bytes(n, sizeof(T))));
  if (prepare) HIPCHK(lu_solve_prepare<T>(LU, n, lda, ipiv, as<T>(ctx->work), ctx->stream));
  if (nrhs <= SOLVE_CHUNK) {
    HIPCHK(lu_solve<T>(LU, n, lda, B, nrhs, as<T>(ctx->work), ctx->lu_ctl, ctx->stream));
    return BEAT_OK;
  }
  HIPCHK(ensure(ctx->chunk, sizeof(T) * 2 * n * SOLVE_CHUNK));
  for (int c0 = 0; c0 < nrhs; c0 += SOLVE_CHUNK) {
    const int nc = std::min(SOLVE_CHUNK, nrhs - c0);
    hipLaunchKernelGGL((extract_cols_kernel<T>), dim3(blocks_for(n * nc)), dim3(256), 0, ctx->stream, (const T*)B, n,
                       nrhs, c0, nc, as<T>(ctx->chunk));
    HIPCHK(hipGetLastError());
    HIPCHK(lu_solve<T>(LU, n, lda, as<T>(ctx->chunk), nc, as<T>(ctx->work), ctx->lu_ctl, ctx->stream));
    hipLaunchKernelGGL((insert_cols_kernel<T>), dim3(blocks_for(n * nc)), dim3(256), 0, ctx->stream,
                       (const T*)as<T>(ctx->chunk), n, nrhs, c0, nc, B);
    HIPCHK(hipGetLastError());
  }
  return BEAT_OK;
}

// First solve after a factorization that carried B's columns (lu_factor with
// nc = n + nrhs): the factorization did the forward solve, only the backward one
// runs. Prepares the per-factorization solve data as solve_all(prepare) does.
template <typename T>
int solve_eliminated(beat_ctx* ctx, const T* LU, int64_t n, int64_t lda, const int* ipiv, T* B, int nrhs) {
  HIPCHK(ensure(ctx->work, lu_solve_work_bytes(n, sizeof(T))));
  HIPCHK(lu_solve_prepare<T>(LU, n, lda, ipiv, as<T>(ctx->work), ctx->stream));
  HIPCHK(lu_solve_eliminated<T>(LU, n, lda, B, nrhs, as<T>(ctx->work), ctx->lu_ctl, ctx->stream));
  return BEAT_OK;
}

// r = b - A x for all columns; A64 row-major.
int residual_all(beat_ctx* ctx, const double* A, int64_t n, int64_t lda, const double* x, const double* b, double* r,
                 int nrhs) {
  HIPCHK(ensure(ctx->rpart, residual64_part_bytes(n)));
  double* part = as<double>(ctx->rpart);
  if (nrhs <= SOLVE_CHUNK) {
    HIPCHK(residual64(A, n, lda, x, b, r, nrhs, part, !ctx->shared_device, ctx->stream));
    return BEAT_OK;
  }
  HIPCHK(ensure(ctx->chunk, sizeof(double) * 2 * n * SOLVE_CHUNK * 3));
  double* cx = as<double>(ctx->chunk);
  double* cb = cx + 2 * n * SOLVE_CHUNK;
  double* cr = cb + 2 * n * SOLVE_CHUNK;
  for (int c0 = 0; c0 < nrhs; c0 += SOLVE_CHUNK) {
    const int nc = std::min(SOLVE_CHUNK, nrhs - c0);
    hipLaunchKernelGGL((extract_cols_kernel<double>), dim3(blocks_for(n * nc)), dim3(256), 0, ctx->stream, x, n, nrhs,
                       c0, nc, cx);
    hipLaunchKernelGGL((extract_cols_kernel<double>), dim3(blocks_for(n * nc)), dim3(256), 0, ctx->stream, b, n, nrhs,
                       c0, nc, cb);
    HIPCHK(residual64(A, n, lda, cx, cb, cr, nc, part, !ctx->shared_device, ctx->stream));
    hipLaunchKernelGGL((insert_cols_kernel<double>), dim3(blocks_for(n * nc)), dim3(256), 0, ctx->stream,
                       (const double*)cr, n, nrhs, c0, nc, r);
    HIPCHK(hipGetLastError());
  }
  return BEAT_OK;
}

int col_norms(beat_ctx* ctx, const double* X, int64_t n, int nrhs, std::vector<double>& out) {
  HIPCHK(ensure(ctx->norms, sizeof(double) * (nrhs + 64 * nrhs)));
  HIPCHK(column_norms64(X, n, nrhs, as<double>(ctx->norms), ctx->stream));
  out.resize(nrhs);
  HIPCHK(hipMemcpyAsync(out.data(), ctx->norms.p, sizeof(double) * nrhs, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return BEAT_OK;
}

// Norm slots of refine_loop's first step in ctx->norms: the column norms of b,
// of the initial residual and of the corrected one (each nrhs norms followed by
// 64*nrhs partial sums), then the step's record: bn, err, ne (nrhs doubles each)
// and the two column masks (nrhs ints each).
size_t refine_norms_bytes(int nrhs) { return sizeof(double) * (3 * (size_t)(nrhs + 64 * nrhs) + 4 * (size_t)nrhs); }

// The first correction step's decisions, made on the device with the host
// loop's rules so that the step runs without a wait for the host.
// phase 0: err = relative errors of the initial residual, m1 = columns above
// rtol (the columns the step corrects); phase 1: ne = errors after the step,
// m2 = columns whose error fell to a finite smaller value, all zero when no
// column was above rtol (the host loop then runs no step).
__global__ void refine_first_kernel(const double* bn, const double* rn, int nrhs, double rtol, double* rec, int phase) {
  double* cbn = rec;
  double* err = rec + nrhs;
  double* ne = err + nrhs;
  int* m1 = reinterpret_cast<int*>(ne + nrhs);
  int* m2 = m1 + nrhs;
  bool run = false;
  if (phase == 1)
    for (int j = 0; j < nrhs; ++j) run |= m1[j] != 0;
  for (int j = threadIdx.x; j < nrhs; j += blockDim.x) {
    const double b = phase == 0 ? bn[j] : cbn[j];
    const double e = b == 0 ? (rn[j] == 0 ? 0.0 : INFINITY) : rn[j] / b;
    if (phase == 0) {
      cbn[j] = b;
      err[j] = e;
      m1[j] = e <= rtol ? 0 : 1;
    } else {
      ne[j] = e;
      m2[j] = (run && isfinite(e) && e < err[j]) ? 1 : 0;
    }
  }
}

struct Refine {
  int steps = 0, status = BEAT_REFINE_NONE, working = BEAT_REFINE_NONE, hist_len = 0;
  std::vector<double> working_res, returned_res, history;
};

// beat_refine_cpu_lu's loop (BeatEngineCpuRefinement.jl:5-78) with device data:
// LU32 factors of the working matrix, A64/b64 the residual operator and rhs.
// x64 (n x nrhs) holds the initial complex64 solution widened; on return it holds
// the accepted working solution. returned_round: score the complex64-rounded
// solution (the reference's public type) or the complex128 working solution.
int refine_loop(beat_ctx* ctx, const float* LU32, int64_t lda32, const int* ipiv, const double* A64, int64_t lda64,
                const double* b64, double* x64, int64_t n, int nrhs, int max_steps, double rtol, bool returned_round,
                Refine& out) {
  HIPCHK(ensure(ctx->r64, sizeof(double) * 2 * n * nrhs));
  HIPCHK(ensure(ctx->n64, sizeof(double) * 2 * n * nrhs));
  HIPCHK(ensure(ctx->rn64, sizeof(double) * 2 * n * nrhs));
  HIPCHK(ensure(ctx->d32, sizeof(float) * 2 * n * nrhs + sizeof(int) * 64 + 4096));
  double* r = as<double>(ctx->r64);
  double* nx = as<double>(ctx->n64);
  double* nr = as<double>(ctx->rn64);
  float* d = as<float>(ctx->d32);
  std::vector<int> mask(nrhs);
  HIPCHK(ensure(ctx->lusmall, sizeof(int) * nrhs + 256));
  int* dmask = as<int>(ctx->lusmall);
  HIPCHK(ensure(ctx->norms, refine_norms_bytes(nrhs)));
  const int64_t slot = nrhs + 64 * nrhs;
  double* nb = as<double>(ctx->norms);
  double* nr0 = nb + slot;
  double* nr1 = nr0 + slot;
  double* rec = nr1 + slot;
  std::vector<double> bn;
  auto errors_of = [&](const std::vector<double>& nrm) {
    std::vector<double> e(nrhs);
    for (int j = 0; j < nrhs; ++j)
      e[j] = bn[j] == 0 ? (nrm[j] == 0 ? 0.0 : INFINITY) : nrm[j] / bn[j];
    return e;
  };
  auto all_le = [&](const std::vector<double>& e) {
    for (double v : e)
      if (!(v <= rtol)) return false;
    return true;
  };
  // Initial residual and the first correction step are queued without a wait:
  // the step's column mask and acceptance come from refine_first_kernel, and
  // the host reads the record once, after the step.
  if (residual_all(ctx, A64, n, lda64, x64, b64, r, nrhs) != BEAT_OK) return BEAT_ERR_HIP;
  HIPCHK(column_norms64(b64, n, nrhs, nb, ctx->stream));
  HIPCHK(column_norms64(r, n, nrhs, nr0, ctx->stream));
  hipLaunchKernelGGL(refine_first_kernel, dim3(1), dim3(64), 0, ctx->stream, (const double*)nb, (const double*)nr0,
                     nrhs, rtol, rec, 0);
  HIPCHK(hipGetLastError());
  const int* m1 = reinterpret_cast<const int*>(rec + 3 * nrhs);
  const int* m2 = m1 + nrhs;
  if (max_steps >= 1) {
    hipLaunchKernelGGL(masked_downcast_kernel, dim3(blocks_for(n * nrhs)), dim3(256), 0, ctx->stream, (const double*)r,
                       d, n, nrhs, m1);
    HIPCHK(hipGetLastError());
    if (solve_all<float>(ctx, LU32, n, lda32, ipiv, d, nrhs, false) != BEAT_OK) return BEAT_ERR_HIP;
    hipLaunchKernelGGL(add_delta_kernel, dim3(grid_cap(2 * n * nrhs)), dim3(256), 0, ctx->stream, (const double*)x64,
                       (const float*)d, nx, 2 * n * nrhs);
    HIPCHK(hipGetLastError());
    if (residual_all(ctx, A64, n, lda64, nx, b64, nr, nrhs) != BEAT_OK) return BEAT_ERR_HIP;
    HIPCHK(column_norms64(nr, n, nrhs, nr1, ctx->stream));
    hipLaunchKernelGGL(refine_first_kernel, dim3(1), dim3(64), 0, ctx->stream, (const double*)nb, (const double*)nr1,
                       nrhs, rtol, rec, 1);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(accept_kernel, dim3(blocks_for(n * nrhs)), dim3(256), 0, ctx->stream, x64, (const double*)nx, r,
                       (const double*)nr, n, nrhs, m2);
    HIPCHK(hipGetLastError());
  }
  std::vector<double> hrec(4 * nrhs);
  HIPCHK(hipMemcpyAsync(hrec.data(), rec, sizeof(double) * 4 * nrhs, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  bn.assign(hrec.begin(), hrec.begin() + nrhs);
  std::vector<double> errors(hrec.begin() + nrhs, hrec.begin() + 2 * nrhs);
  out.history = errors;
  out.hist_len = 1;
  int steps = 0, status = BEAT_REFINE_STEP_LIMIT;
  bool done = false;
  if (max_steps >= 1) {
    if (all_le(errors)) {
      status = BEAT_REFINE_CONVERGED;
      done = true;
    } else {
      const double* ne = hrec.data() + 2 * nrhs;
      const int* acc = reinterpret_cast<const int*>(hrec.data() + 3 * nrhs) + nrhs;
      bool improved = false;
      for (int j = 0; j < nrhs; ++j)
        if (acc[j]) {
          errors[j] = ne[j];
          improved = true;
        }
      steps = 1;
      out.history.insert(out.history.end(), errors.begin(), errors.end());
      out.hist_len += 1;
      if (!improved) {
        status = BEAT_REFINE_STAGNATED;
        done = true;
      }
    }
  }
  for (int step = 2; step <= max_steps && !done; ++step) {
    if (all_le(errors)) {
      status = BEAT_REFINE_CONVERGED;
      break;
    }
    for (int j = 0; j < nrhs; ++j) mask[j] = errors[j] <= rtol ? 0 : 1;
    HIPCHK(hipMemcpyAsync(dmask, mask.data(), sizeof(int) * nrhs, hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(masked_downcast_kernel, dim3(blocks_for(n * nrhs)), dim3(256), 0, ctx->stream, (const double*)r,
                       d, n, nrhs, (const int*)dmask);
    HIPCHK(hipGetLastError());
    if (solve_all<float>(ctx, LU32, n, lda32, ipiv, d, nrhs, false) != BEAT_OK) return BEAT_ERR_HIP;
    hipLaunchKernelGGL(add_delta_kernel, dim3(grid_cap(2 * n * nrhs)), dim3(256), 0, ctx->stream, (const double*)x64,
                       (const float*)d, nx, 2 * n * nrhs);
    HIPCHK(hipGetLastError());
    if (residual_all(ctx, A64, n, lda64, nx, b64, nr, nrhs) != BEAT_OK) return BEAT_ERR_HIP;
    std::vector<double> nrn;
    if (col_norms(ctx, nr, n, nrhs, nrn) != BEAT_OK) return BEAT_ERR_HIP;
    std::vector<double> ne = errors_of(nrn);
    bool improved = false;
    for (int j = 0; j < nrhs; ++j) {
      mask[j] = (std::isfinite(ne[j]) && ne[j] < errors[j]) ? 1 : 0;
      if (mask[j]) {
        errors[j] = ne[j];
        improved = true;
      }
    }
    HIPCHK(hipMemcpyAsync(dmask, mask.data(), sizeof(int) * nrhs, hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(accept_kernel, dim3(blocks_for(n * nrhs)), dim3(256), 0, ctx->stream, x64, (const double*)nx, r,
                       (const double*)nr, n, nrhs, (const int*)dmask);
    HIPCHK(hipGetLastError());
    steps = step;
    out.history.insert(out.history.end(), errors.begin(), errors.end());
    out.hist_len += 1;
    if (!improved) {
      status = BEAT_REFINE_STAGNATED;
      break;
    }
  }
  if (all_le(errors)) status = BEAT_REFINE_CONVERGED;
  std::vector<double> returned = errors;
  if (returned_round) {
    hipLaunchKernelGGL(round32_kernel, dim3(grid_cap(2 * n * nrhs)), dim3(256), 0, ctx->stream, (const double*)x64, nx,
                       2 * n * nrhs);
    HIPCHK(hipGetLastError());
    if (residual_all(ctx, A64, n, lda64, nx, b64, nr, nrhs) != BEAT_OK) return BEAT_ERR_HIP;
    std::vector<double> rr;
    if (col_norms(ctx, nr, n, nrhs, rr) != BEAT_OK) return BEAT_ERR_HIP;
    returned = errors_of(rr);
  }
  bool nonfinite_w = false, nonfinite_r = false;
  for (double v : errors) nonfinite_w |= !std::isfinite(v);
  for (double v : returned) nonfinite_r |= !std::isfinite(v);
  const int working = nonfinite_w ? BEAT_REFINE_NONFINITE : status;
  int st;
  if (nonfinite_r) st = BEAT_REFINE_NONFINITE;
  else if (all_le(returned)) st = BEAT_REFINE_CONVERGED;
  else if (working == BEAT_REFINE_CONVERGED) st = BEAT_REFINE_ROUNDING_LIMITED;
  else st = working;
  out.steps = steps;
  out.status = st;
  out.working = working;
  out.working_res = errors;
  out.returned_res = returned;
  return BEAT_OK;
}

void fill_report(beat_refine_report* rep, const Refine& R, double rtol, int nrhs) {
  if (!rep) return;
  rep->steps = R.steps;
  rep->status = R.status;
  rep->working_status = R.working;
  rep->history_len = R.hist_len;
  rep->rtol = rtol;
  if (rep->working_residuals)
    for (int j = 0; j < nrhs; ++j) rep->working_residuals[j] = R.working_res[j];
  if (rep->returned_residuals)
    for (int j = 0; j < nrhs; ++j) rep->returned_residuals[j] = R.returned_res[j];
  if (rep->history)
    for (size_t i = 0; i < R.history.size(); ++i) rep->history[i] = R.history[i];
}

template <typename T>
T* mesh_rec(const beat_mesh* m);
template <>
double* mesh_rec<double>(const beat_mesh* m) {
  return m->rec64;
}
template <>
float* mesh_rec<float>(const beat_mesh* m) {
  return m->rec32;
}

// The trial tile partition of the regular pass: wide, or narrow (beside a
// factorization).
template <typename T>
void set_trial_tiles(const beat_mesh* m, bool narrow, RegularArgs<T>& ra) {
  ra.narrow = narrow ? 1 : 0;
  ra.flocs = narrow ? m->flocn : m->flocs;
  ra.stile_start = narrow ? m->ntile_start : m->stile_start;
  ra.stile_count = narrow ? m->ntile_count : m->stile_count;
  ra.stile_nv = narrow ? m->ntile_nv : m->stile_nv;
  ra.stile_verts = narrow ? m->ntile_verts : m->stile_verts;
  ra.nstiles = (int)(narrow ? m->h.ntile_start.size() : m->h.stile_start.size());
}

// Geometry, tiles, rule, wavenumber and images of the regular pass.
template <typename T>
void regular_geometry(const beat_ctx* ctx, const beat_mesh* m, double k, RegularArgs<T>& ra) {
  ra.rec = mesh_rec<T>(m);
  ra.fverts = m->fverts;
  ra.floc = m->floc;
  ra.tile_start = m->tile_start;
  ra.tile_count = m->tile_count;
  ra.tile_nv = m->tile_nv;
  ra.tile_verts = m->tile_verts;
  ra.ntiles = (int)m->h.tile_start.size();
  set_trial_tiles(m, false, ra);
  ra.qtile_start = m->qtile_start;
  ra.qtile_count = m->qtile_count;
  ra.qtile_nv = m->qtile_nv;
  ra.qtile_verts = m->qtile_verts;
  ra.flocq = m->flocq;
  ra.tile_qfirst = m->tile_qfirst;
  ra.unit_start = m->sym_unit_start;
  ra.frot = m->frot;
  ra.nqtiles = (int)m->h.qtile_start.size();
  ra.nunits = m->h.sym_unit_start.empty() ? 0 : m->h.sym_unit_start.back();
  ra.symmetric = ctx->regular_symmetric ? 1 : 0;
  ra.k = (T)k;
  ra.krev = (T)(k / (2.0 * M_PI));
  ra.ksteps = phase_steps_per_unit(k);
  ra.trig = ctx->trig;
  for (int i = 0; i < 6; ++i) {
    const bool on = i < m->h.nq;
    const double xi = on ? m->h.rule_pts[2 * i] : 0, eta = on ? m->h.rule_pts[2 * i + 1] : 0;
    ra.w[i] = on ? (T)m->h.rule_w[i] : T(0);
    ra.phi[i][0] = (T)(1.0 - xi - eta);
    ra.phi[i][1] = (T)xi;
    ra.phi[i][2] = (T)eta;
    ra.pw[i][0] = (T)((1.0 - xi - eta) * (on ? m->h.rule_w[i] : 0.0));
    ra.pw[i][1] = (T)(xi * (on ? m->h.rule_w[i] : 0.0));
    ra.pw[i][2] = (T)(eta * (on ? m->h.rule_w[i] : 0.0));
  }
  ra.nimg = m->h.nimg;
  for (int x = 0; x < MAX_IMAGES; ++x)
    for (int c = 0; c < 3; ++c) {
      const bool on = x < m->h.nimg;
      ra.img[x][c] = on ? (T)m->h.img[x].s[c] : T(1);
      ra.img[x][3 + c] = on ? (T)(m->h.img[x].det * m->h.img[x].s[c]) : T(1);
    }
  ra.one_per_cu = 0;
}

// Pairs, rules, wavenumber and reflection table of the singular pass.
template <typename TO>
void singular_geometry(const beat_ctx* ctx, const beat_mesh* m, double k, SingularArgs<TO>& sa) {
  sa.trig = ctx->trig;
  sa.ksteps = phase_steps_per_unit(k);
  sa.wide = !(std::fabs(k) * m->diameter < SINCOS_TABLE_LIMIT) ? 1 : 0;
  sa.rec = m->rec64;
  sa.verts = m->verts;
  sa.fverts = m->fverts;
  sa.pairs = m->spairs;
  sa.npairs = (int64_t)m->h.spairs.size();
  sa.rule_off = m->rule_off;
  sa.rule_cnt = m->rule_cnt;
  sa.rule_data = m->rule_data;
  sa.k = k;
  for (int x = 0; x <= MAX_IMAGES; ++x)
    for (int c = 0; c < 3; ++c) {
      const bool on = x >= 1 && x <= m->h.nimg;
      sa.img[x][c] = on ? m->h.img[x - 1].s[c] : 1.0;
      sa.img[x][3 + c] = on ? m->h.img[x - 1].det * m->h.img[x - 1].s[c] : 1.0;
    }
}

// Assemble A (row-major, lda) and B (n x nrhs row-major) in precision T from the
// internal-order q64 (and q32 for T = float).
// st: the stream; ev_reg0 / ev_reg1 (may be null) bracket the regular passes;
// one_per_cu: run the regular kernel at one workgroup per CU; narrow: its narrow
// trial tiles (the assembly runs beside a factorization).
template <typename T>
int assemble(beat_ctx* ctx, const beat_mesh* m, double k, double er, double ei, const double* q64, const T* qT,
             int nrhs, T* A, int64_t lda, T* B, hipStream_t st, hipEvent_t ev_reg0, hipEvent_t ev_reg1,
             bool one_per_cu = false, bool narrow = false, int tiles = -1) {
  const int64_t n = m->h.nv;
  HIPCHK(launch_fill<T>(A, 2 * n * lda, T(0), st));
  HIPCHK(launch_fill<T>(B, 2 * n * nrhs, T(0), st));
  RegularArgs<T> ra;
  regular_geometry<T>(ctx, m, k, ra);
  ra.A = A;
  ra.lda = lda;
  ra.B = B;
  ra.q = qT;
  ra.nrhs = nrhs;
  ra.eta_re = (T)er;
  ra.eta_im = (T)ei;
  ra.one_per_cu = one_per_cu ? 1 : 0;
  // the tuned tiles, when the autotuner has chosen them
  if (tiles >= 0) {
    ra.symmetric = tiles == 0 ? 1 : 0;
    narrow = tiles == 2;
  }
  if (ev_reg0) HIPCHK(hipEventRecord(ev_reg0, st));
  const bool wide = !(std::fabs(k) * m->diameter < SINCOS_TABLE_LIMIT);
  set_trial_tiles(m, narrow, ra);
  HIPCHK(launch_regular<T>(ra, m->h.nq, er == 0.0, wide, false, st));
  set_trial_tiles(m, false, ra);  // image passes: the wide partition
  if (m->h.nimg > 0) HIPCHK(launch_regular<T>(ra, m->h.nq, er == 0.0, wide, true, st));
  if (ev_reg1) HIPCHK(hipEventRecord(ev_reg1, st));
  SingularArgs<T> sa;
  singular_geometry<T>(ctx, m, k, sa);
  sa.A = A;
  sa.lda = lda;
  sa.B = B;
  sa.q = q64;
  sa.nrhs = nrhs;
  sa.eta_re = er;
  sa.eta_im = ei;
  HIPCHK(launch_singular<T>(sa, st));
  IdentityArgs<T> ia;
  ia.rec = m->rec64;
  ia.fverts = m->fverts;
  ia.nf = m->h.nf;
  ia.A = A;
  ia.lda = lda;
  ia.B = B;
  ia.q = q64;
  ia.nrhs = nrhs;
  ia.eta_re = er;
  ia.eta_im = ei;
  for (int p = 0; p < 3; ++p) {
    ia.pd[p] = m->h.id_pd[p];
    for (int b = 0; b < 3; ++b) ia.pp[p][b] = m->h.id_pp[p][b];
  }
  HIPCHK(launch_identity<T>(ia, st));
  HIPCHK(launch_row_scale<T>(A, lda, n, B, nrhs, m->wrows, m->wvals, (int64_t)m->h.weighted_rows.size(), st));
  return BEAT_OK;
}

// The four operators S, K' (n x ldf) and D, H (n x ldd), row-major, internal
// numbering, in precision T: regular and image passes, touching and image-touching
// pairs, orbit row weights.
template <typename T>
int assemble_operators(beat_ctx* ctx, const beat_mesh* m, double k, T* S, T* D, T* K, T* H, int64_t ldd,
                       int64_t ldf) {
  const int64_t n = m->h.nv;
  hipStream_t s = ctx->stream;
  HIPCHK(launch_fill<T>(S, 2 * n * ldf, T(0), s));
  HIPCHK(launch_fill<T>(K, 2 * n * ldf, T(0), s));
  HIPCHK(launch_fill<T>(D, 2 * n * ldd, T(0), s));
  HIPCHK(launch_fill<T>(H, 2 * n * ldd, T(0), s));
  OperatorArgs<T> oa;
  regular_geometry<T>(ctx, m, k, oa.g);
  oa.g.A = oa.g.B = nullptr;
  oa.g.q = nullptr;
  oa.g.lda = 0;
  oa.g.nrhs = 0;
  oa.g.eta_re = oa.g.eta_im = T(0);
  oa.S = S;
  oa.D = D;
  oa.K = K;
  oa.H = H;
  oa.ldd = ldd;
  oa.ldf = ldf;
  const bool wide = !(std::fabs(k) * m->diameter < SINCOS_TABLE_LIMIT);
  HIPCHK(launch_operators<T>(oa, m->h.nq, wide, false, s));
  if (m->h.nimg > 0) HIPCHK(launch_operators<T>(oa, m->h.nq, wide, true, s));
  SingularOperatorArgs<T> so;
  singular_geometry<T>(ctx, m, k, so.g);
  so.g.A = so.g.B = nullptr;
  so.g.q = nullptr;
  so.g.lda = 0;
  so.g.nrhs = 0;
  so.g.eta_re = so.g.eta_im = 0;
  so.S = S;
  so.D = D;
  so.K = K;
  so.H = H;
  so.ldd = ldd;
  so.ldf = ldf;
  HIPCHK(launch_singular_operators<T>(so, s));
  const int64_t nw = (int64_t)m->h.weighted_rows.size();
  HIPCHK(launch_row_scale<T>(D, ldd, n, S, (int)ldf, m->wrows, m->wvals, nw, s));
  HIPCHK(launch_row_scale<T>(H, ldd, n, K, (int)ldf, m->wrows, m->wvals, nw, s));
  return BEAT_OK;
}

int check_mesh(beat_ctx* ctx, const beat_mesh* m) {
  if (!ctx) return BEAT_ERR_ARG;
  if (!m || m->ctx != ctx) return fail(ctx, BEAT_ERR_ARG, "mesh does not belong to this context");
  return BEAT_OK;
}

// Device buffers of one system in flight: q (internal order), the assembled
// matrix and right-hand sides. a32 / b32 hold the complex64 system of the FP32
// path, or the complex64 copy factored by the refined path.
struct SystemBufs {
  DevBuf *q64, *q32, *a64, *b64, *a32, *b32;
  // the matrix the FP64 and FP32 paths factor: a copy of the assembled one (a64,
  // a32), which stays as assembled for a second factorization with exchanges
  // (the refined path factors a32, its copy of a64)
  DevBuf* af;
};

// Timing events of one system: assembly start, regular passes, assembly end.
struct AsmEvents {
  hipEvent_t start, reg0, reg1, end;
};

// Stage 1 (stream st): gather q_dev (caller column-major complex128) into internal
// order and assemble the system of the precision: complex128 for FP64 and the
// refined path (plus the complex64 copy the refined path factors), complex64 for
// FP32 (the complex128 system rounded once, or the complex64 kernels).
int stage_assemble(beat_ctx* ctx, const beat_mesh* m, int precision, double k, double er, double ei,
                   const double* q_dev, int nrhs, SystemBufs& sb, hipStream_t st, const AsmEvents& ev,
                   bool one_per_cu = false, bool narrow = false) {
  const int64_t n = m->h.nv, nf = m->h.nf, lda = lda_for(n);
  const int tiles = tuned(ctx, precision, n).asm_tile;
  HIPCHK(hipEventRecord(ev.start, st));
  HIPCHK(ensure(*sb.q64, sizeof(double) * 2 * nf * nrhs));
  hipLaunchKernelGGL((gather_q_kernel<double>), dim3(blocks_for(nf * nrhs)), dim3(256), 0, st, q_dev, nf, nrhs,
                     (const int32_t*)m->face_perm, as<double>(*sb.q64));
  HIPCHK(hipGetLastError());
  if (precision == BEAT_PREC_FP32 && !ctx->fp32_assembly_fp64) {
    HIPCHK(ensure(*sb.q32, sizeof(float) * 2 * nf * nrhs));
    HIPCHK(launch_downcast(as<double>(*sb.q64), as<float>(*sb.q32), 2 * nf * nrhs, st));
    HIPCHK(ensure(*sb.a32, sizeof(float) * 2 * n * lda));
    HIPCHK(ensure(*sb.b32, sizeof(float) * 2 * n * nrhs));
    if (int rc = assemble<float>(ctx, m, k, er, ei, as<double>(*sb.q64), as<float>(*sb.q32), nrhs, as<float>(*sb.a32),
                                 lda, as<float>(*sb.b32), st, ev.reg0, ev.reg1, one_per_cu, narrow, tiles))
      return rc;
    HIPCHK(ensure(*sb.af, sizeof(float) * 2 * n * lda));
    HIPCHK(hipMemcpyAsync(sb.af->p, sb.a32->p, sizeof(float) * 2 * n * lda, hipMemcpyDeviceToDevice, st));
  } else if (precision == BEAT_PREC_FP32) {
    // the complex64 system as the complex128 assembly rounded once: the FP64
    // regular kernel (table phase, no transcendental instructions) outruns the
    // FP32 one, and the rounded matrix is at least as accurate
    HIPCHK(ensure(*sb.a64, sizeof(double) * 2 * n * lda));
    HIPCHK(ensure(*sb.b64, sizeof(double) * 2 * n * nrhs));
    if (int rc = assemble<double>(ctx, m, k, er, ei, as<double>(*sb.q64), as<double>(*sb.q64), nrhs,
                                  as<double>(*sb.a64), lda, as<double>(*sb.b64), st, ev.reg0, ev.reg1, one_per_cu,
                                  narrow, tiles))
      return rc;
    HIPCHK(ensure(*sb.a32, sizeof(float) * 2 * n * lda));
    HIPCHK(ensure(*sb.b32, sizeof(float) * 2 * n * nrhs));
    HIPCHK(launch_downcast(as<double>(*sb.a64), as<float>(*sb.a32), 2 * n * lda, st));
    HIPCHK(launch_downcast(as<double>(*sb.b64), as<float>(*sb.b32), 2 * n * nrhs, st));
    HIPCHK(ensure(*sb.af, sizeof(float) * 2 * n * lda));
    HIPCHK(hipMemcpyAsync(sb.af->p, sb.a32->p, sizeof(float) * 2 * n * lda, hipMemcpyDeviceToDevice, st));
  } else {
    HIPCHK(ensure(*sb.a64, sizeof(double) * 2 * n * lda));
    HIPCHK(ensure(*sb.b64, sizeof(double) * 2 * n * nrhs));
    if (int rc = assemble<double>(ctx, m, k, er, ei, as<double>(*sb.q64), as<double>(*sb.q64), nrhs,
                                  as<double>(*sb.a64), lda, as<double>(*sb.b64), st, ev.reg0, ev.reg1, one_per_cu,
                                  narrow, tiles))
      return rc;
    if (precision == BEAT_PREC_FP64) {
      HIPCHK(ensure(*sb.af, sizeof(double) * 2 * n * lda));
      HIPCHK(hipMemcpyAsync(sb.af->p, sb.a64->p, sizeof(double) * 2 * n * lda, hipMemcpyDeviceToDevice, st));
    }
    if (precision == BEAT_PREC_FP32_REFINED) {
      // the complex64 copy the refined path factors, made here so that in a sweep
      // it is off the factorization's stream
      HIPCHK(ensure(*sb.a32, sizeof(float) * 2 * n * lda));
      HIPCHK(ensure(*sb.b32, sizeof(float) * 2 * n * nrhs));
      HIPCHK(launch_downcast(as<double>(*sb.a64), as<float>(*sb.a32), 2 * n * lda, st));
      HIPCHK(launch_downcast(as<double>(*sb.b64), as<float>(*sb.b32), 2 * n * nrhs, st));
    }
  }
  HIPCHK(hipEventRecord(ev.end, st));
  return BEAT_OK;
}

// Stage 2 (context stream, after the assembly): factor once, solve every column
// (and refine), write the pressures (caller column-major) to p_dev, then wait for
// the stream and report. The assembled buffers of sb are consumed.
int stage_solve(beat_ctx* ctx, const beat_mesh* m, int precision, int nrhs, SystemBufs& sb, double* p_dev,
                const beat_solve_options* opts, const AsmEvents& asm_ev, beat_solve_report* rep) {
  const int64_t n = m->h.nv;
  const int64_t lda = lda_for(n);
  hipStream_t s = ctx->stream;
  int max_steps = opts ? opts->max_steps : 3;
  double rtol = opts ? opts->rtol : 1e-10;
  HIPCHK(hipStreamWaitEvent(s, asm_ev.end, 0));
  LuWork lw;
  if (ensure_lu_work<double>(ctx, n, lw) != BEAT_OK) return BEAT_ERR_HIP;
  const Tuning tu = tuned(ctx, precision, n);
  if (tu.lu_layout == 1) {  // tuned: one stream for this precision and size class
    lw.side = nullptr;
    lw.inb = nullptr;
  }
  lw.gemm_tile = std::max<int>(0, tu.gemm_tile);
  int panels = 0, info = 0, factorizations = 0;
  Refine R;
  // the right-hand sides ride along in the padding columns of the factored
  // matrix when they fit: the factorization then does the first forward solve
  const bool carry = nrhs <= SOLVE_CHUNK && n + nrhs <= lda;
  const int64_t nc = carry ? n + nrhs : 0;
  // First the speculative factorization (no row exchanges, every column checked
  // against partial pivoting's choice); if a column would have exchanged rows,
  // the matrix is taken again as assembled and factored with the exchange panels.
  for (int attempt = 0; attempt < 2; ++attempt) {
    lw.nopiv = attempt == 0;
    if (attempt == 1) {
      // the factored buffer back to the assembled matrix
      if (precision == BEAT_PREC_FP64)
        HIPCHK(hipMemcpyAsync(sb.af->p, sb.a64->p, sizeof(double) * 2 * n * lda, hipMemcpyDeviceToDevice, s));
      else if (precision == BEAT_PREC_FP32)
        HIPCHK(hipMemcpyAsync(sb.af->p, sb.a32->p, sizeof(float) * 2 * n * lda, hipMemcpyDeviceToDevice, s));
      else
        HIPCHK(launch_downcast(as<double>(*sb.a64), as<float>(*sb.a32), 2 * n * lda, s));
    }
    ++factorizations;
    HIPCHK(hipEventRecord(ctx->ev[1], s));
    if (precision == BEAT_PREC_FP64) {
      double* F = as<double>(*sb.af);
      if (carry) HIPCHK(lu_append_rhs<double>(F, n, lda, as<double>(*sb.b64), nrhs, s));
      HIPCHK(lu_factor<double>(F, n, lda, lw, LU_NB, s, &panels, nc));
      HIPCHK(hipEventRecord(ctx->ev[2], s));
      // the solution overwrites b64 (kept for a second attempt: lu_append_rhs copied it)
      HIPCHK(ensure(ctx->x64, sizeof(double) * 2 * n * nrhs));
      double* X = as<double>(ctx->x64);
      HIPCHK(hipMemcpyAsync(X, sb.b64->p, sizeof(double) * 2 * n * nrhs, hipMemcpyDeviceToDevice, s));
      int rc = carry ? solve_eliminated<double>(ctx, F, n, lda, lw.ipiv, X, nrhs)
                     : solve_all<double>(ctx, F, n, lda, lw.ipiv, X, nrhs, true);
      if (rc) return rc;
      hipLaunchKernelGGL((scatter_p_kernel<double>), dim3(blocks_for(n * nrhs)), dim3(256), 0, s, (const double*)X, n,
                         nrhs, (const int32_t*)m->vert_inv, p_dev);
      HIPCHK(hipGetLastError());
    } else if (precision == BEAT_PREC_FP32) {
      float* F = as<float>(*sb.af);
      if (carry) HIPCHK(lu_append_rhs<float>(F, n, lda, as<float>(*sb.b32), nrhs, s));
      HIPCHK(lu_factor<float>(F, n, lda, lw, LU_NB, s, &panels, nc));
      HIPCHK(hipEventRecord(ctx->ev[2], s));
      HIPCHK(ensure(ctx->d32, sizeof(float) * 2 * n * nrhs + sizeof(int) * 64 + 4096));
      float* X = as<float>(ctx->d32);
      HIPCHK(hipMemcpyAsync(X, sb.b32->p, sizeof(float) * 2 * n * nrhs, hipMemcpyDeviceToDevice, s));
      int rc = carry ? solve_eliminated<float>(ctx, F, n, lda, lw.ipiv, X, nrhs)
                     : solve_all<float>(ctx, F, n, lda, lw.ipiv, X, nrhs, true);
      if (rc) return rc;
      hipLaunchKernelGGL((scatter_p_kernel<float>), dim3(blocks_for(n * nrhs)), dim3(256), 0, s, (const float*)X, n,
                         nrhs, (const int32_t*)m->vert_inv, p_dev);
      HIPCHK(hipGetLastError());
    } else {
      HIPCHK(ensure(ctx->x64, sizeof(double) * 2 * n * nrhs));
      if (carry) HIPCHK(lu_append_rhs<float>(as<float>(*sb.a32), n, lda, as<float>(*sb.b32), nrhs, s));
      // these factors only precondition the complex128 refinement: the trailing
      // updates may use the three-real-product complex form
      lw.gauss3 = true;
      HIPCHK(lu_factor<float>(as<float>(*sb.a32), n, lda, lw, LU_NB, s, &panels, nc));
      HIPCHK(hipEventRecord(ctx->ev[2], s));
      // the first solution in x64 through d32 (b32 stays for a second attempt; the
      // refinement's residuals are complex128, from b64)
      HIPCHK(ensure(ctx->d32, sizeof(float) * 2 * n * nrhs + sizeof(int) * 64 + 4096));
      float* X = as<float>(ctx->d32);
      HIPCHK(hipMemcpyAsync(X, sb.b32->p, sizeof(float) * 2 * n * nrhs, hipMemcpyDeviceToDevice, s));
      int rc = carry ? solve_eliminated<float>(ctx, as<float>(*sb.a32), n, lda, lw.ipiv, X, nrhs)
                     : solve_all<float>(ctx, as<float>(*sb.a32), n, lda, lw.ipiv, X, nrhs, true);
      if (rc) return rc;
      hipLaunchKernelGGL(upcast_kernel, dim3(grid_cap(2 * n * nrhs)), dim3(256), 0, s, (const float*)X,
                         as<double>(ctx->x64), 2 * n * nrhs);
      HIPCHK(hipGetLastError());
      rc = refine_loop(ctx, as<float>(*sb.a32), lda, lw.ipiv, as<double>(*sb.a64), lda, as<double>(*sb.b64),
                       as<double>(ctx->x64), n, nrhs, max_steps, rtol, false, R);
      if (rc) return rc;
      hipLaunchKernelGGL((scatter_p_kernel<double>), dim3(blocks_for(n * nrhs)), dim3(256), 0, s,
                         (const double*)as<double>(ctx->x64), n, nrhs, (const int32_t*)m->vert_inv, p_dev);
      HIPCHK(hipGetLastError());
    }
    HIPCHK(hipEventRecord(ctx->ev[3], s));
    // the factorization's zero-pivot report, wait control and speculation check,
    // after the stream drained
    HIPCHK(hipStreamSynchronize(s));
    bool exchange = false;
    if (int rc = lu_outcome(ctx, lw, &info, &exchange)) return rc;
    if (!exchange) break;
  }
  if (rep) {
    float ta, tr, t12, t23;
    HIPCHK(hipEventElapsedTime(&ta, asm_ev.start, asm_ev.end));
    HIPCHK(hipEventElapsedTime(&tr, asm_ev.reg0, asm_ev.reg1));
    HIPCHK(hipEventElapsedTime(&t12, ctx->ev[1], ctx->ev[2]));
    HIPCHK(hipEventElapsedTime(&t23, ctx->ev[2], ctx->ev[3]));
    rep->regular_kernel_s = tr * 1e-3;
    rep->assembly_s = ta * 1e-3;
    rep->factor_s = t12 * 1e-3;
    rep->solve_s = t23 * 1e-3;
    rep->factorization_count = factorizations;
    rep->lu_panels = panels;
    rep->regular_pairs = m->regular_pairs;
    rep->singular_pairs = (int64_t)m->h.spairs.size();
    if (precision == BEAT_PREC_FP32_REFINED) {
      fill_report(&rep->refine, R, rtol, nrhs);
    } else {
      rep->refine.steps = 0;
      rep->refine.status = BEAT_REFINE_NONE;
      rep->refine.working_status = BEAT_REFINE_NONE;
      rep->refine.history_len = 0;
      rep->refine.rtol = rtol;
    }
  }
  if (info != 0) return fail(ctx, BEAT_ERR_SINGULAR, "LU found an exactly zero pivot at column " + std::to_string(info));
  return BEAT_OK;
}

int check_solve_options(beat_ctx* ctx, int precision, const beat_solve_options* opts) {
  const int max_steps = opts ? opts->max_steps : 3;
  const double rtol = opts ? opts->rtol : 1e-10;
  if (max_steps < 1 || max_steps > 8) return fail(ctx, BEAT_ERR_ARG, "max_steps must be in 1..8");
  if (!(rtol > 0) || !std::isfinite(rtol)) return fail(ctx, BEAT_ERR_ARG, "rtol must be finite and positive");
  if (precision != BEAT_PREC_FP64 && precision != BEAT_PREC_FP32 && precision != BEAT_PREC_FP32_REFINED)
    return fail(ctx, BEAT_ERR_ARG, "unknown precision");
  return BEAT_OK;
}

// q_dev: caller column-major complex128 (device); p_dev: caller column-major out.
int solve_impl(beat_ctx* ctx, const beat_mesh* m, int precision, double k, double er, double ei, const double* q_dev,
               int nrhs, double* p_dev, const beat_solve_options* opts, beat_solve_report* rep) {
  if (int rc = check_solve_options(ctx, precision, opts)) return rc;
  SystemBufs sb{&ctx->q64, &ctx->q32, &ctx->a64, &ctx->b64, &ctx->a32, &ctx->b32, &ctx->af};
  const AsmEvents ev{ctx->ev[0], ctx->ev[4], ctx->ev[5], ctx->ev[6]};
  if (int rc = stage_assemble(ctx, m, precision, k, er, ei, q_dev, nrhs, sb, ctx->stream, ev)) return rc;
  return stage_solve(ctx, m, precision, nrhs, sb, p_dev, opts, ev, rep);
}

// Jobs (mesh, wavenumber, coupling, q) with two systems in flight: the assembly
// of job i + 1 runs on the assembly stream while job i is factored and solved on
// the context stream; a system's buffers are reassembled only after its solve has
// finished. q_dev / p_dev: per job a column-major block (nf_j x nrhs, nv_j x nrhs)
// at the offsets q_off / p_off (complex entries).
int batch_impl(beat_ctx* ctx, int64_t njobs, const beat_mesh* const* meshes, int precision, const double* ks,
               const double* ers, const double* eis, const double* q_dev, const int64_t* q_off, int nrhs, double* p_dev,
               const int64_t* p_off, const beat_solve_options* opts, beat_solve_report* reps) {
  if (int rc = check_solve_options(ctx, precision, opts)) return rc;
  if (njobs <= 0) return BEAT_OK;
  SystemBufs sb[2];
  for (int s2 = 0; s2 < 2; ++s2) {
    sb[s2] = SystemBufs{&ctx->sw_q64[s2], &ctx->sw_q32[s2], &ctx->sw_a64[s2], &ctx->sw_b64[s2],
                        &ctx->sw_a32[s2], &ctx->sw_b32[s2], &ctx->sw_af[s2]};
  }
  // size both systems for the largest job up front (no reallocation while in flight)
  int64_t nmax = 0, fmax = 0;
  for (int64_t i = 0; i < njobs; ++i) {
    nmax = std::max<int64_t>(nmax, meshes[i]->h.nv);
    fmax = std::max<int64_t>(fmax, meshes[i]->h.nf);
  }
  for (int s2 = 0; s2 < 2; ++s2) {
    HIPCHK(ensure(*sb[s2].q64, sizeof(double) * 2 * fmax * nrhs));
    if (precision == BEAT_PREC_FP32) HIPCHK(ensure(*sb[s2].q32, sizeof(float) * 2 * fmax * nrhs));
    if (precision != BEAT_PREC_FP64) {
      HIPCHK(ensure(*sb[s2].a32, sizeof(float) * 2 * nmax * lda_for(nmax)));
      HIPCHK(ensure(*sb[s2].b32, sizeof(float) * 2 * nmax * nrhs));
    }
    if (precision != BEAT_PREC_FP32) {
      HIPCHK(ensure(*sb[s2].a64, sizeof(double) * 2 * nmax * lda_for(nmax)));
      HIPCHK(ensure(*sb[s2].b64, sizeof(double) * 2 * nmax * nrhs));
    }
    if (precision != BEAT_PREC_FP32_REFINED)
      HIPCHK(ensure(*sb[s2].af, (precision == BEAT_PREC_FP64 ? sizeof(double) : sizeof(float)) * 2 * nmax *
                                    lda_for(nmax)));
  }
  {
    // the factorization, solve and refinement workspaces too
    LuWork lw;
    if (ensure_lu_work<double>(ctx, nmax, lw) != BEAT_OK) return BEAT_ERR_HIP;
    HIPCHK(ensure(ctx->work, lu_solve_work_bytes(nmax, sizeof(double))));
    if (precision == BEAT_PREC_FP32_REFINED) {
      HIPCHK(ensure(ctx->x64, sizeof(double) * 2 * nmax * nrhs));
      HIPCHK(ensure(ctx->r64, sizeof(double) * 2 * nmax * nrhs));
      HIPCHK(ensure(ctx->rpart, residual64_part_bytes(nmax)));
      HIPCHK(ensure(ctx->n64, sizeof(double) * 2 * nmax * nrhs));
      HIPCHK(ensure(ctx->rn64, sizeof(double) * 2 * nmax * nrhs));
      HIPCHK(ensure(ctx->d32, sizeof(float) * 2 * nmax * nrhs + sizeof(int) * 64 + 4096));
      HIPCHK(ensure(ctx->norms, refine_norms_bytes(nrhs)));
      HIPCHK(ensure(ctx->lusmall, sizeof(int) * nrhs + 256));
    }
  }
  AsmEvents ev[2];
  for (int s2 = 0; s2 < 2; ++s2)
    ev[s2] = AsmEvents{ctx->sw_ev[s2][0], ctx->sw_ev[s2][1], ctx->sw_ev[s2][2], ctx->sw_ev[s2][3]};
  auto enqueue = [&](int64_t i) -> int {
    const int s2 = (int)(i & 1);
    // the system's previous solve (job i - 2) must be done with its buffers
    if (i >= 2) HIPCHK(hipStreamWaitEvent(ctx->asm_stream, ctx->sw_free[s2], 0));
    // the first job's assembly has nothing beside it: every slot, wide tiles;
    // later ones run beside a factorization: narrow tiles
    const signed char occ = tuned(ctx, precision, meshes[i]->h.nv).asm_occ;
    const bool one_per_cu = i > 0 && (occ < 0 ? ctx->sweep_asm_one_per_cu : occ == 1);
    return stage_assemble(ctx, meshes[i], precision, ks[i], ers[i], eis[i], q_dev + 2 * q_off[i], nrhs, sb[s2],
                          ctx->asm_stream, ev[s2], one_per_cu, i > 0);
  };
  // the reference's marker-file cancellation, checked before each job
  // (cancel_requested, BeatEngineCompiledDriver.jl:1199-1200,1270)
  auto cancelled = [&]() { return !ctx->cancel_path.empty() && access(ctx->cancel_path.c_str(), F_OK) == 0; };
  if (reps)
    for (int64_t i = 0; i < njobs; ++i) reps[i].factorization_count = 0;
  if (cancelled()) return fail(ctx, BEAT_ERR_CANCELLED, "cancelled after 0 frequencies");
  if (int rc = enqueue(0)) return rc;
  int result = BEAT_OK;
  for (int64_t i = 0; i < njobs; ++i) {
    const int s2 = (int)(i & 1);
    if (i > 0 && cancelled()) {
      HIPCHK(hipStreamSynchronize(ctx->asm_stream));
      return fail(ctx, BEAT_ERR_CANCELLED, "cancelled after " + std::to_string(i) + " frequencies");
    }
    if (i + 1 < njobs)
      if (int rc = enqueue(i + 1)) return rc;
    ctx->shared_device = i + 1 < njobs;
    int rc = stage_solve(ctx, meshes[i], precision, nrhs, sb[s2], p_dev + 2 * p_off[i], opts, ev[s2],
                         reps ? reps + i : nullptr);
    ctx->shared_device = false;
    if (rc != BEAT_OK && rc != BEAT_ERR_SINGULAR) {
      (void)hipStreamSynchronize(ctx->asm_stream);
      return rc;
    }
    if (rc == BEAT_ERR_SINGULAR) result = rc;
    HIPCHK(hipEventRecord(ctx->sw_free[s2], ctx->stream));
  }
  HIPCHK(hipStreamSynchronize(ctx->asm_stream));
  return result;
}

// ---- grouped batches: many small systems factored, solved and refined together ----
//
// Jobs are taken in groups of up to GROUP_MAX. A group's systems are assembled one
// by one (eager launches: the wavenumber is in their arguments), then one
// captured hipGraph does the rest for the whole group: the batched speculative LU
// (one launch per kernel and leaf for every system), the backward substitutions
// of the carried right-hand sides, and for the refined path the complex128
// refinement with its per-column acceptance decided on the device. The graph
// depends only on the group's shape (sizes, precision, options) and buffers, so
// groups of the same shape (the frequencies of one mesh) replay it. A system
// whose speculative factorization would have exchanged rows is solved again on
// its own by the per-job path.
constexpr int GROUP_MAX = 32;
constexpr int64_t GROUP_NMAX = 4096;  // larger systems keep the per-job pipeline

// Device state of one system's refinement (beat_refine_cpu_lu's loop,
// BeatEngineCpuRefinement.jl:5-78, with its decisions taken on the device).
struct RefineState {
  int steps, status, done, skip, hist_len, working, pad[2];
  int mask[8], acc[8];
  double bn[8], err[8], hist[9 * 8];
};

struct BatchJobDesc {
  int64_t n;
  const int32_t* vert_inv;
  double* p_out;  // caller column-major n x nrhs (eager scatter only)
};

__device__ __forceinline__ bool all_le(const double* e, int nrhs, double rtol) {
  for (int d = 0; d < nrhs; ++d)
    if (!(e[d] <= rtol)) return false;
  return true;
}

// mode 0: initial residual norms -> errors, history row 0, first decision;
// mode 1: after correction step `step` -> per-column acceptance, history, decision;
// mode 2: final statuses. One thread per system. norms: [bn | rn] per system (8 + 8).
__global__ void refine_control_kernel(RefineState* st, const double* norms, int nsys, int nrhs, double rtol, int mode,
                                      int step) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nsys) return;
  RefineState& S = st[j];
  const double* bn = norms + 16 * j;
  const double* rn = bn + 8;
  auto error_of = [&](int d, double nr) {
    return S.bn[d] == 0 ? (nr == 0 ? 0.0 : INFINITY) : nr / S.bn[d];
  };
  if (mode == 0) {
    for (int d = 0; d < 8; ++d) {
      S.bn[d] = d < nrhs ? bn[d] : 0.0;
      S.err[d] = d < nrhs ? error_of(d, rn[d]) : 0.0;
      S.hist[d] = S.err[d];
      S.mask[d] = d < nrhs && !(S.err[d] <= rtol);
      S.acc[d] = 0;
    }
    S.hist_len = 1;
    S.steps = 0;
    S.status = BEAT_REFINE_STEP_LIMIT;
    S.done = 0;
    if (all_le(S.err, nrhs, rtol)) {
      S.status = BEAT_REFINE_CONVERGED;
      S.done = 1;
    }
    S.skip = S.done;
    return;
  }
  if (mode == 1) {
    if (S.skip) return;
    bool improved = false;
    for (int d = 0; d < nrhs; ++d) {
      const double ne = error_of(d, rn[d]);
      S.acc[d] = isfinite(ne) && ne < S.err[d];
      if (S.acc[d]) {
        S.err[d] = ne;
        improved = true;
      }
    }
    S.steps = step;
    for (int d = 0; d < 8; ++d) S.hist[S.hist_len * 8 + d] = S.err[d];
    S.hist_len += 1;
    if (!improved) {
      S.status = BEAT_REFINE_STAGNATED;
      S.done = 1;
    } else if (all_le(S.err, nrhs, rtol)) {
      S.status = BEAT_REFINE_CONVERGED;
      S.done = 1;
    }
    for (int d = 0; d < nrhs; ++d) S.mask[d] = !(S.err[d] <= rtol);
    return;
  }
  if (all_le(S.err, nrhs, rtol)) S.status = BEAT_REFINE_CONVERGED;
  bool nonfinite = false;
  for (int d = 0; d < nrhs; ++d) nonfinite |= !isfinite(S.err[d]);
  S.working = nonfinite ? BEAT_REFINE_NONFINITE : S.status;
  // the returned solution is the complex128 working one: returned = working errors
  if (nonfinite) S.status = BEAT_REFINE_NONFINITE;
}

// the next step runs only for systems whose refinement has not finished
__global__ void refine_advance_kernel(RefineState* st, int nsys) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < nsys) st[j].skip = st[j].done;
}

// Column 2-norms of each system's vector block (norms[16 j + off + d]); one
// workgroup per system. Skipped with the system's skip flag (when given).
__global__ __launch_bounds__(256) void colnorms_batched_kernel(const BatchVec<double>* vec, int nrhs, double* norms,
                                                               int off) {
  const BatchVec<double> V = vec[blockIdx.x];
  if (V.skip && *V.skip) return;
  __shared__ double red[8][256];
  const int t = threadIdx.x;
  double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int64_t r = t; r < V.n; r += 256)
#pragma unroll
    for (int d = 0; d < 8; ++d)
      if (d < nrhs) {
        const double a = V.v[2 * (r * V.ld + d)], b = V.v[2 * (r * V.ld + d) + 1];
        acc[d] += a * a + b * b;
      }
#pragma unroll
  for (int d = 0; d < 8; ++d) red[d][t] = acc[d];
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (t < o)
#pragma unroll
      for (int d = 0; d < 8; ++d) red[d][t] += red[d][t + o];
    __syncthreads();
  }
  if (t < nrhs) norms[16 * blockIdx.x + off + t] = sqrt(red[t][0]);
}

// Elementwise work on every system's n x nrhs block (blockIdx.y: the system):
//  0  dst64 = (double) src32            (the first solution, from the carried columns)
//  1  dst32 = mask ? (float) src64 : 0  (the correction's right-hand side)
//  2  dst64 = a64 + (double) src32      (corrected solution)
//  3  a64 = b64, c64 = d64 where accepted (accepted columns of solution and residual)
//  4  dst64 = src64 (complex128 copy)
template <int OP>
__global__ void batch_elementwise_kernel(const BatchVec<double>* d64, const BatchVec<float>* s32,
                                         const BatchVec<double>* a64, const BatchVec<double>* b64,
                                         const BatchVec<double>* c64, const BatchVec<double>* e64,
                                         const RefineState* st, int nrhs) {
  const int j = blockIdx.y;
  if (st && OP != 0 && OP != 4 && st[j].skip) return;
  const int64_t n = (OP == 1 || OP == 0) ? s32[j].n : d64[j].n;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * nrhs) return;
  const int64_t r = i / nrhs;
  const int d = (int)(i - r * nrhs);
  if (OP == 0) {
    const BatchVec<float> S = s32[j];
    const BatchVec<double> D = d64[j];
    D.v[2 * (r * D.ld + d)] = (double)S.v[2 * (r * S.ld + d)];
    D.v[2 * (r * D.ld + d) + 1] = (double)S.v[2 * (r * S.ld + d) + 1];
  } else if (OP == 1) {
    const BatchVec<double> R = d64[j];
    const BatchVec<float> D = s32[j];
    const bool on = st[j].mask[d] != 0;
    D.v[2 * (r * D.ld + d)] = on ? (float)R.v[2 * (r * R.ld + d)] : 0.f;
    D.v[2 * (r * D.ld + d) + 1] = on ? (float)R.v[2 * (r * R.ld + d) + 1] : 0.f;
  } else if (OP == 2) {
    const BatchVec<double> D = d64[j], X = a64[j];
    const BatchVec<float> S = s32[j];
    D.v[2 * (r * D.ld + d)] = X.v[2 * (r * X.ld + d)] + (double)S.v[2 * (r * S.ld + d)];
    D.v[2 * (r * D.ld + d) + 1] = X.v[2 * (r * X.ld + d) + 1] + (double)S.v[2 * (r * S.ld + d) + 1];
  } else if (OP == 3) {
    if (!st[j].acc[d]) return;
    const BatchVec<double> X = d64[j], NX = a64[j], R = b64[j], NR = c64[j];
    X.v[2 * (r * X.ld + d)] = NX.v[2 * (r * NX.ld + d)];
    X.v[2 * (r * X.ld + d) + 1] = NX.v[2 * (r * NX.ld + d) + 1];
    R.v[2 * (r * R.ld + d)] = NR.v[2 * (r * NR.ld + d)];
    R.v[2 * (r * R.ld + d) + 1] = NR.v[2 * (r * NR.ld + d) + 1];
  } else {
    const BatchVec<double> D = d64[j], S = e64[j];
    D.v[2 * (r * D.ld + d)] = S.v[2 * (r * S.ld + d)];
    D.v[2 * (r * D.ld + d) + 1] = S.v[2 * (r * S.ld + d) + 1];
  }
}

// complex64 variant of OP 0 for the plain complex64 path: dst64 = (double) src32
// is OP 0; the complex128 path copies with OP 4. Scatter of each system's internal
// n x nrhs solution to the caller's column-major block.
__global__ void batch_scatter_kernel(const BatchVec<double>* x, const BatchJobDesc* jobs, int nrhs) {
  const int j = blockIdx.y;
  const BatchVec<double> X = x[j];
  const BatchJobDesc J = jobs[j];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= J.n * nrhs) return;
  const int64_t v = i % J.n;
  const int d = (int)(i / J.n);
  const int64_t r = J.vert_inv[v];
  J.p_out[2 * i] = X.v[2 * (r * X.ld + d)];
  J.p_out[2 * i + 1] = X.v[2 * (r * X.ld + d) + 1];
}

// Group layout: per system the strides of every buffer kind (complex entries), so
// the descriptors of a shape never change while the buffers stay.
struct GroupLayout {
  int J = 0;
  int64_t nmax = 0, nfmax = 0, lda = 0;
  size_t mat, vec, qv, linv;
};

int batch_grouped_impl(beat_ctx* ctx, int64_t njobs, const beat_mesh* const* meshes, int precision, const double* ks,
                       const double* ers, const double* eis, const double* q_dev, const int64_t* q_off, int nrhs,
                       double* p_dev, const int64_t* p_off, const beat_solve_options* opts, beat_solve_report* reps) {
  if (int rc = check_solve_options(ctx, precision, opts)) return rc;
  const int max_steps = opts ? opts->max_steps : 3;
  const double rtol = opts ? opts->rtol : 1e-10;
  hipStream_t s = ctx->stream, sa = ctx->asm_stream;
  const bool refined = precision == BEAT_PREC_FP32_REFINED, fp32 = precision == BEAT_PREC_FP32;
  // buffers sized once for the largest group of this batch, two slots: group
  // g + 1 is assembled in one (assembly stream) while group g's graph runs on
  // the other (context stream)
  GroupLayout G;
  for (int64_t i = 0; i < njobs; ++i) {
    G.nmax = std::max<int64_t>(G.nmax, meshes[i]->h.nv);
    G.nfmax = std::max<int64_t>(G.nfmax, meshes[i]->h.nf);
  }
  G.J = (int)std::min<int64_t>(GROUP_MAX, njobs);
  G.lda = lda_for(G.nmax + nrhs);
  G.mat = (size_t)G.nmax * G.lda;
  G.vec = (size_t)G.nmax * nrhs;
  G.qv = (size_t)G.nfmax * nrhs;
  G.linv = (size_t)((G.nmax + 63) / 64) * 64 * 64;
  const size_t J = (size_t)G.J, SJ = 2 * J;  // systems over both slots
  const bool want64 = !fp32, want32 = fp32 || refined;
  // captured graphs hold these buffers' addresses: any reallocation drops them all
  const void* before[] = {ctx->g_a64.p, ctx->g_a32.p, ctx->g_b64.p, ctx->g_b32.p, ctx->g_q64.p, ctx->g_q32.p,
                          ctx->g_linv.p, ctx->g_vec.p, ctx->g_state.p, ctx->g_desc.p};
  HIPCHK(ensure(ctx->g_q64, sizeof(double) * 2 * G.qv * SJ));
  if (fp32) HIPCHK(ensure(ctx->g_q32, sizeof(float) * 2 * G.qv * SJ));
  if (want64) HIPCHK(ensure(ctx->g_a64, sizeof(double) * 2 * G.mat * SJ));
  if (want64) HIPCHK(ensure(ctx->g_b64, sizeof(double) * 2 * G.vec * SJ));
  if (want32) HIPCHK(ensure(ctx->g_a32, sizeof(float) * 2 * G.mat * SJ));
  if (want32) HIPCHK(ensure(ctx->g_b32, sizeof(float) * 2 * G.vec * SJ));
  HIPCHK(ensure(ctx->g_linv, sizeof(double) * 2 * G.linv * SJ));
  // vectors (complex128 n x nrhs): x, r, nx, nr, out; d (complex64)
  HIPCHK(ensure(ctx->g_vec, sizeof(double) * 2 * G.vec * SJ * 5 + sizeof(float) * 2 * G.vec * SJ + 4096));
  HIPCHK(ensure(ctx->g_state, sizeof(RefineState) * SJ + sizeof(int) * 4 * SJ + sizeof(double) * 16 * SJ + 4096));
  // per slot: 2 system arrays, 12 vector arrays, job descriptors
  const size_t desc_slot = 2 * J * sizeof(BatchSystem<double>) + 12 * J * sizeof(BatchVec<double>) +
                           J * sizeof(BatchJobDesc) + 256;
  HIPCHK(ensure(ctx->g_desc, 2 * desc_slot + 256));
  {
    const void* after[] = {ctx->g_a64.p, ctx->g_a32.p, ctx->g_b64.p, ctx->g_b32.p, ctx->g_q64.p, ctx->g_q32.p,
                           ctx->g_linv.p, ctx->g_vec.p, ctx->g_state.p, ctx->g_desc.p};
    if (!std::equal(std::begin(before), std::end(before), std::begin(after))) {
      HIPCHK(hipStreamSynchronize(s));
      for (auto& g : ctx->g_graphs) HIPCHK(hipGraphExecDestroy(g.second));
      ctx->g_graphs.clear();
    }
  }
  double* v64 = as<double>(ctx->g_vec);
  double* vx = v64;
  double* vr = vx + 2 * G.vec * SJ;
  double* vnx = vr + 2 * G.vec * SJ;
  double* vnr = vnx + 2 * G.vec * SJ;
  double* vout = vnr + 2 * G.vec * SJ;
  float* vd = reinterpret_cast<float*>(vout + 2 * G.vec * SJ);
  RefineState* state_all = reinterpret_cast<RefineState*>(ctx->g_state.p);
  int* lustat_all = reinterpret_cast<int*>(state_all + SJ);
  double* norms_all = reinterpret_cast<double*>(lustat_all + 4 * SJ);

  // one slot's device descriptor arrays
  struct SlotDesc {
    BatchSystem<double>* sys64;
    BatchSystem<float>* sys32;
    BatchVec<double> *dX, *dR, *dNX, *dNR, *dB, *dOut, *dRs, *dNXs, *dNRs, *dCar64;
    BatchVec<float> *dD, *dCar32;
    BatchJobDesc* jobs;
    RefineState* state;
    int* lustat;
    double* norms;
  };
  auto slot_desc = [&](int sl) {
    SlotDesc D;
    char* dp = reinterpret_cast<char*>(ctx->g_desc.p) + (size_t)sl * desc_slot;
    D.sys64 = reinterpret_cast<BatchSystem<double>*>(dp);
    D.sys32 = reinterpret_cast<BatchSystem<float>*>(D.sys64 + J);
    BatchVec<double>* dv = reinterpret_cast<BatchVec<double>*>(D.sys32 + J);
    D.dX = dv;
    D.dR = dv + J;
    D.dNX = dv + 2 * J;
    D.dNR = dv + 3 * J;
    D.dB = dv + 4 * J;
    D.dOut = dv + 5 * J;
    D.dRs = dv + 6 * J;
    D.dNXs = dv + 7 * J;
    D.dNRs = dv + 8 * J;
    D.dCar64 = dv + 9 * J;
    D.dD = reinterpret_cast<BatchVec<float>*>(dv + 10 * J);
    D.dCar32 = D.dD + J;
    D.jobs = reinterpret_cast<BatchJobDesc*>(dv + 12 * J);
    D.state = state_all + (size_t)sl * J;
    D.lustat = lustat_all + (size_t)sl * 4 * J;
    D.norms = norms_all + (size_t)sl * 16 * J;
    return D;
  };
  struct HostGroup {
    int64_t g0 = 0;
    int nj = 0;
    int64_t gmax = 0, gncmax = 0;
    std::string key;
    std::vector<BatchSystem<double>> h64;
    std::vector<BatchSystem<float>> h32;
    std::vector<BatchVec<double>> hv;
    std::vector<BatchVec<float>> hd;
    std::vector<BatchJobDesc> hj;
  };
  // descriptors and assembly of group g in slot sl (assembly stream)
  auto prepare = [&](int64_t g0, int sl, HostGroup& H) -> int {
    const SlotDesc D = slot_desc(sl);
    H.g0 = g0;
    H.nj = (int)std::min<int64_t>(G.J, njobs - g0);
    H.gmax = H.gncmax = 0;
    H.h64.assign(J, BatchSystem<double>{});
    H.h32.assign(J, BatchSystem<float>{});
    H.hv.assign(10 * J, BatchVec<double>{});
    H.hd.assign(2 * J, BatchVec<float>{});
    H.hj.assign(J, BatchJobDesc{});
    // the graph's shape: options, strides, slot and the group's sizes (the buffers
    // stay: a reallocation above drops every graph)
    H.key = std::to_string(precision) + "/" + std::to_string(nrhs) + "/" + std::to_string(max_steps) + "/" +
            std::to_string(rtol) + "/" + std::to_string(H.nj) + "/" + std::to_string(G.lda) + "/" +
            std::to_string(G.mat) + "/" + std::to_string(G.vec) + "/" + std::to_string(G.linv) + "/" +
            std::to_string(G.J) + "/slot" + std::to_string(sl);
    for (int jj = 0; jj < H.nj; ++jj) {
      const beat_mesh* m = meshes[g0 + jj];
      const int64_t n = m->h.nv;
      const size_t js = (size_t)sl * J + jj;  // system index over both slots
      H.gmax = std::max(H.gmax, n);
      H.gncmax = std::max(H.gncmax, n + nrhs);
      H.key += ":" + std::to_string(n);
      int* stj = D.lustat + 4 * jj;
      H.h64[jj] = BatchSystem<double>{as<double>(ctx->g_a64) + 2 * G.mat * js, n, G.lda, n + nrhs,
                                      as<double>(ctx->g_linv) + 2 * G.linv * js, stj};
      H.h32[jj] = BatchSystem<float>{as<float>(ctx->g_a32) + 2 * G.mat * js, n, G.lda, n + nrhs,
                                     as<float>(ctx->g_linv) + 2 * G.linv * js, stj};
      const int* skip = &D.state[jj].skip;
      auto v = [&](double* base, const int* sk) { return BatchVec<double>{base + 2 * G.vec * js, n, nrhs, sk}; };
      H.hv[jj] = v(vx, nullptr);
      H.hv[J + jj] = v(vr, nullptr);
      H.hv[2 * J + jj] = v(vnx, skip);
      H.hv[3 * J + jj] = v(vnr, skip);
      H.hv[4 * J + jj] = BatchVec<double>{as<double>(ctx->g_b64) + 2 * G.vec * js, n, nrhs, nullptr};
      H.hv[5 * J + jj] = v(vout, nullptr);
      H.hv[6 * J + jj] = v(vr, skip);
      H.hv[7 * J + jj] = v(vnx, skip);
      H.hv[8 * J + jj] = v(vnr, skip);
      H.hv[9 * J + jj] = BatchVec<double>{as<double>(ctx->g_a64) + 2 * (G.mat * js + n), n, G.lda, nullptr};
      H.hd[jj] = BatchVec<float>{vd + 2 * G.vec * js, n, nrhs, skip};
      H.hd[J + jj] = BatchVec<float>{as<float>(ctx->g_a32) + 2 * (G.mat * js + n), n, G.lda, nullptr};
      H.hj[jj] = BatchJobDesc{n, m->vert_inv, p_dev + 2 * p_off[g0 + jj]};
    }
    H.key += "/tile" + std::to_string(std::max<int>(0, tuned(ctx, precision, H.gmax).gemm_tile));
    HIPCHK(hipMemcpyAsync(D.sys64, H.h64.data(), sizeof(BatchSystem<double>) * J, hipMemcpyHostToDevice, sa));
    HIPCHK(hipMemcpyAsync(D.sys32, H.h32.data(), sizeof(BatchSystem<float>) * J, hipMemcpyHostToDevice, sa));
    HIPCHK(hipMemcpyAsync(D.dX, H.hv.data(), sizeof(BatchVec<double>) * 10 * J, hipMemcpyHostToDevice, sa));
    HIPCHK(hipMemcpyAsync(D.dD, H.hd.data(), sizeof(BatchVec<float>) * 2 * J, hipMemcpyHostToDevice, sa));
    HIPCHK(hipMemcpyAsync(D.jobs, H.hj.data(), sizeof(BatchJobDesc) * J, hipMemcpyHostToDevice, sa));
    HIPCHK(hipEventRecord(ctx->sw_ev[sl][1], sa));
    for (int jj = 0; jj < H.nj; ++jj) {
      const int64_t i = g0 + jj;
      const size_t js = (size_t)sl * J + jj;
      const beat_mesh* m = meshes[i];
      const int64_t n = m->h.nv, nf = m->h.nf;
      double* q64 = as<double>(ctx->g_q64) + 2 * G.qv * js;
      hipLaunchKernelGGL((gather_q_kernel<double>), dim3(blocks_for(nf * nrhs)), dim3(256), 0, sa, q_dev + 2 * q_off[i],
                         nf, nrhs, (const int32_t*)m->face_perm, q64);
      HIPCHK(hipGetLastError());
      if (fp32) {
        float* q32 = as<float>(ctx->g_q32) + 2 * G.qv * js;
        float* A = H.h32[jj].A;
        float* B = as<float>(ctx->g_b32) + 2 * G.vec * js;
        HIPCHK(launch_downcast(q64, q32, 2 * nf * nrhs, sa));
        if (int rc = assemble<float>(ctx, m, ks[i], ers[i], eis[i], q64, q32, nrhs, A, G.lda, B, sa, nullptr, nullptr,
                                     false, false, tuned(ctx, precision, n).asm_tile))
          return rc;
        HIPCHK(lu_append_rhs<float>(A, n, G.lda, B, nrhs, sa));
      } else {
        double* A = H.h64[jj].A;
        double* B = as<double>(ctx->g_b64) + 2 * G.vec * js;
        if (int rc = assemble<double>(ctx, m, ks[i], ers[i], eis[i], q64, q64, nrhs, A, G.lda, B, sa, nullptr, nullptr,
                                      false, false, tuned(ctx, precision, n).asm_tile))
          return rc;
        if (refined) {
          float* A32 = H.h32[jj].A;
          float* B32 = as<float>(ctx->g_b32) + 2 * G.vec * js;
          HIPCHK(launch_downcast(A, A32, 2 * n * G.lda, sa));
          HIPCHK(launch_downcast(B, B32, 2 * n * nrhs, sa));
          HIPCHK(lu_append_rhs<float>(A32, n, G.lda, B32, nrhs, sa));
        } else {
          HIPCHK(lu_append_rhs<double>(A, n, G.lda, B, nrhs, sa));
        }
      }
    }
    HIPCHK(hipEventRecord(ctx->sw_ev[sl][2], sa));
    return BEAT_OK;
  };
  // the factor / solve / refine graph of group H in slot sl, then the scatter (context stream)
  auto launch = [&](int sl, const HostGroup& H) -> int {
    const SlotDesc D = slot_desc(sl);
    const int nj = H.nj;
    HIPCHK(hipStreamWaitEvent(s, ctx->sw_ev[sl][2], 0));
    HIPCHK(hipEventRecord(ctx->sw_ev[sl][0], s));
    hipGraphExec_t exec = nullptr;
    auto it = ctx->g_graphs.find(H.key);
    if (it != ctx->g_graphs.end()) {
      exec = it->second;
    } else {
      hipGraph_t graph = nullptr;
      HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
      hipError_t ce = hipSuccess;
      auto ok = [&](hipError_t e) {
        if (ce == hipSuccess) ce = e;
      };
      const int64_t gmax = H.gmax, gncmax = H.gncmax;
      const int tile = std::max<int>(0, tuned(ctx, precision, gmax).gemm_tile);
      ok(launch_fill<int>(D.lustat, 4 * (int64_t)nj, 0, s));
      const unsigned tb = (unsigned)((nj + 63) / 64);
      const dim3 ge((unsigned)((gmax * nrhs + 255) / 256), (unsigned)nj);
      if (refined) {
        // complex64 speculative LU (three-product updates), first solution
        ok(lu_factor_batched<float>(D.sys32, nj, gmax, gncmax, true, tile, s));
        ok(tri_solve_batched<float>(D.sys32, D.dCar32, nj, nrhs, false, true, s));
        hipLaunchKernelGGL((batch_elementwise_kernel<0>), ge, dim3(256), 0, s, D.dX, D.dCar32, nullptr, nullptr,
                           nullptr, nullptr, nullptr, nrhs);
        // complex128 refinement with the device-side decisions
        hipLaunchKernelGGL(colnorms_batched_kernel, dim3((unsigned)nj), dim3(256), 0, s, D.dB, nrhs, D.norms, 0);
        ok(residual64_batched(D.sys64, D.dX, D.dB, D.dR, nj, gmax, nrhs, s));
        hipLaunchKernelGGL(colnorms_batched_kernel, dim3((unsigned)nj), dim3(256), 0, s, D.dR, nrhs, D.norms, 8);
        hipLaunchKernelGGL(refine_control_kernel, dim3(tb), dim3(64), 0, s, D.state, D.norms, nj, nrhs, rtol, 0, 0);
        for (int step = 1; step <= max_steps; ++step) {
          hipLaunchKernelGGL((batch_elementwise_kernel<1>), ge, dim3(256), 0, s, D.dRs, D.dD, nullptr, nullptr,
                             nullptr, nullptr, D.state, nrhs);
          ok(tri_solve_batched<float>(D.sys32, D.dD, nj, nrhs, true, true, s));
          hipLaunchKernelGGL((batch_elementwise_kernel<2>), ge, dim3(256), 0, s, D.dNXs, D.dD, D.dX, nullptr, nullptr,
                             nullptr, D.state, nrhs);
          ok(residual64_batched(D.sys64, D.dNXs, D.dB, D.dNRs, nj, gmax, nrhs, s));
          hipLaunchKernelGGL(colnorms_batched_kernel, dim3((unsigned)nj), dim3(256), 0, s, D.dNRs, nrhs, D.norms, 8);
          hipLaunchKernelGGL(refine_control_kernel, dim3(tb), dim3(64), 0, s, D.state, D.norms, nj, nrhs, rtol, 1,
                             step);
          hipLaunchKernelGGL((batch_elementwise_kernel<3>), ge, dim3(256), 0, s, D.dX, nullptr, D.dNX, D.dR, D.dNR,
                             nullptr, D.state, nrhs);
          hipLaunchKernelGGL(refine_advance_kernel, dim3(tb), dim3(64), 0, s, D.state, nj);
        }
        hipLaunchKernelGGL(refine_control_kernel, dim3(tb), dim3(64), 0, s, D.state, D.norms, nj, nrhs, rtol, 2, 0);
        hipLaunchKernelGGL((batch_elementwise_kernel<4>), ge, dim3(256), 0, s, D.dOut, nullptr, nullptr, nullptr,
                           nullptr, D.dX, nullptr, nrhs);
      } else if (fp32) {
        ok(lu_factor_batched<float>(D.sys32, nj, gmax, gncmax, false, tile, s));
        ok(tri_solve_batched<float>(D.sys32, D.dCar32, nj, nrhs, false, true, s));
        hipLaunchKernelGGL((batch_elementwise_kernel<0>), ge, dim3(256), 0, s, D.dOut, D.dCar32, nullptr, nullptr,
                           nullptr, nullptr, nullptr, nrhs);
      } else {
        ok(lu_factor_batched<double>(D.sys64, nj, gmax, gncmax, false, tile, s));
        ok(tri_solve_batched<double>(D.sys64, D.dCar64, nj, nrhs, false, true, s));
        hipLaunchKernelGGL((batch_elementwise_kernel<4>), ge, dim3(256), 0, s, D.dOut, nullptr, nullptr, nullptr,
                           nullptr, D.dCar64, nullptr, nrhs);
      }
      ok(hipGetLastError());
      hipError_t ee = hipStreamEndCapture(s, &graph);
      if (ce != hipSuccess || ee != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        return fail(ctx, BEAT_ERR_HIP,
                    std::string("batch graph capture: ") + hipGetErrorString(ce != hipSuccess ? ce : ee));
      }
      hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      if (ie != hipSuccess)
        return fail(ctx, BEAT_ERR_HIP, std::string("batch graph instantiate: ") + hipGetErrorString(ie));
      ctx->g_graphs[H.key] = exec;
    }
    HIPCHK(hipGraphLaunch(exec, s));
    // the pressures to the caller's blocks
    hipLaunchKernelGGL(batch_scatter_kernel, dim3((unsigned)((H.gmax * nrhs + 255) / 256), (unsigned)nj), dim3(256), 0,
                       s, (const BatchVec<double>*)D.dOut, (const BatchJobDesc*)D.jobs, nrhs);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(ctx->sw_ev[sl][3], s));
    HIPCHK(hipEventRecord(ctx->sw_free[sl], s));
    return BEAT_OK;
  };
  int result = BEAT_OK;
  // reports and the per-job re-solve of exchange-flagged systems of group H
  auto finish = [&](int sl, const HostGroup& H) -> int {
    const SlotDesc D = slot_desc(sl);
    const int nj = H.nj;
    HIPCHK(hipEventSynchronize(ctx->sw_ev[sl][3]));
    std::vector<RefineState> hs(J);
    std::vector<int> hl(4 * J);
    HIPCHK(hipMemcpy(hs.data(), D.state, sizeof(RefineState) * nj, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(hl.data(), D.lustat, sizeof(int) * 4 * nj, hipMemcpyDeviceToHost));
    float ta = 0, tf = 0;
    HIPCHK(hipEventElapsedTime(&ta, ctx->sw_ev[sl][1], ctx->sw_ev[sl][2]));
    HIPCHK(hipEventElapsedTime(&tf, ctx->sw_ev[sl][0], ctx->sw_ev[sl][3]));
    for (int jj = 0; jj < nj; ++jj) {
      const int64_t i = H.g0 + jj;
      const bool exchange = hl[4 * jj + 3] != 0 || ctx->debug_force_exchange;
      if (exchange) {
        // partial pivoting would exchange rows here: this job again on its own
        const beat_mesh* m = meshes[i];
        SystemBufs sb{&ctx->q64, &ctx->q32, &ctx->a64, &ctx->b64, &ctx->a32, &ctx->b32, &ctx->af};
        const AsmEvents ev{ctx->ev[0], ctx->ev[4], ctx->ev[5], ctx->ev[6]};
        if (int rc = stage_assemble(ctx, m, precision, ks[i], ers[i], eis[i], q_dev + 2 * q_off[i], nrhs, sb, s, ev))
          return rc;
        int rc = stage_solve(ctx, m, precision, nrhs, sb, p_dev + 2 * p_off[i], opts, ev, reps ? reps + i : nullptr);
        if (rc == BEAT_ERR_SINGULAR) result = rc;
        else if (rc) return rc;
        if (reps) reps[i].factorization_count = 2;
        continue;
      }
      if (hl[4 * jj] != 0) {
        result = BEAT_ERR_SINGULAR;
        fail(ctx, BEAT_ERR_SINGULAR, "LU found an exactly zero pivot at column " + std::to_string(hl[4 * jj]));
      }
      if (!reps) continue;
      beat_solve_report& R = reps[i];
      const beat_mesh* m = meshes[i];
      R.assembly_s = ta * 1e-3 / nj;  // the group's assembly, shared evenly
      R.regular_kernel_s = 0;
      R.factor_s = tf * 1e-3 / nj;  // the group's graph and scatter, shared evenly
      R.solve_s = 0;
      R.factorization_count = 1;
      R.lu_panels = (int)((m->h.nv + 63) / 64);
      R.regular_pairs = m->regular_pairs;
      R.singular_pairs = (int64_t)m->h.spairs.size();
      if (refined) {
        const RefineState& S = hs[jj];
        R.refine.steps = S.steps;
        R.refine.status = S.status;
        R.refine.working_status = S.working;
        R.refine.history_len = S.hist_len;
        R.refine.rtol = rtol;
        if (R.refine.working_residuals)
          for (int d = 0; d < nrhs; ++d) R.refine.working_residuals[d] = S.err[d];
        if (R.refine.returned_residuals)
          for (int d = 0; d < nrhs; ++d) R.refine.returned_residuals[d] = S.err[d];
        if (R.refine.history)
          for (int e = 0; e < S.hist_len; ++e)
            for (int d = 0; d < nrhs; ++d) R.refine.history[e * nrhs + d] = S.hist[e * 8 + d];
      } else {
        R.refine.steps = 0;
        R.refine.status = BEAT_REFINE_NONE;
        R.refine.working_status = BEAT_REFINE_NONE;
        R.refine.history_len = 0;
        R.refine.rtol = rtol;
      }
    }
    return BEAT_OK;
  };
  const int64_t ngroups = (njobs + G.J - 1) / G.J;
  HostGroup hg[2];
  // the assembly stream starts after everything earlier on the context stream
  HIPCHK(hipEventRecord(ctx->sw_free[1], s));
  HIPCHK(hipStreamWaitEvent(sa, ctx->sw_free[1], 0));
  if (int rc = prepare(0, 0, hg[0])) return rc;
  for (int64_t g = 0; g < ngroups; ++g) {
    const int sl = (int)(g & 1);
    if (g + 1 < ngroups) {
      // slot 1 - sl is free once group g - 1's graph and scatter are done
      if (g >= 1) HIPCHK(hipStreamWaitEvent(sa, ctx->sw_free[1 - sl], 0));
      if (int rc = prepare((g + 1) * G.J, 1 - sl, hg[1 - sl])) {
        (void)hipStreamSynchronize(sa);
        return rc;
      }
    }
    if (int rc = launch(sl, hg[sl])) {
      (void)hipStreamSynchronize(sa);
      return rc;
    }
    if (int rc = finish(sl, hg[sl])) {
      (void)hipStreamSynchronize(sa);
      return rc;
    }
  }
  HIPCHK(hipStreamSynchronize(sa));
  HIPCHK(hipStreamSynchronize(s));
  return result;
}

// A frequency sweep on one mesh: consecutive q and p blocks.
int sweep_impl(beat_ctx* ctx, const beat_mesh* m, int precision, int64_t nfreq, const double* ks, const double* ers,
               const double* eis, const double* q_dev, int nrhs, double* p_dev, const beat_solve_options* opts,
               beat_solve_report* reps) {
  std::vector<const beat_mesh*> ms(std::max<int64_t>(nfreq, 0), m);
  std::vector<int64_t> qo(ms.size()), po(ms.size());
  for (size_t i = 0; i < ms.size(); ++i) {
    qo[i] = (int64_t)i * m->h.nf * nrhs;
    po[i] = (int64_t)i * m->h.nv * nrhs;
  }
  return batch_impl(ctx, nfreq, ms.data(), precision, ks, ers, eis, q_dev, qo.data(), nrhs, p_dev, po.data(), opts,
                    reps);
}

template <typename T>
hipError_t upload(T** dst, const std::vector<T>& v) {
  if (v.empty()) {
    *dst = nullptr;
    return hipSuccess;
  }
  hipError_t e = hipMalloc((void**)dst, sizeof(T) * v.size());
  if (e != hipSuccess) return e;
  return hipMemcpy(*dst, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice);
}


// Near-panel correction of the gradient (correct_field_gradient!,
// BeatEngineFieldGradient.jl:45-73): list the (point, face image) pairs within 3
// face diameters (count, then fill), then integrate each pair adaptively.
template <typename T>
int near_panel_correction(beat_ctx* ctx, const beat_mesh* m, const T* pts, int64_t npts, const T* p, const T* q,
                          int nrhs, double k, T* out) {
  hipStream_t s = ctx->stream;
  HIPCHK(ensure(ctx->grad, 256));
  unsigned long long* count = as<unsigned long long>(ctx->grad);
  int* failed = reinterpret_cast<int*>(count + 1);
  HIPCHK(hipMemsetAsync(ctx->grad.p, 0, 16, s));
  HIPCHK(gradient_near_pairs<T>(m->verts, m->fverts, m->h.nf, m->h.nimg, m->h.img, pts, npts, count, nullptr, 0, false,
                                s));
  unsigned long long npairs = 0;
  HIPCHK(hipMemcpyAsync(&npairs, count, sizeof(npairs), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  if (npairs == 0) return BEAT_OK;
  if (npairs > (1ull << 31)) return fail(ctx, BEAT_ERR_ARG, "too many near-panel gradient pairs");
  HIPCHK(ensure(ctx->grad, 256 + sizeof(int2) * npairs));
  count = as<unsigned long long>(ctx->grad);
  failed = reinterpret_cast<int*>(count + 1);
  int2* list = reinterpret_cast<int2*>(reinterpret_cast<char*>(ctx->grad.p) + 256);
  HIPCHK(hipMemsetAsync(ctx->grad.p, 0, 16, s));
  HIPCHK(gradient_near_pairs<T>(m->verts, m->fverts, m->h.nf, m->h.nimg, m->h.img, pts, npts, count, list, npairs, true,
                                s));
  HIPCHK(gradient_correction<T>(m->verts, m->fverts, m->rec64, m->h.nimg, m->h.img, list, (int64_t)npairs, pts, p, q,
                                nrhs, k, m->rule4, out, failed, s));
  int nfail = 0;
  HIPCHK(hipMemcpyAsync(&nfail, failed, sizeof(int), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  if (nfail > 0)
    return fail(ctx, BEAT_ERR_ARG, "BEM gradient sample is too close to a boundary for reliable evaluation.");
  return BEAT_OK;
}

// Field (gradient false: npts x nrhs) or value and gradient (gradient true:
// npts x 4 x nrhs) at the caller's points, column-major complex128.
int field_impl(beat_ctx* ctx, const beat_mesh* m, int precision, double k, const double* pressure, const double* q,
               int64_t nrhs, const double* pts, int64_t npts, double* out, bool gradient) {
  if (int rc = check_mesh(ctx, m)) return rc;
  if (nrhs < 1 || npts < 0 || (npts > 0 && (!pressure || !q || !pts || !out))) return fail(ctx, BEAT_ERR_ARG, "bad arguments");
  if (precision != BEAT_PREC_FP64 && precision != BEAT_PREC_FP32) return fail(ctx, BEAT_ERR_ARG, "precision");
  if (npts == 0) return BEAT_OK;
  std::lock_guard<std::mutex> lock(ctx->mu);
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  const int64_t n = m->h.nv, nf = m->h.nf, ns = nf * m->h.nq * (1 + m->h.nimg);
  // stage: pressure (n x nrhs, caller col-major), q (nf x nrhs) -> internal row-major
  HIPCHK(ensure(ctx->qin, sizeof(double) * 2 * (nf + n) * nrhs));
  double* qin = as<double>(ctx->qin);
  double* pin = qin + 2 * nf * nrhs;
  HIPCHK(hipMemcpyAsync(qin, q, sizeof(double) * 2 * nf * nrhs, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(pin, pressure, sizeof(double) * 2 * n * nrhs, hipMemcpyHostToDevice, s));
  HIPCHK(ensure(ctx->q64, sizeof(double) * 2 * (nf + n) * nrhs));
  double* qint = as<double>(ctx->q64);
  double* pint = qint + 2 * nf * nrhs;
  hipLaunchKernelGGL((gather_q_kernel<double>), dim3(blocks_for(nf * nrhs)), dim3(256), 0, s, (const double*)qin, nf,
                     (int)nrhs, (const int32_t*)m->face_perm, qint);
  // pressure: internal row v takes caller row vert_perm[v]; gather with the vertex permutation
  std::vector<int32_t> vp(m->h.vert_perm);
  HIPCHK(ensure(ctx->lusmall, sizeof(int32_t) * n + 256));
  HIPCHK(hipMemcpyAsync(ctx->lusmall.p, vp.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL((gather_q_kernel<double>), dim3(blocks_for(n * nrhs)), dim3(256), 0, s, (const double*)pin, n,
                     (int)nrhs, (const int32_t*)as<int32_t>(ctx->lusmall), pint);
  HIPCHK(hipGetLastError());
  const size_t es = precision == BEAT_PREC_FP64 ? sizeof(double) : sizeof(float);
  HIPCHK(ensure(ctx->fld_src, es * (ns * 6 + 4 * ns * nrhs + 2 * (nf + n) * nrhs)));
  HIPCHK(ensure(ctx->fld_pts, sizeof(double) * 3 * npts));
  const int ncomp = gradient ? 4 : 1;
  HIPCHK(ensure(ctx->fld_out, sizeof(double) * 2 * npts * nrhs * ncomp));
  HIPCHK(hipMemcpyAsync(ctx->fld_pts.p, pts, sizeof(double) * 3 * npts, hipMemcpyHostToDevice, s));
  std::vector<double> res((size_t)2 * npts * nrhs * ncomp);
  if (precision == BEAT_PREC_FP64) {
    double* sp = as<double>(ctx->fld_src);
    double* sn = sp + 3 * ns;
    double* pw = sn + 3 * ns;
    double* qw = pw + 2 * ns * nrhs;
    HIPCHK(field_sources<double>(m->rec64, m->fverts, nf, m->h.nq, m->rule3, pint, qint, (int)nrhs, m->h.nimg, m->h.img,
                                 sp, sn, pw, qw, s));
    if (gradient) {
      HIPCHK(field_gradient_eval<double>(sp, sn, pw, qw, ns, (int)nrhs, as<double>(ctx->fld_pts), npts,
                                         as<double>(ctx->fld_out), k, s));
      if (int rc = near_panel_correction<double>(ctx, m, as<double>(ctx->fld_pts), npts, pint, qint, (int)nrhs, k,
                                                 as<double>(ctx->fld_out)))
        return rc;
    } else {
      HIPCHK(field_eval<double>(sp, sn, pw, qw, ns, (int)nrhs, as<double>(ctx->fld_pts), npts,
                                as<double>(ctx->fld_out), k, s));
    }
    HIPCHK(hipMemcpyAsync(res.data(), ctx->fld_out.p, sizeof(double) * res.size(), hipMemcpyDeviceToHost, s));
  } else {
    float* sp = as<float>(ctx->fld_src);
    float* sn = sp + 3 * ns;
    float* pw = sn + 3 * ns;
    float* qw = pw + 2 * ns * nrhs;
    float* p32 = qw + 2 * ns * nrhs;
    float* q32 = p32 + 2 * n * nrhs;
    HIPCHK(launch_downcast(pint, p32, 2 * n * nrhs, s));
    HIPCHK(launch_downcast(qint, q32, 2 * nf * nrhs, s));
    // points in float
    float* pts32 = reinterpret_cast<float*>(as<double>(ctx->fld_pts) + 0);
    std::vector<float> hp(3 * npts);
    for (int64_t i = 0; i < 3 * npts; ++i) hp[i] = (float)pts[i];
    HIPCHK(hipMemcpyAsync(pts32, hp.data(), sizeof(float) * hp.size(), hipMemcpyHostToDevice, s));
    float* o32 = as<float>(ctx->fld_out);
    HIPCHK(field_sources<float>(m->rec64, m->fverts, nf, m->h.nq, m->rule3, p32, q32, (int)nrhs, m->h.nimg, m->h.img,
                                sp, sn, pw, qw, s));
    if (gradient) {
      HIPCHK(field_gradient_eval<float>(sp, sn, pw, qw, ns, (int)nrhs, pts32, npts, o32, (float)k, s));
      if (int rc = near_panel_correction<float>(ctx, m, pts32, npts, p32, q32, (int)nrhs, k, o32)) return rc;
    } else {
      HIPCHK(field_eval<float>(sp, sn, pw, qw, ns, (int)nrhs, pts32, npts, o32, (float)k, s));
    }
    std::vector<float> r32(res.size());
    HIPCHK(hipMemcpyAsync(r32.data(), o32, sizeof(float) * r32.size(), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    std::copy(r32.begin(), r32.end(), res.begin());
  }
  HIPCHK(hipStreamSynchronize(s));
  // internal (point, rhs, component) -> caller (point, component, rhs), column-major
  for (int64_t d = 0; d < nrhs; ++d)
    for (int c = 0; c < ncomp; ++c)
      for (int64_t i = 0; i < npts; ++i) {
        const size_t src = 2 * ((size_t)(i * nrhs + d) * ncomp + c);
        const size_t dst = 2 * (size_t)(i + npts * (c + (int64_t)ncomp * d));
        out[dst] = res[src];
        out[dst + 1] = res[src + 1];
      }
  return BEAT_OK;
}

// Bytes of the no-allocation scratch a field evaluation needs, in the order the
// buffers are laid out (each rounded to 256 bytes).
size_t noalloc_field_bytes(int64_t nv, int64_t nf, int64_t ns, int64_t nrhs, int64_t npts, bool fp32) {
  auto r = [](size_t b) { return (b + 255) / 256 * 256; };
  const size_t e = fp32 ? sizeof(float) : sizeof(double);
  size_t b = r(sizeof(double) * 2 * nf * nrhs) + r(sizeof(double) * 2 * nv * nrhs);  // q, p internal (complex128)
  if (fp32) b += r(e * 2 * nf * nrhs) + r(e * 2 * nv * nrhs) + r(e * 3 * npts);     // complex64 copies, points
  b += r(e * 3 * ns) + r(e * 3 * ns) + r(e * 2 * ns * nrhs) + r(e * 2 * ns * nrhs);  // sources
  b += r(e * 2 * npts * nrhs);                                                       // accumulated field
  return b;
}

// beat_field_eval on device pointers with every buffer in the scratch array.
int field_noalloc_impl(beat_ctx* ctx, const beat_mesh* m, int precision, double k, const double* p_dev,
                       const double* q_dev, int64_t nrhs, const double* pts_dev, int64_t npts, double* out_dev) {
  if (int rc = check_mesh(ctx, m)) return rc;
  if (nrhs < 1 || npts < 0 || (npts > 0 && (!p_dev || !q_dev || !pts_dev || !out_dev)))
    return fail(ctx, BEAT_ERR_ARG, "bad arguments");
  if (precision != BEAT_PREC_FP64 && precision != BEAT_PREC_FP32) return fail(ctx, BEAT_ERR_ARG, "precision");
  if (npts == 0) return BEAT_OK;
  const int64_t n = m->h.nv, nf = m->h.nf, ns = nf * m->h.nq * (1 + m->h.nimg);
  const bool fp32 = precision == BEAT_PREC_FP32;
  const size_t need = noalloc_field_bytes(n, nf, ns, nrhs, npts, fp32);
  if (need > NOALLOC_SCRATCH_BYTES)
    return fail(ctx, BEAT_ERR_ARG,
                "beat_field_eval_device_noalloc needs " + std::to_string(need) + " bytes of scratch for this mesh, " +
                    std::to_string(nrhs) + " right-hand sides and " + std::to_string(npts) + " points; it holds " +
                    std::to_string(NOALLOC_SCRATCH_BYTES) + " (use beat_field_eval, or fewer points per call)");
  std::lock_guard<std::mutex> lock(ctx->mu);
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  unsigned char* base = ctx->scratch;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    unsigned char* p = base + off;
    off += (bytes + 255) / 256 * 256;
    return p;
  };
  double* qint = reinterpret_cast<double*>(take(sizeof(double) * 2 * nf * nrhs));
  double* pint = reinterpret_cast<double*>(take(sizeof(double) * 2 * n * nrhs));
  hipLaunchKernelGGL((gather_q_kernel<double>), dim3(blocks_for(nf * nrhs)), dim3(256), 0, s, q_dev, nf, (int)nrhs,
                     (const int32_t*)m->face_perm, qint);
  hipLaunchKernelGGL(p_to_internal_kernel, dim3(blocks_for(n * nrhs)), dim3(256), 0, s, p_dev, n, (int)nrhs,
                     (const int32_t*)m->vert_inv, pint);
  HIPCHK(hipGetLastError());
  if (!fp32) {
    double* sp = reinterpret_cast<double*>(take(sizeof(double) * 3 * ns));
    double* sn = reinterpret_cast<double*>(take(sizeof(double) * 3 * ns));
    double* pw = reinterpret_cast<double*>(take(sizeof(double) * 2 * ns * nrhs));
    double* qw = reinterpret_cast<double*>(take(sizeof(double) * 2 * ns * nrhs));
    double* o = reinterpret_cast<double*>(take(sizeof(double) * 2 * npts * nrhs));
    HIPCHK(field_sources<double>(m->rec64, m->fverts, nf, m->h.nq, m->rule3, pint, qint, (int)nrhs, m->h.nimg, m->h.img,
                                 sp, sn, pw, qw, s));
    // field_eval clears o (a fill kernel) before accumulating into it
    HIPCHK(field_eval<double>(sp, sn, pw, qw, ns, (int)nrhs, pts_dev, npts, o, k, s));
    hipLaunchKernelGGL((cols_out_kernel<double, double>), dim3(blocks_for(npts * nrhs)), dim3(256), 0, s,
                       (const double*)o, npts, (int)nrhs, out_dev);
  } else {
    float* qq = reinterpret_cast<float*>(take(sizeof(float) * 2 * nf * nrhs));
    float* pp = reinterpret_cast<float*>(take(sizeof(float) * 2 * n * nrhs));
    float* pts32 = reinterpret_cast<float*>(take(sizeof(float) * 3 * npts));
    HIPCHK(launch_downcast(pint, pp, 2 * n * nrhs, s));
    HIPCHK(launch_downcast(qint, qq, 2 * nf * nrhs, s));
    HIPCHK(launch_downcast(pts_dev, pts32, 3 * npts, s));
    float* sp = reinterpret_cast<float*>(take(sizeof(float) * 3 * ns));
    float* sn = reinterpret_cast<float*>(take(sizeof(float) * 3 * ns));
    float* pw = reinterpret_cast<float*>(take(sizeof(float) * 2 * ns * nrhs));
    float* qw = reinterpret_cast<float*>(take(sizeof(float) * 2 * ns * nrhs));
    float* o = reinterpret_cast<float*>(take(sizeof(float) * 2 * npts * nrhs));
    HIPCHK(field_sources<float>(m->rec64, m->fverts, nf, m->h.nq, m->rule3, pp, qq, (int)nrhs, m->h.nimg, m->h.img, sp,
                                sn, pw, qw, s));
    HIPCHK(field_eval<float>(sp, sn, pw, qw, ns, (int)nrhs, pts32, npts, o, (float)k, s));
    hipLaunchKernelGGL((cols_out_kernel<float, double>), dim3(blocks_for(npts * nrhs)), dim3(256), 0, s,
                       (const float*)o, npts, (int)nrhs, out_dev);
  }
  HIPCHK(hipGetLastError());
  if (off > NOALLOC_SCRATCH_BYTES) return fail(ctx, BEAT_ERR_HIP, "no-allocation scratch layout overflow");
  HIPCHK(hipStreamSynchronize(s));
  return BEAT_OK;
}

}  // namespace

extern "C" {

// The LU's look-ahead stream carries the far trailing updates, the factorization's
// largest kernels; without a limit they fill every CU and the next leaf's panel
// kernels on the main stream wait for a CU until they drain. The side stream
// therefore runs on a CU mask that keeps r/8 of the CUs free (BEAT_LU_SIDE_RESERVE
// = r, default 1; 0: every CU): bit i is cleared when ((i & 7) - ((i >> 5) & 7)) & 7
// < r, which keeps 4r CUs of every XCD free whether the mask's bits run XCD by
// XCD or interleave over the XCDs.
static hipError_t create_side_stream(beat_ctx* c) {
  int r = 1;
  if (const char* v = std::getenv("BEAT_LU_SIDE_RESERVE")) r = std::max(0, std::min(7, std::atoi(v)));
  int ncu = 0;
  if (r > 0 && hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c->device) == hipSuccess &&
      ncu >= 64) {
    std::vector<uint32_t> mask((size_t)(ncu + 31) / 32, 0u);
    for (int i = 0; i < ncu; ++i)
      if ((((i & 7) - ((i >> 5) & 7)) & 7) >= r) mask[(size_t)i / 32] |= 1u << (i % 32);
    c->side_reserve = r;
    return hipExtStreamCreateWithCUMask(&c->side, (uint32_t)mask.size(), mask.data());
  }
  c->side_reserve = 0;
  return hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking);
}

int beat_ctx_create(int device, beat_ctx** out) {
  if (!out) return BEAT_ERR_ARG;
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) return BEAT_ERR_ARG;
  if (hipSetDevice(device) != hipSuccess) return BEAT_ERR_HIP;
  beat_ctx* c = new beat_ctx();
  c->device = device;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return BEAT_ERR_HIP;
  }
  for (auto& e : c->ev)
    if (hipEventCreate(&e) != hipSuccess) {
      delete c;
      return BEAT_ERR_HIP;
    }
  for (auto& row : c->sw_ev)
    for (auto& e : row)
      if (hipEventCreate(&e) != hipSuccess) {
        delete c;
        return BEAT_ERR_HIP;
      }
  for (auto& e : c->sw_free)
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
      delete c;
      return BEAT_ERR_HIP;
    }
  if (hipStreamCreateWithFlags(&c->asm_stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return BEAT_ERR_HIP;
  }
  if (create_side_stream(c) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_fact, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_far, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_far2, hipEventDisableTiming) != hipSuccess ||
      hipStreamCreateWithFlags(&c->inb, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_panel, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_inb, hipEventDisableTiming) != hipSuccess) {
    delete c;
    return BEAT_ERR_HIP;
  }
  {
    double t[2 * TRIG_N];
    fill_trig_table(t);
    if (hipMalloc((void**)&c->trig, sizeof(t)) != hipSuccess ||
        hipMemcpy(c->trig, t, sizeof(t), hipMemcpyHostToDevice) != hipSuccess) {
      delete c;
      return BEAT_ERR_HIP;
    }
  }
  // the no-allocation entry point's scratch on this device (the device is current)
  if (hipGetSymbolAddress(reinterpret_cast<void**>(&c->scratch), HIP_SYMBOL(scratch)) != hipSuccess || !c->scratch) {
    delete c;
    return BEAT_ERR_HIP;
  }
  if (const char* v = std::getenv("BEAT_SWEEP_ASM_ONE_PER_CU")) c->sweep_asm_one_per_cu = v[0] != '0';
  if (const char* v = std::getenv("BEAT_REGULAR_SYMMETRIC")) c->regular_symmetric = v[0] != '0';
  if (const char* v = std::getenv("BEAT_FP32_ASSEMBLY")) c->fp32_assembly_fp64 = std::strcmp(v, "native") != 0;
  if (const char* v = std::getenv("BEAT_BATCH_WORKERS")) c->batch_workers = std::max(1, std::atoi(v));
  if (const char* v = std::getenv("BEAT_BATCH_GROUPED")) c->batch_grouped = v[0] != '0';
  if (const char* v = std::getenv("BEAT_LU_TAIL_ROWS")) c->lu_tail_rows = std::max(0, std::atoi(v));
  // BEAT_LU_LOOKAHEAD=0 runs the whole factorization in order on one stream
  if (const char* v = std::getenv("BEAT_LU_LOOKAHEAD"))
    if (v[0] == '0') {
      (void)hipStreamDestroy(c->side);
      (void)hipStreamDestroy(c->inb);
      c->side = nullptr;
      c->inb = nullptr;
    }
  *out = c;
  return BEAT_OK;
}

void beat_ctx_destroy(beat_ctx* c) {
  if (!c) return;
  for (beat_ctx* a : c->aux) beat_ctx_destroy(a);
  c->aux.clear();
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  for (auto& g : c->g_graphs) (void)hipGraphExecDestroy(g.second);
  c->g_graphs.clear();
  DevBuf* bufs[] = {&c->a64, &c->a32, &c->b64, &c->b32, &c->q64, &c->q32, &c->x64, &c->r64, &c->n64,
                    &c->rn64, &c->xn64, &c->d32, &c->work, &c->chunk, &c->norms, &c->ipiv, &c->lusmall, &c->qin,
                    &c->pout, &c->fld_src, &c->fld_out, &c->fld_pts, &c->grad, &c->adj, &c->af,
                    &c->g_a64, &c->g_a32, &c->g_b64, &c->g_b32, &c->g_q64, &c->g_q32, &c->g_linv, &c->g_vec,
                    &c->g_state, &c->g_desc, &c->rpart};
  for (DevBuf* b : bufs)
    if (b->p) (void)hipFree(b->p);
  if (c->asm_stream) (void)hipStreamSynchronize(c->asm_stream);
  for (int s2 = 0; s2 < 2; ++s2) {
    DevBuf* sw[] = {&c->sw_q64[s2], &c->sw_q32[s2], &c->sw_a64[s2], &c->sw_b64[s2], &c->sw_a32[s2], &c->sw_b32[s2], &c->sw_af[s2]};
    for (DevBuf* b : sw)
      if (b->p) (void)hipFree(b->p);
    for (auto& e : c->sw_ev[s2])
      if (e) (void)hipEventDestroy(e);
    if (c->sw_free[s2]) (void)hipEventDestroy(c->sw_free[s2]);
  }
  if (c->asm_stream) (void)hipStreamDestroy(c->asm_stream);
  for (auto& e : c->ev)
    if (e) (void)hipEventDestroy(e);
  if (c->side) (void)hipStreamDestroy(c->side);
  if (c->ev_fact) (void)hipEventDestroy(c->ev_fact);
  if (c->ev_far) (void)hipEventDestroy(c->ev_far);
  if (c->ev_far2) (void)hipEventDestroy(c->ev_far2);
  if (c->inb) (void)hipStreamDestroy(c->inb);
  if (c->ev_panel) (void)hipEventDestroy(c->ev_panel);
  if (c->ev_inb) (void)hipEventDestroy(c->ev_inb);
  if (c->trig) (void)hipFree(c->trig);
  (void)hipStreamDestroy(c->stream);
  delete c;
}

const char* beat_last_error(const beat_ctx* c) { return c ? c->err.c_str() : "null context"; }

int beat_ctx_set_cancel_path(beat_ctx* ctx, const char* path) {
  if (!ctx) return BEAT_ERR_ARG;
  std::lock_guard<std::mutex> lock(ctx->mu);
  ctx->cancel_path = path ? path : "";
  return BEAT_OK;
}

int beat_ctx_debug_abort_waits(beat_ctx* ctx, int on) {
  if (!ctx) return BEAT_ERR_ARG;
  std::lock_guard<std::mutex> lock(ctx->mu);
  ctx->debug_abort_waits = on ? 1 : 0;
  return BEAT_OK;
}

int beat_ctx_debug_force_exchange(beat_ctx* ctx, int on) {
  if (!ctx) return BEAT_ERR_ARG;
  std::lock_guard<std::mutex> lock(ctx->mu);
  ctx->debug_force_exchange = on != 0;
  return BEAT_OK;
}

int beat_ctx_regular_symmetric(beat_ctx* ctx, int on) {
  if (!ctx) return BEAT_ERR_ARG;
  std::lock_guard<std::mutex> lock(ctx->mu);
  ctx->regular_symmetric = on != 0;
  return BEAT_OK;
}

int beat_ctx_reclaim(beat_ctx* c) {
  if (!c) return BEAT_ERR_ARG;
  for (beat_ctx* a : c->aux)
    if (int rc = beat_ctx_reclaim(a)) return rc;
  std::lock_guard<std::mutex> lock(c->mu);
  beat_ctx* ctx = c;
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipStreamSynchronize(c->asm_stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (c->side) HIPCHK(hipStreamSynchronize(c->side));
  if (c->inb) HIPCHK(hipStreamSynchronize(c->inb));
  for (auto& g : c->g_graphs) HIPCHK(hipGraphExecDestroy(g.second));
  c->g_graphs.clear();
  DevBuf* bufs[] = {&c->a64, &c->a32, &c->b64, &c->b32, &c->q64, &c->q32, &c->x64, &c->r64, &c->n64, &c->rn64,
                    &c->xn64, &c->d32, &c->work, &c->chunk, &c->norms, &c->ipiv, &c->lusmall, &c->qin, &c->pout,
                    &c->fld_src, &c->fld_out, &c->fld_pts, &c->grad, &c->adj, &c->af,
                    &c->g_a64, &c->g_a32, &c->g_b64, &c->g_b32, &c->g_q64, &c->g_q32, &c->g_linv, &c->g_vec,
                    &c->g_state, &c->g_desc, &c->rpart};
  for (DevBuf* b : bufs) {
    if (b->p) HIPCHK(hipFree(b->p));
    b->p = nullptr;
    b->bytes = 0;
  }
  for (int s2 = 0; s2 < 2; ++s2) {
    DevBuf* sw[] = {&c->sw_q64[s2], &c->sw_q32[s2], &c->sw_a64[s2], &c->sw_b64[s2], &c->sw_a32[s2], &c->sw_b32[s2], &c->sw_af[s2]};
    for (DevBuf* b : sw) {
      if (b->p) HIPCHK(hipFree(b->p));
      b->p = nullptr;
      b->bytes = 0;
    }
  }
  return BEAT_OK;
}

int beat_ctx_memory_info(beat_ctx* c, int64_t* info) {
  if (!c || !info) return BEAT_ERR_ARG;
  std::lock_guard<std::mutex> lock(c->mu);
  beat_ctx* ctx = c;
  HIPCHK(hipSetDevice(c->device));
  size_t used = 0;
  const DevBuf* bufs[] = {&c->a64, &c->a32, &c->b64, &c->b32, &c->q64, &c->q32, &c->x64, &c->r64, &c->n64, &c->rn64,
                          &c->xn64, &c->d32, &c->work, &c->chunk, &c->norms, &c->ipiv, &c->lusmall, &c->qin,
                          &c->pout, &c->fld_src, &c->fld_out, &c->fld_pts, &c->grad, &c->adj, &c->af,
                    &c->g_a64, &c->g_a32, &c->g_b64, &c->g_b32, &c->g_q64, &c->g_q32, &c->g_linv, &c->g_vec,
                    &c->g_state, &c->g_desc, &c->rpart};
  for (const DevBuf* b : bufs) used += b->bytes;
  for (int s2 = 0; s2 < 2; ++s2) {
    const DevBuf* sw[] = {&c->sw_q64[s2], &c->sw_q32[s2], &c->sw_a64[s2], &c->sw_b64[s2], &c->sw_a32[s2],
                          &c->sw_b32[s2], &c->sw_af[s2]};
    for (const DevBuf* b : sw) used += b->bytes;
  }
  size_t free_b = 0, total_b = 0;
  HIPCHK(hipMemGetInfo(&free_b, &total_b));
  info[0] = (int64_t)used;
  info[1] = (int64_t)free_b;
  info[2] = (int64_t)total_b;
  return BEAT_OK;
}

int64_t beat_capabilities(const beat_ctx* ctx, char* buf, int64_t len) {
  std::string arch = "gfx950";
  if (ctx) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, ctx->device) == hipSuccess) arch = prop.gcnArchName;
  }
  const std::string j =
      std::string("{\"library\":\"beat_bem_hip\",\"abi_version\":1,\"arch\":\"") + arch +
      "\",\"precisions\":[\"float64\",\"float32_refined\",\"float32\"]"
      ",\"burton_miller_assembly\":[\"direct_system\",\"operator_matrices\"]"
      ",\"symmetry_modes\":[\"off\",\"x\",\"xy\",\"ground\"]"
      ",\"field_quantities\":[\"pressure\",\"pressure_gradient_v1\"]"
      ",\"phasor_conventions\":[\"exp(-i omega t)\",\"exp(+i omega t)\"]"
      ",\"cancellation\":\"marker_file\",\"linear_solver\":\"hip_mfma_lu\"}";
  if (buf && len > 0) {
    const size_t n = std::min<size_t>((size_t)len - 1, j.size());
    std::memcpy(buf, j.data(), n);
    buf[n] = 0;
  }
  return (int64_t)j.size() + 1;
}

int beat_mesh_upload_symmetric(beat_ctx* ctx, const double* verts, int64_t nv, const int32_t* faces, int64_t nf,
                               int quad_order, int singular_order, int symmetry, beat_mesh** out) {
  if (!ctx || !out || !verts || !faces) return fail(ctx, BEAT_ERR_ARG, "null argument");
  std::lock_guard<std::mutex> lock(ctx->mu);
  *out = nullptr;
  HIPCHK(hipSetDevice(ctx->device));
  beat_mesh* m = new beat_mesh();
  m->ctx = ctx;
  std::string msg = build_host_mesh(verts, nv, faces, nf, quad_order, singular_order, symmetry, m->h);
  if (!msg.empty()) {
    delete m;
    return fail(ctx, BEAT_ERR_ARG, msg);
  }
  std::vector<float> rec32(m->h.rec.begin(), m->h.rec.end());
  std::vector<double> rule3;
  for (int q = 0; q < m->h.nq; ++q) rule3.insert(rule3.end(), {m->h.rule_pts[2 * q], m->h.rule_pts[2 * q + 1], m->h.rule_w[q]});
  hipError_t e = hipSuccess;
  if (e == hipSuccess) e = upload(&m->rec64, m->h.rec);
  if (e == hipSuccess) e = upload(&m->rec32, rec32);
  if (e == hipSuccess) e = upload(&m->verts, m->h.verts);
  if (e == hipSuccess) e = upload(&m->fverts, m->h.fverts);
  if (e == hipSuccess) e = upload(&m->floc, m->h.floc);
  if (e == hipSuccess) e = upload(&m->tile_start, m->h.tile_start);
  if (e == hipSuccess) e = upload(&m->tile_count, m->h.tile_count);
  if (e == hipSuccess) e = upload(&m->tile_nv, m->h.tile_nv);
  if (e == hipSuccess) e = upload(&m->tile_verts, m->h.tile_verts);
  if (e == hipSuccess) e = upload(&m->stile_start, m->h.stile_start);
  if (e == hipSuccess) e = upload(&m->stile_count, m->h.stile_count);
  if (e == hipSuccess) e = upload(&m->stile_nv, m->h.stile_nv);
  if (e == hipSuccess) e = upload(&m->stile_verts, m->h.stile_verts);
  if (e == hipSuccess) e = upload(&m->flocs, m->h.flocs);
  if (e == hipSuccess) e = upload(&m->ntile_start, m->h.ntile_start);
  if (e == hipSuccess) e = upload(&m->ntile_count, m->h.ntile_count);
  if (e == hipSuccess) e = upload(&m->ntile_nv, m->h.ntile_nv);
  if (e == hipSuccess) e = upload(&m->ntile_verts, m->h.ntile_verts);
  if (e == hipSuccess) e = upload(&m->flocn, m->h.flocn);
  if (e == hipSuccess) e = upload(&m->qtile_start, m->h.qtile_start);
  if (e == hipSuccess) e = upload(&m->qtile_count, m->h.qtile_count);
  if (e == hipSuccess) e = upload(&m->qtile_nv, m->h.qtile_nv);
  if (e == hipSuccess) e = upload(&m->qtile_verts, m->h.qtile_verts);
  if (e == hipSuccess) e = upload(&m->flocq, m->h.flocq);
  if (e == hipSuccess) e = upload(&m->tile_qfirst, m->h.tile_qfirst);
  if (e == hipSuccess) e = upload(&m->sym_unit_start, m->h.sym_unit_start);
  if (e == hipSuccess) e = upload(&m->frot, m->h.frot);
  if (e == hipSuccess) e = upload(&m->spairs, m->h.spairs);
  if (e == hipSuccess) e = upload(&m->rule_off, m->h.rule_off);
  if (e == hipSuccess) e = upload(&m->rule_cnt, m->h.rule_cnt);
  if (e == hipSuccess) e = upload(&m->rule_data, m->h.rule_data);
  if (e == hipSuccess) e = upload(&m->rule3, rule3);
  {
    double rp[12], rw[6];
    triangle_rule(4, rp, rw);
    std::vector<double> r4;
    for (int q = 0; q < 6; ++q) r4.insert(r4.end(), {rp[2 * q], rp[2 * q + 1], rw[q]});
    if (e == hipSuccess) e = upload(&m->rule4, r4);
  }
  if (e == hipSuccess) e = upload(&m->face_perm, m->h.face_perm);
  if (e == hipSuccess) e = upload(&m->vert_inv, m->h.vert_inv);
  if (e == hipSuccess) e = upload(&m->wrows, m->h.weighted_rows);
  if (e == hipSuccess) e = upload(&m->wvals, m->h.row_weights);
  if (e != hipSuccess) {
    beat_mesh_destroy(m);
    return fail(ctx, BEAT_ERR_HIP, std::string("mesh upload: ") + hipGetErrorString(e));
  }
  m->regular_pairs = nf * nf - (int64_t)(m->h.spairs.size() - m->h.n_image_singular) + m->h.nimg * nf * nf;
  {
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int64_t v = 0; v < nv; ++v)
      for (int c = 0; c < 3; ++c) {
        lo[c] = std::fmin(lo[c], verts[3 * v + c]);
        hi[c] = std::fmax(hi[c], verts[3 * v + c]);
      }
    // a reflected trial face lies in the mirrored box: the union spans at most
    // twice the largest coordinate magnitude per axis
    double d2 = 0;
    for (int c = 0; c < 3; ++c) {
      const double span = m->h.nimg > 0 ? 2.0 * std::fmax(std::fabs(lo[c]), std::fabs(hi[c])) : hi[c] - lo[c];
      d2 += span * span;
    }
    m->diameter = std::sqrt(d2);
  }
  *out = m;
  return BEAT_OK;
}

int beat_mesh_upload(beat_ctx* ctx, const double* verts, int64_t nv, const int32_t* faces, int64_t nf,
                     int quad_order, int singular_order, beat_mesh** out) {
  return beat_mesh_upload_symmetric(ctx, verts, nv, faces, nf, quad_order, singular_order, 0, out);
}

void beat_mesh_destroy(beat_mesh* m) {
  if (!m) return;
  void* ptrs[] = {m->rec64, m->rec32, m->verts, m->fverts, m->floc, m->tile_start, m->tile_count, m->tile_nv,
                  m->tile_verts, m->stile_start, m->stile_count, m->stile_nv, m->stile_verts, m->flocs,
                  m->ntile_start, m->ntile_count, m->ntile_nv, m->ntile_verts, m->flocn, m->qtile_start,
                  m->qtile_count, m->qtile_nv, m->qtile_verts, m->flocq, m->tile_qfirst, m->sym_unit_start, m->frot, m->spairs,
                  m->rule_off, m->rule_cnt, m->rule_data, m->rule3, m->rule4, m->face_perm, m->vert_inv, m->wrows,
                  m->wvals};
  if (m->ctx) (void)hipSetDevice(m->ctx->device);
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  delete m;
}

int beat_mesh_counts(const beat_mesh* m, int64_t* counts) {
  if (!m || !counts) return BEAT_ERR_ARG;
  counts[0] = m->regular_pairs;
  counts[1] = m->h.n_coincident;
  counts[2] = m->h.n_edge;
  counts[3] = m->h.n_vertex;
  counts[4] = (int64_t)m->h.tile_start.size();
  counts[5] = m->h.nq;
  return BEAT_OK;
}

int beat_mesh_symmetry(const beat_mesh* m, int64_t* info) {
  if (!m || !info) return BEAT_ERR_ARG;
  info[0] = m->h.symmetry;
  info[1] = m->h.nimg;
  info[2] = m->h.n_image_singular;
  info[3] = (int64_t)m->h.weighted_rows.size();
  return BEAT_OK;
}

int beat_bm_assemble(beat_ctx* ctx, const beat_mesh* m, int precision, double k, double er, double ei, const double* q,
                     int64_t nrhs, double* a_out, double* b_out) {
  if (int rc = check_mesh(ctx, m)) return rc;
  if (!q || !a_out || !b_out || nrhs < 1) return fail(ctx, BEAT_ERR_ARG, "bad arguments");
  if (precision != BEAT_PREC_FP64 && precision != BEAT_PREC_FP32) return fail(ctx, BEAT_ERR_ARG, "precision");
  std::lock_guard<std::mutex> lock(ctx->mu);
  HIPCHK(hipSetDevice(ctx->device));
  const int64_t n = m->h.nv, nf = m->h.nf, lda = lda_for(n);
  hipStream_t s = ctx->stream;
  HIPCHK(ensure(ctx->qin, sizeof(double) * 2 * nf * nrhs));
  HIPCHK(hipMemcpyAsync(ctx->qin.p, q, sizeof(double) * 2 * nf * nrhs, hipMemcpyHostToDevice, s));
  HIPCHK(ensure(ctx->q64, sizeof(double) * 2 * nf * nrhs));
  hipLaunchKernelGGL((gather_q_kernel<double>), dim3(blocks_for(nf * nrhs)), dim3(256), 0, s,
                     (const double*)as<double>(ctx->qin), nf, (int)nrhs, (const int32_t*)m->face_perm,
                     as<double>(ctx->q64));
  HIPCHK(hipGetLastError());
  std::vector<double> A((size_t)2 * n * lda), B((size_t)2 * n * nrhs);
  if (precision == BEAT_PREC_FP64) {
    HIPCHK(ensure(ctx->a64, sizeof(double) * 2 * n * lda));
    HIPCHK(ensure(ctx->b64, sizeof(double) * 2 * n * nrhs));
    if (int rc = assemble<double>(ctx, m, k, er, ei, as<double>(ctx->q64), as<double>(ctx->q64), (int)nrhs,
                                  as<double>(ctx->a64), lda, as<double>(ctx->b64), s, nullptr, nullptr))
    return rc;
    HIPCHK(hipMemcpyAsync(A.data(), ctx->a64.p, sizeof(double) * A.size(), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(B.data(), ctx->b64.p, sizeof(double) * B.size(), hipMemcpyDeviceToHost, s));
  } else {
    HIPCHK(ensure(ctx->q32, sizeof(float) * 2 * nf * nrhs));
    HIPCHK(launch_downcast(as<double>(ctx->q64), as<float>(ctx->q32), 2 * nf * nrhs, s));
    HIPCHK(ensure(ctx->a32, sizeof(float) * 2 * n * lda));
    HIPCHK(ensure(ctx->b32, sizeof(float) * 2 * n * nrhs));
    if (int rc = assemble<float>(ctx, m, k, er, ei, as<double>(ctx->q64), as<float>(ctx->q32), (int)nrhs,
                                 as<float>(ctx->a32), lda, as<float>(ctx->b32), s, nullptr, nullptr))
    return rc;
    std::vector<float> A32(A.size()), B32(B.size());
    HIPCHK(hipMemcpyAsync(A32.data(), ctx->a32.p, sizeof(float) * A32.size(), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(B32.data(), ctx->b32.p, sizeof(float) * B32.size(), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    std::copy(A32.begin(), A32.end(), A.begin());
    std::copy(B32.begin(), B32.end(), B.begin());
  }
  HIPCHK(hipStreamSynchronize(s));
  const auto& inv = m->h.vert_inv;
  for (int64_t c = 0; c < n; ++c)
    for (int64_t r = 0; r < n; ++r) {
      const size_t src = 2 * ((size_t)inv[r] * lda + inv[c]);
      a_out[2 * (r + c * n)] = A[src];
      a_out[2 * (r + c * n) + 1] = A[src + 1];
    }
  for (int64_t d = 0; d < nrhs; ++d)
    for (int64_t r = 0; r < n; ++r) {
      const size_t src = 2 * ((size_t)inv[r] * nrhs + d);
      b_out[2 * (r + d * n)] = B[src];
      b_out[2 * (r + d * n) + 1] = B[src + 1];
    }
  return BEAT_OK;
}

int beat_operators_assemble(beat_ctx* ctx, const beat_mesh* m, int precision, double k, double* s_out, double* d_out,
                            double* k_out, double* h_out) {
  if (int rc = check_mesh(ctx, m)) return rc;
  if (!s_out || !d_out || !k_out || !h_out) return fail(ctx, BEAT_ERR_ARG, "bad arguments");
  if (precision != BEAT_PREC_FP64 && precision != BEAT_PREC_FP32) return fail(ctx, BEAT_ERR_ARG, "precision");
  std::lock_guard<std::mutex> lock(ctx->mu);
  HIPCHK(hipSetDevice(ctx->device));
  const int64_t n = m->h.nv, nf = m->h.nf, ldd = lda_for(n), ldf = lda_for(nf);
  hipStream_t s = ctx->stream;
  const size_t elems = (size_t)2 * (2 * n * ldd + 2 * n * ldf);
  std::vector<double> host(elems);
  if (precision == BEAT_PREC_FP64) {
    HIPCHK(ensure(ctx->a64, sizeof(double) * elems));
    double* base = as<double>(ctx->a64);
    double *S = base, *K = S + 2 * n * ldf, *D = K + 2 * n * ldf, *H = D + 2 * n * ldd;
    if (int rc = assemble_operators<double>(ctx, m, k, S, D, K, H, ldd, ldf)) return rc;
    HIPCHK(hipMemcpyAsync(host.data(), base, sizeof(double) * elems, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
  } else {
    HIPCHK(ensure(ctx->a32, sizeof(float) * elems));
    float* base = as<float>(ctx->a32);
    float *S = base, *K = S + 2 * n * ldf, *D = K + 2 * n * ldf, *H = D + 2 * n * ldd;
    if (int rc = assemble_operators<float>(ctx, m, k, S, D, K, H, ldd, ldf)) return rc;
    std::vector<float> h32(elems);
    HIPCHK(hipMemcpyAsync(h32.data(), base, sizeof(float) * elems, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    std::copy(h32.begin(), h32.end(), host.begin());
  }
  const double *S = host.data(), *K = S + 2 * n * ldf, *D = K + 2 * n * ldf, *H = D + 2 * n * ldd;
  const auto& vinv = m->h.vert_inv;
  std::vector<int32_t> finv(nf);
  for (int64_t i = 0; i < nf; ++i) finv[m->h.face_perm[i]] = (int32_t)i;
  for (int64_t c = 0; c < n; ++c)
    for (int64_t r = 0; r < n; ++r) {
      const size_t src = 2 * ((size_t)vinv[r] * ldd + vinv[c]), dst = 2 * (size_t)(r + c * n);
      d_out[dst] = D[src];
      d_out[dst + 1] = D[src + 1];
      h_out[dst] = H[src];
      h_out[dst + 1] = H[src + 1];
    }
  for (int64_t f = 0; f < nf; ++f)
    for (int64_t r = 0; r < n; ++r) {
      const size_t src = 2 * ((size_t)vinv[r] * ldf + finv[f]), dst = 2 * (size_t)(r + f * n);
      s_out[dst] = S[src];
      s_out[dst + 1] = S[src + 1];
      k_out[dst] = K[src];
      k_out[dst + 1] = K[src + 1];
    }
  return BEAT_OK;
}

int beat_bm_solve(beat_ctx* ctx, const beat_mesh* m, int precision, double k, double er, double ei, const double* q,
                  int64_t nrhs, double* pressure_out, const beat_solve_options* opts, beat_solve_report* rep) {
  if (int rc = check_mesh(ctx, m)) return rc;
  if (!q || !pressure_out || nrhs < 1 || nrhs > 4096) return fail(ctx, BEAT_ERR_ARG, "bad arguments");
  std::lock_guard<std::mutex> lock(ctx->mu);
  HIPCHK(hipSetDevice(ctx->device));
  const int64_t n = m->h.nv, nf = m->h.nf;
  HIPCHK(ensure(ctx->qin, sizeof(double) * 2 * nf * nrhs));
  HIPCHK(ensure(ctx->pout, sizeof(double) * 2 * n * nrhs));
  HIPCHK(hipMemcpyAsync(ctx->qin.p, q, sizeof(double) * 2 * nf * nrhs, hipMemcpyHostToDevice, ctx->stream));
  int rc = solve_impl(ctx, m, precision, k, er, ei, as<double>(ctx->qin), (int)nrhs, as<double>(ctx->pout), opts, rep);
  if (rc != BEAT_OK && rc != BEAT_ERR_SINGULAR) return rc;
  HIPCHK(hipMemcpyAsync(pressure_out, ctx->pout.p, sizeof(double) * 2 * n * nrhs, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return rc;
}

int beat_bm_solve_device(beat_ctx* ctx, const beat_mesh* m, int precision, double k, double er, double ei,
                         const void* q_dev, int64_t nrhs, void* p_dev, const beat_solve_options* opts,
                         beat_solve_report* rep) {
  if (int rc = check_mesh(ctx, m)) return rc;
  if (!q_dev || !p_dev || nrhs < 1 || nrhs > 4096) return fail(ctx, BEAT_ERR_ARG, "bad arguments");
  std::lock_guard<std::mutex> lock(ctx->mu);
  HIPCHK(hipSetDevice(ctx->device));
  return solve_impl(ctx, m, precision, k, er, ei, (const double*)q_dev, (int)nrhs, (double*)p_dev, opts, rep);
}

int beat_bm_sweep_device(beat_ctx* ctx, const beat_mesh* m, int precision, int64_t nfreq, const double* k,
                         const double* eta_re, const double* eta_im, const void* q_dev, int64_t nrhs, void* p_dev,
                         const beat_solve_options* opts, beat_solve_report* reps) {
  if (int rc = check_mesh(ctx, m)) return rc;
  if (nfreq < 0 || (nfreq > 0 && (!k || !eta_re || !eta_im || !q_dev || !p_dev)) || nrhs < 1 || nrhs > 4096)
    return fail(ctx, BEAT_ERR_ARG, "bad arguments");
  std::lock_guard<std::mutex> lock(ctx->mu);
  HIPCHK(hipSetDevice(ctx->device));
  return sweep_impl(ctx, m, precision, nfreq, k, eta_re, eta_im, (const double*)q_dev, (int)nrhs, (double*)p_dev, opts,
                    reps);
}

int beat_bm_sweep(beat_ctx* ctx, const beat_mesh* m, int precision, int64_t nfreq, const double* k,
                  const double* eta_re, const double* eta_im, const double* q, int64_t nrhs, double* pressure_out,
                  const beat_solve_options* opts, beat_solve_report* reps) {
  if (int rc = check_mesh(ctx, m)) return rc;
  if (nfreq < 0 || (nfreq > 0 && (!k || !eta_re || !eta_im || !q || !pressure_out)) || nrhs < 1 || nrhs > 4096)
    return fail(ctx, BEAT_ERR_ARG, "bad arguments");
  if (nfreq == 0) return BEAT_OK;
  std::lock_guard<std::mutex> lock(ctx->mu);
  HIPCHK(hipSetDevice(ctx->device));
  const int64_t n = m->h.nv, nf = m->h.nf;
  HIPCHK(ensure(ctx->qin, sizeof(double) * 2 * nf * nrhs * nfreq));
  HIPCHK(ensure(ctx->pout, sizeof(double) * 2 * n * nrhs * nfreq));
  HIPCHK(hipMemcpyAsync(ctx->qin.p, q, sizeof(double) * 2 * nf * nrhs * nfreq, hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  int rc = sweep_impl(ctx, m, precision, nfreq, k, eta_re, eta_im, as<double>(ctx->qin), (int)nrhs,
                      as<double>(ctx->pout), opts, reps);
  // a cancelled sweep still returns the frequencies it solved
  if (rc != BEAT_OK && rc != BEAT_ERR_SINGULAR && rc != BEAT_ERR_CANCELLED) return rc;
  HIPCHK(hipMemcpyAsync(pressure_out, ctx->pout.p, sizeof(double) * 2 * n * nrhs * nfreq, hipMemcpyDeviceToHost,
                        ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return rc;
}


int beat_adjoint_spl_gradient(beat_ctx* ctx, const beat_mesh* m, int64_t nfreq, const double* k, const double* eta_re,
                              const double* eta_im, const double* q, const double* pts, int64_t npts, double target_db,
                              double* loss_out, double* spl_out, double* grad_out) {
  if (int rc = check_mesh(ctx, m)) return rc;
  if (nfreq < 1 || npts < 1 || !k || !eta_re || !eta_im || !q || !pts || !loss_out || !grad_out)
    return fail(ctx, BEAT_ERR_ARG, "bad arguments");
  if (m->h.nimg != 0) return fail(ctx, BEAT_ERR_ARG, "the adjoint gradient supports meshes without symmetry images");
  std::lock_guard<std::mutex> lock(ctx->mu);
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  const int64_t n = m->h.nv, nf = m->h.nf, lda = lda_for(n);
  const int nq = m->h.nq;
  // device layout (doubles): gV n*3 | seeds nf*34 | fi nf*38 | mu 2n | rhs 2n | v 2n | cw 2npts | pts 3npts | zero 2n
  const size_t o_gv = 0, o_sd = o_gv + 3 * n, o_fi = o_sd + 34 * nf, o_mu = o_fi + 38 * nf, o_rhs = o_mu + 2 * n,
               o_v = o_rhs + 2 * n, o_cw = o_v + 2 * n, o_pts = o_cw + 2 * npts, o_zero = o_pts + 3 * npts,
               o_end = o_zero + 2 * n;
  HIPCHK(ensure(ctx->adj, sizeof(double) * o_end));
  double* base = as<double>(ctx->adj);
  double *gV = base + o_gv, *seeds = base + o_sd, *fi = base + o_fi, *mu = base + o_mu, *rhs = base + o_rhs,
         *v = base + o_v, *cw = base + o_cw, *dpts = base + o_pts, *zero = base + o_zero;
  HIPCHK(launch_fill<double>(gV, 3 * n, 0.0, s));
  HIPCHK(launch_fill<double>(zero, 2 * n, 0.0, s));
  HIPCHK(hipMemcpyAsync(dpts, pts, sizeof(double) * 3 * npts, hipMemcpyHostToDevice, s));
  HIPCHK(ensure(ctx->qin, sizeof(double) * 2 * nf));
  HIPCHK(ensure(ctx->q64, sizeof(double) * 2 * nf));
  HIPCHK(ensure(ctx->a64, sizeof(double) * 2 * n * lda));
  HIPCHK(ensure(ctx->b64, sizeof(double) * 2 * n));
  const int64_t ns = nf * nq;
  HIPCHK(ensure(ctx->fld_src, sizeof(double) * (ns * 6 + 4 * ns)));
  HIPCHK(ensure(ctx->fld_out, sizeof(double) * 2 * npts));
  IdentityBlocks idb;
  for (int a = 0; a < 3; ++a) {
    idb.pd[a] = m->h.id_pd[a];
    for (int b = 0; b < 3; ++b) idb.pp[a][b] = m->h.id_pp[a][b];
  }
  double loss = 0;
  std::vector<double> u(2 * npts), c(2 * npts);
  const double ln10 = std::log(10.0);
  for (int64_t f = 0; f < nfreq; ++f) {
    // forward: complex128 system, LU, solve
    HIPCHK(hipMemcpyAsync(ctx->qin.p, q + 2 * f * nf, sizeof(double) * 2 * nf, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL((gather_q_kernel<double>), dim3(blocks_for(nf)), dim3(256), 0, s,
                       (const double*)as<double>(ctx->qin), nf, 1, (const int32_t*)m->face_perm, as<double>(ctx->q64));
    HIPCHK(hipGetLastError());
    const double* qd = as<double>(ctx->q64);
    double* A = as<double>(ctx->a64);
    double* p = as<double>(ctx->b64);
    if (int rc = assemble<double>(ctx, m, k[f], eta_re[f], eta_im[f], qd, qd, 1, A, lda, p, s, nullptr, nullptr))
      return rc;
    LuWork lw;
    if (ensure_lu_work<double>(ctx, n, lw) != BEAT_OK) return BEAT_ERR_HIP;
    int panels = 0;
    HIPCHK(lu_factor<double>(A, n, lda, lw, LU_NB, s, &panels));
    if (int rc = solve_all<double>(ctx, A, n, lda, lw.ipiv, p, 1, true)) return rc;
    // field at the observation points
    double* sp = as<double>(ctx->fld_src);
    double *sn = sp + 3 * ns, *pw = sn + 3 * ns, *qw = pw + 2 * ns;
    HIPCHK(field_sources<double>(m->rec64, m->fverts, nf, nq, m->rule3, p, qd, 1, 0, m->h.img, sp, sn, pw, qw, s));
    HIPCHK(field_eval<double>(sp, sn, pw, qw, ns, 1, dpts, npts, as<double>(ctx->fld_out), k[f], s));
    HIPCHK(hipMemcpyAsync(u.data(), ctx->fld_out.p, sizeof(double) * 2 * npts, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    int info = 0;
    if (int rc = lu_outcome(ctx, lw, &info)) return rc;
    if (info != 0) return fail(ctx, BEAT_ERR_SINGULAR, "LU found an exactly zero pivot at column " + std::to_string(info));
    // loss and c = conj(dL/du*) = 2 (S - target) (10 / ln 10) conj(u) / |u|^2
    for (int64_t d = 0; d < npts; ++d) {
      const double ur = u[2 * d], ui = u[2 * d + 1], a2 = ur * ur + ui * ui;
      const double spl = 20.0 * std::log10(std::sqrt(a2) / 20e-6);
      if (spl_out) spl_out[d + f * npts] = spl;
      loss += (spl - target_db) * (spl - target_db);
      const double g = 2.0 * (spl - target_db) * (10.0 / ln10) / a2;
      c[2 * d] = g * ur;
      c[2 * d + 1] = -g * ui;
    }
    if (!std::isfinite(loss)) return fail(ctx, BEAT_ERR_ARG, "the field vanishes at an observation point");
    HIPCHK(hipMemcpyAsync(cw, c.data(), sizeof(double) * 2 * npts, hipMemcpyHostToDevice, s));
    // field part and the adjoint right-hand side F^T c
    HIPCHK(launch_fill<double>(seeds, 34 * nf, 0.0, s));
    HIPCHK(launch_fill<double>(rhs, 2 * n, 0.0, s));
    HIPCHK(adjoint_face_interp(m->rec64, m->fverts, nf, nq, m->rule3, zero, p, qd, fi, s));
    HIPCHK(adjoint_field(m->rec64, m->fverts, nf, nq, m->rule3, fi, dpts, cw, npts, k[f], seeds, rhs, s));
    // A^T mu = F^T c with the same factors
    HIPCHK(hipMemcpyAsync(v, rhs, sizeof(double) * 2 * n, hipMemcpyDeviceToDevice, s));
    HIPCHK(lu_solve_transposed<double>(A, n, lda, v, mu, as<double>(ctx->work), s));
    // system part and the chain to the vertices
    HIPCHK(adjoint_face_interp(m->rec64, m->fverts, nf, nq, m->rule3, mu, p, qd, fi, s));
    HIPCHK(adjoint_pairs(m->rec64, m->verts, m->fverts, nf, nq, m->rule3, m->spairs, (int64_t)m->h.spairs.size(),
                         m->rule_off, m->rule_cnt, m->rule_data, fi, mu, p, k[f], eta_re[f], eta_im[f], seeds, gV, s));
    HIPCHK(adjoint_chain(m->verts, m->fverts, nf, nq, m->rule3, seeds, mu, p, qd, eta_re[f], eta_im[f], idb, gV, s));
  }
  std::vector<double> g(3 * n);
  HIPCHK(hipMemcpyAsync(g.data(), gV, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  const auto& inv = m->h.vert_inv;
  for (int64_t vtx = 0; vtx < n; ++vtx)
    for (int cc = 0; cc < 3; ++cc) grad_out[3 * vtx + cc] = 2.0 * g[3 * (size_t)inv[vtx] + cc];
  *loss_out = loss;
  return BEAT_OK;
}

int beat_autotune(beat_ctx* ctx, const beat_mesh* m, int precision, double k, int64_t nrhs, double* timings,
                  int* choice) {
  if (int rc = check_mesh(ctx, m)) return rc;
  if (nrhs < 1 || nrhs > 64 || !(k > 0) || !choice) return fail(ctx, BEAT_ERR_ARG, "bad arguments");
  if (int rc = check_solve_options(ctx, precision, nullptr)) return rc;
  std::lock_guard<std::mutex> lock(ctx->mu);
  HIPCHK(hipSetDevice(ctx->device));
  const int64_t n = m->h.nv, nf = m->h.nf;
  Tuning& tu = tuned(ctx, precision, n);
  constexpr int NJ = 8;  // jobs of the trial batches
  // synthetic drive: unit normal velocity on every face and right-hand side
  std::vector<double> qh(2 * nf * nrhs * NJ);
  for (int64_t i = 0; i < nf * nrhs * NJ; ++i) {
    qh[2 * i] = 0.0;
    qh[2 * i + 1] = 1.0 + 0.01 * (double)(i % 7);
  }
  HIPCHK(ensure(ctx->qin, sizeof(double) * qh.size()));
  HIPCHK(ensure(ctx->pout, sizeof(double) * 2 * n * nrhs * NJ));
  HIPCHK(hipMemcpy(ctx->qin.p, qh.data(), sizeof(double) * qh.size(), hipMemcpyHostToDevice));
  const double eta_re = 0.0, eta_im = 1.0 / k;
  // one solve at a time: the assembly's and the factorization's times from the
  // solve's own events, best of two after a warm-up
  auto run_stages = [&](double* asm_best, double* fac_best) -> int {
    *asm_best = *fac_best = INFINITY;
    for (int r = 0; r < 3; ++r) {
      beat_solve_report rep{};
      if (int rc = solve_impl(ctx, m, precision, k, eta_re, eta_im, as<double>(ctx->qin), (int)nrhs,
                              as<double>(ctx->pout), nullptr, &rep))
        return rc;
      if (r > 0) {
        *asm_best = std::min(*asm_best, rep.assembly_s);
        *fac_best = std::min(*fac_best, rep.factor_s);
      }
    }
    return BEAT_OK;
  };
  // a batch of NJ (or 3) jobs on this mesh, wall time, best of two after a warm-up
  std::vector<const beat_mesh*> ms(NJ, m);
  std::vector<double> ks(NJ, k), ers(NJ, eta_re), eis(NJ, eta_im);
  std::vector<int64_t> qo(NJ), po(NJ);
  for (int i = 0; i < NJ; ++i) {
    qo[i] = i * nf * nrhs;
    po[i] = i * n * nrhs;
  }
  auto run_batch = [&](int nj, bool grouped, double* best) -> int {
    *best = INFINITY;
    for (int r = 0; r < 3; ++r) {
      const auto t0 = std::chrono::steady_clock::now();
      int rc = grouped ? batch_grouped_impl(ctx, nj, ms.data(), precision, ks.data(), ers.data(), eis.data(),
                                            as<double>(ctx->qin), qo.data(), (int)nrhs, as<double>(ctx->pout),
                                            po.data(), nullptr, nullptr)
                       : batch_impl(ctx, nj, ms.data(), precision, ks.data(), ers.data(), eis.data(),
                                    as<double>(ctx->qin), qo.data(), (int)nrhs, as<double>(ctx->pout), po.data(),
                                    nullptr, nullptr);
      if (rc) return rc;
      const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      if (r > 0) *best = std::min(*best, dt);
    }
    return BEAT_OK;
  };
  double t[4 + GEMM_TILES + 3];
  for (double& v : t) v = INFINITY;
  double unused;
  // the regular pass's tiles: the assembly alone
  for (int v = 0; v < 3; ++v) {
    tu.asm_tile = (signed char)v;
    if (int rc = run_stages(&t[4 + GEMM_TILES + v], &unused)) return rc;
  }
  tu.asm_tile = (signed char)(std::min_element(t + 4 + GEMM_TILES, t + 4 + GEMM_TILES + 3) - (t + 4 + GEMM_TILES));
  // factorization stream layout: the factorization alone
  for (int layout = 0; layout < 2; ++layout) {
    tu.lu_layout = (signed char)layout;
    if (layout == 0 && (!ctx->side || !ctx->inb)) continue;
    if (int rc = run_stages(&unused, &t[layout])) return rc;
  }
  tu.lu_layout = t[1] < t[0] ? 1 : 0;
  // the trailing updates' tile shape: the factorization alone, or, for systems the
  // grouped batches take, a grouped batch of NJ systems (its batched factorization)
  const bool grouped = ctx->batch_grouped && ctx->cancel_path.empty() && n <= GROUP_NMAX && nrhs <= SOLVE_CHUNK;
  for (int v = 0; v < GEMM_TILES; ++v) {
    tu.gemm_tile = (signed char)v;
    int rc = grouped ? run_batch(NJ, true, &t[4 + v]) : run_stages(&unused, &t[4 + v]);
    if (rc) return rc;
  }
  tu.gemm_tile = (signed char)(std::min_element(t + 4, t + 4 + GEMM_TILES) - (t + 4));
  // assembly occupancy inside a (per-job pipeline) batch of three solves
  for (int occ = 0; occ < 2; ++occ) {
    tu.asm_occ = (signed char)occ;
    if (int rc = run_batch(3, false, &t[2 + occ])) return rc;
  }
  tu.asm_occ = t[3] < t[2] ? 1 : 0;
  choice[0] = tu.lu_layout;
  choice[1] = tu.asm_occ;
  choice[2] = tu.gemm_tile;
  choice[3] = tu.asm_tile;
  if (timings)
    for (int i = 0; i < 4 + GEMM_TILES + 3; ++i) timings[i] = t[i];
  return BEAT_OK;
}

int beat_ctx_set_tuning(beat_ctx* ctx, int precision, int64_t nv, const int* choice) {
  if (!ctx) return BEAT_ERR_ARG;
  if (nv < 1 || !choice) return fail(ctx, BEAT_ERR_ARG, "bad arguments");
  if (int rc = check_solve_options(ctx, precision, nullptr)) return rc;
  const int lim[4] = {2, 2, GEMM_TILES, 3};
  for (int i = 0; i < 4; ++i)
    if (choice[i] < -1 || choice[i] >= lim[i]) return fail(ctx, BEAT_ERR_ARG, "tuning choice out of range");
  std::lock_guard<std::mutex> lock(ctx->mu);
  Tuning& tu = tuned(ctx, precision, nv);
  tu.lu_layout = (signed char)choice[0];
  tu.asm_occ = (signed char)choice[1];
  tu.gemm_tile = (signed char)choice[2];
  tu.asm_tile = (signed char)choice[3];
  return BEAT_OK;
}

int beat_bm_batch(beat_ctx* ctx, int64_t njobs, const beat_mesh* const* meshes, int precision, const double* k,
                  const double* eta_re, const double* eta_im, const double* q, int64_t nrhs, double* pressure_out,
                  const beat_solve_options* opts, beat_solve_report* reps) {
  if (!ctx) return BEAT_ERR_ARG;
  if (njobs < 0 || (njobs > 0 && (!meshes || !k || !eta_re || !eta_im || !q || !pressure_out)) || nrhs < 1 ||
      nrhs > 4096)
    return fail(ctx, BEAT_ERR_ARG, "bad arguments");
  if (njobs == 0) return BEAT_OK;
  std::vector<int64_t> qo(njobs), po(njobs);
  int64_t qt = 0, pt = 0;
  for (int64_t i = 0; i < njobs; ++i) {
    if (int rc = check_mesh(ctx, meshes[i])) return rc;
    qo[i] = qt;
    po[i] = pt;
    qt += meshes[i]->h.nf * nrhs;
    pt += meshes[i]->h.nv * nrhs;
  }
  std::lock_guard<std::mutex> lock(ctx->mu);
  HIPCHK(hipSetDevice(ctx->device));
  HIPCHK(ensure(ctx->qin, sizeof(double) * 2 * qt));
  HIPCHK(ensure(ctx->pout, sizeof(double) * 2 * pt));
  HIPCHK(hipMemcpyAsync(ctx->qin.p, q, sizeof(double) * 2 * qt, hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  {
    // systems small enough go through the grouped path: batched factorizations,
    // solves and refinement, replayed as one graph per group shape
    bool grouped = nrhs <= SOLVE_CHUNK && ctx->batch_grouped && ctx->cancel_path.empty();
    for (int64_t i = 0; i < njobs && grouped; ++i) grouped = meshes[i]->h.nv <= GROUP_NMAX;
    if (grouped) {
      int rc = batch_grouped_impl(ctx, njobs, meshes, precision, k, eta_re, eta_im, as<double>(ctx->qin), qo.data(),
                                  (int)nrhs, as<double>(ctx->pout), po.data(), opts, reps);
      if (rc != BEAT_OK && rc != BEAT_ERR_SINGULAR) return rc;
      HIPCHK(hipMemcpyAsync(pressure_out, ctx->pout.p, sizeof(double) * 2 * pt, hipMemcpyDeviceToHost, ctx->stream));
      HIPCHK(hipStreamSynchronize(ctx->stream));
      return rc;
    }
  }
  // Small systems leave most of the device idle, so several contexts run jobs at
  // once. Their factorizations and triangular solves keep all their workgroups
  // waiting on each other; the number of contexts is bounded so that those
  // workgroups of every context together never exceed the CU count, which lets
  // each of them find a CU once the other (finite) kernels drain.
  int64_t nmax = 0;
  for (int64_t i = 0; i < njobs; ++i) nmax = std::max<int64_t>(nmax, meshes[i]->h.nv);
  int cus = 0;
  HIPCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
  const int64_t spin = (nmax + 63) / 64;  // workgroups of one factorization panel or block solve
  int workers = (int)std::min<int64_t>({(int64_t)ctx->batch_workers, (int64_t)cus / std::max<int64_t>(spin, 1),
                                        njobs});
  if (workers < 1) workers = 1;
  while ((int)ctx->aux.size() < workers - 1) {
    beat_ctx* a = nullptr;
    if (beat_ctx_create(ctx->device, &a) != BEAT_OK) return fail(ctx, BEAT_ERR_HIP, "could not create a batch context");
    ctx->aux.push_back(a);
  }
  std::vector<beat_ctx*> cs{ctx};
  for (int w = 1; w < workers; ++w) {
    beat_ctx* a = ctx->aux[w - 1];
    a->cancel_path = ctx->cancel_path;
    a->debug_abort_waits = ctx->debug_abort_waits;
    a->debug_force_exchange = ctx->debug_force_exchange;
    a->regular_symmetric = ctx->regular_symmetric;
    a->fp32_assembly_fp64 = ctx->fp32_assembly_fp64;
    a->sweep_asm_one_per_cu = ctx->sweep_asm_one_per_cu;
    std::memcpy(a->tune, ctx->tune, sizeof(ctx->tune));
    cs.push_back(a);
  }
  // worker w takes jobs w, w + workers, ... (interleaved sizes)
  std::vector<int> rcs(workers, BEAT_OK);
  auto run = [&](int w) {
    beat_ctx* c = cs[w];
    if (hipSetDevice(c->device) != hipSuccess) {
      rcs[w] = BEAT_ERR_HIP;
      return;
    }
    std::vector<const beat_mesh*> ms;
    std::vector<double> kk, er, ei;
    std::vector<int64_t> qq, pp, idx;
    for (int64_t i = w; i < njobs; i += workers) {
      ms.push_back(meshes[i]);
      kk.push_back(k[i]);
      er.push_back(eta_re[i]);
      ei.push_back(eta_im[i]);
      qq.push_back(qo[i]);
      pp.push_back(po[i]);
      idx.push_back(i);
    }
    std::vector<beat_solve_report> rr(reps ? ms.size() : 0);
    if (reps)
      for (size_t j = 0; j < ms.size(); ++j) rr[j] = reps[idx[j]];
    rcs[w] = batch_impl(c, (int64_t)ms.size(), ms.data(), precision, kk.data(), er.data(), ei.data(),
                        as<double>(ctx->qin), qq.data(), (int)nrhs, as<double>(ctx->pout), pp.data(), opts,
                        reps ? rr.data() : nullptr);
    if (reps)
      for (size_t j = 0; j < ms.size(); ++j) reps[idx[j]] = rr[j];
  };
  if (workers == 1) {
    run(0);
  } else {
    std::vector<std::thread> th;
    for (int w = 1; w < workers; ++w) th.emplace_back(run, w);
    run(0);
    for (auto& t : th) t.join();
  }
  int rc = BEAT_OK;
  for (int w = 0; w < workers; ++w) {
    const int r = rcs[w];
    if (r == BEAT_OK) continue;
    if (w > 0 && r != BEAT_OK) ctx->err = cs[w]->err;
    if (rc == BEAT_OK || (rc == BEAT_ERR_SINGULAR || rc == BEAT_ERR_CANCELLED)) rc = r;
  }
  if (rc != BEAT_OK && rc != BEAT_ERR_SINGULAR && rc != BEAT_ERR_CANCELLED) return rc;
  HIPCHK(hipMemcpyAsync(pressure_out, ctx->pout.p, sizeof(double) * 2 * pt, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return rc;
}

int beat_field_eval(beat_ctx* ctx, const beat_mesh* m, int precision, double k, const double* pressure,
                    const double* q, int64_t nrhs, const double* pts, int64_t npts, double* out) {
  return field_impl(ctx, m, precision, k, pressure, q, nrhs, pts, npts, out, false);
}

int beat_field_eval_device_noalloc(beat_ctx* ctx, const beat_mesh* m, int precision, double k, const void* pressure_dev,
                                   const void* q_dev, int64_t nrhs, const void* pts_dev, int64_t npts, void* out_dev) {
  return field_noalloc_impl(ctx, m, precision, k, static_cast<const double*>(pressure_dev),
                            static_cast<const double*>(q_dev), nrhs, static_cast<const double*>(pts_dev), npts,
                            static_cast<double*>(out_dev));
}

int beat_field_gradient_eval(beat_ctx* ctx, const beat_mesh* m, int precision, double k, const double* pressure,
                             const double* q, int64_t nrhs, const double* pts, int64_t npts, double* out) {
  return field_impl(ctx, m, precision, k, pressure, q, nrhs, pts, npts, out, true);
}

int beat_dense_lu_solve(beat_ctx* ctx, int precision, const double* a, int64_t n, const double* b, int64_t nrhs,
                        double* x) {
  if (!ctx) return BEAT_ERR_ARG;
  if (!a || !b || !x || n < 1 || nrhs < 1) return fail(ctx, BEAT_ERR_ARG, "bad arguments");
  if (precision != BEAT_PREC_FP64 && precision != BEAT_PREC_FP32) return fail(ctx, BEAT_ERR_ARG, "precision");
  std::lock_guard<std::mutex> lock(ctx->mu);
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  const int64_t lda = lda_for(n);
  HIPCHK(ensure(ctx->qin, sizeof(double) * 2 * (n * n + n * nrhs)));
  double* ain = as<double>(ctx->qin);
  double* bin = ain + 2 * n * n;
  HIPCHK(hipMemcpyAsync(ain, a, sizeof(double) * 2 * n * n, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(bin, b, sizeof(double) * 2 * n * nrhs, hipMemcpyHostToDevice, s));
  LuWork lw;
  if (ensure_lu_work<double>(ctx, n, lw) != BEAT_OK) return BEAT_ERR_HIP;
  HIPCHK(ensure(ctx->pout, sizeof(double) * 2 * n * nrhs));
  int panels = 0, info = 0;
  // speculative factorization first; with exchanges when a column needs them
  for (int attempt = 0; attempt < 2; ++attempt) {
  lw.nopiv = attempt == 0;
  if (precision == BEAT_PREC_FP64) {
    HIPCHK(ensure(ctx->a64, sizeof(double) * 2 * n * lda));
    HIPCHK(ensure(ctx->b64, sizeof(double) * 2 * n * nrhs));
    hipLaunchKernelGGL((transpose_in_kernel<double>), dim3(blocks_for(n * n)), dim3(256), 0, s, (const double*)ain, n,
                       lda, as<double>(ctx->a64));
    hipLaunchKernelGGL((cols_in_kernel<double, double>), dim3(blocks_for(n * nrhs)), dim3(256), 0, s,
                       (const double*)bin, n, (int)nrhs, as<double>(ctx->b64));
    HIPCHK(hipGetLastError());
    HIPCHK(lu_factor<double>(as<double>(ctx->a64), n, lda, lw, LU_NB, s, &panels));
    if (int rc = solve_all<double>(ctx, as<double>(ctx->a64), n, lda, lw.ipiv, as<double>(ctx->b64), (int)nrhs, true))
    return rc;
    hipLaunchKernelGGL((cols_out_kernel<double, double>), dim3(blocks_for(n * nrhs)), dim3(256), 0, s,
                       (const double*)as<double>(ctx->b64), n, (int)nrhs, as<double>(ctx->pout));
  } else {
    HIPCHK(ensure(ctx->a32, sizeof(float) * 2 * n * lda));
    HIPCHK(ensure(ctx->b32, sizeof(float) * 2 * n * nrhs));
    hipLaunchKernelGGL((transpose_in_kernel<float>), dim3(blocks_for(n * n)), dim3(256), 0, s, (const double*)ain, n,
                       lda, as<float>(ctx->a32));
    hipLaunchKernelGGL((cols_in_kernel<double, float>), dim3(blocks_for(n * nrhs)), dim3(256), 0, s, (const double*)bin,
                       n, (int)nrhs, as<float>(ctx->b32));
    HIPCHK(hipGetLastError());
    HIPCHK(lu_factor<float>(as<float>(ctx->a32), n, lda, lw, LU_NB, s, &panels));
    if (int rc = solve_all<float>(ctx, as<float>(ctx->a32), n, lda, lw.ipiv, as<float>(ctx->b32), (int)nrhs, true))
    return rc;
    hipLaunchKernelGGL((cols_out_kernel<float, double>), dim3(blocks_for(n * nrhs)), dim3(256), 0, s,
                       (const float*)as<float>(ctx->b32), n, (int)nrhs, as<double>(ctx->pout));
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(x, ctx->pout.p, sizeof(double) * 2 * n * nrhs, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  bool exchange = false;
  if (int rc = lu_outcome(ctx, lw, &info, &exchange)) return rc;
  if (!exchange) break;
  }
  if (info != 0) return fail(ctx, BEAT_ERR_SINGULAR, "LU found an exactly zero pivot at column " + std::to_string(info));
  return BEAT_OK;
}

extern "C++" {
namespace {
// Factors of lu_factor in LAPACK getrf form: the exchanges of every later outer
// block applied to the earlier blocks' columns, then column-major complex128 and
// 1-based pivots.
template <typename T>
int dense_factor_out(beat_ctx* ctx, const T* LU, int64_t n, int64_t lda, const int* ipiv_dev, double* lu_out,
                     int32_t* ipiv_out) {
  std::vector<T> h((size_t)2 * n * lda);
  std::vector<int> piv((size_t)n);
  HIPCHK(hipMemcpyAsync(h.data(), LU, sizeof(T) * h.size(), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipMemcpyAsync(piv.data(), ipiv_dev, sizeof(int) * n, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  for (int64_t k = 0; k < n; ++k) {
    const int64_t p = piv[k], cend = k / LU_BLOCK_COLS * LU_BLOCK_COLS;
    if (p < k || p >= n) return fail(ctx, BEAT_ERR_HIP, "pivot index out of range");
    if (p != k)
      for (int64_t c = 0; c < 2 * cend; ++c) std::swap(h[2 * k * lda + c], h[2 * p * lda + c]);
    ipiv_out[k] = (int32_t)(p + 1);
  }
  for (int64_t c = 0; c < n; ++c)
    for (int64_t r = 0; r < n; ++r) {
      lu_out[2 * (c * n + r)] = (double)h[2 * (r * lda + c)];
      lu_out[2 * (c * n + r) + 1] = (double)h[2 * (r * lda + c) + 1];
    }
  return BEAT_OK;
}
}  // namespace
}  // extern "C++"

int beat_dense_lu_factor(beat_ctx* ctx, int precision, const double* a, int64_t n, double* lu_out, int32_t* ipiv_out,
                         int32_t* info_out) {
  if (!ctx) return BEAT_ERR_ARG;
  if (!a || !lu_out || !ipiv_out || !info_out || n < 1) return fail(ctx, BEAT_ERR_ARG, "bad arguments");
  if (precision != BEAT_PREC_FP64 && precision != BEAT_PREC_FP32) return fail(ctx, BEAT_ERR_ARG, "precision");
  std::lock_guard<std::mutex> lock(ctx->mu);
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  const int64_t lda = lda_for(n);
  HIPCHK(ensure(ctx->qin, sizeof(double) * 2 * n * n));
  double* ain = as<double>(ctx->qin);
  HIPCHK(hipMemcpyAsync(ain, a, sizeof(double) * 2 * n * n, hipMemcpyHostToDevice, s));
  LuWork lw;
  if (ensure_lu_work<double>(ctx, n, lw) != BEAT_OK) return BEAT_ERR_HIP;
  int panels = 0, rc, info = 0;
  // speculative factorization first (its factors are getrf's whenever its check
  // passes); with exchanges when a column needs them
  for (int attempt = 0; attempt < 2; ++attempt) {
    lw.nopiv = attempt == 0;
    if (precision == BEAT_PREC_FP64) {
      HIPCHK(ensure(ctx->a64, sizeof(double) * 2 * n * lda));
      hipLaunchKernelGGL((transpose_in_kernel<double>), dim3(blocks_for(n * n)), dim3(256), 0, s, (const double*)ain,
                         n, lda, as<double>(ctx->a64));
      HIPCHK(hipGetLastError());
      HIPCHK(lu_factor<double>(as<double>(ctx->a64), n, lda, lw, LU_NB, s, &panels));
    } else {
      HIPCHK(ensure(ctx->a32, sizeof(float) * 2 * n * lda));
      hipLaunchKernelGGL((transpose_in_kernel<float>), dim3(blocks_for(n * n)), dim3(256), 0, s, (const double*)ain, n,
                         lda, as<float>(ctx->a32));
      HIPCHK(hipGetLastError());
      HIPCHK(lu_factor<float>(as<float>(ctx->a32), n, lda, lw, LU_NB, s, &panels));
    }
    HIPCHK(hipStreamSynchronize(s));
    bool exchange = false;
    if ((rc = lu_outcome(ctx, lw, &info, &exchange)) != BEAT_OK) return rc;
    if (!exchange) break;
  }
  rc = precision == BEAT_PREC_FP64
           ? dense_factor_out<double>(ctx, as<double>(ctx->a64), n, lda, lw.ipiv, lu_out, ipiv_out)
           : dense_factor_out<float>(ctx, as<float>(ctx->a32), n, lda, lw.ipiv, lu_out, ipiv_out);
  if (rc != BEAT_OK) return rc;
  *info_out = info;
  return BEAT_OK;
}

int beat_refine_lu(beat_ctx* ctx, const float* matrix, int64_t n, const float* rhs, int64_t nrhs,
                   const double* residual_matrix, const double* residual_rhs, int max_steps, double rtol, float* x_out,
                   beat_refine_report* rep) {
  if (!ctx) return BEAT_ERR_ARG;
  if (max_steps < 1 || max_steps > 3) return fail(ctx, BEAT_ERR_ARG, "max_steps must be in 1:3");
  if (!(rtol > 0) || !std::isfinite(rtol)) return fail(ctx, BEAT_ERR_ARG, "rtol must be finite and positive");
  if (!matrix || !rhs || !x_out || n < 1 || nrhs < 1) return fail(ctx, BEAT_ERR_ARG, "bad arguments");
  std::lock_guard<std::mutex> lock(ctx->mu);
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  const int64_t lda = lda_for(n);
  // host staging: widen the complex64 inputs where no residual operator is given
  std::vector<double> A64(2 * n * n), b64(2 * n * nrhs);
  for (int64_t i = 0; i < 2 * n * n; ++i) A64[i] = residual_matrix ? residual_matrix[i] : (double)matrix[i];
  for (int64_t i = 0; i < 2 * n * nrhs; ++i) b64[i] = residual_rhs ? residual_rhs[i] : (double)rhs[i];
  HIPCHK(ensure(ctx->qin, sizeof(double) * 2 * (n * n + n * nrhs) + sizeof(float) * 2 * (n * n + n * nrhs)));
  double* ain = as<double>(ctx->qin);
  double* bin = ain + 2 * n * n;
  float* mat32 = reinterpret_cast<float*>(bin + 2 * n * nrhs);
  float* rhs32 = mat32 + 2 * n * n;
  HIPCHK(hipMemcpyAsync(ain, A64.data(), sizeof(double) * A64.size(), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(bin, b64.data(), sizeof(double) * b64.size(), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(mat32, matrix, sizeof(float) * 2 * n * n, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(rhs32, rhs, sizeof(float) * 2 * n * nrhs, hipMemcpyHostToDevice, s));
  HIPCHK(ensure(ctx->a64, sizeof(double) * 2 * n * lda));
  HIPCHK(ensure(ctx->a32, sizeof(float) * 2 * n * lda));
  HIPCHK(ensure(ctx->b64, sizeof(double) * 2 * n * nrhs));
  HIPCHK(ensure(ctx->b32, sizeof(float) * 2 * n * nrhs));
  HIPCHK(ensure(ctx->x64, sizeof(double) * 2 * n * nrhs));
  hipLaunchKernelGGL((transpose_in_kernel<double>), dim3(blocks_for(n * n)), dim3(256), 0, s, (const double*)ain, n, lda,
                     as<double>(ctx->a64));
  hipLaunchKernelGGL((transpose_in_f32_kernel<float>), dim3(blocks_for(n * n)), dim3(256), 0, s, (const float*)mat32, n,
                     lda, as<float>(ctx->a32));
  hipLaunchKernelGGL((cols_in_kernel<double, double>), dim3(blocks_for(n * nrhs)), dim3(256), 0, s, (const double*)bin,
                     n, (int)nrhs, as<double>(ctx->b64));
  hipLaunchKernelGGL((cols_in_kernel<float, float>), dim3(blocks_for(n * nrhs)), dim3(256), 0, s, (const float*)rhs32, n,
                     (int)nrhs, as<float>(ctx->b32));
  HIPCHK(hipGetLastError());
  LuWork lw;
  if (ensure_lu_work<double>(ctx, n, lw) != BEAT_OK) return BEAT_ERR_HIP;
  int panels = 0;
  HIPCHK(lu_factor<float>(as<float>(ctx->a32), n, lda, lw, LU_NB, s, &panels));
  if (int rc = solve_all<float>(ctx, as<float>(ctx->a32), n, lda, lw.ipiv, as<float>(ctx->b32), (int)nrhs, true))
    return rc;
  hipLaunchKernelGGL(upcast_kernel, dim3(grid_cap(2 * n * nrhs)), dim3(256), 0, s, (const float*)as<float>(ctx->b32),
                     as<double>(ctx->x64), 2 * n * nrhs);
  HIPCHK(hipGetLastError());
  Refine R;
  if (int rc = refine_loop(ctx, as<float>(ctx->a32), lda, lw.ipiv, as<double>(ctx->a64), lda, as<double>(ctx->b64),
                           as<double>(ctx->x64), n, (int)nrhs, max_steps, rtol, true, R))
    return rc;
  HIPCHK(ensure(ctx->pout, sizeof(float) * 2 * n * nrhs));
  hipLaunchKernelGGL((cols_out_kernel<double, float>), dim3(blocks_for(n * nrhs)), dim3(256), 0, s,
                     (const double*)as<double>(ctx->x64), n, (int)nrhs, as<float>(ctx->pout));
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(x_out, ctx->pout.p, sizeof(float) * 2 * n * nrhs, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  int info = 0;
  if (int rc = lu_outcome(ctx, lw, &info)) return rc;
  fill_report(rep, R, rtol, (int)nrhs);
  return BEAT_OK;
}

}  // extern "C"

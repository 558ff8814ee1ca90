// Blocked right-looking complex LU with partial pivoting on gfx950, and the
// multi-RHS triangular solves that use it.
//
// Storage: row-major interleaved complex, leading dimension lda. Per panel of nb
// columns:
//   panel_kernel     cooperative grid; each workgroup holds a block of panel rows
//                    in LDS; per column a grid-wide FP64 |re|+|im| pivot search
//                    (the LAPACK izamax measure), swap, scale and rank-1 update
//   laswp_kernel     row swaps of the trailing columns
//   gemm kernels     U12 = L11^-1 A12 with the leaf inverse the panel formed,
//                    and A22 -= L21 U12, as real GEMMs on MFMA: complex operands
//                    are expanded to 2x2 real blocks [[br, bi], [-bi, br]]
//                    while staged (the 4-real-multiply complex product)
// The factorization and solve kernels raise their wave priority above the
// assembly kernels' (which a sweep runs beside them); the panel, the critical
// path, raises its own higher still.
// Columns left of the panel are never swapped; the forward solve applies each
// panel's swaps to the right-hand sides just before that panel's block, which is
// equivalent to LAPACK's getrs on fully swapped factors.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <type_traits>
#include <vector>

#include "bem_common.h"
#include "kernels.h"

namespace beat {

namespace {

constexpr int PNB = 64;   // max panel width
constexpr int RPW = 64;   // panel rows per workgroup
constexpr int MAXBW = LU_BLOCK_COLS; // outer block width (columns factored before the wide trailing update)
constexpr int SLOT_WORDS = 16;  // one exchange slot per 128-byte line

template <typename T>
struct PanelArgs {
  T* A;
  int64_t n, lda, j0;
  int nb;
  int* ipiv;
  unsigned long long* slots;  // 2 x maxg slots of SLOT_WORDS words: (tag << 32 | row), tagged value
  unsigned long long* slot_rows;  // 2 x maxg x PNB tagged elements
  unsigned long long* slot_diag;  // 2 x PNB tagged elements
  int maxg;
  int* info;
  int* ctl;  // wait control: [0] abort word, [1] give-up record (see spin_expired)
  T* linv;  // PNB x PNB row-major: inverse of the leaf's unit-lower diagonal block
};

__device__ __forceinline__ void st_u64(unsigned long long* p, unsigned long long v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned long long ld_u64(const unsigned long long* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// complex element store/load through 64-bit atomics (float complex = one word)
__device__ __forceinline__ void st_cplx(float* p, float re, float im) {
  st_u64(reinterpret_cast<unsigned long long*>(p), ((unsigned long long)__float_as_uint(im) << 32) | __float_as_uint(re));
}
__device__ __forceinline__ void ld_cplx(const float* p, float& re, float& im) {
  const unsigned long long v = ld_u64(reinterpret_cast<const unsigned long long*>(p));
  re = __uint_as_float((unsigned)(v & 0xffffffffu));
  im = __uint_as_float((unsigned)(v >> 32));
}
__device__ __forceinline__ void st_cplx(double* p, double re, double im) {
  st_u64(reinterpret_cast<unsigned long long*>(p), (unsigned long long)__double_as_longlong(re));
  st_u64(reinterpret_cast<unsigned long long*>(p) + 1, (unsigned long long)__double_as_longlong(im));
}
__device__ __forceinline__ void ld_cplx(const double* p, double& re, double& im) {
  re = __longlong_as_double((long long)ld_u64(reinterpret_cast<const unsigned long long*>(p)));
  im = __longlong_as_double((long long)ld_u64(reinterpret_cast<const unsigned long long*>(p) + 1));
}
__device__ __forceinline__ void st_flag(int* p, int v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
__device__ __forceinline__ int ld_flag(const int* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }

// Sum over the four lanes of a quad (DPP quad permutations [1,0,3,2] and
// [2,3,0,1]; every lane receives the total).
template <int CTRL>
__device__ __forceinline__ float dpp_move(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xF, 0xF, false));
}
template <int CTRL>
__device__ __forceinline__ double dpp_move(double v) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_mov_dpp((int)(unsigned)(b & 0xffffffffull), CTRL, 0xF, 0xF, false);
  const unsigned hi = (unsigned)__builtin_amdgcn_mov_dpp((int)(unsigned)(b >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
template <typename T>
__device__ __forceinline__ T quad_sum(T v) {
  v += dpp_move<0xB1>(v);
  return v + dpp_move<0x4E>(v);
}

// Slot u (0..15) of column quarter kq in a PNB-wide block: pairs of adjacent
// columns interleaved over the four quarters, c = 8 (u / 2) + 2 kq + u % 2, so that
// the four threads of a row read 64 (complex64) or 128 (complex128) contiguous
// bytes per vector load instead of each touching its own cache line.
__device__ __forceinline__ int qcol(int u, int kq) { return 8 * (u >> 1) + 2 * kq + (u & 1); }

// This thread's 16 entries of a block row in slot order (row: the block's first
// column, 16-byte aligned); columns at or beyond ncols, and all when !ok, are zero.
__device__ __forceinline__ void load_row16(const float* row, int ncols, int kq, bool ok, float* vr, float* vi) {
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int c = 8 * j + 2 * kq;
    float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
    if (ok && c + 1 < ncols) {
      v = *reinterpret_cast<const float4*>(row + 2 * c);
    } else if (ok && c < ncols) {
      v.x = row[2 * c];
      v.y = row[2 * c + 1];
    }
    vr[2 * j] = v.x;
    vi[2 * j] = v.y;
    vr[2 * j + 1] = v.z;
    vi[2 * j + 1] = v.w;
  }
}
__device__ __forceinline__ void load_row16(const double* row, int ncols, int kq, bool ok, double* vr, double* vi) {
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int c = 8 * j + 2 * kq;
    double2 v0 = make_double2(0.0, 0.0), v1 = make_double2(0.0, 0.0);
    if (ok && c < ncols) v0 = *reinterpret_cast<const double2*>(row + 2 * c);
    if (ok && c + 1 < ncols) v1 = *reinterpret_cast<const double2*>(row + 2 * c + 2);
    vr[2 * j] = v0.x;
    vi[2 * j] = v0.y;
    vr[2 * j + 1] = v1.x;
    vi[2 * j + 1] = v1.y;
  }
}

// Bounded cross-workgroup waits. Every spin loop of the factorization and of the
// triangular solves counts its rounds; every 64th round it reads the abort word
// ctl[0] and the real-time counter (s_memrealtime, 100 MHz). A wait gives up when
// the abort word is set or after SPIN_LIMIT_TICKS (4 s, far beyond any wait of a
// resident grid, whose longest is the placement of a workgroup behind finite
// kernels of other streams); a wait that times out records its kernel in ctl[1]
// and sets the abort word, so that every other wait of the call gives up within
// 64 rounds. The kernels then leave their loops together and end normally; the
// host reads ctl after the call and reports the abandoned wait as an error.
constexpr unsigned long long SPIN_LIMIT_TICKS = 400000000ull;
enum SpinSite : int { SPIN_PANEL_SLOT = 1, SPIN_PANEL_ROW = 2, SPIN_TRI_FWD = 4, SPIN_TRI_BWD = 8 };
struct Spin {
  unsigned long long t0 = 0;
  unsigned n = 0;
};
__device__ __forceinline__ bool spin_expired(int* ctl, Spin& sp, int site) {
  if ((++sp.n & 63u) != 0) return false;
  if (ld_flag(ctl) != 0) return true;
  const unsigned long long t = __builtin_amdgcn_s_memrealtime();
  if (sp.t0 == 0) {
    sp.t0 = t;
    return false;
  }
  if (t - sp.t0 < SPIN_LIMIT_TICKS) return false;
  __hip_atomic_fetch_or(ctl + 1, site, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  st_flag(ctl, 1);
  return true;
}

// Panel rows published to other workgroups: every 32-bit half of a complex
// element's parts travels in its own 64-bit word as (tag << 32 | bits), TW words
// per element (complex64 2, complex128 4). A reader checks the tag of each word it
// loads, so the writer needs no fence between the row and the slot announcing it.
template <typename T>
struct TaggedRow {
  static constexpr int TW = sizeof(T) / 2;
};
__device__ __forceinline__ void st_tagged(unsigned long long* q, float re, float im, unsigned tag) {
  const unsigned long long t = (unsigned long long)tag << 32;
  st_u64(q, t | __float_as_uint(re));
  st_u64(q + 1, t | __float_as_uint(im));
}
__device__ __forceinline__ void st_tagged(unsigned long long* q, double re, double im, unsigned tag) {
  const unsigned long long t = (unsigned long long)tag << 32;
  const unsigned long long a = (unsigned long long)__double_as_longlong(re), b = (unsigned long long)__double_as_longlong(im);
  st_u64(q, t | (a & 0xffffffffull));
  st_u64(q + 1, t | (a >> 32));
  st_u64(q + 2, t | (b & 0xffffffffull));
  st_u64(q + 3, t | (b >> 32));
}
// The element of TW loaded words (TaggedRow), zero words giving zero.
__device__ __forceinline__ void untag(const unsigned long long* w, float& re, float& im) {
  re = __uint_as_float((unsigned)w[0]);
  im = __uint_as_float((unsigned)w[1]);
}
__device__ __forceinline__ void untag(const unsigned long long* w, double& re, double& im) {
  re = __longlong_as_double((long long)((w[1] << 32) | (w[0] & 0xffffffffull)));
  im = __longlong_as_double((long long)((w[3] << 32) | (w[2] & 0xffffffffull)));
}
// Loads one published element, re-reading until every word carries the tag;
// false when the wait gave up (the element is then zero).
__device__ __forceinline__ bool ld_tagged(const unsigned long long* q, float& re, float& im, unsigned tag, int* ctl,
                                          int site = SPIN_PANEL_ROW) {
  Spin sp;
  for (;;) {
    const unsigned long long a = ld_u64(q), b = ld_u64(q + 1);
    if ((unsigned)(a >> 32) == tag && (unsigned)(b >> 32) == tag) {
      re = __uint_as_float((unsigned)a);
      im = __uint_as_float((unsigned)b);
      return true;
    }
    if (spin_expired(ctl, sp, site)) {
      re = im = 0.f;
      return false;
    }
    __builtin_amdgcn_s_sleep(1);
  }
}
__device__ __forceinline__ bool ld_tagged(const unsigned long long* q, double& re, double& im, unsigned tag, int* ctl,
                                          int site = SPIN_PANEL_ROW) {
  Spin sp;
  for (;;) {
    const unsigned long long a = ld_u64(q), b = ld_u64(q + 1), c = ld_u64(q + 2), d = ld_u64(q + 3);
    if ((unsigned)(a >> 32) == tag && (unsigned)(b >> 32) == tag && (unsigned)(c >> 32) == tag &&
        (unsigned)(d >> 32) == tag) {
      re = __longlong_as_double((long long)((b << 32) | (a & 0xffffffffull)));
      im = __longlong_as_double((long long)((d << 32) | (c & 0xffffffffull)));
      return true;
    }
    if (spin_expired(ctl, sp, site)) {
      re = im = 0.0;
      return false;
    }
    __builtin_amdgcn_s_sleep(1);
  }
}

// Wavefront reductions of 32-bit keys (full wavefront active): DPP exchanges
// within each row of 16 lanes (quad_perm [1,0,3,2], quad_perm [2,3,0,1],
// row_half_mirror, row_mirror), then the four row results by readlane. The
// result is wavefront-uniform.
__device__ __forceinline__ unsigned umax32(unsigned a, unsigned b) { return a > b ? a : b; }
__device__ __forceinline__ unsigned umin32(unsigned a, unsigned b) { return a < b ? a : b; }
template <bool MAX>
__device__ __forceinline__ unsigned wave_reduce_u32(unsigned x) {
  auto op = [](unsigned a, unsigned b) { return MAX ? umax32(a, b) : umin32(a, b); };
  x = op(x, (unsigned)__builtin_amdgcn_mov_dpp((int)x, 0xB1, 0xF, 0xF, false));
  x = op(x, (unsigned)__builtin_amdgcn_mov_dpp((int)x, 0x4E, 0xF, 0xF, false));
  x = op(x, (unsigned)__builtin_amdgcn_mov_dpp((int)x, 0x141, 0xF, 0xF, false));
  x = op(x, (unsigned)__builtin_amdgcn_mov_dpp((int)x, 0x140, 0xF, 0xF, false));
  const unsigned a = (unsigned)__builtin_amdgcn_readlane((int)x, 0);
  const unsigned b = (unsigned)__builtin_amdgcn_readlane((int)x, 16);
  const unsigned c = (unsigned)__builtin_amdgcn_readlane((int)x, 32);
  const unsigned d = (unsigned)__builtin_amdgcn_readlane((int)x, 48);
  return op(op(a, b), op(c, d));
}

// Pivot choice over the wavefront: the largest non-negative value v among the
// lanes with valid set, ties to the smallest row (LAPACK's izamax takes the
// first maximum). Compares the high words, then the low words of the tied
// lanes, then takes the smallest row among exact ties. Returns row -1 when no
// lane is valid.
__device__ __forceinline__ void wave_pick(double v, int row, bool valid, double& vbest, int& rbest) {
  const unsigned long long bits = (unsigned long long)__double_as_longlong(v);
  const unsigned h = valid ? (unsigned)(bits >> 32) + 1u : 0u;
  const unsigned H = wave_reduce_u32<true>(h);
  const unsigned lo = h == H ? (unsigned)bits : 0u;
  const unsigned Lo = wave_reduce_u32<true>(lo);
  const bool tie = h == H && (unsigned)bits == Lo && valid;
  const unsigned R = wave_reduce_u32<false>(tie ? (unsigned)row : 0xffffffffu);
  rbest = H == 0 ? -1 : (int)R;
  vbest = __longlong_as_double((long long)(((unsigned long long)(H - 1u) << 32) | Lo));
}

// Four wavefronts per workgroup; lane r of every wavefront stands for panel row
// r0 + r. The panel block is kept column-major in LDS (col[c][r], rows padded to
// RPW + 1) so that the column operations are bank-conflict free. Column jj runs
// in three phases separated by workgroup barriers:
//  A1  wavefront 0 applies column jj - 1's elimination to column jj, picks the
//      workgroup's pivot candidate and copies the candidate row and the diagonal
//      row into registers;
//  A2  wavefront 0 publishes the copies and runs the cross-workgroup exchange
//      while wavefronts 1-3 apply column jj - 1's elimination to the columns
//      > jj (so the elimination is hidden behind the exchange latency). The
//      copies lack that elimination in their columns > jj; the readers complete
//      them from the multiplier in the copy and the previous pivot row;
//  B   wavefront 0 exchanges the rows, forms the multipliers of column jj and
//      keeps the pivot row for the next step.
//
// Cross-workgroup exchange per column, without a shared counter: every
// workgroup writes its candidate row (and the diagonal row, if it owns it) as
// tagged words, then the slot's two words: the candidate's FP64 magnitude as two
// tagged 32-bit halves (tag << 32 | half), so candidates of different workgroups
// compare exactly, as izamax compares them (-1: no candidate). Readers poll both
// words of every slot until both carry the column's tag and take the first
// maximum in workgroup order, then load the winning row, re-reading words until
// each carries the tag. The row index itself never travels: only the winning
// workgroup needs it (it is its own candidate), and that workgroup records the
// pivot and swaps with the diagonal row. Every poll is
// bounded (spin_expired): a wait that gives up sets the workgroup's abort flag in
// LDS, and after the phase's barrier all four wavefronts leave the column loop
// together. Slots sit on their own 128-byte lines and alternate between two
// buffers by column parity; a workgroup can only overwrite a buffer after every
// workgroup has published the following column, that is after all reads of the
// buffer ended. Slot words and published rows are cleared before each
// factorization, and the tag of a column is its index + 1.
// Leaf inverse: complex64 builds it during the factorization (wavefront 1 forms
// row jj - 1 of X = L11^-1 in phase A2 of column jj, from row jj - 1 of L, final
// once column jj - 1's swap is done), in its own LDS array; complex128 keeps the
// LDS for occupancy and inverts after the last column.
template <typename T>
struct PanelInv {
  static constexpr bool inc = sizeof(T) == 4;
};

template <typename T>
__global__ __launch_bounds__(256) void panel_kernel(const PanelArgs<T> a) {
  constexpr int LDR = RPW + 1;
  constexpr int U = 4;  // unroll of the LDS-bound loops (independent reads in flight)
  constexpr bool INC = PanelInv<T>::inc;
  constexpr int TW = TaggedRow<T>::TW;
  __shared__ T col[PNB * LDR * 2];
  __shared__ T piv[PNB * 2];
  __shared__ T part[4 * PNB * 2];
  __shared__ T xs[INC ? PNB * PNB * 2 : 2];  // X[r][m], lane m
  __shared__ int gave_up;  // a cross-workgroup wait of this workgroup gave up
  // the panel is the factorization's critical path: its waves issue first on
  // CUs shared with trailing-update GEMMs
  __builtin_amdgcn_s_setprio(3);
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, g = blockIdx.x, G = gridDim.x;
  const int64_t r0 = a.j0 + (int64_t)g * RPW;
  const int nr = (int)min((int64_t)RPW, a.n - r0);
  const int nb = a.nb;
  if (tid == 0) gave_up = 0;
  // stage the block: lane = column, wavefront w takes rows w, w + 4, ...
  for (int rb = wv; rb < nr; rb += 16) {
    T vr[4], vi[4];
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (rb + 4 * u < nr && lane < nb) {
        const T* src = a.A + 2 * ((r0 + rb + 4 * u) * a.lda + a.j0 + lane);
        vr[u] = src[0];
        vi[u] = src[1];
      }
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (rb + 4 * u < nr && lane < nb) {
        col[2 * (lane * LDR + rb + 4 * u)] = vr[u];
        col[2 * (lane * LDR + rb + 4 * u) + 1] = vi[u];
      }
  }
  __syncthreads();
  // X row r = row r of L11^-1: X(r,m) = [m == r] - sum_{k<r} L(r,k) X(k,m), lane m
  auto inverse_row = [&](int r) {
    T sr = 0, si = 0;
    int k = 0;
    for (; k + 8 <= r; k += 8) {
      T lr[8], li[8], xr[8], xi[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        lr[u] = col[2 * ((k + u) * LDR + r)];
        li[u] = col[2 * ((k + u) * LDR + r) + 1];
        xr[u] = xs[2 * ((k + u) * PNB + lane)];
        xi[u] = xs[2 * ((k + u) * PNB + lane) + 1];
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        sr += lr[u] * xr[u] - li[u] * xi[u];
        si += lr[u] * xi[u] + li[u] * xr[u];
      }
    }
    for (; k < r; ++k) {
      const T lr = col[2 * (k * LDR + r)], li = col[2 * (k * LDR + r) + 1];
      const T xr = xs[2 * (k * PNB + lane)], xi = xs[2 * (k * PNB + lane) + 1];
      sr += lr * xr - li * xi;
      si += lr * xi + li * xr;
    }
    xs[2 * (r * PNB + lane)] = (lane == r ? T(1) : T(0)) - sr;
    xs[2 * (r * PNB + lane) + 1] = -si;
  };
  for (int jj = 0; jj < nb; ++jj) {
    const int64_t cg = a.j0 + jj;
    const unsigned tag = (unsigned)(cg + 1);
    const int par = jj & 1;
    const bool own_diag = cg >= r0 && cg < r0 + nr;
    // column jj - 1's elimination (rows >= cg) is applied during this step: a
    // zero pivot leaves a zero multiplier column, so it is skipped
    const bool upd = jj > 0 && !(piv[2 * (jj - 1)] == T(0) && piv[2 * (jj - 1) + 1] == T(0));
    const bool below = lane < nr && r0 + lane >= cg;
    // A1 (wavefront 0): bring column jj up to date, pick this workgroup's
    // candidate, and copy the candidate row and the diagonal row (lane = column)
    // before the other wavefronts touch them. Columns > jj of the copies still
    // lack column jj - 1's elimination.
    T cr = 0, ci = 0, dr = 0, di = 0;
    int row = -1;
    double v = 0.0;
    if (wv == 0) {
      T xr = 0, xi = 0;
      if (below) {
        xr = col[2 * (jj * LDR + lane)];
        xi = col[2 * (jj * LDR + lane) + 1];
        if (upd) {
          const T lr = col[2 * ((jj - 1) * LDR + lane)], li = col[2 * ((jj - 1) * LDR + lane) + 1];
          const T ur = piv[2 * jj], ui = piv[2 * jj + 1];
          xr -= lr * ur - li * ui;
          xi -= lr * ui + li * ur;
          col[2 * (jj * LDR + lane)] = xr;
          col[2 * (jj * LDR + lane) + 1] = xi;
        }
      }
      wave_pick(below ? fabs((double)xr) + fabs((double)xi) : 0.0, (int)(r0 + lane), below, v, row);
      const int lr = (int)(row - r0), ld = (int)(cg - r0);
      const T sxr = __shfl(xr, lr, 64), sxi = __shfl(xi, lr, 64);
      const T dxr = __shfl(xr, own_diag ? ld : 0, 64), dxi = __shfl(xi, own_diag ? ld : 0, 64);
      if (lane < nb) {
        cr = lane == jj ? sxr : col[2 * (lane * LDR + lr)];
        ci = lane == jj ? sxi : col[2 * (lane * LDR + lr) + 1];
        if (own_diag) {
          dr = lane == jj ? dxr : col[2 * (lane * LDR + ld)];
          di = lane == jj ? dxi : col[2 * (lane * LDR + ld) + 1];
        }
      }
    }
    __syncthreads();
    // A2: wavefront 0 runs the cross-workgroup exchange while wavefronts 1-3
    // apply column jj - 1's elimination to the columns > jj (and, complex64,
    // wavefront 1 forms row jj - 1 of the leaf inverse).
    T pr = 0, pim = 0, qr = 0, qi = 0;
    int pw = -1;                   // the workgroup holding the pivot row
    bool mine = false, swap = true;  // this workgroup holds it; it is not row cg
    if (wv == 0) {
      if (lane < nb) st_tagged(a.slot_rows + TW * (((int64_t)par * a.maxg + g) * PNB + lane), cr, ci, tag);
      if (own_diag && lane < nb) st_tagged(a.slot_diag + TW * (par * PNB + lane), dr, di, tag);
      if (lane == 0) {
        // no candidate in this workgroup (every row above the diagonal): -1
        const double vp = row >= 0 ? v : -1.0;
        unsigned long long* slot = a.slots + SLOT_WORDS * (par * a.maxg + g);
        const unsigned long long vb = (unsigned long long)__double_as_longlong(vp), tt = (unsigned long long)tag << 32;
        st_u64(slot, tt | (vb & 0xffffffffull));
        st_u64(slot + 1, tt | (vb >> 32));
      }
      // gather every slot of this column: lanes poll disjoint slots (up to four
      // each per round); every round re-reads all of a lane's pending slots at
      // once, so one round trip covers them all. A lane visits its slots in
      // increasing workgroup order and keeps the first maximum; workgroups hold
      // increasing row ranges, so the smallest workgroup among equal maxima holds
      // the first maximum row, as izamax chooses
      double bv = -1.0;
      int bw = -1;
      bool quit = false;
      for (int base = lane; base < G && !quit; base += 4 * 64) {
        bool pend[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) pend[u] = base + 64 * u < G;
        bool any = true;
        Spin sp;
        while (any) {
          unsigned long long lo[4], hi[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            lo[u] = hi[u] = 0;
            if (pend[u]) {
              const unsigned long long* slot = a.slots + SLOT_WORDS * (par * a.maxg + base + 64 * u);
              lo[u] = ld_u64(slot);
              hi[u] = ld_u64(slot + 1);
            }
          }
          any = false;
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            if (!pend[u]) continue;
            if ((lo[u] >> 32) == tag && (hi[u] >> 32) == tag) {
              pend[u] = false;
              const double vi = __longlong_as_double((long long)((hi[u] << 32) | (lo[u] & 0xffffffffull)));
              if (vi >= 0.0 && (bw < 0 || vi > bv)) {
                bv = vi;
                bw = base + 64 * u;
              }
            } else {
              any = true;
            }
          }
          if (any) {
            if (spin_expired(a.ctl, sp, SPIN_PANEL_SLOT)) {
              quit = true;
              break;
            }
            __builtin_amdgcn_s_sleep(1);
          }
        }
      }
      // a lane whose wait gave up leaves the exchange to the whole wavefront
      bool stop = __ballot(quit) != 0;
      if (!stop) {
        double pv;
        wave_pick(bv, bw, bw >= 0, pv, pw);
        stop = pw < 0;  // every slot was read, so a candidate exists: pw < 0 only on a give-up
      }
      // the winning workgroup's candidate is its own row
      mine = pw == g;
      swap = !(mine && row == cg);
      if (!stop) {
        bool ok = true;
        if (lane < nb) {
          ok = ld_tagged(a.slot_rows + TW * (((int64_t)par * a.maxg + pw) * PNB + lane), pr, pim, tag, a.ctl);
          if (mine && swap && ok) ok = ld_tagged(a.slot_diag + TW * (par * PNB + lane), qr, qi, tag, a.ctl);
        }
        stop = __ballot(!ok) != 0;
      }
      if (stop && lane == 0) gave_up = 1;
      // complete the copies: x[c] -= x[jj - 1] u[c] for c > jj (u = pivot row jj - 1)
      if (upd && !stop) {
        const T plr = __shfl(pr, jj - 1, 64), pli = __shfl(pim, jj - 1, 64);
        const T qlr = __shfl(qr, jj - 1, 64), qli = __shfl(qi, jj - 1, 64);
        if (lane > jj && lane < nb) {
          const T ur = piv[2 * lane], ui = piv[2 * lane + 1];
          pr -= plr * ur - pli * ui;
          pim -= plr * ui + pli * ur;
          qr -= qlr * ur - qli * ui;
          qi -= qlr * ui + qli * ur;
        }
      }
    } else {
      if (upd && below) {
        const T lr = col[2 * ((jj - 1) * LDR + lane)], li = col[2 * ((jj - 1) * LDR + lane) + 1];
        // this wavefront's columns c > jj with c = jj + wv (mod 3)
        int c = jj + wv;
        for (; c + 3 * (U - 1) < nb; c += 3 * U) {
          T xr[U], xi[U], ur[U], ui[U];
#pragma unroll
          for (int u = 0; u < U; ++u) {
            xr[u] = col[2 * ((c + 3 * u) * LDR + lane)];
            xi[u] = col[2 * ((c + 3 * u) * LDR + lane) + 1];
            ur[u] = piv[2 * (c + 3 * u)];
            ui[u] = piv[2 * (c + 3 * u) + 1];
          }
#pragma unroll
          for (int u = 0; u < U; ++u) {
            col[2 * ((c + 3 * u) * LDR + lane)] = xr[u] - (lr * ur[u] - li * ui[u]);
            col[2 * ((c + 3 * u) * LDR + lane) + 1] = xi[u] - (lr * ui[u] + li * ur[u]);
          }
        }
        for (; c < nb; c += 3) {
          const T ur = piv[2 * c], ui = piv[2 * c + 1];
          col[2 * (c * LDR + lane)] -= lr * ur - li * ui;
          col[2 * (c * LDR + lane) + 1] -= lr * ui + li * ur;
        }
      }
      if (INC && g == 0 && wv == 1 && jj > 0 && a.linv) inverse_row(jj - 1);
    }
    __syncthreads();
    if (gave_up) break;  // uniform: read after the barrier by every wavefront
    // B (wavefront 0): row exchange, multipliers of column jj, pivot row to piv
    if (wv == 0) {
      if (lane < nb) {
        piv[2 * lane] = pr;
        piv[2 * lane + 1] = pim;
      }
      if (mine && lane == 0) a.ipiv[cg] = row;
      if (swap && lane < nb) {
        if (mine) {
          col[2 * (lane * LDR + (row - r0))] = qr;
          col[2 * (lane * LDR + (row - r0)) + 1] = qi;
        }
        if (own_diag) {
          col[2 * (lane * LDR + (cg - r0))] = pr;
          col[2 * (lane * LDR + (cg - r0)) + 1] = pim;
        }
      }
      const T vr = __shfl(pr, jj, 64), vi = __shfl(pim, jj, 64);
      const T qjr = __shfl(qr, jj, 64), qji = __shfl(qi, jj, 64);  // row p's new column-jj entry
      if (vr == T(0) && vi == T(0)) {
        if (g == 0 && lane == 0 && *a.info == 0) *a.info = (int)cg + 1;
      } else if (lane < nr && r0 + lane > cg) {
        const T den = vr * vr + vi * vi;
        const T ir = vr / den, ii = -vi / den;
        const bool at_p = mine && r0 + lane == row;
        const T ar = at_p ? qjr : col[2 * (jj * LDR + lane)], ai = at_p ? qji : col[2 * (jj * LDR + lane) + 1];
        col[2 * (jj * LDR + lane)] = ar * ir - ai * ii;
        col[2 * (jj * LDR + lane) + 1] = ar * ii + ai * ir;
      }
    }
    __syncthreads();
  }
  if (gave_up) return;  // the factors are abandoned; the host reports the give-up
  for (int r = wv; r < nr; r += 4)
    if (lane < nb) {
      T* dst = a.A + 2 * ((r0 + r) * a.lda + a.j0 + lane);
      dst[0] = col[2 * (lane * LDR + r)];
      dst[1] = col[2 * (lane * LDR + r) + 1];
    }
  if (INC) {
    if (g == 0 && a.linv) {
      if (wv == 1) inverse_row(nb - 1);
      __syncthreads();
      for (int r = wv; r < nb; r += 4)
        if (lane < nb) {
          a.linv[2 * (r * PNB + lane)] = xs[2 * (r * PNB + lane)];
          a.linv[2 * (r * PNB + lane) + 1] = xs[2 * (r * PNB + lane) + 1];
        }
    }
    return;
  }
  // The first workgroup holds the leaf's diagonal block (already stored back):
  // invert its unit-lower part in place, columns right to left as in LAPACK's
  // trti2. With the strict upper part of col cleared and its diagonal set to 1,
  // row i of X forms X(i,j) = -sum_{k=j+1}^{nb-1} col[k][i] L(k,j) over a uniform
  // k range: col[k][i] holds X(i,k) for j < k < i, 1 for k = i and 0 for k > i,
  // and L(k,j) = col[j][k] is still intact while column j is being formed.
  // Wavefront w sums the terms k = w (mod 4); wavefront 0 adds the four parts.
  if (g == 0 && a.linv) {
    const int i = lane;
    __syncthreads();
    if (wv == 0 && i < nb)
      for (int c = i; c < nb; ++c) {
        col[2 * (c * LDR + i)] = c == i ? T(1) : T(0);
        col[2 * (c * LDR + i) + 1] = T(0);
      }
    __syncthreads();
    for (int j = nb - 2; j >= 0; --j) {
      T sr = 0, si = 0;
      int k = j + 1 + ((wv - (j + 1)) & 3);
      for (; k + 4 * (U - 1) < nb; k += 4 * U) {
        T xr[U], xi[U], lr[U], li[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          xr[u] = col[2 * ((k + 4 * u) * LDR + i)];
          xi[u] = col[2 * ((k + 4 * u) * LDR + i) + 1];
          lr[u] = col[2 * (j * LDR + k + 4 * u)];
          li[u] = col[2 * (j * LDR + k + 4 * u) + 1];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
          sr += xr[u] * lr[u] - xi[u] * li[u];
          si += xr[u] * li[u] + xi[u] * lr[u];
        }
      }
      for (; k < nb; k += 4) {
        const T xr = col[2 * (k * LDR + i)], xi = col[2 * (k * LDR + i) + 1];
        const T lr = col[2 * (j * LDR + k)], li = col[2 * (j * LDR + k) + 1];
        sr += xr * lr - xi * li;
        si += xr * li + xi * lr;
      }
      part[2 * (wv * PNB + i)] = sr;
      part[2 * (wv * PNB + i) + 1] = si;
      __syncthreads();
      if (wv == 0 && i > j && i < nb) {
        col[2 * (j * LDR + i)] = -(part[2 * i] + part[2 * (PNB + i)] + part[2 * (2 * PNB + i)] + part[2 * (3 * PNB + i)]);
        col[2 * (j * LDR + i) + 1] =
            -(part[2 * i + 1] + part[2 * (PNB + i) + 1] + part[2 * (2 * PNB + i) + 1] + part[2 * (3 * PNB + i) + 1]);
      }
      __syncthreads();
    }
    for (int r = wv; r < nb; r += 4)
      if (lane < nb) {
        a.linv[2 * (r * PNB + lane)] = col[2 * (lane * LDR + r)];
        a.linv[2 * (r * PNB + lane) + 1] = col[2 * (lane * LDR + r) + 1];
      }
  }
}

// Row exchanges ipiv[k] <-> k for k in [k0, k1), in order (k1 - k0 <= MAXBW),
// on columns [c0, c1), as one permutation: the rows that change are k0..k1-1 and
// the targets ipiv[k] >= k1; the final content of row r is the original row
// src(r) found by applying the exchanges to r in reverse order. Each workgroup
// composes the list itself, then gathers all moved values of its 16 columns into
// registers before any is written back (so the in-place update is race-free).
constexpr int SWP_COLS = 16;
template <typename T>
__global__ __launch_bounds__(256) void laswp_kernel(T* A, int64_t lda, int64_t c0, int64_t c1, const int* ipiv,
                                                    int64_t k0, int64_t k1) {
  // in a sweep the factorization and solves run beside the next system's
  // assembly: their waves issue ahead of the assembly's
  __builtin_amdgcn_s_setprio(2);
  __shared__ int sdst[2 * MAXBW], ssrc[2 * MAXBW];
  __shared__ int4 spiv[MAXBW / 4];
  const int tid = threadIdx.x;
  const int K = (int)(k1 - k0);
  // the exchanges, padded to a multiple of 4 with no-ops (p = k)
  for (int k = tid; k < MAXBW; k += 256) reinterpret_cast<int*>(spiv)[k] = k < K ? ipiv[k0 + k] : (int)(k0 + k);
  __syncthreads();
  // thread t composes entries t and t + 256 (rows k0 + t, and the target of
  // exchange t when it lies beyond k1), walking the exchanges in reverse
  int q[2], dst[2];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int e = tid + 256 * h;
    dst[h] = -1;
    if (e < K) dst[h] = (int)(k0 + e);
    else if (e >= MAXBW && e - MAXBW < K) {
      const int p = reinterpret_cast<const int*>(spiv)[e - MAXBW];
      if (p >= k1) dst[h] = p;
    }
    q[h] = dst[h];
  }
  for (int k4 = (K + 3) / 4 - 1; k4 >= 0; --k4) {
    const int4 p4 = spiv[k4];
    const int ps[4] = {p4.x, p4.y, p4.z, p4.w};
#pragma unroll
    for (int u = 3; u >= 0; --u) {
      const int k = (int)k0 + 4 * k4 + u;
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        if (q[h] == k) q[h] = ps[u];
        else if (q[h] == ps[u]) q[h] = k;
      }
    }
  }
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int e = tid + 256 * h;
    sdst[e] = q[h] == dst[h] ? -1 : dst[h];  // unchanged rows are skipped
    ssrc[e] = q[h];
  }
  __syncthreads();
  const int cl = tid % SWP_COLS, eg = tid / SWP_COLS;
  const int64_t c = c0 + (int64_t)blockIdx.x * SWP_COLS + cl;
  const bool act = c < c1;
  constexpr int NE = 2 * MAXBW / (256 / SWP_COLS);
  T vr[NE], vi[NE];
#pragma unroll
  for (int j = 0; j < NE; ++j) {
    const int e = eg + j * (256 / SWP_COLS);
    vr[j] = vi[j] = T(0);
    if (act && sdst[e] >= 0) {
      const T* s = A + 2 * ((int64_t)ssrc[e] * lda + c);
      vr[j] = s[0];
      vi[j] = s[1];
    }
  }
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
#pragma unroll
  for (int j = 0; j < NE; ++j) {
    const int e = eg + j * (256 / SWP_COLS);
    if (act && sdst[e] >= 0) {
      T* d = A + 2 * ((int64_t)sdst[e] * lda + c);
      d[0] = vr[j];
      d[1] = vi[j];
    }
  }
}

typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef double f64x4 __attribute__((ext_vector_type(4)));

// C[m x n] (-)= A[m x kc] B[kc x n] for row-major interleaved complex matrices
// (leading dimensions in complex elements), as one real GEMM on MFMA: the right
// operand's complex entries are expanded to 2x2 real blocks [[br, bi], [-bi, br]]
// when read from LDS. Workgroup tile BM rows x 64 complex columns, 4 waves of
// BM/2 x 64 real outputs; K staged through LDS in chunks of KB complex with the
// next chunk prefetched into registers while the current one is multiplied.
//   complex64:  v_mfma_f32_32x32x2_f32, lane l holds A[l%32][l/32], B[l/32][l%32],
//               output item v of lane l is D[8(v/4) + 4(l/32) + v%4][l%32]
//   complex128: v_mfma_f64_16x16x4_f64, lane l holds A[l%16][l/16], B[l/16][l%16],
//               output item v of lane l is D[l/16 + 4v][l%16]
// G3 (complex64): the three-real-product form of the complex product (Gauss):
// T1 = Ar Br, T2 = Ai Bi, T3 = (Ar + Ai)(Br + Bi), then Cr = T1 - T2 and
// Ci = T3 - T1 - T2. Three 32x32x2 products per two steps of k instead of four;
// normwise stable, with Ci's error bounded by |A||B| rather than componentwise.
template <typename T>
struct GemmCfg;
template <>
struct GemmCfg<float> {
  static constexpr int VEC = 4;  // scalars per 16-byte vector
};
template <>
struct GemmCfg<double> {
  static constexpr int VEC = 2;
};

template <typename T, bool SET, int BM, int KB, bool G3 = false>
// B may alias C (U12 = L11^-1 A12 in place): every workgroup then reads all of
// its rows' K before its epilogue writes them, which needs the whole of m in one
// row tile (m <= BM, checked by gemm_launch), and B carries no __restrict__.
__device__ __forceinline__ void zgemm_tile(const T* __restrict__ A, int64_t lda, const T* B, int64_t ldb, T* C,
                                           int64_t ldc, int64_t m, int64_t n, int kc, int64_t row0, int64_t col0) {
  constexpr int BN = 64;
  constexpr int VEC = GemmCfg<T>::VEC;
  constexpr int AP = 2 * KB + VEC;  // padded row lengths in scalars
  constexpr int BP = 2 * BN + VEC;
  constexpr int AV = BM * 2 * KB / VEC / 256;  // vectors per thread
  constexpr int BV = KB * 2 * BN / VEC / 256;
  typedef T vec __attribute__((ext_vector_type(VEC)));
  __shared__ __attribute__((aligned(16))) T As[BM * AP];
  __shared__ __attribute__((aligned(16))) T Bs[KB * BP];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  vec ra[AV], rb[BV];

  // whole tiles of 16-byte aligned rows (the factorization's case): vector loads
  // with no per-vector checks
  const bool whole_rows = row0 + BM <= m && col0 + BN <= n && ((uintptr_t)A & 15) == 0 &&
                          ((uintptr_t)B & 15) == 0 && (lda & (VEC / 2 - 1)) == 0 && (ldb & (VEC / 2 - 1)) == 0;
  auto load = [&](int k0) {
    if (whole_rows && k0 + KB <= kc) {
#pragma unroll
      for (int it = 0; it < AV; ++it) {
        const int idx = tid + 256 * it;
        const int r = idx / (2 * KB / VEC), cv = idx % (2 * KB / VEC);
        ra[it] = *reinterpret_cast<const vec*>(A + 2 * ((row0 + r) * lda + k0 + cv * VEC / 2));
      }
#pragma unroll
      for (int it = 0; it < BV; ++it) {
        const int idx = tid + 256 * it;
        const int r = idx / (2 * BN / VEC), cv = idx % (2 * BN / VEC);
        rb[it] = *reinterpret_cast<const vec*>(B + 2 * ((int64_t)(k0 + r) * ldb + col0 + cv * VEC / 2));
      }
      return;
    }
#pragma unroll
    for (int it = 0; it < AV; ++it) {
      const int idx = tid + 256 * it;
      const int r = idx / (2 * KB / VEC), cv = idx % (2 * KB / VEC);
      const int64_t gr = row0 + r;
      const int kcx = k0 + cv * VEC / 2;  // first complex column of this vector
      const T* src = A + 2 * (gr * lda + kcx);
      vec v;
      if (gr < m && kcx + VEC / 2 <= kc && ((uintptr_t)src & 15) == 0) {
        v = *reinterpret_cast<const vec*>(src);
      } else {
#pragma unroll
        for (int e = 0; e < VEC; ++e) v[e] = (gr < m && kcx + e / 2 < kc) ? src[e] : T(0);
      }
      ra[it] = v;
    }
#pragma unroll
    for (int it = 0; it < BV; ++it) {
      const int idx = tid + 256 * it;
      const int r = idx / (2 * BN / VEC), cv = idx % (2 * BN / VEC);
      const int gk = k0 + r;
      const int64_t gc = col0 + cv * VEC / 2;
      const T* src = B + 2 * ((int64_t)gk * ldb + gc);
      vec v;
      if (gk < kc && gc + VEC / 2 <= n && ((uintptr_t)src & 15) == 0) {
        v = *reinterpret_cast<const vec*>(src);
      } else {
#pragma unroll
        for (int e = 0; e < VEC; ++e) v[e] = (gk < kc && gc + e / 2 < n) ? src[e] : T(0);
      }
      rb[it] = v;
    }
  };
  auto stash = [&]() {
#pragma unroll
    for (int it = 0; it < AV; ++it) {
      const int idx = tid + 256 * it;
      const int r = idx / (2 * KB / VEC), cv = idx % (2 * KB / VEC);
      *reinterpret_cast<vec*>(&As[r * AP + cv * VEC]) = ra[it];
    }
#pragma unroll
    for (int it = 0; it < BV; ++it) {
      const int idx = tid + 256 * it;
      const int r = idx / (2 * BN / VEC), cv = idx % (2 * BN / VEC);
      *reinterpret_cast<vec*>(&Bs[r * BP + cv * VEC]) = rb[it];
    }
  };

  constexpr int WR = BM / 2;  // rows per wave (two waves along the rows, two along the columns)
  if constexpr (sizeof(T) == 4 && G3) {
    // each wave: WR rows (MI blocks of 32) x 32 complex columns of T1, T2, T3;
    // lane (h, l) takes step kk + h of k, row l of each block and complex column l
    static_assert(KB % 2 == 0, "G3 tiles: even K chunks");
    constexpr int MI = WR / 32;
    f32x16 t1[MI], t2[MI], t3[MI];
#pragma unroll
    for (int mi = 0; mi < MI; ++mi)
#pragma unroll
      for (int v = 0; v < 16; ++v) t1[mi][v] = t2[mi][v] = t3[mi][v] = 0.f;
    const int h = lane >> 5, l32 = lane & 31;
    const int jc = wn * 32 + l32;
    load(0);
    for (int k0 = 0; k0 < kc; k0 += KB) {
      stash();
      __syncthreads();
      if (k0 + KB < kc) load(k0 + KB);
#pragma unroll
      for (int kk = 0; kk < KB; kk += 2) {
        const int kx = kk + h;
        const float2 bv = *reinterpret_cast<const float2*>(&Bs[kx * BP + 2 * jc]);
        const float bs = bv.x + bv.y;
#pragma unroll
        for (int mi = 0; mi < MI; ++mi) {
          const float2 av = *reinterpret_cast<const float2*>(&As[(wm * WR + mi * 32 + l32) * AP + 2 * kx]);
          t1[mi] = __builtin_amdgcn_mfma_f32_32x32x2f32(av.x, bv.x, t1[mi], 0, 0, 0);
          t2[mi] = __builtin_amdgcn_mfma_f32_32x32x2f32(av.y, bv.y, t2[mi], 0, 0, 0);
          t3[mi] = __builtin_amdgcn_mfma_f32_32x32x2f32(av.x + av.y, bs, t3[mi], 0, 0, 0);
        }
      }
      __syncthreads();
    }
    const int64_t gc = col0 + jc;
#pragma unroll
    for (int mi = 0; mi < MI; ++mi)
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int64_t gr = row0 + wm * WR + mi * 32 + 8 * (v >> 2) + 4 * h + (v & 3);
        if (gr < m && gc < n) {
          const float cr = t1[mi][v] - t2[mi][v], ci = t3[mi][v] - t1[mi][v] - t2[mi][v];
          float2* dst = reinterpret_cast<float2*>(C + 2 * (gr * ldc + gc));
          if (SET) {
            *dst = make_float2(cr, ci);
          } else {
            const float2 o = *dst;
            *dst = make_float2(o.x - cr, o.y - ci);
          }
        }
      }
  } else if constexpr (sizeof(T) == 4) {
    constexpr int MI = WR / 32;
    f32x16 acc[MI][2];
#pragma unroll
    for (int i = 0; i < MI; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int v = 0; v < 16; ++v) acc[i][j][v] = 0.f;
    const int h = lane >> 5, l32 = lane & 31;
    load(0);
    for (int k0 = 0; k0 < kc; k0 += KB) {
      stash();
      __syncthreads();
      if (k0 + KB < kc) load(k0 + KB);
#pragma unroll
      for (int kk = 0; kk < KB; ++kk) {
        float av[MI], bv[2];
#pragma unroll
        for (int mi = 0; mi < MI; ++mi) av[mi] = As[(wm * WR + mi * 32 + l32) * AP + 2 * kk + h];
#pragma unroll
        for (int ni = 0; ni < 2; ++ni) {
          const int j = wn * 64 + ni * 32 + l32;
          const int jc = j >> 1, part = j & 1;
          const float br = Bs[kk * BP + 2 * jc], bi = Bs[kk * BP + 2 * jc + 1];
          bv[ni] = h == 0 ? (part == 0 ? br : bi) : (part == 0 ? -bi : br);
        }
#pragma unroll
        for (int mi = 0; mi < MI; ++mi)
#pragma unroll
          for (int ni = 0; ni < 2; ++ni)
            acc[mi][ni] = __builtin_amdgcn_mfma_f32_32x32x2f32(av[mi], bv[ni], acc[mi][ni], 0, 0, 0);
      }
      __syncthreads();
    }
#pragma unroll
    for (int mi = 0; mi < MI; ++mi)
#pragma unroll
      for (int ni = 0; ni < 2; ++ni)
#pragma unroll
        for (int v = 0; v < 16; ++v) {
          const int r = wm * WR + mi * 32 + 8 * (v >> 2) + 4 * h + (v & 3);
          const int j = wn * 64 + ni * 32 + l32;
          const int64_t gr = row0 + r, gcr = 2 * col0 + j;
          if (gr < m && gcr < 2 * n) {
            if (SET) C[gr * 2 * ldc + gcr] = acc[mi][ni][v];
            else C[gr * 2 * ldc + gcr] -= acc[mi][ni][v];
          }
        }
  } else {
    constexpr int MI = WR / 16;
    f64x4 acc[MI][4];
#pragma unroll
    for (int i = 0; i < MI; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[i][j][v] = 0.0;
    const int kq = lane >> 4, l16 = lane & 15;
    load(0);
    for (int k0 = 0; k0 < kc; k0 += KB) {
      stash();
      __syncthreads();
      if (k0 + KB < kc) load(k0 + KB);
#pragma unroll 2
      for (int kk = 0; kk < KB; kk += 2) {
        const int kcx = kk + (kq >> 1), hp = kq & 1;
        double av[MI], bv[4];
#pragma unroll
        for (int mi = 0; mi < MI; ++mi) av[mi] = As[(wm * WR + mi * 16 + l16) * AP + 2 * kcx + hp];
#pragma unroll
        for (int ni = 0; ni < 4; ++ni) {
          const int j = wn * 64 + ni * 16 + l16;
          const int jc = j >> 1, part = j & 1;
          const double br = Bs[kcx * BP + 2 * jc], bi = Bs[kcx * BP + 2 * jc + 1];
          bv[ni] = hp == 0 ? (part == 0 ? br : bi) : (part == 0 ? -bi : br);
        }
#pragma unroll
        for (int mi = 0; mi < MI; ++mi)
#pragma unroll
          for (int ni = 0; ni < 4; ++ni)
            acc[mi][ni] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[mi], bv[ni], acc[mi][ni], 0, 0, 0);
      }
      __syncthreads();
    }
#pragma unroll
    for (int mi = 0; mi < MI; ++mi)
#pragma unroll
      for (int ni = 0; ni < 4; ++ni)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int r = wm * WR + mi * 16 + kq + 4 * v;
          const int j = wn * 64 + ni * 16 + l16;
          const int64_t gr = row0 + r, gcr = 2 * col0 + j;
          if (gr < m && gcr < 2 * n) {
            if (SET) C[gr * 2 * ldc + gcr] = acc[mi][ni][v];
            else C[gr * 2 * ldc + gcr] -= acc[mi][ni][v];
          }
        }
  }
}

template <typename T, bool SET, int BM, int KB, bool G3 = false>
__global__ __launch_bounds__(256) void zgemm_kernel(const T* __restrict__ A, int64_t lda, const T* B,
                                                    int64_t ldb, T* C, int64_t ldc, int64_t m, int64_t n, int kc,
                                                    int behind) {
  // in a sweep the factorization and solves run beside the next system's
  // assembly: their waves issue ahead of the assembly's; the look-ahead update
  // (behind) issues after the factorization's critical path
  if (behind) __builtin_amdgcn_s_setprio(1);
  else __builtin_amdgcn_s_setprio(2);
  zgemm_tile<T, SET, BM, KB, G3>(A, lda, B, ldb, C, ldc, m, n, kc, (int64_t)blockIdx.y * BM, (int64_t)blockIdx.x * 64);
}

// Leaf t's trailing products for every system of a batch: UPPER, U12 = L11^-1 A12
// in place over the columns [j0 + b, nc); else A22 -= L21 U12 over the same
// columns and the rows below the leaf. blockIdx.z: the system; BM x 64-complex
// tiles with K chunks of KB (the tile shape chosen per size class).
template <typename T, bool UPPER, bool G3, int BM, int KB>
__global__ __launch_bounds__(256) void zgemm_batched_kernel(const BatchSystem<T>* sys, int t) {
  const BatchSystem<T> S = sys[blockIdx.z];
  const int64_t j0 = (int64_t)t * PNB;
  if (j0 >= S.n) return;
  const int b = (int)min((int64_t)PNB, S.n - j0);
  const int64_t cols = S.nc - j0 - b, rows = UPPER ? b : S.n - j0 - b;
  const int64_t row0 = (int64_t)blockIdx.y * BM, col0 = (int64_t)blockIdx.x * 64;
  if (cols <= 0 || rows <= 0 || row0 >= rows || col0 >= cols) return;
  T* const U12 = S.A + 2 * (j0 * S.lda + j0 + b);
  if (UPPER)
    zgemm_tile<T, true, BM, KB, G3>(S.linv + 2 * (int64_t)t * PNB * PNB, PNB, U12, S.lda, U12, S.lda, rows, cols, b,
                                    row0, col0);
  else
    zgemm_tile<T, false, BM, KB, G3>(S.A + 2 * ((j0 + b) * S.lda + j0), S.lda, U12, S.lda,
                                     S.A + 2 * ((j0 + b) * S.lda + j0 + b), S.lda, rows, cols, b, row0, col0);
}

// 16 x 16 matrix-core tiles of the solves (one wavefront, K = 4 per step):
// A[i][k] in lane i + 16 k, B[k][j] in lane j + 16 k; accumulator element v of
// lane l is row mfma_row<T>(l / 16, v), column l % 16.
typedef float f32x4 __attribute__((ext_vector_type(4)));
template <typename T>
struct Acc4;
template <>
struct Acc4<float> {
  typedef f32x4 type;
};
template <>
struct Acc4<double> {
  typedef f64x4 type;
};
__device__ __forceinline__ f32x4 mfma16(float a, float b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f64x4 mfma16(double a, double b, f64x4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
template <typename T>
__device__ __forceinline__ int mfma_row(int q, int v) {
  return sizeof(T) == 4 ? 4 * q + v : q + 4 * v;
}
// 16 consecutive complex entries (p 16-byte aligned); entries at or beyond nv,
// and all when !ok, are zero (and not read)
__device__ __forceinline__ void load_seg16(const float* p, int nv, bool ok, float* vr, float* vi) {
  if (ok && nv >= 16) {  // a whole segment: no per-vector checks
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const float4 v = *reinterpret_cast<const float4*>(p + 4 * j);
      vr[2 * j] = v.x;
      vi[2 * j] = v.y;
      vr[2 * j + 1] = v.z;
      vi[2 * j + 1] = v.w;
    }
    return;
  }
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
    if (ok && 2 * j + 1 < nv) {
      v = *reinterpret_cast<const float4*>(p + 4 * j);
    } else if (ok && 2 * j < nv) {
      v.x = p[4 * j];
      v.y = p[4 * j + 1];
    }
    vr[2 * j] = v.x;
    vi[2 * j] = v.y;
    vr[2 * j + 1] = v.z;
    vi[2 * j + 1] = v.w;
  }
}
__device__ __forceinline__ void load_seg16(const double* p, int nv, bool ok, double* vr, double* vi) {
  if (ok && nv >= 16) {  // a whole segment: no per-vector checks
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      const double2 v = *reinterpret_cast<const double2*>(p + 2 * j);
      vr[j] = v.x;
      vi[j] = v.y;
    }
    return;
  }
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    double2 v = make_double2(0.0, 0.0);
    if (ok && j < nv) v = *reinterpret_cast<const double2*>(p + 2 * j);
    vr[j] = v.x;
    vi[j] = v.y;
  }
}
// Triangular substitutions of one system's vector block by one workgroup (512
// threads): forward L y = v (unit lower, below the diagonal of A), then backward
// U x = y, block of 64 rows by block. Per block the diagonal part is solved by
// one wavefront per right-hand side (lane = row, the solved entry broadcast by
// shuffle) from the block staged in LDS; the other rows then subtract the block's
// contribution on the matrix cores (update_rows).
template <typename T, bool FWD, bool BWD>
__global__ __launch_bounds__(512) void tri_wg_kernel(const BatchSystem<T>* sys, const BatchVec<T>* vec, int nrhs) {
  const BatchSystem<T> S = sys[blockIdx.x];
  const BatchVec<T> V = vec[blockIdx.x];
  if (V.skip && *V.skip) return;
  __shared__ T blk[PNB][PNB + 1][2];
  __shared__ T ys[PNB][8][2];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int64_t n = S.n, lda = S.lda, ld = V.ld;
  T* v = V.v;
  const T* A = S.A;
  const int nblk = (int)((n + PNB - 1) / PNB);
  typedef typename Acc4<T>::type acc4;
  // rows [r0, r1) subtract the block's contribution: A[rows, jb..jb+nb) ys, by
  // 16-row tiles on the matrix cores, tile u on wavefront u % 8 (lanes as in
  // tri_solve_kernel: row lane % 16, columns 16 (lane / 16) .. + 15 of the block;
  // accumulator columns: the right-hand sides' real parts, then their imaginary
  // parts; columns of ys beyond nrhs only reach accumulator columns not stored)
  auto update_rows = [&](int64_t r0, int64_t r1, int64_t jb, int nb) {
    const int l16 = lane & 15, mq = lane >> 4, d = l16 & 7, part = l16 >> 3;
    const int nv = nb - 16 * mq;
    for (int64_t tr = r0 + 16 * (int64_t)wave; tr < r1; tr += 16 * 8) {
      const int64_t ar = tr + l16;
      T mr[16], mi[16];
      load_seg16(A + 2 * ((ar < r1 ? ar : r0) * lda + jb + 16 * mq), nv, ar < r1, mr, mi);
      acc4 c;
#pragma unroll
      for (int e = 0; e < 4; ++e) c[e] = 0;
#pragma unroll
      for (int j = 0; j < 16; ++j) {
        const int k = 16 * mq + j;
        const T zr = ys[k][d][0], zi = ys[k][d][1];
        c = mfma16(mr[j], part ? zi : zr, c);
        c = mfma16(mi[j], part ? zr : -zi, c);
      }
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int64_t row = tr + mfma_row<T>(mq, e);
        if (row < r1 && d < nrhs) v[2 * (row * ld + d) + part] -= c[e];
      }
    }
  };
  if (FWD)
    for (int kb = 0; kb < nblk; ++kb) {
      const int64_t jb = (int64_t)kb * PNB;
      const int nb = (int)min((int64_t)PNB, n - jb);
      for (int e = t; e < PNB * PNB; e += 512) {
        const int i = e / PNB, j = e - i * PNB;
        const bool in = i < nb && j < i;
        blk[i][j][0] = in ? A[2 * ((jb + i) * lda + jb + j)] : T(0);
        blk[i][j][1] = in ? A[2 * ((jb + i) * lda + jb + j) + 1] : T(0);
      }
      __syncthreads();
      if (wave < nrhs) {
        const int d = wave;
        T yr = lane < nb ? v[2 * ((jb + lane) * ld + d)] : T(0), yi = lane < nb ? v[2 * ((jb + lane) * ld + d) + 1] : T(0);
        for (int j = 0; j < nb; ++j) {
          const T zr = __shfl(yr, j, 64), zi = __shfl(yi, j, 64);
          if (lane > j) {
            const T lr = blk[lane][j][0], li = blk[lane][j][1];
            yr -= lr * zr - li * zi;
            yi -= lr * zi + li * zr;
          }
        }
        if (lane < nb) {
          v[2 * ((jb + lane) * ld + d)] = yr;
          v[2 * ((jb + lane) * ld + d) + 1] = yi;
        }
        ys[lane][d][0] = lane < nb ? yr : T(0);
        ys[lane][d][1] = lane < nb ? yi : T(0);
      }
      __syncthreads();
      update_rows(jb + nb, n, jb, nb);
      __syncthreads();
    }
  if (BWD)
    for (int kb = nblk - 1; kb >= 0; --kb) {
      const int64_t jb = (int64_t)kb * PNB;
      const int nb = (int)min((int64_t)PNB, n - jb);
      for (int e = t; e < PNB * PNB; e += 512) {
        const int i = e / PNB, j = e - i * PNB;
        const bool in = i < nb && j >= i && j < nb;
        blk[i][j][0] = in ? A[2 * ((jb + i) * lda + jb + j)] : T(0);
        blk[i][j][1] = in ? A[2 * ((jb + i) * lda + jb + j) + 1] : T(0);
      }
      __syncthreads();
      if (wave < nrhs) {
        const int d = wave;
        T xr = lane < nb ? v[2 * ((jb + lane) * ld + d)] : T(0), xi = lane < nb ? v[2 * ((jb + lane) * ld + d) + 1] : T(0);
        for (int j = nb - 1; j >= 0; --j) {
          if (lane == j) {
            const T ur = blk[j][j][0], ui = blk[j][j][1];
            const T den = ur * ur + ui * ui;
            const T qr = (xr * ur + xi * ui) / den, qi = (xi * ur - xr * ui) / den;
            xr = qr;
            xi = qi;
          }
          const T zr = __shfl(xr, j, 64), zi = __shfl(xi, j, 64);
          if (lane < j) {
            const T ur = blk[lane][j][0], ui = blk[lane][j][1];
            xr -= ur * zr - ui * zi;
            xi -= ur * zi + ui * zr;
          }
        }
        if (lane < nb) {
          v[2 * ((jb + lane) * ld + d)] = xr;
          v[2 * ((jb + lane) * ld + d) + 1] = xi;
        }
        ys[lane][d][0] = lane < nb ? xr : T(0);
        ys[lane][d][1] = lane < nb ? xi : T(0);
      }
      __syncthreads();
      update_rows(0, jb, jb, nb);
      __syncthreads();
    }
}

// The trailing-update tile shapes: BM rows x 64 complex columns, K staged KB at a
// time. Shape 0 (64 x 16) is the default; the others are candidates for the
// autotuner (beat_autotune), which keeps the fastest per size class.
template <typename T, bool SET, bool G3, int BM, int KB>
static void gemm_go(const T* A, int64_t lda, const T* B, int64_t ldb, T* C, int64_t ldc, int64_t m, int64_t n, int kc,
                    hipStream_t s, bool behind) {
  dim3 grid((unsigned)((n + 63) / 64), (unsigned)((m + BM - 1) / BM));
  hipLaunchKernelGGL((zgemm_kernel<T, SET, BM, KB, G3>), grid, dim3(256), 0, s, A, lda, B, ldb, C, ldc, m, n, kc,
                     behind ? 1 : 0);
}

template <typename T, bool SET, bool G3>
static void gemm_shape(int tile, const T* A, int64_t lda, const T* B, int64_t ldb, T* C, int64_t ldc, int64_t m,
                       int64_t n, int kc, hipStream_t s, bool behind) {
  switch (tile) {
    case 1: gemm_go<T, SET, G3, 64, 32>(A, lda, B, ldb, C, ldc, m, n, kc, s, behind); break;
    case 2: gemm_go<T, SET, G3, 128, 16>(A, lda, B, ldb, C, ldc, m, n, kc, s, behind); break;
    case 3: gemm_go<T, SET, G3, 128, 32>(A, lda, B, ldb, C, ldc, m, n, kc, s, behind); break;
    default: gemm_go<T, SET, G3, 64, 16>(A, lda, B, ldb, C, ldc, m, n, kc, s, behind); break;
  }
}

template <typename T, bool SET>
hipError_t gemm_launch(const T* A, int64_t lda, const T* B, int64_t ldb, T* C, int64_t ldc, int64_t m, int64_t n,
                       int kc, hipStream_t s, bool behind, bool g3, int tile) {
  if (m <= 0 || n <= 0 || kc <= 0) return hipSuccess;
  if (tile < 0 || tile >= GEMM_TILES) return hipErrorInvalidValue;
  // in place (B == C): one row tile only, so no workgroup writes rows another reads
  if (static_cast<const void*>(B) == static_cast<const void*>(C) && (m > 64 || ldb != ldc)) return hipErrorInvalidValue;
  // Measured on gfx950 at the LU's update shapes of a full factorization, 64-row
  // tiles with K chunks of 16 are the fastest in both precisions (complex64 ~100
  // TFLOP/s at m = n = 10^4, K = 256; complex128 ~42), and keep the LDS small
  // enough that a panel workgroup fits on a CU that runs the update's workgroups
  // at full occupancy; the autotuner may choose another shape per size class.
  if constexpr (sizeof(T) == 4) {
    if (g3) {
      gemm_shape<T, SET, true>(tile, A, lda, B, ldb, C, ldc, m, n, kc, s, behind);
      return hipGetLastError();
    }
  }
  gemm_shape<T, SET, false>(tile, A, lda, B, ldb, C, ldc, m, n, kc, s, behind);
  return hipGetLastError();
}

template <typename T>
hipError_t gemm_sub(const T* A, int64_t lda, const T* B, int64_t ldb, T* C, int64_t ldc, int64_t m, int64_t n,
                    int kc, hipStream_t s, bool behind = false, bool g3 = false, int tile = 0) {
  return gemm_launch<T, false>(A, lda, B, ldb, C, ldc, m, n, kc, s, behind, g3, tile);
}

template <typename T>
hipError_t gemm_set(const T* A, int64_t lda, const T* B, int64_t ldb, T* C, int64_t ldc, int64_t m, int64_t n,
                    int kc, hipStream_t s, bool behind = false, bool g3 = false, int tile = 0) {
  return gemm_launch<T, true>(A, lda, B, ldb, C, ldc, m, n, kc, s, behind, g3, tile);
}

// ---- solves ----------------------------------------------------------------------

// The chain matrices of the block triangular solves (leaf = PNB rows; chain
// position p of leaf i: i forward, np - 1 - i backward). Per leaf, direction
// (blockIdx.y / 2: L forward, U backward) and kind (blockIdx.y % 2) one
// triangular solve with the leaf's diagonal block and 64 right-hand-side
// columns, one workgroup each: kind 0 the inverse D_i (X = I), kind 1 the link
// W_i = D_i M_i,k to the leaf k one position earlier on the chain (zero for the
// first). L's rows are taken in final order through the gather table of leaf
// k's outer block. Results go to mats + ((dir * 2 + kind) np + i) PNB^2
// (row-major blocks). The diagonal block is staged in LDS (identity beyond the
// leaf's extent). All columns are solved at once: four lanes per column, lane kq
// holding the column's rows q = kq (mod 4) in registers, each row's sum over the
// solved rows split over the four lanes and combined by two exchanges.
template <typename T>
__global__ __launch_bounds__(256) void solve_blocks_kernel(const T* LU, int64_t n, int64_t lda, const int* G, int bw,
                                                           int np, T* mats) {
  constexpr int DL = PNB + 1;
  __shared__ T As[2 * PNB * DL];  // the diagonal block
  const int i = blockIdx.x, dir = blockIdx.y >> 1, kind = blockIdx.y & 1, t = threadIdx.x, c = t >> 2, kq = t & 3;
  const bool lower = dir == 0;
  const int64_t base = (int64_t)i * PNB;
  const int nbi = (int)min((int64_t)PNB, n - base);
  for (int idx = t; idx < PNB * PNB; idx += 256) {
    const int r = idx / PNB, q = idx % PNB;
    T vr = r == q ? T(1) : T(0), vi = 0;
    if (r < nbi && q < nbi) {
      vr = LU[2 * ((base + r) * lda + base + q)];
      vi = LU[2 * ((base + r) * lda + base + q) + 1];
    }
    As[2 * (r * DL + q)] = vr;
    As[2 * (r * DL + q) + 1] = vi;
  }
  // the right-hand side's column c, rows 4 s + kq in slot s (b), and the solution
  // (x, zero until its row is solved, so that the row sums need no mask)
  T br[16], bi[16], xr[16], xi[16];
#pragma unroll
  for (int sl = 0; sl < 16; ++sl) xr[sl] = xi[sl] = 0;
  if (kind == 0) {
#pragma unroll
    for (int sl = 0; sl < 16; ++sl) {
      br[sl] = 4 * sl + kq == c ? T(1) : T(0);
      bi[sl] = 0;
    }
  } else {
    const int k = lower ? i - 1 : i + 1;  // the leaf one position earlier on the chain
    const bool has = lower ? k >= 0 : k < np;
    const int64_t c0 = (int64_t)k * PNB;  // its first column
    const int nb = has ? (int)min((int64_t)PNB, n - c0) : 0;
    const int* Gb = G + (lower && has ? (c0 / bw) * n + base : 0);
    // the rows first (independent loads), then the entries
    int64_t row[16];
#pragma unroll
    for (int sl = 0; sl < 16; ++sl) {
      const int q = 4 * sl + kq;
      row[sl] = (q < nbi && c < nb) ? (lower ? (int64_t)Gb[q] : base + q) : -1;
    }
#pragma unroll
    for (int sl = 0; sl < 16; ++sl) {
      T vr = 0, vi = 0;
      if (row[sl] >= 0) {
        vr = LU[2 * (row[sl] * lda + c0 + c)];
        vi = LU[2 * (row[sl] * lda + c0 + c) + 1];
      }
      br[sl] = vr;
      bi[sl] = vi;
    }
  }
  __syncthreads();
  // substitution over the rows (runtime loop, the slots unrolled): the row's sum
  // over the solved part split over the quad, the solved row's slot written by
  // selection
  for (int step = 0; step < PNB; ++step) {
    const int r = lower ? step : PNB - 1 - step;
    T sr = 0, si = 0;
#pragma unroll
    for (int sl = 0; sl < 16; ++sl) {
      const int q = 4 * sl + kq;
      const T dr = As[2 * (r * DL + q)], di = As[2 * (r * DL + q) + 1];
      sr += dr * xr[sl] - di * xi[sl];
      si += dr * xi[sl] + di * xr[sl];
    }
    sr = quad_sum(sr);
    si = quad_sum(si);
    T rr = 0, ri = 0;
#pragma unroll
    for (int sl = 0; sl < 16; ++sl) {
      const bool own = 4 * sl + kq == r;
      rr = own ? br[sl] : rr;
      ri = own ? bi[sl] : ri;
    }
    rr -= sr;
    ri -= si;
    if (!lower) {
      const T pr = As[2 * (r * DL + r)], pi = As[2 * (r * DL + r) + 1];
      const T den = pr * pr + pi * pi;
      const T vr = (rr * pr + ri * pi) / den, vi = (ri * pr - rr * pi) / den;
      rr = vr;
      ri = vi;
    }
#pragma unroll
    for (int sl = 0; sl < 16; ++sl) {
      const bool own = 4 * sl + kq == r;
      xr[sl] = own ? rr : xr[sl];
      xi[sl] = own ? ri : xi[sl];
    }
  }
  T* out = mats + 2 * (((int64_t)(dir * 2 + kind) * np + i) * PNB * PNB);
#pragma unroll
  for (int sl = 0; sl < 16; ++sl) {
    const int q = 4 * sl + kq;
    out[2 * (q * PNB + c)] = xr[sl];
    out[2 * (q * PNB + c) + 1] = xi[sl];
  }
}

// Row gather tables of the forward solve. The L columns of outer block J are
// stored in the row order after block J's own exchanges (later blocks never
// swap them), while the solve runs in the final order. For final row r,
// G[J][r] is the row it occupied after block J's exchanges, found by undoing the
// exchanges of blocks J+1.. in reverse; perm[r] is its original row.
// Two steps: each block's exchanges composed into one sparse map R[J] (row after
// the block's exchanges -> row before them; identity elsewhere), then per row a
// walk through the nJ maps (instead of through all n exchanges).
__global__ __launch_bounds__(256) void block_map_kernel(const int* ipiv, int64_t n, int bw, int* R) {
  __shared__ int sp[MAXBW];
  const int tid = threadIdx.x, J = blockIdx.x;
  const int64_t k0 = (int64_t)J * bw, k1 = min(n, k0 + bw);
  const int K = (int)(k1 - k0);
  for (int k = tid; k < K; k += blockDim.x) sp[k] = ipiv[k0 + k];
  __syncthreads();
  // entry e < K: row k0 + e; entry MAXBW + e: the target of exchange e beyond k1
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int e = tid + 256 * h;
    int dst = -1;
    if (e < K) dst = (int)(k0 + e);
    else if (e >= MAXBW && e - MAXBW < K && sp[e - MAXBW] >= k1) dst = sp[e - MAXBW];
    if (dst < 0) continue;
    int q = dst;
    for (int k = K - 1; k >= 0; --k) {
      const int kk = (int)k0 + k, p = sp[k];
      if (q == kk) q = p;
      else if (q == p) q = kk;
    }
    R[(int64_t)J * n + dst] = q;
  }
}

__global__ __launch_bounds__(256) void identity_maps_kernel(int64_t n, int nJ, int* R) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (int64_t)nJ * n;
       i += (int64_t)gridDim.x * blockDim.x)
    R[i] = (int)(i % n);
}

__global__ __launch_bounds__(256) void gather_table_kernel(const int* R, int64_t n, int nJ, int* G, int* perm) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  int q = (int)r;
  for (int J = nJ - 1; J >= 0; --J) {
    G[(int64_t)J * n + r] = q;
    q = R[(int64_t)J * n + q];
  }
  perm[r] = q;
}

// dst (n x 8, zero-padded) = rows perm[r] of src (n x nrhs)
template <typename T>
__global__ void gather_rhs_kernel(const T* src, int64_t n, int nrhs, const int* perm, T* dst) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n * 8) return;
  const int64_t r = idx >> 3;
  const int d = (int)(idx & 7);
  T re = 0, im = 0;
  if (d < nrhs) {
    const int64_t s = perm[r];
    re = src[2 * (s * nrhs + d)];
    im = src[2 * (s * nrhs + d) + 1];
  }
  dst[2 * idx] = re;
  dst[2 * idx + 1] = im;
}

template <typename T>
struct TriArgs {
  const T* LU;
  int64_t n, lda;
  const T* mats;      // this direction's chain matrices: D, W x np blocks of PNB x PNB (solve_blocks_kernel)
  const int* gather;  // forward: outer-block row gather tables (nJ x n)
  int bw;             // outer block width of the factorization
  const T* src;       // n x 8: forward P b, backward y
  T* pub;             // n x 8: block results, written with coherent stores for the other workgroups
  int* flags;         // per leaf: epoch once the leaf's result is published
  unsigned long long* tpub;  // n x 8 results as epoch-tagged words (TaggedRow), read by the next leaf on the chain
  int* ctl;           // wait control of the factorization (abort word, give-up record)
  int epoch;
  int np;
  T* out;             // backward: caller's n x nrhs solution
  int nrhs;
};

// One part (real or imaginary) of a published element: its TW / 2 tagged words
// (TaggedRow), and its plain word.
__device__ __forceinline__ void st_tagged_part(unsigned long long* q, float v, unsigned tag) {
  st_u64(q, ((unsigned long long)tag << 32) | __float_as_uint(v));
}
__device__ __forceinline__ void st_tagged_part(unsigned long long* q, double v, unsigned tag) {
  const unsigned long long t = (unsigned long long)tag << 32, b = (unsigned long long)__double_as_longlong(v);
  st_u64(q, t | (b & 0xffffffffull));
  st_u64(q + 1, t | (b >> 32));
}
__device__ __forceinline__ void st_part(float* p, float v) {
  __hip_atomic_store(reinterpret_cast<unsigned*>(p), __float_as_uint(v), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_part(double* p, double v) {
  st_u64(reinterpret_cast<unsigned long long*>(p), (unsigned long long)__double_as_longlong(v));
}

// Persistent block triangular solve with 8 right-hand sides (zero-padded): one
// workgroup per leaf of PNB rows; workgroup p takes chain position p (leaf p
// forward, L y = P b; leaf np - 1 - p backward, U x = y) and waits only for
// earlier positions (launch_chain). Every leaf but the neighbour on the chain is
// folded into a_p = src_p - sum_k M_pk x_k while the chain is still further
// away, then z_p = D_p a_p; when the neighbour's result n arrives, the leaf's own
// is z_p - W_p n with the link matrix W_p = D_p M_p,nb of solve_blocks_kernel,
// so one block product sits on the chain per leaf. A result is published twice:
// as epoch-tagged words, which the next leaf on the chain polls directly (no
// flag round trip on the chain), and as plain words announced by a flag after
// the stores have completed, which the folds of later leaves read.
// All block products run on the matrix cores: wave w computes rows 16w..16w+15 of
// a PNB x 8 complex product as a 16 x 16 real tile (columns: the real parts of
// the 8 right-hand sides, then their imaginary parts) over K = 2 PNB real terms,
// the lane holding its row's 16 consecutive columns 16 (lane / 16) .. of the
// left factor; the fold accumulates in the tile, which is published from the
// accumulator, one part per lane. The next fold block's rows are loaded before
// its flag is awaited.
template <typename T, bool FWD>
__global__ __launch_bounds__(256) void tri_solve_kernel(const TriArgs<T> a) {
  // in a sweep the factorization and solves run beside the next system's
  // assembly: their waves issue ahead of the assembly's
  __builtin_amdgcn_s_setprio(2);
  constexpr int YS = PNB * 8 + 6;  // staged block result; 2-complex pad per 16-row group
  constexpr int TW = TaggedRow<T>::TW;
  constexpr int64_t BB = (int64_t)PNB * PNB;
  typedef typename Acc4<T>::type acc4;
  __shared__ T ys[2 * YS];
  __shared__ T bs[2 * PNB * 8];
  __shared__ int gave_up;
  const int t = threadIdx.x;
  const int wv = t >> 6, l16 = t & 15, mq = (t >> 4) & 3;  // wave, tile column / row, K group
  const int d = l16 & 7, part = l16 >> 3;                   // this lane's tile column: right-hand side, part
  if (t == 0) gave_up = 0;
  const int p = (int)blockIdx.x;
  auto leaf = [&](int pos) { return FWD ? pos : a.np - 1 - pos; };
  const int i = leaf(p);
  const int64_t base = (int64_t)i * PNB;
  const int nbi = (int)min((int64_t)PNB, a.n - base);
  const int site = FWD ? SPIN_TRI_FWD : SPIN_TRI_BWD;
  const int ar = 16 * wv + l16;  // this lane's row of a left factor
  // the accumulator starts from the right-hand side
  acc4 ca;
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int rr = 16 * wv + mfma_row<T>(mq, v);
    ca[v] = rr < nbi ? a.src[2 * ((base + rr) * 8 + d) + part] : T(0);
  }
  // this lane's columns of its rows of D and W: complex64 at once (the registers
  // are there), complex128 once the earlier leaves are folded in
  T dr[16], di[16], wr[16], wi[16];
  auto load_dw = [&]() {
    const int64_t o = 2 * ((int64_t)i * BB + (int64_t)ar * PNB + 16 * mq);
    load_seg16(a.mats + o, 16, true, dr, di);
    if (p > 0) load_seg16(a.mats + 2 * a.np * BB + o, 16, true, wr, wi);
  };
  if constexpr (sizeof(T) == 4) load_dw();
  auto stage = [&](int kk, int dd, T zr, T zi) {
    const int idx = kk * 8 + dd;
    ys[2 * (idx + (kk >> 4) * 2)] = zr;
    ys[2 * (idx + (kk >> 4) * 2) + 1] = zi;
  };
  // waits for leaf k's flag, then stages its published result in ys
  auto await_stage = [&](int k) -> bool {
    if (t == 0) {
      Spin sp;
      while (ld_flag(a.flags + k) != a.epoch) {
        if (spin_expired(a.ctl, sp, site)) {
          gave_up = 1;
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
    }
    __syncthreads();
    if (gave_up) return false;  // uniform after the barrier; the host reports the give-up
    const int64_t kb = (int64_t)k * PNB;
    const int nbk = (int)min((int64_t)PNB, a.n - kb);
    T zr[2], zi[2];
#pragma unroll
    for (int e = 0; e < 2; ++e) {  // both entries' loads first
      const int idx = t + 256 * e, kk = idx >> 3, dd = idx & 7;
      zr[e] = zi[e] = 0;
      if (nbk == PNB || kk < nbk) ld_cplx(a.pub + 2 * ((kb + kk) * 8 + dd), zr[e], zi[e]);
    }
#pragma unroll
    for (int e = 0; e < 2; ++e) {
      const int idx = t + 256 * e;
      stage(idx >> 3, idx & 7, zr[e], zi[e]);
    }
    __syncthreads();
    return true;
  };
  // stages the chain neighbour's tagged result: each thread polls the words of its
  // two entries (rows t/8 and t/8 + 32, right-hand side t%8) until all carry the
  // epoch, with the next poll's loads issued before the current ones are checked
  auto await_tagged = [&](int k) -> bool {
    const int64_t kb = (int64_t)k * PNB;
    const int nbk = (int)min((int64_t)PNB, a.n - kb);
    const int kk0 = t >> 3, kk1 = kk0 + 32, dd = t & 7;
    const bool v0 = kk0 < nbk, v1 = kk1 < nbk;
    const unsigned long long* q0 = a.tpub + ((kb + kk0) * 8 + dd) * TW;
    const unsigned long long* q1 = a.tpub + ((kb + kk1) * 8 + dd) * TW;
    const unsigned tag = (unsigned)a.epoch;
    unsigned long long a0[TW], a1[TW], b0[TW], b1[TW];
    auto issue = [&](unsigned long long* x0, unsigned long long* x1) {
#pragma unroll
      for (int e = 0; e < TW; ++e) {
        x0[e] = v0 ? ld_u64(q0 + e) : 0ull;
        x1[e] = v1 ? ld_u64(q1 + e) : 0ull;
      }
    };
    auto ready = [&](const unsigned long long* x0, const unsigned long long* x1) {
      bool ok = true;
#pragma unroll
      for (int e = 0; e < TW; ++e)
        ok = ok && (!v0 || (unsigned)(x0[e] >> 32) == tag) && (!v1 || (unsigned)(x1[e] >> 32) == tag);
      return ok;
    };
    auto take = [&](const unsigned long long* x0, const unsigned long long* x1) {
      T zr, zi;
      untag(x0, zr, zi);
      stage(kk0, dd, zr, zi);
      untag(x1, zr, zi);
      stage(kk1, dd, zr, zi);
    };
    Spin sp;
    issue(a0, a1);
    for (;;) {
      issue(b0, b1);
      if (ready(a0, a1)) {
        take(a0, a1);
        break;
      }
      if (spin_expired(a.ctl, sp, site)) {
        gave_up = 1;
        break;
      }
      issue(a0, a1);
      if (ready(b0, b1)) {
        take(b0, b1);
        break;
      }
      if (spin_expired(a.ctl, sp, site)) {
        gave_up = 1;
        break;
      }
    }
    __syncthreads();
    return !gave_up;
  };
  // c += s M[tile rows, :] Z for a staged PNB x 8 block Z (pad: the ys layout)
  // and sign s: step j takes columns j, 16 + j, 32 + j, 48 + j
  auto tile_product = [&](acc4& c, const T* mr, const T* mi, const T* zs, bool pad, T sgn) {
    const bool imc = part != 0;
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      const int k = 16 * mq + j;
      const int o = k * 8 + d + (pad ? mq * 2 : 0);
      const T zr = sgn * zs[2 * o], zi = sgn * zs[2 * o + 1];
      c = mfma16(mr[j], imc ? zi : zr, c);
      c = mfma16(mi[j], imc ? zr : -zi, c);
    }
  };
  // off the chain: every leaf but the neighbour
  for (int s = 0; s + 1 < p; ++s) {
    const int k = leaf(s);
    const int64_t kb = (int64_t)k * PNB;
    const int nv = (int)min((int64_t)PNB, a.n - kb) - 16 * mq;  // valid columns of this lane's segment
    T mr[16], mi[16];
    {
      const bool ok = ar < nbi;
      const int64_t srow = FWD ? (ok ? (int64_t)a.gather[(kb / a.bw) * a.n + base + ar] : 0) : base + ar;
      load_seg16(a.LU + 2 * (srow * a.lda + kb + 16 * mq), nv, ok, mr, mi);
    }
    if (!await_stage(k)) return;
    // each block's product summed on its own, then added: rounding grows with the
    // 2 PNB terms of a block and the number of blocks, not with their product
    acc4 tb;
#pragma unroll
    for (int v = 0; v < 4; ++v) tb[v] = 0;
    tile_product(tb, mr, mi, ys, true, T(-1));
    ca += tb;
  }
  if constexpr (sizeof(T) == 8) load_dw();
  // z = D a: the accumulator staged as the right operand
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int rr = 16 * wv + mfma_row<T>(mq, v);
    bs[2 * (rr * 8 + d) + part] = ca[v];
  }
  __syncthreads();
  acc4 c;
#pragma unroll
  for (int v = 0; v < 4; ++v) c[v] = 0;
  tile_product(c, dr, di, bs, false, T(1));
  // on the chain: z - W n
  if (p > 0) {
    if (!await_tagged(leaf(p - 1))) return;
    tile_product(c, wr, wi, ys, true, T(-1));
  }
  // publish: lane l16 < 8 the real part of right-hand side l16, the others the
  // imaginary part of l16 - 8; the tagged words first
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int rr = 16 * wv + mfma_row<T>(mq, v);
    if (rr < nbi) st_tagged_part(a.tpub + ((base + rr) * 8 + d) * TW + part * (TW / 2), c[v], (unsigned)a.epoch);
  }
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int rr = 16 * wv + mfma_row<T>(mq, v);
    if (rr < nbi) {
      st_part(a.pub + 2 * ((base + rr) * 8 + d) + part, c[v]);
      if (!FWD && d < a.nrhs) a.out[2 * ((base + rr) * a.nrhs + d) + part] = c[v];
    }
  }
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (t == 0) st_flag(a.flags + i, a.epoch);
}

// r = b - A x, complex128, 8 right-hand sides (zero-padded). Reading A once is
// the cost, so lanes run along the row (coalesced 1 KB per wavefront load): each
// wavefront accumulates 4 rows x 8 right-hand sides over k = lane (mod 64), with
// x staged in LDS ([d][k], conflict-free) in chunks shared by the workgroup's 8
// wavefronts (32 rows). The 64 partial sums per lane are then reduced across the
// wavefront by recursive halving (63 exchanges), leaving sum number l in lane l.
constexpr int RES_WAVES = 8, RES_ROWS = 4, RES_KC = 128;
constexpr int RES_SPLIT_MAX = 8;  // k ranges per row block of the single-system residual

template <int N>
__device__ __forceinline__ void halve_step(double* v, int lane) {
  const bool hi = (lane & N) != 0;
#pragma unroll
  for (int j = 0; j < N; ++j) {
    const double keep = hi ? v[N + j] : v[j];
    const double send = hi ? v[j] : v[N + j];
    v[j] = keep + __shfl_xor(send, N, 64);
  }
}

// xld / bld / rld: row strides of x, b and r (nrhs for the single-system form)
__device__ __forceinline__ void residual64_body(const double* A, int64_t n, int64_t lda, const double* x, int64_t xld,
                                                const double* b, int64_t bld, double* r, int64_t rld, int nrhs,
                                                int64_t block) {
  __shared__ double xs[8 * RES_KC * 2];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t row0 = (block * RES_WAVES + wave) * RES_ROWS;
  double acc[RES_ROWS * 8 * 2];
#pragma unroll
  for (int j = 0; j < RES_ROWS * 8 * 2; ++j) acc[j] = 0;
  constexpr int SUB = RES_KC / 64;
  for (int64_t k0 = 0; k0 < n; k0 += RES_KC) {
    const int kc = (int)min((int64_t)RES_KC, n - k0);
    // this chunk's A values first (all loads in flight), then x into LDS
    double ar[RES_ROWS][SUB], ai[RES_ROWS][SUB];
#pragma unroll
    for (int q = 0; q < RES_ROWS; ++q)
#pragma unroll
      for (int s = 0; s < SUB; ++s) {
        const int kk = lane + 64 * s;
        const int64_t row = row0 + q;
        ar[q][s] = ai[q][s] = 0;
        if (row < n && kk < kc) {
          ar[q][s] = A[2 * (row * lda + k0 + kk)];
          ai[q][s] = A[2 * (row * lda + k0 + kk) + 1];
        }
      }
    __syncthreads();
    for (int idx = tid; idx < RES_KC * 8; idx += 64 * RES_WAVES) {
      const int kk = idx >> 3, d = idx & 7;
      double xr = 0, xi = 0;
      if (kk < kc && d < nrhs) {
        xr = x[2 * ((k0 + kk) * xld + d)];
        xi = x[2 * ((k0 + kk) * xld + d) + 1];
      }
      xs[2 * (d * RES_KC + kk)] = xr;
      xs[2 * (d * RES_KC + kk) + 1] = xi;
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < SUB; ++s) {
      const int kk = lane + 64 * s;
      double xr[8], xi[8];
#pragma unroll
      for (int d = 0; d < 8; ++d) {
        xr[d] = xs[2 * (d * RES_KC + kk)];
        xi[d] = xs[2 * (d * RES_KC + kk) + 1];
      }
#pragma unroll
      for (int q = 0; q < RES_ROWS; ++q)
#pragma unroll
        for (int d = 0; d < 8; ++d) {
          acc[(q * 8 + d) * 2] = fma(ar[q][s], xr[d], fma(-ai[q][s], xi[d], acc[(q * 8 + d) * 2]));
          acc[(q * 8 + d) * 2 + 1] = fma(ar[q][s], xi[d], fma(ai[q][s], xr[d], acc[(q * 8 + d) * 2 + 1]));
        }
    }
  }
  halve_step<32>(acc, lane);
  halve_step<16>(acc, lane);
  halve_step<8>(acc, lane);
  halve_step<4>(acc, lane);
  halve_step<2>(acc, lane);
  halve_step<1>(acc, lane);
  // lane l holds sum number l = (q * 8 + d) * 2 + (0: re, 1: im)
  const int q = lane >> 4, d = (lane >> 1) & 7, c = lane & 1;
  const int64_t row = row0 + q;
  if (row < n && d < nrhs) r[2 * (row * rld + d) + c] = b[2 * (row * bld + d) + c] - acc[0];
}

__global__ __launch_bounds__(512) void residual64_batched_kernel(const BatchSystem<double>* sys,
                                                                 const BatchVec<double>* x, const BatchVec<double>* b,
                                                                 const BatchVec<double>* r, int nrhs) {
  const int j = blockIdx.y;
  const BatchSystem<double> S = sys[j];
  const BatchVec<double> X = x[j], B = b[j], R = r[j];
  if (R.skip && *R.skip) return;
  if ((int64_t)blockIdx.x * RES_WAVES * RES_ROWS >= S.n) return;
  residual64_body(S.A, S.n, S.lda, X.v, X.ld, B.v, B.ld, R.v, R.ld, nrhs, (int64_t)blockIdx.x);
}

// The single-system residual r = b - A x over a (row block, k range) grid: the
// workgroups of one row block split its columns into grid.y ranges of 64-column
// chunks, so that the grid fills whole rounds of the CUs (one workgroup of 8
// wavefronts per CU), and each writes its partial sums A x to part[range]; a
// second kernel forms b - sum. Within a workgroup, the next chunk's A values
// (registers) and x values (the other LDS buffer) are loaded while the current
// chunk is multiplied; one barrier per chunk. Lanes run along the row as in
// residual64_body (one 1 KB load per row and chunk per wavefront).
constexpr int RS_SUB = 1, RS_KC = 64 * RS_SUB;  // columns per chunk: RS_SUB per lane
__global__ __launch_bounds__(512) void residual64_split_kernel(const double* A, int64_t n, int64_t lda,
                                                               const double* x, int nrhs, double* part) {
  __builtin_amdgcn_s_setprio(2);
  __shared__ double xs[2][8 * RS_KC * 2];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t row0 = ((int64_t)blockIdx.x * RES_WAVES + wave) * RES_ROWS;
  const int64_t nch = (n + RS_KC - 1) / RS_KC;
  const int64_t per = (nch + gridDim.y - 1) / gridDim.y;
  const int64_t c0 = (int64_t)blockIdx.y * per, c1 = min(nch, c0 + per);
  double acc[RES_ROWS * 8 * 2];
#pragma unroll
  for (int j = 0; j < RES_ROWS * 8 * 2; ++j) acc[j] = 0;
  // column lane + 64 s of a chunk; thread tid stages the x entries (k = tid / 8 +
  // 64 s, right-hand side tid % 8)
  const int xk = tid >> 3, xd = tid & 7;
  auto load_a = [&](int64_t c, double (*ar)[RS_SUB], double (*ai)[RS_SUB]) {
#pragma unroll
    for (int q = 0; q < RES_ROWS; ++q)
#pragma unroll
      for (int sb = 0; sb < RS_SUB; ++sb) {
        const int64_t row = row0 + q, k = c * RS_KC + lane + 64 * sb;
        ar[q][sb] = ai[q][sb] = 0;
        if (row < n && k < n) {
          const double2 v = *reinterpret_cast<const double2*>(A + 2 * (row * lda + k));
          ar[q][sb] = v.x;
          ai[q][sb] = v.y;
        }
      }
  };
  auto load_x = [&](int64_t c, double2* v) {
#pragma unroll
    for (int sb = 0; sb < RS_SUB; ++sb) {
      const int64_t k = c * RS_KC + xk + 64 * sb;
      v[sb] = make_double2(0.0, 0.0);
      if (k < n && xd < nrhs) v[sb] = *reinterpret_cast<const double2*>(x + 2 * (k * nrhs + xd));
    }
  };
  auto put_x = [&](int buf, const double2* v) {
#pragma unroll
    for (int sb = 0; sb < RS_SUB; ++sb) {
      xs[buf][2 * (xd * RS_KC + xk + 64 * sb)] = v[sb].x;
      xs[buf][2 * (xd * RS_KC + xk + 64 * sb) + 1] = v[sb].y;
    }
  };
  double ar[RES_ROWS][RS_SUB], ai[RES_ROWS][RS_SUB], nr[RES_ROWS][RS_SUB], ni[RES_ROWS][RS_SUB];
  if (c0 < c1) {
    double2 v[RS_SUB];
    load_a(c0, ar, ai);
    load_x(c0, v);
    put_x(0, v);
  }
  __syncthreads();
  for (int64_t c = c0; c < c1; ++c) {
    const int buf = (int)((c - c0) & 1);
    const bool more = c + 1 < c1;
    double2 xn[RS_SUB];
    if (more) {
      load_a(c + 1, nr, ni);
      load_x(c + 1, xn);
    }
#pragma unroll
    for (int sb = 0; sb < RS_SUB; ++sb) {
      double xr[8], xi[8];
#pragma unroll
      for (int d = 0; d < 8; ++d) {
        xr[d] = xs[buf][2 * (d * RS_KC + lane + 64 * sb)];
        xi[d] = xs[buf][2 * (d * RS_KC + lane + 64 * sb) + 1];
      }
#pragma unroll
      for (int q = 0; q < RES_ROWS; ++q)
#pragma unroll
        for (int d = 0; d < 8; ++d) {
          acc[(q * 8 + d) * 2] = fma(ar[q][sb], xr[d], fma(-ai[q][sb], xi[d], acc[(q * 8 + d) * 2]));
          acc[(q * 8 + d) * 2 + 1] = fma(ar[q][sb], xi[d], fma(ai[q][sb], xr[d], acc[(q * 8 + d) * 2 + 1]));
        }
    }
    if (more) put_x(buf ^ 1, xn);
    __syncthreads();
#pragma unroll
    for (int q = 0; q < RES_ROWS; ++q)
#pragma unroll
      for (int sb = 0; sb < RS_SUB; ++sb) {
        ar[q][sb] = nr[q][sb];
        ai[q][sb] = ni[q][sb];
      }
  }
  halve_step<32>(acc, lane);
  halve_step<16>(acc, lane);
  halve_step<8>(acc, lane);
  halve_step<4>(acc, lane);
  halve_step<2>(acc, lane);
  halve_step<1>(acc, lane);
  // lane l holds sum number l = (q * 8 + d) * 2 + (0: re, 1: im)
  const int q = lane >> 4, d = (lane >> 1) & 7, cpart = lane & 1;
  const int64_t row = row0 + q;
  if (row < n) part[((int64_t)blockIdx.y * n + row) * 16 + d * 2 + cpart] = acc[0];
}

// The same residual in small workgroups, for a GPU shared with other streams'
// long kernels (a sweep's next assembly): 4 wavefronts of 2 rows x 8 right-hand
// sides each (164 VGPRs), so that a workgroup fits the slot one retired
// workgroup of the other kernel leaves on a CU, where residual64_split_kernel's
// 8 wavefronts at 220 VGPRs wait for a whole CU to drain (in the L5 sweep the
// solve and refinement take 4.5 against 8.4 ms; alone it is slower, 615 against
// 419 us). Each wavefront keeps the next two chunks' A values in flight (three
// register buffers, rotated by unrolling) while the current chunk is
// multiplied; x in two LDS buffers, one barrier per chunk; one k range.
constexpr int RS_ROWS = 2, RS_WAVES = 4;  // rows per wavefront, wavefronts per workgroup
__global__ __launch_bounds__(64 * RS_WAVES) void residual64_narrow_kernel(const double* A, int64_t n, int64_t lda,
                                                                         const double* x, int nrhs, double* part) {
  __builtin_amdgcn_s_setprio(2);
  __shared__ double xs[2][8 * RS_KC * 2];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t row0 = ((int64_t)blockIdx.x * RS_WAVES + wave) * RS_ROWS;
  const int64_t nch = (n + RS_KC - 1) / RS_KC;
  const int64_t per = (nch + gridDim.y - 1) / gridDim.y;
  const int64_t c0 = (int64_t)blockIdx.y * per, c1 = min(nch, c0 + per);
  double acc[RS_ROWS * 8 * 2];
#pragma unroll
  for (int j = 0; j < RS_ROWS * 8 * 2; ++j) acc[j] = 0;
  // column lane + 64 s of a chunk; thread tid stages the x entries (k = tid / 8 +
  // 32 e, right-hand side tid % 8)
  constexpr int XE = RS_KC * 8 / (64 * RS_WAVES);
  const int xk = tid >> 3, xd = tid & 7;
  auto load_a = [&](int64_t c, double (*ar)[RS_SUB], double (*ai)[RS_SUB]) {
#pragma unroll
    for (int q = 0; q < RS_ROWS; ++q)
#pragma unroll
      for (int sb = 0; sb < RS_SUB; ++sb) {
        const int64_t row = row0 + q, k = c * RS_KC + lane + 64 * sb;
        ar[q][sb] = ai[q][sb] = 0;
        if (row < n && k < n) {
          const double2 v = *reinterpret_cast<const double2*>(A + 2 * (row * lda + k));
          ar[q][sb] = v.x;
          ai[q][sb] = v.y;
        }
      }
  };
  auto load_x = [&](int64_t c, double2* v) {
#pragma unroll
    for (int e = 0; e < XE; ++e) {
      const int64_t k = c * RS_KC + xk + 8 * RS_WAVES * e;
      v[e] = make_double2(0.0, 0.0);
      if (k < n && xd < nrhs) v[e] = *reinterpret_cast<const double2*>(x + 2 * (k * nrhs + xd));
    }
  };
  auto put_x = [&](int buf, const double2* v) {
#pragma unroll
    for (int e = 0; e < XE; ++e) {
      xs[buf][2 * (xd * RS_KC + xk + 8 * RS_WAVES * e)] = v[e].x;
      xs[buf][2 * (xd * RS_KC + xk + 8 * RS_WAVES * e) + 1] = v[e].y;
    }
  };
  // A in three register buffers, rotated by unrolling: while chunk c is
  // multiplied, chunks c + 1 and c + 2 are in flight; x in two LDS buffers
  double pr[RS_ROWS][RS_SUB], pi[RS_ROWS][RS_SUB], qr[RS_ROWS][RS_SUB], qi[RS_ROWS][RS_SUB];
  double sr[RS_ROWS][RS_SUB], si[RS_ROWS][RS_SUB];
  if (c0 < c1) {
    double2 v[XE];
    load_a(c0, pr, pi);
    load_x(c0, v);
    put_x(0, v);
  }
  if (c0 + 1 < c1) load_a(c0 + 1, qr, qi);
  if (c0 + 2 < c1) load_a(c0 + 2, sr, si);
  __syncthreads();
  // chunk c from buffer (ar, ai): the next chunk's x staged during its products
  auto step = [&](int64_t c, double (*ar)[RS_SUB], double (*ai)[RS_SUB]) {
    const int buf = (int)((c - c0) & 1);
    const bool more = c + 1 < c1;
    double2 xn[XE];
    if (more) load_x(c + 1, xn);
#pragma unroll
    for (int sb = 0; sb < RS_SUB; ++sb)
#pragma unroll
      for (int d = 0; d < 8; ++d) {
        const double xr = xs[buf][2 * (d * RS_KC + lane + 64 * sb)];
        const double xi = xs[buf][2 * (d * RS_KC + lane + 64 * sb) + 1];
#pragma unroll
        for (int q = 0; q < RS_ROWS; ++q) {
          acc[(q * 8 + d) * 2] = fma(ar[q][sb], xr, fma(-ai[q][sb], xi, acc[(q * 8 + d) * 2]));
          acc[(q * 8 + d) * 2 + 1] = fma(ar[q][sb], xi, fma(ai[q][sb], xr, acc[(q * 8 + d) * 2 + 1]));
        }
      }
    if (more) put_x(buf ^ 1, xn);
    __syncthreads();
  };
  for (int64_t c = c0; c < c1; c += 3) {
    step(c, pr, pi);
    if (c + 3 < c1) load_a(c + 3, pr, pi);
    if (c + 1 < c1) {
      step(c + 1, qr, qi);
      if (c + 4 < c1) load_a(c + 4, qr, qi);
    }
    if (c + 2 < c1) {
      step(c + 2, sr, si);
      if (c + 5 < c1) load_a(c + 5, sr, si);
    }
  }
  // the two half-wavefronts' sums first, then recursive halving over the 32
  // sums: lanes l and l + 32 end with sum number l % 32 = (q * 8 + d) * 2 + part
#pragma unroll
  for (int j = 0; j < RS_ROWS * 16; ++j) acc[j] += __shfl_xor(acc[j], 32, 64);
  halve_step<16>(acc, lane);
  halve_step<8>(acc, lane);
  halve_step<4>(acc, lane);
  halve_step<2>(acc, lane);
  halve_step<1>(acc, lane);
  const int q = (lane >> 4) & 1, d = (lane >> 1) & 7, cpart = lane & 1;
  const int64_t row = row0 + q;
  if (lane < 32 && row < n) part[((int64_t)blockIdx.y * n + row) * 16 + d * 2 + cpart] = acc[0];
}

// r = b - sum over the k ranges of the partial products
__global__ void residual64_finish_kernel(const double* part, int ks, int64_t n, const double* b, double* r, int nrhs) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n * nrhs * 2) return;
  const int64_t row = idx / (nrhs * 2);
  const int e = (int)(idx % (nrhs * 2));  // d * 2 + (re, im)
  double s = 0;
  for (int j = 0; j < ks; ++j) s += part[((int64_t)j * n + row) * 16 + e];
  r[idx] = b[idx] - s;
}

__global__ __launch_bounds__(256) void colnorm_partial_kernel(const double* X, int64_t n, int nrhs, double* part) {
  __shared__ double red[256];
  const int tid = threadIdx.x;
  for (int d = 0; d < nrhs; ++d) {
    double s = 0;
    for (int64_t r = (int64_t)blockIdx.x * 256 + tid; r < n; r += (int64_t)gridDim.x * 256) {
      const double a = X[2 * (r * nrhs + d)], b = X[2 * (r * nrhs + d) + 1];
      s += a * a + b * b;
    }
    red[tid] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
      if (tid < o) red[tid] += red[tid + o];
      __syncthreads();
    }
    if (tid == 0) part[(int64_t)blockIdx.x * nrhs + d] = red[0];
    __syncthreads();
  }
}

__global__ void colnorm_final_kernel(const double* part, int nparts, int nrhs, double* out) {
  const int d = threadIdx.x;
  if (d >= nrhs) return;
  double s = 0;
  for (int p = 0; p < nparts; ++p) s += part[(int64_t)p * nrhs + d];
  out[d] = sqrt(s);
}

// ---- speculative panels: no row exchange, verified --------------------------------
// Partial pivoting exchanges rows only where a row below the diagonal holds a
// larger |re| + |im| in the current column than the diagonal. Burton-Miller
// Galerkin matrices have a dominant diagonal, and the factorization keeps it, so
// LAPACK's getrf makes no exchange on them. The speculative panel factors a leaf
// without exchanges and checks, at every column, izamax's choice on every row
// (the diagonal must be the first maximum of its column); the check is exact, so
// whenever it passes the factors are those of partial pivoting with the identity
// pivot sequence. No workgroup waits for another. A failed check raises a flag
// (LuWork::ctl[2]) and the caller factors the matrix again with the exchange
// panels.

template <int I, int N, typename F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr (I < N) {
    f(std::integral_constant<int, I>{});
    static_for<I + 1, N>(f);
  }
}

// lane K of every group of four lanes to the whole group (DPP quad_perm [K,K,K,K])
template <int K>
__device__ __forceinline__ float quad_bcast(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), K * 0x55, 0xF, 0xF, false));
}
template <int K>
__device__ __forceinline__ double quad_bcast(double v) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_mov_dpp((int)(unsigned)(b & 0xffffffffull), K * 0x55, 0xF, 0xF, false);
  const unsigned hi = (unsigned)__builtin_amdgcn_mov_dpp((int)(unsigned)(b >> 32), K * 0x55, 0xF, 0xF, false);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

template <typename T>
__device__ __forceinline__ double cabs1(T re, T im) {
  return fabs((double)re) + fabs((double)im);
}

// Multiplier arithmetic of both speculative kernels: the pivot's reciprocal
// (conj(u) / |u|^2) once per column, then products, as getf2 scales a column by
// 1 / u. A zero pivot gives zero multipliers (the column is left as it is).
template <typename T>
__device__ __forceinline__ void pivot_reciprocal(T ur, T ui, T& ir, T& ii) {
  const T den = ur * ur + ui * ui;
  T rd;
  if constexpr (sizeof(T) == 4) {
    // the hardware reciprocal and one Newton step (within an ulp of 1 / den; the
    // complex64 factors are those of a preconditioner or of the plain complex64
    // solve), instead of the longer correctly rounded division on each column's
    // critical path
    const float y = __builtin_amdgcn_rcpf(den);
    rd = den != 0.f ? fmaf(fmaf(-den, y, 1.f), y, y) : 0.f;
  } else {
    rd = den != T(0) ? T(1) / den : T(0);
  }
  ir = ur * rd;
  ii = -ui * rd;
}

// Diagonal block of a leaf (one workgroup): the b x b block at (j0, j0) factored
// in place without exchanges, and the inverse of its unit-lower factor (the leaf
// inverse of the U12 products) formed alongside by applying the same row
// operations to the identity. Thread t owns row t / 4 and, of it, the columns
// c = 4s + t % 4 (s = 0..15) in registers; row j travels through LDS (two buffers
// by step parity, one barrier per step), the multiplier of a row through its four
// lanes. Rows and columns beyond b are the identity. The empty asm statements
// with register operands make each step's updates happen in that step (the
// compiler would otherwise defer them to their first use and keep every step's
// operands live).
template <typename T, bool LINV = true>
__device__ __forceinline__ void diag_nopiv_body(T* A, int64_t lda, int64_t j0, int b, T* linv, int* info, int* ctl) {
  __shared__ T prow[2][4][16][2];  // row j of U by column class [parity][c % 4][c / 4][re, im]
  __shared__ T plow[2][4][16][2];  // row j of L^-1 (LINV)
  __shared__ T prcp[2][2];         // 1 / u_jj
  const int t = threadIdx.x, r = t >> 2, q = t & 3;
  T xr[16], xi[16], yr[16], yi[16];
  {
    const T* src = A + 2 * ((j0 + min(r, b - 1)) * lda + j0 + q);  // rows beyond b: an in-bounds row, masked
#pragma unroll
    for (int s = 0; s < 16; ++s) {
      xr[s] = src[8 * s];
      xi[s] = src[8 * s + 1];
    }
  }
  // (every load issued first: the asm statements order what follows them)
#pragma unroll
  for (int s = 0; s < 16; ++s) {
    const int c = 4 * s + q;
    const bool in = r < b && c < b;
    xr[s] = in ? xr[s] : (r == c ? T(1) : T(0));
    xi[s] = in ? xi[s] : T(0);
    yr[s] = r == c ? T(1) : T(0);
    yi[s] = T(0);
    __asm__ __volatile__("" : "+v"(xr[s]), "+v"(xi[s]), "+v"(yr[s]), "+v"(yi[s]));
  }
  bool viol = false;
  int zero = 0;
  static_for<0, PNB>([&](auto J) {
    constexpr int j = decltype(J)::value, S = j >> 2, k = j & 3, par = j & 1;
    if (r == j) {
#pragma unroll
      for (int s = S; s < 16; ++s) {
        prow[par][q][s][0] = xr[s];
        prow[par][q][s][1] = xi[s];
      }
      if constexpr (LINV)
#pragma unroll
        for (int s = 0; s <= S; ++s) {
          plow[par][q][s][0] = yr[s];
          plow[par][q][s][1] = yi[s];
        }
      if (q == k) pivot_reciprocal(xr[S], xi[S], prcp[par][0], prcp[par][1]);
    }
    __syncthreads();
    const T ur = prow[par][k][S][0], ui = prow[par][k][S][1];
    const T ir = prcp[par][0], ii = prcp[par][1];
    // the lane of class k holds column j of the row
    const T mr = quad_bcast<k>(xr[S] * ir - xi[S] * ii), mi = quad_bcast<k>(xr[S] * ii + xi[S] * ir);
    if (r > j) {
      if (q == k) {
        viol |= cabs1(xr[S], xi[S]) > cabs1(ur, ui);
        xr[S] = mr;
        xi[S] = mi;
      }
#pragma unroll
      for (int s = S; s < 16; ++s)
        if (s > S || q > k) {
          const T pr = prow[par][q][s][0], pi = prow[par][q][s][1];
          xr[s] -= mr * pr - mi * pi;
          xi[s] -= mr * pi + mi * pr;
        }
      if constexpr (LINV)
#pragma unroll
        for (int s = 0; s <= S; ++s)
          if (s < S || q <= k) {
            const T pr = plow[par][q][s][0], pi = plow[par][q][s][1];
            yr[s] -= mr * pr - mi * pi;
            yi[s] -= mr * pi + mi * pr;
          }
    }
#pragma unroll
    for (int s = S; s < 16; ++s) __asm__ __volatile__("" : "+v"(xr[s]), "+v"(xi[s]));
    if constexpr (LINV)
#pragma unroll
      for (int s = 0; s <= S; ++s) __asm__ __volatile__("" : "+v"(yr[s]), "+v"(yi[s]));
    if (ur == T(0) && ui == T(0) && zero == 0 && j < b) zero = j + 1;
  });
#pragma unroll
  for (int s = 0; s < 16; ++s) {
    const int c = 4 * s + q;
    if (r < b && c < b) {
      T* dst = A + 2 * ((j0 + r) * lda + j0 + c);
      dst[0] = xr[s];
      dst[1] = xi[s];
    }
    if (LINV) {
      linv[2 * (r * PNB + c)] = yr[s];
      linv[2 * (r * PNB + c) + 1] = yi[s];
    }
  }
  if (__ballot(viol) != 0 && (t & 63) == 0) __hip_atomic_fetch_or(ctl + 2, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (t == 0 && zero != 0 && *info == 0) *info = (int)j0 + zero;
}

template <typename T, bool LINV>
__global__ __launch_bounds__(256) void diag_nopiv_kernel(T* A, int64_t lda, int64_t j0, int b, T* linv, int* info,
                                                         int* ctl) {
  __builtin_amdgcn_s_setprio(3);
  diag_nopiv_body<T, LINV>(A, lda, j0, b, linv, info, ctl);
}

// The inverse of the leaf's unit-lower factor L11 (strictly below the diagonal of
// the factored diagonal block), by the row operations of the elimination applied
// to the identity: step j broadcasts row j of the inverse, final after steps
// 0..j-1, and every row i > j subtracts L11[i][j] times it. Threads as in
// diag_nopiv_body; L11[i][j] reaches the row's four lanes by quad broadcast.
template <typename T>
__device__ __forceinline__ void linv_body(const T* A, int64_t lda, int64_t j0, int b, T* linv) {
  __shared__ T plow[2][4][16][2];  // row j of L^-1 by column class [parity][c % 4][c / 4][re, im]
  const int t = threadIdx.x, r = t >> 2, q = t & 3;
  T lr[16], li[16], yr[16], yi[16];
  {
    const T* src = A + 2 * ((j0 + min(r, b - 1)) * lda + j0 + q);
#pragma unroll
    for (int s = 0; s < 16; ++s) {
      lr[s] = src[8 * s];
      li[s] = src[8 * s + 1];
    }
  }
#pragma unroll
  for (int s = 0; s < 16; ++s) {
    const int c = 4 * s + q;
    const bool in = r < b && c < r;  // the strictly lower part
    lr[s] = in ? lr[s] : T(0);
    li[s] = in ? li[s] : T(0);
    yr[s] = r == c ? T(1) : T(0);
    yi[s] = T(0);
    __asm__ __volatile__("" : "+v"(lr[s]), "+v"(li[s]), "+v"(yr[s]), "+v"(yi[s]));
  }
  static_for<0, PNB>([&](auto J) {
    constexpr int j = decltype(J)::value, S = j >> 2, k = j & 3, par = j & 1;
    if (r == j)
#pragma unroll
      for (int s = 0; s <= S; ++s) {
        plow[par][q][s][0] = yr[s];
        plow[par][q][s][1] = yi[s];
      }
    __syncthreads();
    const T mr = quad_bcast<k>(lr[S]), mi = quad_bcast<k>(li[S]);
    if (r > j)
#pragma unroll
      for (int s = 0; s <= S; ++s)
        if (s < S || q <= k) {
          const T pr = plow[par][q][s][0], pi = plow[par][q][s][1];
          yr[s] -= mr * pr - mi * pi;
          yi[s] -= mr * pi + mi * pr;
        }
#pragma unroll
    for (int s = 0; s <= S; ++s) __asm__ __volatile__("" : "+v"(yr[s]), "+v"(yi[s]));
  });
#pragma unroll
  for (int s = 0; s < 16; ++s) {
    const int c = 4 * s + q;
    linv[2 * (r * PNB + c)] = yr[s];
    linv[2 * (r * PNB + c) + 1] = yi[s];
  }
}

// Rows below the diagonal block of a leaf: L21 = A21 U11^-1 by substitution
// (U11 staged in LDS by column class; no barrier inside the column loop), each
// column checked against izamax's choice: no row's |re| + |im| may exceed the
// diagonal's. 64 rows per workgroup, threads laid out as in diag_nopiv_kernel.
template <typename T>
__device__ __forceinline__ void below_nopiv_body(T* A, int64_t n, int64_t lda, int64_t j0, int b, int* ctl,
                                                 int64_t rblock) {
  __shared__ T us[PNB][4][16][2];  // U11 row j by column class
  __shared__ T rcp[PNB][2];        // 1 / u_jj
  const int t = threadIdx.x, r = t >> 2, q = t & 3;
  const int64_t row = j0 + b + rblock * PNB + r;
  T xr[16], xi[16];
  {
    const T* src = A + 2 * ((row < n ? row : j0) * lda + j0 + q);  // rows beyond n: an in-bounds row, masked
#pragma unroll
    for (int s = 0; s < 16; ++s) {
      xr[s] = src[8 * s];
      xi[s] = src[8 * s + 1];
    }
  }
  // U11 (identity beyond the upper part): every load issued first, from in-bounds
  // addresses, then selected and staged (a load per iteration followed by its use
  // would put 16 memory round trips in a row)
  T gr[PNB * PNB / 256], gi[PNB * PNB / 256];
#pragma unroll
  for (int it = 0; it < PNB * PNB / 256; ++it) {
    const int e = t + 256 * it, i = e >> 6, c = e & 63;
    const int64_t o = 2 * ((j0 + min(i, b - 1)) * lda + j0 + min(c, b - 1));
    gr[it] = A[o];
    gi[it] = A[o + 1];
  }
#pragma unroll
  for (int it = 0; it < PNB * PNB / 256; ++it) {
    const int e = t + 256 * it, i = e >> 6, c = e & 63;
    const bool in = i < b && c < b && c >= i;
    const T vr = in ? gr[it] : (i == c ? T(1) : T(0)), vi = in ? gi[it] : T(0);
    us[i][c & 3][c >> 2][0] = vr;
    us[i][c & 3][c >> 2][1] = vi;
    if (i == c) pivot_reciprocal(vr, vi, rcp[i][0], rcp[i][1]);
  }
#pragma unroll
  for (int s = 0; s < 16; ++s) {
    const bool in = row < n && 4 * s + q < b;
    xr[s] = in ? xr[s] : T(0);
    xi[s] = in ? xi[s] : T(0);
    __asm__ __volatile__("" : "+v"(xr[s]), "+v"(xi[s]));  // the masks applied here, not at first use
  }
  __syncthreads();
  bool viol = false;
  static_for<0, PNB>([&](auto J) {
    constexpr int j = decltype(J)::value, S = j >> 2, k = j & 3;
    const T ur = us[j][k][S][0], ui = us[j][k][S][1];
    const T ir = rcp[j][0], ii = rcp[j][1];
    const T mr = quad_bcast<k>(xr[S] * ir - xi[S] * ii), mi = quad_bcast<k>(xr[S] * ii + xi[S] * ir);
    if (q == k) {
      viol |= cabs1(xr[S], xi[S]) > cabs1(ur, ui);
      xr[S] = mr;
      xi[S] = mi;
    }
#pragma unroll
    for (int s = S; s < 16; ++s)
      if (s > S || q > k) {
        const T pr = us[j][q][s][0], pi = us[j][q][s][1];
        xr[s] -= mr * pr - mi * pi;
        xi[s] -= mr * pi + mi * pr;
      }
#pragma unroll
    for (int s = S; s < 16; ++s) __asm__ __volatile__("" : "+v"(xr[s]), "+v"(xi[s]));
    __asm__ __volatile__("" ::: "memory");  // and the next step's LDS reads stay in that step
  });
  if (row < n) {
    T* dst = A + 2 * (row * lda + j0 + q);
#pragma unroll
    for (int s = 0; s < 16; ++s)
      if (4 * s + q < b) {
        dst[8 * s] = xr[s];
        dst[8 * s + 1] = xi[s];
      }
  }
  if (__ballot(viol) != 0 && (t & 63) == 0) __hip_atomic_fetch_or(ctl + 2, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The rows below the diagonal block, one workgroup per 64 rows, and one more
// workgroup (the last) that forms the leaf inverse L11^-1 beside them.
template <typename T>
__global__ __launch_bounds__(256) void below_nopiv_kernel(T* A, int64_t n, int64_t lda, int64_t j0, int b, int* ctl,
                                                          T* linv) {
  __builtin_amdgcn_s_setprio(3);
  if (blockIdx.x == gridDim.x - 1) {
    linv_body(A, lda, j0, b, linv);
    return;
  }
  below_nopiv_body(A, n, lda, j0, b, ctl, (int64_t)blockIdx.x);
}

// Leaf t of every system of a batch (systems with fewer columns exit at once).
template <typename T>
__global__ __launch_bounds__(256) void diag_nopiv_batched_kernel(const BatchSystem<T>* sys, int t) {
  const BatchSystem<T> S = sys[blockIdx.x];
  const int64_t j0 = (int64_t)t * PNB;
  if (j0 >= S.n) return;
  const int b = (int)min((int64_t)PNB, S.n - j0);
  diag_nopiv_body(S.A, S.lda, j0, b, S.linv + 2 * (int64_t)t * PNB * PNB, S.status, S.status + 1);
}
template <typename T>
__global__ __launch_bounds__(256) void below_nopiv_batched_kernel(const BatchSystem<T>* sys, int t) {
  const BatchSystem<T> S = sys[blockIdx.y];
  const int64_t j0 = (int64_t)t * PNB;
  if (j0 >= S.n) return;
  const int b = (int)min((int64_t)PNB, S.n - j0);
  if ((int64_t)blockIdx.x * PNB >= S.n - j0 - b) return;
  below_nopiv_body(S.A, S.n, S.lda, j0, b, S.status + 1, (int64_t)blockIdx.x);
}

}  // namespace

// Grids whose workgroups wait on each other must be resident all at once. A
// plain launch on the caller's stream keeps such a kernel on that stream's
// hardware queue (no cross-queue dependency before and after it); it is used
// when the grid fits on the device at the kernel's occupancy, so every
// workgroup is placed once the (finite, non-waiting) work of other streams that
// holds resources drains. Larger grids go through the cooperative launch, which
// checks residency itself.
__global__ void iota_kernel(int* p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = (int)i;
}

template <typename Args>
static hipError_t launch_resident(const void* kernel, void (*fn)(Args), int grid, int block, Args& args,
                                  hipStream_t s) {
  static std::atomic<int> cap[16];
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  int c = dev < 16 ? cap[dev].load() : 0;
  if (c <= 0) {
    int cus = 0, per = 0;
    if ((e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess) return e;
    if ((e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, block, 0)) != hipSuccess) return e;
    c = cus * per;
    if (dev < 16) cap[dev].store(c);
  }
  if (grid <= c) {
    hipLaunchKernelGGL(fn, dim3(grid), dim3(block), 0, s, args);
    return hipGetLastError();
  }
  void* p[] = {&args};
  return hipLaunchCooperativeKernel(kernel, dim3(grid), dim3(block), p, 0, s);
}

template <typename T>
static hipError_t launch_panel(T* A, int64_t n, int64_t lda, int64_t j0, int b, LuWork& w, T* linv, hipStream_t s) {
  const int G = (int)((n - j0 + RPW - 1) / RPW);
  if (G > w.maxg) return hipErrorInvalidValue;
  PanelArgs<T> pa;
  pa.A = A;
  pa.n = n;
  pa.lda = lda;
  pa.j0 = j0;
  pa.nb = b;
  pa.ipiv = w.ipiv;
  pa.slots = w.slots;
  pa.slot_rows = reinterpret_cast<unsigned long long*>(w.slot_rows);
  pa.slot_diag = reinterpret_cast<unsigned long long*>(w.slot_diag);
  pa.maxg = w.maxg;
  pa.info = w.info;
  pa.ctl = w.ctl;
  pa.linv = linv;
  return launch_resident<PanelArgs<T>>(reinterpret_cast<const void*>(&panel_kernel<T>), panel_kernel<T>, G, 256, pa, s);
}

// The triangular solves wait only for workgroups of lower index, which the
// dispatcher places first: the lowest unfinished one is always resident, so the
// grid need not be.
template <typename Args>
static hipError_t launch_chain(void (*fn)(Args), int grid, Args& args, hipStream_t s) {
  hipLaunchKernelGGL(fn, dim3(grid), dim3(256), 0, s, args);
  return hipGetLastError();
}

// The speculative leaf: diagonal block, then the rows below it.
template <typename T>
static hipError_t launch_nopiv_leaf(T* A, int64_t n, int64_t lda, int64_t j0, int b, LuWork& w, T* linv,
                                    hipStream_t s) {
  const int64_t below = n - j0 - b;
  if (below <= 0) {
    // the last leaf: its inverse in the diagonal kernel
    hipLaunchKernelGGL((diag_nopiv_kernel<T, true>), dim3(1), dim3(256), 0, s, A, lda, j0, b, linv, w.info, w.ctl);
    return hipGetLastError();
  }
  // the inverse off the chain, in an extra workgroup of the rows-below kernel
  hipLaunchKernelGGL((diag_nopiv_kernel<T, false>), dim3(1), dim3(256), 0, s, A, lda, j0, b, linv, w.info, w.ctl);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((below_nopiv_kernel<T>), dim3((unsigned)((below + PNB - 1) / PNB + 1)), dim3(256), 0, s, A, n,
                     lda, j0, b, w.ctl, linv);
  return hipGetLastError();
}

template <typename T>
static hipError_t launch_laswp(T* A, int64_t lda, int64_t c0, int64_t c1, const int* ipiv, int64_t k0, int64_t k1,
                               hipStream_t s) {
  if (c1 <= c0 || k1 <= k0) return hipSuccess;
  for (int64_t kb = k0; kb < k1; kb += MAXBW) {
    hipLaunchKernelGGL((laswp_kernel<T>), dim3((unsigned)((c1 - c0 + SWP_COLS - 1) / SWP_COLS)), dim3(256), 0, s, A, lda,
                       c0, c1, ipiv, kb, std::min<int64_t>(k1, kb + MAXBW));
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

// Outer blocks of MAXBW columns factored as nb-wide leaves; the block's swaps,
// U12 by leaf inverses and one K = MAXBW GEMM then update the trailing matrix.
//
// Streams (when the context has them; otherwise everything runs in order on s):
//  s      panels, each leaf's update of the next leaf's columns (swaps, U12,
//         GEMM: the only in-block work between two panels), and the block's
//         update of the next block's columns;
//  inb    the rest of each leaf's in-block work, concurrent with the next panel:
//         its swaps on the block's earlier columns and its update of the leaves
//         after the next one. Updates of one column range stay in leaf order: s
//         waits for inb's work of leaf t-1 before it applies leaf t to leaf t+1.
//  side   block J's update of the columns beyond the next block (look-ahead),
//         while s updates the next block's columns and factors it; s waits for
//         that far update before it updates the block after.
// Leaf inverses alternate between two buffers by block parity, so a far update
// still reading block J's inverses never sees block J+1's.
// Columns [n, nc) (nc <= lda) are right-hand sides carried along: every block's
// swaps, U12 and trailing update reach them as they reach the trailing columns,
// so on return they hold L^-1 P b, the forward solve done by the factorization.
template <typename T>
hipError_t lu_factor(T* A, int64_t n, int64_t lda, LuWork& w, int nb, hipStream_t s, int* panels_out, int64_t nc) {
  if (nb > PNB) nb = PNB;
  const int bw = MAXBW;
  const int64_t ncol = nc > n ? nc : n;
  if (ncol > lda) return hipErrorInvalidValue;
  const bool g3 = w.gauss3;  // complex64 updates in the three-real-product form
  const bool nopiv = w.nopiv;  // speculative leaves: no exchanges (row swaps are no-ops)
  const int tile = w.gemm_tile;  // trailing-update tile shape
  if (tile < 0 || tile >= GEMM_TILES) return hipErrorInvalidValue;
  hipError_t e;
  // zero-pivot report and wait control (info, abort word, give-up record)
  if ((e = launch_fill<int>(w.info, 4, 0, s)) != hipSuccess) return e;
  if (w.abort_init && (e = launch_fill<int>(w.ctl, 1, w.abort_init, s)) != hipSuccess) return e;
  if (w.nopiv && w.force_exchange && (e = launch_fill<int>(w.ctl + 2, 1, 1, s)) != hipSuccess) return e;
  // pivots start as no-op exchanges: columns a give-up leaves unfactored then swap
  // nothing in the row exchanges and gather tables that follow
  hipLaunchKernelGGL(iota_kernel, dim3((unsigned)std::min<int64_t>(1024, (n + 255) / 256)), dim3(256), 0, s, w.ipiv, n);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  // slot words and published rows carry column tags: clear both
  if ((e = launch_fill<int>(reinterpret_cast<int*>(w.slots), 2 * SLOT_WORDS * 2 * (int64_t)w.maxg, 0, s)) != hipSuccess)
    return e;
  if ((e = launch_fill<int>(reinterpret_cast<int*>(w.slot_rows), (int64_t)2 * w.maxg * PNB * 4 * 2, 0, s)) != hipSuccess)
    return e;
  if ((e = launch_fill<int>(reinterpret_cast<int*>(w.slot_diag), (int64_t)2 * PNB * 4 * 2, 0, s)) != hipSuccess) return e;
  int panels = 0;
  const T* cA = A;
  T* linv = reinterpret_cast<T*>(w.linv);
  const bool look_ahead = w.side != nullptr;
  const bool split = w.inb != nullptr;
  auto leaf_inv = [&](int64_t j0, int64_t J) {
    return linv + 2 * ((((J / bw) & 1) * (bw / nb)) + (j0 - J) / nb) * PNB * PNB;
  };
  // leaf [j0, j0 + b) of block J: its update of columns [c0, c1) (swaps, U12, GEMM)
  auto leaf_update = [&](int64_t J, int64_t j0, int b, int64_t c0, int64_t c1, hipStream_t st) -> hipError_t {
    if (c1 <= c0) return hipSuccess;
    hipError_t r;
    if (!nopiv && (r = launch_laswp<T>(A, lda, c0, c1, w.ipiv, j0, j0 + b, st)) != hipSuccess) return r;
    r = gemm_set<T>(leaf_inv(j0, J), PNB, cA + 2 * (j0 * lda + c0), lda, A + 2 * (j0 * lda + c0), lda, b, c1 - c0, b,
                    st, false, g3, tile);
    if (r != hipSuccess) return r;
    return gemm_sub<T>(cA + 2 * ((j0 + b) * lda + j0), lda, cA + 2 * (j0 * lda + c0), lda, A + 2 * ((j0 + b) * lda + c0),
                       lda, n - j0 - b, c1 - c0, b, st, false, g3, tile);
  };
  // block [J, Je)'s update of columns [c0, c1): swaps, U12 by leaf inverses, GEMM
  auto update = [&](int64_t J, int64_t Je, int64_t c0, int64_t c1, hipStream_t st) -> hipError_t {
    if (c1 <= c0) return hipSuccess;
    const bool behind = look_ahead && st == w.side;  // the look-ahead's far update
    hipError_t r;
    if (!nopiv && (r = launch_laswp<T>(A, lda, c0, c1, w.ipiv, J, Je, st)) != hipSuccess) return r;
    for (int64_t j0 = J; j0 < Je; j0 += nb) {
      const int b = (int)std::min<int64_t>(nb, Je - j0);
      r = gemm_set<T>(leaf_inv(j0, J), PNB, cA + 2 * (j0 * lda + c0), lda, A + 2 * (j0 * lda + c0), lda, b, c1 - c0,
                      b, st, behind, g3, tile);
      if (r != hipSuccess) return r;
      if (j0 + b < Je) {
        r = gemm_sub<T>(cA + 2 * ((j0 + b) * lda + j0), lda, cA + 2 * (j0 * lda + c0), lda,
                        A + 2 * ((j0 + b) * lda + c0), lda, Je - j0 - b, c1 - c0, b, st, behind, g3, tile);
        if (r != hipSuccess) return r;
      }
    }
    return gemm_sub<T>(cA + 2 * (Je * lda + J), lda, cA + 2 * (J * lda + c0), lda, A + 2 * (Je * lda + c0), lda,
                       n - Je, c1 - c0, (int)(Je - J), st, behind, g3, tile);
  };
  // In the factorization's tail (at most w.tail_rows trailing rows after a block)
  // the look-ahead gains nothing, its updates being small, while every
  // cross-stream event puts tens of microseconds of latency on the critical path:
  // the tail's blocks run in order on s.
  // The side stream's far update of a block waits only for the block's factors on
  // s (ev_fact, recorded right after them), not for s to have seen the previous
  // far update end: the far updates follow each other on the side stream without
  // a cross-stream hop between them. s waits for each far update (alternate
  // blocks record ev_far and ev_far2, so that the wait names that update and not
  // the next one) before its own update of the same columns.
  bool far_pending = false;  // a far update on the side stream that s has not yet waited for
  hipEvent_t far_ev = w.ev_far, far_next = w.ev_far2;
  for (int64_t J = 0; J < n; J += bw) {
    const int64_t Je = std::min<int64_t>(n, J + bw);
    const bool tail = n - Je <= w.tail_rows;
    const bool split_b = split && !tail;
    for (int64_t j0 = J; j0 < Je; j0 += nb) {
      const int b = (int)std::min<int64_t>(nb, Je - j0);
      const int64_t nx = std::min<int64_t>(Je, j0 + b + nb);  // end of the next leaf's columns
      e = nopiv ? launch_nopiv_leaf<T>(A, n, lda, j0, b, w, leaf_inv(j0, J), s)
                : launch_panel<T>(A, n, lda, j0, b, w, leaf_inv(j0, J), s);
      if (e != hipSuccess) return e;
      ++panels;
      if (!split_b) {
        if (!nopiv && (e = launch_laswp<T>(A, lda, J, j0, w.ipiv, j0, j0 + b, s)) != hipSuccess) return e;
        if ((e = leaf_update(J, j0, b, j0 + b, Je, s)) != hipSuccess) return e;
        continue;
      }
      if ((e = hipEventRecord(w.ev_panel, s)) != hipSuccess) return e;
      if (j0 > J && nx > j0 + b)  // inb's work of the previous leaf touched these columns
        if ((e = hipStreamWaitEvent(s, w.ev_inb, 0)) != hipSuccess) return e;
      if ((e = leaf_update(J, j0, b, j0 + b, nx, s)) != hipSuccess) return e;
      if ((e = hipStreamWaitEvent(w.inb, w.ev_panel, 0)) != hipSuccess) return e;
      if (!nopiv && (e = launch_laswp<T>(A, lda, J, j0, w.ipiv, j0, j0 + b, w.inb)) != hipSuccess) return e;
      if ((e = leaf_update(J, j0, b, nx, Je, w.inb)) != hipSuccess) return e;
      if ((e = hipEventRecord(w.ev_inb, w.inb)) != hipSuccess) return e;
    }
    if (split_b && (e = hipStreamWaitEvent(s, w.ev_inb, 0)) != hipSuccess) return e;
    const int64_t Kn = std::min<int64_t>(n, Je + bw);
    const bool far = Je < n && look_ahead && !tail && Kn < ncol;
    if (far) {
      // this block's far update (columns beyond the next block) on the side stream,
      // behind the previous one there
      if ((e = hipEventRecord(w.ev_fact, s)) != hipSuccess) return e;
      if ((e = hipStreamWaitEvent(w.side, w.ev_fact, 0)) != hipSuccess) return e;
      if ((e = update(J, Je, Kn, ncol, w.side)) != hipSuccess) return e;
      if ((e = hipEventRecord(far_next, w.side)) != hipSuccess) return e;
    }
    // the previous block's far update (of every column beyond this block) comes
    // before any update of those columns on s
    if (far_pending) {
      if ((e = hipStreamWaitEvent(s, far_ev, 0)) != hipSuccess) return e;
      far_pending = false;
    }
    if (far) {
      std::swap(far_ev, far_next);
      far_pending = true;
    }
    if (Je >= n) {
      // the last block's update of the carried right-hand sides
      if (ncol > n && (e = update(J, Je, n, ncol, s)) != hipSuccess) return e;
      continue;
    }
    if (look_ahead && !tail) {
      if ((e = update(J, Je, Je, Kn, s)) != hipSuccess) return e;
    } else {
      if ((e = update(J, Je, Je, ncol, s)) != hipSuccess) return e;
    }
  }
  if (far_pending && (e = hipStreamWaitEvent(s, far_ev, 0)) != hipSuccess) return e;
  if (panels_out) *panels_out = panels;
  return hipSuccess;
}

// Solve workspace (complex T units, then ints): the chain matrices (per direction
// and leaf the inverse D and the link W, PNB x PNB each), three n x 8 blocks
// (gathered rhs, forward results, backward results), the forward and backward
// results as tagged words (twice a complex entry's size), then the row gather
// tables (nJ x n), the total row permutation (n), the forward and backward flags
// (np each) and the per-block row maps the tables are built from (nJ x n).
struct SolveLayout {
  int np, nJ;
  size_t mat, b0, yf, xb, tf, tb, cend;  // complex offsets
};
static SolveLayout solve_layout(int64_t n) {
  SolveLayout L;
  L.np = (int)((n + PNB - 1) / PNB);
  L.nJ = (int)((n + MAXBW - 1) / MAXBW);
  // chain matrices: 2 directions x 2 kinds (solve_blocks_kernel) x np blocks
  L.mat = 0;
  L.b0 = L.mat + (size_t)4 * L.np * PNB * PNB;
  L.yf = L.b0 + (size_t)n * 8;
  L.xb = L.yf + (size_t)n * 8;
  // tagged results (TaggedRow: twice a complex entry's size), forward and backward
  L.tf = L.xb + (size_t)n * 8;
  L.tb = L.tf + (size_t)2 * n * 8;
  L.cend = L.tb + (size_t)2 * n * 8;
  return L;
}

size_t lu_solve_work_bytes(int64_t n, size_t elem) {
  const SolveLayout L = solve_layout(n);
  return elem * 2 * L.cend + 256 + sizeof(int) * (2 * (size_t)L.nJ * n + n + 2 * (size_t)L.np) + 256;
}

template <typename T>
static int* solve_ints(T* work, const SolveLayout& L) {
  const uintptr_t p = reinterpret_cast<uintptr_t>(work + 2 * L.cend);
  return reinterpret_cast<int*>((p + 255) & ~uintptr_t(255));
}

static int next_epoch() {
  static std::atomic<int> counter{0};
  int e = counter.fetch_add(1) + 1;
  return (e & 0x7fffffff) == 0 ? 1 : (e & 0x7fffffff);
}

template <typename T>
hipError_t lu_solve_prepare(const T* LU, int64_t n, int64_t lda, const int* ipiv, T* work, hipStream_t s) {
  const SolveLayout L = solve_layout(n);
  int* G = solve_ints(work, L);
  int* perm = G + (size_t)L.nJ * n;
  int* flags = perm + n;
  hipError_t e;
  int* R = perm + n + 2 * L.np;  // after the flags: nJ x n block maps
  hipLaunchKernelGGL(identity_maps_kernel, dim3((unsigned)std::min<int64_t>(4096, ((int64_t)L.nJ * n + 255) / 256)),
                     dim3(256), 0, s, n, L.nJ, R);
  hipLaunchKernelGGL(block_map_kernel, dim3((unsigned)L.nJ), dim3(256), 0, s, ipiv, n, MAXBW, R);
  hipLaunchKernelGGL(gather_table_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, (const int*)R, n, L.nJ, G,
                     perm);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL((solve_blocks_kernel<T>), dim3((unsigned)L.np, 4), dim3(256), 0, s, LU, n, lda, (const int*)G,
                     MAXBW, L.np, work + 2 * L.mat);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  // tagged results: no word may carry a later solve's epoch by accident
  if ((e = launch_fill<int>(reinterpret_cast<int*>(work + 2 * L.tf), (int64_t)4 * 2 * n * 8 * sizeof(T) / sizeof(int),
                            0, s)) != hipSuccess)
    return e;
  return launch_fill<int>(flags, 2 * (int64_t)L.np, 0, s);
}

template <typename T>
hipError_t lu_solve(const T* LU, int64_t n, int64_t lda, T* B, int nrhs, T* work, int* ctl, hipStream_t s) {
  if (nrhs > 8 || nrhs < 1) return hipErrorInvalidValue;
  const SolveLayout L = solve_layout(n);
  int* G = solve_ints(work, L);
  int* perm = G + (size_t)L.nJ * n;
  int* flags = perm + n;
  hipLaunchKernelGGL((gather_rhs_kernel<T>), dim3((unsigned)((n * 8 + 255) / 256)), dim3(256), 0, s, (const T*)B, n,
                     nrhs, (const int*)perm, work + 2 * L.b0);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const int epoch = next_epoch();
  TriArgs<T> ta;
  ta.LU = LU;
  ta.n = n;
  ta.lda = lda;
  ta.bw = MAXBW;
  ta.epoch = epoch;
  ta.ctl = ctl;
  ta.np = L.np;
  ta.nrhs = nrhs;
  ta.out = B;
  // forward
  ta.mats = work + 2 * L.mat;
  ta.gather = G;
  ta.src = work + 2 * L.b0;
  ta.pub = work + 2 * L.yf;
  ta.flags = flags;
  ta.tpub = reinterpret_cast<unsigned long long*>(work + 2 * L.tf);
  e = launch_chain(tri_solve_kernel<T, true>, L.np, ta, s);
  if (e != hipSuccess) return e;
  // backward
  ta.mats = work + 2 * (L.mat + (size_t)2 * L.np * PNB * PNB);
  ta.gather = nullptr;
  ta.src = work + 2 * L.yf;
  ta.pub = work + 2 * L.xb;
  ta.flags = flags + L.np;
  ta.tpub = reinterpret_cast<unsigned long long*>(work + 2 * L.tb);
  return launch_chain(tri_solve_kernel<T, false>, L.np, ta, s);
}

// Right-hand sides carried by the factorization: in (n x nrhs) to the LU's
// columns [n, n + nrhs) before it, and L^-1 P b from there to the backward
// solve's input (n x 8, zero-padded) after it.
template <typename T>
__global__ void rhs_to_cols_kernel(const T* B, int64_t n, int nrhs, T* A, int64_t lda) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n * nrhs) return;
  const int64_t r = idx / nrhs;
  const int d = (int)(idx % nrhs);
  A[2 * (r * lda + n + d)] = B[2 * idx];
  A[2 * (r * lda + n + d) + 1] = B[2 * idx + 1];
}

template <typename T>
__global__ void cols_to_rhs8_kernel(const T* A, int64_t n, int64_t lda, int nrhs, T* Y) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n * 8) return;
  const int64_t r = idx >> 3;
  const int d = (int)(idx & 7);
  Y[2 * idx] = d < nrhs ? A[2 * (r * lda + n + d)] : T(0);
  Y[2 * idx + 1] = d < nrhs ? A[2 * (r * lda + n + d) + 1] : T(0);
}

template <typename T>
hipError_t lu_append_rhs(T* A, int64_t n, int64_t lda, const T* B, int nrhs, hipStream_t s) {
  if (nrhs < 1 || n + nrhs > lda) return hipErrorInvalidValue;
  hipLaunchKernelGGL((rhs_to_cols_kernel<T>), dim3((unsigned)((n * nrhs + 255) / 256)), dim3(256), 0, s, B, n, nrhs,
                     A, lda);
  return hipGetLastError();
}

template <typename T>
hipError_t lu_solve_eliminated(const T* LU, int64_t n, int64_t lda, T* B, int nrhs, T* work, int* ctl,
                               hipStream_t s) {
  if (nrhs > 8 || nrhs < 1 || n + nrhs > lda) return hipErrorInvalidValue;
  const SolveLayout L = solve_layout(n);
  int* G = solve_ints(work, L);
  int* flags = G + (size_t)L.nJ * n + n;
  hipLaunchKernelGGL((cols_to_rhs8_kernel<T>), dim3((unsigned)((n * 8 + 255) / 256)), dim3(256), 0, s, LU, n, lda,
                     nrhs, work + 2 * L.yf);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  TriArgs<T> ta;
  ta.LU = LU;
  ta.n = n;
  ta.lda = lda;
  ta.bw = MAXBW;
  ta.epoch = next_epoch();
  ta.ctl = ctl;
  ta.np = L.np;
  ta.nrhs = nrhs;
  ta.out = B;
  ta.mats = work + 2 * (L.mat + (size_t)2 * L.np * PNB * PNB);
  ta.gather = nullptr;
  ta.src = work + 2 * L.yf;
  ta.pub = work + 2 * L.xb;
  ta.flags = flags + L.np;
  ta.tpub = reinterpret_cast<unsigned long long*>(work + 2 * L.tb);
  return launch_chain(tri_solve_kernel<T, false>, L.np, ta, s);
}

template hipError_t lu_append_rhs<float>(float*, int64_t, int64_t, const float*, int, hipStream_t);
template hipError_t lu_append_rhs<double>(double*, int64_t, int64_t, const double*, int, hipStream_t);
template hipError_t lu_solve_eliminated<float>(const float*, int64_t, int64_t, float*, int, float*, int*, hipStream_t);
template hipError_t lu_solve_eliminated<double>(const double*, int64_t, int64_t, double*, int, double*, int*,
                                                hipStream_t);

size_t residual64_part_bytes(int64_t n) { return sizeof(double) * 16 * (size_t)RES_SPLIT_MAX * (size_t)n; }

hipError_t residual64(const double* A, int64_t n, int64_t lda, const double* x, const double* b, double* r, int nrhs,
                      double* part, bool split, hipStream_t s) {
  if (nrhs > 8 || nrhs < 1) return hipErrorInvalidValue;
  if (n <= 0) return hipSuccess;
  int ks = 1;
  if (split) {
    // alone on the GPU: residual64_split_kernel with k ranges per row block, the
    // count (up to RES_SPLIT_MAX) whose grid comes closest to whole rounds of one
    // workgroup per CU, keeping at least 8 chunks per range; the fewest ranges
    // among equals
    constexpr int rows = RES_WAVES * RES_ROWS;
    const int64_t nblk = (n + rows - 1) / rows;
    static std::atomic<int> cus_cached{0};
    int cus = cus_cached.load();
    if (cus <= 0) {
      int dev = 0;
      if (hipGetDevice(&dev) != hipSuccess ||
          hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
        cus = 256;
      cus_cached.store(cus);
    }
    const int64_t nch = (n + RS_KC - 1) / RS_KC;
    double best = 0;
    for (int k = 1; k <= RES_SPLIT_MAX && nch / k >= 8; ++k) {
      const double wg = (double)(nblk * k), rounds = std::ceil(wg / cus), eff = wg / (rounds * cus);
      if (eff > best + 1e-3) {
        best = eff;
        ks = k;
      }
    }
    hipLaunchKernelGGL(residual64_split_kernel, dim3((unsigned)nblk, (unsigned)ks), dim3(64 * RES_WAVES), 0, s, A, n,
                       lda, x, nrhs, part);
  } else {
    // beside other streams' long kernels (a sweep's next assembly): the small
    // workgroups of residual64_narrow_kernel, one k range
    constexpr int rows = RS_WAVES * RS_ROWS;
    const int64_t nblk = (n + rows - 1) / rows;
    hipLaunchKernelGGL(residual64_narrow_kernel, dim3((unsigned)nblk, 1), dim3(64 * RS_WAVES), 0, s, A, n, lda, x,
                       nrhs, part);
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(residual64_finish_kernel, dim3((unsigned)((n * nrhs * 2 + 255) / 256)), dim3(256), 0, s,
                     (const double*)part, ks, n, b, r, nrhs);
  return hipGetLastError();
}

hipError_t column_norms64(const double* X, int64_t n, int nrhs, double* out, hipStream_t s) {
  // out must hold nrhs + 64*nrhs doubles: norms first, then the partial sums
  const int parts = 64;
  double* part = out + nrhs;
  hipLaunchKernelGGL(colnorm_partial_kernel, dim3(parts), dim3(256), 0, s, X, n, nrhs, part);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(colnorm_final_kernel, dim3(1), dim3(64), 0, s, (const double*)part, parts, nrhs, out);
  return hipGetLastError();
}

template hipError_t lu_factor<float>(float*, int64_t, int64_t, LuWork&, int, hipStream_t, int*, int64_t);
template hipError_t lu_factor<double>(double*, int64_t, int64_t, LuWork&, int, hipStream_t, int*, int64_t);

// Solve A^T x = c with the factors of lu_factor (A = P L U, so U^T L^T P^T x = c),
// one workgroup of 1024 threads: U^T z = c forward and L^T y = z backward by
// 64-column blocks (diagonal block in LDS, solved by one wavefront; the rest of
// the vector updated from contiguous rows of U and of L in its final row order,
// read through the gather tables of lu_solve_prepare), then x[perm[r]] = y[r].
// v (n complex) holds c on entry and is overwritten.
template <typename T>
__global__ __launch_bounds__(1024) void lu_solve_t_kernel(const T* LU, int64_t n, int64_t lda, const int* G,
                                                          const int* perm, T* v, T* x) {
  __shared__ T blk[2 * PNB * PNB];
  __shared__ T zb[2 * PNB];
  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  // forward: U^T z = c
  for (int64_t jb = 0; jb < n; jb += PNB) {
    const int nb = (int)min((int64_t)PNB, n - jb);
    for (int e = t; e < PNB * PNB; e += blockDim.x) {
      const int i = e / PNB, j = e - i * PNB;  // blk[i][j] = U[jb+i][jb+j]
      T re = 0, im = 0;
      if (i < nb && j < nb) {
        re = LU[2 * ((jb + i) * lda + jb + j)];
        im = LU[2 * ((jb + i) * lda + jb + j) + 1];
      }
      blk[2 * e] = re;
      blk[2 * e + 1] = im;
    }
    __syncthreads();
    if (wave == 0) {
      T vr = lane < nb ? v[2 * (jb + lane)] : T(0), vi = lane < nb ? v[2 * (jb + lane) + 1] : T(0);
      for (int j = 0; j < nb; ++j) {
        // z_j = v_j / U_jj on lane j, broadcast
        T zr = 0, zi = 0;
        if (lane == j) {
          const T dr = blk[2 * (j * PNB + j)], di = blk[2 * (j * PNB + j) + 1];
          const T den = dr * dr + di * di;
          zr = (vr * dr + vi * di) / den;
          zi = (vi * dr - vr * di) / den;
          vr = zr;
          vi = zi;
        }
        zr = __shfl(zr, j, 64);
        zi = __shfl(zi, j, 64);
        if (lane > j && lane < nb) {
          const T ur = blk[2 * (j * PNB + lane)], ui = blk[2 * (j * PNB + lane) + 1];
          vr -= ur * zr - ui * zi;
          vi -= ur * zi + ui * zr;
        }
      }
      if (lane < nb) {
        zb[2 * lane] = vr;
        zb[2 * lane + 1] = vi;
        v[2 * (jb + lane)] = vr;
        v[2 * (jb + lane) + 1] = vi;
      }
    }
    __syncthreads();
    for (int64_t i = jb + nb + t; i < n; i += blockDim.x) {
      T ar = 0, ai = 0;
      for (int j = 0; j < nb; ++j) {
        const T ur = LU[2 * ((jb + j) * lda + i)], ui = LU[2 * ((jb + j) * lda + i) + 1];
        const T zr = zb[2 * j], zi = zb[2 * j + 1];
        ar += ur * zr - ui * zi;
        ai += ur * zi + ui * zr;
      }
      v[2 * i] -= ar;
      v[2 * i + 1] -= ai;
    }
    __syncthreads();
  }
  // backward: L^T y = z, L unit lower in its final row order: L[r][c] is stored in
  // row G[c / MAXBW][r]
  const int64_t nblk = (n + PNB - 1) / PNB;
  for (int64_t b = nblk - 1; b >= 0; --b) {
    const int64_t jb = b * PNB;
    const int nb = (int)min((int64_t)PNB, n - jb);
    const int* Gj = G + (jb / MAXBW) * n;
    for (int e = t; e < PNB * PNB; e += blockDim.x) {
      const int i = e / PNB, j = e - i * PNB;  // blk[i][j] = L[jb+i][jb+j], i > j
      T re = 0, im = 0;
      if (i < nb && j < i) {
        const int64_t sr = Gj[jb + i];
        re = LU[2 * (sr * lda + jb + j)];
        im = LU[2 * (sr * lda + jb + j) + 1];
      }
      blk[2 * e] = re;
      blk[2 * e + 1] = im;
    }
    __syncthreads();
    if (wave == 0) {
      T vr = lane < nb ? v[2 * (jb + lane)] : T(0), vi = lane < nb ? v[2 * (jb + lane) + 1] : T(0);
      for (int j = nb - 1; j >= 0; --j) {
        const T yr = __shfl(vr, j, 64), yi = __shfl(vi, j, 64);
        if (lane < j) {
          const T lr = blk[2 * (j * PNB + lane)], li = blk[2 * (j * PNB + lane) + 1];
          vr -= lr * yr - li * yi;
          vi -= lr * yi + li * yr;
        }
      }
      if (lane < nb) {
        zb[2 * lane] = vr;
        zb[2 * lane + 1] = vi;
        v[2 * (jb + lane)] = vr;
        v[2 * (jb + lane) + 1] = vi;
      }
    }
    __syncthreads();
    for (int64_t i = t; i < jb; i += blockDim.x) {
      const int* Gi = G + (i / MAXBW) * n;
      T ar = 0, ai = 0;
      for (int j = 0; j < nb; ++j) {
        const int64_t sr = Gi[jb + j];
        const T lr = LU[2 * (sr * lda + i)], li = LU[2 * (sr * lda + i) + 1];
        const T yr = zb[2 * j], yi = zb[2 * j + 1];
        ar += lr * yr - li * yi;
        ai += lr * yi + li * yr;
      }
      v[2 * i] -= ar;
      v[2 * i + 1] -= ai;
    }
    __syncthreads();
  }
  for (int64_t r = t; r < n; r += blockDim.x) {
    const int64_t o = perm[r];
    x[2 * o] = v[2 * r];
    x[2 * o + 1] = v[2 * r + 1];
  }
}

template <typename T>
hipError_t lu_solve_transposed(const T* LU, int64_t n, int64_t lda, T* v, T* x, const T* work, hipStream_t s) {
  const SolveLayout L = solve_layout(n);
  const int* G = solve_ints(const_cast<T*>(work), L);
  const int* perm = G + (size_t)L.nJ * n;
  hipLaunchKernelGGL((lu_solve_t_kernel<T>), dim3(1), dim3(1024), 0, s, LU, n, lda, G, perm, v, x);
  return hipGetLastError();
}

template hipError_t lu_solve_transposed<float>(const float*, int64_t, int64_t, float*, float*, const float*,
                                               hipStream_t);
template hipError_t lu_solve_transposed<double>(const double*, int64_t, int64_t, double*, double*, const double*,
                                                hipStream_t);

template <typename T, bool G3, int BM, int KB>
static void gemm_batched_go(const BatchSystem<T>* sys, int nsys, int t, int64_t cols, int64_t below, hipStream_t s) {
  const dim3 gu((unsigned)((cols + 63) / 64), 1, (unsigned)nsys);
  const dim3 gs((unsigned)((cols + 63) / 64), (unsigned)std::max<int64_t>(1, (below + BM - 1) / BM), (unsigned)nsys);
  hipLaunchKernelGGL((zgemm_batched_kernel<T, true, G3, BM, KB>), gu, dim3(256), 0, s, sys, t);
  if (below > 0) hipLaunchKernelGGL((zgemm_batched_kernel<T, false, G3, BM, KB>), gs, dim3(256), 0, s, sys, t);
}

template <typename T, bool G3>
static void gemm_batched_shape(int tile, const BatchSystem<T>* sys, int nsys, int t, int64_t cols, int64_t below,
                               hipStream_t s) {
  switch (tile) {
    case 1: gemm_batched_go<T, G3, 64, 32>(sys, nsys, t, cols, below, s); break;
    case 2: gemm_batched_go<T, G3, 128, 16>(sys, nsys, t, cols, below, s); break;
    case 3: gemm_batched_go<T, G3, 128, 32>(sys, nsys, t, cols, below, s); break;
    default: gemm_batched_go<T, G3, 64, 16>(sys, nsys, t, cols, below, s); break;
  }
}

template <typename T>
hipError_t lu_factor_batched(const BatchSystem<T>* sys, int nsys, int64_t nmax, int64_t ncmax, bool gauss3,
                             int tile, hipStream_t s) {
  if (nsys <= 0 || nmax <= 0) return hipSuccess;
  if (tile < 0 || tile >= GEMM_TILES) return hipErrorInvalidValue;
  const int leaves = (int)((nmax + PNB - 1) / PNB);
  hipError_t e;
  for (int t = 0; t < leaves; ++t) {
    const int64_t j0 = (int64_t)t * PNB;
    const int64_t below = std::max<int64_t>(0, nmax - j0 - PNB);
    const int64_t cols = std::max<int64_t>(0, ncmax - j0 - 1);
    hipLaunchKernelGGL((diag_nopiv_batched_kernel<T>), dim3((unsigned)nsys), dim3(256), 0, s, sys, t);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if (below > 0) {
      hipLaunchKernelGGL((below_nopiv_batched_kernel<T>), dim3((unsigned)((below + PNB - 1) / PNB), (unsigned)nsys),
                         dim3(256), 0, s, sys, t);
      if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    if (cols <= 0) continue;
    if (sizeof(T) == 4 && gauss3) gemm_batched_shape<T, true>(tile, sys, nsys, t, cols, below, s);
    else gemm_batched_shape<T, false>(tile, sys, nsys, t, cols, below, s);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  return hipSuccess;
}

template <typename T>
hipError_t tri_solve_batched(const BatchSystem<T>* sys, const BatchVec<T>* vec, int nsys, int nrhs, bool fwd, bool bwd,
                             hipStream_t s) {
  if (nsys <= 0) return hipSuccess;
  if (nrhs < 1 || nrhs > 8) return hipErrorInvalidValue;
  if (fwd && bwd) hipLaunchKernelGGL((tri_wg_kernel<T, true, true>), dim3((unsigned)nsys), dim3(512), 0, s, sys, vec, nrhs);
  else if (fwd) hipLaunchKernelGGL((tri_wg_kernel<T, true, false>), dim3((unsigned)nsys), dim3(512), 0, s, sys, vec, nrhs);
  else if (bwd) hipLaunchKernelGGL((tri_wg_kernel<T, false, true>), dim3((unsigned)nsys), dim3(512), 0, s, sys, vec, nrhs);
  return hipGetLastError();
}

hipError_t residual64_batched(const BatchSystem<double>* sys, const BatchVec<double>* x, const BatchVec<double>* b,
                              const BatchVec<double>* r, int nsys, int64_t nmax, int nrhs, hipStream_t s) {
  if (nsys <= 0) return hipSuccess;
  if (nrhs > 8) return hipErrorInvalidValue;
  constexpr int rows = RES_WAVES * RES_ROWS;
  hipLaunchKernelGGL(residual64_batched_kernel, dim3((unsigned)((nmax + rows - 1) / rows), (unsigned)nsys),
                     dim3(64 * RES_WAVES), 0, s, sys, x, b, r, nrhs);
  return hipGetLastError();
}

template hipError_t lu_factor_batched<float>(const BatchSystem<float>*, int, int64_t, int64_t, bool, int,
                                            hipStream_t);
template hipError_t lu_factor_batched<double>(const BatchSystem<double>*, int, int64_t, int64_t, bool, int,
                                             hipStream_t);
template hipError_t tri_solve_batched<float>(const BatchSystem<float>*, const BatchVec<float>*, int, int, bool, bool,
                                             hipStream_t);
template hipError_t tri_solve_batched<double>(const BatchSystem<double>*, const BatchVec<double>*, int, int, bool, bool,
                                              hipStream_t);

template hipError_t lu_solve_prepare<float>(const float*, int64_t, int64_t, const int*, float*, hipStream_t);
template hipError_t lu_solve_prepare<double>(const double*, int64_t, int64_t, const int*, double*, hipStream_t);
template hipError_t lu_solve<float>(const float*, int64_t, int64_t, float*, int, float*, int*, hipStream_t);
template hipError_t lu_solve<double>(const double*, int64_t, int64_t, double*, int, double*, int*, hipStream_t);

}  // namespace beat

// Fused Burton-Miller assembly on gfx950.
//
//   A = 0.5 M_p1p1 - D + eta H          (N x N, P1 x P1)
//   B = (-S - eta (K' + 0.5 M_p1dp0)) q (N x nrhs)
//
// per element pair and quadrature point pair, the algebra of the reference's
// _beat_cpu_bm_regular_pair_blocks / _beat_cpu_bm_pair_blocks
// (src/beat_engine/julia_local/src/BeatEngineCpuBurtonMiller.jl:62-167):
//   lhs_s = -G' (r.n_y / r) w - eta k^2 (n_x.n_y) G w
//   rhs_s = -G w - eta G' (-r.n_x / r) w,     G' = G (-1/r + i k)
//   LHS_ab += phi_a(x) phi_b(y) lhs_s,  RHS_a += phi_a(x) rhs_s,
//   LHS_ab += (curl_a . curl_b) eta sum(G w)
// Regular pairs (no shared vertex) run in bm_regular_kernel: one workgroup per
// (test tile, trial tile), each lane owns one test face, waves stride over the
// trial faces; 3x3 blocks land in an LDS tile keyed by (test vertex slot, trial
// vertex slot) and one global atomic per touched entry flushes it. Touching
// pairs run in bm_singular_kernel in FP64 over the Duffy rules, one wavefront
// per pair with a wavefront reduction.
#include <hip/hip_runtime.h>

#include "bem_common.h"
#include "bem_math.h"
#include "kernels.h"

namespace beat {

namespace {

template <typename T>
__device__ __forceinline__ void lds_add(T* p, T v) {
  atomicAdd(p, v);
}

template <typename T>
__device__ __forceinline__ void gadd(T* p, T v) {
  unsafeAtomicAdd(p, v);
}

// sin and cos of k*r. FP32: the hardware sin/cos take revolutions; the phase is
// reduced to [-0.5, 0.5] revolutions first. FP64: the pi/64 table form on the
// phase in table steps (r times k 64/pi) when the host has bounded |k| r below
// its limit (WIDE false), else the pi/2 form with the library fallback.
template <bool WIDE>
__device__ __forceinline__ void phase_sincos(float r, float krev, double ksteps, float kk, const double* trig,
                                             float& s, float& c) {
  (void)ksteps;
  (void)kk;
  (void)trig;
  float t = r * krev;
  t = t - rintf(t);
  s = __builtin_amdgcn_sinf(t);
  c = __builtin_amdgcn_cosf(t);
}
template <bool WIDE>
__device__ __forceinline__ void phase_sincos(double r, double krev, double ksteps, double kk, const double* trig,
                                             double& s, double& c) {
  (void)krev;
  if (WIDE) sincos_phase(kk * r, &s, &c);
  else sincos_steps(r * ksteps, trig, &s, &c);
}

// The phase split in two, for software pipelining: phase_begin forms what needs
// no table value and issues the table read; phase_end finishes sin and cos.
template <typename T>
struct PhaseStage {
  T ir, y, S, C;  // 1/r; the table path: the remainder in steps and the table entry
};
template <bool WIDE>
__device__ __forceinline__ void phase_begin(float r, float krev, double ksteps, float kk, const double* trig,
                                            PhaseStage<float>& o) {
  phase_sincos<WIDE>(r, krev, ksteps, kk, trig, o.S, o.C);
}
template <bool WIDE>
__device__ __forceinline__ void phase_end(const PhaseStage<float>& o, float& s, float& c) {
  s = o.S;
  c = o.C;
}
template <bool WIDE>
__device__ __forceinline__ void phase_begin(double r, double krev, double ksteps, double kk, const double* trig,
                                            PhaseStage<double>& o) {
  if (WIDE) {
    phase_sincos<WIDE>(r, krev, ksteps, kk, trig, o.S, o.C);
  } else {
    // sincos_steps, split: t = r k 64/pi, n = rint(t), the remainder u = t - n
    const double t = r * ksteps;
    const double n = rint(t);
    o.y = t - n;
    const int m = (int)n & (TRIG_N - 1);
    o.S = trig[2 * m];
    o.C = trig[2 * m + 1];
  }
}
template <bool WIDE>
__device__ __forceinline__ void phase_end(const PhaseStage<double>& o, double& s, double& c) {
  if (WIDE) {
    s = o.S;
    c = o.C;
  } else {
    sincos_rotate_steps(o.y, o.S, o.C, &s, &c);
  }
}

__device__ __forceinline__ float inv_sqrt(float x) { return __builtin_amdgcn_rsqf(x); }
// 1/sqrt(x): the hardware estimate (relative error 5.2e-8 measured on gfx950 over
// r^2 in [1e-8, 1e2]) and one Newton step (measured 4.1e-15), far inside the
// 1e-12 agreement of the FP64 operators
__device__ __forceinline__ double inv_sqrt(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  const double e = fma(-(x * y), y, 1.0);
  return fma(0.5 * y, e, y);
}

}  // namespace

// One workgroup per (test tile, trial tile): 4 waves, lane = test face, each wave
// takes every 4th trial face. Trial-face data are read with wave-uniform
// addresses (scalar loads); 3x3 blocks go to the LDS tile accA[test vertex slot]
// [trial vertex slot] and the P1 x DP0 blocks to accC[test vertex slot][trial
// face]; both are flushed once with global atomics. Two workgroups fit per CU.
template <typename T, int NQ, bool PURE, bool WIDE, bool IMAGE, int TFS = TILE_FS, int TVS = TILE_VS>
__global__ __launch_bounds__(256, 2) void bm_regular_kernel(const RegularArgs<T> a) {
  // rows padded by one scalar: the 64 lanes add into rows tl (varying) at the same
  // column, which without the pad would all fall into one LDS bank
  constexpr int LDA_ = 2 * TVS + 1, LDC_ = 2 * TFS + 1;
  __shared__ T accA[TILE_V * LDA_];
  __shared__ T accC[TILE_V * LDC_];
  __shared__ T xs[NQ * 3 * 64];  // test points, [point*3 + c][lane]
  __shared__ T sphw[6 * 8];      // per rule point: phi_0..2, w, phi_0..2 w, unused
  __shared__ double trig[2 * TRIG_N];  // (sin, cos)(m pi/64)
  __shared__ int grow[TILE_V], gcol[TVS];

  const int64_t total = (int64_t)a.ntiles * a.nstiles;
  const int64_t per = (total + 7) / 8;
  int64_t bid = blockIdx.x;
  // image passes: the grid holds one copy of the tile-pair range per transform
  T sg[6] = {T(1), T(1), T(1), T(1), T(1), T(1)};
  if (IMAGE) {
    const int xi = (int)(bid / (per * 8));
    bid -= (int64_t)xi * per * 8;
#pragma unroll
    for (int x = 0; x < MAX_IMAGES; ++x)
      if (x == xi)
#pragma unroll
        for (int c = 0; c < 6; ++c) sg[c] = a.img[x][c];
  }
  const int64_t L = (bid & 7) * per + (bid >> 3);
  if (L >= total) return;
  // Each XCD walks its range of tile pairs test tile by test tile, streaming every
  // trial tile's records (5.2 MB FP64 at 20480 faces, more than an XCD's L2);
  // odd test tiles walk the trial tiles backwards, so they start on the ones the
  // previous test tile read last, which are still in L2.
  const int tt = (int)(L / a.nstiles), sl = (int)(L % a.nstiles);
  const int st = (tt & 1) ? a.nstiles - 1 - sl : sl;
  const int tid = threadIdx.x;
  const int t0 = a.tile_start[tt], nT = a.tile_count[tt], nvT = a.tile_nv[tt];
  const int s0 = a.stile_start[st], nS = a.stile_count[st], nvS = a.stile_nv[st];
  const T* __restrict__ rec = a.rec;
  const int32_t* __restrict__ fverts = a.fverts;
  const uint8_t* __restrict__ flocs = a.flocs;

  for (int i = tid; i < TILE_V * LDA_; i += blockDim.x) accA[i] = T(0);
  for (int i = tid; i < TILE_V * LDC_; i += blockDim.x) accC[i] = T(0);
  if (tid < TILE_V) grow[tid] = a.tile_verts[(int64_t)tt * TILE_V + tid];
  if (tid < TVS) gcol[tid] = a.stile_verts[(int64_t)st * TVS + tid];
  for (int i = tid; i < 2 * TRIG_N; i += blockDim.x) trig[i] = a.trig[i];
  if (tid == 0) {
#pragma unroll
    for (int i = 0; i < 6; ++i) {
#pragma unroll
      for (int b = 0; b < 3; ++b) {
        sphw[8 * i + b] = a.phi[i][b];
        sphw[8 * i + 4 + b] = a.phi[i][b] * a.w[i];
      }
      sphw[8 * i + 3] = a.w[i];
      sphw[8 * i + 7] = T(0);
    }
  }

  const int lane = tid & 63, wave = tid >> 6;
  const bool have_test = lane < nT;
  const int tf = t0 + (have_test ? lane : 0);
  const T* tr = rec + (int64_t)tf * FREC;
  if (wave == 0)
    for (int i = 0; i < NQ * 3; ++i) xs[i * 64 + lane] = tr[i];
  __syncthreads();
  const T nT0 = tr[REC_N], nT1 = tr[REC_N + 1], nT2 = tr[REC_N + 2];
  const T dT = tr[REC_D], areaT = tr[REC_AREA];
  const int tv0 = fverts[(int64_t)tf * 3], tv1 = fverts[(int64_t)tf * 3 + 1], tv2 = fverts[(int64_t)tf * 3 + 2];
  const int tl0 = a.floc[(int64_t)tf * 3], tl1 = a.floc[(int64_t)tf * 3 + 1], tl2 = a.floc[(int64_t)tf * 3 + 2];

  const T k = a.k, krev = a.krev;
  const T er = a.eta_re, ei = a.eta_im;
  const int wu = __builtin_amdgcn_readfirstlane(wave);

  for (int sj = wu; sj < nS; sj += 4) {
    const int64_t sf = s0 + sj;
    const T* sr = rec + sf * FREC;
    T y[NQ][3];
#pragma unroll
    for (int j = 0; j < NQ; ++j)
#pragma unroll
      for (int c = 0; c < 3; ++c) y[j][c] = IMAGE ? sr[3 * j + c] * sg[c] : sr[3 * j + c];
    // the reflected plane keeps its offset: (s v1).(s n) = v1.n
    const T nS0 = IMAGE ? sr[REC_N] * sg[0] : sr[REC_N];
    const T nS1 = IMAGE ? sr[REC_N + 1] * sg[1] : sr[REC_N + 1];
    const T nS2 = IMAGE ? sr[REC_N + 2] * sg[2] : sr[REC_N + 2];
    const T dS = sr[REC_D];
    const T areaS = sr[REC_AREA];
    const int sv0 = fverts[sf * 3], sv1 = fverts[sf * 3 + 1], sv2 = fverts[sf * 3 + 2];
    const bool adj = (tv0 == sv0) | (tv0 == sv1) | (tv0 == sv2) | (tv1 == sv0) | (tv1 == sv1) | (tv1 == sv2) |
                     (tv2 == sv0) | (tv2 == sv1) | (tv2 == sv2);
    if (!have_test || (!IMAGE && adj)) continue;
    const T cw = T(4) * areaT * areaS * T(INV_FOUR_PI);
    const T nprod = nT0 * nS0 + nT1 * nS1 + nT2 * nS2;
    const T kk2n = k * k * nprod;
    // Every accumulated quantity is linear in G w, w = cw w_i w_j, so the loop
    // works with G = e^{ikr} / r and the weights sit in the factors instead: w_j
    // in the trial basis (sphw columns 4..6) and in PW, cw w_i in the test basis.
    // (r.n_x) = y.n_x - d_x is a per-trial-point constant; (r.n_y) = d_y - x.n_y
    // is a per-test-point constant (both faces are flat). PW[j] = -(r.n_x) w_j,
    // times eta_im when eta = i eta_im.
    T PW[NQ];
#pragma unroll
    for (int j = 0; j < NQ; ++j) {
      const T pt = (dT - (y[j][0] * nT0 + y[j][1] * nT1 + y[j][2] * nT2)) * sphw[8 * j + 3];
      PW[j] = PURE ? ei * pt : pt;
    }
    T Lr[3][3], Li[3][3], Rr[3], Ri[3], csr = 0, csi = 0;
#pragma unroll
    for (int p = 0; p < 3; ++p) {
      Rr[p] = Ri[p] = 0;
#pragma unroll
      for (int b = 0; b < 3; ++b) Lr[p][b] = Li[p][b] = 0;
    }
#pragma unroll 1
    for (int i = 0; i < NQ; ++i) {
      const T x0 = xs[(3 * i) * 64 + lane], x1 = xs[(3 * i + 1) * 64 + lane], x2 = xs[(3 * i + 2) * 64 + lane];
      const T PS = dS - (x0 * nS0 + x1 * nS1 + x2 * nS2);
      const T cwi = cw * sphw[8 * i + 3];
      // M: sum_j phi_b(y_j) w_j lhs_s; E - Gs: sum_j w_j rhs_s, split as the
      // eta G' term (E) and the G term (Gs, also the curl term's sum)
      T Mr[3] = {0, 0, 0}, Mi[3] = {0, 0, 0}, Er = 0, Ei = 0, Gsr = 0, Gsi = 0;
#pragma unroll
      for (int j = 0; j < NQ; ++j) {
        const T rx = y[j][0] - x0, ry = y[j][1] - x1, rz = y[j][2] - x2;
        const T r2 = rx * rx + ry * ry + rz * rz;
        const T ir = inv_sqrt(r2);
        const T r = r2 * ir;
        T s, c;
        phase_sincos<WIDE>(r, krev, a.ksteps, k, trig, s, c);
        const T Gr = c * ir, Gi = s * ir;
        const T Dr = -ir * Gr - k * Gi, Di = -ir * Gi + k * Gr;  // G'
        const T t = PS * ir;  // (r.n_y)/r
        const T wj = sphw[8 * j + 3];
        T lr, li;
        if (PURE) {
          // eta = i ei
          const T kap = ei * kk2n, eu = PW[j] * ir;  // eu = ei w_j (-(r.n_x)/r)
          lr = -t * Dr + kap * Gi;
          li = -t * Di - kap * Gr;
          Er += eu * Di;
          Ei -= eu * Dr;
        } else {
          const T knr = er * kk2n, kni = ei * kk2n;
          lr = -t * Dr - (knr * Gr - kni * Gi);
          li = -t * Di - (knr * Gi + kni * Gr);
          const T u = PW[j] * ir;  // w_j (-(r.n_x)/r)
          const T eur = er * u, eui = ei * u;
          Er -= eur * Dr - eui * Di;
          Ei -= eur * Di + eui * Dr;
        }
        Gsr += wj * Gr;
        Gsi += wj * Gi;
#pragma unroll
        for (int b = 0; b < 3; ++b) {
          const T ph = sphw[8 * j + 4 + b];
          Mr[b] += ph * lr;
          Mi[b] += ph * li;
        }
      }
      const T Sr = Er - Gsr, Si = Ei - Gsi;
      csr += cwi * Gsr;
      csi += cwi * Gsi;
#pragma unroll
      for (int p = 0; p < 3; ++p) {
        const T ph = cwi * sphw[8 * i + p];
        Rr[p] += ph * Sr;
        Ri[p] += ph * Si;
#pragma unroll
        for (int b = 0; b < 3; ++b) {
          Lr[p][b] += ph * Mr[b];
          Li[p][b] += ph * Mi[b];
        }
      }
    }
    // hypersingular curl term: (curl_a . curl_b) * eta * sum(G w)
    const T ecr = er * csr - ei * csi, eci = er * csi + ei * csr;
    const int tl[3] = {tl0, tl1, tl2};
    const int sl[3] = {flocs[sf * 3], flocs[sf * 3 + 1], flocs[sf * 3 + 2]};
#pragma unroll
    for (int p = 0; p < 3; ++p) {
      const T c0 = tr[REC_CURL + 3 * p], c1 = tr[REC_CURL + 3 * p + 1], c2 = tr[REC_CURL + 3 * p + 2];
#pragma unroll
      for (int b = 0; b < 3; ++b) {
        const T b0 = sr[REC_CURL + 3 * b], b1 = sr[REC_CURL + 3 * b + 1], b2 = sr[REC_CURL + 3 * b + 2];
        const T cc = IMAGE ? c0 * (b0 * sg[3]) + c1 * (b1 * sg[4]) + c2 * (b2 * sg[5]) : c0 * b0 + c1 * b1 + c2 * b2;
        T* dst = accA + tl[p] * LDA_ + 2 * sl[b];
        lds_add(dst, Lr[p][b] + cc * ecr);
        lds_add(dst + 1, Li[p][b] + cc * eci);
      }
      T* dc = accC + tl[p] * LDC_ + 2 * sj;
      lds_add(dc, Rr[p]);
      lds_add(dc + 1, Ri[p]);
    }
  }
  __syncthreads();

  // flush the P1 x P1 block
  for (int idx = tid; idx < nvT * nvS; idx += blockDim.x) {
    const int r = idx / nvS, c = idx - r * nvS;
    const T vr = accA[r * LDA_ + 2 * c], vi = accA[r * LDA_ + 2 * c + 1];
    if (vr != T(0) || vi != T(0)) {
      T* dst = a.A + 2 * ((int64_t)grow[r] * a.lda + gcol[c]);
      gadd(dst, vr);
      gadd(dst + 1, vi);
    }
  }
  // flush the right-hand sides: B[row, d] += sum_s C[row, s] q[s, d]
  for (int idx = tid; idx < nvT * a.nrhs; idx += blockDim.x) {
    const int r = idx / a.nrhs, d = idx - r * a.nrhs;
    T br = 0, bi = 0;
    for (int s = 0; s < nS; ++s) {
      const T cr = accC[r * LDC_ + 2 * s], ci = accC[r * LDC_ + 2 * s + 1];
      const T* qq = a.q + 2 * ((int64_t)(s0 + s) * a.nrhs + d);
      const T qr = qq[0], qi = qq[1];
      br += cr * qr - ci * qi;
      bi += cr * qi + ci * qr;
    }
    if (br != T(0) || bi != T(0)) {
      T* dst = a.B + 2 * ((int64_t)grow[r] * a.nrhs + d);
      gadd(dst, br);
      gadd(dst + 1, bi);
    }
  }
}

// ---- symmetric regular pass ----------------------------------------------------
// The regular pairs over unordered tile pairs. A unit is (test tile t, sub-tile
// s) with s from t's first sub-tile on; sub-tiles nest in the test tiles (at most
// TILE_FS_N faces / TILE_VS_N vertices). Lane = test face T of tile t; waves stride
// over the faces S of sub-tile s. When s lies in a later test tile, one evaluation
// of r, G = e^{ikr} / r and G' = G (-1/r + i k) per point pair serves both orders:
// (T test, S trial) into fA / fC and (S test, T trial) into rA / rC. With r = y - x,
// x on T and y on S, the two normal factors swap roles between the orders:
//   a = (d_S - x.n_S) / r = (r.n_S) / r   forward (r.n_y)/r, reverse -(r'.n_x')/r'
//   b = (d_T - y.n_T) / r = -(r.n_T) / r  forward -(r.n_x)/r, reverse (r'.n_y')/r'
// (r' = x - y is the reverse pair's y' - x'). When s lies in t itself, both faces
// of a pair are test faces of the unit: the pair is taken once, from its lower face.
// Basis sums: on each side the first basis slot accumulates the sum over the three
// functions (phi_0 + phi_1 + phi_2 = 1), and the first slot's own entry is the
// difference, formed once per face pair. The test side's slots run over the face's
// vertices in a per-face rotated order (the mesh set-up chooses the rotations).
// LDS pointers keep their address space through the call, so that the atomics on
// them are known not to touch the face records (which then stay scalar loads)
template <typename T>
using lds_ptr = __attribute__((address_space(3))) T*;

template <typename T, int NQ, bool PURE, bool WIDE>
__device__ __forceinline__ void sym_unit(const RegularArgs<T> a, int tt, int st, bool diag, lds_ptr<T> fA,
                                         lds_ptr<T> fC, lds_ptr<T> rA, lds_ptr<T> rC, lds_ptr<const T> xs,
                                         lds_ptr<const T> sphw, const double* trig) {
  constexpr int LF = 2 * TILE_VS_N + 1, LR = 2 * TILE_V + 1;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int t0 = a.tile_start[tt], nT = a.tile_count[tt];
  const int s0 = a.qtile_start[st], nS = a.qtile_count[st];
  const T* __restrict__ rec = a.rec;
  const int32_t* __restrict__ fverts = a.fverts;
  const uint8_t* __restrict__ flocq = a.flocq;
  const bool have_test = lane < nT;
  const int tf = t0 + (have_test ? lane : 0);
  const T* tr = rec + (int64_t)tf * FREC;
  const T nT0 = tr[REC_N], nT1 = tr[REC_N + 1], nT2 = tr[REC_N + 2];
  const T dT = tr[REC_D], areaT = tr[REC_AREA];
  const int tv0 = fverts[(int64_t)tf * 3], tv1 = fverts[(int64_t)tf * 3 + 1], tv2 = fverts[(int64_t)tf * 3 + 2];
  // test-side slot k stands for vertex tp[k] = (k + rot) % 3 of the face: the
  // faces of a tile that share a vertex mostly hold it in different slots, so one
  // scatter instruction's lanes mostly add to distinct LDS rows
  const int rot = a.frot[tf];
  const int tp[3] = {rot, rot == 2 ? 0 : rot + 1, rot == 0 ? 2 : rot - 1};
  const int tl[3] = {a.floc[(int64_t)tf * 3 + tp[0]], a.floc[(int64_t)tf * 3 + tp[1]],
                     a.floc[(int64_t)tf * 3 + tp[2]]};
  const T k = a.k, krev = a.krev;
  const T er = a.eta_re, ei = a.eta_im;
  const int wu = __builtin_amdgcn_readfirstlane(wave);

  for (int sj = wu; sj < nS; sj += 4) {
    const int64_t sf = __builtin_amdgcn_readfirstlane(s0 + sj);
    const T* sr = rec + sf * FREC;
    T y[NQ][3];
#pragma unroll
    for (int j = 0; j < NQ; ++j)
#pragma unroll
      for (int c = 0; c < 3; ++c) y[j][c] = sr[3 * j + c];
    const T nS0 = sr[REC_N], nS1 = sr[REC_N + 1], nS2 = sr[REC_N + 2];
    const T dS = sr[REC_D], areaS = sr[REC_AREA];
    const int sv0 = fverts[sf * 3], sv1 = fverts[sf * 3 + 1], sv2 = fverts[sf * 3 + 2];
    const bool adj = (tv0 == sv0) | (tv0 == sv1) | (tv0 == sv2) | (tv1 == sv0) | (tv1 == sv1) | (tv1 == sv2) |
                     (tv2 == sv0) | (tv2 == sv1) | (tv2 == sv2);
    if (!have_test || adj || (diag && tf >= sf)) continue;
    const T cw = T(4) * areaT * areaS * T(INV_FOUR_PI);
    const T kk2n = k * k * (nT0 * nS0 + nT1 * nS1 + nT2 * nS2);
    const T knr = er * kk2n, kni = ei * kk2n;
    // QW[j] = d_T - y_j.n_T: b times r
    T QW[NQ];
#pragma unroll
    for (int j = 0; j < NQ; ++j) QW[j] = dT - (y[j][0] * nT0 + y[j][1] * nT1 + y[j][2] * nT2);
    // forward [test basis][trial basis] and right-hand side [test basis]; reverse
    // [trial-face basis][test-face basis] and [trial-face basis]; slot 0 = sum
    T Lfr[3][3], Lfi[3][3], Rfr[3], Rfi[3];
    T Lrr[3][3], Lri[3][3], Rrr[3], Rri[3];
    T csr = 0, csi = 0;
#pragma unroll
    for (int p = 0; p < 3; ++p) {
      Rfr[p] = Rfi[p] = Rrr[p] = Rri[p] = 0;
#pragma unroll
      for (int b = 0; b < 3; ++b) Lfr[p][b] = Lfi[p][b] = Lrr[p][b] = Lri[p][b] = 0;
    }
    // the next test point's coordinates are read one point ahead
    T nx0 = xs[lane], nx1 = xs[64 + lane], nx2 = xs[128 + lane];
#pragma unroll 1
    for (int i = 0; i < NQ; ++i) {
      const T x0 = nx0, x1 = nx1, x2 = nx2;
      if (i + 1 < NQ) {
        nx0 = xs[(3 * i + 3) * 64 + lane];
        nx1 = xs[(3 * i + 4) * 64 + lane];
        nx2 = xs[(3 * i + 5) * 64 + lane];
      }
      const T PS = dS - (x0 * nS0 + x1 * nS1 + x2 * nS2);  // a times r
      const T cwi = cw * sphw[8 * i + 3];
      // Per trial point three complex quantities, each summed against the three
      // trial basis weights (slot 0 the rule weight, slots 1 and 2 the weight times
      // xi, eta): G = e^{ikr}/r, E = G' / r = G (-1/r^2 + i k / r) and b E (b = the
      // trial face's normal projection d_T - y.n_T). Every block of both orders is
      // linear in these sums:
      //   forward lhs   -a E - K        (a = d_S - x.n_S is constant over y)
      //   reverse lhs   -b E - K        K = eta k^2 (n_T.n_S) G
      //   reverse rhs   -G - eta a E
      //   forward rhs   -G - eta b E    (slot 0)
      T Ar[3] = {0, 0, 0}, Ai[3] = {0, 0, 0}, Gr3[3] = {0, 0, 0}, Gi3[3] = {0, 0, 0};
      T Br[3] = {0, 0, 0}, Bi[3] = {0, 0, 0};
      // per point pair, first the distance and the phase table entry, one trial
      // point ahead of the rest (the table read is in flight meanwhile)
      auto stage = [&](int j, PhaseStage<T>& o) {
        const T rx = y[j][0] - x0, ry = y[j][1] - x1, rz = y[j][2] - x2;
        const T r2 = rx * rx + ry * ry + rz * rz;
        o.ir = inv_sqrt(r2);
        phase_begin<WIDE>(r2 * o.ir, krev, a.ksteps, k, trig, o);
      };
      PhaseStage<T> cur, nxt;
      stage(0, cur);
#pragma unroll
      for (int j = 0; j < NQ; ++j) {
        if (j + 1 < NQ) stage(j + 1, nxt);
        const T ir = cur.ir;
        T s, c;
        phase_end<WIDE>(cur, s, c);
        cur = nxt;
        const T Gr = c * ir, Gi = s * ir;
        const T ir2 = ir * ir, kir = k * ir;
        const T Er = -ir2 * Gr - kir * Gi, Ei = kir * Gr - ir2 * Gi;
        const T QEr = QW[j] * Er, QEi = QW[j] * Ei;
        // rule constants: compile-time, so they live in no register across the loop
        const double wjd = TriRule<NQ>::w[j], p1d = wjd * TriRule<NQ>::xi[j], p2d = wjd * TriRule<NQ>::eta[j];
        const T wb[3] = {(T)wjd, (T)p1d, (T)p2d};
#pragma unroll
        for (int b = 0; b < 3; ++b) {
          Gr3[b] += wb[b] * Gr;
          Gi3[b] += wb[b] * Gi;
          Ar[b] += wb[b] * Er;
          Ai[b] += wb[b] * Ei;
          Br[b] += wb[b] * QEr;
          Bi[b] += wb[b] * QEi;
        }
      }
      // the blocks' trial-side sums from the three basis sums
      T Mfr[3], Mfi[3], Mrr[3], Mri[3], Frr[3], Fri[3], Sr, Si;
      if (PURE) {
        // eta = i ei: K = i kni G
        const T eps = ei * PS;
#pragma unroll
        for (int b = 0; b < 3; ++b) {
          Mfr[b] = kni * Gi3[b] - PS * Ar[b];
          Mfi[b] = -(kni * Gr3[b]) - PS * Ai[b];
          Mrr[b] = kni * Gi3[b] - Br[b];
          Mri[b] = -(kni * Gr3[b]) - Bi[b];
          Frr[b] = eps * Ai[b] - Gr3[b];
          Fri[b] = -(eps * Ar[b]) - Gi3[b];
        }
        Sr = ei * Bi[0] - Gr3[0];
        Si = -(ei * Br[0]) - Gi3[0];
      } else {
        const T epr = er * PS, epi = ei * PS;
#pragma unroll
        for (int b = 0; b < 3; ++b) {
          const T Kr = knr * Gr3[b] - kni * Gi3[b], Ki = knr * Gi3[b] + kni * Gr3[b];
          Mfr[b] = -(PS * Ar[b]) - Kr;
          Mfi[b] = -(PS * Ai[b]) - Ki;
          Mrr[b] = -Br[b] - Kr;
          Mri[b] = -Bi[b] - Ki;
          Frr[b] = epi * Ai[b] - epr * Ar[b] - Gr3[b];
          Fri[b] = -(epr * Ai[b]) - epi * Ar[b] - Gi3[b];
        }
        Sr = ei * Bi[0] - er * Br[0] - Gr3[0];
        Si = -(er * Bi[0]) - ei * Br[0] - Gi3[0];
      }
      csr += cwi * Gr3[0];
      csi += cwi * Gi3[0];
      const T ph[3] = {cwi, cwi * sphw[8 * i + tp[1]], cwi * sphw[8 * i + tp[2]]};
#pragma unroll
      for (int p = 0; p < 3; ++p) {
        Rfr[p] += ph[p] * Sr;
        Rfi[p] += ph[p] * Si;
#pragma unroll
        for (int b = 0; b < 3; ++b) {
          Lfr[p][b] += ph[p] * Mfr[b];
          Lfi[p][b] += ph[p] * Mfi[b];
        }
      }
#pragma unroll
      for (int p = 0; p < 3; ++p) {
        Rrr[p] += cwi * Frr[p];
        Rri[p] += cwi * Fri[p];
#pragma unroll
        for (int b = 0; b < 3; ++b) {
          Lrr[p][b] += ph[b] * Mrr[p];
          Lri[p][b] += ph[b] * Mri[p];
        }
      }
    }
    // slot 0 held the basis sums: phi_0's entries are the differences
#pragma unroll
    for (int p = 0; p < 3; ++p) {
      Lfr[p][0] -= Lfr[p][1] + Lfr[p][2];
      Lfi[p][0] -= Lfi[p][1] + Lfi[p][2];
    }
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      Lfr[0][b] -= Lfr[1][b] + Lfr[2][b];
      Lfi[0][b] -= Lfi[1][b] + Lfi[2][b];
    }
    Rfr[0] -= Rfr[1] + Rfr[2];
    Rfi[0] -= Rfi[1] + Rfi[2];
#pragma unroll
    for (int p = 0; p < 3; ++p) {
      Lrr[p][0] -= Lrr[p][1] + Lrr[p][2];
      Lri[p][0] -= Lri[p][1] + Lri[p][2];
    }
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      Lrr[0][b] -= Lrr[1][b] + Lrr[2][b];
      Lri[0][b] -= Lri[1][b] + Lri[2][b];
    }
    Rrr[0] -= Rrr[1] + Rrr[2];
    Rri[0] -= Rri[1] + Rri[2];
    // hypersingular curl term: (curl_a . curl_b) * eta * sum(G w), the same for both orders
    const T ecr = er * csr - ei * csi, eci = er * csi + ei * csr;
    const int sl[3] = {flocq[sf * 3], flocq[sf * 3 + 1], flocq[sf * 3 + 2]};
#pragma unroll
    for (int p = 0; p < 3; ++p) {
      const T* tc = tr + REC_CURL + 3 * tp[p];
      const T c0 = tc[0], c1 = tc[1], c2 = tc[2];
#pragma unroll
      for (int b = 0; b < 3; ++b) {
        const T b0 = sr[REC_CURL + 3 * b], b1 = sr[REC_CURL + 3 * b + 1], b2 = sr[REC_CURL + 3 * b + 2];
        const T cc = c0 * b0 + c1 * b1 + c2 * b2;
        lds_ptr<T> dst = fA + tl[p] * LF + 2 * sl[b];
        atomicAdd((T*)dst, Lfr[p][b] + cc * ecr);
        atomicAdd((T*)(dst + 1), Lfi[p][b] + cc * eci);
        lds_ptr<T> dr = rA + sl[b] * LR + 2 * tl[p];
        atomicAdd((T*)dr, Lrr[b][p] + cc * ecr);
        atomicAdd((T*)(dr + 1), Lri[b][p] + cc * eci);
      }
      lds_ptr<T> dc = fC + tl[p] * LF + 2 * sj;
      atomicAdd((T*)dc, Rfr[p]);
      atomicAdd((T*)(dc + 1), Rfi[p]);
      lds_ptr<T> drc = rC + sl[p] * LR + 2 * lane;
      atomicAdd((T*)drc, Rrr[p]);
      atomicAdd((T*)(drc + 1), Rri[p]);
    }
  }
}

// One workgroup per unit of the symmetric pass (4 waves); the units of one test
// tile are consecutive, each XCD takes a contiguous range of them, and odd test
// tiles walk their sub-tiles backwards (the trial records the previous tile read
// last are still in L2). Flushes as in bm_regular_kernel, the reverse blocks
// transposed: rows of sub-tile vertices, columns of test-tile vertices / faces.
template <typename T, int NQ, bool PURE, bool WIDE>
__global__ __launch_bounds__(256, 2) void bm_regular_sym_kernel(const RegularArgs<T> a) {
  constexpr int LF = 2 * TILE_VS_N + 1, LR = 2 * TILE_V + 1;
  __shared__ T fA[TILE_V * LF];
  __shared__ T fC[TILE_V * LF];
  __shared__ T rA[TILE_VS_N * LR];
  __shared__ T rC[TILE_VS_N * LR];
  __shared__ T xs[NQ * 3 * 64];
  __shared__ T sphw[6 * 8];
  __shared__ double trig[2 * TRIG_N];
  __shared__ int grow[TILE_V], gcol[TILE_VS_N];

  const int64_t total = a.nunits;
  const int64_t per = (total + 7) / 8;
  const int64_t bid = blockIdx.x;
  const int64_t L = (bid & 7) * per + (bid >> 3);
  if (L >= total) return;
  // the test tile: the last one whose first unit is at most L
  int lo = 0, hi = a.ntiles;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (a.unit_start[mid] <= L) lo = mid;
    else hi = mid;
  }
  // (wave-uniform: the trial records below are then scalar loads)
  const int tt = __builtin_amdgcn_readfirstlane(lo);
  const int q0 = a.tile_qfirst[tt], q1 = a.tile_qfirst[tt + 1];
  const int off = (int)(L - a.unit_start[tt]);
  const int st = __builtin_amdgcn_readfirstlane((tt & 1) ? a.nqtiles - 1 - off : q0 + off);
  const bool diag = st < q1;
  const int tid = threadIdx.x;
  const int t0 = a.tile_start[tt], nT = a.tile_count[tt], nvT = a.tile_nv[tt];
  const int s0 = a.qtile_start[st], nS = a.qtile_count[st], nvS = a.qtile_nv[st];

  for (int i = tid; i < TILE_V * LF; i += blockDim.x) fA[i] = fC[i] = T(0);
  for (int i = tid; i < TILE_VS_N * LR; i += blockDim.x) rA[i] = rC[i] = T(0);
  if (tid < TILE_V) grow[tid] = a.tile_verts[(int64_t)tt * TILE_V + tid];
  if (tid < TILE_VS_N) gcol[tid] = a.qtile_verts[(int64_t)st * TILE_VS_N + tid];
  for (int i = tid; i < 2 * TRIG_N; i += blockDim.x) trig[i] = a.trig[i];
  if (tid == 0) {
#pragma unroll
    for (int i = 0; i < 6; ++i) {
#pragma unroll
      for (int b = 0; b < 3; ++b) {
        sphw[8 * i + b] = a.phi[i][b];
        sphw[8 * i + 4 + b] = a.phi[i][b] * a.w[i];
      }
      sphw[8 * i + 3] = a.w[i];
      sphw[8 * i + 7] = T(0);
    }
  }
  {
    const int lane = tid & 63;
    const int tf = t0 + (lane < nT ? lane : 0);
    if ((tid >> 6) == 0)
      for (int i = 0; i < NQ * 3; ++i) xs[i * 64 + lane] = a.rec[(int64_t)tf * FREC + i];
  }
  __syncthreads();
  const lds_ptr<T> lfA = (lds_ptr<T>)fA, lfC = (lds_ptr<T>)fC, lrA = (lds_ptr<T>)rA, lrC = (lds_ptr<T>)rC;
  const lds_ptr<const T> lxs = (lds_ptr<const T>)xs, lsp = (lds_ptr<const T>)sphw;
  sym_unit<T, NQ, PURE, WIDE>(a, tt, st, diag, lfA, lfC, lrA, lrC, lxs, lsp, trig);
  __syncthreads();

  // forward P1 x P1 block: rows of the test tile, columns of the sub-tile
  for (int idx = tid; idx < nvT * nvS; idx += blockDim.x) {
    const int r = idx / nvS, c = idx - r * nvS;
    const T vr = fA[r * LF + 2 * c], vi = fA[r * LF + 2 * c + 1];
    if (vr != T(0) || vi != T(0)) {
      T* dst = a.A + 2 * ((int64_t)grow[r] * a.lda + gcol[c]);
      gadd(dst, vr);
      gadd(dst + 1, vi);
    }
  }
  // forward right-hand sides: B[row, d] += sum_s fC[row, s] q[s, d]
  for (int idx = tid; idx < nvT * a.nrhs; idx += blockDim.x) {
    const int r = idx / a.nrhs, d = idx - r * a.nrhs;
    T br = 0, bi = 0;
    for (int s = 0; s < nS; ++s) {
      const T cr = fC[r * LF + 2 * s], ci = fC[r * LF + 2 * s + 1];
      const T* qq = a.q + 2 * ((int64_t)(s0 + s) * a.nrhs + d);
      br += cr * qq[0] - ci * qq[1];
      bi += cr * qq[1] + ci * qq[0];
    }
    if (br != T(0) || bi != T(0)) {
      T* dst = a.B + 2 * ((int64_t)grow[r] * a.nrhs + d);
      gadd(dst, br);
      gadd(dst + 1, bi);
    }
  }
  // reverse P1 x P1 block: rows of the sub-tile, columns of the test tile
  for (int idx = tid; idx < nvS * nvT; idx += blockDim.x) {
    const int r = idx / nvT, c = idx - r * nvT;
    const T vr = rA[r * LR + 2 * c], vi = rA[r * LR + 2 * c + 1];
    if (vr != T(0) || vi != T(0)) {
      T* dst = a.A + 2 * ((int64_t)gcol[r] * a.lda + grow[c]);
      gadd(dst, vr);
      gadd(dst + 1, vi);
    }
  }
  // reverse right-hand sides: B[sub-tile row, d] += sum_t rC[row, t] q[t, d]
  for (int idx = tid; idx < nvS * a.nrhs; idx += blockDim.x) {
    const int r = idx / a.nrhs, d = idx - r * a.nrhs;
    T br = 0, bi = 0;
    for (int t = 0; t < nT; ++t) {
      const T cr = rC[r * LR + 2 * t], ci = rC[r * LR + 2 * t + 1];
      const T* qq = a.q + 2 * ((int64_t)(t0 + t) * a.nrhs + d);
      br += cr * qq[0] - ci * qq[1];
      bi += cr * qq[1] + ci * qq[0];
    }
    if (br != T(0) || bi != T(0)) {
      T* dst = a.B + 2 * ((int64_t)gcol[r] * a.nrhs + d);
      gadd(dst, br);
      gadd(dst + 1, bi);
    }
  }
}

// Sum over the 16 lanes of a DPP row (every lane receives the total): rotations
// by 8, 4, 2 and 1 within the row.
template <int N>
__device__ __forceinline__ double row_ror(double v) {
  const long long b = __builtin_bit_cast(long long, v);
  const int lo = __builtin_amdgcn_mov_dpp((int)b, 0x120 + N, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), 0x120 + N, 0xF, 0xF, false);
  return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ double row_sum16(double v) {
  v += row_ror<8>(v);
  v += row_ror<4>(v);
  v += row_ror<2>(v);
  v += row_ror<1>(v);
  return v;
}

// Touching pairs over the remapped Duffy rules, FP64 arithmetic. A group of 16
// lanes (one DPP row) per pair, four pairs per wavefront; the host orders the
// pairs by rule size, so a wavefront's pairs mostly share their point count. The
// lanes of a group stride over the rule points; 1/r is the hardware reciprocal
// square root with one Newton step and the phase the pi/64 table form (the pi/2
// form when the host has not bounded |k| r below the table's limit).
template <typename TO, bool WIDE>
__global__ __launch_bounds__(256) void bm_singular_kernel(const SingularArgs<TO> a) {
  __shared__ double trig[2 * TRIG_N];
  if (!WIDE) {
    for (int i = threadIdx.x; i < 2 * TRIG_N; i += blockDim.x) trig[i] = a.trig[i];
    __syncthreads();
  }
  const int sub = threadIdx.x & 15;
  const int64_t pidx = (int64_t)blockIdx.x * (blockDim.x >> 4) + (threadIdx.x >> 4);
  if (pidx >= a.npairs) return;  // whole rows: the row sums below stay within live groups
  const SingularPair P = a.pairs[pidx];
  const double* tr = a.rec + (int64_t)P.test * FREC;
  const double* sr = a.rec + (int64_t)P.trial * FREC;
  // trial-face reflection of an image pair (identity signs otherwise)
  double sg[6];
#pragma unroll
  for (int c = 0; c < 6; ++c) sg[c] = a.img[0][c];
#pragma unroll
  for (int x = 1; x <= MAX_IMAGES; ++x)
    if (P.xform == x)
#pragma unroll
      for (int c = 0; c < 6; ++c) sg[c] = a.img[x][c];
  // vertex 0 and the edge vectors of both faces
  double t0v[3], t1[3], t2[3], s0v[3], s1[3], s2[3];
  int tvi[3], svi[3];
  for (int v = 0; v < 3; ++v) {
    tvi[v] = a.fverts[(int64_t)P.test * 3 + v];
    svi[v] = a.fverts[(int64_t)P.trial * 3 + v];
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const double ta = a.verts[3 * (int64_t)tvi[0] + c], tb = a.verts[3 * (int64_t)tvi[1] + c],
                 tc = a.verts[3 * (int64_t)tvi[2] + c];
    const double sa = a.verts[3 * (int64_t)svi[0] + c] * sg[c], sb = a.verts[3 * (int64_t)svi[1] + c] * sg[c],
                 sc = a.verts[3 * (int64_t)svi[2] + c] * sg[c];
    t0v[c] = ta;
    t1[c] = tb - ta;
    t2[c] = tc - ta;
    s0v[c] = sa;
    s1[c] = sb - sa;
    s2[c] = sc - sa;
  }
  const double nT[3] = {tr[REC_N], tr[REC_N + 1], tr[REC_N + 2]};
  const double nS[3] = {sr[REC_N] * sg[0], sr[REC_N + 1] * sg[1], sr[REC_N + 2] * sg[2]};
  const double k = a.k, er = a.eta_re, ei = a.eta_im;
  const double knr = er * (k * k * P.nprod), kni = ei * (k * k * P.nprod);
  const int off = a.rule_off[P.rule], cnt = a.rule_cnt[P.rule];
  double acc[26];
#pragma unroll
  for (int i = 0; i < 26; ++i) acc[i] = 0;
  for (int q = sub; q < cnt; q += 16) {
    const double* rd = a.rule_data + 5 * (int64_t)(off + q);
    const double tx = rd[0], ty = rd[1], sx = rd[2], sy = rd[3];
    double x[3], y[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      x[c] = fma(tx, t1[c], fma(ty, t2[c], t0v[c]));
      y[c] = fma(sx, s1[c], fma(sy, s2[c], s0v[c]));
    }
    const double rx = y[0] - x[0], ry = y[1] - x[1], rz = y[2] - x[2];
    const double r2 = rx * rx + ry * ry + rz * rz;
    if (r2 == 0.0) continue;
    const double ir = inv_sqrt(r2);
    const double radius = r2 * ir;
    double s, c;
    if (WIDE) sincos_phase(k * radius, &s, &c);
    else sincos_steps(radius * a.ksteps, trig, &s, &c);
    const double scale = INV_FOUR_PI * ir;
    const double Gr = c * scale, Gi = s * scale;
    const double gr = Gr * -ir - Gi * k, gi = Gr * k + Gi * -ir;  // G (-1/r + i k)
    const double w = rd[4] * P.jac;
    const double wgr = Gr * w, wgi = Gi * w;
    const double tdot = (rx * nS[0] + ry * nS[1] + rz * nS[2]) * ir;
    const double sdot = -(rx * nT[0] + ry * nT[1] + rz * nT[2]) * ir;
    // lhs_s = -grad*tdot*w - kn*wg ; rhs_s = -wg - eta*grad*sdot*w
    const double tw = tdot * w, sw = sdot * w;
    const double lr = -gr * tw - (knr * wgr - kni * wgi);
    const double li = -gi * tw - (knr * wgi + kni * wgr);
    const double egr = er * gr - ei * gi, egi = er * gi + ei * gr;
    const double hr = -wgr - egr * sw, hi = -wgi - egi * sw;
    const double pt[3] = {1.0 - tx - ty, tx, ty}, ps[3] = {1.0 - sx - sy, sx, sy};
    // phi_b(y) lhs, then the test basis
    double plr[3], pli[3];
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      plr[b] = ps[b] * lr;
      pli[b] = ps[b] * li;
    }
#pragma unroll
    for (int p = 0; p < 3; ++p) {
      acc[18 + 2 * p] += pt[p] * hr;
      acc[19 + 2 * p] += pt[p] * hi;
#pragma unroll
      for (int b = 0; b < 3; ++b) {
        acc[2 * (3 * p + b)] += pt[p] * plr[b];
        acc[2 * (3 * p + b) + 1] += pt[p] * pli[b];
      }
    }
    acc[24] += wgr;
    acc[25] += wgi;
  }
#pragma unroll
  for (int i = 0; i < 26; ++i) acc[i] = row_sum16(acc[i]);
  // eta * sum(G w) times the curl products
  const double ecr = er * acc[24] - ei * acc[25], eci = er * acc[25] + ei * acc[24];
  for (int e = sub; e < 9 + 3 * a.nrhs; e += 16) {
    if (e < 9) {
      const int p = e / 3, b = e % 3;
      const double cc = tr[REC_CURL + 3 * p] * (sr[REC_CURL + 3 * b] * sg[3]) +
                        tr[REC_CURL + 3 * p + 1] * (sr[REC_CURL + 3 * b + 1] * sg[4]) +
                        tr[REC_CURL + 3 * p + 2] * (sr[REC_CURL + 3 * b + 2] * sg[5]);
      double vr = 0, vi = 0;
#pragma unroll
      for (int i = 0; i < 9; ++i)
        if (i == e) {
          vr = acc[2 * i] + cc * ecr;
          vi = acc[2 * i + 1] + cc * eci;
        }
      TO* dst = a.A + 2 * ((int64_t)tvi[p] * a.lda + svi[b]);
      gadd(dst, (TO)vr);
      gadd(dst + 1, (TO)vi);
    } else {
      const int f = e - 9, p = f / a.nrhs, d = f - p * a.nrhs;
      double hr = 0, hi = 0;
#pragma unroll
      for (int i = 0; i < 3; ++i)
        if (i == p) {
          hr = acc[18 + 2 * i];
          hi = acc[19 + 2 * i];
        }
      const double* qq = a.q + 2 * ((int64_t)P.trial * a.nrhs + d);
      const double vr = hr * qq[0] - hi * qq[1], vi = hr * qq[1] + hi * qq[0];
      TO* dst = a.B + 2 * ((int64_t)tvi[p] * a.nrhs + d);
      gadd(dst, (TO)vr);
      gadd(dst + 1, (TO)vi);
    }
  }
}


// ---- four separate operators (operator_matrices) -----------------------------------
// _beat_cpu_accumulate_pair! / _beat_cpu_accumulate_regular_pair_signed!
// (BeatEngineCpuAssembly.jl:151-218,278-373): per point pair, with G' = G (-1/r + i k),
//   S[a, s]  += phi_a G w              K'[a, s] += phi_a G' (-r.n_x / r) w
//   D[a, b]  += phi_a phi_b G' (r.n_y / r) w
//   H[a, b]  += (curl_a . curl_b - k^2 phi_a phi_b (n_x . n_y)) G w
// Same tiling as bm_regular_kernel; four LDS tiles (D, H by trial vertex slot, S, K'
// by trial face), so one workgroup per CU in FP64.
template <typename T, int NQ, bool WIDE, bool IMAGE>
__global__ __launch_bounds__(256, 1) void bm_operators_kernel(const OperatorArgs<T> o) {
  constexpr int LDA_ = 2 * TILE_VS + 1, LDC_ = 2 * TILE_FS + 1;
  __shared__ T accD[TILE_V * LDA_];
  __shared__ T accH[TILE_V * LDA_];
  __shared__ T accS[TILE_V * LDC_];
  __shared__ T accK[TILE_V * LDC_];
  __shared__ T xs[NQ * 3 * 64];
  __shared__ T sphw[6 * 4];
  __shared__ double trig[2 * TRIG_N];
  __shared__ int grow[TILE_V], gcol[TILE_VS];
  const RegularArgs<T>& a = o.g;

  const int64_t total = (int64_t)a.ntiles * a.nstiles;
  const int64_t per = (total + 7) / 8;
  int64_t bid = blockIdx.x;
  T sg[6] = {T(1), T(1), T(1), T(1), T(1), T(1)};
  if (IMAGE) {
    const int xi = (int)(bid / (per * 8));
    bid -= (int64_t)xi * per * 8;
#pragma unroll
    for (int x = 0; x < MAX_IMAGES; ++x)
      if (x == xi)
#pragma unroll
        for (int c = 0; c < 6; ++c) sg[c] = a.img[x][c];
  }
  const int64_t L = (bid & 7) * per + (bid >> 3);
  if (L >= total) return;
  const int tt = (int)(L / a.nstiles), st = (int)(L % a.nstiles);
  const int tid = threadIdx.x;
  const int t0 = a.tile_start[tt], nT = a.tile_count[tt], nvT = a.tile_nv[tt];
  const int s0 = a.stile_start[st], nS = a.stile_count[st], nvS = a.stile_nv[st];
  const T* __restrict__ rec = a.rec;
  const int32_t* __restrict__ fverts = a.fverts;
  const uint8_t* __restrict__ flocs = a.flocs;

  for (int i = tid; i < TILE_V * LDA_; i += blockDim.x) accD[i] = accH[i] = T(0);
  for (int i = tid; i < TILE_V * LDC_; i += blockDim.x) accS[i] = accK[i] = T(0);
  if (tid < TILE_V) grow[tid] = a.tile_verts[(int64_t)tt * TILE_V + tid];
  if (tid < TILE_VS) gcol[tid] = a.stile_verts[(int64_t)st * TILE_VS + tid];
  for (int i = tid; i < 2 * TRIG_N; i += blockDim.x) trig[i] = a.trig[i];
  if (tid == 0) {
#pragma unroll
    for (int i = 0; i < 6; ++i) {
      sphw[4 * i] = a.phi[i][0];
      sphw[4 * i + 1] = a.phi[i][1];
      sphw[4 * i + 2] = a.phi[i][2];
      sphw[4 * i + 3] = a.w[i];
    }
  }
  const int lane = tid & 63, wave = tid >> 6;
  const bool have_test = lane < nT;
  const int tf = t0 + (have_test ? lane : 0);
  const T* tr = rec + (int64_t)tf * FREC;
  if (wave == 0)
    for (int i = 0; i < NQ * 3; ++i) xs[i * 64 + lane] = tr[i];
  __syncthreads();
  const T nT0 = tr[REC_N], nT1 = tr[REC_N + 1], nT2 = tr[REC_N + 2];
  const T dT = tr[REC_D], areaT = tr[REC_AREA];
  const int tv0 = fverts[(int64_t)tf * 3], tv1 = fverts[(int64_t)tf * 3 + 1], tv2 = fverts[(int64_t)tf * 3 + 2];
  const int tl[3] = {a.floc[(int64_t)tf * 3], a.floc[(int64_t)tf * 3 + 1], a.floc[(int64_t)tf * 3 + 2]};
  const T k = a.k, krev = a.krev;
  const int wu = __builtin_amdgcn_readfirstlane(wave);

  for (int sj = wu; sj < nS; sj += 4) {
    const int64_t sf = s0 + sj;
    const int sv0 = fverts[sf * 3], sv1 = fverts[sf * 3 + 1], sv2 = fverts[sf * 3 + 2];
    const bool adj = (tv0 == sv0) | (tv0 == sv1) | (tv0 == sv2) | (tv1 == sv0) | (tv1 == sv1) | (tv1 == sv2) |
                     (tv2 == sv0) | (tv2 == sv1) | (tv2 == sv2);
    if (!have_test || (!IMAGE && adj)) continue;
    const T* sr = rec + sf * FREC;
    T y[NQ][3];
#pragma unroll
    for (int j = 0; j < NQ; ++j)
#pragma unroll
      for (int c = 0; c < 3; ++c) y[j][c] = IMAGE ? sr[3 * j + c] * sg[c] : sr[3 * j + c];
    const T nS0 = IMAGE ? sr[REC_N] * sg[0] : sr[REC_N];
    const T nS1 = IMAGE ? sr[REC_N + 1] * sg[1] : sr[REC_N + 1];
    const T nS2 = IMAGE ? sr[REC_N + 2] * sg[2] : sr[REC_N + 2];
    const T dS = sr[REC_D];
    const T cw = T(4) * areaT * sr[REC_AREA] * T(INV_FOUR_PI);
    const T nprod = nT0 * nS0 + nT1 * nS1 + nT2 * nS2;
    T PT[NQ];
#pragma unroll
    for (int j = 0; j < NQ; ++j) PT[j] = dT - (y[j][0] * nT0 + y[j][1] * nT1 + y[j][2] * nT2);
    T Dr[3][3], Di[3][3], Hr[3][3], Hi[3][3], Sr[3], Si[3], Kr[3], Ki[3], csr = 0, csi = 0;
#pragma unroll
    for (int p = 0; p < 3; ++p) {
      Sr[p] = Si[p] = Kr[p] = Ki[p] = 0;
#pragma unroll
      for (int b = 0; b < 3; ++b) Dr[p][b] = Di[p][b] = Hr[p][b] = Hi[p][b] = 0;
    }
#pragma unroll 1
    for (int i = 0; i < NQ; ++i) {
      const T x0 = xs[(3 * i) * 64 + lane], x1 = xs[(3 * i + 1) * 64 + lane], x2 = xs[(3 * i + 2) * 64 + lane];
      const T PS = dS - (x0 * nS0 + x1 * nS1 + x2 * nS2);
      const T cwi = cw * sphw[4 * i + 3];
      T mDr[3] = {0, 0, 0}, mDi[3] = {0, 0, 0}, mHr[3] = {0, 0, 0}, mHi[3] = {0, 0, 0};
      T gr = 0, gi = 0, kr = 0, ki = 0;
#pragma unroll
      for (int j = 0; j < NQ; ++j) {
        const T rx = y[j][0] - x0, ry = y[j][1] - x1, rz = y[j][2] - x2;
        const T r2 = rx * rx + ry * ry + rz * rz;
        const T ir = inv_sqrt(r2);
        const T r = r2 * ir;
        T s, c;
        phase_sincos<WIDE>(r, krev, a.ksteps, k, trig, s, c);
        const T g = ir * (cwi * sphw[4 * j + 3]);
        const T Gr = c * g, Gi = s * g;                            // G w
        const T Pr = -ir * Gr - k * Gi, Pi = -ir * Gi + k * Gr;  // G' w
        const T t = PS * ir;     // (r.n_y)/r
        const T u = PT[j] * ir;  // -(r.n_x)/r
        gr += Gr;
        gi += Gi;
        kr += Pr * u;
        ki += Pi * u;
#pragma unroll
        for (int b = 0; b < 3; ++b) {
          const T ph = sphw[4 * j + b];
          mDr[b] += ph * (Pr * t);
          mDi[b] += ph * (Pi * t);
          mHr[b] += ph * Gr;
          mHi[b] += ph * Gi;
        }
      }
      csr += gr;
      csi += gi;
#pragma unroll
      for (int p = 0; p < 3; ++p) {
        const T ph = sphw[4 * i + p];
        Sr[p] += ph * gr;
        Si[p] += ph * gi;
        Kr[p] += ph * kr;
        Ki[p] += ph * ki;
#pragma unroll
        for (int b = 0; b < 3; ++b) {
          Dr[p][b] += ph * mDr[b];
          Di[p][b] += ph * mDi[b];
          Hr[p][b] += ph * mHr[b];
          Hi[p][b] += ph * mHi[b];
        }
      }
    }
    const T kn = -k * k * nprod;
    const int sl[3] = {flocs[sf * 3], flocs[sf * 3 + 1], flocs[sf * 3 + 2]};
#pragma unroll
    for (int p = 0; p < 3; ++p) {
      const T c0 = tr[REC_CURL + 3 * p], c1 = tr[REC_CURL + 3 * p + 1], c2 = tr[REC_CURL + 3 * p + 2];
#pragma unroll
      for (int b = 0; b < 3; ++b) {
        const T b0 = sr[REC_CURL + 3 * b], b1 = sr[REC_CURL + 3 * b + 1], b2 = sr[REC_CURL + 3 * b + 2];
        const T cc = IMAGE ? c0 * (b0 * sg[3]) + c1 * (b1 * sg[4]) + c2 * (b2 * sg[5]) : c0 * b0 + c1 * b1 + c2 * b2;
        const int o2 = tl[p] * LDA_ + 2 * sl[b];
        lds_add(accD + o2, Dr[p][b]);
        lds_add(accD + o2 + 1, Di[p][b]);
        lds_add(accH + o2, cc * csr + kn * Hr[p][b]);
        lds_add(accH + o2 + 1, cc * csi + kn * Hi[p][b]);
      }
      const int oc = tl[p] * LDC_ + 2 * sj;
      lds_add(accS + oc, Sr[p]);
      lds_add(accS + oc + 1, Si[p]);
      lds_add(accK + oc, Kr[p]);
      lds_add(accK + oc + 1, Ki[p]);
    }
  }
  __syncthreads();
  for (int idx = tid; idx < nvT * nvS; idx += blockDim.x) {
    const int r = idx / nvS, c = idx - r * nvS;
    const int64_t dst = 2 * ((int64_t)grow[r] * o.ldd + gcol[c]);
    const T dr = accD[r * LDA_ + 2 * c], di = accD[r * LDA_ + 2 * c + 1];
    if (dr != T(0) || di != T(0)) {
      gadd(o.D + dst, dr);
      gadd(o.D + dst + 1, di);
    }
    const T hr = accH[r * LDA_ + 2 * c], hi = accH[r * LDA_ + 2 * c + 1];
    if (hr != T(0) || hi != T(0)) {
      gadd(o.H + dst, hr);
      gadd(o.H + dst + 1, hi);
    }
  }
  for (int idx = tid; idx < nvT * nS; idx += blockDim.x) {
    const int r = idx / nS, c = idx - r * nS;
    const int64_t dst = 2 * ((int64_t)grow[r] * o.ldf + s0 + c);
    const T sr_ = accS[r * LDC_ + 2 * c], si_ = accS[r * LDC_ + 2 * c + 1];
    if (sr_ != T(0) || si_ != T(0)) {
      gadd(o.S + dst, sr_);
      gadd(o.S + dst + 1, si_);
    }
    const T kr_ = accK[r * LDC_ + 2 * c], ki_ = accK[r * LDC_ + 2 * c + 1];
    if (kr_ != T(0) || ki_ != T(0)) {
      gadd(o.K + dst, kr_);
      gadd(o.K + dst + 1, ki_);
    }
  }
}

// Touching pairs of the four operators over the Duffy (or image delta) rules, FP64,
// one wavefront per pair.
template <typename TO>
__global__ __launch_bounds__(256) void bm_singular_operators_kernel(const SingularOperatorArgs<TO> o) {
  const SingularArgs<TO>& a = o.g;
  const int lane = threadIdx.x & 63;
  const int64_t pidx = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (pidx >= a.npairs) return;
  const SingularPair P = a.pairs[pidx];
  const double* tr = a.rec + (int64_t)P.test * FREC;
  const double* sr = a.rec + (int64_t)P.trial * FREC;
  double sg[6];
#pragma unroll
  for (int c = 0; c < 6; ++c) sg[c] = a.img[0][c];
#pragma unroll
  for (int x = 1; x <= MAX_IMAGES; ++x)
    if (P.xform == x)
#pragma unroll
      for (int c = 0; c < 6; ++c) sg[c] = a.img[x][c];
  double tv[3][3], sv[3][3];
  int tvi[3], svi[3];
  for (int v = 0; v < 3; ++v) {
    tvi[v] = a.fverts[(int64_t)P.test * 3 + v];
    svi[v] = a.fverts[(int64_t)P.trial * 3 + v];
    for (int c = 0; c < 3; ++c) {
      tv[v][c] = a.verts[3 * (int64_t)tvi[v] + c];
      sv[v][c] = a.verts[3 * (int64_t)svi[v] + c] * sg[c];
    }
  }
  const double nT[3] = {tr[REC_N], tr[REC_N + 1], tr[REC_N + 2]};
  const double nS[3] = {sr[REC_N] * sg[0], sr[REC_N + 1] * sg[1], sr[REC_N + 2] * sg[2]};
  const double k = a.k;
  const int off = a.rule_off[P.rule], cnt = a.rule_cnt[P.rule];
  // acc: D 0..17, Hphi 18..35 (sum phi phi G w), S 36..41, K 42..47, sum G w 48..49
  double acc[50];
#pragma unroll
  for (int i = 0; i < 50; ++i) acc[i] = 0;
  for (int q = lane; q < cnt; q += 64) {
    const double* rd = a.rule_data + 5 * (int64_t)(off + q);
    const double tx = rd[0], ty = rd[1], sx = rd[2], sy = rd[3];
    const double ta = 1.0 - tx - ty, sa = 1.0 - sx - sy;
    double x[3], y[3];
    for (int c = 0; c < 3; ++c) {
      x[c] = ta * tv[0][c] + tx * tv[1][c] + ty * tv[2][c];
      y[c] = sa * sv[0][c] + sx * sv[1][c] + sy * sv[2][c];
    }
    const double rx = y[0] - x[0], ry = y[1] - x[1], rz = y[2] - x[2];
    const double radius = sqrt(rx * rx + ry * ry + rz * rz);
    if (radius == 0.0) continue;
    const double ir = 1.0 / radius;
    double s, c;
    sincos_phase(k * radius, &s, &c);
    const double w = rd[4] * P.jac;
    const double Gr = c * INV_FOUR_PI * ir * w, Gi = s * INV_FOUR_PI * ir * w;  // G w
    const double Pr = Gr * -ir - Gi * k, Pi = Gr * k + Gi * -ir;                // G' w
    const double tdot = (rx * nS[0] + ry * nS[1] + rz * nS[2]) * ir;
    const double sdot = -(rx * nT[0] + ry * nT[1] + rz * nT[2]) * ir;
    const double pt[3] = {ta, tx, ty}, ps[3] = {sa, sx, sy};
    for (int p = 0; p < 3; ++p) {
      acc[36 + 2 * p] += pt[p] * Gr;
      acc[37 + 2 * p] += pt[p] * Gi;
      acc[42 + 2 * p] += pt[p] * (Pr * sdot);
      acc[43 + 2 * p] += pt[p] * (Pi * sdot);
      for (int b = 0; b < 3; ++b) {
        const double bp = pt[p] * ps[b];
        acc[2 * (3 * p + b)] += bp * (Pr * tdot);
        acc[2 * (3 * p + b) + 1] += bp * (Pi * tdot);
        acc[18 + 2 * (3 * p + b)] += bp * Gr;
        acc[19 + 2 * (3 * p + b)] += bp * Gi;
      }
    }
    acc[48] += Gr;
    acc[49] += Gi;
  }
#pragma unroll
  for (int i = 0; i < 50; ++i) {
    double v = acc[i];
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    acc[i] = v;
  }
  const double kn = -k * k * P.nprod;
  if (lane < 9) {
    const int p = lane / 3, b = lane % 3;
    const double cc = tr[REC_CURL + 3 * p] * (sr[REC_CURL + 3 * b] * sg[3]) +
                      tr[REC_CURL + 3 * p + 1] * (sr[REC_CURL + 3 * b + 1] * sg[4]) +
                      tr[REC_CURL + 3 * p + 2] * (sr[REC_CURL + 3 * b + 2] * sg[5]);
    double dr = 0, di = 0, hr = 0, hi = 0;
#pragma unroll
    for (int i = 0; i < 9; ++i)
      if (i == lane) {
        dr = acc[2 * i];
        di = acc[2 * i + 1];
        hr = cc * acc[48] + kn * acc[18 + 2 * i];
        hi = cc * acc[49] + kn * acc[19 + 2 * i];
      }
    const int64_t dst = 2 * ((int64_t)tvi[p] * o.ldd + svi[b]);
    gadd(o.D + dst, (TO)dr);
    gadd(o.D + dst + 1, (TO)di);
    gadd(o.H + dst, (TO)hr);
    gadd(o.H + dst + 1, (TO)hi);
  } else if (lane < 12) {
    const int p = lane - 9;
    double s_r = 0, s_i = 0, k_r = 0, k_i = 0;
#pragma unroll
    for (int i = 0; i < 3; ++i)
      if (i == p) {
        s_r = acc[36 + 2 * i];
        s_i = acc[37 + 2 * i];
        k_r = acc[42 + 2 * i];
        k_i = acc[43 + 2 * i];
      }
    const int64_t dst = 2 * ((int64_t)tvi[p] * o.ldf + P.trial);
    gadd(o.S + dst, (TO)s_r);
    gadd(o.S + dst + 1, (TO)s_i);
    gadd(o.K + dst, (TO)k_r);
    gadd(o.K + dst + 1, (TO)k_i);
  }
}

// Identity blocks: A += 0.5 M_p1p1, B += -0.5 eta M_p1dp0 q (BeatEngineCpuBurtonMiller.jl:450-453).
template <typename TO>
__global__ void bm_identity_kernel(const IdentityArgs<TO> a) {
  const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= a.nf) return;
  const double area = a.rec[f * FREC + REC_AREA];
  int v[3];
  for (int p = 0; p < 3; ++p) v[p] = a.fverts[3 * f + p];
  for (int p = 0; p < 3; ++p) {
    for (int b = 0; b < 3; ++b) gadd(a.A + 2 * ((int64_t)v[p] * a.lda + v[b]), (TO)(0.5 * (area * a.pp[p][b])));
    const double m = area * a.pd[p];
    // (-0.5 eta) * m * q
    const double cr = -0.5 * a.eta_re * m, ci = -0.5 * a.eta_im * m;
    for (int d = 0; d < a.nrhs; ++d) {
      const double* qq = a.q + 2 * (f * a.nrhs + d);
      gadd(a.B + 2 * ((int64_t)v[p] * a.nrhs + d), (TO)(cr * qq[0] - ci * qq[1]));
      gadd(a.B + 2 * ((int64_t)v[p] * a.nrhs + d) + 1, (TO)(cr * qq[1] + ci * qq[0]));
    }
  }
}

// Orbit weights of the rows on a mirror plane (the reference scales the assembled
// rows, BeatEngineCpuBurtonMiller.jl:443-449, and its identity blocks carry the
// same weights, BeatEngineCore.jl:1187-1214): one workgroup per listed row.
template <typename TO>
__global__ void row_scale_kernel(TO* A, int64_t lda, int64_t n, TO* B, int nrhs, const int32_t* rows, const double* w) {
  const int64_t r = rows[blockIdx.x];
  const TO s = (TO)w[blockIdx.x];
  TO* a = A + 2 * r * lda;
  for (int64_t i = threadIdx.x; i < 2 * n; i += blockDim.x) a[i] *= s;
  TO* b = B + 2 * r * nrhs;
  for (int i = threadIdx.x; i < 2 * nrhs; i += blockDim.x) b[i] *= s;
}

template <typename T>
__global__ void fill_kernel(T* p, int64_t n, T v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ void downcast_kernel(const double* src, float* dst, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = (float)src[i];
}

// ---- launchers -------------------------------------------------------------------

// Dynamic LDS that lifts a workgroup's LDS above half of a CU's 160 KB, so that
// one workgroup of the kernel runs per CU (0 if the query fails: every slot).
template <typename K>
static unsigned one_per_cu_pad(K kernel) {
  hipFuncAttributes at;
  if (hipFuncGetAttributes(&at, reinterpret_cast<const void*>(kernel)) != hipSuccess) return 0u;
  constexpr size_t half = 80 * 1024;
  return at.sharedSizeBytes < half + 1024 ? (unsigned)(half + 1024 - at.sharedSizeBytes) : 0u;
}

template <typename T, int NQ, bool WIDE, bool IMAGE, int TFS, int TVS>
static void launch_regular_t(const RegularArgs<T>& a, bool pure, dim3 grid, dim3 block, hipStream_t s) {
  if (pure) {
    const auto k = bm_regular_kernel<T, NQ, true, WIDE, IMAGE, TFS, TVS>;
    hipLaunchKernelGGL(k, grid, block, a.one_per_cu ? one_per_cu_pad(k) : 0u, s, a);
  } else {
    const auto k = bm_regular_kernel<T, NQ, false, WIDE, IMAGE, TFS, TVS>;
    hipLaunchKernelGGL(k, grid, block, a.one_per_cu ? one_per_cu_pad(k) : 0u, s, a);
  }
}

// the kernel's tile shape follows the trial partition the arguments carry (the
// image passes always carry the wide one)
template <typename T, int NQ, bool WIDE, bool IMAGE>
static void launch_regular_q(const RegularArgs<T>& a, bool pure, dim3 grid, dim3 block, hipStream_t s) {
  if constexpr (!IMAGE) {
    if (a.narrow) {
      launch_regular_t<T, NQ, WIDE, IMAGE, TILE_FS_N, TILE_VS_N>(a, pure, grid, block, s);
      return;
    }
  }
  launch_regular_t<T, NQ, WIDE, IMAGE, TILE_FS, TILE_VS>(a, pure, grid, block, s);
}

template <typename T, bool WIDE, bool IMAGE>
static void launch_regular_nq(const RegularArgs<T>& a, int nq, bool pure, dim3 grid, dim3 block, hipStream_t s) {
  if (nq == 6) launch_regular_q<T, 6, WIDE, IMAGE>(a, pure, grid, block, s);
  else if (nq == 3) launch_regular_q<T, 3, WIDE, IMAGE>(a, pure, grid, block, s);
  else launch_regular_q<T, 1, WIDE, IMAGE>(a, pure, grid, block, s);
}

template <typename T, bool IMAGE>
static void launch_regular_range(const RegularArgs<T>& a, int nq, bool pure, bool wide, dim3 grid, dim3 block,
                                 hipStream_t s) {
  // the FP32 phase needs no range bound
  if constexpr (sizeof(T) == 8) {
    if (wide) {
      launch_regular_nq<T, true, IMAGE>(a, nq, pure, grid, block, s);
      return;
    }
  }
  (void)wide;
  launch_regular_nq<T, false, IMAGE>(a, nq, pure, grid, block, s);
}

template <typename T, int NQ, bool WIDE>
static void launch_regular_sym_q(const RegularArgs<T>& a, bool pure, dim3 grid, hipStream_t s) {
  if (pure) {
    const auto k = bm_regular_sym_kernel<T, NQ, true, WIDE>;
    hipLaunchKernelGGL(k, grid, dim3(256), a.one_per_cu ? one_per_cu_pad(k) : 0u, s, a);
  } else {
    const auto k = bm_regular_sym_kernel<T, NQ, false, WIDE>;
    hipLaunchKernelGGL(k, grid, dim3(256), a.one_per_cu ? one_per_cu_pad(k) : 0u, s, a);
  }
}

template <typename T, bool WIDE>
static void launch_regular_sym_nq(const RegularArgs<T>& a, int nq, bool pure, dim3 grid, hipStream_t s) {
  if (nq == 6) launch_regular_sym_q<T, 6, WIDE>(a, pure, grid, s);
  else if (nq == 3) launch_regular_sym_q<T, 3, WIDE>(a, pure, grid, s);
  else launch_regular_sym_q<T, 1, WIDE>(a, pure, grid, s);
}

// the symmetric pass: one workgroup per (test tile, sub-tile) unit
template <typename T>
static hipError_t launch_regular_sym(const RegularArgs<T>& a, int nq, bool pure, bool wide, hipStream_t s) {
  if (!a.unit_start || !a.tile_qfirst || a.nqtiles < 1 || a.nunits < 1) return hipErrorInvalidValue;
  const int64_t blocks = (a.nunits + 7) / 8 * 8;
  if (blocks > 0x7fffffff) return hipErrorInvalidValue;
  const dim3 grid((unsigned)blocks);
  if constexpr (sizeof(T) == 8) {
    if (wide) {
      launch_regular_sym_nq<T, true>(a, nq, pure, grid, s);
      return hipGetLastError();
    }
  }
  launch_regular_sym_nq<T, false>(a, nq, pure, grid, s);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_regular(const RegularArgs<T>& a, int nq, bool pure, bool wide, bool image, hipStream_t s) {
  if (!image && a.symmetric) return launch_regular_sym<T>(a, nq, pure, wide, s);
  if (image && a.narrow) return hipErrorInvalidValue;
  const int64_t total = (int64_t)a.ntiles * a.nstiles;
  const int64_t per = (total + 7) / 8;
  if (image && (a.nimg < 1 || a.nimg > MAX_IMAGES)) return hipErrorInvalidValue;
  const int64_t blocks = per * 8 * (image ? a.nimg : 1);
  if (blocks > 0x7fffffff) return hipErrorInvalidValue;
  dim3 grid((unsigned)blocks), block(256);
  if (image) launch_regular_range<T, true>(a, nq, pure, wide, grid, block, s);
  else launch_regular_range<T, false>(a, nq, pure, wide, grid, block, s);
  return hipGetLastError();
}

template <typename T, int NQ, bool IMAGE>
static void launch_operators_nq(const OperatorArgs<T>& a, bool wide, dim3 grid, dim3 block, hipStream_t s) {
  if constexpr (sizeof(T) == 8) {
    if (wide) {
      hipLaunchKernelGGL((bm_operators_kernel<T, NQ, true, IMAGE>), grid, block, 0, s, a);
      return;
    }
  }
  hipLaunchKernelGGL((bm_operators_kernel<T, NQ, false, IMAGE>), grid, block, 0, s, a);
}

template <typename T>
hipError_t launch_operators(const OperatorArgs<T>& a, int nq, bool wide, bool image, hipStream_t s) {
  const int64_t total = (int64_t)a.g.ntiles * a.g.nstiles;
  const int64_t per = (total + 7) / 8;
  if (image && (a.g.nimg < 1 || a.g.nimg > MAX_IMAGES)) return hipErrorInvalidValue;
  const int64_t blocks = per * 8 * (image ? a.g.nimg : 1);
  if (blocks > 0x7fffffff) return hipErrorInvalidValue;
  dim3 grid((unsigned)blocks), block(256);
  if (nq == 6) {
    if (image) launch_operators_nq<T, 6, true>(a, wide, grid, block, s);
    else launch_operators_nq<T, 6, false>(a, wide, grid, block, s);
  } else if (nq == 3) {
    if (image) launch_operators_nq<T, 3, true>(a, wide, grid, block, s);
    else launch_operators_nq<T, 3, false>(a, wide, grid, block, s);
  } else {
    if (image) launch_operators_nq<T, 1, true>(a, wide, grid, block, s);
    else launch_operators_nq<T, 1, false>(a, wide, grid, block, s);
  }
  return hipGetLastError();
}

template <typename TO>
hipError_t launch_singular_operators(const SingularOperatorArgs<TO>& a, hipStream_t s) {
  if (a.g.npairs == 0) return hipSuccess;
  dim3 grid((unsigned)((a.g.npairs + 3) / 4)), block(256);
  hipLaunchKernelGGL((bm_singular_operators_kernel<TO>), grid, block, 0, s, a);
  return hipGetLastError();
}

template <typename TO>
hipError_t launch_singular(const SingularArgs<TO>& a, hipStream_t s) {
  if (a.npairs == 0) return hipSuccess;
  if (!a.wide && !a.trig) return hipErrorInvalidValue;
  // 16 pairs per workgroup of 256 (a 16-lane group each)
  dim3 grid((unsigned)((a.npairs + 15) / 16)), block(256);
  if (a.wide) hipLaunchKernelGGL((bm_singular_kernel<TO, true>), grid, block, 0, s, a);
  else hipLaunchKernelGGL((bm_singular_kernel<TO, false>), grid, block, 0, s, a);
  return hipGetLastError();
}

template <typename TO>
hipError_t launch_identity(const IdentityArgs<TO>& a, hipStream_t s) {
  dim3 grid((unsigned)((a.nf + 255) / 256)), block(256);
  hipLaunchKernelGGL((bm_identity_kernel<TO>), grid, block, 0, s, a);
  return hipGetLastError();
}

template <typename TO>
hipError_t launch_row_scale(TO* A, int64_t lda, int64_t n, TO* B, int nrhs, const int32_t* rows, const double* w,
                            int64_t nrows, hipStream_t s) {
  if (nrows <= 0) return hipSuccess;
  hipLaunchKernelGGL((row_scale_kernel<TO>), dim3((unsigned)nrows), dim3(256), 0, s, A, lda, n, B, nrhs, rows, w);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_fill(T* p, int64_t n, T v, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  int64_t blocks = std::min<int64_t>((n + 255) / 256, 8192);
  hipLaunchKernelGGL((fill_kernel<T>), dim3((unsigned)blocks), dim3(256), 0, s, p, n, v);
  return hipGetLastError();
}

hipError_t launch_downcast(const double* src, float* dst, int64_t n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  int64_t blocks = std::min<int64_t>((n + 255) / 256, 8192);
  hipLaunchKernelGGL(downcast_kernel, dim3((unsigned)blocks), dim3(256), 0, s, src, dst, n);
  return hipGetLastError();
}

template hipError_t launch_regular<float>(const RegularArgs<float>&, int, bool, bool, bool, hipStream_t);
template hipError_t launch_regular<double>(const RegularArgs<double>&, int, bool, bool, bool, hipStream_t);
template hipError_t launch_row_scale<float>(float*, int64_t, int64_t, float*, int, const int32_t*, const double*,
                                            int64_t, hipStream_t);
template hipError_t launch_row_scale<double>(double*, int64_t, int64_t, double*, int, const int32_t*, const double*,
                                             int64_t, hipStream_t);
template hipError_t launch_operators<float>(const OperatorArgs<float>&, int, bool, bool, hipStream_t);
template hipError_t launch_operators<double>(const OperatorArgs<double>&, int, bool, bool, hipStream_t);
template hipError_t launch_singular_operators<float>(const SingularOperatorArgs<float>&, hipStream_t);
template hipError_t launch_singular_operators<double>(const SingularOperatorArgs<double>&, hipStream_t);
template hipError_t launch_singular<float>(const SingularArgs<float>&, hipStream_t);
template hipError_t launch_singular<double>(const SingularArgs<double>&, hipStream_t);
template hipError_t launch_identity<float>(const IdentityArgs<float>&, hipStream_t);
template hipError_t launch_identity<double>(const IdentityArgs<double>&, hipStream_t);
template hipError_t launch_fill<float>(float*, int64_t, float, hipStream_t);
template hipError_t launch_fill<double>(double*, int64_t, double, hipStream_t);
template hipError_t launch_fill<int>(int*, int64_t, int, hipStream_t);

}  // namespace beat

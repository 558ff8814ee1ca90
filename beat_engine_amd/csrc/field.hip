// Exterior field from the solved traces (the reference's
// evaluate_galerkin_field_cpu, src/beat_engine/julia_local/src/BeatEngineCpuField.jl:1-49):
//   p(x) = sum_s [ G (ik - 1/r) (r.n_s)/r  p_s w_s  -  G q_s w_s ],  r = y_s - x
// over the quadrature sources y_s of every face, and with a symmetry mode of
// every face's images (build_field_evaluation_cache, BeatEngineCore.jl:1770-1810:
// reflected points and normals, the same weighted trace values). field_sources
// forms the weighted source values once per call; field_kernel stages sources
// through LDS and splits the source range across workgroups, each adding its
// partial sums.
#include <hip/hip_runtime.h>

#include "bem_common.h"
#include "kernels.h"

namespace beat {

namespace {

constexpr int FIELD_PTS = 128;   // points per workgroup (one per thread)
constexpr int FIELD_CHUNK = 128; // sources staged per LDS round
constexpr int FIELD_RHS = 8;     // right-hand sides per pass

struct FieldImages {
  int s[MAX_IMAGES + 1][3];  // point signs per transform; row 0 the identity
};

template <typename T>
__global__ void field_sources_kernel(const double* rec, const int32_t* fverts, int64_t nf, int nq,
                                     const double* __restrict__ rule, const T* p, const T* q, int nrhs, int ntr,
                                     const FieldImages im, T* sp, T* sn, T* pw, T* qw) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nb = nf * nq;
  if (s >= nb * ntr) return;
  const int x = (int)(s / nb);
  const int64_t sb = s - (int64_t)x * nb;
  const int64_t f = sb / nq;
  const int qi = (int)(sb - f * nq);
  double sg[3] = {1.0, 1.0, 1.0};
#pragma unroll
  for (int t = 1; t <= MAX_IMAGES; ++t)
    if (t == x)
#pragma unroll
      for (int c = 0; c < 3; ++c) sg[c] = (double)im.s[t][c];
  const double* r = rec + f * FREC;
  for (int c = 0; c < 3; ++c) {
    sp[3 * s + c] = (T)(r[3 * qi + c] * sg[c]);
    sn[3 * s + c] = (T)(r[REC_N + c] * sg[c]);
  }
  const double xi = rule[3 * qi], eta = rule[3 * qi + 1], wq = rule[3 * qi + 2];
  const double w = wq * (2.0 * r[REC_AREA]);
  const double ph[3] = {1.0 - xi - eta, xi, eta};
  for (int d = 0; d < nrhs; ++d) {
    double ar = 0, ai = 0;
    for (int a = 0; a < 3; ++a) {
      const int64_t v = fverts[3 * f + a];
      ar += ph[a] * (double)p[2 * (v * nrhs + d)];
      ai += ph[a] * (double)p[2 * (v * nrhs + d) + 1];
    }
    pw[2 * (s * nrhs + d)] = (T)(ar * w);
    pw[2 * (s * nrhs + d) + 1] = (T)(ai * w);
    qw[2 * (s * nrhs + d)] = (T)((double)q[2 * (f * nrhs + d)] * w);
    qw[2 * (s * nrhs + d) + 1] = (T)((double)q[2 * (f * nrhs + d) + 1] * w);
  }
}

__device__ __forceinline__ void green_sc(float kr, float& s, float& c) {
  float t = kr * 0.15915494309189535f;
  t = t - rintf(t);
  s = __builtin_amdgcn_sinf(t);
  c = __builtin_amdgcn_cosf(t);
}
__device__ __forceinline__ void green_sc(double kr, double& s, double& c) { sincos(kr, &s, &c); }

template <typename T>
__global__ __launch_bounds__(FIELD_PTS) void field_kernel(const T* sp, const T* sn, const T* pw, const T* qw,
                                                          int64_t nsrc, int nrhs, int d0, const T* pts, int64_t npts,
                                                          T* out, T k, int64_t chunk) {
  __shared__ T ss[FIELD_CHUNK * 6];
  __shared__ T sv[FIELD_CHUNK * FIELD_RHS * 4];
  const int tid = threadIdx.x;
  const int64_t pi = (int64_t)blockIdx.x * FIELD_PTS + tid;
  const int64_t s_begin = (int64_t)blockIdx.y * chunk;
  const int64_t s_end = min(nsrc, s_begin + chunk);
  const int nd = min(FIELD_RHS, nrhs - d0);
  T x0 = 0, x1 = 0, x2 = 0;
  if (pi < npts) {
    x0 = pts[3 * pi];
    x1 = pts[3 * pi + 1];
    x2 = pts[3 * pi + 2];
  }
  T ar[FIELD_RHS], ai[FIELD_RHS];
#pragma unroll
  for (int d = 0; d < FIELD_RHS; ++d) ar[d] = ai[d] = 0;
  const T c4 = T(INV_FOUR_PI);
  for (int64_t b = s_begin; b < s_end; b += FIELD_CHUNK) {
    const int n = (int)min((int64_t)FIELD_CHUNK, s_end - b);
    __syncthreads();
    for (int i = tid; i < n; i += blockDim.x) {
      for (int c = 0; c < 3; ++c) {
        ss[6 * i + c] = sp[3 * (b + i) + c];
        ss[6 * i + 3 + c] = sn[3 * (b + i) + c];
      }
      for (int d = 0; d < nd; ++d) {
        const int64_t o = 2 * ((b + i) * nrhs + d0 + d);
        sv[4 * (i * FIELD_RHS + d)] = pw[o];
        sv[4 * (i * FIELD_RHS + d) + 1] = pw[o + 1];
        sv[4 * (i * FIELD_RHS + d) + 2] = qw[o];
        sv[4 * (i * FIELD_RHS + d) + 3] = qw[o + 1];
      }
    }
    __syncthreads();
    if (pi < npts) {
      for (int i = 0; i < n; ++i) {
        const T rx = ss[6 * i] - x0, ry = ss[6 * i + 1] - x1, rz = ss[6 * i + 2] - x2;
        const T r = sqrt(rx * rx + ry * ry + rz * rz);
        if (r == T(0)) continue;
        const T ir = T(1) / r;
        T s, c;
        green_sc(k * r, s, c);
        const T gs = c4 * ir;
        const T Gr = c * gs, Gi = s * gs;
        const T proj = (rx * ss[6 * i + 3] + ry * ss[6 * i + 4] + rz * ss[6 * i + 5]) * ir;
        // G (ik - 1/r) proj
        const T dr = (-Gr * ir - Gi * k) * proj, di = (Gr * k - Gi * ir) * proj;
#pragma unroll
        for (int d = 0; d < FIELD_RHS; ++d) {
          if (d >= nd) break;
          const T pr = sv[4 * (i * FIELD_RHS + d)], pim = sv[4 * (i * FIELD_RHS + d) + 1];
          const T qr = sv[4 * (i * FIELD_RHS + d) + 2], qim = sv[4 * (i * FIELD_RHS + d) + 3];
          ar[d] += dr * pr - di * pim - (Gr * qr - Gi * qim);
          ai[d] += dr * pim + di * pr - (Gr * qim + Gi * qr);
        }
      }
    }
  }
  if (pi < npts)
#pragma unroll
    for (int d = 0; d < FIELD_RHS; ++d)
      if (d < nd) {
        unsafeAtomicAdd(out + 2 * (pi * nrhs + d0 + d), ar[d]);
        unsafeAtomicAdd(out + 2 * (pi * nrhs + d0 + d) + 1, ai[d]);
      }
}

// ---- value and observation gradient (pressure_gradient_v1) -------------------------
// field_value_gradient (BeatEngineFieldGradient.jl:2-13): with r = y - x, u = r / |r|,
// a = -1/r + i k, b = 3/r^2 - k^2 - 3 i k / r and proj = u.n,
//   value    = G (a proj p - q)
//   gradient = G (a q u - p (b proj u + (a / r) n))
// summed over the same weighted sources as the field (pw = p w, qw = q w).
constexpr int GRAD_RHS = 4;  // right-hand sides per pass (4 complex outputs each)

template <typename T>
__global__ __launch_bounds__(FIELD_PTS) void field_gradient_kernel(const T* sp, const T* sn, const T* pw, const T* qw,
                                                                   int64_t nsrc, int nrhs, int d0, const T* pts,
                                                                   int64_t npts, T* out, T k, int64_t chunk) {
  __shared__ T ss[FIELD_CHUNK * 6];
  __shared__ T sv[FIELD_CHUNK * GRAD_RHS * 4];
  const int tid = threadIdx.x;
  const int64_t pi = (int64_t)blockIdx.x * FIELD_PTS + tid;
  const int64_t s_begin = (int64_t)blockIdx.y * chunk;
  const int64_t s_end = min(nsrc, s_begin + chunk);
  const int nd = min(GRAD_RHS, nrhs - d0);
  T x0 = 0, x1 = 0, x2 = 0;
  if (pi < npts) {
    x0 = pts[3 * pi];
    x1 = pts[3 * pi + 1];
    x2 = pts[3 * pi + 2];
  }
  T acc[GRAD_RHS][4][2];
#pragma unroll
  for (int d = 0; d < GRAD_RHS; ++d)
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[d][c][0] = acc[d][c][1] = 0;
  const T c4 = T(INV_FOUR_PI);
  for (int64_t b0 = s_begin; b0 < s_end; b0 += FIELD_CHUNK) {
    const int n = (int)min((int64_t)FIELD_CHUNK, s_end - b0);
    __syncthreads();
    for (int i = tid; i < n; i += blockDim.x) {
      for (int c = 0; c < 3; ++c) {
        ss[6 * i + c] = sp[3 * (b0 + i) + c];
        ss[6 * i + 3 + c] = sn[3 * (b0 + i) + c];
      }
      for (int d = 0; d < nd; ++d) {
        const int64_t o = 2 * ((b0 + i) * nrhs + d0 + d);
        sv[4 * (i * GRAD_RHS + d)] = pw[o];
        sv[4 * (i * GRAD_RHS + d) + 1] = pw[o + 1];
        sv[4 * (i * GRAD_RHS + d) + 2] = qw[o];
        sv[4 * (i * GRAD_RHS + d) + 3] = qw[o + 1];
      }
    }
    __syncthreads();
    if (pi < npts) {
      for (int i = 0; i < n; ++i) {
        const T rx = ss[6 * i] - x0, ry = ss[6 * i + 1] - x1, rz = ss[6 * i + 2] - x2;
        const T r = sqrt(rx * rx + ry * ry + rz * rz);
        const T ir = T(1) / r;
        const T ux = rx * ir, uy = ry * ir, uz = rz * ir;
        const T nx = ss[6 * i + 3], ny = ss[6 * i + 4], nz = ss[6 * i + 5];
        T s, c;
        green_sc(k * r, s, c);
        const T gs = c4 * ir;
        const T Gr = c * gs, Gi = s * gs;
        const T proj = ux * nx + uy * ny + uz * nz;
        // G a = G (-1/r + i k)
        const T Gar = -Gr * ir - Gi * k, Gai = Gr * k - Gi * ir;
        // G b proj, b = (3/r^2 - k^2) - i (3 k / r)
        const T br = T(3) * ir * ir - k * k, bi = -T(3) * k * ir;
        const T Gbr = (Gr * br - Gi * bi) * proj, Gbi = (Gr * bi + Gi * br) * proj;
        // per component: Bv = G (b proj u + (a / r) n) = Gb u + (Ga / r) n
        const T Bxr = Gbr * ux + Gar * ir * nx, Bxi = Gbi * ux + Gai * ir * nx;
        const T Byr = Gbr * uy + Gar * ir * ny, Byi = Gbi * uy + Gai * ir * ny;
        const T Bzr = Gbr * uz + Gar * ir * nz, Bzi = Gbi * uz + Gai * ir * nz;
        const T Apr = Gar * proj, Api = Gai * proj;  // G a proj
#pragma unroll
        for (int d = 0; d < GRAD_RHS; ++d) {
          if (d >= nd) break;
          const T pr = sv[4 * (i * GRAD_RHS + d)], pim = sv[4 * (i * GRAD_RHS + d) + 1];
          const T qr = sv[4 * (i * GRAD_RHS + d) + 2], qim = sv[4 * (i * GRAD_RHS + d) + 3];
          acc[d][0][0] += Apr * pr - Api * pim - (Gr * qr - Gi * qim);
          acc[d][0][1] += Apr * pim + Api * pr - (Gr * qim + Gi * qr);
          const T gqr = Gar * qr - Gai * qim, gqi = Gar * qim + Gai * qr;  // G a q
          acc[d][1][0] += gqr * ux - (pr * Bxr - pim * Bxi);
          acc[d][1][1] += gqi * ux - (pr * Bxi + pim * Bxr);
          acc[d][2][0] += gqr * uy - (pr * Byr - pim * Byi);
          acc[d][2][1] += gqi * uy - (pr * Byi + pim * Byr);
          acc[d][3][0] += gqr * uz - (pr * Bzr - pim * Bzi);
          acc[d][3][1] += gqi * uz - (pr * Bzi + pim * Bzr);
        }
      }
    }
  }
  if (pi < npts)
#pragma unroll
    for (int d = 0; d < GRAD_RHS; ++d)
      if (d < nd)
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          T* o = out + 2 * ((pi * nrhs + d0 + d) * 4 + c);
          unsafeAtomicAdd(o, acc[d][c][0]);
          unsafeAtomicAdd(o + 1, acc[d][c][1]);
        }
}

// Near-panel correction (correct_field_gradient!, BeatEngineFieldGradient.jl:45-73):
// for every point and every face image whose centroid lies within 3 diameters,
// add the FP64 adaptive integral minus the 6-point rule in T. near_pairs_kernel
// lists the (point, face image) pairs; correction_kernel integrates one pair per
// thread. The integrand is linear in (p, q), so each pair forms the moments
//   Mp[c][a] = int phi_a Fp_c,  Mq[c] = int Fq_c   (Fp, Fq the p and q parts)
// once and applies them to every right-hand side.
struct GradImages {
  int n;                        // transforms including the identity
  int s[MAX_IMAGES + 1][3];     // point signs; row 0 the identity
};

template <typename T>
__device__ __forceinline__ void face_image(const double* verts, const int32_t* fverts, int64_t f, const int* sg,
                                           T v[3][3]) {
  for (int a = 0; a < 3; ++a) {
    const int64_t vi = fverts[3 * f + a];
    for (int c = 0; c < 3; ++c) v[a][c] = (T)verts[3 * vi + c] * (T)sg[c];
  }
}

template <typename T>
__device__ __forceinline__ void pick_signs(const GradImages& im, int t, int sg[3]) {
  sg[0] = sg[1] = sg[2] = 1;
#pragma unroll
  for (int x = 1; x <= MAX_IMAGES; ++x)
    if (x == t)
#pragma unroll
      for (int c = 0; c < 3; ++c) sg[c] = im.s[x][c];
}

template <typename T>
__device__ __forceinline__ T dist3(const T a[3], const T b[3]) {
  const T d0 = a[0] - b[0], d1 = a[1] - b[1], d2 = a[2] - b[2];
  return sqrt(d0 * d0 + d1 * d1 + d2 * d2);
}

// fill = false: count the near pairs into *count; fill = true: also store them
// (point, face * ntr + transform) while the count stays below cap.
template <typename T>
__global__ void near_pairs_kernel(const double* verts, const int32_t* fverts, int64_t nf, const GradImages im,
                                  const T* pts, int64_t npts, unsigned long long* count, int2* list,
                                  unsigned long long cap, bool fill) {
  const int64_t nft = nf * im.n;
  const int64_t total = npts * nft;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = idx / nft;
    const int64_t ft = idx - i * nft;
    const int64_t f = ft / im.n;
    const int t = (int)(ft - f * im.n);
    int sg[3];
    pick_signs<T>(im, t, sg);
    T v[3][3];
    face_image<T>(verts, fverts, f, sg, v);
    T cen[3];
    for (int c = 0; c < 3; ++c) cen[c] = (v[0][c] + v[1][c] + v[2][c]) / T(3);
    const T diam = max(max(dist3(v[0], v[1]), dist3(v[1], v[2])), dist3(v[2], v[0]));
    const T x[3] = {pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]};
    if (!(dist3(x, cen) <= T(3) * diam)) continue;
    const unsigned long long slot = atomicAdd(count, 1ull);
    if (fill && slot < cap) list[slot] = make_int2((int)i, (int)ft);
  }
}

// moments of one triangle (vertices v, normal n, jac |cross|) under the 6-point rule,
// accumulated with the root-face barycentric basis phi(l) = B0 + l.x B1 + l.y B2
template <typename R>
__device__ void gradient_moments(const R x[3], const R v[3][3], const R n[3], R k, const R* rule, const R bary[3][2],
                                 R Mp[4][3][2], R Mq[4][2]) {
  const R e1[3] = {v[1][0] - v[0][0], v[1][1] - v[0][1], v[1][2] - v[0][2]};
  const R e2[3] = {v[2][0] - v[0][0], v[2][1] - v[0][1], v[2][2] - v[0][2]};
  const R cr[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
  const R jac = sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
  const R c4 = R(INV_FOUR_PI);
  for (int j = 0; j < 6; ++j) {
    const R xi = rule[3 * j], eta = rule[3 * j + 1], w = rule[3 * j + 2] * jac;
    const R l0 = R(1) - xi - eta;
    R y[3];
    for (int c = 0; c < 3; ++c) y[c] = l0 * v[0][c] + xi * v[1][c] + eta * v[2][c];
    // root-face basis at y: barycentric coordinates of the sub-triangle's vertices
    const R rx = l0 * bary[0][0] + xi * bary[1][0] + eta * bary[2][0];
    const R ry = l0 * bary[0][1] + xi * bary[1][1] + eta * bary[2][1];
    const R phi[3] = {R(1) - rx - ry, rx, ry};
    const R d0 = y[0] - x[0], d1 = y[1] - x[1], d2 = y[2] - x[2];
    const R r = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
    const R ir = R(1) / r;
    const R u[3] = {d0 * ir, d1 * ir, d2 * ir};
    R s, c;
    green_sc(k * r, s, c);
    const R Gr = c * c4 * ir * w, Gi = s * c4 * ir * w;  // G w
    const R proj = u[0] * n[0] + u[1] * n[1] + u[2] * n[2];
    const R Gar = -Gr * ir - Gi * k, Gai = Gr * k - Gi * ir;
    const R br = R(3) * ir * ir - k * k, bi = -R(3) * k * ir;
    const R Gbr = (Gr * br - Gi * bi) * proj, Gbi = (Gr * bi + Gi * br) * proj;
    // p part: value G a proj, gradient -(G b proj u + (G a / r) n)
    R Fp[4][2], Fq[4][2];
    Fp[0][0] = Gar * proj;
    Fp[0][1] = Gai * proj;
    Fq[0][0] = -Gr;
    Fq[0][1] = -Gi;
    for (int cc = 0; cc < 3; ++cc) {
      Fp[1 + cc][0] = -(Gbr * u[cc] + Gar * ir * n[cc]);
      Fp[1 + cc][1] = -(Gbi * u[cc] + Gai * ir * n[cc]);
      Fq[1 + cc][0] = Gar * u[cc];
      Fq[1 + cc][1] = Gai * u[cc];
    }
    for (int cc = 0; cc < 4; ++cc) {
      for (int a = 0; a < 3; ++a) {
        Mp[cc][a][0] += phi[a] * Fp[cc][0];
        Mp[cc][a][1] += phi[a] * Fp[cc][1];
      }
      Mq[cc][0] += Fq[cc][0];
      Mq[cc][1] += Fq[cc][1];
    }
  }
}

// out: npts x nrhs x 4 complex T (internal order); p n x nrhs, q F x nrhs row-major
// complex T; rule4: the 6-point rule (xi, eta, w) in double; *failed counts pairs
// beyond depth 12.
template <typename T>
__global__ __launch_bounds__(64) void gradient_correction_kernel(const double* verts, const int32_t* fverts,
                                                                 const double* rec, const GradImages im,
                                                                 const int2* list, int64_t npairs, const T* pts,
                                                                 const T* p, const T* q, int nrhs, double k,
                                                                 const double* rule4, T* out, int* failed) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= npairs) return;
  const int2 pr = list[e];
  const int64_t i = pr.x, ft = pr.y;
  const int64_t f = ft / im.n;
  const int t = (int)(ft - f * im.n);
  int sg[3];
  pick_signs<T>(im, t, sg);
  // coarse: the 6-point rule on the face in T
  T vT[3][3], nT[3], xT[3], ruleT[18];
  face_image<T>(verts, fverts, f, sg, vT);
  for (int c = 0; c < 3; ++c) {
    nT[c] = (T)rec[f * FREC + REC_N + c] * (T)sg[c];
    xT[c] = pts[3 * i + c];
  }
  for (int j = 0; j < 18; ++j) ruleT[j] = (T)rule4[j];
  const T baryT[3][2] = {{T(0), T(0)}, {T(1), T(0)}, {T(0), T(1)}};
  T McP[4][3][2] = {}, McQ[4][2] = {};
  gradient_moments<T>(xT, vT, nT, (T)k, ruleT, baryT, McP, McQ);
  // fine: adaptive subdivision in FP64, depth-first over the path of child digits
  double v[3][3], n[3], x[3];
  face_image<double>(verts, fverts, f, sg, v);
  for (int c = 0; c < 3; ++c) {
    n[c] = rec[f * FREC + REC_N + c] * sg[c];
    x[c] = (double)pts[3 * i + c];
  }
  double MfP[4][3][2] = {}, MfQ[4][2] = {};
  uint32_t path = 0;  // two bits per level, level 1 in the low bits
  int depth = 0;
  bool bad = false;
  for (;;) {
    // node vertices: replay the path from the root (barycentric and physical)
    double b[3][2] = {{0, 0}, {1, 0}, {0, 1}};
    double w[3][3];
    for (int a = 0; a < 3; ++a)
      for (int c = 0; c < 3; ++c) w[a][c] = v[a][c];
    for (int l = 0; l < depth; ++l) {
      const int dgt = (path >> (2 * l)) & 3;
      double mb[3][2], mw[3][3];  // midpoints ab, bc, ca
      for (int c = 0; c < 2; ++c) {
        mb[0][c] = (b[0][c] + b[1][c]) / 2;
        mb[1][c] = (b[1][c] + b[2][c]) / 2;
        mb[2][c] = (b[2][c] + b[0][c]) / 2;
      }
      for (int c = 0; c < 3; ++c) {
        mw[0][c] = (w[0][c] + w[1][c]) / 2;
        mw[1][c] = (w[1][c] + w[2][c]) / 2;
        mw[2][c] = (w[2][c] + w[0][c]) / 2;
      }
      // children: (a, ab, ca), (ab, b, bc), (ca, bc, c), (ab, bc, ca)
      double nb[3][2], nw[3][3];
      for (int c = 0; c < 2; ++c) {
        nb[0][c] = dgt == 0 ? b[0][c] : dgt == 1 ? mb[0][c] : dgt == 2 ? mb[2][c] : mb[0][c];
        nb[1][c] = dgt == 0 ? mb[0][c] : dgt == 1 ? b[1][c] : dgt == 2 ? mb[1][c] : mb[1][c];
        nb[2][c] = dgt == 0 ? mb[2][c] : dgt == 1 ? mb[1][c] : dgt == 2 ? b[2][c] : mb[2][c];
      }
      for (int c = 0; c < 3; ++c) {
        nw[0][c] = dgt == 0 ? w[0][c] : dgt == 1 ? mw[0][c] : dgt == 2 ? mw[2][c] : mw[0][c];
        nw[1][c] = dgt == 0 ? mw[0][c] : dgt == 1 ? w[1][c] : dgt == 2 ? mw[1][c] : mw[1][c];
        nw[2][c] = dgt == 0 ? mw[2][c] : dgt == 1 ? mw[1][c] : dgt == 2 ? w[2][c] : mw[2][c];
      }
      for (int a = 0; a < 3; ++a) {
        for (int c = 0; c < 2; ++c) b[a][c] = nb[a][c];
        for (int c = 0; c < 3; ++c) w[a][c] = nw[a][c];
      }
    }
    double cen[3];
    for (int c = 0; c < 3; ++c) cen[c] = (w[0][c] + w[1][c] + w[2][c]) / 3;
    const double diam = fmax(fmax(dist3(w[0], w[1]), dist3(w[1], w[2])), dist3(w[2], w[0]));
    const double dist = dist3(x, cen);
    bool leaf = true;
    if (diam <= 0.4 * dist && fabs(k) * diam <= 0.7) {
      gradient_moments<double>(x, w, n, k, rule4, b, MfP, MfQ);
    } else if (depth < 12) {
      ++depth;  // first child: digit 0
      path &= ~(3u << (2 * (depth - 1)));
      leaf = false;
    } else {
      bad = true;
      break;
    }
    if (leaf) {
      // next sibling, climbing past exhausted levels
      while (depth > 0 && ((path >> (2 * (depth - 1))) & 3) == 3) --depth;
      if (depth == 0) break;
      path += 1u << (2 * (depth - 1));
    }
  }
  if (bad) {
    atomicAdd(failed, 1);
    return;
  }
  for (int d = 0; d < nrhs; ++d) {
    double pv[3][2];
    for (int a = 0; a < 3; ++a) {
      const int64_t vi = fverts[3 * f + a];
      pv[a][0] = (double)p[2 * (vi * nrhs + d)];
      pv[a][1] = (double)p[2 * (vi * nrhs + d) + 1];
    }
    const double qr = (double)q[2 * (f * nrhs + d)], qi = (double)q[2 * (f * nrhs + d) + 1];
    for (int c = 0; c < 4; ++c) {
      double cr = 0, ci = 0;
      for (int a = 0; a < 3; ++a) {
        const double mr = MfP[c][a][0] - (double)McP[c][a][0], mi = MfP[c][a][1] - (double)McP[c][a][1];
        cr += pv[a][0] * mr - pv[a][1] * mi;
        ci += pv[a][0] * mi + pv[a][1] * mr;
      }
      const double mr = MfQ[c][0] - (double)McQ[c][0], mi = MfQ[c][1] - (double)McQ[c][1];
      cr += qr * mr - qi * mi;
      ci += qr * mi + qi * mr;
      T* o = out + 2 * ((i * nrhs + d) * 4 + c);
      unsafeAtomicAdd(o, (T)cr);
      unsafeAtomicAdd(o + 1, (T)ci);
    }
  }
}

}  // namespace

template <typename T>
hipError_t field_sources(const double* rec, const int32_t* fverts, int64_t nf, int nq, const double* rule, const T* p,
                         const T* q, int nrhs, int nimg, const ImageXf* img, T* src_pts, T* src_nrm, T* pw, T* qw,
                         hipStream_t s) {
  if (nimg < 0 || nimg > MAX_IMAGES) return hipErrorInvalidValue;
  FieldImages im;
  for (int t = 0; t <= MAX_IMAGES; ++t)
    for (int c = 0; c < 3; ++c) im.s[t][c] = (t >= 1 && t <= nimg) ? img[t - 1].s[c] : 1;
  const int64_t ns = nf * nq * (1 + nimg);
  hipLaunchKernelGGL((field_sources_kernel<T>), dim3((unsigned)((ns + 255) / 256)), dim3(256), 0, s, rec, fverts, nf,
                     nq, rule, p, q, nrhs, 1 + nimg, im, src_pts, src_nrm, pw, qw);
  return hipGetLastError();
}

template <typename T>
hipError_t field_eval(const T* src_pts, const T* src_nrm, const T* pw, const T* qw, int64_t nsrc, int nrhs,
                      const T* pts, int64_t npts, T* out, T k, hipStream_t s) {
  if (npts <= 0) return hipSuccess;
  hipError_t e = launch_fill<T>(out, 2 * npts * nrhs, T(0), s);
  if (e != hipSuccess) return e;
  const int64_t pblocks = (npts + FIELD_PTS - 1) / FIELD_PTS;
  // enough source splits to give ~1024 workgroups
  int64_t splits = std::max<int64_t>(1, std::min<int64_t>((1024 + pblocks - 1) / pblocks, (nsrc + 1023) / 1024));
  const int64_t chunk = (nsrc + splits - 1) / splits;
  splits = (nsrc + chunk - 1) / chunk;
  for (int d0 = 0; d0 < nrhs; d0 += FIELD_RHS) {
    hipLaunchKernelGGL((field_kernel<T>), dim3((unsigned)pblocks, (unsigned)splits), dim3(FIELD_PTS), 0, s, src_pts,
                       src_nrm, pw, qw, nsrc, nrhs, d0, pts, npts, out, k, chunk);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  return hipSuccess;
}

template <typename T>
hipError_t field_gradient_eval(const T* src_pts, const T* src_nrm, const T* pw, const T* qw, int64_t nsrc, int nrhs,
                               const T* pts, int64_t npts, T* out, T k, hipStream_t s) {
  if (npts <= 0) return hipSuccess;
  hipError_t e = launch_fill<T>(out, 2 * 4 * npts * nrhs, T(0), s);
  if (e != hipSuccess) return e;
  const int64_t pblocks = (npts + FIELD_PTS - 1) / FIELD_PTS;
  int64_t splits = std::max<int64_t>(1, std::min<int64_t>((1024 + pblocks - 1) / pblocks, (nsrc + 1023) / 1024));
  const int64_t chunk = (nsrc + splits - 1) / splits;
  splits = (nsrc + chunk - 1) / chunk;
  for (int d0 = 0; d0 < nrhs; d0 += GRAD_RHS) {
    hipLaunchKernelGGL((field_gradient_kernel<T>), dim3((unsigned)pblocks, (unsigned)splits), dim3(FIELD_PTS), 0, s,
                       src_pts, src_nrm, pw, qw, nsrc, nrhs, d0, pts, npts, out, k, chunk);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  return hipSuccess;
}

static GradImages grad_images(int nimg, const ImageXf* img) {
  GradImages im;
  im.n = 1 + nimg;
  for (int t = 0; t <= MAX_IMAGES; ++t)
    for (int c = 0; c < 3; ++c) im.s[t][c] = (t >= 1 && t <= nimg) ? img[t - 1].s[c] : 1;
  return im;
}

template <typename T>
hipError_t gradient_near_pairs(const double* verts, const int32_t* fverts, int64_t nf, int nimg, const ImageXf* img,
                               const T* pts, int64_t npts, unsigned long long* count, int2* list,
                               unsigned long long cap, bool fill, hipStream_t s) {
  if (nimg < 0 || nimg > MAX_IMAGES) return hipErrorInvalidValue;
  const GradImages im = grad_images(nimg, img);
  const int64_t total = npts * nf * im.n;
  const unsigned blocks = (unsigned)std::max<int64_t>(1, std::min<int64_t>((total + 255) / 256, 65536));
  hipLaunchKernelGGL((near_pairs_kernel<T>), dim3(blocks), dim3(256), 0, s, verts, fverts, nf, im, pts, npts, count,
                     list, cap, fill);
  return hipGetLastError();
}

template <typename T>
hipError_t gradient_correction(const double* verts, const int32_t* fverts, const double* rec, int nimg,
                               const ImageXf* img, const int2* list, int64_t npairs, const T* pts, const T* p,
                               const T* q, int nrhs, double k, const double* rule4, T* out, int* failed,
                               hipStream_t s) {
  if (npairs <= 0) return hipSuccess;
  if (nimg < 0 || nimg > MAX_IMAGES) return hipErrorInvalidValue;
  const GradImages im = grad_images(nimg, img);
  hipLaunchKernelGGL((gradient_correction_kernel<T>), dim3((unsigned)((npairs + 63) / 64)), dim3(64), 0, s, verts,
                     fverts, rec, im, list, npairs, pts, p, q, nrhs, k, rule4, out, failed);
  return hipGetLastError();
}

#define BEAT_GRAD_INST(T)                                                                                           \
  template hipError_t field_gradient_eval<T>(const T*, const T*, const T*, const T*, int64_t, int, const T*, int64_t, \
                                             T*, T, hipStream_t);                                                   \
  template hipError_t gradient_near_pairs<T>(const double*, const int32_t*, int64_t, int, const ImageXf*, const T*,   \
                                             int64_t, unsigned long long*, int2*, unsigned long long, bool,          \
                                             hipStream_t);                                                          \
  template hipError_t gradient_correction<T>(const double*, const int32_t*, const double*, int, const ImageXf*,       \
                                             const int2*, int64_t, const T*, const T*, const T*, int, double,         \
                                             const double*, T*, int*, hipStream_t);
BEAT_GRAD_INST(float)
BEAT_GRAD_INST(double)
#undef BEAT_GRAD_INST

template hipError_t field_sources<float>(const double*, const int32_t*, int64_t, int, const double*, const float*,
                                         const float*, int, int, const ImageXf*, float*, float*, float*, float*,
                                         hipStream_t);
template hipError_t field_sources<double>(const double*, const int32_t*, int64_t, int, const double*, const double*,
                                          const double*, int, int, const ImageXf*, double*, double*, double*, double*,
                                          hipStream_t);
template hipError_t field_eval<float>(const float*, const float*, const float*, const float*, int64_t, int,
                                      const float*, int64_t, float*, float, hipStream_t);
template hipError_t field_eval<double>(const double*, const double*, const double*, const double*, int64_t, int,
                                       const double*, int64_t, double*, double, hipStream_t);

}  // namespace beat

// Adjoint shape gradient of a far-field SPL loss (FP64).
//
// For one frequency, with the Burton-Miller system A(x) p = b(x) (b = B(x) q, q
// fixed per face) and the field u = F(x) p + g(x) at the observation points,
// L = sum_d (SPL(u_d) - target)^2, SPL = 20 log10(|u| / 20e-6) (the reference's
// spl_for_points, BeatEngineDriver.jl:657-663). With w_d = dL/du_d* and
// A^T mu = F^T conj(w) (the forward LU reused, lu_solve_transposed):
//   dL = 2 Re( conj(w)^T (dF p + dg) - mu^T (dA p - db) ).
// Neither dA nor dF is formed. Psi = mu^T (A p - b) is a sum over element pairs of
//   sum_ij W_ij I_ij,   I = m (P lhs - q rhs) + eta C G
// (m, P: mu and p interpolated at the two points, C = (sum_a mu_a curl_a^t) .
// (sum_b p_b curl_b^s), lhs / rhs the fused Burton-Miller scalars of
// BeatEngineCpuBurtonMiller.jl:62-112), whose derivatives with respect to the
// points, normals, curls and areas are contracted here on the fly; the field part
// likewise. Seeds for the per-face quantities (regular points, area, normal, the
// two curl contractions) are chained to the vertices once per face.
#include <hip/hip_runtime.h>

#include "bem_common.h"
#include "kernels.h"

namespace beat {

namespace {

struct cplx {
  double r, i;
};
__device__ __forceinline__ cplx cmk(double r, double i) { return {r, i}; }
__device__ __forceinline__ cplx operator+(cplx a, cplx b) { return {a.r + b.r, a.i + b.i}; }
__device__ __forceinline__ cplx operator-(cplx a, cplx b) { return {a.r - b.r, a.i - b.i}; }
__device__ __forceinline__ cplx operator-(cplx a) { return {-a.r, -a.i}; }
__device__ __forceinline__ cplx operator*(cplx a, cplx b) { return {a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r}; }
__device__ __forceinline__ cplx operator*(double s, cplx a) { return {s * a.r, s * a.i}; }
__device__ __forceinline__ cplx ld(const double* p) { return {p[0], p[1]}; }

// per-face interpolants (FI_STRIDE doubles): m_i (mu at the regular points), P_i (p
// at the regular points), cmu = sum_a mu_a curl_a, cp = sum_a p_a curl_a, q
constexpr int FI_M = 0, FI_P = 12, FI_CMU = 24, FI_CP = 30, FI_Q = 36, FI_STRIDE = 38;
// per-face seeds (SD_STRIDE doubles): regular points (real parts, 6 x 3), area,
// normal (3), d/d(cmu) (3 complex), d/d(cp) (3 complex)
constexpr int SD_P = 0, SD_A = 18, SD_N = 19, SD_CMU = 22, SD_CP = 28, SD_STRIDE = 34;

__global__ void face_interp_kernel(const double* rec, const int32_t* fverts, int64_t nf, int nq, const double* rule,
                                   const double* mu, const double* p, const double* q, double* fi) {
  const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  const double* r = rec + f * FREC;
  int v[3];
  cplx mv[3], pv[3];
  for (int a = 0; a < 3; ++a) {
    v[a] = fverts[3 * f + a];
    mv[a] = ld(mu + 2 * v[a]);
    pv[a] = ld(p + 2 * v[a]);
  }
  double* o = fi + f * FI_STRIDE;
  for (int i = 0; i < 6; ++i) {
    cplx m = {0, 0}, pp = {0, 0};
    if (i < nq) {
      const double xi = rule[3 * i], eta = rule[3 * i + 1];
      const double ph[3] = {1.0 - xi - eta, xi, eta};
      for (int a = 0; a < 3; ++a) {
        m = m + ph[a] * mv[a];
        pp = pp + ph[a] * pv[a];
      }
    }
    o[FI_M + 2 * i] = m.r;
    o[FI_M + 2 * i + 1] = m.i;
    o[FI_P + 2 * i] = pp.r;
    o[FI_P + 2 * i + 1] = pp.i;
  }
  for (int c = 0; c < 3; ++c) {
    cplx cm = {0, 0}, cpp = {0, 0};
    for (int a = 0; a < 3; ++a) {
      const double cu = r[REC_CURL + 3 * a + c];
      cm = cm + cu * mv[a];
      cpp = cpp + cu * pv[a];
    }
    o[FI_CMU + 2 * c] = cm.r;
    o[FI_CMU + 2 * c + 1] = cm.i;
    o[FI_CP + 2 * c] = cpp.r;
    o[FI_CP + 2 * c + 1] = cpp.i;
  }
  o[FI_Q] = q[2 * f];
  o[FI_Q + 1] = q[2 * f + 1];
}

// Integrand and its derivatives for one point pair (x on the test face, y on the
// trial face): I, grad_r I (r = y - x), dI/dn_t, dI/dn_s, dI/dC.
struct PairTerms {
  cplx I, gr[3], gnt[3], gns[3], gC;
};

__device__ __forceinline__ void pair_terms(const double x[3], const double y[3], const double nt[3], const double ns[3],
                                           cplx m, cplx P, cplx q, cplx C, double k, cplx eta, PairTerms& o) {
  const double rv[3] = {y[0] - x[0], y[1] - x[1], y[2] - x[2]};
  const double rho = sqrt(rv[0] * rv[0] + rv[1] * rv[1] + rv[2] * rv[2]);
  const double ir = 1.0 / rho;
  const double u[3] = {rv[0] * ir, rv[1] * ir, rv[2] * ir};
  double s, c;
  sincos(k * rho, &s, &c);
  const cplx G = cmk(c * INV_FOUR_PI * ir, s * INV_FOUR_PI * ir);
  const cplx ikr = cmk(-ir, k);       // ik - 1/r
  const cplx G1 = G * ikr;            // dG/dr
  const cplx G2 = G * (ikr * ikr + cmk(ir * ir, 0.0));
  const cplx h = ir * G1;             // G' / r
  const cplx hp = ir * G2 - (ir * ir) * G1;
  const double a = rv[0] * ns[0] + rv[1] * ns[1] + rv[2] * ns[2];
  const double b = rv[0] * nt[0] + rv[1] * nt[1] + rv[2] * nt[2];
  const double cn = nt[0] * ns[0] + nt[1] * ns[1] + nt[2] * ns[2];
  const double k2 = k * k;
  const cplx mP = m * P, mq = m * q;
  // I = m P (-h a - eta k^2 c G) + m q G - m q eta h b + eta C G
  o.I = mP * (-(a * h) - (k2 * cn) * (eta * G)) + mq * G - (b * (mq * (eta * h))) + eta * (C * G);
  const cplx alpha = -(a * (mP * hp)) - (k2 * cn) * (mP * (eta * G1)) + mq * G1 - (b * (mq * (eta * hp))) +
                     eta * (C * G1);
  const cplx bs = -(mP * h), bt = -(mq * (eta * h));
  const cplx ekG = (k2) * (eta * G);
  for (int d = 0; d < 3; ++d) {
    o.gr[d] = u[d] * alpha + ns[d] * bs + nt[d] * bt;
    o.gns[d] = mP * (-(rv[d] * h) - nt[d] * ekG);
    o.gnt[d] = -(ns[d] * (mP * ekG)) - rv[d] * (mq * (eta * h));
  }
  o.gC = eta * G;
}

__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Regular pairs, one role per pass. TEST: one wavefront per test face, lanes over
// the trial faces; the seeds of the test face's points, normal, area and mu-curl.
// Otherwise the roles swap (trial face owned; the seeds of its points, normal, area
// and p-curl). Pairs sharing a vertex are left to the singular kernel.
template <bool TEST>
__global__ __launch_bounds__(64) void adj_regular_kernel(const double* rec, const int32_t* fverts, int64_t nf, int nq,
                                                         const double* rule, const double* fi, double k, double er,
                                                         double ei, double sgn, double* seeds) {
  const int64_t f = blockIdx.x;
  const int lane = threadIdx.x;
  const cplx eta = cmk(er, ei);
  const double* ro = rec + f * FREC;
  const double* fo = fi + f * FI_STRIDE;
  const int ov[3] = {fverts[3 * f], fverts[3 * f + 1], fverts[3 * f + 2]};
  const double no[3] = {ro[REC_N], ro[REC_N + 1], ro[REC_N + 2]};
  const double areaO = ro[REC_AREA];
  double gP[6][3], gN[3] = {0, 0, 0};
  cplx gC[3] = {{0, 0}, {0, 0}, {0, 0}}, sWI = {0, 0};
  for (int i = 0; i < 6; ++i)
    for (int d = 0; d < 3; ++d) gP[i][d] = 0;
  for (int64_t g = lane; g < nf; g += 64) {
    const int gv0 = fverts[3 * g], gv1 = fverts[3 * g + 1], gv2 = fverts[3 * g + 2];
    bool adj = false;
    for (int a = 0; a < 3; ++a) adj |= (ov[a] == gv0) | (ov[a] == gv1) | (ov[a] == gv2);
    if (adj) continue;
    const double* rg = rec + g * FREC;
    const double* fg = fi + g * FI_STRIDE;
    const double ng[3] = {rg[REC_N], rg[REC_N + 1], rg[REC_N + 2]};
    const double W0 = 4.0 * areaO * rg[REC_AREA];
    // roles: test face t, trial face s
    const double* rt = TEST ? ro : rg;
    const double* rs = TEST ? rg : ro;
    const double* ft = TEST ? fo : fg;
    const double* fs = TEST ? fg : fo;
    const double* nt = TEST ? no : ng;
    const double* ns = TEST ? ng : no;
    cplx C = {0, 0};
    for (int d = 0; d < 3; ++d) C = C + ld(ft + FI_CMU + 2 * d) * ld(fs + FI_CP + 2 * d);
    const cplx qs = ld(fs + FI_Q);
    cplx gCsum = {0, 0};
    // outer loop over the owned face's points (unrolled: constant seed indices),
    // inner loop over the other face's points
#pragma unroll
    for (int oi = 0; oi < 6; ++oi) {
      if (oi >= nq) break;
      double acc[3] = {0, 0, 0};
      for (int ot = 0; ot < nq; ++ot) {
        const int i = TEST ? oi : ot, j = TEST ? ot : oi;
        const double x[3] = {rt[3 * i], rt[3 * i + 1], rt[3 * i + 2]};
        const double y[3] = {rs[3 * j], rs[3 * j + 1], rs[3 * j + 2]};
        PairTerms pt;
        pair_terms(x, y, nt, ns, ld(ft + FI_M + 2 * i), ld(fs + FI_P + 2 * j), qs, C, k, eta, pt);
        const double W = W0 * rule[3 * i + 2] * rule[3 * j + 2];
        const double sd = TEST ? -W : W;  // d/dx = -grad_r, d/dy = +grad_r
        for (int d = 0; d < 3; ++d) {
          acc[d] += sd * pt.gr[d].r;
          gN[d] += W * (TEST ? pt.gnt[d].r : pt.gns[d].r);
        }
        gCsum = gCsum + W * pt.gC;
        sWI = sWI + W * pt.I;
      }
      for (int d = 0; d < 3; ++d) gP[oi][d] += acc[d];
    }
    // d Psi / d(own curl contraction) = sum W eta G times the other face's contraction
    for (int d = 0; d < 3; ++d) gC[d] = gC[d] + gCsum * ld((TEST ? fs + FI_CP : ft + FI_CMU) + 2 * d);
  }
  double* o = seeds + f * SD_STRIDE;
  for (int i = 0; i < 6; ++i)
    for (int d = 0; d < 3; ++d) {
      const double v = wave_sum(gP[i][d]);
      if (lane == 0) o[SD_P + 3 * i + d] += sgn * v;
    }
  for (int d = 0; d < 3; ++d) {
    const double vn = wave_sum(gN[d]);
    const double cr = wave_sum(gC[d].r), ci = wave_sum(gC[d].i);
    if (lane == 0) {
      o[SD_N + d] += sgn * vn;
      o[(TEST ? SD_CMU : SD_CP) + 2 * d] += sgn * cr;
      o[(TEST ? SD_CMU : SD_CP) + 2 * d + 1] += sgn * ci;
    }
  }
  const double sr = wave_sum(sWI.r);
  if (lane == 0) o[SD_A] += sgn * sr / areaO;
}

// Touching pairs over their Duffy rules, one wavefront per pair: point derivatives
// go straight to the vertices (the Duffy points are fixed combinations of them).
__global__ __launch_bounds__(256) void adj_singular_kernel(const double* rec, const double* verts,
                                                           const int32_t* fverts, const SingularPair* pairs,
                                                           int64_t npairs, const int32_t* rule_off,
                                                           const int32_t* rule_cnt, const double* rule_data,
                                                           const double* fi, const double* mu, const double* p,
                                                           double k, double er, double ei, double sgn, double* seeds,
                                                           double* gV) {
  const int lane = threadIdx.x & 63;
  const int64_t pidx = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (pidx >= npairs) return;
  const SingularPair P = pairs[pidx];
  if (P.xform != 0) return;  // image pairs: not differentiated
  const cplx eta = cmk(er, ei);
  const double* rt = rec + (int64_t)P.test * FREC;
  const double* rs = rec + (int64_t)P.trial * FREC;
  const double* ft = fi + (int64_t)P.test * FI_STRIDE;
  const double* fs = fi + (int64_t)P.trial * FI_STRIDE;
  double tv[3][3], sv[3][3];
  int tvi[3], svi[3];
  cplx mt[3], ps[3];
  for (int v = 0; v < 3; ++v) {
    tvi[v] = fverts[(int64_t)P.test * 3 + v];
    svi[v] = fverts[(int64_t)P.trial * 3 + v];
    mt[v] = ld(mu + 2 * tvi[v]);
    ps[v] = ld(p + 2 * svi[v]);
    for (int c = 0; c < 3; ++c) {
      tv[v][c] = verts[3 * (int64_t)tvi[v] + c];
      sv[v][c] = verts[3 * (int64_t)svi[v] + c];
    }
  }
  const double nt[3] = {rt[REC_N], rt[REC_N + 1], rt[REC_N + 2]};
  const double ns[3] = {rs[REC_N], rs[REC_N + 1], rs[REC_N + 2]};
  cplx C = {0, 0};
  for (int d = 0; d < 3; ++d) C = C + ld(ft + FI_CMU + 2 * d) * ld(fs + FI_CP + 2 * d);
  const cplx qs = ld(fs + FI_Q);
  const int off = rule_off[P.rule], cnt = rule_cnt[P.rule];
  double gT[3][3] = {}, gS[3][3] = {}, gNt[3] = {0, 0, 0}, gNs[3] = {0, 0, 0};
  cplx gCs = {0, 0}, sWI = {0, 0};
  for (int q = lane; q < cnt; q += 64) {
    const double* rd = rule_data + 5 * (int64_t)(off + q);
    const double pt_[3] = {1.0 - rd[0] - rd[1], rd[0], rd[1]};
    const double pss[3] = {1.0 - rd[2] - rd[3], rd[2], rd[3]};
    double x[3], y[3];
    cplx m = {0, 0}, Pp = {0, 0};
    for (int c = 0; c < 3; ++c) {
      x[c] = pt_[0] * tv[0][c] + pt_[1] * tv[1][c] + pt_[2] * tv[2][c];
      y[c] = pss[0] * sv[0][c] + pss[1] * sv[1][c] + pss[2] * sv[2][c];
    }
    for (int a = 0; a < 3; ++a) {
      m = m + pt_[a] * mt[a];
      Pp = Pp + pss[a] * ps[a];
    }
    const double r0 = y[0] - x[0], r1 = y[1] - x[1], r2 = y[2] - x[2];
    if (r0 * r0 + r1 * r1 + r2 * r2 == 0.0) continue;
    PairTerms pt;
    pair_terms(x, y, nt, ns, m, Pp, qs, C, k, eta, pt);
    const double W = rd[4] * P.jac;
    for (int d = 0; d < 3; ++d) {
      for (int a = 0; a < 3; ++a) {
        gT[a][d] -= W * pt_[a] * pt.gr[d].r;
        gS[a][d] += W * pss[a] * pt.gr[d].r;
      }
      gNt[d] += W * pt.gnt[d].r;
      gNs[d] += W * pt.gns[d].r;
    }
    gCs = gCs + W * pt.gC;
    sWI = sWI + W * pt.I;
  }
  for (int a = 0; a < 3; ++a)
    for (int d = 0; d < 3; ++d) {
      const double t = wave_sum(gT[a][d]), u = wave_sum(gS[a][d]);
      if (lane == 0) {
        atomicAdd(gV + 3 * (int64_t)tvi[a] + d, sgn * t);
        atomicAdd(gV + 3 * (int64_t)svi[a] + d, sgn * u);
      }
    }
  double* ot = seeds + (int64_t)P.test * SD_STRIDE;
  double* os = seeds + (int64_t)P.trial * SD_STRIDE;
  const cplx gc = {wave_sum(gCs.r), wave_sum(gCs.i)};
  for (int d = 0; d < 3; ++d) {
    const double a1 = wave_sum(gNt[d]), a2 = wave_sum(gNs[d]);
    if (lane == 0) {
      atomicAdd(ot + SD_N + d, sgn * a1);
      atomicAdd(os + SD_N + d, sgn * a2);
      const cplx cm = gc * ld(fs + FI_CP + 2 * d), cpp = gc * ld(ft + FI_CMU + 2 * d);
      atomicAdd(ot + SD_CMU + 2 * d, sgn * cm.r);
      atomicAdd(ot + SD_CMU + 2 * d + 1, sgn * cm.i);
      atomicAdd(os + SD_CP + 2 * d, sgn * cpp.r);
      atomicAdd(os + SD_CP + 2 * d + 1, sgn * cpp.i);
    }
  }
  const double swr = wave_sum(sWI.r);
  if (lane == 0) {
    atomicAdd(ot + SD_A, sgn * swr / rt[REC_AREA]);
    atomicAdd(os + SD_A, sgn * swr / rs[REC_AREA]);
  }
}

// Field part: one thread per face. For every source point i and observation point
// d, J = c_d W_i [h (r.n) P_i - G q] (r = y_i - x_d, c_d = conj(w_d)): the seeds of
// the face's points, normal and area; and the adjoint right-hand side
// (F^T c)_v += phi_a(i) W_i sum_d c_d h (r.n).
__global__ void adj_field_kernel(const double* rec, const int32_t* fverts, int64_t nf, int nq, const double* rule,
                                 const double* fi, const double* pts, const double* cw, int64_t npts, double k,
                                 double sgn, double* seeds, double* rhs) {
  const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  const double* r = rec + f * FREC;
  const double* fo = fi + f * FI_STRIDE;
  const double n[3] = {r[REC_N], r[REC_N + 1], r[REC_N + 2]};
  const double area = r[REC_AREA];
  const cplx q = ld(fo + FI_Q);
  double* o = seeds + f * SD_STRIDE;
  double gN[3] = {0, 0, 0}, sJ = 0;
  for (int i = 0; i < nq; ++i) {
    const double W = rule[3 * i + 2] * 2.0 * area;
    const double y[3] = {r[3 * i], r[3 * i + 1], r[3 * i + 2]};
    const cplx P = ld(fo + FI_P + 2 * i);
    double gp[3] = {0, 0, 0};
    cplx T = {0, 0};
    for (int64_t d = 0; d < npts; ++d) {
      const double rv[3] = {y[0] - pts[3 * d], y[1] - pts[3 * d + 1], y[2] - pts[3 * d + 2]};
      const double rho = sqrt(rv[0] * rv[0] + rv[1] * rv[1] + rv[2] * rv[2]);
      const double ir = 1.0 / rho;
      const double u[3] = {rv[0] * ir, rv[1] * ir, rv[2] * ir};
      double s, c;
      sincos(k * rho, &s, &c);
      const cplx G = cmk(c * INV_FOUR_PI * ir, s * INV_FOUR_PI * ir);
      const cplx ikr = cmk(-ir, k);
      const cplx G1 = G * ikr;
      const cplx G2 = G * (ikr * ikr + cmk(ir * ir, 0.0));
      const cplx h = ir * G1;
      const cplx hp = ir * G2 - (ir * ir) * G1;
      const double rn = rv[0] * n[0] + rv[1] * n[1] + rv[2] * n[2];
      const cplx cd = ld(cw + 2 * d);
      const cplx J = cd * (rn * (h * P) - G * q);
      const cplx al = cd * (rn * (hp * P) - G1 * q);  // along u
      const cplx be = cd * (h * P);                   // along n, and r for d/dn
      for (int e = 0; e < 3; ++e) {
        gp[e] += W * (u[e] * al.r + n[e] * be.r);
        gN[e] += W * rv[e] * be.r;
      }
      sJ += W * J.r;
      T = T + (W * rn) * (cd * h);
    }
    for (int e = 0; e < 3; ++e) o[SD_P + 3 * i + e] += sgn * gp[e];
    const double xi = rule[3 * i], eta = rule[3 * i + 1];
    const double ph[3] = {1.0 - xi - eta, xi, eta};
    for (int a = 0; a < 3; ++a) {
      const int64_t v = fverts[3 * f + a];
      atomicAdd(rhs + 2 * v, ph[a] * T.r);
      atomicAdd(rhs + 2 * v + 1, ph[a] * T.i);
    }
  }
  for (int e = 0; e < 3; ++e) o[SD_N + e] += sgn * gN[e];
  o[SD_A] += sgn * sJ / area;
}

// Forward-mode dual number of one direction.
struct dual {
  double v, d;
};
__device__ __forceinline__ dual operator+(dual a, dual b) { return {a.v + b.v, a.d + b.d}; }
__device__ __forceinline__ dual operator-(dual a, dual b) { return {a.v - b.v, a.d - b.d}; }
__device__ __forceinline__ dual operator-(dual a) { return {-a.v, -a.d}; }
__device__ __forceinline__ dual operator*(dual a, dual b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
__device__ __forceinline__ dual operator/(dual a, dual b) { return {a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v)}; }
__device__ __forceinline__ dual dsqrt(dual a) {
  const double s = sqrt(a.v);
  return {s, a.d / (2.0 * s)};
}

// Face geometry as the mesh setup forms it (area, unit normal, surface curls),
// in dual numbers.
__device__ void face_geometry(const dual v[3][3], dual& area, dual n[3], dual curl[3][3]) {
  dual e1[3], e2[3];
  for (int c = 0; c < 3; ++c) {
    e1[c] = v[1][c] - v[0][c];
    e2[c] = v[2][c] - v[0][c];
  }
  const dual cr[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
  const dual nn = dsqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
  area = nn * dual{0.5, 0.0};
  for (int c = 0; c < 3; ++c) n[c] = cr[c] / nn;
  const dual g11 = e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2];
  const dual g12 = e1[0] * e2[0] + e1[1] * e2[1] + e1[2] * e2[2];
  const dual g22 = e2[0] * e2[0] + e2[1] * e2[1] + e2[2] * e2[2];
  const dual det = g11 * g22 - g12 * g12;
  const dual i11 = g22 / det, i12 = -g12 / det, i22 = g11 / det;
  dual gr[3][3];
  for (int c = 0; c < 3; ++c) {
    const dual l1 = e1[c] * i11 + e2[c] * i12, l2 = e1[c] * i12 + e2[c] * i22;
    gr[0][c] = -l1 - l2;
    gr[1][c] = l1;
    gr[2][c] = l2;
  }
  for (int a = 0; a < 3; ++a) {
    curl[a][0] = n[1] * gr[a][2] - n[2] * gr[a][1];
    curl[a][1] = n[2] * gr[a][0] - n[0] * gr[a][2];
    curl[a][2] = n[0] * gr[a][1] - n[1] * gr[a][0];
  }
}

// Per face: the identity blocks' area seed, then every seed chained to the three
// vertices (points through the P1 basis, area / normal / curls by forward-mode
// derivatives in each of the nine vertex coordinates).
__global__ void adj_chain_kernel(const double* verts, const int32_t* fverts, int64_t nf, int nq, const double* rule,
                                 const double* seeds, const double* mu, const double* p, const double* q, double er,
                                 double ei, double id_sgn, const IdentityBlocks idb, double* gV) {
  const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  const double* s = seeds + f * SD_STRIDE;
  int vi[3];
  cplx mv[3], pv[3];
  for (int a = 0; a < 3; ++a) {
    vi[a] = fverts[3 * f + a];
    mv[a] = ld(mu + 2 * vi[a]);
    pv[a] = ld(p + 2 * vi[a]);
  }
  // identity: Psi += area (1/2 sum mu_a pp_ab p_b + 1/2 eta sum mu_a pd_a q)
  cplx idv = {0, 0};
  const cplx eta = cmk(er, ei), qf = ld(q + 2 * f);
  for (int a = 0; a < 3; ++a) {
    for (int b = 0; b < 3; ++b) idv = idv + (0.5 * idb.pp[a][b]) * (mv[a] * pv[b]);
    idv = idv + (0.5 * idb.pd[a]) * (mv[a] * (eta * qf));
  }
  const double gA = s[SD_A] + id_sgn * idv.r;
  double g[3][3] = {};
  for (int i = 0; i < nq; ++i) {
    const double xi = rule[3 * i], et = rule[3 * i + 1];
    const double ph[3] = {1.0 - xi - et, xi, et};
    for (int a = 0; a < 3; ++a)
      for (int c = 0; c < 3; ++c) g[a][c] += ph[a] * s[SD_P + 3 * i + c];
  }
  // curl seeds: d Psi / d curl_a = mu_a d/d(cmu) + p_a d/d(cp)
  cplx gcurl[3][3];
  for (int a = 0; a < 3; ++a)
    for (int c = 0; c < 3; ++c) gcurl[a][c] = mv[a] * ld(s + SD_CMU + 2 * c) + pv[a] * ld(s + SD_CP + 2 * c);
#pragma unroll
  for (int a0 = 0; a0 < 3; ++a0)
#pragma unroll
    for (int c0 = 0; c0 < 3; ++c0) {
      dual v[3][3];
      for (int a = 0; a < 3; ++a)
        for (int c = 0; c < 3; ++c) v[a][c] = {verts[3 * (int64_t)vi[a] + c], (a == a0 && c == c0) ? 1.0 : 0.0};
      dual area, n[3], curl[3][3];
      face_geometry(v, area, n, curl);
      double d = gA * area.d;
      for (int c = 0; c < 3; ++c) d += s[SD_N + c] * n[c].d;
      for (int a = 0; a < 3; ++a)
        for (int c = 0; c < 3; ++c) d += gcurl[a][c].r * curl[a][c].d;
      g[a0][c0] += d;
    }
  for (int a = 0; a < 3; ++a)
    for (int c = 0; c < 3; ++c) atomicAdd(gV + 3 * (int64_t)vi[a] + c, g[a][c]);
}

}  // namespace

hipError_t adjoint_face_interp(const double* rec, const int32_t* fverts, int64_t nf, int nq, const double* rule,
                               const double* mu, const double* p, const double* q, double* fi, hipStream_t s) {
  hipLaunchKernelGGL(face_interp_kernel, dim3((unsigned)((nf + 255) / 256)), dim3(256), 0, s, rec, fverts, nf, nq, rule,
                     mu, p, q, fi);
  return hipGetLastError();
}

hipError_t adjoint_field(const double* rec, const int32_t* fverts, int64_t nf, int nq, const double* rule,
                         const double* fi, const double* pts, const double* cw, int64_t npts, double k, double* seeds,
                         double* rhs, hipStream_t s) {
  hipLaunchKernelGGL(adj_field_kernel, dim3((unsigned)((nf + 63) / 64)), dim3(64), 0, s, rec, fverts, nf, nq, rule, fi,
                     pts, cw, npts, k, 1.0, seeds, rhs);
  return hipGetLastError();
}

hipError_t adjoint_pairs(const double* rec, const double* verts, const int32_t* fverts, int64_t nf, int nq,
                         const double* rule, const SingularPair* pairs, int64_t npairs, const int32_t* rule_off,
                         const int32_t* rule_cnt, const double* rule_data, const double* fi, const double* mu,
                         const double* p, double k, double er, double ei, double* seeds, double* gV, hipStream_t s) {
  // Psi enters the loss derivative with a minus sign
  hipLaunchKernelGGL((adj_regular_kernel<true>), dim3((unsigned)nf), dim3(64), 0, s, rec, fverts, nf, nq, rule, fi, k, er,
                     ei, -1.0, seeds);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((adj_regular_kernel<false>), dim3((unsigned)nf), dim3(64), 0, s, rec, fverts, nf, nq, rule, fi, k,
                     er, ei, -1.0, seeds);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if (npairs > 0) {
    hipLaunchKernelGGL(adj_singular_kernel, dim3((unsigned)((npairs + 3) / 4)), dim3(256), 0, s, rec, verts, fverts,
                       pairs, npairs, rule_off, rule_cnt, rule_data, fi, mu, p, k, er, ei, -1.0, seeds, gV);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t adjoint_chain(const double* verts, const int32_t* fverts, int64_t nf, int nq, const double* rule,
                         const double* seeds, const double* mu, const double* p, const double* q, double er, double ei,
                         const IdentityBlocks& idb, double* gV, hipStream_t s) {
  hipLaunchKernelGGL(adj_chain_kernel, dim3((unsigned)((nf + 127) / 128)), dim3(128), 0, s, verts, fverts, nf, nq, rule,
                     seeds, mu, p, q, er, ei, -1.0, idb, gV);
  return hipGetLastError();
}

}  // namespace beat

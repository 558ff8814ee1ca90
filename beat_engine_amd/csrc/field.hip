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
    for (int d = 0; d < nd; ++d) {
      unsafeAtomicAdd(out + 2 * (pi * nrhs + d0 + d), ar[d]);
      unsafeAtomicAdd(out + 2 * (pi * nrhs + d0 + d) + 1, ai[d]);
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

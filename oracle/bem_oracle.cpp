// CPU ORACLE — TEST INFRASTRUCTURE ONLY.
//
// A C++17 restatement of the reference's CPU Burton-Miller path, used as the
// checker for the HIP library and as the CPU baseline that bench.py times.
// Only tests/, the smoke check and bench.py's cpu_baseline leg may load it;
// the product library never links or calls it.
//
// Every function names the reference file:line it restates (paths relative to
// src/beat_engine/julia_local/). The arithmetic follows the reference's order
// of operations; summation order across threads follows its colour groups.
//
// Parity pins (tests/test_oracle_*.py):
//   rules          exactness on monomials, closed-form weight sums
//   duffy rules    brute-force double integrals over touching pairs
//   kernels        finite differences of G, closed-form mass blocks
//   fused system   equals the four-operator system (reference invariant)
//   whole path     analytic pulsating sphere and the reference's golden
//                  per-frequency norms (results/baseline_sweep_cpu.json)
#include <algorithm>
#include <array>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <map>
#include <set>
#include <tuple>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

template <class T> struct V3 {
  T x, y, z;
};
template <class T> inline V3<T> operator-(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> inline V3<T> operator+(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> inline V3<T> operator*(T s, V3<T> a) { return {s * a.x, s * a.y, s * a.z}; }
template <class T> inline T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T> inline V3<T> cross(V3<T> a, V3<T> b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class T> inline T norm(V3<T> a) { return std::sqrt(dot(a, a)); }

template <class T> struct P2 {
  T x, y;
};

// BoundaryMesh constructor: JL/src/BeatEngineCore.jl:344-364.
template <class T> struct Mesh {
  std::vector<V3<T>> verts;
  std::vector<std::array<int, 3>> faces;
  std::vector<V3<T>> normals, centroids;
  std::vector<T> areas;
};

template <class T> Mesh<T> make_mesh(const double* xyz, int64_t nv, const int32_t* f, int64_t nf) {
  Mesh<T> m;
  m.verts.resize(nv);
  for (int64_t i = 0; i < nv; ++i) m.verts[i] = {T(xyz[3 * i]), T(xyz[3 * i + 1]), T(xyz[3 * i + 2])};
  m.faces.resize(nf);
  m.normals.resize(nf);
  m.centroids.resize(nf);
  m.areas.resize(nf);
  for (int64_t e = 0; e < nf; ++e) {
    m.faces[e] = {f[3 * e], f[3 * e + 1], f[3 * e + 2]};
    V3<T> v1 = m.verts[f[3 * e]], v2 = m.verts[f[3 * e + 1]], v3 = m.verts[f[3 * e + 2]];
    V3<T> c = cross(v2 - v1, v3 - v1);
    T n = norm(c);
    m.centroids[e] = {(v1.x + v2.x + v3.x) / T(3), (v1.y + v2.y + v3.y) / T(3), (v1.z + v2.z + v3.z) / T(3)};
    m.areas[e] = n / T(2);
    m.normals[e] = {c.x / n, c.y / n, c.z / n};
  }
  return m;
}

// triangle_rule: JL/src/BeatEngineCore.jl:660-692 (literals kept verbatim).
template <class T> void triangle_rule(int order, std::vector<P2<T>>& p, std::vector<T>& w) {
  p.clear();
  w.clear();
  if (order <= 1) {
    p.push_back({T(1) / T(3), T(1) / T(3)});
    w.push_back(T(0.5));
  } else if (order == 4) {
    p = {{T(0.4459484909159651), T(0.4459484909159651)}, {T(0.0915762135097710), T(0.0915762135097700)},
         {T(0.1081030181680700), T(0.4459484909159651)}, {T(0.4459484909159651), T(0.1081030181680700)},
         {T(0.8168475729804590), T(0.0915762135097700)}, {T(0.0915762135097710), T(0.8168475729804580)}};
    const T raw[6] = {T(0.2233815896780110), T(0.1099517436553220), T(0.2233815896780110),
                      T(0.2233815896780110), T(0.1099517436553220), T(0.1099517436553220)};
    for (int i = 0; i < 6; ++i) w.push_back(T(0.5) * raw[i]);
  } else {
    p = {{T(1) / T(6), T(1) / T(6)}, {T(2) / T(3), T(1) / T(6)}, {T(1) / T(6), T(2) / T(3)}};
    w = {T(1) / T(6), T(1) / T(6), T(1) / T(6)};
  }
}

// gauss_rule_1d on [0,1]: JL/src/BeatEngineCore.jl:713-740 (orders 1-4).
template <class T> void gauss_rule(int order, std::vector<T>& x, std::vector<T>& w) {
  if (order == 1) {
    x = {T(0.5)};
    w = {T(1.0)};
  } else if (order == 2) {
    T a = T(0.5) / std::sqrt(T(3.0));
    x = {T(0.5) - a, T(0.5) + a};
    w = {T(0.5), T(0.5)};
  } else if (order == 3) {
    T a = std::sqrt(T(3.0) / T(5.0)) / T(2.0);
    x = {T(0.5) - a, T(0.5), T(0.5) + a};
    w = {T(5.0) / T(18.0), T(4.0) / T(9.0), T(5.0) / T(18.0)};
  } else {
    T s65 = std::sqrt(T(6.0) / T(5.0));
    T x1 = std::sqrt(T(3.0) / T(7.0) - T(2.0) / T(7.0) * s65) / T(2.0);
    T x2 = std::sqrt(T(3.0) / T(7.0) + T(2.0) / T(7.0) * s65) / T(2.0);
    T w1 = (T(18.0) + std::sqrt(T(30.0))) / T(72.0);
    T w2 = (T(18.0) - std::sqrt(T(30.0))) / T(72.0);
    x = {T(0.5) - x2, T(0.5) - x1, T(0.5) + x1, T(0.5) + x2};
    w = {w2, w1, w1, w2};
  }
}

template <class T> struct Duffy {
  std::vector<P2<T>> tp, sp;
  std::vector<T> w;
  void add(T tx, T ty, T sx, T sy, T wt) {  // append_duffy_point!: JL/src/BeatEngineCore.jl:798-802
    tp.push_back({tx - ty, ty});
    sp.push_back({sx - sy, sy});
    w.push_back(wt);
  }
};

// duffy_rule: JL/src/BeatEngineCore.jl:742-796. kind 1 coincident, 2 edge, 3 vertex.
template <class T> Duffy<T> duffy_rule(int order, int kind) {
  std::vector<T> xr, wr;
  gauss_rule(order, xr, wr);
  std::vector<P2<T>> tpts;
  std::vector<T> tw;
  for (size_t i = 0; i < xr.size(); ++i)
    for (size_t j = 0; j < xr.size(); ++j) {
      tpts.push_back({xr[j], xr[i]});
      tw.push_back(wr[i] * wr[j]);
    }
  Duffy<T> d;
  const T one = T(1.0);
  for (size_t a = 0; a < tpts.size(); ++a)
    for (size_t b = 0; b < tpts.size(); ++b) {
      T xsi = tpts[a].x, e1 = tpts[a].y, e2 = tpts[b].x, e3 = tpts[b].y;
      T e12 = e1 * e2, e123 = e12 * e3;
      T bw = tw[a] * tw[b];
      if (kind == 1) {
        T wt = bw * (xsi * xsi * xsi) * (e1 * e1) * e2;
        d.add(xsi, xsi * (one - e1 + e12), xsi * (one - e123), xsi * (one - e1), wt);
        d.add(xsi * (one - e123), xsi * (one - e1), xsi, xsi * (one - e1 + e12), wt);
        d.add(xsi, xsi * (e1 - e12 + e123), xsi * (one - e12), xsi * (e1 - e12), wt);
        d.add(xsi * (one - e12), xsi * (e1 - e12), xsi, xsi * (e1 - e12 + e123), wt);
        d.add(xsi * (one - e123), xsi * (e1 - e123), xsi, xsi * (e1 - e12), wt);
        d.add(xsi, xsi * (e1 - e12), xsi * (one - e123), xsi * (e1 - e123), wt);
      } else if (kind == 2) {
        T wt = bw * (xsi * xsi * xsi) * (e1 * e1);
        d.add(xsi, xsi * e1 * e3, xsi * (one - e12), xsi * e1 * (one - e2), wt);
        d.add(xsi, xsi * e1, xsi * (one - e123), xsi * e1 * e2 * (one - e3), wt * e2);
        d.add(xsi * (one - e12), xsi * e1 * (one - e2), xsi, xsi * e123, wt * e2);
        d.add(xsi * (one - e123), xsi * e12 * (one - e3), xsi, xsi * e1, wt * e2);
        d.add(xsi * (one - e123), xsi * e1 * (one - e2 * e3), xsi, xsi * e12, wt * e2);
      } else {
        T wt = bw * (xsi * xsi * xsi) * e2;
        d.add(xsi, xsi * e1, xsi * e2, xsi * e2 * e3, wt);
        d.add(xsi * e2, xsi * e2 * e3, xsi, xsi * e1, wt);
      }
    }
  return d;
}

// remap_shared_vertex / remap_shared_edge: JL/src/BeatEngineCore.jl:804-826 (1-based vertex ids).
template <class T> P2<T> remap_vertex(P2<T> p, int v) {
  if (v == 1) return p;
  if (v == 2) return {T(1.0) - p.x - p.y, p.y};
  return {p.x, T(1.0) - p.x - p.y};
}
template <class T> P2<T> remap_edge(P2<T> p, int s1, int s2) {
  const P2<T> ref[3] = {{T(0), T(0)}, {T(1), T(0)}, {T(0), T(1)}};
  int rem = 6 - s1 - s2;
  P2<T> v0 = ref[s1 - 1], v1 = ref[s2 - 1], v2 = ref[rem - 1];
  return {v0.x + p.x * (v1.x - v0.x) + p.y * (v2.x - v0.x), v0.y + p.x * (v1.y - v0.y) + p.y * (v2.y - v0.y)};
}

struct AdjInfo {
  int kind;  // 0 regular, 1 coincident, 2 edge, 3 vertex
  int ta, tb, sa, sb;  // 1-based local vertex ids
};

// adjacency_info: JL/src/BeatEngineCore.jl:598-622.
inline AdjInfo adjacency_info(const std::array<int, 3>& a, const std::array<int, 3>& b) {
  int sa[9], sb[9], n = 0;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      if (a[i] == b[j]) {
        sa[n] = i + 1;
        sb[n] = j + 1;
        ++n;
      }
  if (n >= 3) return {1, 1, 2, 1, 2};
  if (n == 2) {
    if (sb[1] < sb[0]) {
      std::swap(sa[0], sa[1]);
      std::swap(sb[0], sb[1]);
    }
    return {2, sa[0], sa[1], sb[0], sb[1]};
  }
  if (n == 1) return {3, sa[0], 0, sb[0], 0};
  return {0, 0, 0, 0, 0};
}

inline bool adjacent(const std::array<int, 3>& a, const std::array<int, 3>& b) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      if (a[i] == b[j]) return true;
  return false;
}

template <class T> struct SingularPair {
  int test, trial, rule;
  T jac, nprod;
};

// build_singular_correction_cache: JL/src/BeatEngineCore.jl:958-985,994-1089.
template <class T> struct SingularCache {
  std::vector<std::vector<SingularPair<T>>> by_test;
  std::vector<Duffy<T>> rules;
  std::vector<std::array<int, 5>> keys;
};

// singular_orientation_key / rule_for_singular_orientation! / remapped_duffy_rule:
// JL/src/BeatEngineCore.jl:994-1047 (rules deduplicated by orientation key).
template <class T>
int rule_for_orientation(SingularCache<T>& c, std::map<std::array<int, 5>, int>& rule_index, const Duffy<T> base[3],
                         const AdjInfo& info) {
  std::array<int, 5> key;
  if (info.kind == 1) key = {1, 0, 0, 0, 0};
  else if (info.kind == 2) key = {2, info.ta, info.tb, info.sa, info.sb};
  else key = {3, info.ta, 0, info.sa, 0};
  auto it = rule_index.find(key);
  if (it != rule_index.end()) return it->second;
  const Duffy<T>& b = base[info.kind - 1];
  Duffy<T> r;
  r.w = b.w;
  for (size_t q = 0; q < b.w.size(); ++q) {
    P2<T> a = b.tp[q], s2 = b.sp[q];
    if (info.kind == 2) {
      a = remap_edge(a, key[1], key[2]);
      s2 = remap_edge(s2, key[3], key[4]);
    } else if (info.kind == 3) {
      a = remap_vertex(a, key[1]);
      s2 = remap_vertex(s2, key[3]);
    }
    r.tp.push_back(a);
    r.sp.push_back(s2);
  }
  c.rules.push_back(r);
  c.keys.push_back(key);
  const int ri = (int)c.rules.size() - 1;
  rule_index[key] = ri;
  return ri;
}

template <class T> SingularCache<T> build_singular_cache(const Mesh<T>& m, int order) {
  const int nf = (int)m.faces.size();
  std::map<int, std::vector<int>> v2e;
  for (int e = 0; e < nf; ++e)
    for (int v : m.faces[e]) v2e[v].push_back(e);
  Duffy<T> base[3] = {duffy_rule<T>(order, 1), duffy_rule<T>(order, 2), duffy_rule<T>(order, 3)};
  SingularCache<T> c;
  c.by_test.resize(nf);
  std::map<std::array<int, 5>, int> rule_index;
  for (int t = 0; t < nf; ++t) {
    std::vector<int> cand;
    std::set<int> seen;
    for (int v : m.faces[t])
      for (int s : v2e[v])
        if (!seen.count(s)) {
          cand.push_back(s);
          seen.insert(s);
        }
    for (int s : cand) {
      AdjInfo info = adjacency_info(m.faces[t], m.faces[s]);
      if (info.kind == 0) continue;
      const int ri = rule_for_orientation(c, rule_index, base, info);
      T jac = (T(2.0) * m.areas[t]) * (T(2.0) * m.areas[s]);
      c.by_test[t].push_back({t, s, ri, jac, dot(m.normals[t], m.normals[s])});
    }
  }
  return c;
}

// ---- symmetry images ------------------------------------------------------------
// normalized_symmetry_mode / symmetry_transforms / reflect_point / reflect_normal /
// reflect_curl: JL/src/BeatEngineCore.jl:448-495. Modes: 0 off, 1 x, 2 xy, 3 ground.
struct Xf {
  int s[3];
  int det;
};
inline std::vector<Xf> image_transforms(int mode) {
  std::vector<Xf> t;
  if (mode == 1) {
    t.push_back({{-1, 1, 1}, -1});
  } else if (mode == 2) {
    t.push_back({{-1, 1, 1}, -1});
    t.push_back({{1, -1, 1}, -1});
    t.push_back({{-1, -1, 1}, 1});
  } else if (mode == 3) {
    t.push_back({{1, -1, 1}, -1});  // rigid_ground_transform
  }
  return t;
}
template <class T> V3<T> reflect(const Xf& x, V3<T> p) { return {T(x.s[0]) * p.x, T(x.s[1]) * p.y, T(x.s[2]) * p.z}; }
template <class T> V3<T> reflect_curl(const Xf& x, V3<T> c) {
  return {T(x.det * x.s[0]) * c.x, T(x.det * x.s[1]) * c.y, T(x.det * x.s[2]) * c.z};
}

// symmetry_plane_tolerance / p1_symmetry_orbit_weights: JL/src/BeatEngineCore.jl:497-517,563-580.
template <class T> std::vector<T> orbit_weights(const Mesh<T>& m, int mode) {
  std::vector<T> w(m.verts.size(), T(1));
  if ((mode != 1 && mode != 2) || m.verts.empty()) return w;
  V3<T> lo = m.verts[0], hi = m.verts[0];
  for (auto& v : m.verts) {
    lo = {std::min(lo.x, v.x), std::min(lo.y, v.y), std::min(lo.z, v.z)};
    hi = {std::max(hi.x, v.x), std::max(hi.y, v.y), std::max(hi.z, v.z)};
  }
  const T tol = std::max(T(1e-9), norm(hi - lo) * T(1e-6));
  for (size_t i = 0; i < m.verts.size(); ++i) {
    T wt = T(1);
    if (std::abs(m.verts[i].x) <= tol) wt *= T(2);
    if (mode == 2 && std::abs(m.verts[i].y) <= tol) wt *= T(2);
    w[i] = wt;
  }
  return w;
}

// geometry_key / geometric_adjacency_info: JL/src/BeatEngineCore.jl:624-658.
template <class T> std::array<long long, 3> geometry_key(V3<T> p, T tol) {
  return {(long long)std::nearbyint(p.x / tol), (long long)std::nearbyint(p.y / tol),
          (long long)std::nearbyint(p.z / tol)};
}
template <class T> AdjInfo geometric_adjacency_info(const V3<T> a[3], const V3<T> b[3], T tol) {
  int sa[9], sb[9], n = 0;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      if (norm(a[i] - b[j]) <= tol) {
        sa[n] = i + 1;
        sb[n] = j + 1;
        ++n;
      }
  if (n == 3) return {1, sa[0], sa[1], sb[0], sb[1]};
  if (n == 2) {
    if (sb[1] < sb[0]) {
      std::swap(sa[0], sa[1]);
      std::swap(sb[0], sb[1]);
    }
    return {2, sa[0], sa[1], sb[0], sb[1]};
  }
  if (n == 1) return {3, sa[0], 0, sb[0], 0};
  return {0, 0, 0, 0, 0};
}

// image_singular_candidates + _beat_cpu_image_singular_cache: JL/src/BeatEngineCore.jl:1139-1166,
// JL/src/BeatEngineCpuAssembly.jl:488-523 (test face touching a reflected trial face).
template <class T> SingularCache<T> build_image_singular_cache(const Mesh<T>& m, int order, const Xf& x) {
  const T tol = T(1e-8);
  const int nf = (int)m.faces.size();
  std::map<std::array<long long, 3>, std::vector<int>> by_key;
  for (int t = 0; t < nf; ++t)
    for (int v : m.faces[t]) by_key[geometry_key(m.verts[v], tol)].push_back(t);
  std::vector<std::pair<int, int>> cand;
  std::set<std::pair<int, int>> seen;
  for (int s = 0; s < nf; ++s)
    for (int v : m.faces[s]) {
      auto it = by_key.find(geometry_key(reflect(x, m.verts[v]), tol));
      if (it == by_key.end()) continue;
      for (int t : it->second)
        if (seen.insert({t, s}).second) cand.push_back({t, s});
    }
  Duffy<T> base[3] = {duffy_rule<T>(order, 1), duffy_rule<T>(order, 2), duffy_rule<T>(order, 3)};
  SingularCache<T> c;
  c.by_test.resize(nf);
  std::map<std::array<int, 5>, int> rule_index;
  for (auto& ts : cand) {
    const int t = ts.first, s = ts.second;
    V3<T> tv[3], sv[3];
    for (int i = 0; i < 3; ++i) {
      tv[i] = m.verts[m.faces[t][i]];
      sv[i] = reflect(x, m.verts[m.faces[s][i]]);
    }
    AdjInfo info = geometric_adjacency_info(tv, sv, tol);
    if (info.kind == 0) continue;
    const int ri = rule_for_orientation(c, rule_index, base, info);
    T jac = (T(2.0) * m.areas[t]) * (T(2.0) * m.areas[s]);
    c.by_test[t].push_back({t, s, ri, jac, dot(m.normals[t], reflect(x, m.normals[s]))});
  }
  return c;
}

// surface_gradients / surface_curls: JL/src/BeatEngineCore.jl:856-873.
template <class T> std::array<V3<T>, 3> surface_curls(const V3<T> v[3], V3<T> n) {
  V3<T> e1 = v[1] - v[0], e2 = v[2] - v[0];
  T g11 = dot(e1, e1), g12 = dot(e2, e1), g22 = dot(e2, e2);
  T det = g11 * g22 - g12 * g12;
  // inv(gram) for the symmetric 2x2 Gram matrix
  T i11 = g22 / det, i12 = -g12 / det, i22 = g11 / det;
  // lift = jac * inv(gram): columns l1 = e1*i11 + e2*i12, l2 = e1*i12 + e2*i22
  V3<T> l1 = {e1.x * i11 + e2.x * i12, e1.y * i11 + e2.y * i12, e1.z * i11 + e2.z * i12};
  V3<T> l2 = {e1.x * i12 + e2.x * i22, e1.y * i12 + e2.y * i22, e1.z * i12 + e2.z * i22};
  V3<T> g0 = {-l1.x - l2.x, -l1.y - l2.y, -l1.z - l2.z};
  return {cross(n, g0), cross(n, l1), cross(n, l2)};
}

// BeatCpuElementData + BeatCpuRegularQuadratureData: JL/src/BeatEngineCpuAssembly.jl:8-21,45-90.
template <class T> struct Elem {
  V3<T> v[3];
  V3<T> n;
  std::array<V3<T>, 3> curl;
  std::array<int, 3> p1;
  T area;
  std::vector<V3<T>> qp;                 // regular quadrature points
  std::vector<std::array<T, 3>> qb;      // p1 basis at those points
};

template <class T> V3<T> local_to_global(const V3<T> v[3], P2<T> p) {  // JL/src/BeatEngineCore.jl:849-852
  T a = T(1) - p.x - p.y;
  return {a * v[0].x + p.x * v[1].x + p.y * v[2].x, a * v[0].y + p.x * v[1].y + p.y * v[2].y,
          a * v[0].z + p.x * v[1].z + p.y * v[2].z};
}
template <class T> std::array<T, 3> p1_values(P2<T> p) { return {T(1) - p.x - p.y, p.x, p.y}; }

template <class T>
std::vector<Elem<T>> element_data(const Mesh<T>& m, const std::vector<P2<T>>& rp) {
  std::vector<Elem<T>> el(m.faces.size());
  for (size_t e = 0; e < m.faces.size(); ++e) {
    Elem<T>& E = el[e];
    for (int i = 0; i < 3; ++i) E.v[i] = m.verts[m.faces[e][i]];
    E.n = m.normals[e];
    E.curl = surface_curls(E.v, E.n);
    E.p1 = m.faces[e];
    E.area = m.areas[e];
    for (auto p : rp) {
      E.qp.push_back(local_to_global(E.v, p));
      E.qb.push_back(p1_values(p));
    }
  }
  return el;
}

// _beat_cpu_reflect_element_data / _beat_cpu_reflect_regular_quadrature_data:
// JL/src/BeatEngineCpuAssembly.jl:61-100 (same dofs, reflected geometry).
template <class T> Elem<T> reflect_elem(const Elem<T>& e, const Xf& x) {
  Elem<T> r = e;
  for (int i = 0; i < 3; ++i) r.v[i] = reflect(x, e.v[i]);
  r.n = reflect(x, e.n);
  for (int a = 0; a < 3; ++a) r.curl[a] = reflect_curl(x, e.curl[a]);
  for (auto& p : r.qp) p = reflect(x, p);
  return r;
}

// _beat_cpu_element_color_groups: JL/src/BeatEngineCpuAssembly.jl:102-130.
template <class T> std::vector<std::vector<int>> color_groups(const Mesh<T>& m) {
  std::map<int, std::set<int>> vcol;
  std::vector<std::vector<int>> groups;
  for (int e = 0; e < (int)m.faces.size(); ++e) {
    std::set<int> used;
    for (int v : m.faces[e]) {
      auto it = vcol.find(v);
      if (it != vcol.end()) used.insert(it->second.begin(), it->second.end());
    }
    int c = 1;
    while (used.count(c)) ++c;
    while ((int)groups.size() < c) groups.emplace_back();
    groups[c - 1].push_back(e);
    for (int v : m.faces[e]) vcol[v].insert(c);
  }
  return groups;
}

template <class T> inline void green(T radius, T inv_radius, T k, T& gr, T& gi) {  // _beat_cpu_green: CpuAssembly.jl:1-6
  T phase = k * radius;
  T scale = (T(1) / (T(4.0) * T(M_PI))) * inv_radius;
  T s = std::sin(phase), c = std::cos(phase);
  gr = c * scale;
  gi = s * scale;
}

typedef std::complex<double> cd;

// One Burton-Miller pair over arbitrary point lists: the algebra of
// _beat_cpu_bm_regular_pair_blocks (JL/src/BeatEngineCpuBurtonMiller.jl:62-112) and
// _beat_cpu_bm_pair_blocks (:116-167). xs/ys are the test/trial points, tb/sb the
// p1 basis at them, w the full weight per point pair (already times jac).
template <class T> struct BMBlock {
  std::complex<T> lhs[3][3];
  std::complex<T> rhs[3];
};

template <class T>
void bm_points(const Elem<T>& te, const Elem<T>& se, T k, std::complex<T> eta, T nprod, int nq,
               const V3<T>* xs, const std::array<T, 3>* tb, const V3<T>* ys, const std::array<T, 3>* sb,
               const T* w, BMBlock<T>& out) {
  typedef std::complex<T> C;
  for (int a = 0; a < 3; ++a) {
    out.rhs[a] = C(0);
    for (int b = 0; b < 3; ++b) out.lhs[a][b] = C(0);
  }
  C curl_total(0);
  C k2n = eta * C(k * k * nprod);
  for (int q = 0; q < nq; ++q) {
    V3<T> r = ys[q] - xs[q];
    T radius = norm(r);
    if (radius == T(0)) continue;
    T ir = T(1) / radius;
    T gr, gi;
    green(radius, ir, k, gr, gi);
    C G(gr, gi);
    C grad = G * C(-ir, k);
    T weight = w[q];
    C wg = G * weight;
    C lhs_s = -grad * (dot(r, se.n) * ir) * weight - k2n * wg;
    C rhs_s = -wg - eta * grad * (-dot(r, te.n) * ir) * weight;
    curl_total += wg;
    for (int a = 0; a < 3; ++a) {
      T tv = tb[q][a];
      out.rhs[a] += tv * rhs_s;
      for (int b = 0; b < 3; ++b) out.lhs[a][b] += (tv * sb[q][b]) * lhs_s;
    }
  }
  curl_total *= eta;
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) out.lhs[a][b] += dot(te.curl[a], se.curl[b]) * curl_total;
}

// Regular pair over the tensor rule, test point outer / trial point inner as in
// JL/src/BeatEngineCpuBurtonMiller.jl:77-104.
template <class T>
void bm_regular_pair(const Elem<T>& te, const Elem<T>& se, const std::vector<T>& rw, T k, std::complex<T> eta,
                     BMBlock<T>& out) {
  const int n = (int)rw.size();
  V3<T> xs[36], ys[36];
  std::array<T, 3> tb[36], sb[36];
  T w[36];
  T jac = T(4.0) * te.area * se.area;
  int q = 0;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      xs[q] = te.qp[i];
      tb[q] = te.qb[i];
      ys[q] = se.qp[j];
      sb[q] = se.qb[j];
      w[q] = rw[i] * rw[j] * jac;
      ++q;
    }
  bm_points(te, se, k, eta, dot(te.n, se.n), q, xs, tb, ys, sb, w, out);
}

template <class T>
void bm_singular_pair(const Elem<T>& te, const Elem<T>& se, const SingularPair<T>& p, const Duffy<T>& d, T k,
                      std::complex<T> eta, BMBlock<T>& out) {
  const int nq = (int)d.w.size();
  std::vector<V3<T>> xs(nq), ys(nq);
  std::vector<std::array<T, 3>> tb(nq), sb(nq);
  std::vector<T> w(nq);
  for (int q = 0; q < nq; ++q) {
    tb[q] = p1_values(d.tp[q]);
    sb[q] = p1_values(d.sp[q]);
    xs[q] = local_to_global(te.v, d.tp[q]);
    ys[q] = local_to_global(se.v, d.sp[q]);
    w[q] = d.w[q] * p.jac;
  }
  bm_points(te, se, k, eta, p.nprod, nq, xs.data(), tb.data(), ys.data(), sb.data(), w.data(), out);
}

// l2_identity_element_matrix at unit area: JL/src/BeatEngineCore.jl:905-923.
template <class T> void identity_blocks(const std::vector<P2<T>>& rp, const std::vector<T>& rw, T area, T pp[3][3], T pd[3]) {
  T jac = T(2.0) * area;
  for (int a = 0; a < 3; ++a) {
    pd[a] = 0;
    for (int b = 0; b < 3; ++b) pp[a][b] = 0;
  }
  for (size_t q = 0; q < rw.size(); ++q) {
    auto v = p1_values(rp[q]);
    for (int a = 0; a < 3; ++a) {
      pd[a] += v[a] * T(1.0) * rw[q] * jac;
      for (int b = 0; b < 3; ++b) pp[a][b] += v[a] * v[b] * rw[q] * jac;
    }
  }
}

int g_threads = 0;
void set_threads() {
#ifdef _OPENMP
  if (g_threads > 0) omp_set_num_threads(g_threads);
#endif
}

// assemble_burton_miller_neumann_system_cpu: JL/src/BeatEngineCpuBurtonMiller.jl:276-469.
// A is N x N column-major, B is N x nrhs column-major, q is F x nrhs column-major
// (all complex interleaved). test_lo/test_hi restrict the regular passes to a
// range of test elements (the bounded CPU-baseline sample); the full system uses
// [0, F). sym: 0 off, 1 x, 2 xy, 3 ground.
template <class T>
void bm_assemble(const double* xyz, int64_t nv, const int32_t* f, int64_t nf, int quad_order, int sing_order,
                 double k_signed, double eta_re, double eta_im, const std::complex<T>* q, int64_t nrhs,
                 std::complex<T>* A, std::complex<T>* B, int test_lo, int test_hi, int do_singular,
                 int do_identity, int sym) {
  typedef std::complex<T> C;
  set_threads();
  Mesh<T> m = make_mesh<T>(xyz, nv, f, nf);
  std::vector<P2<T>> rp;
  std::vector<T> rw;
  triangle_rule<T>(quad_order, rp, rw);
  auto el = element_data(m, rp);
  auto groups = color_groups(m);
  T k = T(k_signed);
  C eta{T(eta_re), T(eta_im)};
  const int64_t N = nv;
  std::memset((void*)A, 0, sizeof(C) * N * N);
  std::memset((void*)B, 0, sizeof(C) * N * nrhs);
  auto scatter = [&](const Elem<T>& te, const Elem<T>& se, int sdof, const BMBlock<T>& b, T sign = T(1)) {
    // _beat_cpu_bm_scatter!: JL/src/BeatEngineCpuBurtonMiller.jl:26-51 (sign -1: the subtracting form)
    for (int a = 0; a < 3; ++a) {
      int row = te.p1[a];
      for (int64_t d = 0; d < nrhs; ++d) B[row + d * N] += sign * b.rhs[a] * q[sdof + d * nf];
      for (int c = 0; c < 3; ++c) A[row + (int64_t)se.p1[c] * N] += sign * b.lhs[a][c];
    }
  };
  const std::vector<Xf> xfs = image_transforms(sym);
  std::vector<std::vector<Elem<T>>> iel(xfs.size());
  for (size_t x = 0; x < xfs.size(); ++x)
    for (auto& e : el) iel[x].push_back(reflect_elem(e, xfs[x]));
  for (auto& g : groups) {
#pragma omp parallel for schedule(dynamic, 1)
    for (int gi = 0; gi < (int)g.size(); ++gi) {
      int t = g[gi];
      if (t < test_lo || t >= test_hi) continue;
      BMBlock<T> b;
      for (int s = 0; s < nf; ++s) {
        if (adjacent(el[t].p1, el[s].p1)) continue;
        bm_regular_pair(el[t], el[s], rw, k, eta, b);
        scatter(el[t], el[s], s, b);
      }
    }
  }
  // image passes: every pair, the trial element reflected (_beat_cpu_bm_regular_image_test!, :190-210)
  for (size_t x = 0; x < xfs.size(); ++x)
    for (auto& g : groups) {
#pragma omp parallel for schedule(dynamic, 1)
      for (int gi = 0; gi < (int)g.size(); ++gi) {
        int t = g[gi];
        if (t < test_lo || t >= test_hi) continue;
        BMBlock<T> b;
        for (int s = 0; s < nf; ++s) {
          bm_regular_pair(el[t], iel[x][s], rw, k, eta, b);
          scatter(el[t], iel[x][s], s, b);
        }
      }
    }
  if (do_singular) {
    SingularCache<T> sc = build_singular_cache(m, sing_order);
    for (auto& g : groups) {
#pragma omp parallel for schedule(dynamic, 1)
      for (int gi = 0; gi < (int)g.size(); ++gi) {
        int t = g[gi];
        BMBlock<T> b;
        for (auto& p : sc.by_test[t]) {
          bm_singular_pair(el[p.test], el[p.trial], p, sc.rules[p.rule], k, eta, b);
          scatter(el[p.test], el[p.trial], p.trial, b);
        }
      }
    }
    // image-singular deltas: Duffy value minus the regular value the image pass added
    // (_beat_cpu_bm_image_singular_delta_test!, :232-265)
    for (size_t x = 0; x < xfs.size(); ++x) {
      SingularCache<T> ic = build_image_singular_cache(m, sing_order, xfs[x]);
      for (auto& g : groups) {
#pragma omp parallel for schedule(dynamic, 1)
        for (int gi = 0; gi < (int)g.size(); ++gi) {
          int t = g[gi];
          BMBlock<T> b;
          for (auto& p : ic.by_test[t]) {
            bm_singular_pair(el[p.test], iel[x][p.trial], p, ic.rules[p.rule], k, eta, b);
            scatter(el[p.test], iel[x][p.trial], p.trial, b);
            bm_regular_pair(el[p.test], iel[x][p.trial], rw, k, eta, b);
            scatter(el[p.test], iel[x][p.trial], p.trial, b, T(-1));
          }
        }
      }
    }
  }
  // orbit row weights scale the operator part (:443-449); the identity blocks carry them too
  const std::vector<T> ow = orbit_weights(m, sym);
  if (sym != 0)
    for (int64_t r = 0; r < N; ++r) {
      for (int64_t c = 0; c < N; ++c) A[r + c * N] *= ow[r];
      for (int64_t d = 0; d < nrhs; ++d) B[r + d * N] *= ow[r];
    }
  if (do_identity) {
    // lhs += 0.5 I_p1p1; rhs += (I_p1dp0 q) * (-0.5 eta): BeatEngineCpuBurtonMiller.jl:450-453
    for (int64_t e = 0; e < nf; ++e) {
      T pp[3][3], pd[3];
      identity_blocks(rp, rw, m.areas[e], pp, pd);
      for (int a = 0; a < 3; ++a) {
        int row = m.faces[e][a];
        for (int c = 0; c < 3; ++c) A[row + (int64_t)m.faces[e][c] * N] += C(T(0.5)) * C(pp[a][c] * ow[row]);
        for (int64_t d = 0; d < nrhs; ++d)
          B[row + d * N] += (C(pd[a] * ow[row]) * q[e + d * nf]) * (-C(T(0.5)) * eta);
      }
    }
  }
}

// assemble_regular_galerkin_operators_cpu with the singular pass
// (JL/src/BeatEngineCpuAssembly.jl:151-218,278-373,739-1031): S, K' are N x F,
// D, H are N x N, column-major complex.
template <class T>
void operators(const double* xyz, int64_t nv, const int32_t* f, int64_t nf, int quad_order, int sing_order,
               double k_signed, std::complex<T>* S, std::complex<T>* D, std::complex<T>* K, std::complex<T>* H,
               int sym) {
  typedef std::complex<T> C;
  set_threads();
  Mesh<T> m = make_mesh<T>(xyz, nv, f, nf);
  std::vector<P2<T>> rp;
  std::vector<T> rw;
  triangle_rule<T>(quad_order, rp, rw);
  auto el = element_data(m, rp);
  auto groups = color_groups(m);
  T k = T(k_signed), k2 = k * k;
  const int64_t N = nv, F = nf;
  std::memset((void*)S, 0, sizeof(C) * N * F);
  std::memset((void*)K, 0, sizeof(C) * N * F);
  std::memset((void*)D, 0, sizeof(C) * N * N);
  std::memset((void*)H, 0, sizeof(C) * N * N);
  // _beat_cpu_accumulate_pair! over a point list
  auto pair = [&](const Elem<T>& te, const Elem<T>& se, int sdof, T nprod, int nq, const V3<T>* xs,
                  const std::array<T, 3>* tb, const V3<T>* ys, const std::array<T, 3>* sb, const T* w,
                  T sign = T(1)) {
    C sl[3] = {}, ad[3] = {}, dl[3][3] = {}, hy[3][3] = {};
    T cp[3][3];
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) cp[a][b] = dot(te.curl[a], se.curl[b]);
    for (int qq = 0; qq < nq; ++qq) {
      V3<T> r = ys[qq] - xs[qq];
      T radius = norm(r);
      if (radius == T(0)) continue;
      T ir = T(1) / radius, gr, gi;
      green(radius, ir, k, gr, gi);
      C G(gr, gi);
      C grad = G * C(-ir, k);
      T tdot = dot(r, se.n) * ir;
      T sdot = -dot(r, te.n) * ir;
      C dv = grad * tdot, av = grad * sdot;
      T weight = w[qq];
      C wg = G * weight, wd = dv * weight, wa = av * weight;
      for (int a = 0; a < 3; ++a) {
        T tv = tb[qq][a];
        sl[a] += tv * wg;
        ad[a] += tv * wa;
        for (int b = 0; b < 3; ++b) {
          T bp = tv * sb[qq][b];
          dl[a][b] += bp * wd;
          hy[a][b] += (cp[a][b] - k2 * bp * nprod) * wg;
        }
      }
    }
    for (int a = 0; a < 3; ++a) {
      int row = te.p1[a];
      S[row + sdof * N] += sign * sl[a];
      K[row + sdof * N] += sign * ad[a];
      for (int b = 0; b < 3; ++b) {
        D[row + (int64_t)se.p1[b] * N] += sign * dl[a][b];
        H[row + (int64_t)se.p1[b] * N] += sign * hy[a][b];
      }
    }
  };
  const int n = (int)rw.size();
  // regular pair of te with se over the tensor rule
  auto regular = [&](const Elem<T>& te, const Elem<T>& se, int sdof, T sign) {
    V3<T> xs[36], ys[36];
    std::array<T, 3> tb[36], sb[36];
    T w[36];
    T jac = T(4.0) * te.area * se.area;
    int qq = 0;
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        xs[qq] = te.qp[i];
        tb[qq] = te.qb[i];
        ys[qq] = se.qp[j];
        sb[qq] = se.qb[j];
        w[qq] = rw[i] * rw[j] * jac;
        ++qq;
      }
    pair(te, se, sdof, dot(te.n, se.n), qq, xs, tb, ys, sb, w, sign);
  };
  // Duffy pair
  auto duffy_pair = [&](const Elem<T>& te, const Elem<T>& se, const SingularPair<T>& p, const Duffy<T>& d) {
    const int nq = (int)d.w.size();
    std::vector<V3<T>> xs(nq), ys(nq);
    std::vector<std::array<T, 3>> tb(nq), sb(nq);
    std::vector<T> w(nq);
    for (int qq = 0; qq < nq; ++qq) {
      tb[qq] = p1_values(d.tp[qq]);
      sb[qq] = p1_values(d.sp[qq]);
      xs[qq] = local_to_global(te.v, d.tp[qq]);
      ys[qq] = local_to_global(se.v, d.sp[qq]);
      w[qq] = d.w[qq] * p.jac;
    }
    pair(te, se, p.trial, p.nprod, nq, xs.data(), tb.data(), ys.data(), sb.data(), w.data());
  };
  const std::vector<Xf> xfs = image_transforms(sym);
  std::vector<std::vector<Elem<T>>> iel(xfs.size());
  for (size_t x = 0; x < xfs.size(); ++x)
    for (auto& e : el) iel[x].push_back(reflect_elem(e, xfs[x]));
  for (auto& g : groups) {
#pragma omp parallel for schedule(dynamic, 1)
    for (int gi = 0; gi < (int)g.size(); ++gi) {
      int t = g[gi];
      for (int s = 0; s < (int)F; ++s) {
        if (adjacent(el[t].p1, el[s].p1)) continue;
        regular(el[t], el[s], s, T(1));
      }
    }
  }
  SingularCache<T> sc = build_singular_cache(m, sing_order);
  for (auto& g : groups) {
#pragma omp parallel for schedule(dynamic, 1)
    for (int gi = 0; gi < (int)g.size(); ++gi) {
      int t = g[gi];
      for (auto& p : sc.by_test[t]) duffy_pair(el[p.test], el[p.trial], p, sc.rules[p.rule]);
    }
  }
  // image passes and image-singular deltas (JL/src/BeatEngineCpuAssembly.jl:571-640,827-960)
  for (size_t x = 0; x < xfs.size(); ++x) {
    for (auto& g : groups) {
#pragma omp parallel for schedule(dynamic, 1)
      for (int gi = 0; gi < (int)g.size(); ++gi) {
        int t = g[gi];
        for (int s = 0; s < (int)F; ++s) regular(el[t], iel[x][s], s, T(1));
      }
    }
    SingularCache<T> ic = build_image_singular_cache(m, sing_order, xfs[x]);
    for (auto& g : groups) {
#pragma omp parallel for schedule(dynamic, 1)
      for (int gi = 0; gi < (int)g.size(); ++gi) {
        int t = g[gi];
        for (auto& p : ic.by_test[t]) {
          duffy_pair(el[p.test], iel[x][p.trial], p, ic.rules[p.rule]);
          regular(el[p.test], iel[x][p.trial], p.trial, T(-1));
        }
      }
    }
  }
  // _beat_cpu_apply_operator_p1_row_weights!: JL/src/BeatEngineCpuAssembly.jl:674-680
  if (sym != 0) {
    const std::vector<T> ow = orbit_weights(m, sym);
    for (int64_t r = 0; r < N; ++r) {
      for (int64_t c = 0; c < F; ++c) {
        S[r + c * N] *= ow[r];
        K[r + c * N] *= ow[r];
      }
      for (int64_t c = 0; c < N; ++c) {
        D[r + c * N] *= ow[r];
        H[r + c * N] *= ow[r];
      }
    }
  }
}

// evaluate_galerkin_field_cpu (scalar kernel): JL/src/BeatEngineCpuField.jl:1-49 with
// the source cache of build_field_evaluation_cache: JL/src/BeatEngineCore.jl:1770-1810.
template <class T>
void field(const double* xyz, int64_t nv, const int32_t* f, int64_t nf, int quad_order, double k_signed,
           const std::complex<T>* p, const std::complex<T>* q, int64_t nrhs, const double* pts, int64_t np,
           std::complex<T>* out, int sym) {
  typedef std::complex<T> C;
  set_threads();
  Mesh<T> m = make_mesh<T>(xyz, nv, f, nf);
  std::vector<P2<T>> rp;
  std::vector<T> rw;
  triangle_rule<T>(quad_order, rp, rw);
  const int nq = (int)rw.size();
  // sources per element for the identity, then each image (symmetry_transforms order)
  std::vector<Xf> xfs = {{{1, 1, 1}, 1}};
  for (auto& x : image_transforms(sym)) xfs.push_back(x);
  const int nt = (int)xfs.size();
  const int64_t ns = nf * nt * nq;
  std::vector<V3<T>> sy(ns), sn(ns);
  std::vector<T> sw(ns);
  std::vector<std::array<T, 3>> sb(ns);
  for (int64_t e = 0; e < nf; ++e) {
    T jac = T(2.0) * m.areas[e];
    for (int x = 0; x < nt; ++x) {
      V3<T> v[3] = {reflect(xfs[x], m.verts[m.faces[e][0]]), reflect(xfs[x], m.verts[m.faces[e][1]]),
                    reflect(xfs[x], m.verts[m.faces[e][2]])};
      for (int qq = 0; qq < nq; ++qq) {
        const int64_t s = (e * nt + x) * nq + qq;
        sy[s] = local_to_global(v, rp[qq]);
        sn[s] = reflect(xfs[x], m.normals[e]);
        sw[s] = rw[qq] * jac;
        sb[s] = p1_values(rp[qq]);
      }
    }
  }
  T k = T(k_signed);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < np; ++i) {
    V3<T> x = {T(pts[3 * i]), T(pts[3 * i + 1]), T(pts[3 * i + 2])};
    for (int64_t d = 0; d < nrhs; ++d) {
      C pot(0);
      for (int64_t s = 0; s < ns; ++s) {
        V3<T> r = sy[s] - x;
        T radius = norm(r);
        if (radius == T(0)) continue;
        T phase = k * radius;
        T gs = T(1) / (T(4.0) * T(M_PI) * radius);
        C G(std::cos(phase) * gs, std::sin(phase) * gs);
        T proj = dot(r, sn[s]) / radius;
        C dl = G * (C(T(0), k) - T(1) / radius) * proj;
        int64_t e = s / (nq * nt);
        const std::array<int, 3>& fc = m.faces[e];
        C ps = (sb[s][0] * p[fc[0] + d * nv] + sb[s][1] * p[fc[1] + d * nv] + sb[s][2] * p[fc[2] + d * nv]) * sw[s];
        C qs = q[e + d * nf] * sw[s];
        pot += dl * ps - G * qs;
      }
      out[i + d * np] = pot;
    }
  }
}

// ---- field value and observation gradient -----------------------------------------
// field_value_gradient: JL/src/BeatEngineFieldGradient.jl:2-13. Value and x-gradient
// of the integrand of D p - S q at source y with normal n (r = y - x, signed k).
template <class T> void field_value_gradient(V3<T> x, V3<T> y, V3<T> n, std::complex<T> p, std::complex<T> q, T k,
                                             std::complex<T> out[4]) {
  typedef std::complex<T> C;
  V3<T> r = y - x;
  T radius = norm(r);
  V3<T> u = {r.x / radius, r.y / radius, r.z / radius};
  C green = std::exp(C(T(0), k * radius)) / (T(4.0 * M_PI) * radius);
  C a(-T(1) / radius, k);
  C b(T(3) / (radius * radius) - k * k, -T(3) * k / radius);
  T proj = dot(u, n);
  out[0] = green * (a * proj * p - q);
  const C aq = a * q, bp = b * proj, ar = a / radius;
  out[1] = green * (aq * u.x - p * (bp * u.x + ar * n.x));
  out[2] = green * (aq * u.y - p * (bp * u.y + ar * n.y));
  out[3] = green * (aq * u.z - p * (bp * u.z + ar * n.z));
}

// _gradient_triangle: JL/src/BeatEngineFieldGradient.jl:15-24 (jac = |cross|).
template <class T> void gradient_triangle(V3<T> x, const V3<T> v[3], V3<T> n, const std::complex<T> p[3],
                                          std::complex<T> q, T k, const std::vector<P2<T>>& rp,
                                          const std::vector<T>& rw, std::complex<T> acc[4]) {
  const T jac = norm(cross(v[1] - v[0], v[2] - v[0]));
  for (size_t j = 0; j < rw.size(); ++j) {
    const T b0 = T(1) - rp[j].x - rp[j].y, b1 = rp[j].x, b2 = rp[j].y;
    const V3<T> y = local_to_global(v, rp[j]);
    std::complex<T> f[4];
    field_value_gradient(x, y, n, b0 * p[0] + b1 * p[1] + b2 * p[2], q, k, f);
    for (int d = 0; d < 4; ++d) acc[d] += f[d] * (jac * rw[j]);
  }
}

// _adaptive_gradient_triangle: JL/src/BeatEngineFieldGradient.jl:26-43. Returns false
// where depth 12 is reached without meeting the size/distance criterion.
inline bool adaptive_gradient_triangle(V3<double> x, const V3<double> v[3], V3<double> n,
                                       const std::complex<double> p[3], std::complex<double> q, double k,
                                       const std::vector<P2<double>>& rp, const std::vector<double>& rw, int depth,
                                       std::complex<double> acc[4]) {
  const V3<double> c = {(v[0].x + v[1].x + v[2].x) / 3, (v[0].y + v[1].y + v[2].y) / 3, (v[0].z + v[1].z + v[2].z) / 3};
  const double diameter = std::max({norm(v[0] - v[1]), norm(v[1] - v[2]), norm(v[2] - v[0])});
  const double distance = norm(x - c);
  if (diameter <= 0.4 * distance && std::abs(k) * diameter <= 0.7) {
    gradient_triangle<double>(x, v, n, p, q, k, rp, rw, acc);
    return true;
  }
  if (depth >= 12) return false;
  auto mid = [](V3<double> a, V3<double> b) { return V3<double>{(a.x + b.x) / 2, (a.y + b.y) / 2, (a.z + b.z) / 2}; };
  const V3<double> ab = mid(v[0], v[1]), bc = mid(v[1], v[2]), ca = mid(v[2], v[0]);
  const std::complex<double> pab = (p[0] + p[1]) / 2.0, pbc = (p[1] + p[2]) / 2.0, pca = (p[2] + p[0]) / 2.0;
  const V3<double> t0[3] = {v[0], ab, ca}, t1[3] = {ab, v[1], bc}, t2[3] = {ca, bc, v[2]}, t3[3] = {ab, bc, ca};
  const std::complex<double> p0[3] = {p[0], pab, pca}, p1[3] = {pab, p[1], pbc}, p2[3] = {pca, pbc, p[2]},
                             p3[3] = {pab, pbc, pca};
  std::complex<double> s[4][4] = {};
  bool ok = adaptive_gradient_triangle(x, t0, n, p0, q, k, rp, rw, depth + 1, s[0]) &&
            adaptive_gradient_triangle(x, t1, n, p1, q, k, rp, rw, depth + 1, s[1]) &&
            adaptive_gradient_triangle(x, t2, n, p2, q, k, rp, rw, depth + 1, s[2]) &&
            adaptive_gradient_triangle(x, t3, n, p3, q, k, rp, rw, depth + 1, s[3]);
  for (int d = 0; d < 4; ++d) acc[d] += s[0][d] + s[1][d] + s[2][d] + s[3][d];
  return ok;
}

// evaluate_galerkin_field_gradient_cpu with correct_field_gradient!:
// JL/src/BeatEngineFieldGradient.jl:45-92. Values (np x 4 x nrhs, column-major:
// pressure, d/dx, d/dy, d/dz) in precision T from the cache sources (rule of
// quad_order, images of the symmetry mode), then for every face (and image)
// whose centroid lies within 3 diameters of the point the FP64 adaptive integral
// minus the order-4 rule in T. Returns the number of (point, face) pairs that hit
// the depth limit (the reference raises an error for them).
template <class T>
int64_t field_gradient(const double* xyz, int64_t nv, const int32_t* f, int64_t nf, int quad_order, double k_signed,
                       const std::complex<T>* p, const std::complex<T>* q, int64_t nrhs, const double* pts, int64_t np,
                       std::complex<T>* out, int sym) {
  typedef std::complex<T> C;
  set_threads();
  Mesh<T> m = make_mesh<T>(xyz, nv, f, nf);
  std::vector<P2<T>> rp, rp4;
  std::vector<T> rw, rw4;
  triangle_rule<T>(quad_order, rp, rw);
  triangle_rule<T>(4, rp4, rw4);
  std::vector<P2<double>> fp;
  std::vector<double> fw;
  triangle_rule<double>(4, fp, fw);
  const int nq = (int)rw.size();
  std::vector<Xf> xfs = {{{1, 1, 1}, 1}};
  for (auto& x : image_transforms(sym)) xfs.push_back(x);
  const int nt = (int)xfs.size();
  const int64_t ns = nf * nt * nq;
  std::vector<V3<T>> sy(ns), sn(ns);
  std::vector<T> sw(ns);
  std::vector<std::array<T, 3>> sb(ns);
  for (int64_t e = 0; e < nf; ++e) {
    T jac = T(2.0) * m.areas[e];
    for (int x = 0; x < nt; ++x) {
      V3<T> v[3] = {reflect(xfs[x], m.verts[m.faces[e][0]]), reflect(xfs[x], m.verts[m.faces[e][1]]),
                    reflect(xfs[x], m.verts[m.faces[e][2]])};
      for (int qq = 0; qq < nq; ++qq) {
        const int64_t s = (e * nt + x) * nq + qq;
        sy[s] = local_to_global(v, rp[qq]);
        sn[s] = reflect(xfs[x], m.normals[e]);
        sw[s] = rw[qq] * jac;
        sb[s] = p1_values(rp[qq]);
      }
    }
  }
  const T k = T(k_signed);
  int64_t failed = 0;
#pragma omp parallel for schedule(dynamic, 4) reduction(+ : failed)
  for (int64_t i = 0; i < np; ++i) {
    const V3<T> x = {T(pts[3 * i]), T(pts[3 * i + 1]), T(pts[3 * i + 2])};
    const V3<double> xd = {double(x.x), double(x.y), double(x.z)};
    for (int64_t d = 0; d < nrhs; ++d) {
      C acc[4] = {};
      for (int64_t s = 0; s < ns; ++s) {
        const int64_t e = s / (nq * nt);
        const std::array<int, 3>& fc = m.faces[e];
        const C ps = sb[s][0] * p[fc[0] + d * nv] + sb[s][1] * p[fc[1] + d * nv] + sb[s][2] * p[fc[2] + d * nv];
        C fv[4];
        field_value_gradient(x, sy[s], sn[s], ps, q[e + d * nf], k, fv);
        for (int c = 0; c < 4; ++c) acc[c] += fv[c] * sw[s];
      }
      std::complex<double> corr[4] = {};
      for (int64_t e = 0; e < nf; ++e)
        for (int t = 0; t < nt; ++t) {
          V3<T> v[3];
          V3<double> vd[3];
          for (int a = 0; a < 3; ++a) {
            v[a] = reflect(xfs[t], m.verts[m.faces[e][a]]);
            vd[a] = {double(v[a].x), double(v[a].y), double(v[a].z)};
          }
          const V3<T> c = {(v[0].x + v[1].x + v[2].x) / T(3), (v[0].y + v[1].y + v[2].y) / T(3),
                           (v[0].z + v[1].z + v[2].z) / T(3)};
          const T diameter = std::max({norm(v[0] - v[1]), norm(v[1] - v[2]), norm(v[2] - v[0])});
          if (!(norm(x - c) <= T(3) * diameter)) continue;
          const V3<T> n = reflect(xfs[t], m.normals[e]);
          const V3<double> nd = {double(n.x), double(n.y), double(n.z)};
          const C pv[3] = {p[m.faces[e][0] + d * nv], p[m.faces[e][1] + d * nv], p[m.faces[e][2] + d * nv]};
          C coarse[4] = {};
          gradient_triangle<T>(x, v, n, pv, q[e + d * nf], k, rp4, rw4, coarse);
          const std::complex<double> pd[3] = {std::complex<double>(pv[0]), std::complex<double>(pv[1]),
                                              std::complex<double>(pv[2])};
          std::complex<double> fine[4] = {};
          if (!adaptive_gradient_triangle(xd, vd, nd, pd, std::complex<double>(q[e + d * nf]), double(k), fp, fw, 0,
                                          fine))
            ++failed;
          for (int c2 = 0; c2 < 4; ++c2) corr[c2] += fine[c2] - std::complex<double>(coarse[c2]);
        }
      // the FP64 correction is added to the precision-T value (values[i,d] += correction[d])
      for (int c = 0; c < 4; ++c) out[i + c * np + d * np * 4] = C(std::complex<double>(acc[c]) + corr[c]);
    }
  }
  return failed;
}

}  // namespace

extern "C" {

void oracle_set_threads(int n) { g_threads = n; }

int oracle_triangle_rule(int order, double* pts, double* w) {
  std::vector<P2<double>> p;
  std::vector<double> ww;
  triangle_rule<double>(order, p, ww);
  for (size_t i = 0; i < ww.size(); ++i) {
    pts[2 * i] = p[i].x;
    pts[2 * i + 1] = p[i].y;
    w[i] = ww[i];
  }
  return (int)ww.size();
}

int oracle_gauss_rule(int order, double* x, double* w) {
  std::vector<double> xx, ww;
  gauss_rule<double>(order, xx, ww);
  for (size_t i = 0; i < ww.size(); ++i) {
    x[i] = xx[i];
    w[i] = ww[i];
  }
  return (int)ww.size();
}

// Base Duffy rule (no orientation remap); test/trial points n x 2.
int oracle_duffy_rule(int order, int kind, double* tp, double* sp, double* w) {
  Duffy<double> d = duffy_rule<double>(order, kind);
  if (tp)
    for (size_t i = 0; i < d.w.size(); ++i) {
      tp[2 * i] = d.tp[i].x;
      tp[2 * i + 1] = d.tp[i].y;
      sp[2 * i] = d.sp[i].x;
      sp[2 * i + 1] = d.sp[i].y;
      w[i] = d.w[i];
    }
  return (int)d.w.size();
}

// Singular pair list in the reference's order: (test, trial, kind, rule) per pair.
int64_t oracle_singular_pairs(const double* xyz, int64_t nv, const int32_t* f, int64_t nf, int order, int32_t* out) {
  Mesh<double> m = make_mesh<double>(xyz, nv, f, nf);
  SingularCache<double> c = build_singular_cache(m, order);
  int64_t n = 0;
  for (auto& v : c.by_test)
    for (auto& p : v) {
      if (out) {
        out[4 * n] = p.test;
        out[4 * n + 1] = p.trial;
        out[4 * n + 2] = c.keys[p.rule][0];
        out[4 * n + 3] = p.rule;
      }
      ++n;
    }
  return n;
}

// sym (all entry points below): 0 off, 1 x, 2 xy, 3 ground.
void oracle_bm_assemble_f64(const double* xyz, int64_t nv, const int32_t* f, int64_t nf, int qo, int so, double k,
                            double eta_re, double eta_im, const double* q, int64_t nrhs, double* A, double* B,
                            int test_lo, int test_hi, int do_singular, int do_identity, int sym) {
  bm_assemble<double>(xyz, nv, f, nf, qo, so, k, eta_re, eta_im, (const cd*)q, nrhs, (cd*)A, (cd*)B, test_lo, test_hi,
                      do_singular, do_identity, sym);
}

void oracle_bm_assemble_f32(const double* xyz, int64_t nv, const int32_t* f, int64_t nf, int qo, int so, double k,
                            double eta_re, double eta_im, const float* q, int64_t nrhs, float* A, float* B,
                            int test_lo, int test_hi, int do_singular, int do_identity, int sym) {
  typedef std::complex<float> cf;
  bm_assemble<float>(xyz, nv, f, nf, qo, so, k, eta_re, eta_im, (const cf*)q, nrhs, (cf*)A, (cf*)B, test_lo, test_hi,
                     do_singular, do_identity, sym);
}

void oracle_operators_f64(const double* xyz, int64_t nv, const int32_t* f, int64_t nf, int qo, int so, double k,
                          double* S, double* D, double* K, double* H, int sym) {
  operators<double>(xyz, nv, f, nf, qo, so, k, (cd*)S, (cd*)D, (cd*)K, (cd*)H, sym);
}

void oracle_operators_f32(const double* xyz, int64_t nv, const int32_t* f, int64_t nf, int qo, int so, double k,
                          float* S, float* D, float* K, float* H, int sym) {
  typedef std::complex<float> cf;
  operators<float>(xyz, nv, f, nf, qo, so, k, (cf*)S, (cf*)D, (cf*)K, (cf*)H, sym);
}

// assemble_l2_identity_matrix (:p1,:p1) N x N and (:p1,:dp0) N x F, real, column-major,
// rows scaled by the orbit weights (JL/src/BeatEngineCore.jl:1187-1214).
void oracle_identity_f64(const double* xyz, int64_t nv, const int32_t* f, int64_t nf, int qo, double* Mpp,
                         double* Mpd, int sym) {
  Mesh<double> m = make_mesh<double>(xyz, nv, f, nf);
  std::vector<P2<double>> rp;
  std::vector<double> rw;
  triangle_rule<double>(qo, rp, rw);
  std::memset(Mpp, 0, sizeof(double) * nv * nv);
  std::memset(Mpd, 0, sizeof(double) * nv * nf);
  for (int64_t e = 0; e < nf; ++e) {
    double pp[3][3], pd[3];
    identity_blocks(rp, rw, m.areas[e], pp, pd);
    for (int a = 0; a < 3; ++a) {
      Mpd[m.faces[e][a] + e * nv] += pd[a];
      for (int b = 0; b < 3; ++b) Mpp[m.faces[e][a] + (int64_t)m.faces[e][b] * nv] += pp[a][b];
    }
  }
  const std::vector<double> ow = orbit_weights(m, sym);
  for (int64_t r = 0; r < nv; ++r) {
    for (int64_t c = 0; c < nv; ++c) Mpp[r + c * nv] *= ow[r];
    for (int64_t c = 0; c < nf; ++c) Mpd[r + c * nv] *= ow[r];
  }
}

void oracle_field_f64(const double* xyz, int64_t nv, const int32_t* f, int64_t nf, int qo, double k, const double* p,
                      const double* q, int64_t nrhs, const double* pts, int64_t np, double* out, int sym) {
  field<double>(xyz, nv, f, nf, qo, k, (const cd*)p, (const cd*)q, nrhs, pts, np, (cd*)out, sym);
}

void oracle_field_f32(const double* xyz, int64_t nv, const int32_t* f, int64_t nf, int qo, double k, const float* p,
                      const float* q, int64_t nrhs, const double* pts, int64_t np, float* out, int sym) {
  typedef std::complex<float> cf;
  field<float>(xyz, nv, f, nf, qo, k, (const cf*)p, (const cf*)q, nrhs, pts, np, (cf*)out, sym);
}

// Field value and gradient (out: np x 4 x nrhs complex, column-major); returns the
// count of (point, face) pairs beyond the adaptive depth limit.
int64_t oracle_field_gradient_f64(const double* xyz, int64_t nv, const int32_t* f, int64_t nf, int qo, double k,
                                  const double* p, const double* q, int64_t nrhs, const double* pts, int64_t np,
                                  double* out, int sym) {
  return field_gradient<double>(xyz, nv, f, nf, qo, k, (const cd*)p, (const cd*)q, nrhs, pts, np, (cd*)out, sym);
}
int64_t oracle_field_gradient_f32(const double* xyz, int64_t nv, const int32_t* f, int64_t nf, int qo, double k,
                                  const float* p, const float* q, int64_t nrhs, const double* pts, int64_t np,
                                  float* out, int sym) {
  typedef std::complex<float> cf;
  return field_gradient<float>(xyz, nv, f, nf, qo, k, (const cf*)p, (const cf*)q, nrhs, pts, np, (cf*)out, sym);
}

// p1_symmetry_orbit_weights per vertex.
void oracle_orbit_weights(const double* xyz, int64_t nv, const int32_t* f, int64_t nf, int sym, double* out) {
  Mesh<double> m = make_mesh<double>(xyz, nv, f, nf);
  const std::vector<double> ow = orbit_weights(m, sym);
  for (int64_t i = 0; i < nv; ++i) out[i] = ow[i];
}

// Image-singular pairs of image transform `image` (0-based, symmetry_transforms order
// without the identity): (test, trial, kind, rule) per pair in the reference's order.
int64_t oracle_image_singular_pairs(const double* xyz, int64_t nv, const int32_t* f, int64_t nf, int order, int sym,
                                    int image, int32_t* out) {
  Mesh<double> m = make_mesh<double>(xyz, nv, f, nf);
  const std::vector<Xf> xfs = image_transforms(sym);
  if (image < 0 || image >= (int)xfs.size()) return -1;
  SingularCache<double> c = build_image_singular_cache(m, order, xfs[image]);
  int64_t n = 0;
  for (auto& v : c.by_test)
    for (auto& p : v) {
      if (out) {
        out[4 * n] = p.test;
        out[4 * n + 1] = p.trial;
        out[4 * n + 2] = c.keys[p.rule][0];
        out[4 * n + 3] = p.rule;
      }
      ++n;
    }
  return n;
}

// Surface curls per face (3 x 3 per face), for checks of the P1 gradient lift.
void oracle_surface_curls(const double* xyz, int64_t nv, const int32_t* f, int64_t nf, double* out) {
  Mesh<double> m = make_mesh<double>(xyz, nv, f, nf);
  for (int64_t e = 0; e < nf; ++e) {
    V3<double> v[3] = {m.verts[m.faces[e][0]], m.verts[m.faces[e][1]], m.verts[m.faces[e][2]]};
    auto c = surface_curls(v, m.normals[e]);
    for (int a = 0; a < 3; ++a) {
      out[9 * e + 3 * a] = c[a].x;
      out[9 * e + 3 * a + 1] = c[a].y;
      out[9 * e + 3 * a + 2] = c[a].z;
    }
  }
}

}  // extern "C"

extern "C" {
// Physical points and weights (jac included) of singular pair `index` in the
// reference's pair order, for brute-force checks of the remapped Duffy rules.
int64_t oracle_singular_pair_points(const double* xyz, int64_t nv, const int32_t* f, int64_t nf, int order,
                                    int64_t index, double* x, double* y, double* w) {
  Mesh<double> m = make_mesh<double>(xyz, nv, f, nf);
  SingularCache<double> c = build_singular_cache(m, order);
  int64_t n = 0;
  for (auto& v : c.by_test)
    for (auto& p : v) {
      if (n++ != index) continue;
      const Duffy<double>& d = c.rules[p.rule];
      V3<double> tv[3] = {m.verts[m.faces[p.test][0]], m.verts[m.faces[p.test][1]], m.verts[m.faces[p.test][2]]};
      V3<double> sv[3] = {m.verts[m.faces[p.trial][0]], m.verts[m.faces[p.trial][1]], m.verts[m.faces[p.trial][2]]};
      for (size_t q = 0; q < d.w.size(); ++q) {
        if (x) {
          V3<double> a = local_to_global(tv, d.tp[q]), b = local_to_global(sv, d.sp[q]);
          x[3 * q] = a.x; x[3 * q + 1] = a.y; x[3 * q + 2] = a.z;
          y[3 * q] = b.x; y[3 * q + 1] = b.y; y[3 * q + 2] = b.z;
          w[q] = d.w[q] * p.jac;
        }
      }
      return (int64_t)d.w.size();
    }
  return -1;
}
}

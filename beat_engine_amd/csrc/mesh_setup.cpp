// Per-mesh preprocessing on the host, done once per mesh and reused across
// frequencies: face and vertex reordering along a Morton curve, face tiles for
// the assembly kernel, per-face geometry records, the touching-pair list and
// the Duffy rules for it, and the unit-area identity blocks.
//
// The quadrature constants and Duffy maps follow the reference's definitions
// (src/beat_engine/julia_local/src/BeatEngineCore.jl): triangle_rule :660-692,
// gauss_rule_1d :713-740, duffy_rule :742-802, remap_shared_* :804-826,
// adjacency_info :598-622, build_singular_correction_cache :1049-1089,
// surface_gradients/curls :856-873, l2_identity_element_matrix :905-923.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <map>
#include <numeric>
#include <set>

#include "bem_common.h"

namespace beat {

namespace {

struct P2 {
  double x, y;
};

// 6-point degree-4 rule with the reference's truncated literals (weights x 0.5).
void rule6(double* p, double* w) {
  for (int i = 0; i < 6; ++i) {
    p[2 * i] = TriRule<6>::xi[i];
    p[2 * i + 1] = TriRule<6>::eta[i];
    w[i] = TriRule<6>::w[i];
  }
}

void gauss01(int order, std::vector<double>& x, std::vector<double>& w) {
  if (order == 1) {
    x = {0.5};
    w = {1.0};
  } else if (order == 2) {
    double a = 0.5 / std::sqrt(3.0);
    x = {0.5 - a, 0.5 + a};
    w = {0.5, 0.5};
  } else if (order == 3) {
    double a = std::sqrt(3.0 / 5.0) / 2.0;
    x = {0.5 - a, 0.5, 0.5 + a};
    w = {5.0 / 18.0, 4.0 / 9.0, 5.0 / 18.0};
  } else {
    double s = std::sqrt(6.0 / 5.0);
    double x1 = std::sqrt(3.0 / 7.0 - 2.0 / 7.0 * s) / 2.0;
    double x2 = std::sqrt(3.0 / 7.0 + 2.0 / 7.0 * s) / 2.0;
    double w1 = (18.0 + std::sqrt(30.0)) / 72.0;
    double w2 = (18.0 - std::sqrt(30.0)) / 72.0;
    x = {0.5 - x2, 0.5 - x1, 0.5 + x1, 0.5 + x2};
    w = {w2, w1, w1, w2};
  }
}

struct Rule {
  std::vector<P2> tp, sp;
  std::vector<double> w;
  void add(double tx, double ty, double sx, double sy, double wt) {
    // points are stored as (x - y, y) of the sub-map coordinates
    tp.push_back({tx - ty, ty});
    sp.push_back({sx - sy, sy});
    w.push_back(wt);
  }
};

// Sauter-Schwab style sub-maps: kind 1 coincident (6), 2 edge (5), 3 vertex (2).
Rule duffy(int order, int kind) {
  std::vector<double> g, gw;
  gauss01(order, g, gw);
  const size_t n = g.size();
  std::vector<P2> tens;
  std::vector<double> tw;
  for (size_t i = 0; i < n; ++i)
    for (size_t j = 0; j < n; ++j) {
      tens.push_back({g[j], g[i]});
      tw.push_back(gw[i] * gw[j]);
    }
  Rule r;
  for (size_t a = 0; a < tens.size(); ++a)
    for (size_t b = 0; b < tens.size(); ++b) {
      const double s = tens[a].x, e1 = tens[a].y, e2 = tens[b].x, e3 = tens[b].y;
      const double e12 = e1 * e2, e123 = e12 * e3, base = tw[a] * tw[b];
      if (kind == 1) {
        const double wt = base * (s * s * s) * (e1 * e1) * e2;
        r.add(s, s * (1.0 - e1 + e12), s * (1.0 - e123), s * (1.0 - e1), wt);
        r.add(s * (1.0 - e123), s * (1.0 - e1), s, s * (1.0 - e1 + e12), wt);
        r.add(s, s * (e1 - e12 + e123), s * (1.0 - e12), s * (e1 - e12), wt);
        r.add(s * (1.0 - e12), s * (e1 - e12), s, s * (e1 - e12 + e123), wt);
        r.add(s * (1.0 - e123), s * (e1 - e123), s, s * (e1 - e12), wt);
        r.add(s, s * (e1 - e12), s * (1.0 - e123), s * (e1 - e123), wt);
      } else if (kind == 2) {
        const double wt = base * (s * s * s) * (e1 * e1);
        r.add(s, s * e1 * e3, s * (1.0 - e12), s * e1 * (1.0 - e2), wt);
        r.add(s, s * e1, s * (1.0 - e123), s * e1 * e2 * (1.0 - e3), wt * e2);
        r.add(s * (1.0 - e12), s * e1 * (1.0 - e2), s, s * e123, wt * e2);
        r.add(s * (1.0 - e123), s * e12 * (1.0 - e3), s, s * e1, wt * e2);
        r.add(s * (1.0 - e123), s * e1 * (1.0 - e2 * e3), s, s * e12, wt * e2);
      } else {
        const double wt = base * (s * s * s) * e2;
        r.add(s, s * e1, s * e2, s * e2 * e3, wt);
        r.add(s * e2, s * e2 * e3, s, s * e1, wt);
      }
    }
  return r;
}

P2 remap_vertex(P2 p, int v) {  // v: 1-based shared vertex
  if (v == 1) return p;
  if (v == 2) return {1.0 - p.x - p.y, p.y};
  return {p.x, 1.0 - p.x - p.y};
}

P2 remap_edge(P2 p, int a, int b) {  // a, b: 1-based shared vertices
  const P2 ref[3] = {{0, 0}, {1, 0}, {0, 1}};
  const int c = 6 - a - b;
  const P2 v0 = ref[a - 1], v1 = ref[b - 1], v2 = ref[c - 1];
  return {v0.x + p.x * (v1.x - v0.x) + p.y * (v2.x - v0.x), v0.y + p.x * (v1.y - v0.y) + p.y * (v2.y - v0.y)};
}

inline uint32_t spread10(uint32_t v) {
  v &= 0x3ff;
  v = (v | (v << 16)) & 0x30000ff;
  v = (v | (v << 8)) & 0x300f00f;
  v = (v | (v << 4)) & 0x30c30c3;
  v = (v | (v << 2)) & 0x9249249;
  return v;
}

}  // namespace

int triangle_rule(int order, double* pts, double* w) {
  if (order <= 1) {
    pts[0] = TriRule<1>::xi[0];
    pts[1] = TriRule<1>::eta[0];
    w[0] = TriRule<1>::w[0];
    return 1;
  }
  if (order == 4) {
    rule6(pts, w);
    return 6;
  }
  for (int i = 0; i < 3; ++i) {
    pts[2 * i] = TriRule<3>::xi[i];
    pts[2 * i + 1] = TriRule<3>::eta[i];
    w[i] = TriRule<3>::w[i];
  }
  return 3;
}

std::string build_host_mesh(const double* V, int64_t nv, const int32_t* F, int64_t nf, int quad_order,
                            int singular_order, int symmetry, HostMesh& m) {
  if (nv < 3 || nf < 1) return "mesh needs at least 3 vertices and 1 face";
  if (symmetry < SYM_OFF || symmetry > SYM_GROUND) return "symmetry must be 0 (off), 1 (x), 2 (xy) or 3 (ground)";
  if (nv > (int64_t)1 << 30 || nf > (int64_t)1 << 30) return "mesh too large";
  if (singular_order < 1 || singular_order > 4) return "singular_order must be in 1..4";
  for (int64_t e = 0; e < 3 * nf; ++e)
    if (F[e] < 0 || F[e] >= nv) return "face vertex index out of range";
  m.nv = nv;
  m.nf = nf;
  m.quad_order = quad_order;
  m.singular_order = singular_order;
  double rp[12], rw[6];
  m.nq = triangle_rule(quad_order, rp, rw);
  m.rule_pts.assign(rp, rp + 2 * m.nq);
  m.rule_w.assign(rw, rw + m.nq);

  // --- per caller face geometry
  std::vector<double> cen(3 * nf), nrm(3 * nf), area(nf);
  for (int64_t e = 0; e < nf; ++e) {
    const double* a = V + 3 * F[3 * e];
    const double* b = V + 3 * F[3 * e + 1];
    const double* c = V + 3 * F[3 * e + 2];
    double e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    double e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
    double cr[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
    double n = std::sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
    if (!(n > 0)) return "degenerate face " + std::to_string(e);
    area[e] = n / 2.0;
    for (int k = 0; k < 3; ++k) {
      nrm[3 * e + k] = cr[k] / n;
      cen[3 * e + k] = (a[k] + b[k] + c[k]) / 3.0;
    }
  }

  // --- Morton order of face centroids
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (int64_t e = 0; e < nf; ++e)
    for (int k = 0; k < 3; ++k) {
      lo[k] = std::min(lo[k], cen[3 * e + k]);
      hi[k] = std::max(hi[k], cen[3 * e + k]);
    }
  double span = std::max({hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2], 1e-300});
  std::vector<uint64_t> key(nf);
  for (int64_t e = 0; e < nf; ++e) {
    uint32_t q[3];
    for (int k = 0; k < 3; ++k) q[k] = (uint32_t)std::min(1023.0, std::floor((cen[3 * e + k] - lo[k]) / span * 1024.0));
    key[e] = ((uint64_t)(spread10(q[0]) | (spread10(q[1]) << 1) | (spread10(q[2]) << 2)) << 32) | (uint64_t)e;
  }
  std::sort(key.begin(), key.end());
  m.face_perm.resize(nf);
  for (int64_t i = 0; i < nf; ++i) m.face_perm[i] = (int32_t)(key[i] & 0xffffffffu);

  // --- vertex numbering: first appearance along the face order
  m.vert_inv.assign(nv, -1);
  m.vert_perm.clear();
  for (int64_t i = 0; i < nf; ++i)
    for (int k = 0; k < 3; ++k) {
      int v = F[3 * m.face_perm[i] + k];
      if (m.vert_inv[v] < 0) {
        m.vert_inv[v] = (int32_t)m.vert_perm.size();
        m.vert_perm.push_back(v);
      }
    }
  for (int64_t v = 0; v < nv; ++v)  // vertices no face touches keep an identity row
    if (m.vert_inv[v] < 0) {
      m.vert_inv[v] = (int32_t)m.vert_perm.size();
      m.vert_perm.push_back((int32_t)v);
    }
  m.verts.resize(3 * nv);
  for (int64_t i = 0; i < nv; ++i)
    for (int k = 0; k < 3; ++k) m.verts[3 * i + k] = V[3 * m.vert_perm[i] + k];
  m.fverts.resize(3 * nf);
  for (int64_t i = 0; i < nf; ++i)
    for (int k = 0; k < 3; ++k) m.fverts[3 * i + k] = m.vert_inv[F[3 * m.face_perm[i] + k]];

  // --- tiles: test tiles (TILE_F faces / TILE_V vertices) and trial tiles
  // (TILE_FS / TILE_VS, and the narrow TILE_FS_N / TILE_VS_N), all greedy along
  // the face order
  // faces [f0, f1) are cut into tiles appended to the arrays
  auto tiles_in = [&](int64_t f0, int64_t f1, int maxf, int maxv, std::vector<int32_t>& st,
                      std::vector<int32_t>& cnt, std::vector<int32_t>& tnv, std::vector<int32_t>& tverts,
                      std::vector<uint8_t>& loc) {
    int64_t i = f0;
    while (i < f1) {
      int start = (int)i, count = 0;
      std::vector<int32_t> tv;
      while (i < f1 && count < maxf) {
        int add = 0;
        for (int k = 0; k < 3; ++k) {
          int v = m.fverts[3 * i + k];
          bool have = std::find(tv.begin(), tv.end(), v) != tv.end();
          bool dup = false;
          for (int kk = 0; kk < k; ++kk) dup |= m.fverts[3 * i + kk] == v;
          if (!have && !dup) ++add;
        }
        if ((int)tv.size() + add > maxv) break;
        for (int k = 0; k < 3; ++k) {
          int v = m.fverts[3 * i + k];
          auto it = std::find(tv.begin(), tv.end(), v);
          if (it == tv.end()) {
            tv.push_back(v);
            it = tv.end() - 1;
          }
          loc[3 * i + k] = (uint8_t)(it - tv.begin());
        }
        ++count;
        ++i;
      }
      st.push_back(start);
      cnt.push_back(count);
      tnv.push_back((int)tv.size());
      tv.resize(maxv, tv.empty() ? 0 : tv[0]);
      tverts.insert(tverts.end(), tv.begin(), tv.end());
    }
  };
  auto tiles = [&](int maxf, int maxv, std::vector<int32_t>& st, std::vector<int32_t>& cnt,
                   std::vector<int32_t>& tnv, std::vector<int32_t>& tverts, std::vector<uint8_t>& loc) {
    st.clear();
    cnt.clear();
    tnv.clear();
    tverts.clear();
    loc.assign(3 * nf, 0);
    tiles_in(0, nf, maxf, maxv, st, cnt, tnv, tverts, loc);
  };
  tiles(TILE_F, TILE_V, m.tile_start, m.tile_count, m.tile_nv, m.tile_verts, m.floc);
  tiles(TILE_FS, TILE_VS, m.stile_start, m.stile_count, m.stile_nv, m.stile_verts, m.flocs);
  tiles(TILE_FS_N, TILE_VS_N, m.ntile_start, m.ntile_count, m.ntile_nv, m.ntile_verts, m.flocn);
  // sub-tiles nested in the test tiles, and the unit ranges of the symmetric pass
  {
    const size_t nt = m.tile_start.size();
    m.qtile_start.clear();
    m.qtile_count.clear();
    m.qtile_nv.clear();
    m.qtile_verts.clear();
    m.flocq.assign(3 * nf, 0);
    m.tile_qfirst.assign(nt + 1, 0);
    for (size_t t = 0; t < nt; ++t) {
      m.tile_qfirst[t] = (int32_t)m.qtile_start.size();
      tiles_in(m.tile_start[t], (int64_t)m.tile_start[t] + m.tile_count[t], TILE_FS_N, TILE_VS_N, m.qtile_start,
               m.qtile_count, m.qtile_nv, m.qtile_verts, m.flocq);
    }
    const int32_t nq = (int32_t)m.qtile_start.size();
    m.tile_qfirst[nt] = nq;
    m.sym_unit_start.assign(nt + 1, 0);
    for (size_t t = 0; t < nt; ++t) m.sym_unit_start[t + 1] = m.sym_unit_start[t] + (nq - m.tile_qfirst[t]);
    // test-side slot rotations: greedily, face by face, the rotation that puts its
    // vertices into the slots where the tile's faces so far hold them least often
    m.frot.assign(nf, 0);
    for (size_t t = 0; t < nt; ++t) {
      int cnt[TILE_V][3] = {};
      const int64_t f0 = m.tile_start[t], f1 = f0 + m.tile_count[t];
      for (int64_t f = f0; f < f1; ++f) {
        int best = 0, best_cost = 1 << 30;
        for (int r = 0; r < 3; ++r) {
          int cost = 0;
          for (int k = 0; k < 3; ++k) cost += cnt[m.floc[3 * f + (k + r) % 3]][k];
          if (cost < best_cost) {
            best_cost = cost;
            best = r;
          }
        }
        m.frot[f] = (uint8_t)best;
        for (int k = 0; k < 3; ++k) ++cnt[m.floc[3 * f + (k + best) % 3]][k];
      }
    }
  }

  // --- face records (internal order)
  m.rec.assign((size_t)FREC * nf, 0.0);
  for (int64_t i = 0; i < nf; ++i) {
    const int64_t e = m.face_perm[i];
    const double* v[3] = {V + 3 * F[3 * e], V + 3 * F[3 * e + 1], V + 3 * F[3 * e + 2]};
    double* r = &m.rec[(size_t)FREC * i];
    for (int q = 0; q < m.nq; ++q) {
      const double xi = m.rule_pts[2 * q], eta = m.rule_pts[2 * q + 1], a = 1.0 - xi - eta;
      for (int k = 0; k < 3; ++k) r[3 * q + k] = a * v[0][k] + xi * v[1][k] + eta * v[2][k];
    }
    const double* n = &nrm[3 * e];
    for (int k = 0; k < 3; ++k) r[REC_N + k] = n[k];
    // surface curls n x grad(phi_a); grads through the inverse Gram matrix
    double e1[3], e2[3];
    for (int k = 0; k < 3; ++k) {
      e1[k] = v[1][k] - v[0][k];
      e2[k] = v[2][k] - v[0][k];
    }
    const double g11 = e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2];
    const double g12 = e1[0] * e2[0] + e1[1] * e2[1] + e1[2] * e2[2];
    const double g22 = e2[0] * e2[0] + e2[1] * e2[1] + e2[2] * e2[2];
    const double det = g11 * g22 - g12 * g12;
    const double i11 = g22 / det, i12 = -g12 / det, i22 = g11 / det;
    double gr[3][3];
    for (int k = 0; k < 3; ++k) {
      const double l1 = e1[k] * i11 + e2[k] * i12, l2 = e1[k] * i12 + e2[k] * i22;
      gr[0][k] = -l1 - l2;
      gr[1][k] = l1;
      gr[2][k] = l2;
    }
    for (int a = 0; a < 3; ++a) {
      r[REC_CURL + 3 * a + 0] = n[1] * gr[a][2] - n[2] * gr[a][1];
      r[REC_CURL + 3 * a + 1] = n[2] * gr[a][0] - n[0] * gr[a][2];
      r[REC_CURL + 3 * a + 2] = n[0] * gr[a][1] - n[1] * gr[a][0];
    }
    r[REC_AREA] = area[e];
    r[REC_D] = v[0][0] * n[0] + v[0][1] * n[1] + v[0][2] * n[2];
  }

  // --- touching pairs and their remapped Duffy rules (reference order per test face)
  std::vector<int32_t> internal_of(nf);
  for (int64_t i = 0; i < nf; ++i) internal_of[m.face_perm[i]] = (int32_t)i;
  std::vector<std::vector<int32_t>> v2e(nv);
  for (int64_t e = 0; e < nf; ++e)
    for (int k = 0; k < 3; ++k) v2e[F[3 * e + k]].push_back((int32_t)e);
  Rule base[3] = {duffy(singular_order, 1), duffy(singular_order, 2), duffy(singular_order, 3)};
  std::map<std::array<int, 5>, int> rule_index;
  m.rule_off.clear();
  m.rule_cnt.clear();
  m.rule_kind.clear();
  m.rule_data.clear();
  m.spairs.clear();
  m.n_coincident = m.n_edge = m.n_vertex = 0;
  for (int64_t t = 0; t < nf; ++t) {
    std::vector<int32_t> cand;
    for (int k = 0; k < 3; ++k)
      for (int32_t s : v2e[F[3 * t + k]])
        if (std::find(cand.begin(), cand.end(), s) == cand.end()) cand.push_back(s);
    for (int32_t s : cand) {
      int sa[9], sb[9], ns = 0;
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
          if (F[3 * t + i] == F[3 * s + j]) {
            sa[ns] = i + 1;
            sb[ns] = j + 1;
            ++ns;
          }
      if (ns == 0) continue;
      std::array<int, 5> k5;
      int kind;
      if (ns >= 3) {
        kind = 1;
        k5 = {1, 0, 0, 0, 0};
      } else if (ns == 2) {
        if (sb[1] < sb[0]) {
          std::swap(sa[0], sa[1]);
          std::swap(sb[0], sb[1]);
        }
        kind = 2;
        k5 = {2, sa[0], sa[1], sb[0], sb[1]};
      } else {
        kind = 3;
        k5 = {3, sa[0], 0, sb[0], 0};
      }
      auto it = rule_index.find(k5);
      int ri;
      if (it == rule_index.end()) {
        const Rule& b = base[kind - 1];
        ri = (int)m.rule_off.size();
        m.rule_off.push_back((int32_t)(m.rule_data.size() / 5));
        m.rule_cnt.push_back((int32_t)b.w.size());
        m.rule_kind.push_back(kind);
        for (size_t q = 0; q < b.w.size(); ++q) {
          P2 a = b.tp[q], c = b.sp[q];
          if (kind == 2) {
            a = remap_edge(a, k5[1], k5[2]);
            c = remap_edge(c, k5[3], k5[4]);
          } else if (kind == 3) {
            a = remap_vertex(a, k5[1]);
            c = remap_vertex(c, k5[3]);
          }
          m.rule_data.insert(m.rule_data.end(), {a.x, a.y, c.x, c.y, b.w[q]});
        }
        rule_index[k5] = ri;
      } else {
        ri = it->second;
      }
      SingularPair p;
      p.test = internal_of[t];
      p.trial = internal_of[s];
      p.rule = ri;
      p.xform = 0;
      p.jac = (2.0 * area[t]) * (2.0 * area[s]);
      p.nprod = nrm[3 * t] * nrm[3 * s] + nrm[3 * t + 1] * nrm[3 * s + 1] + nrm[3 * t + 2] * nrm[3 * s + 2];
      m.spairs.push_back(p);
      if (kind == 1) ++m.n_coincident;
      else if (kind == 2) ++m.n_edge;
      else ++m.n_vertex;
    }
  }

  // --- symmetry images: pairs whose reflected trial face touches the test face
  // get the Duffy value minus the regular value the image pass adds for them
  // (image_singular_candidates, BeatEngineCore.jl:1139-1166, geometric_adjacency_info
  // :624-658, _beat_cpu_image_singular_cache, BeatEngineCpuAssembly.jl:488-523,
  // _beat_cpu_bm_image_singular_delta_test!, BeatEngineCpuBurtonMiller.jl:232-265).
  // The delta rule of an orientation key is the remapped Duffy rule followed by
  // the regular product rule with negated weights, so one pass over it forms the
  // difference.
  m.symmetry = symmetry;
  m.nimg = image_transforms(symmetry, m.img);
  m.n_image_singular = 0;
  m.weighted_rows.clear();
  m.row_weights.clear();
  if (m.nimg > 0) {
    const double gtol = 1e-8;
    auto key_of = [&](const double* p, const int* sg) {
      return std::array<long long, 3>{(long long)std::nearbyint(sg[0] * p[0] / gtol),
                                      (long long)std::nearbyint(sg[1] * p[1] / gtol),
                                      (long long)std::nearbyint(sg[2] * p[2] / gtol)};
    };
    const int ident[3] = {1, 1, 1};
    std::map<std::array<long long, 3>, std::vector<int32_t>> by_key;
    for (int64_t t = 0; t < nf; ++t)
      for (int k = 0; k < 3; ++k) by_key[key_of(V + 3 * F[3 * t + k], ident)].push_back((int32_t)t);
    for (int x = 0; x < m.nimg; ++x) {
      const ImageXf& X = m.img[x];
      std::map<std::array<int, 5>, int> delta_index;
      std::set<std::pair<int32_t, int32_t>> seen;
      for (int64_t s = 0; s < nf; ++s)
        for (int k = 0; k < 3; ++k) {
          auto it = by_key.find(key_of(V + 3 * F[3 * s + k], X.s));
          if (it == by_key.end()) continue;
          for (int32_t t : it->second) {
            if (!seen.insert({t, (int32_t)s}).second) continue;
            int sa[9], sb[9], ns = 0;
            for (int i = 0; i < 3; ++i)
              for (int j = 0; j < 3; ++j) {
                const double* a = V + 3 * F[3 * t + i];
                const double* b = V + 3 * F[3 * s + j];
                const double d0 = a[0] - X.s[0] * b[0], d1 = a[1] - X.s[1] * b[1], d2 = a[2] - X.s[2] * b[2];
                if (std::sqrt(d0 * d0 + d1 * d1 + d2 * d2) <= gtol) {
                  sa[ns] = i + 1;
                  sb[ns] = j + 1;
                  ++ns;
                }
              }
            if (ns == 0) continue;
            std::array<int, 5> k5;
            int kind;
            if (ns >= 3) {
              kind = 1;
              k5 = {1, 0, 0, 0, 0};
            } else if (ns == 2) {
              if (sb[1] < sb[0]) {
                std::swap(sa[0], sa[1]);
                std::swap(sb[0], sb[1]);
              }
              kind = 2;
              k5 = {2, sa[0], sa[1], sb[0], sb[1]};
            } else {
              kind = 3;
              k5 = {3, sa[0], 0, sb[0], 0};
            }
            auto dit = delta_index.find(k5);
            int ri;
            if (dit == delta_index.end()) {
              const Rule& b = base[kind - 1];
              ri = (int)m.rule_off.size();
              m.rule_off.push_back((int32_t)(m.rule_data.size() / 5));
              m.rule_cnt.push_back((int32_t)(b.w.size() + (size_t)m.nq * m.nq));
              m.rule_kind.push_back(kind);
              for (size_t q = 0; q < b.w.size(); ++q) {
                P2 a = b.tp[q], c = b.sp[q];
                if (kind == 2) {
                  a = remap_edge(a, k5[1], k5[2]);
                  c = remap_edge(c, k5[3], k5[4]);
                } else if (kind == 3) {
                  a = remap_vertex(a, k5[1]);
                  c = remap_vertex(c, k5[3]);
                }
                m.rule_data.insert(m.rule_data.end(), {a.x, a.y, c.x, c.y, b.w[q]});
              }
              for (int i = 0; i < m.nq; ++i)
                for (int j = 0; j < m.nq; ++j)
                  m.rule_data.insert(m.rule_data.end(), {m.rule_pts[2 * i], m.rule_pts[2 * i + 1], m.rule_pts[2 * j],
                                                         m.rule_pts[2 * j + 1], -(m.rule_w[i] * m.rule_w[j])});
              delta_index[k5] = ri;
            } else {
              ri = dit->second;
            }
            SingularPair p;
            p.test = internal_of[t];
            p.trial = internal_of[s];
            p.rule = ri;
            p.xform = x + 1;
            p.jac = (2.0 * area[t]) * (2.0 * area[s]);
            p.nprod = nrm[3 * t] * (X.s[0] * nrm[3 * s]) + nrm[3 * t + 1] * (X.s[1] * nrm[3 * s + 1]) +
                      nrm[3 * t + 2] * (X.s[2] * nrm[3 * s + 2]);
            m.spairs.push_back(p);
            ++m.n_image_singular;
          }
        }
    }
  }
  // pairs by rule size, largest first, within the ordinary and the image section
  // (stable; every consumer is order-free): the singular kernel runs several pairs
  // per wavefront, which then mostly share their point count
  {
    auto by_size = [&](const SingularPair& x, const SingularPair& y) {
      return m.rule_cnt[x.rule] > m.rule_cnt[y.rule];
    };
    const auto mid = m.spairs.end() - m.n_image_singular;
    std::stable_sort(m.spairs.begin(), mid, by_size);
    std::stable_sort(mid, m.spairs.end(), by_size);
  }
  // P1 orbit weights: rows on a mirror plane stand for one physical dof shared
  // with its image (symmetry_plane_tolerance / p1_symmetry_orbit_weights,
  // BeatEngineCore.jl:497-517,563-580); the rigid ground has none.
  if (symmetry == SYM_X || symmetry == SYM_XY) {
    double blo[3] = {V[0], V[1], V[2]}, bhi[3] = {V[0], V[1], V[2]};
    for (int64_t v = 0; v < nv; ++v)
      for (int k = 0; k < 3; ++k) {
        blo[k] = std::min(blo[k], V[3 * v + k]);
        bhi[k] = std::max(bhi[k], V[3 * v + k]);
      }
    const double dg = std::sqrt((bhi[0] - blo[0]) * (bhi[0] - blo[0]) + (bhi[1] - blo[1]) * (bhi[1] - blo[1]) +
                                (bhi[2] - blo[2]) * (bhi[2] - blo[2]));
    const double ptol = std::max(1e-9, dg * 1e-6);
    for (int64_t v = 0; v < nv; ++v) {
      double w = 1.0;
      if (std::fabs(V[3 * v]) <= ptol) w *= 2.0;
      if (symmetry == SYM_XY && std::fabs(V[3 * v + 1]) <= ptol) w *= 2.0;
      if (w != 1.0) {
        m.weighted_rows.push_back(m.vert_inv[v]);
        m.row_weights.push_back(w);
      }
    }
  }

  // --- unit-area identity blocks from the rule (jac = 2 A)
  for (int a = 0; a < 3; ++a) {
    m.id_pd[a] = 0;
    for (int b = 0; b < 3; ++b) m.id_pp[a][b] = 0;
  }
  for (int q = 0; q < m.nq; ++q) {
    const double xi = m.rule_pts[2 * q], eta = m.rule_pts[2 * q + 1];
    const double ph[3] = {1.0 - xi - eta, xi, eta};
    for (int a = 0; a < 3; ++a) {
      m.id_pd[a] += ph[a] * m.rule_w[q] * 2.0;
      for (int b = 0; b < 3; ++b) m.id_pp[a][b] += ph[a] * ph[b] * m.rule_w[q] * 2.0;
    }
  }
  return "";
}

}  // namespace beat

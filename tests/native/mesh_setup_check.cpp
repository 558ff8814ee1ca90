// Host check of the library's per-mesh preprocessing (build_host_mesh, compiled
// for the host by the CPU test): reads a mesh, builds it with a symmetry mode and
// prints what the tests compare with the oracle.
//   input  (stdin): nv nf quad_order singular_order symmetry, then nv x 3
//          vertices and nf x 3 faces
//   output: "pair test trial kind xform" per touching pair (caller face ids),
//           "rule index count weight_sum" per rule,
//           "weight vertex w" per orbit-weighted row (caller vertex ids),
//           "images n" and "done"
#include <cstdio>
#include <vector>

#include "../../beat_engine_amd/csrc/bem_common.h"

int main() {
  long long nv, nf;
  int qo, so, sym;
  if (std::scanf("%lld %lld %d %d %d", &nv, &nf, &qo, &so, &sym) != 5) return 2;
  std::vector<double> V(3 * nv);
  std::vector<int32_t> F(3 * nf);
  for (auto& v : V)
    if (std::scanf("%lf", &v) != 1) return 2;
  for (auto& f : F)
    if (std::scanf("%d", &f) != 1) return 2;
  beat::HostMesh m;
  const std::string msg = beat::build_host_mesh(V.data(), nv, F.data(), nf, qo, so, sym, m);
  if (!msg.empty()) {
    std::printf("error %s\n", msg.c_str());
    return 1;
  }
  for (const auto& p : m.spairs)
    std::printf("pair %d %d %d %d\n", m.face_perm[p.test], m.face_perm[p.trial], m.rule_kind[p.rule], p.xform);
  for (size_t r = 0; r < m.rule_off.size(); ++r) {
    double ws = 0;
    for (int q = 0; q < m.rule_cnt[r]; ++q) ws += m.rule_data[5 * (size_t)(m.rule_off[r] + q) + 4];
    std::printf("rule %zu %d %.17g\n", r, m.rule_cnt[r], ws);
  }
  for (size_t i = 0; i < m.weighted_rows.size(); ++i)
    std::printf("weight %d %.17g\n", m.vert_perm[m.weighted_rows[i]], m.row_weights[i]);
  std::printf("images %d\ndone\n", m.nimg);
  return 0;
}

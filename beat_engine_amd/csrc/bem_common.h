// Shared definitions for the gfx950 Burton-Miller library.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

namespace beat {

// Face tiles: up to TILE_F faces that touch at most TILE_V distinct vertices.
constexpr int TILE_F = 64;
constexpr int TILE_V = 64;
// Trial tiles of the regular assembly: up to TILE_FS faces touching at most
// TILE_VS distinct vertices (a second partition of the same face order). A third,
// narrow partition (TILE_FS_N / TILE_VS_N) halves the regular kernel's LDS, so
// that a factorization's panel workgroup fits on a CU beside two assembly
// workgroups when a sweep assembles the next system during a factorization.
constexpr int TILE_FS = 32;
constexpr int TILE_VS = 32;
constexpr int TILE_FS_N = 16;
constexpr int TILE_VS_N = 16;
// Per-face geometry record (FREC scalars): regular quadrature points (q*3+c, up
// to 6 points), unit normal, surface curls (a*3+c), area, plane offset v1.n.
constexpr int FREC = 32;
constexpr int REC_N = 18;
constexpr int REC_CURL = 21;
constexpr int REC_AREA = 30;
constexpr int REC_D = 31;

constexpr double INV_FOUR_PI = 0.079577471545947667884441881686257181;

// Singular (touching) pair: internal test/trial face, rule, image transform of
// the trial face (0: none), jacobian and normal product.
struct SingularPair {
  int32_t test, trial, rule, xform;
  double jac, nprod;
};

// Symmetry modes and their image transforms (the reference's
// symmetry_image_transforms, BeatEngineCore.jl:448-495): a point maps to
// (s0 x, s1 y, s2 z), a normal likewise, a surface curl to det * (s0, s1, s2) c.
constexpr int SYM_OFF = 0, SYM_X = 1, SYM_XY = 2, SYM_GROUND = 3;
constexpr int MAX_IMAGES = 3;
struct ImageXf {
  int s[3];
  int det;
};
// image transforms of a mode, in the reference's order; returns their count
inline int image_transforms(int mode, ImageXf* out) {
  if (mode == SYM_X) {
    out[0] = {{-1, 1, 1}, -1};
    return 1;
  }
  if (mode == SYM_XY) {
    out[0] = {{-1, 1, 1}, -1};
    out[1] = {{1, -1, 1}, -1};
    out[2] = {{-1, -1, 1}, 1};
    return 3;
  }
  if (mode == SYM_GROUND) {
    out[0] = {{1, -1, 1}, -1};
    return 1;
  }
  return 0;
}

// The regular triangle rules on the reference triangle (area 1/2): per point the
// (xi, eta) coordinates and the weight. One point (order 1), three points (orders
// 2 and 3) and the six-point table of order 4, as BeatEngineCore.jl tabulates them.
// Shared by the host set-up and, as compile-time constants, the regular kernels.
template <int NQ>
struct TriRule;
template <>
struct TriRule<1> {
  static constexpr double xi[1] = {1.0 / 3.0};
  static constexpr double eta[1] = {1.0 / 3.0};
  static constexpr double w[1] = {0.5};
};
template <>
struct TriRule<3> {
  static constexpr double xi[3] = {1.0 / 6.0, 2.0 / 3.0, 1.0 / 6.0};
  static constexpr double eta[3] = {1.0 / 6.0, 1.0 / 6.0, 2.0 / 3.0};
  static constexpr double w[3] = {1.0 / 6.0, 1.0 / 6.0, 1.0 / 6.0};
};
template <>
struct TriRule<6> {
  static constexpr double xi[6] = {0.4459484909159651, 0.0915762135097710, 0.1081030181680700,
                                   0.4459484909159651, 0.8168475729804590, 0.0915762135097710};
  static constexpr double eta[6] = {0.4459484909159651, 0.0915762135097700, 0.4459484909159651,
                                    0.1081030181680700, 0.0915762135097700, 0.8168475729804580};
  static constexpr double w[6] = {0.5 * 0.2233815896780110, 0.5 * 0.1099517436553220, 0.5 * 0.2233815896780110,
                                  0.5 * 0.2233815896780110, 0.5 * 0.1099517436553220, 0.5 * 0.1099517436553220};
};

// Host-side mesh after preprocessing; all arrays in internal (tile) order
// except where noted.
struct HostMesh {
  int64_t nv = 0, nf = 0;
  int quad_order = 4, singular_order = 4, nq = 6;
  std::vector<double> rule_pts;  // nq x 2
  std::vector<double> rule_w;    // nq
  std::vector<int32_t> face_perm;  // internal face -> caller face
  std::vector<int32_t> vert_perm;  // internal vertex -> caller vertex
  std::vector<int32_t> vert_inv;   // caller vertex -> internal vertex
  std::vector<int32_t> fverts;     // 3 per internal face: internal vertex ids
  std::vector<uint8_t> floc;       // 3 per internal face: vertex slot in its tile
  std::vector<double> rec;         // FREC per internal face
  std::vector<double> verts;       // 3 per internal vertex
  std::vector<int32_t> tile_start, tile_count, tile_nv;
  std::vector<int32_t> tile_verts;  // TILE_V per tile
  std::vector<int32_t> stile_start, stile_count, stile_nv;
  std::vector<int32_t> stile_verts;  // TILE_VS per trial tile
  std::vector<uint8_t> flocs;        // 3 per internal face: vertex slot in its trial tile
  std::vector<int32_t> ntile_start, ntile_count, ntile_nv;
  std::vector<int32_t> ntile_verts;  // TILE_VS_N per narrow trial tile
  std::vector<uint8_t> flocn;        // 3 per internal face: vertex slot in its narrow trial tile
  // sub-tiles of the symmetric regular pass: each test tile cut greedily into
  // runs of at most TILE_FS_N faces / TILE_VS_N vertices, so that sub-tiles nest
  // in test tiles; tile_qfirst[t] is test tile t's first sub-tile (ntiles + 1
  // entries); sym_unit_start[t] is the first (test tile, sub-tile) unit of test
  // tile t, which pairs it with every sub-tile from its own first one on
  std::vector<int32_t> qtile_start, qtile_count, qtile_nv;
  std::vector<int32_t> qtile_verts;  // TILE_VS_N per sub-tile
  std::vector<uint8_t> flocq;        // 3 per internal face: vertex slot in its sub-tile
  std::vector<int32_t> tile_qfirst;
  std::vector<int64_t> sym_unit_start;
  // per internal face, the rotation r (0..2) of the order in which the symmetric
  // pass visits its test-side basis (slot k stands for vertex (k + r) % 3): chosen
  // per test tile so that the faces sharing a vertex hold it in different slots,
  // which spreads the LDS atomics of one instruction over distinct addresses
  std::vector<uint8_t> frot;
  std::vector<SingularPair> spairs;
  std::vector<int32_t> rule_off, rule_cnt, rule_kind;
  std::vector<double> rule_data;  // 5 per rule point: test xi, eta, trial xi, eta, weight
  int64_t n_coincident = 0, n_edge = 0, n_vertex = 0;
  // symmetry: image transforms, the image-singular pairs among spairs (xform > 0,
  // placed after the ordinary pairs) and the P1 orbit weights of the rows that
  // lie on a symmetry plane (internal vertex ids)
  int symmetry = SYM_OFF, nimg = 0;
  ImageXf img[MAX_IMAGES];
  int64_t n_image_singular = 0;
  std::vector<int32_t> weighted_rows;
  std::vector<double> row_weights;
  double id_pp[3][3];  // unit-area P1 x P1 mass block from the rule
  double id_pd[3];     // unit-area P1 x DP0 block
};

std::string build_host_mesh(const double* verts, int64_t nv, const int32_t* faces, int64_t nf, int quad_order,
                            int singular_order, int symmetry, HostMesh& out);

// Reference rules (restated from BeatEngineCore.jl, see mesh_setup.cpp).
int triangle_rule(int order, double* pts, double* w);

}  // namespace beat

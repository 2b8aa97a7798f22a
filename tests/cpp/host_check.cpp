// host_check.cpp -- CPU-only checks of the solver's host code (the tile planner pbs_tiles.cpp and the
// mesh utilities pbs_mesh.cpp), built with g++ from those sources alone so that it can also run
// under AddressSanitizer / UndefinedBehaviorSanitizer.  No GPU, no HIP runtime.
//
// Tile plan: for every tile the corner slots form a permutation of the tile's 4 * tet_cnt slots;
// a simulated gather (each local node sums its contiguous run of slots) visits, for every local
// node, exactly that node's corners in (tet, corner) order, so its FP32 sums are bitwise the
// sequential sums; owned + halo partials add up to the direct per-node sums; halo partial slots
// are a permutation grouped by owner node.
//
//   usage: host_check <golden dir>      (exit status 0 = all checks passed)
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "pbs_host.h"
#include "pbs_tiles.h"
#include "phybs.h"

namespace pbs {
int set_error(int code, const std::string &msg) {
  std::fprintf(stderr, "  (error %d: %s)\n", code, msg.c_str());
  return code;
}
}  // namespace pbs

static int g_fail = 0;
#define CHECK(cond, ...)                                  \
  do {                                                    \
    if (!(cond)) {                                        \
      std::fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__); \
      std::fprintf(stderr, __VA_ARGS__);                  \
      std::fprintf(stderr, "\n");                         \
      ++g_fail;                                           \
      return;                                             \
    }                                                     \
  } while (0)

// Checks one plan against the mesh it was built from (tets in the caller's ids).
static void check_plan(const char *name, int nn, int nt, const std::vector<int> &tets, const pbs::TilePlan &P) {
  CHECK(P.nn == nn && P.nt == nt, "%s: sizes", name);
  std::vector<int> seen_tet(nt, 0);
  std::mt19937 rng(1234);
  std::uniform_real_distribution<float> U(-1.0f, 1.0f);
  // per-corner values (solver tet order), direct per-node sums in (tet, corner) order
  std::vector<float> val(4 * (size_t)nt);
  for (auto &v : val) v = U(rng);
  std::vector<double> direct(nn, 0.0), viaplan(nn, 0.0);
  for (int t = 0; t < nt; ++t)
    for (int c = 0; c < 4; ++c) direct[P.tets_new[4 * (size_t)t + c]] += val[4 * (size_t)t + c];
  std::vector<int> hslot_seen(P.halo_total, 0);
  for (int b = 0; b < P.ntiles; ++b) {
    const pbs::TileInfo &ti = P.tiles[b];
    const int L = ti.own_cnt + ti.halo_cnt, cnt = ti.tet_cnt;
    CHECK(L <= P.lcap && cnt <= P.tcap, "%s tile %d: caps", name, b);
    auto global = [&](int l) { return l < ti.own_cnt ? ti.own_start + l : P.halo[ti.halo_off + l - ti.own_cnt]; };
    // slot of every corner; slots a permutation of [0, 4 cnt)
    std::vector<int> slot_corner(4 * (size_t)cnt, -1);
    std::vector<std::vector<int>> runs(L);  // per local node: corners in (tet, corner) order
    for (int s = 0; s < cnt; ++s) {
      const int t = ti.tet_start + s;
      CHECK(t < nt && !seen_tet[t], "%s tile %d: tet %d twice or out of range", name, b, t);
      seen_tet[t] = 1;
      const uint32_t *w = &P.tet_lp[4 * (size_t)t];
      const int lid[4] = {(int)(w[0] & 0xffff), (int)(w[0] >> 16), (int)(w[1] & 0xffff), (int)(w[1] >> 16)};
      const int pos[4] = {(int)(w[2] & 0xffff), (int)(w[2] >> 16), (int)(w[3] & 0xffff), (int)(w[3] >> 16)};
      int smallest = nn;
      for (int c = 0; c < 4; ++c) {
        CHECK(lid[c] < L, "%s tile %d: local id", name, b);
        CHECK(global(lid[c]) == P.tets_new[4 * (size_t)t + c], "%s tile %d: local id maps to the wrong node", name, b);
        CHECK(pos[c] < 4 * cnt && slot_corner[pos[c]] < 0, "%s tile %d: slot reused or out of range", name, b);
        slot_corner[pos[c]] = 4 * t + c;
        runs[lid[c]].push_back(4 * t + c);
        smallest = std::min(smallest, P.tets_new[4 * (size_t)t + c]);
      }
      CHECK(smallest >= ti.own_start && smallest < ti.own_start + ti.own_cnt, "%s tile %d: tet not owned", name, b);
    }
    // simulated gather: local node l sums slots tile_off[l] .. tile_off[l+1]-1 in order
    const uint16_t *off = &P.tile_off[ti.off_start];
    for (int l = 0; l < L; ++l) {
      const int len = off[l + 1] - off[l];
      CHECK(len == (int)runs[l].size(), "%s tile %d: run length %d vs %zu", name, b, len, runs[l].size());
      float seq = 0.0f, sim = 0.0f;
      for (int j = 0; j < len; ++j) {
        const int sl = off[l] + j;
        CHECK(sl < 4 * cnt && slot_corner[sl] == runs[l][j], "%s tile %d: slot of run entry %d", name, b, j);
        sim += val[slot_corner[sl]];
        seq += val[runs[l][j]];
      }
      CHECK(std::memcmp(&sim, &seq, sizeof(float)) == 0, "%s tile %d: gather sum differs from the run order", name, b);
      viaplan[global(l)] += sim;
      if (l >= ti.own_cnt) {
        const int hs = ti.halo_off + (l - ti.own_cnt);
        CHECK(hs < (int)P.halo_total && !hslot_seen[hs], "%s tile %d: halo entry", name, b);
        hslot_seen[hs] = 1;
      }
    }
    CHECK(off[L] == 4 * cnt, "%s tile %d: runs do not cover the tile's slots", name, b);
  }
  for (int t = 0; t < nt; ++t) CHECK(seen_tet[t], "%s: tet %d in no tile", name, t);
  double worst = 0.0;
  for (int n = 0; n < nn; ++n) worst = std::max(worst, std::fabs(viaplan[n] - direct[n]));
  CHECK(worst <= 1e-4, "%s: owned + halo partials differ from the direct sums by %g", name, worst);
  // halo partial slots: a permutation, contiguous per owner node
  std::vector<int> slot_seen(P.halo_total, 0);
  for (long h = 0; h < P.halo_total; ++h) {
    const int s = P.halo_slot[h], n = P.halo[h];
    CHECK(s >= P.bnd_off[n] && s < P.bnd_off[n + 1] && !slot_seen[s], "%s: halo slot %ld", name, h);
    slot_seen[s] = 1;
  }
  std::printf("  %-28s tiles %6d  lcap %4d  tcap %4d  halo %8ld  ok\n", name, P.ntiles, P.lcap, P.tcap, P.halo_total);
}

static void check_bar(int nx, int ny, int nz, int tile) {
  const int nn = (nx + 1) * (ny + 1) * (nz + 1), nt = 6 * nx * ny * nz;
  std::vector<float> nodes(3 * (size_t)nn);
  std::vector<int> tets(4 * (size_t)nt);
  if (pbs_mesh_bar(nx, ny, nz, 0.05f, nodes.data(), tets.data()) != PBS_OK) {
    std::fprintf(stderr, "FAIL pbs_mesh_bar\n");
    ++g_fail;
    return;
  }
  pbs::TilePlan P;
  std::string msg;
  if (!pbs::build_tile_plan(nn, nodes.data(), nt, tets.data(), tile, P, msg)) {
    std::fprintf(stderr, "FAIL plan %dx%dx%d tile %d: %s\n", nx, ny, nz, tile, msg.c_str());
    ++g_fail;
    return;
  }
  char name[64];
  std::snprintf(name, sizeof(name), "bar %dx%dx%d tile %d", nx, ny, nz, tile);
  check_plan(name, nn, nt, tets, P);
}

static void check_tetgen(const std::string &prefix, int tile) {
  int nn = 0, nt = 0, nf = 0;
  if (pbs_tetgen_count(prefix.c_str(), &nn, &nt, &nf) != PBS_OK) {
    std::fprintf(stderr, "FAIL tetgen_count %s\n", prefix.c_str());
    ++g_fail;
    return;
  }
  std::vector<float> nodes(3 * (size_t)nn);
  std::vector<int> tets(4 * (size_t)nt), faces(3 * (size_t)std::max(nf, 1));
  if (pbs_tetgen_load(prefix.c_str(), 0.1f, nodes.data(), tets.data(), nf ? faces.data() : nullptr) != PBS_OK) {
    std::fprintf(stderr, "FAIL tetgen_load %s\n", prefix.c_str());
    ++g_fail;
    return;
  }
  // the integer adjacency structures (sizes from the first call, then filled)
  const int64_t nnz = pbs_csr_pattern(nn, nt, tets.data(), nullptr, nullptr, nullptr, nullptr);
  if (nnz <= 0) {
    std::fprintf(stderr, "FAIL csr_pattern %s\n", prefix.c_str());
    ++g_fail;
    return;
  }
  std::vector<int> rowptr(3 * (size_t)nn + 1), col(nnz), diag(3 * (size_t)nn), kid(48 * (size_t)nt);
  pbs_csr_pattern(nn, nt, tets.data(), rowptr.data(), col.data(), diag.data(), kid.data());
  std::vector<int> ioff(nn + 1), ient(4 * (size_t)nt);
  pbs_incidence(nn, nt, tets.data(), ioff.data(), ient.data());
  if (nf > 0) {
    const int64_t nring = pbs_face_ring(nn, nf, faces.data(), nullptr, nullptr);
    if (nring > 0) {
      std::vector<int64_t> roff(nf + 1);
      std::vector<int> ring(nring);
      pbs_face_ring(nn, nf, faces.data(), roff.data(), ring.data());
    }
  }
  pbs::TilePlan P;
  std::string msg;
  if (!pbs::build_tile_plan(nn, nodes.data(), nt, tets.data(), tile, P, msg)) {
    std::fprintf(stderr, "FAIL plan %s: %s\n", prefix.c_str(), msg.c_str());
    ++g_fail;
    return;
  }
  check_plan((prefix.substr(prefix.rfind('/') + 1) + " tile " + std::to_string(tile)).c_str(), nn, nt, tets, P);
}

static void check_batch() {
  // instances of 1 tet, a 2x1x3 bar and a 4x3x5 bar side by side
  std::vector<float> nodes;
  std::vector<int> tets, node_off{0}, tet_off{0};
  auto add = [&](int nx, int ny, int nz) {
    const int nn = (nx + 1) * (ny + 1) * (nz + 1), nt = 6 * nx * ny * nz;
    std::vector<float> x(3 * (size_t)nn);
    std::vector<int> t(4 * (size_t)nt);
    pbs_mesh_bar(nx, ny, nz, 0.05f, x.data(), t.data());
    const int base = node_off.back();
    nodes.insert(nodes.end(), x.begin(), x.end());
    for (int v : t) tets.push_back(v + base);
    node_off.push_back(base + nn);
    tet_off.push_back(tet_off.back() + nt);
  };
  {
    const float x[12] = {0, 0, 0, 0.05f, 0, 0, 0, 0.05f, 0, 0, 0, 0.05f};
    nodes.insert(nodes.end(), x, x + 12);
    for (int v : {0, 1, 2, 3}) tets.push_back(v);
    node_off.push_back(4);
    tet_off.push_back(1);
  }
  add(2, 1, 3);
  add(4, 3, 5);
  const int ninst = (int)node_off.size() - 1;
  // one tile size for all instances, then one per instance (tiles never exceed their instance's)
  const std::vector<int> uniform(ninst, 32), mixed = {16, 8, 64};
  for (const std::vector<int> *tiles : {&uniform, &mixed}) {
    pbs::TilePlan P;
    std::string msg;
    if (!pbs::build_batch_plan(ninst, node_off.data(), tet_off.data(), nodes.data(), tets.data(), tiles->data(), P,
                               msg)) {
      std::fprintf(stderr, "FAIL batch plan: %s\n", msg.c_str());
      ++g_fail;
      return;
    }
    check_plan(tiles == &uniform ? "batch of 3 instances tile 32" : "batch of 3 instances tiles 16/8/64",
               node_off.back(), tet_off.back(), tets, P);
    for (int b = 0; b < P.ntiles; ++b) {
      const pbs::TileInfo &ti = P.tiles[b];
      if (!(ti.inst >= 0 && ti.inst < ninst && ti.own_start >= P.inst_node_off[ti.inst] &&
            ti.own_start + ti.own_cnt <= P.inst_node_off[ti.inst + 1])) {
        std::fprintf(stderr, "FAIL batch tile %d mixes instances\n", b);
        ++g_fail;
        return;
      }
      if (ti.tet_cnt > (*tiles)[ti.inst]) {
        std::fprintf(stderr, "FAIL batch tile %d has %d tets, above its instance's tile size %d\n", b, ti.tet_cnt,
                     (*tiles)[ti.inst]);
        ++g_fail;
        return;
      }
    }
    if (P.tile_tets != *std::max_element(tiles->begin(), tiles->end()) || P.inst_tile_tets != *tiles) {
      std::fprintf(stderr, "FAIL batch plan tile sizes\n");
      ++g_fail;
      return;
    }
  }
}

int main(int argc, char **argv) {
  const std::string golden = argc > 1 ? argv[1] : "tests/golden";
  std::printf("host_check: tile plans and mesh utilities\n");
  check_bar(1, 1, 1, 16);
  for (int tile : {16, 64, 256, 512}) check_bar(4, 4, 16, tile);
  check_bar(12, 12, 50, 256);
  check_bar(16, 16, 32, 128);
  check_bar(3, 7, 2, 33);
  for (const char *m : {"model_tet/tet.1", "model_tet2/tet.2", "model_box/box.1", "model_bar/bar.1"})
    for (int tile : {64, 256}) check_tetgen(golden + "/" + m, tile);
  check_batch();
  if (g_fail) {
    std::printf("host_check: %d failure(s)\n", g_fail);
    return 1;
  }
  std::printf("host_check: all passed\n");
  return 0;
}

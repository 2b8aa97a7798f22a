// pbs_tiles.cpp -- node renumbering (Morton order of the rest positions) and the node-tile
// partition used by the fused descent kernel.  Host-side, run once per solver.
#include <algorithm>
#include <cmath>
#include <numeric>
#include <string>

#include "pbs_tiles.h"

namespace pbs {

static inline uint64_t spread3(uint64_t v) {  // 21 bits -> every third bit
  v &= 0x1fffffULL;
  v = (v | (v << 32)) & 0x1f00000000ffffULL;
  v = (v | (v << 16)) & 0x1f0000ff0000ffULL;
  v = (v | (v << 8)) & 0x100f00f00f00f00fULL;
  v = (v | (v << 4)) & 0x10c30c30c30c30c3ULL;
  v = (v | (v << 2)) & 0x1249249249249249ULL;
  return v;
}

bool build_tile_plan(int nn, const float *rest, int nt, const int *tets, int tile_nodes, TilePlan &P,
                     std::string &msg) {
  P = TilePlan();
  P.nn = nn;
  P.nt = nt;
  P.tile_nodes = tile_nodes;
  // ---- node order: Morton code of positions quantised to the shortest tet edge ----
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (int i = 0; i < nn; ++i)
    for (int d = 0; d < 3; ++d) {
      lo[d] = std::min(lo[d], (double)rest[3 * i + d]);
      hi[d] = std::max(hi[d], (double)rest[3 * i + d]);
    }
  double ext = std::max({hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2], 1e-30});
  double hq = 1e300;
  static const int E[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
  for (int t = 0; t < nt; ++t)
    for (auto &e : E) {
      const float *a = rest + 3 * tets[4 * t + e[0]], *b = rest + 3 * tets[4 * t + e[1]];
      double l = std::sqrt((double)(a[0] - b[0]) * (a[0] - b[0]) + (double)(a[1] - b[1]) * (a[1] - b[1]) +
                           (double)(a[2] - b[2]) * (a[2] - b[2]));
      if (l > 0) hq = std::min(hq, l);
    }
  if (!(hq < 1e300)) hq = ext / 1024.0;
  hq = std::max(hq, ext / 2097151.0);
  std::vector<uint64_t> code(nn);
  for (int i = 0; i < nn; ++i) {
    uint64_t q[3];
    for (int d = 0; d < 3; ++d) q[d] = (uint64_t)std::llround(((double)rest[3 * i + d] - lo[d]) / hq);
    code[i] = spread3(q[0]) | (spread3(q[1]) << 1) | (spread3(q[2]) << 2);
  }
  P.perm.resize(nn);
  std::iota(P.perm.begin(), P.perm.end(), 0);
  std::stable_sort(P.perm.begin(), P.perm.end(), [&](int a, int b) { return code[a] < code[b]; });
  P.iperm.resize(nn);
  for (int i = 0; i < nn; ++i) P.iperm[P.perm[i]] = i;
  // ---- tet order: by smallest new node id, then original id (corner order kept) ----
  std::vector<int> tkey(nt);
  for (int t = 0; t < nt; ++t) {
    int m = nn;
    for (int j = 0; j < 4; ++j) m = std::min(m, P.iperm[tets[4 * t + j]]);
    tkey[t] = m;
  }
  P.tperm.resize(nt);
  std::iota(P.tperm.begin(), P.tperm.end(), 0);
  std::stable_sort(P.tperm.begin(), P.tperm.end(), [&](int a, int b) { return tkey[a] < tkey[b]; });
  P.itperm.resize(nt);
  for (int t = 0; t < nt; ++t) P.itperm[P.tperm[t]] = t;
  P.tets_new.resize(4 * (size_t)nt);
  for (int t = 0; t < nt; ++t)
    for (int j = 0; j < 4; ++j) P.tets_new[4 * (size_t)t + j] = P.iperm[tets[4 * (size_t)P.tperm[t] + j]];
  // ---- setup incidence: per new node, 4*new_tet + corner in ascending original tet id ----
  P.inc_off.assign(nn + 1, 0);
  for (int t = 0; t < nt; ++t)
    for (int j = 0; j < 4; ++j) P.inc_off[P.iperm[tets[4 * t + j]] + 1]++;
  for (int i = 0; i < nn; ++i) P.inc_off[i + 1] += P.inc_off[i];
  P.inc.resize(4 * (size_t)nt);
  {
    std::vector<int> fill(nn, 0);
    for (int t = 0; t < nt; ++t)
      for (int j = 0; j < 4; ++j) {
        int n = P.iperm[tets[4 * t + j]];
        P.inc[P.inc_off[n] + fill[n]++] = 4 * P.itperm[t] + j;
      }
  }
  // ---- tiles: owned nodes, owned tets (smallest node owned), halo nodes, chunk incidence ----
  P.ntiles = (nn + tile_nodes - 1) / tile_nodes;
  P.tiles.resize(P.ntiles);
  P.tet_local.assign(4 * (size_t)nt, 0);
  std::vector<int> loc(nn, -1);
  std::vector<std::vector<int>> bnd_lists(nn);
  int t0 = 0;
  for (int b = 0; b < P.ntiles; ++b) {
    TileInfo &ti = P.tiles[b];
    ti.own_start = b * tile_nodes;
    ti.own_cnt = std::min(tile_nodes, nn - ti.own_start);
    const int own_end = ti.own_start + ti.own_cnt;
    int t1 = t0;
    while (t1 < nt && tkey[P.tperm[t1]] < own_end) ++t1;
    ti.tet_start = t0;
    ti.tet_cnt = t1 - t0;
    ti.halo_off = (int)P.halo.size();
    std::vector<int> hl;
    for (int t = t0; t < t1; ++t)
      for (int j = 0; j < 4; ++j) {
        int n = P.tets_new[4 * (size_t)t + j];
        if (n >= own_end) hl.push_back(n);
      }
    std::sort(hl.begin(), hl.end());
    hl.erase(std::unique(hl.begin(), hl.end()), hl.end());
    ti.halo_cnt = (int)hl.size();
    if (ti.own_cnt + ti.halo_cnt > 65535) {
      msg = "tile has more than 65535 local nodes: use a smaller tile";
      return false;
    }
    for (int k = 0; k < ti.halo_cnt; ++k) {
      loc[hl[k]] = ti.own_cnt + k;
      P.halo.push_back(hl[k]);
      bnd_lists[hl[k]].push_back(ti.halo_off + k);
    }
    for (int t = t0; t < t1; ++t)
      for (int j = 0; j < 4; ++j) {
        int n = P.tets_new[4 * (size_t)t + j];
        P.tet_local[4 * (size_t)t + j] = (uint16_t)(n < own_end ? n - ti.own_start : loc[n]);
      }
    // incidence blocks per chunk of CHUNK_TETS tets
    ti.chunk_first = (int)P.chunk_off.size();
    ti.nchunk = (ti.tet_cnt + CHUNK_TETS - 1) / CHUNK_TETS;
    std::vector<std::pair<int, int>> ent;  // (local node, slot<<2|corner)
    for (int c = 0; c < ti.nchunk; ++c) {
      ent.clear();
      const int cs = t0 + c * CHUNK_TETS, ce = std::min(t1, cs + CHUNK_TETS);
      for (int t = cs; t < ce; ++t)
        for (int j = 0; j < 4; ++j) ent.push_back({P.tet_local[4 * (size_t)t + j], ((t - cs) << 2) | j});
      std::sort(ent.begin(), ent.end());
      std::vector<uint16_t> nodes, offs;
      for (size_t i = 0; i < ent.size(); ++i)
        if (i == 0 || ent[i].first != ent[i - 1].first) {
          nodes.push_back((uint16_t)ent[i].first);
          offs.push_back((uint16_t)i);
        }
      offs.push_back((uint16_t)ent.size());
      const size_t start = P.chunk_inc.size();
      P.chunk_off.push_back((uint32_t)start);
      P.chunk_inc.push_back((uint16_t)nodes.size());
      for (uint16_t v : nodes) P.chunk_inc.push_back(v);
      for (uint16_t v : offs) P.chunk_inc.push_back(v);
      for (auto &e : ent) P.chunk_inc.push_back((uint16_t)e.second);
      while (P.chunk_inc.size() % 8) P.chunk_inc.push_back(0);
      if (P.chunk_inc.size() - start > (size_t)CHUNK_WORDS_MAX) {
        msg = "chunk incidence block exceeds 8 KiB";
        return false;
      }
    }
    for (int k = 0; k < ti.halo_cnt; ++k) loc[hl[k]] = -1;
    P.lcap = std::max(P.lcap, ti.own_cnt + ti.halo_cnt);
    t0 = t1;
  }
  P.chunk_off.push_back((uint32_t)P.chunk_inc.size());  // end of the last block
  P.halo_total = (long)P.halo.size();
  P.bnd_off.assign(nn + 1, 0);
  for (int n = 0; n < nn; ++n) P.bnd_off[n + 1] = P.bnd_off[n] + (int)bnd_lists[n].size();
  P.bnd.reserve(P.bnd_off[nn]);
  for (int n = 0; n < nn; ++n)
    for (int v : bnd_lists[n]) P.bnd.push_back(v);
  return true;
}

}  // namespace pbs

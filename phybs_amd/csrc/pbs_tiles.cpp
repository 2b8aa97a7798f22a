// pbs_tiles.cpp -- node renumbering (Morton order of the rest positions), tet ordering and the
// tile partition used by the tile pass of the descent.  Host-side, run once per solver.
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

bool build_tile_plan(int nn, const float *rest, int nt, const int *tets, int tile_tets, TilePlan &P,
                     std::string &msg) {
  P = TilePlan();
  P.nn = nn;
  P.nt = nt;
  P.tile_tets = tile_tets;
  if (tile_tets < 1 || tile_tets > 16383) {
    msg = "tile_tets must be in [1, 16383]";
    return false;
  }
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
  // ---- tiles: owned node runs cut at tile_tets owned tets, at max(64, tile_tets / 2) local
  // nodes (owned + halo) and at kMaxOwn owned nodes (which bounds the LDS of meshes with many
  // tet-less nodes); halo nodes; contribution runs ----
  constexpr int kMaxOwn = 1024;
  std::vector<int> own_tets(nn + 1, 0);  // tets per smallest node
  for (int t = 0; t < nt; ++t) own_tets[tkey[t]]++;
  P.tet_lp.assign(4 * (size_t)nt, 0);
  std::vector<int> loc(nn, -1);
  std::vector<std::vector<int>> bnd_lists(nn);
  std::vector<uint16_t> lid(4 * (size_t)std::max(tile_tets, 1));
  std::vector<std::pair<int, int>> ent;  // (local node, 4*slot + corner)
  std::vector<int> first_tet(nn + 1, 0);  // tets with smallest node n: [first_tet[n], first_tet[n+1])
  for (int n = 0; n < nn; ++n) first_tet[n + 1] = first_tet[n] + own_tets[n];
  // local node cap: tiles whose tets reach unusually many distinct nodes are cut short
  const int lmax = std::max(64, tile_tets / 2);
  std::vector<int> stamp(nn, -1);
  int n0 = 0, t0 = 0, ntile = 0;
  while (n0 < nn) {
    int n1 = n0, cnt = 0, nloc = 0;
    while (n1 < nn) {
      if (n1 > n0 && (cnt + own_tets[n1] > tile_tets || n1 - n0 >= kMaxOwn)) break;
      // distinct nodes that adding n1 brings (n1 itself and its tets' corners)
      int add = 0;
      const int tag = 2 * ntile + 1;  // provisional mark; committed marks are 2 * ntile
      auto probe = [&](int n) {
        if (stamp[n] != 2 * ntile && stamp[n] != tag) {
          stamp[n] = tag;
          ++add;
        }
      };
      probe(n1);
      for (int t = first_tet[n1]; t < first_tet[n1 + 1]; ++t)
        for (int j = 0; j < 4; ++j) probe(P.tets_new[4 * (size_t)t + j]);
      const bool take = n1 == n0 || nloc + add <= lmax;
      auto settle = [&](int n) {
        if (stamp[n] == tag) stamp[n] = take ? 2 * ntile : -1;
      };
      settle(n1);
      for (int t = first_tet[n1]; t < first_tet[n1 + 1]; ++t)
        for (int j = 0; j < 4; ++j) settle(P.tets_new[4 * (size_t)t + j]);
      if (!take) break;
      nloc += add;
      cnt += own_tets[n1++];
    }
    ++ntile;
    if (cnt > tile_tets) {
      msg = "a node is the smallest node of more tets than tile_tets";
      return false;
    }
    TileInfo ti{};
    ti.own_start = n0;
    ti.own_cnt = n1 - n0;
    ti.tet_start = t0;
    ti.tet_cnt = cnt;
    const int t1 = t0 + cnt;
    ti.halo_off = (int)P.halo.size();
    std::vector<int> hl;
    for (int t = t0; t < t1; ++t)
      for (int j = 0; j < 4; ++j) {
        int n = P.tets_new[4 * (size_t)t + j];
        if (n >= n1) hl.push_back(n);
      }
    std::sort(hl.begin(), hl.end());
    hl.erase(std::unique(hl.begin(), hl.end()), hl.end());
    ti.halo_cnt = (int)hl.size();
    const int L = ti.own_cnt + ti.halo_cnt;
    if (L + 1 > 65535) {
      msg = "tile has more than 65534 local nodes: use smaller tiles";
      return false;
    }
    for (int k = 0; k < ti.halo_cnt; ++k) {
      loc[hl[k]] = ti.own_cnt + k;
      P.halo.push_back(hl[k]);
      bnd_lists[hl[k]].push_back(ti.halo_off + k);
    }
    ent.clear();
    for (int t = t0; t < t1; ++t)
      for (int j = 0; j < 4; ++j) {
        int n = P.tets_new[4 * (size_t)t + j];
        int l = n < n1 ? n - n0 : loc[n];
        lid[4 * (size_t)(t - t0) + j] = (uint16_t)l;
        ent.push_back({l, 4 * (t - t0) + j});
      }
    std::sort(ent.begin(), ent.end());
    ti.off_start = (int)P.tile_off.size();
    {
      size_t e = 0;
      for (int i = 0; i <= L; ++i) {
        while (e < ent.size() && ent[e].first < i) ++e;
        P.tile_off.push_back((uint16_t)e);
      }
    }
    std::vector<uint16_t> pos(4 * (size_t)cnt);
    for (size_t e = 0; e < ent.size(); ++e) pos[ent[e].second] = (uint16_t)e;
    for (int t = t0; t < t1; ++t) {
      const uint16_t *l = &lid[4 * (size_t)(t - t0)], *q = &pos[4 * (size_t)(t - t0)];
      uint32_t *w = &P.tet_lp[4 * (size_t)t];
      w[0] = (uint32_t)l[0] | ((uint32_t)l[1] << 16);
      w[1] = (uint32_t)l[2] | ((uint32_t)l[3] << 16);
      w[2] = (uint32_t)q[0] | ((uint32_t)q[1] << 16);
      w[3] = (uint32_t)q[2] | ((uint32_t)q[3] << 16);
    }
    for (int k = 0; k < ti.halo_cnt; ++k) loc[hl[k]] = -1;
    P.lcap = std::max(P.lcap, L);
    P.tcap = std::max(P.tcap, cnt);
    P.tiles.push_back(ti);
    n0 = n1;
    t0 = t1;
  }
  P.ntiles = (int)P.tiles.size();
  P.ninst = 1;
  P.inst_node_off = {0, nn};
  P.inst_tet_off = {0, nt};
  P.inst_tile_off = {0, P.ntiles};
  P.halo_total = (long)P.halo.size();
  P.bnd_off.assign(nn + 1, 0);
  for (int n = 0; n < nn; ++n) P.bnd_off[n + 1] = P.bnd_off[n] + (int)bnd_lists[n].size();
  P.halo_slot.assign(P.halo.size(), 0);
  for (int n = 0; n < nn; ++n)
    for (size_t j = 0; j < bnd_lists[n].size(); ++j) P.halo_slot[bnd_lists[n][j]] = P.bnd_off[n] + (int)j;
  return true;
}

bool build_batch_plan(int ninst, const int *node_off, const int *tet_off, const float *rest, const int *tets,
                      int tile_tets, TilePlan &P, std::string &msg) {
  const std::vector<int> tt(std::max(ninst, 0), tile_tets);
  return build_batch_plan(ninst, node_off, tet_off, rest, tets, tt.data(), P, msg);
}

bool build_batch_plan(int ninst, const int *node_off, const int *tet_off, const float *rest, const int *tets,
                      const int *inst_tile_tets, TilePlan &P, std::string &msg) {
  P = TilePlan();
  P.ninst = ninst;
  P.nn = node_off[ninst];
  P.nt = tet_off[ninst];
  P.inst_tile_tets.assign(inst_tile_tets, inst_tile_tets + ninst);
  P.tile_tets = ninst > 0 ? *std::max_element(inst_tile_tets, inst_tile_tets + ninst) : 0;
  P.inst_node_off.assign(node_off, node_off + ninst + 1);
  P.inst_tet_off.assign(tet_off, tet_off + ninst + 1);
  P.inst_tile_off.assign(1, 0);
  P.bnd_off.assign(1, 0);
  std::vector<int> lt;
  for (int i = 0; i < ninst; ++i) {
    const int n0 = node_off[i], n1 = node_off[i + 1], t0 = tet_off[i], t1 = tet_off[i + 1];
    if (n1 < n0 || t1 < t0) {
      msg = "instance offsets must not decrease";
      return false;
    }
    lt.resize(4 * (size_t)(t1 - t0));
    for (size_t k = 0; k < lt.size(); ++k) {
      const int g = tets[4 * (size_t)t0 + k];
      if (g < n0 || g >= n1) {
        msg = "instance " + std::to_string(i) + ": tet node outside the instance's node range";
        return false;
      }
      lt[k] = g - n0;
    }
    TilePlan Q;
    if (!build_tile_plan(n1 - n0, rest + 3 * (size_t)n0, t1 - t0, lt.data(), inst_tile_tets[i], Q, msg)) {
      msg = "instance " + std::to_string(i) + ": " + msg;
      return false;
    }
    const int tiles0 = (int)P.tiles.size(), off0 = (int)P.tile_off.size();
    const long halo0 = (long)P.halo.size();
    for (int v : Q.perm) P.perm.push_back(v + n0);
    for (int v : Q.iperm) P.iperm.push_back(v + n0);
    for (int v : Q.tperm) P.tperm.push_back(v + t0);
    for (int v : Q.itperm) P.itperm.push_back(v + t0);
    for (int v : Q.tets_new) P.tets_new.push_back(v + n0);
    const int inc0 = P.inc_off.empty() ? 0 : P.inc_off.back();
    if (P.inc_off.empty()) P.inc_off.push_back(0);
    for (int n = 0; n < n1 - n0; ++n) P.inc_off.push_back(inc0 + Q.inc_off[n + 1]);
    for (int v : Q.inc) P.inc.push_back(v + 4 * t0);
    for (TileInfo ti : Q.tiles) {
      ti.own_start += n0;
      ti.halo_off += (int)halo0;
      ti.tet_start += t0;
      ti.off_start += off0;
      ti.inst = i;
      P.tiles.push_back(ti);
    }
    for (int v : Q.halo) P.halo.push_back(v + n0);
    for (uint16_t v : Q.tile_off) P.tile_off.push_back(v);
    for (uint32_t v : Q.tet_lp) P.tet_lp.push_back(v);
    const int b0 = P.bnd_off.back();
    for (int n = 0; n < n1 - n0; ++n) P.bnd_off.push_back(b0 + Q.bnd_off[n + 1]);
    for (int v : Q.halo_slot) P.halo_slot.push_back(v + b0);
    P.lcap = std::max(P.lcap, Q.lcap);
    P.tcap = std::max(P.tcap, Q.tcap);
    P.inst_tile_off.push_back(tiles0 + Q.ntiles);
  }
  if (P.inc_off.empty()) P.inc_off.push_back(0);
  P.ntiles = (int)P.tiles.size();
  P.halo_total = (long)P.halo.size();
  return true;
}

}  // namespace pbs

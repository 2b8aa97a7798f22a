// pbs_tiles.h -- host-side partition of a tet mesh for the tiled descent kernels.
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

namespace pbs {

// Nodes are renumbered in Morton order and tets sorted by their smallest new node id.  Tile b
// owns the nodes [own_start, own_start + own_cnt) and the tets whose smallest node it owns --
// a contiguous tet range [tet_start, tet_start + tet_cnt) -- so every tet is evaluated once.
// Tiles are cut so that each holds at most tile_tets tets, 1024 owned nodes and, unless one
// node's tets alone exceed it, max(64, tile_tets / 2) local nodes.  A tile's tets also reach "halo"
// nodes owned by later tiles; the tile writes one partial force per owned and per halo node,
// and the node pass adds a node's partials in tile order.  Halo partial slots are numbered by
// owner node, so each node's partials are contiguous and the node pass reads them coalesced.
//
// Inside a tile, node i (local id: owned nodes first, then halo nodes) receives the corner
// contributions run [off[i], off[i+1]) of the tile's contribution array, ordered by tet, then
// corner.  Each tet carries its corners' local node ids and the position of each corner in that
// array, so the tet pass writes every corner force straight into its slot and the gather sums
// contiguous runs.
struct TileInfo {
  int own_start, own_cnt;  // owned nodes
  int halo_off, halo_cnt;  // halo node ids: halo[halo_off .. +halo_cnt); partial slots likewise
  int tet_start, tet_cnt;  // owned tets (solver order)
  int off_start;           // run offsets: tile_off[off_start .. off_start + own_cnt + halo_cnt]
  int inst;                // mesh instance (batched solvers; 0 otherwise)
};

struct TilePlan {
  int nn = 0, nt = 0;
  // instances (independent meshes of a batched solver) own contiguous node, tet and tile ranges
  int ninst = 1;
  std::vector<int> inst_node_off, inst_tet_off, inst_tile_off;
  int tile_tets = 0, ntiles = 0;
  std::vector<int> inst_tile_tets;  // per instance: its tile size
  int lcap = 0;                     // max local nodes (owned + halo) over tiles
  int tcap = 0;                     // max tets over tiles
  long halo_total = 0;              // sum of halo counts = number of halo partial slots
  std::vector<int> perm, iperm;     // node: new -> old, old -> new
  std::vector<int> tperm, itperm;   // tet:  new -> old, old -> new
  std::vector<int> tets_new;        // 4 per tet, new order, new node ids
  std::vector<int> inc_off, inc;    // per new node: 4*new_tet + corner, in ascending original tet id
  std::vector<TileInfo> tiles;
  std::vector<int> halo;            // new node ids
  std::vector<uint16_t> tile_off;   // per tile: local node count + 1 run offsets
  std::vector<uint32_t> tet_lp;     // 4 words per tet: local ids (c0 | c1 << 16, c2 | c3 << 16),
                                    // contribution positions (likewise)
  std::vector<int> bnd_off;         // per node: its halo partial slots [bnd_off[n], bnd_off[n+1]), tile order
  std::vector<int> halo_slot;       // per halo entry (tile, halo node): its partial slot
};

// Builds the plan.  rest: 3N AoS; tets: 4T.  tile_tets: max tets per tile (1 .. 16383).
// Returns false (with msg) if a tile exceeds the 16-bit local index limits.
bool build_tile_plan(int nn, const float *rest, int nt, const int *tets, int tile_tets, TilePlan &plan,
                     std::string &msg);
// Batched: ninst meshes, instance i owning nodes [node_off[i], node_off[i+1]) and tets
// [tet_off[i], tet_off[i+1]) of the concatenated arrays (tets hold global node ids inside their
// instance's range).  Each instance is planned on its own with its tile size inst_tile_tets[i];
// the plans are concatenated (plan.tile_tets = the largest of them).
bool build_batch_plan(int ninst, const int *node_off, const int *tet_off, const float *rest, const int *tets,
                      const int *inst_tile_tets, TilePlan &plan, std::string &msg);
// The same with one tile size for every instance.
bool build_batch_plan(int ninst, const int *node_off, const int *tet_off, const float *rest, const int *tets,
                      int tile_tets, TilePlan &plan, std::string &msg);

}  // namespace pbs

// pbs_tiles.h -- host-side partition of a tet mesh for the tiled descent kernels.
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

namespace pbs {

// Nodes are renumbered in Morton order and tets sorted by their smallest new node id.  Tile b
// owns the nodes [own_start, own_start + own_cnt) and the tets whose smallest node it owns --
// a contiguous tet range [tet_start, tet_start + tet_cnt) -- so every tet is evaluated once.
// Its tets also reach "halo" nodes owned by later tiles; the tile writes a partial force for
// each owned and each halo node, and the node pass adds a node's partials in tile order.
struct TileInfo {
  int own_start, own_cnt;   // owned nodes
  int halo_off, halo_cnt;   // halo node ids: halo[halo_off .. +halo_cnt); partial slots likewise
  int tet_start, tet_cnt;   // owned tets (solver order)
  int chunk_first, nchunk;  // incidence blocks of the tile's tet chunks: chunk_off[chunk_first + c]
};

// Incidence block of one chunk of CHUNK_TETS consecutive tets of a tile (16-bit words, padded to
// a multiple of 8 words): [n_nodes, node_0 .. node_{n-1}, off_0 .. off_n, entry ...], where
// node_i is a tile-local node id, entries off_i .. off_{i+1}-1 (counted from the first entry)
// belong to node_i, and an entry is (tet slot in chunk << 2) | corner, ascending per node.
constexpr int CHUNK_TETS = 256;
constexpr int CHUNK_WORDS_MAX = 4096;  // 8 KiB: two 16-byte vectors per thread of a 256-thread block

struct TilePlan {
  int nn = 0, nt = 0;
  int tile_nodes = 0, ntiles = 0;
  int lcap = 0;                     // max local nodes (owned + halo) over tiles
  long halo_total = 0;              // sum of halo counts = number of halo partial slots
  std::vector<int> perm, iperm;     // node: new -> old, old -> new
  std::vector<int> tperm, itperm;   // tet:  new -> old, old -> new
  std::vector<int> tets_new;        // 4 per tet, new order, new node ids
  std::vector<int> inc_off, inc;    // per new node: 4*new_tet + corner, in ascending original tet id
  std::vector<TileInfo> tiles;
  std::vector<int> halo;            // new node ids
  std::vector<uint16_t> tet_local;  // 4 per tet (new order): tile-local ids of its corners
  std::vector<uint32_t> chunk_off;  // per chunk: offset (16-bit words, multiple of 8) into chunk_inc
  std::vector<uint16_t> chunk_inc;  // incidence blocks
  std::vector<int> bnd_off, bnd;    // per node: halo partial slots holding part of its force, tile order
};

// Builds the plan.  rest: 3N AoS; tets: 4T.  tile_nodes: owned nodes per tile.
// Returns false (with msg) if a tile exceeds the 16-bit local id limit.
bool build_tile_plan(int nn, const float *rest, int nt, const int *tets, int tile_nodes, TilePlan &plan,
                     std::string &msg);

}  // namespace pbs

// Tiled neighbor list: the full list of a block of particles (a tile)
// rewritten against the tile's own set of neighbors.
//
// A tile is a run of T = 512 or 256 columns.  Built by the per-tile walk
// (k_tile_walk.hip) the columns are the tile's cell-list slots reordered by
// list length, column s belonging to particle order[s]; built from a plain
// list (k_tiles.hip) column i is particle i.  The build collects the distinct
// neighbors of the tile's particles into jset[t][0..count[t]) and stores
// every list entry as 16 x its rank in that set (the byte offset of the
// neighbor's staged position in LDS), two entries per 32-bit word: word m of
// column s (at list[m * pitch + s]) holds entries 2m (low half) and 2m + 1
// (high half).  Columns are padded with 16 x count[t] (one past the last
// rank) up to their wavefront's longest list rounded up to 8.  The force
// kernel stages the tile's neighbor positions in LDS once (slot count[t]
// holds a dummy) and gathers from LDS, and the list stream shrinks from 4 to
// 2 bytes per entry.
#pragma once

#include <cstdint>

#include "azp_internal.h"
#include "cells.h"

namespace azp
{

// Most LDS position slots per tile (distinct neighbors + the dummy slot):
// 3968 int4 = 62 KB, leaving room for the per-type-pair tables of up to 6
// types within a 64 KB workgroup allocation.
constexpr uint32_t kTileCap = 3968;

struct TileList
    {
    DeviceArray<uint32_t> count; // [ntiles]
    DeviceArray<uint32_t> jset;  // [ntiles * kTileCap]
    DeviceArray<uint32_t> list;  // [words * pitch]
    DeviceArray<uint32_t> flag;  // [0] largest distinct count above capacity, [1] largest count
    DeviceArray<uint32_t> nn;        // per-tile walk: list length of column s
    DeviceArray<uint32_t> order_buf; // per-tile walk: particle of column s
    DeviceArray<uint32_t> col;       // per-tile walk: column of particle i
    // Tiles of the per-tile walk: a tile's columns are its cell-list slots
    // reordered by list length (longest first); column s belongs to particle
    // order[s] and has nn[s] entries.  Null for tiles of consecutive
    // particles (column i is particle i, counts in the list's n_neigh).
    const uint32_t* order = nullptr;
    uint32_t size = 0;           // columns per tile (0: no tiles)
    uint32_t slots = 0;          // LDS slots the force kernel needs: largest count + 1
    uint32_t ntiles = 0;
    bool valid = false;
    };

// Builds tiles of `tile_size` (256 or 512) particles from the full list in
// particle-major rows (entry k of particle i at d_rows[i * max_neighbors + k],
// max_neighbors a multiple of 8); `pitch` is the column pitch of the tiled
// list.
// Returns AZP_OK and sets t.valid when every tile fits kTileCap - 1 distinct
// neighbors; t.valid is false otherwise (the caller keeps the plain list).
// Synchronizes `stream` (reads two flag words through `h_flag`, pinned).
int tiles_build(TileList& t,
                uint32_t tile_size,
                uint32_t N,
                const uint32_t* d_n_neigh,
                const uint32_t* d_rows,
                uint32_t pitch,
                uint32_t max_neighbors,
                uint32_t* h_flag,
                hipStream_t stream);

// Inputs and outputs of the fused walk (k_tile_walk.hip): the cell list
// (cell_of, cell_rank, start, members, the cell-ordered fixed-point copy and
// rank_of) and the list radii as in the plain walk; per-particle counts;
// 16-bit slot-major rows (entry k of slot s at rows[s * max_neighbors + k],
// LDS candidate indices) as scratch; the slot-ordered tiled list as output.
struct TileWalkArgs
    {
    const double4* pos;
    uint32_t N;
    DevBox box;
    float fx_scale[3]; // L / 2^32 per axis
    CellGrid g;
    const uint32_t* cell_of;   // geometric cell of each particle
    const uint32_t* cell_rank; // Morton rank of each geometric cell
    const uint32_t* rank_cell; // geometric cell of each Morton rank
    const uint32_t* start;     // slots of rank r: start[r] .. start[r + 1]
    const uint32_t* members;
    const int4* cell_fixed;
    const uint32_t* rank_of;
    const double* r_listsq;
    const float* r_listsq_lo;
    const float* r_listsq_hi;
    double r_list_max;
    uint32_t ntypes;
    uint32_t* n_neigh;    // by particle
    uint32_t* tile_nn;    // by column (set by tiles_walk)
    uint32_t* tile_order; // particle of each column (set by tiles_walk)
    uint32_t* tile_col;   // column of each particle (set by tiles_walk)
    uint16_t* rows;       // by slot
    uint32_t max_neighbors; // a multiple of 8
    // [0] largest count above max_neighbors, [2] bad type (1) / position (2),
    // [3] a tile did not fit (1: cells, 2: candidates, 4: neighbor set,
    // 8: own cells), [4] largest neighbor-set size, [5] tiles walked with
    // FP32 offsets from their anchor
    uint32_t* flags;
    uint32_t* tile_count;
    uint32_t* tile_jset;
    uint32_t* tlist;
    uint32_t pitch;
    };

// Enqueues the fused walk with tiles of `tile_size` (256 or 512) particles
// into t's arrays (sized here); the caller reads args.flags afterwards and
// sets t.size / t.slots / t.valid when flags[3] is clear.
int tiles_walk(TileList& t, uint32_t tile_size, const TileWalkArgs& args, hipStream_t stream);

// Most particle types tiles_walk serves.
constexpr uint32_t kTileWalkMaxTypes = 8;

// Enqueues the expansion of a valid tiled list of the per-tile walk into the
// plain column list (d_nlist[k * pitch + i], padding entries hold i).
int tiles_to_plain(const TileList& t, uint32_t N, const uint32_t* d_n_neigh, uint32_t pitch, uint32_t* d_nlist,
                   hipStream_t stream);

} // namespace azp

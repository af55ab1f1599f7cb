// Tiled neighbor list: the full list of a block of consecutive particles
// (a tile) rewritten against the tile's own set of neighbors.
//
// For tile t (particles [t T, (t + 1) T), T = 256 or 512, in space-filling
// curve order) the build collects the distinct neighbor indices of all its
// particles, sorted ascending, into jset[t][0..count[t]) and stores every
// list entry as a 16-bit index into that set, two entries per 32-bit word:
// word m of particle i (at list[m * pitch + i]) holds entries 2m (low half)
// and 2m + 1 (high half).  Columns are padded with the tile's own count
// (count[t], one past its last rank) up to the wavefront's longest list
// rounded up to 8.  The force kernel stages the tile's neighbor positions in
// LDS once (slot count[t] holds a dummy) and gathers from LDS, and the list
// stream shrinks from 4 to 2 bytes per entry.
#pragma once

#include <cstdint>

#include "azp_internal.h"

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
    uint32_t size = 0;           // particles per tile (0: no tiles)
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

} // namespace azp

// Fused neighbor walk and tile build.  Tiles are runs of T = 512 or 256
// consecutive slots of the cell list, whose cells are stored in Morton order,
// so a tile covers a compact block of space whatever order the particles
// are in; slot s holds particle members[s].  One workgroup per tile writes
// the tiled list (tiles.h) straight from the cell list:
//
//   1. union:   the distinct cells of the tile's particles and their 27
//               neighbor cells, collected in an LDS hash and sorted by Morton
//               rank; their member counts give each cell's offset in LDS,
//   2. stage:   the members' fixed-point positions (the cell-ordered copy,
//               16 bytes each, contiguous per cell) are copied into LDS, so
//               the tile's candidates are read from global memory once,
//   3. walk:    each lane tests its particle against the candidates of its
//               27 cells from LDS (the same FP32 test with the FP64 decision
//               near r_list as the plain walk, so the pair set is the same)
//               and appends the candidate's LDS index to its own row of
//               16-bit entries; then every listed candidate is marked from
//               the finished rows,
//   4. compact: an exclusive scan of the marks gives each neighbor its rank
//               in the tile's neighbor set (Morton cell order), whose
//               particle indices go to jset,
//   5. columns: each lane rereads its row and writes ranks two per word in
//               column s of the tiled list, padded with the tile's count.
//
// A tile whose cells hold more candidates than LDS does, or whose neighbor
// set exceeds kTileCap - 1, reports it; the host then uses the plain walk.
#include "cells.h"
#include "device_common.h"
#include "tiles.h"

namespace azp
{

template<uint32_t T> struct TileWalkShape;
// candidates staged per tile: 96 KB (T = 512) or 64 KB (T = 256) of LDS
template<> struct TileWalkShape<512>
    {
    static constexpr uint32_t cand = 6144;
    };
template<> struct TileWalkShape<256>
    {
    static constexpr uint32_t cand = 4096;
    };

constexpr uint32_t kUnionCells = 512; // distinct cells around one tile
constexpr uint32_t kCellHash = 1024;  // open-addressing table of those cells
// Per-lane ring of 16 list entries (32 bytes) padded to 40 bytes, so that the
// lanes' 16-bit appends fall in different LDS banks.
constexpr uint32_t kRingWords = 10;
// The tile's own cells are the cells of Morton rank first .. last of its
// slots; each gets a 27-entry table of its neighbor cells' candidate ranges.
constexpr uint32_t kOwnCells = 128;
constexpr uint32_t kNoRange = 0xffffffffu; // a repeated stencil cell (grids narrower than 3)
// Candidates tested per step of the walk (two steps' reads in flight).
constexpr uint32_t kWalkBatch = 4;
constexpr uint16_t kUsed = 0x8000u;    // mark bit beside a 15-bit rank

__device__ __forceinline__ uint32_t cell_hash(uint32_t c)
    {
    return (c * 2654435761u) >> 22; // 10 bits
    }

// Hash entries: cell index in the high word; once the offsets are known the
// low word holds the cell's candidate range in LDS (first | end << 16).
constexpr uint64_t kNoEntry = ~0ull;

// Inserts cell c; returns true when this call added it.  *full is set when
// the table has no free slot left for it.
__device__ __forceinline__ bool cell_insert(uint64_t* tab, uint32_t c, bool* full)
    {
    const uint64_t key = (uint64_t)c << 32;
    uint32_t h = cell_hash(c);
    for (uint32_t p = 0; p < kCellHash; ++p)
        {
        const uint64_t prev = atomicCAS(reinterpret_cast<unsigned long long*>(&tab[h]), kNoEntry, key);
        if (prev == kNoEntry)
            return true;
        if ((uint32_t)(prev >> 32) == c)
            return false;
        h = (h + 1u) & (kCellHash - 1u);
        }
    *full = true;
    return false;
    }

// Slot of cell c (present in the table).
__device__ __forceinline__ uint32_t cell_slot(const uint64_t* tab, uint32_t c)
    {
    uint32_t h = cell_hash(c);
    for (uint32_t p = 0; p < kCellHash; ++p)
        {
        if ((uint32_t)(tab[h] >> 32) == c)
            return h;
        h = (h + 1u) & (kCellHash - 1u);
        }
    return 0u; // not reached: every looked-up cell was inserted
    }

// Index of the union cell whose candidate range holds m (ranges ascending,
// the cell's own range non-empty): the largest u < U with off[u] <= m.
__device__ __forceinline__ uint32_t union_cell_of(const uint32_t* off, uint32_t U, uint32_t m)
    {
    uint32_t lo = 0, hi = U;
    while (hi - lo > 1u)
        {
        const uint32_t mid = (lo + hi) >> 1;
        if (off[mid] <= m)
            lo = mid;
        else
            hi = mid;
        }
    return lo;
    }

// Largest value over the 64 lanes of the wavefront (every lane gets it).
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v)
    {
    for (int off = 32; off > 0; off >>= 1)
        v = max(v, (uint32_t)__shfl_xor((int)v, off));
    return v;
    }

// Exclusive scan of one value per thread over the T-thread workgroup.
template<uint32_t T> __device__ __forceinline__ uint32_t tile_scan(uint32_t v, uint32_t* s_wave, uint32_t& total)
    {
    const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
    uint32_t x = v;
    for (uint32_t off = 1; off < 64; off <<= 1)
        {
        const uint32_t y = (uint32_t)__shfl_up((int)x, off);
        if (lane >= off)
            x += y;
        }
    if (lane == 63u)
        s_wave[w] = x;
    __syncthreads();
    uint32_t base = 0, tot = 0;
    for (uint32_t k = 0; k < T / 64; ++k)
        {
        const uint32_t sw = s_wave[k];
        base += k < w ? sw : 0u;
        tot += sw;
        }
    __syncthreads();
    total = tot;
    return base + x - v;
    }

struct TileWalkLds
    {
    const int4* cand;
    uint16_t* mark;
    const uint32_t* stencil; // 27 candidate ranges (first | end << 16) per own cell
    const uint32_t* off;
    const uint32_t* gs; // first cell-list slot of each union cell
    const float2* lohi; // r_list^2 (1 -+ 1e-5) per type pair (several types)
    uint32_t U;
    };

// the fused walk's type-pair table lives in LDS
constexpr uint32_t kTileWalkTypes = kTileWalkMaxTypes;

// How a tile's candidates are staged and tested.
enum WalkMode
    {
    kWalkFixed = 0,      // wrapped fixed-point positions, r^2 in physical units
    kWalkFixedCubic = 1, // the same, r^2 in fixed-point units (cubic box)
    kWalkRelative = 2,   // FP32 offsets from a tile anchor in physical units
    };

// Lists particle i against the staged candidates; returns its neighbor
// count.  Entries are LDS candidate indices, appended to the lane's 16-entry
// LDS ring and stored to its row eight at a time (16 bytes).  Candidates are
// read four at a time (the staging array has room for the three past a
// cell's end); the per-candidate tests are branch-free, and only the rare
// candidates near r_list take the FP64 branch.
//
// kWalkRelative: every staged position is its minimum-image offset from the
// tile's anchor in FP32 (physical units), and the lane's own position is its
// own staged entry, so a displacement is one FP32 subtraction per axis.  The
// tile uses it only when all its candidate cells lie within L/4 of the
// anchor on every axis, where the difference of two such offsets is the
// minimum image, and within 8 r_list, where the FP32 error of r^2 stays
// inside the 1e-5 band.
template<bool SINGLE, int MODE>
__device__ uint32_t tile_walk_one(const TileWalkArgs& a,
                                  const TileWalkLds& s,
                                  uint32_t slot,
                                  uint32_t i,
                                  uint32_t own,
                                  uint32_t* ring)
    {
    const double4 pi = a.pos[i];
    const uint32_t ti = checked_type(pi.w, a.ntypes);
    const bool finite = isfinite(pi.x) && isfinite(pi.y) && isfinite(pi.z);
    if (ti == a.ntypes || !finite)
        {
        atomicOr(a.flags + 2, finite ? 1u : 2u);
        return 0;
        }
    constexpr bool CUBIC = MODE == kWalkFixedCubic;
    constexpr bool REL = MODE == kWalkRelative;
    const int32_t fxi = to_fixed(pi.x, a.box.Linv[0]);
    const int32_t fyi = to_fixed(pi.y, a.box.Linv[1]);
    const int32_t fzi = to_fixed(pi.z, a.box.Linv[2]);
    const float sx = a.fx_scale[0], sy = a.fx_scale[1], sz = a.fx_scale[2];
    const float unit2 = CUBIC ? 1.0f / (sx * sx) : 1.0f; // r^2 -> fixed-point units^2
    // the lane's 27 neighbor cells in stencil_axis order; the own cell is the
    // entry with offset 0 on every axis: the middle one, or the first on an
    // axis one cell wide (a box of exactly 2 r_list), whose other entries
    // repeat it and are empty
    const uint32_t* stencil = s.stencil + own * 27u;
    const uint32_t own_k = (a.g.n[0] > 1 ? 1u : 0u) + (a.g.n[1] > 1 ? 3u : 0u) + (a.g.n[2] > 1 ? 9u : 0u);
    const uint32_t own_first = stencil[own_k] & 0xffffu;
    const uint32_t self = own_first + (slot - s.gs[union_cell_of(s.off, s.U, own_first)]);
    const int4 me = s.cand[self];
    const float rxi = __int_as_float(me.x), ryi = __int_as_float(me.y), rzi = __int_as_float(me.z);
    const double rl0 = a.r_listsq[0];
    // wave-uniform, held in scalar registers: the candidate loop then never
    // waits on the vector memory counter (which the row stores also use)
    const float lo0 = __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, a.r_listsq_lo[0])))
                      * unit2;
    const float hi0 = __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, a.r_listsq_hi[0])))
                      * unit2;
    uint16_t* ring16 = reinterpret_cast<uint16_t*>(ring);
    uint16_t* row = a.rows + (size_t)slot * a.max_neighbors;
    uint32_t count = 0;
    // a completed group read from the ring, not yet stored to the row
    bool pend = false;
    uint4 pend_data = make_uint4(0u, 0u, 0u, 0u);
    uint32_t pend_at = 0;
    auto store_pending = [&]() {
        if (pend)
            *reinterpret_cast<uint4*>(row + pend_at) = pend_data;
        pend = false;
    };
    // the next cell's range and first candidates are read one step ahead
    // one batch: four candidates from m0 (entries past the cell's end masked).
    // The flags are named scalars, not arrays: an array of flags becomes a
    // packed byte vector in VGPRs, while scalars stay lane masks.
    static_assert(kWalkBatch == 4, "the batch below is written out for four candidates");
    // CHECKED false: the wavefront knows that every lane's four candidates lie
    // inside its cell and that the cell is not the lane's own (no range or
    // self test per candidate)
    auto batch = [&](const int4 (&pc)[kWalkBatch], uint32_t m0, uint32_t e, auto checked) {
        constexpr bool CHECKED = decltype(checked)::value;
        // offsets from m0: candidate q is in the range when q < rem and is
        // the lane's own particle when q == dself (which wraps to a large
        // value when self < m0)
        const uint32_t rem = e - m0;
        const uint32_t dself = self - m0;
        auto test = [&](const int4& c, uint32_t q, bool& hit, bool& band) {
            const uint32_t tj = (uint32_t)c.w;
            // with one type, a candidate of an invalid type can only exist
            // when its own lane reports it, which fails the build
            const bool valid = (!CHECKED || ((q < rem) & (q != dself))) & (SINGLE || tj < a.ntypes);
            const uint32_t tp = SINGLE ? 0u : ti * a.ntypes + min(tj, a.ntypes - 1);
            const float lo = SINGLE ? lo0 : s.lohi[tp].x * unit2;
            const float hi = SINGLE ? hi0 : s.lohi[tp].y * unit2;
            float dx, dy, dz;
            if constexpr (REL)
                {
                dx = __int_as_float(c.x) - rxi;
                dy = __int_as_float(c.y) - ryi;
                dz = __int_as_float(c.z) - rzi;
                }
            else
                {
                dx = (float)fixed_delta(fxi, c.x) * (CUBIC ? 1.0f : sx);
                dy = (float)fixed_delta(fyi, c.y) * (CUBIC ? 1.0f : sy);
                dz = (float)fixed_delta(fzi, c.z) * (CUBIC ? 1.0f : sz);
                }
            // FP32 rounding here is far inside the 1e-5 band
            const float rsq = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
            const bool below = rsq < lo;
            hit = valid & below;
            band = valid & !below & (rsq <= hi);
        };
        // The FP64 predicate of the host oracle for a candidate FP32 cannot
        // decide; only the loads are lane-divergent.
        auto retest = [&](const int4& c, uint32_t q, bool band) {
            double rsq64 = 0.0, rl = 0.0;
            if (band)
                {
                const uint32_t m = m0 + q;
                const uint32_t u = union_cell_of(s.off, s.U, m);
                const uint32_t j = a.members[s.gs[u] + (m - s.off[u])];
                const double4 p64 = a.pos[j];
                const uint32_t tp = SINGLE ? 0u : ti * a.ntypes + (uint32_t)c.w;
                rl = SINGLE ? rl0 : a.r_listsq[tp];
                const double qx = min_image_1d(pi.x - p64.x, a.box.L[0], a.box.Linv[0]);
                const double qy = min_image_1d(pi.y - p64.y, a.box.L[1], a.box.Linv[1]);
                const double qz = min_image_1d(pi.z - p64.z, a.box.L[2], a.box.Linv[2]);
                rsq64 = rsq3_nofma(qx, qy, qz);
                }
            return band & (rsq64 <= rl);
        };
        bool h0, h1, h2, h3, b0, b1, b2, b3;
        test(pc[0], 0u, h0, b0);
        test(pc[1], 1u, h1, b1);
        test(pc[2], 2u, h2, b2);
        test(pc[3], 3u, h3, b3);
        // taken by the whole wavefront when any lane has such a candidate
        if (__ballot(b0 | b1 | b2 | b3) != 0ull)
            {
            h0 = h0 | retest(pc[0], 0u, b0);
            h1 = h1 | retest(pc[1], 1u, b1);
            h2 = h2 | retest(pc[2], 2u, b2);
            h3 = h3 | retest(pc[3], 3u, b3);
            }
        // Append the hits in candidate order.  Every candidate is written to
        // the ring slot after the last hit; a miss is overwritten by the next
        // candidate, so the ring needs no per-candidate branch.
        uint32_t pos = count;
        auto put = [&](uint32_t q, bool h) {
            ring16[pos & 15u] = (uint16_t)(m0 + q);
            pos += h ? 1u : 0u;
        };
        put(0u, h0);
        put(1u, h1);
        put(2u, h2);
        put(3u, h3);
        // A completed group of eight leaves as one 16-byte store.  The ring
        // is read here, the store is issued after the next batch's reads
        // (store_pending): waiting for this read right away would also wait
        // for every LDS access issued before it.
        if (((pos ^ count) & ~7u) != 0u && count < a.max_neighbors)
            {
            const uint2* h = reinterpret_cast<const uint2*>(ring + ((count >> 1) & 4u));
            const uint2 w0 = h[0], w1 = h[1];
            pend_data = make_uint4(w0.x, w0.y, w1.x, w1.y);
            pend_at = count & ~7u;
            pend = true;
            }
        count = pos;
    };
    auto load_batch = [&](int4 (&pc)[kWalkBatch], uint32_t m0) {
#pragma unroll
        for (uint32_t q = 0; q < kWalkBatch; ++q)
            pc[q] = s.cand[m0 + q];
    };
    // Two batches per iteration in two register sets: the reads of the next
    // batch are issued before the current one is tested (the scheduling
    // barriers keep them there), so LDS latency overlaps the tests.
    uint32_t range_next = stencil[0];
    for (uint32_t k = 0; k < 27u; ++k)
        {
        const uint32_t range = range_next;
        range_next = stencil[min(k + 1u, 26u)];
        const uint32_t b = range & 0xffffu, e = range >> 16;
        int4 pa[kWalkBatch], pb[kWalkBatch];
        load_batch(pa, b);
        for (uint32_t m0 = b; m0 < e; m0 += 2 * kWalkBatch)
            {
            load_batch(pb, m0 + kWalkBatch);
            __builtin_amdgcn_sched_barrier(0);
            store_pending();
            // both batches whole for every lane still in its cell, and not
            // the own cell: the tests need no masks
            if (k != own_k && __ballot(m0 + 2 * kWalkBatch > e) == 0ull)
                {
                batch(pa, m0, e, std::false_type{});
                load_batch(pa, m0 + 2 * kWalkBatch);
                __builtin_amdgcn_sched_barrier(0);
                store_pending();
                batch(pb, m0 + kWalkBatch, e, std::false_type{});
                continue;
                }
            batch(pa, m0, e, std::true_type{});
            if (m0 + kWalkBatch < e)
                {
                load_batch(pa, m0 + 2 * kWalkBatch);
                __builtin_amdgcn_sched_barrier(0);
                store_pending();
                batch(pb, m0 + kWalkBatch, e, std::true_type{});
                }
            }
        }
    store_pending();
    if ((count & 7u) != 0u && count < a.max_neighbors)
        {
        const uint2* h = reinterpret_cast<const uint2*>(ring + ((count >> 1) & 4u));
        const uint2 w0 = h[0], w1 = h[1];
        *reinterpret_cast<uint4*>(row + (count & ~7u)) = make_uint4(w0.x, w0.y, w1.x, w1.y);
        }
    return count;
    }

template<uint32_t T, bool SINGLE> __global__ void __launch_bounds__(T) k_tile_walk(const TileWalkArgs a)
    {
    constexpr uint32_t C = TileWalkShape<T>::cand;
    __shared__ int4 s_cand[C + 3 * kWalkBatch];
    __shared__ uint16_t s_mark[C];
    __shared__ uint64_t s_tab[kCellHash];
    __shared__ uint64_t s_ucell[kUnionCells];
    __shared__ uint32_t s_off[kUnionCells + 1];
    __shared__ uint32_t s_gs[kUnionCells];
    __shared__ uint32_t s_ring[T * kRingWords];
    __shared__ uint32_t s_stencil[kOwnCells * 27];
    __shared__ uint32_t s_wave[T / 64];
    __shared__ float2 s_lohi[SINGLE ? 1 : kTileWalkTypes * kTileWalkTypes];
    __shared__ uint32_t s_n, s_bad, s_rfirst, s_rlast, s_ext[3];

    const uint32_t tid = threadIdx.x, lane = tid & 63u;
    const uint32_t t = xcd_logical_block();
    const uint32_t slot = t * T + tid;
    const bool active = slot < a.N;
    const uint32_t nx = (uint32_t)a.g.n[0], ny = (uint32_t)a.g.n[1], nz = (uint32_t)a.g.n[2];

    for (uint32_t k = tid; k < kCellHash; k += T)
        s_tab[k] = kNoEntry;
    if (!SINGLE)
        for (uint32_t k = tid; k < a.ntypes * a.ntypes; k += T)
            s_lohi[k] = make_float2(a.r_listsq_lo[k], a.r_listsq_hi[k]);
    if (tid == 0)
        {
        s_n = 0u;
        s_bad = 0u;
        s_ext[0] = s_ext[1] = s_ext[2] = 0u;
        }
    __syncthreads();

    // 1. the tile's own cells: Morton ranks first .. last of its slots
    const uint32_t i = active ? a.members[slot] : 0u;
    const uint32_t c = active ? a.cell_of[i] : 0u;
    const uint32_t crank = active ? a.cell_rank[c] : 0u;
    if (tid == 0)
        s_rfirst = crank;
    if (active && (tid == T - 1 || slot == a.N - 1))
        s_rlast = crank;
    __syncthreads();
    const uint32_t r_first = s_rfirst;
    const uint32_t D = s_rlast - r_first + 1u;
    if (D > kOwnCells)
        {
        if (tid == 0)
            atomicOr(a.flags + 3, 8u);
        return; // D is the same for every thread
        }
    // 2. their neighbor cells (periodic) make the union; the stencil table
    // keeps each neighbor's cell index until the ranges are known, and marks
    // the offsets that repeat a cell on grids narrower than 3
    for (uint32_t q = tid; q < D * 27u; q += T)
        {
        const uint32_t oc = a.rank_cell[r_first + q / 27u];
        const uint32_t sidx = q % 27u;
        const uint32_t dx = sidx % 3u, dy = (sidx / 3u) % 3u, dz = sidx / 9u; // offset + 1
        if (dx > min(nx, 3u) - 1u || dy > min(ny, 3u) - 1u || dz > min(nz, 3u) - 1u)
            {
            s_stencil[q] = kNoRange;
            continue;
            }
        const uint32_t ox = oc % nx, oy = (oc / nx) % ny, oz = oc / (nx * ny);
        const uint32_t x = (ox + nx + dx - 1u) % nx;
        const uint32_t y = (oy + ny + dy - 1u) % ny;
        const uint32_t z = (oz + nz + dz - 1u) % nz;
        const uint32_t nb = (z * ny + y) * nx + x;
        s_stencil[q] = nb;
        bool full = false;
        if (cell_insert(s_tab, nb, &full))
            {
            const uint32_t k = atomicAdd(&s_n, 1u);
            if (k < kUnionCells)
                s_ucell[k] = ((uint64_t)a.cell_rank[nb] << 32) | nb;
            }
        if (full)
            s_bad = 1u;
        }
    __syncthreads();
    const uint32_t U = s_n;
    if (s_bad != 0u || U > kUnionCells)
        {
        if (tid == 0)
            atomicOr(a.flags + 3, 1u);
        return; // the whole workgroup leaves (U and s_bad are shared)
        }

    // Morton order: bitonic sort over the next power of two
    uint32_t P2 = 1;
    while (P2 < U)
        P2 <<= 1;
    for (uint32_t k = U + tid; k < P2; k += T)
        s_ucell[k] = ~0ull;
    __syncthreads();
    if (P2 <= T)
        {
        // one key per thread; stages with a stride below 64 exchange keys
        // within the wavefront (no barrier), the others through LDS
        uint64_t key = tid < P2 ? s_ucell[tid] : ~0ull;
        for (uint32_t size = 2; size <= P2; size <<= 1)
            for (uint32_t stride = size >> 1; stride > 0; stride >>= 1)
                {
                uint64_t other;
                if (stride >= 64u)
                    {
                    __syncthreads();
                    if (tid < P2)
                        s_ucell[tid] = key;
                    __syncthreads();
                    other = tid < P2 ? s_ucell[tid ^ stride] : ~0ull;
                    }
                else
                    {
                    const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)key, (int)stride);
                    const uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(key >> 32), (int)stride);
                    other = ((uint64_t)hi << 32) | lo;
                    }
                // ascending blocks keep the smaller key at the lower position
                const bool lower = (tid & stride) == 0u, up = (tid & size) == 0u;
                key = (lower == up) ? min(key, other) : max(key, other);
                }
        __syncthreads();
        if (tid < P2)
            s_ucell[tid] = key;
        __syncthreads();
        }
    else
        {
        for (uint32_t size = 2; size <= P2; size <<= 1)
            for (uint32_t stride = size >> 1; stride > 0; stride >>= 1)
                {
                for (uint32_t e = tid; e < P2; e += T)
                    {
                    const uint32_t p = e ^ stride;
                    if (p > e)
                        {
                        const bool up = (e & size) == 0u;
                        const uint64_t x = s_ucell[e], y = s_ucell[p];
                        if ((x > y) == up)
                            {
                            s_ucell[e] = y;
                            s_ucell[p] = x;
                            }
                        }
                    }
                __syncthreads();
                }
        }

    // 2. LDS offsets of the union cells' members; the anchor of the FP32
    // offsets is the tile's middle particle
    const uint32_t mid = t * T + (min(T, a.N - t * T) >> 1);
    const uint32_t anchor_cell = a.cell_of[a.members[mid]];
    constexpr uint32_t QC = kUnionCells / T;
    uint32_t cnt[QC];
    uint32_t sum = 0;
#pragma unroll
    for (uint32_t q = 0; q < QC; ++q)
        {
        const uint32_t u = tid * QC + q;
        cnt[q] = 0u;
        if (u < U)
            {
            const uint32_t r = (uint32_t)(s_ucell[u] >> 32);
            const uint32_t b = a.start[r];
            s_gs[u] = b;
            cnt[q] = a.start[r + 1] - b;
            // periodic distance in cells from the anchor's cell, per axis
            const uint32_t cu = (uint32_t)s_ucell[u], ca = anchor_cell;
            const uint32_t d[3] = {cu % nx > ca % nx ? cu % nx - ca % nx : ca % nx - cu % nx,
                                   (cu / nx) % ny > (ca / nx) % ny ? (cu / nx) % ny - (ca / nx) % ny
                                                                   : (ca / nx) % ny - (cu / nx) % ny,
                                   cu / (nx * ny) > ca / (nx * ny) ? cu / (nx * ny) - ca / (nx * ny)
                                                                   : ca / (nx * ny) - cu / (nx * ny)};
            atomicMax(&s_ext[0], min(d[0], nx - d[0]));
            atomicMax(&s_ext[1], min(d[1], ny - d[1]));
            atomicMax(&s_ext[2], min(d[2], nz - d[2]));
            }
        sum += cnt[q];
        }
    uint32_t Mc;
    uint32_t run = tile_scan<T>(sum, s_wave, Mc);
#pragma unroll
    for (uint32_t q = 0; q < QC; ++q)
        {
        const uint32_t u = tid * QC + q;
        if (u < U)
            {
            const uint32_t cu = (uint32_t)s_ucell[u];
            s_off[u] = run;
            // candidate range in the hash entry (Mc <= C < 2^16 is checked below
            // before any lookup)
            s_tab[cell_slot(s_tab, cu)] = ((uint64_t)cu << 32) | (run & 0xffffu) | ((run + cnt[q]) << 16);
            }
        run += cnt[q];
        }
    if (tid == 0)
        s_off[U] = Mc;
    if (Mc > C)
        {
        if (tid == 0)
            atomicOr(a.flags + 3, 2u);
        return; // Mc is the same for every thread
        }
    __syncthreads();
    // FP32 offsets from the anchor (the tile's middle particle) when every
    // candidate cell is within L/4 of the anchor's cell on every axis (so
    // differences of offsets are minimum images) and within 8 r_list (so the
    // FP32 error of r^2, below 6e-8 (13.9 R/r + 6.5) r^2 for offsets up to R,
    // stays under the 1e-5 band)
    const double reach = 8.0 * a.r_list_max;
    const bool rel = (s_ext[0] + 1u) * 4u < nx && (s_ext[1] + 1u) * 4u < ny && (s_ext[2] + 1u) * 4u < nz
                     && (s_ext[0] + 1u) * a.box.L[0] <= reach * nx && (s_ext[1] + 1u) * a.box.L[1] <= reach * ny
                     && (s_ext[2] + 1u) * a.box.L[2] <= reach * nz;
    const int4 anchor = a.cell_fixed[mid];
    const float sx = a.fx_scale[0], sy = a.fx_scale[1], sz = a.fx_scale[2];
    auto staged = [&](int4 v) {
        return rel ? make_int4(__float_as_int((float)fixed_delta(v.x, anchor.x) * sx),
                               __float_as_int((float)fixed_delta(v.y, anchor.y) * sy),
                               __float_as_int((float)fixed_delta(v.z, anchor.z) * sz), v.w)
                   : v;
    };
    // stage the candidates: one wavefront per cell, four cells' loads in
    // flight at a time (cells past 64 members finish in a second loop)
    constexpr uint32_t NW = T / 64;
    auto stage_load = [&](uint32_t u) {
        return (u < U && lane < s_off[u + 1] - s_off[u]) ? a.cell_fixed[s_gs[u] + lane] : make_int4(0, 0, 0, 0);
    };
    auto stage_store = [&](uint32_t u, int4 v) {
        if (u >= U)
            return;
        const uint32_t o = s_off[u], n = s_off[u + 1] - o;
        if (lane < n)
            s_cand[o + lane] = staged(v);
        for (uint32_t k = lane + 64; k < n; k += 64)
            s_cand[o + k] = staged(a.cell_fixed[s_gs[u] + k]);
    };
    for (uint32_t u0 = tid >> 6; u0 < U; u0 += 4 * NW)
        {
        const int4 v0 = stage_load(u0), v1 = stage_load(u0 + NW);
        const int4 v2 = stage_load(u0 + 2 * NW), v3 = stage_load(u0 + 3 * NW);
        stage_store(u0, v0);
        stage_store(u0 + NW, v1);
        stage_store(u0 + 2 * NW, v2);
        stage_store(u0 + 3 * NW, v3);
        }
    for (uint32_t m = tid; m < Mc; m += T)
        s_mark[m] = 0u;
    for (uint32_t q = tid; q < D * 27u; q += T)
        {
        const uint32_t nb = s_stencil[q];
        s_stencil[q] = nb == kNoRange ? 0u : (uint32_t)s_tab[cell_slot(s_tab, nb)];
        }
    __syncthreads();

    // 3. walk
    uint32_t count = 0;
    if (active)
        {
        const TileWalkLds s = {s_cand, s_mark, s_stencil, s_off, s_gs, s_lohi, U};
        uint32_t* ring = s_ring + tid * kRingWords;
        const bool cubic = a.fx_scale[0] == a.fx_scale[1] && a.fx_scale[0] == a.fx_scale[2];
        const uint32_t own = crank - r_first;
        count = rel     ? tile_walk_one<SINGLE, kWalkRelative>(a, s, slot, i, own, ring)
                : cubic ? tile_walk_one<SINGLE, kWalkFixedCubic>(a, s, slot, i, own, ring)
                        : tile_walk_one<SINGLE, kWalkFixed>(a, s, slot, i, own, ring);
        a.n_neigh[i] = min(count, a.max_neighbors);
        if (count > a.max_neighbors)
            atomicMax(a.flags, count);
        // mark the listed candidates, reading the row back in 16-byte chunks
        // (once per entry here instead of once per tested candidate above)
        const uint32_t nl = min(count, a.max_neighbors);
        const uint4* row = reinterpret_cast<const uint4*>(a.rows + (size_t)slot * a.max_neighbors);
        for (uint32_t k0 = 0; k0 < nl; k0 += 8)
            {
            const uint4 ch = row[k0 / 8];
            const uint32_t w4[4] = {ch.x, ch.y, ch.z, ch.w};
#pragma unroll
            for (uint32_t w = 0; w < 4; ++w)
                {
                if (k0 + 2 * w < nl)
                    s_mark[w4[w] & 0xffffu] = 1u;
                if (k0 + 2 * w + 1 < nl)
                    s_mark[w4[w] >> 16] = 1u;
                }
            }
        }
    __syncthreads();

    // 4. ranks of the listed candidates: in-place exclusive scan of the marks
    constexpr uint32_t QM = C / T;
    uint32_t msum = 0;
#pragma unroll
    for (uint32_t q = 0; q < QM; ++q)
        {
        const uint32_t m = tid * QM + q;
        msum += (m < Mc && s_mark[m] != 0u) ? 1u : 0u;
        }
    uint32_t M;
    uint32_t r = tile_scan<T>(msum, s_wave, M);
#pragma unroll
    for (uint32_t q = 0; q < QM; ++q)
        {
        const uint32_t m = tid * QM + q;
        if (m < Mc)
            {
            const bool used = s_mark[m] != 0u;
            s_mark[m] = (uint16_t)(used ? (kUsed | r) : r);
            r += used ? 1u : 0u;
            }
        }
    if (M > kTileCap - 1)
        {
        if (tid == 0)
            atomicOr(a.flags + 3, 4u);
        return; // M is the same for every thread
        }
    __syncthreads();
    for (uint32_t m = tid; m < Mc; m += T)
        {
        const uint32_t v = s_mark[m];
        if (v & kUsed)
            {
            const uint32_t lo = union_cell_of(s_off, U, m);
            a.tile_jset[(size_t)t * kTileCap + (v & ~(uint32_t)kUsed)] = a.members[s_gs[lo] + (m - s_off[lo])];
            }
        }
    if (tid == 0)
        {
        a.tile_count[t] = M;
        atomicMax(a.flags + 4, M);
        if (rel)
            atomicAdd(a.flags + 5, 1u);
        }

    // 5. columns ordered by list length (longest first, ties by slot), so the
    // pair kernel's 64-lane wavefronts run lists of nearly equal length: a
    // bitonic sort of (length, slot) keys, one per thread, whose stages with
    // a stride below 64 exchange keys within the wavefront (no barrier); the
    // staging space (free now) carries the others and the slots' particles
    // and lengths
    const uint32_t nn = active ? min(count, a.max_neighbors) : 0u;
    uint32_t* s_sort = reinterpret_cast<uint32_t*>(s_cand);
    uint32_t* s_who = s_sort + T;
    uint32_t* s_len = s_who + T;
    s_who[tid] = i;
    s_len[tid] = nn;
    uint32_t key = (min(nn, 0xffffu) << 16) | (0xffffu - tid);
    for (uint32_t size = 2; size <= T; size <<= 1)
        for (uint32_t stride = size >> 1; stride > 0; stride >>= 1)
            {
            uint32_t other;
            if (stride >= 64u)
                {
                __syncthreads();
                s_sort[tid] = key;
                __syncthreads();
                other = s_sort[tid ^ stride];
                }
            else
                other = (uint32_t)__shfl_xor((int)key, (int)stride);
            // the lower position of a pair keeps the larger key in a
            // descending block, the smaller in an ascending one
            const bool lower = (tid & stride) == 0u, down = (tid & size) == 0u;
            key = (lower == down) ? max(key, other) : min(key, other);
            }
    // column t T + tid holds the slot of the tid-th key: its writer is this
    // thread, so consecutive lanes store consecutive columns
    const uint32_t cslot = 0xffffu - (key & 0xffffu);
    __syncthreads();
    const uint32_t ci = s_who[cslot], cnn = s_len[cslot];
    const bool cactive = t * T + cslot < a.N;
    // the column wavefront's longest list, padded to whole steps of 8
    const uint32_t P = (wave_max_u32(cactive ? cnn : 0u) + 7u) & ~7u;
    if (!cactive)
        return;
    const uint32_t col = t * T + tid;
    a.tile_nn[col] = cnn;
    a.tile_order[col] = ci;
    a.tile_col[ci] = col;
    // the row is reread in 16-byte chunks (8 entries), four in flight
    const uint4* row = reinterpret_cast<const uint4*>(a.rows + (size_t)(t * T + cslot) * a.max_neighbors);
    auto chunk_load = [&](uint32_t kc) { return kc < cnn ? row[kc / 8] : make_uint4(0u, 0u, 0u, 0u); };
    auto rank_of_entry = [&](uint32_t k, uint32_t word, uint32_t half) {
        const uint32_t m = (word >> (16u * half)) & 0xffffu;
        return (k < cnn ? (uint32_t)(s_mark[m] & ~kUsed) : M) << 4; // stored as 16 x rank
    };
    auto chunk_store = [&](uint32_t kc, uint4 ch) {
        if (kc >= P)
            return;
        const uint32_t w4[4] = {ch.x, ch.y, ch.z, ch.w};
#pragma unroll
        for (uint32_t w = 0; w < 4; ++w)
            a.tlist[(size_t)((kc >> 1) + w) * a.pitch + col] =
                rank_of_entry(kc + 2 * w, w4[w], 0u) | (rank_of_entry(kc + 2 * w + 1, w4[w], 1u) << 16);
    };
    for (uint32_t k0 = 0; k0 < P; k0 += 32)
        {
        const uint4 c0 = chunk_load(k0), c1 = chunk_load(k0 + 8);
        const uint4 c2 = chunk_load(k0 + 16), c3 = chunk_load(k0 + 24);
        chunk_store(k0, c0);
        chunk_store(k0 + 8, c1);
        chunk_store(k0 + 16, c2);
        chunk_store(k0 + 24, c3);
        }
    }

// Plain column list from a tiled list of the per-tile walk: lane i (a
// particle) reads its column tile_col[i]; entry k is jset[tile][rank],
// padding entries hold i up to the longest list of the 64 particles rounded
// up to 8.
__global__ void __launch_bounds__(64) k_tiles_to_plain(uint32_t N,
                                                       const uint32_t* __restrict__ n_neigh,
                                                       const uint32_t* __restrict__ tile_col,
                                                       uint32_t tile_size,
                                                       const uint32_t* __restrict__ tile_jset,
                                                       const uint32_t* __restrict__ tlist,
                                                       uint32_t pitch,
                                                       uint32_t* __restrict__ nlist)
    {
    const uint32_t i = blockIdx.x * 64u + threadIdx.x;
    const bool active = i < N;
    const uint32_t nn = active ? n_neigh[i] : 0u;
    uint32_t wmax = nn;
    for (int off = 32; off > 0; off >>= 1)
        wmax = max(wmax, (uint32_t)__shfl_xor((int)wmax, off));
    const uint32_t P = (wmax + 7u) & ~7u;
    if (!active)
        return;
    const uint32_t s = tile_col[i];
    const uint32_t* js = tile_jset + (size_t)(s / tile_size) * kTileCap;
    for (uint32_t k = 0; k < P; k += 2)
        {
        uint32_t w = 0;
        if (k < nn)
            w = tlist[(size_t)(k >> 1) * pitch + s];
        nlist[(size_t)k * pitch + i] = k < nn ? js[(w & 0xffffu) >> 4] : i;
        nlist[(size_t)(k + 1) * pitch + i] = k + 1 < nn ? js[w >> 20] : i;
        }
    }

int tiles_walk(TileList& t, uint32_t tile_size, const TileWalkArgs& args, hipStream_t s)
    {
    if (tile_size != 256 && tile_size != 512)
        return set_error(AZP_ERR_INVALID, "tile size %u must be 256 or 512", tile_size);
    if (args.ntypes == 0 || args.ntypes > kTileWalkTypes)
        return set_error(AZP_ERR_UNSUPPORTED, "the per-tile walk serves 1 to %u particle types", kTileWalkTypes);
    const uint32_t ntiles = (args.N + tile_size - 1) / tile_size;
    const uint32_t words = (((args.max_neighbors + 7u) & ~7u) + 1u) / 2u;
    AZP_TRY(t.count.reserve(ntiles));
    AZP_TRY(t.jset.reserve((size_t)ntiles * kTileCap));
    AZP_TRY(t.list.reserve((size_t)words * args.pitch));
    AZP_TRY(t.nn.reserve(args.N));
    AZP_TRY(t.order_buf.reserve(args.N));
    AZP_TRY(t.col.reserve(args.N));
    TileWalkArgs a = args;
    a.tile_nn = t.nn.ptr;
    a.tile_order = t.order_buf.ptr;
    a.tile_col = t.col.ptr;
    a.tile_count = t.count.ptr;
    a.tile_jset = t.jset.ptr;
    a.tlist = t.list.ptr;
    const bool single = a.ntypes == 1;
    if (tile_size == 512)
        {
        if (single)
            hipLaunchKernelGGL((k_tile_walk<512, true>), dim3(ntiles), dim3(512), 0, s, a);
        else
            hipLaunchKernelGGL((k_tile_walk<512, false>), dim3(ntiles), dim3(512), 0, s, a);
        }
    else
        {
        if (single)
            hipLaunchKernelGGL((k_tile_walk<256, true>), dim3(ntiles), dim3(256), 0, s, a);
        else
            hipLaunchKernelGGL((k_tile_walk<256, false>), dim3(ntiles), dim3(256), 0, s, a);
        }
    AZP_LAUNCH_CHECK();
    t.ntiles = ntiles;
    t.order = t.order_buf.ptr;
    return AZP_OK;
    }

int tiles_to_plain(const TileList& t, uint32_t N, const uint32_t* d_n_neigh, uint32_t pitch, uint32_t* d_nlist,
                   hipStream_t s)
    {
    if (N == 0)
        return AZP_OK;
    if (!t.valid || !t.order)
        return set_error(AZP_ERR_STATE, "no tiled list of the per-tile walk to expand");
    hipLaunchKernelGGL(k_tiles_to_plain, dim3((N + 63) / 64), dim3(64), 0, s, N, d_n_neigh, t.col.ptr, t.size,
                       t.jset.ptr, t.list.ptr, pitch, d_nlist);
    AZP_LAUNCH_CHECK();
    return AZP_OK;
    }

} // namespace azp

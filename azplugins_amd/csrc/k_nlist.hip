// Cell list and full neighbor list on gfx950.
//
// Replaces the upstream CellListGPU + NeighborListGPUBinned pair that the
// azplugins pair classes reach through their `nlist` argument (pair.py:110-111).
//
// Build (all kernels, no host round trip except the overflow / rebuild flags):
//   1. bin:     cell of each particle, atomically counted per cell
//   2. scan:    exclusive prefix sum of the counts (three launches)
//   3. scatter: members written into per-cell ranges, then each cell's members
//               sorted by index so the list order is deterministic
//   4. walk:    one lane per particle scans its 27 neighbor cells (duplicates
//               removed for grids narrower than 3 cells) over a cell-ordered
//               fixed-point copy of the positions and appends j with
//               rsq <= r_list(ti,tj)^2 — FP32 on exact integer displacements,
//               with every candidate near r_list decided by the same FP64
//               predicate, evaluated without contraction, as the host oracle,
//               so the pair sets agree exactly.  Each lane appends to its own
//               particle-major row (its successive 4-byte writes fill whole
//               lines; scattered column writes cost 8x in DRAM traffic).
//               Lanes that find more than max_neighbors entries keep counting
//               and report the largest count; the host grows the list and
//               reruns the walk.
//   5. columns: rows transposed to nlist[k * pitch + i] (coalesced for
//               one-lane-per-particle consumers), padded per 64 particles.
// With tiles enabled, steps 4 and 5 are replaced by the fused per-tile walk
// (k_tile_walk.hip), which writes the tiled list directly; the plain list is
// expanded from it only when a consumer asks for it.
// Rebuild check: a particle moved farther than r_buff/2 since the last build.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <vector>

#include "cells.h"
#include "step.h"
#include "tiles.h"

namespace azp
{

__global__ void k_fill_u32(uint32_t* p, uint32_t v, size_t n)
    {
    const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n)
        p[k] = v;
    }

// cell_of[i] is the geometric cell index; the cell's members occupy slots
// start[cell_rank[c]] .. start[cell_rank[c] + 1] (Morton order of cells).
// The count is aggregated per wavefront: one atomic per distinct cell among
// its 64 particles (neighbors in memory usually share a cell).
__global__ void k_bin(const double4* __restrict__ pos,
                      uint32_t N,
                      DevBox box,
                      CellGrid g,
                      const uint32_t* __restrict__ cell_rank,
                      uint32_t* __restrict__ cell_of,
                      uint32_t* __restrict__ slot_of,
                      uint32_t* __restrict__ count)
    {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t lane = threadIdx.x & 63u;
    const bool active = i < N;
    uint32_t r = 0xffffffffu;
    if (active)
        {
        const double4 p = pos[i];
        const int cx = cell_coord(p.x, box.L[0], box.Linv[0], g.n[0]);
        const int cy = cell_coord(p.y, box.L[1], box.Linv[1], g.n[1]);
        const int cz = cell_coord(p.z, box.L[2], box.Linv[2], g.n[2]);
        const uint32_t c = (uint32_t)((cz * g.n[1] + cy) * g.n[0] + cx);
        cell_of[i] = c;
        r = cell_rank[c];
        }
    uint64_t pending = __ballot(active);
    while (pending != 0ull)
        {
        const uint32_t leader = (uint32_t)__ffsll((unsigned long long)pending) - 1u;
        const uint32_t lr = (uint32_t)__shfl((int)r, (int)leader);
        const uint64_t same = __ballot(active && r == lr);
        uint32_t base = 0;
        if (lane == leader)
            base = atomicAdd(&count[lr], (uint32_t)__popcll(same));
        base = (uint32_t)__shfl((int)base, (int)leader);
        if (active && r == lr)
            slot_of[i] = base + (uint32_t)__popcll(same & ((1ull << lane) - 1ull));
        pending &= ~same;
        }
    }

__global__ void k_scatter(uint32_t N,
                          const uint32_t* __restrict__ cell_of,
                          const uint32_t* __restrict__ cell_rank,
                          const uint32_t* __restrict__ slot_of,
                          const uint32_t* __restrict__ start,
                          uint32_t* __restrict__ members)
    {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N)
        return;
    members[start[cell_rank[cell_of[i]]] + slot_of[i]] = i;
    }

// Each cell's members sorted by particle index, so that the lists do not
// depend on the order the binning atomics took: one wavefront per cell ranks
// up to 64 members by counting the smaller ones; a longer cell is
// insertion-sorted by its first lane.
__global__ void __launch_bounds__(256) k_sort_cells(uint32_t ncell,
                                                    const uint32_t* __restrict__ start,
                                                    uint32_t* __restrict__ members)
    {
    const uint32_t c = blockIdx.x * 4u + (threadIdx.x >> 6);
    const uint32_t lane = threadIdx.x & 63u;
    if (c >= ncell)
        return;
    const uint32_t b = start[c], e = start[c + 1], n = e - b;
    if (n <= 64u)
        {
        const uint32_t v = lane < n ? members[b + lane] : 0xffffffffu;
        uint32_t r = 0;
        for (uint32_t k = 0; k < n; ++k)
            r += (uint32_t)__shfl((int)v, (int)k) < v ? 1u : 0u;
        if (lane < n)
            members[b + r] = v;
        return;
        }
    if (lane != 0)
        return;
    for (uint32_t k = b + 1; k < e; ++k)
        {
        const uint32_t v = members[k];
        uint32_t m = k;
        while (m > b && members[m - 1] > v)
            {
            members[m] = members[m - 1];
            --m;
            }
        members[m] = v;
        }
    }

// Cell-ordered copy of the positions for the walk: slot m (the m-th member in
// cell order) holds the member's position as wrapped 32-bit fixed point and,
// in .w, its checked type (ntypes for an invalid type or a non-finite
// position, which is then never listed); rank_of[j] = m.
__global__ void k_cell_fixed(uint32_t N,
                             const uint32_t* __restrict__ members,
                             const double4* __restrict__ pos,
                             DevBox box,
                             uint32_t ntypes,
                             int4* __restrict__ cell_fixed,
                             uint32_t* __restrict__ rank_of)
    {
    const uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= N)
        return;
    const uint32_t j = members[m];
    const double4 p = pos[j];
    const bool finite = isfinite(p.x) && isfinite(p.y) && isfinite(p.z);
    const uint32_t t = finite ? checked_type(p.w, ntypes) : ntypes;
    cell_fixed[m] = finite ? make_int4(to_fixed(p.x, box.Linv[0]), to_fixed(p.y, box.Linv[1]),
                                       to_fixed(p.z, box.Linv[2]), (int)t)
                           : make_int4(0, 0, 0, (int)t);
    rank_of[j] = m;
    }

struct WalkArgs
    {
    const double4* pos;
    uint32_t N;
    DevBox box;
    float fx_scale[3]; // L / 2^32 per axis
    CellGrid g;
    const uint32_t* cell_of;
    const uint32_t* cell_rank;
    const uint32_t* start;
    const uint32_t* members;
    const int4* cell_fixed;
    const uint32_t* rank_of;
    const double* r_listsq;
    const float* r_listsq_lo; // r_listsq (1 - 1e-5) in FP32
    const float* r_listsq_hi; // r_listsq (1 + 1e-5) in FP32
    uint32_t ntypes;
    uint32_t* n_neigh;
    uint32_t* rows; // particle-major staging: entry k of particle i at rows[i * max_neighbors + k]
    uint32_t max_neighbors;
    uint32_t* overflow; // [0] largest count seen when > max_neighbors, [2] bad type seen
    };

// Eight buffered entries -> one 32-byte store (rows are 32-byte aligned:
// max_neighbors is a multiple of 8).
__device__ __forceinline__ void store_chunk(uint32_t* dst, const uint32_t* lbuf)
    {
    uint4* d = reinterpret_cast<uint4*>(dst);
    d[0] = make_uint4(lbuf[0], lbuf[1], lbuf[2], lbuf[3]);
    d[1] = make_uint4(lbuf[4], lbuf[5], lbuf[6], lbuf[7]);
    }

// Lists particle i's neighbors and returns how many it has.  Candidates are
// read from the cell-ordered fixed-point copy (16 bytes, contiguous per
// cell) and tested in FP32 on the exact integer displacement; a candidate
// whose FP32 r^2 lies within a relative 1e-5 of r_list^2 is decided by the
// FP64 predicate of the host oracle (minimum image and r^2 without
// contraction, rsq <= r_list^2), so the pair set is the FP64 one.  Outside
// the band the FP32 value is within ~1e-6 relative of the FP64 one.
template<bool SINGLE> __device__ uint32_t walk_one(const WalkArgs& a, uint32_t i, uint32_t* lbuf)
    {
    const double4 pi = a.pos[i];
    const uint32_t ti = checked_type(pi.w, a.ntypes);
    const bool finite = isfinite(pi.x) && isfinite(pi.y) && isfinite(pi.z);
    if (ti == a.ntypes || !finite)
        {
        // no neighbors for a particle without a valid type or position; reported to the host
        atomicOr(a.overflow + 2, finite ? 1u : 2u);
        return 0;
        }
    const int32_t fxi = to_fixed(pi.x, a.box.Linv[0]);
    const int32_t fyi = to_fixed(pi.y, a.box.Linv[1]);
    const int32_t fzi = to_fixed(pi.z, a.box.Linv[2]);
    const uint32_t self = a.rank_of[i];
    const uint32_t c = a.cell_of[i];
    const int cx = (int)(c % (uint32_t)a.g.n[0]);
    const int cy = (int)((c / (uint32_t)a.g.n[0]) % (uint32_t)a.g.n[1]);
    const int cz = (int)(c / ((uint32_t)a.g.n[0] * (uint32_t)a.g.n[1]));
    // neighbor cells per axis, deduplicated when the grid is narrower than 3
    int ox[3], oy[3], oz[3];
    const int nx = stencil_axis(cx, a.g.n[0], ox);
    const int ny = stencil_axis(cy, a.g.n[1], oy);
    const int nz = stencil_axis(cz, a.g.n[2], oz);
    const double rl0 = a.r_listsq[0];
    const float lo0 = a.r_listsq_lo[0], hi0 = a.r_listsq_hi[0];
    uint32_t count = 0;
    for (int iz = 0; iz < nz; ++iz)
        for (int iy = 0; iy < ny; ++iy)
            for (int ix = 0; ix < nx; ++ix)
                {
                const uint32_t nb = a.cell_rank[(oz[iz] * a.g.n[1] + oy[iy]) * a.g.n[0] + ox[ix]];
                const uint32_t b = a.start[nb], e = a.start[nb + 1];
                // candidates 4 at a time: their loads are issued together
                for (uint32_t m0 = b; m0 < e; m0 += 4)
                    {
                    int4 pc[4];
#pragma unroll
                    for (uint32_t u = 0; u < 4; ++u)
                        pc[u] = a.cell_fixed[min(m0 + u, e - 1)];
#pragma unroll
                    for (uint32_t u = 0; u < 4; ++u)
                    {
                    const uint32_t m = m0 + u;
                    const int4 pj = pc[u];
                    const uint32_t tj = (uint32_t)pj.w;
                    if (m >= e || m == self || tj >= a.ntypes)
                        continue;
                    const uint32_t tp = SINGLE ? 0u : ti * a.ntypes + tj;
                    const float lo = SINGLE ? lo0 : a.r_listsq_lo[tp];
                    const float hi = SINGLE ? hi0 : a.r_listsq_hi[tp];
                    const float dx = (float)fixed_delta(fxi, pj.x) * a.fx_scale[0];
                    const float dy = (float)fixed_delta(fyi, pj.y) * a.fx_scale[1];
                    const float dz = (float)fixed_delta(fzi, pj.z) * a.fx_scale[2];
                    const float rsq = dx * dx + dy * dy + dz * dz;
                    bool in = rsq < lo;
                    if (!in && rsq <= hi)
                        {
                        const uint32_t j = a.members[m];
                        const double4 p64 = a.pos[j];
                        const double rl = SINGLE ? rl0 : a.r_listsq[tp];
                        const double ex = min_image_1d(pi.x - p64.x, a.box.L[0], a.box.Linv[0]);
                        const double ey = min_image_1d(pi.y - p64.y, a.box.L[1], a.box.Linv[1]);
                        const double ez = min_image_1d(pi.z - p64.z, a.box.L[2], a.box.Linv[2]);
                        in = rsq3_nofma(ex, ey, ez) <= rl;
                        }
                    if (in)
                        {
                        // entries gather in the lane's LDS slots and leave
                        // as whole 32-byte chunks of its row
                        lbuf[count & 7u] = a.members[m];
                        if ((count & 7u) == 7u && count < a.max_neighbors)
                            store_chunk(a.rows + (size_t)i * a.max_neighbors + (count & ~7u), lbuf);
                        ++count;
                        }
                    }
                    }
                }
    // the last partial chunk (slots past the count are never read)
    if ((count & 7u) != 0u && count < a.max_neighbors)
        store_chunk(a.rows + (size_t)i * a.max_neighbors + (count & ~7u), lbuf);
    return count;
    }

// One lane per particle.  Each lane writes its entries to its own row of the
// particle-major staging array (successive writes of a lane fill whole cache
// lines); k_list_columns then transposes into the column layout.
template<bool SINGLE> __global__ void __launch_bounds__(256) k_walk(const WalkArgs a)
    {
    __shared__ uint32_t s_buf[256 * 8];
    const uint32_t i = xcd_logical_block() * blockDim.x + threadIdx.x;
    if (i >= a.N)
        return;
    const uint32_t count = walk_one<SINGLE>(a, i, s_buf + threadIdx.x * 8u);
    // a list longer than the capacity is truncated (and reported), so
    // consumers never read past the allocation
    a.n_neigh[i] = min(count, a.max_neighbors);
    if (count > a.max_neighbors)
        atomicMax(a.overflow, count);
    }

// Particle-major rows -> column layout nlist[k * pitch + i], one wavefront per
// 64 consecutive particles: each row k of the group is one contiguous 256-byte
// store.  Columns are padded with the particle's own index up to the group's
// longest list rounded up to 8, so consumers can run a wave-uniform loop over
// batches of 8 entries (padding entries have j == i).
__global__ void __launch_bounds__(64) k_list_columns(uint32_t N,
                                                     const uint32_t* __restrict__ n_neigh,
                                                     const uint32_t* __restrict__ rows,
                                                     uint32_t max_neighbors,
                                                     uint32_t* __restrict__ nlist,
                                                     uint32_t pitch)
    {
    const uint32_t i = blockIdx.x * 64u + threadIdx.x;
    const bool active = i < N;
    const uint32_t count = active ? n_neigh[i] : 0u;
    uint32_t wmax = count;
    for (int off = 32; off > 0; off >>= 1)
        wmax = max(wmax, (uint32_t)__shfl_xor((int)wmax, off));
    // a multiple of 8 (max_neighbors is one)
    const uint32_t padded = min((wmax + 7u) & ~7u, max_neighbors);
    if (!active)
        return;
    // the row is read in 32-byte chunks, two chunks in flight
    const uint4* row = reinterpret_cast<const uint4*>(rows + (size_t)i * max_neighbors);
    for (uint32_t k0 = 0; k0 < padded; k0 += 16)
        {
        const bool two = k0 + 8 < padded;
        uint4 c[4];
        c[0] = row[k0 / 4];
        c[1] = row[k0 / 4 + 1];
        if (two)
            {
            c[2] = row[k0 / 4 + 2];
            c[3] = row[k0 / 4 + 3];
            }
        const uint32_t v[16] = {c[0].x, c[0].y, c[0].z, c[0].w, c[1].x, c[1].y, c[1].z, c[1].w,
                                c[2].x, c[2].y, c[2].z, c[2].w, c[3].x, c[3].y, c[3].z, c[3].w};
#pragma unroll
        for (uint32_t u = 0; u < 16; ++u)
            if (u < 8 || two)
                nlist[(size_t)(k0 + u) * pitch + i] = k0 + u < count ? v[u] : i;
        }
    }

// Sets *flag when some particle moved farther than sqrt(thresh_sq).
__global__ void k_displacement(const double4* __restrict__ pos,
                               const double4* __restrict__ last,
                               uint32_t N,
                               DevBox box,
                               double thresh_sq,
                               uint32_t* __restrict__ flag)
    {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N)
        return;
    const uint32_t bits = displaced(pos[i], last[i], box, thresh_sq);
    if (bits != 0u)
        atomicOr(flag, bits);
    }

__global__ void k_copy_d4(const double4* __restrict__ in, double4* __restrict__ out, uint32_t N)
    {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N)
        out[i] = in[i];
    }

} // namespace azp

using namespace azp;

struct azp_nlist
    {
    double r_buff = 0.4;
    uint32_t ntypes = 0;
    std::vector<double> r_cut;
    bool params_dirty = true;

    DeviceArray<double> r_listsq;
    DeviceArray<float> r_listsq_lo, r_listsq_hi;
    DeviceArray<int4> cell_fixed;
    DeviceArray<uint32_t> rank_of;
    DeviceArray<uint32_t> n_neigh;
    DeviceArray<uint32_t> list;
    DeviceArray<uint32_t> rows; // particle-major staging of the walk
    DeviceArray<double4> last_pos;
    DeviceArray<uint32_t> cell_of, slot_of, count, start, members, scan_sums;
    DeviceArray<uint32_t> cell_rank; // Morton rank of each cell of the grid below
    DeviceArray<uint32_t> rank_cell; // its inverse
    int rank_grid[3] = {0, 0, 0};
    DeviceArray<uint32_t> flags; // [0] overflow, [1] rebuild, [2] bad type, [3] tile misfit, [4] largest tile set,
                                 // [5] tiles walked in FP32 offsets, [6] the step kernel's check
    uint32_t* h_flags = nullptr;  // pinned, 8 words ([4], [5] tile build flags)

    bool want_tiles = false;
    uint32_t tile_mode = 0; // 1: 512 else 256; 256 / 512: that size only
    TileList tiles;
    // The fused walk writes only the tiled list; the plain list is expanded
    // from it when a consumer asks for it (plain_ready false until then).
    bool plain_ready = true;
    bool fused = false; // the last build used the fused walk
    uint32_t relative_tiles = 0; // of which tiles were walked in FP32 offsets
    hipStream_t last_stream = nullptr;
    uint32_t walk_skip[2] = {0, 0}; // builds to skip the fused walk of 512 / 256 after a tile did not fit

    uint32_t pitch = 0;
    uint32_t max_neighbors = 0;
    uint32_t built_N = 0;
    azp_box built_box = {0, 0, 0};
    bool have_list = false;
    uint64_t n_builds = 0;
    // flags[6] holds the rebuild check of these positions, made by the step
    // after the last force (k_step) as (checked_seq << 2) | bits; the next
    // update_compute_step reads it.  Each step check takes a new number, so
    // the word is cleared only on first use and when the numbers wrap.
    const double* checked_pos = nullptr;
    uint32_t step_seq = 0;
    uint32_t checked_seq = 0;

    ~azp_nlist()
        {
        if (h_flags)
            (void)hipHostFree(h_flags);
        }

    double r_list_max() const
        {
        double m = 0.0;
        for (double rc : r_cut)
            if (rc > 0.0)
                m = std::max(m, rc + r_buff);
        return m;
        }
    };

static int nlist_build(azp_nlist* nl, const azp_particles* p, hipStream_t s)
    {
    // until this build succeeds there is no list: an error return leaves the
    // list invalid, so the next update rebuilds it instead of comparing the
    // displacement against the failed build's positions
    nl->have_list = false;
    nl->built_N = 0;
    nl->tiles.valid = false;
    nl->checked_pos = nullptr;
    const uint32_t N = p->N;
    const DevBox box = make_box(p->box);
    const double rmax = nl->r_list_max();
    const double L[3] = {p->box.Lx, p->box.Ly, p->box.Lz};
    const CellGrid g = cell_grid_for(L, rmax, N);
    const size_t ncell = (size_t)g.n[0] * g.n[1] * g.n[2];

    AZP_TRY(nl->cell_of.reserve(N));
    AZP_TRY(nl->slot_of.reserve(N));
    AZP_TRY(nl->count.reserve(ncell));
    AZP_TRY(nl->start.reserve(ncell + 1));
    AZP_TRY(nl->scan_sums.reserve(scan_chunks((uint32_t)ncell)));
    AZP_TRY(nl->members.reserve(N));
    AZP_TRY(nl->cell_fixed.reserve(N));
    AZP_TRY(nl->rank_of.reserve(N));
    AZP_TRY(nl->n_neigh.reserve(N));
    AZP_TRY(nl->last_pos.reserve(N));
    AZP_TRY(nl->flags.reserve(7));

    if (g.n[0] != nl->rank_grid[0] || g.n[1] != nl->rank_grid[1] || g.n[2] != nl->rank_grid[2])
        {
        std::vector<uint32_t> rank, inv(ncell);
        morton_cell_rank(g, rank);
        for (size_t k = 0; k < ncell; ++k)
            inv[rank[k]] = (uint32_t)k;
        AZP_TRY(nl->cell_rank.reserve(ncell));
        AZP_TRY(nl->rank_cell.reserve(ncell));
        AZP_HIP(hipMemcpyAsync(nl->cell_rank.ptr, rank.data(), ncell * sizeof(uint32_t), hipMemcpyHostToDevice, s));
        AZP_HIP(hipMemcpyAsync(nl->rank_cell.ptr, inv.data(), ncell * sizeof(uint32_t), hipMemcpyHostToDevice, s));
        AZP_HIP(hipStreamSynchronize(s)); // the host vectors go out of scope
        for (int k = 0; k < 3; ++k)
            nl->rank_grid[k] = g.n[k];
        }

    const unsigned B = 256;
    const double4* pos = reinterpret_cast<const double4*>(p->d_pos);
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(ncell, B)), dim3(B), 0, s, nl->count.ptr, 0u, ncell);
    hipLaunchKernelGGL(k_bin, dim3(grid_for(N, B)), dim3(B), 0, s, pos, N, box, g, nl->cell_rank.ptr,
                       nl->cell_of.ptr, nl->slot_of.ptr, nl->count.ptr);
    enqueue_scan(nl->count.ptr, (uint32_t)ncell, nl->scan_sums.ptr, nl->start.ptr, nullptr, s);
    hipLaunchKernelGGL(k_scatter, dim3(grid_for(N, B)), dim3(B), 0, s, N, nl->cell_of.ptr, nl->cell_rank.ptr,
                       nl->slot_of.ptr, nl->start.ptr, nl->members.ptr);
    hipLaunchKernelGGL(k_sort_cells, dim3(grid_for(ncell, 4)), dim3(256), 0, s, (uint32_t)ncell,
                       nl->start.ptr, nl->members.ptr);
    hipLaunchKernelGGL(k_cell_fixed, dim3(grid_for(N, B)), dim3(B), 0, s, N, nl->members.ptr, pos, box, nl->ntypes,
                       nl->cell_fixed.ptr, nl->rank_of.ptr);
    AZP_LAUNCH_CHECK();

    // initial capacity from the density
    const double vol = L[0] * L[1] * L[2];
    if (nl->max_neighbors == 0)
        {
        const double est = (double)N / vol * 4.18879020478639 * rmax * rmax * rmax;
        nl->max_neighbors = (uint32_t)std::min(1.3 * est + 16.0, (double)N);
        nl->max_neighbors = (nl->max_neighbors + 7u) & ~7u;
        if (nl->max_neighbors == 0)
            nl->max_neighbors = 8;
        }
    nl->pitch = wave_pitch(N);

    for (int attempt = 0; attempt < 4; ++attempt)
        {
        AZP_TRY(nl->list.reserve((size_t)nl->max_neighbors * nl->pitch));
        AZP_TRY(nl->rows.reserve((size_t)nl->max_neighbors * N));
        nl->tiles.valid = false;
        nl->plain_ready = true;
        nl->fused = false;
        if (nl->want_tiles && nl->ntypes <= kTileWalkMaxTypes)
            {
            // fused walk: the tiled list straight from the cell list, 512-
            // particle tiles, else 256
            bool grown = false, done = false;
            for (uint32_t ts : {512u, 256u})
                {
                if (nl->tile_mode != 1 && nl->tile_mode != ts)
                    continue;
                uint32_t& skip = nl->walk_skip[ts == 512u ? 0 : 1];
                if (skip > 0)
                    {
                    --skip;
                    continue;
                    }
                hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, nl->flags.ptr, 0u, (size_t)6);
                TileWalkArgs ta;
                ta.pos = pos;
                ta.N = N;
                ta.box = box;
                for (int k = 0; k < 3; ++k)
                    ta.fx_scale[k] = (float)(box.L[k] / 4294967296.0);
                ta.g = g;
                ta.cell_of = nl->cell_of.ptr;
                ta.cell_rank = nl->cell_rank.ptr;
                ta.rank_cell = nl->rank_cell.ptr;
                ta.start = nl->start.ptr;
                ta.members = nl->members.ptr;
                ta.cell_fixed = nl->cell_fixed.ptr;
                ta.rank_of = nl->rank_of.ptr;
                ta.r_listsq = nl->r_listsq.ptr;
                ta.r_listsq_lo = nl->r_listsq_lo.ptr;
                ta.r_listsq_hi = nl->r_listsq_hi.ptr;
                ta.r_list_max = rmax;
                ta.ntypes = nl->ntypes;
                ta.n_neigh = nl->n_neigh.ptr;
                ta.rows = reinterpret_cast<uint16_t*>(nl->rows.ptr);
                ta.max_neighbors = nl->max_neighbors;
                ta.flags = nl->flags.ptr;
                ta.pitch = nl->pitch;
                AZP_TRY(tiles_walk(nl->tiles, ts, ta, s));
                // the positions of this build, queued before the flags are
                // read so that no host round trip separates it from the walk
                // (a failed attempt copies them again)
                hipLaunchKernelGGL(k_copy_d4, dim3(grid_for(N, B)), dim3(B), 0, s, pos, nl->last_pos.ptr, N);
                AZP_LAUNCH_CHECK();
                AZP_HIP(hipMemcpyAsync(nl->h_flags, nl->flags.ptr, 6 * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
                AZP_HIP(hipStreamSynchronize(s));
                if (nl->h_flags[2] & 2u)
                    return set_error(AZP_ERR_INVALID, "a particle position is not finite (NaN or inf)");
                if (nl->h_flags[2] & 1u)
                    return set_error(AZP_ERR_INVALID,
                                     "a particle type index is outside [0, %u) (pos.w must hold the type)",
                                     nl->ntypes);
                if (nl->h_flags[0] != 0u)
                    {
                    grown = true;
                    break;
                    }
                if (nl->h_flags[3] == 0u)
                    {
                    nl->tiles.size = ts;
                    nl->tiles.slots = nl->h_flags[4] + 1u;
                    nl->tiles.valid = true;
                    nl->plain_ready = false;
                    nl->fused = true;
                    nl->relative_tiles = nl->h_flags[5];
                    done = true;
                    break;
                    }
                skip = 8; // a tile did not fit: plain walk for the next builds
                }
            if (grown)
                {
                nl->max_neighbors = ((uint32_t)(nl->h_flags[0] * 1.1) + 8u) & ~7u;
                continue;
                }
            if (done)
                {
                nl->built_N = N;
                nl->built_box = p->box;
                nl->have_list = true;
                nl->n_builds++;
                return AZP_OK;
                }
            }
        hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, nl->flags.ptr, 0u, (size_t)3);
        WalkArgs wa;
        wa.pos = pos;
        wa.N = N;
        wa.box = box;
        for (int k = 0; k < 3; ++k)
            wa.fx_scale[k] = (float)(box.L[k] / 4294967296.0);
        wa.g = g;
        wa.cell_of = nl->cell_of.ptr;
        wa.cell_rank = nl->cell_rank.ptr;
        wa.start = nl->start.ptr;
        wa.members = nl->members.ptr;
        wa.cell_fixed = nl->cell_fixed.ptr;
        wa.rank_of = nl->rank_of.ptr;
        wa.r_listsq = nl->r_listsq.ptr;
        wa.r_listsq_lo = nl->r_listsq_lo.ptr;
        wa.r_listsq_hi = nl->r_listsq_hi.ptr;
        wa.ntypes = nl->ntypes;
        wa.n_neigh = nl->n_neigh.ptr;
        wa.rows = nl->rows.ptr;
        wa.max_neighbors = nl->max_neighbors;
        wa.overflow = nl->flags.ptr;
        if (nl->ntypes == 1)
            hipLaunchKernelGGL(k_walk<true>, dim3(grid_for(N, B)), dim3(B), 0, s, wa);
        else
            hipLaunchKernelGGL(k_walk<false>, dim3(grid_for(N, B)), dim3(B), 0, s, wa);
        hipLaunchKernelGGL(k_list_columns, dim3(grid_for(N, 64)), dim3(64), 0, s, N, nl->n_neigh.ptr, nl->rows.ptr,
                           nl->max_neighbors, nl->list.ptr, nl->pitch);
        AZP_LAUNCH_CHECK();
        AZP_HIP(hipMemcpyAsync(nl->h_flags, nl->flags.ptr, 3 * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        AZP_HIP(hipStreamSynchronize(s));
        if (nl->h_flags[2] & 2u)
            return set_error(AZP_ERR_INVALID, "a particle position is not finite (NaN or inf)");
        if (nl->h_flags[2] & 1u)
            return set_error(AZP_ERR_INVALID,
                             "a particle type index is outside [0, %u) (pos.w must hold the type)",
                             nl->ntypes);
        const uint32_t need = nl->h_flags[0];
        if (need == 0)
            {
            hipLaunchKernelGGL(k_copy_d4, dim3(grid_for(N, B)), dim3(B), 0, s, pos, nl->last_pos.ptr, N);
            AZP_LAUNCH_CHECK();
            if (nl->want_tiles)
                {
                // tiles from the plain list: 512-particle tiles, else 256;
                // without either the plain list serves
                for (uint32_t ts : {512u, 256u})
                    {
                    if (nl->tile_mode != 1 && nl->tile_mode != ts)
                        continue;
                    AZP_TRY(tiles_build(nl->tiles, ts, N, nl->n_neigh.ptr, nl->rows.ptr, nl->pitch,
                                        nl->max_neighbors, nl->h_flags + 4, s));
                    if (nl->tiles.valid)
                        break;
                    }
                }
            nl->built_N = N;
            nl->built_box = p->box;
            nl->have_list = true;
            nl->n_builds++;
            return AZP_OK;
            }
        nl->max_neighbors = ((uint32_t)(need * 1.1) + 8u) & ~7u;
        }
    return set_error(AZP_ERR_STATE, "neighbor list did not converge after growing its capacity");
    }

extern "C" int azp_nlist_create(azp_nlist** out)
    {
    if (!out)
        return set_error(AZP_ERR_INVALID, "null out");
    azp_nlist* nl = new azp_nlist();
    if (hipHostMalloc((void**)&nl->h_flags, 8 * sizeof(uint32_t), hipHostMallocCoherent) != hipSuccess)
        {
        nl->h_flags = nullptr;
        delete nl;
        return set_error(AZP_ERR_NOMEM, "pinned host allocation failed");
        }
    *out = nl;
    return AZP_OK;
    }

extern "C" int azp_nlist_destroy(azp_nlist* nl)
    {
    delete nl;
    return AZP_OK;
    }

extern "C" int azp_nlist_set_r_buff(azp_nlist* nl, double r_buff)
    {
    if (!nl || !(r_buff >= 0.0))
        return set_error(AZP_ERR_INVALID, "r_buff must be >= 0");
    if (r_buff != nl->r_buff)
        nl->params_dirty = true;
    nl->r_buff = r_buff;
    return AZP_OK;
    }

extern "C" int azp_nlist_set_r_cut(azp_nlist* nl, uint32_t ntypes, const double* r_cut)
    {
    if (!nl || ntypes == 0 || !r_cut)
        return set_error(AZP_ERR_INVALID, "bad r_cut arguments");
    std::vector<double> rc(r_cut, r_cut + (size_t)ntypes * ntypes);
    for (uint32_t a = 0; a < ntypes; ++a)
        for (uint32_t b = 0; b < ntypes; ++b)
            if (rc[a * ntypes + b] != rc[b * ntypes + a])
                return set_error(AZP_ERR_INVALID, "r_cut must be symmetric");
    if (ntypes != nl->ntypes || rc != nl->r_cut)
        nl->params_dirty = true;
    nl->ntypes = ntypes;
    nl->r_cut = rc;
    return AZP_OK;
    }

// Value the rebuild flag never takes (it holds bits 1 and 2 only).
constexpr uint32_t kFlagPending = 0xa5a5a5a5u;

// azp_nlist_update; with use_checked, a rebuild check of p->d_pos that a
// step made after the last update (nlist_step_checked) stands in for the
// displacement kernel.  Any update consumes that check.
static int nlist_update(azp_nlist* nl, const azp_particles* p, int force_rebuild, int* rebuilt, void* stream,
                        bool use_checked)
    {
    if (rebuilt)
        *rebuilt = 0;
    if (!nl || !p)
        return set_error(AZP_ERR_INVALID, "null argument");
    if (nl->ntypes == 0)
        return set_error(AZP_ERR_STATE, "set r_cut before building the neighbor list");
    if (!box_valid(p->box))
        return set_error(AZP_ERR_BOX, "box lengths must be positive");
    if (p->N > 0 && !p->d_pos)
        return set_error(AZP_ERR_INVALID, "null positions");
    const double rmax = nl->r_list_max();
    const double L[3] = {p->box.Lx, p->box.Ly, p->box.Lz};
    for (int k = 0; k < 3; ++k)
        if (L[k] < 2.0 * rmax)
            return set_error(AZP_ERR_BOX,
                             "box length %g is smaller than twice the neighbor range %g: particles "
                             "would interact with their own images",
                             L[k],
                             rmax);
    hipStream_t s = (hipStream_t)stream;
    nl->last_stream = s;
    if (nl->params_dirty)
        {
        const size_t nt2 = (size_t)nl->ntypes * nl->ntypes;
        std::vector<double> rl(nt2);
        for (size_t k = 0; k < nt2; ++k)
            {
            const double r = nl->r_cut[k] + nl->r_buff;
            rl[k] = nl->r_cut[k] > 0.0 ? r * r : 0.0;
            }
        std::vector<float> lo(nt2), hi(nt2);
        for (size_t k = 0; k < nt2; ++k)
            {
            // excluded type pairs (r_cut <= 0) can never pass either test
            lo[k] = rl[k] > 0.0 ? (float)(rl[k] * (1.0 - 1e-5)) : -1.0f;
            hi[k] = rl[k] > 0.0 ? (float)(rl[k] * (1.0 + 1e-5)) : -1.0f;
            }
        AZP_TRY(nl->r_listsq.reserve(nt2));
        AZP_TRY(nl->r_listsq_lo.reserve(nt2));
        AZP_TRY(nl->r_listsq_hi.reserve(nt2));
        AZP_HIP(hipMemcpyAsync(nl->r_listsq.ptr, rl.data(), nt2 * sizeof(double), hipMemcpyHostToDevice, s));
        AZP_HIP(hipMemcpyAsync(nl->r_listsq_lo.ptr, lo.data(), nt2 * sizeof(float), hipMemcpyHostToDevice, s));
        AZP_HIP(hipMemcpyAsync(nl->r_listsq_hi.ptr, hi.data(), nt2 * sizeof(float), hipMemcpyHostToDevice, s));
        // the host vectors go out of scope: wait for the copies
        AZP_HIP(hipStreamSynchronize(s));
        }
    bool need = force_rebuild || !nl->have_list || nl->params_dirty || p->N != nl->built_N
                || p->box.Lx != nl->built_box.Lx || p->box.Ly != nl->built_box.Ly
                || p->box.Lz != nl->built_box.Lz;
    // the step after the last force may have checked these positions already
    const bool prechecked = use_checked && nl->checked_pos != nullptr && nl->checked_pos == p->d_pos;
    nl->checked_pos = nullptr;
    if (!need && p->N > 0)
        {
        if (!prechecked)
            {
            const unsigned B = 256;
            hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, nl->flags.ptr + 1, 0u, (size_t)1);
            const double half = 0.5 * nl->r_buff;
            hipLaunchKernelGGL(k_displacement, dim3(grid_for(p->N, B)), dim3(B), 0, s,
                               reinterpret_cast<const double4*>(p->d_pos), nl->last_pos.ptr, p->N,
                               make_box(p->box), half * half, nl->flags.ptr + 1);
            AZP_LAUNCH_CHECK();
            }
        // The host polls the pinned word for the copied flag (a sentinel until
        // it lands) instead of waiting for the stream: the step's critical
        // path then skips the wake-up of a blocking synchronization.  After
        // 2 ms it falls back to synchronizing the stream.
        const uint32_t w = prechecked ? 6u : 1u;
        volatile uint32_t* hflag = nl->h_flags + w;
        *hflag = kFlagPending;
        std::atomic_thread_fence(std::memory_order_seq_cst);
        AZP_HIP(hipMemcpyAsync(nl->h_flags + w, nl->flags.ptr + w, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        const auto t0 = std::chrono::steady_clock::now();
        while (*hflag == kFlagPending)
            {
            if (std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(2))
                {
                AZP_HIP(hipStreamSynchronize(s));
                break;
                }
            }
        std::atomic_thread_fence(std::memory_order_seq_cst);
        uint32_t bits = nl->h_flags[w];
        if (prechecked)
            bits = (bits >> 2) == nl->checked_seq ? (bits & 3u) : 0u;
        if (bits & 2u)
            return set_error(AZP_ERR_INVALID, "a particle position is not finite (NaN or inf)");
        need = bits != 0;
        }
    nl->params_dirty = false;
    if (!need)
        return AZP_OK;
    if (p->N == 0)
        {
        nl->fused = false;
        nl->plain_ready = true;
        nl->tiles.valid = false;
        nl->built_N = 0;
        nl->built_box = p->box;
        nl->have_list = true;
        nl->pitch = 0;
        AZP_TRY(nl->n_neigh.reserve(1));
        AZP_TRY(nl->list.reserve(1));
        if (rebuilt)
            *rebuilt = 1;
        return AZP_OK;
        }
    AZP_TRY(nlist_build(nl, p, s));
    if (rebuilt)
        *rebuilt = 1;
    return AZP_OK;
    }

extern "C" int azp_nlist_update(azp_nlist* nl, const azp_particles* p, int force_rebuild, int* rebuilt, void* stream)
    {
    return nlist_update(nl, p, force_rebuild, rebuilt, stream, false);
    }

// Expands the tiled list into the plain one on stream s if that has not
// happened since the last build.
static int nlist_expand(const azp_nlist* cnl, hipStream_t s)
    {
    azp_nlist* nl = const_cast<azp_nlist*>(cnl);
    if (nl->plain_ready)
        return AZP_OK;
    AZP_TRY(tiles_to_plain(nl->tiles, nl->built_N, nl->n_neigh.ptr, nl->pitch, nl->list.ptr, s));
    nl->plain_ready = true;
    return AZP_OK;
    }

namespace azp
{
int nlist_step_check(azp_nlist* nl, hipStream_t s, StepKernelArgs* k)
    {
    if (!nl || !nl->have_list)
        return set_error(AZP_ERR_STATE, "neighbor list not built");
    // a new check number per step; the word is cleared on first use and
    // before the numbers would pass 2^30 (the flag holds seq << 2)
    if (nl->step_seq == 0 || nl->step_seq >= (1u << 30) - 1u)
        {
        hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, nl->flags.ptr + 6, 0u, (size_t)1);
        AZP_LAUNCH_CHECK();
        nl->step_seq = 0;
        }
    ++nl->step_seq;
    k->last = nl->last_pos.ptr;
    k->flag = nl->flags.ptr + 6;
    k->seq = nl->step_seq;
    const double half = 0.5 * nl->r_buff;
    k->thresh_sq = half * half;
    return AZP_OK;
    }

void nlist_step_checked(azp_nlist* nl, const double* d_pos)
    {
    nl->checked_pos = d_pos;
    nl->checked_seq = nl->step_seq;
    }

int nlist_update_checked(azp_nlist* nl, const azp_particles* p, int force_rebuild, int* rebuilt, void* stream)
    {
    return nlist_update(nl, p, force_rebuild, rebuilt, stream, true);
    }

// The list arrays for work enqueued on stream s.  With need_plain false the
// plain list is returned only if it is current (else null: a tiled consumer).
int nlist_get_on_stream(const azp_nlist* nl,
                        hipStream_t s,
                        bool need_plain,
                        const uint32_t** d_n_neigh,
                        const uint32_t** d_nlist,
                        uint32_t* pitch,
                        uint32_t* max_neighbors)
    {
    if (!nl)
        return set_error(AZP_ERR_INVALID, "null nlist");
    if (!nl->have_list)
        return set_error(AZP_ERR_STATE, "neighbor list not built");
    if (need_plain)
        AZP_TRY(nlist_expand(nl, s));
    if (d_n_neigh)
        *d_n_neigh = nl->n_neigh.ptr;
    if (d_nlist)
        *d_nlist = nl->plain_ready ? nl->list.ptr : nullptr;
    if (pitch)
        *pitch = nl->pitch;
    if (max_neighbors)
        *max_neighbors = nl->max_neighbors;
    return AZP_OK;
    }
} // namespace azp

extern "C" int azp_nlist_get(const azp_nlist* nl,
                             const uint32_t** d_n_neigh,
                             const uint32_t** d_nlist,
                             uint32_t* pitch,
                             uint32_t* max_neighbors)
    {
    if (!nl)
        return set_error(AZP_ERR_INVALID, "null nlist");
    if (!nl->have_list)
        return set_error(AZP_ERR_STATE, "neighbor list not built");
    if (!nl->plain_ready)
        {
        // expanded on the build's stream, finished before the pointers are handed out
        AZP_TRY(nlist_expand(nl, nl->last_stream));
        AZP_HIP(hipStreamSynchronize(nl->last_stream));
        }
    return nlist_get_on_stream(nl, nl->last_stream, true, d_n_neigh, d_nlist, pitch, max_neighbors);
    }

extern "C" int azp_nlist_set_tiled(azp_nlist* nl, int mode)
    {
    if (!nl)
        return set_error(AZP_ERR_INVALID, "null nlist");
    if (mode != 0 && mode != 1 && mode != 256 && mode != 512)
        return set_error(AZP_ERR_INVALID, "tile mode %d must be 0, 1, 256 or 512", mode);
    if ((uint32_t)mode != nl->tile_mode)
        nl->params_dirty = true; // next update rebuilds (and tiles) the list
    if (mode == 0 && nl->have_list && !nl->plain_ready)
        {
        // the plain list stays usable until that rebuild
        AZP_TRY(nlist_expand(nl, nl->last_stream));
        AZP_HIP(hipStreamSynchronize(nl->last_stream));
        }
    nl->tile_mode = (uint32_t)mode;
    nl->want_tiles = mode != 0;
    if (!nl->want_tiles)
        nl->tiles.valid = false;
    return AZP_OK;
    }

extern "C" int azp_nlist_get_tiles(const azp_nlist* nl,
                                   const uint32_t** d_tile_count,
                                   const uint32_t** d_tile_jset,
                                   const uint32_t** d_tile_list,
                                   uint32_t* tile_size,
                                   uint32_t* tile_slots)
    {
    if (!nl)
        return set_error(AZP_ERR_INVALID, "null nlist");
    const bool ok = nl->want_tiles && nl->tiles.valid && nl->have_list;
    if (d_tile_count)
        *d_tile_count = ok ? nl->tiles.count.ptr : nullptr;
    if (d_tile_jset)
        *d_tile_jset = ok ? nl->tiles.jset.ptr : nullptr;
    if (d_tile_list)
        *d_tile_list = ok ? nl->tiles.list.ptr : nullptr;
    if (tile_size)
        *tile_size = ok ? nl->tiles.size : 0u;
    if (tile_slots)
        *tile_slots = ok ? nl->tiles.slots : 0u;
    return AZP_OK;
    }

extern "C" int azp_nlist_get_tile_order(const azp_nlist* nl, const uint32_t** d_tile_order, const uint32_t** d_tile_nn)
    {
    if (!nl)
        return set_error(AZP_ERR_INVALID, "null nlist");
    const bool ok = nl->want_tiles && nl->tiles.valid && nl->have_list && nl->tiles.order;
    if (d_tile_order)
        *d_tile_order = ok ? nl->tiles.order : nullptr;
    if (d_tile_nn)
        *d_tile_nn = ok ? nl->tiles.nn.ptr : nullptr;
    return AZP_OK;
    }

extern "C" int azp_nlist_walk_kind(const azp_nlist* nl, int* kind)
    {
    if (!nl || !kind)
        return set_error(AZP_ERR_INVALID, "null argument");
    *kind = nl->fused ? 1 : 0;
    return AZP_OK;
    }

extern "C" int azp_nlist_tile_stats(const azp_nlist* nl, uint32_t* ntiles, uint32_t* relative_tiles)
    {
    if (!nl || !ntiles || !relative_tiles)
        return set_error(AZP_ERR_INVALID, "null argument");
    const bool ok = nl->fused && nl->tiles.valid;
    *ntiles = ok ? nl->tiles.ntiles : 0u;
    *relative_tiles = ok ? nl->relative_tiles : 0u;
    return AZP_OK;
    }

extern "C" int azp_nlist_num_builds(const azp_nlist* nl, uint64_t* n)
    {
    if (!nl || !n)
        return set_error(AZP_ERR_INVALID, "null argument");
    *n = nl->n_builds;
    return AZP_OK;
    }

extern "C" int azp_nlist_download(const azp_nlist* nl,
                                  uint32_t N,
                                  uint32_t* h_n_neigh,
                                  uint32_t* h_nlist,
                                  uint64_t nlist_capacity)
    {
    if (!nl || !h_n_neigh || !h_nlist)
        return set_error(AZP_ERR_INVALID, "null argument");
    if (!nl->have_list)
        return set_error(AZP_ERR_STATE, "neighbor list not built");
    if (N != nl->built_N)
        return set_error(AZP_ERR_INVALID, "N %u differs from the list's %u", N, nl->built_N);
    const uint64_t need = (uint64_t)nl->max_neighbors * nl->pitch;
    if (nlist_capacity < need)
        return set_error(AZP_ERR_INVALID, "host list needs %llu entries", (unsigned long long)need);
    if (N == 0)
        return AZP_OK;
    AZP_TRY(nlist_expand(nl, nl->last_stream));
    AZP_HIP(hipDeviceSynchronize());
    AZP_HIP(hipMemcpy(h_n_neigh, nl->n_neigh.ptr, (size_t)N * sizeof(uint32_t), hipMemcpyDeviceToHost));
    AZP_HIP(hipMemcpy(h_nlist, nl->list.ptr, (size_t)need * sizeof(uint32_t), hipMemcpyDeviceToHost));
    return AZP_OK;
    }

// ---- particle sort ------------------------------------------------------
// Particles reordered along the Morton curve of a grid of bins so that
// particles close in space are close in memory (the role of HOOMD's
// SFCPackTuner, which sorts along a Hilbert curve).  The bins reuse the cell
// list's kernels: bin, scan, scatter, then each bin's members by index.

struct azp_sorter
    {
    DeviceArray<uint32_t> cell_of, slot_of, count, start, scan_sums, cell_rank;
    int rank_grid[3] = {0, 0, 0};
    };

namespace azp
{

template<typename T>
__global__ void k_gather_rows(const T* __restrict__ src, T* __restrict__ dst, const uint32_t* __restrict__ order,
                              uint32_t N)
    {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= N)
        return;
    const uint32_t j = order[k];
    if (j < N)
        dst[k] = src[j];
    }

struct Row32
    {
    uint4 a, b;
    };

} // namespace azp

extern "C" int azp_sorter_create(azp_sorter** out)
    {
    if (!out)
        return set_error(AZP_ERR_INVALID, "null out");
    *out = new azp_sorter();
    return AZP_OK;
    }

extern "C" int azp_sorter_destroy(azp_sorter* so)
    {
    delete so;
    return AZP_OK;
    }

extern "C" int azp_sorter_order(azp_sorter* so, const azp_particles* p, uint32_t grid, uint32_t* d_order, void* stream)
    {
    if (!so || !p)
        return set_error(AZP_ERR_INVALID, "null argument");
    if (!box_valid(p->box))
        return set_error(AZP_ERR_BOX, "box lengths must be positive");
    if (grid == 0 || grid > 1024)
        return set_error(AZP_ERR_INVALID, "grid must be in [1, 1024], got %u", grid);
    const uint32_t N = p->N;
    if (N == 0)
        return AZP_OK;
    if (!p->d_pos || !d_order)
        return set_error(AZP_ERR_INVALID, "null positions or order");
    hipStream_t s = (hipStream_t)stream;
    const double L[3] = {p->box.Lx, p->box.Ly, p->box.Lz};
    const double w = std::max(L[0], std::max(L[1], L[2])) / grid;
    // bins of width >= w, about 4 per particle at most
    const CellGrid g = cell_grid_for(L, w, N);
    const size_t nbin = (size_t)g.n[0] * g.n[1] * g.n[2];
    AZP_TRY(so->cell_of.reserve(N));
    AZP_TRY(so->slot_of.reserve(N));
    AZP_TRY(so->count.reserve(nbin));
    AZP_TRY(so->start.reserve(nbin + 1));
    AZP_TRY(so->scan_sums.reserve(scan_chunks((uint32_t)nbin)));
    if (g.n[0] != so->rank_grid[0] || g.n[1] != so->rank_grid[1] || g.n[2] != so->rank_grid[2])
        {
        std::vector<uint32_t> rank;
        morton_cell_rank(g, rank);
        AZP_TRY(so->cell_rank.reserve(nbin));
        AZP_HIP(hipMemcpyAsync(so->cell_rank.ptr, rank.data(), nbin * sizeof(uint32_t), hipMemcpyHostToDevice, s));
        AZP_HIP(hipStreamSynchronize(s)); // the host vector goes out of scope
        for (int k = 0; k < 3; ++k)
            so->rank_grid[k] = g.n[k];
        }
    const unsigned B = 256;
    const double4* pos = reinterpret_cast<const double4*>(p->d_pos);
    const DevBox box = make_box(p->box);
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(nbin, B)), dim3(B), 0, s, so->count.ptr, 0u, nbin);
    hipLaunchKernelGGL(k_bin, dim3(grid_for(N, B)), dim3(B), 0, s, pos, N, box, g, so->cell_rank.ptr,
                       so->cell_of.ptr, so->slot_of.ptr, so->count.ptr);
    enqueue_scan(so->count.ptr, (uint32_t)nbin, so->scan_sums.ptr, so->start.ptr, nullptr, s);
    hipLaunchKernelGGL(k_scatter, dim3(grid_for(N, B)), dim3(B), 0, s, N, so->cell_of.ptr, so->cell_rank.ptr,
                       so->slot_of.ptr, so->start.ptr, d_order);
    hipLaunchKernelGGL(k_sort_cells, dim3(grid_for(nbin, 4)), dim3(256), 0, s, (uint32_t)nbin, so->start.ptr,
                       d_order);
    AZP_LAUNCH_CHECK();
    return AZP_OK;
    }

extern "C" int azp_gather_rows(const void* d_src,
                               void* d_dst,
                               const uint32_t* d_order,
                               uint32_t N,
                               uint32_t row_bytes,
                               void* stream)
    {
    if (N == 0)
        return AZP_OK;
    if (!d_src || !d_dst || !d_order)
        return set_error(AZP_ERR_INVALID, "null argument");
    if (d_src == d_dst)
        return set_error(AZP_ERR_INVALID, "source and destination must differ");
    hipStream_t s = (hipStream_t)stream;
    const unsigned B = 256;
    const dim3 grid(grid_for(N, B));
    switch (row_bytes)
        {
        case 4:
            hipLaunchKernelGGL(k_gather_rows<uint32_t>, grid, dim3(B), 0, s, (const uint32_t*)d_src,
                               (uint32_t*)d_dst, d_order, N);
            break;
        case 8:
            hipLaunchKernelGGL(k_gather_rows<uint2>, grid, dim3(B), 0, s, (const uint2*)d_src, (uint2*)d_dst,
                               d_order, N);
            break;
        case 16:
            hipLaunchKernelGGL(k_gather_rows<uint4>, grid, dim3(B), 0, s, (const uint4*)d_src, (uint4*)d_dst,
                               d_order, N);
            break;
        case 32:
            hipLaunchKernelGGL(k_gather_rows<Row32>, grid, dim3(B), 0, s, (const Row32*)d_src, (Row32*)d_dst,
                               d_order, N);
            break;
        default:
            return set_error(AZP_ERR_INVALID, "row_bytes must be 4, 8, 16 or 32, got %u", row_bytes);
        }
    AZP_LAUNCH_CHECK();
    return AZP_OK;
    }

// Cell-list helpers shared by the single-system and the batched neighbor
// list builds.  Both builds must bin and test pairs identically, so the
// arithmetic here is evaluated without contraction whatever the translation
// unit's flags are.
#pragma once

#include <algorithm>
#include <cmath>
#include <utility>
#include <vector>

#include "device_common.h"

namespace azp
{

struct CellGrid
    {
    int n[3];
    double w[3];
    };

// Cells of width >= r_list (with a 1e-9 relative margin) per axis, at most
// 1024 per axis and about 4 per particle (dilute systems).
inline CellGrid cell_grid_for(const double L[3], double rmax, uint32_t N)
    {
    CellGrid g;
    size_t ncell = 1;
    for (int k = 0; k < 3; ++k)
        {
        int n = (rmax > 0.0) ? (int)std::floor(L[k] / (rmax * (1.0 + 1e-9))) : 1;
        n = std::max(1, std::min(n, 1024));
        g.n[k] = n;
        g.w[k] = L[k] / n;
        ncell *= (size_t)n;
        }
    while (ncell > 4 * (size_t)std::max<uint32_t>(N, 64) + 64)
        {
        int kk = 0;
        for (int k = 1; k < 3; ++k)
            if (g.n[k] > g.n[kk])
                kk = k;
        g.n[kk] = std::max(1, g.n[kk] / 2);
        g.w[kk] = L[kk] / g.n[kk];
        ncell = (size_t)g.n[0] * g.n[1] * g.n[2];
        }
    return g;
    }

// Rank of every cell (geometric index (z * ny + y) * nx + x) along the
// Morton (Z-order) curve of its coordinates.  The cell list stores cells in
// this order, so any run of consecutive slots covers a compact block of
// space whatever order the particles come in.
inline void morton_cell_rank(const CellGrid& g, std::vector<uint32_t>& rank)
    {
    const size_t n = (size_t)g.n[0] * g.n[1] * g.n[2];
    auto spread = [](uint64_t v) {
        v &= 0x1fffffull;
        v = (v | (v << 32)) & 0x1f00000000ffffull;
        v = (v | (v << 16)) & 0x1f0000ff0000ffull;
        v = (v | (v << 8)) & 0x100f00f00f00f00full;
        v = (v | (v << 4)) & 0x10c30c30c30c30c3ull;
        v = (v | (v << 2)) & 0x1249249249249249ull;
        return v;
    };
    std::vector<std::pair<uint64_t, uint32_t>> key(n);
    size_t c = 0;
    for (int z = 0; z < g.n[2]; ++z)
        for (int y = 0; y < g.n[1]; ++y)
            for (int x = 0; x < g.n[0]; ++x, ++c)
                key[c] = {spread((uint64_t)x) | (spread((uint64_t)y) << 1) | (spread((uint64_t)z) << 2),
                          (uint32_t)c};
    std::sort(key.begin(), key.end());
    rank.resize(n);
    for (size_t r = 0; r < n; ++r)
        rank[key[r].second] = (uint32_t)r;
    }

// Cell index along one axis for x in [-L/2, L/2) on an n-cell grid.
__device__ __forceinline__ int cell_coord(double x, double L, double Linv, int n)
    {
#pragma clang fp contract(off)
    const double f = (x + 0.5 * L) * Linv;
    int c = (int)floor(f * (double)n);
    c %= n;
    if (c < 0)
        c += n;
    return c;
    }

// Type index of a particle, or ntypes when pos.w is not a valid type index.
__device__ __forceinline__ uint32_t checked_type(double w, uint32_t ntypes)
    {
    return (w >= 0.0 && w < (double)ntypes) ? (uint32_t)w : ntypes;
    }

// The neighbor cells of cell c along one axis (c-1, c, c+1 with periodic
// wrap), with duplicates removed when the grid is narrower than 3 cells.
// Returns how many distinct cells were written to out.
__device__ __forceinline__ int stencil_axis(int c, int n, int out[3])
    {
    int m = 0;
    for (int d = -1; d <= 1; ++d)
        {
        const int x = (c + d + n) % n;
        bool dup = false;
        for (int k = 0; k < m; ++k)
            dup |= (out[k] == x);
        if (!dup)
            out[m++] = x;
        }
    return m;
    }

// ---------------------------------------------------------------------------
// Exclusive scan of n uint32 counts in three launches: per-chunk sums, a scan
// of the chunk sums by one workgroup, then each chunk's local scan plus its
// offset.  Chunks are kScanChunk = 4 x 1024 counts; up to 1024 chunks per
// pass of the middle kernel, which loops over larger inputs.  out[n] = total.
// Kernels return at once when `gate` is non-null and its bit 0 is clear.
// ---------------------------------------------------------------------------
constexpr uint32_t kScanThreads = 1024;
constexpr uint32_t kScanPerThread = 4;
constexpr uint32_t kScanChunk = kScanThreads * kScanPerThread;

inline uint32_t scan_chunks(uint32_t n)
    {
    return (n + kScanChunk - 1) / kScanChunk;
    }

__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* s, uint32_t& total)
    {
    // Hillis-Steele over kScanThreads values in LDS
    const uint32_t t = threadIdx.x;
    s[t] = v;
    __syncthreads();
    for (uint32_t off = 1; off < kScanThreads; off <<= 1)
        {
        const uint32_t w = (t >= off) ? s[t - off] : 0u;
        __syncthreads();
        s[t] += w;
        __syncthreads();
        }
    total = s[kScanThreads - 1];
    const uint32_t incl = s[t];
    __syncthreads();
    return incl - v;
    }

static __global__ void __launch_bounds__(kScanThreads)
    k_scan_sums(const uint32_t* __restrict__ in, uint32_t n, uint32_t* __restrict__ sums, const uint32_t* gate)
    {
    if (gate && (gate[0] & 1u) == 0u)
        return;
    __shared__ uint32_t s[kScanThreads];
    const uint32_t base = blockIdx.x * kScanChunk + threadIdx.x * kScanPerThread;
    uint32_t v = 0;
    for (uint32_t k = 0; k < kScanPerThread; ++k)
        if (base + k < n)
            v += in[base + k];
    uint32_t total;
    (void)block_exclusive_scan(v, s, total);
    if (threadIdx.x == 0)
        sums[blockIdx.x] = total;
    }

static __global__ void __launch_bounds__(kScanThreads)
    k_scan_offsets(uint32_t* __restrict__ sums, uint32_t nchunks, const uint32_t* gate)
    {
    if (gate && (gate[0] & 1u) == 0u)
        return;
    __shared__ uint32_t s[kScanThreads];
    uint32_t carry = 0;
    for (uint32_t b0 = 0; b0 < nchunks; b0 += kScanThreads)
        {
        const uint32_t k = b0 + threadIdx.x;
        const uint32_t v = k < nchunks ? sums[k] : 0u;
        uint32_t total;
        const uint32_t ex = block_exclusive_scan(v, s, total);
        if (k < nchunks)
            sums[k] = carry + ex;
        carry += total;
        }
    }

static __global__ void __launch_bounds__(kScanThreads) k_scan_apply(const uint32_t* __restrict__ in,
                                                                   uint32_t n,
                                                                   const uint32_t* __restrict__ offsets,
                                                                   uint32_t* __restrict__ out,
                                                                   const uint32_t* gate)
    {
    if (gate && (gate[0] & 1u) == 0u)
        return;
    __shared__ uint32_t s[kScanThreads];
    const uint32_t base = blockIdx.x * kScanChunk + threadIdx.x * kScanPerThread;
    uint32_t v[kScanPerThread];
    uint32_t sum = 0;
    for (uint32_t k = 0; k < kScanPerThread; ++k)
        {
        v[k] = base + k < n ? in[base + k] : 0u;
        sum += v[k];
        }
    uint32_t total;
    uint32_t run = offsets[blockIdx.x] + block_exclusive_scan(sum, s, total);
    for (uint32_t k = 0; k < kScanPerThread; ++k)
        {
        if (base + k < n)
            out[base + k] = run;
        run += v[k];
        }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == kScanThreads - 1)
        out[n] = offsets[blockIdx.x] + total;
    }

// Enqueues the exclusive scan of in[0..n) into out[0..n] (n >= 1); `sums`
// holds scan_chunks(n) values.
inline void enqueue_scan(const uint32_t* in, uint32_t n, uint32_t* sums, uint32_t* out, const uint32_t* gate, hipStream_t s)
    {
    const uint32_t nc = scan_chunks(n);
    hipLaunchKernelGGL(k_scan_sums, dim3(nc), dim3(kScanThreads), 0, s, in, n, sums, gate);
    hipLaunchKernelGGL(k_scan_offsets, dim3(1), dim3(kScanThreads), 0, s, sums, nc, gate);
    hipLaunchKernelGGL(k_scan_apply, dim3(nc), dim3(kScanThreads), 0, s, in, n, sums, out, gate);
    }

} // namespace azp

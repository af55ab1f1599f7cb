// Build of the tiled neighbor list (see tiles.h): one workgroup per tile.
//
//   1. every lane inserts its particle's neighbor indices into an LDS hash
//      set (open addressing, linear probing); insertion stops once the set
//      holds more than the tile capacity, which is then reported,
//   2. the set is compacted and bitonic-sorted in LDS (deterministic order,
//      ascending indices so the force kernel's staging reads are near
//      contiguous), written to jset, and each key's hash slot records its
//      rank,
//   3. every lane rewrites its list entries as ranks, two per word.
#include "device_common.h"
#include "tiles.h"

namespace azp
{

constexpr uint32_t kHashSlots = 7680;   // load factor <= 0.6 at capacity
constexpr uint32_t kSortSlots = 4096;   // >= kTileCap, power of two
constexpr uint32_t kEmpty = 0xffffffffu;
constexpr uint32_t kBatch = 8;          // list words loaded per batch (the padding granule)

__device__ __forceinline__ uint32_t tile_hash(uint32_t j)
    {
    return (uint32_t)(((uint64_t)(j * 2654435761u) * kHashSlots) >> 32);
    }

__device__ __forceinline__ void tile_insert(uint32_t* keys, uint32_t* count, uint32_t j)
    {
    uint32_t h = tile_hash(j);
    for (uint32_t probe = 0; probe < kHashSlots; ++probe)
        {
        const uint32_t prev = atomicCAS(&keys[h], kEmpty, j);
        if (prev == kEmpty)
            {
            atomicAdd(count, 1u);
            return;
            }
        if (prev == j)
            return;
        h = (h + 1 == kHashSlots) ? 0u : h + 1;
        }
    }

__device__ __forceinline__ uint32_t tile_find(const uint32_t* keys, uint32_t j)
    {
    uint32_t h = tile_hash(j);
    for (uint32_t probe = 0; probe < kHashSlots; ++probe)
        {
        if (keys[h] == j)
            return h;
        h = (h + 1 == kHashSlots) ? 0u : h + 1;
        }
    return 0u; // not reached: every listed index was inserted
    }

// Eight entries k0 .. k0 + 7 of particle i from its particle-major row.
__device__ __forceinline__ void row_chunk(const uint32_t* rows, uint32_t stride, uint32_t i, uint32_t k0, uint32_t v[8])
    {
    const uint4* r = reinterpret_cast<const uint4*>(rows + (size_t)i * stride + k0);
    const uint4 a = r[0], b = r[1];
    v[0] = a.x;
    v[1] = a.y;
    v[2] = a.z;
    v[3] = a.w;
    v[4] = b.x;
    v[5] = b.y;
    v[6] = b.z;
    v[7] = b.w;
    }

__global__ void __launch_bounds__(512) k_tile_build(uint32_t N,
                                                    const uint32_t* __restrict__ n_neigh,
                                                    const uint32_t* __restrict__ rows,
                                                    uint32_t stride,
                                                    uint32_t pitch,
                                                    uint32_t* __restrict__ tile_count,
                                                    uint32_t* __restrict__ tile_jset,
                                                    uint32_t* __restrict__ tlist,
                                                    uint32_t* __restrict__ overflow)
    {
    __shared__ uint32_t keys[kHashSlots];
    __shared__ uint16_t rank[kHashSlots];
    __shared__ uint32_t sorted[kSortSlots];
    __shared__ uint32_t s_count, s_next;
    const uint32_t T = blockDim.x;
    const uint32_t tid = threadIdx.x;
    const uint32_t t = blockIdx.x;
    for (uint32_t k = tid; k < kHashSlots; k += T)
        keys[k] = kEmpty;
    if (tid == 0)
        {
        s_count = 0u;
        s_next = 0u;
        }
    __syncthreads();

    const uint32_t i = t * T + tid;
    const uint32_t nn = i < N ? n_neigh[i] : 0u;
    volatile uint32_t* vcount = &s_count;
    // list words are loaded 8 at a time (independent loads in flight), then inserted
    for (uint32_t k0 = 0; k0 < nn; k0 += kBatch)
        {
        if (*vcount >= kTileCap)
            break;
        uint32_t jj[kBatch];
        row_chunk(rows, stride, i, k0, jj);
#pragma unroll
        for (uint32_t u = 0; u < kBatch; ++u)
            if (k0 + u < nn)
                tile_insert(keys, &s_count, jj[u]);
        }
    __syncthreads();
    const uint32_t M = s_count;
    if (M > kTileCap - 1)
        {
        if (tid == 0)
            atomicMax(overflow, M);
        return; // the whole workgroup leaves (M is the same for every lane)
        }
    if (tid == 0)
        atomicMax(overflow + 1, M);

    for (uint32_t k = tid; k < kHashSlots; k += T)
        {
        const uint32_t key = keys[k];
        if (key != kEmpty)
            sorted[atomicAdd(&s_next, 1u)] = key;
        }
    for (uint32_t k = M + tid; k < kSortSlots; k += T)
        sorted[k] = kEmpty;
    __syncthreads();
    for (uint32_t size = 2; size <= kSortSlots; size <<= 1)
        for (uint32_t stride = size >> 1; stride > 0; stride >>= 1)
            {
            for (uint32_t e = tid; e < kSortSlots; e += T)
                {
                const uint32_t p = e ^ stride;
                if (p > e)
                    {
                    const bool up = (e & size) == 0u;
                    const uint32_t a = sorted[e], b = sorted[p];
                    if ((a > b) == up)
                        {
                        sorted[e] = b;
                        sorted[p] = a;
                        }
                    }
                }
            __syncthreads();
            }
    for (uint32_t m = tid; m < M; m += T)
        {
        const uint32_t key = sorted[m];
        tile_jset[(size_t)t * kTileCap + m] = key;
        rank[tile_find(keys, key)] = (uint16_t)m;
        }
    if (tid == 0)
        tile_count[t] = M;
    __syncthreads();

    // entries as ranks, two per word, padded to the wavefront's longest list
    // rounded up to 8
    uint32_t wmax = nn;
    for (int off = 32; off > 0; off >>= 1)
        wmax = max(wmax, (uint32_t)__shfl_xor((int)wmax, off));
    const uint32_t P = (wmax + 7u) & ~7u;
    if (i >= N)
        return;
    for (uint32_t k0 = 0; k0 < P; k0 += kBatch)
        {
        uint32_t jj[kBatch];
        if (k0 < nn)
            row_chunk(rows, stride, i, k0, jj);
        uint32_t e[kBatch];
#pragma unroll
        for (uint32_t u = 0; u < kBatch; ++u)
            e[u] = (k0 + u < nn ? (uint32_t)rank[tile_find(keys, jj[u])] : M) << 4; // 16 x rank
#pragma unroll
        for (uint32_t w = 0; w < kBatch / 2; ++w)
            tlist[(size_t)((k0 >> 1) + w) * pitch + i] = e[2 * w] | (e[2 * w + 1] << 16);
        }
    }

__global__ void k_tile_flag_clear(uint32_t* flag)
    {
    if (blockIdx.x == 0 && threadIdx.x < 2)
        flag[threadIdx.x] = 0u;
    }

int tiles_build(TileList& t,
                uint32_t tile_size,
                uint32_t N,
                const uint32_t* d_n_neigh,
                const uint32_t* d_rows,
                uint32_t pitch,
                uint32_t max_neighbors,
                uint32_t* h_flag,
                hipStream_t s)
    {
    t.valid = false;
    t.size = 0;
    t.order = nullptr;
    if (tile_size != 256 && tile_size != 512)
        return set_error(AZP_ERR_INVALID, "tile size %u must be 256 or 512", tile_size);
    if (N == 0)
        return AZP_OK;
    const uint32_t ntiles = (N + tile_size - 1) / tile_size;
    // words per column: the padded length (<= max_neighbors rounded up to 8) over 2
    const uint32_t words = (((max_neighbors + 7u) & ~7u) + 1u) / 2u;
    AZP_TRY(t.count.reserve(ntiles));
    AZP_TRY(t.jset.reserve((size_t)ntiles * kTileCap));
    AZP_TRY(t.list.reserve((size_t)words * pitch));
    AZP_TRY(t.flag.reserve(2));
    hipLaunchKernelGGL(k_tile_flag_clear, dim3(1), dim3(64), 0, s, t.flag.ptr);
    hipLaunchKernelGGL(k_tile_build, dim3(ntiles), dim3(tile_size), 0, s, N, d_n_neigh, d_rows, max_neighbors, pitch,
                       t.count.ptr, t.jset.ptr, t.list.ptr, t.flag.ptr);
    AZP_LAUNCH_CHECK();
    AZP_HIP(hipMemcpyAsync(h_flag, t.flag.ptr, 2 * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    AZP_HIP(hipStreamSynchronize(s));
    if (h_flag[0] != 0u)
        return AZP_OK; // some tile has too many distinct neighbors
    t.ntiles = ntiles;
    t.size = tile_size;
    t.slots = h_flag[1] + 1u;
    t.valid = true;
    return AZP_OK;
    }

} // namespace azp

// Device-side building blocks: minimum image, precision-generic fast math,
// the Philox4x32-10 generator with the HOOMD-style seed/counter packing.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "azp_internal.h"

namespace azp
{

// ---------------------------------------------------------------------------
// Minimum image in FP64: img = rint(d * Linv); d -= L * img.
// Evaluated without contraction, as the host oracle evaluates it.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double min_image_1d(double d, double L, double Linv)
    {
#pragma clang fp contract(off)
    const double img = rint(d * Linv);
    return d - L * img;
    }

// The same minimum image with one rounding: for positions inside the box
// |rint(d / L)| <= 2, so L rint(d / L) is exact and the fused form rounds
// d - L rint(d / L) as the two-step form does; the pair and DPD kernels use
// it, bit-identical to the oracle's geometry.
__device__ __forceinline__ double min_image_fused(double d, double L, double Linv)
    {
    return __builtin_fma(-L, rint(d * Linv), d);
    }

// The largest x, y, z and -x, -y, -z (in FP32) of the positions a tile kernel
// stages, reduced over its workgroup; wrap_axes() then tells, per axis,
// whether a difference of two of them may need the periodic image.
struct TileExtent
    {
    float m[6];
    __device__ TileExtent()
        {
#pragma unroll
        for (uint32_t c = 0; c < 6; ++c)
            m[c] = -__builtin_huge_valf();
        }
    __device__ void add(double x, double y, double z)
        {
        const float fx = (float)x, fy = (float)y, fz = (float)z;
        m[0] = fmaxf(m[0], fx);
        m[1] = fmaxf(m[1], fy);
        m[2] = fmaxf(m[2], fz);
        m[3] = fmaxf(m[3], -fx);
        m[4] = fmaxf(m[4], -fy);
        m[5] = fmaxf(m[5], -fz);
        }
    // each wavefront's maxima to s_ext[wave * 8 + c]; before a barrier
    __device__ void store_wave(float* s_ext) const
        {
#pragma unroll
        for (uint32_t c = 0; c < 6; ++c)
            {
            float v = m[c];
            for (int off = 32; off > 0; off >>= 1)
                v = fmaxf(v, __shfl_xor(v, off));
            if ((threadIdx.x & 63u) == 0)
                s_ext[(threadIdx.x >> 6) * 8u + c] = v;
            }
        }
    };

// After the barrier: bit a set when the workgroup's positions span 0.45 L or
// more on axis a.  A clear bit means every difference d of two of them has
// |d / L| < 0.46, rint(d / L) = 0, and the minimum image on that axis is d
// (up to the sign of a zero difference).  Wave-uniform (a scalar register).
__device__ __forceinline__ uint32_t wrap_axes(const float* s_ext, const double* L)
    {
    float m[6];
#pragma unroll
    for (uint32_t c = 0; c < 6; ++c)
        m[c] = -__builtin_huge_valf();
    for (uint32_t w = 0; w < blockDim.x / 64u; ++w)
        {
#pragma unroll
        for (uint32_t c = 0; c < 6; ++c)
            m[c] = fmaxf(m[c], s_ext[w * 8u + c]);
        }
    uint32_t wrap = 0u;
#pragma unroll
    for (uint32_t c = 0; c < 3; ++c)
        {
        if (!(m[c] + m[c + 3] < 0.45f * (float)L[c]))
            wrap |= 1u << c;
        }
    return __builtin_amdgcn_readfirstlane(wrap);
    }

// Bytes of LDS the per-wavefront maxima take (8 floats for each of up to 8
// wavefronts).
constexpr size_t kExtentBytes = 8 * 8 * sizeof(float);

__device__ __forceinline__ double rsq3_nofma(double x, double y, double z)
    {
#pragma clang fp contract(off)
    return x * x + y * y + z * z;
    }

// ---------------------------------------------------------------------------
// XCD-aware workgroup order.  gfx950 deals workgroups round-robin to its 8
// XCDs (workgroup b runs on XCD b % 8), and each XCD has its own L2.  This
// maps hardware workgroup b to a logical workgroup so that every XCD owns one
// contiguous eighth of the logical range: with particles in space-filling-curve
// order, the neighbor gathers of an XCD stay within a compact region that fits
// its L2.  The map is a bijection of [0, gridDim.x).
// ---------------------------------------------------------------------------
constexpr uint32_t kNumXcd = 8;

__device__ __forceinline__ uint32_t xcd_logical_block()
    {
    const uint32_t nb = gridDim.x;
    const uint32_t b = blockIdx.x;
    const uint32_t x = b % kNumXcd;
    const uint32_t idx = b / kNumXcd;
    const uint32_t q = nb / kNumXcd;
    const uint32_t r = nb % kNumXcd;
    return x * q + min(x, r) + idx;
    }

// ---------------------------------------------------------------------------
// Precision-generic math.  FP64 keeps the IEEE operations of the reference
// expressions; FP32 uses the hardware rcp/sqrt/exp/log units (about 1 ulp).
// ---------------------------------------------------------------------------
// 1 / x correctly rounded: the hardware reciprocal, two Newton steps and a
// final correction, the steps of the compiler's IEEE division of 1 by x
// without its operand scaling (exact powers of two, which change no bit of
// the result for a normal x); the fixup step gives 1/0, 1/inf and NaN their
// IEEE values.
__device__ __forceinline__ double f_rcp(double x)
    {
    double r = __builtin_amdgcn_rcp(x);
    r = __builtin_fma(r, __builtin_fma(-x, r, 1.0), r);
    r = __builtin_fma(r, __builtin_fma(-x, r, 1.0), r);
    const double q = __builtin_fma(__builtin_fma(-x, r, 1.0), r, r);
    return __builtin_amdgcn_div_fixup(q, x, 1.0);
    }
__device__ __forceinline__ float f_rcp(float x)
    {
    return __builtin_amdgcn_rcpf(x);
    }
__device__ __forceinline__ double f_div(double a, double b)
    {
    return a / b;
    }
__device__ __forceinline__ float f_div(float a, float b)
    {
    return a * __builtin_amdgcn_rcpf(b);
    }
__device__ __forceinline__ double f_sqrt(double x)
    {
    return sqrt(x);
    }
__device__ __forceinline__ float f_sqrt(float x)
    {
    return __builtin_amdgcn_sqrtf(x);
    }
__device__ __forceinline__ double f_rsqrt(double x)
    {
    return f_rcp(sqrt(x)); // the same bits as 1.0 / sqrt(x)
    }
__device__ __forceinline__ float f_rsqrt(float x)
    {
    return __builtin_amdgcn_rsqf(x);
    }
__device__ __forceinline__ double f_exp(double x)
    {
    return exp(x);
    }
__device__ __forceinline__ float f_exp(float x)
    {
    return __expf(x);
    }
__device__ __forceinline__ double f_log(double x)
    {
    return log(x);
    }
__device__ __forceinline__ float f_log(float x)
    {
    return __logf(x);
    }
__device__ __forceinline__ double f_max(double a, double b)
    {
    return fmax(a, b);
    }
__device__ __forceinline__ float f_max(float a, float b)
    {
    return fmaxf(a, b);
    }

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11): 10 rounds of two
// 32x32->64 multiplies with Weyl key bumps.  Each 32x32->64 product maps to
// one v_mad_u64_u32.
// ---------------------------------------------------------------------------
struct u32x4
    {
    uint32_t v[4];
    };

__device__ __forceinline__ u32x4 philox4x32_10(u32x4 c, uint32_t k0, uint32_t k1)
    {
#pragma unroll
    for (int round = 0; round < 10; ++round)
        {
        if (round > 0)
            {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
            }
        const uint64_t p0 = (uint64_t)0xD2511F53u * (uint64_t)c.v[0];
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * (uint64_t)c.v[2];
        u32x4 n;
        n.v[0] = (uint32_t)(p1 >> 32) ^ c.v[1] ^ k0;
        n.v[1] = (uint32_t)p1;
        n.v[2] = (uint32_t)(p0 >> 32) ^ c.v[3] ^ k1;
        n.v[3] = (uint32_t)p0;
        c = n;
        }
    return c;
    }

// RNG identifier of the DPD general-weight evaluator (RNGIdentifiers.h:23).
constexpr uint8_t kRngDPDGeneralWeight = 200;
// Stream used by this library's Langevin integrator (not a reference stream).
constexpr uint8_t kRngLangevin = 210;

// Seed(id, timestep, seed) -> Philox key, as in upstream hoomd/RandomNumbers.h:
//   key0 = id << 24 | seed << 8 | (timestep >> 32) & 0xff, key1 = timestep low word.
__device__ __forceinline__ void hoomd_seed_key(uint8_t id,
                                               uint64_t timestep,
                                               uint16_t seed,
                                               uint32_t& k0,
                                               uint32_t& k1)
    {
    k0 = ((uint32_t)id << 24) | ((uint32_t)seed << 8) | (uint32_t)((timestep >> 32) & 0xffu);
    k1 = (uint32_t)(timestep & 0xffffffffu);
    }

// UniformDistribution<double>(-1, 1) of the first Philox block:
//   u64 = out0 << 32 | out1;  U = u64 * 2^-64 + 2^-65;  alpha = -1 + 2 U.
// Every step after the u64 -> double conversion is exact (powers of two).
__device__ __forceinline__ double uniform_m1_1(const u32x4& u)
    {
#pragma clang fp contract(off)
    const uint64_t u64 = ((uint64_t)u.v[0] << 32) | (uint64_t)u.v[1];
    const double U = (double)u64 * (1.0 / 18446744073709551616.0) + (0.5 / 18446744073709551616.0);
    return -1.0 + 2.0 * U;
    }

// DPD alpha for the (unordered) tag pair at (seed, timestep):
// DPDPairEvaluatorGeneralWeight.h:213-233.  The evaluator stores the
// timestep as 32 bits (:122,:273), so its high word is always zero.
__device__ __forceinline__ double dpd_alpha(uint16_t seed, uint32_t timestep, uint32_t ti, uint32_t tj)
    {
    const uint32_t oi = ti > tj ? tj : ti;
    const uint32_t oj = ti > tj ? ti : tj;
    uint32_t k0, k1;
    hoomd_seed_key(kRngDPDGeneralWeight, (uint64_t)timestep, seed, k0, k1);
    u32x4 c;
    c.v[0] = oi;
    c.v[1] = oj;
    c.v[2] = 0u;
    c.v[3] = 0u;
    return uniform_m1_1(philox4x32_10(c, k0, k1));
    }

// Fixed-point position helpers for the mixed form: coordinate x in
// [-L/2, L/2) maps to the wrapped int32 round(x / L * 2^32).  Differences of
// two such ints (mod 2^32) are minimum-image displacements in units of L/2^32.
__device__ __forceinline__ int32_t to_fixed(double x, double Linv)
    {
    double f = x * Linv;           // in [-1/2, 1/2)
    f = f - rint(f);               // wrap anything slightly outside
    double s = rint(f * 4294967296.0);
    long long v = (long long)s;    // in [-2^31, 2^31]
    return (int32_t)(uint32_t)(unsigned long long)v;
    }

__device__ __forceinline__ int32_t fixed_delta(int32_t a, int32_t b)
    {
    return (int32_t)((uint32_t)a - (uint32_t)b);
    }

// Stages `count` entries of a tile's neighbor set into LDS: each lane of the
// workgroup takes slots threadIdx.x, + blockDim.x, ..., K at a time, with the
// K index loads and then the K data loads all issued before the first store,
// so a tile's staging waits on two memory latencies per K slots instead of
// two per slot (the whole workgroup waits for its staging at a barrier).
// load(j) reads neighbor j's data, store(m, v) writes slot m.
template<uint32_t K, class Load, class Store>
__device__ __forceinline__ void stage_slots(uint32_t count, const uint32_t* js, Load&& load, Store&& store)
    {
    const uint32_t bs = blockDim.x;
    for (uint32_t base = threadIdx.x; base < count; base += K * bs)
        {
        uint32_t j[K];
#pragma unroll
        for (uint32_t k = 0; k < K; ++k)
            {
            const uint32_t m = base + k * bs;
            j[k] = m < count ? js[m] : 0u;
            }
        decltype(load(0u)) v[K];
#pragma unroll
        for (uint32_t k = 0; k < K; ++k)
            {
            if (base + k * bs < count)
                v[k] = load(j[k]);
            }
#pragma unroll
        for (uint32_t k = 0; k < K; ++k)
            {
            const uint32_t m = base + k * bs;
            if (m < count)
                store(m, v[k]);
            }
        }
    }

} // namespace azp

// Isotropic pair-force kernels over a full neighbor list (gfx950): the
// kernel template and its launch helpers, shared by the translation units
// that instantiate the column-layout and tiled forms (k_pair.hip) and the
// head-list form (k_pair_head.hip).
//
// Replaces the launcher instantiated per evaluator at
// PotentialPairGPUKernel.cu.inc:25-28 (gpu_compute_pair_forces<Eval>).
//
// Layout: one lane per particle i, walking its full neighbor list
// nlist[k * pitch + i] — consecutive lanes read consecutive words, so the
// list stream is fully coalesced.  Per-type-pair constants (cutoff, hoisted
// energy shift, parameter products) live in LDS; with one particle type they
// are held in registers.  Force, energy and the six virial components are
// accumulated in FP64 registers and written once (no atomics; every lane owns
// its particle).
//
// Two precision forms:
//   plain (FP64):  FP64 positions, min image, evaluator and accumulation.
//   mixed:         positions as wrapped 32-bit fixed point (16-byte gather),
//                  dx exact in integers then FP32, FP32 evaluator; every 32
//                  list entries are summed in FP32 and added to FP64
//                  accumulators.  Pairs whose FP32 rsq lies within a relative
//                  band of 1e-5 around rcutsq are re-tested on the FP64
//                  positions, so the in/out cutoff decision is the FP64 one.
#pragma once

#include <algorithm>
#include <type_traits>

#include "evaluators.h"
#include "tiles.h"

namespace azp
{

constexpr float kCutBand = 1e-5f;
// Mixed form: list entries summed in FP32 between additions to the FP64
// accumulators (a multiple of the batch of 8).
constexpr uint32_t kFlushEntries = 32;

struct PairKernelArgs
    {
    const double4* pos;
    const int4* pos_fixed;
    uint32_t N;
    DevBox box;
    float fx_scale[3];    // L / 2^32 per axis
    double fx_scale64[3]; // the same in FP64
    float fx_scale2[3];   // (L / 2^32)^2 per axis
    const uint32_t* n_neigh;
    const uint32_t* nlist;
    uint32_t nlist_pitch;
    const void* params;
    const double* rcutsq;
    const double* ronsq;
    uint32_t ntypes;
    int shift_mode;
    double4* force;
    double* virial;
    uint32_t vpitch;
    // tiled list (mixed form): per-tile neighbor sets and 16-bit entries;
    // for tiles of cell-list slots, column s is particle tile_order[s] with
    // tile_nn[s] entries (both null for tiles of consecutive particles)
    const uint32_t* tile_count;
    const uint32_t* tile_jset;
    const uint32_t* tile_list;
    const uint32_t* tile_order;
    const uint32_t* tile_nn;
    uint32_t tile_slots; // LDS position slots of the tiled kernel (the tables follow them)
    // head-list layout: particle i's entries at nlist[head_list[i] + k], k < n_neigh[i]
    const uint64_t* head_list;
    // mixed form: add every entry's terms to the FP64 accumulators
    bool acc64;
    };

// Position (double4 x, y, z, type) -> wrapped 32-bit fixed point of the box.
__device__ __forceinline__ int4 fixed4(const double4& p, const DevBox& box)
    {
    return make_int4(to_fixed(p.x, box.Linv[0]), to_fixed(p.y, box.Linv[1]), to_fixed(p.z, box.Linv[2]), (int)p.w);
    }

template<typename R> struct TypePairExtra
    {
    double rcsq64;
    R rcsq;
    R rcsq_lo; // rcsq (1 - band)
    R rcsq_hi; // rcsq (1 + band)
    R ronsq;
    };

template<class Eval, typename R>
__device__ __forceinline__ void prepare_pair(const PairKernelArgs& a,
                                             uint32_t t,
                                             typename Eval::template Prep<R>& q,
                                             TypePairExtra<R>& x)
    {
    const auto* params = static_cast<const typename Eval::param_type*>(a.params);
    const double rc = a.rcutsq[t];
    const double ron = a.ronsq ? a.ronsq[t] : 0.0;
    const bool shift
        = a.shift_mode == AZP_SHIFT_SHIFT || (a.shift_mode == AZP_SHIFT_XPLOR && ron > rc);
    q = Eval::template prepare<R>(params[t], rc, shift);
    x.rcsq64 = rc;
    x.rcsq = (R)rc;
    x.rcsq_lo = (R)(rc * (1.0 - (double)kCutBand));
    x.rcsq_hi = (R)(rc * (1.0 + (double)kCutBand));
    x.ronsq = (R)ron;
    }

// xplor smoothing (upstream PotentialPair, restated in the oracle):
// S = (rc2 - r2)^2 (rc2 + 2 r2 - 3 ron2) / (rc2 - ron2)^3 for ron2 <= r2 < rc2.
template<typename R>
__device__ __forceinline__ void apply_xplor(R rsq, R rcutsq, R ronsq, R& fdivr, R& e)
    {
    if (rsq >= ronsq && rsq < rcutsq)
        {
        const R old_e = e, old_f = fdivr;
        const R denom = (rcutsq - ronsq) * (rcutsq - ronsq) * (rcutsq - ronsq);
        const R denom_inv = f_rcp(denom);
        const R d = rsq - rcutsq;
        const R S = d * d * (rcutsq + R(2) * rsq - R(3) * ronsq) * denom_inv;
        const R dS_dr_divr = R(12) * (rsq - ronsq) * d * denom_inv;
        e = old_e * S;
        fdivr = S * old_f - dS_dr_divr * old_e;
        }
    }

struct Acc
    {
    double fx = 0, fy = 0, fz = 0, e = 0;
    double v0 = 0, v1 = 0, v2 = 0, v3 = 0, v4 = 0, v5 = 0;
    };

template<bool VIRIAL>
__device__ __forceinline__ void accumulate(Acc& acc, double fd, double dx, double dy, double dz, double e)
    {
    const double px = fd * dx, py = fd * dy, pz = fd * dz;
    acc.fx += px;
    acc.fy += py;
    acc.fz += pz;
    acc.e += e;
    if (VIRIAL)
        {
        acc.v0 = fma(px, dx, acc.v0);
        acc.v1 = fma(px, dy, acc.v1);
        acc.v2 = fma(px, dz, acc.v2);
        acc.v3 = fma(py, dy, acc.v3);
        acc.v4 = fma(py, dz, acc.v4);
        acc.v5 = fma(pz, dz, acc.v5);
        }
    }

// Mixed form: the terms of two batches of list entries are summed in FP32 (at
// most 16 terms), then each such sum is added to the FP64 accumulators.
struct Acc32
    {
    float fx = 0, fy = 0, fz = 0, e = 0;
    float v0 = 0, v1 = 0, v2 = 0, v3 = 0, v4 = 0, v5 = 0;
    };

template<bool VIRIAL>
__device__ __forceinline__ void accumulate32(Acc32& acc, float fd, float dx, float dy, float dz, float e)
    {
    const float px = fd * dx, py = fd * dy, pz = fd * dz;
    acc.fx += px;
    acc.fy += py;
    acc.fz += pz;
    acc.e += e;
    if (VIRIAL)
        {
        acc.v0 = fmaf(px, dx, acc.v0);
        acc.v1 = fmaf(px, dy, acc.v1);
        acc.v2 = fmaf(px, dz, acc.v2);
        acc.v3 = fmaf(py, dy, acc.v3);
        acc.v4 = fmaf(py, dz, acc.v4);
        acc.v5 = fmaf(pz, dz, acc.v5);
        }
    }

template<bool VIRIAL>
__device__ __forceinline__ void flush32(Acc& acc, const Acc32& b)
    {
    acc.fx += (double)b.fx;
    acc.fy += (double)b.fy;
    acc.fz += (double)b.fz;
    acc.e += (double)b.e;
    if (VIRIAL)
        {
        acc.v0 += (double)b.v0;
        acc.v1 += (double)b.v1;
        acc.v2 += (double)b.v2;
        acc.v3 += (double)b.v3;
        acc.v4 += (double)b.v4;
        acc.v5 += (double)b.v5;
        }
    }

template<bool VIRIAL>
__device__ __forceinline__ void store(const PairKernelArgs& a, uint32_t i, const Acc& acc)
    {
    a.force[i] = make_double4(acc.fx, acc.fy, acc.fz, 0.5 * acc.e);
    if (VIRIAL)
        {
        const size_t p = a.vpitch;
        a.virial[0 * p + i] = 0.5 * acc.v0;
        a.virial[1 * p + i] = 0.5 * acc.v1;
        a.virial[2 * p + i] = 0.5 * acc.v2;
        a.virial[3 * p + i] = 0.5 * acc.v3;
        a.virial[4 * p + i] = 0.5 * acc.v4;
        a.virial[5 * p + i] = 0.5 * acc.v5;
        }
    }

// FP64 displacement with minimum image, no contraction (matches the oracle).
__device__ __forceinline__ double exact_rsq(const PairKernelArgs& a, uint32_t i, uint32_t j)
    {
    const double4 pi = a.pos[i];
    const double4 pj = a.pos[j];
    const double dx = min_image_1d(pi.x - pj.x, a.box.L[0], a.box.Linv[0]);
    const double dy = min_image_1d(pi.y - pj.y, a.box.L[1], a.box.Linv[1]);
    const double dz = min_image_1d(pi.z - pj.z, a.box.L[2], a.box.Linv[2]);
    return rsq3_nofma(dx, dy, dz);
    }

// Neighbor entries processed per batch: the batch's list words are loaded
// first, then all its gathers are issued, then the pair terms are evaluated,
// so each lane keeps U gathers in flight.  The neighbor list pads every
// wavefront's columns with the particle's own index to a multiple of 8, so the
// batch loop is wave-uniform and padding entries are masked by j == i.
template<bool MIXED> struct Unroll
    {
    static constexpr uint32_t value = MIXED ? 8 : 4;
    };

__device__ __forceinline__ uint32_t wave_max(uint32_t v)
    {
    for (int off = 32; off > 0; off >>= 1)
        v = max(v, (uint32_t)__shfl_xor((int)v, off));
    return v;
    }

__device__ __forceinline__ uint32_t wave_min(uint32_t v)
    {
    for (int off = 32; off > 0; off >>= 1)
        v = min(v, (uint32_t)__shfl_xor((int)v, off));
    return v;
    }

// U consecutive entries of one lane's head-list row (U a multiple of 4) as
// 16-byte loads; a row starts at any 4-byte boundary, and the copy lets the
// compiler pick the loads that alignment allows.
template<uint32_t U> __device__ __forceinline__ void head_words(const uint32_t* p, uint32_t (&jj)[U])
    {
    static_assert(U % 4 == 0, "whole 16-byte groups");
#pragma unroll
    for (uint32_t u = 0; u < U; u += 4)
        {
        uint4 w;
        __builtin_memcpy(&w, p + u, sizeof(w));
        jj[u] = w.x;
        jj[u + 1] = w.y;
        jj[u + 2] = w.z;
        jj[u + 3] = w.w;
        }
    }

// Accumulators of the mixed form are kept in integer-grid units of the
// fixed-point positions (dx = n * L / 2^32): sum f n and sum f n_a n_b are
// formed from the integer displacements (as FP32) and scaled once at the end.
template<bool VIRIAL>
__device__ __forceinline__ void store_scaled(const PairKernelArgs& a, uint32_t i, const Acc& acc)
    {
    const double sx = a.fx_scale64[0], sy = a.fx_scale64[1], sz = a.fx_scale64[2];
    a.force[i] = make_double4(acc.fx * sx, acc.fy * sy, acc.fz * sz, 0.5 * acc.e);
    if (VIRIAL)
        {
        const size_t p = a.vpitch;
        a.virial[0 * p + i] = 0.5 * acc.v0 * (sx * sx);
        a.virial[1 * p + i] = 0.5 * acc.v1 * (sx * sy);
        a.virial[2 * p + i] = 0.5 * acc.v2 * (sx * sz);
        a.virial[3 * p + i] = 0.5 * acc.v3 * (sy * sy);
        a.virial[4 * p + i] = 0.5 * acc.v4 * (sy * sz);
        a.virial[5 * p + i] = 0.5 * acc.v5 * (sz * sz);
        }
    }

// Branchless evaluation of one list entry (FP64 form): padding entries and
// pairs beyond the cutoff are evaluated at a safe distance and contribute zero.
template<class Eval, bool VIRIAL, bool XPLOR>
__device__ __forceinline__ void term_fp64(const typename Eval::template Prep<double>& q,
                                          const TypePairExtra<double>& x,
                                          bool valid,
                                          double dx,
                                          double dy,
                                          double dz,
                                          Acc& acc)
    {
    const double rsq = dx * dx + dy * dy + dz * dz;
    const bool in = valid && rsq < x.rcsq64 && Eval::active(q);
    const double rs = in ? rsq : x.rcsq64;
    double fdivr, e;
    Eval::compute(q, rs, fdivr, e);
    if constexpr (XPLOR)
        apply_xplor<double>(rs, x.rcsq64, x.ronsq, fdivr, e);
    fdivr = in ? fdivr : 0.0;
    e = in ? e : 0.0;
    accumulate<VIRIAL>(acc, fdivr, dx, dy, dz, e);
    }

// One list entry of particle i in the mixed form.  `valid` is false for
// padding; `jkey` is the neighbor's index, or its rank in the tile's neighbor
// set `jmap` (tiled list), needed only for the FP64 re-test near the cutoff.
template<class Eval, bool VIRIAL, bool XPLOR, bool CUBIC>
__device__ __forceinline__ void term_mixed(const PairKernelArgs& a,
                                           const typename Eval::template Prep<float>& q,
                                           const TypePairExtra<float>& x,
                                           bool valid,
                                           uint32_t i,
                                           uint32_t jkey,
                                           const uint32_t* jmap,
                                           int32_t ix,
                                           int32_t iy,
                                           int32_t iz,
                                           Acc32& acc)
    {
    const float fx = (float)ix, fy = (float)iy, fz = (float)iz;
    float rsq;
    if constexpr (CUBIC)
        rsq = (fx * fx + fy * fy + fz * fz) * a.fx_scale2[0];
    else
        rsq = fx * fx * a.fx_scale2[0] + fy * fy * a.fx_scale2[1] + fz * fz * a.fx_scale2[2];
    // two compares give the three outcomes (inside, the FP64 band, outside) as
    // lane masks from the compare builtins (float <)
    const uint64_t below_lo = __builtin_amdgcn_fcmpf(rsq, x.rcsq_lo, 4);
    const uint64_t below_hi = __builtin_amdgcn_fcmpf(rsq, x.rcsq_hi, 4);
    bool in = valid && __builtin_amdgcn_inverse_ballot_w64(below_lo);
    double r64 = -1.0;
    if (valid && __builtin_amdgcn_inverse_ballot_w64(below_hi & ~below_lo))
        {
        // FP32 cannot decide this one: apply the FP64 cutoff test
        const uint32_t j = jmap ? jmap[jkey] : jkey;
        r64 = exact_rsq(a, i, j);
        in = r64 < x.rcsq64;
        rsq = (float)r64;
        }
    in = in && Eval::active(q);
    float fdivr, e;
    if constexpr (HasMasked<Eval>::value && !XPLOR)
        {
        // masked entries at r^2 = +inf (the evaluator's terms vanish there)
        Eval::compute_masked(q, in ? rsq : __builtin_huge_valf(), in, fdivr, e);
        accumulate32<VIRIAL>(acc, fdivr, fx, fy, fz, e);
        return;
        }
    const float rs = in ? rsq : x.rcsq;
    if constexpr (NeedsRsq64<Eval>::value)
        {
        if (r64 < 0.0)
            {
            const double ex = (double)ix * a.fx_scale64[0];
            const double ey = (double)iy * a.fx_scale64[1];
            const double ez = (double)iz * a.fx_scale64[2];
            r64 = ex * ex + ey * ey + ez * ez;
            }
        Eval::compute_precise(q, rs, in ? r64 : x.rcsq64, fdivr, e);
        }
    else
        {
        Eval::compute(q, rs, fdivr, e);
        }
    if constexpr (XPLOR)
        apply_xplor<float>(rs, x.rcsq, x.ronsq, fdivr, e);
    fdivr = in ? fdivr : 0.0f;
    e = in ? e : 0.0f;
    accumulate32<VIRIAL>(acc, fdivr, fx, fy, fz, e);
    }

// LDS bytes of the per-type-pair tables, rounded up to 16 (the tile's
// positions follow them).
template<class Eval, typename R> __host__ __device__ constexpr size_t table_bytes(uint32_t nt2)
    {
    return (nt2 * (sizeof(typename Eval::template Prep<R>) + sizeof(TypePairExtra<R>)) + 15) & ~(size_t)15;
    }

// LDS the tiled kernel may use for its type-pair tables and neighbor slots.
constexpr size_t kTileLds = 64 * 1024;

// LDS of the FP64 form over the tiled list: 32-byte position slots, up to the
// 160 KiB a gfx950 workgroup may use.
constexpr size_t kTileLds64 = 160 * 1024;

template<class Eval> static bool tile_fits(uint32_t ntypes, uint32_t slots)
    {
    const size_t tables = ntypes == 1 ? 0 : table_bytes<Eval, float>(ntypes * ntypes);
    return tables + (size_t)slots * sizeof(int4) <= kTileLds;
    }

// Bytes of one staged neighbor in the FP64 form over the tiled list: x, y, z
// with one particle type (so that two workgroups of a tile at the 1M workload
// fit in a CU's LDS), x, y, z, type otherwise.
__host__ __device__ constexpr size_t fp64_slot_bytes(bool single)
    {
    return single ? 3 * sizeof(double) : sizeof(double4);
    }

template<class Eval> static bool tile_fits64(uint32_t ntypes, uint32_t slots)
    {
    const size_t tables = ntypes == 1 ? 0 : table_bytes<Eval, double>(ntypes * ntypes);
    return tables + kExtentBytes + (size_t)slots * fp64_slot_bytes(ntypes == 1) <= kTileLds64;
    }


// Neighbor-list layouts a kernel instance reads.
constexpr int kLayoutColumn = 0; // nlist[k * pitch + i], padded with i
constexpr int kLayoutTiled = 1;  // per-tile 16-bit entries (mixed form)
constexpr int kLayoutHead = 2;   // HOOMD's head list: nlist[head_list[i] + k], k < n_neigh[i]

// ACC64 (mixed form only): every list entry's FP32 terms are added to the
// FP64 accumulators as they come, instead of being summed in FP32 over 32
// entries first.
template<class Eval, bool MIXED, bool VIRIAL, bool SINGLE, bool XPLOR, bool CUBIC, int LAYOUT, bool ACC64>
__global__ void __launch_bounds__(512) k_pair(const PairKernelArgs a)
    {
    constexpr bool TILED = LAYOUT == kLayoutTiled;
    constexpr bool HEAD = LAYOUT == kLayoutHead;
    using R = std::conditional_t<MIXED, float, double>;
    using Prep = typename Eval::template Prep<R>;
    constexpr uint32_t U = Unroll<MIXED>::value;
    extern __shared__ double smem[];
    const uint32_t nt = a.ntypes;
    const uint32_t nt2 = nt * nt;
    // tiled: the neighbor positions sit at LDS offset 0 (a list entry,
    // 16 x rank, is their byte address in the mixed form's 16-byte slots, and
    // 3/2 or 2 x it in the FP64 form's 24- or 32-byte slots) and the
    // type-pair tables after them
    constexpr size_t kSlotBytes = MIXED ? sizeof(int4) : fp64_slot_bytes(SINGLE);
    constexpr size_t kExtent = (TILED && !MIXED) ? kExtentBytes : 0;
    [[maybe_unused]] float* s_ext = reinterpret_cast<float*>(reinterpret_cast<char*>(smem) + (TILED ? a.tile_slots * kSlotBytes : 0));
    Prep* s_prep = reinterpret_cast<Prep*>(reinterpret_cast<char*>(smem) + (TILED ? a.tile_slots * kSlotBytes : 0)
                                           + kExtent);
    TypePairExtra<R>* s_x = reinterpret_cast<TypePairExtra<R>*>(s_prep + nt2);

    Prep q0;
    TypePairExtra<R> x0;
    if (SINGLE)
        {
        prepare_pair<Eval, R>(a, 0, q0, x0);
        }
    else
        {
        for (uint32_t t = threadIdx.x; t < nt2; t += blockDim.x)
            prepare_pair<Eval, R>(a, t, s_prep[t], s_x[t]);
        }

    const uint32_t lb = xcd_logical_block();
    int4* s_pos = nullptr;
    const uint32_t* tile_js = nullptr;
    uint32_t tile_m = 0;
    if constexpr (TILED && MIXED)
        {
        // stage the tile's neighbor positions, converted to fixed point here
        // (block lb is tile lb); every lane of the workgroup takes part before
        // any wavefront may leave.  Slot M is the padding entries' dummy.
        s_pos = reinterpret_cast<int4*>(smem);
        tile_js = a.tile_jset + (size_t)lb * kTileCap;
        tile_m = a.tile_count[lb];
        stage_slots<8>(
            tile_m, tile_js, [&](uint32_t j) { return a.pos[j]; },
            [&](uint32_t m, const double4& p) { s_pos[m] = fixed4(p, a.box); });
        if (threadIdx.x == 0)
            s_pos[tile_m] = make_int4(0, 0, 0, 0);
        }
    if constexpr (TILED && !MIXED)
        {
        // FP64 form: the FP64 positions themselves (x, y, z in 24-byte slots
        // with one type, x, y, z, type in 32-byte slots otherwise)
        tile_js = a.tile_jset + (size_t)lb * kTileCap;
        tile_m = a.tile_count[lb];
        // extent of the staged positions and of the tile's own particles
        TileExtent ext;
        auto track = [&](const double4& p) { ext.add(p.x, p.y, p.z); };
        if constexpr (SINGLE)
            {
            double* s_pos3 = reinterpret_cast<double*>(smem);
            stage_slots<8>(
                tile_m, tile_js, [&](uint32_t j) { return a.pos[j]; },
                [&](uint32_t m, const double4& p) {
                    s_pos3[3 * m] = p.x;
                    s_pos3[3 * m + 1] = p.y;
                    s_pos3[3 * m + 2] = p.z;
                    track(p);
                });
            if (threadIdx.x == 0)
                s_pos3[3 * tile_m] = s_pos3[3 * tile_m + 1] = s_pos3[3 * tile_m + 2] = 0.0;
            }
        else
            {
            double4* s_pos64 = reinterpret_cast<double4*>(smem);
            stage_slots<8>(
                tile_m, tile_js, [&](uint32_t j) { return a.pos[j]; },
                [&](uint32_t m, const double4& p) {
                    s_pos64[m] = p;
                    track(p);
                });
            if (threadIdx.x == 0)
                s_pos64[tile_m] = make_double4(0.0, 0.0, 0.0, 0.0);
            }
        const uint32_t i0 = lb * blockDim.x + threadIdx.x;
        if (i0 < a.N)
            track(a.pos[a.tile_order != nullptr ? a.tile_order[i0] : i0]);
        ext.store_wave(s_ext);
        }
    if (!SINGLE || TILED)
        __syncthreads();
    // FP64 over the tiled list, bit a set: some pair of the tile may need
    // the periodic image on axis a (the tile's positions span 0.45 L or more
    // on it); on the other axes the minimum image is the difference itself
    [[maybe_unused]] uint32_t wrap = 7u;
    if constexpr (TILED && !MIXED)
        wrap = wrap_axes(s_ext, a.box.L);

    const uint32_t i = lb * blockDim.x + threadIdx.x;
    // wavefronts entirely past N leave together; lanes past N in the last
    // wavefront shadow particle N-1 (same column, results not stored)
    if (lb * blockDim.x + (threadIdx.x & ~63u) >= a.N)
        return;
    const uint32_t ie = min(i, a.N - 1);
    // slot-ordered tiles: column ie belongs to particle pe
    const bool by_slot = TILED && a.tile_order != nullptr;
    const uint32_t pe = by_slot ? a.tile_order[ie] : ie;
    const uint32_t nn = by_slot ? a.tile_nn[ie] : a.n_neigh[ie];
    // wave-uniform trip count, held in a scalar register
    const uint32_t P = __builtin_amdgcn_readfirstlane((wave_max(nn) + U - 1) & ~(U - 1));
    const size_t stride = a.nlist_pitch;

    Acc acc;
    if constexpr (!MIXED && TILED)
        {
        // FP64 over the tiled list: the FP64 positions from LDS; the same
        // terms in the same order as over the plain list, so the results are
        // bit-identical to the plain form over the same entries
        const double4 pi = a.pos[pe];
        const uint32_t ti = (uint32_t)pi.w;
        const uint32_t col_off = ie * 4u;
        const uint32_t tile_m16 = tile_m << 4; // the padding entry
        const char* s_base = reinterpret_cast<const char*>(smem);
        // below F every lane of the wavefront has U real entries per step
        const uint32_t F = __builtin_amdgcn_readfirstlane(wave_min(nn) & ~(U - 1));
        auto step = [&](uint32_t k, auto padded) {
            constexpr bool PADDED = decltype(padded)::value;
            uint32_t w[U / 2];
#pragma unroll
            for (uint32_t u = 0; u < U / 2; ++u)
                {
                const char* row = reinterpret_cast<const char*>(a.tile_list + (size_t)((k >> 1) + u) * stride);
                w[u] = *reinterpret_cast<const uint32_t*>(row + col_off);
                }
            uint32_t e[U];
#pragma unroll
            for (uint32_t u = 0; u < U / 2; ++u)
                {
                e[2 * u] = w[u] & 0xffffu;
                e[2 * u + 1] = w[u] >> 16;
                }
            double4 pj[U];
#pragma unroll
            for (uint32_t u = 0; u < U; ++u)
                {
                if constexpr (SINGLE)
                    {
                    const double* p3 = reinterpret_cast<const double*>(s_base + e[u] + (e[u] >> 1));
                    pj[u] = make_double4(p3[0], p3[1], p3[2], 0.0);
                    }
                else
                    {
                    pj[u] = *reinterpret_cast<const double4*>(s_base + 2u * e[u]);
                    }
                }
            double dx[U], dy[U], dz[U];
#pragma unroll
            for (uint32_t u = 0; u < U; ++u)
                {
                dx[u] = pi.x - pj[u].x;
                dy[u] = pi.y - pj[u].y;
                dz[u] = pi.z - pj[u].z;
                }
            // the images, on the axes the tile needs them (a scalar branch)
            if (wrap & 1u)
                {
#pragma unroll
                for (uint32_t u = 0; u < U; ++u)
                    dx[u] = min_image_fused(dx[u], a.box.L[0], a.box.Linv[0]);
                }
            if (wrap & 2u)
                {
#pragma unroll
                for (uint32_t u = 0; u < U; ++u)
                    dy[u] = min_image_fused(dy[u], a.box.L[1], a.box.Linv[1]);
                }
            if (wrap & 4u)
                {
#pragma unroll
                for (uint32_t u = 0; u < U; ++u)
                    dz[u] = min_image_fused(dz[u], a.box.L[2], a.box.Linv[2]);
                }
#pragma unroll
            for (uint32_t u = 0; u < U; ++u)
                {
                const uint32_t tp = SINGLE ? 0u : ti * nt + (uint32_t)pj[u].w;
                term_fp64<Eval, VIRIAL, XPLOR>(SINGLE ? q0 : s_prep[tp], SINGLE ? x0 : s_x[tp],
                                               !PADDED || e[u] != tile_m16, dx[u], dy[u], dz[u], acc);
                }
        };
        uint32_t k = 0;
        for (; k < F; k += U)
            step(k, std::false_type{});
        for (; k < P; k += U)
            step(k, std::true_type{});
        if (i < a.N)
            store<VIRIAL>(a, pe, acc);
        }
    else if constexpr (!MIXED)
        {
        const double4 pi = a.pos[ie];
        const uint32_t ti = (uint32_t)pi.w;
        const uint32_t* hrow = HEAD ? a.nlist + a.head_list[ie] : nullptr;
        // head list: below F every lane has U entries left in its row, read as
        // 16-byte vectors (one lane's row is contiguous; lanes' rows are not)
        const uint32_t F = HEAD ? __builtin_amdgcn_readfirstlane(wave_min(nn) & ~(U - 1)) : 0u;
        auto step = [&](uint32_t k, auto whole) {
            uint32_t jj[U];
            if constexpr (HEAD && decltype(whole)::value)
                head_words<U>(hrow + k, jj);
            else
                {
#pragma unroll
                for (uint32_t u = 0; u < U; ++u)
                    {
                    if constexpr (HEAD)
                        jj[u] = k + u < nn ? hrow[k + u] : ie; // rows are not padded
                    else
                        jj[u] = (a.nlist + (size_t)(k + u) * stride)[ie];
                    }
                }
            double4 pj[U];
#pragma unroll
            for (uint32_t u = 0; u < U; ++u)
                pj[u] = a.pos[jj[u]];
#pragma unroll
            for (uint32_t u = 0; u < U; ++u)
                {
                const double dx = min_image_fused(pi.x - pj[u].x, a.box.L[0], a.box.Linv[0]);
                const double dy = min_image_fused(pi.y - pj[u].y, a.box.L[1], a.box.Linv[1]);
                const double dz = min_image_fused(pi.z - pj[u].z, a.box.L[2], a.box.Linv[2]);
                const uint32_t tp = SINGLE ? 0u : ti * nt + (uint32_t)pj[u].w;
                term_fp64<Eval, VIRIAL, XPLOR>(SINGLE ? q0 : s_prep[tp], SINGLE ? x0 : s_x[tp],
                                               jj[u] != ie, dx, dy, dz, acc);
                }
        };
        uint32_t k = 0;
        for (; k < F; k += U)
            step(k, std::true_type{});
        for (; k < P; k += U)
            step(k, std::false_type{});
        if (i < a.N)
            store<VIRIAL>(a, i, acc);
        }
    else if constexpr (TILED)
        {
        const int4 pi = fixed4(a.pos[pe], a.box);
        const uint32_t ti = (uint32_t)pi.w;
        const uint32_t col_off = ie * 4u;
        const uint32_t tile_m16 = tile_m << 4; // the padding entry
        // below F every lane of the wavefront has U real entries per step:
        // those steps skip the padding test
        const uint32_t F = __builtin_amdgcn_readfirstlane(wave_min(nn) & ~(U - 1));
        Acc32 part;
        auto step = [&](uint32_t k, auto padded) {
            constexpr bool PADDED = decltype(padded)::value;
            uint32_t w[U / 2];
#pragma unroll
            for (uint32_t u = 0; u < U / 2; ++u)
                {
                const char* row = reinterpret_cast<const char*>(a.tile_list + (size_t)((k >> 1) + u) * stride);
                w[u] = *reinterpret_cast<const uint32_t*>(row + col_off);
                }
            uint32_t e[U];
#pragma unroll
            for (uint32_t u = 0; u < U / 2; ++u)
                {
                e[2 * u] = w[u] & 0xffffu;
                e[2 * u + 1] = w[u] >> 16;
                }
            int4 pj[U];
#pragma unroll
            for (uint32_t u = 0; u < U; ++u)
                pj[u] = *reinterpret_cast<const int4*>(reinterpret_cast<const char*>(s_pos) + e[u]);
#pragma unroll
            for (uint32_t u = 0; u < U; ++u)
                {
                const uint32_t tp = SINGLE ? 0u : ti * nt + (uint32_t)pj[u].w;
                term_mixed<Eval, VIRIAL, XPLOR, CUBIC>(a, SINGLE ? q0 : s_prep[tp], SINGLE ? x0 : s_x[tp],
                                                       !PADDED || e[u] != tile_m16, pe, e[u] >> 4, tile_js,
                                                       fixed_delta(pi.x, pj[u].x), fixed_delta(pi.y, pj[u].y),
                                                       fixed_delta(pi.z, pj[u].z), part);
                if (ACC64)
                    {
                    flush32<VIRIAL>(acc, part);
                    part = Acc32();
                    }
                }
            if (!ACC64 && (((k + U) & (kFlushEntries - 1)) == 0 || k + U >= P))
                {
                flush32<VIRIAL>(acc, part);
                part = Acc32();
                }
        };
        uint32_t k = 0;
        for (; k < F; k += U)
            step(k, std::false_type{});
        for (; k < P; k += U)
            step(k, std::true_type{});
        if (i < a.N)
            store_scaled<VIRIAL>(a, pe, acc);
        }
    else
        {
        const int4 pi = a.pos_fixed[ie];
        const uint32_t ti = (uint32_t)pi.w;
        // 32-bit byte offsets from wave-uniform bases (global loads with an
        // SGPR base and a VGPR offset: no per-entry 64-bit address math)
        const uint32_t col_off = ie * 4u;
        const char* pf_base = reinterpret_cast<const char*>(a.pos_fixed);
        const uint32_t* hrow = HEAD ? a.nlist + a.head_list[ie] : nullptr;
        // head list: below F every lane has U entries left in its row, read as
        // 16-byte vectors
        const uint32_t F = HEAD ? __builtin_amdgcn_readfirstlane(wave_min(nn) & ~(U - 1)) : 0u;
        Acc32 part;
        auto step = [&](uint32_t k, auto whole) {
            uint32_t jj[U];
            if constexpr (HEAD && decltype(whole)::value)
                head_words<U>(hrow + k, jj);
            else
                {
#pragma unroll
                for (uint32_t u = 0; u < U; ++u)
                    {
                    if constexpr (HEAD)
                        {
                        jj[u] = k + u < nn ? hrow[k + u] : ie; // rows are not padded
                        }
                    else
                        {
                        const char* row = reinterpret_cast<const char*>(a.nlist + (size_t)(k + u) * stride);
                        jj[u] = *reinterpret_cast<const uint32_t*>(row + col_off);
                        }
                    }
                }
            int4 pj[U];
#pragma unroll
            for (uint32_t u = 0; u < U; ++u)
                pj[u] = *reinterpret_cast<const int4*>(pf_base + (jj[u] << 4));
#pragma unroll
            for (uint32_t u = 0; u < U; ++u)
                {
                const uint32_t tp = SINGLE ? 0u : ti * nt + (uint32_t)pj[u].w;
                term_mixed<Eval, VIRIAL, XPLOR, CUBIC>(a, SINGLE ? q0 : s_prep[tp], SINGLE ? x0 : s_x[tp],
                                                       jj[u] != ie, ie, jj[u], nullptr, fixed_delta(pi.x, pj[u].x),
                                                       fixed_delta(pi.y, pj[u].y), fixed_delta(pi.z, pj[u].z),
                                                       part);
                if (ACC64)
                    {
                    flush32<VIRIAL>(acc, part);
                    part = Acc32();
                    }
                }
            if (!ACC64 && (((k + U) & (kFlushEntries - 1)) == 0 || k + U >= P))
                {
                flush32<VIRIAL>(acc, part);
                part = Acc32();
                }
        };
        uint32_t k = 0;
        for (; k < F; k += U)
            step(k, std::true_type{});
        for (; k < P; k += U)
            step(k, std::false_type{});
        if (i < a.N)
            store_scaled<VIRIAL>(a, i, acc);
        }
    }


template<class Eval, bool MIXED, bool VIRIAL, bool SINGLE, bool XPLOR, bool CUBIC, int LAYOUT>
static int launch_t(const PairKernelArgs& ka, unsigned block, unsigned slots, hipStream_t stream)
    {
    constexpr bool TILED = LAYOUT == kLayoutTiled;
    using R = std::conditional_t<MIXED, float, double>;
    const size_t nt2 = (size_t)ka.ntypes * ka.ntypes;
    size_t shmem = SINGLE ? 0 : table_bytes<Eval, R>((uint32_t)nt2);
    PairKernelArgs kb = ka;
    kb.tile_slots = TILED ? slots : 0u;
    if (TILED)
        shmem += (size_t)slots * (MIXED ? sizeof(int4) : fp64_slot_bytes(SINGLE)) + (MIXED ? 0 : kExtentBytes);
    const size_t cap = (TILED && !MIXED) ? kTileLds64 : kTileLds;
    if (shmem > cap)
        return set_error(AZP_ERR_UNSUPPORTED, "too many particle types (%u) for LDS staging", ka.ntypes);
    if (shmem > 64 * 1024)
        {
        // more than 64 KiB of dynamic LDS (gfx950 gives a workgroup 160 KiB)
        static bool raised = false;
        if (!raised)
            {
            // only the FP64 tiled form (ACC64 false) stages more than 64 KiB
            const hipError_t e = hipFuncSetAttribute(
                reinterpret_cast<const void*>(&k_pair<Eval, MIXED, VIRIAL, SINGLE, XPLOR, CUBIC, LAYOUT, false>),
                hipFuncAttributeMaxDynamicSharedMemorySize, (int)cap);
            if (e != hipSuccess)
                return set_error(AZP_ERR_HIP, "hipFuncSetAttribute (dynamic LDS %zu bytes) failed: %s", cap,
                                 hipGetErrorString(e));
            raised = true;
            }
        }
    const unsigned grid = grid_for(ka.N, block);
    if (MIXED && ka.acc64)
        hipLaunchKernelGGL((k_pair<Eval, MIXED, VIRIAL, SINGLE, XPLOR, CUBIC, LAYOUT, MIXED>), dim3(grid), dim3(block),
                           shmem, stream, kb);
    else
        hipLaunchKernelGGL((k_pair<Eval, MIXED, VIRIAL, SINGLE, XPLOR, CUBIC, LAYOUT, false>), dim3(grid), dim3(block),
                           shmem, stream, kb);
    AZP_LAUNCH_CHECK();
    return AZP_OK;
    }

template<class Eval, bool MIXED, bool VIRIAL, bool SINGLE, int LAYOUT>
static int launch_mode(const PairKernelArgs& ka, unsigned block, unsigned slots, hipStream_t s)
    {
    const bool xplor = ka.shift_mode == AZP_SHIFT_XPLOR;
    if constexpr (MIXED)
        {
        const bool cubic = ka.fx_scale2[0] == ka.fx_scale2[1] && ka.fx_scale2[0] == ka.fx_scale2[2];
        if (xplor)
            return cubic ? launch_t<Eval, true, VIRIAL, SINGLE, true, true, LAYOUT>(ka, block, slots, s)
                         : launch_t<Eval, true, VIRIAL, SINGLE, true, false, LAYOUT>(ka, block, slots, s);
        return cubic ? launch_t<Eval, true, VIRIAL, SINGLE, false, true, LAYOUT>(ka, block, slots, s)
                     : launch_t<Eval, true, VIRIAL, SINGLE, false, false, LAYOUT>(ka, block, slots, s);
        }
    else
        {
        return xplor ? launch_t<Eval, false, VIRIAL, SINGLE, true, false, LAYOUT>(ka, block, slots, s)
                     : launch_t<Eval, false, VIRIAL, SINGLE, false, false, LAYOUT>(ka, block, slots, s);
        }
    }

template<class Eval, bool MIXED, int LAYOUT>
static int launch_virial(const PairKernelArgs& ka, bool virial, unsigned block, unsigned slots, hipStream_t s)
    {
    const bool single = ka.ntypes == 1;
    if (virial)
        return single ? launch_mode<Eval, MIXED, true, true, LAYOUT>(ka, block, slots, s)
                      : launch_mode<Eval, MIXED, true, false, LAYOUT>(ka, block, slots, s);
    return single ? launch_mode<Eval, MIXED, false, true, LAYOUT>(ka, block, slots, s)
                  : launch_mode<Eval, MIXED, false, false, LAYOUT>(ka, block, slots, s);
    }

// One layout's kernels for the evaluator code: the per-layout translation
// units (k_pair.hip column, k_pair_tiled.hip tiled, k_pair_head.hip head list)
// each instantiate this for their layout.
template<int LAYOUT>
static int launch_pair_evaluator(const PairKernelArgs& ka,
                                 int evaluator,
                                 bool mixed,
                                 bool virial,
                                 unsigned block,
                                 unsigned slots,
                                 hipStream_t s)
    {
    if (mixed && LAYOUT != kLayoutTiled && !ka.pos_fixed)
        return set_error(AZP_ERR_INVALID, "the mixed form over the plain list needs fixed-point positions");
    switch (evaluator)
        {
    case AZP_EVAL_HERTZ:
        return mixed ? launch_virial<EvalHertz, true, LAYOUT>(ka, virial, block, slots, s)
                     : launch_virial<EvalHertz, false, LAYOUT>(ka, virial, block, slots, s);
    case AZP_EVAL_EXPANDED_YUKAWA:
        return mixed ? launch_virial<EvalExpandedYukawa, true, LAYOUT>(ka, virial, block, slots, s)
                     : launch_virial<EvalExpandedYukawa, false, LAYOUT>(ka, virial, block, slots, s);
    case AZP_EVAL_PERTURBED_LJ:
        return mixed ? launch_virial<EvalPerturbedLJ, true, LAYOUT>(ka, virial, block, slots, s)
                     : launch_virial<EvalPerturbedLJ, false, LAYOUT>(ka, virial, block, slots, s);
    case AZP_EVAL_COLLOID:
        if (mixed)
            return set_error(AZP_ERR_UNSUPPORTED,
                             "Colloid runs in the plain FP64 form only (its colloid-colloid "
                             "branch cancels catastrophically in FP32)");
        return launch_virial<EvalColloid, false, LAYOUT>(ka, virial, block, slots, s);
    case AZP_EVAL_DPD_GENERAL_WEIGHT:
        // PotentialPairConservative<DPDPairEvaluatorGeneralWeight>
        return mixed ? launch_virial<EvalDPDConservative, true, LAYOUT>(ka, virial, block, slots, s)
                     : launch_virial<EvalDPDConservative, false, LAYOUT>(ka, virial, block, slots, s);
    default:
        return set_error(AZP_ERR_INVALID, "evaluator %d is not an isotropic pair evaluator", evaluator);
        }
    }

// Defined in k_pair.hip, k_pair_tiled.hip and k_pair_head.hip.  The tiled
// form runs workgroups of `block` = the tile size over `slots` LDS slots.
int launch_pair_column(const PairKernelArgs& ka, int evaluator, bool mixed, bool virial, unsigned block,
                       hipStream_t s);
int launch_pair_tiled(const PairKernelArgs& ka, int evaluator, bool mixed, bool virial, unsigned tile,
                      unsigned slots, hipStream_t s);
int launch_pair_head(const PairKernelArgs& ka, int evaluator, bool mixed, bool virial, unsigned block,
                     hipStream_t s);
// Whether the tiled kernel of the evaluator fits its positions and tables in LDS.
bool pair_tile_fits(int evaluator, uint32_t ntypes, uint32_t tile_slots, bool mixed);

} // namespace azp

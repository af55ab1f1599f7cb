// Batched replicas: many independent periodic systems of uneven size held in
// one set of concatenated arrays and advanced together.
//
// This is the batching form of the anisotropic pair path: the same
// TwoPatchMorse pair term as AnisoPotentialPair<AnisoPairEvaluatorTwoPatchMorse>
// (AnisoPotentialPairGPUKernel.cu.inc:21-25, evaluator
// AnisoPairEvaluatorTwoPatchMorse.h:127-216), applied to R replicas at once.
//
// * One cell list over all replicas: replica r owns a contiguous range of
//   global cells (its own grid, cells at least r_list wide) and a contiguous
//   range of particles; the walk lists each particle's neighbors inside its
//   own replica with the same FP64 predicate as the single-system build.
// * One persistent force launch: a grid sized to the chip pulls 64-particle
//   chunks from a work counter, so replicas of 500 and of 20,000 particles
//   share the launch without tail effects.  Block size, chunks per fetch and
//   workgroups per CU come from an autotuner that times the candidates on the
//   live batch.
// * Time steps (Langevin, translation and rotation of spherical tops) are
//   captured once into a HIP graph and replayed.  Everything that depends on
//   data is decided on the device: the rebuild check flags each replica that
//   needs a new list (and a global "any" flag); the rebuild kernels return at
//   once when no replica is flagged and otherwise touch only the flagged
//   replicas' particles and cells (the others keep their lists, counts and
//   cell ranges).  Counters and flags are cleared by fill kernels (no memset
//   nodes in the graph).
#include <algorithm>
#include <cmath>
#include <vector>

#include "cells.h"
#include "tpm.h"

namespace azp
{

struct RepInfo
    {
    double L[3];
    double Linv[3];
    float fx_scale[3];
    int n[3];           // cell grid
    uint32_t cell_base; // first global cell
    uint32_t first;     // first particle
    uint32_t count;     // particles
    double kT;          // Langevin temperature
    };

// Gate of the rebuild kernels: they run when gate is null or its bit 0 is set.
__device__ __forceinline__ bool gated_off(const uint32_t* gate)
    {
    return gate && (gate[0] & 1u) == 0u;
    }

__global__ void k_b_fill(uint32_t* p, uint32_t v, uint32_t n, const uint32_t* gate)
    {
    if (gated_off(gate))
        return;
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n)
        p[k] = v;
    }

__global__ void k_b_fill64(unsigned long long* p, unsigned long long v)
    {
    if (blockIdx.x == 0 && threadIdx.x == 0)
        p[0] = v;
    }

// Clears the cell counts of the flagged replicas.
__global__ void k_b_clear_counts(uint32_t* __restrict__ count,
                                 uint32_t ncell,
                                 const uint32_t* __restrict__ cell_rep,
                                 const uint32_t* __restrict__ rep_flag,
                                 const uint32_t* gate)
    {
    if (gated_off(gate))
        return;
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < ncell && rep_flag[cell_rep[c]])
        count[c] = 0u;
    }

__global__ void k_b_bin(const double4* __restrict__ pos,
                        uint32_t N,
                        const uint32_t* __restrict__ rep_of,
                        const RepInfo* __restrict__ reps,
                        const uint32_t* __restrict__ rep_flag,
                        uint32_t* __restrict__ cell_of,
                        uint32_t* __restrict__ slot_of,
                        uint32_t* __restrict__ count,
                        const uint32_t* gate)
    {
    if (gated_off(gate))
        return;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N)
        return;
    const uint32_t r = rep_of[i];
    if (!rep_flag[r])
        return;
    const double4 p = pos[i];
    const RepInfo& R = reps[r];
    const int cx = cell_coord(p.x, R.L[0], R.Linv[0], R.n[0]);
    const int cy = cell_coord(p.y, R.L[1], R.Linv[1], R.n[1]);
    const int cz = cell_coord(p.z, R.L[2], R.Linv[2], R.n[2]);
    const uint32_t c = R.cell_base + (uint32_t)((cz * R.n[1] + cy) * R.n[0] + cx);
    cell_of[i] = c;
    slot_of[i] = atomicAdd(&count[c], 1u);
    }

__global__ void k_b_scatter(uint32_t N,
                            const uint32_t* __restrict__ rep_of,
                            const uint32_t* __restrict__ rep_flag,
                            const uint32_t* __restrict__ cell_of,
                            const uint32_t* __restrict__ slot_of,
                            const uint32_t* __restrict__ start,
                            uint32_t* __restrict__ members,
                            const uint32_t* gate)
    {
    if (gated_off(gate))
        return;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N || !rep_flag[rep_of[i]])
        return;
    members[start[cell_of[i]] + slot_of[i]] = i;
    }

// Each cell's members in increasing particle index (a deterministic order).
// One wavefront per cell: a member's place is the number of the cell's
// members with a smaller index (indices are distinct), so after copying the
// cell to LDS every lane writes its members straight to their places.  A
// cell with more members than the copy holds is sorted by the wavefront's
// first lane (insertion sort).  Every exit is uniform over the wavefront and
// lanes exchange data only through their wavefront's LDS row.
constexpr uint32_t kSortWaves = 4;
constexpr uint32_t kSortCap = 256;

__global__ void __launch_bounds__(256) k_b_sort_cells(uint32_t ncell,
                                                      const uint32_t* __restrict__ cell_rep,
                                                      const uint32_t* __restrict__ rep_flag,
                                                      const uint32_t* __restrict__ start,
                                                      uint32_t* __restrict__ members,
                                                      const uint32_t* gate)
    {
    __shared__ uint32_t s_m[kSortWaves][kSortCap];
    if (gated_off(gate))
        return;
    const uint32_t w = threadIdx.x >> 6;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t c = blockIdx.x * kSortWaves + w;
    if (c >= ncell || !rep_flag[cell_rep[c]])
        return;
    const uint32_t b = start[c];
    const uint32_t n = start[c + 1] - b;
    if (n < 2)
        return;
    if (n > kSortCap)
        {
        if (lane == 0)
            for (uint32_t k = b + 1; k < b + n; ++k)
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
        return;
        }
    uint32_t* sm = s_m[w];
    for (uint32_t k = lane; k < n; k += 64)
        sm[k] = members[b + k];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    for (uint32_t k = lane; k < n; k += 64)
        {
        const uint32_t v = sm[k];
        uint32_t rank = 0;
        for (uint32_t t = 0; t < n; ++t)
            rank += sm[t] < v ? 1u : 0u;
        members[b + rank] = v;
        }
    }

struct BatchWalkArgs
    {
    const double4* pos;
    uint32_t N;
    const uint32_t* rep_of;
    const RepInfo* reps;
    const uint32_t* cell_of;
    const uint32_t* start;
    const uint32_t* members;
    const double* r_listsq;
    uint32_t ntypes;
    uint32_t* n_neigh;
    uint32_t* nlist;
    uint32_t pitch;
    uint32_t max_neighbors;
    uint32_t* flags; // [0] largest count beyond capacity, [2] bad type (1) / non-finite (2)
    const uint32_t* gate;
    const uint32_t* rep_flag;
    };

// Particle i's neighbors inside its own replica (same cell order, predicate
// and rounding as the single-system walk).
__device__ uint32_t b_walk_one(const BatchWalkArgs& a, uint32_t i, const double* __restrict__ r_listsq)
    {
    const double4 pi = a.pos[i];
    const uint32_t ti = checked_type(pi.w, a.ntypes);
    const bool finite = isfinite(pi.x) && isfinite(pi.y) && isfinite(pi.z);
    if (ti == a.ntypes || !finite)
        {
        atomicOr(a.flags + 2, finite ? 1u : 2u);
        return 0;
        }
    const RepInfo& R = a.reps[a.rep_of[i]];
    // the box in registers: the list stores below may not be assumed apart
    // from the replica table, which would otherwise be reloaded per candidate
    const double L0 = R.L[0], L1 = R.L[1], L2 = R.L[2];
    const double Li0 = R.Linv[0], Li1 = R.Linv[1], Li2 = R.Linv[2];
    const uint32_t lc = a.cell_of[i] - R.cell_base;
    const int cx = (int)(lc % (uint32_t)R.n[0]);
    const int cy = (int)((lc / (uint32_t)R.n[0]) % (uint32_t)R.n[1]);
    const int cz = (int)(lc / ((uint32_t)R.n[0] * (uint32_t)R.n[1]));
    int ox[3], oy[3], oz[3];
    const int nx = stencil_axis(cx, R.n[0], ox);
    const int ny = stencil_axis(cy, R.n[1], oy);
    const int nz = stencil_axis(cz, R.n[2], oz);
    uint32_t count = 0;
    auto consider = [&](uint32_t j, const double4& pj) {
        if (j == i)
            return;
        const uint32_t tj = checked_type(pj.w, a.ntypes);
        if (tj == a.ntypes)
            return;
        const double rl = r_listsq[ti * a.ntypes + tj];
        if (!(rl > 0.0))
            return;
        const double dx = min_image_1d(pi.x - pj.x, L0, Li0);
        const double dy = min_image_1d(pi.y - pj.y, L1, Li1);
        const double dz = min_image_1d(pi.z - pj.z, L2, Li2);
        const double rsq = rsq3_nofma(dx, dy, dz);
        if (rsq <= rl)
            {
            if (count < a.max_neighbors)
                a.nlist[(size_t)count * a.pitch + i] = j;
            ++count;
            }
    };
    for (int iz = 0; iz < nz; ++iz)
        for (int iy = 0; iy < ny; ++iy)
            for (int ix = 0; ix < nx; ++ix)
                {
                const uint32_t nb = R.cell_base + (uint32_t)((oz[iz] * R.n[1] + oy[iy]) * R.n[0] + ox[ix]);
                const uint32_t b = a.start[nb], e = a.start[nb + 1];
                // four candidates at a time: their member and position
                // loads are in flight together; they are tested in order
                uint32_t m = b;
                for (; m + 4 <= e; m += 4)
                    {
                    uint32_t jj[4];
                    double4 pp[4];
#pragma unroll
                    for (int u = 0; u < 4; ++u)
                        jj[u] = a.members[m + u];
#pragma unroll
                    for (int u = 0; u < 4; ++u)
                        pp[u] = a.pos[jj[u]];
#pragma unroll
                    for (int u = 0; u < 4; ++u)
                        consider(jj[u], pp[u]);
                    }
                for (; m < e; ++m)
                    {
                    const uint32_t j = a.members[m];
                    consider(j, a.pos[j]);
                    }
                }
    return count;
    }

// One lane per particle; each wavefront pads its columns with the lane's own
// index up to its longest list rounded up to 8 (as the single-system walk).
__global__ void __launch_bounds__(256) k_b_walk(const BatchWalkArgs a)
    {
    // the list radii (up to 8 types) in LDS
    __shared__ double s_rl[64];
    if (gated_off(a.gate))
        return;
    const uint32_t nt2 = a.ntypes * a.ntypes;
    const bool staged = nt2 <= 64u;
    if (staged)
        {
        for (uint32_t t = threadIdx.x; t < nt2; t += blockDim.x)
            s_rl[t] = a.r_listsq[t];
        __syncthreads();
        }
    const uint32_t i = xcd_logical_block() * blockDim.x + threadIdx.x;
    const bool active = i < a.N && a.rep_flag[a.rep_of[i]] != 0u;
    const uint32_t count = active ? b_walk_one(a, i, staged ? s_rl : a.r_listsq) : 0u;
    if (active)
        {
        // a list longer than the capacity is truncated (and reported), so
        // consumers never read past the allocation
        a.n_neigh[i] = min(count, a.max_neighbors);
        if (count > a.max_neighbors)
            atomicMax(a.flags, count);
        }
    uint32_t wmax = count;
    for (int off = 32; off > 0; off >>= 1)
        wmax = max(wmax, (uint32_t)__shfl_xor((int)wmax, off));
    const uint32_t padded = (wmax + 7u) & ~7u;
    if (active && wmax <= a.max_neighbors)
        for (uint32_t k = count; k < padded; ++k)
            a.nlist[(size_t)k * a.pitch + i] = i;
    }

// Snapshot of the positions of the last build of each rebuilt replica, and
// the counters of rebuild events and of replica rebuilds.
__global__ void k_b_copy_last(const double4* __restrict__ pos,
                              double4* __restrict__ last,
                              uint32_t N,
                              const uint32_t* __restrict__ rep_of,
                              const RepInfo* __restrict__ reps,
                              const uint32_t* __restrict__ rep_flag,
                              unsigned long long* __restrict__ builds,
                              const uint32_t* gate)
    {
    if (gated_off(gate))
        return;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0)
        atomicAdd(builds, 1ull);
    if (i >= N)
        return;
    const uint32_t r = rep_of[i];
    if (!rep_flag[r])
        return;
    last[i] = pos[i];
    if (i == reps[r].first)
        atomicAdd(builds + 1, 1ull);
    }

// Rebuild flags: replica r is flagged (and the global flags[1]) when one of
// its particles moved farther than r_buff / 2 since its last build, or has a
// position that is not finite (also reported in flags[2], bit 1).
__global__ void k_b_check(const double4* __restrict__ pos,
                          const double4* __restrict__ last,
                          uint32_t N,
                          const uint32_t* __restrict__ rep_of,
                          const RepInfo* __restrict__ reps,
                          double thresh_sq,
                          uint32_t* __restrict__ rep_flag,
                          uint32_t* __restrict__ flags)
    {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N)
        return;
    const uint32_t r = rep_of[i];
    const double4 p = pos[i];
    if (!(isfinite(p.x) && isfinite(p.y) && isfinite(p.z)))
        {
        rep_flag[r] = 1u;
        flags[1] = 1u;
        atomicOr(flags + 2, 2u);
        return;
        }
    const double4 q = last[i];
    const RepInfo& R = reps[r];
    const double dx = min_image_1d(p.x - q.x, R.L[0], R.Linv[0]);
    const double dy = min_image_1d(p.y - q.y, R.L[1], R.Linv[1]);
    const double dz = min_image_1d(p.z - q.z, R.L[2], R.Linv[2]);
    if (rsq3_nofma(dx, dy, dz) > thresh_sq)
        {
        rep_flag[r] = 1u;
        flags[1] = 1u;
        }
    }

// Mixed form: positions to fixed point in each replica's box, and the
// FP32 director of every particle (read by its neighbors instead of the
// quaternion).
__global__ void k_b_pack_fixed(const double4* __restrict__ pos,
                               const double4* __restrict__ orientation,
                               uint32_t N,
                               const uint32_t* __restrict__ rep_of,
                               const RepInfo* __restrict__ reps,
                               int4* __restrict__ out,
                               float4* __restrict__ dirs)
    {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N)
        return;
    const double4 p = pos[i];
    const RepInfo& R = reps[rep_of[i]];
    out[i] = make_int4(to_fixed(p.x, R.Linv[0]), to_fixed(p.y, R.Linv[1]), to_fixed(p.z, R.Linv[2]), (int)p.w);
    float n[3];
    director_of<float>(orientation[i], n);
    dirs[i] = make_float4(n[0], n[1], n[2], 0.0f);
    }

struct BatchForceArgs
    {
    const double4* pos;
    const int4* pos_fixed;
    const double4* orientation;
    uint32_t N;
    const uint32_t* rep_of;
    const RepInfo* reps;
    const uint32_t* n_neigh;
    const uint32_t* nlist;
    uint32_t pitch;
    const azp_param_two_patch_morse* params;
    const double* rcutsq;
    uint32_t ntypes;
    int shift_mode;
    const float4* dirs; // mixed form: FP32 directors
    double4* force;
    double4* torque;
    double* virial;
    uint32_t vpitch;
    uint32_t* counter; // work counter, cleared before the launch
    uint32_t n_chunks; // ceil(N / 64)
    uint32_t grab;     // chunks taken per fetch
    };

// Persistent TwoPatchMorse forces over all replicas.  Each wavefront fetches
// `grab` 64-particle chunks at a time from the work counter until the counter
// passes the last chunk, so every wavefront reaches the exit.
template<bool MIXED, bool VIRIAL, bool SINGLE>
__global__ void __launch_bounds__(512) k_b_forces(const BatchForceArgs a)
    {
    using R = std::conditional_t<MIXED, float, double>;
    extern __shared__ double smem[];
    TPMPrep<R>* s_q = reinterpret_cast<TPMPrep<R>*>(smem);
    const uint32_t nt = a.ntypes;
    TPMPrep<R> q0;
    if (SINGLE)
        {
        q0 = tpm_prepare<R>(a.params[0], a.rcutsq[0], a.shift_mode == AZP_SHIFT_SHIFT);
        }
    else
        {
        for (uint32_t t = threadIdx.x; t < nt * nt; t += blockDim.x)
            s_q[t] = tpm_prepare<R>(a.params[t], a.rcutsq[t], a.shift_mode == AZP_SHIFT_SHIFT);
        __syncthreads();
        }
    const uint32_t lane = threadIdx.x & 63u;
    for (;;)
        {
        uint32_t c0 = 0;
        if (lane == 0)
            c0 = atomicAdd(a.counter, a.grab);
        c0 = (uint32_t)__shfl((int)c0, 0);
        if (c0 >= a.n_chunks)
            break;
        const uint32_t c1 = min(c0 + a.grab, a.n_chunks);
        for (uint32_t c = c0; c < c1; ++c)
            {
            const uint32_t i = c * 64u + lane;
            const bool active = i < a.N;
            const uint32_t ie = min(i, a.N - 1);
            const RepInfo& rep = a.reps[a.rep_of[ie]];
            LaneBox box;
#pragma unroll
            for (int k = 0; k < 3; ++k)
                {
                box.L[k] = rep.L[k];
                box.Linv[k] = rep.Linv[k];
                box.fx_scale[k] = rep.fx_scale[k];
                }
            const AnisoSelf<R> me = aniso_self<MIXED>(a.pos, a.pos_fixed, a.orientation, ie);
            AnisoAcc acc;
            const uint32_t nn = active ? a.n_neigh[i] : 0u;
            const uint32_t trips = __builtin_amdgcn_readfirstlane((tpm_wave_max(nn) + 3u) & ~3u);
            tpm_list<MIXED, VIRIAL, SINGLE, MIXED>(ie, nn, trips, a.nlist + ie, a.pitch, a.pos, a.pos_fixed, a.orientation,
                                                   a.dirs, me, box, nt, q0, s_q, acc);
            if (active)
                aniso_store<VIRIAL>(a.force, a.torque, a.virial, a.vpitch, i, acc);
            }
        }
    }

__device__ __forceinline__ void b_wrap(double& x, double L)
    {
    if (x >= 0.5 * L)
        x -= L;
    else if (x < -0.5 * L)
        x += L;
    }

// First half step: v += dt a / 2, x += dt v (wrapped into the replica's box);
// L += dt tau / 2, q <- rotation by omega dt (spherical top, omega = L / I).
__global__ void k_b_step_one(double4* __restrict__ pos,
                             double4* __restrict__ vel,
                             const double4* __restrict__ accel,
                             double4* __restrict__ orientation,
                             double4* __restrict__ angmom,
                             const double4* __restrict__ torque,
                             uint32_t N,
                             const uint32_t* __restrict__ rep_of,
                             const RepInfo* __restrict__ reps,
                             double dt,
                             double inertia)
    {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N)
        return;
    const RepInfo& R = reps[rep_of[i]];
    double4 p = pos[i];
    double4 v = vel[i];
    const double4 a = accel[i];
    v.x += 0.5 * dt * a.x;
    v.y += 0.5 * dt * a.y;
    v.z += 0.5 * dt * a.z;
    p.x += dt * v.x;
    p.y += dt * v.y;
    p.z += dt * v.z;
    b_wrap(p.x, R.L[0]);
    b_wrap(p.y, R.L[1]);
    b_wrap(p.z, R.L[2]);
    pos[i] = p;
    vel[i] = v;

    double4 L = angmom[i];
    const double4 t = torque[i];
    L.x += 0.5 * dt * t.x;
    L.y += 0.5 * dt * t.y;
    L.z += 0.5 * dt * t.z;
    angmom[i] = L;
    const double wx = L.x / inertia, wy = L.y / inertia, wz = L.z / inertia;
    const double w = sqrt(wx * wx + wy * wy + wz * wz);
    if (w == 0.0)
        return;
    const double half = 0.5 * w * dt;
    const double s = sin(half) / w;
    const double r0 = cos(half), r1 = s * wx, r2 = s * wy, r3 = s * wz;
    const double4 q = orientation[i]; // (s, x, y, z)
    double4 o;
    o.x = r0 * q.x - r1 * q.y - r2 * q.z - r3 * q.w;
    o.y = r0 * q.y + r1 * q.x + r2 * q.w - r3 * q.z;
    o.z = r0 * q.z - r1 * q.w + r2 * q.x + r3 * q.y;
    o.w = r0 * q.w + r1 * q.z - r2 * q.y + r3 * q.x;
    const double n = 1.0 / sqrt(o.x * o.x + o.y * o.y + o.z * o.z + o.w * o.w);
    o.x *= n;
    o.y *= n;
    o.z *= n;
    o.w *= n;
    orientation[i] = o;
    }

// Three uniform(-1, 1) variates for (particle, replica, stream) at timestep t.
__device__ __forceinline__ void b_noise(uint16_t seed, uint64_t t, uint32_t local, uint32_t rep, uint32_t stream, double u[3])
    {
    uint32_t k0, k1;
    hoomd_seed_key(kRngLangevin, t, seed, k0, k1);
    u32x4 c;
    c.v[0] = local;
    c.v[1] = rep;
    c.v[2] = stream;
    c.v[3] = 0u;
    const u32x4 r0 = philox4x32_10(c, k0, k1);
    c.v[3] = 1u;
    const u32x4 r1 = philox4x32_10(c, k0, k1);
    u32x4 x;
    x.v[0] = r0.v[0];
    x.v[1] = r0.v[1];
    u[0] = uniform_m1_1(x);
    x.v[0] = r0.v[2];
    x.v[1] = r0.v[3];
    u[1] = uniform_m1_1(x);
    x.v[0] = r1.v[0];
    x.v[1] = r1.v[1];
    u[2] = uniform_m1_1(x);
    }

// Second half step with the Langevin thermostat at the replica's kT:
// a = (F - gamma v + sqrt(6 gamma kT / dt) u) / m, v += dt a / 2 (a, noise
// included, also drives the next first half step);
// L += dt tau / 2 + dt (-gamma_r L / I + sqrt(6 gamma_r kT / dt) u').
__global__ void k_b_step_two(double4* __restrict__ vel,
                             double4* __restrict__ accel,
                             const double4* __restrict__ force,
                             double4* __restrict__ angmom,
                             const double4* __restrict__ torque,
                             uint32_t N,
                             const uint32_t* __restrict__ rep_of,
                             const RepInfo* __restrict__ reps,
                             double dt,
                             double gamma,
                             double gamma_r,
                             double inertia,
                             uint16_t seed,
                             const unsigned long long* __restrict__ timestep)
    {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N)
        return;
    const uint32_t r = rep_of[i];
    const RepInfo& R = reps[r];
    const uint64_t t = timestep[0];
    const uint32_t local = i - R.first;
    double u[3];
    double4 v = vel[i];
    const double4 f = force[i];
    b_noise(seed, t, local, r, 0u, u);
    const double ct = sqrt(6.0 * gamma * R.kT / dt);
    const double minv = 1.0 / v.w;
    const double4 a = make_double4((f.x - gamma * v.x + ct * u[0]) * minv,
                                   (f.y - gamma * v.y + ct * u[1]) * minv,
                                   (f.z - gamma * v.z + ct * u[2]) * minv,
                                   0.0);
    v.x += 0.5 * dt * a.x;
    v.y += 0.5 * dt * a.y;
    v.z += 0.5 * dt * a.z;
    vel[i] = v;
    accel[i] = a;

    double4 L = angmom[i];
    const double4 tq = torque[i];
    b_noise(seed, t, local, r, 1u, u);
    // the rotational thermostat acts once per step (the first half step uses
    // the pair torque only), so its drag and noise take the full dt
    const double cr = sqrt(6.0 * gamma_r * R.kT / dt);
    L.x += 0.5 * dt * tq.x + dt * (-gamma_r * L.x / inertia + cr * u[0]);
    L.y += 0.5 * dt * tq.y + dt * (-gamma_r * L.y / inertia + cr * u[1]);
    L.z += 0.5 * dt * tq.z + dt * (-gamma_r * L.z / inertia + cr * u[2]);
    angmom[i] = L;
    }

// a = F / m (start of a run, no thermostat terms).
__global__ void k_b_accel(double4* __restrict__ accel, const double4* __restrict__ force, const double4* __restrict__ vel, uint32_t N)
    {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N)
        return;
    const double4 f = force[i];
    const double minv = 1.0 / vel[i].w;
    accel[i] = make_double4(f.x * minv, f.y * minv, f.z * minv, 0.0);
    }

__global__ void k_b_tick(unsigned long long* timestep)
    {
    if (blockIdx.x == 0 && threadIdx.x == 0)
        timestep[0] += 1ull;
    }

// Per-replica sums (one workgroup per replica, fixed order):
// out[4 r + 0] translational KE, [1] rotational KE, [2] potential energy,
// [3] virial trace (xx + yy + zz).
__global__ void __launch_bounds__(256) k_b_thermo(const double4* __restrict__ vel,
                                                  const double4* __restrict__ angmom,
                                                  const double4* __restrict__ force,
                                                  const double* __restrict__ virial,
                                                  uint32_t vpitch,
                                                  const RepInfo* __restrict__ reps,
                                                  double inertia,
                                                  double* __restrict__ out)
    {
    const RepInfo& R = reps[blockIdx.x];
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (uint32_t i = R.first + threadIdx.x; i < R.first + R.count; i += blockDim.x)
        {
        if (vel)
            {
            const double4 v = vel[i];
            acc[0] += 0.5 * v.w * (v.x * v.x + v.y * v.y + v.z * v.z);
            }
        if (angmom)
            {
            const double4 L = angmom[i];
            acc[1] += 0.5 * (L.x * L.x + L.y * L.y + L.z * L.z) / inertia;
            }
        acc[2] += force[i].w;
        if (virial)
            acc[3] += virial[i] + virial[3 * (size_t)vpitch + i] + virial[5 * (size_t)vpitch + i];
        }
    __shared__ double s[4][256];
    for (int f = 0; f < 4; ++f)
        s[f][threadIdx.x] = acc[f];
    __syncthreads();
    for (uint32_t off = blockDim.x / 2; off > 0; off >>= 1)
        {
        if (threadIdx.x < off)
            for (int f = 0; f < 4; ++f)
                s[f][threadIdx.x] += s[f][threadIdx.x + off];
        __syncthreads();
        }
    if (threadIdx.x < 4)
        out[4 * (size_t)blockIdx.x + threadIdx.x] = s[threadIdx.x][0];
    }

} // namespace azp

using namespace azp;

namespace
{
constexpr uint32_t kFlagOverflow = 0;
constexpr uint32_t kFlagRebuild = 1;
constexpr uint32_t kFlagBad = 2;
constexpr uint32_t kNumFlags = 4;
} // namespace

struct azp_batch
    {
    // layout
    uint32_t R = 0;
    uint32_t N = 0;
    uint32_t ntypes = 0;
    int evaluator = AZP_EVAL_TWO_PATCH_MORSE;
    std::vector<uint32_t> offsets; // R + 1
    std::vector<azp_box> boxes;
    std::vector<double> kT;

    // interaction
    std::vector<azp_param_two_patch_morse> params;
    std::vector<uint8_t> params_set;
    std::vector<double> r_cut;
    double r_buff = 0.4;
    int shift_mode = AZP_SHIFT_NONE;
    int precision = AZP_PREC_MIXED;
    bool compute_virial = true;
    bool tables_dirty = true;

    // Langevin
    double dt = 0.005, gamma = 1.0, gamma_r = 1.0, inertia = 1.0;
    uint16_t seed = 0;

    // device state
    DeviceArray<RepInfo> d_reps;
    DeviceArray<uint32_t> rep_of;
    DeviceArray<azp_param_two_patch_morse> d_params;
    DeviceArray<double> d_rcutsq, d_rlistsq;
    DeviceArray<uint32_t> cell_of, slot_of, cell_count, cell_start, members, cell_rep, scan_sums;
    DeviceArray<uint32_t> n_neigh, list, flags, counter, rep_flag;
    DeviceArray<double4> last_pos;
    DeviceArray<float4> dirs;
    DeviceArray<unsigned long long> d_timestep, d_builds; // builds: [0] rebuild events, [1] replica rebuilds
    uint32_t* h_flags = nullptr; // pinned
    uint32_t ncell = 0;
    uint32_t pitch = 0;
    uint32_t max_neighbors = 0;
    bool layout_uploaded = false;
    bool have_list = false;
    bool prepared = false;

    azp_batch_arrays arr = {};
    bool bound = false;

    azp_batch_tuning tune = {256, 1, 4, 256, 0.0};
    int num_cu = 0;

    // graphs of `graph_chunk` steps and of one step
    hipStream_t own = nullptr;
    hipEvent_t ev_in = nullptr, ev_out = nullptr;
    hipGraphExec_t exec_chunk = nullptr, exec_one = nullptr;
    uint32_t graph_chunk = 10;
    uint64_t graph_launches = 0;

    ~azp_batch()
        {
        drop_graphs();
        if (ev_in)
            (void)hipEventDestroy(ev_in);
        if (ev_out)
            (void)hipEventDestroy(ev_out);
        if (own)
            (void)hipStreamDestroy(own);
        if (h_flags)
            (void)hipHostFree(h_flags);
        }

    void drop_graphs()
        {
        if (exec_chunk)
            (void)hipGraphExecDestroy(exec_chunk);
        if (exec_one)
            (void)hipGraphExecDestroy(exec_one);
        exec_chunk = exec_one = nullptr;
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

namespace
{

int batch_check_ready(const azp_batch* b, bool for_run)
    {
    if (!b->bound)
        return set_error(AZP_ERR_STATE, "bind the particle arrays before computing");
    for (uint32_t t = 0; t < b->ntypes * b->ntypes; ++t)
        if (!b->params_set[t])
            return set_error(AZP_ERR_STATE, "parameters of type pair (%u, %u) are not set", t / b->ntypes,
                             t % b->ntypes);
    if (b->r_list_max() <= 0.0)
        return set_error(AZP_ERR_STATE, "set r_cut before computing");
    const double rmax = b->r_list_max();
    for (uint32_t r = 0; r < b->R; ++r)
        {
        const azp_box& bx = b->boxes[r];
        if (bx.Lx < 2.0 * rmax || bx.Ly < 2.0 * rmax || bx.Lz < 2.0 * rmax)
            return set_error(AZP_ERR_BOX,
                             "replica %u box (%g, %g, %g) is smaller than twice the neighbor range %g", r, bx.Lx,
                             bx.Ly, bx.Lz, rmax);
        }
    const azp_batch_arrays& a = b->arr;
    if (!a.d_pos || !a.d_orientation || !a.d_force || !a.d_torque)
        return set_error(AZP_ERR_INVALID, "positions, orientations, force and torque arrays are required");
    if (b->precision == AZP_PREC_MIXED && !a.d_pos_fixed)
        return set_error(AZP_ERR_INVALID, "mixed precision needs d_pos_fixed");
    if (b->compute_virial && (!a.d_virial || a.virial_pitch < b->N))
        return set_error(AZP_ERR_INVALID, "virial array missing or pitch %u < N %u", a.virial_pitch, b->N);
    if (for_run && (!a.d_vel || !a.d_angmom || !a.d_accel))
        return set_error(AZP_ERR_INVALID, "time steps need velocity, angular momentum and acceleration arrays");
    return AZP_OK;
    }

// Uploads the replica table, per-particle replica index, cutoffs and params.
int batch_upload_tables(azp_batch* b, hipStream_t s)
    {
    if (!b->tables_dirty && b->layout_uploaded)
        return AZP_OK;
    const double rmax = b->r_list_max();
    std::vector<RepInfo> reps(b->R);
    uint32_t cells = 0;
    for (uint32_t r = 0; r < b->R; ++r)
        {
        const double L[3] = {b->boxes[r].Lx, b->boxes[r].Ly, b->boxes[r].Lz};
        const uint32_t n_r = b->offsets[r + 1] - b->offsets[r];
        const CellGrid g = cell_grid_for(L, rmax, n_r);
        RepInfo& ri = reps[r];
        for (int k = 0; k < 3; ++k)
            {
            ri.L[k] = L[k];
            ri.Linv[k] = 1.0 / L[k];
            ri.fx_scale[k] = (float)(L[k] / 4294967296.0);
            ri.n[k] = g.n[k];
            }
        ri.cell_base = cells;
        ri.first = b->offsets[r];
        ri.count = n_r;
        ri.kT = b->kT[r];
        const uint64_t nc = (uint64_t)g.n[0] * g.n[1] * g.n[2];
        if ((uint64_t)cells + nc > 0xffffffffull / 2)
            return set_error(AZP_ERR_INVALID, "too many cells in the batch");
        cells += (uint32_t)nc;
        }
    b->ncell = cells;
    std::vector<uint32_t> cell_rep(cells);
    for (uint32_t r = 0; r < b->R; ++r)
        {
        const uint32_t nc = (uint32_t)reps[r].n[0] * reps[r].n[1] * reps[r].n[2];
        std::fill(cell_rep.begin() + reps[r].cell_base, cell_rep.begin() + reps[r].cell_base + nc, r);
        }
    AZP_TRY(b->d_reps.reserve(b->R));
    AZP_TRY(b->cell_rep.reserve(cells));
    AZP_TRY(b->rep_flag.reserve(b->R));
    AZP_HIP(hipMemcpyAsync(b->d_reps.ptr, reps.data(), b->R * sizeof(RepInfo), hipMemcpyHostToDevice, s));
    AZP_HIP(hipMemcpyAsync(b->cell_rep.ptr, cell_rep.data(), cells * sizeof(uint32_t), hipMemcpyHostToDevice, s));
    if (!b->layout_uploaded)
        {
        std::vector<uint32_t> rep_of(b->N);
        for (uint32_t r = 0; r < b->R; ++r)
            for (uint32_t i = b->offsets[r]; i < b->offsets[r + 1]; ++i)
                rep_of[i] = r;
        AZP_TRY(b->rep_of.reserve(b->N));
        AZP_HIP(hipMemcpyAsync(b->rep_of.ptr, rep_of.data(), b->N * sizeof(uint32_t), hipMemcpyHostToDevice, s));
        }
    const size_t nt2 = (size_t)b->ntypes * b->ntypes;
    std::vector<double> rcsq(nt2), rlsq(nt2);
    for (size_t k = 0; k < nt2; ++k)
        {
        const double rc = b->r_cut[k];
        rcsq[k] = rc > 0.0 ? rc * rc : 0.0;
        const double rl = rc + b->r_buff;
        rlsq[k] = rc > 0.0 ? rl * rl : 0.0;
        }
    AZP_TRY(b->d_params.reserve(nt2));
    AZP_TRY(b->d_rcutsq.reserve(nt2));
    AZP_TRY(b->d_rlistsq.reserve(nt2));
    AZP_HIP(hipMemcpyAsync(b->d_params.ptr, b->params.data(), nt2 * sizeof(azp_param_two_patch_morse),
                           hipMemcpyHostToDevice, s));
    AZP_HIP(hipMemcpyAsync(b->d_rcutsq.ptr, rcsq.data(), nt2 * sizeof(double), hipMemcpyHostToDevice, s));
    AZP_HIP(hipMemcpyAsync(b->d_rlistsq.ptr, rlsq.data(), nt2 * sizeof(double), hipMemcpyHostToDevice, s));
    // host arrays above are released at return: wait for the copies
    AZP_HIP(hipStreamSynchronize(s));

    AZP_TRY(b->cell_of.reserve(b->N));
    AZP_TRY(b->slot_of.reserve(b->N));
    AZP_TRY(b->members.reserve(b->N));
    AZP_TRY(b->n_neigh.reserve(b->N));
    AZP_TRY(b->last_pos.reserve(b->N));
    AZP_TRY(b->dirs.reserve(b->N));
    AZP_TRY(b->cell_count.reserve(b->ncell));
    AZP_TRY(b->cell_start.reserve((size_t)b->ncell + 1));
    AZP_TRY(b->scan_sums.reserve(scan_chunks(b->ncell)));
    AZP_TRY(b->flags.reserve(kNumFlags));
    AZP_TRY(b->counter.reserve(1));
    if (!b->d_timestep.ptr)
        {
        AZP_TRY(b->d_timestep.reserve(1));
        AZP_TRY(b->d_builds.reserve(2));
        hipLaunchKernelGGL(k_b_fill64, dim3(1), dim3(64), 0, s, b->d_timestep.ptr, 0ull);
        hipLaunchKernelGGL(k_b_fill64, dim3(1), dim3(64), 0, s, b->d_builds.ptr, 0ull);
        hipLaunchKernelGGL(k_b_fill64, dim3(1), dim3(64), 0, s, b->d_builds.ptr + 1, 0ull);
        hipLaunchKernelGGL(k_b_fill, dim3(1), dim3(64), 0, s, b->flags.ptr, 0u, kNumFlags, nullptr);
        AZP_LAUNCH_CHECK();
        }
    if (b->max_neighbors == 0)
        {
        // capacity from the densest replica, with margin for fluctuations
        double dens = 0.0;
        for (uint32_t r = 0; r < b->R; ++r)
            dens = std::max(dens, (b->offsets[r + 1] - b->offsets[r])
                                      / (b->boxes[r].Lx * b->boxes[r].Ly * b->boxes[r].Lz));
        const double est = dens * 4.18879020478639 * rmax * rmax * rmax;
        b->max_neighbors = ((uint32_t)std::min(1.5 * est + 32.0, (double)b->N + 8.0) + 7u) & ~7u;
        }
    b->pitch = wave_pitch(b->N);
    AZP_TRY(b->list.reserve((size_t)b->max_neighbors * b->pitch));
    b->layout_uploaded = true;
    b->tables_dirty = false;
    b->have_list = false;
    b->drop_graphs();
    return AZP_OK;
    }

// Enqueues the gated rebuild (bin, scan, scatter, sort, walk, snapshot).
void enqueue_rebuild(azp_batch* b, hipStream_t s)
    {
    const unsigned B = 256;
    const uint32_t N = b->N;
    const uint32_t* gate = b->flags.ptr + kFlagRebuild;
    const uint32_t* rf = b->rep_flag.ptr;
    const double4* pos = reinterpret_cast<const double4*>(b->arr.d_pos);
    hipLaunchKernelGGL(k_b_clear_counts, dim3(grid_for(b->ncell, B)), dim3(B), 0, s, b->cell_count.ptr, b->ncell,
                       b->cell_rep.ptr, rf, gate);
    hipLaunchKernelGGL(k_b_bin, dim3(grid_for(N, B)), dim3(B), 0, s, pos, N, b->rep_of.ptr, b->d_reps.ptr, rf,
                       b->cell_of.ptr, b->slot_of.ptr, b->cell_count.ptr, gate);
    // the scan covers every replica; the counts of replicas not rebuilt are
    // unchanged, so are their cell ranges
    enqueue_scan(b->cell_count.ptr, b->ncell, b->scan_sums.ptr, b->cell_start.ptr, gate, s);
    hipLaunchKernelGGL(k_b_scatter, dim3(grid_for(N, B)), dim3(B), 0, s, N, b->rep_of.ptr, rf, b->cell_of.ptr,
                       b->slot_of.ptr, b->cell_start.ptr, b->members.ptr, gate);
    hipLaunchKernelGGL(k_b_sort_cells, dim3((b->ncell + kSortWaves - 1) / kSortWaves), dim3(64 * kSortWaves), 0, s,
                       b->ncell, b->cell_rep.ptr, rf, b->cell_start.ptr, b->members.ptr, gate);
    BatchWalkArgs wa;
    wa.pos = pos;
    wa.N = N;
    wa.rep_of = b->rep_of.ptr;
    wa.reps = b->d_reps.ptr;
    wa.cell_of = b->cell_of.ptr;
    wa.start = b->cell_start.ptr;
    wa.members = b->members.ptr;
    wa.r_listsq = b->d_rlistsq.ptr;
    wa.ntypes = b->ntypes;
    wa.n_neigh = b->n_neigh.ptr;
    wa.nlist = b->list.ptr;
    wa.pitch = b->pitch;
    wa.max_neighbors = b->max_neighbors;
    wa.flags = b->flags.ptr;
    wa.gate = gate;
    wa.rep_flag = rf;
    const unsigned wb = b->tune.walk_block ? b->tune.walk_block : 256;
    hipLaunchKernelGGL(k_b_walk, dim3(grid_for(N, wb)), dim3(wb), 0, s, wa);
    hipLaunchKernelGGL(k_b_copy_last, dim3(grid_for(N, B)), dim3(B), 0, s, pos, b->last_pos.ptr, N, b->rep_of.ptr,
                       b->d_reps.ptr, rf, b->d_builds.ptr, gate);
    }

// Rebuild decision on the device: every replica flagged when forced, else
// the per-replica displacement check.
void enqueue_check(azp_batch* b, bool force, hipStream_t s)
    {
    const unsigned B = 256;
    const uint32_t v = force ? 1u : 0u;
    hipLaunchKernelGGL(k_b_fill, dim3(1), dim3(64), 0, s, b->flags.ptr + kFlagRebuild, v, 1u, nullptr);
    hipLaunchKernelGGL(k_b_fill, dim3(grid_for(b->R, B)), dim3(B), 0, s, b->rep_flag.ptr, v, b->R, nullptr);
    if (!force)
        {
        const double half = 0.5 * b->r_buff;
        hipLaunchKernelGGL(k_b_check, dim3(grid_for(b->N, B)), dim3(B), 0, s,
                           reinterpret_cast<const double4*>(b->arr.d_pos), b->last_pos.ptr, b->N, b->rep_of.ptr,
                           b->d_reps.ptr, half * half, b->rep_flag.ptr, b->flags.ptr);
        }
    }

template<bool MIXED, bool VIRIAL, bool SINGLE>
int launch_forces_t(azp_batch* b, const azp_batch_tuning& t, hipStream_t s)
    {
    using R = std::conditional_t<MIXED, float, double>;
    BatchForceArgs fa;
    fa.pos = reinterpret_cast<const double4*>(b->arr.d_pos);
    fa.pos_fixed = reinterpret_cast<const int4*>(b->arr.d_pos_fixed);
    fa.orientation = reinterpret_cast<const double4*>(b->arr.d_orientation);
    fa.N = b->N;
    fa.rep_of = b->rep_of.ptr;
    fa.reps = b->d_reps.ptr;
    fa.n_neigh = b->n_neigh.ptr;
    fa.nlist = b->list.ptr;
    fa.pitch = b->pitch;
    fa.params = b->d_params.ptr;
    fa.rcutsq = b->d_rcutsq.ptr;
    fa.ntypes = b->ntypes;
    fa.shift_mode = b->shift_mode;
    fa.dirs = b->dirs.ptr;
    fa.force = reinterpret_cast<double4*>(b->arr.d_force);
    fa.torque = reinterpret_cast<double4*>(b->arr.d_torque);
    fa.virial = b->arr.d_virial;
    fa.vpitch = b->arr.virial_pitch;
    fa.counter = b->counter.ptr;
    fa.n_chunks = (b->N + 63u) / 64u;
    fa.grab = std::max(1u, t.grab);
    const unsigned block = t.block_size;
    const unsigned waves_per_block = block / 64u;
    unsigned grid = (fa.n_chunks + waves_per_block - 1) / waves_per_block;
    if (t.workgroups_per_cu > 0)
        grid = std::min<unsigned>(grid, (unsigned)b->num_cu * t.workgroups_per_cu);
    grid = std::max(grid, 1u);
    const size_t shmem = SINGLE ? 0 : (size_t)b->ntypes * b->ntypes * sizeof(TPMPrep<R>);
    hipLaunchKernelGGL(k_b_fill, dim3(1), dim3(64), 0, s, b->counter.ptr, 0u, 1u, nullptr);
    hipLaunchKernelGGL((k_b_forces<MIXED, VIRIAL, SINGLE>), dim3(grid), dim3(block), shmem, s, fa);
    return AZP_OK;
    }

int enqueue_forces(azp_batch* b, const azp_batch_tuning& t, hipStream_t s)
    {
    const unsigned block = t.block_size;
    if (block < 64 || block > 512 || block % 64 != 0)
        return set_error(AZP_ERR_INVALID, "force block size %u must be a multiple of 64 in [64, 512]", block);
    if ((size_t)b->ntypes * b->ntypes * sizeof(TPMPrep<double>) > 64 * 1024)
        return set_error(AZP_ERR_UNSUPPORTED, "too many particle types for LDS staging");
    if (b->precision == AZP_PREC_MIXED)
        {
        hipLaunchKernelGGL(k_b_pack_fixed, dim3(grid_for(b->N, 256)), dim3(256), 0, s,
                           reinterpret_cast<const double4*>(b->arr.d_pos),
                           reinterpret_cast<const double4*>(b->arr.d_orientation), b->N, b->rep_of.ptr,
                           b->d_reps.ptr, reinterpret_cast<int4*>(b->arr.d_pos_fixed), b->dirs.ptr);
        }
    const bool single = b->ntypes == 1;
    const bool mixed = b->precision == AZP_PREC_MIXED;
    const bool vir = b->compute_virial;
    if (mixed)
        {
        if (vir)
            return single ? launch_forces_t<true, true, true>(b, t, s) : launch_forces_t<true, true, false>(b, t, s);
        return single ? launch_forces_t<true, false, true>(b, t, s) : launch_forces_t<true, false, false>(b, t, s);
        }
    if (vir)
        return single ? launch_forces_t<false, true, true>(b, t, s) : launch_forces_t<false, true, false>(b, t, s);
    return single ? launch_forces_t<false, false, true>(b, t, s) : launch_forces_t<false, false, false>(b, t, s);
    }

// One MD step: first half step, rebuild check (+ gated rebuild), forces,
// Langevin second half step, timestep + 1.  Host-free: capturable.
int enqueue_step(azp_batch* b, hipStream_t s)
    {
    const unsigned B = 256;
    const uint32_t N = b->N;
    const azp_batch_arrays& a = b->arr;
    hipLaunchKernelGGL(k_b_step_one, dim3(grid_for(N, B)), dim3(B), 0, s, reinterpret_cast<double4*>(a.d_pos),
                       reinterpret_cast<double4*>(a.d_vel), reinterpret_cast<const double4*>(a.d_accel),
                       reinterpret_cast<double4*>(a.d_orientation), reinterpret_cast<double4*>(a.d_angmom),
                       reinterpret_cast<const double4*>(a.d_torque), N, b->rep_of.ptr, b->d_reps.ptr, b->dt,
                       b->inertia);
    enqueue_check(b, false, s);
    enqueue_rebuild(b, s);
    AZP_TRY(enqueue_forces(b, b->tune, s));
    hipLaunchKernelGGL(k_b_step_two, dim3(grid_for(N, B)), dim3(B), 0, s, reinterpret_cast<double4*>(a.d_vel),
                       reinterpret_cast<double4*>(a.d_accel), reinterpret_cast<const double4*>(a.d_force),
                       reinterpret_cast<double4*>(a.d_angmom), reinterpret_cast<const double4*>(a.d_torque), N,
                       b->rep_of.ptr, b->d_reps.ptr, b->dt, b->gamma, b->gamma_r, b->inertia, b->seed,
                       b->d_timestep.ptr);
    hipLaunchKernelGGL(k_b_tick, dim3(1), dim3(64), 0, s, b->d_timestep.ptr);
    return AZP_OK;
    }

// Reads the sticky flags; grows the list on overflow (invalidating graphs).
int batch_read_flags(azp_batch* b, hipStream_t s, bool* overflowed)
    {
    *overflowed = false;
    AZP_HIP(hipMemcpyAsync(b->h_flags, b->flags.ptr, kNumFlags * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    AZP_HIP(hipStreamSynchronize(s));
    const uint32_t bad = b->h_flags[kFlagBad];
    if (bad)
        {
        // reported once; the list built from such data is not kept
        b->have_list = false;
        hipLaunchKernelGGL(k_b_fill, dim3(1), dim3(64), 0, s, b->flags.ptr + kFlagBad, 0u, 1u, nullptr);
        AZP_LAUNCH_CHECK();
        if (bad & 2u)
            return set_error(AZP_ERR_INVALID, "a particle position is not finite (NaN or inf)");
        return set_error(AZP_ERR_INVALID, "a particle type index is outside [0, %u)", b->ntypes);
        }
    const uint32_t need = b->h_flags[kFlagOverflow];
    if (need > 0)
        {
        *overflowed = true;
        b->max_neighbors = ((uint32_t)(need * 1.2) + 8u) & ~7u;
        b->drop_graphs();
        b->have_list = false;
        AZP_TRY(b->list.reserve((size_t)b->max_neighbors * b->pitch));
        hipLaunchKernelGGL(k_b_fill, dim3(1), dim3(64), 0, s, b->flags.ptr + kFlagOverflow, 0u, 1u, nullptr);
        AZP_LAUNCH_CHECK();
        }
    return AZP_OK;
    }

struct StreamScope
    {
    // Work runs on the batch's own non-blocking stream, ordered after the
    // caller's stream on entry and before it on exit.
    azp_batch* b;
    hipStream_t caller;
    int enter(azp_batch* bb, void* stream)
        {
        b = bb;
        caller = (hipStream_t)stream;
        if (!b->own)
            {
            AZP_HIP(hipStreamCreateWithFlags(&b->own, hipStreamNonBlocking));
            AZP_HIP(hipEventCreateWithFlags(&b->ev_in, hipEventDisableTiming));
            AZP_HIP(hipEventCreateWithFlags(&b->ev_out, hipEventDisableTiming));
            }
        AZP_HIP(hipEventRecord(b->ev_in, caller));
        AZP_HIP(hipStreamWaitEvent(b->own, b->ev_in, 0));
        return AZP_OK;
        }
    int leave()
        {
        AZP_HIP(hipEventRecord(b->ev_out, b->own));
        AZP_HIP(hipStreamWaitEvent(caller, b->ev_out, 0));
        return AZP_OK;
        }
    };

int batch_device_init(azp_batch* b)
    {
    if (b->num_cu == 0)
        {
        int dev = 0;
        AZP_HIP(hipGetDevice(&dev));
        AZP_HIP(hipDeviceGetAttribute(&b->num_cu, hipDeviceAttributeMultiprocessorCount, dev));
        if (b->num_cu <= 0)
            b->num_cu = 1;
        }
    if (!b->h_flags)
        {
        if (hipHostMalloc((void**)&b->h_flags, kNumFlags * sizeof(uint32_t), hipHostMallocDefault) != hipSuccess)
            {
            b->h_flags = nullptr;
            return set_error(AZP_ERR_NOMEM, "pinned host allocation failed");
            }
        }
    return AZP_OK;
    }

// Forces at the current positions, rebuilding the list when needed; grows the
// list and repeats when a particle has more neighbors than the capacity.
int batch_compute(azp_batch* b, bool force_rebuild, hipStream_t s)
    {
    AZP_TRY(batch_device_init(b));
    AZP_TRY(batch_upload_tables(b, s));
    bool force = force_rebuild || !b->have_list;
    for (int attempt = 0; attempt < 4; ++attempt)
        {
        enqueue_check(b, force, s);
        enqueue_rebuild(b, s);
        AZP_TRY(enqueue_forces(b, b->tune, s));
        AZP_LAUNCH_CHECK();
        bool overflowed = false;
        AZP_TRY(batch_read_flags(b, s, &overflowed));
        if (!overflowed)
            {
            b->have_list = true;
            return AZP_OK;
            }
        force = true;
        }
    return set_error(AZP_ERR_STATE, "neighbor list did not converge after growing its capacity");
    }

int capture(azp_batch* b, uint32_t steps, hipGraphExec_t* out)
    {
    hipGraph_t g = nullptr;
    AZP_HIP(hipStreamBeginCapture(b->own, hipStreamCaptureModeThreadLocal));
    int st = AZP_OK;
    for (uint32_t k = 0; k < steps && st == AZP_OK; ++k)
        st = enqueue_step(b, b->own);
    const hipError_t e = hipStreamEndCapture(b->own, &g);
    if (st != AZP_OK)
        {
        if (g)
            (void)hipGraphDestroy(g);
        return st;
        }
    if (e != hipSuccess)
        return set_error(AZP_ERR_HIP, "graph capture failed: %s", hipGetErrorString(e));
    const hipError_t ie = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ie != hipSuccess)
        return set_error(AZP_ERR_HIP, "graph instantiation failed: %s", hipGetErrorString(ie));
    return AZP_OK;
    }

} // namespace

extern "C" int azp_batch_create(azp_batch** out,
                                uint32_t n_replicas,
                                const uint32_t* sizes,
                                const azp_box* boxes,
                                uint32_t ntypes,
                                int evaluator)
    {
    if (!out)
        return set_error(AZP_ERR_INVALID, "null out");
    *out = nullptr;
    if (n_replicas == 0 || !sizes || !boxes)
        return set_error(AZP_ERR_INVALID, "a batch needs at least one replica with sizes and boxes");
    if (ntypes == 0 || ntypes > 64)
        return set_error(AZP_ERR_INVALID, "ntypes must be in [1, 64]");
    if (evaluator != AZP_EVAL_TWO_PATCH_MORSE)
        return set_error(AZP_ERR_UNSUPPORTED, "batched replicas run the TwoPatchMorse evaluator");
    uint64_t total = 0;
    for (uint32_t r = 0; r < n_replicas; ++r)
        {
        if (sizes[r] == 0)
            return set_error(AZP_ERR_INVALID, "replica %u is empty", r);
        if (!box_valid(boxes[r]))
            return set_error(AZP_ERR_BOX, "replica %u box lengths must be positive", r);
        total += sizes[r];
        }
    if (total > 0x7fffff00ull)
        return set_error(AZP_ERR_INVALID, "too many particles in one batch");
    azp_batch* b = new azp_batch();
    b->R = n_replicas;
    b->N = (uint32_t)total;
    b->ntypes = ntypes;
    b->evaluator = evaluator;
    b->offsets.resize(n_replicas + 1);
    b->offsets[0] = 0;
    for (uint32_t r = 0; r < n_replicas; ++r)
        b->offsets[r + 1] = b->offsets[r] + sizes[r];
    b->boxes.assign(boxes, boxes + n_replicas);
    b->kT.assign(n_replicas, 1.0);
    b->params.assign((size_t)ntypes * ntypes, azp_param_two_patch_morse{});
    b->params_set.assign((size_t)ntypes * ntypes, 0);
    b->r_cut.assign((size_t)ntypes * ntypes, 0.0);
    *out = b;
    return AZP_OK;
    }

extern "C" int azp_batch_destroy(azp_batch* b)
    {
    delete b;
    return AZP_OK;
    }

extern "C" int azp_batch_size(const azp_batch* b, uint32_t* n_replicas, uint32_t* n_total)
    {
    if (!b)
        return set_error(AZP_ERR_INVALID, "null batch");
    if (n_replicas)
        *n_replicas = b->R;
    if (n_total)
        *n_total = b->N;
    return AZP_OK;
    }

extern "C" int azp_batch_set_params(azp_batch* b, uint32_t ti, uint32_t tj, const double* values)
    {
    if (!b || !values)
        return set_error(AZP_ERR_INVALID, "null argument");
    if (ti >= b->ntypes || tj >= b->ntypes)
        return set_error(AZP_ERR_INVALID, "type index out of range");
    azp_param_two_patch_morse p;
    AZP_TRY(azp_param_pack(b->evaluator, values, &p));
    b->params[ti * b->ntypes + tj] = p;
    b->params[tj * b->ntypes + ti] = p;
    b->params_set[ti * b->ntypes + tj] = b->params_set[tj * b->ntypes + ti] = 1;
    b->tables_dirty = true;
    return AZP_OK;
    }

extern "C" int azp_batch_set_r_cut(azp_batch* b, uint32_t ti, uint32_t tj, double r_cut)
    {
    if (!b)
        return set_error(AZP_ERR_INVALID, "null batch");
    if (ti >= b->ntypes || tj >= b->ntypes || !(r_cut >= 0.0))
        return set_error(AZP_ERR_INVALID, "bad r_cut arguments");
    b->r_cut[ti * b->ntypes + tj] = b->r_cut[tj * b->ntypes + ti] = r_cut;
    b->tables_dirty = true;
    b->max_neighbors = 0;
    return AZP_OK;
    }

extern "C" int azp_batch_set_r_buff(azp_batch* b, double r_buff)
    {
    if (!b || !(r_buff >= 0.0))
        return set_error(AZP_ERR_INVALID, "r_buff must be >= 0");
    b->r_buff = r_buff;
    b->tables_dirty = true;
    b->max_neighbors = 0;
    return AZP_OK;
    }

extern "C" int azp_batch_set_mode(azp_batch* b, int shift_mode, int precision, int compute_virial)
    {
    if (!b)
        return set_error(AZP_ERR_INVALID, "null batch");
    if (shift_mode != AZP_SHIFT_NONE && shift_mode != AZP_SHIFT_SHIFT)
        return set_error(AZP_ERR_UNSUPPORTED, "anisotropic pairs accept modes none and shift");
    if (precision != AZP_PREC_FP64 && precision != AZP_PREC_MIXED)
        return set_error(AZP_ERR_INVALID, "unknown precision %d", precision);
    if (b->shift_mode != shift_mode || b->precision != precision || b->compute_virial != (compute_virial != 0))
        b->drop_graphs();
    b->shift_mode = shift_mode;
    b->precision = precision;
    b->compute_virial = compute_virial != 0;
    return AZP_OK;
    }

extern "C" int azp_batch_set_langevin(azp_batch* b,
                                      const double* kT,
                                      double dt,
                                      double gamma,
                                      double gamma_r,
                                      double inertia,
                                      uint16_t seed)
    {
    if (!b || !kT)
        return set_error(AZP_ERR_INVALID, "null argument");
    if (!(dt > 0.0) || !(gamma >= 0.0) || !(gamma_r >= 0.0) || !(inertia > 0.0))
        return set_error(AZP_ERR_INVALID, "need dt > 0, gamma >= 0, gamma_r >= 0 and inertia > 0");
    for (uint32_t r = 0; r < b->R; ++r)
        if (!(kT[r] >= 0.0))
            return set_error(AZP_ERR_INVALID, "replica %u kT must be >= 0", r);
    b->kT.assign(kT, kT + b->R);
    b->dt = dt;
    b->gamma = gamma;
    b->gamma_r = gamma_r;
    b->inertia = inertia;
    b->seed = seed;
    b->tables_dirty = true;
    return AZP_OK;
    }

extern "C" int azp_batch_bind(azp_batch* b, const azp_batch_arrays* arrays)
    {
    if (!b || !arrays)
        return set_error(AZP_ERR_INVALID, "null argument");
    b->arr = *arrays;
    b->bound = true;
    b->prepared = false;
    b->have_list = false;
    b->drop_graphs();
    return AZP_OK;
    }

extern "C" int azp_batch_set_timestep(azp_batch* b, uint64_t timestep, void* stream)
    {
    if (!b)
        return set_error(AZP_ERR_INVALID, "null batch");
    AZP_TRY(batch_device_init(b));
    StreamScope sc;
    AZP_TRY(sc.enter(b, stream));
    AZP_TRY(batch_upload_tables(b, b->own));
    hipLaunchKernelGGL(k_b_fill64, dim3(1), dim3(64), 0, b->own, b->d_timestep.ptr, (unsigned long long)timestep);
    AZP_LAUNCH_CHECK();
    return sc.leave();
    }

extern "C" int azp_batch_compute(azp_batch* b, int force_rebuild, void* stream)
    {
    if (!b)
        return set_error(AZP_ERR_INVALID, "null batch");
    AZP_TRY(batch_check_ready(b, false));
    AZP_TRY(batch_device_init(b));
    StreamScope sc;
    AZP_TRY(sc.enter(b, stream));
    AZP_TRY(batch_compute(b, force_rebuild != 0, b->own));
    return sc.leave();
    }

extern "C" int azp_batch_run(azp_batch* b, uint32_t nsteps, int use_graph, void* stream)
    {
    if (!b)
        return set_error(AZP_ERR_INVALID, "null batch");
    AZP_TRY(batch_check_ready(b, true));
    AZP_TRY(batch_device_init(b));
    StreamScope sc;
    AZP_TRY(sc.enter(b, stream));
    hipStream_t s = b->own;
    if (!b->prepared || b->tables_dirty || !b->have_list)
        {
        AZP_TRY(batch_compute(b, !b->have_list, s));
        hipLaunchKernelGGL(k_b_accel, dim3(grid_for(b->N, 256)), dim3(256), 0, s,
                           reinterpret_cast<double4*>(b->arr.d_accel), reinterpret_cast<const double4*>(b->arr.d_force),
                           reinterpret_cast<const double4*>(b->arr.d_vel), b->N);
        AZP_LAUNCH_CHECK();
        b->prepared = true;
        }
    uint32_t left = nsteps;
    if (use_graph)
        {
        if (left >= b->graph_chunk && !b->exec_chunk)
            AZP_TRY(capture(b, b->graph_chunk, &b->exec_chunk));
        while (left >= b->graph_chunk)
            {
            AZP_HIP(hipGraphLaunch(b->exec_chunk, s));
            b->graph_launches++;
            left -= b->graph_chunk;
            }
        if (left > 0 && !b->exec_one)
            AZP_TRY(capture(b, 1, &b->exec_one));
        for (; left > 0; --left)
            {
            AZP_HIP(hipGraphLaunch(b->exec_one, s));
            b->graph_launches++;
            }
        }
    else
        {
        for (; left > 0; --left)
            AZP_TRY(enqueue_step(b, s));
        AZP_LAUNCH_CHECK();
        }
    bool overflowed = false;
    AZP_TRY(batch_read_flags(b, s, &overflowed));
    AZP_TRY(sc.leave());
    if (overflowed)
        return set_error(AZP_ERR_STATE,
                         "a particle had more neighbors than the list capacity during the run; the capacity is now "
                         "%u and the run's last steps used truncated lists: repeat them",
                         b->max_neighbors);
    return AZP_OK;
    }

extern "C" int azp_batch_thermo(azp_batch* b, double* d_out, void* stream)
    {
    if (!b || !d_out)
        return set_error(AZP_ERR_INVALID, "null argument");
    if (!b->bound || !b->layout_uploaded)
        return set_error(AZP_ERR_STATE, "compute forces before reducing");
    StreamScope sc;
    AZP_TRY(sc.enter(b, stream));
    hipLaunchKernelGGL(k_b_thermo, dim3(b->R), dim3(256), 0, b->own, reinterpret_cast<const double4*>(b->arr.d_vel),
                       reinterpret_cast<const double4*>(b->arr.d_angmom),
                       reinterpret_cast<const double4*>(b->arr.d_force), b->compute_virial ? b->arr.d_virial : nullptr,
                       b->arr.virial_pitch, b->d_reps.ptr, b->inertia, d_out);
    AZP_LAUNCH_CHECK();
    return sc.leave();
    }

extern "C" int azp_batch_get_tuning(const azp_batch* b, azp_batch_tuning* out)
    {
    if (!b || !out)
        return set_error(AZP_ERR_INVALID, "null argument");
    *out = b->tune;
    return AZP_OK;
    }

extern "C" int azp_batch_set_tuning(azp_batch* b, const azp_batch_tuning* t)
    {
    if (!b || !t)
        return set_error(AZP_ERR_INVALID, "null argument");
    if (t->block_size < 64 || t->block_size > 512 || t->block_size % 64 != 0)
        return set_error(AZP_ERR_INVALID, "block size must be a multiple of 64 in [64, 512]");
    if (t->grab == 0 || t->grab > 64)
        return set_error(AZP_ERR_INVALID, "grab must be in [1, 64]");
    if (t->workgroups_per_cu > 64)
        return set_error(AZP_ERR_INVALID, "workgroups_per_cu must be in [0, 64]");
    if (t->walk_block != 64 && t->walk_block != 128 && t->walk_block != 256)
        return set_error(AZP_ERR_INVALID, "walk block must be 64, 128 or 256");
    b->tune = *t;
    b->drop_graphs();
    return AZP_OK;
    }

extern "C" int azp_batch_autotune(azp_batch* b, azp_batch_tuning* out, void* stream)
    {
    if (!b)
        return set_error(AZP_ERR_INVALID, "null batch");
    AZP_TRY(batch_check_ready(b, false));
    AZP_TRY(batch_device_init(b));
    StreamScope sc;
    AZP_TRY(sc.enter(b, stream));
    hipStream_t s = b->own;
    AZP_TRY(batch_compute(b, !b->have_list, s));
    hipEvent_t e0, e1;
    AZP_HIP(hipEventCreate(&e0));
    AZP_HIP(hipEventCreate(&e1));
    auto time_it = [&](auto&& enqueue, int reps, float* ms) -> int {
        AZP_TRY(enqueue()); // warm-up
        AZP_HIP(hipEventRecord(e0, s));
        for (int k = 0; k < reps; ++k)
            AZP_TRY(enqueue());
        AZP_HIP(hipEventRecord(e1, s));
        AZP_HIP(hipEventSynchronize(e1));
        AZP_HIP(hipEventElapsedTime(ms, e0, e1));
        *ms /= (float)reps;
        return AZP_OK;
    };
    int st = AZP_OK;
    azp_batch_tuning best = b->tune;
    float best_ms = 1e30f;
    // force kernel: block size x chunks per fetch x persistent grid size
    const unsigned blocks[] = {64, 128, 256, 512};
    const unsigned grabs[] = {1, 2, 4, 8};
    const unsigned wgs[] = {0, 1, 2, 4, 8, 16};
    for (unsigned blk : blocks)
        for (unsigned gr : grabs)
            for (unsigned wg : wgs)
                {
                if (st != AZP_OK)
                    break;
                if (wg > 0 && wg * blk / 64 > 32) // more waves than a CU holds
                    continue;
                if (wg == 0 && gr > 1)
                    continue;
                azp_batch_tuning t = b->tune;
                t.block_size = blk;
                t.grab = gr;
                t.workgroups_per_cu = wg;
                float ms = 0.0f;
                st = time_it([&] { return enqueue_forces(b, t, s); }, 3, &ms);
                if (st == AZP_OK && ms < best_ms)
                    {
                    best_ms = ms;
                    best = t;
                    best.force_us = 1e3 * (double)ms;
                    }
                }
    // neighbor walk: workgroup size (gate forced on)
    float best_walk = 1e30f;
    const unsigned walks[] = {64, 128, 256};
    for (unsigned wb : walks)
        {
        if (st != AZP_OK)
            break;
        b->tune.walk_block = wb;
        float ms = 0.0f;
        st = time_it(
            [&] {
                enqueue_check(b, true, s);
                enqueue_rebuild(b, s);
                return AZP_OK;
            },
            2, &ms);
        if (st == AZP_OK && ms < best_walk)
            {
            best_walk = ms;
            best.walk_block = wb;
            }
        }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    AZP_TRY(st);
    b->tune = best;
    b->drop_graphs();
    // leave the outputs computed with the chosen configuration
    AZP_TRY(batch_compute(b, true, s));
    if (out)
        *out = best;
    return sc.leave();
    }

extern "C" int azp_batch_stats(azp_batch* b, azp_batch_stats_t* out)
    {
    if (!b || !out)
        return set_error(AZP_ERR_INVALID, "null argument");
    out->max_neighbors = b->max_neighbors;
    out->n_cells = b->ncell;
    out->graph_launches = b->graph_launches;
    out->timestep = 0;
    out->n_builds = 0;
    out->replica_builds = 0;
    if (b->d_timestep.ptr && b->own)
        {
        AZP_HIP(hipStreamSynchronize(b->own));
        unsigned long long v[3] = {0, 0, 0};
        AZP_HIP(hipMemcpy(&v[0], b->d_timestep.ptr, sizeof(unsigned long long), hipMemcpyDeviceToHost));
        AZP_HIP(hipMemcpy(&v[1], b->d_builds.ptr, 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost));
        out->timestep = v[0];
        out->n_builds = v[1];
        out->replica_builds = v[2];
        }
    return AZP_OK;
    }

extern "C" int azp_batch_nlist_download(azp_batch* b, uint32_t* h_n_neigh, uint32_t* h_nlist, uint64_t capacity,
                                        uint32_t* pitch, uint32_t* max_neighbors)
    {
    if (!b || !h_n_neigh || !h_nlist)
        return set_error(AZP_ERR_INVALID, "null argument");
    if (!b->have_list)
        return set_error(AZP_ERR_STATE, "neighbor list not built");
    const uint64_t need = (uint64_t)b->max_neighbors * b->pitch;
    if (pitch)
        *pitch = b->pitch;
    if (max_neighbors)
        *max_neighbors = b->max_neighbors;
    if (capacity < need)
        return set_error(AZP_ERR_INVALID, "host list needs %llu entries", (unsigned long long)need);
    AZP_HIP(hipDeviceSynchronize());
    AZP_HIP(hipMemcpy(h_n_neigh, b->n_neigh.ptr, (size_t)b->N * sizeof(uint32_t), hipMemcpyDeviceToHost));
    AZP_HIP(hipMemcpy(h_nlist, b->list.ptr, (size_t)need * sizeof(uint32_t), hipMemcpyDeviceToHost));
    return AZP_OK;
    }

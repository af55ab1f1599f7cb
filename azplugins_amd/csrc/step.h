// Per-particle pieces of a velocity Verlet / Langevin step and of the
// neighbor-list rebuild check, shared by the integrator kernels
// (k_integrate.hip), the rebuild check (k_nlist.hip) and the step that
// follows a pair force with the next step's rebuild check folded in
// (k_step): the FP64 operations and their order are fixed here, with
// contraction off, so every caller takes the same trajectory and the same
// rebuild decisions bit for bit.
#pragma once

#include "device_common.h"

namespace azp
{

// Wraps x into [-L/2, L/2) and counts the crossing in img.
__device__ __forceinline__ void wrap_1d(double& x, int& img, double L)
    {
    if (x >= 0.5 * L)
        {
        x -= L;
        ++img;
        }
    else if (x < -0.5 * L)
        {
        x += L;
        --img;
        }
    }

// The second half step of step `timestep` and the first half step of the
// next for one particle, from its net force (fx, fy, fz): Langevin adds the
// drag -gamma v and a uniform random force of variance 2 gamma kT / dt per
// component, keyed by (seed, timestep, tag).  v (x, y, z, mass), the position
// p and the image im are updated.
__device__ __forceinline__ void vv_two_one_particle(double4& p,
                                                    double4& v,
                                                    int4& im,
                                                    double fx,
                                                    double fy,
                                                    double fz,
                                                    uint32_t tag,
                                                    const DevBox& box,
                                                    double dt,
                                                    int langevin,
                                                    double gamma,
                                                    double kT,
                                                    uint16_t seed,
                                                    uint64_t timestep)
    {
#pragma clang fp contract(off)
    if (langevin)
        {
        uint32_t k0, k1;
        hoomd_seed_key(kRngLangevin, timestep, seed, k0, k1);
        u32x4 c;
        c.v[0] = tag;
        c.v[1] = 0u;
        c.v[2] = 0u;
        c.v[3] = 0u;
        const u32x4 r0 = philox4x32_10(c, k0, k1);
        c.v[3] = 1u;
        const u32x4 r1 = philox4x32_10(c, k0, k1);
        u32x4 a, b, d;
        a.v[0] = r0.v[0];
        a.v[1] = r0.v[1];
        b.v[0] = r0.v[2];
        b.v[1] = r0.v[3];
        d.v[0] = r1.v[0];
        d.v[1] = r1.v[1];
        const double coeff = sqrt(6.0 * gamma * kT / dt);
        fx += -gamma * v.x + coeff * uniform_m1_1(a);
        fy += -gamma * v.y + coeff * uniform_m1_1(b);
        fz += -gamma * v.z + coeff * uniform_m1_1(d);
        }
    const double minv = 1.0 / v.w;
    const double ax = fx * minv, ay = fy * minv, az = fz * minv;
    // second half of step t
    v.x += 0.5 * dt * ax;
    v.y += 0.5 * dt * ay;
    v.z += 0.5 * dt * az;
    // first half of step t + 1
    v.x += 0.5 * dt * ax;
    v.y += 0.5 * dt * ay;
    v.z += 0.5 * dt * az;
    p.x += dt * v.x;
    p.y += dt * v.y;
    p.z += dt * v.z;
    wrap_1d(p.x, im.x, box.L[0]);
    wrap_1d(p.y, im.y, box.L[1]);
    wrap_1d(p.z, im.z, box.L[2]);
    }

// The rebuild check of one particle against its position at the last build:
// 2 when the position is not finite, 1 when it moved farther than
// sqrt(thresh_sq) (minimum image), else 0.
__device__ __forceinline__ uint32_t displaced(const double4& p, const double4& q, const DevBox& box, double thresh_sq)
    {
    if (!(isfinite(p.x) && isfinite(p.y) && isfinite(p.z)))
        return 2u;
    const double dx = min_image_1d(p.x - q.x, box.L[0], box.Linv[0]);
    const double dy = min_image_1d(p.y - q.y, box.L[1], box.Linv[1]);
    const double dz = min_image_1d(p.z - q.z, box.L[2], box.Linv[2]);
    return rsq3_nofma(dx, dy, dz) > thresh_sq ? 1u : 0u;
    }

// A step run right after a pair force (k_step): positions, velocities and
// images updated in place, and the list's rebuild check of the new positions.
struct StepKernelArgs
    {
    double4* pos;
    double4* vel;
    int4* image;          // may be null
    const uint32_t* tag;  // Langevin keys (may be null for velocity Verlet)
    const double4* last;  // positions at the last list build
    uint32_t* flag;       // rebuild check of the new positions: (seq << 2) | bits as displaced(), by atomicMax
    uint32_t seq;         // this step's check number (never 0; the word needs no clearing between steps)
    double thresh_sq;
    double dt;
    int langevin;
    double gamma;
    double kT;
    uint16_t seed;
    uint64_t timestep;
    };

// The step of particle i (live: i < N) from its net force f, and its rebuild
// check (one atomic per wavefront, when any lane moved too far).  Every lane
// of the wavefront calls this.
__device__ __forceinline__ void step_particle(const StepKernelArgs& s,
                                              const DevBox& box,
                                              uint32_t i,
                                              bool live,
                                              const double4& f)
    {
    uint32_t bits = 0;
    if (live)
        {
        double4 p = s.pos[i];
        double4 v = s.vel[i];
        int4 im = s.image ? s.image[i] : make_int4(0, 0, 0, 0);
        vv_two_one_particle(p, v, im, f.x, f.y, f.z, s.langevin ? s.tag[i] : 0u, box, s.dt, s.langevin, s.gamma,
                            s.kT, s.seed, s.timestep);
        s.pos[i] = p;
        s.vel[i] = v;
        if (s.image)
            s.image[i] = im;
        bits = displaced(p, s.last[i], box, s.thresh_sq);
        }
    const uint64_t moved = __ballot(bits & 1u), bad = __ballot(bits & 2u);
    if ((moved | bad) != 0ull && (threadIdx.x & 63u) == (uint32_t)__builtin_ctzll(moved | bad))
        atomicMax(s.flag, (s.seq << 2) | (moved ? 1u : 0u) | (bad ? 2u : 0u));
    }

} // namespace azp

// Time integration and thermodynamic reductions around the pair-force path.
//
// Velocity Verlet in two half steps (upstream TwoStepConstantVolume without a
// thermostat), a Langevin second half step (drag -gamma v plus a uniform
// random force of variance 2 gamma kT / dt per component, upstream
// TwoStepLangevin's form), and rotation of spherical tops (isotropic moment
// of inertia; space-frame angular momentum; exact rotation by omega dt).
// Deterministic FP64 reductions: fixed grid, one partial per workgroup,
// partials summed in a fixed order by a second kernel.
#include "device_common.h"
#include "step.h"

namespace azp
{

__global__ void k_nve_one(double4* __restrict__ pos,
                          double4* __restrict__ vel,
                          const double4* __restrict__ accel,
                          int4* __restrict__ image,
                          uint32_t N,
                          DevBox box,
                          double dt)
    {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N)
        return;
    double4 p = pos[i];
    double4 v = vel[i];
    const double4 a = accel[i];
    v.x += 0.5 * dt * a.x;
    v.y += 0.5 * dt * a.y;
    v.z += 0.5 * dt * a.z;
    p.x += dt * v.x;
    p.y += dt * v.y;
    p.z += dt * v.z;
    int4 im = image ? image[i] : make_int4(0, 0, 0, 0);
    wrap_1d(p.x, im.x, box.L[0]);
    wrap_1d(p.y, im.y, box.L[1]);
    wrap_1d(p.z, im.z, box.L[2]);
    pos[i] = p;
    vel[i] = v;
    if (image)
        image[i] = im;
    }

// a = F/m (+ Langevin drag and noise); v += dt a / 2
__global__ void k_vv_two(double4* __restrict__ vel,
                         double4* __restrict__ accel,
                         const double4* __restrict__ force,
                         const uint32_t* __restrict__ tag,
                         uint32_t N,
                         double dt,
                         int langevin,
                         double gamma,
                         double kT,
                         uint16_t seed,
                         uint64_t timestep)
    {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N)
        return;
    double4 v = vel[i];
    const double4 f = force[i];
    double fx = f.x, fy = f.y, fz = f.z;
    if (langevin)
        {
        uint32_t k0, k1;
        hoomd_seed_key(kRngLangevin, timestep, seed, k0, k1);
        u32x4 c;
        c.v[0] = tag[i];
        c.v[1] = 0u;
        c.v[2] = 0u;
        c.v[3] = 0u;
        const u32x4 r0 = philox4x32_10(c, k0, k1);
        c.v[3] = 1u;
        const u32x4 r1 = philox4x32_10(c, k0, k1);
        u32x4 a, b, d;
        a.v[0] = r0.v[0]; a.v[1] = r0.v[1];
        b.v[0] = r0.v[2]; b.v[1] = r0.v[3];
        d.v[0] = r1.v[0]; d.v[1] = r1.v[1];
        const double coeff = sqrt(6.0 * gamma * kT / dt);
        fx += -gamma * v.x + coeff * uniform_m1_1(a);
        fy += -gamma * v.y + coeff * uniform_m1_1(b);
        fz += -gamma * v.z + coeff * uniform_m1_1(d);
        }
    const double minv = 1.0 / v.w;
    const double4 acc = make_double4(fx * minv, fy * minv, fz * minv, 0.0);
    v.x += 0.5 * dt * acc.x;
    v.y += 0.5 * dt * acc.y;
    v.z += 0.5 * dt * acc.z;
    vel[i] = v;
    accel[i] = acc;
    }

// The second half step of step t and the first half step of step t + 1 in
// one pass (nothing runs between them): the same FP64 operations in the same
// order as k_vv_two followed by k_nve_one, so the trajectory is bit-identical,
// without storing the intermediate velocity and the acceleration.
__global__ void k_vv_two_one(double4* __restrict__ pos,
                             double4* __restrict__ vel,
                             int4* __restrict__ image,
                             const double4* __restrict__ force,
                             const uint32_t* __restrict__ tag,
                             uint32_t N,
                             DevBox box,
                             double dt,
                             int langevin,
                             double gamma,
                             double kT,
                             uint16_t seed,
                             uint64_t timestep)
    {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N)
        return;
    double4 v = vel[i];
    const double4 f = force[i];
    double4 p = pos[i];
    int4 im = image ? image[i] : make_int4(0, 0, 0, 0);
    vv_two_one_particle(p, v, im, f.x, f.y, f.z, langevin ? tag[i] : 0u, box, dt, langevin, gamma, kT, seed,
                        timestep);
    pos[i] = p;
    vel[i] = v;
    if (image)
        image[i] = im;
    }

// The step of k_vv_two_one from forces already computed, with the list's
// rebuild check of the new positions (so the next update needs no
// displacement kernel).
__global__ void k_step(const double4* __restrict__ force, uint32_t N, DevBox box, StepKernelArgs s)
    {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = i < N;
    step_particle(s, box, i, live, live ? force[i] : make_double4(0.0, 0.0, 0.0, 0.0));
    }

int launch_step(const double4* force, uint32_t N, const DevBox& box, const StepKernelArgs& step, hipStream_t s)
    {
    if (N == 0)
        return AZP_OK;
    const unsigned B = 256;
    hipLaunchKernelGGL(k_step, dim3(grid_for(N, B)), dim3(B), 0, s, force, N, box, step);
    AZP_LAUNCH_CHECK();
    return AZP_OK;
    }

// Spherical tops: L += dt tau / 2, then q <- exp(dt omega / 2) q with omega = L / I.
__global__ void k_rot_one(double4* __restrict__ orientation,
                          double4* __restrict__ angmom,
                          const double4* __restrict__ torque,
                          const double* __restrict__ inertia,
                          uint32_t N,
                          double dt)
    {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N)
        return;
    double4 L = angmom[i];
    const double4 t = torque[i];
    L.x += 0.5 * dt * t.x;
    L.y += 0.5 * dt * t.y;
    L.z += 0.5 * dt * t.z;
    angmom[i] = L;
    const double I = inertia[i];
    if (!(I > 0.0))
        return;
    const double wx = L.x / I, wy = L.y / I, wz = L.z / I;
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

__global__ void k_rot_two(double4* __restrict__ angmom,
                          const double4* __restrict__ torque,
                          const double* __restrict__ inertia,
                          const uint32_t* __restrict__ tag,
                          uint32_t N,
                          double dt,
                          int langevin,
                          double gamma_r,
                          double kT,
                          uint16_t seed,
                          uint64_t timestep)
    {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N)
        return;
    double4 L = angmom[i];
    const double4 t = torque[i];
    double tx = t.x, ty = t.y, tz = t.z;
    if (langevin && inertia[i] > 0.0)
        {
        uint32_t k0, k1;
        hoomd_seed_key(kRngLangevin, timestep, seed, k0, k1);
        u32x4 c;
        c.v[0] = tag[i];
        c.v[1] = 1u;
        c.v[2] = 0u;
        c.v[3] = 0u;
        const u32x4 r0 = philox4x32_10(c, k0, k1);
        c.v[3] = 1u;
        const u32x4 r1 = philox4x32_10(c, k0, k1);
        u32x4 a, b, d;
        a.v[0] = r0.v[0]; a.v[1] = r0.v[1];
        b.v[0] = r0.v[2]; b.v[1] = r0.v[3];
        d.v[0] = r1.v[0]; d.v[1] = r1.v[1];
        // the rotational thermostat acts once per step (the first half step
        // uses the pair torque only), so its drag and noise take the full dt:
        // <L_a^2> = I kT at equilibrium
        const double I = inertia[i];
        const double coeff = sqrt(6.0 * gamma_r * kT / dt);
        tx += 2.0 * (-gamma_r * L.x / I + coeff * uniform_m1_1(a));
        ty += 2.0 * (-gamma_r * L.y / I + coeff * uniform_m1_1(b));
        tz += 2.0 * (-gamma_r * L.z / I + coeff * uniform_m1_1(d));
        }
    L.x += 0.5 * dt * tx;
    L.y += 0.5 * dt * ty;
    L.z += 0.5 * dt * tz;
    angmom[i] = L;
    }

constexpr int kReduceBlocks = 1024;
constexpr int kReduceThreads = 256;
constexpr int kThermoFields = 10;

// fields: 0 translational KE, 1 rotational KE, 2 potential energy,
//         3..8 virial xx xy xz yy yz zz, 9 count
__global__ void __launch_bounds__(kReduceThreads) k_thermo_partial(const double4* __restrict__ vel,
                                                                   const double4* __restrict__ force,
                                                                   const double* __restrict__ virial,
                                                                   uint32_t vpitch,
                                                                   const double4* __restrict__ angmom,
                                                                   const double* __restrict__ inertia,
                                                                   uint32_t N,
                                                                   double* __restrict__ partial)
    {
    double acc[kThermoFields];
#pragma unroll
    for (int f = 0; f < kThermoFields; ++f)
        acc[f] = 0.0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x)
        {
        const double4 v = vel[i];
        acc[0] += 0.5 * v.w * (v.x * v.x + v.y * v.y + v.z * v.z);
        if (angmom && inertia && inertia[i] > 0.0)
            {
            const double4 L = angmom[i];
            acc[1] += 0.5 * (L.x * L.x + L.y * L.y + L.z * L.z) / inertia[i];
            }
        if (force)
            acc[2] += force[i].w;
        if (virial)
            for (int c = 0; c < 6; ++c)
                acc[3 + c] += virial[(size_t)c * vpitch + i];
        acc[9] += 1.0;
        }
    __shared__ double s[kThermoFields][kReduceThreads];
#pragma unroll
    for (int f = 0; f < kThermoFields; ++f)
        s[f][threadIdx.x] = acc[f];
    __syncthreads();
    for (int off = kReduceThreads / 2; off > 0; off >>= 1)
        {
        if ((int)threadIdx.x < off)
            for (int f = 0; f < kThermoFields; ++f)
                s[f][threadIdx.x] += s[f][threadIdx.x + off];
        __syncthreads();
        }
    if (threadIdx.x < kThermoFields)
        partial[(size_t)threadIdx.x * kReduceBlocks + blockIdx.x] = s[threadIdx.x][0];
    }

// Kinetic tensor: fields 0..5 = sum of m v_a v_b for xx xy xz yy yz zz.
constexpr int kKineticFields = 6;
__global__ void __launch_bounds__(kReduceThreads) k_kinetic_partial(const double4* __restrict__ vel,
                                                                    uint32_t N,
                                                                    double* __restrict__ partial)
    {
    double acc[kKineticFields];
#pragma unroll
    for (int f = 0; f < kKineticFields; ++f)
        acc[f] = 0.0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x)
        {
        const double4 v = vel[i];
        acc[0] += v.w * v.x * v.x;
        acc[1] += v.w * v.x * v.y;
        acc[2] += v.w * v.x * v.z;
        acc[3] += v.w * v.y * v.y;
        acc[4] += v.w * v.y * v.z;
        acc[5] += v.w * v.z * v.z;
        }
    __shared__ double s[kKineticFields][kReduceThreads];
#pragma unroll
    for (int f = 0; f < kKineticFields; ++f)
        s[f][threadIdx.x] = acc[f];
    __syncthreads();
    for (int off = kReduceThreads / 2; off > 0; off >>= 1)
        {
        if ((int)threadIdx.x < off)
            for (int f = 0; f < kKineticFields; ++f)
                s[f][threadIdx.x] += s[f][threadIdx.x + off];
        __syncthreads();
        }
    if (threadIdx.x < kKineticFields)
        partial[(size_t)threadIdx.x * kReduceBlocks + blockIdx.x] = s[threadIdx.x][0];
    }

__global__ void __launch_bounds__(kReduceThreads) k_thermo_final(const double* __restrict__ partial,
                                                                 int nfields,
                                                                 double* __restrict__ out)
    {
    __shared__ double s[kReduceThreads];
    for (int f = 0; f < nfields; ++f)
        {
        double a = 0.0;
        for (int b = threadIdx.x; b < kReduceBlocks; b += kReduceThreads)
            a += partial[(size_t)f * kReduceBlocks + b];
        s[threadIdx.x] = a;
        __syncthreads();
        for (int off = kReduceThreads / 2; off > 0; off >>= 1)
            {
            if ((int)threadIdx.x < off)
                s[threadIdx.x] += s[threadIdx.x + off];
            __syncthreads();
            }
        if (threadIdx.x == 0)
            out[f] = s[0];
        __syncthreads();
        }
    }

} // namespace azp

using namespace azp;

extern "C" int azp_gpu_integrate_vv_one(double* d_pos,
                                        double* d_vel,
                                        const double* d_accel,
                                        int32_t* d_image,
                                        uint32_t N,
                                        const azp_box* box,
                                        double dt,
                                        void* stream)
    {
    if (N == 0)
        return AZP_OK;
    if (!d_pos || !d_vel || !d_accel || !box || !box_valid(*box))
        return set_error(AZP_ERR_INVALID, "bad integrator arguments");
    const unsigned B = 256;
    hipLaunchKernelGGL(k_nve_one, dim3(grid_for(N, B)), dim3(B), 0, (hipStream_t)stream,
                       reinterpret_cast<double4*>(d_pos), reinterpret_cast<double4*>(d_vel),
                       reinterpret_cast<const double4*>(d_accel), reinterpret_cast<int4*>(d_image), N,
                       make_box(*box), dt);
    AZP_LAUNCH_CHECK();
    return AZP_OK;
    }

extern "C" int azp_gpu_integrate_vv_two(double* d_vel,
                                        double* d_accel,
                                        const double* d_force,
                                        const uint32_t* d_tag,
                                        uint32_t N,
                                        double dt,
                                        int langevin,
                                        double gamma,
                                        double kT,
                                        uint16_t seed,
                                        uint64_t timestep,
                                        void* stream)
    {
    if (N == 0)
        return AZP_OK;
    if (!d_vel || !d_accel || !d_force || (langevin && !d_tag))
        return set_error(AZP_ERR_INVALID, "bad integrator arguments");
    const unsigned B = 256;
    hipLaunchKernelGGL(k_vv_two, dim3(grid_for(N, B)), dim3(B), 0, (hipStream_t)stream,
                       reinterpret_cast<double4*>(d_vel), reinterpret_cast<double4*>(d_accel),
                       reinterpret_cast<const double4*>(d_force), d_tag, N, dt, langevin, gamma, kT, seed,
                       timestep);
    AZP_LAUNCH_CHECK();
    return AZP_OK;
    }

extern "C" int azp_gpu_integrate_vv_two_one(double* d_pos,
                                            double* d_vel,
                                            int32_t* d_image,
                                            const double* d_force,
                                            const uint32_t* d_tag,
                                            uint32_t N,
                                            const azp_box* box,
                                            double dt,
                                            int langevin,
                                            double gamma,
                                            double kT,
                                            uint16_t seed,
                                            uint64_t timestep,
                                            void* stream)
    {
    if (N == 0)
        return AZP_OK;
    if (!d_pos || !d_vel || !d_force || !box || !box_valid(*box))
        return set_error(AZP_ERR_INVALID, "bad integration arguments");
    if (langevin && !d_tag)
        return set_error(AZP_ERR_INVALID, "the Langevin step needs tags");
    const unsigned B = 256;
    hipLaunchKernelGGL(k_vv_two_one, dim3(grid_for(N, B)), dim3(B), 0, (hipStream_t)stream,
                       reinterpret_cast<double4*>(d_pos), reinterpret_cast<double4*>(d_vel),
                       reinterpret_cast<int4*>(d_image), reinterpret_cast<const double4*>(d_force), d_tag, N,
                       make_box(*box), dt, langevin, gamma, kT, seed, timestep);
    AZP_LAUNCH_CHECK();
    return AZP_OK;
    }

extern "C" int azp_gpu_integrate_rot_one(double* d_orientation,
                                         double* d_angmom,
                                         const double* d_torque,
                                         const double* d_inertia,
                                         uint32_t N,
                                         double dt,
                                         void* stream)
    {
    if (N == 0)
        return AZP_OK;
    if (!d_orientation || !d_angmom || !d_torque || !d_inertia)
        return set_error(AZP_ERR_INVALID, "bad rotation arguments");
    const unsigned B = 256;
    hipLaunchKernelGGL(k_rot_one, dim3(grid_for(N, B)), dim3(B), 0, (hipStream_t)stream,
                       reinterpret_cast<double4*>(d_orientation), reinterpret_cast<double4*>(d_angmom),
                       reinterpret_cast<const double4*>(d_torque), d_inertia, N, dt);
    AZP_LAUNCH_CHECK();
    return AZP_OK;
    }

extern "C" int azp_gpu_integrate_rot_two(double* d_angmom,
                                         const double* d_torque,
                                         const double* d_inertia,
                                         const uint32_t* d_tag,
                                         uint32_t N,
                                         double dt,
                                         int langevin,
                                         double gamma_r,
                                         double kT,
                                         uint16_t seed,
                                         uint64_t timestep,
                                         void* stream)
    {
    if (N == 0)
        return AZP_OK;
    if (!d_angmom || !d_torque || !d_inertia || (langevin && !d_tag))
        return set_error(AZP_ERR_INVALID, "bad rotation arguments");
    const unsigned B = 256;
    hipLaunchKernelGGL(k_rot_two, dim3(grid_for(N, B)), dim3(B), 0, (hipStream_t)stream,
                       reinterpret_cast<double4*>(d_angmom), reinterpret_cast<const double4*>(d_torque),
                       d_inertia, d_tag, N, dt, langevin, gamma_r, kT, seed, timestep);
    AZP_LAUNCH_CHECK();
    return AZP_OK;
    }

extern "C" int azp_thermo_workspace_size(void)
    {
    return kThermoFields * kReduceBlocks;
    }

extern "C" int azp_gpu_thermo_reduce(const double* d_vel,
                                     const double* d_force,
                                     const double* d_virial,
                                     uint32_t virial_pitch,
                                     const double* d_angmom,
                                     const double* d_inertia,
                                     uint32_t N,
                                     double* d_work,
                                     double* d_out,
                                     void* stream)
    {
    if (!d_vel || !d_work || !d_out)
        return set_error(AZP_ERR_INVALID, "bad reduction arguments");
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_thermo_partial, dim3(kReduceBlocks), dim3(kReduceThreads), 0, s,
                       reinterpret_cast<const double4*>(d_vel), reinterpret_cast<const double4*>(d_force),
                       d_virial, virial_pitch, reinterpret_cast<const double4*>(d_angmom), d_inertia, N, d_work);
    hipLaunchKernelGGL(k_thermo_final, dim3(1), dim3(kReduceThreads), 0, s, d_work, kThermoFields, d_out);
    AZP_LAUNCH_CHECK();
    return AZP_OK;
    }

extern "C" int azp_gpu_kinetic_tensor(const double* d_vel, uint32_t N, double* d_work, double* d_out, void* stream)
    {
    if (!d_vel || !d_work || !d_out)
        return set_error(AZP_ERR_INVALID, "bad reduction arguments");
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_kinetic_partial, dim3(kReduceBlocks), dim3(kReduceThreads), 0, s,
                       reinterpret_cast<const double4*>(d_vel), N, d_work);
    hipLaunchKernelGGL(k_thermo_final, dim3(1), dim3(kReduceThreads), 0, s, d_work, kKineticFields, d_out);
    AZP_LAUNCH_CHECK();
    return AZP_OK;
    }

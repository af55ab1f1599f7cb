// Wall potentials: the LJ 9-3 and colloid wall evaluators (reference
// WallEvaluatorLJ93.h:81-143, WallEvaluatorColloid.h:91-201) acting on every
// particle from HOOMD's wall geometries (spheres, cylinders, planes), the
// external-force path the reference drives through PotentialExternalWallGPU
// (PotentialExternalWallGPUKernel.cu.inc:26-27, export_PotentialExternalWall.cc.inc).
//
// One thread per particle, FP64, the walls read from a small device table.
// For each wall the particle's signed distance s (positive on the wall's
// inside) and the unit normal n pointing into the inside give:
//   inside and s > r_extrap:  the evaluator at r^2 = s^2, F = f/r s n
//   s <= r_extrap (r_extrap > 0, inside near the wall or outside):
//       linear extrapolation from r_extrap, F = f(r_extrap) n and
//       E = E(r_extrap) + f(r_extrap) (r_extrap - s)
// with the energy shifted to zero at r_cut, the whole energy on the particle
// and the virial 1/2 (s n) x F.
#include <atomic>
#include <cmath>
#include <vector>

#include "cells.h"
#include "device_common.h"

namespace azp
{

constexpr int kWallLJ93 = 0;
constexpr int kWallColloid = 1;

// Per-type parameters as the reference stores them: LJ93 {sigma_3, A},
// colloid {c_1, c_2, a}; then r_cut^2 and r_extrap.
constexpr int kWallParamStride = 5;
// Geometry rows: sphere {origin, radius, inside, open}, cylinder {origin,
// unit axis, radius, inside, open}, plane {origin, unit normal, open}.
constexpr int kSphereRow = 6;
constexpr int kCylinderRow = 9;
constexpr int kPlaneRow = 7;

// f/r and the energy shifted to zero at r_cut; false outside the cutoff.
template<int KIND> __device__ __forceinline__ bool wall_eval(const double* prm, double rsq, double& fdivr, double& e)
    {
#pragma clang fp contract(off)
    const double rcutsq = prm[3];
    if constexpr (KIND == kWallLJ93)
        {
        const double s3 = prm[0], A = prm[1];
        const double lj1 = (2.0 / 15.0) * A * s3 * s3 * s3;
        const double lj2 = A * s3;
        if (!(rsq < rcutsq && lj1 != 0.0))
            return false;
        const double r2inv = 1.0 / rsq;
        const double r3inv = r2inv * sqrt(r2inv);
        const double r6inv = r3inv * r3inv;
        fdivr = r2inv * r3inv * (9.0 * lj1 * r6inv - 3.0 * lj2);
        e = r3inv * (lj1 * r6inv - lj2);
        const double rc2inv = 1.0 / rcutsq;
        const double rc3inv = rc2inv * sqrt(rc2inv);
        const double rc6inv = rc3inv * rc3inv;
        e -= rc3inv * (lj1 * rc6inv - lj2);
        return true;
        }
    else
        {
        const double c1 = prm[0], c2 = prm[1], a = prm[2];
        if (!(rsq < rcutsq && c1 != 0.0 && a > 0.0))
            return false;
        auto potential = [&](double r2, bool want_force, double& fd) {
            const double r = sqrt(r2);
            const double arinv = a / r;
            const double rmi = 1.0 / (r - a);
            const double rpi = 1.0 / (r + a);
            const double r2mi = rmi * rpi;
            const double rmi2 = rmi * rmi;
            const double rmi6 = rmi2 * rmi2 * rmi2;
            const double rpi2 = rpi * rpi;
            const double rpi6 = rpi2 * rpi2 * rpi2;
            if (want_force)
                {
                const double arinv8 = 8.0 * arinv;
                fd = 6.0 * c1 * ((arinv8 - 1.0) * rmi2 * rmi6 + (arinv8 + 1.0) * rpi2 * rpi6);
                fd -= c2 * (4.0 * a * a * arinv * r2mi * r2mi);
                }
            const double a7 = 7.0 * a;
            double en = c1 * ((a7 - r) * rmi * rmi6 + (a7 + r) * rpi * rpi6);
            en -= c2 * (2.0 * a * r * r2mi + log(rpi / rmi));
            return en;
        };
        double unused = 0.0;
        e = potential(rsq, true, fdivr);
        e -= potential(rcutsq, false, unused);
        return true;
        }
    }

__global__ void k_wall_clear_flag(uint32_t* flag)
    {
    *flag = 0u;
    }

struct WallAcc
    {
    double fx = 0.0, fy = 0.0, fz = 0.0, e = 0.0;
    double v[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    };

template<int KIND>
__device__ __forceinline__ void wall_apply(const double* prm, double s, double nx, double ny, double nz, bool open,
                                           WallAcc& acc)
    {
#pragma clang fp contract(off)
    const double rextrap = prm[4];
    double f = 0.0, e = 0.0; // force along n, energy
    bool hit = false;
    if (rextrap > 0.0 && s <= rextrap)
        {
        double fdivr, ee;
        if (wall_eval<KIND>(prm, rextrap * rextrap, fdivr, ee))
            {
            f = fdivr * rextrap;
            e = ee + f * (rextrap - s);
            hit = true;
            }
        }
    else if (s > 0.0 || (s == 0.0 && !open))
        {
        double fdivr, ee;
        if (wall_eval<KIND>(prm, s * s, fdivr, ee))
            {
            f = fdivr * s;
            e = ee;
            hit = true;
            }
        }
    if (!hit)
        return;
    const double Fx = f * nx, Fy = f * ny, Fz = f * nz;
    acc.fx += Fx;
    acc.fy += Fy;
    acc.fz += Fz;
    acc.e += e;
    const double dx = s * nx, dy = s * ny, dz = s * nz;
    acc.v[0] += 0.5 * dx * Fx;
    acc.v[1] += 0.5 * dx * Fy;
    acc.v[2] += 0.5 * dx * Fz;
    acc.v[3] += 0.5 * dy * Fy;
    acc.v[4] += 0.5 * dy * Fz;
    acc.v[5] += 0.5 * dz * Fz;
    }

template<int KIND>
__global__ void __launch_bounds__(256) k_wall(const double4* __restrict__ pos,
                                              uint32_t N,
                                              const double* __restrict__ params,
                                              uint32_t ntypes,
                                              const double* __restrict__ geom,
                                              uint32_t ns,
                                              uint32_t nc,
                                              uint32_t np,
                                              double4* __restrict__ force,
                                              double* __restrict__ virial,
                                              uint32_t vpitch,
                                              uint32_t* __restrict__ flags)
    {
#pragma clang fp contract(off)
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N)
        return;
    const double4 p = pos[i];
    WallAcc acc;
    const uint32_t t = checked_type(p.w, ntypes);
    if (t < ntypes)
        {
        const double* prm = params + (size_t)t * kWallParamStride;
        const double* g = geom;
        for (uint32_t k = 0; k < ns; ++k, g += kSphereRow)
            {
            const double dx = p.x - g[0], dy = p.y - g[1], dz = p.z - g[2];
            const double d = sqrt(dx * dx + dy * dy + dz * dz);
            const bool inside = g[4] != 0.0;
            const double s = inside ? g[3] - d : d - g[3];
            const double sg = d > 0.0 ? (inside ? -1.0 : 1.0) / d : 0.0;
            wall_apply<KIND>(prm, s, sg * dx, sg * dy, sg * dz, g[5] != 0.0, acc);
            }
        for (uint32_t k = 0; k < nc; ++k, g += kCylinderRow)
            {
            const double dx = p.x - g[0], dy = p.y - g[1], dz = p.z - g[2];
            const double proj = dx * g[3] + dy * g[4] + dz * g[5];
            const double qx = dx - proj * g[3], qy = dy - proj * g[4], qz = dz - proj * g[5];
            const double d = sqrt(qx * qx + qy * qy + qz * qz);
            const bool inside = g[7] != 0.0;
            const double s = inside ? g[6] - d : d - g[6];
            const double sg = d > 0.0 ? (inside ? -1.0 : 1.0) / d : 0.0;
            wall_apply<KIND>(prm, s, sg * qx, sg * qy, sg * qz, g[8] != 0.0, acc);
            }
        for (uint32_t k = 0; k < np; ++k, g += kPlaneRow)
            {
            const double s = (p.x - g[0]) * g[3] + (p.y - g[1]) * g[4] + (p.z - g[2]) * g[5];
            wall_apply<KIND>(prm, s, g[3], g[4], g[5], g[6] != 0.0, acc);
            }
        }
    else
        atomicOr(flags, 1u);
    force[i] = make_double4(acc.fx, acc.fy, acc.fz, acc.e);
    if (virial)
        for (int c = 0; c < 6; ++c)
            virial[(size_t)c * vpitch + i] = acc.v[c];
    }

} // namespace azp

using namespace azp;

// Value the type-error flag never takes (it holds bit 1 only).
constexpr uint32_t kWallFlagPending = 0xa5a5a5a5u;

struct azp_wall
    {
    int kind = kWallLJ93;
    uint32_t ntypes = 0;
    std::vector<double> params; // [ntypes][kWallParamStride]
    std::vector<double> r_cut;  // as given, for readback
    std::vector<char> set;
    std::vector<double> geom;
    uint32_t ns = 0, nc = 0, np = 0;
    bool dirty = true;
    DeviceArray<double> d_params, d_geom;
    DeviceArray<uint32_t> d_flags;
    uint32_t* h_flags = nullptr;
    bool check_pending = false; // a type-error flag copy is queued and not yet read
    bool flag_ready = false;    // the device flag was cleared once after allocation

    ~azp_wall()
        {
        if (h_flags)
            (void)hipHostFree(h_flags);
        }
    };

// Number of schema values of a wall class: LJ93 (A, sigma, r_cut, r_extrap),
// colloid (A, a, sigma, r_cut, r_extrap) — the keys of the reference's
// wall.py TypeParameterDicts, in that order.
static uint32_t wall_nvalues(int kind)
    {
    return kind == kWallLJ93 ? 4u : 5u;
    }

extern "C" int azp_wall_create(azp_wall** out, const char* cpp_class_name, uint32_t ntypes)
    {
    if (!out || !cpp_class_name || ntypes == 0)
        return set_error(AZP_ERR_INVALID, "bad wall arguments");
    std::string name(cpp_class_name);
    if (name.size() > 3 && name.compare(name.size() - 3, 3, "GPU") == 0)
        name.resize(name.size() - 3);
    int kind;
    if (name == "WallsPotentialLJ93")
        kind = kWallLJ93;
    else if (name == "WallsPotentialColloid")
        kind = kWallColloid;
    else
        return set_error(AZP_ERR_INVALID, "unknown wall class %s", cpp_class_name);
    azp_wall* w = new azp_wall();
    w->kind = kind;
    w->ntypes = ntypes;
    w->params.assign((size_t)ntypes * kWallParamStride, 0.0);
    w->r_cut.assign(ntypes, 0.0);
    w->set.assign(ntypes, 0);
    *out = w;
    return AZP_OK;
    }

extern "C" int azp_wall_destroy(azp_wall* w)
    {
    delete w;
    return AZP_OK;
    }

extern "C" int azp_wall_set_params(azp_wall* w, uint32_t type, const double* v)
    {
    if (!w || !v || type >= w->ntypes)
        return set_error(AZP_ERR_INVALID, "bad wall parameter arguments");
    double* p = w->params.data() + (size_t)type * kWallParamStride;
    double r_cut, r_extrap;
    if (w->kind == kWallLJ93)
        {
        // WallEvaluatorLJ93.h:33-38: sigma_3 = sigma^3, A
        const double A = v[0], sigma = v[1];
        p[0] = sigma * sigma * sigma;
        p[1] = A;
        p[2] = 0.0;
        r_cut = v[2];
        r_extrap = v[3];
        }
    else
        {
        // WallEvaluatorColloid.h:34-42: c_1 = A sigma^6 / 7560, c_2 = A / 6, a
        const double A = v[0], a = v[1], sigma = v[2];
        const double sigma_3 = sigma * sigma * sigma;
        p[0] = A * sigma_3 * sigma_3 / 7560.0;
        p[1] = A / 6.0;
        p[2] = a;
        r_cut = v[3];
        r_extrap = v[4];
        }
    if (!(r_cut >= 0.0) || !(r_extrap >= 0.0))
        return set_error(AZP_ERR_INVALID, "r_cut and r_extrap must be >= 0");
    p[3] = r_cut * r_cut;
    p[4] = r_extrap;
    w->r_cut[type] = r_cut;
    w->set[type] = 1;
    w->dirty = true;
    return AZP_OK;
    }

extern "C" int azp_wall_get_params(const azp_wall* w, uint32_t type, double* v)
    {
    if (!w || !v || type >= w->ntypes)
        return set_error(AZP_ERR_INVALID, "bad wall parameter arguments");
    const double* p = w->params.data() + (size_t)type * kWallParamStride;
    if (w->kind == kWallLJ93)
        {
        // asDict: sigma = cbrt(sigma_3) (WallEvaluatorLJ93.h:40-46)
        v[0] = p[1];
        v[1] = std::cbrt(p[0]);
        v[2] = w->r_cut[type];
        v[3] = p[4];
        }
    else
        {
        // asDict: A = 6 c_2, sigma = (7560 c_1 / A)^(1/6) (WallEvaluatorColloid.h:44-54)
        const double A = 6.0 * p[1];
        v[0] = A;
        v[1] = p[2];
        v[2] = std::pow(7560.0 * p[0] / A, 1.0 / 6.0);
        v[3] = w->r_cut[type];
        v[4] = p[4];
        }
    return AZP_OK;
    }

extern "C" int azp_wall_param_nvalues(const azp_wall* w)
    {
    return w ? (int)wall_nvalues(w->kind) : -1;
    }

extern "C" int azp_wall_set_geometry(azp_wall* w,
                                     uint32_t n_spheres,
                                     const double* spheres,
                                     uint32_t n_cylinders,
                                     const double* cylinders,
                                     uint32_t n_planes,
                                     const double* planes)
    {
    if (!w || (n_spheres && !spheres) || (n_cylinders && !cylinders) || (n_planes && !planes))
        return set_error(AZP_ERR_INVALID, "bad wall geometry arguments");
    std::vector<double> g;
    g.reserve((size_t)n_spheres * kSphereRow + (size_t)n_cylinders * kCylinderRow + (size_t)n_planes * kPlaneRow);
    for (uint32_t k = 0; k < n_spheres; ++k)
        {
        const double* s = spheres + (size_t)k * kSphereRow;
        if (!(s[3] > 0.0))
            return set_error(AZP_ERR_INVALID, "sphere %u: radius must be positive", k);
        g.insert(g.end(), s, s + kSphereRow);
        }
    auto unit = [](const double* in, double* out) {
        const double n = std::sqrt(in[0] * in[0] + in[1] * in[1] + in[2] * in[2]);
        if (!(n > 0.0))
            return false;
        for (int c = 0; c < 3; ++c)
            out[c] = in[c] / n;
        return true;
    };
    for (uint32_t k = 0; k < n_cylinders; ++k)
        {
        const double* c = cylinders + (size_t)k * kCylinderRow;
        double row[kCylinderRow];
        for (int q = 0; q < kCylinderRow; ++q)
            row[q] = c[q];
        if (!unit(c + 3, row + 3) || !(c[6] > 0.0))
            return set_error(AZP_ERR_INVALID, "cylinder %u: needs a nonzero axis and a positive radius", k);
        g.insert(g.end(), row, row + kCylinderRow);
        }
    for (uint32_t k = 0; k < n_planes; ++k)
        {
        const double* pl = planes + (size_t)k * kPlaneRow;
        double row[kPlaneRow];
        for (int q = 0; q < kPlaneRow; ++q)
            row[q] = pl[q];
        if (!unit(pl + 3, row + 3))
            return set_error(AZP_ERR_INVALID, "plane %u: needs a nonzero normal", k);
        g.insert(g.end(), row, row + kPlaneRow);
        }
    w->geom.swap(g);
    w->ns = n_spheres;
    w->nc = n_cylinders;
    w->np = n_planes;
    w->dirty = true;
    return AZP_OK;
    }

// Reports the type error the sticky device flag holds and clears the flag, so
// that calls made after the caller fixed the types run normally.
static int wall_report_type_error(azp_wall* w, hipStream_t s)
    {
    hipLaunchKernelGGL(k_wall_clear_flag, dim3(1), dim3(1), 0, s, w->d_flags.ptr);
    AZP_LAUNCH_CHECK();
    return set_error(AZP_ERR_INVALID, "a particle type index is outside [0, %u) (pos.w must hold the type)",
                     w->ntypes);
    }

extern "C" int azp_wall_check(azp_wall* w, void* stream)
    {
    if (!w)
        return set_error(AZP_ERR_INVALID, "null argument");
    if (!w->flag_ready)
        return AZP_OK; // nothing computed yet
    hipStream_t s = (hipStream_t)stream;
    *reinterpret_cast<volatile uint32_t*>(w->h_flags) = kWallFlagPending;
    std::atomic_thread_fence(std::memory_order_seq_cst);
    AZP_HIP(hipMemcpyAsync(w->h_flags, w->d_flags.ptr, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    AZP_HIP(hipStreamSynchronize(s));
    w->check_pending = false;
    if (*reinterpret_cast<volatile uint32_t*>(w->h_flags) & 1u)
        return wall_report_type_error(w, s);
    return AZP_OK;
    }

extern "C" int azp_wall_compute(azp_wall* w, const azp_particles* p, const azp_outputs* out, int compute_virial,
                                void* stream)
    {
    if (!w || !p || !out)
        return set_error(AZP_ERR_INVALID, "null argument");
    for (uint32_t t = 0; t < w->ntypes; ++t)
        if (!w->set[t])
            return set_error(AZP_ERR_STATE, "wall parameters of type %u are not set", t);
    const uint32_t N = p->N;
    if (N == 0)
        return AZP_OK;
    if (!p->d_pos || !out->d_force || (compute_virial && !out->d_virial))
        return set_error(AZP_ERR_INVALID, "null particle or output arrays");
    if (compute_virial && out->virial_pitch < N)
        return set_error(AZP_ERR_INVALID, "virial pitch %u is smaller than N %u", out->virial_pitch, N);
    hipStream_t s = (hipStream_t)stream;
    // The type-error flag, read without waiting: its copy has landed once the
    // stream passed it (a sentinel until then).  The device flag is sticky
    // (kernels only OR into it), so a copy queued by any earlier call sees
    // every error up to its place in the stream.
    if (w->check_pending)
        {
        const uint32_t v = *reinterpret_cast<volatile uint32_t*>(w->h_flags);
        if (v != kWallFlagPending)
            {
            w->check_pending = false;
            if (v & 1u)
                return wall_report_type_error(w, s);
            }
        }
    if (!w->h_flags && hipHostMalloc((void**)&w->h_flags, sizeof(uint32_t), hipHostMallocCoherent) != hipSuccess)
        {
        w->h_flags = nullptr;
        return set_error(AZP_ERR_NOMEM, "pinned host allocation failed");
        }
    if (w->dirty)
        {
        AZP_TRY(w->d_params.reserve(w->params.size()));
        AZP_TRY(w->d_geom.reserve(std::max<size_t>(w->geom.size(), 1)));
        AZP_TRY(w->d_flags.reserve(1));
        AZP_HIP(hipMemcpyAsync(w->d_params.ptr, w->params.data(), w->params.size() * sizeof(double),
                               hipMemcpyHostToDevice, s));
        if (!w->geom.empty())
            AZP_HIP(hipMemcpyAsync(w->d_geom.ptr, w->geom.data(), w->geom.size() * sizeof(double),
                                   hipMemcpyHostToDevice, s));
        AZP_HIP(hipStreamSynchronize(s)); // the host vectors may change after return
        w->dirty = false;
        }
    if (!w->flag_ready)
        {
        hipLaunchKernelGGL(k_wall_clear_flag, dim3(1), dim3(1), 0, s, w->d_flags.ptr);
        w->flag_ready = true;
        }
    const unsigned B = 256;
    auto launch = [&](auto kernel) {
        hipLaunchKernelGGL(kernel, dim3(grid_for(N, B)), dim3(B), 0, s, reinterpret_cast<const double4*>(p->d_pos), N,
                           w->d_params.ptr, w->ntypes, w->d_geom.ptr, w->ns, w->nc, w->np,
                           reinterpret_cast<double4*>(out->d_force), compute_virial ? out->d_virial : nullptr,
                           out->virial_pitch, w->d_flags.ptr);
    };
    if (w->kind == kWallLJ93)
        launch(k_wall<kWallLJ93>);
    else
        launch(k_wall<kWallColloid>);
    AZP_LAUNCH_CHECK();
    if (!w->check_pending)
        {
        *reinterpret_cast<volatile uint32_t*>(w->h_flags) = kWallFlagPending;
        std::atomic_thread_fence(std::memory_order_seq_cst);
        AZP_HIP(hipMemcpyAsync(w->h_flags, w->d_flags.ptr, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        w->check_pending = true;
        }
    return AZP_OK;
    }

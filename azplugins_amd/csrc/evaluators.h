// Device evaluators for the azplugins pair functors.
//
// Each evaluator is split into
//   prepare(param, rcutsq, energy_shift) -> Prep<R>   (once per type pair)
//   active(Prep)                                      (the "potential is off" guard)
//   compute(Prep, rsq, force_divr, pair_eng)          (per pair, rsq < rcutsq known)
// so the kernels can hoist everything that depends only on the type pair
// (cutoff, energy shift, products of parameters) into LDS.  For R = double the
// per-pair arithmetic follows the reference expressions operation by operation;
// for R = float the same expressions run on the FP32 hardware units.
#pragma once

#include <type_traits>

#include "device_common.h"

namespace azp
{

// ---------------------------------------------------------------------------
// Hertz: PairEvaluatorHertz.h:93-110.  No shift branch (U(rc) = 0).
// ---------------------------------------------------------------------------
struct EvalHertz
    {
    using param_type = azp_param_hertz;
    template<typename R> struct Prep
        {
        R epsilon;
        R rcut;
        R rcutinv;
        };
    template<typename R>
    __device__ static Prep<R> prepare(const param_type& p, double rcutsq, bool /*shift*/)
        {
        Prep<R> q;
        const double rcut = sqrt(rcutsq);
        q.epsilon = (R)p.epsilon;
        q.rcut = (R)rcut;
        q.rcutinv = (R)(1.0 / rcut);
        return q;
        }
    template<typename R> __device__ static bool active(const Prep<R>& q)
        {
        return q.epsilon != R(0);
        }
    __device__ static void compute(const Prep<double>& q, double rsq, double& fdivr, double& e)
        {
        const double r = sqrt(rsq);
        const double x = 1.0 - (r / q.rcut);
        const double xsqrt = sqrt(x);
        const double ex3p2 = q.epsilon * x * xsqrt;
        fdivr = 2.5 * ex3p2 / (r * q.rcut);
        e = ex3p2 * x;
        }
    __device__ static void compute(const Prep<float>& q, float rsq, float& fdivr, float& e)
        {
        const float rinv = f_rsqrt(rsq);
        const float r = rsq * rinv;
        // clamp: a pair admitted by the exact FP64 cutoff test may sit a rounding
        // error beyond rc in FP32
        const float x = f_max(1.0f - r * q.rcutinv, 0.0f);
        const float xsqrt = f_sqrt(x);
        const float ex3p2 = q.epsilon * x * xsqrt;
        fdivr = 2.5f * ex3p2 * rinv * q.rcutinv;
        e = ex3p2 * x;
        }
    };

// ---------------------------------------------------------------------------
// ExpandedYukawa: PairEvaluatorExpandedYukawa.h:92-115.
// ---------------------------------------------------------------------------
struct EvalExpandedYukawa
    {
    using param_type = azp_param_expanded_yukawa;
    // r - delta cancels when r ~ delta; the FP32 form needs r^2 in FP64
    static constexpr bool kNeedsRsq64 = true;
    template<typename R> struct Prep
        {
        double delta_sq;
        R epsilon;
        R kappa;
        R delta;
        R eshift;
        };
    template<typename R>
    __device__ static Prep<R> prepare(const param_type& p, double rcutsq, bool shift)
        {
        Prep<R> q;
        q.delta_sq = p.delta * p.delta;
        q.epsilon = (R)p.epsilon;
        q.kappa = (R)p.kappa;
        q.delta = (R)p.delta;
        double es = 0.0;
        if (shift)
            {
            const double rcut = sqrt(rcutsq);
            const double rcut_delta = rcut - p.delta;
            es = p.epsilon * exp(-p.kappa * rcut_delta) / rcut_delta;
            }
        q.eshift = (R)es;
        return q;
        }
    template<typename R> __device__ static bool active(const Prep<R>& q)
        {
        return q.epsilon != R(0);
        }
    template<typename R> __device__ static void compute(const Prep<R>& q, R rsq, R& fdivr, R& e)
        {
        const R r = f_sqrt(rsq);
        const R r_delta = r - q.delta;
        const R r_delta_inv = f_rcp(r_delta);
        e = q.epsilon * f_exp(-q.kappa * r_delta) * r_delta_inv;
        fdivr = f_div(e * (q.kappa + r_delta_inv), r);
        e -= q.eshift;
        }
    // FP32 form with r - delta = (r^2 - delta^2) / (r + delta), r^2 - delta^2 in FP64
    __device__ static void compute_precise(const Prep<float>& q, float rsq, double rsq64, float& fdivr, float& e)
        {
        const float r = f_sqrt(rsq);
        const float r_delta = (float)(rsq64 - q.delta_sq) * f_rcp(r + q.delta);
        const float r_delta_inv = f_rcp(r_delta);
        e = q.epsilon * f_exp(-q.kappa * r_delta) * r_delta_inv;
        fdivr = e * (q.kappa + r_delta_inv) * f_rcp(r);
        e -= q.eshift;
        }
    };

// Evaluators that do not need FP64 r^2 in their FP32 form.
template<class Eval, typename = void> struct NeedsRsq64
    {
    static constexpr bool value = false;
    };
template<class Eval> struct NeedsRsq64<Eval, std::void_t<decltype(Eval::kNeedsRsq64)>>
    {
    static constexpr bool value = Eval::kNeedsRsq64;
    };

// ---------------------------------------------------------------------------
// PerturbedLennardJones: PairEvaluatorPerturbedLennardJones.h:99-103, 117-155.
// ---------------------------------------------------------------------------
struct EvalPerturbedLJ
    {
    using param_type = azp_param_perturbed_lj;
    template<typename R> struct Prep
        {
        R lj1;      // 4 eps sigma^12
        R lj1x12;   // 12 lj1
        R lj2;      // 4 eps sigma^6
        R lj2x6;    // 6 lj2
        R asf;      // attraction_scale_factor
        R rwcasq;   // 2^(1/3) sigma^2
        R wca_shift;
        R eshift;
        };
    template<typename R>
    __device__ static Prep<R> prepare(const param_type& p, double rcutsq, bool shift)
        {
        const double lj1 = p.epsilon_x_4 * p.sigma_6 * p.sigma_6;
        const double lj2 = p.epsilon_x_4 * p.sigma_6;
        const double asf = p.attraction_scale_factor;
        const double wca_shift = p.epsilon_x_4 * (1.0 - asf) / 4.0;
        double es = 0.0;
        if (shift)
            {
            const double rcut2inv = 1.0 / rcutsq;
            const double rcut6inv = rcut2inv * rcut2inv * rcut2inv;
            es = rcut6inv * (lj1 * rcut6inv - lj2);
            if (rcutsq < p.rwcasq)
                es += wca_shift;
            else
                es *= asf;
            }
        Prep<R> q;
        q.lj1 = (R)lj1;
        q.lj1x12 = (R)(12.0 * lj1);
        q.lj2 = (R)lj2;
        q.lj2x6 = (R)(6.0 * lj2);
        q.asf = (R)asf;
        q.rwcasq = (R)p.rwcasq;
        q.wca_shift = (R)wca_shift;
        q.eshift = (R)es;
        return q;
        }
    template<typename R> __device__ static bool active(const Prep<R>& q)
        {
        return q.lj1 != R(0);
        }
    template<typename R> __device__ static void compute(const Prep<R>& q, R rsq, R& fdivr, R& e)
        {
        const R r2inv = f_rcp(rsq);
        const R r6inv = r2inv * r2inv * r2inv;
        fdivr = r2inv * r6inv * (q.lj1x12 * r6inv - q.lj2x6);
        e = r6inv * (q.lj1 * r6inv - q.lj2);
        if (rsq < q.rwcasq)
            {
            e += q.wca_shift;
            }
        else
            {
            fdivr *= q.asf;
            e *= q.asf;
            }
        e -= q.eshift;
        }
    // The same with the cutoff mask folded in (FP32 form): both branches become
    // f *= s, e = e s + c with (s, c) = (1, wca_shift - eshift) inside the WCA
    // radius and (asf, -eshift) outside.  A masked entry arrives with
    // rsq = +inf: 1/rsq is then 0, so f and the LJ part of e are 0, and c is
    // masked to 0.
    static constexpr bool kMasked = true;
    __device__ static void compute_masked(const Prep<float>& q, float rsq, bool in, float& fdivr, float& e)
        {
        const float r2inv = f_rcp(rsq);
        const float r6inv = r2inv * r2inv * r2inv;
        const bool wca = rsq < q.rwcasq;
        const float s = wca ? 1.0f : q.asf;
        const float c = in ? (wca ? q.wca_shift - q.eshift : -q.eshift) : 0.0f;
        const float t = s * r6inv;
        fdivr = (t * r2inv) * fmaf(q.lj1x12, r6inv, -q.lj2x6);
        e = fmaf(t, fmaf(q.lj1, r6inv, -q.lj2), c);
        }
    };

// Evaluators with a fused, masked FP32 form.
template<class Eval, typename = void> struct HasMasked
    {
    static constexpr bool value = false;
    };
template<class Eval> struct HasMasked<Eval, std::void_t<decltype(Eval::kMasked)>>
    {
    static constexpr bool value = Eval::kMasked;
    };

// ---------------------------------------------------------------------------
// DPD conservative part only (PotentialPairConservative<DPDGeneralWeight>):
// DPDPairEvaluatorGeneralWeight.h:165-183.  No A != 0 guard in the reference.
// ---------------------------------------------------------------------------
struct EvalDPDConservative
    {
    using param_type = azp_param_dpd_general_weight;
    template<typename R> struct Prep
        {
        R A;
        R rcutsq;
        R rcutinv;
        R rcut;
        };
    template<typename R>
    __device__ static Prep<R> prepare(const param_type& p, double rcutsq, bool /*shift*/)
        {
        Prep<R> q;
        const double rcutinv = 1.0 / sqrt(rcutsq);
        q.A = (R)p.A;
        q.rcutsq = (R)rcutsq;
        q.rcutinv = (R)rcutinv;
        q.rcut = (R)(1.0 / rcutinv);
        return q;
        }
    template<typename R> __device__ static bool active(const Prep<R>&)
        {
        return true;
        }
    template<typename R> __device__ static void compute(const Prep<R>& q, R rsq, R& fdivr, R& e)
        {
        const R rinv = f_rsqrt(rsq);
        const R r = f_rcp(rinv);
        fdivr = q.A * (rinv - q.rcutinv);
        e = q.A * (q.rcut - r) - R(0.5) * q.A * q.rcutinv * (q.rcutsq - rsq);
        }
    };

// ---------------------------------------------------------------------------
// Colloid: PairEvaluatorColloid.h:101-269.  FP64 only.
// ---------------------------------------------------------------------------
struct EvalColloid
    {
    using param_type = azp_param_colloid;
    template<typename R> struct Prep
        {
        R A;
        R ai;
        R aj;
        R sigma_3;
        R sigma_6;
        R eshift;
        int branch; // 0 solvent-solvent, 1 colloid-solvent, 2 colloid-colloid
        };

    __device__ static double solvent_solvent(const Prep<double>& q, bool force, double& fdivr, double rsq)
        {
        const double r2inv = 1.0 / rsq;
        const double r6inv = r2inv * r2inv * r2inv;
        const double c1 = q.A * q.sigma_6 / 36.0;
        if (force)
            fdivr = 6.0 * c1 * r2inv * r6inv * (2.0 * q.sigma_6 * r6inv - 1.0);
        return c1 * r6inv * (q.sigma_6 * r6inv - 1.0);
        }

    __device__ static double colloid_solvent(const Prep<double>& q, bool force, double& fdivr, double rsq)
        {
        const double a = (q.ai > q.aj) ? q.ai : q.aj;
        const double asq = a * a;
        const double asq_minus_rsq = asq - rsq;
        const double rsqsq = rsq * rsq;
        const double asq_minus_rsq_3 = asq_minus_rsq * asq_minus_rsq * asq_minus_rsq;
        const double asq_minus_rsq_6 = asq_minus_rsq_3 * asq_minus_rsq_3;
        const double fR = q.sigma_3 * q.A * a * asq / asq_minus_rsq_3;
        if (force)
            {
            fdivr = (4.0 / 15.0) * fR
                    * (2.0 * (asq + rsq) * (asq * (5.0 * asq + 22.0 * rsq) + 5.0 * rsqsq) * q.sigma_6
                           / asq_minus_rsq_6
                       - 5.0)
                    / asq_minus_rsq;
            }
        return (2.0 / 9.0) * fR
               * (1.0
                  - (asq * (asq * (asq / 3.0 + 3.0 * rsq) + 4.2 * rsqsq) + rsq * rsqsq) * q.sigma_6
                        / asq_minus_rsq_6);
        }

    __device__ static double colloid_colloid(const Prep<double>& q, bool force, double& fdivr, double rsq)
        {
        const double r = sqrt(rsq);
        const double k0 = q.ai * q.aj;
        const double k1 = q.ai + q.aj;
        const double k2 = q.ai - q.aj;
        const double k3 = k1 + r;
        const double k4 = k1 - r;
        const double k5 = k2 + r;
        const double k6 = k2 - r;
        const double k7 = 1.0 / (k3 * k4);
        const double k8 = 1.0 / (k5 * k6);
        const double k3inv = 1.0 / k3;
        double g0 = k3inv * k3inv;
        g0 *= g0 * g0;
        g0 *= k3inv;
        const double k4inv = 1.0 / k4;
        double g1 = k4inv * k4inv;
        g1 *= g1 * g1;
        g1 *= k4inv;
        const double k5inv = 1.0 / k5;
        double g2 = k5inv * k5inv;
        g2 *= g2 * g2;
        g2 *= k5inv;
        const double k6inv = 1.0 / k6;
        double g3 = k6inv * k6inv;
        g3 *= g3 * g3;
        g3 *= k6inv;
        const double h0 = ((k3 + 5.0 * k1) * k3 + 30.0 * k0) * g0;
        const double h1 = ((k4 + 5.0 * k1) * k4 + 30.0 * k0) * g1;
        const double h2 = ((k5 + 5.0 * k2) * k5 - 30.0 * k0) * g2;
        const double h3 = ((k6 + 5.0 * k2) * k6 - 30.0 * k0) * g3;
        g0 *= 42.0 * k0 / k3 + 6.0 * k1 + k3;
        g1 *= 42.0 * k0 / k4 + 6.0 * k1 + k4;
        g2 *= -42.0 * k0 / k5 + 6.0 * k2 + k5;
        g3 *= -42.0 * k0 / k6 + 6.0 * k2 + k6;
        const double fR = q.A * q.sigma_6 / r / 37800.0;
        double pair_eng = fR * (h0 - h1 - h2 + h3);
        if (force)
            {
            const double dUR = pair_eng / r + 5.0 * fR * (g0 + g1 - g2 - g3);
            const double dUA
                = -q.A / 3.0 * r * ((2.0 * k0 * k7 + 1.0) * k7 + (2.0 * k0 * k8 - 1.0) * k8);
            fdivr = (dUR + dUA) / r;
            }
        pair_eng += q.A / 6.0 * (2.0 * k0 * (k7 + k8) - log(k8 / k7));
        return pair_eng;
        }

    __device__ static double branch_energy(const Prep<double>& q, bool force, double& fdivr, double rsq)
        {
        if (q.branch == 0)
            return solvent_solvent(q, force, fdivr, rsq);
        if (q.branch == 2)
            return colloid_colloid(q, force, fdivr, rsq);
        return colloid_solvent(q, force, fdivr, rsq);
        }

    template<typename R>
    __device__ static Prep<R> prepare(const param_type& p, double rcutsq, bool shift)
        {
        Prep<double> q;
        q.A = p.A;
        q.ai = p.a_1;
        q.aj = p.a_2;
        q.sigma_3 = p.sigma_3;
        q.sigma_6 = p.sigma_3 * p.sigma_3;
        q.branch = (p.a_1 == 0.0 && p.a_2 == 0.0) ? 0 : ((p.a_1 != 0.0 && p.a_2 != 0.0) ? 2 : 1);
        q.eshift = 0.0;
        if (shift && p.A != 0.0)
            {
            double dummy = 0.0;
            q.eshift = branch_energy(q, false, dummy, rcutsq);
            }
        return q;
        }
    template<typename R> __device__ static bool active(const Prep<R>& q)
        {
        return q.A != R(0);
        }
    __device__ static void compute(const Prep<double>& q, double rsq, double& fdivr, double& e)
        {
        e = branch_energy(q, true, fdivr, rsq);
        e -= q.eshift;
        }
    };

} // namespace azp

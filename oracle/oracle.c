/*
 * ORACLE — test infrastructure only (see oracle.h).  Plain C, FP64, one
 * thread, compiled with -ffp-contract=off so that every product and sum is
 * rounded separately, in source order.
 *
 * Reference files restated (paths relative to the reference's src/):
 *   PairEvaluatorHertz.h, PairEvaluatorExpandedYukawa.h,
 *   PairEvaluatorPerturbedLennardJones.h, PairEvaluatorColloid.h,
 *   DPDPairEvaluatorGeneralWeight.h, RNGIdentifiers.h,
 *   AnisoPairEvaluatorTwoPatchMorse.h, pair.py (parameter keys / modes),
 *   WallEvaluatorLJ93.h, WallEvaluatorColloid.h, wall.py (wall potentials).
 * Upstream HOOMD-blue v7.0.1 driver behaviour (PotentialPair,
 * PotentialPairDPDThermo, AnisoPotentialPair, NeighborListBinned, BoxDim,
 * RandomNumbers) is not in the reference tree; it is restated from its
 * documented semantics at the reference's call sites:
 *   export_PotentialPair.cc.inc:26-29, export_PotentialPairDPDThermo.cc.inc:31-39,
 *   export_AnisoPotentialPair.cc.inc:22-26, DPDPairEvaluatorGeneralWeight.h:227-233.
 */
#include "oracle.h"

#include "../azplugins_amd/csrc/portable_math.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* parameter packing                                                         */
/* ------------------------------------------------------------------------ */

int orc_param_nvalues(int evaluator)
{
    switch (evaluator)
        {
    case ORC_HERTZ:
        return 1;
    case ORC_EXPANDED_YUKAWA:
        return 3;
    case ORC_PERTURBED_LJ:
        return 3;
    case ORC_COLLOID:
        return 4;
    case ORC_DPD_GENERAL_WEIGHT:
        return 3;
    case ORC_TWO_PATCH_MORSE:
        return 6;
    default:
        return -1;
        }
}

/* Follows the param_type constructors from a dict:
 *   PairEvaluatorHertz.h:28-31, PairEvaluatorExpandedYukawa.h:28-33,
 *   PairEvaluatorPerturbedLennardJones.h:29-44, PairEvaluatorColloid.h:28-35,
 *   DPDPairEvaluatorGeneralWeight.h:37-42, AnisoPairEvaluatorTwoPatchMorse.h:39-47.
 * Value order = the dict keys in pair.py (:116, :229, :295, :349, :423, :516-521).
 * Parity: pinned (param read-back of the reference tests, tests/test_oracle_golden.py). */
int orc_param_pack(int evaluator, const double* v, double* p)
{
    switch (evaluator)
        {
    case ORC_HERTZ:
        p[0] = v[0];
        return 1;
    case ORC_EXPANDED_YUKAWA:
        p[0] = v[0];
        p[1] = v[1];
        p[2] = v[2];
        return 3;
    case ORC_PERTURBED_LJ:
        {
        const double epsilon = v[0];
        const double sigma = v[1];
        const double sigma_2 = sigma * sigma;
        const double sigma_4 = sigma_2 * sigma_2;
        p[0] = sigma_2 * sigma_4;           /* sigma_6 */
        p[1] = 4.0 * epsilon;               /* epsilon_x_4 */
        p[2] = v[2];                        /* attraction_scale_factor */
        p[3] = pow(2.0, 1.0 / 3.0) * sigma_2; /* rwcasq */
        return 4;
        }
    case ORC_COLLOID:
        p[0] = v[0];
        p[1] = v[1];
        p[2] = v[2];
        p[3] = v[3] * v[3] * v[3]; /* sigma_3 */
        return 4;
    case ORC_DPD_GENERAL_WEIGHT:
        p[0] = v[0];
        p[1] = v[1];
        p[2] = v[2];
        return 3;
    case ORC_TWO_PATCH_MORSE:
        p[0] = v[0];
        p[1] = 1.0 / v[1]; /* M_rinv */
        p[2] = v[2];
        p[3] = v[3];
        p[4] = v[4];
        p[5] = (v[5] != 0.0) ? 1.0 : 0.0;
        return 6;
    default:
        return -1;
        }
}

/* Follows asDict()/toPython(): PairEvaluatorPerturbedLennardJones.h:47-54,
 * PairEvaluatorColloid.h:37-44, AnisoPairEvaluatorTwoPatchMorse.h:50-60. */
int orc_param_unpack(int evaluator, const double* p, double* v)
{
    switch (evaluator)
        {
    case ORC_PERTURBED_LJ:
        v[0] = p[1] / 4.0;
        v[1] = pow(p[0], 1.0 / 6.0);
        v[2] = p[2];
        return 3;
    case ORC_COLLOID:
        v[0] = p[0];
        v[1] = p[1];
        v[2] = p[2];
        v[3] = cbrt(p[3]);
        return 4;
    case ORC_TWO_PATCH_MORSE:
        v[0] = p[0];
        v[1] = 1.0 / p[1];
        v[2] = p[2];
        v[3] = p[3];
        v[4] = p[4];
        v[5] = p[5];
        return 6;
    default:
        {
        const int n = orc_param_nvalues(evaluator);
        for (int k = 0; k < n; ++k)
            v[k] = p[k];
        return n;
        }
        }
}

/* ------------------------------------------------------------------------ */
/* isotropic evaluators                                                      */
/* ------------------------------------------------------------------------ */

/* PairEvaluatorHertz.h:93-110.  Pinned: test_pair.py:174-216. */
static int eval_hertz(double rsq, double rcutsq, const double* p, double* fdivr, double* e)
{
    const double epsilon = p[0];
    if (rsq < rcutsq && epsilon != 0.0)
        {
        const double r = sqrt(rsq);
        const double rcut = sqrt(rcutsq);
        const double x = 1.0 - (r / rcut);
        const double xsqrt = sqrt(x);
        const double ex3p2 = epsilon * x * xsqrt;
        *fdivr = 2.5 * ex3p2 / (r * rcut);
        *e = ex3p2 * x;
        return 1;
        }
    return 0;
}

/* PairEvaluatorExpandedYukawa.h:92-115.  Pinned: test_pair.py:110-172. */
static int
eval_yukawa(double rsq, double rcutsq, const double* p, int shift, double* fdivr, double* e)
{
    const double epsilon = p[0], kappa = p[1], delta = p[2];
    if (rsq < rcutsq && epsilon != 0.0)
        {
        const double r = sqrt(rsq);
        const double r_delta = r - delta;
        const double r_delta_inv = 1.0 / r_delta;
        *e = epsilon * exp(-kappa * r_delta) * r_delta_inv;
        *fdivr = *e * (kappa + r_delta_inv) / r;
        if (shift)
            {
            const double rcut = sqrt(rcutsq);
            const double rcut_delta = rcut - delta;
            *e -= epsilon * exp(-kappa * rcut_delta) / rcut_delta;
            }
        return 1;
        }
    return 0;
}

/* PairEvaluatorPerturbedLennardJones.h:99-103 (ctor) and :117-155.
 * Pinned: test_pair.py:218-306. */
static int eval_plj(double rsq, double rcutsq, const double* p, int shift, double* fdivr, double* e)
{
    const double sigma_6 = p[0], epsilon_x_4 = p[1];
    const double lj1 = epsilon_x_4 * sigma_6 * sigma_6;
    const double lj2 = epsilon_x_4 * sigma_6;
    const double asf = p[2];
    const double rwcasq = p[3];
    const double wca_shift = epsilon_x_4 * (1.0 - asf) / 4.0;
    if (rsq < rcutsq && lj1 != 0)
        {
        const double r2inv = 1.0 / rsq;
        const double r6inv = r2inv * r2inv * r2inv;
        *fdivr = r2inv * r6inv * (12.0 * lj1 * r6inv - 6.0 * lj2);
        *e = r6inv * (lj1 * r6inv - lj2);
        if (rsq < rwcasq)
            {
            *e += wca_shift;
            }
        else
            {
            *fdivr *= asf;
            *e *= asf;
            }
        if (shift)
            {
            const double rcut2inv = 1.0 / rcutsq;
            const double rcut6inv = rcut2inv * rcut2inv * rcut2inv;
            double pair_eng_shift = rcut6inv * (lj1 * rcut6inv - lj2);
            if (rcutsq < rwcasq)
                pair_eng_shift += wca_shift;
            else
                pair_eng_shift *= asf;
            *e -= pair_eng_shift;
            }
        return 1;
        }
    return 0;
}

/* PairEvaluatorColloid.h:101-113 (solvent-solvent). */
static double colloid_ss(const double* p, int force, double* fdivr, double rsq)
{
    const double A = p[0];
    const double sigma_6 = p[3] * p[3];
    const double r2inv = 1.0 / rsq;
    const double r6inv = r2inv * r2inv * r2inv;
    const double c1 = A * sigma_6 / 36.0;
    if (force)
        *fdivr = 6.0 * c1 * r2inv * r6inv * (2.0 * sigma_6 * r6inv - 1.0);
    return c1 * r6inv * (sigma_6 * r6inv - 1.0);
}

/* PairEvaluatorColloid.h:125-152 (colloid-solvent). */
static double colloid_cs(const double* p, int force, double* fdivr, double rsq)
{
    const double A = p[0], ai = p[1], aj = p[2], sigma_3 = p[3];
    const double sigma_6 = sigma_3 * sigma_3;
    const double a = (ai > aj) ? ai : aj;
    const double asq = a * a;
    const double asq_minus_rsq = asq - rsq;
    const double rsqsq = rsq * rsq;
    const double asq_minus_rsq_3 = asq_minus_rsq * asq_minus_rsq * asq_minus_rsq;
    const double asq_minus_rsq_6 = asq_minus_rsq_3 * asq_minus_rsq_3;
    const double fR = sigma_3 * A * a * asq / asq_minus_rsq_3;
    if (force)
        {
        *fdivr = (4.0 / 15.0) * fR
                 * (2.0 * (asq + rsq) * (asq * (5.0 * asq + 22.0 * rsq) + 5.0 * rsqsq) * sigma_6
                        / asq_minus_rsq_6
                    - 5.0)
                 / asq_minus_rsq;
        }
    return (2.0 / 9.0) * fR
           * (1.0
              - (asq * (asq * (asq / 3.0 + 3.0 * rsq) + 4.2 * rsqsq) + rsq * rsqsq) * sigma_6
                    / asq_minus_rsq_6);
}

/* PairEvaluatorColloid.h:164-220 (colloid-colloid, Everaers-Ejtehadi). */
static double colloid_cc(const double* p, int force, double* fdivr, double rsq)
{
    const double A = p[0], ai = p[1], aj = p[2], sigma_3 = p[3];
    const double sigma_6 = sigma_3 * sigma_3;
    const double r = sqrt(rsq);
    const double k0 = ai * aj;
    const double k1 = ai + aj;
    const double k2 = ai - aj;
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

    const double fR = A * sigma_6 / r / 37800.0;
    double pair_eng = fR * (h0 - h1 - h2 + h3);
    if (force)
        {
        const double dUR = pair_eng / r + 5.0 * fR * (g0 + g1 - g2 - g3);
        const double dUA
            = -A / 3.0 * r * ((2.0 * k0 * k7 + 1.0) * k7 + (2.0 * k0 * k8 - 1.0) * k8);
        *fdivr = (dUR + dUA) / r;
        }
    pair_eng += A / 6.0 * (2.0 * k0 * (k7 + k8) - log(k8 / k7));
    return pair_eng;
}

/* PairEvaluatorColloid.h:233-269.  Pinned: test_pair.py:22-73. */
static int
eval_colloid(double rsq, double rcutsq, const double* p, int shift, double* fdivr, double* e)
{
    const double A = p[0], ai = p[1], aj = p[2];
    double dummy = 0.0;
    if (rsq < rcutsq && A != 0.0)
        {
        if (ai == 0.0 && aj == 0.0)
            {
            *e = colloid_ss(p, 1, fdivr, rsq);
            if (shift)
                *e -= colloid_ss(p, 0, &dummy, rcutsq);
            }
        else if (ai != 0.0 && aj != 0.0)
            {
            *e = colloid_cc(p, 1, fdivr, rsq);
            if (shift)
                *e -= colloid_cc(p, 0, &dummy, rcutsq);
            }
        else
            {
            *e = colloid_cs(p, 1, fdivr, rsq);
            if (shift)
                *e -= colloid_cs(p, 0, &dummy, rcutsq);
            }
        return 1;
        }
    return 0;
}

/* DPDPairEvaluatorGeneralWeight.h:165-183 (conservative only, used by
 * PotentialPairConservative). */
static int eval_dpd_cons(double rsq, double rcutsq, const double* p, double* fdivr, double* e)
{
    const double A = p[0];
    if (rsq < rcutsq)
        {
        const double rinv = 1.0 / sqrt(rsq);
        const double r = 1.0 / rinv;
        const double rcutinv = 1.0 / sqrt(rcutsq);
        const double rcut = 1.0 / rcutinv;
        *fdivr = A * (rinv - rcutinv);
        *e = A * (rcut - r) - 0.5 * A * rcutinv * (rcutsq - rsq);
        return 1;
        }
    return 0;
}

int orc_eval_iso(int evaluator,
                 double rsq,
                 double rcutsq,
                 const double* p,
                 int shift,
                 double* fdivr,
                 double* e)
{
    switch (evaluator)
        {
    case ORC_HERTZ:
        return eval_hertz(rsq, rcutsq, p, fdivr, e);
    case ORC_EXPANDED_YUKAWA:
        return eval_yukawa(rsq, rcutsq, p, shift, fdivr, e);
    case ORC_PERTURBED_LJ:
        return eval_plj(rsq, rcutsq, p, shift, fdivr, e);
    case ORC_COLLOID:
        return eval_colloid(rsq, rcutsq, p, shift, fdivr, e);
    case ORC_DPD_GENERAL_WEIGHT:
        return eval_dpd_cons(rsq, rcutsq, p, fdivr, e);
    default:
        return 0;
        }
}

/* ------------------------------------------------------------------------ */
/* Philox4x32-10 and the DPD random variate                                   */
/* ------------------------------------------------------------------------ */

/* Philox4x32 with 10 rounds (Salmon, Moraes, Dror, Shaw, SC'11).
 * Pinned: published Random123 known-answer vectors (tests/golden/philox_kat.json). */
void orc_philox4x32_10(const uint32_t* ctr, const uint32_t* key, uint32_t* out)
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int round = 0; round < 10; ++round)
        {
        if (round > 0)
            {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
            }
        const uint64_t p0 = (uint64_t)0xD2511F53u * (uint64_t)c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * (uint64_t)c2;
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        const uint32_t n1 = (uint32_t)p1;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        const uint32_t n3 = (uint32_t)p0;
        c0 = n0;
        c1 = n1;
        c2 = n2;
        c3 = n3;
        }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

/* DPDPairEvaluatorGeneralWeight.h:213-233 with RNGIdentifiers.h:23 (id 200).
 * The Seed/Counter packing, 64-bit word order and u01 conversion belong to
 * upstream hoomd/RandomNumbers.h (not in the reference tree):
 *   key = { id<<24 | seed<<8 | (t>>32)&0xff, t & 0xffffffff },
 *   ctr = { min(tag_i,tag_j), max(tag_i,tag_j), 0, 0 },
 *   u64 = out0<<32 | out1,  U = u64 * 2^-64 + 2^-65,  alpha = -1 + 2 U.
 * Parity unpinned for the packing (no reference test holds a vector); the
 * Philox core is pinned by KATs and the distribution by moments. */
double orc_dpd_alpha(uint16_t seed, uint32_t timestep, uint32_t tag_i, uint32_t tag_j)
{
    const uint32_t oi = (tag_i > tag_j) ? tag_j : tag_i;
    const uint32_t oj = (tag_i > tag_j) ? tag_i : tag_j;
    const uint64_t t = (uint64_t)timestep;
    uint32_t key[2];
    key[0] = ((uint32_t)200u << 24) | ((uint32_t)seed << 8) | (uint32_t)((t >> 32) & 0xffu);
    key[1] = (uint32_t)(t & 0xffffffffu);
    uint32_t ctr[4] = {oi, oj, 0u, 0u};
    uint32_t u[4];
    orc_philox4x32_10(ctr, key, u);
    const uint64_t u64 = ((uint64_t)u[0] << 32) | (uint64_t)u[1];
    const double U = (double)u64 * (1.0 / 18446744073709551616.0) + (0.5 / 18446744073709551616.0);
    return -1.0 + 2.0 * U;
}

void orc_dpd_alpha_many(uint16_t seed,
                        uint32_t timestep,
                        uint64_t n,
                        const uint32_t* tag_i,
                        const uint32_t* tag_j,
                        double* alpha)
{
    for (uint64_t k = 0; k < n; ++k)
        alpha[k] = orc_dpd_alpha(seed, timestep, tag_i[k], tag_j[k]);
}

double orc_ppow(double x, double y)
{
    return azp_ppow(x, y);
}
double orc_pexp(double x)
{
    return azp_pexp(x);
}
double orc_plog(double x)
{
    return azp_plog(x);
}

/* DPDPairEvaluatorGeneralWeight.h:198-255.  fast::rsqrt(x) is 1/sqrt(x) and
 * fast::pow is pow (HOOMDMath, host side).  portable_pow selects the shared
 * bit-exact pow that the device uses instead of libm pow.
 * Pinned: conservative part by test_pair.py:75-108 (kT = 0); dissipative part
 * by the closed form; random part bit-exact against the device (same packing). */
int orc_eval_dpd(double rsq,
                 double rcutsq,
                 const double* p,
                 uint16_t seed,
                 uint32_t tag_i,
                 uint32_t tag_j,
                 uint32_t timestep,
                 double dt,
                 double kT,
                 double rdotv,
                 int portable_pow,
                 double* fdivr,
                 double* fdivr_cons,
                 double* e)
{
    const double A = p[0], gamma = p[1], s = p[2];
    if (rsq < rcutsq)
        {
        const double rinv = 1.0 / sqrt(rsq);
        const double r = 1.0 / rinv;
        const double rcutinv = 1.0 / sqrt(rcutsq);
        const double rcut = 1.0 / rcutinv;
        const double alpha = orc_dpd_alpha(seed, timestep, tag_i, tag_j);
        *fdivr = A * (rinv - rcutinv);
        *fdivr_cons = *fdivr;
        const double base = 1.0 - r * rcutinv;
        const double w = portable_pow ? azp_ppow(base, 0.5 * s) : pow(base, 0.5 * s);
        const double wR = w * rinv;
        *fdivr -= gamma * wR * wR * rdotv;
        *fdivr += (1.0 / sqrt(dt / (kT * gamma * 6.0))) * wR * alpha;
        *e = A * (rcut - r) - 0.5 * A * rcutinv * (rcutsq - rsq);
        return 1;
        }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* anisotropic: TwoPatchMorse                                               */
/* ------------------------------------------------------------------------ */

static inline double dot3(const double* a, const double* b)
{
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

static inline void cross3(const double* a, const double* b, double* c)
{
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

/* rotate(q, v) = (s^2 - |u|^2) v + 2 s (u x v) + 2 (u.v) u for q = (s, u),
 * upstream VectorMath.h as called at AnisoPairEvaluatorTwoPatchMorse.h:145-146. */
static void rotate3(const double* q, const double* v, double* out)
{
    const double s = q[0];
    const double u[3] = {q[1], q[2], q[3]};
    double c[3];
    cross3(u, v, c);
    const double a = s * s - dot3(u, u);
    const double uv = dot3(u, v);
    for (int k = 0; k < 3; ++k)
        out[k] = a * v[k] + 2.0 * s * c[k] + 2.0 * uv * u[k];
}

/* AnisoPairEvaluatorTwoPatchMorse.h:127-216.  Pinned: test_pair_aniso.py:21-110. */
int orc_eval_two_patch_morse(const double* dr,
                             const double* quat_i,
                             const double* quat_j,
                             double rcutsq,
                             const double* p,
                             int energy_shift,
                             double* force,
                             double* pair_eng,
                             double* torque_i,
                             double* torque_j)
{
    const double M_d = p[0], M_rinv = p[1], r_eq = p[2], omega = p[3], alpha = p[4];
    const int repulsion = p[5] != 0.0;
    const double rsq = dot3(dr, dr);
    if (rsq > rcutsq)
        return 0;
    const double rinv = 1.0 / sqrt(rsq);
    const double r = 1.0 / rinv;
    double unitr[3] = {dr[0] * rinv, dr[1] * rinv, dr[2] * rinv};
    const double ex[3] = {1.0, 0.0, 0.0};
    double n_i[3], n_j[3];
    rotate3(quat_i, ex, n_i);
    rotate3(quat_j, ex, n_j);

    double UMorse = -1.0 * M_d;
    double dUMorse_dr = 0.0;
    if (r > r_eq || repulsion)
        {
        const double Morse_exp = exp(-(r - r_eq) * M_rinv);
        const double one_minus_exp = 1.0 - Morse_exp;
        UMorse = M_d * (one_minus_exp * one_minus_exp - 1.0);
        dUMorse_dr = 2.0 * M_d * M_rinv * Morse_exp * one_minus_exp;
        }

    const double gamma_i = dot3(unitr, n_i);
    const double gamma_i_exp = exp(-omega * (gamma_i * gamma_i - alpha));
    const double Omega_i = 1.0 / (1.0 + gamma_i_exp);
    const double gamma_j = dot3(unitr, n_j);
    const double gamma_j_exp = exp(-omega * (gamma_j * gamma_j - alpha));
    const double Omega_j = 1.0 / (1.0 + gamma_j_exp);

    double e = 0.0;
    e += UMorse * Omega_i * Omega_j;

    const double dU_dr = dUMorse_dr * Omega_i * Omega_j;
    const double dOmegai_dgi = 2.0 * omega * gamma_i * gamma_i_exp * Omega_i * Omega_i;
    const double dOmegaj_dgj = 2.0 * omega * gamma_j * gamma_j_exp * Omega_j * Omega_j;
    const double dU_dgi = dOmegai_dgi * UMorse * Omega_j;
    const double dU_dgj = dOmegaj_dgj * UMorse * Omega_i;

    double mur[3] = {-unitr[0], -unitr[1], -unitr[2]};
    double ci[3], cj[3], n_i_perp[3], n_j_perp[3];
    cross3(unitr, n_i, ci);
    cross3(unitr, n_j, cj);
    cross3(mur, ci, n_i_perp);
    cross3(mur, cj, n_j_perp);

    for (int k = 0; k < 3; ++k)
        {
        force[k] = -dU_dr * unitr[k] - rinv * (dU_dgi * n_i_perp[k] + dU_dgj * n_j_perp[k]);
        torque_i[k] = dU_dgi * ci[k];
        torque_j[k] = dU_dgj * cj[k];
        }

    if (energy_shift)
        {
        const double rcut = sqrt(rcutsq);
        const double Morse_exp_shift = exp(-(rcut - r_eq) * M_rinv);
        const double one_minus_exp_shift = 1.0 - Morse_exp_shift;
        const double UMorse_shift = M_d * (one_minus_exp_shift * one_minus_exp_shift - 1.0);
        e -= UMorse_shift * Omega_i * Omega_j;
        }
    *pair_eng = e;
    return 1;
}

/* ------------------------------------------------------------------------ */
/* box and neighbor lists                                                    */
/* ------------------------------------------------------------------------ */

/* Orthorhombic, fully periodic minimum image: img = rint(d / L); d -= L img
 * (upstream BoxDim::minImage as reached through the pair drivers). */
void orc_min_image(const double* L, double* d)
{
    for (int k = 0; k < 3; ++k)
        {
        const double Linv = 1.0 / L[k];
        const double img = rint(d[k] * Linv);
        d[k] = d[k] - L[k] * img;
        }
}

static inline unsigned type_of(const double* pos4, uint32_t i)
{
    return (unsigned)pos4[4 * (size_t)i + 3];
}

static inline double pair_rsq(const double* pos4, const double* L, uint32_t i, uint32_t j, double* d)
{
    d[0] = pos4[4 * (size_t)i + 0] - pos4[4 * (size_t)j + 0];
    d[1] = pos4[4 * (size_t)i + 1] - pos4[4 * (size_t)j + 1];
    d[2] = pos4[4 * (size_t)i + 2] - pos4[4 * (size_t)j + 2];
    orc_min_image(L, d);
    return d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
}

/* All pairs with rsq <= r_listsq(ti,tj) (upstream NeighborListBinned predicate;
 * pairs whose r_listsq is 0 are excluded).  O(N^2), for small N. */
int64_t orc_nlist_brute(uint32_t N,
                        const double* pos4,
                        const double* L,
                        uint32_t ntypes,
                        const double* r_listsq,
                        int64_t cap,
                        uint32_t* pi,
                        uint32_t* pj)
{
    int64_t n = 0;
    double d[3];
    for (uint32_t i = 0; i < N; ++i)
        {
        const unsigned ti = type_of(pos4, i);
        for (uint32_t j = i + 1; j < N; ++j)
            {
            const double rl = r_listsq[ti * ntypes + type_of(pos4, j)];
            if (rl <= 0.0)
                continue;
            if (pair_rsq(pos4, L, i, j, d) <= rl)
                {
                if (n < cap)
                    {
                    pi[n] = i;
                    pj[n] = j;
                    }
                ++n;
                }
            }
        }
    return (n <= cap) ? n : -n;
}

/* Same predicate, binned into cells of width >= max r_list so only the 27
 * surrounding cells are scanned.  Produces the same pair set as the brute
 * force (order differs). */
int64_t orc_nlist_cell(uint32_t N,
                       const double* pos4,
                       const double* L,
                       uint32_t ntypes,
                       const double* r_listsq,
                       int64_t cap,
                       uint32_t* pi,
                       uint32_t* pj)
{
    double rmax = 0.0;
    for (uint32_t k = 0; k < ntypes * ntypes; ++k)
        if (r_listsq[k] > rmax)
            rmax = r_listsq[k];
    rmax = sqrt(rmax) * (1.0 + 1e-9);
    int nc[3];
    for (int k = 0; k < 3; ++k)
        {
        nc[k] = (int)floor(L[k] / rmax);
        if (nc[k] < 1)
            nc[k] = 1;
        }
    const int64_t ncell = (int64_t)nc[0] * nc[1] * nc[2];
    int64_t* head = (int64_t*)malloc(sizeof(int64_t) * (size_t)ncell);
    int64_t* next = (int64_t*)malloc(sizeof(int64_t) * (size_t)N);
    int64_t* cell_of = (int64_t*)malloc(sizeof(int64_t) * (size_t)N);
    for (int64_t c = 0; c < ncell; ++c)
        head[c] = -1;
    /* insert in reverse so each cell's chain is in increasing index order */
    for (int64_t ii = (int64_t)N - 1; ii >= 0; --ii)
        {
        int ci[3];
        for (int k = 0; k < 3; ++k)
            {
            const double f = (pos4[4 * ii + k] + 0.5 * L[k]) / L[k];
            int c = (int)floor(f * nc[k]);
            c %= nc[k];
            if (c < 0)
                c += nc[k];
            ci[k] = c;
            }
        const int64_t cid = ((int64_t)ci[2] * nc[1] + ci[1]) * nc[0] + ci[0];
        cell_of[ii] = cid;
        next[ii] = head[cid];
        head[cid] = ii;
        }
    int64_t n = 0;
    double d[3];
    for (uint32_t i = 0; i < N; ++i)
        {
        const unsigned ti = type_of(pos4, i);
        const int64_t cid = cell_of[i];
        const int cx = (int)(cid % nc[0]);
        const int cy = (int)((cid / nc[0]) % nc[1]);
        const int cz = (int)(cid / ((int64_t)nc[0] * nc[1]));
        int64_t seen[27];
        int nseen = 0;
        for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx)
                    {
                    const int x = (cx + dx + nc[0]) % nc[0];
                    const int y = (cy + dy + nc[1]) % nc[1];
                    const int z = (cz + dz + nc[2]) % nc[2];
                    const int64_t nb = ((int64_t)z * nc[1] + y) * nc[0] + x;
                    int dup = 0;
                    for (int s = 0; s < nseen; ++s)
                        if (seen[s] == nb)
                            dup = 1;
                    if (dup)
                        continue;
                    seen[nseen++] = nb;
                    for (int64_t j = head[nb]; j >= 0; j = next[j])
                        {
                        if ((uint32_t)j <= i)
                            continue;
                        const double rl = r_listsq[ti * ntypes + type_of(pos4, (uint32_t)j)];
                        if (rl <= 0.0)
                            continue;
                        if (pair_rsq(pos4, L, i, (uint32_t)j, d) <= rl)
                            {
                            if (n < cap)
                                {
                                pi[n] = i;
                                pj[n] = (uint32_t)j;
                                }
                            ++n;
                            }
                        }
                    }
        }
    free(head);
    free(next);
    free(cell_of);
    return (n <= cap) ? n : -n;
}

/* ------------------------------------------------------------------------ */
/* drivers                                                                   */
/* ------------------------------------------------------------------------ */

static void zero_outputs(uint32_t N, double* force4, double* virial, double* torque4)
{
    if (force4)
        memset(force4, 0, sizeof(double) * 4 * (size_t)N);
    if (virial)
        memset(virial, 0, sizeof(double) * 6 * (size_t)N);
    if (torque4)
        memset(torque4, 0, sizeof(double) * 4 * (size_t)N);
}

static inline void add_virial_iso(double* virial, uint32_t N, uint32_t i, double fd2, const double* d)
{
    virial[0 * (size_t)N + i] += fd2 * d[0] * d[0];
    virial[1 * (size_t)N + i] += fd2 * d[0] * d[1];
    virial[2 * (size_t)N + i] += fd2 * d[0] * d[2];
    virial[3 * (size_t)N + i] += fd2 * d[1] * d[1];
    virial[4 * (size_t)N + i] += fd2 * d[1] * d[2];
    virial[5 * (size_t)N + i] += fd2 * d[2] * d[2];
}

/* Upstream PotentialPair<evaluator>::computeForces (CPU, half list, third law)
 * as instantiated at export_PotentialPair.cc.inc:26-29:
 *   energy_shift = mode == shift || (mode == xplor && ronsq > rcutsq);
 *   xplor for ronsq <= rsq < rcutsq:
 *     S = (rc2 - r2)^2 (rc2 + 2 r2 - 3 ron2) / (rc2 - ron2)^3,
 *     force_divr = S f - (dS/dr / r) E,  E = S E;
 *   F_i += dx f, F_j -= dx f, E_i, E_j += E/2, W_i, W_j += (f/2) dx (x) dx.
 * Parity: shift/none pinned (test_pair.py); xplor parity unpinned (no reference
 * test), checked by F = -dU/dr.  Virial pinned by W = sum dx (x) F / 2 and the
 * affine-strain derivative of U. */
/* One pair's force_divr and energy (with shift / xplor) and its
 * displacement d = r_i - r_j (minimum image); 0 when it does not interact. */
static int iso_pair(int evaluator,
                    const double* pos4,
                    const double* L,
                    uint32_t i,
                    uint32_t j,
                    uint32_t ntypes,
                    const double* packed,
                    const double* rcutsq_m,
                    const double* ronsq_m,
                    int shift_mode,
                    double* d,
                    double* fdivr_out,
                    double* e_out)
{
    const int stride = 8; /* packed params: 8 doubles per type pair */
    const double rsq = pair_rsq(pos4, L, i, j, d);
    const unsigned tp = type_of(pos4, i) * ntypes + type_of(pos4, j);
    const double rcutsq = rcutsq_m[tp];
    const double ronsq = ronsq_m ? ronsq_m[tp] : 0.0;
    int energy_shift = 0;
    if (shift_mode == ORC_SHIFT_SHIFT)
        energy_shift = 1;
    else if (shift_mode == ORC_SHIFT_XPLOR && ronsq > rcutsq)
        energy_shift = 1;
    double fdivr = 0.0, e = 0.0;
    if (!orc_eval_iso(evaluator, rsq, rcutsq, packed + stride * tp, energy_shift, &fdivr, &e))
        return 0;
    if (shift_mode == ORC_SHIFT_XPLOR && rsq >= ronsq && rsq < rcutsq)
        {
        const double old_e = e, old_f = fdivr;
        const double denom = (rcutsq - ronsq) * (rcutsq - ronsq) * (rcutsq - ronsq);
        const double denom_inv = 1.0 / denom;
        const double rsq_m_rcsq = rsq - rcutsq;
        const double S = rsq_m_rcsq * rsq_m_rcsq * (rcutsq + 2.0 * rsq - 3.0 * ronsq) * denom_inv;
        const double dS_dr_divr = 12.0 * (rsq - ronsq) * rsq_m_rcsq * denom_inv;
        e = old_e * S;
        fdivr = S * old_f - dS_dr_divr * old_e;
        }
    *fdivr_out = fdivr;
    *e_out = e;
    return 1;
}

void orc_pair_forces(int evaluator,
                     uint32_t N,
                     const double* pos4,
                     const double* L,
                     int64_t npairs,
                     const uint32_t* pair_i,
                     const uint32_t* pair_j,
                     uint32_t ntypes,
                     const double* packed,
                     const double* rcutsq_m,
                     const double* ronsq_m,
                     int shift_mode,
                     double* force4,
                     double* virial)
{
    zero_outputs(N, force4, virial, NULL);
    double d[3];
    for (int64_t k = 0; k < npairs; ++k)
        {
        const uint32_t i = pair_i[k], j = pair_j[k];
        double fdivr, e;
        if (!iso_pair(evaluator, pos4, L, i, j, ntypes, packed, rcutsq_m, ronsq_m, shift_mode, d, &fdivr, &e))
            continue;
        const double fd2 = fdivr * 0.5;
        for (int a = 0; a < 3; ++a)
            {
            force4[4 * (size_t)i + a] += d[a] * fdivr;
            force4[4 * (size_t)j + a] -= d[a] * fdivr;
            }
        force4[4 * (size_t)i + 3] += e * 0.5;
        force4[4 * (size_t)j + 3] += e * 0.5;
        if (virial)
            {
            add_virial_iso(virial, N, i, fd2, d);
            add_virial_iso(virial, N, j, fd2, d);
            }
        }
}

/* Full (directed) CSR list from a half list by counting sort: the neighbors
 * of i are nbrs[offsets[i] .. offsets[i + 1]), in the order the half list
 * meets them.  offsets has N + 1 entries, nbrs 2 npairs. */
void orc_full_list(uint32_t N, int64_t npairs, const uint32_t* pair_i, const uint32_t* pair_j, int64_t* offsets,
                   uint32_t* nbrs)
{
    memset(offsets, 0, sizeof(int64_t) * ((size_t)N + 1));
    for (int64_t k = 0; k < npairs; ++k)
        {
        offsets[pair_i[k] + 1]++;
        offsets[pair_j[k] + 1]++;
        }
    for (uint32_t i = 0; i < N; ++i)
        offsets[i + 1] += offsets[i];
    int64_t* fill = (int64_t*)malloc(sizeof(int64_t) * ((size_t)N + 1));
    memcpy(fill, offsets, sizeof(int64_t) * ((size_t)N + 1));
    for (int64_t k = 0; k < npairs; ++k)
        {
        nbrs[fill[pair_i[k]]++] = pair_j[k];
        nbrs[fill[pair_j[k]]++] = pair_i[k];
        }
    free(fill);
}

/* The same forces over a full (directed) list in CSR form, threaded over
 * particles with OpenMP: particle i sums its own F_i, E_i/2 and W_i over its
 * neighbors (no third law, no shared writes).  This is the CPU path threaded
 * over particle ranges that the CPU baseline reports beside the one-thread
 * half list; per-particle results equal orc_pair_forces up to the order of
 * the FP64 sums. */
void orc_pair_forces_threaded(int evaluator,
                              uint32_t N,
                              const double* pos4,
                              const double* L,
                              const int64_t* offsets,
                              const uint32_t* nbrs,
                              uint32_t ntypes,
                              const double* packed,
                              const double* rcutsq_m,
                              const double* ronsq_m,
                              int shift_mode,
                              double* force4,
                              double* virial,
                              int nthreads)
{
    zero_outputs(N, force4, virial, NULL);
#pragma omp parallel for schedule(dynamic, 256) num_threads(nthreads > 0 ? nthreads : 1)
    for (int64_t ii = 0; ii < (int64_t)N; ++ii)
        {
        const uint32_t i = (uint32_t)ii;
        double f[3] = {0.0, 0.0, 0.0}, en = 0.0, w[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        double d[3];
        for (int64_t k = offsets[i]; k < offsets[i + 1]; ++k)
            {
            double fdivr, e;
            if (!iso_pair(evaluator, pos4, L, i, nbrs[k], ntypes, packed, rcutsq_m, ronsq_m, shift_mode, d, &fdivr,
                          &e))
                continue;
            const double fd2 = fdivr * 0.5;
            for (int a = 0; a < 3; ++a)
                f[a] += d[a] * fdivr;
            en += e * 0.5;
            w[0] += fd2 * d[0] * d[0];
            w[1] += fd2 * d[0] * d[1];
            w[2] += fd2 * d[0] * d[2];
            w[3] += fd2 * d[1] * d[1];
            w[4] += fd2 * d[1] * d[2];
            w[5] += fd2 * d[2] * d[2];
            }
        for (int a = 0; a < 3; ++a)
            force4[4 * (size_t)i + a] = f[a];
        force4[4 * (size_t)i + 3] = en;
        if (virial)
            for (int c = 0; c < 6; ++c)
                virial[c * (size_t)N + i] = w[c];
        }
}

/* Upstream PotentialPairDPDThermo<evaluator>::computeForces (CPU, half list),
 * instantiated at export_PotentialPairDPDThermo.cc.inc:23-24,31-39:
 *   rdotv = dx . (v_i - v_j), seed/tags/timestep passed to the evaluator,
 *   virial from force_divr_cons only (DPDPairEvaluatorGeneralWeight.h:193-194). */
void orc_dpd_forces(uint32_t N,
                    const double* pos4,
                    const double* vel4,
                    const uint32_t* tag,
                    const double* L,
                    int64_t npairs,
                    const uint32_t* pair_i,
                    const uint32_t* pair_j,
                    uint32_t ntypes,
                    const double* packed,
                    const double* rcutsq_m,
                    uint16_t seed,
                    uint32_t timestep,
                    double dt,
                    double kT,
                    int portable_pow,
                    double* force4,
                    double* virial)
{
    zero_outputs(N, force4, virial, NULL);
    const int stride = 8;
    double d[3];
    for (int64_t k = 0; k < npairs; ++k)
        {
        const uint32_t i = pair_i[k], j = pair_j[k];
        const double rsq = pair_rsq(pos4, L, i, j, d);
        const double dv0 = vel4[4 * (size_t)i + 0] - vel4[4 * (size_t)j + 0];
        const double dv1 = vel4[4 * (size_t)i + 1] - vel4[4 * (size_t)j + 1];
        const double dv2 = vel4[4 * (size_t)i + 2] - vel4[4 * (size_t)j + 2];
        const double rdotv = d[0] * dv0 + d[1] * dv1 + d[2] * dv2;
        const unsigned tp = type_of(pos4, i) * ntypes + type_of(pos4, j);
        double fdivr = 0.0, fcons = 0.0, e = 0.0;
        if (!orc_eval_dpd(rsq,
                          rcutsq_m[tp],
                          packed + stride * tp,
                          seed,
                          tag[i],
                          tag[j],
                          timestep,
                          dt,
                          kT,
                          rdotv,
                          portable_pow,
                          &fdivr,
                          &fcons,
                          &e))
            continue;
        const double fd2 = fcons * 0.5;
        for (int a = 0; a < 3; ++a)
            {
            force4[4 * (size_t)i + a] += d[a] * fdivr;
            force4[4 * (size_t)j + a] -= d[a] * fdivr;
            }
        force4[4 * (size_t)i + 3] += e * 0.5;
        force4[4 * (size_t)j + 3] += e * 0.5;
        if (virial)
            {
            add_virial_iso(virial, N, i, fd2, d);
            add_virial_iso(virial, N, j, fd2, d);
            }
        }
}

/* Upstream AnisoPotentialPair<evaluator>::computeForces (CPU, half list),
 * instantiated at export_AnisoPotentialPair.cc.inc:22-26:
 *   F_i += F, tau_i += torque_i, F_j -= F, tau_j += torque_j,
 *   W_ab += dx_b F_a / 2 for (ab) in (xx, xy, xz, yy, yz, zz) on both.
 * Parity: force, energy, torque pinned (test_pair_aniso.py); the off-diagonal
 * virial index order is parity unpinned. */
void orc_aniso_forces(uint32_t N,
                      const double* pos4,
                      const double* quat4,
                      const double* L,
                      int64_t npairs,
                      const uint32_t* pair_i,
                      const uint32_t* pair_j,
                      uint32_t ntypes,
                      const double* packed,
                      const double* rcutsq_m,
                      int shift_mode,
                      double* force4,
                      double* torque4,
                      double* virial)
{
    zero_outputs(N, force4, virial, torque4);
    const int stride = 8;
    double d[3];
    for (int64_t k = 0; k < npairs; ++k)
        {
        const uint32_t i = pair_i[k], j = pair_j[k];
        (void)pair_rsq(pos4, L, i, j, d);
        const unsigned tp = type_of(pos4, i) * ntypes + type_of(pos4, j);
        double F[3], ti_[3], tj_[3], e = 0.0;
        if (!orc_eval_two_patch_morse(d,
                                      quat4 + 4 * (size_t)i,
                                      quat4 + 4 * (size_t)j,
                                      rcutsq_m[tp],
                                      packed + stride * tp,
                                      shift_mode == ORC_SHIFT_SHIFT,
                                      F,
                                      &e,
                                      ti_,
                                      tj_))
            continue;
        for (int a = 0; a < 3; ++a)
            {
            force4[4 * (size_t)i + a] += F[a];
            force4[4 * (size_t)j + a] -= F[a];
            torque4[4 * (size_t)i + a] += ti_[a];
            torque4[4 * (size_t)j + a] += tj_[a];
            }
        force4[4 * (size_t)i + 3] += e * 0.5;
        force4[4 * (size_t)j + 3] += e * 0.5;
        if (virial)
            {
            const double h[3] = {0.5 * F[0], 0.5 * F[1], 0.5 * F[2]};
            const uint32_t ids[2] = {i, j};
            for (int s = 0; s < 2; ++s)
                {
                const uint32_t m = ids[s];
                virial[0 * (size_t)N + m] += d[0] * h[0];
                virial[1 * (size_t)N + m] += d[1] * h[0];
                virial[2 * (size_t)N + m] += d[2] * h[0];
                virial[3 * (size_t)N + m] += d[1] * h[1];
                virial[4 * (size_t)N + m] += d[2] * h[1];
                virial[5 * (size_t)N + m] += d[2] * h[2];
                }
            }
        }
}

int64_t orc_count_pairs_in_cut(uint32_t N,
                               const double* pos4,
                               const double* L,
                               int64_t npairs,
                               const uint32_t* pair_i,
                               const uint32_t* pair_j,
                               uint32_t ntypes,
                               const double* rcutsq_m)
{
    (void)N;
    int64_t n = 0;
    double d[3];
    for (int64_t k = 0; k < npairs; ++k)
        {
        const uint32_t i = pair_i[k], j = pair_j[k];
        const double rsq = pair_rsq(pos4, L, i, j, d);
        if (rsq < rcutsq_m[type_of(pos4, i) * ntypes + type_of(pos4, j)])
            ++n;
        }
    return n;
}

/* Velocity Verlet (upstream TwoStepConstantVolume without a thermostat):
 *   v += dt F / (2m); x += dt v; wrap; rebuild if max displacement > r_buff/2;
 *   F = F(x); v += dt F / (2m). */
int64_t orc_run_nve(int evaluator,
                    uint32_t N,
                    double* pos4,
                    double* vel4,
                    const double* L,
                    uint32_t ntypes,
                    const double* packed,
                    const double* r_cut,
                    double r_buff,
                    int shift_mode,
                    double dt,
                    uint64_t steps,
                    double* force4,
                    double* virial)
{
    const uint32_t nt2 = ntypes * ntypes;
    double* r_listsq = (double*)malloc(sizeof(double) * nt2);
    double* rcutsq = (double*)malloc(sizeof(double) * nt2);
    double* ronsq = (double*)calloc(nt2, sizeof(double));
    for (uint32_t k = 0; k < nt2; ++k)
        {
        const double rl = r_cut[k] + r_buff;
        r_listsq[k] = (r_cut[k] > 0.0) ? rl * rl : 0.0;
        rcutsq[k] = r_cut[k] * r_cut[k];
        }
    double* last = (double*)malloc(sizeof(double) * 4 * (size_t)N);
    int64_t cap = (int64_t)N * 64 + 1024;
    uint32_t* pi = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)cap);
    uint32_t* pj = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)cap);
    int64_t npairs;
    int64_t nbuilds = 0;
    const double trig = 0.25 * r_buff * r_buff;

#define ORC_BUILD()                                                                             \
    do                                                                                          \
        {                                                                                       \
        for (;;)                                                                                \
            {                                                                                   \
            npairs = orc_nlist_cell(N, pos4, L, ntypes, r_listsq, cap, pi, pj);                 \
            if (npairs >= 0)                                                                    \
                break;                                                                          \
            cap = -npairs + 1024;                                                               \
            pi = (uint32_t*)realloc(pi, sizeof(uint32_t) * (size_t)cap);                        \
            pj = (uint32_t*)realloc(pj, sizeof(uint32_t) * (size_t)cap);                        \
            }                                                                                   \
        memcpy(last, pos4, sizeof(double) * 4 * (size_t)N);                                     \
        ++nbuilds;                                                                              \
        } while (0)

    ORC_BUILD();
    orc_pair_forces(evaluator, N, pos4, L, npairs, pi, pj, ntypes, packed, rcutsq, ronsq, shift_mode, force4, virial);
    for (uint64_t s = 0; s < steps; ++s)
        {
        for (uint32_t i = 0; i < N; ++i)
            {
            const double minv = 1.0 / vel4[4 * (size_t)i + 3];
            for (int a = 0; a < 3; ++a)
                {
                vel4[4 * (size_t)i + a] += 0.5 * dt * force4[4 * (size_t)i + a] * minv;
                double x = pos4[4 * (size_t)i + a] + dt * vel4[4 * (size_t)i + a];
                x -= L[a] * rint(x / L[a]);
                pos4[4 * (size_t)i + a] = x;
                }
            }
        double maxsq = 0.0, d[3];
        for (uint32_t i = 0; i < N; ++i)
            {
            for (int a = 0; a < 3; ++a)
                d[a] = pos4[4 * (size_t)i + a] - last[4 * (size_t)i + a];
            orc_min_image(L, d);
            const double dsq = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            if (dsq > maxsq)
                maxsq = dsq;
            }
        if (maxsq > trig)
            ORC_BUILD();
        orc_pair_forces(evaluator, N, pos4, L, npairs, pi, pj, ntypes, packed, rcutsq, ronsq, shift_mode, force4, virial);
        for (uint32_t i = 0; i < N; ++i)
            {
            const double minv = 1.0 / vel4[4 * (size_t)i + 3];
            for (int a = 0; a < 3; ++a)
                vel4[4 * (size_t)i + a] += 0.5 * dt * force4[4 * (size_t)i + a] * minv;
            }
        }
#undef ORC_BUILD
    free(r_listsq);
    free(rcutsq);
    free(ronsq);
    free(last);
    free(pi);
    free(pj);
    return nbuilds;
}

/* ------------------------------------------------------------------------ */
/* wall potentials                                                           */
/* ------------------------------------------------------------------------ */

/* Derived fields from the wall.py schema values (per type):
 *   LJ93 (A, sigma, r_cut, r_extrap): sigma_3 = sigma^3 (WallEvaluatorLJ93.h:33-38)
 *   colloid (A, a, sigma, r_cut, r_extrap): c_1 = A sigma^6 / 7560, c_2 = A / 6
 *   (WallEvaluatorColloid.h:34-42).
 * Energies and f/r: WallEvaluatorLJ93.h:105-131 (lj1 = 2/15 A sigma^9,
 * lj2 = A sigma^3) and WallEvaluatorColloid.h:124-185 (computePotential),
 * both shifted to zero at r_cut as the reference's golden values require
 * (test_wall.py:22-56).  Parity: pinned by those four golden cases,
 * F = -dE/ds by central differences and E(r_cut) = 0 (tests/test_oracle_walls.py). */
static int wall_energy(int kind, const double* v, double rsq, double rcutsq, double* fdivr, double* e)
{
    if (kind == ORC_WALL_LJ93)
        {
        double sigma = v[1], A = v[0];
        double s3 = sigma * sigma * sigma;
        double lj1 = (2.0 / 15.0) * A * s3 * s3 * s3;
        double lj2 = A * s3;
        double r2i, r3i, r6i, c2i, c3i, c6i;
        if (!(rsq < rcutsq && lj1 != 0.0))
            return 0;
        r2i = 1.0 / rsq;
        r3i = r2i * sqrt(r2i);
        r6i = r3i * r3i;
        *fdivr = r2i * r3i * (9.0 * lj1 * r6i - 3.0 * lj2);
        c2i = 1.0 / rcutsq;
        c3i = c2i * sqrt(c2i);
        c6i = c3i * c3i;
        *e = r3i * (lj1 * r6i - lj2) - c3i * (lj1 * c6i - lj2);
        return 1;
        }
    else
        {
        double A = v[0], a = v[1], sigma = v[2];
        double s3 = sigma * sigma * sigma;
        double c1 = A * s3 * s3 / 7560.0, c2 = A / 6.0;
        double r, at;
        int pass;
        if (!(rsq < rcutsq && c1 != 0.0 && a > 0.0))
            return 0;
        *e = 0.0;
        /* pass 0 at r, pass 1 at r_cut (the shift) */
        for (pass = 0; pass < 2; ++pass)
            {
            double x2 = pass == 0 ? rsq : rcutsq;
            double um, up, u2, um2, um6, up2, up6, en;
            r = sqrt(x2);
            at = a / r;
            um = 1.0 / (r - a);
            up = 1.0 / (r + a);
            u2 = um * up;
            um2 = um * um;
            um6 = um2 * um2 * um2;
            up2 = up * up;
            up6 = up2 * up2 * up2;
            if (pass == 0)
                {
                double a8 = 8.0 * at;
                double f = 6.0 * c1 * ((a8 - 1.0) * um2 * um6 + (a8 + 1.0) * up2 * up6);
                f -= c2 * (4.0 * a * a * at * u2 * u2);
                *fdivr = f;
                }
            en = c1 * ((7.0 * a - r) * um * um6 + (7.0 * a + r) * up * up6);
            en -= c2 * (2.0 * a * r * u2 + log(up / um));
            *e += pass == 0 ? en : -en;
            }
        return 1;
        }
}

/* One wall acting on one particle: signed distance s (positive inside),
 * unit normal n into the inside.  Standard evaluation at r = s inside
 * (s > 0, or s == 0 on a closed wall); with r_extrap > 0, every s <= r_extrap
 * takes the linear extrapolation E(r_e) + f(r_e) (r_e - s), F = f(r_e) n —
 * HOOMD's EvaluatorWalls (hoomd/md/EvaluatorWalls.h, not in the reference
 * tree) as documented for hoomd.md.external.wall.  Parity unpinned for
 * r_extrap > 0 and for the virial (1/2 (s n) x F): no reference test covers them. */
static void wall_one(int kind, const double* v, double s, const double* n, int open, double* F, double* E,
                     double* W)
{
    double r_cut = v[kind == ORC_WALL_LJ93 ? 2 : 3];
    double r_e = v[kind == ORC_WALL_LJ93 ? 3 : 4];
    double rcutsq = r_cut * r_cut;
    double fdivr, e, f;
    double d[3];
    int k;
    if (r_e > 0.0 && s <= r_e)
        {
        if (!wall_energy(kind, v, r_e * r_e, rcutsq, &fdivr, &e))
            return;
        f = fdivr * r_e;
        e = e + f * (r_e - s);
        }
    else if (s > 0.0 || (s == 0.0 && !open))
        {
        if (!wall_energy(kind, v, s * s, rcutsq, &fdivr, &e))
            return;
        f = fdivr * s;
        }
    else
        return;
    for (k = 0; k < 3; ++k)
        {
        F[k] += f * n[k];
        d[k] = s * n[k];
        }
    *E += e;
    W[0] += 0.5 * d[0] * (f * n[0]);
    W[1] += 0.5 * d[0] * (f * n[1]);
    W[2] += 0.5 * d[0] * (f * n[2]);
    W[3] += 0.5 * d[1] * (f * n[1]);
    W[4] += 0.5 * d[1] * (f * n[2]);
    W[5] += 0.5 * d[2] * (f * n[2]);
}

/* Geometry (hoomd.wall): sphere {origin, radius, inside, open}: s = R - |x - o|
 * inside (n toward the center), |x - o| - R outside; cylinder {origin, axis,
 * radius, inside, open}: the same with the distance from the axis; plane
 * {origin, normal, open}: s = n . (x - o).  Axes and normals are normalized
 * here.  Positions are used as given (walls do not wrap). */
void orc_wall_forces(int kind,
                     uint32_t N,
                     const double* pos4,
                     uint32_t ntypes,
                     const double* values,
                     uint32_t ns,
                     const double* spheres,
                     uint32_t nc,
                     const double* cylinders,
                     uint32_t np,
                     const double* planes,
                     double* force4,
                     double* virial)
{
    int nv = kind == ORC_WALL_LJ93 ? 4 : 5;
    uint32_t i, k;
    for (i = 0; i < N; ++i)
        {
        const double* x = pos4 + 4 * (size_t)i;
        uint32_t t = (uint32_t)x[3];
        const double* v = values + (size_t)nv * (t < ntypes ? t : 0);
        double F[3] = {0.0, 0.0, 0.0}, E = 0.0, W[6] = {0, 0, 0, 0, 0, 0};
        double n[3];
        int c;
        for (k = 0; k < ns; ++k)
            {
            const double* g = spheres + 6 * (size_t)k;
            double d[3] = {x[0] - g[0], x[1] - g[1], x[2] - g[2]};
            double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
            int inside = g[4] != 0.0;
            double s = inside ? g[3] - r : r - g[3];
            for (c = 0; c < 3; ++c)
                n[c] = r > 0.0 ? (inside ? -d[c] : d[c]) / r : 0.0;
            wall_one(kind, v, s, n, g[5] != 0.0, F, &E, W);
            }
        for (k = 0; k < nc; ++k)
            {
            const double* g = cylinders + 9 * (size_t)k;
            double an = sqrt(g[3] * g[3] + g[4] * g[4] + g[5] * g[5]);
            double ax[3] = {g[3] / an, g[4] / an, g[5] / an};
            double d[3] = {x[0] - g[0], x[1] - g[1], x[2] - g[2]};
            double proj = d[0] * ax[0] + d[1] * ax[1] + d[2] * ax[2];
            double q[3] = {d[0] - proj * ax[0], d[1] - proj * ax[1], d[2] - proj * ax[2]};
            double r = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
            int inside = g[7] != 0.0;
            double s = inside ? g[6] - r : r - g[6];
            for (c = 0; c < 3; ++c)
                n[c] = r > 0.0 ? (inside ? -q[c] : q[c]) / r : 0.0;
            wall_one(kind, v, s, n, g[8] != 0.0, F, &E, W);
            }
        for (k = 0; k < np; ++k)
            {
            const double* g = planes + 7 * (size_t)k;
            double nn = sqrt(g[3] * g[3] + g[4] * g[4] + g[5] * g[5]);
            double s;
            n[0] = g[3] / nn;
            n[1] = g[4] / nn;
            n[2] = g[5] / nn;
            s = (x[0] - g[0]) * n[0] + (x[1] - g[1]) * n[1] + (x[2] - g[2]) * n[2];
            wall_one(kind, v, s, n, g[6] != 0.0, F, &E, W);
            }
        force4[4 * (size_t)i + 0] = F[0];
        force4[4 * (size_t)i + 1] = F[1];
        force4[4 * (size_t)i + 2] = F[2];
        force4[4 * (size_t)i + 3] = E;
        if (virial)
            for (c = 0; c < 6; ++c)
                virial[(size_t)c * N + i] = W[c];
        }
}

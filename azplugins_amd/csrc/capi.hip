// C ABI core: error state, parameter packing, class registry and the pair
// potential objects that stand in for the reference's pybind11 classes
// PotentialPair<Name>GPU, PotentialPairDPDThermo<Name>GPU,
// PotentialPairConservative<Name> and AnisoPotentialPair<Name>GPU
// (module.cc:142-150,171-179; export_*Pair*.cc.inc).
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "azp_internal.h"
#include "step.h"

namespace azp
{

static thread_local std::string g_last_error;

int set_error(int code, const char* fmt, ...)
    {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
    }

void clear_error()
    {
    g_last_error.clear();
    }

int pack_fixed(const double* d_pos, uint32_t N, const azp_box& box, int32_t* d_out, hipStream_t stream);
bool pair_tile_fits(int evaluator, uint32_t ntypes, uint32_t tile_slots, bool mixed);
int nlist_get_on_stream(const azp_nlist* nl,
                        hipStream_t s,
                        bool need_plain,
                        const uint32_t** d_n_neigh,
                        const uint32_t** d_nlist,
                        uint32_t* pitch,
                        uint32_t* max_neighbors);

} // namespace azp

using namespace azp;

extern "C" const char* azp_last_error(void)
    {
    return g_last_error.c_str();
    }

extern "C" int azp_abi_version(void)
    {
    return AZP_ABI_VERSION;
    }

extern "C" int azp_param_nvalues(int evaluator)
    {
    switch (evaluator)
        {
    case AZP_EVAL_HERTZ:
        return 1;
    case AZP_EVAL_EXPANDED_YUKAWA:
    case AZP_EVAL_PERTURBED_LJ:
    case AZP_EVAL_DPD_GENERAL_WEIGHT:
        return 3;
    case AZP_EVAL_COLLOID:
        return 4;
    case AZP_EVAL_TWO_PATCH_MORSE:
        return 6;
    default:
        return -1;
        }
    }

extern "C" int azp_param_size(int evaluator)
    {
    switch (evaluator)
        {
    case AZP_EVAL_HERTZ:
        return (int)sizeof(azp_param_hertz);
    case AZP_EVAL_EXPANDED_YUKAWA:
        return (int)sizeof(azp_param_expanded_yukawa);
    case AZP_EVAL_PERTURBED_LJ:
        return (int)sizeof(azp_param_perturbed_lj);
    case AZP_EVAL_COLLOID:
        return (int)sizeof(azp_param_colloid);
    case AZP_EVAL_DPD_GENERAL_WEIGHT:
        return (int)sizeof(azp_param_dpd_general_weight);
    case AZP_EVAL_TWO_PATCH_MORSE:
        return (int)sizeof(azp_param_two_patch_morse);
    default:
        return -1;
        }
    }

// Dict values (key order of pair.py's TypeParameterDicts) -> stored fields.
extern "C" int azp_param_pack(int evaluator, const double* v, void* pod)
    {
    if (!v || !pod)
        return set_error(AZP_ERR_INVALID, "null argument");
    switch (evaluator)
        {
    case AZP_EVAL_HERTZ: // PairEvaluatorHertz.h:28-31
        {
        azp_param_hertz p;
        p.epsilon = v[0];
        std::memcpy(pod, &p, sizeof(p));
        return AZP_OK;
        }
    case AZP_EVAL_EXPANDED_YUKAWA: // PairEvaluatorExpandedYukawa.h:28-33
        {
        azp_param_expanded_yukawa p;
        p.epsilon = v[0];
        p.kappa = v[1];
        p.delta = v[2];
        std::memcpy(pod, &p, sizeof(p));
        return AZP_OK;
        }
    case AZP_EVAL_PERTURBED_LJ: // PairEvaluatorPerturbedLennardJones.h:35-44
        {
        const double epsilon = v[0];
        const double sigma = v[1];
        const double sigma_2 = sigma * sigma;
        const double sigma_4 = sigma_2 * sigma_2;
        azp_param_perturbed_lj p;
        p.sigma_6 = sigma_2 * sigma_4;
        p.epsilon_x_4 = 4.0 * epsilon;
        p.attraction_scale_factor = v[2];
        p.rwcasq = std::pow(2.0, 1.0 / 3.0) * sigma_2;
        std::memcpy(pod, &p, sizeof(p));
        return AZP_OK;
        }
    case AZP_EVAL_COLLOID: // PairEvaluatorColloid.h:28-35
        {
        azp_param_colloid p;
        p.A = v[0];
        p.a_1 = v[1];
        p.a_2 = v[2];
        const double sigma = v[3];
        p.sigma_3 = sigma * sigma * sigma;
        std::memcpy(pod, &p, sizeof(p));
        return AZP_OK;
        }
    case AZP_EVAL_DPD_GENERAL_WEIGHT: // DPDPairEvaluatorGeneralWeight.h:37-42
        {
        azp_param_dpd_general_weight p;
        p.A = v[0];
        p.gamma = v[1];
        p.s = v[2];
        std::memcpy(pod, &p, sizeof(p));
        return AZP_OK;
        }
    case AZP_EVAL_TWO_PATCH_MORSE: // AnisoPairEvaluatorTwoPatchMorse.h:39-47
        {
        azp_param_two_patch_morse p;
        std::memset(&p, 0, sizeof(p));
        p.M_d = v[0];
        p.M_rinv = 1.0 / v[1];
        p.r_eq = v[2];
        p.omega = v[3];
        p.alpha = v[4];
        p.repulsion = v[5] != 0.0;
        std::memcpy(pod, &p, sizeof(p));
        return AZP_OK;
        }
    default:
        return set_error(AZP_ERR_INVALID, "unknown evaluator %d", evaluator);
        }
    }

// Stored fields -> dict values (asDict / toPython).
extern "C" int azp_param_unpack(int evaluator, const void* pod, double* v)
    {
    if (!v || !pod)
        return set_error(AZP_ERR_INVALID, "null argument");
    switch (evaluator)
        {
    case AZP_EVAL_HERTZ:
        {
        azp_param_hertz p;
        std::memcpy(&p, pod, sizeof(p));
        v[0] = p.epsilon;
        return AZP_OK;
        }
    case AZP_EVAL_EXPANDED_YUKAWA:
        {
        azp_param_expanded_yukawa p;
        std::memcpy(&p, pod, sizeof(p));
        v[0] = p.epsilon;
        v[1] = p.kappa;
        v[2] = p.delta;
        return AZP_OK;
        }
    case AZP_EVAL_PERTURBED_LJ: // PairEvaluatorPerturbedLennardJones.h:47-54
        {
        azp_param_perturbed_lj p;
        std::memcpy(&p, pod, sizeof(p));
        v[0] = p.epsilon_x_4 / 4.0;
        v[1] = std::pow(p.sigma_6, 1.0 / 6.0);
        v[2] = p.attraction_scale_factor;
        return AZP_OK;
        }
    case AZP_EVAL_COLLOID: // PairEvaluatorColloid.h:37-44
        {
        azp_param_colloid p;
        std::memcpy(&p, pod, sizeof(p));
        v[0] = p.A;
        v[1] = p.a_1;
        v[2] = p.a_2;
        v[3] = std::cbrt(p.sigma_3);
        return AZP_OK;
        }
    case AZP_EVAL_DPD_GENERAL_WEIGHT:
        {
        azp_param_dpd_general_weight p;
        std::memcpy(&p, pod, sizeof(p));
        v[0] = p.A;
        v[1] = p.gamma;
        v[2] = p.s;
        return AZP_OK;
        }
    case AZP_EVAL_TWO_PATCH_MORSE: // AnisoPairEvaluatorTwoPatchMorse.h:50-60
        {
        azp_param_two_patch_morse p;
        std::memcpy(&p, pod, sizeof(p));
        v[0] = p.M_d;
        v[1] = 1.0 / p.M_rinv;
        v[2] = p.r_eq;
        v[3] = p.omega;
        v[4] = p.alpha;
        v[5] = p.repulsion ? 1.0 : 0.0;
        return AZP_OK;
        }
    default:
        return set_error(AZP_ERR_INVALID, "unknown evaluator %d", evaluator);
        }
    }

namespace
{
struct ClassEntry
    {
    const char* name;
    int evaluator;
    int driver;
    };
// module.cc:142-150 (CPU names) and :171-179 (GPU names); both resolve to the
// one gfx950 implementation.
const ClassEntry kClasses[] = {
    {"PotentialPairColloid", AZP_EVAL_COLLOID, AZP_DRIVER_PAIR},
    {"PotentialPairExpandedYukawa", AZP_EVAL_EXPANDED_YUKAWA, AZP_DRIVER_PAIR},
    {"PotentialPairHertz", AZP_EVAL_HERTZ, AZP_DRIVER_PAIR},
    {"PotentialPairPerturbedLennardJones", AZP_EVAL_PERTURBED_LJ, AZP_DRIVER_PAIR},
    {"PotentialPairDPDThermoGeneralWeight", AZP_EVAL_DPD_GENERAL_WEIGHT, AZP_DRIVER_DPD_THERMO},
    {"PotentialPairConservativeGeneralWeight", AZP_EVAL_DPD_GENERAL_WEIGHT, AZP_DRIVER_CONSERVATIVE},
    {"AnisoPotentialPairTwoPatchMorse", AZP_EVAL_TWO_PATCH_MORSE, AZP_DRIVER_ANISO},
};
} // namespace

extern "C" int azp_class_lookup(const char* name, int* evaluator, int* driver)
    {
    if (!name)
        return set_error(AZP_ERR_INVALID, "null class name");
    std::string n(name);
    if (n.size() > 3 && n.compare(n.size() - 3, 3, "GPU") == 0)
        n.resize(n.size() - 3);
    for (const ClassEntry& c : kClasses)
        if (n == c.name)
            {
            if (evaluator)
                *evaluator = c.evaluator;
            if (driver)
                *driver = c.driver;
            return AZP_OK;
            }
    return set_error(AZP_ERR_INVALID, "unknown pair class '%s'", name);
    }

struct azp_pair
    {
    int evaluator = -1;
    int driver = -1;
    uint32_t ntypes = 0;
    size_t psize = 0;
    std::vector<unsigned char> params;
    std::vector<char> params_set;
    std::vector<double> r_cut;
    std::vector<double> r_on;
    int shift_mode = AZP_SHIFT_NONE;
    int precision = AZP_PREC_FP64;
    uint16_t seed = 0;
    double dt = 0.0;
    double kT = 0.0;
    bool dirty = true;
    DeviceArray<unsigned char> d_params;
    DeviceArray<double> d_rcutsq;
    DeviceArray<double> d_ronsq;
    DeviceArray<int32_t> d_pos_fixed;
    };

extern "C" int azp_pair_create(azp_pair** out, const char* cpp_class_name, uint32_t ntypes)
    {
    if (!out || ntypes == 0)
        return set_error(AZP_ERR_INVALID, "bad arguments to azp_pair_create");
    int ev, drv;
    AZP_TRY(azp_class_lookup(cpp_class_name, &ev, &drv));
    azp_pair* pp = new azp_pair();
    pp->evaluator = ev;
    pp->driver = drv;
    pp->ntypes = ntypes;
    pp->psize = (size_t)azp_param_size(ev);
    const size_t nt2 = (size_t)ntypes * ntypes;
    pp->params.assign(nt2 * pp->psize, 0);
    pp->params_set.assign(nt2, 0);
    pp->r_cut.assign(nt2, 0.0);
    pp->r_on.assign(nt2, 0.0);
    *out = pp;
    return AZP_OK;
    }

extern "C" int azp_pair_destroy(azp_pair* pp)
    {
    delete pp;
    return AZP_OK;
    }

static int check_types(const azp_pair* pp, uint32_t ti, uint32_t tj)
    {
    if (!pp)
        return set_error(AZP_ERR_INVALID, "null pair object");
    if (ti >= pp->ntypes || tj >= pp->ntypes)
        return set_error(AZP_ERR_INVALID, "type index out of range (%u, %u) for %u types", ti, tj, pp->ntypes);
    return AZP_OK;
    }

extern "C" int azp_pair_set_params(azp_pair* pp, uint32_t ti, uint32_t tj, const double* values)
    {
    AZP_TRY(check_types(pp, ti, tj));
    std::vector<unsigned char> pod(pp->psize);
    AZP_TRY(azp_param_pack(pp->evaluator, values, pod.data()));
    for (int s = 0; s < 2; ++s)
        {
        const size_t k = s == 0 ? (size_t)ti * pp->ntypes + tj : (size_t)tj * pp->ntypes + ti;
        std::memcpy(&pp->params[k * pp->psize], pod.data(), pp->psize);
        pp->params_set[k] = 1;
        }
    pp->dirty = true;
    return AZP_OK;
    }

extern "C" int azp_pair_get_params(const azp_pair* pp, uint32_t ti, uint32_t tj, double* values)
    {
    AZP_TRY(check_types(pp, ti, tj));
    const size_t k = (size_t)ti * pp->ntypes + tj;
    if (!pp->params_set[k])
        return set_error(AZP_ERR_STATE, "params for type pair (%u, %u) are not set", ti, tj);
    return azp_param_unpack(pp->evaluator, &pp->params[k * pp->psize], values);
    }

extern "C" int azp_pair_set_r_cut(azp_pair* pp, uint32_t ti, uint32_t tj, double r_cut)
    {
    AZP_TRY(check_types(pp, ti, tj));
    if (!(r_cut >= 0.0))
        return set_error(AZP_ERR_INVALID, "r_cut must be >= 0");
    pp->r_cut[(size_t)ti * pp->ntypes + tj] = r_cut;
    pp->r_cut[(size_t)tj * pp->ntypes + ti] = r_cut;
    pp->dirty = true;
    return AZP_OK;
    }

extern "C" int azp_pair_set_r_on(azp_pair* pp, uint32_t ti, uint32_t tj, double r_on)
    {
    AZP_TRY(check_types(pp, ti, tj));
    pp->r_on[(size_t)ti * pp->ntypes + tj] = r_on;
    pp->r_on[(size_t)tj * pp->ntypes + ti] = r_on;
    pp->dirty = true;
    return AZP_OK;
    }

extern "C" int azp_pair_get_r_cut(const azp_pair* pp, double* r_cut)
    {
    if (!pp || !r_cut)
        return set_error(AZP_ERR_INVALID, "null argument");
    std::memcpy(r_cut, pp->r_cut.data(), pp->r_cut.size() * sizeof(double));
    return AZP_OK;
    }

extern "C" int azp_pair_set_shift_mode(azp_pair* pp, int mode)
    {
    if (!pp)
        return set_error(AZP_ERR_INVALID, "null pair object");
    if (mode < AZP_SHIFT_NONE || mode > AZP_SHIFT_XPLOR)
        return set_error(AZP_ERR_INVALID, "unknown shift mode %d", mode);
    // pair.py:215 (DPD accepts "none" only); anisotropic pairs accept none/shift
    if (pp->driver == AZP_DRIVER_DPD_THERMO && mode != AZP_SHIFT_NONE)
        return set_error(AZP_ERR_INVALID, "DPDGeneralWeight accepts mode 'none' only");
    if (pp->driver == AZP_DRIVER_ANISO && mode == AZP_SHIFT_XPLOR)
        return set_error(AZP_ERR_INVALID, "anisotropic pairs accept modes 'none' and 'shift'");
    pp->shift_mode = mode;
    return AZP_OK;
    }

extern "C" int azp_pair_set_precision(azp_pair* pp, int precision)
    {
    if (!pp || (precision != AZP_PREC_FP64 && precision != AZP_PREC_MIXED && precision != AZP_PREC_MIXED_ACC64))
        return set_error(AZP_ERR_INVALID, "bad precision");
    if (precision != AZP_PREC_FP64
        && (pp->evaluator == AZP_EVAL_COLLOID || pp->driver == AZP_DRIVER_DPD_THERMO))
        return set_error(AZP_ERR_UNSUPPORTED, "this potential runs in the plain FP64 form only");
    if (precision == AZP_PREC_MIXED_ACC64 && pp->driver == AZP_DRIVER_ANISO)
        return set_error(AZP_ERR_UNSUPPORTED, "the anisotropic pair has the FP64 and mixed forms");
    pp->precision = precision;
    return AZP_OK;
    }

extern "C" int azp_pair_set_thermostat(azp_pair* pp, uint16_t seed, double dt, double kT)
    {
    if (!pp)
        return set_error(AZP_ERR_INVALID, "null pair object");
    pp->seed = seed;
    pp->dt = dt;
    pp->kT = kT;
    return AZP_OK;
    }

static int pair_upload(azp_pair* pp, hipStream_t s)
    {
    const size_t nt2 = (size_t)pp->ntypes * pp->ntypes;
    for (size_t k = 0; k < nt2; ++k)
        if (!pp->params_set[k])
            return set_error(AZP_ERR_STATE,
                             "params for type pair (%zu, %zu) are not set",
                             k / pp->ntypes,
                             k % pp->ntypes);
    std::vector<double> rcsq(nt2), ronsq(nt2);
    for (size_t k = 0; k < nt2; ++k)
        {
        rcsq[k] = pp->r_cut[k] * pp->r_cut[k];
        ronsq[k] = pp->r_on[k] * pp->r_on[k];
        }
    AZP_TRY(pp->d_params.reserve(nt2 * pp->psize));
    AZP_TRY(pp->d_rcutsq.reserve(nt2));
    AZP_TRY(pp->d_ronsq.reserve(nt2));
    AZP_HIP(hipMemcpyAsync(pp->d_params.ptr, pp->params.data(), nt2 * pp->psize, hipMemcpyHostToDevice, s));
    AZP_HIP(hipMemcpyAsync(pp->d_rcutsq.ptr, rcsq.data(), nt2 * sizeof(double), hipMemcpyHostToDevice, s));
    AZP_HIP(hipMemcpyAsync(pp->d_ronsq.ptr, ronsq.data(), nt2 * sizeof(double), hipMemcpyHostToDevice, s));
    AZP_HIP(hipStreamSynchronize(s));
    pp->dirty = false;
    return AZP_OK;
    }

extern "C" int azp_pair_compute(azp_pair* pp,
                                const azp_nlist* nl,
                                const azp_particles* p,
                                const azp_outputs* out,
                                uint64_t timestep,
                                int compute_virial,
                                void* stream)
    {
    if (!pp || !nl || !p || !out)
        return set_error(AZP_ERR_INVALID, "null argument");
    hipStream_t s = (hipStream_t)stream;
    if (pp->dirty)
        AZP_TRY(pair_upload(pp, s));
    azp_pair_args a;
    std::memset(&a, 0, sizeof(a));
    a.particles = *p;
    a.out = *out;
    // the isotropic drivers and DPD run over the tiled list when it is there
    // and their neighbors fit in LDS (16-byte slots in the mixed form, 32-byte
    // FP64 slots in the plain form and for DPD); the anisotropic driver reads
    // the plain list (expanded from the tiled one on first use)
    uint32_t tiled = 0, slots = 0;
    if (pp->driver != AZP_DRIVER_ANISO)
        {
        AZP_TRY(azp_nlist_get_tiles(nl, nullptr, nullptr, nullptr, &tiled, &slots));
        const bool fits = pp->driver == AZP_DRIVER_DPD_THERMO
                              ? dpd_tile_fits(pp->ntypes, slots, tiled)
                              : pair_tile_fits(pp->evaluator, pp->ntypes, slots, pp->precision != AZP_PREC_FP64);
        if (tiled != 0 && !fits)
            tiled = 0;
        }
    AZP_TRY(nlist_get_on_stream(nl, s, tiled == 0, &a.d_n_neigh, &a.d_nlist, &a.nlist_pitch, &a.max_neighbors));
    a.d_rcutsq = pp->d_rcutsq.ptr;
    a.d_ronsq = pp->d_ronsq.ptr;
    a.ntypes = pp->ntypes;
    a.shift_mode = pp->shift_mode;
    a.compute_virial = compute_virial;
    a.precision = pp->precision;
    a.seed = pp->seed;
    a.timestep = timestep;
    a.dt = pp->dt;
    a.kT = pp->kT;
    a.stream = stream;
    // the tiled kernel converts positions itself
    if (pp->precision != AZP_PREC_FP64 && !p->d_pos_fixed && p->N > 0 && tiled == 0)
        {
        AZP_TRY(pp->d_pos_fixed.reserve((size_t)p->N * 4));
        AZP_TRY(pack_fixed(p->d_pos, p->N, p->box, pp->d_pos_fixed.ptr, s));
        a.particles.d_pos_fixed = pp->d_pos_fixed.ptr;
        }
    if (tiled != 0)
        {
        AZP_TRY(azp_nlist_get_tiles(nl, &a.d_tile_count, &a.d_tile_jset, &a.d_tile_list, &a.tile_size,
                                    &a.tile_slots));
        AZP_TRY(azp_nlist_get_tile_order(nl, &a.d_tile_order, &a.d_tile_nn));
        }
    switch (pp->driver)
        {
    case AZP_DRIVER_PAIR:
    case AZP_DRIVER_CONSERVATIVE:
        return azp_gpu_compute_pair_forces(&a, pp->evaluator, pp->d_params.ptr);
    case AZP_DRIVER_DPD_THERMO:
        return azp_gpu_compute_dpd_forces(&a, pp->d_params.ptr);
    case AZP_DRIVER_ANISO:
        return azp_gpu_compute_pair_aniso_forces(&a, pp->evaluator, pp->d_params.ptr);
    default:
        return set_error(AZP_ERR_STATE, "unknown driver");
        }
    }

extern "C" int azp_pair_update_compute(azp_pair* pp,
                                       azp_nlist* nl,
                                       const azp_particles* p,
                                       const azp_outputs* out,
                                       uint64_t timestep,
                                       int compute_virial,
                                       int force_rebuild,
                                       int* rebuilt,
                                       void* stream)
    {
    AZP_TRY(azp_nlist_update(nl, p, force_rebuild, rebuilt, stream));
    return azp_pair_compute(pp, nl, p, out, timestep, compute_virial, stream);
    }

extern "C" int azp_pair_update_compute_step(azp_pair* pp,
                                            azp_nlist* nl,
                                            const azp_particles* p,
                                            const azp_outputs* out,
                                            uint64_t timestep,
                                            int compute_virial,
                                            int force_rebuild,
                                            int* rebuilt,
                                            const azp_step_args* step,
                                            void* stream)
    {
    if (!pp || !nl || !p || !out || !step)
        return set_error(AZP_ERR_INVALID, "null argument");
    if (pp->driver != AZP_DRIVER_PAIR && pp->driver != AZP_DRIVER_CONSERVATIVE)
        return set_error(AZP_ERR_UNSUPPORTED,
                         "the step follows isotropic pair forces only (the DPD and anisotropic drivers have their "
                         "own integration)");
    if (p->N > 0 && (!p->d_pos || !step->d_vel || !out->d_force || (step->langevin && !step->d_tag)))
        return set_error(AZP_ERR_INVALID, "the step needs positions, velocities, forces and (Langevin) tags");
    if (step->d_pos != p->d_pos)
        return set_error(AZP_ERR_INVALID, "step->d_pos must be the positions the forces read (p->d_pos)");
    if (!(step->dt > 0.0))
        return set_error(AZP_ERR_INVALID, "dt must be positive");
    AZP_TRY(nlist_update_checked(nl, p, force_rebuild, rebuilt, stream));
    AZP_TRY(azp_pair_compute(pp, nl, p, out, timestep, compute_virial, stream));
    if (p->N == 0)
        return AZP_OK;
    hipStream_t s = (hipStream_t)stream;
    StepKernelArgs k{};
    k.pos = reinterpret_cast<double4*>(step->d_pos);
    k.vel = reinterpret_cast<double4*>(step->d_vel);
    k.image = reinterpret_cast<int4*>(step->d_image);
    k.tag = step->d_tag;
    AZP_TRY(nlist_step_check(nl, s, &k));
    k.dt = step->dt;
    k.langevin = step->langevin;
    k.gamma = step->gamma;
    k.kT = step->kT;
    k.seed = step->seed;
    k.timestep = step->timestep;
    AZP_TRY(launch_step(reinterpret_cast<const double4*>(out->d_force), p->N, make_box(p->box), k, s));
    nlist_step_checked(nl, p->d_pos);
    return AZP_OK;
    }

/*
 * ORACLE — test infrastructure only.
 *
 * CPU restatement (plain C, FP64, single thread) of azplugins' short-range
 * pair-force path: the isotropic, DPD and anisotropic evaluators and the
 * upstream HOOMD-blue driver semantics (half neighbor list, Newton's third law,
 * energy split 1/2-1/2, per-particle virial), and the wall potentials on
 * HOOMD's wall geometries (their r_extrap extrapolation and virial: parity
 * unpinned).  It is the checker for the HIP
 * product and the "port" CPU baseline.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it.  The product never does.
 *
 * Every function cites the reference file:line it follows.  Parity status per
 * function is stated beside it: "pinned" (by the reference's golden vectors,
 * published known-answer vectors, closed forms or finite differences in
 * tests/) or "parity unpinned".
 */
#ifndef AZP_ORACLE_H
#define AZP_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum
    {
    ORC_HERTZ = 0,
    ORC_EXPANDED_YUKAWA = 1,
    ORC_PERTURBED_LJ = 2,
    ORC_COLLOID = 3,
    ORC_DPD_GENERAL_WEIGHT = 4,
    ORC_TWO_PATCH_MORSE = 5
    };

enum
    {
    ORC_SHIFT_NONE = 0,
    ORC_SHIFT_SHIFT = 1,
    ORC_SHIFT_XPLOR = 2
    };

/* ---- parameter packing: dict values (key order of the reference's Python
 * TypeParameterDict) -> stored param_type fields, and back.  Returns the number
 * of stored doubles (or -1). ---- */
int orc_param_pack(int evaluator, const double* values, double* packed);
int orc_param_unpack(int evaluator, const double* packed, double* values);
int orc_param_nvalues(int evaluator);

/* ---- single-pair evaluators (FP64).  Return 1 if evaluated. ---- */
int orc_eval_iso(int evaluator,
                 double rsq,
                 double rcutsq,
                 const double* packed,
                 int energy_shift,
                 double* force_divr,
                 double* pair_eng);

int orc_eval_dpd(double rsq,
                 double rcutsq,
                 const double* packed,
                 uint16_t seed,
                 uint32_t tag_i,
                 uint32_t tag_j,
                 uint32_t timestep,
                 double dt,
                 double kT,
                 double rdotv,
                 int portable_pow,
                 double* force_divr,
                 double* force_divr_cons,
                 double* pair_eng);

int orc_eval_two_patch_morse(const double* dr,
                             const double* quat_i,
                             const double* quat_j,
                             double rcutsq,
                             const double* packed,
                             int energy_shift,
                             double* force,
                             double* pair_eng,
                             double* torque_i,
                             double* torque_j);

/* ---- Philox4x32-10 and the HOOMD-style DPD random variate ---- */
void orc_philox4x32_10(const uint32_t* ctr, const uint32_t* key, uint32_t* out);
double orc_dpd_alpha(uint16_t seed, uint32_t timestep, uint32_t tag_i, uint32_t tag_j);
void orc_dpd_alpha_many(uint16_t seed,
                        uint32_t timestep,
                        uint64_t n,
                        const uint32_t* tag_i,
                        const uint32_t* tag_j,
                        double* alpha);

/* ---- portable math (shared bit-exact routines) vs libm, for tests ---- */
double orc_ppow(double x, double y);
double orc_pexp(double x);
double orc_plog(double x);

/* ---- box ---- */
void orc_min_image(const double* L, double* d);

/* ---- neighbor lists: half list (each unordered pair once, j > i), CSR ----
 * r_listsq[ti*ntypes+tj] = (r_cut(ti,tj) + r_buff)^2 or 0 to exclude the pair.
 * Returns the number of pairs written, or -(needed) if cap is too small. */
int64_t orc_nlist_brute(uint32_t N,
                        const double* pos4,
                        const double* L,
                        uint32_t ntypes,
                        const double* r_listsq,
                        int64_t cap,
                        uint32_t* pair_i,
                        uint32_t* pair_j);
int64_t orc_nlist_cell(uint32_t N,
                       const double* pos4,
                       const double* L,
                       uint32_t ntypes,
                       const double* r_listsq,
                       int64_t cap,
                       uint32_t* pair_i,
                       uint32_t* pair_j);

/* ---- force drivers over a half pair list (third law) ----
 * Outputs are overwritten: force4[N*4] = (fx, fy, fz, energy),
 * virial[6*N] SoA (xx, xy, xz, yy, yz, zz), torque4[N*4]. */
void orc_pair_forces(int evaluator,
                     uint32_t N,
                     const double* pos4,
                     const double* L,
                     int64_t npairs,
                     const uint32_t* pair_i,
                     const uint32_t* pair_j,
                     uint32_t ntypes,
                     const double* packed_params,
                     const double* rcutsq,
                     const double* ronsq,
                     int shift_mode,
                     double* force4,
                     double* virial);

/* Full (directed) CSR list from a half list (counting sort). */
void orc_full_list(uint32_t N, int64_t npairs, const uint32_t* pair_i, const uint32_t* pair_j, int64_t* offsets,
                   uint32_t* nbrs);

/* Full-list, OpenMP-threaded form of orc_pair_forces (CSR list: neighbors of
 * i are nbrs[offsets[i] .. offsets[i + 1])). */
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
                              int nthreads);

void orc_dpd_forces(uint32_t N,
                    const double* pos4,
                    const double* vel4,
                    const uint32_t* tag,
                    const double* L,
                    int64_t npairs,
                    const uint32_t* pair_i,
                    const uint32_t* pair_j,
                    uint32_t ntypes,
                    const double* packed_params,
                    const double* rcutsq,
                    uint16_t seed,
                    uint32_t timestep,
                    double dt,
                    double kT,
                    int portable_pow,
                    double* force4,
                    double* virial);

void orc_aniso_forces(uint32_t N,
                      const double* pos4,
                      const double* quat4,
                      const double* L,
                      int64_t npairs,
                      const uint32_t* pair_i,
                      const uint32_t* pair_j,
                      uint32_t ntypes,
                      const double* packed_params,
                      const double* rcutsq,
                      int shift_mode,
                      double* force4,
                      double* torque4,
                      double* virial);

/* ---- count of unordered pairs with rsq < rcutsq(ti,tj) ---- */
int64_t orc_count_pairs_in_cut(uint32_t N,
                               const double* pos4,
                               const double* L,
                               int64_t npairs,
                               const uint32_t* pair_i,
                               const uint32_t* pair_j,
                               uint32_t ntypes,
                               const double* rcutsq);

/* ---- NVE velocity Verlet with a Verlet-buffer neighbor list (CPU baseline
 * run of the Hertz config).  Returns number of neighbor-list builds. ---- */
int64_t orc_run_nve(int evaluator,
                    uint32_t N,
                    double* pos4,
                    double* vel4,
                    const double* L,
                    uint32_t ntypes,
                    const double* packed_params,
                    const double* r_cut,
                    double r_buff,
                    int shift_mode,
                    double dt,
                    uint64_t steps,
                    double* force4,
                    double* virial);

#ifdef __cplusplus
}
#endif

/* wall potentials (WallEvaluatorLJ93.h, WallEvaluatorColloid.h on hoomd.wall
 * geometries); values per type in wall.py schema order: LJ93 (A, sigma, r_cut,
 * r_extrap), colloid (A, a, sigma, r_cut, r_extrap).  virial is 6 x N or NULL. */
enum
    {
    ORC_WALL_LJ93 = 0,
    ORC_WALL_COLLOID = 1
    };
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
                     double* virial);

#endif

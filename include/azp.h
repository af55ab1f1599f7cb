/*
 * azp.h — C ABI of the MI355X (gfx950) short-range pair-force library
 * (azplugins_amd/lib/libazp.so).
 *
 * Plain C types only: device pointers, sizes and POD structs.  Every entry
 * point returns AZP_OK (0) or an error code; azp_last_error() returns the
 * message of the last failure on the calling thread.  No exception crosses the
 * boundary.  Device arrays are owned by the caller unless an object (neighbor
 * list, pair potential) says otherwise; the caller's device pointers need only
 * be valid for the duration of the call.  All work is enqueued on the HIP
 * stream passed as `stream` (NULL = the default stream).
 *
 * Particle layout in HBM (structure of arrays, N particles):
 *   pos          double4[N]  x, y, z, type (type index stored as a double)
 *   pos_fixed    int4[N]     x, y, z as wrapped 32-bit fixed point of the box
 *                            (x / L * 2^32), type   (mixed precision only)
 *   vel          double4[N]  vx, vy, vz, mass        (DPD)
 *   tag          uint32[N]   particle tags           (DPD random stream)
 *   orientation  double4[N]  quaternion s, x, y, z   (anisotropic)
 *   force        double4[N]  fx, fy, fz, energy      (output, overwritten)
 *   virial       double[6*pitch] xx, xy, xz, yy, yz, zz, component-major
 *   torque       double4[N]  tx, ty, tz, 0           (anisotropic output)
 *   neighbor list: n_neigh uint32[N]; nlist uint32[max_neighbors * pitch],
 *                  entry k of particle i at nlist[k * pitch + i] (full list);
 *                  each 64-particle group's columns are padded with the
 *                  particle's own index up to its longest list rounded up to 8.
 * The box is orthorhombic and periodic in x, y, z; positions in [-L/2, L/2).
 */
#ifndef AZP_H
#define AZP_H

#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AZP_ABI_VERSION 3

enum azp_status
    {
    AZP_OK = 0,
    AZP_ERR_INVALID = 1,     /* bad argument */
    AZP_ERR_HIP = 2,         /* HIP runtime error */
    AZP_ERR_BOX = 3,         /* box too small for the interaction range */
    AZP_ERR_UNSUPPORTED = 4, /* feature not available for this evaluator */
    AZP_ERR_NOMEM = 5,       /* device allocation failed */
    AZP_ERR_STATE = 6        /* object used before it was configured */
    };

/* Evaluators: the six azplugins functors on the pair path. */
enum azp_evaluator
    {
    AZP_EVAL_HERTZ = 0,             /* PairEvaluatorHertz.h */
    AZP_EVAL_EXPANDED_YUKAWA = 1,   /* PairEvaluatorExpandedYukawa.h */
    AZP_EVAL_PERTURBED_LJ = 2,      /* PairEvaluatorPerturbedLennardJones.h */
    AZP_EVAL_COLLOID = 3,           /* PairEvaluatorColloid.h */
    AZP_EVAL_DPD_GENERAL_WEIGHT = 4, /* DPDPairEvaluatorGeneralWeight.h */
    AZP_EVAL_TWO_PATCH_MORSE = 5    /* AnisoPairEvaluatorTwoPatchMorse.h */
    };

/* Force drivers (the upstream HOOMD templates the evaluators plug into). */
enum azp_driver
    {
    AZP_DRIVER_PAIR = 0,         /* PotentialPair<Eval>[GPU] */
    AZP_DRIVER_DPD_THERMO = 1,   /* PotentialPairDPDThermo<Eval>[GPU] */
    AZP_DRIVER_CONSERVATIVE = 2, /* PotentialPairConservative<Eval> */
    AZP_DRIVER_ANISO = 3         /* AnisoPotentialPair<Eval>[GPU] */
    };

enum azp_shift_mode
    {
    AZP_SHIFT_NONE = 0,
    AZP_SHIFT_SHIFT = 1,
    AZP_SHIFT_XPLOR = 2
    };

enum azp_precision
    {
    AZP_PREC_FP64 = 0,       /* plain form: FP64 everywhere */
    AZP_PREC_MIXED = 1,      /* FP32 pair arithmetic; the terms of 32 list entries are summed in FP32,
                              * each such sum is added to FP64 accumulators */
    AZP_PREC_MIXED_ACC64 = 2 /* FP32 pair arithmetic, every entry's terms added to FP64 accumulators
                              * (isotropic pair launcher only) */
    };

typedef struct
    {
    double Lx, Ly, Lz;
    } azp_box;

/* ---- param_type PODs: same fields, order and alignment as the reference ---- */

/* PairEvaluatorHertz.h:41-47 */
typedef struct
    {
    double epsilon;
    } __attribute__((aligned(8))) azp_param_hertz;

/* PairEvaluatorExpandedYukawa.h:45-53 */
typedef struct
    {
    double epsilon;
    double kappa;
    double delta;
    } __attribute__((aligned(32))) azp_param_expanded_yukawa;

/* PairEvaluatorPerturbedLennardJones.h:57-66 (stored derived fields) */
typedef struct
    {
    double sigma_6;
    double epsilon_x_4;
    double attraction_scale_factor;
    double rwcasq;
    } __attribute__((aligned(32))) azp_param_perturbed_lj;

/* PairEvaluatorColloid.h:48-57 */
typedef struct
    {
    double A;
    double a_1;
    double a_2;
    double sigma_3;
    } __attribute__((aligned(32))) azp_param_colloid;

/* DPDPairEvaluatorGeneralWeight.h:54-62 */
typedef struct
    {
    double A;
    double gamma;
    double s;
    } __attribute__((aligned(32))) azp_param_dpd_general_weight;

/* AnisoPairEvaluatorTwoPatchMorse.h:63-68 */
typedef struct
    {
    double M_d;
    double M_rinv;
    double r_eq;
    double omega;
    double alpha;
    bool repulsion;
    } azp_param_two_patch_morse;

/* ---- particle data and outputs (device pointers) ---- */
typedef struct
    {
    uint32_t N;
    azp_box box;
    const double* d_pos;
    /* optional (mixed form over the plain or head list): when NULL the
     * launchers pack it into library-owned device scratch (one buffer per
     * device, grown on demand, kept until the process ends; launches that
     * overlap on one device share it).  The tiled list does not read it. */
    const int32_t* d_pos_fixed;
    const double* d_vel;
    const uint32_t* d_tag;
    const double* d_orientation;
    } azp_particles;

typedef struct
    {
    double* d_force;
    double* d_virial; /* may be NULL when compute_virial == 0 */
    uint32_t virial_pitch;
    double* d_torque;
    } azp_outputs;

/* ---- kernel launcher arguments (replaces HOOMD's pair_args_t,
 * dpd_pair_args_t and a_pair_args_t) ---- */
typedef struct
    {
    azp_particles particles;
    azp_outputs out;
    const uint32_t* d_n_neigh;
    const uint32_t* d_nlist;
    uint32_t nlist_pitch;
    uint32_t max_neighbors;
    const double* d_rcutsq; /* ntypes^2 */
    const double* d_ronsq;  /* ntypes^2, xplor only (may be NULL otherwise) */
    uint32_t ntypes;
    int32_t shift_mode;
    int32_t compute_virial;
    int32_t precision;
    uint32_t block_size; /* 0: library default / autotuned value */
    /* DPD thermostat */
    uint16_t seed;
    uint64_t timestep;
    double dt;
    double kT;
    void* stream;
    /* optional tiled list (azp_nlist_get_tiles), read by
     * azp_gpu_compute_pair_forces (every form) and azp_gpu_compute_dpd_forces
     * when the tile's neighbors fit in LDS; NULL d_tile_list = use d_nlist
     * (and d_pos_fixed) only.  The tiled mixed kernel converts positions to
     * fixed point itself (d_pos_fixed is not read).  d_tile_order / d_tile_nn
     * (azp_nlist_get_tile_order) for tiles of cell-list slots, NULL for tiles
     * of consecutive particles; d_nlist may be NULL when the tiled list is
     * given and fits. */
    const uint32_t* d_tile_count;
    const uint32_t* d_tile_jset;
    const uint32_t* d_tile_list;
    uint32_t tile_size;
    uint32_t tile_slots;
    const uint32_t* d_tile_order;
    const uint32_t* d_tile_nn;
    /* optional: HOOMD's own list layout (pair_args_t of
     * PotentialPairGPUKernel.cu.inc:25-28).  When non-NULL, particle i's
     * neighbors are d_nlist[d_head_list[i] + k] for k < d_n_neigh[i] (rows
     * unpadded); nlist_pitch is not read and no tiled list may be given.
     * Results equal the column layout's bit for bit (same entries in the
     * same order). */
    const uint64_t* d_head_list;
    } azp_pair_args;

/* ---- library ---- */
const char* azp_last_error(void);
int azp_abi_version(void);

/* ---- parameters: dict values in the reference's Python key order <-> POD ---- */
int azp_param_nvalues(int evaluator);
int azp_param_size(int evaluator);
/* Replaces PairParameters<Name>(pybind11::dict) (e.g. PairEvaluatorPerturbedLennardJones.h:35-44). */
int azp_param_pack(int evaluator, const double* values, void* pod);
/* Replaces param_type::asDict()/toPython() (e.g. PairEvaluatorPerturbedLennardJones.h:47-54). */
int azp_param_unpack(int evaluator, const void* pod, double* values);
/* Registry of the reference's pybind11 class names (module.cc:142-150,171-179),
 * e.g. "PotentialPairPerturbedLennardJonesGPU" -> (AZP_EVAL_PERTURBED_LJ, AZP_DRIVER_PAIR). */
int azp_class_lookup(const char* cpp_class_name, int* evaluator, int* driver);

/* ---- kernel launchers: replace the explicit template instantiations
 * gpu_compute_pair_forces<Eval>        (PotentialPairGPUKernel.cu.inc:25-28),
 * gpu_compute_dpd_forces<Eval>         (PotentialPairDPDThermoGPUKernel.cu.inc:21-24),
 * gpu_compute_pair_aniso_forces<Eval>  (AnisoPotentialPairGPUKernel.cu.inc:21-25).
 * d_params points to ntypes^2 PODs of the evaluator's param type in device memory. */
int azp_gpu_compute_pair_forces(const azp_pair_args* args, int evaluator, const void* d_params);
/* No counterpart in the reference.  The same launch as
 * azp_gpu_compute_pair_forces without any device allocation, for callers
 * that cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Where the launcher must pack
 * fixed-point positions itself (mixed form over the plain or head list with
 * d_pos_fixed NULL) it uses a file-scope device array of 2^20 int4 (16 MiB
 * per device, held until the process ends) and rejects N > 2^20 with
 * AZP_ERR_INVALID; the other forms use no scratch.  Every call on a device
 * shares that array: calls that overlap on one device race. */
int azp_gpu_compute_pair_forces_noalloc(const azp_pair_args* args, int evaluator, const void* d_params);
/* PotentialPairDPDThermoGPU's launch (PotentialPairDPDThermoGPUKernel.cu.inc:21-24):
 * FP64 only.  Over the tiled list when given and its neighbor positions (32
 * bytes a neighbor) fit in LDS, else over d_nlist (column or head-list
 * layout); the results are the same bit for bit. */
int azp_gpu_compute_dpd_forces(const azp_pair_args* args, const void* d_params);
int azp_gpu_compute_pair_aniso_forces(const azp_pair_args* args, int evaluator, const void* d_params);
/* pos (double4) -> pos_fixed (int4) for the mixed-precision form */
int azp_gpu_pack_fixed(const double* d_pos, uint32_t N, const azp_box* box, int32_t* d_pos_fixed, void* stream);
/* DPD random variates alpha_k = U(-1,1) for (tag_i[k], tag_j[k]) at (seed, timestep). */
int azp_gpu_dpd_alpha(uint16_t seed,
                      uint64_t timestep,
                      uint64_t n,
                      const uint32_t* d_tag_i,
                      const uint32_t* d_tag_j,
                      double* d_alpha,
                      void* stream);

/* ---- adjoint of the isotropic pair forces (no counterpart in the reference:
 * the force-matching extension of the path).  For ExpandedYukawa
 * (theta = epsilon, kappa, delta) and PerturbedLennardJones (theta = epsilon,
 * sigma, attraction_scale_factor), modes none/shift, 1..4 types.
 * d_grad:        double4[N] (dL/dF_x, dL/dF_y, dL/dF_z, dL/dE_i)
 * d_grad_pos:    double4[N] dL/dr (x, y, z, 0)
 * d_grad_params: double[ntypes^2 * 3], dL/dtheta per ordered type pair (t_i, t_j)
 *                (the gradient of a symmetric parameter is the (a,b) plus (b,a) entries)
 * d_work:        azp_adjoint_workspace_size(N, ntypes) doubles.
 * args->precision selects FP64 or FP32 pair arithmetic (FP64 accumulation). */
size_t azp_adjoint_workspace_size(uint32_t N, uint32_t ntypes);
int azp_gpu_pair_adjoint(const azp_pair_args* args,
                         int evaluator,
                         const void* d_params,
                         const double* d_grad,
                         double* d_grad_pos,
                         double* d_grad_params,
                         double* d_work,
                         size_t work_len);

/* ---- integration around the force path (replace the upstream
 * TwoStepConstantVolume / TwoStepLangevin kernels reached from the reference
 * tests, test_pair.py:325-326, test_pair_dpd.py:189-200) ----
 * accel: double4[N] (ax, ay, az, 0), image: int4[N] or NULL. */
int azp_gpu_integrate_vv_one(double* d_pos,
                             double* d_vel,
                             const double* d_accel,
                             int32_t* d_image,
                             uint32_t N,
                             const azp_box* box,
                             double dt,
                             void* stream);
/* a = (F [- gamma v + sqrt(6 gamma kT / dt) U(-1,1)]) / m;  v += dt a / 2 */
int azp_gpu_integrate_vv_two(double* d_vel,
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
                             void* stream);
/* azp_gpu_integrate_vv_two at `timestep` followed by azp_gpu_integrate_vv_one
 * in one pass, for the steps where nothing runs between the two (this
 * library's own addition): the same FP64 operations in the same order, so
 * the trajectory is bit-identical; the acceleration is not stored. */
int azp_gpu_integrate_vv_two_one(double* d_pos,
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
                                 void* stream);
/* spherical tops: angmom double4[N] space-frame L, inertia double[N] (isotropic I).
 * rot_one: L += dt tau / 2, then rotate q by omega dt (omega = L / I).
 * rot_two: L += dt tau / 2 [+ dt (-gamma_r L / I + sqrt(6 gamma_r kT / dt) U(-1,1))] */
int azp_gpu_integrate_rot_one(double* d_orientation,
                              double* d_angmom,
                              const double* d_torque,
                              const double* d_inertia,
                              uint32_t N,
                              double dt,
                              void* stream);
int azp_gpu_integrate_rot_two(double* d_angmom,
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
                              void* stream);
/* FP64 sums (deterministic order): out[10] = translational KE, rotational KE,
 * potential energy, virial xx xy xz yy yz zz, N.  d_work holds
 * azp_thermo_workspace_size() doubles. */
int azp_thermo_workspace_size(void);
int azp_gpu_thermo_reduce(const double* d_vel,
                          const double* d_force,
                          const double* d_virial,
                          uint32_t virial_pitch,
                          const double* d_angmom,
                          const double* d_inertia,
                          uint32_t N,
                          double* d_work,
                          double* d_out,
                          void* stream);
/* Kinetic tensor for the pressure tensor (upstream ComputeThermo's
 * Σ m v_a v_b): out[6] = xx xy xz yy yz zz, FP64 in a deterministic order;
 * d_work as for azp_gpu_thermo_reduce. */
int azp_gpu_kinetic_tensor(const double* d_vel, uint32_t N, double* d_work, double* d_out, void* stream);

/* ---- neighbor list (replaces upstream CellListGPU + NeighborListGPUBinned;
 * reached from the pair classes through `nlist`, pair.py:110-111) ---- */
typedef struct azp_nlist azp_nlist;
int azp_nlist_create(azp_nlist** out);
int azp_nlist_destroy(azp_nlist* nl);
int azp_nlist_set_r_buff(azp_nlist* nl, double r_buff);
/* ntypes^2 host array, r_cut <= 0 excludes the type pair */
int azp_nlist_set_r_cut(azp_nlist* nl, uint32_t ntypes, const double* r_cut);
/* Rebuild when forced, when N/box/r_cut changed, or when a particle moved more
 * than r_buff/2 since the last build; grows capacity and retries on overflow. */
int azp_nlist_update(azp_nlist* nl, const azp_particles* p, int force_rebuild, int* rebuilt, void* stream);
int azp_nlist_get(const azp_nlist* nl,
                  const uint32_t** d_n_neigh,
                  const uint32_t** d_nlist,
                  uint32_t* pitch,
                  uint32_t* max_neighbors);
int azp_nlist_num_builds(const azp_nlist* nl, uint64_t* n);
/* Tiled form of the list for the mixed-precision pair kernel (this
 * library's own addition; no reference counterpart): the full list of each
 * block of 512 (or 256) consecutive particles rewritten as 16-bit indices
 * into the block's set of distinct neighbors.  Built with every rebuild
 * once enabled.  mode: 0 off, 1 blocks of 512 else 256, 256 or 512 that
 * size only.  tile_size = 0 when some block has more than 3967 distinct
 * neighbors (the plain list is then used); tile_slots = LDS slots the pair
 * kernel stages (largest distinct count + 1).  While tiles are enabled the
 * build walks the cell list once per block (the block's candidates staged
 * in LDS) and writes only the tiled list; azp_nlist_get / _download then
 * expand the plain list from it on first use.  Blocks whose surrounding
 * cells do not fit in LDS make that build take the plain walk instead. */
int azp_nlist_set_tiled(azp_nlist* nl, int mode);
int azp_nlist_get_tiles(const azp_nlist* nl,
                        const uint32_t** d_tile_count,
                        const uint32_t** d_tile_jset,
                        const uint32_t** d_tile_list,
                        uint32_t* tile_size,
                        uint32_t* tile_slots);
/* Column order of the tiled list.  Tiles from the per-block walk are runs
 * of consecutive cell-list slots (cells in Morton order, so each tile is a
 * compact block of space for any particle order): column s belongs to
 * particle d_tile_order[s] and has d_tile_nn[s] entries.  Both are NULL for
 * tiles built from the plain list, whose column i is particle i with
 * n_neigh[i] entries. */
int azp_nlist_get_tile_order(const azp_nlist* nl, const uint32_t** d_tile_order, const uint32_t** d_tile_nn);
/* How the last build made the list: 1 the per-block walk that writes the
 * tiled list directly, 0 the plain walk. */
int azp_nlist_walk_kind(const azp_nlist* nl, int* kind);
/* Diagnostics of the last per-block walk: its number of blocks, and how many
 * of them tested candidates as FP32 offsets from a block anchor (blocks whose
 * cells lie within L/4 and 8 r_list of the anchor) rather than as fixed-point
 * differences; both 0 after a plain walk. */
int azp_nlist_tile_stats(const azp_nlist* nl, uint32_t* ntiles, uint32_t* relative_tiles);
/* Copy n_neigh[N] and the list (max_neighbors * pitch entries) to host memory
 * (synchronizes the device). */
int azp_nlist_download(const azp_nlist* nl,
                       uint32_t N,
                       uint32_t* h_n_neigh,
                       uint32_t* h_nlist,
                       uint64_t nlist_capacity);

/* ---- particle sort (HOOMD's SFCPackTuner, the particle sorter every
 * simulation of the reference runs [U: hoomd/SFCPackTuner.h, not in the
 * reference tree]; azplugins itself has no counterpart).  The sorter bins
 * the particles on a grid of about `grid` bins along the box's longest
 * axis (coarsened to at most ~4 bins per particle) and writes d_order[k],
 * the particle that moves to index k: bins in Morton order, each bin's
 * members by current index.  d_order is a permutation of [0, N).
 * azp_gather_rows then applies it to any per-particle array:
 * dst[k] = src[order[k]] for rows of row_bytes (4, 8, 16 or 32); src and
 * dst must not overlap.  Enqueued on `stream`; the sorter's buffers are
 * reused between calls. ---- */
typedef struct azp_sorter azp_sorter;
int azp_sorter_create(azp_sorter** out);
int azp_sorter_destroy(azp_sorter* so);
int azp_sorter_order(azp_sorter* so, const azp_particles* p, uint32_t grid, uint32_t* d_order, void* stream);
int azp_gather_rows(const void* d_src, void* d_dst, const uint32_t* d_order, uint32_t N, uint32_t row_bytes,
                    void* stream);

/* ---- wall potentials (replace the reference's WallsPotentialLJ93[GPU] and
 * WallsPotentialColloid[GPU], export_PotentialExternalWall.cc.inc and
 * PotentialExternalWallGPUKernel.cu.inc:26-27, i.e. HOOMD's
 * PotentialExternal<EvaluatorWalls<WallEvaluator*>> with the evaluators of
 * WallEvaluatorLJ93.h / WallEvaluatorColloid.h).  cpp_class_name is
 * "WallsPotentialLJ93" or "WallsPotentialColloid" (a "GPU" suffix is
 * accepted).  Parameters per particle type in the schema order of the
 * reference's wall.py: LJ93 (A, sigma, r_cut, r_extrap), colloid (A, a,
 * sigma, r_cut, r_extrap); stored as the reference's derived fields and
 * read back through its asDict formulas.  Geometry rows (host arrays):
 * sphere {origin x y z, radius, inside, open}, cylinder {origin x y z,
 * axis x y z, radius, inside, open}, plane {origin x y z, normal x y z,
 * open}; inside / open as 0 or 1 (HOOMD's hoomd.wall attributes).  compute
 * writes force.xyz, the energy (shifted to zero at r_cut, all of it on the
 * particle) in force.w and, when asked, the virial 1/2 (s n) x F.  It does
 * not wait for the stream: a particle type outside [0, ntypes) raises a
 * sticky device flag, reported by a later compute once its copy has reached
 * the host, or by azp_wall_check, which waits for the stream (callers run it
 * at the end of a run so that an error of the last step is not lost).  The
 * flag is cleared when an error is reported. ---- */
typedef struct azp_wall azp_wall;
int azp_wall_create(azp_wall** out, const char* cpp_class_name, uint32_t ntypes);
int azp_wall_destroy(azp_wall* w);
int azp_wall_param_nvalues(const azp_wall* w);
int azp_wall_set_params(azp_wall* w, uint32_t type, const double* values);
int azp_wall_get_params(const azp_wall* w, uint32_t type, double* values);
int azp_wall_set_geometry(azp_wall* w,
                          uint32_t n_spheres,
                          const double* spheres,
                          uint32_t n_cylinders,
                          const double* cylinders,
                          uint32_t n_planes,
                          const double* planes);
int azp_wall_compute(azp_wall* w, const azp_particles* p, const azp_outputs* out, int compute_virial, void* stream);
int azp_wall_check(azp_wall* w, void* stream);

/* ---- pair potential objects (replace PotentialPair<Name>GPU,
 * PotentialPairDPDThermo<Name>GPU, AnisoPotentialPair<Name>GPU) ---- */
typedef struct azp_pair azp_pair;
int azp_pair_create(azp_pair** out, const char* cpp_class_name, uint32_t ntypes);
int azp_pair_destroy(azp_pair* pp);
int azp_pair_set_params(azp_pair* pp, uint32_t ti, uint32_t tj, const double* values);
int azp_pair_get_params(const azp_pair* pp, uint32_t ti, uint32_t tj, double* values);
int azp_pair_set_r_cut(azp_pair* pp, uint32_t ti, uint32_t tj, double r_cut);
int azp_pair_set_r_on(azp_pair* pp, uint32_t ti, uint32_t tj, double r_on);
int azp_pair_set_shift_mode(azp_pair* pp, int mode);
int azp_pair_set_precision(azp_pair* pp, int precision);
int azp_pair_set_thermostat(azp_pair* pp, uint16_t seed, double dt, double kT);
/* r_cut matrix (ntypes^2) the neighbor list must cover */
int azp_pair_get_r_cut(const azp_pair* pp, double* r_cut);
int azp_pair_compute(azp_pair* pp,
                     const azp_nlist* nl,
                     const azp_particles* p,
                     const azp_outputs* out,
                     uint64_t timestep,
                     int compute_virial,
                     void* stream);

/* azp_nlist_update followed by azp_pair_compute in one call (this library's
 * own addition): the force launches as soon as the rebuild check's flag
 * arrives, without a return to the caller in between. */
int azp_pair_update_compute(azp_pair* pp,
                            azp_nlist* nl,
                            const azp_particles* p,
                            const azp_outputs* out,
                            uint64_t timestep,
                            int compute_virial,
                            int force_rebuild,
                            int* rebuilt,
                            void* stream);

/* No counterpart in the reference, whose integrators run after the force
 * compute (TwoStepConstantVolume / TwoStepLangevin around PotentialPair).
 * azp_pair_update_compute followed by a caller's MD step in one call, for
 * drivers PAIR and CONSERVATIVE: once the forces are stored, one kernel
 * takes the second half step of `timestep` and the first half step of the
 * next (velocity Verlet; Langevin drag and uniform noise keyed by seed,
 * timestep and tag), updating positions, velocities and images in place
 * with the FP64 operations of azp_gpu_integrate_vv_two_one (bit-identical
 * trajectories), and makes the list's rebuild check of the new positions.
 * The next azp_pair_update_compute_step with the same position array reads
 * that check instead of launching its own; any other list update discards
 * it, and the caller must not write the positions in between except through
 * an update that forces a rebuild. */
typedef struct
    {
    double* d_pos;         /* the array p->d_pos points to (updated) */
    double* d_vel;         /* double4[N] vx, vy, vz, mass (updated) */
    int32_t* d_image;      /* int4[N] image counts (updated; may be NULL) */
    const uint32_t* d_tag; /* tags (Langevin noise; may be NULL for velocity Verlet) */
    double dt;
    int langevin;          /* 0: velocity Verlet, 1: Langevin */
    double gamma;          /* Langevin drag */
    double kT;
    uint16_t seed;
    uint64_t timestep;     /* the step whose second half is taken */
    } azp_step_args;
int azp_pair_update_compute_step(azp_pair* pp,
                                 azp_nlist* nl,
                                 const azp_particles* p,
                                 const azp_outputs* out,
                                 uint64_t timestep,
                                 int compute_virial,
                                 int force_rebuild,
                                 int* rebuilt,
                                 const azp_step_args* step,
                                 void* stream);

/* ---- batched replicas (no single reference counterpart: R independent
 * AnisoPotentialPair<AnisoPairEvaluatorTwoPatchMorse> systems,
 * AnisoPotentialPairGPUKernel.cu.inc:21-25, each with its own periodic box,
 * size and Langevin kT, advanced together) ----
 * Particles of replica r occupy [offset_r, offset_r + size_r) of every array,
 * offsets being the prefix sums of the sizes; positions are in the replica's
 * own box frame.  The arrays stay owned by the caller and must outlive the
 * binding: captured graphs hold their addresses. */
typedef struct azp_batch azp_batch;
typedef struct
    {
    double* d_pos;         /* double4[N_total] x, y, z, type */
    int32_t* d_pos_fixed;  /* int4[N_total], mixed precision (written by the library) */
    double* d_orientation; /* double4[N_total] quaternion s, x, y, z */
    double* d_vel;         /* double4[N_total] vx, vy, vz, mass      (time steps) */
    double* d_angmom;      /* double4[N_total] angular momentum      (time steps) */
    double* d_accel;       /* double4[N_total] acceleration          (time steps) */
    double* d_force;       /* double4[N_total] out: F, energy */
    double* d_torque;      /* double4[N_total] out: torque */
    double* d_virial;      /* double[6 * virial_pitch] out (may be NULL without virial) */
    uint32_t virial_pitch;
    } azp_batch_arrays;
/* Launch configuration of the persistent force kernel and of the walk. */
typedef struct
    {
    uint32_t block_size;        /* force kernel workgroup: 64..512, multiple of 64 */
    uint32_t grab;              /* 64-particle chunks taken per work-queue fetch */
    uint32_t workgroups_per_cu; /* persistent grid = CUs x this; 0 = one wave per chunk */
    uint32_t walk_block;        /* neighbor walk workgroup: 64, 128 or 256 */
    double force_us;            /* force-kernel time measured by the autotuner */
    } azp_batch_tuning;
typedef struct
    {
    uint64_t timestep;
    uint64_t n_builds;       /* list builds (rebuild events) */
    uint64_t replica_builds; /* replica lists rebuilt, summed over builds */
    uint64_t graph_launches;
    uint32_t max_neighbors;
    uint32_t n_cells;
    } azp_batch_stats_t;
/* Host-side validation only; device memory is allocated on first use. */
int azp_batch_create(azp_batch** out,
                     uint32_t n_replicas,
                     const uint32_t* sizes,
                     const azp_box* boxes,
                     uint32_t ntypes,
                     int evaluator);
int azp_batch_destroy(azp_batch* b);
int azp_batch_size(const azp_batch* b, uint32_t* n_replicas, uint32_t* n_total);
int azp_batch_set_params(azp_batch* b, uint32_t ti, uint32_t tj, const double* values);
int azp_batch_set_r_cut(azp_batch* b, uint32_t ti, uint32_t tj, double r_cut);
int azp_batch_set_r_buff(azp_batch* b, double r_buff);
int azp_batch_set_mode(azp_batch* b, int shift_mode, int precision, int compute_virial);
/* kT: host array, one value per replica */
int azp_batch_set_langevin(azp_batch* b,
                           const double* kT,
                           double dt,
                           double gamma,
                           double gamma_r,
                           double inertia,
                           uint16_t seed);
int azp_batch_bind(azp_batch* b, const azp_batch_arrays* arrays);
int azp_batch_set_timestep(azp_batch* b, uint64_t timestep, void* stream);
/* Forces, torques, energies and virials of every replica in one persistent
 * launch (the neighbor list is rebuilt when forced or when a particle moved
 * more than r_buff / 2). */
int azp_batch_compute(azp_batch* b, int force_rebuild, void* stream);
/* nsteps Langevin steps (translation + rotation); use_graph replays captured
 * HIP graphs of 10 steps (and of 1 step for the remainder). */
int azp_batch_run(azp_batch* b, uint32_t nsteps, int use_graph, void* stream);
/* d_out: double[4 * R]: translational KE, rotational KE, energy, virial trace */
int azp_batch_thermo(azp_batch* b, double* d_out, void* stream);
/* Times the launch candidates on the live batch and keeps the fastest. */
int azp_batch_autotune(azp_batch* b, azp_batch_tuning* out, void* stream);
int azp_batch_get_tuning(const azp_batch* b, azp_batch_tuning* out);
int azp_batch_set_tuning(azp_batch* b, const azp_batch_tuning* tuning);
int azp_batch_stats(azp_batch* b, azp_batch_stats_t* out);
int azp_batch_nlist_download(azp_batch* b,
                             uint32_t* h_n_neigh,
                             uint32_t* h_nlist,
                             uint64_t capacity,
                             uint32_t* pitch,
                             uint32_t* max_neighbors);

#ifdef __cplusplus
}
#endif

#endif /* AZP_H */

/*
 * exagoop_amd.h — C ABI of the MI355X-native explicit MPM step.
 *
 * Drop-in boundary for ExaGOOP's MPMParticleContainer hot path.  Every entry
 * point below replaces one reference method or free function (cited per
 * declaration, paths relative to the ExaGOOP source tree).  Plain C types
 * only: pointers, sizes and status codes.  All entry points return
 * MPM_OK (0) or a negative MPM_E* code instead of aborting the process
 * (the reference calls amrex::Abort on the same conditions).
 *
 * Threading: one host thread per context; work is enqueued on the context's
 * HIP stream and is stream-ordered.  Host buffers passed in are borrowed only
 * for the duration of the call.  Device state is owned by the context.
 */
#ifndef EXAGOOP_AMD_H_
#define EXAGOOP_AMD_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* Status codes                                                             */
/* ------------------------------------------------------------------------ */
#define MPM_OK 0
#define MPM_E_ARG -1        /* bad argument / null pointer / size mismatch */
#define MPM_E_BC -2         /* unknown boundary condition (mpm_kernels.H:107-110) */
#define MPM_E_ORDER -3      /* unsupported order_scheme (mpm_particle_grid_ops.cpp:293-296) */
#define MPM_E_MODEL -4      /* invalid constitutive model (mpm_particle_timestep.cpp:76-80) */
#define MPM_E_OOM -5        /* device allocation failed */
#define MPM_E_HIP -6        /* HIP runtime error */
#define MPM_E_NAN -7        /* non-finite state detected */
#define MPM_E_UNSUPPORTED -8 /* combination the reference aborts on (e.g. delta model for fluids) */
#define MPM_E_NODEVICE -9   /* no HIP device */

/* ------------------------------------------------------------------------ */
/* Particle record layout (mpm_specs.H:17-73, 3-D, USE_TEMP off)            */
/* Exchange format is component-major: field f of particle i at [f*np + i]. */
/* ------------------------------------------------------------------------ */
#define MPM_NREAL 46
#define MPM_NINT 3

enum mpm_real_field {
    MPM_R_RADIUS = 0,
    MPM_R_VEL = 1,        /* 1..3  */
    MPM_R_VELP = 4,       /* 4..6  grid velocity v' used for advection */
    MPM_R_STRAINRATE = 7, /* 7..12 packed XX,XY,XZ,YY,YZ,ZZ */
    MPM_R_STRAIN = 13,    /* 13..18 */
    MPM_R_STRESS = 19,    /* 19..24 */
    MPM_R_DEFGRAD = 25,   /* 25..33 row-major 3x3 */
    MPM_R_VOLUME = 34,
    MPM_R_MASS = 35,
    MPM_R_DENSITY = 36,
    MPM_R_JACOBIAN = 37,
    MPM_R_PRESSURE = 38,
    MPM_R_VOLINIT = 39,
    MPM_R_E = 40,
    MPM_R_NU = 41,
    MPM_R_BULK = 42,
    MPM_R_GAMMA = 43,
    MPM_R_VISC = 44,
    MPM_R_YACC = 45
};

enum mpm_int_field { MPM_I_PHASE = 0, MPM_I_RIGID_ID = 1, MPM_I_MODEL = 2 };

/* Nodal component indices (constants.H:36-55). */
enum mpm_nodal_comp {
    MPM_N_MASS = 0,
    MPM_N_VEL = 1,   /* 1..3 */
    MPM_N_FRC = 4,   /* 4..6 */
    MPM_N_DVEL = 7,  /* 7..9 */
    MPM_N_MASS_OLD = 10,
    MPM_N_NCOMP = 20
};

/* Boundary condition codes (constants.H:89-93). */
enum mpm_bc {
    MPM_BC_PERIODIC = 0,
    MPM_BC_NOSLIP = 1,
    MPM_BC_SLIP = 2,
    MPM_BC_PARTIALSLIP = 3,
    MPM_BC_OUTFLOW = 4
};

enum mpm_scheme { MPM_USL = 0, MPM_MUSL = 1 };

/* ------------------------------------------------------------------------ */
/* Problem specification: the subset of MPMspecs (mpm_specs.H:75-220) the    */
/* step reads.  mpm_specs_default() fills the reference defaults.           */
/* ------------------------------------------------------------------------ */
typedef struct mpm_specs {
    double plo[3];
    double phi[3];
    int ncells[3];
    int periodic[3];
    int order_scheme; /* 1 linear, 2 quadratic B-spline, 3 cubic B-spline */
    int num_redist;   /* particle redistribution interval (utilities.cpp:420-434) */
    double gravity[3];
    int external_loads;
    double force_slab_lo[3];
    double force_slab_hi[3];
    double extforce[3];
    double mass_tolerance;
    double alpha_pic_flip;
    int stress_update_scheme; /* MPM_USL / MPM_MUSL */
    int fixed_timestep;
    double timestep;
    double CFL;
    double dt_min_limit;
    double dt_max_limit;
    double applied_strainrate;
    double applied_strainrate_time;
    int calculate_strain_based_on_delta;
    int bclo[3];
    int bchi[3];
    double wall_mu_lo[3];
    double wall_mu_hi[3];
    double wall_vel_lo[9]; /* [dir*3 + comp] */
    double wall_vel_hi[9];
} mpm_specs;

/* Wall-velocity user function, same signature as WallVelUDFFuncPtr
 * (mpm_specs.H:12): f(x, y, z, t, vel[3]). */
typedef void (*mpm_wall_udf_fn)(double, double, double, double, double *);

typedef struct mpm_ctx mpm_ctx;

/* Reference defaults of MPMspecs (mpm_specs.H:78-220). */
void mpm_specs_default(mpm_specs *s);

/* Per-direction order after the small-grid fallback (mpm_init.cpp:180-233):
 * writes order_scheme_directional[3]; returns MPM_E_ORDER for order not in
 * {1,2,3}. */
int mpm_order_directional(const mpm_specs *s, int out_order[3]);

/* ------------------------------------------------------------------------ */
/* Context lifetime                                                         */
/* ------------------------------------------------------------------------ */

/* Replaces Initialise_Domain (mpm_init.cpp:108-299) + the
 * MPMParticleContainer constructor (main.cpp:59): validates specs, picks
 * per-direction orders, allocates nodal storage on the current device. */
int mpm_create(const mpm_specs *specs, mpm_ctx **out);
void mpm_destroy(mpm_ctx *ctx);

/* Replaces InitParticles' hand-off to the container (mpm_init.cpp:942-1167,
 * the particle tile fill + Redistribute).  pos is [3*np], rdata
 * [MPM_NREAL*np], idata [MPM_NINT*np], all component-major. */
int mpm_upload_particles(mpm_ctx *ctx, int64_t np, const double *pos,
                         const double *rdata, const int *idata);
/* Particle input formats.  pos, rdata, idata as in mpm_upload_particles; call
 * once with pos = NULL to get *np, then with arrays of that size.  totals
 * (may be NULL) receives {total mass, total volume} of the phase-0 particles.
 *
 * InitParticles(filename) (mpm_init.cpp:942-1167): the ASCII particle file
 * ("dim: 3", "number_of_material_points: N", a '#' header line, then per
 * particle: phase x y z radius density vx vy vz cm_id, then E nu for models
 * 0 and 2 or Bulk_modulus Gamma_pressure Dynamic_viscosity for model 1);
 * volume = 4/3 pi r^3 with the reference's truncated constants, mass =
 * density volume.  Every condition the reference aborts on is an error. */
int mpm_read_particle_file(const char *path, int64_t *np, double *pos, double *rdata,
                           int *idata, double totals[2]);
/* The same, uploaded into the context (InitParticles + its Redistribute). */
int mpm_load_particle_file(mpm_ctx *ctx, const char *path, double totals[2]);
/* The USE_TEMP build's file (mpm_init.cpp:1097-1108): four more columns per
 * particle after the model's, temperature specific_heat thermal_conductivity
 * heat_source; th [MPM_NTHERMAL * np] receives them (heat flux zero).  The
 * loader needs the thermal module enabled and uploads both records. */
int mpm_read_particle_file_thermal(const char *path, int64_t *np, double *pos, double *rdata,
                                   int *idata, double *th, double totals[2]);
int mpm_load_particle_file_thermal(mpm_ctx *ctx, const char *path, double totals[2]);
/* InitParticles(mincoords, maxcoords, vel, ppc, ...) with generate_particle
 * (mpm_init.cpp:1199-1473): particles of the cells of the specs' domain (x
 * fastest) inside [mincoords, maxcoords]; one at each cell centre, or with
 * multi_per_cell a ppc lattice per cell, where, as in the reference's 3-D
 * build, ppc[2] is ignored and one layer is placed along z (volume = cell
 * volume / (ppc[0] ppc[1])).  Radius (3 V / (4 pi))^(1/3). */
int mpm_autogen_particles(const mpm_specs *s, const double mincoords[3],
                          const double maxcoords[3], const double vel[3], const int ppc[3],
                          double dens, int model, double E, double nu, double K, double gamma,
                          double visc, int multi_per_cell, int64_t *np, double *pos,
                          double *rdata, int *idata, double totals[2]);
/* Writes the ASCII particle file in the layout of the reference's generator
 * (Generate_MPs_Inputfile_Generic.py:1020-1127): tab-separated, integers as
 * such, reals as %.6e; the header's model columns are those of particle 0's
 * model, and each row carries its own model's fields. */
int mpm_write_particle_file(const char *path, int64_t np, const double *pos, const double *rdata,
                            const int *idata);
/* The same with the generator's temperature columns (T spheat thermcond
 * heatsrc, Generate_MPs_Inputfile_Generic.py:1083-1090) from th. */
int mpm_write_particle_file_thermal(const char *path, int64_t np, const double *pos,
                                    const double *rdata, const int *idata, const double *th);

/* Checkpoint / restart.  mpm_write_checkpoint replaces writeCheckpointFile
 * (mpm_particle_outputs.cpp:355-499): dir/Header in the reference's layout
 * without EB ("Checkpoint version: 1", step, output index, time to 17 digits;
 * WriteHeader :311-353; the reader also takes the EB build's variant with
 * EB_generate_max_level before the time) and dir/particles.bin holding every particle field, the ids and
 * the cold fields (the reference stores the particles in AMReX's checkpoint
 * format instead).  mpm_read_checkpoint replaces readCheckpointFile
 * (:537-589): the context takes the particles (ids kept), step count and
 * time; as the reference's driver does (main.cpp:65-69), the caller then runs
 * mpm_init_internal_forces.  No nodal data is stored (the reference stores
 * none).  With the thermal module enabled, dir/thermal.bin holds the thermal
 * fields by upload index (the USE_TEMP build's record carries them), and a
 * context with the module enabled reads them back (MPM_E_ARG when the
 * checkpoint has none). */
int mpm_write_checkpoint(mpm_ctx *ctx, const char *dir, int output_it);
/* writeAsciiFiles (mpm_particle_outputs.cpp:188-196): the live particles as
 * text in "<prefix>_t<time with 6 decimals>" (the reference's name rule; the
 * body's layout, AMReX's WriteAsciiFile in the reference, is the count, then
 * per particle: x y z id cpu, the real fields, the integer fields).  The path
 * is copied to path_out when given. */
int mpm_write_ascii_particles(mpm_ctx *ctx, const char *prefix, double time, char *path_out,
                              int64_t path_len);
int mpm_read_checkpoint(mpm_ctx *ctx, const char *dir, int *output_it);

/* Number of live particles (removed ones excluded). */
int mpm_num_particles(mpm_ctx *ctx, int64_t *np);
/* Copies live particles back in the upload's order (ids preserved):
 * pos/rdata/idata as in upload, id[np] the upload index of each particle. */
int mpm_download_particles(mpm_ctx *ctx, double *pos, double *rdata,
                           int *idata, int64_t *id);
/* One real field (mpm_real_field) of every particle by upload index:
 * out[id] for id < n_upload (the count of the last upload), NaN for removed
 * particles.  A light read-back of the same data mpm_download_particles
 * returns (the particle output of mpm_particle_outputs.cpp:188-334). */
int mpm_get_particle_field(mpm_ctx *ctx, int field, int64_t n_upload, double *out);
/* Copies nodal component comp (mpm_nodal_comp; mpm_nodal_thermal_comp with
 * the thermal module enabled) of the valid nodal box, (n0+1)(n1+1)(n2+1)
 * values, i fastest. */
int mpm_get_nodal(mpm_ctx *ctx, int comp, double *out);
/* Overwrites nodal component comp with in[], laid out as mpm_get_nodal
 * returns it.  In the reference the caller owns the nodal MultiFab
 * (main.cpp:83-187 hands nodaldata to each step function) and may write it
 * between calls; components the build does not store (the rigid-body and
 * normal components) are an error. */
int mpm_set_nodal(mpm_ctx *ctx, int comp, const double *in);

/* ------------------------------------------------------------------------ */
/* MPMParticleContainer methods (mpm_particle_container.H:150-577)          */
/* ------------------------------------------------------------------------ */

/* Reset_Nodaldata_to_Zero (nodal_data_ops.cpp:924-927). */
int mpm_reset_nodal(mpm_ctx *ctx);

/* deposit_onto_grid_momentum (mpm_particle_grid_ops.cpp:333-628), including
 * the v = p/m normalisation.  Flags as in the reference; update_forces 2
 * deposits as 1 and leaves nodal component 15 zero (:398-400, 617-625). */
int mpm_p2g(mpm_ctx *ctx, int update_mass, int update_vel, int update_forces);

/* backup_current_velocity (nodal_data_ops.cpp:57-81). */
int mpm_backup_velocity(mpm_ctx *ctx);

/* Nodal_Time_Update_Momentum (nodal_data_ops.cpp:451-482). */
int mpm_nodal_update(mpm_ctx *ctx, double dt);

/* Apply_Nodal_BCs (utilities.cpp:183-204): wall BCs, UDF wall BCs,
 * level-set BCs, then store_delta_velocity.  t is the time the UDFs are
 * evaluated at. */
int mpm_apply_nodal_bcs(mpm_ctx *ctx, double dt, double t);

/* interpolate_from_grid (mpm_particle_grid_ops.cpp:924-1098). */
int mpm_g2p(mpm_ctx *ctx, int update_vel, int update_strainrate, double dt);

/* moveParticles (mpm_particle_timestep.cpp:226-578) followed by the
 * Redistribute_Fill_Update that follows it in the time loop
 * (main.cpp:142, utilities.cpp:420-434): periodic wrap and removal of
 * particles outside a non-periodic domain when step % num_redist == 0. */
int mpm_move(mpm_ctx *ctx, double dt, int step);

/* Redistribute_Fill_Update (utilities.cpp:420-434) as the time loop calls it
 * at the top of a step (main.cpp:88): when step % num_redist == 0, periodic
 * wrap and removal of particles outside a non-periodic domain (the
 * neighbour images it refreshes are index arithmetic here).  mpm_move runs
 * the same after the move (main.cpp:142). */
int mpm_redistribute(mpm_ctx *ctx, int step);

/* updateVolume (mpm_particle_timestep.cpp:131-182). */
int mpm_update_volume(mpm_ctx *ctx);

/* Calculate_MP_Stress_Strain (utilities.cpp:372-399) → apply_constitutive_model
 * (mpm_particle_container.cpp:43-142) or _delta (:178-272). */
int mpm_stress(mpm_ctx *ctx, double dt, double time);

/* Calculate_time_step (mpm_particle_timestep.cpp:38-111). */
int mpm_calculate_dt(mpm_ctx *ctx, double *dt);

/* Initialise_Internal_Forces (mpm_init.cpp:368-401). */
int mpm_init_internal_forces(mpm_ctx *ctx);

/* nsteps iterations of the time loop body (main.cpp:83-187) without output
 * and diagnostics: dt, P2G, grid update, BCs, USL/MUSL G2P, move,
 * redistribute, volume, stress, time += dt. */
int mpm_step(mpm_ctx *ctx, int nsteps);

/* Current simulation time / step count kept by the context. */
int mpm_get_time(mpm_ctx *ctx, double *time, int64_t *steps, double *last_dt);

/* ------------------------------------------------------------------------ */
/* Boundary data                                                            */
/* ------------------------------------------------------------------------ */

/* Level-set body (mpm_eb.H:176-288, nodal_data_ops.cpp:129-220): phi is the
 * body's signed distance on the valid nodal box ((n0+1)(n1+1)(n2+1), i
 * fastest, lsref = 1).  Up to MPM_MAX_LS_BODIES bodies. */
#define MPM_MAX_LS_BODIES 8
int mpm_add_levelset_body(mpm_ctx *ctx, const double *phi, int bc, double mu,
                          const double wall_vel[3]);

/* Analytic level-set bodies: the signed distance, negative inside the solid,
 * on the valid nodal box in the layout mpm_add_levelset_body takes, of the
 * shapes the reference builds with AMReX EB2 implicit functions
 * (mpm_eb.cpp:212-388).  params:
 *   MPM_LS_SPHERE   {cx, cy, cz, radius}
 *   MPM_LS_PLANE    {px, py, pz, nx, ny, nz}  (solid behind the plane, against n)
 *   MPM_LS_CYLINDER {cx, cy, cz, radius, height, direction}  (height <= 0: unbounded)
 *   MPM_LS_BOX      {lo_x, lo_y, lo_z, hi_x, hi_y, hi_z}
 * fluid_inside (the *_has_fluid_inside inputs) flips the sign: the solid is
 * then the outside of the shape.  EB2's own signed-distance fill is AMReX
 * code absent from the reference tree; these are the shapes' exact
 * distances. */
#define MPM_LS_SPHERE 0
#define MPM_LS_PLANE 1
#define MPM_LS_CYLINDER 2
#define MPM_LS_BOX 3
int mpm_levelset_analytic(const mpm_specs *s, int type, const double *params, int fluid_inside,
                          double *phi);
/* removeParticlesInsideEB (mpm_init.cpp:1520-1596): particles where any
 * body's level set (interpolated as the level-set BC reads it) is below
 * TINYVAL = 1e-20 are removed; *removed (may be NULL) is their number. */
int mpm_remove_particles_inside_levelset(mpm_ctx *ctx, int64_t *removed);

/* Wall-velocity UDF for face (dir, hi) (mpm_specs.H:452-485,
 * nodal_data_ops.cpp:929-1175, mpm_particle_timestep.cpp:417-475).  The
 * context evaluates it on the host at each BC pass as the reference does. */
int mpm_set_wall_udf(mpm_ctx *ctx, int dir, int hi, mpm_wall_udf_fn fn);

/* ------------------------------------------------------------------------ */
/* Diagnostics (mpm_diagnostics.cpp:23-127)                                 */
/* ------------------------------------------------------------------------ */

/* out[0]=TKE, out[1]=TSE, out[2..4]=sum m*v, out[5]=sum m (all particles). */
int mpm_energies(mpm_ctx *ctx, double out[6]);

/* The sums of mpm_energies over particles the caller holds in device memory,
 * allocating nothing.  No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under HIP graph capture, serving
 * loops, harnesses that forbid hipMalloc).  rdata: device array of the
 * MPM_NREAL real fields, component-major with stride `stride` >= np (field f
 * of particle i at rdata[f * stride + i]); id: device array of np ids, a
 * particle with id < 0 is skipped, or NULL (all counted); out: 6 doubles of
 * device memory in the order of mpm_energies; stream: the hipStream_t to
 * enqueue on (NULL = the null stream); the call does not synchronise.
 * Partial sums go to one library-owned device array per GPU, sized for
 * np <= MPM_ENERGIES_NOALLOC_MAX_PARTICLES (larger np: MPM_E_ARG); calls that
 * overlap on one GPU race on it.  The sums run in a fixed order: equal inputs
 * give equal bits. */
#define MPM_ENERGIES_NOALLOC_MAX_PARTICLES (1 << 24)
int mpm_energies_noalloc(int64_t np, int64_t stride, const double *rdata, const int64_t *id,
                         double *out, void *stream);

/* Context options without a reference counterpart (tooling for callers that
 * measure or inspect the step): MPM_OPT_PROFILE records HIP events around
 * each kernel of mpm_step (read back with mpm_last_step_timing);
 * MPM_OPT_KEEP_FORCES stores nodal forces for mpm_get_nodal (default 1);
 * MPM_OPT_SORTED selects the cell-sorted step (default 1) or the unsorted
 * one-thread-per-particle atomic step (0); MPM_OPT_FAST selects the tile
 * kernels of the particle side where they apply (uniform orders, with or
 * without level-set bodies and wall UDF tables; default 1) or the generic
 * sorted kernel (0); MPM_OPT_FUSED_DT lets the tile particle kernel reduce
 * the next step's dt candidates on its way out (default 1) instead of a
 * separate pass over the particles at the top of the next step (0);
 * MPM_OPT_DETERMINISTIC (default 0) makes the sorted step's P2G free of
 * floating-point atomics, so that equal inputs give equal bits run to run:
 * each cell's particles are taken in a fixed order, the four wavefronts of a
 * tile sum private slabs in a fixed order, and each node adds the node blocks
 * of its tiles in a fixed order (an extra pass; the reference deposits with
 * order-dependent atomics, mpm_particle_grid_ops.cpp:484-578).  It applies
 * (thermal deposit and external slab loads included) to particles inside
 * their sorted tile (redistributed every step, or not yet outside the
 * domain), MUSL's re-projection deposit with periodic faces included (its
 * blocks carry the current and the start-of-step momentum, the images take
 * the latter); it allocates one 7 x 343-double node block per cell tile.
 * MPM_OPT_TILE_WAVES: wavefronts per cell tile of the sorted step's P2G and
 * fused particle kernels, 4 or 8 (8: shorter per-tile chains for scenes with
 * fewer than one full tile per compute unit), 0 (default) to choose by the
 * particle count.  The deterministic P2G always runs 4.
 * MPM_OPT_COARSE_BLEND (default 0): the coarse steps of mpm_step_learned
 * blend PIC and FLIP with alpha^nfine instead of alpha in their particle
 * update, the blend of the nfine fine steps a coarse step stands for. */
#define MPM_OPT_PROFILE 1
#define MPM_OPT_KEEP_FORCES 2
#define MPM_OPT_SORTED 3
#define MPM_OPT_FAST 4
#define MPM_OPT_FUSED_DT 5
#define MPM_OPT_DETERMINISTIC 6
#define MPM_OPT_COARSE_BLEND 7
#define MPM_OPT_TILE_WAVES 8
int mpm_set_option(mpm_ctx *ctx, int option, int value);

/* Last HIP error string (static storage). */
const char *mpm_last_error(void);

/* Milliseconds of the last mpm_step call measured with HIP events on the
 * context's stream, and per-kernel average durations (ms) over that call:
 * kernel order is p2g, grid, g2p (fused particle kernel), sort/other. */
int mpm_last_step_timing(mpm_ctx *ctx, double *total_ms, double kernel_ms[4]);

/* ---------------------------------------------------------------------------
 * Learned coarse-step correction.  No counterpart in the reference: an
 * addition beside the plain step, which stays the oracle.  One coarse step
 * (nfine times the CFL dt) takes the place of nfine plain steps; after its
 * nodal update a per-node stencil MLP corrects the nodal velocity:
 *   inputs  27 neighbour nodes (dz, dy, dx in {-1,0,1}, dx fastest) x
 *           (m / m_ref, m v / (m_ref v_ref)) = 108 values, zero outside the box;
 *   network 108 -> 64 -> 64 -> 3, tanh-form GELU, BF16 MFMA operands with
 *           FP32 accumulation; output x dv_ref = velocity correction.
 * Parameters (FP32, MPM_MLP_NPARAMS): W1[108][64], b1[64], W2[64][64],
 * b2[64], W3[64][3], b3[3] with W[in][out].  Callers: the learned-step
 * driver (exagoop_amd/learned.py) and its tests.
 * ------------------------------------------------------------------------- */
#define MPM_MLP_NPARAMS 11331
#define MPM_MLP_IN 128 /* feature row length (108 used, zero padded) */
typedef struct mpm_mlp mpm_mlp;
typedef struct mpm_dataset mpm_dataset;

int mpm_mlp_create(mpm_mlp **out);
void mpm_mlp_destroy(mpm_mlp *m);
/* scales = {m_ref, v_ref, dv_ref}. */
int mpm_mlp_set_params(mpm_mlp *m, const float *params, const double scales[3]);
int mpm_mlp_get_params(mpm_mlp *m, float *params, double scales[3]);
/* Network outputs Y[nrows][3] of host feature rows X[nrows][MPM_MLP_IN]. */
int mpm_mlp_forward(mpm_mlp *m, int64_t nrows, const float *X, float *Y);
/* Mean squared error (1/nrows) sum (Y - T)^2 over host rows and its gradient
 * grad[MPM_MLP_NPARAMS]. */
int mpm_mlp_gradient(mpm_mlp *m, int64_t nrows, const float *X, const float *T, float *grad,
                     double *loss);
/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8, step number t >= 1) with
 * the given gradient. */
int mpm_mlp_adam(mpm_mlp *m, const float *grad, float lr, int t);

/* Training rows on the device: features (BF16) and velocity-correction
 * targets (m/s). */
int mpm_dataset_create(int64_t capacity, mpm_dataset **out);
void mpm_dataset_destroy(mpm_dataset *d);
int mpm_dataset_size(mpm_dataset *d, int64_t *nrows);
int mpm_dataset_add(mpm_dataset *d, int64_t nrows, const float *X, const float *T);
/* Root mean square of the targets' components. */
int mpm_dataset_target_rms(mpm_dataset *d, double *rms);

/* Feature rows X[nrows][MPM_MLP_IN] (BF16 values widened to float) of the
 * listed nodes from the context's current nodal mass and velocity. */
int mpm_learn_features(mpm_ctx *ctx, mpm_mlp *m, int64_t nrows, const int64_t *nodes,
                       float *X);
/* From the context's state: one coarse step (P2G + nodal update with nfine x
 * the CFL dt) gives the features of up to max_rows randomly chosen active
 * nodes (scales from m); the coarse step is discarded and nfine plain steps
 * run (the context advances as with mpm_step).  Each chosen node still
 * active after them adds a row with target = fine nodal velocity - coarse
 * nodal velocity.  Requires the specialised USL step (sorted, redistribution
 * every step, no level set or UDF). */
int mpm_learn_collect(mpm_ctx *ctx, mpm_mlp *m, mpm_dataset *d, int nfine, int max_rows,
                      uint64_t seed);
/* epochs of shuffled minibatches with Adam; targets are divided by dv_ref;
 * loss_out[epoch] = mean loss of the epoch (may be NULL). */
int mpm_mlp_train(mpm_mlp *m, mpm_dataset *d, int epochs, int batch, float lr, uint64_t seed,
                  double *loss_out);
/* ncoarse learned steps of nfine x dt each.  guards = {gp, ge}: with M the
 * active nodal mass, u the mass-weighted rms nodal velocity of the coarse
 * step and s = max(u, |g| dt_coarse), the correction dv must be finite,
 * |sum m dv| <= gp M s (momentum) and the corrected rms velocity
 * <= (1 + ge) s (energy); otherwise the coarse step is dropped and nfine
 * plain steps run instead (*fallbacks counts them).  m == NULL runs the
 * coarse steps without correction. */
int mpm_step_learned(mpm_ctx *ctx, mpm_mlp *m, int ncoarse, int nfine, const double guards[2],
                     int *fallbacks);

/* ---------------------------------------------------------------------------
 * Batched ensembles.  No counterpart in the reference (one scene per run):
 * many independent scenes of uneven size advance together, one persistent
 * kernel launch per step, the steps replayed from a captured HIP graph.
 * Each scene follows the specialised sorted USL step of mpm_step (the time
 * loop body, main.cpp:83-187) and gives the same results as a context of its
 * own, up to the order of the FP64 atomic sums.  Requirements: one order for
 * every direction of every scene, USL, redistribution every step
 * (num_redist = 1), walls at rest, no level set or wall UDF, and one
 * calculate_strain_based_on_delta setting for the whole batch.  The nodal
 * force components are not kept (mpm_batch_get_nodal returns zeros for them).
 * Callers: the ensemble driver (exagoop_amd/ensemble.py), its tests and the
 * benchmark.
 * ------------------------------------------------------------------------- */
typedef struct mpm_batch mpm_batch;

/* Work decomposition of the persistent step (the autotuner's search space):
 * persistent workgroups (multiple of 8; 0 = as many as fit on the device at
 * once), particles per key/scatter task, particles per P2G/particle task
 * (each such task takes a share of the scene's non-empty 4^3-cell tiles),
 * nodes per grid task, steps captured per graph replay. */
typedef struct mpm_batch_tuning {
    int workgroups;
    int sort_chunk;
    int tile_chunk;
    int node_chunk;
    int graph_steps;
} mpm_batch_tuning;

/* nscenes scenes with specs[s] and np[s] >= 1 particles each; allocates all
 * device state on the current device. */
int mpm_batch_create(int nscenes, const mpm_specs *specs, const int64_t *np, mpm_batch **out);
void mpm_batch_destroy(mpm_batch *b);
/* Particles of one scene, as mpm_upload_particles (np must match). */
int mpm_batch_upload(mpm_batch *b, int scene, int64_t np, const double *pos,
                     const double *rdata, const int *idata);
/* Initialise_Internal_Forces and the first Calculate_time_step of every
 * scene (all scenes uploaded). */
int mpm_batch_init(mpm_batch *b);
/* nsteps steps of every scene. */
int mpm_batch_step(mpm_batch *b, int nsteps);
int mpm_batch_num_particles(mpm_batch *b, int scene, int64_t *np);
/* As mpm_download_particles / mpm_get_nodal / mpm_get_time for one scene. */
int mpm_batch_download(mpm_batch *b, int scene, double *pos, double *rdata, int *idata,
                       int64_t *id);
int mpm_batch_get_nodal(mpm_batch *b, int scene, int comp, double *out);
int mpm_batch_get_time(mpm_batch *b, int scene, double *time, int64_t *steps, double *last_dt);
int mpm_batch_get_tuning(mpm_batch *b, mpm_batch_tuning *t);
int mpm_batch_set_tuning(mpm_batch *b, const mpm_batch_tuning *t);
/* Autotuner: times trial_steps steps per candidate decomposition (coordinate
 * search over the mpm_batch_tuning fields) from a snapshot of the state,
 * restores the snapshot after every trial and keeps the fastest: the
 * simulation state is unchanged.  best / best_ms may be NULL. */
int mpm_batch_autotune(mpm_batch *b, int trial_steps, mpm_batch_tuning *best, double *best_ms);
/* nsteps steps with an instrumented launch (no graph; the scenes advance):
 * per stage (keys, scan, scatter, order, p2g, grid, particle; order runs in
 * deterministic batches only), workgroup-milliseconds
 * per step spent in its tasks (busy_ms) and waiting for the previous stage
 * (wait_ms), and its tasks per step; workgroup-milliseconds per step outside
 * tasks (taking tasks, leaving); milliseconds per step.  Any output may be
 * NULL. */
int mpm_batch_profile(mpm_batch *b, int nsteps, double busy_ms[7], double wait_ms[7],
                      int64_t tasks[7], double *other_ms, double *step_ms);
/* MPM_OPT_DETERMINISTIC (the only batch option; before mpm_batch_init): every
 * scene takes the single-scene deterministic step (each cell's particles in
 * index order, atomic-free tile deposits into whole node blocks, each node
 * adding its blocks in a fixed order), so that equal inputs give equal bits
 * run to run and each scene the bits of a deterministic context of its own.
 * Allocates one 7 x 343-double node block per cell tile of every scene. */
int mpm_batch_set_option(mpm_batch *b, int option, int value);
/* Milliseconds per step of the last mpm_batch_step call (HIP events), the
 * particle slots of all scenes and the tasks of one step. */
int mpm_batch_stats(mpm_batch *b, double *step_ms, int64_t *particles, int *tasks);

/* ------------------------------------------------------------------------ */
/* Thermal module: the USE_TEMP build's heat conduction (a context option).  */
/* ------------------------------------------------------------------------ */
/* Particle thermal fields, realData 46..52 of the USE_TEMP build
 * (mpm_specs.H:52-58).  They travel in a record of their own (component-
 * major, by upload index) so the 46-field mechanical record is unchanged. */
#define MPM_NTHERMAL 7
enum mpm_thermal_field {
    MPM_TH_TEMPERATURE = 0,
    MPM_TH_SPECIFIC_HEAT = 1,
    MPM_TH_CONDUCTIVITY = 2,
    MPM_TH_HEAT_FLUX = 3, /* 3..5 */
    MPM_TH_HEAT_SOURCE = 6
};

/* Nodal components of the USE_TEMP build (constants.H:31, 58-63), readable
 * with mpm_get_nodal once the module is enabled. */
enum mpm_nodal_thermal_comp {
    MPM_N_MASS_SPHEAT = 20,      /* sum m c_p N */
    MPM_N_MASS_SPHEAT_TEMP = 21, /* sum m c_p T N */
    MPM_N_TEMPERATURE = 22,
    MPM_N_SOURCE_TEMP = 23,
    MPM_N_DELTA_TEMPERATURE = 24,
    MPM_N_MASS_CONDUCTIVITY = 25, /* sum m k N */
    MPM_N_NCOMP_TEMP = 26
};

/* Temperature wall conditions (constants.H:95-99). */
enum mpm_bc_temp {
    MPM_BC_TEMP_PERIODIC = 0,
    MPM_BC_TEMP_ISOTHERMAL = 1, /* T = T_wall */
    MPM_BC_TEMP_ADIABATIC = 2,  /* T = T of the inward neighbour */
    MPM_BC_TEMP_USERFLUX = 3,   /* T = T_nb + flux dx / k_node */
    MPM_BC_TEMP_CONVECTION = 4  /* T = (T_nb + Bi T_inf) / (1 + Bi), Bi = h dx / k_node */
};

/* The temperature-BC subset of MPMspecs (mpm_specs.H:166-176, 227-234,
 * 488-560): per face type and parameters. */
typedef struct mpm_thermal_specs {
    int bclo[3];
    int bchi[3];
    double T_wall_lo[3], T_wall_hi[3];
    double flux_lo[3], flux_hi[3];
    double h_lo[3], h_hi[3];
    double T_inf_lo[3], T_inf_hi[3];
} mpm_thermal_specs;

/* Reference defaults (mpm_specs.H:227-234, 166-176): periodic faces
 * MPM_BC_TEMP_PERIODIC, all others MPM_BC_TEMP_ADIABATIC, parameters 0. */
void mpm_thermal_specs_default(const mpm_specs *s, mpm_thermal_specs *t);

/* Turns on the thermal stages of mpm_init_internal_forces and mpm_step
 * (main.cpp:99-105, 112-116, 130-135; mpm_init.cpp:417-441) with these wall
 * conditions; the thermal fields of all particles start at zero.  Level-set
 * bodies carry no temperature condition (the reference's default). */
int mpm_enable_thermal(mpm_ctx *ctx, const mpm_thermal_specs *t);
/* Temperature wall user function, the signature of TempUDFFuncPtr
 * (mpm_specs.H:13): f(x, y, z, t, out[2]).  out[0] is the wall temperature
 * (isothermal face), the flux (user-flux face) or h (convection face); out[1]
 * is T_inf (convection face). */
typedef void (*mpm_wall_temp_udf_fn)(double, double, double, double, double *);
/* Sets (fn) or clears (NULL) the temperature function of one face.  It is
 * evaluated at the face's nodes at the start of every step
 * (compute_udf_temp_at_nodes, nodal_data_ops.cpp:1178-1306) and applied after
 * the wall conditions of both temperature BC passes of the step, by the face's
 * temperature BC type (apply_udf_nodal_bcs_temperature, :1308-1446).  The
 * initial pass applies the wall conditions only, as the reference's does. */
int mpm_set_wall_temp_udf(mpm_ctx *ctx, int dir, int hi, mpm_wall_temp_udf_fn fn);
/* Temperature condition of a level-set body (LevelSetBody::temp_bc_int,
 * mpm_eb.H:78-118; mpm_eb.cpp:521-545): the codes and parameters of
 * <name>.temp_bc_type, lset_T_wall, heat_flux, h_conv, T_inf. */
enum mpm_ls_temp_bc {
    MPM_LS_TEMP_ADIABATIC = 0, /* no condition (the default) */
    MPM_LS_TEMP_ISOTHERMAL = 1,
    MPM_LS_TEMP_HEATFLUX = 3,
    MPM_LS_TEMP_CONVECTIVE = 4
};
/* Sets the temperature condition of level-set body `body` (0-based, in the
 * order of mpm_add_levelset_body).  Applied after the wall conditions and
 * the face functions of a temperature BC pass (nodal_levelset_bcs_temperature,
 * nodal_data_ops.cpp:235-350): at nodes inside the body with thermal mass,
 * T = T_wall, or from the neighbour along the normal's dominant direction
 * with the flux or convection formula; the first pass of a MUSL step applies
 * the isothermal bodies only (utilities.cpp:221-240, main.cpp:103-104). */
int mpm_set_levelset_body_temp(mpm_ctx *ctx, int body, int temp_bc, double T_wall,
                               double heat_flux, double h_conv, double T_inf);
/* Thermal fields by upload index: th[f * n_upload + i] for field f of the
 * particle uploaded at index i (n_upload = count of the last upload; the
 * reader's T, spheat, thermcond, heatsrc columns, mpm_init.cpp:1097-1108). */
int mpm_upload_thermal(mpm_ctx *ctx, int64_t n_upload, const double *th);
/* Same layout; NaN for particles removed since the upload. */
int mpm_download_thermal(mpm_ctx *ctx, int64_t n_upload, double *th);

#ifdef __cplusplus
}
#endif

#endif /* EXAGOOP_AMD_H_ */

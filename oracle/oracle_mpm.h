/*
 * oracle_mpm.h — CPU restatement of ExaGOOP's explicit MPM step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker and the
 * timed CPU baseline; it is never linked into or called by the product
 * library (libexagoop_amd.so).  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg load it.
 *
 * It restates the reference with its own data model: AoS particle records
 * (pos + 46 reals + 3 ints + id, mpm_specs.H:17-73), a component-major
 * nodal array with ghost layers (20 comps, constants.H:30-55), periodic
 * ghost-particle images for P2G and ghost-node fill for G2P (the AMReX
 * NeighborParticleContainer / FillBoundary semantics), per-node
 * re-evaluation of basisval_and_grad, and atomic adds under OpenMP.
 * The struct types come from include/exagoop_amd.h so tests can drive the
 * oracle and the HIP library with the same inputs.
 */
#ifndef ORACLE_MPM_H_
#define ORACLE_MPM_H_

#include "../include/exagoop_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_ctx orc_ctx;

int orc_create(const mpm_specs *specs, orc_ctx **out);
void orc_destroy(orc_ctx *ctx);
void orc_set_threads(orc_ctx *ctx, int nthreads);
int orc_upload_particles(orc_ctx *ctx, int64_t np, const double *pos,
                         const double *rdata, const int *idata);
int orc_num_particles(orc_ctx *ctx, int64_t *np);
int orc_download_particles(orc_ctx *ctx, double *pos, double *rdata,
                           int *idata, int64_t *id);
int orc_get_nodal(orc_ctx *ctx, int comp, double *out);

int orc_reset_nodal(orc_ctx *ctx);
int orc_fill_neighbors(orc_ctx *ctx);
int orc_p2g(orc_ctx *ctx, int update_mass, int update_vel, int update_forces);
int orc_backup_velocity(orc_ctx *ctx);
int orc_nodal_update(orc_ctx *ctx, double dt);
int orc_apply_nodal_bcs(orc_ctx *ctx, double dt, double t);
int orc_g2p(orc_ctx *ctx, int update_vel, int update_strainrate, double dt);
int orc_move(orc_ctx *ctx, double dt, int step);
int orc_redistribute(orc_ctx *ctx, int step);
int orc_update_volume(orc_ctx *ctx);
int orc_stress(orc_ctx *ctx, double dt, double time);
int orc_calculate_dt(orc_ctx *ctx, double *dt);
int orc_init_internal_forces(orc_ctx *ctx);
int orc_step(orc_ctx *ctx, int nsteps);
int orc_get_time(orc_ctx *ctx, double *time, int64_t *steps, double *last_dt);
int orc_add_levelset_body(orc_ctx *ctx, const double *phi, int bc, double mu,
                          const double wall_vel[3]);
int orc_set_wall_udf(orc_ctx *ctx, int dir, int hi, mpm_wall_udf_fn fn);
int orc_energies(orc_ctx *ctx, double out[6]);
/* Thermal module (the USE_TEMP build). */
int orc_enable_thermal(orc_ctx *ctx, const mpm_thermal_specs *t);
int orc_upload_thermal(orc_ctx *ctx, int64_t n_upload, const double *th);
int orc_download_thermal(orc_ctx *ctx, int64_t n_upload, double *th);
int orc_set_wall_temp_udf(orc_ctx *ctx, int dir, int hi, mpm_wall_temp_udf_fn fn);
int orc_set_levelset_body_temp(orc_ctx *ctx, int body, int temp_bc, double T_wall,
                               double heat_flux, double h_conv, double T_inf);

/* Pieces exposed for closed-form pinning in tests. */
double orc_quadraticspline_1d(int type, double z);
double orc_quadraticspline_1d_der(int type, double z);
double orc_cubicspline_1d(int type, double z);
double orc_cubicspline_1d_der(int type, double z);
/* basisval_and_grad for one (particle, stencil node): returns N, grad[3]. */
double orc_basisval_and_grad(const int stencil[3], const int cell[3],
                             const double xp[3], const double plo[3],
                             const double dx[3], const int order[3],
                             const int periodic[3], const int lo[3],
                             const int hi[3], double grad[3]);
/* applybc: returns modify_pos, writes relvel_out (preloaded by caller). */
int orc_applybc(const double relvel_in[3], double relvel_out[3], double mu,
                const double normal[3], int bc);

#ifdef __cplusplus
}
#endif

#endif

// mpm_container.hpp — C++ host facade with the MPMParticleContainer method
// surface (mpm_particle_container.H:150-577) and the step-level free
// functions (utilities.H:28-79, nodal_data_ops.H:21-95), forwarding to the
// C ABI of libexagoop_amd.so.  Header-only; errors are thrown as
// std::runtime_error carrying the library's status and message (the
// reference aborts instead).
#pragma once

#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/exagoop_amd.h"

namespace exagoop_amd {

inline void check(int rc, const char *what) {
    if (rc != MPM_OK)
        throw std::runtime_error(std::string(what) + " failed (" + std::to_string(rc) +
                                 "): " + mpm_last_error());
}

class MPMParticleContainer {
  public:
    // Initialise_Domain + the container constructor (mpm_init.cpp:108-299, main.cpp:59).
    explicit MPMParticleContainer(const mpm_specs &specs) : specs_(specs) {
        check(mpm_create(&specs_, &ctx_), "mpm_create");
    }
    ~MPMParticleContainer() { mpm_destroy(ctx_); }
    MPMParticleContainer(const MPMParticleContainer &) = delete;
    MPMParticleContainer &operator=(const MPMParticleContainer &) = delete;

    // InitParticles hand-off: component-major arrays (include/exagoop_amd.h).
    void InitParticles(int64_t np, const double *pos, const double *rdata, const int *idata) {
        check(mpm_upload_particles(ctx_, np, pos, rdata, idata), "mpm_upload_particles");
    }
    int64_t TotalNumberOfParticles() const {
        int64_t n = 0;
        check(mpm_num_particles(ctx_, &n), "mpm_num_particles");
        return n;
    }
    void GetParticles(double *pos, double *rdata, int *idata, int64_t *id) const {
        check(mpm_download_particles(ctx_, pos, rdata, idata, id), "mpm_download_particles");
    }

    // deposit_onto_grid_momentum (mpm_particle_grid_ops.cpp:333-628); gravity,
    // loads and tolerance come from the specs the container was built with.
    void deposit_onto_grid_momentum(int update_mass, int update_vel, int update_forces) {
        check(mpm_p2g(ctx_, update_mass, update_vel, update_forces), "mpm_p2g");
    }
    // interpolate_from_grid (mpm_particle_grid_ops.cpp:924-1098).
    void interpolate_from_grid(int update_vel, int update_strainrate, double dt) {
        check(mpm_g2p(ctx_, update_vel, update_strainrate, dt), "mpm_g2p");
    }
    // apply_constitutive_model[_delta] via Calculate_MP_Stress_Strain
    // (utilities.cpp:372-399): the applied strain rate is selected from time.
    void apply_constitutive_model(double dt, double time) {
        check(mpm_stress(ctx_, dt, time), "mpm_stress");
    }
    // updateVolume (mpm_particle_timestep.cpp:131-182).
    void updateVolume() { check(mpm_update_volume(ctx_), "mpm_update_volume"); }
    // moveParticles + the following Redistribute_Fill_Update.
    void moveParticles(double dt, int step) { check(mpm_move(ctx_, dt, step), "mpm_move"); }
    // Redistribute_Fill_Update at the top of a step (utilities.cpp:420-434).
    void Redistribute_Fill_Update(int step) {
        check(mpm_redistribute(ctx_, step), "mpm_redistribute");
    }
    // Calculate_time_step (mpm_particle_timestep.cpp:38-111).
    double Calculate_time_step() {
        double dt = 0.0;
        check(mpm_calculate_dt(ctx_, &dt), "mpm_calculate_dt");
        return dt;
    }
    // Calculate_Total_Energies + MWA velocity sums (mpm_diagnostics.cpp:23-127).
    void Calculate_Total_Energies(double &TKE, double &TSE) {
        double e[6];
        check(mpm_energies(ctx_, e), "mpm_energies");
        TKE = e[0];
        TSE = e[1];
    }

    // Free functions of the step (nodal_data_ops.H, utilities.H).
    void Reset_Nodaldata_to_Zero() { check(mpm_reset_nodal(ctx_), "mpm_reset_nodal"); }
    void backup_current_velocity() { check(mpm_backup_velocity(ctx_), "mpm_backup_velocity"); }
    void Nodal_Time_Update_Momentum(double dt) {
        check(mpm_nodal_update(ctx_, dt), "mpm_nodal_update");
    }
    void Apply_Nodal_BCs(double dt, double time) {
        check(mpm_apply_nodal_bcs(ctx_, dt, time), "mpm_apply_nodal_bcs");
    }
    void Initialise_Internal_Forces() {
        check(mpm_init_internal_forces(ctx_), "mpm_init_internal_forces");
    }
    // nsteps iterations of main.cpp:83-187 without output.
    void Step(int nsteps) { check(mpm_step(ctx_, nsteps), "mpm_step"); }
    void Time(double &time, int64_t &steps, double &dt) const {
        check(mpm_get_time(ctx_, &time, &steps, &dt), "mpm_get_time");
    }
    std::vector<double> Nodal(int comp, int64_t nn) const {
        std::vector<double> out(nn);
        check(mpm_get_nodal(ctx_, comp, out.data()), "mpm_get_nodal");
        return out;
    }
    void AddLevelSetBody(const double *phi, int bc, double mu, const double wall_vel[3]) {
        check(mpm_add_levelset_body(ctx_, phi, bc, mu, wall_vel), "mpm_add_levelset_body");
    }
    void SetWallUDF(int dir, int hi, mpm_wall_udf_fn fn) {
        check(mpm_set_wall_udf(ctx_, dir, hi, fn), "mpm_set_wall_udf");
    }
    // Thermal module (the USE_TEMP build): the temperature wall conditions of
    // MPMspecs (bclo_temp / bchi_temp and their parameters), then the thermal
    // fields by upload index.
    void EnableThermal(const mpm_thermal_specs &t) {
        check(mpm_enable_thermal(ctx_, &t), "mpm_enable_thermal");
    }
    void SetThermalFields(int64_t n_upload, const double *th) {
        check(mpm_upload_thermal(ctx_, n_upload, th), "mpm_upload_thermal");
    }
    void GetThermalFields(int64_t n_upload, double *th) const {
        check(mpm_download_thermal(ctx_, n_upload, th), "mpm_download_thermal");
    }
    void SetWallTempUDF(int dir, int hi, mpm_wall_temp_udf_fn fn) {
        check(mpm_set_wall_temp_udf(ctx_, dir, hi, fn), "mpm_set_wall_temp_udf");
    }
    mpm_ctx *raw() { return ctx_; }

  private:
    mpm_specs specs_;
    mpm_ctx *ctx_ = nullptr;
};

}  // namespace exagoop_amd

// mpm_ctx.hip — context, device storage and the C ABI of exagoop_amd.h.
//
// The context owns particle SoA buffers (component-major, reference field
// order), the nodal arrays of the valid nodal box, P2G accumulators and the
// device step scalars (dt, time).  Every entry point enqueues on the
// context's stream; dt and time stay on the device so a step needs no host
// round trip unless a wall-velocity UDF is set (the reference evaluates UDFs
// on the host every step too, nodal_data_ops.cpp:929-1059).

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <numeric>
#include <iomanip>
#include <sstream>
#include <string>
#include <sys/stat.h>
#include <cerrno>
#include <fstream>
#include <vector>

#include "mpm_context.h"

using namespace xmpm;

namespace xmpm {

thread_local char g_err[256] = "";

int set_err(int code, const char *what) {
    std::snprintf(g_err, sizeof(g_err), "%s", what);
    return code;
}

int hip_err(hipError_t e, const char *expr) {
    std::snprintf(g_err, sizeof(g_err), "%s: %s", expr, hipGetErrorString(e));
    return e == hipErrorOutOfMemory ? MPM_E_OOM : MPM_E_HIP;
}

int check_launch() {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        std::snprintf(g_err, sizeof(g_err), "kernel launch: %s", hipGetErrorString(e));
        return MPM_E_HIP;
    }
    return MPM_OK;
}

// Validated grid and step parameters of a problem specification
// (Initialise_Domain, mpm_init.cpp:108-299, and the MPMspecs fields the step
// reads).
int specs_params(const mpm_specs *s, GridP &g, PhysP &ph) {
    int od[3];
    int rc = mpm_order_directional(s, od);
    if (rc) return rc;
    for (int d = 0; d < 3; ++d) {
        if (s->ncells[d] < 1 || !(s->phi[d] > s->plo[d]))
            return set_err(MPM_E_ARG, "bad domain");
        if (s->bclo[d] < 0 || s->bclo[d] > MPM_BC_OUTFLOW || s->bchi[d] < 0 ||
            s->bchi[d] > MPM_BC_OUTFLOW)
            return set_err(MPM_E_BC, "unknown boundary condition");
        // applybc aborts on BC_OUTFLOW at the first nodal-BC pass
        // (mpm_kernels.H:107-110).
        if (s->bclo[d] == MPM_BC_OUTFLOW || s->bchi[d] == MPM_BC_OUTFLOW)
            return set_err(MPM_E_BC, "outflow BC is not supported by applybc");
    }
    if (s->stress_update_scheme != MPM_USL && s->stress_update_scheme != MPM_MUSL)
        return set_err(MPM_E_ARG, "stress_update_scheme must be USL or MUSL");
    std::memset(&g, 0, sizeof(g));
    g.uniform = (od[0] == od[1] && od[1] == od[2]);
    for (int d = 0; d < 3; ++d) {
        g.plo[d] = s->plo[d];
        g.phi[d] = s->phi[d];
        g.n[d] = s->ncells[d];
        g.hi[d] = s->ncells[d] - 1;
        g.dx[d] = (s->phi[d] - s->plo[d]) / (double)s->ncells[d];
        g.dxi[d] = 1.0 / g.dx[d];
        g.per[d] = s->periodic[d] ? 1 : 0;
        g.od[d] = od[d];
        double rhi = s->phi[d];
        while ((int)std::floor((rhi - s->plo[d]) * g.dxi[d]) >= g.n[d])
            rhi = std::nextafter(rhi, -INFINITY);
        g.rhi[d] = rhi;
    }
    g.nn = (int64_t)(g.n[0] + 1) * (g.n[1] + 1) * (g.n[2] + 1);
    for (int d = 0; d < 3; ++d) {
        g.ntc[d] = (g.n[d] + kTileCells - 1) / kTileCells;
        g.ntn[d] = g.n[d] / kTileCells + 1;
    }
    std::memset(&ph, 0, sizeof(ph));
    for (int d = 0; d < 3; ++d) {
        ph.grav[d] = s->gravity[d];
        ph.slab_lo[d] = s->force_slab_lo[d];
        ph.slab_hi[d] = s->force_slab_hi[d];
        ph.extforce[d] = s->extforce[d];
        ph.bclo[d] = s->bclo[d];
        ph.bchi[d] = s->bchi[d];
        ph.mu_lo[d] = s->wall_mu_lo[d];
        ph.mu_hi[d] = s->wall_mu_hi[d];
    }
    for (int q = 0; q < 9; ++q) {
        ph.wv_lo[q] = s->wall_vel_lo[q];
        ph.wv_hi[q] = s->wall_vel_hi[q];
    }
    ph.extloads = s->external_loads;
    ph.mass_tol = s->mass_tolerance;
    ph.alpha = s->alpha_pic_flip;
    ph.applied_rate = s->applied_strainrate;
    ph.applied_time = s->applied_strainrate_time;
    ph.delta = s->calculate_strain_based_on_delta == 1;
    ph.cfl = s->CFL;
    ph.dt_min = s->dt_min_limit;
    ph.dt_max = s->dt_max_limit;
    return MPM_OK;
}

bool walls_at_rest(const mpm_specs *s) {
    for (int q = 0; q < 9; ++q)
        if (s->wall_vel_lo[q] != 0.0 || s->wall_vel_hi[q] != 0.0) return false;
    return true;
}

}  // namespace xmpm


namespace {

void free_pstate(PState &P) {
    (void)hipFree(P.pos);
    (void)hipFree(P.r);
    (void)hipFree(P.idata);
    (void)hipFree(P.id);
    P.pos = P.r = nullptr;
    P.idata = nullptr;
    P.id = nullptr;
    P.np = P.cap = 0;
}

void free_particles(mpm_ctx *c) {
    (void)hipFree(c->P.vsnap);
    c->P.vsnap = nullptr;
    free_pstate(c->P);
    free_pstate(c->Q);
    free_pstate(c->R);
    (void)hipFree(c->S.key);
    (void)hipFree(c->S.rank);
    (void)hipFree(c->S.perm);
    (void)hipFree(c->S.pkt);
    c->S.key = c->S.rank = c->S.perm = nullptr;
    c->S.pkt = nullptr;
    c->S.pcap = 0;
}

hipError_t alloc_pstate(PState &P, int64_t cap) {
    hipError_t e;
    if ((e = hipMalloc(&P.pos, 3 * cap * sizeof(double))) != hipSuccess) return e;
    if ((e = hipMalloc(&P.r, (int64_t)MPM_NREAL * cap * sizeof(double))) != hipSuccess) return e;
    if ((e = hipMalloc(&P.idata, (int64_t)MPM_NINT * cap * sizeof(int))) != hipSuccess) return e;
    if ((e = hipMalloc(&P.id, cap * sizeof(int64_t))) != hipSuccess) return e;
    P.cap = cap;
    return hipSuccess;
}

int upload_udf_tables(mpm_ctx *c, double t) {
    // compute_udf_wall_vel_at_nodes (nodal_data_ops.cpp:929-1059), 3-D.
    for (int d = 0; d < 3; ++d)
        for (int h = 0; h < 2; ++h) {
            if (!c->udf_fn[d][h]) continue;
            const int p0 = (d == 0) ? 1 : 0;
            const int p1 = (d == 2) ? 1 : 2;
            std::vector<double> tab((size_t)(c->g.n[p0] + 1) * (c->g.n[p1] + 1) * 3, 0.0);
            for (int j = 0; j <= c->g.n[p0]; ++j)
                for (int k = 0; k <= c->g.n[p1]; ++k) {
                    double co[3] = {0, 0, 0};
                    co[d] = h ? c->s.phi[d] : c->s.plo[d];
                    co[p0] = c->s.plo[p0] + j * c->g.dx[p0];
                    co[p1] = c->s.plo[p1] + k * c->g.dx[p1];
                    double vel[3] = {0, 0, 0};
                    c->udf_fn[d][h](co[0], co[1], co[2], t, vel);
                    const size_t f = ((size_t)j * (c->g.n[p1] + 1) + k) * 3;
                    for (int q = 0; q < 3; ++q) tab[f + q] = vel[q];
                }
            HIPCHK(hipMemcpyAsync(c->udf_dev[d][h], tab.data(), tab.size() * sizeof(double),
                                  hipMemcpyHostToDevice, c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
        }
    return MPM_OK;
}

}  // namespace

extern "C" {

const char *mpm_last_error(void) { return g_err; }

void mpm_specs_default(mpm_specs *s) {
    std::memset(s, 0, sizeof(*s));
    // mpm_specs.H:78-220 and read_mpm_specs defaults (:223-565)
    s->gravity[1] = -9.81;
    s->num_redist = 10;
    s->order_scheme = 1;
    s->mass_tolerance = 1e-8;
    s->alpha_pic_flip = 0.95;
    s->stress_update_scheme = MPM_USL;
    s->CFL = 0.1;
    s->dt_min_limit = 1e-12;
    s->dt_max_limit = 1.0;
    for (int d = 0; d < 3; ++d) {
        s->bclo[d] = MPM_BC_SLIP;
        s->bchi[d] = MPM_BC_SLIP;
        s->wall_mu_lo[d] = 0.1;
        s->wall_mu_hi[d] = 0.1;
    }
}

int mpm_order_directional(const mpm_specs *s, int od[3]) {
    if (!s || !od) return MPM_E_ARG;
    // Initialise_Domain (mpm_init.cpp:180-233)
    for (int d = 0; d < 3; ++d) {
        if (s->order_scheme == 1) {
            od[d] = 1;
        } else if (s->order_scheme == 2 || s->order_scheme == 3) {
            const int nc = s->ncells[d];
            od[d] = (s->periodic[d] == 0) ? ((nc < 5) ? 1 : s->order_scheme)
                                          : ((nc < 3) ? 1 : s->order_scheme);
        } else {
            return set_err(MPM_E_ORDER, "order_scheme must be 1, 2 or 3");
        }
    }
    return MPM_OK;
}

int mpm_create(const mpm_specs *s, mpm_ctx **out) {
    if (!s || !out) return set_err(MPM_E_ARG, "null argument");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return set_err(MPM_E_NODEVICE, "no HIP device");
    GridP g;
    PhysP ph;
    int rc = specs_params(s, g, ph);
    if (rc) return rc;
    mpm_ctx *c = new mpm_ctx();
    c->s = *s;
    c->g = g;
    c->ph = ph;
    for (int d = 0; d < 3; ++d) c->any_per |= (g.per[d] != 0);
    std::memset(&c->ls, 0, sizeof(c->ls));
    std::memset(&c->udf, 0, sizeof(c->udf));
    std::memset(&c->P, 0, sizeof(c->P));
    std::memset(&c->Q, 0, sizeof(c->Q));
    std::memset(&c->N, 0, sizeof(c->N));
    std::memset(&c->S, 0, sizeof(c->S));
    c->skip_ok = walls_at_rest(s);

    auto fail = [&](int code) {
        mpm_destroy(c);
        return code;
    };
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
        return fail(set_err(MPM_E_HIP, "hipStreamCreate failed"));
    for (int q = 0; q < 8; ++q)
        if (hipEventCreate(&c->ev[q]) != hipSuccess) return fail(set_err(MPM_E_HIP, "event"));
    const int64_t nn = g.nn;
    NState &N = c->N;
    if (hipMalloc(&N.acc, 7 * nn * sizeof(double)) != hipSuccess ||
        hipMalloc(&N.mass, nn * sizeof(double)) != hipSuccess ||
        hipMalloc(&N.vel, 3 * nn * sizeof(double)) != hipSuccess ||
        hipMalloc(&N.frc, 3 * nn * sizeof(double)) != hipSuccess ||
        hipMalloc(&N.dvel, 3 * nn * sizeof(double)) != hipSuccess ||
        hipMalloc(&N.mold, nn * sizeof(double)) != hipSuccess ||
        hipMalloc(&c->sc, sizeof(Scalars)) != hipSuccess ||
        hipMalloc(&c->d_energy, 8 * sizeof(double)) != hipSuccess)
        return fail(set_err(MPM_E_OOM, "nodal allocation failed"));
    SortState &S = c->S;
    S.ntiles = g.ntc[0] * g.ntc[1] * g.ntc[2];
    S.nkeys = S.ntiles * kTileCells * kTileCells * kTileCells;
    S.scan_bytes = sort_scan_bytes(S.nkeys);
    const int ntn = g.ntn[0] * g.ntn[1] * g.ntn[2];
    if (hipMalloc(&S.count, (S.nkeys + 1) * sizeof(int)) != hipSuccess ||
        hipMalloc(&S.cell_start, (S.nkeys + 1) * sizeof(int)) != hipSuccess ||
        hipMalloc(&S.scan_tmp, S.scan_bytes > 0 ? S.scan_bytes : 16) != hipSuccess ||
        hipMalloc(&S.tile_epoch, ntn * sizeof(int)) != hipSuccess ||
        hipMalloc(&S.tiles, (S.ntiles + 1) * sizeof(int)) != hipSuccess)
        return fail(set_err(MPM_E_OOM, "sort allocation failed"));
    if (hipMemsetAsync(S.count, 0, (S.nkeys + 1) * sizeof(int), c->stream) != hipSuccess ||
        hipMemsetAsync(S.cell_start, 0, (S.nkeys + 1) * sizeof(int), c->stream) != hipSuccess)
        return fail(set_err(MPM_E_HIP, "sort init failed"));
    {
        std::vector<int> h(ntn, -1000);
        if (hipMemcpy(S.tile_epoch, h.data(), ntn * sizeof(int), hipMemcpyHostToDevice) !=
            hipSuccess)
            return fail(set_err(MPM_E_HIP, "sort init failed"));
    }
    launch_zero(N.acc, 7 * nn, c->stream);
    launch_zero(N.mass, nn, c->stream);
    launch_zero(N.vel, 3 * nn, c->stream);
    launch_zero(N.frc, 3 * nn, c->stream);
    launch_zero(N.dvel, 3 * nn, c->stream);
    launch_zero(N.mold, nn, c->stream);
    launch_zero(reinterpret_cast<double *>(c->sc), sizeof(Scalars) / sizeof(double), c->stream);
    if (hipStreamSynchronize(c->stream) != hipSuccess)
        return fail(set_err(MPM_E_HIP, "init sync failed"));
    *out = c;
    return MPM_OK;
}

void mpm_destroy(mpm_ctx *c) {
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    free_particles(c);
    thermal_free(c);
    (void)hipFree(c->N.acc);
    (void)hipFree(c->N.mass);
    (void)hipFree(c->N.vel);
    (void)hipFree(c->N.frc);
    (void)hipFree(c->N.dvel);
    (void)hipFree(c->N.mold);
    (void)hipFree(c->sc);
    (void)hipFree(c->d_energy);
    (void)hipFree(c->S.count);
    (void)hipFree(c->S.cell_start);
    (void)hipFree(c->S.scan_tmp);
    (void)hipFree(c->S.tile_epoch);
    (void)hipFree(c->S.tiles);
    (void)hipFree(c->det_blk);
    (void)hipFree(c->det_slot);
    (void)hipFree(c->ml_nodes);
    (void)hipFree(c->ml_count);
    (void)hipFree(c->ml_x);
    (void)hipFree(c->ml_feat);
    (void)hipFree(c->ml_y);
    (void)hipFree(c->ml_vc);
    (void)hipFree(c->ml_guard);
    if (c->ml_hflag) (void)hipHostFree(c->ml_hflag);
    (void)hipFree(c->ml_scbak);
    for (int q = 0; q < 2; ++q)
        if (c->ml_ev[q]) (void)hipEventDestroy(c->ml_ev[q]);
    for (double *p : c->ls_bufs) (void)hipFree(p);
    for (int d = 0; d < 3; ++d)
        for (int h = 0; h < 2; ++h) (void)hipFree(c->udf_dev[d][h]);
    for (int q = 0; q < 8; ++q)
        if (c->ev[q]) (void)hipEventDestroy(c->ev[q]);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

}  // extern "C"

namespace {

// Particles into a context: ids[i] the upload index of particle i (NULL:
// 0..np-1); the cold fields by upload index, n_upload of them (NULL: taken
// from rdata / idata, n_upload = np).
int upload_impl(mpm_ctx *c, int64_t np, const double *pos, const double *rdata, const int *idata,
                const int64_t *ids_in, int64_t n_upload, const double *cold_rad,
                const double *cold_yacc, const int *cold_rigid) {
    if (!c || np < 0 || (np > 0 && (!pos || !rdata || !idata)))
        return set_err(MPM_E_ARG, "bad particle arrays");
    for (int64_t i = 0; i < np; ++i) {
        const int m = idata[MPM_I_MODEL * np + i];
        if (m < 0 || m > 2) return set_err(MPM_E_MODEL, "constitutive model must be 0, 1 or 2");
        if (c->s.calculate_strain_based_on_delta == 1 && m != 0 &&
            idata[MPM_I_PHASE * np + i] == 0)
            return set_err(MPM_E_UNSUPPORTED,
                           "delta strain update is implemented for model 0 only");
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    free_particles(c);
    if (np >= (int64_t)1 << 31) return set_err(MPM_E_ARG, "at most 2^31-1 particles");
    const int64_t cap = np > 0 ? np : 1;
    PState &P = c->P;
    HIPCHK(alloc_pstate(P, cap));
    HIPCHK(alloc_pstate(c->Q, cap));
    HIPCHK(hipMalloc(&c->S.key, cap * sizeof(int)));
    HIPCHK(hipMalloc(&c->S.rank, cap * sizeof(int)));
    HIPCHK(hipMalloc(&c->S.perm, cap * sizeof(int)));
    // No packets: the tile deposit gathers the particle fields (slab loads
    // included) through the permutation.
    c->S.pkt = nullptr;
    c->S.pcap = cap;
    if (c->any_per && c->s.stress_update_scheme == MPM_MUSL)
        HIPCHK(hipMalloc(&P.vsnap, 3 * cap * sizeof(double)));
    c->Q.vsnap = P.vsnap;
    P.np = np;
    c->Q.np = np;
    if (cold_rad) {
        c->cold_radius.assign(cold_rad, cold_rad + n_upload);
        c->cold_yacc.assign(cold_yacc, cold_yacc + n_upload);
        c->cold_rigid.assign(cold_rigid, cold_rigid + n_upload);
    } else {
        c->cold_radius.assign(rdata + (size_t)MPM_R_RADIUS * np,
                              rdata + (size_t)MPM_R_RADIUS * np + np);
        c->cold_yacc.assign(rdata + (size_t)MPM_R_YACC * np, rdata + (size_t)MPM_R_YACC * np + np);
        c->cold_rigid.assign(idata + (size_t)MPM_I_RIGID_ID * np,
                             idata + (size_t)MPM_I_RIGID_ID * np + np);
    }
    c->all_dirty = true;
    c->dt_next_valid = false;
    if (np > 0) {
        HIPCHK(hipMemcpy(P.pos, pos, 3 * np * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(P.r, rdata, (int64_t)MPM_NREAL * np * sizeof(double),
                         hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(P.idata, idata, (int64_t)MPM_NINT * np * sizeof(int),
                         hipMemcpyHostToDevice));
        std::vector<int64_t> ids(np);
        if (ids_in) ids.assign(ids_in, ids_in + np);
        else std::iota(ids.begin(), ids.end(), (int64_t)0);
        HIPCHK(hipMemcpy(P.id, ids.data(), np * sizeof(int64_t), hipMemcpyHostToDevice));
        // InitParticles ends with Redistribute() (mpm_init.cpp:1166).
        launch_redistribute(c->g, P, c->stream);
        int rc = check_launch();
        if (rc) return rc;
    }
    c->dirty = false;
    HIPCHK(hipStreamSynchronize(c->stream));
    return thermal_on_upload(c);
}

}  // namespace

extern "C" {

int mpm_upload_particles(mpm_ctx *c, int64_t np, const double *pos, const double *rdata,
                         const int *idata) {
    return upload_impl(c, np, pos, rdata, idata, nullptr, np, nullptr, nullptr, nullptr);
}

// writeCheckpointFile (mpm_particle_outputs.cpp:355-499) and WriteHeader
// (:311-353): dir/Header as the reference writes it; the particles (the
// reference's AMReX particle checkpoint) as dir/particles.bin:
//   "XMPMCHK1", int64 n_live, n_upload, MPM_NREAL, MPM_NINT, then pos[3][n],
//   rdata[MPM_NREAL][n], int32 idata[MPM_NINT][n], int64 id[n] of the live
//   particles by upload index, then the cold fields by upload index:
//   radius[n_upload], yacceleration[n_upload], int32 rigid_body_id[n_upload].
int mpm_write_checkpoint(mpm_ctx *c, const char *dir, int output_it) {
    if (!c || !dir) return set_err(MPM_E_ARG, "null argument");
    if (mkdir(dir, 0755) != 0 && errno != EEXIST)
        return set_err(MPM_E_ARG, "cannot create checkpoint directory");
    double time = 0.0, last_dt = 0.0;
    int64_t steps = 0;
    int rc = mpm_get_time(c, &time, &steps, &last_dt);
    if (rc) return rc;
    {
        std::ofstream h(std::string(dir) + "/Header", std::ios::out | std::ios::trunc);
        if (!h.good()) return set_err(MPM_E_ARG, "cannot write checkpoint Header");
        h.precision(17);
        h << "Checkpoint version: 1\n" << steps << "\n" << output_it << "\n" << time << "\n";
        if (!h.good()) return set_err(MPM_E_ARG, "cannot write checkpoint Header");
    }
    int64_t n = 0;
    rc = mpm_num_particles(c, &n);
    if (rc) return rc;
    std::vector<double> pos(3 * n), r((size_t)MPM_NREAL * n);
    std::vector<int> idata((size_t)MPM_NINT * n);
    std::vector<int64_t> id(n);
    if (n) {
        rc = mpm_download_particles(c, pos.data(), r.data(), idata.data(), id.data());
        if (rc) return rc;
    }
    std::ofstream f(std::string(dir) + "/particles.bin",
                    std::ios::out | std::ios::trunc | std::ios::binary);
    if (!f.good()) return set_err(MPM_E_ARG, "cannot write checkpoint particles");
    const int64_t hdr[4] = {n, (int64_t)c->cold_radius.size(), MPM_NREAL, MPM_NINT};
    f.write("XMPMCHK1", 8);
    f.write(reinterpret_cast<const char *>(hdr), sizeof(hdr));
    f.write(reinterpret_cast<const char *>(pos.data()), pos.size() * sizeof(double));
    f.write(reinterpret_cast<const char *>(r.data()), r.size() * sizeof(double));
    f.write(reinterpret_cast<const char *>(idata.data()), idata.size() * sizeof(int));
    f.write(reinterpret_cast<const char *>(id.data()), id.size() * sizeof(int64_t));
    f.write(reinterpret_cast<const char *>(c->cold_radius.data()),
            c->cold_radius.size() * sizeof(double));
    f.write(reinterpret_cast<const char *>(c->cold_yacc.data()),
            c->cold_yacc.size() * sizeof(double));
    f.write(reinterpret_cast<const char *>(c->cold_rigid.data()),
            c->cold_rigid.size() * sizeof(int));
    if (!f.good()) return set_err(MPM_E_ARG, "cannot write checkpoint particles");
    if (!c->thermal) return MPM_OK;
    // The USE_TEMP build's checkpoint carries realData 46..52 with the rest
    // of the record: here a file of their own, by upload index.
    const int64_t nu = (int64_t)c->cold_radius.size();
    std::vector<double> th((size_t)MPM_NTHERMAL * nu);
    if (nu) {
        rc = mpm_download_thermal(c, nu, th.data());
        if (rc) return rc;
    }
    std::ofstream t(std::string(dir) + "/thermal.bin",
                    std::ios::out | std::ios::trunc | std::ios::binary);
    const int64_t thdr[2] = {nu, MPM_NTHERMAL};
    t.write("XMPMTHM1", 8);
    t.write(reinterpret_cast<const char *>(thdr), sizeof(thdr));
    t.write(reinterpret_cast<const char *>(th.data()), th.size() * sizeof(double));
    if (!t.good()) return set_err(MPM_E_ARG, "cannot write checkpoint thermal fields");
    return MPM_OK;
}

// writeAsciiFiles (mpm_particle_outputs.cpp:188-196, called with 6 digits by
// utilities.cpp:94-99): file "<prefix>_t<time, fixed, 6 decimals>".  The
// body stands in for AMReX's WriteAsciiFile, absent from the reference tree:
// the particle count, then one line per live particle (upload-index order):
// x y z id cpu, the MPM_NREAL real fields in record order, the MPM_NINT
// integer fields; reals to 17 significant digits.
int mpm_write_ascii_particles(mpm_ctx *c, const char *prefix, double time, char *path_out,
                              int64_t path_len) {
    if (!c || !prefix) return set_err(MPM_E_ARG, "null argument");
    std::ostringstream name;
    name << prefix << "_t" << std::fixed << std::setprecision(6) << time;
    const std::string path = name.str();
    if (path_out) {
        if (path_len < (int64_t)path.size() + 1) return set_err(MPM_E_ARG, "path buffer too small");
        std::memcpy(path_out, path.c_str(), path.size() + 1);
    }
    int64_t n = 0;
    int rc = mpm_num_particles(c, &n);
    if (rc) return rc;
    std::vector<double> pos(3 * n), r((size_t)MPM_NREAL * n);
    std::vector<int> idata((size_t)MPM_NINT * n);
    std::vector<int64_t> id(n);
    if (n) {
        rc = mpm_download_particles(c, pos.data(), r.data(), idata.data(), id.data());
        if (rc) return rc;
    }
    FILE *f = std::fopen(path.c_str(), "w");
    if (!f) return set_err(MPM_E_ARG, "cannot open the ASCII particle file");
    std::fprintf(f, "%lld\n", (long long)n);
    for (int64_t i = 0; i < n; ++i) {
        for (int d = 0; d < 3; ++d) std::fprintf(f, "%.17g ", pos[d * n + i]);
        std::fprintf(f, "%lld 0", (long long)id[i]);
        for (int k = 0; k < MPM_NREAL; ++k) std::fprintf(f, " %.17g", r[(size_t)k * n + i]);
        for (int k = 0; k < MPM_NINT; ++k) std::fprintf(f, " %d", idata[(size_t)k * n + i]);
        std::fprintf(f, "\n");
    }
    const bool ok = std::ferror(f) == 0;
    std::fclose(f);
    return ok ? MPM_OK : set_err(MPM_E_ARG, "write failed");
}

// readCheckpointFile (mpm_particle_outputs.cpp:537-589): step, output index
// and time from dir/Header (either Header variant of WriteHeader: with or
// without the EB line), the particles from dir/particles.bin.  As in the
// reference, the caller runs Initialise_Internal_Forces next.  A context with
// the thermal module on needs dir/thermal.bin.
int mpm_read_checkpoint(mpm_ctx *c, const char *dir, int *output_it) {
    if (!c || !dir) return set_err(MPM_E_ARG, "null argument");
    std::ifstream h(std::string(dir) + "/Header");
    if (!h.good()) return set_err(MPM_E_ARG, "cannot open checkpoint Header");
    std::string line;
    std::getline(h, line);
    if (line != "Checkpoint version: 1") return set_err(MPM_E_ARG, "not a checkpoint Header");
    int64_t steps = 0;
    int oit = 0;
    double time = 0.0;
    if (!(h >> steps >> oit)) return set_err(MPM_E_ARG, "bad checkpoint Header");
    // The time is the last line: the reference's default EB build writes
    // EB_generate_max_level (an integer) before it (:381-385), a build
    // without EB (and mpm_write_checkpoint) does not.
    std::vector<std::string> rest;
    while (std::getline(h, line))
        if (line.find_first_not_of(" \t\r") != std::string::npos) rest.push_back(line);
    if (rest.empty() || rest.size() > 2) return set_err(MPM_E_ARG, "bad checkpoint Header");
    {
        std::istringstream ts(rest.back());
        if (!(ts >> time)) return set_err(MPM_E_ARG, "bad checkpoint Header time");
    }
    if (c->thermal) {
        std::ifstream probe(std::string(dir) + "/thermal.bin", std::ios::binary);
        if (!probe.good())
            return set_err(MPM_E_ARG, "the thermal module is on but the checkpoint holds no "
                                      "thermal fields (thermal.bin)");
    }
    std::ifstream f(std::string(dir) + "/particles.bin", std::ios::binary);
    char magic[8];
    int64_t hdr[4];
    if (!f.read(magic, 8) || std::memcmp(magic, "XMPMCHK1", 8) != 0 ||
        !f.read(reinterpret_cast<char *>(hdr), sizeof(hdr)))
        return set_err(MPM_E_ARG, "not a checkpoint particle file");
    const int64_t n = hdr[0], nu = hdr[1];
    if (n < 0 || nu < n || hdr[2] != MPM_NREAL || hdr[3] != MPM_NINT)
        return set_err(MPM_E_ARG, "checkpoint particle layout differs from this build");
    std::vector<double> pos(3 * n), r((size_t)MPM_NREAL * n), crad(nu), cyacc(nu);
    std::vector<int> idata((size_t)MPM_NINT * n), crig(nu);
    std::vector<int64_t> id(n);
    f.read(reinterpret_cast<char *>(pos.data()), pos.size() * sizeof(double));
    f.read(reinterpret_cast<char *>(r.data()), r.size() * sizeof(double));
    f.read(reinterpret_cast<char *>(idata.data()), idata.size() * sizeof(int));
    f.read(reinterpret_cast<char *>(id.data()), id.size() * sizeof(int64_t));
    f.read(reinterpret_cast<char *>(crad.data()), crad.size() * sizeof(double));
    f.read(reinterpret_cast<char *>(cyacc.data()), cyacc.size() * sizeof(double));
    f.read(reinterpret_cast<char *>(crig.data()), crig.size() * sizeof(int));
    if (!f.good()) return set_err(MPM_E_ARG, "truncated checkpoint particle file");
    for (int64_t i = 0; i < n; ++i)
        if (id[i] < 0 || id[i] >= nu) return set_err(MPM_E_ARG, "bad particle id in checkpoint");
    int rc = upload_impl(c, n, pos.data(), r.data(), idata.data(), id.data(), nu, crad.data(),
                         cyacc.data(), crig.data());
    if (rc) return rc;
    std::ifstream t(std::string(dir) + "/thermal.bin", std::ios::binary);
    if (c->thermal && t.good()) {
        int64_t thdr[2];
        if (!t.read(magic, 8) || std::memcmp(magic, "XMPMTHM1", 8) != 0 ||
            !t.read(reinterpret_cast<char *>(thdr), sizeof(thdr)) || thdr[0] != nu ||
            thdr[1] != MPM_NTHERMAL)
            return set_err(MPM_E_ARG, "checkpoint thermal file does not match its particles");
        std::vector<double> th((size_t)MPM_NTHERMAL * nu);
        if (!t.read(reinterpret_cast<char *>(th.data()), th.size() * sizeof(double)))
            return set_err(MPM_E_ARG, "truncated checkpoint thermal file");
        rc = mpm_upload_thermal(c, nu, th.data());
        if (rc) return rc;
    }
    c->steps = steps;
    Scalars sc;
    std::memset(&sc, 0, sizeof(sc));
    sc.time = time;
    sc.dtmin_bits = sc.dtmin_next_bits =
        (unsigned long long)0x7ff0000000000000ull;  // +inf: no candidates yet
    HIPCHK(hipMemcpy(c->sc, &sc, sizeof(Scalars), hipMemcpyHostToDevice));
    c->time_host = time;
    c->last_dt = 0.0;
    c->dt_next_valid = false;
    c->all_dirty = true;
    if (output_it) *output_it = oit;
    return MPM_OK;
}

int mpm_num_particles(mpm_ctx *c, int64_t *np) {
    if (!c || !np) return set_err(MPM_E_ARG, "null argument");
    HIPCHK(hipStreamSynchronize(c->stream));
    std::vector<int64_t> ids(c->P.np);
    if (c->P.np)
        HIPCHK(hipMemcpy(ids.data(), c->P.id, c->P.np * sizeof(int64_t), hipMemcpyDeviceToHost));
    int64_t n = 0;
    for (int64_t v : ids) n += (v >= 0);
    *np = n;
    return MPM_OK;
}

int mpm_download_particles(mpm_ctx *c, double *pos, double *rdata, int *idata, int64_t *id) {
    if (!c) return set_err(MPM_E_ARG, "null context");
    HIPCHK(hipStreamSynchronize(c->stream));
    const int64_t np = c->P.np, cap = c->P.cap;
    if (np == 0) return MPM_OK;
    std::vector<int64_t> ids(np);
    HIPCHK(hipMemcpy(ids.data(), c->P.id, np * sizeof(int64_t), hipMemcpyDeviceToHost));
    std::vector<int64_t> order;
    order.reserve(np);
    for (int64_t i = 0; i < np; ++i)
        if (ids[i] >= 0) order.push_back(i);
    std::sort(order.begin(), order.end(),
              [&](int64_t a, int64_t b) { return ids[a] < ids[b]; });
    const int64_t nl = (int64_t)order.size();
    if (pos) {
        std::vector<double> h(3 * cap);
        HIPCHK(hipMemcpy(h.data(), c->P.pos, 3 * cap * sizeof(double), hipMemcpyDeviceToHost));
        for (int d = 0; d < 3; ++d)
            for (int64_t q = 0; q < nl; ++q) pos[d * nl + q] = h[d * cap + order[q]];
    }
    if (rdata) {
        std::vector<double> h((size_t)MPM_NREAL * cap);
        HIPCHK(hipMemcpy(h.data(), c->P.r, h.size() * sizeof(double), hipMemcpyDeviceToHost));
        for (int f = 0; f < MPM_NREAL; ++f)
            for (int64_t q = 0; q < nl; ++q) rdata[f * nl + q] = h[f * cap + order[q]];
        // cold fields from the host copies, by upload index
        for (int64_t q = 0; q < nl; ++q) {
            rdata[MPM_R_RADIUS * nl + q] = c->cold_radius[ids[order[q]]];
            rdata[MPM_R_YACC * nl + q] = c->cold_yacc[ids[order[q]]];
        }
    }
    if (idata) {
        std::vector<int> h((size_t)MPM_NINT * cap);
        HIPCHK(hipMemcpy(h.data(), c->P.idata, h.size() * sizeof(int), hipMemcpyDeviceToHost));
        for (int f = 0; f < MPM_NINT; ++f)
            for (int64_t q = 0; q < nl; ++q) idata[f * nl + q] = h[f * cap + order[q]];
        for (int64_t q = 0; q < nl; ++q)
            idata[MPM_I_RIGID_ID * nl + q] = c->cold_rigid[ids[order[q]]];
    }
    if (id)
        for (int64_t q = 0; q < nl; ++q) id[q] = ids[order[q]];
    return MPM_OK;
}

int mpm_get_particle_field(mpm_ctx *c, int field, int64_t n_upload, double *out) {
    if (!c || !out || field < 0 || field >= MPM_NREAL) return set_err(MPM_E_ARG, "bad argument");
    if (n_upload != (int64_t)c->cold_radius.size())
        return set_err(MPM_E_ARG, "n_upload differs from the uploaded particle count");
    HIPCHK(hipStreamSynchronize(c->stream));
    const int64_t np = c->P.np, cap = c->P.cap;
    for (int64_t k = 0; k < n_upload; ++k) out[k] = std::numeric_limits<double>::quiet_NaN();
    if (np == 0) return MPM_OK;
    std::vector<int64_t> ids(np);
    HIPCHK(hipMemcpy(ids.data(), c->P.id, np * sizeof(int64_t), hipMemcpyDeviceToHost));
    if (field == MPM_R_RADIUS || field == MPM_R_YACC) {
        const std::vector<double> &cold = field == MPM_R_RADIUS ? c->cold_radius : c->cold_yacc;
        for (int64_t i = 0; i < np; ++i)
            if (ids[i] >= 0) out[ids[i]] = cold[ids[i]];
        return MPM_OK;
    }
    std::vector<double> h(np);
    HIPCHK(hipMemcpy(h.data(), c->P.r + (int64_t)field * cap, np * sizeof(double),
                     hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < np; ++i)
        if (ids[i] >= 0) out[ids[i]] = h[i];
    return MPM_OK;
}

int mpm_get_nodal(mpm_ctx *c, int comp, double *out) {
    const int ncomp = c && c->thermal ? (int)MPM_N_NCOMP_TEMP : (int)MPM_N_NCOMP;
    if (!c || !out || comp < 0 || comp >= ncomp) return set_err(MPM_E_ARG, "bad comp");
    HIPCHK(hipStreamSynchronize(c->stream));
    const int64_t nn = c->g.nn;
    const double *src = nullptr;
    if (comp >= MPM_N_MASS_SPHEAT) src = c->tn + (comp - MPM_N_MASS_SPHEAT) * nn;
    else if (comp == MPM_N_MASS) src = c->N.mass;
    else if (comp >= MPM_N_VEL && comp < MPM_N_VEL + 3) src = c->N.vel + (comp - MPM_N_VEL) * nn;
    else if (comp >= MPM_N_FRC && comp < MPM_N_FRC + 3) src = c->N.frc + (comp - MPM_N_FRC) * nn;
    else if (comp >= MPM_N_DVEL && comp < MPM_N_DVEL + 3)
        src = c->N.dvel + (comp - MPM_N_DVEL) * nn;
    else if (comp == MPM_N_MASS_OLD) src = c->N.mold;
    if (!src) {
        std::memset(out, 0, nn * sizeof(double));
        return MPM_OK;
    }
    HIPCHK(hipMemcpy(out, src, nn * sizeof(double), hipMemcpyDeviceToHost));
    return MPM_OK;
}

int mpm_set_nodal(mpm_ctx *c, int comp, const double *in) {
    const int ncomp = c && c->thermal ? (int)MPM_N_NCOMP_TEMP : (int)MPM_N_NCOMP;
    if (!c || !in || comp < 0 || comp >= ncomp) return set_err(MPM_E_ARG, "bad comp");
    const int64_t nn = c->g.nn;
    double *dst = nullptr;
    if (comp >= MPM_N_MASS_SPHEAT) dst = c->tn + (comp - MPM_N_MASS_SPHEAT) * nn;
    else if (comp == MPM_N_MASS) dst = c->N.mass;
    else if (comp >= MPM_N_VEL && comp < MPM_N_VEL + 3) dst = c->N.vel + (comp - MPM_N_VEL) * nn;
    else if (comp >= MPM_N_FRC && comp < MPM_N_FRC + 3) dst = c->N.frc + (comp - MPM_N_FRC) * nn;
    else if (comp >= MPM_N_DVEL && comp < MPM_N_DVEL + 3)
        dst = c->N.dvel + (comp - MPM_N_DVEL) * nn;
    else if (comp == MPM_N_MASS_OLD) dst = c->N.mold;
    if (!dst) return set_err(MPM_E_ARG, "nodal component is not stored");
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipMemcpy(dst, in, nn * sizeof(double), hipMemcpyHostToDevice));
    c->all_dirty = true;  // the next grid pass visits every node
    return MPM_OK;
}

int mpm_reset_nodal(mpm_ctx *c) {
    if (!c) return set_err(MPM_E_ARG, "null context");
    c->all_dirty = true;
    const int64_t nn = c->g.nn;
    launch_zero(c->N.mass, nn, c->stream);
    launch_zero(c->N.vel, 3 * nn, c->stream);
    launch_zero(c->N.frc, 3 * nn, c->stream);
    launch_zero(c->N.dvel, 3 * nn, c->stream);
    launch_zero(c->N.mold, nn, c->stream);
    return check_launch();
}

int mpm_p2g(mpm_ctx *c, int um, int uv, int uf) {
    if (!c) return set_err(MPM_E_ARG, "null context");
    if (uf < 0 || uf > 2) return set_err(MPM_E_ARG, "update_forces must be 0, 1 or 2");
    c->all_dirty = true;
    // update_forces == 2 also zeroes nodal component 15 and divides it by the
    // mass (:398-400, 617-625), but no deposit writes that component, so it
    // stays zero: the step is that of update_forces == 1 (component 15 is not
    // stored and reads as zero).
    if (uf == 2) uf = 1;
    // The reference zeroes the selected components first (:368-403); the
    // accumulators are already zero.
    launch_p2g(c->g, c->ph, c->P, c->N, um, uv, uf, nullptr, c->stream);
    int flags = (um ? G_FINISH_MASS : 0) | (uv ? G_FINISH_VEL : 0) |
                (uf ? (G_FINISH_FRC | G_KEEP_FRC) : 0);
    if (flags) launch_grid(c->g, c->ph, c->ls, c->udf, c->N, flags, 0.0, nullptr, c->stream);
    return check_launch();
}

int mpm_backup_velocity(mpm_ctx *c) {
    if (!c) return set_err(MPM_E_ARG, "null context");
    c->all_dirty = true;
    launch_grid(c->g, c->ph, c->ls, c->udf, c->N, G_BACKUP, 0.0, nullptr, c->stream);
    return check_launch();
}

int mpm_nodal_update(mpm_ctx *c, double dt) {
    if (!c) return set_err(MPM_E_ARG, "null context");
    c->all_dirty = true;
    launch_grid(c->g, c->ph, c->ls, c->udf, c->N, G_UPDATE, dt, nullptr, c->stream);
    return check_launch();
}

int mpm_apply_nodal_bcs(mpm_ctx *c, double dt, double t) {
    if (!c) return set_err(MPM_E_ARG, "null context");
    c->all_dirty = true;
    if (c->any_udf) {
        int rc = upload_udf_tables(c, t);
        if (rc) return rc;
    }
    launch_grid(c->g, c->ph, c->ls, c->udf, c->N, G_BCS, dt, nullptr, c->stream);
    return check_launch();
}

int mpm_g2p(mpm_ctx *c, int uv, int us, double dt) {
    if (!c) return set_err(MPM_E_ARG, "null context");
    c->dt_next_valid = false;
    int flags = (uv ? P_G2P_VEL : 0) | (us ? P_G2P_GRAD : 0);
    if (flags)
        launch_particle(c->g, c->ph, c->ls, c->udf, c->P, c->N, flags, dt, 0.0, nullptr,
                        c->stream);
    return check_launch();
}

int mpm_move(mpm_ctx *c, double dt, int step) {
    if (!c) return set_err(MPM_E_ARG, "null context");
    c->dt_next_valid = false;
    const int nr = c->s.num_redist > 0 ? c->s.num_redist : 1;
    const bool redist = (step % nr) == 0;
    launch_particle(c->g, c->ph, c->ls, c->udf, c->P, c->N, P_MOVE | (redist ? P_REDIST : 0), dt,
                    0.0, nullptr, c->stream);
    c->dirty = !redist;
    return check_launch();
}

int mpm_redistribute(mpm_ctx *c, int step) {
    if (!c) return set_err(MPM_E_ARG, "null context");
    const int nr = c->s.num_redist > 0 ? c->s.num_redist : 1;
    if (step % nr != 0) return MPM_OK;
    launch_redistribute(c->g, c->P, c->stream);
    c->dirty = false;
    c->dt_next_valid = false;  // removals change the dt reduction's set
    return check_launch();
}

int mpm_update_volume(mpm_ctx *c) {
    if (!c) return set_err(MPM_E_ARG, "null context");
    c->dt_next_valid = false;
    launch_particle(c->g, c->ph, c->ls, c->udf, c->P, c->N, P_VOLUME, 0.0, 0.0, nullptr,
                    c->stream);
    return check_launch();
}

int mpm_stress(mpm_ctx *c, double dt, double time) {
    if (!c) return set_err(MPM_E_ARG, "null context");
    c->dt_next_valid = false;
    launch_particle(c->g, c->ph, c->ls, c->udf, c->P, c->N, P_STRESS, dt, time, nullptr,
                    c->stream);
    return check_launch();
}

int mpm_calculate_dt(mpm_ctx *c, double *dt) {
    if (!c || !dt) return set_err(MPM_E_ARG, "null argument");
    launch_dt(c->g, c->ph, c->P, c->sc, c->s.fixed_timestep == 1, c->s.timestep, c->stream);
    int rc = check_launch();
    if (rc) return rc;
    Scalars h;
    HIPCHK(hipMemcpyAsync(&h, c->sc, sizeof(Scalars), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    *dt = h.dt;
    return MPM_OK;
}

int mpm_init_internal_forces(mpm_ctx *c) {
    if (!c) return set_err(MPM_E_ARG, "null context");
    c->all_dirty = true;
    c->dt_next_valid = false;
    // Initialise_Internal_Forces (mpm_init.cpp:368-401): P2G(1,1,1), backup,
    // G2P(update_vel=0, update_strainrate=1, dt=0), stress(dt=0, applied rate).
    // The sorted deposit (particles outside their cell go through the global
    // path) replaces the atomic one-thread-per-particle deposit when enabled.
    if (c->sorted && c->P.np > 0) {
        det_sync(c);
        launch_sort(c->g, c->ph, c->P, c->S, nullptr, 0, c->stream);
        launch_p2g_tiles(c->g, c->ph, c->P, c->S, c->N, (int)c->steps, 0, 1, c->sc, c->stream);
    } else {
        launch_p2g(c->g, c->ph, c->P, c->N, 1, 1, 1, nullptr, c->stream);
    }
    launch_grid(c->g, c->ph, c->ls, c->udf, c->N,
                G_FINISH_MASS | G_FINISH_VEL | G_FINISH_FRC | G_KEEP_FRC | G_BACKUP, 0.0, nullptr,
                c->stream);
    launch_particle(c->g, c->ph, c->ls, c->udf, c->P, c->N, P_G2P_GRAD | P_STRESS_INIT, 0.0, 0.0,
                    nullptr, c->stream);
    int rc = check_launch();
    if (rc) return rc;
    return thermal_init(c, c->sorted && c->P.np > 0);
}

int mpm_step(mpm_ctx *c, int nsteps) {
    if (!c || nsteps < 0) return set_err(MPM_E_ARG, "bad argument");
    const int nr = c->s.num_redist > 0 ? c->s.num_redist : 1;
    const bool musl = c->s.stress_update_scheme == MPM_MUSL;
    hipStream_t st = c->stream;
    const bool prof = c->profile != 0;
    double kms[4] = {0, 0, 0, 0};
    HIPCHK(hipEventRecord(c->ev[0], st));
    for (int it = 0; it < nsteps; ++it) {
        c->steps++;
        const int step = (int)c->steps;
        const bool redist = (step % nr) == 0;
        // Redistribute_Fill_Update at the top of the step (main.cpp:88).
        // The previous fast particle kernel already reduced this step's dt
        // candidates unless a redistribution (which may remove particles)
        // runs first.
        const bool fused_dt = c->dt_next_valid && !(redist && c->dirty);
        if (redist && c->dirty) launch_redistribute(c->g, c->P, st);
        if (musl && c->P.vsnap) launch_snapshot_vel(c->P, st);
        // The previous step of this call ends here: its dt advances the time
        // inside the dt launch; the last step's advance follows the loop.
        // Without a reduction over the particles, and with no host read of
        // the time before the sort, the sort's first launch computes the dt.
        const bool temp_udf = c->thermal && c->any_tudf;
        const bool fixed_dt = c->s.fixed_timestep == 1;
        DtFold fold;
        if (c->sorted && c->P.np > 0 && !(c->any_udf || temp_udf) && (fixed_dt || fused_dt)) {
            fold.on = 1;
            fold.fixed = fixed_dt ? 1 : 0;
            fold.fixed_dt = c->s.timestep;
            fold.advance = it > 0 ? 1 : 0;
            fold.sc = c->sc;
        } else {
            launch_dt(c->g, c->ph, c->P, c->sc, fixed_dt, c->s.timestep, st, fused_dt ? 1 : 0,
                      it > 0 ? 1 : 0);
        }
        if (c->any_udf || temp_udf) {
            Scalars h;
            HIPCHK(hipMemcpyAsync(&h, c->sc, sizeof(Scalars), hipMemcpyDeviceToHost, st));
            HIPCHK(hipStreamSynchronize(st));
            int rc = c->any_udf ? upload_udf_tables(c, h.time) : MPM_OK;
            if (!rc && temp_udf) rc = thermal_udf_tables(c, h.time);
            if (rc) return rc;
        }
        const bool sorted = c->sorted != 0;
        if (prof) HIPCHK(hipEventRecord(c->ev[1], st));
        if (sorted) {
            det_sync(c);
            launch_sort(c->g, c->ph, c->P, c->S, c->Q.id, 0, st, fold);
        }
        if (prof) HIPCHK(hipEventRecord(c->ev[2], st));
        if (sorted) launch_p2g_tiles(c->g, c->ph, c->P, c->S, c->N, step, 0, nr > 1, c->sc, st);
        else launch_p2g(c->g, c->ph, c->P, c->N, 1, 1, 1, nullptr, st);
        if (prof) HIPCHK(hipEventRecord(c->ev[3], st));
        // Idle node tiles are skipped when their massless state is all-zero;
        // a pass after an out-of-step nodal call visits every node.
        const int *te = (sorted && c->skip_ok && !c->any_udf && !c->all_dirty) ? c->S.tile_epoch
                                                                                : nullptr;
        launch_grid(c->g, c->ph, c->ls, c->udf, c->N, grid_flags_full(c), 0.0, c->sc, st, te,
                    step);
        c->all_dirty = false;
        if (c->thermal) {
            int rc = thermal_grid(c, sorted, nr > 1, st);
            if (rc) return rc;
        }
        if (prof) HIPCHK(hipEventRecord(c->ev[4], st));
        const int tail = P_MOVE | (redist ? P_REDIST : 0) | P_VOLUME | P_STRESS;
        int last = P_G2P_VEL | P_G2P_GRAD | tail;
        if (musl) {
            if (sorted && c->fast && c->g.uniform)
                launch_g2p_vel_tiles(c->g, c->ph, c->P, c->S, c->N, nr > 1 ? 1 : 0, c->sc, st);
            else
                launch_particle(c->g, c->ph, c->ls, c->udf, c->P, c->N, P_G2P_VEL, 0.0, 0.0, c->sc,
                                st);
            // Images of the re-projection P2G carry the velocities of the
            // neighbour fill at the top of the step (updateNeighbors is not
            // called in between, main.cpp:120-140): the tile deposit's image
            // form (P.vsnap set).
            if (!sorted) launch_p2g(c->g, c->ph, c->P, c->N, 0, 1, 0, c->P.vsnap, st);
            else launch_p2g_tiles(c->g, c->ph, c->P, c->S, c->N, step, 1, nr > 1, c->sc, st);
            launch_grid(c->g, c->ph, c->ls, c->udf, c->N, G_FINISH_VEL | G_BCS, 0.0, c->sc, st, te,
                        step);
            last = P_G2P_GRAD | tail;
        }
        if (c->thermal) {
            int rc = thermal_particles(c, musl, sorted ? c->S.perm : nullptr,
                                       sorted ? c->S.cell_start + c->S.nkeys : nullptr, nr > 1,
                                       st);
            if (rc) return rc;
            if (sorted) thermal_swap(c);
        }
        // The tile kernel (uniform order; mixed orders interpolate each
        // velocity component with its own direction's order, in the generic
        // kernel); with redistribution every few steps its form for particles
        // outside their tile (outside the domain) runs.
        if (sorted && c->fast && c->g.uniform) {
            launch_particle_fast(c->g, c->ph, c->ls, c->udf, c->P, c->Q, c->S, c->N, musl ? 0 : 1,
                                 redist ? 1 : 0, nr > 1 ? 1 : 0, c->sc, st);
            std::swap(c->P, c->Q);
            c->dt_next_valid = c->fuse_dt != 0;
        } else if (sorted) {
            launch_particle_sorted(c->g, c->ph, c->ls, c->udf, c->P, c->Q, c->S, c->N, last, c->sc,
                                   st);
            std::swap(c->P, c->Q);
            c->dt_next_valid = false;
        } else {
            launch_particle(c->g, c->ph, c->ls, c->udf, c->P, c->N, last, 0.0, 0.0, c->sc, st);
            c->dt_next_valid = false;
        }
        if (prof) HIPCHK(hipEventRecord(c->ev[6], st));
        c->dirty = !redist;
        int rc = check_launch();
        if (rc) return rc;
        if (prof) {
            HIPCHK(hipEventSynchronize(c->ev[6]));
            float a = 0, b = 0, d = 0, e = 0;
            (void)hipEventElapsedTime(&a, c->ev[2], c->ev[3]);
            (void)hipEventElapsedTime(&b, c->ev[3], c->ev[4]);
            (void)hipEventElapsedTime(&d, c->ev[4], c->ev[6]);
            (void)hipEventElapsedTime(&e, c->ev[1], c->ev[2]);
            kms[0] += a;
            kms[1] += b;
            kms[2] += d;
            kms[3] += e;
        }
    }
    if (nsteps > 0) launch_advance_time(c->sc, st);
    HIPCHK(hipEventRecord(c->ev[5], st));
    HIPCHK(hipEventSynchronize(c->ev[5]));
    float tot = 0;
    (void)hipEventElapsedTime(&tot, c->ev[0], c->ev[5]);
    c->step_ms = tot;
    for (int q = 0; q < 4; ++q) c->kernel_ms[q] = nsteps ? kms[q] / nsteps : 0.0;
    if (c->sorted && nsteps > 0 && c->P.np > 0) {
        int nlive = 0;
        HIPCHK(hipMemcpy(&nlive, c->S.cell_start + c->S.nkeys, sizeof(int), hipMemcpyDeviceToHost));
        c->P.np = c->Q.np = nlive;  // removed particles were dropped by the reorder
    }
    Scalars h;
    HIPCHK(hipMemcpy(&h, c->sc, sizeof(Scalars), hipMemcpyDeviceToHost));
    c->time_host = h.time;
    c->last_dt = h.dt;
    if (h.err) {
        int zero = 0;
        HIPCHK(hipMemcpy(&c->sc->err, &zero, sizeof(int), hipMemcpyHostToDevice));
        if (h.err & 2)
            return set_err(MPM_E_UNSUPPORTED,
                           "a level-set temperature condition chained more than 4 nodes");
        return set_err(MPM_E_UNSUPPORTED, "a particle lay outside its sorted tile");
    }
    return MPM_OK;
}

int mpm_get_time(mpm_ctx *c, double *time, int64_t *steps, double *last_dt) {
    if (!c) return set_err(MPM_E_ARG, "null context");
    HIPCHK(hipStreamSynchronize(c->stream));
    Scalars h;
    HIPCHK(hipMemcpy(&h, c->sc, sizeof(Scalars), hipMemcpyDeviceToHost));
    if (time) *time = h.time;
    if (steps) *steps = c->steps;
    if (last_dt) *last_dt = h.dt;
    return MPM_OK;
}

int mpm_add_levelset_body(mpm_ctx *c, const double *phi, int bc, double mu,
                          const double w[3]) {
    if (!c || !phi) return set_err(MPM_E_ARG, "null argument");
    if (c->ls.nb >= MPM_MAX_LS_BODIES) return set_err(MPM_E_ARG, "too many level-set bodies");
    if (bc < 0 || bc > MPM_BC_OUTFLOW) return set_err(MPM_E_BC, "unknown body BC");
    const GridP &g = c->g;
    const int G0 = g.n[0] + 3, G1 = g.n[1] + 3, G2 = g.n[2] + 3;
    // One ghost layer: 1.0 outside non-periodic faces, periodic image otherwise
    // (nodal_data_ops.cpp:144-153).
    std::vector<double> h((size_t)G0 * G1 * G2, 1.0);
    for (int k = -1; k <= g.n[2] + 1; ++k)
        for (int j = -1; j <= g.n[1] + 1; ++j)
            for (int i = -1; i <= g.n[0] + 1; ++i) {
                int I[3] = {i, j, k}, S[3];
                bool ok = true;
                for (int d = 0; d < 3; ++d) {
                    S[d] = I[d];
                    if (I[d] < 0 || I[d] > g.n[d]) {
                        if (!g.per[d]) ok = false;
                        else S[d] = I[d] < 0 ? I[d] + g.n[d] : I[d] - g.n[d];
                    }
                }
                if (!ok) continue;
                h[((size_t)(k + 1) * G1 + (j + 1)) * G0 + (i + 1)] =
                    phi[((size_t)S[2] * (g.n[1] + 1) + S[1]) * (g.n[0] + 1) + S[0]];
            }
    double *d = nullptr;
    HIPCHK(hipMalloc(&d, h.size() * sizeof(double)));
    HIPCHK(hipMemcpy(d, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
    c->ls_bufs.push_back(d);
    const int b = c->ls.nb++;
    c->ls.phi[b] = d;
    c->ls.bc[b] = bc;
    c->ls.mu[b] = mu;
    for (int q = 0; q < 3; ++q) c->ls.w[b][q] = w ? w[q] : 0.0;
    return MPM_OK;
}

int mpm_remove_particles_inside_levelset(mpm_ctx *c, int64_t *removed) {
    if (!c) return set_err(MPM_E_ARG, "null context");
    int *d = nullptr;
    HIPCHK(hipMalloc(&d, sizeof(int)));
    int h = 0;
    hipError_t e = hipMemcpy(d, &h, sizeof(int), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        launch_remove_inside_ls(c->g, c->ls, c->P, d, c->stream);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess) e = hipMemcpy(&h, d, sizeof(int), hipMemcpyDeviceToHost);
    (void)hipFree(d);
    if (e != hipSuccess) return hip_err(e, "remove particles inside level set");
    // Redistribute() drops them (mpm_init.cpp:1595): every kernel skips a
    // removed particle and the next sorted step leaves it out.
    c->dt_next_valid = false;
    if (removed) *removed = h;
    return MPM_OK;
}

int mpm_set_wall_udf(mpm_ctx *c, int dir, int hi, mpm_wall_udf_fn fn) {
    if (!c || dir < 0 || dir > 2 || hi < 0 || hi > 1) return set_err(MPM_E_ARG, "bad face");
    c->udf_fn[dir][hi] = fn;
    if (fn && !c->udf_dev[dir][hi]) {
        const int p0 = (dir == 0) ? 1 : 0;
        const int p1 = (dir == 2) ? 1 : 2;
        size_t n = (size_t)(c->g.n[p0] + 1) * (c->g.n[p1] + 1) * 3;
        HIPCHK(hipMalloc(&c->udf_dev[dir][hi], n * sizeof(double)));
        HIPCHK(hipMemset(c->udf_dev[dir][hi], 0, n * sizeof(double)));
    }
    c->udf.tab[dir][hi] = fn ? c->udf_dev[dir][hi] : nullptr;
    c->any_udf = false;
    for (int d = 0; d < 3; ++d)
        for (int h = 0; h < 2; ++h) c->any_udf |= (c->udf_fn[d][h] != nullptr);
    return MPM_OK;
}

int mpm_energies(mpm_ctx *c, double out[6]) {
    if (!c || !out) return set_err(MPM_E_ARG, "null argument");
    launch_zero(c->d_energy, 6, c->stream);
    launch_energies(c->P, c->d_energy, c->stream);
    int rc = check_launch();
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(out, c->d_energy, 6 * sizeof(double), hipMemcpyDeviceToHost,
                          c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return MPM_OK;
}

int mpm_set_option(mpm_ctx *c, int option, int value) {
    if (!c) return set_err(MPM_E_ARG, "null context");
    if (option == MPM_OPT_PROFILE) c->profile = value;
    else if (option == MPM_OPT_KEEP_FORCES) c->keep_forces = value;
    else if (option == MPM_OPT_SORTED) {
        c->sorted = value;
        c->all_dirty = true;
        c->dt_next_valid = false;
    } else if (option == MPM_OPT_FAST) {
        c->fast = value;
        c->dt_next_valid = false;
    } else if (option == MPM_OPT_FUSED_DT) {
        c->fuse_dt = value;
        c->dt_next_valid = false;
    } else if (option == MPM_OPT_TILE_WAVES) {
        if (value != 0 && value != 4 && value != 8)
            return set_err(MPM_E_ARG, "tile wavefronts must be 0, 4 or 8");
        c->S.waves = value;
    } else if (option == MPM_OPT_COARSE_BLEND) {
        c->coarse_blend = value ? 1 : 0;
    } else if (option == MPM_OPT_DETERMINISTIC) {
        if (value && !c->det_blk) {
            // one node block per tile (7 components), and the tile stamps
            // (-1: no pass)
            const size_t nb = (size_t)c->S.ntiles * 7 * xmpm::kTNNodes;
            if (hipMalloc(&c->det_blk, nb * sizeof(double)) != hipSuccess ||
                hipMalloc(&c->det_slot, (size_t)c->S.ntiles * sizeof(int64_t)) != hipSuccess) {
                (void)hipFree(c->det_blk);
                c->det_blk = nullptr;
                c->det_slot = nullptr;
                return set_err(MPM_E_OOM, "deterministic P2G buffers");
            }
            HIPCHK(hipMemsetAsync(c->det_slot, 0xff, (size_t)c->S.ntiles * sizeof(int64_t),
                                  c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
        }
        c->det = value ? 1 : 0;
    }
    else return set_err(MPM_E_ARG, "unknown option");
    return MPM_OK;
}

int mpm_last_step_timing(mpm_ctx *c, double *total_ms, double kernel_ms[4]) {
    if (!c) return set_err(MPM_E_ARG, "null context");
    if (total_ms) *total_ms = c->step_ms;
    if (kernel_ms)
        for (int q = 0; q < 4; ++q) kernel_ms[q] = c->kernel_ms[q];
    return MPM_OK;
}

}  // extern "C"

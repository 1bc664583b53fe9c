// mpm_stages.h — one-item bodies of the generic step kernels (one particle's
// deposit, one node's nodal pipeline, one particle's pipeline), shared by
// the per-array kernels of mpm_kernels.hip and the batched ensemble's
// persistent kernel (mpm_batch.hip).
#pragma once

#include "mpm_device.h"

namespace xmpm {

// deposit_onto_grid_momentum scatter (mpm_particle_grid_ops.cpp:416-587).
// One particle's deposit.
template <bool FORCES>
__device__ __forceinline__ void p2g_one(const GridP &g, const PhysP &ph, const PState &P,
                                        int64_t i, double *acc, int um, int uv,
                                        const double *vimg) {
    if (P.id[i] < 0) return;
    if (P.idata[MPM_I_PHASE * P.cap + i] != 0) return;
    const int64_t cap = P.cap;
    double x[3];
    int iv[3], mn[3], mx[3];
    for (int d = 0; d < 3; ++d) {
        x[d] = P.pos[d * cap + i];
        iv[d] = (int)floor((x[d] - g.plo[d]) * g.dxi[d]);
        bounds(iv[d], 0, g.hi[d], g.od[d], g.per[d], mn[d], mx[d]);
    }
    double bv[3][4], bd[3][4];
    for (int d = 0; d < 3; ++d)
        for (int l = mn[d]; l < mx[d]; ++l)
            basis1d(g.od[d], g.per[d], iv[d], l, 0, g.hi[d], x[d], g.plo[d], g.dx[d],
                    bv[d][l - mn[d]], bd[d][l - mn[d]]);
    const double m = P.r[MPM_R_MASS * cap + i];
    const double vel[3] = {P.r[(MPM_R_VEL + 0) * cap + i], P.r[(MPM_R_VEL + 1) * cap + i],
                           P.r[(MPM_R_VEL + 2) * cap + i]};
    double vimgv[3] = {vel[0], vel[1], vel[2]};
    if (vimg)
        for (int d = 0; d < 3; ++d) vimgv[d] = vimg[d * cap + i];
    double T[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    double V = 0.0;
    bool slab = false;
    if (FORCES) {
        int ind = 0;
        for (int a = 0; a < 3; a++)
            for (int b = a; b < 3; b++) {
                double s = P.r[(MPM_R_STRESS + ind) * cap + i];
                T[a * 3 + b] = s;
                T[b * 3 + a] = s;
                ind++;
            }
        V = P.r[MPM_R_VOLUME * cap + i];
        slab = ph.extloads && x[0] > ph.slab_lo[0] && x[0] < ph.slab_hi[0] &&
               x[1] > ph.slab_lo[1] && x[1] < ph.slab_hi[1] && x[2] > ph.slab_lo[2] &&
               x[2] < ph.slab_hi[2];
    }
    const int64_t nn = g.nn;
    for (int n = mn[2]; n < mx[2]; ++n) {
        int tz[2];
        bool iz[2];
        const int cz = targets(iv[2] + n, g.n[2], g.per[2], tz, iz);
        if (!cz) continue;
        for (int mm = mn[1]; mm < mx[1]; ++mm) {
            int ty[2];
            bool iy[2];
            const int cy = targets(iv[1] + mm, g.n[1], g.per[1], ty, iy);
            if (!cy) continue;
            for (int l = mn[0]; l < mx[0]; ++l) {
                int tx[2];
                bool ix[2];
                const int cx = targets(iv[0] + l, g.n[0], g.per[0], tx, ix);
                if (!cx) continue;
                const double vx = bv[0][l - mn[0]], vy = bv[1][mm - mn[1]], vz = bv[2][n - mn[2]];
                const double N = vx * vy * vz;
                double dm = 0.0, dp[3] = {0, 0, 0}, dpi[3] = {0, 0, 0}, df[3] = {0, 0, 0};
                if (um) dm = m * N;
                if (uv) {
                    for (int d = 0; d < 3; ++d) {
                        dp[d] = m * vel[d] * N;
                        dpi[d] = m * vimgv[d] * N;
                    }
                }
                if (FORCES) {
                    const double dx_ = bd[0][l - mn[0]], dy_ = bd[1][mm - mn[1]],
                                 dz_ = bd[2][n - mn[2]];
                    const double gr[3] = {dx_ * vy * vz, vx * dy_ * vz, vx * vy * dz_};
                    double bf[3] = {m * ph.grav[0] * N, m * ph.grav[1] * N, m * ph.grav[2] * N};
                    if (slab)
                        for (int d = 0; d < 3; ++d) bf[d] += ph.extforce[d] * N;
                    for (int a = 0; a < 3; ++a) {
                        double tv = 0.0;
                        for (int b = 0; b < 3; ++b) tv += T[a * 3 + b] * gr[b];
                        df[a] = bf[a] + (-V * tv);
                    }
                }
                for (int qz = 0; qz < cz; ++qz)
                    for (int qy = 0; qy < cy; ++qy)
                        for (int qx = 0; qx < cx; ++qx) {
                            const int64_t node = nidx(g, tx[qx], ty[qy], tz[qz]);
                            const bool image = ix[qx] || iy[qy] || iz[qz];
                            if (um) unsafeAtomicAdd(&acc[node], dm);
                            if (uv) {
                                const double *src = image ? dpi : dp;
                                unsafeAtomicAdd(&acc[nn + node], src[0]);
                                unsafeAtomicAdd(&acc[2 * nn + node], src[1]);
                                unsafeAtomicAdd(&acc[3 * nn + node], src[2]);
                            }
                            if (FORCES) {
                                unsafeAtomicAdd(&acc[4 * nn + node], df[0]);
                                unsafeAtomicAdd(&acc[5 * nn + node], df[1]);
                                unsafeAtomicAdd(&acc[6 * nn + node], df[2]);
                            }
                        }
            }
        }
    }
}

// Nodal pipeline: P2G normalisation (mpm_particle_grid_ops.cpp:590-627),
// backup_current_velocity (nodal_data_ops.cpp:57-81),
// Nodal_Time_Update_Momentum (:451-482), nodal_bcs (:663-766),
// apply_udf_nodal_bcs (:1061-1175), nodal_levelset_bcs (:129-220),
// store_delta_velocity (:368-392).
// One node of the pipeline: grid_one for a node index, grid_at for a node
// whose (I, J, K) the caller knows.
__device__ __forceinline__ void grid_at(const GridP &g, const PhysP &ph, const LsP &ls,
                                        const UdfP &udf, const NState &N, int flags, double dt,
                                        int64_t node, int I, int J, int K, const int *tile_epoch,
                                        int epoch);

__device__ __forceinline__ void grid_one(const GridP &g, const PhysP &ph, const LsP &ls,
                                         const UdfP &udf, const NState &N, int flags, double dt,
                                         int64_t node, const int *tile_epoch, int epoch) {
    const int I = (int)(node % (g.n[0] + 1));
    const int J = (int)((node / (g.n[0] + 1)) % (g.n[1] + 1));
    const int K = (int)(node / ((int64_t)(g.n[0] + 1) * (g.n[1] + 1)));
    grid_at(g, ph, ls, udf, N, flags, dt, node, I, J, K, tile_epoch, epoch);
}

__device__ __forceinline__ void grid_at(const GridP &g, const PhysP &ph, const LsP &ls,
                                        const UdfP &udf, const NState &N, int flags, double dt,
                                        int64_t node, int I, int J, int K, const int *tile_epoch,
                                        int epoch) {
    const int64_t nn = g.nn;
    if (tile_epoch) {
        // Node tiles no particle deposited on this step hold the massless
        // state, which is all-zero when walls are at rest and no UDF is set:
        // written once when a tile goes idle, skipped afterwards.
        const int t = tile_epoch[node_tile(g, I, J, K)];
        if (t != epoch) {
            if (t == epoch - 1 && (flags & G_FINISH_MASS)) {
                N.mass[node] = 0.0;
                N.mold[node] = 0.0;
                for (int d = 0; d < 3; ++d) {
                    N.vel[d * nn + node] = 0.0;
                    N.dvel[d * nn + node] = 0.0;
                    if (flags & G_KEEP_FRC) N.frc[d * nn + node] = 0.0;
                }
            }
            return;
        }
    }
    double m = N.mass[node];
    double v[3] = {0.0, 0.0, 0.0}, dv[3] = {0.0, 0.0, 0.0};
    const bool full = (flags & G_FINISH_MASS) && (flags & G_BACKUP);
    // v is rebuilt from the momentum sums under G_FINISH_VEL and dv is reset
    // in a full pass: only the other cases read them.
    for (int d = 0; d < 3; ++d) {
        if (!(flags & G_FINISH_VEL)) v[d] = N.vel[d * nn + node];
        if (!full) dv[d] = N.dvel[d * nn + node];
    }
    double f[3] = {0, 0, 0};
    double mold = 0.0;
    if (flags & G_FINISH_MASS) {
        m = N.acc[node];
        N.acc[node] = 0.0;
        N.mass[node] = m;
        if (full) {
            // Reset_Nodaldata_to_Zero preceded this P2G: Δv and mass_old of
            // massless nodes stay zero.
            dv[0] = dv[1] = dv[2] = 0.0;
        }
    }
    if (flags & G_FINISH_VEL) {
        for (int d = 0; d < 3; ++d) {
            double p = N.acc[(1 + d) * nn + node];
            N.acc[(1 + d) * nn + node] = 0.0;
            if (m > 0.0) v[d] = (m >= ph.mass_tol) ? p / m : 0.0;
            else v[d] = p;
        }
    }
    if (flags & G_FINISH_FRC) {
        for (int d = 0; d < 3; ++d) {
            f[d] = N.acc[(4 + d) * nn + node];
            N.acc[(4 + d) * nn + node] = 0.0;
        }
        if (flags & G_KEEP_FRC)
            for (int d = 0; d < 3; ++d) N.frc[d * nn + node] = f[d];
    } else if (flags & G_UPDATE) {
        for (int d = 0; d < 3; ++d) f[d] = N.frc[d * nn + node];
    }
    if (flags & G_BACKUP) {
        if (m > 0.0) {
            mold = m;
            for (int d = 0; d < 3; ++d) dv[d] = v[d];
            N.mold[node] = mold;
        } else if (full) {
            N.mold[node] = 0.0;
        }
    }
    if (flags & G_UPDATE) {
        if (m >= ph.mass_tol) {
            for (int d = 0; d < 3; ++d) v[d] += f[d] / m * dt;
        } else {
            v[0] = v[1] = v[2] = 0.0;
        }
    }
    if (flags & G_BCS) {
        const int id3[3] = {I, J, K};
        double rin[3] = {v[0], v[1], v[2]}, rout[3] = {v[0], v[1], v[2]};
        double wv[3] = {0, 0, 0};
        for (int dir = 0; dir < 3; ++dir) {
            int side = -1;
            if (id3[dir] == 0) side = 0;
            else if (id3[dir] == g.hi[dir] + 1) side = 1;
            if (side < 0) continue;
            const double *wvel = side ? ph.wv_hi : ph.wv_lo;
            for (int d = 0; d < 3; ++d) {
                wv[d] = wvel[dir * 3 + d];
                rin[d] -= wv[d];
            }
            double nrm[3] = {0, 0, 0};
            nrm[dir] = side ? -1.0 : 1.0;
            applybc(rin, rout, side ? ph.mu_hi[dir] : ph.mu_lo[dir], nrm,
                    side ? ph.bchi[dir] : ph.bclo[dir]);
            for (int q = 0; q < 3; ++q) v[q] = rout[q] + wv[q];
        }
        for (int dir = 0; dir < 3; ++dir) {
            const double *tab = nullptr;
            int bc = -1;
            if (id3[dir] == 0 && udf.tab[dir][0]) {
                tab = udf.tab[dir][0];
                bc = ph.bclo[dir];
            } else if (id3[dir] == g.hi[dir] + 1 && udf.tab[dir][1]) {
                tab = udf.tab[dir][1];
                bc = ph.bchi[dir];
            }
            if (!tab) continue;
            const int p0 = (dir == 0) ? 1 : 0;
            const int p1 = (dir == 2) ? 1 : 2;
            const int64_t fl = ((int64_t)id3[p0] * (g.n[p1] + 1) + id3[p1]) * 3;
            if (bc == MPM_BC_NOSLIP) {
                for (int q = 0; q < 3; ++q) v[q] = tab[fl + q];
            } else {
                v[dir] = tab[fl + dir];
            }
        }
        if (ls.nb > 0) {
            const double xp[3] = {g.plo[0] + I * g.dx[0], g.plo[1] + J * g.dx[1],
                                  g.plo[2] + K * g.dx[2]};
            for (int b = 0; b < ls.nb; ++b) {
                const double lv = ls_value(g, ls.phi[b], xp);
                if (lv >= 0.0 || m <= 0.0) continue;
                double nrm[3] = {1.0, 0.0, 0.0};
                ls_grad(g, ls.phi[b], xp, nrm);
                double gm = 0.0;
                for (int d = 0; d < 3; d++) gm += nrm[d] * nrm[d];
                gm = sqrt(gm);
                if (!(gm >= 1.0e-10)) continue;
                for (int d = 0; d < 3; d++) nrm[d] /= gm;
                double ri[3], ro[3];
                for (int d = 0; d < 3; d++) {
                    ri[d] = v[d] - ls.w[b][d];
                    ro[d] = ri[d];
                }
                double vn = 0.0;
                for (int d = 0; d < 3; d++) vn += ri[d] * nrm[d];
                if (vn > 0.0) continue;
                applybc(ri, ro, ls.mu[b], nrm, ls.bc[b]);
                for (int d = 0; d < 3; d++) v[d] = ro[d] + ls.w[b][d];
            }
        }
        if (m > 0.0)
            for (int d = 0; d < 3; ++d) dv[d] = v[d] - dv[d];
    }
    for (int d = 0; d < 3; ++d) {
        N.vel[d * nn + node] = v[d];
        N.dvel[d * nn + node] = dv[d];
    }
}

// Per-particle pipeline, in place: interpolate_from_grid
// (mpm_particle_grid_ops.cpp:967-1096), moveParticles
// (mpm_particle_timestep.cpp:329-576), the Redistribute that follows it,
// updateVolume (:144-180) and apply_constitutive_model[_delta]
// (mpm_particle_container.cpp:43-272).
// One particle of the pipeline.
__device__ __forceinline__ void particle_one(const GridP &g, const PhysP &ph, const LsP &ls,
                                             const UdfP &udf, const PState &P, const NState &N,
                                             int flags, double dt, double time, int64_t i) {
    if (P.id[i] < 0) return;
    Rec R;
    load_all(P, i, R);
    if ((flags & (P_G2P_VEL | P_G2P_GRAD)) && R.phase == 0)
        stage_g2p(g, ph, N, R, flags & P_G2P_VEL, flags & P_G2P_GRAD, dt);
    if (flags & P_MOVE) {
        stage_move(g, ph, ls, udf, R, dt);
        if (flags & P_REDIST) stage_redistribute(g, R.x, R.id);
    }
    if ((flags & P_VOLUME) && R.phase == 0) stage_volume(R);
    if ((flags & (P_STRESS | P_STRESS_INIT)) && R.phase == 0)
        stage_stress(ph, R, dt, time, flags & P_STRESS_INIT);
    store_all(P, i, R);
}

}  // namespace xmpm

// mpm_kernels.hip — HIP kernels of the explicit MPM step for gfx950.
//
// P2G: one thread per particle, 1-D spline factors hoisted per particle,
// FP64 hardware atomics into the nodal accumulators.  Periodic directions
// are handled by index arithmetic that reproduces the reference's
// ghost-particle images (a deposit at unwrapped node I lands on I when
// 0 <= I <= n and on I +/- n when that lies in [0, n]).
// Grid: one thread per node, fused normalise / backup / update / BCs / delta.
// Particle: one thread per particle, fused G2P / F / D / move /
// redistribute / volume / stress.

#include "mpm_device.h"

namespace xmpm {

namespace {

constexpr int kBlock = 256;

// ---------------------------------------------------------------------------
__global__ void k_zero(double *p, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = 0.0;
}

// deposit_onto_grid_momentum scatter (mpm_particle_grid_ops.cpp:416-587).
template <bool FORCES>
__global__ __launch_bounds__(kBlock) void k_p2g(GridP g, PhysP ph, PState P, double *acc,
                                                int um, int uv, const double *vimg) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.np) return;
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
__global__ __launch_bounds__(kBlock) void k_grid(GridP g, PhysP ph, LsP ls, UdfP udf,
                                                 NState N, int flags, double dt_val,
                                                 const Scalars *sc, const int *tile_epoch,
                                                 int epoch) {
    const int64_t node = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (node >= g.nn) return;
    const int64_t nn = g.nn;
    const double dt = sc ? sc->dt : dt_val;
    const int I = (int)(node % (g.n[0] + 1));
    const int J = (int)((node / (g.n[0] + 1)) % (g.n[1] + 1));
    const int K = (int)(node / ((int64_t)(g.n[0] + 1) * (g.n[1] + 1)));
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
    double v[3], dv[3];
    for (int d = 0; d < 3; ++d) {
        v[d] = N.vel[d * nn + node];
        dv[d] = N.dvel[d * nn + node];
    }
    double f[3] = {0, 0, 0};
    double mold = 0.0;
    const bool full = (flags & G_FINISH_MASS) && (flags & G_BACKUP);
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
__global__ __launch_bounds__(kBlock) void k_particle(GridP g, PhysP ph, LsP ls, UdfP udf,
                                                     PState P, NState N, int flags,
                                                     double dt_val, double time_val,
                                                     const Scalars *sc) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.np) return;
    if (P.id[i] < 0) return;
    const double dt = sc ? sc->dt : dt_val;
    const double time = sc ? sc->time : time_val;
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

// Calculate_time_step (mpm_particle_timestep.cpp:38-111): per-particle
// dx_min / (c + |v|), min-reduced through the ordered bit pattern of
// positive doubles.
__global__ __launch_bounds__(kBlock) void k_dt(GridP g, PState P, Scalars *sc) {
    __shared__ double red[kBlock];
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double v = __builtin_huge_val();
    if (i < P.np && P.id[i] >= 0 && P.idata[MPM_I_PHASE * P.cap + i] == 0) {
        const int64_t cap = P.cap;
        const double *R = P.r;
        const int m = P.idata[MPM_I_MODEL * cap + i];
        double Cs;
        if (m == 1) {
            Cs = sqrt(R[MPM_R_BULK * cap + i] / R[MPM_R_DENSITY * cap + i]);
        } else {
            const double E = R[MPM_R_E * cap + i], nu = R[MPM_R_NU * cap + i];
            const double lam = E * nu / ((1 + nu) * (1 - 2.0 * nu));
            const double mu = E / (2.0 * (1 + nu));
            Cs = sqrt((lam + 2.0 * mu) / R[MPM_R_DENSITY * cap + i]);
        }
        double vm = 0.0;
        for (int d = 0; d < 3; ++d) vm += R[(MPM_R_VEL + d) * cap + i] * R[(MPM_R_VEL + d) * cap + i];
        vm = sqrt(vm);
        double dxmin = g.dx[0];
        for (int d = 1; d < 3; ++d) dxmin = fmin(dxmin, g.dx[d]);
        v = dxmin / (Cs + vm);
    }
    red[threadIdx.x] = v;
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] = fmin(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double b = red[0];
        if (b < __builtin_huge_val())
            atomicMin(&sc->dtmin_bits, (unsigned long long)__double_as_longlong(b));
    }
}

__global__ void k_dt_init(Scalars *sc) {
    sc->dtmin_bits = (unsigned long long)__double_as_longlong(__builtin_huge_val());
}

__global__ void k_dt_final(Scalars *sc, PhysP ph, int fixed, double fixed_dt, int fused,
                           int advance) {
    // The step about to run accumulates the following step's minimum, so the
    // running value is reset here.  Read and reset are one atomic exchange: a
    // plain load of a uniform address becomes a scalar-cache load, which the
    // hardware does not order against the vector store of the reset.
    const unsigned long long nxt = atomicExch(
        &sc->dtmin_next_bits, (unsigned long long)__double_as_longlong(__builtin_huge_val()));
    const unsigned long long bits = fused ? nxt : sc->dtmin_bits;
    double dt;
    if (fixed) {
        dt = fixed_dt;
    } else {
        dt = __longlong_as_double((long long)bits);
        if (dt == __builtin_huge_val()) dt = 1.79769313486231570815e+308;  // numeric_limits max
        dt = ph.cfl * dt;
        dt = fmax(fmin(dt, ph.dt_max), ph.dt_min);
    }
    // Same exchange for dt: the previous step's dt advances the time
    // (the end of that step, folded into this launch).
    const unsigned long long old = atomicExch(reinterpret_cast<unsigned long long *>(&sc->dt),
                                              (unsigned long long)__double_as_longlong(dt));
    if (advance) sc->time = sc->time + __longlong_as_double((long long)old);
}

__global__ void k_advance_time(Scalars *sc) { sc->time = sc->time + sc->dt; }

// Calculate_Total_Energies / MWA sums (mpm_diagnostics.cpp:23-127).
__global__ __launch_bounds__(kBlock) void k_energies(PState P, double *out) {
    __shared__ double red[6][kBlock];
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double e[6] = {0, 0, 0, 0, 0, 0};
    if (i < P.np && P.id[i] >= 0) {
        const int64_t cap = P.cap;
        const double *R = P.r;
        const double m = R[MPM_R_MASS * cap + i];
        double v2 = 0.0;
        for (int d = 0; d < 3; ++d) v2 += R[(MPM_R_VEL + d) * cap + i] * R[(MPM_R_VEL + d) * cap + i];
        e[0] = 0.5 * m * v2;
        const double *S = R + MPM_R_STRESS * cap + i;
        const double *ep = R + MPM_R_STRAIN * cap + i;
        double se = S[XX * cap] * ep[XX * cap] + S[YY * cap] * ep[YY * cap] + S[ZZ * cap] * ep[ZZ * cap] +
                    2.0 * (S[XY * cap] * ep[XY * cap] + S[YZ * cap] * ep[YZ * cap] +
                           S[XZ * cap] * ep[XZ * cap]);
        e[1] = 0.5 * R[MPM_R_VOLUME * cap + i] * se;
        for (int d = 0; d < 3; ++d) e[2 + d] = m * R[(MPM_R_VEL + d) * cap + i];
        e[5] = m;
    }
    for (int q = 0; q < 6; ++q) red[q][threadIdx.x] = e[q];
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int q = 0; q < 6; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int q = 0; q < 6; ++q) atomicAdd(&out[q], red[q][0]);
}

__global__ void k_snapshot(PState P) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.np) return;
    for (int d = 0; d < 3; ++d) P.vsnap[d * P.cap + i] = P.r[(MPM_R_VEL + d) * P.cap + i];
}

__global__ void k_redistribute(GridP g, PState P) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.np || P.id[i] < 0) return;
    double x[3] = {P.pos[i], P.pos[P.cap + i], P.pos[2 * P.cap + i]};
    int64_t id = P.id[i];
    stage_redistribute(g, x, id);
    for (int d = 0; d < 3; ++d) P.pos[d * P.cap + i] = x[d];
    P.id[i] = id;
}

inline unsigned blocks_for(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

}  // namespace

void launch_zero(double *p, int64_t n, hipStream_t s) {
    if (n <= 0) return;
    unsigned b = blocks_for(n);
    if (b > 65536) b = 65536;
    hipLaunchKernelGGL(k_zero, dim3(b), dim3(kBlock), 0, s, p, n);
}

void launch_p2g(const GridP &g, const PhysP &ph, const PState &P, const NState &N, int um,
                int uv, int uf, const double *vimg, hipStream_t s) {
    if (P.np <= 0) return;
    if (uf)
        hipLaunchKernelGGL(k_p2g<true>, dim3(blocks_for(P.np)), dim3(kBlock), 0, s, g, ph, P,
                           N.acc, um, uv, vimg);
    else
        hipLaunchKernelGGL(k_p2g<false>, dim3(blocks_for(P.np)), dim3(kBlock), 0, s, g, ph, P,
                           N.acc, um, uv, vimg);
}

void launch_grid(const GridP &g, const PhysP &ph, const LsP &ls, const UdfP &udf,
                 const NState &N, int flags, double dt, const Scalars *sc, hipStream_t s,
                 const int *tile_epoch, int epoch) {
    hipLaunchKernelGGL(k_grid, dim3(blocks_for(g.nn)), dim3(kBlock), 0, s, g, ph, ls, udf, N,
                       flags, dt, sc, tile_epoch, epoch);
}

void launch_particle(const GridP &g, const PhysP &ph, const LsP &ls, const UdfP &udf,
                     const PState &P, const NState &N, int flags, double dt, double time,
                     const Scalars *sc, hipStream_t s) {
    if (P.np <= 0) return;
    hipLaunchKernelGGL(k_particle, dim3(blocks_for(P.np)), dim3(kBlock), 0, s, g, ph, ls, udf,
                       P, N, flags, dt, time, sc);
}

void launch_dt(const GridP &g, const PhysP &ph, const PState &P, Scalars *sc, int fixed,
               double fixed_dt, hipStream_t s, int fused, int advance) {
    if (!fixed && !fused) {
        hipLaunchKernelGGL(k_dt_init, dim3(1), dim3(1), 0, s, sc);
        if (P.np > 0)
            hipLaunchKernelGGL(k_dt, dim3(blocks_for(P.np)), dim3(kBlock), 0, s, g, P, sc);
    }
    hipLaunchKernelGGL(k_dt_final, dim3(1), dim3(1), 0, s, sc, ph, fixed, fixed_dt,
                       (fixed || !fused) ? 0 : 1, advance);
}

void launch_advance_time(Scalars *sc, hipStream_t s) {
    hipLaunchKernelGGL(k_advance_time, dim3(1), dim3(1), 0, s, sc);
}

void launch_energies(const PState &P, double *out6, hipStream_t s) {
    if (P.np <= 0) return;
    hipLaunchKernelGGL(k_energies, dim3(blocks_for(P.np)), dim3(kBlock), 0, s, P, out6);
}

void launch_snapshot_vel(const PState &P, hipStream_t s) {
    if (P.np <= 0) return;
    hipLaunchKernelGGL(k_snapshot, dim3(blocks_for(P.np)), dim3(kBlock), 0, s, P);
}

void launch_redistribute(const GridP &g, const PState &P, hipStream_t s) {
    if (P.np <= 0) return;
    hipLaunchKernelGGL(k_redistribute, dim3(blocks_for(P.np)), dim3(kBlock), 0, s, g, P);
}

}  // namespace xmpm

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

#include "mpm_internal.h"
#include "mpm_math.h"

namespace xmpm {

namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ int64_t nidx(const GridP &g, int I, int J, int K) {
    return ((int64_t)K * (g.n[1] + 1) + J) * (g.n[0] + 1) + I;
}

// Deposit targets of an unwrapped node index along one direction.
__device__ __forceinline__ int targets(int I, int n, int per, int t[2], bool img[2]) {
    int c = 0;
    if (I >= 0 && I <= n) {
        t[c] = I;
        img[c] = false;
        ++c;
    }
    if (per) {
        if (I <= 0 && I + n >= 0) {
            t[c] = I + n;
            img[c] = true;
            ++c;
        } else if (I >= n && I - n <= n) {
            t[c] = I - n;
            img[c] = true;
            ++c;
        }
    }
    return c;
}

// Node index read by G2P for an unwrapped index: ghost nodes of periodic
// directions hold the value one period away (FillBoundary); ghost nodes of
// non-periodic directions hold zero.
__device__ __forceinline__ int wrap_read(int I, int n, int per, bool &ok) {
    if (I >= 0 && I <= n) return I;
    if (!per) {
        ok = false;
        return 0;
    }
    const int t = I < 0 ? I + n : I - n;
    if (t < 0 || t > n) {
        ok = false;  // more than one period away: not a ghost node
        return 0;
    }
    return t;
}

__device__ __forceinline__ double nodal_read(const GridP &g, const double *arr, int I, int J,
                                             int K) {
    bool ok = true;
    int a = wrap_read(I, g.n[0], g.per[0], ok);
    int b = wrap_read(J, g.n[1], g.per[1], ok);
    int c = wrap_read(K, g.n[2], g.per[2], ok);
    return ok ? arr[nidx(g, a, b, c)] : 0.0;
}

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

// Level-set value / gradient (mpm_eb.H:176-288), lsref = 1, one ghost layer.
__device__ double ls_value(const GridP &g, const double *phi, const double x[3]) {
    int idx[3];
    for (int d = 0; d < 3; ++d) idx[d] = (int)floor((x[d] - g.plo[d] + kTiny) / g.dx[d]);
    double r[3];
    for (int d = 0; d < 3; ++d) r[d] = (x[d] - (g.plo[d] + idx[d] * g.dx[d])) / g.dx[d];
    const int G0 = g.n[0] + 3, G1 = g.n[1] + 3;
    double v = 0.0;
    for (int corner = 0; corner < 8; ++corner) {
        int off[3] = {corner & 1, (corner >> 1) & 1, (corner >> 2) & 1};
        int ii[3];
        double w = 1.0;
        for (int d = 0; d < 3; ++d) {
            ii[d] = idx[d] + off[d];
            ii[d] = ii[d] < -1 ? -1 : (ii[d] > g.n[d] + 1 ? g.n[d] + 1 : ii[d]);
            w *= (off[d] == 0) ? (1.0 - r[d]) : r[d];
        }
        v += phi[((int64_t)(ii[2] + 1) * G1 + (ii[1] + 1)) * G0 + (ii[0] + 1)] * w;
    }
    return v;
}

__device__ void ls_grad(const GridP &g, const double *phi, const double x[3], double gr[3]) {
    int idx[3];
    for (int d = 0; d < 3; ++d) idx[d] = (int)floor((x[d] - g.plo[d] + kTiny) / g.dx[d]);
    double r[3], dinv[3];
    for (int d = 0; d < 3; ++d) {
        r[d] = (x[d] - (g.plo[d] + idx[d] * g.dx[d])) / g.dx[d];
        dinv[d] = 1.0 / g.dx[d];
    }
    const int G0 = g.n[0] + 3, G1 = g.n[1] + 3;
    gr[0] = gr[1] = gr[2] = 0.0;
    for (int corner = 0; corner < 8; ++corner) {
        int off[3] = {corner & 1, (corner >> 1) & 1, (corner >> 2) & 1};
        int ii[3];
        for (int d = 0; d < 3; ++d) {
            ii[d] = idx[d] + off[d];
            ii[d] = ii[d] < -1 ? -1 : (ii[d] > g.n[d] + 1 ? g.n[d] + 1 : ii[d]);
        }
        const double ph = phi[((int64_t)(ii[2] + 1) * G1 + (ii[1] + 1)) * G0 + (ii[0] + 1)];
        for (int dir = 0; dir < 3; ++dir) {
            double prod = 1.0;
            for (int d = 0; d < 3; ++d) {
                double val = (off[d] == 0) ? (1.0 - r[d]) : r[d];
                double der = (off[d] == 0) ? -dinv[d] : dinv[d];
                prod *= (d == dir) ? der : val;
            }
            gr[dir] += ph * prod;
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
                                                 const Scalars *sc) {
    const int64_t node = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (node >= g.nn) return;
    const int64_t nn = g.nn;
    const double dt = sc ? sc->dt : dt_val;
    const int I = (int)(node % (g.n[0] + 1));
    const int J = (int)((node / (g.n[0] + 1)) % (g.n[1] + 1));
    const int K = (int)(node / ((int64_t)(g.n[0] + 1) * (g.n[1] + 1)));
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

// Per-particle pipeline: interpolate_from_grid (mpm_particle_grid_ops.cpp:
// 967-1096), moveParticles (mpm_particle_timestep.cpp:329-576), the
// Redistribute that follows it, updateVolume (:144-180) and
// apply_constitutive_model[_delta] (mpm_particle_container.cpp:43-272).
__global__ __launch_bounds__(kBlock) void k_particle(GridP g, PhysP ph, LsP ls, UdfP udf,
                                                     PState P, NState N, int flags,
                                                     double dt_val, double time_val,
                                                     const Scalars *sc) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.np) return;
    if (P.id[i] < 0) return;
    const int64_t cap = P.cap;
    const int64_t nn = g.nn;
    const double dt = sc ? sc->dt : dt_val;
    const double time = sc ? sc->time : time_val;
    double *R = P.r;
    const int phase = P.idata[MPM_I_PHASE * cap + i];
    double x[3] = {P.pos[i], P.pos[cap + i], P.pos[2 * cap + i]};

    if ((flags & (P_G2P_VEL | P_G2P_GRAD)) && phase == 0) {
        int iv[3], mn[3], mx[3];
        for (int d = 0; d < 3; ++d) {
            iv[d] = (int)floor((x[d] - g.plo[d]) * g.dxi[d]);
            bounds(iv[d], 0, g.hi[d], g.od[d], g.per[d], mn[d], mx[d]);
        }
        double L[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
        const bool grad = flags & P_G2P_GRAD;
        double bv[3][4], bd[3][4];
        if (grad)
            for (int d = 0; d < 3; ++d)
                for (int l = mn[d]; l < mx[d]; ++l)
                    basis1d(g.od[d], g.per[d], iv[d], l, 0, g.hi[d], x[d], g.plo[d], g.dx[d],
                            bv[d][l - mn[d]], bd[d][l - mn[d]]);
        bool grad_done = false;
        if (flags & P_G2P_VEL) {
            double vals0[3] = {0, 0, 0}, vals1[3] = {0, 0, 0};
            const double alpha = ph.alpha;
            if (g.uniform && g.od[0] >= 2) {
                const int O = g.od[0];
                double wv[3][4];
                for (int d = 0; d < 3; ++d)
                    for (int l = mn[d]; l < mx[d]; ++l)
                        wv[d][l - mn[d]] =
                            interp1d(O, g.per[d], iv[d] + l, 0, g.hi[d], x[d], g.plo[d], g.dx[d]);
                for (int n = mn[2]; n < mx[2]; ++n)
                    for (int mm = mn[1]; mm < mx[1]; ++mm)
                        for (int l = mn[0]; l < mx[0]; ++l) {
                            const int I = iv[0] + l, J = iv[1] + mm, K = iv[2] + n;
                            bool ok = true;
                            const int a = wrap_read(I, g.n[0], g.per[0], ok);
                            const int b = wrap_read(J, g.n[1], g.per[1], ok);
                            const int c = wrap_read(K, g.n[2], g.per[2], ok);
                            const int64_t q = nidx(g, a, b, c);
                            double nv[3], nd[3];
                            for (int d = 0; d < 3; ++d) {
                                nv[d] = ok ? N.vel[d * nn + q] : 0.0;
                                nd[d] = ok ? N.dvel[d * nn + q] : 0.0;
                            }
                            const double w = wv[0][l - mn[0]] * wv[1][mm - mn[1]] * wv[2][n - mn[2]];
                            for (int d = 0; d < 3; ++d) {
                                vals0[d] += w * nv[d];
                                vals1[d] += w * nd[d];
                            }
                            if (grad) {
                                const double vx = bv[0][l - mn[0]], vy = bv[1][mm - mn[1]],
                                             vz = bv[2][n - mn[2]];
                                const double gr[3] = {bd[0][l - mn[0]] * vy * vz,
                                                      vx * bd[1][mm - mn[1]] * vz,
                                                      vx * vy * bd[2][n - mn[2]]};
                                for (int d1 = 0; d1 < 3; ++d1)
                                    for (int d2 = 0; d2 < 3; ++d2) L[d1][d2] += nv[d1] * gr[d2];
                            }
                        }
                grad_done = grad;
            } else {
                for (int dim = 0; dim < 3; ++dim) {
                    const double *A0 = N.vel + dim * nn;
                    const double *A1 = N.dvel + dim * nn;
                    const int O = g.od[dim];
                    if (O == 1) {
                        // bilin_interp_two (interpolants.H:583-656), trilinear in all dirs
                        const int ii = iv[0], jj = iv[1], kk = iv[2];
                        const double a = (x[0] - (g.plo[0] + ii * g.dx[0])) / g.dx[0];
                        const double b = (x[1] - (g.plo[1] + jj * g.dx[1])) / g.dx[1];
                        const double c = (x[2] - (g.plo[2] + kk * g.dx[2])) / g.dx[2];
                        const double w000 = (1.0 - a) * (1.0 - b) * (1.0 - c);
                        const double w100 = a * (1.0 - b) * (1.0 - c);
                        const double w010 = (1.0 - a) * b * (1.0 - c);
                        const double w110 = a * b * (1.0 - c);
                        const double w001 = (1.0 - a) * (1.0 - b) * c;
                        const double w101 = a * (1.0 - b) * c;
                        const double w011 = (1.0 - a) * b * c;
                        const double w111 = a * b * c;
                        for (int q = 0; q < 2; ++q) {
                            const double *A = q ? A1 : A0;
                            double s = w000 * nodal_read(g, A, ii, jj, kk) +
                                       w100 * nodal_read(g, A, ii + 1, jj, kk) +
                                       w010 * nodal_read(g, A, ii, jj + 1, kk) +
                                       w110 * nodal_read(g, A, ii + 1, jj + 1, kk) +
                                       w001 * nodal_read(g, A, ii, jj, kk + 1) +
                                       w101 * nodal_read(g, A, ii + 1, jj, kk + 1) +
                                       w011 * nodal_read(g, A, ii, jj + 1, kk + 1) +
                                       w111 * nodal_read(g, A, ii + 1, jj + 1, kk + 1);
                            if (q) vals1[dim] = s;
                            else vals0[dim] = s;
                        }
                    } else {
                        double wv[3][4];
                        for (int d = 0; d < 3; ++d)
                            for (int l = mn[d]; l < mx[d]; ++l)
                                wv[d][l - mn[d]] = interp1d(O, g.per[d], iv[d] + l, 0, g.hi[d], x[d],
                                                            g.plo[d], g.dx[d]);
                        for (int n = mn[2]; n < mx[2]; ++n)
                            for (int mm = mn[1]; mm < mx[1]; ++mm)
                                for (int l = mn[0]; l < mx[0]; ++l) {
                                    const double w =
                                        wv[0][l - mn[0]] * wv[1][mm - mn[1]] * wv[2][n - mn[2]];
                                    vals0[dim] += w * nodal_read(g, A0, iv[0] + l, iv[1] + mm, iv[2] + n);
                                    vals1[dim] += w * nodal_read(g, A1, iv[0] + l, iv[1] + mm, iv[2] + n);
                                }
                    }
                }
            }
            for (int d = 0; d < 3; ++d) {
                R[(MPM_R_VELP + d) * cap + i] = vals0[d];
                double vo = R[(MPM_R_VEL + d) * cap + i];
                R[(MPM_R_VEL + d) * cap + i] = alpha * vo + alpha * vals1[d] + (1 - alpha) * vals0[d];
            }
        }
        if (grad) {
            if (!grad_done) {
                for (int n = mn[2]; n < mx[2]; ++n)
                    for (int mm = mn[1]; mm < mx[1]; ++mm)
                        for (int l = mn[0]; l < mx[0]; ++l) {
                            const double vx = bv[0][l - mn[0]], vy = bv[1][mm - mn[1]],
                                         vz = bv[2][n - mn[2]];
                            const double gr[3] = {bd[0][l - mn[0]] * vy * vz,
                                                  vx * bd[1][mm - mn[1]] * vz,
                                                  vx * vy * bd[2][n - mn[2]]};
                            for (int d1 = 0; d1 < 3; ++d1) {
                                const double v1 = nodal_read(g, N.vel + d1 * nn, iv[0] + l,
                                                             iv[1] + mm, iv[2] + n);
                                for (int d2 = 0; d2 < 3; ++d2) L[d1][d2] += v1 * gr[d2];
                            }
                        }
            }
            // get_deformation_gradient_tensor (interpolants.H:1732-1787)
            double nrm = 0.0;
            for (int a = 0; a < 3; a++)
                for (int b = 0; b < 3; b++) nrm += L[a][b] * L[a][b];
            if (!(nrm >= 1.0e10)) {
                double Lp[3][3], Fp[3][3];
                for (int a = 0; a < 3; a++)
                    for (int b = 0; b < 3; b++) {
                        Lp[a][b] = (a == b) ? 1.0 + L[a][b] * dt : L[a][b] * dt;
                        Fp[a][b] = R[(MPM_R_DEFGRAD + 3 * a + b) * cap + i];
                    }
                for (int a = 0; a < 3; ++a)
                    for (int b = 0; b < 3; ++b) {
                        double acc = 0.0;
                        for (int q = 0; q < 3; ++q) acc += Lp[a][q] * Fp[q][b];
                        R[(MPM_R_DEFGRAD + a * 3 + b) * cap + i] = acc;
                    }
            }
            int ind = 0;
            for (int d1 = 0; d1 < 3; ++d1)
                for (int d2 = d1; d2 < 3; ++d2) {
                    R[(MPM_R_STRAINRATE + ind) * cap + i] = 0.5 * (L[d1][d2] + L[d2][d1]);
                    ind++;
                }
        }
    }

    if (flags & P_MOVE) {
        for (int d = 0; d < 3; ++d) x[d] += R[(MPM_R_VELP + d) * cap + i] * dt;
        double rin[3], rout[3];
        for (int d = 0; d < 3; ++d) {
            rin[d] = R[(MPM_R_VEL + d) * cap + i];
            rout[d] = rin[d];
        }
        if (ls.nb > 0 && phase == 0) {
            int hit = -1;
            double minphi = kTiny;
            for (int b = 0; b < ls.nb; ++b) {
                double ph_ = ls_value(g, ls.phi[b], x);
                if (ph_ < minphi) {
                    minphi = ph_;
                    hit = b;
                }
            }
            if (hit >= 0) {
                double nrm[3];
                ls_grad(g, ls.phi[hit], x, nrm);
                double gm = 0.0;
                for (int d = 0; d < 3; ++d) gm += nrm[d] * nrm[d];
                gm = sqrt(gm);
                for (int d = 0; d < 3; ++d) nrm[d] /= (gm + kTiny);
                int mod = applybc(rin, rout, ls.mu[hit], nrm, ls.bc[hit]);
                if (mod > 0)
                    for (int d = 0; d < 3; ++d) x[d] += 2.0 * fabs(minphi) * nrm[d];
                for (int d = 0; d < 3; ++d) R[(MPM_R_VEL + d) * cap + i] = rout[d];
            }
        }
        double wv[3] = {0, 0, 0};
        for (int dir = 0; dir < 3; ++dir) {
            int side = -1;
            if (x[dir] < g.plo[dir]) side = 0;
            else if (x[dir] > g.phi[dir]) side = 1;
            if (side < 0) continue;
            const double *wvel = side ? ph.wv_hi : ph.wv_lo;
            for (int d = 0; d < 3; ++d) wv[d] = wvel[dir * 3 + d];
            const double *tab = udf.tab[dir][side];
            if (tab) {
                const int p0 = (dir == 0) ? 1 : 0;
                const int p1 = (dir == 2) ? 1 : 2;
                double f0 = (x[p0] - g.plo[p0]) / g.dx[p0];
                double f1 = (x[p1] - g.plo[p1]) / g.dx[p1];
                f0 = fmax(0.0, fmin((double)g.n[p0], f0));
                f1 = fmax(0.0, fmin((double)g.n[p1], f1));
                const int j0 = (int)f0, j1 = min(j0 + 1, g.n[p0]);
                const int k0 = (int)f1, k1 = min(k0 + 1, g.n[p1]);
                const double wx1 = f0 - j0, wx0 = 1.0 - wx1;
                const double wy1 = f1 - k0, wy0 = 1.0 - wy1;
                const int n1 = g.n[p1] + 1;
                for (int q = 0; q < 3; ++q)
                    wv[q] = wx0 * wy0 * tab[(j0 * n1 + k0) * 3 + q] +
                            wx0 * wy1 * tab[(j0 * n1 + k1) * 3 + q] +
                            wx1 * wy0 * tab[(j1 * n1 + k0) * 3 + q] +
                            wx1 * wy1 * tab[(j1 * n1 + k1) * 3 + q];
            }
            for (int d = 0; d < 3; ++d) rin[d] -= wv[d];
            double nrm[3] = {0, 0, 0};
            nrm[dir] = side ? -1.0 : 1.0;
            int mod = applybc(rin, rout, side ? ph.mu_hi[dir] : ph.mu_lo[dir], nrm,
                              side ? ph.bchi[dir] : ph.bclo[dir]);
            if (mod > 0) x[dir] = 2.0 * (side ? g.phi[dir] : g.plo[dir]) - x[dir];
            for (int q = 0; q < 3; ++q) R[(MPM_R_VEL + q) * cap + i] = rout[q] + wv[q];
        }
        if (flags & P_REDIST) {
            // Redistribute: periodic wrap (AMReX enforcePeriodic) and removal
            // of particles whose cell is outside a non-periodic domain.
            for (int d = 0; d < 3; ++d) {
                if (!g.per[d]) continue;
                const double L_ = g.phi[d] - g.plo[d];
                if (x[d] > g.rhi[d]) {
                    while (x[d] > g.rhi[d]) x[d] -= L_;
                    if (x[d] < g.plo[d]) x[d] = g.plo[d];
                } else if (x[d] < g.plo[d]) {
                    while (x[d] < g.plo[d]) x[d] += L_;
                    if (x[d] > g.rhi[d]) x[d] = g.rhi[d];
                }
            }
            for (int d = 0; d < 3; ++d) {
                if (g.per[d]) continue;
                const int c = (int)floor((x[d] - g.plo[d]) * g.dxi[d]);
                if (c < 0 || c > g.hi[d]) P.id[i] = -1 - P.id[i];
            }
        }
        for (int d = 0; d < 3; ++d) P.pos[d * cap + i] = x[d];
    }

    if ((flags & P_VOLUME) && phase == 0) {
        double F[9];
        for (int q = 0; q < 9; ++q) F[q] = R[(MPM_R_DEFGRAD + q) * cap + i];
        const double det = det3(F);
        R[MPM_R_JACOBIAN * cap + i] = det;
        const double V = R[MPM_R_VOLINIT * cap + i] * det;
        R[MPM_R_VOLUME * cap + i] = V;
        R[MPM_R_DENSITY * cap + i] = R[MPM_R_MASS * cap + i] / V;
    }

    if ((flags & (P_STRESS | P_STRESS_INIT)) && phase == 0) {
        const bool init = flags & P_STRESS_INIT;
        const double sdt = init ? 0.0 : dt;
        const double rate =
            init ? ph.applied_rate : ((time < ph.applied_time) ? ph.applied_rate : 0.0);
        const bool delta = !init && ph.delta;
        const int model = P.idata[MPM_I_MODEL * cap + i];
        double eps[6], D[6];
        for (int q = 0; q < 6; ++q) {
            D[q] = R[(MPM_R_STRAINRATE + q) * cap + i];
            eps[q] = R[(MPM_R_STRAIN + q) * cap + i] + sdt * D[q];
        }
        eps[XX] += sdt * rate;
        eps[YY] += sdt * rate;
        eps[ZZ] += sdt * rate;
        for (int q = 0; q < 6; ++q) R[(MPM_R_STRAIN + q) * cap + i] = eps[q];
        const double E = R[MPM_R_E * cap + i], nu = R[MPM_R_NU * cap + i];
        double s[6];
        if (delta) {
            // apply_constitutive_model_delta: model 0 only (others rejected on the host)
            double de[6];
            for (int q = 0; q < 6; ++q) de[q] = sdt * D[q];
            de[XX] += sdt * rate;
            de[YY] += sdt * rate;
            de[ZZ] += sdt * rate;
            double ds[6];
            linear_elastic(de, ds, E, nu);
            for (int q = 0; q < 6; ++q)
                R[(MPM_R_STRESS + q) * cap + i] = R[(MPM_R_STRESS + q) * cap + i] + ds[q];
        } else {
            if (model == 0) {
                linear_elastic(eps, s, E, nu);
            } else if (model == 1) {
                const double pinf = 0.0;
                const double pr = R[MPM_R_BULK * cap + i] *
                                      (pow(1.0 / R[MPM_R_JACOBIAN * cap + i],
                                           R[MPM_R_GAMMA * cap + i]) -
                                       1.0) +
                                  pinf;
                R[MPM_R_PRESSURE * cap + i] = pr;
                newtonian(D, s, R[MPM_R_VISC * cap + i], pr);
            } else {
                double F[9];
                for (int q = 0; q < 9; ++q) F[q] = R[(MPM_R_DEFGRAD + q) * cap + i];
                neo_hookean(F, s, E, nu);
            }
            for (int q = 0; q < 6; ++q) R[(MPM_R_STRESS + q) * cap + i] = s[q];
        }
    }
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

__global__ void k_dt_final(Scalars *sc, PhysP ph, int fixed, double fixed_dt) {
    if (fixed) {
        sc->dt = fixed_dt;
        return;
    }
    double dt = __longlong_as_double((long long)sc->dtmin_bits);
    if (dt == __builtin_huge_val()) dt = 1.79769313486231570815e+308;  // numeric_limits max
    dt = ph.cfl * dt;
    dt = fmax(fmin(dt, ph.dt_max), ph.dt_min);
    sc->dt = dt;
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
    for (int d = 0; d < 3; ++d) {
        if (!g.per[d]) continue;
        const double L_ = g.phi[d] - g.plo[d];
        if (x[d] > g.rhi[d]) {
            while (x[d] > g.rhi[d]) x[d] -= L_;
            if (x[d] < g.plo[d]) x[d] = g.plo[d];
        } else if (x[d] < g.plo[d]) {
            while (x[d] < g.plo[d]) x[d] += L_;
            if (x[d] > g.rhi[d]) x[d] = g.rhi[d];
        }
    }
    for (int d = 0; d < 3; ++d) {
        P.pos[d * P.cap + i] = x[d];
        if (g.per[d]) continue;
        const int c = (int)floor((x[d] - g.plo[d]) * g.dxi[d]);
        if (c < 0 || c > g.hi[d]) P.id[i] = -1 - P.id[i];
    }
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
                 const NState &N, int flags, double dt, const Scalars *sc, hipStream_t s) {
    hipLaunchKernelGGL(k_grid, dim3(blocks_for(g.nn)), dim3(kBlock), 0, s, g, ph, ls, udf, N,
                       flags, dt, sc);
}

void launch_particle(const GridP &g, const PhysP &ph, const LsP &ls, const UdfP &udf,
                     const PState &P, const NState &N, int flags, double dt, double time,
                     const Scalars *sc, hipStream_t s) {
    if (P.np <= 0) return;
    hipLaunchKernelGGL(k_particle, dim3(blocks_for(P.np)), dim3(kBlock), 0, s, g, ph, ls, udf,
                       P, N, flags, dt, time, sc);
}

void launch_dt(const GridP &g, const PhysP &ph, const PState &P, Scalars *sc, int fixed,
               double fixed_dt, hipStream_t s) {
    if (!fixed) {
        hipLaunchKernelGGL(k_dt_init, dim3(1), dim3(1), 0, s, sc);
        if (P.np > 0)
            hipLaunchKernelGGL(k_dt, dim3(blocks_for(P.np)), dim3(kBlock), 0, s, g, P, sc);
    }
    hipLaunchKernelGGL(k_dt_final, dim3(1), dim3(1), 0, s, sc, ph, fixed, fixed_dt);
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

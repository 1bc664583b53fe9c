// mpm_device.h — device helpers shared by the step kernels: node indexing,
// periodic deposit targets and ghost reads, level-set interpolation, and the
// per-particle stages (G2P, move, redistribute, volume, stress) written over
// a register record so the same code serves the in-place kernels and the
// reordering kernel of the sorted step.
#pragma once

#include "mpm_internal.h"
#include "mpm_math.h"

namespace xmpm {

// The step's dt (Calculate_time_step, mpm_particle_timestep.cpp:38-111) from
// the minimum candidate: fused takes the minimum the previous step's particle
// kernel accumulated, else the one k_dt reduced; fixed takes fixed_dt; the
// result times scale (a coarse step of several fine steps, 1 otherwise).
// advance: the previous dt first advances the time (that step's end).  One
// thread runs it.
__device__ __forceinline__ void dt_final_one(Scalars *sc, const PhysP &ph, int fixed,
                                             double fixed_dt, int fused, int advance,
                                             double scale) {
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
    dt = dt * scale;
    // Same exchange for dt: the previous step's dt advances the time.
    const unsigned long long old = atomicExch(reinterpret_cast<unsigned long long *>(&sc->dt),
                                              (unsigned long long)__double_as_longlong(dt));
    if (advance) sc->time = sc->time + __longlong_as_double((long long)old);
}

__device__ __forceinline__ int64_t nidx(const GridP &g, int I, int J, int K) {
    return ((int64_t)K * (g.n[1] + 1) + J) * (g.n[0] + 1) + I;
}

// Node tile holding node (I, J, K) (valid indices only).
__device__ __forceinline__ int node_tile(const GridP &g, int I, int J, int K) {
    return ((K / kTileCells) * g.ntn[1] + (J / kTileCells)) * g.ntn[0] + (I / kTileCells);
}

// Deposit targets of an unwrapped node index along one direction.  Real
// deposits land on I when 0 <= I <= n; in a periodic direction the
// reference's ghost-particle image deposits on I + n or I - n when that lies
// in [0, n] (a NeighborParticleContainer image at x +/- L whose stencil
// reaches the valid box).  Out-of-range indices of non-periodic directions are
// dropped, as the reference's clip to the nodal box does
// (mpm_particle_grid_ops.cpp:436-445).
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

// targets() without the arrays: the first target in t0, the second (the
// periodic image, when there are two) in t1; returns the count.
__device__ __forceinline__ int targets2(int I, int n, int per, int &t0, int &t1) {
    const bool in = I >= 0 && I <= n;
    int im = -1;
    if (per) {
        if (I <= 0 && I + n >= 0) im = I + n;
        else if (I >= n && I - n <= n) im = I - n;
    }
    t0 = in ? I : im;
    t1 = im;
    return (in ? 1 : 0) + (im >= 0 ? 1 : 0);
}

__device__ __forceinline__ int floor4(int x) { return x >= 0 ? x / 4 : -((3 - x) / 4); }

// Block sources of node index I along one direction for the deterministic
// P2G: the pairs (cell tile t, block offset a) whose block node, unwrapped
// index G = kTileCells t - 1 + a, the tile flush sends to I (targets2): in a
// periodic direction the image G = I - n, then G = I, then the image
// G = I + n; each by increasing t.  img (optional): whether the source is an
// image.
__device__ __forceinline__ int block_sources(int I, int n, int per, int ntc, int t[6], int a[6],
                                             bool *img = nullptr) {
    static_assert(kTileCells == 4, "floor4");
    int c = 0;
    for (int s = 0; s < 3; ++s) {
        if (s != 1 && !per) continue;
        const int G = I + (s - 1) * n;
        int lo = floor4(G - 2), hi = floor4(G + 1);  // 4 t - 1 <= G <= 4 t + 5
        lo = lo < 0 ? 0 : lo;
        hi = hi > ntc - 1 ? ntc - 1 : hi;
        for (int tt = lo; tt <= hi && c < 6; ++tt) {
            t[c] = tt;
            a[c] = G - 4 * tt + 1;
            if (img) img[c] = s != 1;
            ++c;
        }
    }
    return c;
}

// Sum of node (I, J, K)'s entries in the blocks of the tiles stamped with
// bg.pass, in a fixed order (z, then y, then x sources): out[0..NC), NC =
// bg.nc.  Returns whether any block contributed.  SAME_LAUNCH: the stamps
// and blocks were written earlier in the same launch by other workgroups
// (read the stamps with agent-scope atomic loads).  IMG: blocks of 2 NC
// components, the second NC summed from image sources (MUSL's re-projection
// with periodic images).
template <int NC, bool SAME_LAUNCH = false, bool IMG = false>
__device__ __forceinline__ bool blk_gather(const GridP &g, const BlkGather &bg, int I, int J,
                                           int K, double out[NC]) {
    constexpr int kBN = kTileCells + 3;
    constexpr int NB = IMG ? 2 * NC : NC;
    int tx[6], ax[6], ty[6], ay[6], tz[6], az[6];
    bool mx[6], my[6], mz[6];
    const int cx = block_sources(I, g.n[0], g.per[0], g.ntc[0], tx, ax, IMG ? mx : nullptr);
    const int cy = block_sources(J, g.n[1], g.per[1], g.ntc[1], ty, ay, IMG ? my : nullptr);
    const int cz = block_sources(K, g.n[2], g.per[2], g.ntc[2], tz, az, IMG ? mz : nullptr);
#pragma unroll
    for (int c = 0; c < NC; ++c) out[c] = 0.0;
    bool any = false;
    for (int iz = 0; iz < cz; ++iz)
        for (int iy = 0; iy < cy; ++iy)
            for (int ix = 0; ix < cx; ++ix) {
                const int tile = (tz[iz] * g.ntc[1] + ty[iy]) * g.ntc[0] + tx[ix];
                const int64_t e =
                    SAME_LAUNCH ? (int64_t)__hip_atomic_load(
                                      reinterpret_cast<const unsigned long long *>(bg.slot + tile),
                                      __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                                : bg.slot[tile];
                if ((int)(e >> 32) != bg.pass) continue;
                const double *b = bg.blk + (int64_t)(uint32_t)e * NB * kTNNodes +
                                  (az[iz] * kBN + ay[iy]) * kBN + ax[ix];
                const int c0 = (IMG && (mx[ix] || my[iy] || mz[iz])) ? NC : 0;
#pragma unroll
                for (int c = 0; c < NC; ++c) out[c] += b[(c0 + c) * kTNNodes];
                any = true;
            }
    return any;
}

// Node read by G2P for an unwrapped index: periodic ghost nodes hold the
// value one period away (FillBoundary); non-periodic ghost nodes hold zero.
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

// Level-set value / gradient (mpm_eb.H:176-288), lsref = 1, one ghost layer.
__device__ inline double ls_value(const GridP &g, const double *phi, const double x[3]) {
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

__device__ inline void ls_grad(const GridP &g, const double *phi, const double x[3],
                               double gr[3]) {
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

// Hot per-particle state (every field a step reads or writes).
struct Rec {
    double x[3], v[3], vp[3], D[6], eps[6], sig[6], F[9];
    double V, m, rho, J, p, V0, E, nu, K, gam, mu;
    int phase, model;
    int64_t id;
};

// Real fields carried by Rec, as (record offset, reference field) pairs are
// implied by load_all/store_all below; radius, yacceleration and
// rigid_body_id are cold (never read by the step) and are not reordered.
__device__ __forceinline__ void load_all(const PState &P, int64_t j, Rec &R) {
    const int64_t c = P.cap;
    const double *r = P.r;
    for (int d = 0; d < 3; ++d) {
        R.x[d] = P.pos[d * c + j];
        R.v[d] = r[(MPM_R_VEL + d) * c + j];
        R.vp[d] = r[(MPM_R_VELP + d) * c + j];
    }
    for (int q = 0; q < 6; ++q) {
        R.D[q] = r[(MPM_R_STRAINRATE + q) * c + j];
        R.eps[q] = r[(MPM_R_STRAIN + q) * c + j];
        R.sig[q] = r[(MPM_R_STRESS + q) * c + j];
    }
    for (int q = 0; q < 9; ++q) R.F[q] = r[(MPM_R_DEFGRAD + q) * c + j];
    R.V = r[MPM_R_VOLUME * c + j];
    R.m = r[MPM_R_MASS * c + j];
    R.rho = r[MPM_R_DENSITY * c + j];
    R.J = r[MPM_R_JACOBIAN * c + j];
    R.p = r[MPM_R_PRESSURE * c + j];
    R.V0 = r[MPM_R_VOLINIT * c + j];
    R.E = r[MPM_R_E * c + j];
    R.nu = r[MPM_R_NU * c + j];
    R.K = r[MPM_R_BULK * c + j];
    R.gam = r[MPM_R_GAMMA * c + j];
    R.mu = r[MPM_R_VISC * c + j];
    R.phase = P.idata[MPM_I_PHASE * c + j];
    R.model = P.idata[MPM_I_MODEL * c + j];
    R.id = P.id[j];
}

__device__ __forceinline__ void store_all(const PState &P, int64_t i, const Rec &R) {
    const int64_t c = P.cap;
    double *r = P.r;
    for (int d = 0; d < 3; ++d) {
        P.pos[d * c + i] = R.x[d];
        r[(MPM_R_VEL + d) * c + i] = R.v[d];
        r[(MPM_R_VELP + d) * c + i] = R.vp[d];
    }
    for (int q = 0; q < 6; ++q) {
        r[(MPM_R_STRAINRATE + q) * c + i] = R.D[q];
        r[(MPM_R_STRAIN + q) * c + i] = R.eps[q];
        r[(MPM_R_STRESS + q) * c + i] = R.sig[q];
    }
    for (int q = 0; q < 9; ++q) r[(MPM_R_DEFGRAD + q) * c + i] = R.F[q];
    r[MPM_R_VOLUME * c + i] = R.V;
    r[MPM_R_MASS * c + i] = R.m;
    r[MPM_R_DENSITY * c + i] = R.rho;
    r[MPM_R_JACOBIAN * c + i] = R.J;
    r[MPM_R_PRESSURE * c + i] = R.p;
    r[MPM_R_VOLINIT * c + i] = R.V0;
    r[MPM_R_E * c + i] = R.E;
    r[MPM_R_NU * c + i] = R.nu;
    r[MPM_R_BULK * c + i] = R.K;
    r[MPM_R_GAMMA * c + i] = R.gam;
    r[MPM_R_VISC * c + i] = R.mu;
    P.idata[MPM_I_PHASE * c + i] = R.phase;
    P.idata[MPM_I_MODEL * c + i] = R.model;
    P.id[i] = R.id;
}

// interpolate_from_grid for one particle (mpm_particle_grid_ops.cpp:967-1096).
__device__ inline void stage_g2p(const GridP &g, const PhysP &ph, const NState &N, Rec &R,
                                 bool uvel, bool ugrad, double dt) {
    const int64_t nn = g.nn;
    const double *x = R.x;
    int iv[3], mn[3], mx[3];
    for (int d = 0; d < 3; ++d) {
        iv[d] = (int)floor((x[d] - g.plo[d]) * g.dxi[d]);
        bounds(iv[d], 0, g.hi[d], g.od[d], g.per[d], mn[d], mx[d]);
    }
    double L[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    double bv[3][4], bd[3][4];
    if (ugrad)
        for (int d = 0; d < 3; ++d)
            for (int l = mn[d]; l < mx[d]; ++l)
                basis1d(g.od[d], g.per[d], iv[d], l, 0, g.hi[d], x[d], g.plo[d], g.dx[d],
                        bv[d][l - mn[d]], bd[d][l - mn[d]]);
    bool grad_done = false;
    if (uvel) {
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
                        bool ok = true;
                        const int a = wrap_read(iv[0] + l, g.n[0], g.per[0], ok);
                        const int b = wrap_read(iv[1] + mm, g.n[1], g.per[1], ok);
                        const int c = wrap_read(iv[2] + n, g.n[2], g.per[2], ok);
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
                        if (ugrad) {
                            const double vx = bv[0][l - mn[0]], vy = bv[1][mm - mn[1]],
                                         vz = bv[2][n - mn[2]];
                            const double gr[3] = {bd[0][l - mn[0]] * vy * vz,
                                                  vx * bd[1][mm - mn[1]] * vz,
                                                  vx * vy * bd[2][n - mn[2]]};
                            for (int d1 = 0; d1 < 3; ++d1)
                                for (int d2 = 0; d2 < 3; ++d2) L[d1][d2] += nv[d1] * gr[d2];
                        }
                    }
            grad_done = ugrad;
        } else {
            for (int dim = 0; dim < 3; ++dim) {
                const double *A0 = N.vel + dim * nn;
                const double *A1 = N.dvel + dim * nn;
                const int O = g.od[dim];
                if (O == 1) {
                    // bilin_interp_two (interpolants.H:583-656): trilinear in all dirs
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
            R.vp[d] = vals0[d];
            R.v[d] = alpha * R.v[d] + alpha * vals1[d] + (1 - alpha) * vals0[d];
        }
    }
    if (ugrad) {
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
                            const double v1 =
                                nodal_read(g, N.vel + d1 * nn, iv[0] + l, iv[1] + mm, iv[2] + n);
                            for (int d2 = 0; d2 < 3; ++d2) L[d1][d2] += v1 * gr[d2];
                        }
                    }
        }
        // get_deformation_gradient_tensor (interpolants.H:1732-1787)
        double nrm = 0.0;
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) nrm += L[a][b] * L[a][b];
        if (!(nrm >= 1.0e10)) {
            double Lp[3][3], Fp[9];
            for (int q = 0; q < 9; ++q) Fp[q] = R.F[q];
            for (int a = 0; a < 3; a++)
                for (int b = 0; b < 3; b++) Lp[a][b] = (a == b) ? 1.0 + L[a][b] * dt : L[a][b] * dt;
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) {
                    double acc = 0.0;
                    for (int q = 0; q < 3; ++q) acc += Lp[a][q] * Fp[q * 3 + b];
                    R.F[a * 3 + b] = acc;
                }
        }
        int ind = 0;
        for (int d1 = 0; d1 < 3; ++d1)
            for (int d2 = d1; d2 < 3; ++d2) {
                R.D[ind] = 0.5 * (L[d1][d2] + L[d2][d1]);
                ind++;
            }
    }
}

// moveParticles (mpm_particle_timestep.cpp:329-576) for one particle:
// position x and velocity v updated from the grid velocity vp; level-set
// bodies (phase 0 only), then the walls with their velocities or UDF tables.
__device__ inline void move_particle(const GridP &g, const PhysP &ph, const LsP &ls,
                                     const UdfP &udf, double x[3], double v[3],
                                     const double vp[3], int phase, double dt) {
    for (int d = 0; d < 3; ++d) x[d] += vp[d] * dt;
    double rin[3], rout[3];
    for (int d = 0; d < 3; ++d) {
        rin[d] = v[d];
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
            for (int d = 0; d < 3; ++d) v[d] = rout[d];
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
        for (int q = 0; q < 3; ++q) v[q] = rout[q] + wv[q];
    }
}

// moveParticles for the record of one particle.
__device__ inline void stage_move(const GridP &g, const PhysP &ph, const LsP &ls,
                                  const UdfP &udf, Rec &R, double dt) {
    move_particle(g, ph, ls, udf, R.x, R.v, R.vp, R.phase, dt);
}

// Redistribute's effect on one particle: periodic wrap (AMReX
// enforcePeriodic) and removal of particles whose cell lies outside a
// non-periodic domain (the id becomes -1 - id, keeping the upload index).
__device__ __forceinline__ void stage_redistribute(const GridP &g, double x[3], int64_t &id) {
    for (int d = 0; d < 3; ++d) {
        if (!g.per[d]) continue;
        const double L_ = g.phi[d] - g.plo[d];
        // bounded shifts: a non-finite or runaway position is left outside
        // (and removed below only in non-periodic directions)
        if (x[d] > g.rhi[d]) {
            for (int k = 0; k < 64 && x[d] > g.rhi[d]; ++k) x[d] -= L_;
            if (x[d] < g.plo[d]) x[d] = g.plo[d];
        } else if (x[d] < g.plo[d]) {
            for (int k = 0; k < 64 && x[d] < g.plo[d]; ++k) x[d] += L_;
            if (x[d] > g.rhi[d]) x[d] = g.rhi[d];
        }
    }
    for (int d = 0; d < 3; ++d) {
        if (g.per[d]) continue;
        const int c = (int)floor((x[d] - g.plo[d]) * g.dxi[d]);
        if ((c < 0 || c > g.hi[d]) && id >= 0) id = -1 - id;
    }
}

// updateVolume (mpm_particle_timestep.cpp:144-180).
__device__ __forceinline__ void stage_volume(Rec &R) {
    const double det = det3(R.F);
    R.J = det;
    R.V = R.V0 * det;
    R.rho = R.m / R.V;
}

// apply_constitutive_model[_delta] (mpm_particle_container.cpp:43-272).
__device__ inline void stage_stress(const PhysP &ph, Rec &R, double dt, double time, bool init) {
    const double sdt = init ? 0.0 : dt;
    const double rate = init ? ph.applied_rate : ((time < ph.applied_time) ? ph.applied_rate : 0.0);
    const bool delta = !init && ph.delta;
    for (int q = 0; q < 6; ++q) R.eps[q] = R.eps[q] + sdt * R.D[q];
    R.eps[XX] += sdt * rate;
    R.eps[YY] += sdt * rate;
    R.eps[ZZ] += sdt * rate;
    if (delta) {
        double de[6], ds[6];
        for (int q = 0; q < 6; ++q) de[q] = sdt * R.D[q];
        de[XX] += sdt * rate;
        de[YY] += sdt * rate;
        de[ZZ] += sdt * rate;
        linear_elastic(de, ds, R.E, R.nu);
        for (int q = 0; q < 6; ++q) R.sig[q] = R.sig[q] + ds[q];
        return;
    }
    if (R.model == 0) {
        linear_elastic(R.eps, R.sig, R.E, R.nu);
    } else if (R.model == 1) {
        const double pinf = 0.0;
        R.p = R.K * (ool_pow(1.0 / R.J, R.gam) - 1.0) + pinf;
        newtonian(R.D, R.sig, R.mu, R.p);
    } else {
        neo_hookean(R.F, R.sig, R.E, R.nu);
    }
}

}  // namespace xmpm

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

#include "mpm_stages.h"

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
    p2g_one<FORCES>(g, ph, P, i, acc, um, uv, vimg);
}

// Nodal pipeline (see grid_one).
__global__ __launch_bounds__(kBlock) void k_grid(GridP g, PhysP ph, LsP ls, UdfP udf,
                                                 NState N, int flags, double dt_val,
                                                 const Scalars *sc, const int *tile_epoch,
                                                 int epoch) {
    const int64_t node = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (node >= g.nn) return;
    grid_one(g, ph, ls, udf, N, flags, sc ? sc->dt : dt_val, node, tile_epoch, epoch);
}

// Per-particle pipeline (see particle_one).
__global__ __launch_bounds__(kBlock) void k_particle(GridP g, PhysP ph, LsP ls, UdfP udf,
                                                     PState P, NState N, int flags,
                                                     double dt_val, double time_val,
                                                     const Scalars *sc) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.np) return;
    particle_one(g, ph, ls, udf, P, N, flags, sc ? sc->dt : dt_val, sc ? sc->time : time_val, i);
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
                           int advance, double scale) {
    dt_final_one(sc, ph, fixed, fixed_dt, fused, advance, scale);
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

// removeParticlesInsideEB (mpm_init.cpp:1520-1596): a particle where any
// body's level set, interpolated as the level-set BC reads it, is below
// TINYVAL (constants.H:24) is marked removed.
__global__ __launch_bounds__(kBlock) void k_remove_inside_ls(GridP g, LsP ls, PState P,
                                                             int *removed) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.np || P.id[i] < 0) return;
    const double x[3] = {P.pos[i], P.pos[P.cap + i], P.pos[2 * P.cap + i]};
    for (int b = 0; b < ls.nb; ++b) {
        if (ls_value(g, ls.phi[b], x) < 1e-20) {
            P.id[i] = -1;
            atomicAdd(removed, 1);
            break;
        }
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
               double fixed_dt, hipStream_t s, int fused, int advance, double scale) {
    if (!fixed && !fused) {
        hipLaunchKernelGGL(k_dt_init, dim3(1), dim3(1), 0, s, sc);
        if (P.np > 0)
            hipLaunchKernelGGL(k_dt, dim3(blocks_for(P.np)), dim3(kBlock), 0, s, g, P, sc);
    }
    hipLaunchKernelGGL(k_dt_final, dim3(1), dim3(1), 0, s, sc, ph, fixed, fixed_dt,
                       (fixed || !fused) ? 0 : 1, advance, scale);
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

void launch_remove_inside_ls(const GridP &g, const LsP &ls, const PState &P, int *removed,
                             hipStream_t s) {
    if (P.np <= 0 || ls.nb <= 0) return;
    hipLaunchKernelGGL(k_remove_inside_ls, dim3(blocks_for(P.np)), dim3(kBlock), 0, s, g, ls, P,
                       removed);
}

void launch_redistribute(const GridP &g, const PState &P, hipStream_t s) {
    if (P.np <= 0) return;
    hipLaunchKernelGGL(k_redistribute, dim3(blocks_for(P.np)), dim3(kBlock), 0, s, g, P);
}

}  // namespace xmpm

// mpm_sort.hip — the cell-sorted step for gfx950.
//
// Each step the live particles are counting-sorted by cell in tile-major
// order (kTileCells^3-cell tiles, cells x-fastest inside a tile):
//   k_keys     key + rank within the cell (atomic increment returns the rank)
//   scan       exclusive sum of the per-cell counts (hipCUB)
//   k_scatter  perm[start[key] + rank] = current index
// P2G then runs one workgroup per tile: particles are visited cell-major
// across the 64 lanes of a wavefront (lane = cell, so the 64 lanes of one
// ds_add hit 64 distinct nodes), deposits go to an LDS block of
// (kTileCells+3)^3 nodes x 7 components, and the block is flushed with one
// global FP64 atomic per non-zero node component.
// The fused particle kernel gathers each particle through perm and writes it
// to the other buffer in sorted order, so particle state stays cell-sorted
// and the next step's sort is close to the identity.

#include <hipcub/hipcub.hpp>

#include "mpm_device.h"

namespace xmpm {

namespace {

constexpr int kBlock = 256;
constexpr int kTN = kTileCells + 3;          // LDS nodes per edge
constexpr int kTNN = kTN * kTN * kTN;        // LDS nodes per tile
constexpr int kCellsPerTile = kTileCells * kTileCells * kTileCells;

// Blocks are dispatched round-robin over the 8 XCDs; give each XCD a
// contiguous range of tiles so neighbouring tiles (which share halo nodes
// and node reads) meet in the same L2.
__device__ __forceinline__ int xcd_remap(int b, int nb) {
    const int per = nb / 8, rem = nb % 8;
    const int x = b % 8, k = b / 8;
    return x * per + min(x, rem) + k;
}

__device__ __forceinline__ int cell_key(const GridP &g, const int c[3]) {
    const int t[3] = {c[0] / kTileCells, c[1] / kTileCells, c[2] / kTileCells};
    const int tile = (t[2] * g.ntc[1] + t[1]) * g.ntc[0] + t[0];
    const int w = (c[0] % kTileCells) + kTileCells * ((c[1] % kTileCells) +
                                                      kTileCells * (c[2] % kTileCells));
    return tile * kCellsPerTile + w;
}

__global__ __launch_bounds__(kBlock) void k_keys(GridP g, PState P, SortState S) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.np) return;
    int key = S.nkeys;  // removed particles
    if (P.id[i] >= 0) {
        int c[3];
        for (int d = 0; d < 3; ++d) {
            // cell as getParticleCell computes it; a particle outside the
            // domain (not yet redistributed) is binned with the nearest cell
            // and deposits through the global path of the tile kernel.
            int v = (int)floor((P.pos[d * P.cap + i] - g.plo[d]) * g.dxi[d]);
            c[d] = v < 0 ? 0 : (v > g.hi[d] ? g.hi[d] : v);
        }
        key = cell_key(g, c);
    }
    S.key[i] = key;
    S.rank[i] = atomicAdd(&S.count[key], 1);
}

__global__ __launch_bounds__(kBlock) void k_scatter(PState P, SortState S) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.np) return;
    const int key = S.key[i];
    if (key < S.nkeys) S.perm[S.cell_start[key] + S.rank[i]] = (int)i;
    S.count[key] = 0;  // ready for the next sort
}

// One particle's deposit straight to the global accumulators (particles
// outside their tile's LDS block).
__device__ void deposit_global(const GridP &g, const PhysP &ph, const double x[3], double m,
                               const double vel[3], double V, const double T[9], bool forces,
                               bool mass, double *acc, int *tile_epoch, int epoch) {
    int iv[3], mn[3], mx[3];
    for (int d = 0; d < 3; ++d) {
        iv[d] = (int)floor((x[d] - g.plo[d]) * g.dxi[d]);
        bounds(iv[d], 0, g.hi[d], g.od[d], g.per[d], mn[d], mx[d]);
    }
    double bv[3][4], bd[3][4];
    for (int d = 0; d < 3; ++d)
        for (int l = mn[d]; l < mx[d]; ++l)
            basis1d(g.od[d], g.per[d], iv[d], l, 0, g.hi[d], x[d], g.plo[d], g.dx[d],
                    bv[d][l - mn[d]], bd[d][l - mn[d]]);
    const bool slab = ph.extloads && x[0] > ph.slab_lo[0] && x[0] < ph.slab_hi[0] &&
                      x[1] > ph.slab_lo[1] && x[1] < ph.slab_hi[1] && x[2] > ph.slab_lo[2] &&
                      x[2] < ph.slab_hi[2];
    const int64_t nn = g.nn;
    for (int n = mn[2]; n < mx[2]; ++n) {
        int tz[2];
        bool iz[2];
        const int cz = targets(iv[2] + n, g.n[2], g.per[2], tz, iz);
        for (int mm = mn[1]; mm < mx[1]; ++mm) {
            int ty[2];
            bool iy[2];
            const int cy = targets(iv[1] + mm, g.n[1], g.per[1], ty, iy);
            for (int l = mn[0]; l < mx[0]; ++l) {
                int tx[2];
                bool ix[2];
                const int cx = targets(iv[0] + l, g.n[0], g.per[0], tx, ix);
                if (!cx || !cy || !cz) continue;
                const double vx = bv[0][l - mn[0]], vy = bv[1][mm - mn[1]], vz = bv[2][n - mn[2]];
                const double N = vx * vy * vz;
                double val[7];
                val[0] = m * N;
                for (int d = 0; d < 3; ++d) val[1 + d] = m * vel[d] * N;
                for (int d = 0; d < 3; ++d) val[4 + d] = 0.0;
                if (forces) {
                    const double gr[3] = {bd[0][l - mn[0]] * vy * vz, vx * bd[1][mm - mn[1]] * vz,
                                          vx * vy * bd[2][n - mn[2]]};
                    for (int a = 0; a < 3; ++a) {
                        double bf = m * ph.grav[a] * N;
                        if (slab) bf += ph.extforce[a] * N;
                        double tv = 0.0;
                        for (int b = 0; b < 3; ++b) tv += T[a * 3 + b] * gr[b];
                        val[4 + a] = bf + (-V * tv);
                    }
                }
                for (int qz = 0; qz < cz; ++qz)
                    for (int qy = 0; qy < cy; ++qy)
                        for (int qx = 0; qx < cx; ++qx) {
                            const int64_t node = nidx(g, tx[qx], ty[qy], tz[qz]);
                            for (int c = mass ? 0 : 1; c < (forces ? 7 : 4); ++c)
                                unsafeAtomicAdd(&acc[c * nn + node], val[c]);
                            tile_epoch[node_tile(g, tx[qx], ty[qy], tz[qz])] = epoch;
                        }
            }
        }
    }
}

// deposit_onto_grid_momentum (mpm_particle_grid_ops.cpp:333-587) over one
// cell tile.  FORCES: mass, momentum and forces (P2G(1,1,1)); otherwise
// momentum only (the MUSL re-projection P2G(0,1,0)).
template <bool FORCES>
__global__ __launch_bounds__(kBlock) void k_p2g_tile(GridP g, PhysP ph, PState P, SortState S,
                                                     double *acc, int epoch) {
    constexpr int NC = FORCES ? 7 : 3;
    __shared__ double lds[NC * kTNN];
    __shared__ int cs[kCellsPerTile + 1];
    __shared__ int maxc;
    const int tile = xcd_remap(blockIdx.x, gridDim.x);
    const int tid = threadIdx.x;
    if (tid <= kCellsPerTile) cs[tid] = S.cell_start[tile * kCellsPerTile + tid];
    if (tid == 0) maxc = 0;
    __syncthreads();
    if (cs[kCellsPerTile] == cs[0]) return;  // empty tile (uniform over the block)
    for (int e = tid; e < NC * kTNN; e += kBlock) lds[e] = 0.0;
    if (tid < kCellsPerTile) atomicMax(&maxc, cs[tid + 1] - cs[tid]);
    const int tx = tile % g.ntc[0];
    const int ty = (tile / g.ntc[0]) % g.ntc[1];
    const int tz = tile / (g.ntc[0] * g.ntc[1]);
    const int T0[3] = {tx * kTileCells, ty * kTileCells, tz * kTileCells};
    __syncthreads();
    const int64_t cap = P.cap;
    const int cell = tid % kCellsPerTile;
    const int cnt = cs[cell + 1] - cs[cell];
    for (int r = tid / kCellsPerTile; r < maxc; r += kBlock / kCellsPerTile) {
        if (r >= cnt) continue;
        const int64_t j = S.perm[cs[cell] + r];
        if (P.idata[MPM_I_PHASE * cap + j] != 0) continue;
        double x[3];
        int iv[3], mn[3], mx[3];
        bool inside = true;
        for (int d = 0; d < 3; ++d) {
            x[d] = P.pos[d * cap + j];
            iv[d] = (int)floor((x[d] - g.plo[d]) * g.dxi[d]);
            inside &= (iv[d] >= T0[d] && iv[d] < T0[d] + kTileCells);
            bounds(iv[d], 0, g.hi[d], g.od[d], g.per[d], mn[d], mx[d]);
        }
        const double m = P.r[MPM_R_MASS * cap + j];
        double vel[3];
        for (int d = 0; d < 3; ++d) vel[d] = P.r[(MPM_R_VEL + d) * cap + j];
        double T[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        double V = 0.0;
        if (FORCES) {
            int ind = 0;
            for (int a = 0; a < 3; a++)
                for (int b = a; b < 3; b++) {
                    const double s = P.r[(MPM_R_STRESS + ind) * cap + j];
                    T[a * 3 + b] = s;
                    T[b * 3 + a] = s;
                    ind++;
                }
            V = P.r[MPM_R_VOLUME * cap + j];
        }
        if (!inside) {
            deposit_global(g, ph, x, m, vel, V, T, FORCES, FORCES, acc, S.tile_epoch, epoch);
            continue;
        }
        double bv[3][4], bd[3][4];
        for (int d = 0; d < 3; ++d)
            for (int l = mn[d]; l < mx[d]; ++l)
                basis1d(g.od[d], g.per[d], iv[d], l, 0, g.hi[d], x[d], g.plo[d], g.dx[d],
                        bv[d][l - mn[d]], bd[d][l - mn[d]]);
        // Per-particle factors: momentum m v, body force m g (+ slab load),
        // internal-force operator B = -V sigma.
        double mv[3], a[3], B[9];
        const bool slab = FORCES && ph.extloads && x[0] > ph.slab_lo[0] &&
                          x[0] < ph.slab_hi[0] && x[1] > ph.slab_lo[1] && x[1] < ph.slab_hi[1] &&
                          x[2] > ph.slab_lo[2] && x[2] < ph.slab_hi[2];
        for (int d = 0; d < 3; ++d) {
            mv[d] = m * vel[d];
            a[d] = m * ph.grav[d] + (slab ? ph.extforce[d] : 0.0);
        }
        for (int q = 0; q < 9; ++q) B[q] = -V * T[q];
        const int lx = iv[0] - T0[0] + 1, ly = iv[1] - T0[1] + 1, lz = iv[2] - T0[2] + 1;
        for (int n = mn[2]; n < mx[2]; ++n) {
            const double wz = bv[2][n - mn[2]], dwz = bd[2][n - mn[2]];
            for (int mm = mn[1]; mm < mx[1]; ++mm) {
                const double wy = bv[1][mm - mn[1]], dwy = bd[1][mm - mn[1]];
                const double wyz = wy * wz, dwyz = dwy * wz, wydz = wy * dwz;
                const double Mr = m * wyz;
                double Pr[3], C[3], E[3];
                for (int d = 0; d < 3; ++d) {
                    Pr[d] = mv[d] * wyz;
                    if (FORCES) {
                        C[d] = a[d] * wyz + B[d * 3 + 1] * dwyz + B[d * 3 + 2] * wydz;
                        E[d] = B[d * 3 + 0] * wyz;
                    }
                }
                double *row = lds + ((lz + n) * kTN + (ly + mm)) * kTN + lx;
                for (int l = mn[0]; l < mx[0]; ++l) {
                    const double wx = bv[0][l - mn[0]], dwx = bd[0][l - mn[0]];
                    double *p = row + l;
                    if (FORCES) {
                        atomicAdd(p, wx * Mr);
                        atomicAdd(p + 1 * kTNN, wx * Pr[0]);
                        atomicAdd(p + 2 * kTNN, wx * Pr[1]);
                        atomicAdd(p + 3 * kTNN, wx * Pr[2]);
                        atomicAdd(p + 4 * kTNN, wx * C[0] + dwx * E[0]);
                        atomicAdd(p + 5 * kTNN, wx * C[1] + dwx * E[1]);
                        atomicAdd(p + 6 * kTNN, wx * C[2] + dwx * E[2]);
                    } else {
                        atomicAdd(p, wx * Pr[0]);
                        atomicAdd(p + 1 * kTNN, wx * Pr[1]);
                        atomicAdd(p + 2 * kTNN, wx * Pr[2]);
                    }
                }
            }
        }
    }
    __syncthreads();
    // Flush: one global atomic per non-zero node component, periodic image
    // targets as in deposit_global; mark the node tiles that received data.
    const int64_t nn = g.nn;
    for (int e = tid; e < kTNN; e += kBlock) {
        double v[NC];
        bool any = false;
        for (int c = 0; c < NC; ++c) {
            v[c] = lds[c * kTNN + e];
            any |= (v[c] != 0.0);
        }
        if (!any) continue;
        const int ax = e % kTN, ay = (e / kTN) % kTN, az = e / (kTN * kTN);
        int tx_[2], ty_[2], tz_[2];
        bool ix[2], iy[2], iz[2];
        const int cx = targets(T0[0] + ax - 1, g.n[0], g.per[0], tx_, ix);
        const int cy = targets(T0[1] + ay - 1, g.n[1], g.per[1], ty_, iy);
        const int cz = targets(T0[2] + az - 1, g.n[2], g.per[2], tz_, iz);
        for (int qz = 0; qz < cz; ++qz)
            for (int qy = 0; qy < cy; ++qy)
                for (int qx = 0; qx < cx; ++qx) {
                    const int64_t node = nidx(g, tx_[qx], ty_[qy], tz_[qz]);
                    for (int c = 0; c < NC; ++c)
                        if (v[c] != 0.0) unsafeAtomicAdd(&acc[(FORCES ? c : c + 1) * nn + node], v[c]);
                    S.tile_epoch[node_tile(g, tx_[qx], ty_[qy], tz_[qz])] = epoch;
                }
    }
}

// Fused per-particle pipeline of the sorted step: gather particle perm[i]
// from Pin, run the stages, write it to Pout at i.
__global__ __launch_bounds__(kBlock) void k_particle_sorted(GridP g, PhysP ph, LsP ls, UdfP udf,
                                                            PState Pin, PState Pout, SortState S,
                                                            NState N, int flags,
                                                            const Scalars *sc) {
    const int nlive = S.cell_start[S.nkeys];
    const int64_t i = (int64_t)xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (i >= nlive) {
        if (i < Pin.np) Pout.id[i] = -1;  // slot freed by a removed particle
        return;
    }
    const int64_t j = S.perm[i];
    const double dt = sc->dt;
    const double time = sc->time;
    Rec R;
    load_all(Pin, j, R);
    if ((flags & (P_G2P_VEL | P_G2P_GRAD)) && R.phase == 0)
        stage_g2p(g, ph, N, R, flags & P_G2P_VEL, flags & P_G2P_GRAD, dt);
    if (flags & P_MOVE) {
        stage_move(g, ph, ls, udf, R, dt);
        if (flags & P_REDIST) stage_redistribute(g, R.x, R.id);
    }
    if ((flags & P_VOLUME) && R.phase == 0) stage_volume(R);
    if ((flags & P_STRESS) && R.phase == 0) stage_stress(ph, R, dt, time, false);
    store_all(Pout, i, R);
}

inline unsigned blocks_for(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

}  // namespace

size_t sort_scan_bytes(int nkeys) {
    size_t bytes = 0;
    int *dummy = nullptr;
    (void)hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, dummy, dummy, nkeys + 1);
    return bytes;
}

void launch_sort(const GridP &g, const PState &P, SortState &S, hipStream_t s) {
    if (P.np <= 0) return;
    hipLaunchKernelGGL(k_keys, dim3(blocks_for(P.np)), dim3(kBlock), 0, s, g, P, S);
    size_t bytes = S.scan_bytes;
    (void)hipcub::DeviceScan::ExclusiveSum(S.scan_tmp, bytes, S.count, S.cell_start, S.nkeys + 1,
                                           s);
    hipLaunchKernelGGL(k_scatter, dim3(blocks_for(P.np)), dim3(kBlock), 0, s, P, S);
}

void launch_p2g_tiles(const GridP &g, const PhysP &ph, const PState &P, const SortState &S,
                      const NState &N, int epoch, int momentum_only, hipStream_t s) {
    if (P.np <= 0) return;
    const unsigned ntiles = (unsigned)(g.ntc[0] * g.ntc[1] * g.ntc[2]);
    if (momentum_only)
        hipLaunchKernelGGL(k_p2g_tile<false>, dim3(ntiles), dim3(kBlock), 0, s, g, ph, P, S,
                           N.acc, epoch);
    else
        hipLaunchKernelGGL(k_p2g_tile<true>, dim3(ntiles), dim3(kBlock), 0, s, g, ph, P, S, N.acc,
                           epoch);
}

void launch_particle_sorted(const GridP &g, const PhysP &ph, const LsP &ls, const UdfP &udf,
                            const PState &Pin, const PState &Pout, const SortState &S,
                            const NState &N, int flags, const Scalars *sc, hipStream_t s) {
    if (Pin.np <= 0) return;
    hipLaunchKernelGGL(k_particle_sorted, dim3(blocks_for(Pin.np)), dim3(kBlock), 0, s, g, ph,
                       ls, udf, Pin, Pout, S, N, flags, sc);
}

}  // namespace xmpm

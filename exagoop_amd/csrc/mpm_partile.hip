// mpm_partile.hip — the particle-side tile kernels of the cell-sorted step
// for gfx950, built from the bodies of mpm_tiles.h: the fused particle
// pipeline (G2P, move, volume, stress) and MUSL's velocity G2P over the
// non-empty tiles, and their launchers.

#include "mpm_tiles.h"

namespace xmpm {

namespace {

// interpolate_from_grid(update_vel = 1, update_strainrate = 0)
// (mpm_particle_grid_ops.cpp:924-1098), MUSL's first G2P, over the non-empty
// tiles, one workgroup each: the tile's 7^3 node block of v and dv and the
// spline tables staged in LDS (as in particle_tile), one thread per particle
// of the tile: PIC velocity vp and v = alpha (v + dv) + (1 - alpha) vp, in
// place at the particle's slot; phase != 0 unchanged.  OUT: a particle
// outside its tile (outside the domain between redistributions) gathers from
// the nodal arrays; else such a particle raises the error flag.  O: the
// order of every direction (mixed orders interpolate each velocity
// component with its own direction's order: the generic path).
template <int O, bool OUT>
__global__ __launch_bounds__(kBlock) void k_g2p_vel_tile(GridP g, PhysP ph, PState P,
                                                         SortState S, NState N, Scalars *sc) {
    __shared__ PartLds<6> L;
    const int tile = active_tile(S, blockIdx.x);
    if (tile < 0) return;  // uniform over the block
    const int beg = S.cell_start[tile * kCellsPerTile];
    const int end = S.cell_start[tile * kCellsPerTile + kCellsPerTile];
    if (beg == end) return;
    const int tid = threadIdx.x;
    const int tx = tile % g.ntc[0];
    const int ty = (tile / g.ntc[0]) % g.ntc[1];
    const int tz = tile / (g.ntc[0] * g.ntc[1]);
    const int T0[3] = {tx * kTileCells, ty * kTileCells, tz * kTileCells};
    const int64_t nn = g.nn, cap = P.cap;
    for (int e = tid; e < kTNN; e += kBlock) {
        const int a[3] = {e % kTN, (e / kTN) % kTN, e / (kTN * kTN)};
        bool ok = true;
        int I[3];
        for (int d = 0; d < 3; ++d) I[d] = wrap_read(T0[d] - 1 + a[d], g.n[d], g.per[d], ok);
        const int64_t qn = nidx(g, I[0], I[1], I[2]);
        for (int c = 0; c < 3; ++c) {
            L.nodes[e * 6 + c] = ok ? N.vel[c * nn + qn] : 0.0;
            L.nodes[e * 6 + 3 + c] = ok ? N.dvel[c * nn + qn] : 0.0;
        }
    }
    if (O > 1) {
        const int *src = reinterpret_cast<const int *>(&kTab);
        int *dst = reinterpret_cast<int *>(&L.tab);
        for (int e = tid; e < (int)(sizeof(SplineTables) / sizeof(int)); e += kBlock)
            dst[e] = src[e];
    }
    __syncthreads();
    const double alpha = ph.alpha;
    for (int i = beg + tid; i < end; i += kBlock) {
        const int64_t j = S.perm[i];
        if (P.idata[MPM_I_PHASE * cap + j] != 0) continue;
        double x[3], v[3], vals0[3], vals1[3], Lg[9];
        for (int d = 0; d < 3; ++d) {
            x[d] = P.pos[d * cap + j];
            v[d] = P.r[(MPM_R_VEL + d) * cap + j];
        }
        if (!g2p_block<O, true, OUT, false>(g, O > 1 ? L.tab : kTab, L.nodes, T0, x, vals0, vals1,
                                           Lg, &N)) {
            sc->err = 1;
            continue;
        }
        for (int d = 0; d < 3; ++d) {
            P.r[(MPM_R_VELP + d) * cap + j] = vals0[d];
            P.r[(MPM_R_VEL + d) * cap + j] =
                alpha * v[d] + alpha * vals1[d] + (1 - alpha) * vals0[d];
        }
    }
}

// Fused particle pipeline over the non-empty tiles, one workgroup each (see
// particle_tile); the next step's dt candidates go into the running minimum.
template <int O, bool VEL, bool DELTA, int NT, bool OUT, bool GEN>
__global__ __launch_bounds__(NT) void k_particle_tile(GridP g, PhysP ph, LsP ls, UdfP udf,
                                                      PState Pin, PState Pout, SortState S,
                                                      NState N, int redist, Scalars *sc) {
    __shared__ PartLds<VEL ? 6 : 3> lds;
    const int tile = active_tile(S, blockIdx.x);
    if (tile < 0) return;  // uniform over the block
    double cmin = __builtin_huge_val();
    particle_tile<O, VEL, DELTA, NT, OUT, GEN>(g, ph, Pin, Pout, S, N, redist, sc->dt, sc->time,
                                               &sc->err, tile, lds, cmin, &ls, &udf);
    dt_wave_min(cmin, sc);  // every lane reaches this point
}

bool any_udf(const UdfP &udf) {
    for (int d = 0; d < 3; ++d)
        if (udf.tab[d][0] || udf.tab[d][1]) return true;
    return false;
}

}  // namespace

void launch_g2p_vel_tiles(const GridP &g, const PhysP &ph, const PState &P, const SortState &S,
                          const NState &N, int outside, Scalars *sc, hipStream_t s) {
    if (P.np <= 0) return;
    const dim3 gr((unsigned)((S.ntiles + 7) / 8 * 8)), bl(kBlock);
#define XMPM_GV(o)                                                                            \
    do {                                                                                      \
        if (outside)                                                                          \
            hipLaunchKernelGGL((k_g2p_vel_tile<o, true>), gr, bl, 0, s, g, ph, P, S, N, sc);  \
        else                                                                                  \
            hipLaunchKernelGGL((k_g2p_vel_tile<o, false>), gr, bl, 0, s, g, ph, P, S, N, sc); \
    } while (0)
    if (g.od[0] == 1) XMPM_GV(1);
    else if (g.od[0] == 2) XMPM_GV(2);
    else XMPM_GV(3);
#undef XMPM_GV
}

void launch_particle_fast(const GridP &g, const PhysP &ph, const LsP &ls, const UdfP &udf,
                          const PState &Pin, const PState &Pout, const SortState &S,
                          const NState &N, int vel, int redist, int outside, Scalars *sc,
                          hipStream_t s) {
    if (Pin.np <= 0) return;
    const dim3 gr((unsigned)((S.ntiles + 7) / 8 * 8)), bl(kBlock), bl2(2 * kBlock);
    const int O = g.od[0];
    // general: level-set bodies or UDF wall tables in the move
    const bool gen = ls.nb > 0 || any_udf(udf);
    // Few tiles (8 wavefronts per tile) only for the plain form; outside:
    // particles are redistributed every few steps, so one outside the domain
    // may lie outside its tile.
    const bool few = !outside && !gen && (S.waves ? S.waves == 8 : few_tiles(Pin.np));
#define XMPM_PTL(o, v, dl, nt, ou, ge, b)                                                     \
    hipLaunchKernelGGL((k_particle_tile<o, v, dl, nt, ou, ge>), gr, b, 0, s, g, ph, ls, udf,  \
                       Pin, Pout, S, N, redist, sc)
#define XMPM_PT(o, v, dl)                                                                     \
    do {                                                                                      \
        if (few) XMPM_PTL(o, v, dl, 2 * kBlock, false, false, bl2);                           \
        else if (outside && gen) XMPM_PTL(o, v, dl, kBlock, true, true, bl);                  \
        else if (outside) XMPM_PTL(o, v, dl, kBlock, true, false, bl);                        \
        else if (gen) XMPM_PTL(o, v, dl, kBlock, false, true, bl);                            \
        else XMPM_PTL(o, v, dl, kBlock, false, false, bl);                                    \
    } while (0)
#define XMPM_PTO(v, dl)                                                                       \
    do {                                                                                      \
        if (O == 1) XMPM_PT(1, v, dl);                                                        \
        else if (O == 2) XMPM_PT(2, v, dl);                                                   \
        else XMPM_PT(3, v, dl);                                                               \
    } while (0)
    if (ph.delta) {
        if (vel) XMPM_PTO(true, true); else XMPM_PTO(false, true);
    } else {
        if (vel) XMPM_PTO(true, false); else XMPM_PTO(false, false);
    }
#undef XMPM_PTO
#undef XMPM_PT
#undef XMPM_PTL
}

}  // namespace xmpm

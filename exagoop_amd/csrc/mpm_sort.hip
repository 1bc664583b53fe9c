// mpm_sort.hip — the cell-sorted step for gfx950: kernels over one scene
// built from the bodies of mpm_tiles.h (sort, tile P2G), the generic sorted
// particle kernel and their launchers.  The particle-side tile kernels are in
// mpm_partile.hip.

#include <atomic>

#include <hipcub/hipcub.hpp>

#include "mpm_tiles.h"

namespace xmpm {

namespace {

__global__ __launch_bounds__(kBlock) void k_keys(GridP g, PState P, SortState S, PhysP ph,
                                                 DtFold dt) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        S.tiles[0] = 0;  // non-empty tile count, filled by the scatter launch
        if (dt.on) {
            if (dt.bak) *dt.bak = *dt.sc;
            dt_final_one(dt.sc, ph, dt.fixed, dt.fixed_dt, 1, dt.advance, dt.scale);
        }
    }
    keys_wave(g, P, S, i, P.np);
}

// Non-empty tiles into S.tiles[1..], their number into S.tiles[0] (zeroed
// by k_keys), run by the scatter launch's first workgroups (the scan is
// complete there): each workgroup compacts 256 consecutive tiles in order
// and reserves its range with one atomic, so runs of neighbouring tiles stay
// together while the order of the runs depends on scheduling.  Every thread
// of the workgroup must call it (barriers).
__device__ __forceinline__ void tile_list_block(const SortState &S) {
    if ((int64_t)blockIdx.x * kBlock >= S.ntiles) return;  // uniform over the block
    __shared__ int wsum[kBlock / 64];
    __shared__ int base;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int t = blockIdx.x * kBlock + tid;
    const bool f = t < S.ntiles &&
                   S.cell_start[(t + 1) * kCellsPerTile] != S.cell_start[t * kCellsPerTile];
    const unsigned long long b = __ballot(f);
    const int inwave = __popcll(b & ((1ull << lane) - 1ull));
    if (lane == 0) wsum[w] = __popcll(b);
    __syncthreads();
    int off = 0, tot = 0;
    for (int q = 0; q < kBlock / 64; ++q) {
        if (q < w) off += wsum[q];
        tot += wsum[q];
    }
    if (tid == 0) base = tot ? atomicAdd(&S.tiles[0], tot) : 0;
    __syncthreads();
    if (f) S.tiles[1 + base + off + inwave] = t;
}

// perm[start[key] + rank] = i; re-zero the counts; mark output slots past the
// live count as free (removed particles are dropped by the reorder).  With
// packets, also write particle i's P2G packet at its sorted position.
__global__ __launch_bounds__(kBlock) void k_scatter(PState P, SortState S, int64_t *out_id,
                                                    PhysP ph, int packets) {
    tile_list_block(S);  // the first ceil(ntiles / 256) workgroups; every thread
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.np) return;
    scatter_one(P, S, out_id, ph, packets, i);
}

// Fused per-particle pipeline of the sorted step (generic: level sets,
// UDFs, mixed orders): gather particle perm[i] from Pin, run the stages,
// write it to Pout at i.
__global__ __launch_bounds__(kBlock) void k_particle_sorted(GridP g, PhysP ph, LsP ls, UdfP udf,
                                                            PState Pin, PState Pout, SortState S,
                                                            NState N, int flags,
                                                            const Scalars *sc) {
    const int nlive = S.cell_start[S.nkeys];
    const int64_t i = (int64_t)xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (i >= nlive) return;
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

// Deterministic sort: each cell's particles in increasing current index
// (the scatter's ranks come from atomics).  A workgroup takes 256
// consecutive cells, whose particles are one contiguous range of perm: the
// range is staged in LDS with coalesced loads, each thread sorts its cell
// there (insertion sort; cells hold a few particles) and the range is
// written back coalesced.  A range longer than kOrderLds entries is sorted
// in global memory, one thread per cell.
constexpr int kOrderLds = 8192;
__global__ __launch_bounds__(kBlock) void k_cell_order(SortState S) {
    __shared__ int seg[kOrderLds];
    const int c0 = blockIdx.x * kBlock;
    const int c = c0 + threadIdx.x;
    const int cend = min(c0 + kBlock, S.nkeys);
    const int r0 = S.cell_start[c0], r1 = S.cell_start[cend];
    const int b = c < S.nkeys ? S.cell_start[c] : r1;
    const int e = c < S.nkeys ? S.cell_start[c + 1] : r1;
    if (r1 - r0 > kOrderLds) {  // uniform over the block
        for (int i = b + 1; i < e; ++i) {
            const int x = S.perm[i];
            int j = i - 1;
            while (j >= b && S.perm[j] > x) {
                S.perm[j + 1] = S.perm[j];
                --j;
            }
            S.perm[j + 1] = x;
        }
        return;
    }
    for (int i = threadIdx.x; i < r1 - r0; i += kBlock) seg[i] = S.perm[r0 + i];
    __syncthreads();
    for (int i = b + 1 - r0; i < e - r0; ++i) {
        const int x = seg[i];
        int j = i - 1;
        while (j >= b - r0 && seg[j] > x) {
            seg[j + 1] = seg[j];
            --j;
        }
        seg[j + 1] = x;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < r1 - r0; i += kBlock) S.perm[r0 + i] = seg[i];
}

// deposit_onto_grid_momentum over the non-empty tiles, one workgroup each
// (see p2g_tile).  DEFER: particles outside their cell are listed
// (SortState::ocount, rank) for k_deposit_outside; the launch clears the
// counter the next listing launch uses.  IMG: MUSL's re-projection with
// periodic images carrying the top-of-step velocity (P.vsnap).
template <int O, bool FORCES, bool FALLBACK, bool GATHER, bool DET, int NW, bool DEFER,
          bool IMG = false, bool SLAB = false>
__global__ __launch_bounds__(NW * 64) void k_p2g_cells(GridP g, PhysP ph, PState P, SortState S,
                                                       double *acc, int epoch, Scalars *sc) {
    __shared__ P2gLdsOf<FORCES, GATHER, DET, NW, IMG> lds;
    if (DEFER && blockIdx.x == 0 && threadIdx.x == 0) S.ocount[(S.ogen + 1) & 1] = 0;
    const int pos = active_pos(S, blockIdx.x);
    if (pos < 0) return;  // uniform over the block
    p2g_tile<O, FORCES, FALLBACK, GATHER, DET, NW, DEFER, IMG, SLAB>(
        g, ph, P, S, acc, epoch, &sc->err, S.tiles[1 + pos], lds, pos, S.pass);
}

// The deposits of the particles a DEFER tile launch listed, one thread per
// entry (deposit_global, as the inline fallback makes them).  IMG: periodic
// image targets receive the momentum at the top of the step (P.vsnap).
template <bool FORCES, bool GATHER, bool IMG = false>
__global__ __launch_bounds__(kBlock) void k_deposit_outside(GridP g, PhysP ph, PState P,
                                                            SortState S, double *acc, int epoch) {
    // the count as the tile launch's atomics left it (a vector load)
    const int n = __hip_atomic_load(&S.ocount[S.ogen & 1], __ATOMIC_RELAXED,
                                    __HIP_MEMORY_SCOPE_AGENT);
    for (int k = blockIdx.x * kBlock + threadIdx.x; k < n; k += gridDim.x * kBlock) {
        const int64_t e = S.rank[k];
        double x[3], vals[13];
        fallback_vals<FORCES, GATHER>(ph, P, S, e, x, vals);
        if (IMG) {
            const double m = P.r[MPM_R_MASS * P.cap + e];
            double mvi[3];
            for (int d = 0; d < 3; ++d) mvi[d] = m * P.vsnap[d * P.cap + e];
            deposit_global(g, x, vals, FORCES, acc, S.tile_epoch, epoch, mvi);
        } else {
            deposit_global(g, x, vals, FORCES, acc, S.tile_epoch, epoch);
        }
    }
}

// The deterministic P2G's gather pass: node (I, J, K) adds its blocks
// (blk_gather, a fixed order) to the accumulators, which hold only the
// FALLBACK path's global deposits.  Node tiles the tile flush did not mark
// in this epoch received only zeros.
template <int NC, bool IMG = false>
__global__ __launch_bounds__(kBlock) void k_p2g_gather(GridP g, BlkGather bg,
                                                       const int *tile_epoch, int epoch,
                                                       double *acc) {
    const int64_t node = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (node >= g.nn) return;
    const int I = (int)(node % (g.n[0] + 1));
    const int J = (int)((node / (g.n[0] + 1)) % (g.n[1] + 1));
    const int K = (int)(node / ((int64_t)(g.n[0] + 1) * (g.n[1] + 1)));
    if (tile_epoch[node_tile(g, I, J, K)] != epoch) return;
    double s[NC];
    if (!blk_gather<NC, false, IMG>(g, bg, I, J, K, s)) return;
    const int64_t nn = g.nn;
#pragma unroll
    for (int c = 0; c < NC; ++c) acc[(NC == 7 ? c : c + 1) * nn + node] += s[c];
}


inline unsigned blocks_for(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

}  // namespace

// Few tiles: with fewer particles than one full tile (4^3 cells x 8) per
// compute unit, each tile's serial chain sets the kernel time and the tile
// kernels run 8 wavefronts per tile (P2G: a z-slice's row pair per
// wavefront; particle kernel: 512 threads, one pass over a full tile).  With
// more, 4 wavefronts per tile keep more tiles resident per CU (elastic disks,
// 33 k particles: 88.5 -> 70.3 us per step; dam break, 524 k: 240 us with 4
// wavefronts against 282 us with 8).
bool few_tiles(int64_t np) {
    // compute units per device, queried once (contexts on several host
    // threads may ask at once: every writer stores the same value)
    static std::atomic<int> cus[64];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) dev = 0;
    int n = cus[dev].load(std::memory_order_relaxed);
    if (n <= 0) {
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            n <= 0)
            n = 256;
        cus[dev].store(n, std::memory_order_relaxed);
    }
    return np < (int64_t)n * 512;
}


size_t sort_scan_bytes(int nkeys) {
    size_t bytes = 0;
    int *dummy = nullptr;
    (void)hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, dummy, dummy, nkeys + 1);
    return bytes;
}

void launch_sort(const GridP &g, const PhysP &ph, const PState &P, SortState &S,
                 int64_t *out_id, int packets, hipStream_t s, const DtFold &dt) {
    if (P.np <= 0) return;  // (callers fold the dt only with particles)
    S.pkt_valid = packets;
    hipLaunchKernelGGL(k_keys, dim3(blocks_for(P.np)), dim3(kBlock), 0, s, g, P, S, ph, dt);
    size_t bytes = S.scan_bytes;
    (void)hipcub::DeviceScan::ExclusiveSum(S.scan_tmp, bytes, S.count, S.cell_start, S.nkeys + 1,
                                           s);
    // the scatter launch also builds the non-empty tile list
    const unsigned nb = std::max(blocks_for(P.np), (unsigned)((S.ntiles + kBlock - 1) / kBlock));
    hipLaunchKernelGGL(k_scatter, dim3(nb), dim3(kBlock), 0, s, P, S, out_id, ph, packets);
    // deterministic P2G: a fixed particle order in each cell (the packets,
    // written at the scatter's positions, are not reordered: no packets then)
    if (S.blk && !packets)
        hipLaunchKernelGGL(k_cell_order, dim3(blocks_for(S.nkeys)), dim3(kBlock), 0, s, S);
}

void launch_p2g_tiles(const GridP &g, const PhysP &ph, const PState &P, SortState &S,
                      const NState &N, int epoch, int momentum_only, int fallback, Scalars *sc,
                      hipStream_t s) {
    if (P.np <= 0) return;
    const dim3 gr((unsigned)((S.ntiles + 7) / 8 * 8)), bl(kBlock), bl8(8 * 64);
    const bool det = S.blk && !S.pkt_valid;
    const bool few = S.waves ? S.waves == 8 : few_tiles(P.np);
    if (det) ++S.pass;
    // The atomic forms list the particles outside their cell for
    // k_deposit_outside (DEFER = fallback); the deterministic forms deposit
    // them inline.
    S.ocount = sc->ocount;
    // Slab loads: the gathering force forms with the slab term (SLAB, a
    // form of its own so the plain one carries none of its code).
    const bool slab = ph.extloads && !momentum_only;
#define XMPM_PGW(o, f, fb, ga, w, sl)                                                         \
    hipLaunchKernelGGL((k_p2g_cells<o, f, fb, ga, false, w, fb, false, sl>), gr,              \
                       (w) == 8 ? bl8 : bl, 0, s, g, ph, P, S, N.acc, epoch, sc)
#define XMPM_PG(o, f, fb)                                                                     \
    do {                                                                                      \
        if (det && slab)                                                                      \
            hipLaunchKernelGGL((k_p2g_cells<o, f, fb, f, true, 4, false, false, f>), gr, bl,  \
                               0, s, g, ph, P, S, N.acc, epoch, sc);                          \
        else if (det)                                                                         \
            hipLaunchKernelGGL((k_p2g_cells<o, f, fb, f, true, 4, false>), gr, bl, 0, s, g, ph, \
                               P, S, N.acc, epoch, sc);                                       \
        else if (f && !S.pkt_valid && slab) {                                                 \
            if (few) XMPM_PGW(o, f, fb, true, 8, f); else XMPM_PGW(o, f, fb, true, 4, f);     \
        } else if (f && !S.pkt_valid) {                                                       \
            if (few) XMPM_PGW(o, f, fb, true, 8, false);                                      \
            else XMPM_PGW(o, f, fb, true, 4, false);                                          \
        } else {                                                                              \
            if (few) XMPM_PGW(o, f, fb, false, 8, false);                                     \
            else XMPM_PGW(o, f, fb, false, 4, false);                                         \
        }                                                                                     \
    } while (0)
#define XMPM_PGO(f, fb)                                                                       \
    do {                                                                                      \
        if (!g.uniform) XMPM_PG(0, f, fb);                                                    \
        else if (g.od[0] == 1) XMPM_PG(1, f, fb);                                             \
        else if (g.od[0] == 2) XMPM_PG(2, f, fb);                                             \
        else XMPM_PG(3, f, fb);                                                               \
    } while (0)
    // MUSL's re-projection with periodic images (P.vsnap): the atomic form
    // (particles outside their cell deferred, with their image momentum) or
    // the deterministic one (both momentum sums in the blocks).
    const bool img = momentum_only && P.vsnap;
#define XMPM_PGIF(o, fb)                                                                      \
    do {                                                                                      \
        if (det)                                                                              \
            hipLaunchKernelGGL((k_p2g_cells<o, false, fb, false, true, 4, false, true>), gr,  \
                               bl, 0, s, g, ph, P, S, N.acc, epoch, sc);                      \
        else if (few)                                                                         \
            hipLaunchKernelGGL((k_p2g_cells<o, false, fb, false, false, 8, fb, true>), gr,    \
                               bl8, 0, s, g, ph, P, S, N.acc, epoch, sc);                     \
        else                                                                                  \
            hipLaunchKernelGGL((k_p2g_cells<o, false, fb, false, false, 4, fb, true>), gr,    \
                               bl, 0, s, g, ph, P, S, N.acc, epoch, sc);                      \
    } while (0)
#define XMPM_PGI(o)                                                                           \
    do {                                                                                      \
        if (fallback) XMPM_PGIF(o, true); else XMPM_PGIF(o, false);                           \
    } while (0)
    if (img) {
        if (!g.uniform) XMPM_PGI(0);
        else if (g.od[0] == 1) XMPM_PGI(1);
        else if (g.od[0] == 2) XMPM_PGI(2);
        else XMPM_PGI(3);
    } else if (momentum_only) {
        if (fallback) XMPM_PGO(false, true); else XMPM_PGO(false, false);
    } else {
        if (fallback) XMPM_PGO(true, true); else XMPM_PGO(true, false);
    }
#undef XMPM_PGI
#undef XMPM_PGIF
#undef XMPM_PGO
#undef XMPM_PG
#undef XMPM_PGW
    if (fallback && !det) {
        const dim3 go(64);
        if (img)
            hipLaunchKernelGGL((k_deposit_outside<false, false, true>), go, bl, 0, s, g, ph, P,
                               S, N.acc, epoch);
        else if (momentum_only)
            hipLaunchKernelGGL((k_deposit_outside<false, false>), go, bl, 0, s, g, ph, P, S,
                               N.acc, epoch);
        else if (!S.pkt_valid)
            hipLaunchKernelGGL((k_deposit_outside<true, true>), go, bl, 0, s, g, ph, P, S, N.acc,
                               epoch);
        else
            hipLaunchKernelGGL((k_deposit_outside<true, false>), go, bl, 0, s, g, ph, P, S,
                               N.acc, epoch);
        ++S.ogen;
    }
    if (det) {
        BlkGather bg;
        bg.blk = S.blk;
        bg.slot = S.tile_slot;
        bg.pass = S.pass;
        bg.nc = momentum_only ? (img ? 6 : 3) : 7;
        if (img)
            hipLaunchKernelGGL((k_p2g_gather<3, true>), dim3(blocks_for(g.nn)), dim3(kBlock), 0, s,
                               g, bg, S.tile_epoch, epoch, N.acc);
        else if (momentum_only)
            hipLaunchKernelGGL(k_p2g_gather<3>, dim3(blocks_for(g.nn)), dim3(kBlock), 0, s, g, bg,
                               S.tile_epoch, epoch, N.acc);
        else
            hipLaunchKernelGGL(k_p2g_gather<7>, dim3(blocks_for(g.nn)), dim3(kBlock), 0, s, g, bg,
                               S.tile_epoch, epoch, N.acc);
    }
}

void launch_particle_sorted(const GridP &g, const PhysP &ph, const LsP &ls, const UdfP &udf,
                            const PState &Pin, const PState &Pout, const SortState &S,
                            const NState &N, int flags, const Scalars *sc, hipStream_t s) {
    if (Pin.np <= 0) return;
    hipLaunchKernelGGL(k_particle_sorted, dim3(blocks_for(Pin.np)), dim3(kBlock), 0, s, g, ph,
                       ls, udf, Pin, Pout, S, N, flags, sc);
}

}  // namespace xmpm

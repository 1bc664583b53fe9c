// mpm_tiles.h — per-tile and per-particle bodies of the cell-sorted step,
// shared by the one-scene kernels (mpm_sort.hip) and the batched ensemble's
// persistent kernel (mpm_batch.hip).
//
// Each step the live particles are counting-sorted by cell in tile-major
// order (kTileCells^3-cell tiles, cells x-fastest inside a tile):
//   keys_wave    key + rank within the cell (one atomic per run of equal keys
//                in a wavefront; particles arrive nearly sorted)
//   scan         exclusive sum of the per-cell counts
//   scatter_one  perm[start[key] + rank] = current index, P2G packet
// P2G (p2g_tile) takes one tile per 256-thread workgroup.  A wavefront takes
// the 16 cells of one z-slice of the tile, four cells (one x-row) at a time,
// 16 lanes per cell: lane (ky, kz) owns the four stencil nodes along x of one
// (y, z) stencil row and accumulates the cell's particles into registers;
// the 12 1-D weights of a particle are computed once (one per lane) and
// shared by lane shuffles.  Per-row sums go into the tile's LDS node block
// with LDS FP64 atomic adds (the four wavefronts' z-ranges overlap), and the
// block is flushed with one global FP64 atomic per non-zero node component.
// The fused particle body (particle_tile) stages the tile's node block (v,
// dv) in LDS, runs G2P / F / D / move / volume / stress per particle and
// writes the particle to the other buffer in sorted order, so particle state
// stays cell-sorted and the next sort is close to the identity.
//
// The bodies address LDS through the structs below, so a kernel that runs
// several of them can overlay their LDS.
#pragma once

#include "mpm_device.h"

namespace xmpm {

namespace {

constexpr int kBlock = 256;
constexpr int kTN = kTileCells + 3;                    // LDS nodes per tile edge (offsets -1..+2)
constexpr int kTNN = kTN * kTN * kTN;                  // LDS nodes per tile
constexpr int kCellsPerTile = kTileCells * kTileCells * kTileCells;

static_assert(kTileCells == 4 && kBlock == 256, "wavefront/cell mapping assumes 4^3 tiles");
static_assert(kTNN == kTNNodes, "node block size");

// Blocks are dispatched round-robin over the 8 XCDs; give each XCD a
// contiguous range of tiles so neighbouring tiles (which share halo nodes
// and node reads) meet in the same L2.
__device__ __forceinline__ int xcd_remap(int b, int nb) {
    const int per = nb / 8, rem = nb % 8;
    const int x = b % 8, k = b / 8;
    return x * per + min(x, rem) + k;
}

// Workgroup b of a tile-kernel launch takes position p of the non-empty
// tile list: each XCD (b mod 8) gets a contiguous eighth of the list, so the
// work is spread evenly over the XCDs whatever part of the domain holds
// particles, and neighbouring tiles (which share halo nodes) meet in the
// same L2.  Returns -1 for a workgroup without a tile.
__device__ __forceinline__ int active_tile(const SortState &S, int b) {
    const int na = S.tiles[0];
    const int per = (na + 7) >> 3;
    const int x = b & 7, k = b >> 3;
    if (k >= per) return -1;
    const int p = x * per + k;
    return p < na ? S.tiles[1 + p] : -1;
}

// The list position active_tile maps workgroup b to (-1: none).
__device__ __forceinline__ int active_pos(const SortState &S, int b) {
    const int na = S.tiles[0];
    const int per = (na + 7) >> 3;
    const int x = b & 7, k = b >> 3;
    if (k >= per) return -1;
    const int p = x * per + k;
    return p < na ? p : -1;
}

// Accesses to the particle streams of the fused particle pass, which touches
// each field once: non-temporal, so the streams do not evict the nodal data
// the following grid update and sort reuse from the caches.  (The sort's
// packet stream is read back by P2G within the step and stays cached.)
template <typename T>
__device__ __forceinline__ void stream_store(T *p, T v) {
    __builtin_nontemporal_store(v, p);
}

template <typename T>
__device__ __forceinline__ T stream_load(const T *p) {
    return __builtin_nontemporal_load(p);
}

// a[d] for a per-lane d in 0..2 (d = 3 gives a[2]).
template <typename T>
__device__ __forceinline__ T pick3(const T (&a)[3], int d) {
    return d == 0 ? a[0] : (d == 1 ? a[1] : a[2]);
}

__device__ __forceinline__ int cell_key(const GridP &g, const int c[3]) {
    const int t[3] = {c[0] / kTileCells, c[1] / kTileCells, c[2] / kTileCells};
    const int tile = (t[2] * g.ntc[1] + t[1]) * g.ntc[0] + t[0];
    const int w = (c[0] % kTileCells) + kTileCells * ((c[1] % kTileCells) +
                                                      kTileCells * (c[2] % kTileCells));
    return tile * kCellsPerTile + w;
}

// Cell key (tile-major) and rank within the cell of particle i < lim: one
// atomic per run of equal keys in the wavefront (particles arrive nearly
// sorted).  Every lane of the wavefront calls it, i = wavefront base + lane.
__device__ __forceinline__ void keys_wave(const GridP &g, const PState &P, const SortState &S,
                                          int64_t i, int64_t lim) {
    const bool valid = i < lim;
    int key = -1;
    if (valid) {
        key = S.nkeys;  // removed particles
        if (P.id[i] >= 0) {
            int c[3];
            for (int d = 0; d < 3; ++d) {
                // cell as getParticleCell computes it; a particle outside the
                // domain (not yet redistributed) is binned with the nearest
                // cell and deposits through the global path.
                int v = (int)floor((P.pos[d * P.cap + i] - g.plo[d]) * g.dxi[d]);
                c[d] = v < 0 ? 0 : (v > g.hi[d] ? g.hi[d] : v);
            }
            key = cell_key(g, c);
        }
    }
    const int lane = threadIdx.x & 63;
    const int prev = __shfl_up(key, 1, 64);
    const bool head = valid && (lane == 0 || prev != key);
    const unsigned long long heads = __ballot(head);
    const unsigned long long upto = (lane == 63) ? ~0ull : ((2ull << lane) - 1ull);
    const int hl = 63 - __builtin_clzll((heads & upto) | 1ull);  // this run's head lane
    const unsigned long long above = heads & ~upto;
    const int next = above ? __builtin_ctzll(above) : 64;
    int base = 0;
    if (head) {
        int end = next;
        const int64_t wave0 = i - lane;
        if (wave0 + end > lim) end = (int)(lim - wave0);  // lanes past lim are in no run
        base = atomicAdd(&S.count[key], end - lane);
    }
    base = __shfl(base, hl, 64);
    if (!valid) return;
    S.key[i] = key;
    S.rank[i] = base + (lane - hl);
}

// perm[start[key] + rank] = i; re-zero the counts; mark output slots past the
// live count as free (removed particles are dropped by the reorder).  With
// packets, also write particle i's P2G packet at its sorted position.
__device__ __forceinline__ void scatter_one(const PState &P, const SortState &S, int64_t *out_id,
                                            const PhysP &ph, int packets, int64_t i) {
    const int key = S.key[i];
    if (key < S.nkeys) {
        const int slot = S.cell_start[key] + S.rank[i];
        S.perm[slot] = (int)i;
        if (packets) {
            const int64_t cap = P.cap, pc = S.pcap;
            double x[3];
            for (int d = 0; d < 3; ++d) x[d] = P.pos[d * cap + i];
            double v[13];
            for (int e = 0; e < 13; ++e) v[e] = 0.0;
            if (P.idata[MPM_I_PHASE * cap + i] == 0) {
                const double m = P.r[MPM_R_MASS * cap + i];
                v[0] = m;
                for (int d = 0; d < 3; ++d) v[1 + d] = m * P.r[(MPM_R_VEL + d) * cap + i];
                const bool sl = ph.extloads && x[0] > ph.slab_lo[0] && x[0] < ph.slab_hi[0] &&
                                x[1] > ph.slab_lo[1] && x[1] < ph.slab_hi[1] &&
                                x[2] > ph.slab_lo[2] && x[2] < ph.slab_hi[2];
                for (int d = 0; d < 3; ++d) v[4 + d] = m * ph.grav[d] + (sl ? ph.extforce[d] : 0.0);
                const double V = P.r[MPM_R_VOLUME * cap + i];
                for (int e = 0; e < 6; ++e) v[7 + e] = -V * P.r[(MPM_R_STRESS + e) * cap + i];
            }
            for (int d = 0; d < 3; ++d) S.pkt[d * pc + slot] = x[d];
            for (int e = 0; e < 13; ++e) S.pkt[(3 + e) * pc + slot] = v[e];
        }
    }
    S.count[key] = 0;
    if (out_id && i >= S.cell_start[S.nkeys]) out_id[i] = -1;
}

// One particle's deposit straight to the global accumulators (particles
// outside their tile's node block), from its P2G packet values: mass m,
// momentum mv, body force density a and B = -V sigma (xx, xy, xz, yy, yz,
// zz).  Per node: m N, mv N and a N + B grad N.  mv_img (MUSL's re-projection
// with periodic images): the momentum that periodic image targets receive.
__device__ __attribute__((noinline)) void deposit_global(const GridP &g, const double x[3],
                                                         const double *vals, bool forces,
                                                         double *acc, int *tile_epoch, int epoch,
                                                         const double *mv_img = nullptr) {
    const double m = vals[0];
    const double *mv = vals + 1, *a = vals + 4, *B = vals + 7;
    const int sym[9] = {0, 1, 2, 1, 3, 4, 2, 4, 5};
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
                for (int d = 0; d < 3; ++d) val[1 + d] = mv[d] * N;
                for (int d = 0; d < 3; ++d) val[4 + d] = 0.0;
                if (forces) {
                    const double gr[3] = {bd[0][l - mn[0]] * vy * vz, vx * bd[1][mm - mn[1]] * vz,
                                          vx * vy * bd[2][n - mn[2]]};
                    for (int d = 0; d < 3; ++d) {
                        double tv = 0.0;
                        for (int b = 0; b < 3; ++b) tv += B[sym[d * 3 + b]] * gr[b];
                        val[4 + d] = a[d] * N + tv;
                    }
                }
                for (int qz = 0; qz < cz; ++qz)
                    for (int qy = 0; qy < cy; ++qy)
                        for (int qx = 0; qx < cx; ++qx) {
                            const int64_t node = nidx(g, tx[qx], ty[qy], tz[qz]);
                            const bool image = ix[qx] || iy[qy] || iz[qz];
                            for (int c = forces ? 0 : 1; c < (forces ? 7 : 4); ++c) {
                                const double a_ = (mv_img && image && c >= 1 && c <= 3)
                                                      ? mv_img[c - 1] * N
                                                      : val[c];
                                unsafeAtomicAdd(&acc[c * nn + node], a_);
                            }
                            tile_epoch[node_tile(g, tx[qx], ty[qy], tz[qz])] = epoch;
                        }
            }
        }
    }
}

// The four factor pairs (value, derivative) of offsets -1..2 of cell i in
// direction d for P2G: basisval_and_grad (interpolants.H:1592-1632), zero
// for offsets outside compute_bounds (mpm_particle_grid_ops.cpp:255-299).  O
// is the direction's order, or 0 to read it from the grid (mixed orders).
// Away from non-periodic walls every node of a cubic stencil has shape type 3
// and offset k lies in segment 3 - k, so the coefficients are constants;

// the table is read only when a node type or segment differs.
template <int O>
__device__ __forceinline__ void p2g_factors4(const SplineTables &T, const GridP &g, int d, int i,
                                             double x, double f[8]) {
    // d differs between lanes: select among the three uniform values rather
    // than index the parameter block per lane.
    const int od = O ? O : pick3(g.od, d);
    const double dx = pick3(g.dx, d), dxi = pick3(g.dxi, d), plo = pick3(g.plo, d);
    const int hi = pick3(g.hi, d), per = pick3(g.per, d);
    if (od == 1) {
        const double r = (x - (plo + i * dx)) * dxi;
        f[0] = f[1] = f[6] = f[7] = 0.0;
        f[2] = 1.0 - r;
        f[3] = -dxi;
        f[4] = r;
        f[5] = dxi;
        return;
    }
    int mn, mx;
    bounds(i, 0, hi, od, per, mn, mx);
    const bool interior = per || (i >= 3 && i <= hi - 3);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int idx = i + k - 1;
        const double r = (x - (plo + idx * dx)) * dxi;
        double b, dv;
        if (od == 3) {
            const int s = (int)floor(r) + 2;
            if (interior && s == 3 - k) {
                const double *c = kTab.cval[3][3 - k];
                const double *e = kTab.cder[3][3 - k];
                const double r2 = r * r, r3 = r * r * r;
                b = c[0] * r3 + c[1] * r2 + c[2] * r + c[3];
                dv = (e[0] * r2 + e[1] * r + e[2]) * dxi;
            } else {
                const int t = per ? 3 : type_arith(idx, 0, hi);
                b = cubic_val(T, t, r);
                dv = cubic_der(T, t, r) * dxi;
            }
        } else {
            const int t = per ? 3 : type_arith(idx, 0, hi);
            b = quad_val(T, t, r);
            dv = quad_der(T, t, r) * dxi;
        }
        const bool u = (k - 1 >= mn && k - 1 < mx);
        f[2 * k] = u ? b : 0.0;
        f[2 * k + 1] = u ? dv : 0.0;
    }
}

// Position and deposit values (deposit_global's vals) of a particle the tile
// deposit finds outside its cell: FORCES and GATHER, the packet of particle
// slot as scatter_one builds it; FORCES without GATHER, the sorted packet at
// position slot; otherwise (momentum only) m and m v of particle slot.
template <bool FORCES, bool GATHER>
__device__ __forceinline__ void fallback_vals(const PhysP &ph, const PState &P,
                                              const SortState &S, int64_t slot, double x[3],
                                              double vals[13]) {
    const int64_t cap = P.cap, pc = S.pcap;
    if (FORCES && GATHER) {
        const int64_t j = slot;
        for (int d = 0; d < 3; ++d) x[d] = P.pos[d * cap + j];
        const double m = P.r[MPM_R_MASS * cap + j];
        vals[0] = m;
        for (int d = 0; d < 3; ++d) vals[1 + d] = m * P.r[(MPM_R_VEL + d) * cap + j];
        const bool sl = ph.extloads && x[0] > ph.slab_lo[0] && x[0] < ph.slab_hi[0] &&
                        x[1] > ph.slab_lo[1] && x[1] < ph.slab_hi[1] && x[2] > ph.slab_lo[2] &&
                        x[2] < ph.slab_hi[2];
        for (int d = 0; d < 3; ++d) vals[4 + d] = m * ph.grav[d] + (sl ? ph.extforce[d] : 0.0);
        const double V = P.r[MPM_R_VOLUME * cap + j];
        for (int e = 0; e < 6; ++e) vals[7 + e] = -V * P.r[(MPM_R_STRESS + e) * cap + j];
    } else if (FORCES) {
        for (int d = 0; d < 3; ++d) x[d] = S.pkt[d * pc + slot];
        for (int e = 0; e < 13; ++e) vals[e] = S.pkt[(3 + e) * pc + slot];
    } else {
        const int64_t j = slot;
        for (int d = 0; d < 3; ++d) x[d] = P.pos[d * cap + j];
        for (int e = 0; e < 13; ++e) vals[e] = 0.0;
        vals[0] = P.r[MPM_R_MASS * cap + j];
        for (int d = 0; d < 3; ++d) vals[1 + d] = vals[0] * P.r[(MPM_R_VEL + d) * cap + j];
    }
}

// Particle packet of the P2G staging area (doubles): 12 factor pairs
// (value, derivative) for directions x, y, z and offsets -1..2, then m, m v,
// the body force density a = m g (+ slab load) and B = -V sigma (xx, xy, xz,
// yy, yz, zz): packet component c >= 3 of the sorted packet array lands at
// kPkM - 3 + c.  All zero for a slot without a depositing particle.
constexpr int kPkFac = 0, kPkM = 24, kPkMV = 25, kPkA = 28, kPkB = 31, kPk = 40;
constexpr int kSlots = 16;  // packets per wavefront and round: 4 cells x 4 particles
// Packet slot s (cell s >> 2) starts at pk_base(s): the stride of 40 doubles
// keeps 16-byte alignment, the extra 2 doubles per cell put the four cells'
// packets (read by four lane groups at once, each group broadcasting one
// address) in four different LDS bank quads.
constexpr int kPkArea = kSlots * kPk + 2 * (kSlots / 4 - 1);
__device__ __forceinline__ int pk_base(int s) { return s * kPk + 2 * (s >> 2); }

__device__ __forceinline__ int sym_idx(int a, int b) {
    // XX XY XZ / XY YY YZ / XZ YZ ZZ
    return (a == 0) ? b : (a == 1 ? (b == 0 ? 1 : b + 2) : (b == 0 ? 2 : b + 3));
}

// One staging lane's share of a packet: the position component of its
// direction and up to four value components.
struct StageRaw {
    double xd;
    double val[4];
    int64_t slot;  // sorted position (packets), else particle index
    int phase;     // gathered forces deposit: the particle's phase (part-0 lane)
    bool act;
};

// deposit_onto_grid_momentum (mpm_particle_grid_ops.cpp:333-587) over one
// cell tile.  FORCES: mass, momentum and forces (P2G(1,1,1)), gathered from
// the particle state through perm (GATHER, the one-scene step) or read from
// the sorted packets the sort wrote (the batched ensemble); otherwise
// momentum only (the MUSL re-projection P2G(0,1,0)) through perm.
//
// Wavefront w takes the 16 cells of z-slice w of the tile, one row of four
// cells (x = 0..3) per value of cy.  Each row's particles are staged in rounds
// of four per cell: 16 packets are built by the 64 lanes (lane part d < 3
// evaluates the four factor pairs of direction d; every part copies a share
// of the value components) into the wavefront's LDS staging area, the next
// round's loads already in flight.  Then 16 lanes per cell, lane (ky, kz)
// owning the four stencil nodes along x of one (y, z) stencil row, read the
// packets and accumulate into registers; offsets compute_bounds excludes
// have zero factors (they only address nodes outside the domain, which the
// reference clips).  After a row the sums go into the tile's LDS node block
// with LDS FP64 atomic adds: one block instead of a private slab per
// wavefront keeps the LDS at 42 KB, three workgroups per CU (the register
// count allows three waves per SIMD).  At the end the block is flushed with
// one global FP64 atomic per non-zero node component.  FALLBACK: a particle
// outside its cell (not redistributed this step) deposits through
// deposit_global, or (DEFER) is listed for k_deposit_outside, which keeps
// deposit_global's registers and stack out of the tile kernel; without
// FALLBACK such a particle raises the error flag.

// LDS of one P2G tile (NC accumulated components).  DET false: the tile's
// node block, shared by the wavefronts (LDS atomic adds).  DET true: one
// block of kZN nodes per wavefront (its cells' four z-levels), plain adds,
// summed in wavefront order at the flush.
constexpr int kZN = kTN * kTN * 4;
// Permutation entries of a tile staged in LDS by the P2G forms that read the
// particle state through the permutation (a tile with more particles reads
// the rest from global memory).
constexpr int kPermLds = 1024;
template <int NC, bool DET = false, int NPERM = 1, int NW = 4>
struct P2gLds {
    double slab[DET ? 4 * NC * kZN : NC * kTNN];
    double pk[NW][kPkArea];      // staging area per wavefront
    int cs[kCellsPerTile + 1];
    int perm[NPERM];
    SplineTables tab;
};
// The LDS of p2g_tile<O, FORCES, FALLBACK, GATHER, DET, NW, DEFER, IMG>.
template <bool FORCES, bool GATHER, bool DET, int NW = 4, bool IMG = false>
using P2gLdsOf =
    P2gLds<FORCES ? 7 : (IMG ? 6 : 3), DET, (GATHER || !FORCES) ? kPermLds : 1, NW>;

// LDS of one particle tile (NCV node components).
template <int NCV>
struct PartLds {
    double nodes[kTNN * NCV];
    SplineTables tab;
};

// DET: the deterministic form (SortState::blk): private wavefront slabs and
// the whole node block written to list position pos of S.blk, stamped with
// pass in S.tile_slot; no atomics (but those of the FALLBACK global path).
// NW wavefronts per tile: 4 (one z-slice each) or 8 (a z-slice's first or
// second pair of cell rows each, half the chain of rounds per wavefront).
// SLAB: external slab loads in the gathering force deposit (the slot's
// position shared by its staging lanes).
// IMG (momentum only, MUSL's re-projection with periodic faces): a deposit
// that lands on a periodic image node carries the velocity of the particle
// at the top of the step (P.vsnap), as the reference's neighbour particles
// do; the tile accumulates both momentum sums and the flush picks one per
// target.
template <int O, bool FORCES, bool FALLBACK, bool GATHER = false, bool DET = false, int NW = 4,
          bool DEFER = false, bool IMG = false, bool SLAB = false>
__device__ __forceinline__ void p2g_tile(const GridP &g, const PhysP &ph, const PState &P,
                                         const SortState &S, double *acc, int epoch, int *err,
                                         int tile, P2gLdsOf<FORCES, GATHER, DET, NW, IMG> &L,
                                         int pos = 0, int pass = 0) {
    static_assert(NW == 4 || (NW == 8 && !DET), "8 wavefronts per tile: atomic form only");
    static_assert(!IMG || (!FORCES && (DET || DEFER == FALLBACK)),
                  "image momentum: the re-projection (atomic form: fallback deferred)");
    static_assert(!SLAB || (FORCES && GATHER), "slab loads: the gathering force deposit");
    constexpr int NT = NW * 64;
    constexpr int NC = FORCES ? 7 : (IMG ? 6 : 3);
    constexpr int kSlabN = DET ? 4 * NC * kZN : NC * kTNN;
    auto &slab = L.slab;
    auto &pk = L.pk;
    auto &cs = L.cs;
    auto &tab = L.tab;
    const int tid = threadIdx.x;
    // Forms that read through the permutation stage the tile's entries in
    // LDS, so a round's field reads wait on no global read of their own.
    constexpr bool kPermL = GATHER || !FORCES;
    const int pbeg = S.cell_start[tile * kCellsPerTile];
    const int pnum = kPermL ? min(S.cell_start[tile * kCellsPerTile + kCellsPerTile] - pbeg,
                                  kPermLds)
                            : 0;
    __syncthreads();  // the workgroup's previous tile is done with the LDS
    if (tid <= kCellsPerTile) cs[tid] = S.cell_start[tile * kCellsPerTile + tid];
    if (kPermL)
        for (int e = tid; e < pnum; e += NT) L.perm[e] = S.perm[pbeg + e];
    if (O != 1) {
        const int *src = reinterpret_cast<const int *>(&kTab);
        int *dst = reinterpret_cast<int *>(&tab);
        for (int e = tid; e < (int)(sizeof(SplineTables) / sizeof(int)); e += NT)
            dst[e] = src[e];
    }
    for (int e = tid; e < kSlabN; e += NT) slab[e] = 0.0;
    __syncthreads();
    if (cs[kCellsPerTile] == cs[0]) return;  // empty tile (uniform over the block)
    const int wave = tid >> 6, lane = tid & 63;
    double *pkw = pk[wave];
    const int tx = tile % g.ntc[0];
    const int ty = (tile / g.ntc[0]) % g.ntc[1];
    const int tz = tile / (g.ntc[0] * g.ntc[1]);
    const int T0[3] = {tx * kTileCells, ty * kTileCells, tz * kTileCells};
    const int64_t cap = P.cap, pc = S.pcap;
    const SplineTables &TT = (O != 1) ? tab : kTab;
    const int cz = wave & 3;
    // the wavefront's cell rows [r0, r1) of its z-slice
    const int r0 = (NW == 4) ? 0 : 2 * (wave >> 2);
    const int r1 = (NW == 4) ? kTileCells : r0 + 2;
    // accumulation role: group q (cell x inside the tile), lane (ky, kz)
    const int q = lane >> 4;
    const int gl = lane & 15;
    const int ky = gl & 3, kz = gl >> 2;
    // staging role: packet slot s (cell x = s >> 2, particle s & 3 of the
    // round), part 0..3; part d < 3 also evaluates direction d's factors.
    // Value components 3..15 of a packet are split 4/3/3/3 over the parts.
    const int s_slot = lane & 15, s_part = lane >> 4;
    const int s_cell = s_slot >> 2, s_t = s_slot & 3;
    const int v_first = (s_part == 0) ? 3 : 4 + 3 * s_part;
    const int v_n =
        FORCES ? ((s_part == 0) ? 4 : 3) : ((s_part == 0) ? 4 : ((IMG && s_part == 1) ? 3 : 0));
    // largest cell count of each row over the wavefront's four cells
    int mc[kTileCells];
#pragma unroll
    for (int r = 0; r < kTileCells; ++r) {
        const int c0 = kTileCells * (r + kTileCells * cz) + q;
        int m = cs[c0 + 1] - cs[c0];
        m = max(m, __shfl_xor(m, 16, 64));
        m = max(m, __shfl_xor(m, 32, 64));
        mc[r] = m;
    }
    auto load = [&](int ry, int base) {
        StageRaw R;
        const int c0 = kTileCells * (ry + kTileCells * cz) + s_cell;
        const int sbeg = cs[c0], scnt = cs[c0 + 1] - sbeg;
        const int t = base + s_t;
        R.act = t < scnt;
        R.xd = 0.0;
        for (int e = 0; e < 4; ++e) R.val[e] = 0.0;
        R.slot = 0;
        R.phase = 0;
        if (R.act) {
            if (FORCES) {
                R.slot = sbeg + t;
                if (s_part < 3) R.xd = S.pkt[s_part * pc + R.slot];
                for (int e = 0; e < 4; ++e)
                    if (e < v_n) R.val[e] = S.pkt[(v_first + e) * pc + R.slot];
            } else {
                const int o = sbeg + t - pbeg;
                const int64_t j = o < pnum ? L.perm[o] : S.perm[sbeg + t];
                R.slot = j;
                if (P.idata[MPM_I_PHASE * cap + j] != 0) {
                    R.act = false;
                } else {
                    if (s_part < 3) R.xd = P.pos[s_part * cap + j];
                    if (s_part == 0) {
                        R.val[0] = P.r[MPM_R_MASS * cap + j];
                        for (int d = 0; d < 3; ++d)
                            R.val[1 + d] = R.val[0] * P.r[(MPM_R_VEL + d) * cap + j];
                    }
                    if (IMG && s_part == 1) {
                        // m times the velocity at the top of the step: the
                        // packet's slots kPkA.. (no forces in this form)
                        const double m = P.r[MPM_R_MASS * cap + j];
                        for (int d = 0; d < 3; ++d) R.val[d] = m * P.vsnap[d * cap + j];
                    }
                }
            }
        }
        return R;
    };
    // GATHER: the permutation entry of this lane's slot in round (ry, base),
    // -1 for an empty slot or past the last round; the packet's inputs of
    // particle j: part 0 m and v, part 1 m, parts 2 and 3 V and three stress
    // components (the products are formed at staging).
    auto fetch_j = [&](int ry, int base) -> int {
        if (ry >= r1) return -1;
        const int c0 = kTileCells * (ry + kTileCells * cz) + s_cell;
        const int sbeg = cs[c0], scnt = cs[c0 + 1] - sbeg;
        const int t = base + s_t;
        if (t >= scnt) return -1;
        const int o = sbeg + t - pbeg;
        return o < pnum ? L.perm[o] : S.perm[sbeg + t];
    };
    // The phase is read by the slot's part-0 lane only (R.phase) and shared
    // at staging; no field read waits for it.
    auto load_g = [&](int j) {
        StageRaw R;
        R.act = j >= 0;
        R.xd = 0.0;
        for (int e = 0; e < 4; ++e) R.val[e] = 0.0;
        R.slot = j;
        R.phase = 0;
        if (R.act) {
            if (s_part == 0) R.phase = P.idata[MPM_I_PHASE * cap + j];
            if (s_part < 3) R.xd = P.pos[s_part * cap + j];
            if (s_part <= 1) {
                R.val[0] = P.r[MPM_R_MASS * cap + j];
                if (s_part == 0)
                    for (int d = 0; d < 3; ++d) R.val[1 + d] = P.r[(MPM_R_VEL + d) * cap + j];
            } else {
                R.val[0] = P.r[MPM_R_VOLUME * cap + j];
                const int e0 = (s_part == 2) ? 0 : 3;
                for (int e = 0; e < 3; ++e) R.val[1 + e] = P.r[(MPM_R_STRESS + e0 + e) * cap + j];
            }
        }
        return R;
    };
    auto mc_of = [&](int r) {
        return r == 0 ? mc[0] : (r == 1 ? mc[1] : (r == 2 ? mc[2] : mc[3]));
    };
    auto advance = [&](int &ry, int &base) {
        base += 4;
        while (ry < r1 && base >= mc_of(ry)) {
            ++ry;
            base = 0;
        }
    };
    int ry = r0, base = -4;
    advance(ry, base);
    double a4[4][NC];
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
        for (int c = 0; c < NC; ++c) a4[k][c] = 0.0;
    StageRaw cur;
    if (ry < r1) cur = (FORCES && GATHER) ? load_g(fetch_j(ry, base)) : load(ry, base);
    while (ry < r1) {
        int nry = ry, nbase = base;
        advance(nry, nbase);
        StageRaw nxt;
        if (nry < r1)
            nxt = (FORCES && GATHER) ? load_g(fetch_j(nry, nbase)) : load(nry, nbase);
        // ---- staging: this round's 16 packets
        if (FORCES && GATHER) {
            // a particle with phase != 0 deposits nothing (the packet path's
            // zero values); the phase comes from the slot's part-0 lane
            const int ph0 = __shfl(cur.phase, s_slot, 64);
            if (ph0 != 0) cur.act = false;
        }
        {
            const int ic = (s_part == 0) ? T0[0] + s_cell
                                         : ((s_part == 1) ? T0[1] + ry : T0[2] + cz);
            const bool out =
                s_part < 3 && cur.act &&
                (int)floor((cur.xd - pick3(g.plo, s_part)) * pick3(g.dxi, s_part)) != ic;
            const unsigned long long ob = __ballot(out);
            const bool slot_out = ((ob >> s_slot) | (ob >> (s_slot + 16)) |
                                   (ob >> (s_slot + 32))) & 1ull;
            bool use = cur.act;
            if (slot_out) {
                if (FALLBACK) {
                    if (s_part == 3 && cur.act) {
                        if (DEFER) {
                            // listed for the deposit after the tile pass
                            const int k = atomicAdd(&S.ocount[S.ogen & 1], 1);
                            S.rank[k] = (int)cur.slot;
                        } else {
                            double x[3], vals[13];
                            fallback_vals<FORCES, GATHER>(ph, P, S, cur.slot, x, vals);
                            if (IMG) {
                                double mvi[3];
                                for (int d = 0; d < 3; ++d)
                                    mvi[d] = vals[0] * P.vsnap[d * cap + cur.slot];
                                deposit_global(g, x, vals, FORCES, acc, S.tile_epoch, epoch, mvi);
                            } else {
                                deposit_global(g, x, vals, FORCES, acc, S.tile_epoch, epoch);
                            }
                        }
                    }
                } else {
                    *err = 1;
                }
                use = false;
            }
            double *pw = pkw + pk_base(s_slot);
            if (s_part < 3) {
                double f[8];
                if (use) {
                    p2g_factors4<O>(TT, g, s_part, ic, cur.xd, f);
                } else {
#pragma unroll
                    for (int e = 0; e < 8; ++e) f[e] = 0.0;
                }
#pragma unroll
                for (int e = 0; e < 8; ++e) pw[kPkFac + s_part * 8 + e] = f[e];
            }
            if (FORCES && GATHER) {
                // packet values from the gathered fields, with the arithmetic
                // of scatter_one: m, m v (part 0); m g plus the slab load
                // (part 1; the slot's position from its part 0..2 lanes);
                // -V sigma (parts 2, 3)
                bool sl = false;
                if (SLAB) {
                    const double px = __shfl(cur.xd, s_slot, 64);
                    const double py = __shfl(cur.xd, s_slot + 16, 64);
                    const double pz = __shfl(cur.xd, s_slot + 32, 64);
                    sl = px > ph.slab_lo[0] && px < ph.slab_hi[0] && py > ph.slab_lo[1] &&
                         py < ph.slab_hi[1] && pz > ph.slab_lo[2] && pz < ph.slab_hi[2];
                }
                double pv[4];
                if (s_part == 0) {
                    pv[0] = cur.val[0];
                    for (int d = 0; d < 3; ++d) pv[1 + d] = cur.val[0] * cur.val[1 + d];
                } else if (s_part == 1) {
                    for (int d = 0; d < 3; ++d)
                        pv[d] = cur.val[0] * ph.grav[d] + (sl ? ph.extforce[d] : 0.0);
                    pv[3] = 0.0;
                } else {
                    for (int e = 0; e < 3; ++e) pv[e] = -cur.val[0] * cur.val[1 + e];
                    pv[3] = 0.0;
                }
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    if (e < v_n) pw[kPkM - 3 + v_first + e] = use ? pv[e] : 0.0;
            } else {
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    if (e < v_n) pw[kPkM - 3 + v_first + e] = use ? cur.val[e] : 0.0;
            }
        }
        // The staging area is private to the wavefront and LDS operations of
        // one wavefront execute in issue order: the compiler only has to keep
        // the packet writes ahead of the reads below.
        asm volatile("" ::: "memory");
        // ---- accumulation: particle base + tt of each of the row's cells
        const int nt = min(4, mc_of(ry) - base);
        for (int tt = 0; tt < nt; ++tt) {
#pragma clang fp contract(fast)
            const double *pr = pkw + pk_base(q * 4 + tt);
            double wx[4], dwx[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                wx[k] = pr[kPkFac + 2 * k];
                dwx[k] = pr[kPkFac + 2 * k + 1];
            }
            const double wy = pr[kPkFac + 8 + 2 * ky], dwy = pr[kPkFac + 8 + 2 * ky + 1];
            const double wz = pr[kPkFac + 16 + 2 * kz], dwz = pr[kPkFac + 16 + 2 * kz + 1];
            const double wyz = wy * wz;
            if (FORCES) {
                const double Mr = pr[kPkM] * wyz;
                const double dwyz = dwy * wz, wydz = wy * dwz;
                double Pr[3], C[3], E[3];
                for (int d = 0; d < 3; ++d) {
                    Pr[d] = pr[kPkMV + d] * wyz;
                    C[d] = pr[kPkA + d] * wyz + pr[kPkB + sym_idx(d, 1)] * dwyz +
                           pr[kPkB + sym_idx(d, 2)] * wydz;
                    E[d] = pr[kPkB + sym_idx(d, 0)] * wyz;
                }
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    a4[k][0] += wx[k] * Mr;
                    for (int d = 0; d < 3; ++d) a4[k][1 + d] += wx[k] * Pr[d];
                    // two fused multiply-adds into the accumulator
                    for (int d = 0; d < 3; ++d) {
                        a4[k][4 + d] += wx[k] * C[d];
                        a4[k][4 + d] += dwx[k] * E[d];
                    }
                }
            } else {
                double Pr[3];
                for (int d = 0; d < 3; ++d) Pr[d] = pr[kPkMV + d] * wyz;
#pragma unroll
                for (int k = 0; k < 4; ++k)
                    for (int d = 0; d < 3; ++d) a4[k][d] += wx[k] * Pr[d];
                if (IMG) {
                    double Pi[3];
                    for (int d = 0; d < 3; ++d) Pi[d] = pr[kPkA + d] * wyz;
#pragma unroll
                    for (int k = 0; k < 4; ++k)
                        for (int d = 0; d < 3; ++d) a4[k][3 + d] += wx[k] * Pi[d];
                }
            }
        }
        asm volatile("" ::: "memory");  // reads of this round before the next writes
        if (nry != ry) {
            // Row done: lane (q, ky, kz) adds its four nodes (x = q + k,
            // y = ry + ky, z = cz + kz).  The other wavefronts' slices reach
            // the same z-levels: atomics into the shared block, or (DET) into
            // this wavefront's own slab (z-level kz).  Lanes q and q + 1 share
            // three nodes, so even the private slab takes atomic adds (plain
            // read-add-writes of one lane may be batched past another lane's);
            // one wavefront's LDS instructions run in issue order and, for one
            // k, its lanes' nodes are distinct, so the order of the adds to a
            // node is fixed by the code.
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                if (DET) {
                    double *sw = slab + wave * NC * kZN;
                    const int idx = (kz * kTN + (ry + ky)) * kTN + (q + k);
#pragma unroll
                    for (int c = 0; c < NC; ++c) {
                        __hip_atomic_fetch_add(&sw[c * kZN + idx], a4[k][c], __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_WAVEFRONT);
                        a4[k][c] = 0.0;
                    }
                } else {
                    const int idx = ((cz + kz) * kTN + (ry + ky)) * kTN + (q + k);
#pragma unroll
                    for (int c = 0; c < NC; ++c) {
                        __hip_atomic_fetch_add(&slab[c * kTNN + idx], a4[k][c], __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_WORKGROUP);
                        a4[k][c] = 0.0;
                    }
                }
            }
        }
        ry = nry;
        base = nbase;
        cur = nxt;
    }
    __syncthreads();
    // Flush the node block: one global atomic per non-zero node component;
    // periodic image targets as in deposit_global; mark the node tiles that
    // received data.  DET: the wavefront slabs summed in wavefront order,
    // the whole block stored at list position pos (the gather pass adds it
    // to the nodes), the node tiles marked alike.
    const int64_t nn = g.nn;
    double *blk = DET ? S.blk + (int64_t)pos * NC * kTNN : nullptr;
    if (DET && tid == 0) S.tile_slot[tile] = ((int64_t)pass << 32) | (uint32_t)pos;
    for (int e = tid; e < kTNN; e += NT) {
        const int ax = e % kTN, ay = (e / kTN) % kTN, az = e / (kTN * kTN);
        double v[NC];
        if (DET) {
            for (int c = 0; c < NC; ++c) v[c] = 0.0;
            const int w0 = az > 3 ? az - 3 : 0, w1 = az < 3 ? az : 3;
            for (int w = w0; w <= w1; ++w)
                for (int c = 0; c < NC; ++c)
                    v[c] += slab[w * NC * kZN + c * kZN + ((az - w) * kTN + ay) * kTN + ax];
            for (int c = 0; c < NC; ++c) blk[c * kTNN + e] = v[c];
        } else {
            for (int c = 0; c < NC; ++c) v[c] = slab[c * kTNN + e];
        }
        bool any = false;
        for (int c = 0; c < NC; ++c) any |= (v[c] != 0.0);
        if (!any) continue;
        int tx0, tx1, ty0, ty1, tz0, tz1;
        const int cx = targets2(T0[0] + ax - 1, g.n[0], g.per[0], tx0, tx1);
        const int cy = targets2(T0[1] + ay - 1, g.n[1], g.per[1], ty0, ty1);
        const int cz2 = targets2(T0[2] + az - 1, g.n[2], g.per[2], tz0, tz1);
#pragma unroll
        for (int qz = 0; qz < 2; ++qz)
#pragma unroll
            for (int qy = 0; qy < 2; ++qy)
#pragma unroll
                for (int qx = 0; qx < 2; ++qx) {
                    if (qz >= cz2 || qy >= cy || qx >= cx) continue;
                    const int I = qx ? tx1 : tx0, J = qy ? ty1 : ty0, K = qz ? tz1 : tz0;
                    if (IMG && !DET) {
                        // a periodic image target: the second target, or the
                        // first when the block node lies outside the domain
                        const bool image = qx || qy || qz || I != T0[0] + ax - 1 ||
                                           J != T0[1] + ay - 1 || K != T0[2] + az - 1;
                        const int64_t node = nidx(g, I, J, K);
                        for (int d = 0; d < 3; ++d) {
                            const double val = v[image ? 3 + d : d];
                            if (val != 0.0) unsafeAtomicAdd(&acc[(1 + d) * nn + node], val);
                        }
                    } else if (!DET) {
                        const int64_t node = nidx(g, I, J, K);
                        for (int c = 0; c < NC; ++c)
                            if (v[c] != 0.0)
                                unsafeAtomicAdd(&acc[(FORCES ? c : c + 1) * nn + node], v[c]);
                    }
                    S.tile_epoch[node_tile(g, I, J, K)] = epoch;
                }
    }
}

// Calculate_time_step's per-particle candidate (mpm_particle_timestep.cpp:
// 51-102) from the state a step leaves behind.
__device__ __forceinline__ double dt_candidate(const GridP &g, int model, double K, double E,
                                               double nu, double rho, double vmag2) {
    double Cs;
    if (model == 1) {
        Cs = sqrt(K / rho);
    } else {
        const double lam = E * nu / ((1 + nu) * (1 - 2.0 * nu));
        const double mu = E / (2.0 * (1 + nu));
        Cs = sqrt((lam + 2.0 * mu) / rho);
    }
    const double vm = sqrt(vmag2);
    double dxmin = g.dx[0];
    for (int d = 1; d < 3; ++d) dxmin = fmin(dxmin, g.dx[d]);
    return dxmin / (Cs + vm);
}

// Wavefront min and one atomic-min per wavefront into the next step's
// running minimum.  Every lane of the wavefront must call it.
__device__ __forceinline__ void dt_wave_min(double c, Scalars *sc) {
    for (int off = 32; off > 0; off >>= 1) c = fmin(c, __shfl_xor(c, off, 64));
    if ((threadIdx.x & 63) == 0 && c < __builtin_huge_val())
        atomicMin(&sc->dtmin_next_bits, (unsigned long long)__double_as_longlong(c));
}

// 1-D factors of stencil offset l of cell i in one direction: value b and
// derivative d of basisval_and_grad (interpolants.H:1592-1632) and the
// *_interp_two weight w (interpolants.H:943-1005, 1295-1365), all zero for
// an offset outside compute_bounds.  r is formed with the reciprocal cell
// size, so the basis and interpolation weights of one node type coincide.
template <int O>
__device__ __forceinline__ void tile_factors(const SplineTables &T, int per, int i, int l,
                                             int hi, double x, double plo, double dx,
                                             double dxi, bool use, double &w, double &b,
                                             double &d) {
    if (O == 1) {
        const double r = (x - (plo + i * dx)) * dxi;
        b = (l == 0) ? 1.0 - r : r;
        d = (l == 0) ? -dxi : dxi;
        w = b;
    } else {
        const int idx = i + l;
        const int ta = per ? 3 : type_arith(idx, 0, hi);
        const int tc = per ? 3 : type_chain(idx, 0, hi);
        const double r = (x - (plo + idx * dx)) * dxi;
        if (O == 2) {
            b = quad_val(T, ta, r);
            d = quad_der(T, ta, r) * dxi;
            w = (ta == tc) ? b : quad_val(T, tc, r);
        } else {
            b = cubic_val(T, ta, r);
            d = cubic_der(T, ta, r) * dxi;
            w = (ta == tc) ? b : cubic_val(T, tc, r);
        }
    }
    if (!use) w = b = d = 0.0;
}

// G2P of one particle from the LDS node block of its tile: PIC velocity
// vals0 (and FLIP increment vals1 when VEL) and the velocity gradient L
// (row-major).  The z offsets run as a loop and the W x W (y, x) offsets
// of a plane are unrolled; the x and y factors stay in registers.  A
// particle outside the tile (outside the domain between redistributions:
// the sort bins it with the nearest cell): OUT, the same sums node by node
// from the nodal arrays N (rolled loops, factors per node); else false,
// with zero sums.  GRAD false: no velocity gradient (L stays zero).
template <int O, bool VEL, bool OUT = false, bool GRAD = true>
__device__ __forceinline__ bool g2p_block(const GridP &g, const SplineTables &T,
                                          const double *nodes, const int T0[3],
                                          const double x[3], double vals0[3], double vals1[3],
                                          double L[9], const NState *N = nullptr) {
#pragma clang fp contract(fast)
    constexpr int W = (O == 1) ? 2 : 4;
    constexpr int L0 = (O == 1) ? 0 : -1;
    constexpr int NCV = VEL ? 6 : 3;
    int iv[3], mn[3], mx[3];
    bool inside = true;
    for (int d = 0; d < 3; ++d) {
        iv[d] = (int)floor((x[d] - g.plo[d]) * g.dxi[d]);
        bounds(iv[d], 0, g.hi[d], O, g.per[d], mn[d], mx[d]);
        inside &= (iv[d] >= T0[d] && iv[d] < T0[d] + kTileCells);
    }
    for (int d = 0; d < 3; ++d) vals0[d] = vals1[d] = 0.0;
    for (int q = 0; q < 9; ++q) L[q] = 0.0;
    if (!inside) {
        if (!OUT) return false;
        const int64_t nn = g.nn;
#pragma unroll 1
        for (int lz = mn[2]; lz < mx[2]; ++lz) {
            double wz, bz, dz;
            tile_factors<O>(T, g.per[2], iv[2], lz, g.hi[2], x[2], g.plo[2], g.dx[2], g.dxi[2],
                            true, wz, bz, dz);
#pragma unroll 1
            for (int ly = mn[1]; ly < mx[1]; ++ly) {
                double wy, by, dy;
                tile_factors<O>(T, g.per[1], iv[1], ly, g.hi[1], x[1], g.plo[1], g.dx[1],
                                g.dxi[1], true, wy, by, dy);
                const double wyz = wy * wz;
                const double p0 = by * bz, p1 = dy * bz, p2 = by * dz;
#pragma unroll 1
                for (int lx = mn[0]; lx < mx[0]; ++lx) {
                    double wx, bx, dxv;
                    tile_factors<O>(T, g.per[0], iv[0], lx, g.hi[0], x[0], g.plo[0], g.dx[0],
                                    g.dxi[0], true, wx, bx, dxv);
                    bool ok = true;
                    const int I = wrap_read(iv[0] + lx, g.n[0], g.per[0], ok);
                    const int J = wrap_read(iv[1] + ly, g.n[1], g.per[1], ok);
                    const int K = wrap_read(iv[2] + lz, g.n[2], g.per[2], ok);
                    if (!ok) continue;
                    const int64_t qn = nidx(g, I, J, K);
                    const double gr0 = dxv * p0, gr1 = bx * p1, gr2 = bx * p2;
                    if (GRAD)
                        for (int a = 0; a < 3; ++a) {
                            const double nv = N->vel[a * nn + qn];
                            L[a * 3 + 0] += nv * gr0;
                            L[a * 3 + 1] += nv * gr1;
                            L[a * 3 + 2] += nv * gr2;
                        }
                    if (VEL) {
                        const double w = wx * wyz;
                        for (int a = 0; a < 3; ++a) {
                            vals0[a] += w * N->vel[a * nn + qn];
                            vals1[a] += w * N->dvel[a * nn + qn];
                        }
                    }
                }
            }
        }
        return true;
    }
    double fw[2][W], fb[2][W], fd[2][W];
#pragma unroll
    for (int d = 0; d < 2; ++d)
#pragma unroll
        for (int k = 0; k < W; ++k) {
            const int l = L0 + k;
            tile_factors<O>(T, g.per[d], iv[d], l, g.hi[d], x[d], g.plo[d], g.dx[d], g.dxi[d],
                            l >= mn[d] && l < mx[d], fw[d][k], fb[d][k], fd[d][k]);
        }
    const double *blk =
        nodes + (((iv[2] - T0[2] + L0 + 1) * kTN + (iv[1] - T0[1] + L0 + 1)) * kTN +
                 (iv[0] - T0[0] + L0 + 1)) * NCV;
#pragma unroll 1
    for (int kz = 0; kz < W; ++kz) {
        const int l = L0 + kz;
        double wz, bz, dz;
        tile_factors<O>(T, g.per[2], iv[2], l, g.hi[2], x[2], g.plo[2], g.dx[2], g.dxi[2],
                        l >= mn[2] && l < mx[2], wz, bz, dz);
        const double *pz = blk + kz * (kTN * kTN * NCV);
#pragma unroll
        for (int ky = 0; ky < W; ++ky) {
            const double wyz = fw[1][ky] * wz;
            const double p0 = fb[1][ky] * bz, p1 = fd[1][ky] * bz, p2 = fb[1][ky] * dz;
            const double *py = pz + ky * (kTN * NCV);
#pragma unroll
            for (int kx = 0; kx < W; ++kx) {
                const double *q = py + kx * NCV;
                const double gr0 = fd[0][kx] * p0, gr1 = fb[0][kx] * p1, gr2 = fb[0][kx] * p2;
                if (GRAD)
                    for (int a = 0; a < 3; ++a) {
                        const double nv = q[a];
                        L[a * 3 + 0] += nv * gr0;
                        L[a * 3 + 1] += nv * gr1;
                        L[a * 3 + 2] += nv * gr2;
                    }
                if (VEL) {
                    const double w = fw[0][kx] * wyz;
                    for (int a = 0; a < 3; ++a) {
                        vals0[a] += w * q[a];
                        vals1[a] += w * q[3 + a];
                    }
                }
            }
        }
    }
    return true;
}

// Fused per-particle pipeline over one cell tile: the tile's 7^3 node block
// of v (and dv) and the spline coefficient tables are staged in LDS, then
// each thread takes one particle of the tile: G2P (g2p_block), F, D, move +
// wall BCs, redistribution, volume, stress, written in sorted order.  The
// next step's dt candidates are min-reduced into cmin.  O: the order of
// every direction.  GEN: level-set bodies and UDF wall tables in the move
// (ls, udf); otherwise walls with constant velocities only.
template <int O, bool VEL, bool DELTA, int NT = kBlock, bool OUT = false, bool GEN = false>
__device__ __forceinline__ void particle_tile(const GridP &g, const PhysP &ph, const PState &Pin,
                                              const PState &Pout, const SortState &S,
                                              const NState &N, int redist, double dt,
                                              double time, int *err, int tile,
                                              PartLds<VEL ? 6 : 3> &L, double &cmin,
                                              const LsP *ls = nullptr,
                                              const UdfP *udf = nullptr) {
    constexpr int NCV = VEL ? 6 : 3;
    auto &nodes = L.nodes;
    auto &tab = L.tab;
    const int tid = threadIdx.x;
    __syncthreads();  // the workgroup's previous tile is done with the LDS
    const int beg = S.cell_start[tile * kCellsPerTile];
    const int end = S.cell_start[tile * kCellsPerTile + kCellsPerTile];
    if (beg == end) return;  // uniform over the block
    const int tx = tile % g.ntc[0];
    const int ty = (tile / g.ntc[0]) % g.ntc[1];
    const int tz = tile / (g.ntc[0] * g.ntc[1]);
    const int T0[3] = {tx * kTileCells, ty * kTileCells, tz * kTileCells};
    const int64_t nn = g.nn;
    for (int e = tid; e < kTNN; e += NT) {
        const int a[3] = {e % kTN, (e / kTN) % kTN, e / (kTN * kTN)};
        bool ok = true;
        int I[3];
        for (int d = 0; d < 3; ++d) I[d] = wrap_read(T0[d] - 1 + a[d], g.n[d], g.per[d], ok);
        const int64_t qn = nidx(g, I[0], I[1], I[2]);
        for (int c = 0; c < 3; ++c) {
            nodes[e * NCV + c] = ok ? N.vel[c * nn + qn] : 0.0;
            if (VEL) nodes[e * NCV + 3 + c] = ok ? N.dvel[c * nn + qn] : 0.0;
        }
    }
    if (O > 1) {
        static_assert(sizeof(SplineTables) % sizeof(int) == 0, "table copy by words");
        const int *src = reinterpret_cast<const int *>(&kTab);
        int *dst = reinterpret_cast<int *>(&tab);
        for (int e = tid; e < (int)(sizeof(SplineTables) / sizeof(int)); e += NT)
            dst[e] = src[e];
    }
    __syncthreads();
    const int64_t ci = Pin.cap, co = Pout.cap;
    const int niter = (end - beg + NT - 1) / NT;
    // The permutation entry and the position of a thread's next particle are
    // read one iteration ahead, before the current particle's writes (reads
    // and writes complete in issue order, so a read issued after the writes
    // would wait for them).
    int64_t jn = 0;
    double xn[3] = {0.0, 0.0, 0.0};
    if (beg + tid < end) {
        jn = S.perm[beg + tid];
        for (int d = 0; d < 3; ++d) xn[d] = stream_load(&Pin.pos[d * ci + jn]);
    }
    for (int it = 0; it < niter; ++it) {
        const int64_t i = beg + (int64_t)it * NT + tid;
        if (i >= end) continue;
        const int64_t j = jn;
        double x[3], v[3], vp[3], D[6], F[9];
        for (int d = 0; d < 3; ++d) x[d] = xn[d];
        const bool has_next = i + NT < end;
        // unconditional (a clamped index), so no branch waits for it
        const int jnext = S.perm[has_next ? i + NT : i];
        // The gather waits for x only if x is read first: reads complete in
        // issue order.
        asm volatile("" ::: "memory");
        const int phase = stream_load(&Pin.idata[MPM_I_PHASE * ci + j]);
        const bool rigid = phase != 0;
        // Every remaining read of the particle, in one batch ahead of its
        // first write and of the G2P gather, which hides the batch's
        // latency: Pin and Pout never overlap, but the compiler cannot know
        // that, so a read placed after a write would wait a memory round
        // trip of its own.
        for (int d = 0; d < 3; ++d) v[d] = stream_load(&Pin.r[(MPM_R_VEL + d) * ci + j]);
        if (!VEL)
            for (int d = 0; d < 3; ++d) vp[d] = stream_load(&Pin.r[(MPM_R_VELP + d) * ci + j]);
        for (int q = 0; q < 9; ++q) F[q] = stream_load(&Pin.r[(MPM_R_DEFGRAD + q) * ci + j]);
        int64_t id = stream_load(&Pin.id[j]);
        const int model = stream_load(&Pin.idata[MPM_I_MODEL * ci + j]);
        const double m = stream_load(&Pin.r[MPM_R_MASS * ci + j]);
        const double V0 = stream_load(&Pin.r[MPM_R_VOLINIT * ci + j]);
        const double E = stream_load(&Pin.r[MPM_R_E * ci + j]);
        const double nu = stream_load(&Pin.r[MPM_R_NU * ci + j]);
        const double K = stream_load(&Pin.r[MPM_R_BULK * ci + j]);
        const double gam = stream_load(&Pin.r[MPM_R_GAMMA * ci + j]);
        const double mu = stream_load(&Pin.r[MPM_R_VISC * ci + j]);
        double eps[6], sig[6];
        for (int q = 0; q < 6; ++q) eps[q] = stream_load(&Pin.r[(MPM_R_STRAIN + q) * ci + j]);
        double J = 0.0, V = 0.0, rho = 0.0, p = 0.0;
        if (DELTA)
            for (int q = 0; q < 6; ++q) sig[q] = stream_load(&Pin.r[(MPM_R_STRESS + q) * ci + j]);
        // The gather runs for every particle (rigid ones are few and drop
        // the result), so no branch on the phase waits for the phase's read
        // ahead of the rest of the batch.
        double vals0[3], vals1[3], L[9];
        // Every particle lies in its tile when particles are redistributed
        // each step; between redistributions (OUT) a particle outside the
        // domain gathers from the nodal arrays.
        const bool in_tile =
            g2p_block<O, VEL, OUT>(g, O > 1 ? tab : kTab, nodes, T0, x, vals0, vals1, L, &N);
        if (!rigid) {
            if (!in_tile) *err = 1;
            if (VEL) {
                const double alpha = ph.alpha;
                for (int d = 0; d < 3; ++d) {
                    vp[d] = vals0[d];
                    v[d] = alpha * v[d] + alpha * vals1[d] + (1 - alpha) * vals0[d];
                }
            }
            // get_deformation_gradient_tensor (interpolants.H:1732-1787)
            double nrm = 0.0;
            for (int q = 0; q < 9; q++) nrm += L[q] * L[q];
            if (!(nrm >= 1.0e10)) {
                double Fp[9];
                for (int q = 0; q < 9; ++q) Fp[q] = F[q];
                for (int a = 0; a < 3; ++a)
                    for (int b = 0; b < 3; ++b) {
                        double acc = 0.0;
                        for (int q = 0; q < 3; ++q) {
                            const double lp = (a == q) ? 1.0 + L[a * 3 + q] * dt : L[a * 3 + q] * dt;
                            acc += lp * Fp[q * 3 + b];
                        }
                        F[a * 3 + b] = acc;
                    }
            }
            int ind = 0;
            for (int d1 = 0; d1 < 3; ++d1)
                for (int d2 = d1; d2 < 3; ++d2) {
                    D[ind] = 0.5 * (L[d1 * 3 + d2] + L[d2 * 3 + d1]);
                    ind++;
                }
        }
        // Reads that depend on the phase or the model come after the
        // gather: in the batch above, the branch on the model would wait for
        // the model's read, and registers shared between the rigid-only
        // values and the gather's accumulators would make the gather wait
        // for most of the batch.
        if (rigid) {
            if (VEL)
                for (int d = 0; d < 3; ++d) vp[d] = stream_load(&Pin.r[(MPM_R_VELP + d) * ci + j]);
            for (int q = 0; q < 6; ++q) D[q] = stream_load(&Pin.r[(MPM_R_STRAINRATE + q) * ci + j]);
            J = stream_load(&Pin.r[MPM_R_JACOBIAN * ci + j]);
            V = stream_load(&Pin.r[MPM_R_VOLUME * ci + j]);
            rho = stream_load(&Pin.r[MPM_R_DENSITY * ci + j]);
            if (!DELTA)
                for (int q = 0; q < 6; ++q) sig[q] = stream_load(&Pin.r[(MPM_R_STRESS + q) * ci + j]);
        }
        if (rigid || model != 1) p = stream_load(&Pin.r[MPM_R_PRESSURE * ci + j]);
        if (has_next) {
            jn = jnext;
            for (int d = 0; d < 3; ++d) xn[d] = stream_load(&Pin.pos[d * ci + jn]);
        }
        double *R = Pout.r;
        // moveParticles (mpm_particle_timestep.cpp:329-576; GEN: with
        // level-set bodies and UDF wall tables), then Redistribute.  Each
        // output is stored as soon as it is final to keep live ranges short.
        if (GEN) {
            move_particle(g, ph, *ls, *udf, x, v, vp, phase, dt);
            if (redist) stage_redistribute(g, x, id);
        } else {
            for (int d = 0; d < 3; ++d) x[d] += vp[d] * dt;
            double rin[3], rout[3], wv[3] = {0, 0, 0};
            for (int d = 0; d < 3; ++d) rin[d] = rout[d] = v[d];
            for (int dir = 0; dir < 3; ++dir) {
                int side = -1;
                if (x[dir] < g.plo[dir]) side = 0;
                else if (x[dir] > g.phi[dir]) side = 1;
                if (side < 0) continue;
                const double *wvel = side ? ph.wv_hi : ph.wv_lo;
                for (int d = 0; d < 3; ++d) {
                    wv[d] = wvel[dir * 3 + d];
                    rin[d] -= wv[d];
                }
                double nrm_[3] = {0, 0, 0};
                nrm_[dir] = side ? -1.0 : 1.0;
                const int mod = applybc(rin, rout, side ? ph.mu_hi[dir] : ph.mu_lo[dir], nrm_,
                                        side ? ph.bchi[dir] : ph.bclo[dir]);
                if (mod > 0) x[dir] = 2.0 * (side ? g.phi[dir] : g.plo[dir]) - x[dir];
                for (int q = 0; q < 3; ++q) v[q] = rout[q] + wv[q];
            }
            if (redist) stage_redistribute(g, x, id);
        }
        for (int d = 0; d < 3; ++d) {
            stream_store(&Pout.pos[d * co + i], x[d]);
            stream_store(&R[(MPM_R_VEL + d) * co + i], v[d]);
            stream_store(&R[(MPM_R_VELP + d) * co + i], vp[d]);
        }
        stream_store(&Pout.id[i], id);
        stream_store(&Pout.idata[MPM_I_PHASE * co + i], phase);
        const double vmag2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
        for (int q = 0; q < 6; ++q) stream_store(&R[(MPM_R_STRAINRATE + q) * co + i], D[q]);
        // updateVolume and the constitutive update
        stream_store(&R[MPM_R_MASS * co + i], m);
        stream_store(&R[MPM_R_VOLINIT * co + i], V0);
        stream_store(&Pout.idata[MPM_I_MODEL * co + i], model);
        if (!rigid) {
            J = det3(F);
            V = V0 * J;
            rho = m / V;
        }
        stream_store(&R[MPM_R_VOLUME * co + i], V);
        stream_store(&R[MPM_R_DENSITY * co + i], rho);
        stream_store(&R[MPM_R_JACOBIAN * co + i], J);
        for (int q = 0; q < 9; ++q) stream_store(&R[(MPM_R_DEFGRAD + q) * co + i], F[q]);
        stream_store(&R[MPM_R_E * co + i], E);
        stream_store(&R[MPM_R_NU * co + i], nu);
        stream_store(&R[MPM_R_BULK * co + i], K);
        // The dt candidate depends only on state final at this point.
        if (id >= 0 && !rigid) cmin = fmin(cmin, dt_candidate(g, model, K, E, nu, rho, vmag2));
        stream_store(&R[MPM_R_GAMMA * co + i], gam);
        stream_store(&R[MPM_R_VISC * co + i], mu);
        if (!rigid) {
            const double rate = (time < ph.applied_time) ? ph.applied_rate : 0.0;
            for (int q = 0; q < 6; ++q) eps[q] = eps[q] + dt * D[q];
            eps[XX] += dt * rate;
            eps[YY] += dt * rate;
            eps[ZZ] += dt * rate;
            for (int q = 0; q < 6; ++q) stream_store(&R[(MPM_R_STRAIN + q) * co + i], eps[q]);
            if (DELTA) {
                double de[6], ds[6];
                for (int q = 0; q < 6; ++q) de[q] = dt * D[q];
                de[XX] += dt * rate;
                de[YY] += dt * rate;
                de[ZZ] += dt * rate;
                linear_elastic(de, ds, E, nu);
                for (int q = 0; q < 6; ++q) sig[q] = sig[q] + ds[q];
            } else if (model == 2) {
                // (the Tait branch goes last: F, eps, E, nu are dead there)
                neo_hookean(F, sig, E, nu);
            } else if (model == 0) {
                linear_elastic(eps, sig, E, nu);
            } else {
                const double pinf = 0.0;
                p = K * (ool_pow(1.0 / J, gam) - 1.0) + pinf;
                newtonian(D, sig, mu, p);
            }
        } else {
            for (int q = 0; q < 6; ++q) stream_store(&R[(MPM_R_STRAIN + q) * co + i], eps[q]);
        }
        for (int q = 0; q < 6; ++q) stream_store(&R[(MPM_R_STRESS + q) * co + i], sig[q]);
        stream_store(&R[MPM_R_PRESSURE * co + i], p);
    }
}

}  // namespace

}  // namespace xmpm

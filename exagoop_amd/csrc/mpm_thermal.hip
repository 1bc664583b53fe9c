// mpm_thermal.hip — the thermal module (the USE_TEMP build's heat
// conduction) on the device: thermal P2G, the nodal temperature update, the
// temperature wall conditions in the reference's serial order, and thermal
// G2P, run by mpm_step / mpm_init_internal_forces when enabled.
//
// Layout: the particle thermal fields (realData 46..52 of the USE_TEMP
// build) in a component-major record of their own, stride P.cap, in the same
// slot order as the particle state (the sorted step's G2P writes them by
// sorted position into the second buffer, as the particle kernel does the
// mechanical record); the nodal components 20..25 (constants.H:58-63) in one
// array of 6 * nn, component-major.  One thread per particle or node: the
// module is a few small HBM-bound passes next to the mechanical step.

#include <cmath>
#include <limits>
#include <vector>

#include "mpm_context.h"
#include "mpm_device.h"
#include "mpm_tiles.h"

using namespace xmpm;

namespace {

constexpr int kThreads = 256;
// Nodal thermal components, offsets in units of nn (comps 20..25).
enum { TN_MSP = 0, TN_MSPT = 1, TN_T = 2, TN_SRC = 3, TN_DT = 4, TN_MC = 5, TN_COMPS = 6 };

inline int blocks(int64_t n) { return (int)((n + kThreads - 1) / kThreads); }

// deposit_onto_grid_temperature (mpm_particle_grid_ops.cpp:656-884) of
// particle slot i (phase 0) with global atomics.  Deposits land on the valid
// nodes and, in periodic directions, on the node one period away (the
// reference's ghost-particle images), as the momentum deposit does.
template <bool MT, bool SRC>
__device__ void th_deposit_one(const GridP &g, const PState &P, const double *th, double *tn,
                               int64_t i) {
    const int64_t cap = P.cap;
    double x[3];
    int iv[3], mn[3], mx[3];
    for (int d = 0; d < 3; ++d) {
        x[d] = P.pos[d * cap + i];
        iv[d] = (int)floor((x[d] - g.plo[d]) * g.dxi[d]);
        bounds(iv[d], 0, g.hi[d], g.od[d], g.per[d], mn[d], mx[d]);
    }
    const double m = P.r[MPM_R_MASS * cap + i], V = P.r[MPM_R_VOLUME * cap + i];
    const double T = th[MPM_TH_TEMPERATURE * cap + i];
    const double cp = th[MPM_TH_SPECIFIC_HEAT * cap + i];
    const double kc = th[MPM_TH_CONDUCTIVITY * cap + i];
    const double Q = th[MPM_TH_HEAT_SOURCE * cap + i];
    const double q[3] = {th[(MPM_TH_HEAT_FLUX + 0) * cap + i], th[(MPM_TH_HEAT_FLUX + 1) * cap + i],
                         th[(MPM_TH_HEAT_FLUX + 2) * cap + i]};
    const int64_t nn = g.nn;
    // The 1-D factors are evaluated per node: rolled loops, no per-lane
    // arrays, few registers.
#pragma unroll 1
    for (int n = mn[2]; n < mx[2]; ++n) {
        int tz[2];
        bool iz[2];
        const int cz = targets(iv[2] + n, g.n[2], g.per[2], tz, iz);
        double vz, dz;
        basis1d(g.od[2], g.per[2], iv[2], n, 0, g.hi[2], x[2], g.plo[2], g.dx[2], vz, dz);
#pragma unroll 1
        for (int mm = mn[1]; mm < mx[1]; ++mm) {
            int ty[2];
            bool iy[2];
            const int cy = targets(iv[1] + mm, g.n[1], g.per[1], ty, iy);
            double vy, dy;
            basis1d(g.od[1], g.per[1], iv[1], mm, 0, g.hi[1], x[1], g.plo[1], g.dx[1], vy, dy);
#pragma unroll 1
            for (int l = mn[0]; l < mx[0]; ++l) {
                int tx[2];
                bool ix[2];
                const int cx = targets(iv[0] + l, g.n[0], g.per[0], tx, ix);
                if (!cx || !cy || !cz) continue;
                double vx, dxv;
                basis1d(g.od[0], g.per[0], iv[0], l, 0, g.hi[0], x[0], g.plo[0], g.dx[0], vx,
                        dxv);
                const double N = vx * vy * vz;
                double a0 = 0.0, a1 = 0.0, a2 = 0.0, s = 0.0;
                if (MT) {
                    a0 = m * cp * N;
                    a1 = m * cp * T * N;
                    a2 = m * kc * N;
                }
                if (SRC) {
                    const double gr[3] = {dxv * vy * vz, vx * dy * vz, vx * vy * dz};
                    double net = 0.0;
                    for (int d = 0; d < 3; ++d) net += q[d] * gr[d];
                    s = net * V + V * Q * N;
                }
                for (int qz = 0; qz < cz; ++qz)
                    for (int qy = 0; qy < cy; ++qy)
                        for (int qx = 0; qx < cx; ++qx) {
                            const int64_t node = nidx(g, tx[qx], ty[qy], tz[qz]);
                            if (MT) {
                                unsafeAtomicAdd(&tn[TN_MSP * nn + node], a0);
                                unsafeAtomicAdd(&tn[TN_MSPT * nn + node], a1);
                                unsafeAtomicAdd(&tn[TN_MC * nn + node], a2);
                            }
                            if (SRC) unsafeAtomicAdd(&tn[TN_SRC * nn + node], s);
                        }
            }
        }
    }
}

// One thread per particle slot.
template <bool MT, bool SRC>
__global__ __launch_bounds__(kThreads) void k_th_p2g(GridP g, PState P, const double *th,
                                                     double *tn) {
    const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (i >= P.np) return;
    if (P.id[i] < 0) return;
    if (P.idata[MPM_I_PHASE * P.cap + i] != 0) return;
    th_deposit_one<MT, SRC>(g, P, th, tn, i);
}

// The same deposit over the cell-sorted state, one workgroup per non-empty
// tile, with the momentum deposit's scheme (p2g_tile): wavefront w takes the
// 16 cells of z-slice w, in rounds of four particles per cell; the 64 lanes
// stage 16 particles (lane part d < 3 evaluates direction d's four factor
// pairs, parts 0, 1 and 3 read the thermal fields through the permutation)
// into the wavefront's LDS staging area, then lane (ky, kz) of each cell's
// 16 lanes accumulates the four nodes along x of stencil row (ky, kz) in
// registers, adds a finished cell row into the tile's LDS node block (LDS
// FP64 atomics) and the block goes to the nodes with one global FP64 atomic
// per non-zero node component.  Four sums instead of the momentum deposit's
// seven: m c_p N, m c_p T N, m k N and the source V Q N + V q.grad N, with
// the momentum deposit's factors and order of operations (the source as its
// first force component with a = V Q and B's first row V q).  FALLBACK: a
// particle outside its cell (outside the domain between redistributions) is
// listed, as the momentum tile deposit lists them (SortState::ocount, rank),
// for k_th_deposit_outside; without it such a particle raises the error
// flag.
constexpr int kThNC = 4;     // accumulated sums: m c_p, m c_p T, m k, source
constexpr int kThPkV = 24;   // packet: 24 factor pairs, then m c_p, m c_p T, m k, V Q, V q
constexpr int kThPk = 32;    // packet stride (doubles)
constexpr int kThPkArea = kSlots * kThPk + 2 * (kSlots / 4 - 1);
// Packet slot s starts at th_pk_base(s): two extra doubles per cell put the
// four cells' packets, read by four lane groups at once, in different LDS
// bank quads.
__device__ __forceinline__ int th_pk_base(int s) { return s * kThPk + 2 * (s >> 2); }
// DET: one block of kZN nodes per wavefront (its cells' four z-levels),
// summed in wavefront order at the flush, as the deterministic momentum
// deposit does.
template <bool DET>
struct ThP2gLds {
    double slab[DET ? 4 * kThNC * kZN : kThNC * kTNN];
    double pk[4][kThPkArea];
    int cs[kCellsPerTile + 1];
    int perm[kPermLds];
    SplineTables tab;
};

template <int O, bool MT, bool SRC, bool FALLBACK, bool DET>
__global__ __launch_bounds__(kBlock) void k_th_p2g_cells(GridP g, PState P, SortState S,
                                                         const double *th, double *tn,
                                                         int *err) {
    static_assert(kBlock == 256, "four wavefronts per tile, one z-slice each");
    constexpr int kSlabN = DET ? 4 * kThNC * kZN : kThNC * kTNN;
    __shared__ ThP2gLds<DET> L;
    if (FALLBACK && blockIdx.x == 0 && threadIdx.x == 0) S.ocount[(S.ogen + 1) & 1] = 0;
    const int pos = active_pos(S, blockIdx.x);
    if (pos < 0) return;  // uniform over the block
    const int tile = S.tiles[1 + pos];
    const int tid = threadIdx.x;
    const int pbeg = S.cell_start[tile * kCellsPerTile];
    const int pnum = min(S.cell_start[tile * kCellsPerTile + kCellsPerTile] - pbeg, kPermLds);
    if (tid <= kCellsPerTile) L.cs[tid] = S.cell_start[tile * kCellsPerTile + tid];
    for (int e = tid; e < pnum; e += kBlock) L.perm[e] = S.perm[pbeg + e];
    if (O != 1) {
        const int *src = reinterpret_cast<const int *>(&kTab);
        int *dst = reinterpret_cast<int *>(&L.tab);
        for (int e = tid; e < (int)(sizeof(SplineTables) / sizeof(int)); e += kBlock)
            dst[e] = src[e];
    }
    for (int e = tid; e < kSlabN; e += kBlock) L.slab[e] = 0.0;
    __syncthreads();
    if (L.cs[kCellsPerTile] == L.cs[0]) return;  // empty tile (uniform over the block)
    const int wave = tid >> 6, lane = tid & 63;
    double *pkw = L.pk[wave];
    const int tx = tile % g.ntc[0];
    const int ty = (tile / g.ntc[0]) % g.ntc[1];
    const int tz = tile / (g.ntc[0] * g.ntc[1]);
    const int T0[3] = {tx * kTileCells, ty * kTileCells, tz * kTileCells};
    const int64_t cap = P.cap;
    const SplineTables &TT = (O != 1) ? L.tab : kTab;
    const int cz = wave;
    // accumulation role: group q (cell x inside the tile), lane (ky, kz)
    const int q = lane >> 4;
    const int ky = lane & 3, kz = (lane >> 2) & 3;
    // staging role: packet slot s (cell x = s >> 2, particle s & 3 of the
    // round), part 0..3
    const int s_slot = lane & 15, s_part = lane >> 4;
    const int s_cell = s_slot >> 2, s_t = s_slot & 3;
    // largest cell count of each row over the wavefront's four cells
    int mc[kTileCells];
#pragma unroll
    for (int r = 0; r < kTileCells; ++r) {
        const int c0 = kTileCells * (r + kTileCells * cz) + q;
        int m = L.cs[c0 + 1] - L.cs[c0];
        m = max(m, __shfl_xor(m, 16, 64));
        m = max(m, __shfl_xor(m, 32, 64));
        mc[r] = m;
    }
    auto mc_of = [&](int r) {
        return r == 0 ? mc[0] : (r == 1 ? mc[1] : (r == 2 ? mc[2] : mc[3]));
    };
    auto advance = [&](int &ry, int &base) {
        base += 4;
        while (ry < kTileCells && base >= mc_of(ry)) {
            ++ry;
            base = 0;
        }
    };
    // The permutation entry of this lane's slot in round (ry, base), -1 for
    // an empty slot; then the slot's inputs: part 0 the phase, m, c_p, T and
    // k; part 1 V and Q; part 3 V and q; parts 0..2 one position component.
    auto fetch_j = [&](int ry, int base) -> int {
        const int c0 = kTileCells * (ry + kTileCells * cz) + s_cell;
        const int sbeg = L.cs[c0], scnt = L.cs[c0 + 1] - sbeg;
        const int t = base + s_t;
        if (t >= scnt) return -1;
        const int o = sbeg + t - pbeg;
        return o < pnum ? L.perm[o] : S.perm[sbeg + t];
    };
    struct Raw {
        double xd, val[4];
        int j, phase;
    };
    auto load = [&](int j) {
        Raw R;
        R.j = j;
        R.phase = 0;
        R.xd = 0.0;
        for (int e = 0; e < 4; ++e) R.val[e] = 0.0;
        if (j >= 0) {
            if (s_part == 0) R.phase = P.idata[MPM_I_PHASE * cap + j];
            if (s_part < 3) R.xd = P.pos[s_part * cap + j];
            if (MT && s_part == 0) {
                R.val[0] = P.r[MPM_R_MASS * cap + j];
                R.val[1] = th[MPM_TH_SPECIFIC_HEAT * cap + j];
                R.val[2] = th[MPM_TH_TEMPERATURE * cap + j];
                R.val[3] = th[MPM_TH_CONDUCTIVITY * cap + j];
            }
            if (SRC && s_part == 1) {
                R.val[0] = P.r[MPM_R_VOLUME * cap + j];
                R.val[1] = th[MPM_TH_HEAT_SOURCE * cap + j];
            }
            if (SRC && s_part == 3) {
                R.val[0] = P.r[MPM_R_VOLUME * cap + j];
                for (int d = 0; d < 3; ++d) R.val[1 + d] = th[(MPM_TH_HEAT_FLUX + d) * cap + j];
            }
        }
        return R;
    };
    int ry = 0, base = -4;
    advance(ry, base);
    double a4[4][kThNC];
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
        for (int c = 0; c < kThNC; ++c) a4[k][c] = 0.0;
    Raw cur;
    if (ry < kTileCells) cur = load(fetch_j(ry, base));
    while (ry < kTileCells) {
        int nry = ry, nbase = base;
        advance(nry, nbase);
        Raw nxt;
        if (nry < kTileCells) nxt = load(fetch_j(nry, nbase));
        // ---- staging: this round's 16 packets.  A particle with phase != 0
        // deposits nothing; the phase comes from the slot's part-0 lane.
        {
            const int ph0 = __shfl(cur.phase, s_slot, 64);
            const bool act = cur.j >= 0 && ph0 == 0;
            const int ic = (s_part == 0) ? T0[0] + s_cell
                                         : ((s_part == 1) ? T0[1] + ry : T0[2] + cz);
            const bool out =
                s_part < 3 && act &&
                (int)floor((cur.xd - pick3(g.plo, s_part)) * pick3(g.dxi, s_part)) != ic;
            const unsigned long long ob = __ballot(out);
            const bool slot_out = ((ob >> s_slot) | (ob >> (s_slot + 16)) |
                                   (ob >> (s_slot + 32))) & 1ull;
            bool use = act;
            if (slot_out) {
                if (FALLBACK) {
                    if (s_part == 3 && act) {
                        const int k = atomicAdd(&S.ocount[S.ogen & 1], 1);
                        S.rank[k] = cur.j;
                    }
                } else {
                    *err = 1;
                }
                use = false;
            }
            double *pw = pkw + th_pk_base(s_slot);
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
            // packet values with the arithmetic of the reference's deposit
            // terms: m c_p, m c_p T, m k (part 0); V Q (part 1); V q (part 3)
            if (MT && s_part == 0) {
                const double mcp = cur.val[0] * cur.val[1];
                pw[kThPkV + 0] = use ? mcp : 0.0;
                pw[kThPkV + 1] = use ? mcp * cur.val[2] : 0.0;
                pw[kThPkV + 2] = use ? cur.val[0] * cur.val[3] : 0.0;
            }
            if (SRC && s_part == 1) pw[kThPkV + 3] = use ? cur.val[0] * cur.val[1] : 0.0;
            if (SRC && s_part == 3)
                for (int d = 0; d < 3; ++d)
                    pw[kThPkV + 4 + d] = use ? cur.val[0] * cur.val[1 + d] : 0.0;
        }
        // The staging area is private to the wavefront and LDS operations of
        // one wavefront execute in issue order: the compiler only has to keep
        // the packet writes ahead of the reads below.
        asm volatile("" ::: "memory");
        // ---- accumulation: particle base + tt of each of the row's cells
        const int nt = min(4, mc_of(ry) - base);
        for (int tt = 0; tt < nt; ++tt) {
#pragma clang fp contract(fast)
            const double *pr = pkw + th_pk_base(q * 4 + tt);
            double wx[4], dwx[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                wx[k] = pr[kPkFac + 2 * k];
                dwx[k] = pr[kPkFac + 2 * k + 1];
            }
            const double wy = pr[kPkFac + 8 + 2 * ky], dwy = pr[kPkFac + 8 + 2 * ky + 1];
            const double wz = pr[kPkFac + 16 + 2 * kz], dwz = pr[kPkFac + 16 + 2 * kz + 1];
            const double wyz = wy * wz;
            if (MT) {
                double Mr[3];
                for (int c = 0; c < 3; ++c) Mr[c] = pr[kThPkV + c] * wyz;
#pragma unroll
                for (int k = 0; k < 4; ++k)
                    for (int c = 0; c < 3; ++c) a4[k][c] += wx[k] * Mr[c];
            }
            if (SRC) {
                const double dwyz = dwy * wz, wydz = wy * dwz;
                const double C = pr[kThPkV + 3] * wyz + pr[kThPkV + 5] * dwyz +
                                 pr[kThPkV + 6] * wydz;
                const double E = pr[kThPkV + 4] * wyz;
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    // two fused multiply-adds into the accumulator
                    a4[k][3] += wx[k] * C;
                    a4[k][3] += dwx[k] * E;
                }
            }
        }
        asm volatile("" ::: "memory");  // reads of this round before the next writes
        if (nry != ry) {
            // Row done: lane (q, ky, kz) adds its four nodes (x = q + k,
            // y = ry + ky, z = cz + kz) into the tile's block.
#pragma unroll
            for (int k = 0; k < 4; ++k) {
#pragma unroll
                for (int c = 0; c < kThNC; ++c) {
                    if ((c < 3 && !MT) || (c == 3 && !SRC)) continue;
                    if (DET) {
                        // this wavefront's own block (z-level kz); lanes q and
                        // q + 1 share nodes, so still atomic adds, in an order
                        // the code fixes
                        const int idx = (kz * kTN + (ry + ky)) * kTN + (q + k);
                        __hip_atomic_fetch_add(&L.slab[wave * kThNC * kZN + c * kZN + idx],
                                               a4[k][c], __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_WAVEFRONT);
                    } else {
                        const int idx = ((cz + kz) * kTN + (ry + ky)) * kTN + (q + k);
                        __hip_atomic_fetch_add(&L.slab[c * kTNN + idx], a4[k][c],
                                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    }
                    a4[k][c] = 0.0;
                }
            }
        }
        ry = nry;
        base = nbase;
        cur = nxt;
    }
    __syncthreads();
    if (DET) {
        // The wavefront blocks summed in wavefront order, the whole node
        // block stored at list position pos and stamped with the pass (the
        // gather pass adds it to the nodes).
        double *blk = S.blk + (int64_t)pos * kThNC * kTNN;
        if (tid == 0) S.tile_slot[tile] = ((int64_t)S.pass << 32) | (uint32_t)pos;
        for (int e = tid; e < kTNN; e += kBlock) {
            const int ax = e % kTN, ay = (e / kTN) % kTN, az = e / (kTN * kTN);
            const int w0 = az > 3 ? az - 3 : 0, w1 = az < 3 ? az : 3;
            for (int c = 0; c < kThNC; ++c) {
                double v = 0.0;
                for (int w = w0; w <= w1; ++w)
                    v += L.slab[w * kThNC * kZN + c * kZN + ((az - w) * kTN + ay) * kTN + ax];
                blk[c * kTNN + e] = v;
            }
        }
        return;
    }
    // Flush the node block: one global atomic per non-zero node component,
    // periodic image targets as in deposit_global.
    const int64_t nn = g.nn;
    const int tnc[kThNC] = {TN_MSP, TN_MSPT, TN_MC, TN_SRC};
    for (int e = tid; e < kTNN; e += kBlock) {
        const int ax = e % kTN, ay = (e / kTN) % kTN, az = e / (kTN * kTN);
        double v[kThNC];
        bool any = false;
        for (int c = 0; c < kThNC; ++c) {
            const bool on = c < 3 ? MT : SRC;
            v[c] = on ? L.slab[c * kTNN + e] : 0.0;
            any |= (v[c] != 0.0);
        }
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
                    const int64_t node =
                        nidx(g, qx ? tx1 : tx0, qy ? ty1 : ty0, qz ? tz1 : tz0);
                    for (int c = 0; c < kThNC; ++c)
                        if (v[c] != 0.0) unsafeAtomicAdd(&tn[tnc[c] * nn + node], v[c]);
                }
    }
}

// The deterministic deposit's gather pass: node (I, J, K) adds its tile
// blocks (blk_gather, a fixed order) to the thermal components, which hold
// only the deferred fallback's deposits.
template <bool MT, bool SRC>
__global__ __launch_bounds__(kThreads) void k_th_gather(GridP g, BlkGather bg, double *tn) {
    const int64_t node = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (node >= g.nn) return;
    const int I = (int)(node % (g.n[0] + 1));
    const int J = (int)((node / (g.n[0] + 1)) % (g.n[1] + 1));
    const int K = (int)(node / ((int64_t)(g.n[0] + 1) * (g.n[1] + 1)));
    double s[kThNC];
    if (!blk_gather<kThNC>(g, bg, I, J, K, s)) return;
    const int64_t nn = g.nn;
    if (MT) {
        tn[TN_MSP * nn + node] += s[0];
        tn[TN_MSPT * nn + node] += s[1];
        tn[TN_MC * nn + node] += s[2];
    }
    if (SRC) tn[TN_SRC * nn + node] += s[3];
}

// The deposits of the particles a FALLBACK k_th_p2g_cells launch listed, one
// thread per entry.
template <bool MT, bool SRC>
__global__ __launch_bounds__(kBlock) void k_th_deposit_outside(GridP g, PState P, SortState S,
                                                               const double *th, double *tn) {
    // the count as the tile launch's atomics left it (a vector load)
    const int n = __hip_atomic_load(&S.ocount[S.ogen & 1], __ATOMIC_RELAXED,
                                    __HIP_MEMORY_SCOPE_AGENT);
    for (int k = blockIdx.x * kBlock + threadIdx.x; k < n; k += gridDim.x * kBlock)
        th_deposit_one<MT, SRC>(g, P, th, tn, S.rank[k]);
}

// FabArray::SumBoundary over periodic direction d (AMReX, called at
// mpm_particle_grid_ops.cpp:848-853): the two face nodes of each periodic
// pair both take their sum.  One thread per pair; directions in turn.
__global__ __launch_bounds__(kThreads) void k_th_sum_boundary(GridP g, double *tn, int d,
                                                              int mass_temp, int source) {
    const int a = d == 0 ? 1 : 0, b = d == 2 ? 1 : 2;
    const int64_t na = g.n[a] + 1, nb = g.n[b] + 1;
    const int64_t t = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (t >= na * nb) return;
    int X[3];
    X[a] = (int)(t % na);
    X[b] = (int)(t / na);
    X[d] = 0;
    const int64_t lo = nidx(g, X[0], X[1], X[2]);
    X[d] = g.n[d];
    const int64_t hi = nidx(g, X[0], X[1], X[2]);
    const int64_t nn = g.nn;
    const int comps[4] = {TN_MSP, TN_MSPT, TN_MC, TN_SRC};
    for (int c = 0; c < 4; ++c) {
        if (c < 3 ? !mass_temp : !source) continue;
        double *v = tn + comps[c] * nn;
        const double s = v[lo] + v[hi];
        v[lo] = s;
        v[hi] = s;
    }
}

// The nodal passes between the deposit and the wall conditions, fused per
// node: the deposit's normalisation (mpm_particle_grid_ops.cpp:857-881; a
// node without thermal mass keeps its temperature), backup_current_temperature
// (nodal_data_ops.cpp:95-114) and, in a time step,
// Nodal_Time_Update_Temperature (nodal_data_ops.cpp:500-527).
template <bool UPDATE>
__global__ __launch_bounds__(kThreads) void k_th_nodal(GridP g, double tol, double *tn,
                                                       const Scalars *sc) {
    const int64_t q = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    const int64_t nn = g.nn;
    if (q >= nn) return;
    const double msp = tn[TN_MSP * nn + q];
    double T = tn[TN_T * nn + q];
    if (msp > 0.0) {
        T = (msp >= tol) ? tn[TN_MSPT * nn + q] / msp : 0.0;
        tn[TN_DT * nn + q] = T;
    }
    if (UPDATE) {
        if (msp >= tol) T += tn[TN_SRC * nn + q] / msp * sc->dt;
        else T = 0.0;
    }
    tn[TN_T * nn + q] = T;
}

__device__ __forceinline__ bool on_face(const GridP &g, int I, int J, int K) {
    return I == 0 || I == g.n[0] || J == 0 || J == g.n[1] || K == 0 || K == g.n[2];
}

// nodal_bcs_temperature (nodal_data_ops.cpp:786-910): the value a wall node
// ends with when a serial ParallelFor sweeps the nodal box (k outermost, i
// fastest) updating the temperature in place.  A node reads its inward
// neighbour's value: the updated one when the neighbour comes earlier in the
// sweep (a hi-face neighbour that is itself a wall node), else the value
// before the pass.  An updated neighbour lies on fewer hi faces than the node
// (its index in that direction is n - 1), so the recursion is at most three
// deep.  T holds the values before the pass.
template <int D>
__device__ double th_wall(const GridP &g, const mpm_thermal_specs &t, const double *T,
                          const double *MC, const double *M, int I, int J, int K) {
    const int X[3] = {I, J, K};
    const int64_t q = nidx(g, I, J, K);
    double v = T[q];
    bool applied = false;
    for (int d = 0; d < 3; ++d) {
        const bool lo = X[d] == 0, hi = X[d] == g.n[d];
        if (!lo && !hi) continue;
        const int bc = lo ? t.bclo[d] : t.bchi[d];
        if (bc == MPM_BC_TEMP_PERIODIC) {
            applied = true;
            continue;
        }
        if (bc == MPM_BC_TEMP_ISOTHERMAL) {
            v = lo ? t.T_wall_lo[d] : t.T_wall_hi[d];
            applied = true;
            continue;
        }
        if (bc < MPM_BC_TEMP_ADIABATIC || bc > MPM_BC_TEMP_CONVECTION) continue;
        if (bc == MPM_BC_TEMP_ADIABATIC && applied) continue;
        int Y[3] = {I, J, K};
        Y[d] += lo ? 1 : -1;
        double tnb = T[nidx(g, Y[0], Y[1], Y[2])];
        if constexpr (D > 0) {
            if (hi && on_face(g, Y[0], Y[1], Y[2]))
                tnb = th_wall<D - 1>(g, t, T, MC, M, Y[0], Y[1], Y[2]);
        }
        if (bc == MPM_BC_TEMP_ADIABATIC) {
            v = tnb;
        } else {
            const double mk = MC[q], m = M[q];
            const double k_node = (m > 0.0) ? mk / m : 1.0;
            if (bc == MPM_BC_TEMP_USERFLUX) {
                const double fl = lo ? t.flux_lo[d] : t.flux_hi[d];
                v = tnb + fl * g.dx[d] / k_node;
            } else {
                const double hc = lo ? t.h_lo[d] : t.h_hi[d];
                const double tinf = lo ? t.T_inf_lo[d] : t.T_inf_hi[d];
                const double Bi = hc * g.dx[d] / k_node;
                v = (tnb + Bi * tinf) / (1.0 + Bi);
            }
        }
        applied = true;
    }
    return v;
}

// Device face tables of the temperature wall functions (null: no function).
struct ThUdf {
    const double *tab[3][2];
};

__device__ __forceinline__ bool on_udf_face(const GridP &g, const ThUdf &u, int I, int J, int K) {
    const int X[3] = {I, J, K};
    for (int d = 0; d < 3; ++d)
        if ((X[d] == 0 && u.tab[d][0]) || (X[d] == g.n[d] && u.tab[d][1])) return true;
    return false;
}

// apply_udf_nodal_bcs_temperature (nodal_data_ops.cpp:1308-1446): the value a
// node on a face with a function ends with after the second serial sweep,
// which applies the face's temperature BC type with the function's values
// (no precedence between faces: the last direction wins).  Neighbours as in
// th_wall; T holds the values after the wall conditions.
template <int D>
__device__ double th_udf(const GridP &g, const mpm_thermal_specs &t, const ThUdf &u,
                         const double *T, const double *MC, const double *M, int I, int J,
                         int K) {
    const int X[3] = {I, J, K};
    const int64_t q = nidx(g, I, J, K);
    double v = T[q];
    for (int dir = 0; dir < 3; ++dir) {
        const double *tab = nullptr;
        int bc = -1;
        bool lo = false;
        if (X[dir] == 0 && u.tab[dir][0]) {
            tab = u.tab[dir][0];
            bc = t.bclo[dir];
            lo = true;
        } else if (X[dir] == g.n[dir] && u.tab[dir][1]) {
            tab = u.tab[dir][1];
            bc = t.bchi[dir];
        }
        if (!tab) continue;
        const int p0 = (dir == 0) ? 1 : 0;
        const int p1 = (dir == 2) ? 1 : 2;
        const int64_t flat = ((int64_t)X[p0] * (g.n[p1] + 1) + X[p1]) * 2;
        const double v0 = tab[flat], v1 = tab[flat + 1];
        if (bc == MPM_BC_TEMP_ISOTHERMAL) {
            v = v0;
            continue;
        }
        if (bc != MPM_BC_TEMP_USERFLUX && bc != MPM_BC_TEMP_CONVECTION) continue;
        int Y[3] = {I, J, K};
        Y[dir] += lo ? 1 : -1;
        double tnb = T[nidx(g, Y[0], Y[1], Y[2])];
        if constexpr (D > 0) {
            if (!lo && on_udf_face(g, u, Y[0], Y[1], Y[2]))
                tnb = th_udf<D - 1>(g, t, u, T, MC, M, Y[0], Y[1], Y[2]);
        }
        const double mk = MC[q], m = M[q];
        const double k_node = (m > 0.0) ? mk / m : 1.0;
        if (bc == MPM_BC_TEMP_USERFLUX) {
            v = tnb + v0 * g.dx[dir] / k_node;
        } else {
            const double Bi = v0 * g.dx[dir] / k_node;
            v = (tnb + Bi * v1) / (1.0 + Bi);
        }
    }
    return v;
}

// One level-set body's temperature condition.
struct ThBody {
    const double *phi;
    int bc;
    double T_wall, q, h, T_inf;
};

// The dominant-direction neighbour of a node that the body's condition
// changes, or false when it changes nothing there.
__device__ inline bool th_ls_active(const GridP &g, const ThBody &b, const double *MSP, int I,
                                    int J, int K, int Y[3], int &dom) {
    const int64_t q = nidx(g, I, J, K);
    const double x[3] = {g.plo[0] + I * g.dx[0], g.plo[1] + J * g.dx[1], g.plo[2] + K * g.dx[2]};
    if (ls_value(g, b.phi, x) >= 0.0 || MSP[q] <= 0.0) return false;
    double nrm[3] = {1.0, 0.0, 0.0};
    ls_grad(g, b.phi, x, nrm);
    const double gm = sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
    if (gm < 1.0e-10) return false;
    for (int d = 0; d < 3; ++d) nrm[d] /= gm;
    dom = 0;
    for (int d = 1; d < 3; ++d)
        if (fabs(nrm[d]) > fabs(nrm[dom])) dom = d;
    Y[0] = I;
    Y[1] = J;
    Y[2] = K;
    Y[dom] += (nrm[dom] > 0.0) ? 1 : -1;
    if (b.bc == MPM_LS_TEMP_ISOTHERMAL) return true;
    return Y[dom] >= -1 && Y[dom] <= g.n[dom] + 1;
}

// nodal_levelset_bcs_temperature (nodal_data_ops.cpp:235-350): the value a
// node ends with after one body's serial in-place sweep.  The neighbour read
// is the updated one when it comes earlier in the sweep (the normal's
// dominant component is negative); a ghost neighbour holds zero in this
// pass (the reset of the step, no fill since).  Chains deeper than D set
// bit 2 of *err.
template <int D>
__device__ double th_ls(const GridP &g, const ThBody &b, const double *T, const double *MSP,
                        const double *MC, const double *M, int I, int J, int K, int *err) {
    const int64_t q = nidx(g, I, J, K);
    int Y[3], dom = 0;
    if (!th_ls_active(g, b, MSP, I, J, K, Y, dom)) return T[q];
    if (b.bc == MPM_LS_TEMP_ISOTHERMAL) return b.T_wall;
    double tnb = 0.0;
    if (Y[dom] >= 0 && Y[dom] <= g.n[dom]) {
        tnb = T[nidx(g, Y[0], Y[1], Y[2])];
        const bool earlier = Y[dom] < (dom == 0 ? I : (dom == 1 ? J : K));
        if (earlier) {
            if constexpr (D > 0) {
                tnb = th_ls<D - 1>(g, b, T, MSP, MC, M, Y[0], Y[1], Y[2], err);
            } else {
                int Z[3], dz = 0;
                if (th_ls_active(g, b, MSP, Y[0], Y[1], Y[2], Z, dz)) atomicOr(err, 2);
            }
        }
    }
    const double mk = MC[q], m = M[q];
    const double k_node = (m > 0.0) ? mk / m : 1.0;
    if (b.bc == MPM_LS_TEMP_HEATFLUX) return tnb + b.q * g.dx[dom] / k_node;
    if (b.bc == MPM_LS_TEMP_CONVECTIVE) {
        const double Bi = b.h * g.dx[dom] / k_node;
        return (tnb + Bi * b.T_inf) / (1.0 + Bi);
    }
    return T[q];
}

__device__ __forceinline__ void node_ijk(const GridP &g, int64_t q, int &I, int &J, int &K) {
    const int64_t n0 = g.n[0] + 1, n1 = g.n[1] + 1;
    I = (int)(q % n0);
    J = (int)((q / n0) % n1);
    K = (int)(q / (n0 * n1));
}

// Wall values into scr (wall nodes only); the temperatures stay untouched
// until every wall value is known.
__global__ __launch_bounds__(kThreads) void k_th_walls(GridP g, mpm_thermal_specs t,
                                                       const double *tn, const double *mass,
                                                       double *scr) {
    const int64_t q = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (q >= g.nn) return;
    int I, J, K;
    node_ijk(g, q, I, J, K);
    if (!on_face(g, I, J, K)) return;
    const int64_t nn = g.nn;
    scr[q] = th_wall<3>(g, t, tn + TN_T * nn, tn + TN_MC * nn, mass, I, J, K);
}

// Function values of the nodes on faces with a temperature function into scr.
__global__ __launch_bounds__(kThreads) void k_th_udf_walls(GridP g, mpm_thermal_specs t, ThUdf u,
                                                           const double *tn, const double *mass,
                                                           double *scr) {
    const int64_t q = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (q >= g.nn) return;
    int I, J, K;
    node_ijk(g, q, I, J, K);
    if (!on_udf_face(g, u, I, J, K)) return;
    const int64_t nn = g.nn;
    scr[q] = th_udf<3>(g, t, u, tn + TN_T * nn, tn + TN_MC * nn, mass, I, J, K);
}

// One body's values into scr, every node.
__global__ __launch_bounds__(kThreads) void k_th_ls_walls(GridP g, ThBody b, const double *tn,
                                                          const double *mass, double *scr,
                                                          int *err) {
    const int64_t q = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (q >= g.nn) return;
    int I, J, K;
    node_ijk(g, q, I, J, K);
    const int64_t nn = g.nn;
    scr[q] = th_ls<4>(g, b, tn + TN_T * nn, tn + TN_MSP * nn, tn + TN_MC * nn, mass, I, J, K,
                      err);
}

// The values of scr into the temperature: mode 0 the wall nodes, 1 the nodes
// on faces with a function, 2 every node.
__global__ __launch_bounds__(kThreads) void k_th_apply_walls(GridP g, ThUdf u, int mode,
                                                             double *tn, const double *scr) {
    const int64_t q = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    const int64_t nn = g.nn;
    if (q >= nn) return;
    int I, J, K;
    node_ijk(g, q, I, J, K);
    const bool take = mode == 2 || (mode == 1 ? on_udf_face(g, u, I, J, K) : on_face(g, I, J, K));
    if (take) tn[TN_T * nn + q] = scr[q];
}

// store_delta_temperature (nodal_data_ops.cpp:409-430).
__global__ __launch_bounds__(kThreads) void k_th_delta(GridP g, double *tn) {
    const int64_t q = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    const int64_t nn = g.nn;
    if (q >= nn) return;
    if (tn[TN_MSP * nn + q] > 0.0) tn[TN_DT * nn + q] = tn[TN_T * nn + q] - tn[TN_DT * nn + q];
}

// interpolate_from_grid_temperature (mpm_particle_grid_ops.cpp:1129-1270):
// temperature += the nodal change by bilin_interp / quadratic_interp /
// cubic_interp (interpolants.H:513-565, 680-862, 1029-1210, chosen by the
// x-direction order for all directions); heat flux = -k grad T.  Thread i
// reads slot perm[i] (the sorted step) or i, and writes out slot i.
template <bool UT, bool UQ>
__global__ __launch_bounds__(kThreads) void k_th_g2p(GridP g, PState P, const int *perm,
                                                     const int *count, const double *thin,
                                                     double *thout, const double *tn) {
    const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    const int64_t n = count ? (int64_t)*count : P.np;
    if (i >= n) return;
    const int64_t j = perm ? (int64_t)perm[i] : i;
    const int64_t cap = P.cap, nn = g.nn;
    double f[MPM_NTHERMAL];
    for (int c = 0; c < MPM_NTHERMAL; ++c) f[c] = thin[c * cap + j];
    if (P.id[j] >= 0 && P.idata[MPM_I_PHASE * cap + j] == 0) {
        double x[3];
        int iv[3], mn[3], mx[3];
        for (int d = 0; d < 3; ++d) {
            x[d] = P.pos[d * cap + j];
            iv[d] = (int)floor((x[d] - g.plo[d]) * g.dxi[d]);
            bounds(iv[d], 0, g.hi[d], g.od[d], g.per[d], mn[d], mx[d]);
        }
        if (UT) {
            const double *DT = tn + TN_DT * nn;
            const int O = g.od[0];
            double value;
            if (O == 1) {
                const int ii = iv[0], jj = iv[1], kk = iv[2];
                const double a = (x[0] - (g.plo[0] + ii * g.dx[0])) / g.dx[0];
                const double b = (x[1] - (g.plo[1] + jj * g.dx[1])) / g.dx[1];
                const double c = (x[2] - (g.plo[2] + kk * g.dx[2])) / g.dx[2];
                value = (1.0 - a) * (1.0 - b) * (1.0 - c) * nodal_read(g, DT, ii, jj, kk);
                value += a * (1.0 - b) * (1.0 - c) * nodal_read(g, DT, ii + 1, jj, kk);
                value += (1.0 - a) * b * (1.0 - c) * nodal_read(g, DT, ii, jj + 1, kk);
                value += a * b * (1.0 - c) * nodal_read(g, DT, ii + 1, jj + 1, kk);
                value += (1.0 - a) * (1.0 - b) * c * nodal_read(g, DT, ii, jj, kk + 1);
                value += a * (1.0 - b) * c * nodal_read(g, DT, ii + 1, jj, kk + 1);
                value += (1.0 - a) * b * c * nodal_read(g, DT, ii, jj + 1, kk + 1);
                value += a * b * c * nodal_read(g, DT, ii + 1, jj + 1, kk + 1);
            } else {
                double wv[3][4];
                for (int d = 0; d < 3; ++d)
                    for (int l = mn[d]; l < mx[d]; ++l)
                        wv[d][l - mn[d]] =
                            interp1d(O, g.per[d], iv[d] + l, 0, g.hi[d], x[d], g.plo[d], g.dx[d]);
                value = 0.0;
                for (int nz = mn[2]; nz < mx[2]; ++nz)
                    for (int ny = mn[1]; ny < mx[1]; ++ny)
                        for (int nx = mn[0]; nx < mx[0]; ++nx)
                            value += wv[0][nx - mn[0]] * wv[1][ny - mn[1]] * wv[2][nz - mn[2]] *
                                     nodal_read(g, DT, iv[0] + nx, iv[1] + ny, iv[2] + nz);
            }
            f[MPM_TH_TEMPERATURE] += value;
        }
        if (UQ) {
            const double *T = tn + TN_T * nn;
            double bv[3][4], bd[3][4];
            for (int d = 0; d < 3; ++d)
                for (int l = mn[d]; l < mx[d]; ++l)
                    basis1d(g.od[d], g.per[d], iv[d], l, 0, g.hi[d], x[d], g.plo[d], g.dx[d],
                            bv[d][l - mn[d]], bd[d][l - mn[d]]);
            double gT[3] = {0.0, 0.0, 0.0};
            for (int nz = mn[2]; nz < mx[2]; ++nz)
                for (int ny = mn[1]; ny < mx[1]; ++ny)
                    for (int nx = mn[0]; nx < mx[0]; ++nx) {
                        const double vx = bv[0][nx - mn[0]], vy = bv[1][ny - mn[1]],
                                     vz = bv[2][nz - mn[2]];
                        const double gr[3] = {bd[0][nx - mn[0]] * vy * vz,
                                              vx * bd[1][ny - mn[1]] * vz,
                                              vx * vy * bd[2][nz - mn[2]]};
                        const double Tn = nodal_read(g, T, iv[0] + nx, iv[1] + ny, iv[2] + nz);
                        for (int d = 0; d < 3; ++d) gT[d] += Tn * gr[d];
                    }
            for (int d = 0; d < 3; ++d)
                f[MPM_TH_HEAT_FLUX + d] = -f[MPM_TH_CONDUCTIVITY] * gT[d];
        }
    }
    for (int c = 0; c < MPM_NTHERMAL; ++c) thout[c * cap + i] = f[c];
}

// The thermal deposit into tn (the components the flags name, on top of what
// they hold), then SumBoundary.  tiled: the particle state was sorted this
// step (perm, cell starts and tile list valid), so the deposit goes through
// the tiled momentum deposit (fallback: particles may lie outside their
// tile); otherwise one thread per particle with global atomics.
// interpolate_from_grid_temperature over the cell-sorted state, one
// workgroup per non-empty tile (uniform order): the tile's node block of T
// and delta T is staged in LDS, each thread takes particles of the tile and
// gathers both with the tile factors of the mechanical G2P.  A particle
// outside its tile (not redistributed this step) reads the nodes from global
// memory.  Sorted position i
// reads slot perm[i] and writes slot i (sorted) or perm[i] (in place).
template <int O, bool UT, bool UQ>
__global__ __launch_bounds__(kBlock) void k_th_g2p_tile(GridP g, PState P, SortState S,
                                                        const double *thin, double *thout,
                                                        int sorted_out, const double *tn) {
    __shared__ double nodes[kTNN * 2];
    __shared__ SplineTables tab;
    const int tile = active_tile(S, blockIdx.x);
    if (tile < 0) return;  // uniform over the block
    const int beg = S.cell_start[tile * kCellsPerTile];
    const int end = S.cell_start[tile * kCellsPerTile + kCellsPerTile];
    if (beg == end) return;
    const int tx = tile % g.ntc[0];
    const int ty = (tile / g.ntc[0]) % g.ntc[1];
    const int tz = tile / (g.ntc[0] * g.ntc[1]);
    const int T0[3] = {tx * kTileCells, ty * kTileCells, tz * kTileCells};
    const int64_t nn = g.nn, cap = P.cap;
    const double *Tg = tn + TN_T * nn, *DTg = tn + TN_DT * nn;
    for (int e = threadIdx.x; e < kTNN; e += kBlock) {
        const int a[3] = {e % kTN, (e / kTN) % kTN, e / (kTN * kTN)};
        bool ok = true;
        int I[3];
        for (int d = 0; d < 3; ++d) I[d] = wrap_read(T0[d] - 1 + a[d], g.n[d], g.per[d], ok);
        const int64_t q = nidx(g, I[0], I[1], I[2]);
        nodes[2 * e] = ok ? Tg[q] : 0.0;
        nodes[2 * e + 1] = ok ? DTg[q] : 0.0;
    }
    if (O > 1) {
        const int *src = reinterpret_cast<const int *>(&kTab);
        int *dst = reinterpret_cast<int *>(&tab);
        for (int e = threadIdx.x; e < (int)(sizeof(SplineTables) / sizeof(int)); e += kBlock)
            dst[e] = src[e];
    }
    __syncthreads();
    for (int i = beg + (int)threadIdx.x; i < end; i += kBlock) {
        const int64_t j = S.perm[i];
        double f[MPM_NTHERMAL];
        for (int c = 0; c < MPM_NTHERMAL; ++c) f[c] = thin[c * cap + j];
        if (P.id[j] >= 0 && P.idata[MPM_I_PHASE * cap + j] == 0) {
            double x[3];
            int iv[3], mn[3], mx[3];
            bool inside = true;
            for (int d = 0; d < 3; ++d) {
                x[d] = P.pos[d * cap + j];
                iv[d] = (int)floor((x[d] - g.plo[d]) * g.dxi[d]);
                bounds(iv[d], 0, g.hi[d], O, g.per[d], mn[d], mx[d]);
                inside &= iv[d] >= T0[d] && iv[d] < T0[d] + kTileCells;
            }
            // The mechanical G2P's factors (tile_factors: interpolation
            // weight w, basis b and its derivative d of each offset, zero
            // outside compute_bounds) and one pass over the stencil for both
            // sums (z, then y, then x).  The nodes come
            // from the tile's block in LDS, or from global memory for a
            // particle outside its tile (not redistributed this step).
            constexpr int W = (O == 1) ? 2 : 4;
            constexpr int L0 = (O == 1) ? 0 : -1;
            const SplineTables &TT = (O > 1) ? tab : kTab;
            double value = 0.0, gT[3] = {0.0, 0.0, 0.0};
            if (inside) {
                double fw[2][W], fb[2][W], fd[2][W];
#pragma unroll
                for (int d = 0; d < 2; ++d)
#pragma unroll
                    for (int k = 0; k < W; ++k) {
                        const int l = L0 + k;
                        tile_factors<O>(TT, g.per[d], iv[d], l, g.hi[d], x[d], g.plo[d], g.dx[d],
                                        g.dxi[d], l >= mn[d] && l < mx[d], fw[d][k], fb[d][k],
                                        fd[d][k]);
                    }
                const double *blk =
                    nodes + 2 * (((iv[2] - T0[2] + L0 + 1) * kTN + (iv[1] - T0[1] + L0 + 1)) * kTN +
                                 (iv[0] - T0[0] + L0 + 1));
#pragma unroll 1
                for (int kz = 0; kz < W; ++kz) {
                    const int lz = L0 + kz;
                    double wz, bz, dz;
                    tile_factors<O>(TT, g.per[2], iv[2], lz, g.hi[2], x[2], g.plo[2], g.dx[2],
                                    g.dxi[2], lz >= mn[2] && lz < mx[2], wz, bz, dz);
#pragma unroll
                    for (int ky = 0; ky < W; ++ky) {
                        // y and z factors first: products of x and y factors
                        // alone would be loop invariant and held in registers
                        // across the z loop.
                        const double wyz = fw[1][ky] * wz;
                        const double p0 = fb[1][ky] * bz, p1 = fd[1][ky] * bz,
                                     p2 = fb[1][ky] * dz;
#pragma unroll
                        for (int kx = 0; kx < W; ++kx) {
                            const double *q = blk + 2 * ((kz * kTN + ky) * kTN + kx);
                            if (UT) value += fw[0][kx] * wyz * q[1];
                            if (UQ) {
                                const double gr[3] = {fd[0][kx] * p0, fb[0][kx] * p1,
                                                      fb[0][kx] * p2};
                                for (int d = 0; d < 3; ++d) gT[d] += q[0] * gr[d];
                            }
                        }
                    }
                }
            } else {
                // Outside the tile (rare): the same factors and sums, node
                // by node, from global memory.
#pragma unroll 1
                for (int lz = mn[2]; lz < mx[2]; ++lz) {
                    double wz, bz, dz;
                    tile_factors<O>(TT, g.per[2], iv[2], lz, g.hi[2], x[2], g.plo[2], g.dx[2],
                                    g.dxi[2], true, wz, bz, dz);
#pragma unroll 1
                    for (int ly = mn[1]; ly < mx[1]; ++ly) {
                        double wy, by, dy;
                        tile_factors<O>(TT, g.per[1], iv[1], ly, g.hi[1], x[1], g.plo[1],
                                        g.dx[1], g.dxi[1], true, wy, by, dy);
                        const double wyz = wy * wz;
                        const double p0 = by * bz, p1 = dy * bz, p2 = by * dz;
#pragma unroll 1
                        for (int lx = mn[0]; lx < mx[0]; ++lx) {
                            double wx, bx, dx;
                            tile_factors<O>(TT, g.per[0], iv[0], lx, g.hi[0], x[0], g.plo[0],
                                            g.dx[0], g.dxi[0], true, wx, bx, dx);
                            const int I = iv[0] + lx, J = iv[1] + ly, K = iv[2] + lz;
                            if (UT) value += wx * wyz * nodal_read(g, DTg, I, J, K);
                            if (UQ) {
                                const double Tn = nodal_read(g, Tg, I, J, K);
                                const double gr[3] = {dx * p0, bx * p1, bx * p2};
                                for (int d = 0; d < 3; ++d) gT[d] += Tn * gr[d];
                            }
                        }
                    }
                }
            }
            if (UT) f[MPM_TH_TEMPERATURE] += value;
            if (UQ)
                for (int d = 0; d < 3; ++d)
                    f[MPM_TH_HEAT_FLUX + d] = -f[MPM_TH_CONDUCTIVITY] * gT[d];
        }
        const int64_t o = sorted_out ? (int64_t)i : j;
        for (int c = 0; c < MPM_NTHERMAL; ++c) thout[c * cap + o] = f[c];
    }
}

void launch_th_p2g(mpm_ctx *c, bool mass_temp, bool source, bool tiled, int fallback,
                   hipStream_t st) {
    const GridP &g = c->g;
    const int64_t np = c->P.np, nn = g.nn;
    if (np > 0 && tiled) {
        const dim3 gr((unsigned)((c->S.ntiles + 7) / 8 * 8)), bl(kBlock);
        const int O = g.uniform ? g.od[0] : 0;
        int *err = &c->sc->err;
#define XMPM_THPD(o, mt, src, fb)                                                             \
    do {                                                                                      \
        if (det)                                                                              \
            k_th_p2g_cells<o, mt, src, fb, true><<<gr, bl, 0, st>>>(g, c->P, c->S, c->th,     \
                                                                     c->tn, err);             \
        else                                                                                  \
            k_th_p2g_cells<o, mt, src, fb, false><<<gr, bl, 0, st>>>(g, c->P, c->S, c->th,    \
                                                                      c->tn, err);            \
    } while (0)
#define XMPM_THP(o, mt, src)                                                                  \
    do {                                                                                      \
        if (fallback) XMPM_THPD(o, mt, src, true); else XMPM_THPD(o, mt, src, false);         \
    } while (0)
#define XMPM_THPO(o)                                                                          \
    do {                                                                                      \
        if (mass_temp && source) XMPM_THP(o, true, true);                                     \
        else if (mass_temp) XMPM_THP(o, true, false);                                         \
        else XMPM_THP(o, false, true);                                                        \
    } while (0)
        c->S.ocount = c->sc->ocount;
        // MPM_OPT_DETERMINISTIC: tile blocks and a gather pass in a fixed
        // order (the stamps of this pass)
        const bool det = c->S.blk != nullptr;
        if (det) ++c->S.pass;
        if (O == 0) XMPM_THPO(0);
        else if (O == 1) XMPM_THPO(1);
        else if (O == 2) XMPM_THPO(2);
        else XMPM_THPO(3);
#undef XMPM_THPO
#undef XMPM_THP
#undef XMPM_THPD
        if (fallback) {
            const dim3 go(64);
            if (mass_temp && source)
                k_th_deposit_outside<true, true><<<go, bl, 0, st>>>(g, c->P, c->S, c->th, c->tn);
            else if (mass_temp)
                k_th_deposit_outside<true, false><<<go, bl, 0, st>>>(g, c->P, c->S, c->th, c->tn);
            else
                k_th_deposit_outside<false, true><<<go, bl, 0, st>>>(g, c->P, c->S, c->th, c->tn);
            ++c->S.ogen;
        }
        if (det) {
            BlkGather bg;
            bg.blk = c->S.blk;
            bg.slot = c->S.tile_slot;
            bg.pass = c->S.pass;
            bg.nc = kThNC;
            if (mass_temp && source)
                k_th_gather<true, true><<<blocks(nn), kThreads, 0, st>>>(g, bg, c->tn);
            else if (mass_temp)
                k_th_gather<true, false><<<blocks(nn), kThreads, 0, st>>>(g, bg, c->tn);
            else
                k_th_gather<false, true><<<blocks(nn), kThreads, 0, st>>>(g, bg, c->tn);
        }
    } else if (np > 0) {
        if (mass_temp && source)
            k_th_p2g<true, true><<<blocks(np), kThreads, 0, st>>>(g, c->P, c->th, c->tn);
        else if (mass_temp)
            k_th_p2g<true, false><<<blocks(np), kThreads, 0, st>>>(g, c->P, c->th, c->tn);
        else if (source)
            k_th_p2g<false, true><<<blocks(np), kThreads, 0, st>>>(g, c->P, c->th, c->tn);
    }
    for (int d = 0; d < 3; ++d) {
        if (!g.per[d]) continue;
        const int a = d == 0 ? 1 : 0, b = d == 2 ? 1 : 2;
        const int64_t pairs = (int64_t)(g.n[a] + 1) * (g.n[b] + 1);
        k_th_sum_boundary<<<blocks(pairs), kThreads, 0, st>>>(g, c->tn, d, mass_temp ? 1 : 0,
                                                              source ? 1 : 0);
    }
}

enum { LS_NONE = 0, LS_ALL = 1, LS_DIRICHLET = 2 };

// Apply_Nodal_BCs_Temperature (utilities.cpp:221-240): the wall conditions,
// then (udf) the face functions, then the level-set bodies in order (ls:
// none, all, or the isothermal ones only); then, when delta is set,
// store_delta_temperature.
void launch_th_walls(mpm_ctx *c, bool delta, bool udf, int ls, hipStream_t st) {
    const int64_t nn = c->g.nn;
    ThUdf u{};
    const bool any = udf && c->any_tudf;
    for (int d = 0; d < 3; ++d)
        for (int h = 0; h < 2; ++h) u.tab[d][h] = (any && c->tudf_fn[d][h]) ? c->tudf_dev[d][h] : nullptr;
    k_th_walls<<<blocks(nn), kThreads, 0, st>>>(c->g, c->ts, c->tn, c->N.mass, c->tscr);
    k_th_apply_walls<<<blocks(nn), kThreads, 0, st>>>(c->g, u, 0, c->tn, c->tscr);
    if (any) {
        k_th_udf_walls<<<blocks(nn), kThreads, 0, st>>>(c->g, c->ts, u, c->tn, c->N.mass,
                                                        c->tscr);
        k_th_apply_walls<<<blocks(nn), kThreads, 0, st>>>(c->g, u, 1, c->tn, c->tscr);
    }
    for (int b = 0; ls != LS_NONE && b < c->ls.nb; ++b) {
        const int bc = c->ls_tbc[b];
        if (bc == MPM_LS_TEMP_ADIABATIC) continue;
        if (ls == LS_DIRICHLET && bc != MPM_LS_TEMP_ISOTHERMAL) continue;
        const ThBody body{c->ls.phi[b], bc, c->ls_tpar[b][0], c->ls_tpar[b][1], c->ls_tpar[b][2],
                          c->ls_tpar[b][3]};
        k_th_ls_walls<<<blocks(nn), kThreads, 0, st>>>(c->g, body, c->tn, c->N.mass, c->tscr,
                                                       &c->sc->err);
        k_th_apply_walls<<<blocks(nn), kThreads, 0, st>>>(c->g, u, 2, c->tn, c->tscr);
    }
    if (delta) k_th_delta<<<blocks(nn), kThreads, 0, st>>>(c->g, c->tn);
}

// Thermal G2P.  With perm (the state sorted this step) and a uniform order,
// the tiled kernel; sorted_out: write sorted positions (else in place).
void launch_th_g2p_tiles(mpm_ctx *c, bool ut, bool uq, bool sorted_out, double *out,
                         hipStream_t st) {
    const GridP &g = c->g;
    const dim3 gr((unsigned)((c->S.ntiles + 7) / 8 * 8)), bl(kBlock);
    const int so = sorted_out ? 1 : 0;
#define XMPM_TG(o)                                                                            \
    do {                                                                                      \
        if (ut && uq)                                                                         \
            k_th_g2p_tile<o, true, true><<<gr, bl, 0, st>>>(g, c->P, c->S, c->th, out, so,    \
                                                            c->tn);                           \
        else if (ut)                                                                          \
            k_th_g2p_tile<o, true, false><<<gr, bl, 0, st>>>(g, c->P, c->S, c->th, out, so,   \
                                                             c->tn);                          \
        else                                                                                  \
            k_th_g2p_tile<o, false, true><<<gr, bl, 0, st>>>(g, c->P, c->S, c->th, out, so,   \
                                                             c->tn);                          \
    } while (0)
    if (g.od[0] == 1) XMPM_TG(1);
    else if (g.od[0] == 2) XMPM_TG(2);
    else XMPM_TG(3);
#undef XMPM_TG
}

void launch_th_g2p(mpm_ctx *c, bool ut, bool uq, const int *perm, const int *count,
                   double *out, hipStream_t st) {
    const int64_t np = c->P.np;
    if (np <= 0) return;
    if (ut && uq)
        k_th_g2p<true, true><<<blocks(np), kThreads, 0, st>>>(c->g, c->P, perm, count, c->th, out,
                                                              c->tn);
    else if (ut)
        k_th_g2p<true, false><<<blocks(np), kThreads, 0, st>>>(c->g, c->P, perm, count, c->th,
                                                               out, c->tn);
    else
        k_th_g2p<false, true><<<blocks(np), kThreads, 0, st>>>(c->g, c->P, perm, count, c->th,
                                                               out, c->tn);
}

int check_specs(const mpm_thermal_specs *t) {
    for (int d = 0; d < 3; ++d)
        if (t->bclo[d] < MPM_BC_TEMP_PERIODIC || t->bclo[d] > MPM_BC_TEMP_CONVECTION ||
            t->bchi[d] < MPM_BC_TEMP_PERIODIC || t->bchi[d] > MPM_BC_TEMP_CONVECTION)
            return set_err(MPM_E_BC, "temperature BC code must be 0..4");
    return MPM_OK;
}

// Particle thermal buffers for the current capacity, zeroed.
int alloc_particle_thermal(mpm_ctx *c) {
    (void)hipFree(c->th);
    (void)hipFree(c->th_alt);
    c->th = c->th_alt = nullptr;
    const int64_t cap = c->P.cap > 0 ? c->P.cap : 1;
    const size_t bytes = (size_t)MPM_NTHERMAL * cap * sizeof(double);
    if (hipMalloc(&c->th, bytes) != hipSuccess || hipMalloc(&c->th_alt, bytes) != hipSuccess)
        return set_err(MPM_E_OOM, "thermal particle buffers");
    launch_zero(c->th, (int64_t)MPM_NTHERMAL * cap, c->stream);
    launch_zero(c->th_alt, (int64_t)MPM_NTHERMAL * cap, c->stream);
    c->th_cap = cap;
    return check_launch();
}

// Upload index of every particle slot (host copy).
int slot_ids(mpm_ctx *c, std::vector<int64_t> &ids) {
    ids.resize(c->P.np);
    HIPCHK(hipStreamSynchronize(c->stream));
    if (c->P.np > 0)
        HIPCHK(hipMemcpy(ids.data(), c->P.id, c->P.np * sizeof(int64_t), hipMemcpyDeviceToHost));
    return MPM_OK;
}

}  // namespace

namespace xmpm {

int thermal_on_upload(mpm_ctx *c) {
    if (!c->thermal) return MPM_OK;
    int rc = alloc_particle_thermal(c);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(c->stream));
    return MPM_OK;
}

// The thermal block of the initial pass (mpm_init.cpp:417-441) on freshly
// zeroed thermal nodal data (mpm_init.cpp:280-281): deposit (mass terms, no
// source), backup, walls, delta, G2P of temperature and heat flux.
int thermal_init(mpm_ctx *c, bool tiled) {
    if (!c->thermal) return MPM_OK;
    hipStream_t st = c->stream;
    const int64_t nn = c->g.nn;
    launch_zero(c->tn, TN_COMPS * nn, st);
    launch_th_p2g(c, true, false, tiled, 1, st);
    k_th_nodal<false><<<blocks(nn), kThreads, 0, st>>>(c->g, c->s.mass_tolerance, c->tn, nullptr);
    launch_th_walls(c, true, false, LS_NONE, st);
    launch_th_g2p(c, true, true, nullptr, nullptr, c->th, st);
    return check_launch();
}

// The grid half of a step's thermal stages (main.cpp:99-105), after the
// momentum grid pass has finished the nodal mass: Reset_Nodaldata_to_Zero's
// share of the thermal components, deposit, normalise + backup + update,
// walls, delta.
int thermal_grid(mpm_ctx *c, bool tiled, int fallback, hipStream_t st) {
    const int64_t nn = c->g.nn;
    launch_zero(c->tn, TN_COMPS * nn, st);
    launch_th_p2g(c, true, true, tiled, fallback, st);
    k_th_nodal<true><<<blocks(nn), kThreads, 0, st>>>(c->g, c->s.mass_tolerance, c->tn, c->sc);
    launch_th_walls(c, true, true, c->s.stress_update_scheme == MPM_MUSL ? LS_DIRICHLET : LS_ALL,
                    st);
    return check_launch();
}

// The particle half, before the particle kernel moves the particles:
// USL G2P of temperature and heat flux (main.cpp:112-114); MUSL G2P of
// temperature, the source-only re-deposit, the walls again and G2P of the
// heat flux (main.cpp:130-135).  With perm (the sorted step) the last G2P
// writes the second buffer by sorted position; thermal_swap follows the
// particle buffers' swap.
int thermal_particles(mpm_ctx *c, bool musl, const int *perm, const int *count,
                      int fallback, hipStream_t st) {
    double *out = perm ? c->th_alt : c->th;
    const bool tiles = perm != nullptr && c->g.uniform;
    if (!musl) {
        if (tiles) launch_th_g2p_tiles(c, true, true, true, out, st);
        else launch_th_g2p(c, true, true, perm, count, out, st);
        return check_launch();
    }
    const int64_t nn = c->g.nn;
    if (tiles) launch_th_g2p_tiles(c, true, false, false, c->th, st);
    else launch_th_g2p(c, true, false, nullptr, nullptr, c->th, st);
    // P2G_Temperature(reset = 1, mass_temp = 0, source = 1): the reset
    // clears the thermal mass, its temperature sum and the source.
    launch_zero(c->tn + TN_MSP * nn, 2 * nn, st);
    launch_zero(c->tn + TN_SRC * nn, nn, st);
    launch_th_p2g(c, false, true, perm != nullptr, fallback, st);
    launch_th_walls(c, false, true, LS_ALL, st);
    if (tiles) launch_th_g2p_tiles(c, false, true, true, out, st);
    else launch_th_g2p(c, false, true, perm, count, out, st);
    return check_launch();
}

void thermal_swap(mpm_ctx *c) { std::swap(c->th, c->th_alt); }

// compute_udf_temp_at_nodes (nodal_data_ops.cpp:1178-1306), 3-D.
int thermal_udf_tables(mpm_ctx *c, double t) {
    for (int d = 0; d < 3; ++d)
        for (int h = 0; h < 2; ++h) {
            if (!c->tudf_fn[d][h]) continue;
            const int p0 = (d == 0) ? 1 : 0;
            const int p1 = (d == 2) ? 1 : 2;
            std::vector<double> tab((size_t)(c->g.n[p0] + 1) * (c->g.n[p1] + 1) * 2, 0.0);
            for (int j = 0; j <= c->g.n[p0]; ++j)
                for (int k = 0; k <= c->g.n[p1]; ++k) {
                    double co[3] = {0, 0, 0};
                    co[d] = h ? c->s.phi[d] : c->s.plo[d];
                    co[p0] = c->s.plo[p0] + j * c->g.dx[p0];
                    co[p1] = c->s.plo[p1] + k * c->g.dx[p1];
                    double out[2] = {0.0, 0.0};
                    c->tudf_fn[d][h](co[0], co[1], co[2], t, out);
                    const size_t f = ((size_t)j * (c->g.n[p1] + 1) + k) * 2;
                    tab[f] = out[0];
                    tab[f + 1] = out[1];
                }
            HIPCHK(hipMemcpyAsync(c->tudf_dev[d][h], tab.data(), tab.size() * sizeof(double),
                                  hipMemcpyHostToDevice, c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
        }
    return MPM_OK;
}

void thermal_free(mpm_ctx *c) {
    for (int d = 0; d < 3; ++d)
        for (int h = 0; h < 2; ++h) {
            (void)hipFree(c->tudf_dev[d][h]);
            c->tudf_dev[d][h] = nullptr;
        }
    (void)hipFree(c->th);
    (void)hipFree(c->th_alt);
    (void)hipFree(c->tn);
    (void)hipFree(c->tscr);
    c->th = c->th_alt = c->tn = c->tscr = nullptr;
}

}  // namespace xmpm

extern "C" {

void mpm_thermal_specs_default(const mpm_specs *s, mpm_thermal_specs *t) {
    if (!t) return;
    *t = mpm_thermal_specs{};
    for (int d = 0; d < 3; ++d) {
        const int code = (s && s->periodic[d]) ? MPM_BC_TEMP_PERIODIC : MPM_BC_TEMP_ADIABATIC;
        t->bclo[d] = code;
        t->bchi[d] = code;
    }
}

int mpm_enable_thermal(mpm_ctx *c, const mpm_thermal_specs *t) {
    if (!c || !t) return set_err(MPM_E_ARG, "null argument");
    int rc = check_specs(t);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(c->stream));
    const int64_t nn = c->g.nn;
    if (!c->tn) {
        if (hipMalloc(&c->tn, TN_COMPS * nn * sizeof(double)) != hipSuccess ||
            hipMalloc(&c->tscr, nn * sizeof(double)) != hipSuccess)
            return set_err(MPM_E_OOM, "thermal nodal buffers");
    }
    launch_zero(c->tn, TN_COMPS * nn, c->stream);
    launch_zero(c->tscr, nn, c->stream);
    c->ts = *t;
    c->thermal = 1;
    rc = alloc_particle_thermal(c);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(c->stream));
    return check_launch();
}

int mpm_set_wall_temp_udf(mpm_ctx *c, int dir, int hi, mpm_wall_temp_udf_fn fn) {
    if (!c || dir < 0 || dir > 2 || hi < 0 || hi > 1) return set_err(MPM_E_ARG, "bad face");
    if (fn && !c->tudf_dev[dir][hi]) {
        const int p0 = (dir == 0) ? 1 : 0;
        const int p1 = (dir == 2) ? 1 : 2;
        const int64_t n = (int64_t)(c->g.n[p0] + 1) * (c->g.n[p1] + 1) * 2;
        if (hipMalloc(&c->tudf_dev[dir][hi], n * sizeof(double)) != hipSuccess)
            return set_err(MPM_E_OOM, "temperature function table");
        launch_zero(c->tudf_dev[dir][hi], n, c->stream);
        HIPCHK(hipStreamSynchronize(c->stream));
    }
    c->tudf_fn[dir][hi] = fn;
    c->any_tudf = false;
    for (int d = 0; d < 3; ++d)
        for (int h = 0; h < 2; ++h) c->any_tudf |= (c->tudf_fn[d][h] != nullptr);
    return check_launch();
}

int mpm_set_levelset_body_temp(mpm_ctx *c, int body, int temp_bc, double T_wall,
                               double heat_flux, double h_conv, double T_inf) {
    if (!c || body < 0 || body >= c->ls.nb) return set_err(MPM_E_ARG, "no such level-set body");
    if (temp_bc != MPM_LS_TEMP_ADIABATIC && temp_bc != MPM_LS_TEMP_ISOTHERMAL &&
        temp_bc != MPM_LS_TEMP_HEATFLUX && temp_bc != MPM_LS_TEMP_CONVECTIVE)
        return set_err(MPM_E_BC, "level-set temperature condition must be 0, 1, 3 or 4");
    c->ls_tbc[body] = temp_bc;
    c->ls_tpar[body][0] = T_wall;
    c->ls_tpar[body][1] = heat_flux;
    c->ls_tpar[body][2] = h_conv;
    c->ls_tpar[body][3] = T_inf;
    return MPM_OK;
}

int mpm_upload_thermal(mpm_ctx *c, int64_t n, const double *th) {
    if (!c || n < 0 || (n > 0 && !th)) return set_err(MPM_E_ARG, "bad argument");
    if (!c->thermal) return set_err(MPM_E_ARG, "thermal module not enabled");
    if (n != (int64_t)c->cold_radius.size())
        return set_err(MPM_E_ARG, "n_upload must be the count of the last particle upload");
    std::vector<int64_t> ids;
    int rc = slot_ids(c, ids);
    if (rc) return rc;
    const int64_t cap = c->th_cap, np = c->P.np;
    std::vector<double> h((size_t)MPM_NTHERMAL * cap, 0.0);
    for (int64_t s = 0; s < np; ++s) {
        const int64_t id = ids[s];
        if (id < 0) continue;
        for (int f = 0; f < MPM_NTHERMAL; ++f) h[(size_t)f * cap + s] = th[(size_t)f * n + id];
    }
    HIPCHK(hipMemcpy(c->th, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
    return MPM_OK;
}

int mpm_download_thermal(mpm_ctx *c, int64_t n, double *th) {
    if (!c || n < 0 || (n > 0 && !th)) return set_err(MPM_E_ARG, "bad argument");
    if (!c->thermal) return set_err(MPM_E_ARG, "thermal module not enabled");
    if (n != (int64_t)c->cold_radius.size())
        return set_err(MPM_E_ARG, "n_upload must be the count of the last particle upload");
    std::vector<int64_t> ids;
    int rc = slot_ids(c, ids);
    if (rc) return rc;
    const int64_t cap = c->th_cap, np = c->P.np;
    std::vector<double> h((size_t)MPM_NTHERMAL * cap);
    HIPCHK(hipMemcpy(h.data(), c->th, h.size() * sizeof(double), hipMemcpyDeviceToHost));
    std::fill(th, th + (size_t)MPM_NTHERMAL * n, std::numeric_limits<double>::quiet_NaN());
    for (int64_t s = 0; s < np; ++s) {
        const int64_t id = ids[s];
        if (id < 0 || id >= n) continue;
        for (int f = 0; f < MPM_NTHERMAL; ++f) th[(size_t)f * n + id] = h[(size_t)f * cap + s];
    }
    return MPM_OK;
}

}  // extern "C"

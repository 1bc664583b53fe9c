// mpm_internal.h — device-side state and kernel parameter blocks shared by
// the kernels (mpm_kernels.hip) and the context / C ABI (mpm_ctx.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/exagoop_amd.h"

namespace xmpm {

// Grid geometry (one box covering the domain, lo = 0).
struct GridP {
    double plo[3], phi[3], dx[3], dxi[3];
    double rhi[3];  // largest position whose cell is n-1 (periodic wrap bound)
    int n[3];       // cells per direction
    int hi[3];      // n - 1
    int per[3];
    int od[3];      // per-direction order after the small-grid fallback
    int uniform;    // all od equal
    int64_t nn;     // nodes (n0+1)(n1+1)(n2+1)
    int ntc[3];     // cell tiles per direction (kTileCells cells each)
    int ntn[3];     // node tiles per direction (n / kTileCells + 1)
};

// Cells per tile edge for the sorted step: particles are binned into
// kTileCells^3-cell tiles; a P2G workgroup owns one tile and accumulates
// into an LDS block of (kTileCells + 3)^3 nodes (stencil offsets -1..+2).
constexpr int kTileCells = 4;
constexpr int kTNNodes = (kTileCells + 3) * (kTileCells + 3) * (kTileCells + 3);  // P2G node block

// Step physics and wall data (MPMspecs subset).
struct PhysP {
    double grav[3];
    int extloads;
    double slab_lo[3], slab_hi[3], extforce[3];
    double mass_tol;
    double alpha;
    int bclo[3], bchi[3];
    double mu_lo[3], mu_hi[3];
    double wv_lo[9], wv_hi[9];
    double applied_rate, applied_time;
    int delta;
    double cfl, dt_min, dt_max;
};

struct LsP {
    int nb;
    const double *phi[MPM_MAX_LS_BODIES];  // (n0+3)(n1+3)(n2+3), one ghost layer
    int bc[MPM_MAX_LS_BODIES];
    double mu[MPM_MAX_LS_BODIES];
    double w[MPM_MAX_LS_BODIES][3];
};

struct UdfP {
    const double *tab[3][2];  // per face, ((n[p0]+1)(n[p1]+1)) x 3, or null
};

// Particle SoA: component-major with stride cap.
struct PState {
    double *pos;    // 3 * cap
    double *r;      // MPM_NREAL * cap
    int *idata;     // MPM_NINT * cap
    int64_t *id;    // cap; < 0 = removed
    double *vsnap;  // 3 * cap velocity at the last neighbour fill (MUSL images)
    int64_t np;
    int64_t cap;
};

// Nodal data on the valid nodal box, component-major with stride nn.
struct NState {
    double *acc;   // 7 * nn: P2G accumulators m, p0..2, f0..2 (kept zero between uses)
    double *mass;  // nn
    double *vel;   // 3 * nn
    double *frc;   // 3 * nn
    double *dvel;  // 3 * nn
    double *mold;  // nn
};

// Device scalars of the step.
struct Scalars {
    double dt;
    double time;
    unsigned long long dtmin_bits;  // running min of the per-particle dt (positive doubles)
    unsigned long long dtmin_next_bits;  // same, accumulated by the step for the next step
    int err;
    int pad;
    // Particles the tile deposit found outside their cell (outside the
    // domain between redistributions), listed for the deposit after the tile
    // pass: counters of alternate launches (SortState::ogen).
    int ocount[2];
};

// Grid-kernel stage flags.
enum {
    G_FINISH_MASS = 1,
    G_FINISH_VEL = 2,
    G_FINISH_FRC = 4,
    G_BACKUP = 8,
    G_UPDATE = 16,
    G_BCS = 32,
    G_KEEP_FRC = 64
};

// Particle-kernel stage flags.
enum {
    P_G2P_VEL = 1,
    P_G2P_GRAD = 2,
    P_MOVE = 4,
    P_REDIST = 8,
    P_VOLUME = 16,
    P_STRESS = 32,
    P_STRESS_INIT = 64  // apply_constitutive_model(dt=0, applied rate) of the init pass
};

// The node blocks of a deterministic P2G pass (SortState::blk): nc
// components per block (7 with forces, 3 momentum only), tiles stamped with
// pass in slot.
struct BlkGather {
    const double *blk = nullptr;
    const int64_t *slot = nullptr;
    int pass = 0;
    int nc = 0;
};

// Launchers (mpm_kernels.hip).
void launch_zero(double *p, int64_t n, hipStream_t s);
void launch_p2g(const GridP &g, const PhysP &ph, const PState &P, const NState &N, int um,
                int uv, int uf, const double *vimg, hipStream_t s);
void launch_grid(const GridP &g, const PhysP &ph, const LsP &ls, const UdfP &udf,
                 const NState &N, int flags, double dt, const Scalars *sc, hipStream_t s,
                 const int *tile_epoch = nullptr, int epoch = 0);
void launch_particle(const GridP &g, const PhysP &ph, const LsP &ls, const UdfP &udf,
                     const PState &P, const NState &N, int flags, double dt, double time,
                     const Scalars *sc, hipStream_t s);
// fused: take the minimum the previous step's particle kernel accumulated;
// advance: first add the previous dt to the time (the previous step's end);
// scale: multiplies the final dt (a coarse step of that many fine steps).
void launch_dt(const GridP &g, const PhysP &ph, const PState &P, Scalars *sc, int fixed,
               double fixed_dt, hipStream_t s, int fused = 0, int advance = 0,
               double scale = 1.0);
void launch_advance_time(Scalars *sc, hipStream_t s);
void launch_energies(const PState &P, double *out6, hipStream_t s);
void launch_snapshot_vel(const PState &P, hipStream_t s);
void launch_redistribute(const GridP &g, const PState &P, hipStream_t s);
// Marks particles inside any level-set body removed; adds their count to *removed.
void launch_remove_inside_ls(const GridP &g, const LsP &ls, const PState &P, int *removed,
                             hipStream_t s);

// Sorted step (mpm_sort.hip).
struct SortState {
    int *count;       // nkeys + 1 (last slot: removed particles)
    int *cell_start;  // nkeys + 1, exclusive scan of count
    int *key;         // cap
    int *rank;        // cap
    int *perm;        // cap: sorted position -> current index
    void *scan_tmp;
    size_t scan_bytes;
    int nkeys;
    int *tile_epoch;  // node tiles: last epoch a deposit landed there
    // P2G packets by sorted position, component-major with stride pcap:
    // x[3], m, m v[3], a[3] = m g (+ slab load), B[6] = -V sigma (xx, xy,
    // xz, yy, yz, zz); all but x zero for a particle with phase != 0.
    double *pkt;
    int64_t pcap;
    // the last sort wrote the packets (otherwise P2G gathers the particle
    // fields through perm)
    int pkt_valid;
    // Non-empty cell tiles after the sort: tiles[0] = count, then the tile
    // indices (ntiles + 1 entries), in runs of increasing index whose order
    // depends on scheduling.
    int *tiles;
    int ntiles;
    // Deterministic P2G (MPM_OPT_DETERMINISTIC; blk null when off).  The
    // sort orders each cell's particles by their current index, the tile
    // kernel writes the node block of the tile at list position p whole to
    // blk[(p * NC + c) * kTNN + e] (NC = 7 with forces, 3 without) and
    // tile_slot[t] = (pass << 32) | p; a gather pass then sums every node's
    // blocks in a fixed order.  pass counts the deterministic P2G launches.
    double *blk = nullptr;
    int64_t *tile_slot = nullptr;
    int pass = 0;
    // Wavefronts per tile of the tile kernels (MPM_OPT_TILE_WAVES): 4 or 8,
    // 0 to choose by the particle count.
    int waves = 0;
    // The tile deposit's list of particles outside their cell: entries in
    // rank (free between the scatter and the next sort), their count in
    // ocount[ogen & 1] (Scalars::ocount); ogen counts the launches that list.
    int *ocount = nullptr;
    int ogen = 0;
};
constexpr int kPktComps = 16;
size_t sort_scan_bytes(int nkeys);
// out_id: id array of the buffer the step writes into (slots past the live
// count are marked free), or null.
// packets: also build the P2G packets of the sorted particles (needed with
// external slab loads; without them P2G can gather the fields through perm).
// The step's dt, computed by the sort's first launch when it needs no
// reduction over the particles (fixed dt, or the minimum the previous
// particle kernel accumulated): the arguments of launch_dt.
struct DtFold {
    int on = 0;
    int fixed = 0;
    double fixed_dt = 0.0;
    int advance = 0;
    double scale = 1.0;
    Scalars *sc = nullptr;
    Scalars *bak = nullptr;  // *sc is copied here first (the learned step's undo), when set
};
void launch_sort(const GridP &g, const PhysP &ph, const PState &P, SortState &S,
                 int64_t *out_id, int packets, hipStream_t s, const DtFold &dt = DtFold());
// fallback: particles may lie outside their tile (not redistributed every
// step); they deposit straight to global memory.  Without it such a particle
// sets Scalars::err.  With S.blk set (deterministic P2G) the tiles write
// their node blocks instead of adding them to N.acc, and a gather pass adds
// each node's blocks to N.acc in a fixed order.
void launch_p2g_tiles(const GridP &g, const PhysP &ph, const PState &P, SortState &S,
                      const NState &N, int epoch, int momentum_only, int fallback, Scalars *sc,
                      hipStream_t s);
void launch_particle_sorted(const GridP &g, const PhysP &ph, const LsP &ls, const UdfP &udf,
                            const PState &Pin, const PState &Pout, const SortState &S,
                            const NState &N, int flags, const Scalars *sc, hipStream_t s);
// Few tiles: fewer particles than one full tile (4^3 cells x 8) per compute
// unit; the tile kernels then run 8 wavefronts per tile.
bool few_tiles(int64_t np);
// MUSL's first G2P (velocities only) over the tiles, in place (uniform
// order, the state sorted this step; outside: as launch_particle_fast).
void launch_g2p_vel_tiles(const GridP &g, const PhysP &ph, const PState &P, const SortState &S,
                          const NState &N, int outside, Scalars *sc, hipStream_t s);
// The fused particle kernel over the tiles (uniform order, the state sorted
// this step; level-set bodies and UDF wall tables); also accumulates the next
// step's dt minimum.  outside: particles are redistributed every few steps
// (a particle outside the domain lies outside its tile and gathers from the
// nodal arrays).
void launch_particle_fast(const GridP &g, const PhysP &ph, const LsP &ls, const UdfP &udf,
                          const PState &Pin, const PState &Pout, const SortState &S,
                          const NState &N, int vel, int redist, int outside, Scalars *sc,
                          hipStream_t s);

}  // namespace xmpm

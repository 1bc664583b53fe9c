// mpm_batch.hip — batched ensembles: many independent scenes of uneven size
// advanced by ONE persistent kernel launch per step, the steps replayed from
// a captured HIP graph.  No counterpart in the reference (one scene per run);
// each scene runs the sorted USL step of mpm_step (main.cpp:83-187) built
// from the same bodies (mpm_tiles.h, mpm_stages.h).
//
// Work of one step, per scene, in six stages (seven when deterministic):
//   KEYS     cell key + rank of each particle          (particle chunks)
//   SCAN     exclusive scan of the cell counts, list of non-empty tiles
//   SCATTER  sorted permutation + P2G packets          (particle chunks)
//   ORDER    (deterministic) each cell's particles by index (tile shares)
//   P2G      tile deposits                              (tile shares)
//   GRID     nodal pipeline over the active node tiles   (node-tile chunks)
//   PART     fused particle pipeline, sorted write, dt candidates (tile shares)
// Deterministic batches (MPM_OPT_DETERMINISTIC) run the single-scene
// step's atomic-free P2G: the scatter writes no packets and P2G gathers the
// fields through the permutation, ORDER fixes the order of each cell's
// particles, the tile deposits write whole node blocks stamped with the step
// number, and GRID first adds each node's blocks in a fixed order.
// The tasks sit in eight queues, stage-major: every task of stage k of a
// scene comes before every task of stage k + 1 of that scene, and a scene's
// tasks are all in one queue.  A workgroup takes tasks from its own queue
// (workgroup b runs on XCD b mod 8, so a scene's data meets in one L2) and
// then from the others.  A task first waits until every task of the previous
// stage of its scene has finished (a per-scene counter): those tasks were
// dequeued earlier from the same queue, by workgroups that are running, so
// the wait always ends whatever the residency.  The last particle task of a
// scene ends its step (time += dt, next dt from the candidates, counters
// reset) and the last workgroup out rewinds the queues, so consecutive
// launches need no host work or memset and can be captured in a graph.
// Waits are bounded: past the limit the kernel sets an error flag and goes on
// (every index it then reads is still in range), and the call reports it.

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstring>
#include <numeric>
#include <vector>

#include "mpm_context.h"
#include "mpm_stages.h"
#include "mpm_tiles.h"

namespace xmpm {

namespace {

enum { B_KEYS = 0, B_SCAN, B_SCATTER, B_ORDER, B_P2G, B_GRID, B_PART, B_NSTAGE };
constexpr int kQueues = 8;
constexpr int kScanChunk = 16 * kBlock;  // cell counts per scan round
constexpr int kGridFlags = G_FINISH_MASS | G_FINISH_VEL | G_FINISH_FRC | G_BACKUP | G_UPDATE |
                           G_BCS;

// One scene as the kernel sees it (read only).  P[b] views the scene's
// particles in buffer b: pointers offset to the scene's first slot, the
// stride (cap) of the batch.  Particles live in P[steps & 1].
struct BScene {
    GridP g;
    PhysP ph;
    PState P[2];
    NState N;
    SortState S;
    Scalars *sc;
    double fixed_dt;
    int fixed;
    int ntask[B_NSTAGE];
    int64_t np;  // particle slots of the scene (live + removed)
};

// Per-scene progress within a launch (zero between launches) and the
// scene's completed steps.
struct BSync {
    int done[B_NSTAGE];
    int step;
    int pad;
    double last_dt;  // dt of the last completed step
};

struct BHead {
    unsigned v;
    unsigned pad[15];  // one 64-byte line per queue head
};

struct BCtl {
    BHead head[kQueues];
    unsigned exited;
    int err;  // a bounded wait ran out
    int pad[14];
    // Instrumented launches: per stage, wall-clock ticks of task work and of
    // the wait for the previous stage, summed over tasks; tasks run; ticks
    // spent outside tasks (taking tasks, leaving).
    unsigned long long busy[B_NSTAGE], wait[B_NSTAGE], ntasks[B_NSTAGE], other;
};

struct BQueues {
    const int4 *task;  // (scene, stage, part, nparts)
    int off[kQueues + 1];
};

struct ScanLds {
    int wsum[kBlock / 64];
    int tsum[kBlock / 64];
};

// Permutation entries of a tile the ORDER stage sorts in LDS (a longer range
// is sorted in global memory).
constexpr int kOrderTileLds = 4096;

template <bool DET>
union BLds;
template <>
union BLds<false> {
    P2gLds<7> p2g;
    PartLds<6> part;
    ScanLds scan;
};
template <>
union BLds<true> {
    P2gLdsOf<true, true, true> p2g;
    PartLds<6> part;
    ScanLds scan;
    int order[kOrderTileLds];
};

// Each cell of a tile gets its particles in increasing current index (the
// order k_cell_order gives the single-scene deterministic step): the tile's
// contiguous range of the permutation is staged in LDS and one thread per
// cell insertion-sorts its run.
__device__ void order_tile(const SortState &S, int tile, int *seg) {
    const int tid = threadIdx.x;
    const int c0 = tile * kCellsPerTile;
    const int r0 = S.cell_start[c0], r1 = S.cell_start[c0 + kCellsPerTile];
    const int b = tid < kCellsPerTile ? S.cell_start[c0 + tid] : r1;
    const int e = tid < kCellsPerTile ? S.cell_start[c0 + tid + 1] : r1;
    __syncthreads();  // the workgroup's previous tile is done with the LDS
    if (r1 - r0 > kOrderTileLds) {  // uniform over the block
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
    for (int i = tid; i < r1 - r0; i += kBlock) seg[i] = S.perm[r0 + i];
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
    for (int i = tid; i < r1 - r0; i += kBlock) S.perm[r0 + i] = seg[i];
}

// Wait until *ctr >= need, then acquire.  Bounded: past `limit` polls the
// error flag is set, and once it is set every wait of the launch gives up at
// its next check, so a launch always ends.
__device__ __forceinline__ void wait_count(int *ctr, int need, BCtl *ctl, unsigned limit) {
    unsigned n = 0;
    while (__hip_atomic_load(ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < need) {
        __builtin_amdgcn_s_sleep(2);
        ++n;
        if (n > limit) {
            atomicOr(&ctl->err, 1);
            break;
        }
        if ((n & 255) == 0 &&
            __hip_atomic_load(&ctl->err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            break;
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}

__device__ __forceinline__ double ld_double(double *p) {
    return __longlong_as_double((long long)__hip_atomic_load(
        reinterpret_cast<unsigned long long *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
}

__device__ __forceinline__ void st_double(double *p, double v) {
    __hip_atomic_store(reinterpret_cast<unsigned long long *>(p),
                       (unsigned long long)__double_as_longlong(v), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
}

// Exclusive sum of the scene's cell counts (count[0..nkeys], the last entry
// counts removed particles) into cell_start, and the non-empty tiles in
// increasing order into tiles[1..] with their number in tiles[0].  Rounds of
// 4096 counts, 16 consecutive ones per thread (four 16-byte loads, the next
// round's in flight); the four threads of a 64-cell tile add their sums for
// the tile flag.  count and cell_start are padded to whole rounds.
__device__ void scan_scene(const SortState &S, ScanLds &L) {
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int n = S.nkeys + 1;
    const int rounds = (n + kScanChunk - 1) / kScanChunk;
    int carry = 0, tcarry = 0;
    const int4 *src = reinterpret_cast<const int4 *>(S.count) + 4 * tid;
    int4 cur[4], nxt[4];
    for (int k = 0; k < 4; ++k) cur[k] = src[k];
    for (int r = 0; r < rounds; ++r) {
        if (r + 1 < rounds)
            for (int k = 0; k < 4; ++k) nxt[k] = src[(r + 1) * (kScanChunk / 4) + k];
        int v[16];
        for (int k = 0; k < 4; ++k) {
            v[4 * k] = cur[k].x;
            v[4 * k + 1] = cur[k].y;
            v[4 * k + 2] = cur[k].z;
            v[4 * k + 3] = cur[k].w;
        }
        int sum = 0;
        for (int k = 0; k < 16; ++k) {
            const int c = v[k];
            v[k] = sum;  // thread-local exclusive prefix
            sum += c;
        }
        int q = sum + __shfl_xor(sum, 1, 64);
        q += __shfl_xor(q, 2, 64);
        const int tile = r * (kScanChunk / 64) + (tid >> 2);
        const bool flag = (tid & 3) == 0 && tile < S.ntiles && q > 0;
        int x = sum;  // wavefront inclusive scan of the thread sums
        for (int off = 1; off < 64; off <<= 1) {
            const int y = __shfl_up(x, off, 64);
            if (lane >= off) x += y;
        }
        const unsigned long long fb = __ballot(flag);
        if (lane == 63) L.wsum[w] = x;
        if (lane == 0) L.tsum[w] = __popcll(fb);
        __syncthreads();
        int woff = 0, wtot = 0, toff = 0, ttot = 0;
        for (int k = 0; k < kBlock / 64; ++k) {
            if (k < w) {
                woff += L.wsum[k];
                toff += L.tsum[k];
            }
            wtot += L.wsum[k];
            ttot += L.tsum[k];
        }
        const int base = carry + woff + x - sum;
        int4 *dst = reinterpret_cast<int4 *>(S.cell_start) + r * (kScanChunk / 4) + 4 * tid;
        for (int k = 0; k < 4; ++k)
            dst[k] = make_int4(base + v[4 * k], base + v[4 * k + 1], base + v[4 * k + 2],
                               base + v[4 * k + 3]);
        if (flag) S.tiles[1 + tcarry + toff + __popcll(fb & ((1ull << lane) - 1ull))] = tile;
        carry += wtot;
        tcarry += ttot;
        __syncthreads();  // wsum / tsum reused by the next round
        for (int k = 0; k < 4; ++k) cur[k] = nxt[k];
    }
    if (tid == 0) S.tiles[0] = tcarry;
}

// End of a scene's step, by the thread that finished its last particle task:
// the time advances by the step's dt and the next dt is taken from the
// candidates the particle tasks reduced (k_dt_final's arithmetic); then the
// stage counters are rewound and the step count advances.
__device__ void finish_scene(const BScene &sd, BSync &sy) {
    Scalars *sc = sd.sc;
    const unsigned long long nb = atomicExch(
        &sc->dtmin_next_bits, (unsigned long long)__double_as_longlong(__builtin_huge_val()));
    const double dt_now = ld_double(&sc->dt);
    double dt;
    if (sd.fixed) {
        dt = sd.fixed_dt;
    } else {
        dt = __longlong_as_double((long long)nb);
        if (dt == __builtin_huge_val()) dt = 1.79769313486231570815e+308;  // numeric_limits max
        dt = sd.ph.cfl * dt;
        dt = fmax(fmin(dt, sd.ph.dt_max), sd.ph.dt_min);
    }
    st_double(&sc->time, ld_double(&sc->time) + dt_now);
    st_double(&sc->dt, dt);
    st_double(&sy.last_dt, dt_now);
    for (int k = 0; k < B_NSTAGE; ++k)
        __hip_atomic_store(&sy.done[k], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_fetch_add(&sy.step, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ int64_t share(int64_t n, int part, int nparts) {
    return n * part / nparts;
}

template <int O, bool DELTA, bool DET>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(2))) void k_batch_step(const BScene *__restrict__ scenes,
                                                       BSync *sync, BQueues Q, BCtl *ctl,
                                                       unsigned spin_limit, int prof) {
    __shared__ BLds<DET> lds;
    __shared__ int4 task_s;
    const int tid = threadIdx.x;
    const int home = blockIdx.x & (kQueues - 1);
    unsigned dry = 0;  // queues this workgroup found empty (thread 0)
    unsigned long long t_end = prof ? wall_clock64() : 0;  // end of the last task (thread 0)
    unsigned long long other = 0;
    for (;;) {
        if (tid == 0) {
            int4 t = make_int4(-1, 0, 0, 0);
            for (int k = 0; k < kQueues; ++k) {
                const int q = (home + k) & (kQueues - 1);
                if (dry & (1u << q)) continue;
                const int n = Q.off[q + 1] - Q.off[q];
                if (n > 0) {
                    const unsigned p = atomicAdd(&ctl->head[q].v, 1u);
                    if (p < (unsigned)n) {
                        t = Q.task[Q.off[q] + p];
                        break;
                    }
                }
                dry |= 1u << q;
            }
            task_s = t;
        }
        __syncthreads();
        const int4 t = task_s;
        if (t.x < 0) break;  // every queue drained (uniform)
        const int s = t.x, stage = t.y, part = t.z, nparts = t.w;
        const BScene &sd = scenes[s];
        unsigned long long t_start = 0, t_work = 0;
        if (prof && tid == 0) t_start = wall_clock64();
        if (stage > 0) {
            // the previous stage (ORDER runs only in deterministic batches)
            const int prev = (!DET && stage == B_P2G) ? B_SCATTER : stage - 1;
            if (tid == 0) wait_count(&sync[s].done[prev], sd.ntask[prev], ctl, spin_limit);
            __syncthreads();
        }
        if (prof && tid == 0) t_work = wall_clock64();
        const int step = __hip_atomic_load(&sync[s].step, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
        const int epoch = step + 1;  // number of the step being taken
        const PState &Pin = sd.P[step & 1];
        const PState &Pout = sd.P[(step & 1) ^ 1];
        const SortState &S = sd.S;
        if (stage == B_KEYS) {
            const int64_t a = share(sd.np, part, nparts), b = share(sd.np, part + 1, nparts);
            for (int64_t i0 = a; i0 < b; i0 += kBlock) keys_wave(sd.g, Pin, S, i0 + tid, b);
        } else if (stage == B_SCAN) {
            scan_scene(S, lds.scan);
        } else if (stage == B_SCATTER) {
            const int64_t a = share(sd.np, part, nparts), b = share(sd.np, part + 1, nparts);
            for (int64_t i = a + tid; i < b; i += kBlock)
                scatter_one(Pin, S, Pout.id, sd.ph, DET ? 0 : 1, i);
        } else if (DET && stage == B_ORDER) {
            const int nt = S.tiles[0];
            const int a = (int)share(nt, part, nparts), b = (int)share(nt, part + 1, nparts);
            if constexpr (DET)
                for (int k = a; k < b; ++k) order_tile(S, S.tiles[1 + k], lds.order);
        } else if (stage == B_P2G) {
            const int nt = S.tiles[0];
            const int a = (int)share(nt, part, nparts), b = (int)share(nt, part + 1, nparts);
            // deterministic: block of list position k stamped with the step
            for (int k = a; k < b; ++k)
                p2g_tile<O, true, false, DET, DET>(sd.g, sd.ph, Pin, S, sd.N.acc, epoch,
                                                   &sd.sc->err, S.tiles[1 + k], lds.p2g, k, epoch);
        } else if (stage == B_GRID) {
            LsP ls;
            ls.nb = 0;
            UdfP udf;
            for (int d = 0; d < 3; ++d) udf.tab[d][0] = udf.tab[d][1] = nullptr;
            // Node tiles (4^3 nodes), one per wavefront: a tile idle since
            // before the previous step already holds the zero massless state
            // and is passed over with one read.
            const double dt = sd.sc->dt;
            const GridP &g = sd.g;
            const int ntx = g.ntn[0], nty = g.ntn[1];
            const int ntn = ntx * nty * g.ntn[2];
            const int a = (int)share(ntn, part, nparts), b = (int)share(ntn, part + 1, nparts);
            const int lane = tid & 63;
            for (int t0 = a; t0 < b; t0 += kBlock / 64) {
                const int t = t0 + (tid >> 6);
                if (t >= b) continue;
                // (written by this launch's P2G tasks: a vector load, never a
                // scalar-cache hit)
                const int e = __hip_atomic_load(&S.tile_epoch[t], __ATOMIC_RELAXED,
                                                __HIP_MEMORY_SCOPE_AGENT);
                if (e != epoch && e != epoch - 1) continue;
                const int I = (t % ntx) * kTileCells + (lane & 3);
                const int J = ((t / ntx) % nty) * kTileCells + ((lane >> 2) & 3);
                const int K = (t / (ntx * nty)) * kTileCells + (lane >> 4);
                if (I > g.n[0] || J > g.n[1] || K > g.n[2]) continue;
                const int64_t node = nidx(g, I, J, K);
                if (DET && e == epoch) {
                    // the node's blocks, in a fixed order, into the
                    // accumulators (which hold no other deposit)
                    BlkGather bg;
                    bg.blk = S.blk;
                    bg.slot = S.tile_slot;
                    bg.pass = epoch;
                    bg.nc = 7;
                    double sm[7];
                    if (blk_gather<7, true>(g, bg, I, J, K, sm))
                        for (int c = 0; c < 7; ++c) sd.N.acc[c * g.nn + node] += sm[c];
                }
                grid_at(g, sd.ph, ls, udf, sd.N, kGridFlags, dt, node, I, J, K, S.tile_epoch,
                        epoch);
            }
        } else {
            const int nt = S.tiles[0];
            const int a = (int)share(nt, part, nparts), b = (int)share(nt, part + 1, nparts);
            const double dt = sd.sc->dt, time = sd.sc->time;
            double cmin = __builtin_huge_val();
            for (int k = a; k < b; ++k)
                particle_tile<O, true, DELTA>(sd.g, sd.ph, Pin, Pout, S, sd.N, 1, dt, time,
                                              &sd.sc->err, S.tiles[1 + k], lds.part, cmin);
            dt_wave_min(cmin, sd.sc);
        }
        __syncthreads();  // the task's writes are done (and task_s is free)
        if (tid == 0) {
            if (prof) {
                const unsigned long long t1 = wall_clock64();
                atomicAdd(&ctl->wait[stage], t_work - t_start);
                atomicAdd(&ctl->busy[stage], t1 - t_work);
                atomicAdd(&ctl->ntasks[stage], 1ull);
                other += t_start - t_end;
                t_end = t1;
            }
            const int prev = __hip_atomic_fetch_add(&sync[s].done[stage], 1, __ATOMIC_ACQ_REL,
                                                    __HIP_MEMORY_SCOPE_AGENT);
            if (stage == B_PART && prev + 1 == sd.ntask[B_PART]) finish_scene(sd, sync[s]);
        }
    }
    // Every task is dequeued; the last workgroup out has seen all finish.
    if (tid == 0) {
        if (prof) atomicAdd(&ctl->other, other + (wall_clock64() - t_end));
        const unsigned e = __hip_atomic_fetch_add(&ctl->exited, 1u, __ATOMIC_ACQ_REL,
                                                  __HIP_MEMORY_SCOPE_AGENT);
        if (e + 1 == gridDim.x) {
            for (int q = 0; q < kQueues; ++q)
                __hip_atomic_store(&ctl->head[q].v, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&ctl->exited, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

__global__ void k_fill_int(int *p, int64_t n, int v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

typedef void (*BatchKernel)(const BScene *, BSync *, BQueues, BCtl *, unsigned, int);

template <bool DET>
BatchKernel batch_kernel_of(int O, int delta) {
    if (delta) {
        if (O == 1) return k_batch_step<1, true, DET>;
        if (O == 2) return k_batch_step<2, true, DET>;
        return k_batch_step<3, true, DET>;
    }
    if (O == 1) return k_batch_step<1, false, DET>;
    if (O == 2) return k_batch_step<2, false, DET>;
    return k_batch_step<3, false, DET>;
}

BatchKernel batch_kernel(int O, int delta, int det) {
    return det ? batch_kernel_of<true>(O, delta) : batch_kernel_of<false>(O, delta);
}

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

}  // namespace

}  // namespace xmpm

using namespace xmpm;

struct mpm_batch {
    int ns = 0;
    int O = 2;
    int delta = 0;
    std::vector<mpm_specs> specs;
    std::vector<BScene> scn;     // host copy of the scene table
    std::vector<int64_t> np, off;
    int64_t cap = 0;             // particle slots of all scenes (buffer stride)
    std::vector<char> uploaded;
    std::vector<std::vector<double>> cold_radius, cold_yacc;
    std::vector<std::vector<int>> cold_rigid;
    char *arena = nullptr;       // all step state, one allocation
    size_t arena_bytes = 0;
    BSync *sync = nullptr;
    BCtl *ctl = nullptr;
    Scalars *sc = nullptr;
    BScene *d_scn = nullptr;
    int4 *d_queue = nullptr;
    int queue_cap = 0;
    BQueues q{};
    int ntasks = 0;
    void *scan_tmp = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    int gsteps = 0;
    mpm_batch_tuning tune{};
    int max_wgs = 0;
    unsigned spin_limit = 1u << 18;  // polls per wait (a step waits microseconds)
    bool inited = false;
    int64_t steps = 0;
    double step_ms = 0.0;
    // Deterministic step (MPM_OPT_DETERMINISTIC): requested, and the node
    // blocks and tile stamps of every scene (allocated by the option).
    int det = 0;
    char *det_arena = nullptr;
    std::vector<size_t> det_slot_off;  // per scene, bytes into det_arena
};

namespace {

void drop_graph(mpm_batch *b) {
    if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->gexec = nullptr;
    b->graph = nullptr;
    b->gsteps = 0;
}

// Task queues of a decomposition: scenes spread over the queues by particle
// count (largest first onto the lightest queue), each queue stage-major with
// its scenes largest first.
int build_queues(mpm_batch *b) {
    const mpm_batch_tuning &t = b->tune;
    std::vector<int> order(b->ns);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return b->np[x] > b->np[y]; });
    std::vector<std::vector<int>> qs(kQueues);
    std::vector<int64_t> load(kQueues, 0);
    for (int s : order) {
        const int q = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        qs[q].push_back(s);
        load[q] += b->np[s] + b->scn[s].g.nn / 8;
    }
    for (int s = 0; s < b->ns; ++s) {
        BScene &sd = b->scn[s];
        const int64_t n = b->np[s];
        const int ks = (int)std::max<int64_t>(1, (n + t.sort_chunk - 1) / t.sort_chunk);
        const int kt = (int)std::max<int64_t>(1, (n + t.tile_chunk - 1) / t.tile_chunk);
        const int kg = (int)std::max<int64_t>(1, (sd.g.nn + t.node_chunk - 1) / t.node_chunk);
        sd.ntask[B_KEYS] = ks;
        sd.ntask[B_SCAN] = 1;
        sd.ntask[B_SCATTER] = ks;
        sd.ntask[B_ORDER] = b->det ? kt : 0;
        sd.ntask[B_P2G] = kt;
        sd.ntask[B_GRID] = kg;
        sd.ntask[B_PART] = kt;
    }
    std::vector<int4> tasks;
    b->q.off[0] = 0;
    for (int q = 0; q < kQueues; ++q) {
        for (int stg = 0; stg < B_NSTAGE; ++stg)
            for (int s : qs[q]) {
                const int n = b->scn[s].ntask[stg];
                for (int p = 0; p < n; ++p) tasks.push_back(make_int4(s, stg, p, n));
            }
        b->q.off[q + 1] = (int)tasks.size();
    }
    b->ntasks = (int)tasks.size();
    if (b->ntasks > b->queue_cap) {
        (void)hipFree(b->d_queue);
        b->d_queue = nullptr;
        HIPCHK(hipMalloc(&b->d_queue, tasks.size() * sizeof(int4)));
        b->queue_cap = b->ntasks;
    }
    HIPCHK(hipMemcpyAsync(b->d_queue, tasks.data(), tasks.size() * sizeof(int4),
                          hipMemcpyHostToDevice, b->st));
    HIPCHK(hipMemcpyAsync(b->d_scn, b->scn.data(), b->scn.size() * sizeof(BScene),
                          hipMemcpyHostToDevice, b->st));
    HIPCHK(hipStreamSynchronize(b->st));
    b->q.task = b->d_queue;
    drop_graph(b);
    return MPM_OK;
}

int launch_steps(mpm_batch *b, int n, int prof = 0) {
    BatchKernel k = batch_kernel(b->O, b->delta, b->det);
    for (int i = 0; i < n; ++i)
        hipLaunchKernelGGL(k, dim3(b->tune.workgroups), dim3(kBlock), 0, b->st, b->d_scn, b->sync,
                           b->q, b->ctl, b->spin_limit, prof);
    return check_launch();
}

// Flags left by the last steps: a bounded wait that ran out, a particle
// outside its sorted tile.  Cleared after reading.
int check_flags(mpm_batch *b) {
    int err = 0;
    HIPCHK(hipMemcpy(&err, &b->ctl->err, sizeof(int), hipMemcpyDeviceToHost));
    std::vector<Scalars> h(b->ns);
    HIPCHK(hipMemcpy(h.data(), b->sc, b->ns * sizeof(Scalars), hipMemcpyDeviceToHost));
    int tile_err = 0;
    for (int s = 0; s < b->ns; ++s) tile_err |= h[s].err;
    if (err) {
        int zero = 0;
        HIPCHK(hipMemcpy(&b->ctl->err, &zero, sizeof(int), hipMemcpyHostToDevice));
        return set_err(MPM_E_HIP, "batched step: a stage wait ran out of its bound");
    }
    if (tile_err) {
        for (int s = 0; s < b->ns; ++s) h[s].err = 0;
        HIPCHK(hipMemcpy(b->sc, h.data(), b->ns * sizeof(Scalars), hipMemcpyHostToDevice));
        return set_err(MPM_E_UNSUPPORTED, "batched step: a particle lay outside its sorted tile");
    }
    return MPM_OK;
}

// Resident workgroups of the step kernel over the device (a multiple of the
// queue count): the persistent launch's default size.
int batch_occupancy(mpm_batch *b) {
    int per_cu = 0;
    hipDeviceProp_t prop;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, reinterpret_cast<const void *>(batch_kernel(b->O, b->delta, b->det)),
            kBlock, 0) != hipSuccess ||
        hipGetDeviceProperties(&prop, 0) != hipSuccess)
        return set_err(MPM_E_HIP, "occupancy query failed");
    b->max_wgs = std::max(kQueues, per_cu * prop.multiProcessorCount / kQueues * kQueues);
    return MPM_OK;
}

// Marks every deterministic block stamp empty (-1): at set-up and after an
// arena restore, which rewinds the step numbers the stamps compare with.
int clear_det_stamps(mpm_batch *b) {
    if (!b->det_arena) return MPM_OK;
    for (int s = 0; s < b->ns; ++s)
        HIPCHK(hipMemsetAsync(b->det_arena + b->det_slot_off[s], 0xff,
                              (size_t)b->scn[s].S.ntiles * sizeof(int64_t), b->st));
    return MPM_OK;
}

int valid_tuning(const mpm_batch *b, mpm_batch_tuning &t) {
    if (t.workgroups == 0) t.workgroups = b->max_wgs;
    if (t.workgroups < kQueues || t.workgroups % kQueues || t.workgroups > 65536 ||
        t.sort_chunk < kBlock || t.tile_chunk < 1 || t.node_chunk < kBlock || t.graph_steps < 1 ||
        t.graph_steps > 1000)
        return set_err(MPM_E_ARG, "bad batch tuning");
    return MPM_OK;
}

}  // namespace

extern "C" {

int mpm_batch_create(int ns, const mpm_specs *specs, const int64_t *np, mpm_batch **out) {
    if (!out || ns < 1 || !specs || !np) return set_err(MPM_E_ARG, "bad argument");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return set_err(MPM_E_NODEVICE, "no HIP device");
    mpm_batch *b = new mpm_batch();
    auto fail = [&](int code) {
        mpm_batch_destroy(b);
        return code;
    };
    b->ns = ns;
    b->specs.assign(specs, specs + ns);
    b->scn.resize(ns);
    b->np.assign(np, np + ns);
    b->off.resize(ns);
    b->uploaded.assign(ns, 0);
    b->cold_radius.resize(ns);
    b->cold_yacc.resize(ns);
    b->cold_rigid.resize(ns);
    for (int s = 0; s < ns; ++s) {
        const mpm_specs *sp = specs + s;
        BScene &sd = b->scn[s];
        std::memset(&sd, 0, sizeof(sd));
        int rc = specs_params(sp, sd.g, sd.ph);
        if (rc) return fail(rc);
        if (!sd.g.uniform) return fail(set_err(MPM_E_UNSUPPORTED, "batch: one order per scene"));
        if (s == 0) {
            b->O = sd.g.od[0];
            b->delta = sd.ph.delta;
        }
        if (sd.g.od[0] != b->O) return fail(set_err(MPM_E_UNSUPPORTED, "batch: one order"));
        if (sd.ph.delta != b->delta)
            return fail(set_err(MPM_E_UNSUPPORTED, "batch: one strain-update setting"));
        if (sp->stress_update_scheme != MPM_USL)
            return fail(set_err(MPM_E_UNSUPPORTED, "batch: USL only"));
        if (sp->num_redist != 1)
            return fail(set_err(MPM_E_UNSUPPORTED, "batch: redistribution every step"));
        if (!walls_at_rest(sp)) return fail(set_err(MPM_E_UNSUPPORTED, "batch: walls at rest"));
        if (np[s] < 1 || np[s] >= ((int64_t)1 << 31))
            return fail(set_err(MPM_E_ARG, "batch: particle count out of range"));
        sd.fixed = sp->fixed_timestep == 1;
        sd.fixed_dt = sp->timestep;
        sd.np = np[s];
        b->off[s] = b->cap;
        b->cap += (np[s] + 63) / 64 * 64;
    }
    if (hipStreamCreateWithFlags(&b->st, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&b->ev[0]) != hipSuccess || hipEventCreate(&b->ev[1]) != hipSuccess)
        return fail(set_err(MPM_E_HIP, "stream/event creation failed"));
    // Arena layout: particle buffers and sort arrays with the batch stride,
    // then per-scene sort, nodal and scalar blocks.
    const int64_t cap = b->cap;
    size_t at = 0;
    auto take = [&](size_t bytes) {
        const size_t o = at;
        at = align256(at + bytes);
        return o;
    };
    size_t o_pos[2], o_r[2], o_idata[2], o_id[2];
    for (int k = 0; k < 2; ++k) {
        o_pos[k] = take(3 * cap * sizeof(double));
        o_r[k] = take((size_t)MPM_NREAL * cap * sizeof(double));
        o_idata[k] = take((size_t)MPM_NINT * cap * sizeof(int));
        o_id[k] = take(cap * sizeof(int64_t));
    }
    const size_t o_key = take(cap * sizeof(int)), o_rank = take(cap * sizeof(int)),
                 o_perm = take(cap * sizeof(int)),
                 o_pkt = take((size_t)kPktComps * cap * sizeof(double));
    std::vector<size_t> o_count(ns), o_start(ns), o_tiles(ns), o_te(ns), o_nod(ns);
    size_t scan_bytes = 0;
    for (int s = 0; s < ns; ++s) {
        const GridP &g = b->scn[s].g;
        const int ntiles = g.ntc[0] * g.ntc[1] * g.ntc[2];
        const int nkeys = ntiles * kTileCells * kTileCells * kTileCells;
        const size_t padded = (size_t)(nkeys + 1 + kScanChunk - 1) / kScanChunk * kScanChunk;
        o_count[s] = take(padded * sizeof(int));
        o_start[s] = take(padded * sizeof(int));
        o_tiles[s] = take((ntiles + 1) * sizeof(int));
        o_te[s] = take((size_t)g.ntn[0] * g.ntn[1] * g.ntn[2] * sizeof(int));
        o_nod[s] = take(18 * g.nn * sizeof(double));
        scan_bytes = std::max(scan_bytes, sort_scan_bytes(nkeys));
    }
    const size_t o_sc = take(ns * sizeof(Scalars)), o_sync = take(ns * sizeof(BSync)),
                 o_ctl = take(sizeof(BCtl));
    b->arena_bytes = at;
    if (hipMalloc(&b->arena, b->arena_bytes) != hipSuccess ||
        hipMalloc(&b->scan_tmp, std::max<size_t>(scan_bytes, 16)) != hipSuccess ||
        hipMalloc(&b->d_scn, ns * sizeof(BScene)) != hipSuccess)
        return fail(set_err(MPM_E_OOM, "batch allocation failed"));
    if (hipMemsetAsync(b->arena, 0, b->arena_bytes, b->st) != hipSuccess)
        return fail(set_err(MPM_E_HIP, "batch init failed"));
    char *A = b->arena;
    b->sc = reinterpret_cast<Scalars *>(A + o_sc);
    b->sync = reinterpret_cast<BSync *>(A + o_sync);
    b->ctl = reinterpret_cast<BCtl *>(A + o_ctl);
    for (int s = 0; s < ns; ++s) {
        BScene &sd = b->scn[s];
        const int64_t o = b->off[s];
        for (int k = 0; k < 2; ++k) {
            PState &P = sd.P[k];
            P.pos = reinterpret_cast<double *>(A + o_pos[k]) + o;
            P.r = reinterpret_cast<double *>(A + o_r[k]) + o;
            P.idata = reinterpret_cast<int *>(A + o_idata[k]) + o;
            P.id = reinterpret_cast<int64_t *>(A + o_id[k]) + o;
            P.vsnap = nullptr;
            P.np = b->np[s];
            P.cap = cap;
        }
        SortState &S = sd.S;
        S.ntiles = sd.g.ntc[0] * sd.g.ntc[1] * sd.g.ntc[2];
        S.nkeys = S.ntiles * kTileCells * kTileCells * kTileCells;
        S.count = reinterpret_cast<int *>(A + o_count[s]);
        S.cell_start = reinterpret_cast<int *>(A + o_start[s]);
        S.key = reinterpret_cast<int *>(A + o_key) + o;
        S.rank = reinterpret_cast<int *>(A + o_rank) + o;
        S.perm = reinterpret_cast<int *>(A + o_perm) + o;
        S.scan_tmp = b->scan_tmp;
        S.scan_bytes = sort_scan_bytes(S.nkeys);
        S.tile_epoch = reinterpret_cast<int *>(A + o_te[s]);
        S.pkt = reinterpret_cast<double *>(A + o_pkt) + o;
        S.pcap = cap;
        S.tiles = reinterpret_cast<int *>(A + o_tiles[s]);
        const int64_t nn = sd.g.nn;
        double *nod = reinterpret_cast<double *>(A + o_nod[s]);
        sd.N.acc = nod;
        sd.N.mass = nod + 7 * nn;
        sd.N.vel = nod + 8 * nn;
        sd.N.frc = nod + 11 * nn;
        sd.N.dvel = nod + 14 * nn;
        sd.N.mold = nod + 17 * nn;
        sd.sc = b->sc + s;
        const int ntn = sd.g.ntn[0] * sd.g.ntn[1] * sd.g.ntn[2];
        hipLaunchKernelGGL(k_fill_int, dim3((ntn + 255) / 256), dim3(256), 0, b->st, S.tile_epoch,
                           (int64_t)ntn, -1000);
    }
    int rc = check_launch();
    if (rc) return fail(rc);
    rc = batch_occupancy(b);
    if (rc) return fail(rc);
    b->tune.workgroups = b->max_wgs;
    b->tune.sort_chunk = 4096;
    b->tune.tile_chunk = 2048;
    b->tune.node_chunk = 32768;
    b->tune.graph_steps = 10;
    rc = build_queues(b);
    if (rc) return fail(rc);
    *out = b;
    return MPM_OK;
}

void mpm_batch_destroy(mpm_batch *b) {
    if (!b) return;
    if (b->st) (void)hipStreamSynchronize(b->st);
    drop_graph(b);
    (void)hipFree(b->arena);
    (void)hipFree(b->det_arena);
    (void)hipFree(b->scan_tmp);
    (void)hipFree(b->d_scn);
    (void)hipFree(b->d_queue);
    for (int k = 0; k < 2; ++k)
        if (b->ev[k]) (void)hipEventDestroy(b->ev[k]);
    if (b->st) (void)hipStreamDestroy(b->st);
    delete b;
}

int mpm_batch_upload(mpm_batch *b, int s, int64_t np, const double *pos, const double *rdata,
                     const int *idata) {
    if (!b || s < 0 || s >= b->ns || !pos || !rdata || !idata)
        return set_err(MPM_E_ARG, "bad argument");
    if (np != b->np[s]) return set_err(MPM_E_ARG, "batch: particle count differs from create");
    const mpm_specs &sp = b->specs[s];
    for (int64_t i = 0; i < np; ++i) {
        const int m = idata[MPM_I_MODEL * np + i];
        if (m < 0 || m > 2) return set_err(MPM_E_MODEL, "constitutive model must be 0, 1 or 2");
        if (sp.calculate_strain_based_on_delta == 1 && m != 0 && idata[MPM_I_PHASE * np + i] == 0)
            return set_err(MPM_E_UNSUPPORTED,
                           "delta strain update is implemented for model 0 only");
    }
    HIPCHK(hipStreamSynchronize(b->st));
    const PState &P = b->scn[s].P[0];
    const int64_t cap = b->cap;
    for (int d = 0; d < 3; ++d)
        HIPCHK(hipMemcpy(P.pos + d * cap, pos + d * np, np * sizeof(double),
                         hipMemcpyHostToDevice));
    for (int f = 0; f < MPM_NREAL; ++f)
        HIPCHK(hipMemcpy(P.r + f * cap, rdata + f * np, np * sizeof(double),
                         hipMemcpyHostToDevice));
    for (int f = 0; f < MPM_NINT; ++f)
        HIPCHK(hipMemcpy(P.idata + f * cap, idata + f * np, np * sizeof(int),
                         hipMemcpyHostToDevice));
    std::vector<int64_t> ids(np);
    std::iota(ids.begin(), ids.end(), (int64_t)0);
    HIPCHK(hipMemcpy(P.id, ids.data(), np * sizeof(int64_t), hipMemcpyHostToDevice));
    b->cold_radius[s].assign(rdata + (size_t)MPM_R_RADIUS * np,
                             rdata + (size_t)MPM_R_RADIUS * np + np);
    b->cold_yacc[s].assign(rdata + (size_t)MPM_R_YACC * np, rdata + (size_t)MPM_R_YACC * np + np);
    b->cold_rigid[s].assign(idata + (size_t)MPM_I_RIGID_ID * np,
                            idata + (size_t)MPM_I_RIGID_ID * np + np);
    // InitParticles ends with Redistribute() (mpm_init.cpp:1166).
    launch_redistribute(b->scn[s].g, P, b->st);
    int rc = check_launch();
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(b->st));
    b->uploaded[s] = 1;
    b->inited = false;
    return MPM_OK;
}

int mpm_batch_init(mpm_batch *b) {
    if (!b) return set_err(MPM_E_ARG, "null batch");
    if (b->steps != 0) return set_err(MPM_E_ARG, "batch: init before the first step");
    for (int s = 0; s < b->ns; ++s)
        if (!b->uploaded[s]) return set_err(MPM_E_ARG, "batch: a scene has no particles");
    LsP ls;
    std::memset(&ls, 0, sizeof(ls));
    UdfP udf;
    std::memset(&udf, 0, sizeof(udf));
    for (int s = 0; s < b->ns; ++s) {
        BScene &sd = b->scn[s];
        SortState S = sd.S;
        // Initialise_Internal_Forces (mpm_init.cpp:368-401) as mpm_ctx runs it,
        // then Calculate_time_step for the first step.  Deterministic: the
        // single-scene deterministic sort and P2G, their block stamps below
        // every step number of the persistent kernel.
        S.pass = -(1 << 30);
        launch_sort(sd.g, sd.ph, sd.P[0], S, nullptr, b->det ? 0 : 1, b->st);
        launch_p2g_tiles(sd.g, sd.ph, sd.P[0], S, sd.N, 0, 0, 1, sd.sc, b->st);
        launch_grid(sd.g, sd.ph, ls, udf, sd.N,
                    G_FINISH_MASS | G_FINISH_VEL | G_FINISH_FRC | G_KEEP_FRC | G_BACKUP, 0.0,
                    nullptr, b->st);
        launch_particle(sd.g, sd.ph, ls, udf, sd.P[0], sd.N, P_G2P_GRAD | P_STRESS_INIT, 0.0, 0.0,
                        nullptr, b->st);
        launch_dt(sd.g, sd.ph, sd.P[0], sd.sc, sd.fixed, sd.fixed_dt, b->st, 0, 0);
        int rc = check_launch();
        if (rc) return rc;
    }
    HIPCHK(hipStreamSynchronize(b->st));
    int rc = check_flags(b);
    if (rc) return rc;
    b->inited = true;
    return MPM_OK;
}

int mpm_batch_step(mpm_batch *b, int nsteps) {
    if (!b || nsteps < 0) return set_err(MPM_E_ARG, "bad argument");
    if (!b->inited) return set_err(MPM_E_ARG, "batch: mpm_batch_init first");
    if (nsteps == 0) return MPM_OK;
    const int G = b->tune.graph_steps;
    if (nsteps >= G && (!b->gexec || b->gsteps != G)) {
        drop_graph(b);
        HIPCHK(hipStreamBeginCapture(b->st, hipStreamCaptureModeThreadLocal));
        int rc = launch_steps(b, G);
        hipGraph_t gr = nullptr;
        hipError_t e = hipStreamEndCapture(b->st, &gr);
        if (rc) return rc;
        if (e != hipSuccess) return hip_err(e, "hipStreamEndCapture");
        b->graph = gr;
        HIPCHK(hipGraphInstantiate(&b->gexec, b->graph, nullptr, nullptr, 0));
        b->gsteps = G;
    }
    HIPCHK(hipEventRecord(b->ev[0], b->st));
    int done = 0;
    for (; done + G <= nsteps; done += G) HIPCHK(hipGraphLaunch(b->gexec, b->st));
    int rc = launch_steps(b, nsteps - done);
    if (rc) return rc;
    HIPCHK(hipEventRecord(b->ev[1], b->st));
    HIPCHK(hipEventSynchronize(b->ev[1]));
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, b->ev[0], b->ev[1]);
    b->step_ms = ms / nsteps;
    b->steps += nsteps;
    return check_flags(b);
}

int mpm_batch_num_particles(mpm_batch *b, int s, int64_t *np) {
    if (!b || s < 0 || s >= b->ns || !np) return set_err(MPM_E_ARG, "bad argument");
    HIPCHK(hipStreamSynchronize(b->st));
    const PState &P = b->scn[s].P[b->steps & 1];
    std::vector<int64_t> ids(b->np[s]);
    HIPCHK(hipMemcpy(ids.data(), P.id, ids.size() * sizeof(int64_t), hipMemcpyDeviceToHost));
    int64_t n = 0;
    for (int64_t v : ids) n += (v >= 0);
    *np = n;
    return MPM_OK;
}

int mpm_batch_download(mpm_batch *b, int s, double *pos, double *rdata, int *idata,
                       int64_t *id) {
    if (!b || s < 0 || s >= b->ns) return set_err(MPM_E_ARG, "bad argument");
    HIPCHK(hipStreamSynchronize(b->st));
    const PState &P = b->scn[s].P[b->steps & 1];
    const int64_t n = b->np[s], cap = b->cap;
    std::vector<int64_t> ids(n);
    HIPCHK(hipMemcpy(ids.data(), P.id, n * sizeof(int64_t), hipMemcpyDeviceToHost));
    std::vector<int64_t> order;
    order.reserve(n);
    for (int64_t i = 0; i < n; ++i)
        if (ids[i] >= 0) order.push_back(i);
    std::sort(order.begin(), order.end(), [&](int64_t x, int64_t y) { return ids[x] < ids[y]; });
    const int64_t nl = (int64_t)order.size();
    std::vector<double> h(n);
    if (pos)
        for (int d = 0; d < 3; ++d) {
            HIPCHK(hipMemcpy(h.data(), P.pos + d * cap, n * sizeof(double),
                             hipMemcpyDeviceToHost));
            for (int64_t k = 0; k < nl; ++k) pos[d * nl + k] = h[order[k]];
        }
    if (rdata) {
        for (int f = 0; f < MPM_NREAL; ++f) {
            HIPCHK(hipMemcpy(h.data(), P.r + f * cap, n * sizeof(double), hipMemcpyDeviceToHost));
            for (int64_t k = 0; k < nl; ++k) rdata[f * nl + k] = h[order[k]];
        }
        for (int64_t k = 0; k < nl; ++k) {
            rdata[MPM_R_RADIUS * nl + k] = b->cold_radius[s][ids[order[k]]];
            rdata[MPM_R_YACC * nl + k] = b->cold_yacc[s][ids[order[k]]];
        }
    }
    if (idata) {
        std::vector<int> hi(n);
        for (int f = 0; f < MPM_NINT; ++f) {
            HIPCHK(hipMemcpy(hi.data(), P.idata + f * cap, n * sizeof(int),
                             hipMemcpyDeviceToHost));
            for (int64_t k = 0; k < nl; ++k) idata[f * nl + k] = hi[order[k]];
        }
        for (int64_t k = 0; k < nl; ++k)
            idata[MPM_I_RIGID_ID * nl + k] = b->cold_rigid[s][ids[order[k]]];
    }
    if (id)
        for (int64_t k = 0; k < nl; ++k) id[k] = ids[order[k]];
    return MPM_OK;
}

int mpm_batch_get_nodal(mpm_batch *b, int s, int comp, double *out) {
    if (!b || s < 0 || s >= b->ns || !out) return set_err(MPM_E_ARG, "bad argument");
    if (comp < 0 || comp >= MPM_N_NCOMP) return set_err(MPM_E_ARG, "bad nodal component");
    HIPCHK(hipStreamSynchronize(b->st));
    const NState &N = b->scn[s].N;
    const int64_t nn = b->scn[s].g.nn;
    const double *src = nullptr;
    if (comp == MPM_N_MASS) src = N.mass;
    else if (comp >= MPM_N_VEL && comp < MPM_N_VEL + 3) src = N.vel + (comp - MPM_N_VEL) * nn;
    else if (comp >= MPM_N_DVEL && comp < MPM_N_DVEL + 3) src = N.dvel + (comp - MPM_N_DVEL) * nn;
    else if (comp == MPM_N_MASS_OLD) src = N.mold;
    if (!src) {
        std::memset(out, 0, nn * sizeof(double));
        return MPM_OK;
    }
    HIPCHK(hipMemcpy(out, src, nn * sizeof(double), hipMemcpyDeviceToHost));
    return MPM_OK;
}

int mpm_batch_get_time(mpm_batch *b, int s, double *time, int64_t *steps, double *last_dt) {
    if (!b || s < 0 || s >= b->ns) return set_err(MPM_E_ARG, "bad argument");
    HIPCHK(hipStreamSynchronize(b->st));
    Scalars h;
    HIPCHK(hipMemcpy(&h, b->sc + s, sizeof(Scalars), hipMemcpyDeviceToHost));
    BSync y;
    HIPCHK(hipMemcpy(&y, b->sync + s, sizeof(BSync), hipMemcpyDeviceToHost));
    if (time) *time = h.time;
    if (steps) *steps = b->steps;
    if (last_dt) *last_dt = y.last_dt;
    return MPM_OK;
}

int mpm_batch_get_tuning(mpm_batch *b, mpm_batch_tuning *t) {
    if (!b || !t) return set_err(MPM_E_ARG, "bad argument");
    *t = b->tune;
    return MPM_OK;
}

int mpm_batch_set_tuning(mpm_batch *b, const mpm_batch_tuning *t) {
    if (!b || !t) return set_err(MPM_E_ARG, "bad argument");
    mpm_batch_tuning v = *t;
    int rc = valid_tuning(b, v);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(b->st));
    b->tune = v;
    return build_queues(b);
}

int mpm_batch_autotune(mpm_batch *b, int trial_steps, mpm_batch_tuning *best_out,
                       double *best_ms) {
    if (!b || trial_steps < 1) return set_err(MPM_E_ARG, "bad argument");
    if (!b->inited) return set_err(MPM_E_ARG, "batch: mpm_batch_init first");
    HIPCHK(hipStreamSynchronize(b->st));
    char *snap = nullptr;
    HIPCHK(hipMalloc(&snap, b->arena_bytes));
    auto done = [&](int rc) {
        (void)hipStreamSynchronize(b->st);
        (void)hipFree(snap);
        return rc;
    };
    if (hipMemcpyAsync(snap, b->arena, b->arena_bytes, hipMemcpyDeviceToDevice, b->st) !=
        hipSuccess)
        return done(set_err(MPM_E_HIP, "autotune snapshot failed"));
    auto trial = [&](const mpm_batch_tuning &t, double &ms) -> int {
        b->tune = t;
        int rc = build_queues(b);
        if (rc) return rc;
        HIPCHK(hipMemcpyAsync(b->arena, snap, b->arena_bytes, hipMemcpyDeviceToDevice, b->st));
        rc = clear_det_stamps(b);
        if (rc) return rc;
        rc = launch_steps(b, 1);  // warm-up
        if (rc) return rc;
        HIPCHK(hipEventRecord(b->ev[0], b->st));
        rc = launch_steps(b, trial_steps);
        if (rc) return rc;
        HIPCHK(hipEventRecord(b->ev[1], b->st));
        HIPCHK(hipEventSynchronize(b->ev[1]));
        float f = 0.0f;
        (void)hipEventElapsedTime(&f, b->ev[0], b->ev[1]);
        ms = f / trial_steps;
        return check_flags(b);
    };
    mpm_batch_tuning best = b->tune;
    double bms = 0.0;
    int rc = trial(best, bms);
    if (rc) return done(rc);
    std::vector<int> wg;
    for (int f : {1, 2, 3, 4}) wg.push_back(std::max(kQueues, b->max_wgs * f / 4 / kQueues * kQueues));
    wg.push_back(b->max_wgs * 2);
    const std::vector<int> sc = {1024, 2048, 4096, 8192, 16384, 32768};
    const std::vector<int> tc = {512, 1024, 2048, 4096, 8192, 16384};
    const std::vector<int> nc = {4096, 16384, 65536, 262144};
    for (int pass = 0; pass < 2; ++pass)
        for (int p = 0; p < 4; ++p) {
            const std::vector<int> &vals = (p == 0) ? tc : (p == 1 ? sc : (p == 2 ? nc : wg));
            for (int v : vals) {
                mpm_batch_tuning t = best;
                int &field = (p == 0) ? t.tile_chunk
                                      : (p == 1 ? t.sort_chunk : (p == 2 ? t.node_chunk : t.workgroups));
                if (field == v) continue;
                field = v;
                double ms = 0.0;
                rc = trial(t, ms);
                if (rc) return done(rc);
                if (ms < bms * 0.99) {
                    bms = ms;
                    best = t;
                }
            }
        }
    b->tune = best;
    rc = build_queues(b);
    if (rc) return done(rc);
    if (hipMemcpyAsync(b->arena, snap, b->arena_bytes, hipMemcpyDeviceToDevice, b->st) !=
        hipSuccess)
        return done(set_err(MPM_E_HIP, "autotune restore failed"));
    rc = clear_det_stamps(b);
    if (rc) return done(rc);
    if (best_out) *best_out = best;
    if (best_ms) *best_ms = bms;
    return done(MPM_OK);
}

int mpm_batch_profile(mpm_batch *b, int nsteps, double busy_ms[7], double wait_ms[7],
                      int64_t tasks[7], double *other_ms, double *step_ms) {
    if (!b || nsteps < 1) return set_err(MPM_E_ARG, "bad argument");
    if (!b->inited) return set_err(MPM_E_ARG, "batch: mpm_batch_init first");
    int khz = 0;
    HIPCHK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, 0));
    if (khz <= 0) return set_err(MPM_E_HIP, "no wall clock rate");
    const size_t off = offsetof(BCtl, busy);
    HIPCHK(hipMemsetAsync(reinterpret_cast<char *>(b->ctl) + off, 0, sizeof(BCtl) - off, b->st));
    HIPCHK(hipEventRecord(b->ev[0], b->st));
    int rc = launch_steps(b, nsteps, 1);
    if (rc) return rc;
    HIPCHK(hipEventRecord(b->ev[1], b->st));
    HIPCHK(hipEventSynchronize(b->ev[1]));
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, b->ev[0], b->ev[1]);
    b->steps += nsteps;
    BCtl h;
    HIPCHK(hipMemcpy(&h, b->ctl, sizeof(BCtl), hipMemcpyDeviceToHost));
    const double tick_ms = 1.0 / khz;
    for (int k = 0; k < B_NSTAGE; ++k) {
        if (busy_ms) busy_ms[k] = h.busy[k] * tick_ms / nsteps;
        if (wait_ms) wait_ms[k] = h.wait[k] * tick_ms / nsteps;
        if (tasks) tasks[k] = (int64_t)(h.ntasks[k] / nsteps);
    }
    if (other_ms) *other_ms = h.other * tick_ms / nsteps;
    if (step_ms) *step_ms = ms / nsteps;
    return check_flags(b);
}

int mpm_batch_set_option(mpm_batch *b, int option, int value) {
    if (!b) return set_err(MPM_E_ARG, "null batch");
    if (option != MPM_OPT_DETERMINISTIC)
        return set_err(MPM_E_ARG, "batch: the only option is MPM_OPT_DETERMINISTIC");
    if (b->steps != 0 || b->inited)
        return set_err(MPM_E_ARG, "batch: set options before mpm_batch_init");
    HIPCHK(hipStreamSynchronize(b->st));
    if (value && !b->det_arena) {
        // per scene: one 7 x 343-double node block per cell tile (the list
        // position of a non-empty tile selects it) and the tile stamps
        size_t at = 0;
        std::vector<size_t> o_blk(b->ns);
        b->det_slot_off.assign(b->ns, 0);
        for (int s = 0; s < b->ns; ++s) {
            const size_t nt = (size_t)b->scn[s].S.ntiles;
            o_blk[s] = at;
            at = align256(at + nt * 7 * kTNNodes * sizeof(double));
            b->det_slot_off[s] = at;
            at = align256(at + nt * sizeof(int64_t));
        }
        if (hipMalloc(&b->det_arena, at) != hipSuccess) {
            b->det_arena = nullptr;
            return set_err(MPM_E_OOM, "batch: deterministic node blocks");
        }
        for (int s = 0; s < b->ns; ++s) {
            SortState &S = b->scn[s].S;
            S.blk = reinterpret_cast<double *>(b->det_arena + o_blk[s]);
            S.tile_slot = reinterpret_cast<int64_t *>(b->det_arena + b->det_slot_off[s]);
        }
        int rc = clear_det_stamps(b);
        if (rc) return rc;
    }
    b->det = value ? 1 : 0;
    int rc = batch_occupancy(b);
    if (rc) return rc;
    if (b->tune.workgroups > b->max_wgs) b->tune.workgroups = b->max_wgs;
    return build_queues(b);
}

int mpm_batch_stats(mpm_batch *b, double *step_ms, int64_t *particles, int *tasks) {
    if (!b) return set_err(MPM_E_ARG, "null batch");
    if (step_ms) *step_ms = b->step_ms;
    if (particles) {
        int64_t n = 0;
        for (int64_t v : b->np) n += v;
        *particles = n;
    }
    if (tasks) *tasks = b->ntasks;
    return MPM_OK;
}

}  // extern "C"

// mpm_learn.hip — the learned coarse-step correction behind the C ABI:
// network and dataset objects, training-row collection from the plain
// solver, training, and the learned step with its guards.
//
// A learned step is the specialised USL step (sort, packed P2G, nodal
// update, fused particle kernel) run once with nfine x the CFL dt, with the
// network's velocity correction added to the nodal velocity (and to the
// velocity increment the FLIP blend uses) between the nodal update and the
// particle kernel.  The guards look at the corrected nodal field only, so a
// rejected coarse step has not touched the particles: it is dropped and the
// plain steps run from the same state.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <random>
#include <vector>

#include "mpm_context.h"
#include "mpm_mlp.h"

using namespace xmpm;

struct mpm_mlp {
    MlpDev d{};
    MlpScales sc{1.0, 1.0, 1.0};
    hipStream_t stream = nullptr;
    __bf16 *x = nullptr;  // host-row staging
    float *t = nullptr;
    int *idx = nullptr;
    float *y = nullptr;
    int64_t cap = 0;
    int adam_t = 0;  // Adam steps taken so far
};

struct mpm_dataset {
    __bf16 *x = nullptr;  // [cap][128]
    float *t = nullptr;   // [cap][3], m/s
    int *count = nullptr; // device row count
    int64_t cap = 0;
};

namespace {

constexpr int kMlCounts = 16;  // ints of mpm_ctx::ml_count

// float -> bfloat16 bits, round to nearest even (the conversion the device
// applies to the same values).
uint16_t bf16_bits(float x) {
    uint32_t u;
    std::memcpy(&u, &x, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) {
        if (u & 0x007fffffu) u |= 0x00400000u;
        return (uint16_t)(u >> 16);
    }
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

int mlp_stage(mpm_mlp *m, int64_t rows) {
    const int64_t pad = (rows + kMRows - 1) / kMRows * kMRows;
    if (pad <= m->cap) return MPM_OK;
    (void)hipFree(m->x);
    (void)hipFree(m->t);
    (void)hipFree(m->idx);
    (void)hipFree(m->y);
    m->x = nullptr;
    m->t = nullptr;
    m->idx = nullptr;
    m->y = nullptr;
    m->cap = 0;
    HIPCHK(hipMalloc(&m->x, pad * kMInPad * sizeof(__bf16)));
    HIPCHK(hipMalloc(&m->t, pad * kMOut * sizeof(float)));
    HIPCHK(hipMalloc(&m->idx, pad * sizeof(int)));
    HIPCHK(hipMalloc(&m->y, pad * kMOut * sizeof(float)));
    m->cap = pad;
    return MPM_OK;
}

// Host rows -> BF16 rows on the device (padding rows zero).
int upload_rows(mpm_mlp *m, int64_t nrows, const float *X) {
    const int64_t pad = (nrows + kMRows - 1) / kMRows * kMRows;
    std::vector<uint16_t> h(pad * kMInPad, 0);
    for (int64_t i = 0; i < nrows * kMInPad; ++i) h[i] = bf16_bits(X[i]);
    HIPCHK(hipMemcpyAsync(m->x, h.data(), h.size() * 2, hipMemcpyHostToDevice, m->stream));
    HIPCHK(hipStreamSynchronize(m->stream));
    return MPM_OK;
}

int ctx_ml_buffers(mpm_ctx *c) {
    const int64_t nn = c->g.nn;
    const int64_t pad = (nn + kMRows - 1) / kMRows * kMRows;
    if (c->ml_cap >= pad) return MPM_OK;
    HIPCHK(hipMalloc(&c->ml_nodes, 2 * pad * sizeof(int)));  // active list, selected rows
    // [0] active count, [1] selected, [2] appended, [3] guard flag, [8] the
    // other step parity's active count
    HIPCHK(hipMalloc(&c->ml_count, kMlCounts * sizeof(int)));
    HIPCHK(hipMemset(c->ml_count, 0, kMlCounts * sizeof(int)));
    HIPCHK(hipMalloc(&c->ml_x, pad * kMInPad * sizeof(__bf16)));
    HIPCHK(hipMalloc(&c->ml_feat, nn * 8));
    HIPCHK(hipMalloc(&c->ml_y, pad * kMOut * sizeof(float)));
    HIPCHK(hipMalloc(&c->ml_vc, pad * 3 * sizeof(double)));
    HIPCHK(hipMalloc(&c->ml_guard, kInferBlocks * 8 * sizeof(double)));
    // the per-step decision is read into pinned memory: a copy into pageable
    // memory goes through a staging buffer
    if (!c->ml_hflag) HIPCHK(hipHostMalloc(&c->ml_hflag, 4 * sizeof(int), hipHostMallocDefault));
    if (!c->ml_scbak) HIPCHK(hipMalloc(&c->ml_scbak, sizeof(Scalars)));
    for (int q = 0; q < 2; ++q)
        if (!c->ml_ev[q]) HIPCHK(hipEventCreateWithFlags(&c->ml_ev[q], hipEventDisableTiming));
    c->ml_cap = pad;
    return MPM_OK;
}

// The third particle buffer of the speculative learned step (the size of P;
// freed and re-made with the particles).
int ctx_spec_buffer(mpm_ctx *c) {
    if (c->R.pos && c->R.cap == c->P.cap) return MPM_OK;
    (void)hipFree(c->R.pos);
    (void)hipFree(c->R.r);
    (void)hipFree(c->R.idata);
    (void)hipFree(c->R.id);
    c->R = PState();
    const int64_t cap = c->P.cap;
    HIPCHK(hipMalloc(&c->R.pos, 3 * cap * sizeof(double)));
    HIPCHK(hipMalloc(&c->R.r, (int64_t)MPM_NREAL * cap * sizeof(double)));
    HIPCHK(hipMalloc(&c->R.idata, (int64_t)MPM_NINT * cap * sizeof(int)));
    HIPCHK(hipMalloc(&c->R.id, cap * sizeof(int64_t)));
    c->R.cap = cap;
    c->R.np = c->P.np;
    c->R.vsnap = c->P.vsnap;
    return MPM_OK;
}

// The learned path runs the specialised USL step only.
int learned_ok(const mpm_ctx *c) {
    if (c->s.stress_update_scheme != MPM_USL)
        return set_err(MPM_E_UNSUPPORTED, "the learned step runs the USL scheme only");
    const int nr = c->s.num_redist > 0 ? c->s.num_redist : 1;
    if (!c->sorted || !c->fast || nr != 1 || !c->g.uniform || c->ls.nb > 0 || c->any_udf)
        return set_err(MPM_E_UNSUPPORTED,
                       "the learned step needs the specialised sorted step (redistribution "
                       "every step, uniform order, no level set or wall UDF)");
    if (c->thermal)
        return set_err(MPM_E_UNSUPPORTED, "the learned step has no thermal stages");
    if (c->P.np <= 0) return set_err(MPM_E_ARG, "no particles");
    return MPM_OK;
}

// Coarse step up to the nodal update: dt = nfine x the CFL dt, sort, P2G,
// nodal update + BCs.  Leaves the step counter advanced.  advance: the
// previous step's time advance is still due (it is folded into the dt
// launch).
// bak: copy the step scalars there before the dt launch changes them.
// Step parameters of a coarse step's particle update: with
// MPM_OPT_COARSE_BLEND the PIC/FLIP blend is alpha^nfine, the blend of the
// nfine fine steps the coarse step stands for (each fine step keeps the
// fraction alpha of the particle's own velocity deviation from the grid
// field; the blend is mpm_particle_grid_ops.cpp:1030-1034).
PhysP coarse_ph(const mpm_ctx *c, int nfine) {
    PhysP p = c->ph;
    if (c->coarse_blend) {
        double a = 1.0;
        for (int k = 0; k < nfine; ++k) a *= c->ph.alpha;
        p.alpha = a;
    }
    return p;
}

int coarse_begin(mpm_ctx *c, int nfine, int advance = 0, const PState *in = nullptr,
                 const PState *out = nullptr, Scalars *bak = nullptr) {
    hipStream_t st = c->stream;
    const PState &Pin = in ? *in : c->P;
    const PState &Pout = out ? *out : c->Q;
    c->steps++;
    const int step = (int)c->steps;
    const bool fixed_dt = c->s.fixed_timestep == 1;
    DtFold fold;  // the dt in the sort's first launch when it needs no reduction
    if (Pin.np > 0 && (fixed_dt || c->dt_next_valid)) {
        fold.on = 1;
        fold.fixed = fixed_dt ? 1 : 0;
        fold.fixed_dt = c->s.timestep;
        fold.advance = advance;
        fold.scale = (double)nfine;
        fold.sc = c->sc;
        fold.bak = bak;
    } else {
        if (bak)
            HIPCHK(hipMemcpyAsync(bak, c->sc, sizeof(Scalars), hipMemcpyDeviceToDevice, st));
        launch_dt(c->g, c->ph, Pin, c->sc, fixed_dt, c->s.timestep, st,
                  c->dt_next_valid ? 1 : 0, advance, (double)nfine);
    }
    det_sync(c);
    launch_sort(c->g, c->ph, Pin, c->S, Pout.id, 0, st, fold);
    launch_p2g_tiles(c->g, c->ph, Pin, c->S, c->N, step, 0, 0, c->sc, st);
    const int *te = (c->skip_ok && !c->all_dirty) ? c->S.tile_epoch : nullptr;
    launch_grid(c->g, c->ph, c->ls, c->udf, c->N, grid_flags_full(c), 0.0, c->sc, st, te, step);
    c->all_dirty = false;
    return check_launch();
}

// The particle kernel of the coarse step and the time advance.
int coarse_finish(mpm_ctx *c, int nfine) {
    hipStream_t st = c->stream;
    launch_particle_fast(c->g, coarse_ph(c, nfine), c->ls, c->udf, c->P, c->Q, c->S, c->N, 1, 1,
                         0, c->sc, st);
    std::swap(c->P, c->Q);
    c->dt_next_valid = c->fuse_dt != 0;
    c->dirty = false;
    launch_advance_time(c->sc, st);
    return check_launch();
}

// Drop a coarse step begun by coarse_begin (particles untouched).
void coarse_drop(mpm_ctx *c) {
    c->steps--;
    c->dt_next_valid = false;  // its dt launch consumed the fused candidates
}

int active_nodes(mpm_ctx *c, int *n) {
    launch_active_nodes(c->g, c->N, c->ph.mass_tol, c->ml_nodes, c->ml_count, nullptr, c->stream);
    HIPCHK(hipMemcpyAsync(n, c->ml_count, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return check_launch();
}

int finish_call(mpm_ctx *c) {
    HIPCHK(hipStreamSynchronize(c->stream));
    int nlive = 0;
    HIPCHK(hipMemcpy(&nlive, c->S.cell_start + c->S.nkeys, sizeof(int), hipMemcpyDeviceToHost));
    c->P.np = c->Q.np = nlive;
    Scalars h;
    HIPCHK(hipMemcpy(&h, c->sc, sizeof(Scalars), hipMemcpyDeviceToHost));
    c->time_host = h.time;
    c->last_dt = h.dt;
    if (h.err) {
        int zero = 0;
        HIPCHK(hipMemcpy(&c->sc->err, &zero, sizeof(int), hipMemcpyHostToDevice));
        return set_err(MPM_E_UNSUPPORTED, "a particle lay outside its sorted tile");
    }
    return MPM_OK;
}

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

// Keep active node r with probability p (hash of seed and node): compact
// the kept nodes to the front of sel, their coarse velocity into vc.
__global__ void k_select_rows(const int *active, int nact, double p, uint64_t seed, NState N,
                              int64_t nn, int *sel, double *vc, int *count) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nact) return;
    const int q = active[r];
    const double u = (double)(mix64(seed ^ (uint64_t)q) >> 11) * 0x1.0p-53;
    if (u >= p) return;
    const int k = atomicAdd(count, 1);
    sel[k] = q;
    for (int d = 0; d < 3; ++d) vc[3 * k + d] = N.vel[d * nn + q];
}

// Rows whose node is still active after the fine steps: append the feature
// row and target = fine - coarse nodal velocity.
__global__ void k_append_rows(const int *sel, int nsel, const __bf16 *X, const double *vc,
                              NState N, int64_t nn, double tol, mpm_dataset d) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nsel) return;
    const int q = sel[r];
    if (!(N.mass[q] >= tol)) return;
    const int k = atomicAdd(d.count, 1);
    if (k >= d.cap) return;
    const uint4 *src = reinterpret_cast<const uint4 *>(X + (int64_t)r * kMInPad);
    uint4 *dst = reinterpret_cast<uint4 *>(d.x + (int64_t)k * kMInPad);
    for (int e = 0; e < kMInPad * 2 / 16; ++e) dst[e] = src[e];
    for (int c = 0; c < 3; ++c) d.t[3 * k + c] = (float)(N.vel[c * nn + q] - vc[3 * r + c]);
}

// Guard decision (see mpm_step_learned): flag = 1 keeps the corrected coarse
// step.  One workgroup sums the inference's kInferBlocks partial rows in a
// fixed order (thread t: rows t, t + 256, ...; then a fixed tree), so equal
// states give equal decisions.  hflag (pinned host memory, or null): the
// decision is also stored there.
__global__ __launch_bounds__(256) void k_guard_decide(const double *part, const Scalars *sc,
                                                      double gmag, double gp, double ge, int *flag,
                                                      int *hflag) {
    __shared__ double red[7][256];
    const int t = threadIdx.x;
    for (int k = 0; k < 7; ++k) {
        double x = 0.0;
        for (int b = t; b < kInferBlocks; b += 256) x += part[b * 8 + k];
        red[k][t] = x;
    }
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (t < h)
            for (int k = 0; k < 7; ++k) red[k][t] += red[k][t + h];
        __syncthreads();
    }
    if (t != 0) return;
    double gs[7];
    for (int k = 0; k < 7; ++k) gs[k] = red[k][0];
    const double dt = sc->dt;
    const double M = gs[0];
    const double u = M > 0 ? sqrt(gs[4] / M) : 0.0;
    const double vs = fmax(u, gmag * dt);
    const double dp = sqrt(gs[1] * gs[1] + gs[2] * gs[2] + gs[3] * gs[3]);
    const double unew = M > 0 ? sqrt(gs[5] / M) : 0.0;
    const bool ok = gs[6] == 0.0 && isfinite(unew) && dp <= gp * M * vs + 1e-300 &&
                    unew <= (1.0 + ge) * vs + 1e-300;
    flag[0] = ok ? 1 : 0;
    if (hflag) {
        *hflag = ok ? 1 : 0;
        __threadfence_system();
    }
}

// Add the correction to the nodal velocity and to the increment used by
// FLIP, when the guard kept it.
__global__ void k_apply_dev(const int *nodes, const int *count, const float *Y, double dv_ref,
                            NState N, int64_t nn, const int *flag) {
    if (!*flag) return;
    const int n = *count;
    for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x) {
        const int q = nodes[r];
        for (int d = 0; d < 3; ++d) {
            const double dv = (double)Y[3 * r + d] * dv_ref;
            N.vel[d * nn + q] += dv;
            N.dvel[d * nn + q] += dv;
        }
    }
}

__global__ void k_zero_d(double *p, int n) {
    const int i = threadIdx.x;
    if (i < n) p[i] = 0.0;
}

__global__ void k_zero_f(float *p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0f;
}

unsigned blocks256(int64_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

extern "C" {

int mpm_mlp_create(mpm_mlp **out) {
    if (!out) return set_err(MPM_E_ARG, "null argument");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return set_err(MPM_E_NODEVICE, "no HIP device");
    mpm_mlp *m = new mpm_mlp();
    auto fail = [&](int rc) {
        mpm_mlp_destroy(m);
        return rc;
    };
    MlpDev &d = m->d;
    if (hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess ||
        hipMalloc(&d.params, kMParams * sizeof(float)) != hipSuccess ||
        hipMalloc(&d.m, kMParams * sizeof(float)) != hipSuccess ||
        hipMalloc(&d.v, kMParams * sizeof(float)) != hipSuccess ||
        hipMalloc(&d.grad, kMParams * sizeof(float)) != hipSuccess ||
        hipMalloc(&d.part, (size_t)kMTrainParts * kMPartStride * sizeof(float)) != hipSuccess ||
        hipMalloc(&d.w1t, kMH * kMInPad * sizeof(__bf16)) != hipSuccess ||
        hipMalloc(&d.w2tp, kMH * kMH * sizeof(__bf16)) != hipSuccess ||
        hipMalloc(&d.w2p, kMH * kMH * sizeof(__bf16)) != hipSuccess ||
        hipMalloc(&d.w3tp, 32 * kMH * sizeof(__bf16)) != hipSuccess ||
        hipMalloc(&d.w3, kMH * 16 * sizeof(__bf16)) != hipSuccess ||
        hipMalloc(&d.loss, sizeof(double)) != hipSuccess)
        return fail(set_err(MPM_E_OOM, "network allocation failed"));
    // zero everything once: padding lanes of the operand copies stay zero
    hipLaunchKernelGGL(k_zero_f, dim3(blocks256(kMParams)), dim3(256), 0, m->stream, d.params,
                       kMParams);
    hipLaunchKernelGGL(k_zero_f, dim3(blocks256(kMParams)), dim3(256), 0, m->stream, d.m, kMParams);
    hipLaunchKernelGGL(k_zero_f, dim3(blocks256(kMParams)), dim3(256), 0, m->stream, d.v, kMParams);
    hipLaunchKernelGGL(k_zero_f, dim3(blocks256(kMParams)), dim3(256), 0, m->stream, d.grad,
                       kMParams);
    hipLaunchKernelGGL(k_zero_f, dim3(blocks256(kMH * kMInPad / 2)), dim3(256), 0, m->stream,
                       reinterpret_cast<float *>(d.w1t), kMH * kMInPad / 2);
    hipLaunchKernelGGL(k_zero_f, dim3(blocks256(kMH * kMH / 2)), dim3(256), 0, m->stream,
                       reinterpret_cast<float *>(d.w2tp), kMH * kMH / 2);
    hipLaunchKernelGGL(k_zero_f, dim3(blocks256(kMH * kMH / 2)), dim3(256), 0, m->stream,
                       reinterpret_cast<float *>(d.w2p), kMH * kMH / 2);
    hipLaunchKernelGGL(k_zero_f, dim3(blocks256(32 * kMH / 2)), dim3(256), 0, m->stream,
                       reinterpret_cast<float *>(d.w3tp), 32 * kMH / 2);
    hipLaunchKernelGGL(k_zero_f, dim3(blocks256(kMH * 16 / 2)), dim3(256), 0, m->stream,
                       reinterpret_cast<float *>(d.w3), kMH * 16 / 2);
    hipLaunchKernelGGL(k_zero_d, dim3(1), dim3(64), 0, m->stream, d.loss, 1);
    if (hipStreamSynchronize(m->stream) != hipSuccess)
        return fail(set_err(MPM_E_HIP, "network init failed"));
    *out = m;
    return MPM_OK;
}

void mpm_mlp_destroy(mpm_mlp *m) {
    if (!m) return;
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    MlpDev &d = m->d;
    for (void *p : {(void *)d.params, (void *)d.m, (void *)d.v, (void *)d.grad, (void *)d.part,
                    (void *)d.w1t, (void *)d.w2tp, (void *)d.w2p, (void *)d.w3tp, (void *)d.w3,
                    (void *)d.loss,
                    (void *)m->x, (void *)m->t, (void *)m->idx, (void *)m->y})
        (void)hipFree(p);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
}

int mpm_mlp_set_params(mpm_mlp *m, const float *params, const double scales[3]) {
    if (!m || !params || !scales) return set_err(MPM_E_ARG, "null argument");
    if (!(scales[0] > 0 && scales[1] > 0 && scales[2] > 0))
        return set_err(MPM_E_ARG, "scales must be positive");
    m->sc = MlpScales{scales[0], scales[1], scales[2]};
    HIPCHK(hipMemcpyAsync(m->d.params, params, kMParams * sizeof(float), hipMemcpyHostToDevice,
                          m->stream));
    launch_mlp_pack(m->d, m->stream);
    HIPCHK(hipStreamSynchronize(m->stream));
    return check_launch();
}

int mpm_mlp_get_params(mpm_mlp *m, float *params, double scales[3]) {
    if (!m) return set_err(MPM_E_ARG, "null argument");
    if (params) {
        HIPCHK(hipMemcpyAsync(params, m->d.params, kMParams * sizeof(float),
                              hipMemcpyDeviceToHost, m->stream));
        HIPCHK(hipStreamSynchronize(m->stream));
    }
    if (scales) {
        scales[0] = m->sc.m_ref;
        scales[1] = m->sc.v_ref;
        scales[2] = m->sc.dv_ref;
    }
    return MPM_OK;
}

int mpm_mlp_forward(mpm_mlp *m, int64_t nrows, const float *X, float *Y) {
    if (!m || nrows < 0 || (nrows && (!X || !Y))) return set_err(MPM_E_ARG, "bad argument");
    if (nrows == 0) return MPM_OK;
    if (nrows > (1 << 30)) return set_err(MPM_E_ARG, "too many rows");
    int rc = mlp_stage(m, nrows);
    if (rc) return rc;
    if ((rc = upload_rows(m, nrows, X))) return rc;
    launch_mlp_infer(m->d, m->x, (int)nrows, m->y, m->stream);
    if ((rc = check_launch())) return rc;
    HIPCHK(hipMemcpyAsync(Y, m->y, nrows * kMOut * sizeof(float), hipMemcpyDeviceToHost,
                          m->stream));
    HIPCHK(hipStreamSynchronize(m->stream));
    return MPM_OK;
}

int mpm_mlp_gradient(mpm_mlp *m, int64_t nrows, const float *X, const float *T, float *grad,
                     double *loss) {
    if (!m || nrows <= 0 || !X || !T || !grad) return set_err(MPM_E_ARG, "bad argument");
    if (nrows > (1 << 30)) return set_err(MPM_E_ARG, "too many rows");
    int rc = mlp_stage(m, nrows);
    if (rc) return rc;
    if ((rc = upload_rows(m, nrows, X))) return rc;
    std::vector<int> idx(nrows);
    std::iota(idx.begin(), idx.end(), 0);
    HIPCHK(hipMemcpyAsync(m->idx, idx.data(), nrows * sizeof(int), hipMemcpyHostToDevice,
                          m->stream));
    HIPCHK(hipMemcpyAsync(m->t, T, nrows * kMOut * sizeof(float), hipMemcpyHostToDevice,
                          m->stream));
    hipLaunchKernelGGL(k_zero_d, dim3(1), dim3(64), 0, m->stream, m->d.loss, 1);
    const int nparts = launch_mlp_train(m->d, m->x, m->t, m->idx, (int)nrows, 1.0f, m->stream);
    launch_mlp_reduce(m->d, nparts, m->stream);
    if ((rc = check_launch())) return rc;
    double l = 0.0;
    HIPCHK(hipMemcpyAsync(grad, m->d.grad, kMParams * sizeof(float), hipMemcpyDeviceToHost,
                          m->stream));
    HIPCHK(hipMemcpyAsync(&l, m->d.loss, sizeof(double), hipMemcpyDeviceToHost, m->stream));
    hipLaunchKernelGGL(k_zero_f, dim3(blocks256(kMParams)), dim3(256), 0, m->stream, m->d.grad,
                       kMParams);
    HIPCHK(hipStreamSynchronize(m->stream));
    if (loss) *loss = l / (double)nrows;
    return check_launch();
}

int mpm_mlp_adam(mpm_mlp *m, const float *grad, float lr, int t) {
    if (!m || !grad || t < 1) return set_err(MPM_E_ARG, "bad argument");
    HIPCHK(hipMemcpyAsync(m->d.grad, grad, kMParams * sizeof(float), hipMemcpyHostToDevice,
                          m->stream));
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f - std::pow(b1, (float)t), c2 = 1.0f - std::pow(b2, (float)t);
    launch_mlp_adam(m->d, 0, lr, b1, b2, 1e-8f, c1, c2, m->stream);
    HIPCHK(hipStreamSynchronize(m->stream));
    return check_launch();
}

int mpm_dataset_create(int64_t capacity, mpm_dataset **out) {
    if (!out || capacity <= 0 || capacity > (1 << 30)) return set_err(MPM_E_ARG, "bad capacity");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return set_err(MPM_E_NODEVICE, "no HIP device");
    mpm_dataset *d = new mpm_dataset();
    const int64_t pad = (capacity + kMRows - 1) / kMRows * kMRows;
    if (hipMalloc(&d->x, pad * kMInPad * sizeof(__bf16)) != hipSuccess ||
        hipMalloc(&d->t, pad * kMOut * sizeof(float)) != hipSuccess ||
        hipMalloc(&d->count, sizeof(int)) != hipSuccess) {
        mpm_dataset_destroy(d);
        return set_err(MPM_E_OOM, "dataset allocation failed");
    }
    d->cap = capacity;
    int zero = 0;
    if (hipMemcpy(d->count, &zero, sizeof(int), hipMemcpyHostToDevice) != hipSuccess) {
        mpm_dataset_destroy(d);
        return set_err(MPM_E_HIP, "dataset init failed");
    }
    *out = d;
    return MPM_OK;
}

void mpm_dataset_destroy(mpm_dataset *d) {
    if (!d) return;
    (void)hipFree(d->x);
    (void)hipFree(d->t);
    (void)hipFree(d->count);
    delete d;
}

int mpm_dataset_size(mpm_dataset *d, int64_t *nrows) {
    if (!d || !nrows) return set_err(MPM_E_ARG, "null argument");
    int n = 0;
    HIPCHK(hipMemcpy(&n, d->count, sizeof(int), hipMemcpyDeviceToHost));
    *nrows = std::min<int64_t>(n, d->cap);
    return MPM_OK;
}

int mpm_dataset_add(mpm_dataset *d, int64_t nrows, const float *X, const float *T) {
    if (!d || nrows < 0 || (nrows && (!X || !T))) return set_err(MPM_E_ARG, "bad argument");
    int64_t n = 0;
    int rc = mpm_dataset_size(d, &n);
    if (rc) return rc;
    if (n + nrows > d->cap) return set_err(MPM_E_ARG, "dataset full");
    std::vector<uint16_t> h(nrows * kMInPad);
    for (int64_t i = 0; i < nrows * kMInPad; ++i) h[i] = bf16_bits(X[i]);
    HIPCHK(hipMemcpy(d->x + n * kMInPad, h.data(), h.size() * 2, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(d->t + n * kMOut, T, nrows * kMOut * sizeof(float), hipMemcpyHostToDevice));
    const int nn = (int)(n + nrows);
    HIPCHK(hipMemcpy(d->count, &nn, sizeof(int), hipMemcpyHostToDevice));
    return MPM_OK;
}

int mpm_dataset_target_rms(mpm_dataset *d, double *rms) {
    if (!d || !rms) return set_err(MPM_E_ARG, "null argument");
    int64_t n = 0;
    int rc = mpm_dataset_size(d, &n);
    if (rc) return rc;
    std::vector<float> t(n * kMOut);
    if (n) HIPCHK(hipMemcpy(t.data(), d->t, t.size() * sizeof(float), hipMemcpyDeviceToHost));
    double s = 0.0;
    for (float v : t) s += (double)v * v;
    *rms = n ? std::sqrt(s / (double)(n * kMOut)) : 0.0;
    return MPM_OK;
}

int mpm_learn_features(mpm_ctx *c, mpm_mlp *m, int64_t nrows, const int64_t *nodes,
                       float *X) {
    if (!c || !m || nrows < 0 || (nrows && (!nodes || !X))) return set_err(MPM_E_ARG, "bad argument");
    if (nrows == 0) return MPM_OK;
    if (nrows > c->g.nn) return set_err(MPM_E_ARG, "too many rows");
    std::vector<int> q(nrows);
    for (int64_t r = 0; r < nrows; ++r) {
        if (nodes[r] < 0 || nodes[r] >= c->g.nn) return set_err(MPM_E_ARG, "node out of range");
        q[r] = (int)nodes[r];
    }
    int rc = ctx_ml_buffers(c);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(c->ml_nodes, q.data(), nrows * sizeof(int), hipMemcpyHostToDevice,
                          c->stream));
    launch_mlp_features(c->g, c->N, c->ml_nodes, (int)nrows, m->sc, c->ml_x, c->stream);
    if ((rc = check_launch())) return rc;
    std::vector<uint16_t> h(nrows * kMInPad);
    HIPCHK(hipMemcpyAsync(h.data(), c->ml_x, h.size() * 2, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int64_t i = 0; i < nrows * kMInPad; ++i) {
        const uint32_t u = (uint32_t)h[i] << 16;
        std::memcpy(&X[i], &u, 4);
    }
    return MPM_OK;
}

int mpm_learn_collect(mpm_ctx *c, mpm_mlp *m, mpm_dataset *d, int nfine, int max_rows,
                      uint64_t seed) {
    if (!c || !m || !d || nfine < 1 || max_rows < 1) return set_err(MPM_E_ARG, "bad argument");
    int rc = learned_ok(c);
    if (rc) return rc;
    if ((rc = ctx_ml_buffers(c))) return rc;
    hipStream_t st = c->stream;
    if ((rc = coarse_begin(c, nfine))) return rc;
    int nact = 0;
    if ((rc = active_nodes(c, &nact))) return rc;
    const double p = nact > 0 ? std::min(1.0, (double)max_rows / (double)nact) : 0.0;
    int *cnt = c->ml_count + 1;
    hipLaunchKernelGGL(k_zero_f, dim3(1), dim3(64), 0, st, reinterpret_cast<float *>(cnt), 1);
    int *sel = c->ml_nodes + c->ml_cap;  // selected rows
    if (nact > 0)
        hipLaunchKernelGGL(k_select_rows, dim3(blocks256(nact)), dim3(256), 0, st, c->ml_nodes,
                           nact, p, seed, c->N, c->g.nn, sel, c->ml_vc, cnt);
    int nsel = 0;
    HIPCHK(hipMemcpyAsync(&nsel, cnt, sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    launch_mlp_features(c->g, c->N, sel, nsel, m->sc, c->ml_x, st);
    if ((rc = check_launch())) return rc;
    coarse_drop(c);
    if ((rc = mpm_step(c, nfine))) return rc;
    if (nsel > 0)
        hipLaunchKernelGGL(k_append_rows, dim3(blocks256(nsel)), dim3(256), 0, st, sel, nsel,
                           c->ml_x, c->ml_vc, c->N, c->g.nn, c->ph.mass_tol, *d);
    HIPCHK(hipStreamSynchronize(st));
    return check_launch();
}

int mpm_mlp_train(mpm_mlp *m, mpm_dataset *d, int epochs, int batch, float lr, uint64_t seed,
                  double *loss_out) {
    if (!m || !d || epochs < 0 || batch < kMRows) return set_err(MPM_E_ARG, "bad argument");
    int64_t n = 0;
    int rc = mpm_dataset_size(d, &n);
    if (rc) return rc;
    if (n == 0) return set_err(MPM_E_ARG, "empty dataset");
    if ((rc = mlp_stage(m, n))) return rc;
    std::mt19937_64 rng(seed);
    std::vector<int> perm(n);
    std::iota(perm.begin(), perm.end(), 0);
    const float inv_t = (float)(1.0 / m->sc.dv_ref);
    const float b1 = 0.9f, b2 = 0.999f;
    int &t = m->adam_t;
    for (int ep = 0; ep < epochs; ++ep) {
        std::shuffle(perm.begin(), perm.end(), rng);
        HIPCHK(hipMemcpyAsync(m->idx, perm.data(), n * sizeof(int), hipMemcpyHostToDevice,
                              m->stream));
        hipLaunchKernelGGL(k_zero_d, dim3(1), dim3(64), 0, m->stream, m->d.loss, 1);
        for (int64_t off = 0; off < n; off += batch) {
            const int nb = (int)std::min<int64_t>(batch, n - off);
            const int nparts = launch_mlp_train(m->d, d->x, d->t, m->idx + off, nb, inv_t,
                                                m->stream);
            ++t;
            const float c1 = 1.0f - std::pow(b1, (float)t), c2 = 1.0f - std::pow(b2, (float)t);
            launch_mlp_adam(m->d, nparts, lr, b1, b2, 1e-8f, c1, c2, m->stream);
        }
        if ((rc = check_launch())) return rc;
        double l = 0.0;
        HIPCHK(hipMemcpyAsync(&l, m->d.loss, sizeof(double), hipMemcpyDeviceToHost, m->stream));
        HIPCHK(hipStreamSynchronize(m->stream));
        if (loss_out) loss_out[ep] = l / (double)n;
    }
    return MPM_OK;
}

int mpm_step_learned(mpm_ctx *c, mpm_mlp *m, int ncoarse, int nfine, const double guards[2],
                     int *fallbacks) {
    if (!c || ncoarse < 0 || nfine < 1 || !guards) return set_err(MPM_E_ARG, "bad argument");
    if (fallbacks) *fallbacks = 0;
    if (ncoarse == 0) return MPM_OK;
    int rc = learned_ok(c);
    if (rc) return rc;
    if ((rc = ctx_ml_buffers(c))) return rc;
    hipStream_t st = c->stream;
    const double gmag = std::sqrt(c->ph.grav[0] * c->ph.grav[0] + c->ph.grav[1] * c->ph.grav[1] +
                                  c->ph.grav[2] * c->ph.grav[2]);
    int *flag = c->ml_count + 3;
    int *slot[2] = {c->ml_count, c->ml_count + 8};  // active counts by step parity
    hipLaunchKernelGGL(k_zero_f, dim3(1), dim3(64), 0, st, reinterpret_cast<float *>(slot[0]), 1);
    int fb = 0;
    bool pending = false;  // an accepted step's time advance is still due
    if (!m) {  // uncorrected coarse steps: nothing to decide, no round trip
        for (int it = 0; it < ncoarse; ++it) {
            if ((rc = coarse_begin(c, nfine, pending ? 1 : 0))) return rc;
            pending = false;
            if ((rc = coarse_finish(c, nfine))) return rc;
        }
        if (fallbacks) *fallbacks = 0;
        return finish_call(c);
    }
    if ((rc = ctx_spec_buffer(c))) return rc;
    // One coarse step from particles `in` into `out`: correction, guard and
    // the particle kernel enqueued back to back; the decision kernel also
    // stores the decision in the host slot of the step's parity, with an
    // event behind it.  A rejected step leaves `in` untouched.  The list
    // launch clears the other parity's count for the next step.
    auto enqueue = [&](int it, int advance, const PState &in, const PState &out,
                       Scalars *bak) -> int {
        int r = coarse_begin(c, nfine, advance, &in, &out, bak);
        if (r) return r;
        int *cnt = slot[it & 1];
        launch_active_nodes_rolling(c->g, c->N, c->ph.mass_tol, c->ml_nodes, cnt,
                                    slot[(it + 1) & 1], c->ml_feat, m->sc, st);
        launch_mlp_infer_nodes(m->d, c->g, c->N, c->ml_feat, c->ml_nodes, cnt, m->sc, c->ml_y,
                               c->ml_guard, st);
        hipLaunchKernelGGL(k_guard_decide, dim3(1), dim3(256), 0, st, c->ml_guard, c->sc, gmag,
                           guards[0], guards[1], flag, c->ml_hflag + (it & 1));
        HIPCHK(hipEventRecord(c->ml_ev[it & 1], st));
        hipLaunchKernelGGL(k_apply_dev, dim3(512), dim3(256), 0, st, c->ml_nodes, cnt, c->ml_y,
                           m->sc.dv_ref, c->N, c->g.nn, flag);
        launch_particle_fast(c->g, coarse_ph(c, nfine), c->ls, c->udf, in, out, c->S, c->N, 1,
                             1, 0, c->sc, st);
        return check_launch();
    };
    // Speculation: step it + 1 (from Q into R) is enqueued as if step it
    // (from P into Q) were accepted, before its decision is read, so the GPU
    // does not idle across the round trip.  Accepted: the buffers rotate and
    // step it + 1 is in flight.  Rejected: the speculative step is drained
    // and undone (the step scalars from the copy its dt launch took first,
    // the step counter, the active counts; P, its input of record, was not
    // written), and the plain fine steps run from P.
    bool inflight = false;  // step it is enqueued (from P into Q)
    for (int it = 0; it < ncoarse; ++it) {
        if (!inflight) {
            if ((rc = enqueue(it, pending ? 1 : 0, c->P, c->Q, nullptr))) return rc;
            pending = false;
        }
        const bool spec = it + 1 < ncoarse;
        if (spec) {
            c->dt_next_valid = c->fuse_dt != 0;  // as after an accepted step
            if ((rc = enqueue(it + 1, 1, c->Q, c->R, c->ml_scbak))) return rc;
        }
        HIPCHK(hipEventSynchronize(c->ml_ev[it & 1]));
        if (c->ml_hflag[it & 1]) {
            c->dt_next_valid = c->fuse_dt != 0;
            c->dirty = false;
            if (spec) {
                // step it + 1 reads the new P and writes the new Q
                const PState p0 = c->P;
                c->P = c->Q;
                c->Q = c->R;
                c->R = p0;
                inflight = true;
            } else {
                std::swap(c->P, c->Q);
                pending = true;
                inflight = false;
            }
        } else {
            if (spec) {
                HIPCHK(hipStreamSynchronize(st));
                HIPCHK(hipMemcpyAsync(c->sc, c->ml_scbak, sizeof(Scalars),
                                      hipMemcpyDeviceToDevice, st));
                hipLaunchKernelGGL(k_zero_f, dim3(1), dim3(64), 0, st,
                                   reinterpret_cast<float *>(slot[0]), 1);
                hipLaunchKernelGGL(k_zero_f, dim3(1), dim3(64), 0, st,
                                   reinterpret_cast<float *>(slot[1]), 1);
                c->steps--;  // the speculative step's coarse_begin
            }
            coarse_drop(c);  // also drops the dt candidates of the rejected kernel
            ++fb;
            if ((rc = mpm_step(c, nfine))) return rc;
            inflight = false;
        }
    }
    if (pending) launch_advance_time(c->sc, st);
    if (fallbacks) *fallbacks = fb;
    return finish_call(c);
}

}  // extern "C"

// mpm_context.h — the context behind the C ABI (mpm_ctx.hip) and the
// internal helpers the entry points share (error text, launch checks).
#pragma once

#include <vector>

#include "mpm_internal.h"

struct mpm_ctx {
    using GridP = xmpm::GridP;
    using PhysP = xmpm::PhysP;
    using LsP = xmpm::LsP;
    using UdfP = xmpm::UdfP;
    using PState = xmpm::PState;
    using NState = xmpm::NState;
    using SortState = xmpm::SortState;
    using Scalars = xmpm::Scalars;

    mpm_specs s;
    GridP g;
    PhysP ph;
    LsP ls;
    UdfP udf;
    PState P;      // current particle buffers
    PState Q;      // second buffer the sorted step writes into
    PState R;      // third buffer: the learned step's speculative next step writes into it
    NState N;
    SortState S;
    // Cold particle fields (never read by the step), by upload index.
    std::vector<double> cold_radius, cold_yacc;
    std::vector<int> cold_rigid;
    bool skip_ok = true;    // idle node tiles hold the all-zero massless state
    bool all_dirty = true;  // next step's grid pass must visit every node
    int sorted = 1;         // use the cell-sorted step
    int fast = 1;           // use the specialised particle kernel where it applies
    int fuse_dt = 1;        // the specialised kernel reduces the next step's dt
    int coarse_blend = 0;   // learned coarse step blends PIC/FLIP with alpha^nfine
    bool dt_next_valid = false;  // the last step reduced the next step's dt candidates
    Scalars *sc = nullptr;
    double *d_energy = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t ev[8];
    std::vector<double *> ls_bufs;
    double *udf_dev[3][2] = {{nullptr, nullptr}, {nullptr, nullptr}, {nullptr, nullptr}};
    mpm_wall_udf_fn udf_fn[3][2] = {{nullptr, nullptr}, {nullptr, nullptr}, {nullptr, nullptr}};
    bool any_udf = false;
    bool any_per = false;
    bool dirty = false;  // particles moved since the last redistribute
    int64_t steps = 0;
    double time_host = 0.0;
    double last_dt = 0.0;
    double step_ms = 0.0;
    double kernel_ms[4] = {0, 0, 0, 0};
    int profile = 0;
    int keep_forces = 1;
    // Deterministic P2G (MPM_OPT_DETERMINISTIC): requested, and its buffers
    // (allocated by the option, kept until destroy); det_sync hands them to
    // S where the mode applies.
    int det = 0;
    double *det_blk = nullptr;
    int64_t *det_slot = nullptr;
    // Learned-correction work buffers (allocated by its first call, nodal sized).
    int *ml_nodes = nullptr;      // active node list / selected rows
    int *ml_count = nullptr;      // [0] active count, [1] selected, [2] appended, [3] guard flag,
                                  // [8] alternate active count
    __bf16 *ml_x = nullptr;       // feature rows, padded to 32 rows
    void *ml_feat = nullptr;      // every node's four BF16 features (nn x 8 bytes)
    float *ml_y = nullptr;        // network outputs [row][3]
    double *ml_vc = nullptr;      // coarse nodal velocity of selected rows [row][3]
    double *ml_guard = nullptr;   // guard sums
    int *ml_hflag = nullptr;      // pinned host copies of the guard flag, by step parity
    Scalars *ml_scbak = nullptr;  // step scalars before a speculative learned step
    hipEvent_t ml_ev[2] = {nullptr, nullptr};  // a step's flag copy done, by step parity
    int64_t ml_cap = 0;
    // Thermal module (mpm_enable_thermal): particle fields in slot order,
    // MPM_NTHERMAL * th_cap each (th_alt: the sorted step's output), nodal
    // components 20..25 (6 * nn) and the wall-value scratch (nn).
    int thermal = 0;
    mpm_thermal_specs ts{};
    double *th = nullptr;
    double *th_alt = nullptr;
    int64_t th_cap = 0;
    double *tn = nullptr;
    double *tscr = nullptr;
    // Temperature wall functions and their per-step face tables (2 values per
    // face node).
    mpm_wall_temp_udf_fn tudf_fn[3][2] = {{nullptr, nullptr}, {nullptr, nullptr}, {nullptr, nullptr}};
    double *tudf_dev[3][2] = {{nullptr, nullptr}, {nullptr, nullptr}, {nullptr, nullptr}};
    bool any_tudf = false;
    // Temperature conditions of the level-set bodies (MPM_LS_TEMP_*) and their
    // T_wall, heat flux, h, T_inf.
    int ls_tbc[MPM_MAX_LS_BODIES] = {};
    double ls_tpar[MPM_MAX_LS_BODIES][4] = {};
};

namespace xmpm {

// The deterministic P2G applies to the whole sorted step (the thermal
// deposit has a deterministic form of its own; the tile deposit gathers
// slab loads itself, so no packets are written at the scatter's atomic
// positions).
inline void det_sync(mpm_ctx *c) {
    const bool on = c->det != 0;
    c->S.blk = on ? c->det_blk : nullptr;
    c->S.tile_slot = on ? c->det_slot : nullptr;
}

// Record the message of a failed call and return its status.
int set_err(int code, const char *what);
int hip_err(hipError_t e, const char *expr);
// Status of the last kernel launch.
int check_launch();
// Validated grid and step parameters of a problem specification.
int specs_params(const mpm_specs *s, GridP &g, PhysP &ph);
// Every wall velocity of the specification is zero (idle node tiles then
// hold the all-zero massless state).
bool walls_at_rest(const mpm_specs *s);

// Thermal stages (mpm_thermal.hip).  thermal_on_upload re-creates the
// particle thermal buffers (zeroed) for new particle buffers; thermal_init is
// the initial pass's thermal block; thermal_grid the step's nodal half (after
// the momentum grid pass); thermal_particles the step's G2P half (before the
// particle kernel; perm / count: the sorted order, or null); thermal_swap
// follows the swap of the particle buffers.  tiled / a non-null perm: the
// state was cell-sorted in this step or pass, so the deposits go through the
// tiled deposit (epoch and fallback as for the momentum deposit).
int thermal_on_upload(mpm_ctx *c);
int thermal_init(mpm_ctx *c, bool tiled);
int thermal_grid(mpm_ctx *c, bool tiled, int fallback, hipStream_t st);
int thermal_particles(mpm_ctx *c, bool musl, const int *perm, const int *count,
                      int fallback, hipStream_t st);
void thermal_swap(mpm_ctx *c);
// The temperature wall functions evaluated at time t into the face tables.
int thermal_udf_tables(mpm_ctx *c, double t);
void thermal_free(mpm_ctx *c);

inline int grid_flags_full(const mpm_ctx *c) {
    return G_FINISH_MASS | G_FINISH_VEL | G_FINISH_FRC | G_BACKUP | G_UPDATE | G_BCS |
           (c->keep_forces ? G_KEEP_FRC : 0);
}

}  // namespace xmpm

#define HIPCHK(expr)                                             \
    do {                                                         \
        hipError_t e_ = (expr);                                  \
        if (e_ != hipSuccess) return ::xmpm::hip_err(e_, #expr); \
    } while (0)

// mpm_mlp.h — device state and launchers of the learned coarse-step
// correction: a per-node stencil MLP (27 neighbour nodes x (mass, momentum)
// = 108 inputs -> 64 -> 64 -> 3, tanh-form GELU) whose output corrects the
// nodal velocity of one coarse step (nfine x the CFL dt) so that it stands
// in for nfine plain steps.  BF16 operands on MFMA with FP32 accumulation;
// FP32 master parameters and Adam moments; FP64 everywhere the step state
// is touched.  No counterpart in the reference (an addition beside the
// plain step, which stays the oracle).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mpm_internal.h"

namespace xmpm {

constexpr int kMIn = 108, kMInPad = 128, kMH = 64, kMOut = 3;
constexpr int kOffW1 = 0, kOffB1 = kOffW1 + kMIn * kMH, kOffW2 = kOffB1 + kMH,
              kOffB2 = kOffW2 + kMH * kMH, kOffW3 = kOffB2 + kMH,
              kOffB3 = kOffW3 + kMH * kMOut, kMParams = kOffB3 + kMOut;
static_assert(kMParams == 11331, "parameter count");
constexpr int kMRows = 32;  // rows (nodes) per wavefront tile

// Device copies of one network.  params/m/v/grad: FP32 [kMParams] in the
// layout of include/exagoop_amd.h (W[in][out]).  The BF16 operand copies:
//   w1t [64][128]  W1 transposed, inputs padded to 128 with zeros
//   w2t [64][64]   W2 transposed (forward B operand)
//   w2  [64][64]   W2 as stored (backward B operand)
//   w3t [32][64]   W3 transposed, outputs padded to 32
//   w3  [64][16]   W3 as stored, outputs padded to 16
struct MlpDev {
    float *params, *m, *v, *grad;
    __bf16 *w1t, *w2t, *w2, *w3t, *w3;
    double *loss;  // running loss sum of a training pass
};

struct MlpScales {
    double m_ref, v_ref, dv_ref;
};

// Feature rows X[r][128] (BF16) of the listed nodes from the nodal mass and
// velocity; rows past nrows up to the next multiple of 32 are zeroed.
void launch_mlp_features(const GridP &g, const NState &N, const int *nodes, int nrows,
                         MlpScales sc, __bf16 *X, hipStream_t s);
// Y[r][3] = network(X[r]) for r < nrows (X padded to a multiple of 32 rows).
void launch_mlp_infer(const MlpDev &w, const __bf16 *X, int nrows, float *Y, hipStream_t s);
// The same with the row count read on the device (*count), no host round trip.
void launch_mlp_features_dev(const GridP &g, const NState &N, const int *nodes, const int *count,
                             MlpScales sc, __bf16 *X, hipStream_t s);
void launch_mlp_infer_dev(const MlpDev &w, const __bf16 *X, const int *count, float *Y,
                          hipStream_t s);
// Gradient of the mean squared error over the rows idx[0..nrows) of (X, T)
// (T scaled by inv_t) added into w.grad; the loss sum into w.loss.
void launch_mlp_train(const MlpDev &w, const __bf16 *X, const float *T, const int *idx,
                      int nrows, float inv_t, hipStream_t s);
// One Adam step from w.grad (then zeroed); refreshes the BF16 copies.
void launch_mlp_adam(const MlpDev &w, float lr, float beta1, float beta2, float eps, float c1,
                     float c2, hipStream_t s);
// BF16 copies from params (after an upload).
void launch_mlp_pack(const MlpDev &w, hipStream_t s);

// Active node list: nodes with mass >= tol, in index order, count in *count.
void launch_active_nodes(const GridP &g, const NState &N, double tol, int *list, int *count,
                         int *scratch, hipStream_t s);
// The same without clearing *count first (it must be zero on entry); the
// launch zeroes *next_count for the following step's list instead.
void launch_active_nodes_rolling(const GridP &g, const NState &N, double tol, int *list,
                                 int *count, int *next_count, hipStream_t s);

}  // namespace xmpm

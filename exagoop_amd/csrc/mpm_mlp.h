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
constexpr int kInferBlocks = 1024;  // workgroups of the learned step's inference
// Training partial rows: kMParams gradient entries, the loss sum, padding
// to a multiple of 64 floats; at most kMTrainParts rows (one per wavefront).
constexpr int kMPartStride = 11392, kMTrainParts = 512;
static_assert(kMPartStride >= kMParams + 1 && kMPartStride % 64 == 0, "partial row");

// Device copies of one network.  params/m/v/grad: FP32 [kMParams] in the
// layout of include/exagoop_amd.h (W[in][out]).  The BF16 operand copies,
// with swap23(i) = i with bits 2 and 3 exchanged (the slot order in which a
// transposed 32x32 MFMA result holds its columns):
//   w1t  [64][128]  w1t[n][k]  = W1[k][n], inputs padded to 128 with zeros
//   w2tp [64][64]   w2tp[n][s] = W2[swap23(s)][n]   (forward B operand)
//   w2p  [64][64]   w2p[k][s]  = W2[k][swap23(s)]   (backward B operand)
//   w3tp [32][64]   w3tp[n][s] = W3[swap23(s)][n], outputs padded to 32
//   w3   [64][16]   w3[k][n]   = W3[k][n], outputs padded to 16
// part: [kMTrainParts][kMPartStride] FP32 partial gradient rows.
struct MlpDev {
    float *params, *m, *v, *grad, *part;
    __bf16 *w1t, *w2tp, *w2p, *w3tp, *w3;
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
// Gradient of the mean squared error over the rows idx[0..nrows) of (X, T)
// (T scaled by inv_t) as partial rows in w.part, in a fixed order (no
// atomics); returns the number of rows written.
int launch_mlp_train(const MlpDev &w, const __bf16 *X, const float *T, const int *idx, int nrows,
                     float inv_t, hipStream_t s);
// w.grad = the sum of the first nparts partial rows, in row order; their loss
// sum added to *w.loss.
void launch_mlp_reduce(const MlpDev &w, int nparts, hipStream_t s);
// One Adam step; refreshes the BF16 copies.  The gradient is the sum of the
// first nparts partial rows (their loss sum added to *w.loss) or, with
// nparts == 0, w.grad (then zeroed).
void launch_mlp_adam(const MlpDev &w, int nparts, float lr, float beta1, float beta2, float eps,
                     float c1, float c2, hipStream_t s);
// BF16 copies from params (after an upload).
void launch_mlp_pack(const MlpDev &w, hipStream_t s);

// Y[r][3] = network(features of nodes[r]) for r < *count, the rows
// gathered in the kernel from feat (every node's four BF16 features, 8 bytes,
// written by launch_active_nodes_rolling; no row buffer), and the guard sums
// of dv = Y dv_ref as kInferBlocks partial rows sums[b * 8 + k], k < 7: m,
// m dv (3), m |v|^2, m |v + dv|^2, non-finite count (every row written).
void launch_mlp_infer_nodes(const MlpDev &w, const GridP &g, const NState &N, const void *feat,
                            const int *nodes, const int *count, MlpScales sc, float *Y,
                            double *sums, hipStream_t s);

// Active node list: nodes with mass >= tol, in index order within each
// workgroup's span of nodes (spans in arbitrary order), count in *count.
void launch_active_nodes(const GridP &g, const NState &N, double tol, int *list, int *count,
                         int *scratch, hipStream_t s);
// The same without clearing *count first (it must be zero on entry); the
// launch zeroes *next_count for the following step's list instead, and
// writes every node's four BF16 features (m / m_ref, m v / (m_ref v_ref), the
// conversions of launch_mlp_features) to feat (nn x 8 bytes).
void launch_active_nodes_rolling(const GridP &g, const NState &N, double tol, int *list,
                                 int *count, int *next_count, void *feat, MlpScales sc,
                                 hipStream_t s);

}  // namespace xmpm

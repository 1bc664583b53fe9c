// mpm_mlp.hip — kernels of the learned coarse-step correction for gfx950.
//
// One wavefront works on a tile of 32 rows (nodes) with
// v_mfma_f32_32x32x16_bf16 and keeps the whole tile in registers.  Operand
// layouts of that instruction, lane l, h = l >> 5:
//   A (32 x 16):  row l & 31, k = 8 h + j, j = 0..7 (8 BF16 per lane)
//   B (16 x 32):  col l & 31, k = 8 h + j
//   C (32 x 32):  col l & 31, row = 8 (r >> 2) + 4 h + (r & 3), r = 0..15
//
// A C tile used as an operand.  Registers r = 8 q + j of lane l hold rows
// 16 q + 8 (j >> 2) + 4 h + (j & 3) = swap23(16 q + 8 h + j) of column l & 31,
// where swap23 exchanges bits 2 and 3 of an index.  So a C tile, taken 8
// registers at a time, is an operand whose lane index is its column and
// whose k slot s stands for row swap23(s):
//   * as the k = rows operand of a weight gradient (dW = H^T dZ sums over
//     rows), where both operands come from C tiles with the same slot order;
//   * transposed: A = C tile, B = the permuted identity (B[s][n] = 1 where
//     swap23(s) == n) gives C^T in C layout, exactly (each product is a
//     BF16 value times 1).  Its registers then hold the tile's columns in
//     slot order swap23, so the weights that multiply it are stored with
//     their k index permuted by swap23 (the operand copies below).
// The forward layers produce C tiles with lanes over features and registers
// over rows; each activation is transposed once (two MFMAs per 32x32 tile)
// for the next layer and used as it is for the weight gradient.  No LDS is
// touched inside a tile.
//
// Forward: Z1 = X W1 + b1, H1 = gelu(Z1); Z2 = H1 W2 + b2, H2 = gelu(Z2);
// Y = H2 W3 + b3.  Backward of L = (1/N) sum (Y - T)^2:
// dY = 2 (Y - T) / N; dZ2 = (dY W3^T) gelu'(Z2); dZ1 = (dZ2 W2^T) gelu'(Z1);
// dW3 = H2^T dY, dW2 = H1^T dZ2, dW1 = X^T dZ1, db = column sums.
// X, the weights, H1, H2, dY, dZ2, dZ1 enter the products as BF16 (round to
// nearest even); products accumulate in FP32.
//
// Training sums the gradient without atomics, in a fixed order: each
// wavefront writes the gradient of its tiles to a partial row of its own,
// and the Adam (or reduce) kernel sums the partial rows in a fixed order.
// The gradient of a batch is therefore the same bits on every run.
#include <hip/hip_runtime.h>

#include "mpm_mlp.h"

namespace xmpm {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kWave = 64;
constexpr int kTrainWaves = 4;  // wavefronts per training workgroup

__device__ __forceinline__ bf16x8 ld8(const __bf16 *p) {
    return *reinterpret_cast<const bf16x8 *>(p);
}
__device__ __forceinline__ int crow(int r, int lane) {
    return 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
}
__device__ __forceinline__ int swap23(int s) {
    return (s & ~12) | ((s << 1) & 8) | ((s >> 1) & 4);
}
__device__ __forceinline__ f32x16 mfma(bf16x8 a, bf16x8 b, f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f32x16 zero16() {
    f32x16 z;
#pragma unroll
    for (int r = 0; r < 16; ++r) z[r] = 0.0f;
    return z;
}
// Registers 8 q .. 8 q + 7 of a C tile as an operand (BF16, round to nearest even).
__device__ __forceinline__ bf16x8 half8(const f32x16 &c, int q) {
    bf16x8 o;
#pragma unroll
    for (int j = 0; j < 8; ++j) o[j] = (__bf16)c[8 * q + j];
    return o;
}
// B operand of an identity over a 32-slot block, k step q: 1 where the slot
// s = 16 q + 8 h + j equals want(n), n = lane & 31; want = swap23 (PERM) or n.
__device__ __forceinline__ bf16x8 ident(int q, int lane, bool perm) {
    const int n = lane & 31, want = perm ? swap23(n) : n, base = 16 * q + 8 * (lane >> 5);
    bf16x8 o;
#pragma unroll
    for (int j = 0; j < 8; ++j) o[j] = (__bf16)(base + j == want ? 1.0f : 0.0f);
    return o;
}
// Transpose of a C tile given as its two BF16 halves, in C layout.
__device__ __forceinline__ f32x16 transpose(bf16x8 lo, bf16x8 hi, bf16x8 p0, bf16x8 p1) {
    return mfma(hi, p1, mfma(lo, p0, zero16()));
}

// tanh-form GELU and its derivative from one tanh, FP32.  tanh u =
// 1 - 2 / (1 + e^(2u)) with the hardware exp2 and reciprocal (about 1e-7
// absolute error, far below the BF16 rounding of the activations): no
// branches, where the library tanhf branches per element and the 64
// activations of a lane would split the kernel into hundreds of blocks.
// e^(2u) overflows to +inf for large u (tanh = 1) and underflows to 0 for
// very negative u (tanh = -1).
__device__ __forceinline__ float gelu_t(float z) {
    const float u = 0.7978845608f * (z + 0.044715f * z * z * z);
    const float e = __builtin_amdgcn_exp2f(2.8853900817779268f * u);  // 2 log2(e) u
    return 1.0f - 2.0f * __builtin_amdgcn_rcpf(1.0f + e);
}
__device__ __forceinline__ float gelu_of(float z, float t) { return 0.5f * z * (1.0f + t); }
__device__ __forceinline__ float gelu_d_of(float z, float t) {
    return 0.5f * (1.0f + t) + 0.5f * z * (1.0f - t * t) * 0.7978845608f * (1.0f + 0.134145f * z * z);
}

// The X operand of a tile: lane l reads row src (lane's row l & 31; src < 0
// is a zero row, read from row 0 and replaced), features 16 ks + 8 h .. + 7.
// Loads and selects, no branches.
__device__ __forceinline__ void load_x(const __bf16 *X, int64_t src, int lane, bf16x8 xa[8]) {
    const __bf16 *p = X + (src >= 0 ? src : 0) * kMInPad + 8 * (lane >> 5);
#pragma unroll
    for (int ks = 0; ks < kMInPad / 16; ++ks) {
        const bf16x8 v = ld8(p + ks * 16);
#pragma unroll
        for (int j = 0; j < 8; ++j) xa[ks][j] = src >= 0 ? v[j] : (__bf16)0.0f;
    }
}

// Forward pass of one tile in registers.  h1[nt][q], h2[nt][q]: the
// activations as C tiles (lanes: features 32 nt + (l & 31); registers: rows),
// in BF16 halves; gd1, gd2: gelu' of the pre-activations in the same layout
// (TRAIN only); y: the output C tile (lanes: outputs, registers: rows).
template <bool TRAIN>
__device__ __forceinline__ void forward_tile(const MlpDev &w, const bf16x8 xa[8], const float b1[2],
                                             const float b2[2], float b3, bf16x8 p0, bf16x8 p1,
                                             bf16x8 h1[2][2], bf16x8 h2[2][2], f32x16 gd1[2],
                                             f32x16 gd2[2], f32x16 &y, int lane) {
    const int col = lane & 31, kh = 8 * (lane >> 5);
    f32x16 h1t[2], h2t[2];
#pragma unroll
    for (int nt = 0; nt < 2; ++nt) {
        f32x16 acc = zero16();
#pragma unroll
        for (int ks = 0; ks < kMInPad / 16; ++ks)
            acc = mfma(xa[ks], ld8(w.w1t + (nt * 32 + col) * kMInPad + ks * 16 + kh), acc);
        f32x16 h;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const float z = acc[r] + b1[nt];
            const float t = gelu_t(z);
            h[r] = gelu_of(z, t);
            if (TRAIN) gd1[nt][r] = gelu_d_of(z, t);
        }
        h1[nt][0] = half8(h, 0);
        h1[nt][1] = half8(h, 1);
        h1t[nt] = transpose(h1[nt][0], h1[nt][1], p0, p1);
    }
#pragma unroll
    for (int nt = 0; nt < 2; ++nt) {
        f32x16 acc = zero16();
#pragma unroll
        for (int ks = 0; ks < kMH / 16; ++ks)
            acc = mfma(half8(h1t[ks >> 1], ks & 1), ld8(w.w2tp + (nt * 32 + col) * kMH + ks * 16 + kh),
                       acc);
        f32x16 h;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const float z = acc[r] + b2[nt];
            const float t = gelu_t(z);
            h[r] = gelu_of(z, t);
            if (TRAIN) gd2[nt][r] = gelu_d_of(z, t);
        }
        h2[nt][0] = half8(h, 0);
        h2[nt][1] = half8(h, 1);
        h2t[nt] = transpose(h2[nt][0], h2[nt][1], p0, p1);
    }
    f32x16 acc = zero16();
#pragma unroll
    for (int ks = 0; ks < kMH / 16; ++ks)
        acc = mfma(half8(h2t[ks >> 1], ks & 1), ld8(w.w3tp + col * kMH + ks * 16 + kh), acc);
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[r] += b3;
    y = acc;
}

__device__ __forceinline__ void load_biases(const MlpDev &w, int col, float b1[2], float b2[2],
                                            float &b3) {
#pragma unroll
    for (int nt = 0; nt < 2; ++nt) {
        b1[nt] = w.params[kOffB1 + nt * 32 + col];
        b2[nt] = w.params[kOffB2 + nt * 32 + col];
    }
    b3 = col < kMOut ? w.params[kOffB3 + col] : 0.0f;
}

// ---------------------------------------------------------------- inference

__global__ __launch_bounds__(256) void k_mlp_infer(MlpDev w, const __bf16 *X, int nrows,
                                                   float *Y) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, col = lane & 31;
    const int ntiles = (nrows + kMRows - 1) / kMRows;
    const bf16x8 p0 = ident(0, lane, true), p1 = ident(1, lane, true);
    float b1[2], b2[2], b3;
    load_biases(w, col, b1, b2, b3);
    for (int t = blockIdx.x * 4 + wave; t < ntiles; t += gridDim.x * 4) {
        // X is padded to whole tiles: every row of the tile is readable
        bf16x8 xa[8];
        load_x(X, (int64_t)t * kMRows + col, lane, xa);
        bf16x8 h1[2][2], h2[2][2];
        f32x16 y;
        forward_tile<false>(w, xa, b1, b2, b3, p0, p1, h1, h2, nullptr, nullptr, y, lane);
        if (col < kMOut)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int row = t * kMRows + crow(r, lane);
                if (row < nrows) Y[(int64_t)row * kMOut + col] = y[r];
            }
    }
}

// ---------------------------------------------------------------- training

// One tile's gradient into the partial row: stored (ADD false, the
// wavefront's first tile) or added to what the row holds.
template <bool ADD>
__device__ __forceinline__ void train_tile(const MlpDev &w, const __bf16 *X, const float *T,
                                           const int *idx, int nrows, float inv_t, float inv_n,
                                           int t, const float b1[2], const float b2[2], float b3,
                                           bf16x8 p0, bf16x8 p1, bf16x8 i0, bf16x8 i1, float *row,
                                           int lane) {
    const int col = lane & 31, kh = 8 * (lane >> 5);
    auto put = [&](int e, float x) { row[e] = ADD ? row[e] + x : x; };
    const int gr = t * kMRows + col;  // this lane's row of the tile
    const int src = gr < nrows ? idx[gr] : -1;
    bf16x8 xa[8];
    load_x(X, src, lane, xa);
    bf16x8 h1[2][2], h2[2][2];
    f32x16 gd1[2], gd2[2], y;
    forward_tile<true>(w, xa, b1, b2, b3, p0, p1, h1, h2, gd1, gd2, y, lane);
    // dY = 2 (Y - T) / N (C layout: lanes outputs, registers rows)
    f32x16 dy;
    float lsum = 0.0f, db3 = 0.0f;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        // unconditional load (row 0, output 2 stand in for padding) and a
        // select, so that the tile stays one basic block
        const int s = __shfl(src, crow(r, lane), kWave);
        const float tv = T[(int64_t)(s >= 0 ? s : 0) * kMOut + (col < kMOut ? col : kMOut - 1)];
        const float e = y[r] - tv * inv_t;
        const bool ok = col < kMOut && s >= 0;
        lsum += ok ? e * e : 0.0f;
        const float d = ok ? 2.0f * e * inv_n : 0.0f;
        dy[r] = (float)(__bf16)d;
        db3 += dy[r];
    }
    const bf16x8 dyb0 = half8(dy, 0), dyb1 = half8(dy, 1);
    // dW3 = H2^T dY
#pragma unroll
    for (int m = 0; m < 2; ++m) {
        const f32x16 g = mfma(h2[m][1], dyb1, mfma(h2[m][0], dyb0, zero16()));
        if (col < kMOut)
#pragma unroll
            for (int r = 0; r < 16; ++r) put(kOffW3 + (m * 32 + crow(r, lane)) * kMOut + col, g[r]);
    }
    db3 += __shfl_xor(db3, 32, kWave);
    if (lane < kMOut) put(kOffB3 + lane, db3);
    for (int off = 32; off > 0; off >>= 1) lsum += __shfl_xor(lsum, off, kWave);
    if (lane == 0) put(kMParams, lsum);
    // dZ2 = (dY W3^T) gelu'(Z2): dY transposed (lanes rows), outputs in
    // the slots of k step 0
    const bf16x8 dya = half8(transpose(dyb0, dyb1, p0, p1), 0);
    bf16x8 dz2[2][2];
#pragma unroll
    for (int nt = 0; nt < 2; ++nt) {
        const f32x16 acc = mfma(dya, ld8(w.w3 + (nt * 32 + col) * 16 + kh), zero16());
        f32x16 d;
        float db = 0.0f;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            d[r] = (float)(__bf16)(acc[r] * gd2[nt][r]);
            db += d[r];
        }
        db += __shfl_xor(db, 32, kWave);
        if (lane < 32) put(kOffB2 + nt * 32 + col, db);
        dz2[nt][0] = half8(d, 0);
        dz2[nt][1] = half8(d, 1);
    }
    // dW2 = H1^T dZ2
#pragma unroll
    for (int m = 0; m < 2; ++m)
#pragma unroll
        for (int nt = 0; nt < 2; ++nt) {
            const f32x16 g = mfma(h1[m][1], dz2[nt][1], mfma(h1[m][0], dz2[nt][0], zero16()));
#pragma unroll
            for (int r = 0; r < 16; ++r)
                put(kOffW2 + (m * 32 + crow(r, lane)) * kMH + nt * 32 + col, g[r]);
        }
    // dZ1 = (dZ2 W2^T) gelu'(Z1)
    f32x16 dz2t[2];
#pragma unroll
    for (int nt = 0; nt < 2; ++nt) dz2t[nt] = transpose(dz2[nt][0], dz2[nt][1], p0, p1);
    bf16x8 dz1[2][2];
#pragma unroll
    for (int nt = 0; nt < 2; ++nt) {
        f32x16 acc = zero16();
#pragma unroll
        for (int ks = 0; ks < kMH / 16; ++ks)
            acc = mfma(half8(dz2t[ks >> 1], ks & 1),
                       ld8(w.w2p + (nt * 32 + col) * kMH + ks * 16 + kh), acc);
        f32x16 d;
        float db = 0.0f;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            d[r] = (float)(__bf16)(acc[r] * gd1[nt][r]);
            db += d[r];
        }
        db += __shfl_xor(db, 32, kWave);
        if (lane < 32) put(kOffB1 + nt * 32 + col, db);
        dz1[nt][0] = half8(d, 0);
        dz1[nt][1] = half8(d, 1);
    }
    // dW1 = X^T dZ1: X as C tiles (lanes features, registers rows) from
    // the natural identity; input rows 108..127 are padding
#pragma unroll
    for (int m = 0; m < kMInPad / 32; ++m) {
        const f32x16 xc = mfma(xa[2 * m + 1], i1, mfma(xa[2 * m], i0, zero16()));
        const bf16x8 x0 = half8(xc, 0), x1 = half8(xc, 1);
#pragma unroll
        for (int nt = 0; nt < 2; ++nt) {
            const f32x16 g = mfma(x1, dz1[nt][1], mfma(x0, dz1[nt][0], zero16()));
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int in = m * 32 + crow(r, lane);
                if (in < kMIn) put(kOffW1 + in * kMH + nt * 32 + col, g[r]);
            }
        }
    }
}

// Wavefront v of the launch takes tile t0 + v of the batch rows
// idx[0..nrows) and writes its gradient to partial row v: part[v][0..kMParams),
// the loss sum (FP32) at part[v][kMParams]; stored (ADD false) or added to
// what the row holds (ADD true, the launches after the first for batches of
// more than kMTrainParts tiles).  Each gradient block is written as soon as
// it is computed, and every entry of a row belongs to one lane.
template <bool ADD>
__global__ __launch_bounds__(kTrainWaves * kWave) void k_mlp_train(MlpDev w, const __bf16 *X,
                                                                   const float *T, const int *idx,
                                                                   int nrows, float inv_t,
                                                                   float inv_n, int t0) {
    const int lane = threadIdx.x & 63, col = lane & 31;
    // wavefront-uniform (SGPR) so that the partial row is a scalar base and
    // every store an SGPR base + 32-bit offset
    const int gw = __builtin_amdgcn_readfirstlane(blockIdx.x * kTrainWaves + (threadIdx.x >> 6));
    const int ntiles = (nrows + kMRows - 1) / kMRows;
    if (t0 + gw >= ntiles) return;
    const bf16x8 p0 = ident(0, lane, true), p1 = ident(1, lane, true);
    const bf16x8 i0 = ident(0, lane, false), i1 = ident(1, lane, false);
    float b1[2], b2[2], b3;
    load_biases(w, col, b1, b2, b3);
    train_tile<ADD>(w, X, T, idx, nrows, inv_t, inv_n, t0 + gw, b1, b2, b3, p0, p1, i0, i1,
                    w.part + (int64_t)gw * kMPartStride, lane);
}

// Sum of the first nparts partial rows at entry e: block (64 entries) x 4
// quarters; quarter q sums rows q, q + 4, ... in order, then the quarters
// are added in order.  Returns the sum in quarter 0 (others: 0).
__device__ __forceinline__ float part_sum(const float *part, int nparts, int e, float *lds) {
    const int q = threadIdx.y;
    float s = 0.0f;
    if (e <= kMParams) {
        int b = q;
        for (; b + 12 < nparts; b += 16) {
            const float a0 = part[(int64_t)b * kMPartStride + e];
            const float a1 = part[(int64_t)(b + 4) * kMPartStride + e];
            const float a2 = part[(int64_t)(b + 8) * kMPartStride + e];
            const float a3 = part[(int64_t)(b + 12) * kMPartStride + e];
            s = (((s + a0) + a1) + a2) + a3;
        }
        for (; b < nparts; b += 4) s += part[(int64_t)b * kMPartStride + e];
    }
    lds[q * 64 + threadIdx.x] = s;
    __syncthreads();
    if (q == 0) s = ((lds[threadIdx.x] + lds[64 + threadIdx.x]) + lds[128 + threadIdx.x]) +
                    lds[192 + threadIdx.x];
    return s;
}

__global__ __launch_bounds__(256) void k_mlp_reduce(MlpDev w, int nparts) {
    __shared__ float lds[256];
    const int e = blockIdx.x * 64 + threadIdx.x;
    const float s = part_sum(w.part, nparts, e, lds);
    if (threadIdx.y != 0 || e > kMParams) return;
    if (e == kMParams)
        *w.loss += (double)s;  // one thread; stream order makes the running sum
    else
        w.grad[e] = s;
}

// BF16 operand copies of parameter e (layouts in mpm_mlp.h).
__device__ __forceinline__ void pack_one(const MlpDev &w, int e, float p) {
    const __bf16 b = (__bf16)p;
    if (e < kOffB1) {
        const int k = (e - kOffW1) / kMH, n = (e - kOffW1) % kMH;
        w.w1t[n * kMInPad + k] = b;
    } else if (e >= kOffW2 && e < kOffB2) {
        const int k = (e - kOffW2) / kMH, n = (e - kOffW2) % kMH;
        w.w2tp[n * kMH + swap23(k)] = b;
        w.w2p[k * kMH + swap23(n)] = b;
    } else if (e >= kOffW3 && e < kOffB3) {
        const int k = (e - kOffW3) / kMOut, n = (e - kOffW3) % kMOut;
        w.w3tp[n * kMH + swap23(k)] = b;
        w.w3[k * 16 + n] = b;  // swap23(n) == n for the outputs 0..2
    }
}

__device__ __forceinline__ void adam_one(const MlpDev &w, int e, float gr, float lr, float beta1,
                                         float beta2, float eps, float c1, float c2) {
    const float m = beta1 * w.m[e] + (1.0f - beta1) * gr;
    const float v = beta2 * w.v[e] + (1.0f - beta2) * gr * gr;
    w.m[e] = m;
    w.v[e] = v;
    const float mh = m / c1, vh = v / c2;
    const float p = w.params[e] - lr * mh / (sqrtf(vh) + eps);
    w.params[e] = p;
    pack_one(w, e, p);
}

// One Adam step from the sum of nparts partial rows (their loss sum added to
// *w.loss); blocks of 64 entries x 4 quarters.
__global__ __launch_bounds__(256) void k_mlp_adam_parts(MlpDev w, int nparts, float lr,
                                                        float beta1, float beta2, float eps,
                                                        float c1, float c2) {
    __shared__ float lds[256];
    const int e = blockIdx.x * 64 + threadIdx.x;
    const float gr = part_sum(w.part, nparts, e, lds);
    if (threadIdx.y != 0 || e > kMParams) return;
    if (e == kMParams) {
        *w.loss += (double)gr;
        return;
    }
    adam_one(w, e, gr, lr, beta1, beta2, eps, c1, c2);
}

// One Adam step from w.grad (then zeroed).
__global__ void k_mlp_adam(MlpDev w, float lr, float beta1, float beta2, float eps, float c1,
                           float c2) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= kMParams) return;
    const float gr = w.grad[e];
    w.grad[e] = 0.0f;
    adam_one(w, e, gr, lr, beta1, beta2, eps, c1, c2);
}

__global__ void k_mlp_pack(MlpDev w) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < kMParams) pack_one(w, e, w.params[e]);
}

// ---------------------------------------------------------------- features

// Row r of node nodes[r]: neighbour nb = (dz+1)*9 + (dy+1)*3 + (dx+1),
// columns 4 nb + (0: m / m_ref, 1..3: m v / (m_ref v_ref)); zero outside
// the nodal box and in columns 108..127.  One thread per (row, 8 columns).
__global__ void k_mlp_features(GridP g, NState N, const int *nodes, int nrows,
                               MlpScales sc, __bf16 *X) {
    const int nrows_pad = (nrows + kMRows - 1) / kMRows * kMRows;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
         t < (int64_t)nrows_pad * (kMInPad / 8); t += (int64_t)gridDim.x * blockDim.x) {
    const int r = (int)(t / (kMInPad / 8)), c8 = (int)(t % (kMInPad / 8));
    bf16x8 out;
    const double im = 1.0 / sc.m_ref, imv = 1.0 / (sc.m_ref * sc.v_ref);
    const int nx = g.n[0] + 1, ny = g.n[1] + 1, nz = g.n[2] + 1;
    int i = 0, j = 0, k = 0;
    if (r < nrows) {
        const int64_t q = nodes[r];
        i = (int)(q % nx);
        j = (int)((q / nx) % ny);
        k = (int)(q / ((int64_t)nx * ny));
    }
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const int c = c8 * 8 + e;
        float v = 0.0f;
        if (r < nrows && c < kMIn) {
            const int nb = c >> 2, comp = c & 3;
            const int a = i + (nb % 3) - 1, b = j + (nb / 3) % 3 - 1, cz = k + nb / 9 - 1;
            if (a >= 0 && a < nx && b >= 0 && b < ny && cz >= 0 && cz < nz) {
                const int64_t p = a + (int64_t)nx * (b + (int64_t)ny * cz);
                const double m = N.mass[p];
                v = comp == 0 ? (float)(m * im) : (float)(m * N.vel[(comp - 1) * g.nn + p] * imv);
            }
        }
        out[e] = (__bf16)v;
    }
    *reinterpret_cast<bf16x8 *>(X + (int64_t)r * kMInPad + c8 * 8) = out;
    }
}

// The four BF16 features of one node (m / m_ref, m v / (m_ref v_ref)), in
// the conversions of k_mlp_features, as one 8-byte word.
__device__ __forceinline__ uint2 node_features(const NState &N, int64_t nn, int64_t p, double im,
                                               double imv) {
    const double m = N.mass[p];
    __bf16 f[4];
    f[0] = (__bf16)(float)(m * im);
    for (int d = 0; d < 3; ++d) f[1 + d] = (__bf16)(float)(m * N.vel[d * nn + p] * imv);
    uint2 w;
    __builtin_memcpy(&w, f, sizeof(w));
    return w;
}

// Active nodes (mass >= tol).  A workgroup takes kActiveSpan consecutive
// nodes in 16 coalesced rounds, keeps its flags in a bit mask, reserves its
// run with one atomic and writes the run in node order.  next_count (if
// given) is zeroed for the following step's list.  feat (if given): every
// node's features (node_features), for the inference's gather.
constexpr int kActiveRounds = 16, kActiveSpan = 256 * kActiveRounds;
__global__ __launch_bounds__(256) void k_active_nodes(GridP g, NState N, double tol, int *list,
                                                      int *count, int *next_count, uint2 *feat,
                                                      MlpScales sc) {
    __shared__ int off[kActiveRounds * 4];  // counts, then offsets, of (round, wavefront)
    __shared__ int base;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int64_t q0 = (int64_t)blockIdx.x * kActiveSpan + threadIdx.x;
    if (next_count && blockIdx.x == 0 && threadIdx.x == 0) *next_count = 0;
    unsigned mask = 0;
    const double im = 1.0 / sc.m_ref, imv = 1.0 / (sc.m_ref * sc.v_ref);
#pragma unroll
    for (int k = 0; k < kActiveRounds; ++k) {
        const int64_t q = q0 + 256 * k;
        if (feat && q < g.nn) feat[q] = node_features(N, g.nn, q, im, imv);
        const bool f = q < g.nn && N.mass[q] >= tol;
        mask |= (unsigned)f << k;
        const unsigned long long b = __ballot(f);
        if (lane == 0) off[4 * k + w] = __popcll(b);
    }
    __syncthreads();
    if (w == 0) {  // exclusive scan of the 64 counts (round-major) by one wavefront
        const int c = off[lane];
        int x = c;
        for (int d = 1; d < 64; d <<= 1) {
            const int y = __shfl_up(x, d, 64);
            if (lane >= d) x += y;
        }
        off[lane] = x - c;
        if (lane == 63) base = x ? atomicAdd(count, x) : 0;
    }
    __syncthreads();
    const unsigned long long lt = (1ull << lane) - 1ull;
#pragma unroll
    for (int k = 0; k < kActiveRounds; ++k) {
        const bool f = (mask >> k) & 1u;
        const unsigned long long b = __ballot(f);
        if (f) list[base + off[4 * k + w] + __popcll(b & lt)] = (int)(q0 + 256 * k);
    }
}

// Inference on the listed nodes with the feature rows gathered in the kernel
// from the per-node features k_active_nodes wrote (the rows of
// k_mlp_features, never stored) and
// the guard sums of the corrections added on the way: sums[0] += m,
// sums[1+d] += m dv_d, sums[4] += m |v|^2, sums[5] += m |v + dv|^2,
// sums[6] += count of non-finite dv (dv = Y dv_ref, the FP64 nodal mass m
// and velocity v).  The row count is read on the device.
// Two wavefronts per SIMD: the register budget then holds one tile's 64
// FP64 gathers in flight with the network, without spills.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_mlp_infer_nodes(
    MlpDev w, GridP g, NState N, const uint2 *feat, const int *nodes, const int *count,
    MlpScales sc, float *Y, double *sums) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, col = lane & 31;
    const int h = lane >> 5;
    const int nrows = *count;
    const int ntiles = (nrows + kMRows - 1) / kMRows;
    const bf16x8 p0 = ident(0, lane, true), p1 = ident(1, lane, true);
    float b1[2], b2[2], b3;
    load_biases(w, col, b1, b2, b3);
    const int nx = g.n[0] + 1, ny = g.n[1] + 1, nz = g.n[2] + 1;
    double gm = 0.0, gmdv = 0.0, gv2 = 0.0, gw2 = 0.0, gbad = 0.0;
    for (int t = blockIdx.x * 4 + wave; t < ntiles; t += gridDim.x * 4) {
        const int row = t * kMRows + col;
        const int q = row < nrows ? nodes[row] : -1;
        int i = 0, j = 0, k = 0;
        if (q >= 0) {
            i = q % nx;
            j = (q / nx) % ny;
            k = q / (nx * ny);
        }
        // features 16 ks + 8 h + e: neighbours 4 ks + 2 h and 4 ks + 2 h + 1,
        // columns 4 nb + (0: m / m_ref, 1..3: m v / (m_ref v_ref))
        bf16x8 xa[8];
#pragma unroll
        for (int ks = 0; ks < 8; ++ks) {
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                const int nb = 4 * ks + 2 * h + u;
                const int a = i + nb % 3 - 1, b = j + (nb / 3) % 3 - 1, cz = k + nb / 9 - 1;
                const bool in = q >= 0 && nb < 27 && a >= 0 && a < nx && b >= 0 && b < ny &&
                                cz >= 0 && cz < nz;
                const int64_t p = in ? a + (int64_t)nx * (b + (int64_t)ny * cz) : 0;
                uint2 wv = in ? feat[p] : make_uint2(0u, 0u);
                __bf16 f[4];
                __builtin_memcpy(f, &wv, sizeof(wv));
#pragma unroll
                for (int e = 0; e < 4; ++e) xa[ks][4 * u + e] = f[e];
            }
        }
        bf16x8 h1[2][2], h2[2][2];
        f32x16 y;
        forward_tile<false>(w, xa, b1, b2, b3, p0, p1, h1, h2, nullptr, nullptr, y, lane);
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int rr = t * kMRows + crow(r, lane);
            const int qr = __shfl(q, crow(r, lane), kWave);
            if (col < kMOut && rr < nrows) {
                Y[(int64_t)rr * kMOut + col] = y[r];
                const double m = N.mass[qr], v = N.vel[col * g.nn + qr];
                const double dv = (double)y[r] * sc.dv_ref;
                if (!isfinite(dv)) gbad += 1.0;
                if (col == 0) gm += m;
                gmdv += m * dv;
                gv2 += m * v * v;
                gw2 += m * (v + dv) * (v + dv);
            }
        }
    }
    // wavefront sums: gmdv per column (the two halves), the rest over all lanes
    gmdv += __shfl_xor(gmdv, 32, kWave);
    for (int off = 32; off > 0; off >>= 1) {
        gm += __shfl_xor(gm, off, kWave);
        gv2 += __shfl_xor(gv2, off, kWave);
        gw2 += __shfl_xor(gw2, off, kWave);
        gbad += __shfl_xor(gbad, off, kWave);
    }
    __shared__ double red[7][4];
    if (lane < kMOut) red[1 + lane][wave] = gmdv;
    if (lane == 0) {
        red[0][wave] = gm;
        red[4][wave] = gv2;
        red[5][wave] = gw2;
        red[6][wave] = gbad;
    }
    __syncthreads();
    if (threadIdx.x < 7)
        sums[blockIdx.x * 8 + threadIdx.x] = red[threadIdx.x][0] + red[threadIdx.x][1] +
                                             red[threadIdx.x][2] + red[threadIdx.x][3];
}

__global__ void k_zero_int(int *p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

}  // namespace

void launch_mlp_features(const GridP &g, const NState &N, const int *nodes, int nrows,
                         MlpScales sc, __bf16 *X, hipStream_t s) {
    const int pad = (nrows + kMRows - 1) / kMRows * kMRows;
    if (pad == 0) return;
    const int64_t nt = (int64_t)pad * (kMInPad / 8);
    hipLaunchKernelGGL(k_mlp_features, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, s, g, N,
                       nodes, nrows, sc, X);
}

void launch_mlp_infer(const MlpDev &w, const __bf16 *X, int nrows, float *Y, hipStream_t s) {
    if (nrows <= 0) return;
    const int ntiles = (nrows + kMRows - 1) / kMRows;
    int blocks = (ntiles + 3) / 4;
    if (blocks > 2048) blocks = 2048;
    hipLaunchKernelGGL(k_mlp_infer, dim3(blocks), dim3(256), 0, s, w, X, nrows, Y);
}

int launch_mlp_train(const MlpDev &w, const __bf16 *X, const float *T, const int *idx, int nrows,
                     float inv_t, hipStream_t s) {
    if (nrows <= 0) return 0;
    const int ntiles = (nrows + kMRows - 1) / kMRows;
    // one tile per wavefront, at most kMTrainParts tiles (partial rows) per
    // launch; the launches after the first add into the same rows
    for (int t0 = 0; t0 < ntiles; t0 += kMTrainParts) {
        const int n = ntiles - t0 < kMTrainParts ? ntiles - t0 : kMTrainParts;
        const dim3 grid((n + kTrainWaves - 1) / kTrainWaves), block(kTrainWaves * kWave);
        if (t0 == 0)
            hipLaunchKernelGGL(k_mlp_train<false>, grid, block, 0, s, w, X, T, idx, nrows, inv_t,
                               1.0f / (float)nrows, t0);
        else
            hipLaunchKernelGGL(k_mlp_train<true>, grid, block, 0, s, w, X, T, idx, nrows, inv_t,
                               1.0f / (float)nrows, t0);
    }
    return ntiles < kMTrainParts ? ntiles : kMTrainParts;
}

void launch_mlp_reduce(const MlpDev &w, int nparts, hipStream_t s) {
    hipLaunchKernelGGL(k_mlp_reduce, dim3((kMParams + 1 + 63) / 64), dim3(64, 4), 0, s, w, nparts);
}

void launch_mlp_adam(const MlpDev &w, int nparts, float lr, float beta1, float beta2, float eps,
                     float c1, float c2, hipStream_t s) {
    if (nparts > 0)
        hipLaunchKernelGGL(k_mlp_adam_parts, dim3((kMParams + 1 + 63) / 64), dim3(64, 4), 0, s, w,
                           nparts, lr, beta1, beta2, eps, c1, c2);
    else
        hipLaunchKernelGGL(k_mlp_adam, dim3((kMParams + 255) / 256), dim3(256), 0, s, w, lr, beta1,
                           beta2, eps, c1, c2);
}

void launch_mlp_pack(const MlpDev &w, hipStream_t s) {
    hipLaunchKernelGGL(k_mlp_pack, dim3((kMParams + 255) / 256), dim3(256), 0, s, w);
}

void launch_active_nodes(const GridP &g, const NState &N, double tol, int *list, int *count,
                         int *scratch, hipStream_t s) {
    (void)scratch;
    hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, s, count, 1);
    hipLaunchKernelGGL(k_active_nodes, dim3((unsigned)((g.nn + kActiveSpan - 1) / kActiveSpan)),
                       dim3(256), 0, s, g, N, tol, list, count, (int *)nullptr, (uint2 *)nullptr,
                       MlpScales{1.0, 1.0, 1.0});
}

void launch_active_nodes_rolling(const GridP &g, const NState &N, double tol, int *list,
                                 int *count, int *next_count, void *feat, MlpScales sc,
                                 hipStream_t s) {
    hipLaunchKernelGGL(k_active_nodes, dim3((unsigned)((g.nn + kActiveSpan - 1) / kActiveSpan)),
                       dim3(256), 0, s, g, N, tol, list, count, next_count,
                       reinterpret_cast<uint2 *>(feat), sc);
}

void launch_mlp_infer_nodes(const MlpDev &w, const GridP &g, const NState &N, const void *feat,
                            const int *nodes, const int *count, MlpScales sc, float *Y,
                            double *sums, hipStream_t s) {
    hipLaunchKernelGGL(k_mlp_infer_nodes, dim3(kInferBlocks), dim3(256), 0, s, w, g, N,
                       reinterpret_cast<const uint2 *>(feat), nodes, count, sc, Y, sums);
}

}  // namespace xmpm

// mpm_mlp.hip — kernels of the learned coarse-step correction for gfx950.
//
// One wavefront works on a tile of 32 rows (nodes) with
// v_mfma_f32_32x32x16_bf16.  Operand layouts of that instruction, lane l:
//   A (32 x 16):  row l & 31, k = 8 (l >> 5) + j, j = 0..7 (8 BF16 per lane)
//   B (16 x 32):  col l & 31, k = 8 (l >> 5) + j
//   C (32 x 32):  col l & 31, row = 8 (r >> 2) + 4 (l >> 5) + (r & 3), r = 0..15
// so every operand is read as 16 bytes of k-contiguous data: activations are
// kept row-major [row][k] (A) and weights transposed [n][k] (B); for the
// weight gradients (k = rows) the activations and backward signals are also
// written transposed [feature][row] to LDS.
//
// Forward: Z1 = X W1 + b1, H1 = gelu(Z1); Z2 = H1 W2 + b2, H2 = gelu(Z2);
// Y = H2 W3 + b3.  Backward of L = (1/N) sum (Y - T)^2:
// dY = 2 (Y - T) / N; dZ2 = (dY W3^T) gelu'(Z2); dZ1 = (dZ2 W2^T) gelu'(Z1);
// dW3 = H2^T dY, dW2 = H1^T dZ2, dW1 = X^T dZ1, db = column sums.
// X, the weights, H1, H2, dY, dZ2, dZ1 enter the products as BF16 (round to
// nearest even); products accumulate in FP32.
#include <hip/hip_runtime.h>

#include "mpm_mlp.h"

namespace xmpm {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kWave = 64;

__device__ __forceinline__ bf16x8 ld8(const __bf16 *p) {
    return *reinterpret_cast<const bf16x8 *>(p);
}
// A operand of rows [row][k] with leading dimension ld, k-step ks (16 wide).
__device__ __forceinline__ bf16x8 opA(const __bf16 *base, int ld, int ks, int lane) {
    return ld8(base + (lane & 31) * ld + ks * 16 + 8 * (lane >> 5));
}
// B operand from an [n][k] array, n-tile nt (32 wide), k-step ks.
__device__ __forceinline__ bf16x8 opB(const __bf16 *base, int ld, int ks, int nt, int lane) {
    return ld8(base + (nt * 32 + (lane & 31)) * ld + ks * 16 + 8 * (lane >> 5));
}
__device__ __forceinline__ int crow(int r, int lane) {
    return 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
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

// tanh-form GELU and its derivative, FP32.
__device__ __forceinline__ float gelu(float z) {
    const float t = tanhf(0.7978845608f * (z + 0.044715f * z * z * z));
    return 0.5f * z * (1.0f + t);
}
__device__ __forceinline__ float gelu_d(float z) {
    const float t = tanhf(0.7978845608f * (z + 0.044715f * z * z * z));
    return 0.5f * (1.0f + t) + 0.5f * z * (1.0f - t * t) * 0.7978845608f * (1.0f + 0.134145f * z * z);
}

// LDS of one wavefront (BF16 tiles of 32 rows).
struct WaveLds {
    __bf16 h1[kMRows * kMH];    // H1 [row][64]
    __bf16 h2[kMRows * kMH];    // H2 [row][64]
    __bf16 h1t[kMH * kMRows];   // H1 [feature][row]
    __bf16 h2t[kMH * kMRows];   // H2 [feature][row]
    __bf16 xt[kMInPad * kMRows];  // X [feature][row]
    __bf16 dy[kMRows * 16];     // dY [row][16]
    __bf16 dyt[32 * kMRows];    // dY [out][row], outputs padded to 32
    __bf16 dz2[kMRows * kMH];   // dZ2 [row][64]
    __bf16 dz2t[kMH * kMRows];  // dZ2 [feature][row]
    __bf16 dz1t[kMH * kMRows];  // dZ1 [feature][row]
    float gd1[kMRows * kMH];    // gelu'(Z1) [row][64]
    float gd2[kMRows * kMH];    // gelu'(Z2) [row][64]
};

struct Fwd {
    f32x16 y;
};

// Forward pass of one tile: X rows at xrow (row-major, ld 128), into the
// pre-activations (registers) and H1/H2 in LDS (row-major; transposed too
// when TRAIN).
template <bool TRAIN>
__device__ __forceinline__ void forward_tile(const MlpDev &w, const __bf16 *xrow, __bf16 *h1,
                                             __bf16 *h2, __bf16 *h1t, __bf16 *h2t, float *gd1,
                                             float *gd2, Fwd &F, int lane) {
    const int col = lane & 31;
#pragma unroll
    for (int nt = 0; nt < 2; ++nt) {
        f32x16 acc = zero16();
#pragma unroll
        for (int ks = 0; ks < kMInPad / 16; ++ks)
            acc = mfma(opA(xrow, kMInPad, ks, lane), opB(w.w1t, kMInPad, ks, nt, lane), acc);
        const float b = w.params[kOffB1 + nt * 32 + col];
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const float z = acc[r] + b;
            acc[r] = z;
            const __bf16 h = (__bf16)gelu(z);
            const int row = crow(r, lane);
            h1[row * kMH + nt * 32 + col] = h;
            if (TRAIN) {
                h1t[(nt * 32 + col) * kMRows + row] = h;
                gd1[row * kMH + nt * 32 + col] = gelu_d(z);
            }
        }
    }
    asm volatile("" ::: "memory");  // LDS of one wavefront is ordered; keep the compiler's order
#pragma unroll
    for (int nt = 0; nt < 2; ++nt) {
        f32x16 acc = zero16();
#pragma unroll
        for (int ks = 0; ks < kMH / 16; ++ks)
            acc = mfma(opA(h1, kMH, ks, lane), opB(w.w2t, kMH, ks, nt, lane), acc);
        const float b = w.params[kOffB2 + nt * 32 + col];
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const float z = acc[r] + b;
            acc[r] = z;
            const __bf16 h = (__bf16)gelu(z);
            const int row = crow(r, lane);
            h2[row * kMH + nt * 32 + col] = h;
            if (TRAIN) {
                h2t[(nt * 32 + col) * kMRows + row] = h;
                gd2[row * kMH + nt * 32 + col] = gelu_d(z);
            }
        }
    }
    asm volatile("" ::: "memory");
    f32x16 acc = zero16();
#pragma unroll
    for (int ks = 0; ks < kMH / 16; ++ks)
        acc = mfma(opA(h2, kMH, ks, lane), opB(w.w3t, kMH, ks, 0, lane), acc);
    const float b = col < kMOut ? w.params[kOffB3 + col] : 0.0f;
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[r] += b;
    F.y = acc;
}

// ---------------------------------------------------------------- inference

__global__ __launch_bounds__(256) void k_mlp_infer(MlpDev w, const __bf16 *X, int nrows,
                                                   const int *nrows_dev, float *Y) {
    __shared__ __bf16 h1s[4][kMRows * kMH], h2s[4][kMRows * kMH];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (nrows_dev) nrows = *nrows_dev;
    const int ntiles = (nrows + kMRows - 1) / kMRows;
    for (int t = blockIdx.x * 4 + wave; t < ntiles; t += gridDim.x * 4) {
        Fwd F;
        forward_tile<false>(w, X + (int64_t)t * kMRows * kMInPad, h1s[wave], h2s[wave], nullptr,
                            nullptr, nullptr, nullptr, F, lane);
        const int col = lane & 31;
        if (col < kMOut)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int row = t * kMRows + crow(r, lane);
                if (row < nrows) Y[(int64_t)row * kMOut + col] = F.y[r];
            }
        asm volatile("" ::: "memory");  // next tile's LDS writes after this tile's reads
    }
}

// ---------------------------------------------------------------- training

// Per workgroup (one wavefront): tiles of 32 batch rows; the gradient of the
// workgroup's tiles accumulates in LDS (FP32 atomics) and is added to the
// global gradient once at the end.
__global__ __launch_bounds__(kWave) void k_mlp_train(MlpDev w, const __bf16 *X, const float *T,
                                                     const int *idx, int nrows, float inv_t,
                                                     float inv_n) {
    __shared__ WaveLds L;
    __shared__ __bf16 xtile[kMRows * kMInPad];
    __shared__ float g[kMParams];
    const int lane = threadIdx.x, col = lane & 31;
    for (int e = lane; e < kMParams; e += kWave) g[e] = 0.0f;
    float lsum = 0.0f;
    const int ntiles = (nrows + kMRows - 1) / kMRows;
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        __syncthreads();  // previous tile's LDS reads done
        // gather the tile's rows (rows past nrows are zero)
        for (int e = lane; e < kMRows * (kMInPad / 8); e += kWave) {
            const int row = e / (kMInPad / 8), c8 = e % (kMInPad / 8);
            const int gr = t * kMRows + row;
            bf16x8 v;
            if (gr < nrows) {
                v = ld8(X + (int64_t)idx[gr] * kMInPad + c8 * 8);
            } else {
#pragma unroll
                for (int j = 0; j < 8; ++j) v[j] = (__bf16)0.0f;
            }
            *reinterpret_cast<bf16x8 *>(xtile + row * kMInPad + c8 * 8) = v;
#pragma unroll
            for (int j = 0; j < 8; ++j) L.xt[(c8 * 8 + j) * kMRows + row] = v[j];
        }
        __syncthreads();
        Fwd F;
        forward_tile<true>(w, xtile, L.h1, L.h2, L.h1t, L.h2t, L.gd1, L.gd2, F, lane);
        // dY = 2 (Y - T) / N on valid rows and outputs
        float db3 = 0.0f;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int row = crow(r, lane);
            const int gr = t * kMRows + row;
            float d = 0.0f;
            if (col < kMOut && gr < nrows) {
                const float e = F.y[r] - T[(int64_t)idx[gr] * kMOut + col] * inv_t;
                lsum += e * e;
                d = 2.0f * e * inv_n;
            }
            const __bf16 db = (__bf16)d;
            db3 += (float)db;
            if (col < 16) L.dy[row * 16 + col] = db;
            L.dyt[col * kMRows + row] = db;
        }
        asm volatile("" ::: "memory");
        // dZ2 = (dY W3^T) gelu'(Z2)
        float db2[2] = {0.0f, 0.0f};
#pragma unroll
        for (int nt = 0; nt < 2; ++nt) {
            f32x16 acc = mfma(opA(L.dy, 16, 0, lane), opB(w.w3, 16, 0, nt, lane), zero16());
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int row = crow(r, lane);
                const __bf16 d = (__bf16)(acc[r] * L.gd2[row * kMH + nt * 32 + col]);
                db2[nt] += (float)d;
                L.dz2[row * kMH + nt * 32 + col] = d;
                L.dz2t[(nt * 32 + col) * kMRows + row] = d;
            }
        }
        asm volatile("" ::: "memory");
        // dZ1 = (dZ2 W2^T) gelu'(Z1)
        float db1[2] = {0.0f, 0.0f};
#pragma unroll
        for (int nt = 0; nt < 2; ++nt) {
            f32x16 acc = zero16();
#pragma unroll
            for (int ks = 0; ks < kMH / 16; ++ks)
                acc = mfma(opA(L.dz2, kMH, ks, lane), opB(w.w2, kMH, ks, nt, lane), acc);
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int row = crow(r, lane);
                const __bf16 d = (__bf16)(acc[r] * L.gd1[row * kMH + nt * 32 + col]);
                db1[nt] += (float)d;
                L.dz1t[(nt * 32 + col) * kMRows + row] = d;
            }
        }
        asm volatile("" ::: "memory");
        // weight gradients: K = the tile's 32 rows (two 16-row steps)
#pragma unroll 1
        for (int mt = 0; mt < kMInPad / 32; ++mt)
#pragma unroll 1
            for (int nt = 0; nt < 2; ++nt) {
                f32x16 acc = zero16();
#pragma unroll
                for (int ks = 0; ks < 2; ++ks)
                    acc = mfma(opA(L.xt + mt * 32 * kMRows, kMRows, ks, lane),
                               opB(L.dz1t, kMRows, ks, nt, lane), acc);
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    const int in = mt * 32 + crow(r, lane);
                    if (in < kMIn) atomicAdd(&g[kOffW1 + in * kMH + nt * 32 + col], acc[r]);
                }
            }
#pragma unroll 1
        for (int mt = 0; mt < 2; ++mt)
#pragma unroll 1
            for (int nt = 0; nt < 2; ++nt) {
                f32x16 acc = zero16();
#pragma unroll
                for (int ks = 0; ks < 2; ++ks)
                    acc = mfma(opA(L.h1t + mt * 32 * kMRows, kMRows, ks, lane),
                               opB(L.dz2t, kMRows, ks, nt, lane), acc);
#pragma unroll
                for (int r = 0; r < 16; ++r)
                    atomicAdd(&g[kOffW2 + (mt * 32 + crow(r, lane)) * kMH + nt * 32 + col], acc[r]);
            }
#pragma unroll 1
        for (int mt = 0; mt < 2; ++mt) {
            f32x16 acc = zero16();
#pragma unroll
            for (int ks = 0; ks < 2; ++ks)
                acc = mfma(opA(L.h2t + mt * 32 * kMRows, kMRows, ks, lane),
                           opB(L.dyt, kMRows, ks, 0, lane), acc);
            if (col < kMOut)
#pragma unroll
                for (int r = 0; r < 16; ++r)
                    atomicAdd(&g[kOffW3 + (mt * 32 + crow(r, lane)) * kMOut + col], acc[r]);
        }
        // bias gradients: the two lane halves hold different rows of a column
#pragma unroll
        for (int nt = 0; nt < 2; ++nt) {
            db1[nt] += __shfl_xor(db1[nt], 32, kWave);
            db2[nt] += __shfl_xor(db2[nt], 32, kWave);
            if (lane < 32) {
                atomicAdd(&g[kOffB1 + nt * 32 + col], db1[nt]);
                atomicAdd(&g[kOffB2 + nt * 32 + col], db2[nt]);
            }
        }
        db3 += __shfl_xor(db3, 32, kWave);
        if (lane < kMOut) atomicAdd(&g[kOffB3 + col], db3);
    }
    __syncthreads();
    for (int e = lane; e < kMParams; e += kWave)
        if (g[e] != 0.0f) atomicAdd(&w.grad[e], g[e]);
    for (int off = 32; off > 0; off >>= 1) lsum += __shfl_xor(lsum, off, kWave);
    if (lane == 0 && lsum != 0.0f) atomicAdd(w.loss, (double)lsum);
}

// BF16 operand copies of parameter e.
__device__ __forceinline__ void pack_one(const MlpDev &w, int e, float p) {
    const __bf16 b = (__bf16)p;
    if (e < kOffB1) {
        const int k = (e - kOffW1) / kMH, n = (e - kOffW1) % kMH;
        w.w1t[n * kMInPad + k] = b;
    } else if (e >= kOffW2 && e < kOffB2) {
        const int k = (e - kOffW2) / kMH, n = (e - kOffW2) % kMH;
        w.w2t[n * kMH + k] = b;
        w.w2[k * kMH + n] = b;
    } else if (e >= kOffW3 && e < kOffB3) {
        const int k = (e - kOffW3) / kMOut, n = (e - kOffW3) % kMOut;
        w.w3t[n * kMH + k] = b;
        w.w3[k * 16 + n] = b;
    }
}

__global__ void k_mlp_adam(MlpDev w, float lr, float beta1, float beta2, float eps, float c1,
                           float c2) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= kMParams) return;
    const float gr = w.grad[e];
    const float m = beta1 * w.m[e] + (1.0f - beta1) * gr;
    const float v = beta2 * w.v[e] + (1.0f - beta2) * gr * gr;
    w.m[e] = m;
    w.v[e] = v;
    const float mh = m / c1, vh = v / c2;
    const float p = w.params[e] - lr * mh / (sqrtf(vh) + eps);
    w.params[e] = p;
    w.grad[e] = 0.0f;
    pack_one(w, e, p);
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
                               const int *nrows_dev, MlpScales sc, __bf16 *X) {
    if (nrows_dev) nrows = *nrows_dev;
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

// Active nodes (mass >= tol): one atomic per workgroup reserves its run.
// next_count (if given) is zeroed for the following step's list.
__global__ __launch_bounds__(256) void k_active_nodes(GridP g, NState N, double tol, int *list,
                                                      int *count, int *next_count) {
    __shared__ int wsum[4];
    __shared__ int base;
    const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (next_count && blockIdx.x == 0 && threadIdx.x == 0) *next_count = 0;
    const bool f = q < g.nn && N.mass[q] >= tol;
    const unsigned long long b = __ballot(f);
    const int inwave = __popcll(b & ((1ull << lane) - 1ull));
    if (lane == 0) wsum[w] = __popcll(b);
    __syncthreads();
    int off = 0, tot = 0;
    for (int k = 0; k < 4; ++k) {
        if (k < w) off += wsum[k];
        tot += wsum[k];
    }
    if (threadIdx.x == 0) base = tot ? atomicAdd(count, tot) : 0;
    __syncthreads();
    if (f) list[base + off + inwave] = (int)q;
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
                       nodes, nrows, (const int *)nullptr, sc, X);
}

void launch_mlp_features_dev(const GridP &g, const NState &N, const int *nodes, const int *count,
                             MlpScales sc, __bf16 *X, hipStream_t s) {
    hipLaunchKernelGGL(k_mlp_features, dim3(2048), dim3(256), 0, s, g, N, nodes, 0, count, sc, X);
}

void launch_mlp_infer(const MlpDev &w, const __bf16 *X, int nrows, float *Y, hipStream_t s) {
    if (nrows <= 0) return;
    const int ntiles = (nrows + kMRows - 1) / kMRows;
    int blocks = (ntiles + 3) / 4;
    if (blocks > 1024) blocks = 1024;
    hipLaunchKernelGGL(k_mlp_infer, dim3(blocks), dim3(256), 0, s, w, X, nrows,
                       (const int *)nullptr, Y);
}

void launch_mlp_infer_dev(const MlpDev &w, const __bf16 *X, const int *count, float *Y,
                          hipStream_t s) {
    hipLaunchKernelGGL(k_mlp_infer, dim3(1024), dim3(256), 0, s, w, X, 0, count, Y);
}

void launch_mlp_train(const MlpDev &w, const __bf16 *X, const float *T, const int *idx,
                      int nrows, float inv_t, hipStream_t s) {
    if (nrows <= 0) return;
    const int ntiles = (nrows + kMRows - 1) / kMRows;
    int blocks = (ntiles + 3) / 4;  // four tiles per workgroup
    if (blocks > 512) blocks = 512;
    hipLaunchKernelGGL(k_mlp_train, dim3(blocks), dim3(kWave), 0, s, w, X, T, idx, nrows, inv_t,
                       1.0f / (float)nrows);
}

void launch_mlp_adam(const MlpDev &w, float lr, float beta1, float beta2, float eps, float c1,
                     float c2, hipStream_t s) {
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
    hipLaunchKernelGGL(k_active_nodes, dim3((unsigned)((g.nn + 255) / 256)), dim3(256), 0, s, g,
                       N, tol, list, count, (int *)nullptr);
}

void launch_active_nodes_rolling(const GridP &g, const NState &N, double tol, int *list,
                                 int *count, int *next_count, hipStream_t s) {
    hipLaunchKernelGGL(k_active_nodes, dim3((unsigned)((g.nn + 255) / 256)), dim3(256), 0, s, g,
                       N, tol, list, count, next_count);
}

}  // namespace xmpm

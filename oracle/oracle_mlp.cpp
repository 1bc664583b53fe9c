// oracle_mlp.cpp — CPU restatement of the learned coarse-step correction.
//
// TEST INFRASTRUCTURE ONLY: the checker for the HIP MLP kernels
// (exagoop_amd/csrc/mpm_mlp.hip).  Only tests/ load it; the product never
// does.  The learned correction has no counterpart in the reference: it is
// the stencil MLP this project adds beside the plain step.  Its
// definition (features, network, rounding points, loss, optimiser) lives
// here and in the HIP kernels; each function below states it.
//
// Network: per active node, 27 neighbour nodes (dz, dy, dx in {-1, 0, 1},
// dx fastest) x (m / m_ref, m v / (m_ref v_ref)) = 108 inputs, zero for
// neighbours outside the nodal box, padded with zeros to 128;
// 108 -> 64 -> 64 -> 3 with tanh-form GELU; the output is the nodal velocity
// correction in units of dv_ref.  Parameters (FP32 master copy, in this
// order): W1[108][64], b1[64], W2[64][64], b2[64], W3[64][3], b3[3]
// (W[in][out]), 11 331 values.
//
// Rounding points of the BF16 MFMA path (mode exact = 0): inputs X, all
// weights, the hidden activations H1 and H2 and the backward signals dY,
// dZ2, dZ1 are rounded to bfloat16 (round to nearest even) before they enter
// a product; every product sum and bias is FP32.  exact = 1 skips every
// rounding and computes in double (pins the backward pass against finite
// differences).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

constexpr int kIn = 108, kInPad = 128, kH = 64, kOut = 3;
constexpr int kOffW1 = 0, kOffB1 = kOffW1 + kIn * kH, kOffW2 = kOffB1 + kH,
              kOffB2 = kOffW2 + kH * kH, kOffW3 = kOffB2 + kH, kOffB3 = kOffW3 + kH * kOut,
              kNParams = kOffB3 + kOut;
static_assert(kNParams == 11331, "parameter count");

// float -> bfloat16 -> float, round to nearest even (NaN kept quiet).
float bf16r(float x) {
    uint32_t u;
    std::memcpy(&u, &x, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) {  // inf / nan
        if (u & 0x007fffffu) u |= 0x00400000u;
        u &= 0xffff0000u;
    } else {
        const uint32_t lsb = (u >> 16) & 1u;
        u += 0x7fffu + lsb;
        u &= 0xffff0000u;
    }
    float r;
    std::memcpy(&r, &u, 4);
    return r;
}

struct Mode {
    bool exact;
    double r(double v) const { return exact ? v : (double)bf16r((float)v); }
    // sums are FP32 in the MFMA path, double when exact
    double s(double v) const { return exact ? v : (double)(float)v; }
};

constexpr double kGeluC = 0.7978845608028654;  // sqrt(2 / pi)
constexpr double kGeluA = 0.044715;

double gelu(double z, bool exact) {
    if (exact) return 0.5 * z * (1.0 + std::tanh(kGeluC * (z + kGeluA * z * z * z)));
    const float zf = (float)z;
    const float t = std::tanh(0.7978845608f * (zf + 0.044715f * zf * zf * zf));
    return (double)(0.5f * zf * (1.0f + t));
}
double gelu_d(double z, bool exact) {
    if (exact) {
        const double t = std::tanh(kGeluC * (z + kGeluA * z * z * z));
        return 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t * t) * kGeluC * (1.0 + 3.0 * kGeluA * z * z);
    }
    const float zf = (float)z;
    const float t = std::tanh(0.7978845608f * (zf + 0.044715f * zf * zf * zf));
    return (double)(0.5f * (1.0f + t) +
                    0.5f * zf * (1.0f - t * t) * 0.7978845608f * (1.0f + 0.134145f * zf * zf));
}

struct Net {
    std::vector<double> W1, b1, W2, b2, W3, b3;
    Net(const float *p, const Mode &M)
        : W1(kIn * kH), b1(kH), W2(kH * kH), b2(kH), W3(kH * kOut), b3(kOut) {
        for (int i = 0; i < kIn * kH; ++i) W1[i] = M.r(p[kOffW1 + i]);
        for (int i = 0; i < kH; ++i) b1[i] = p[kOffB1 + i];
        for (int i = 0; i < kH * kH; ++i) W2[i] = M.r(p[kOffW2 + i]);
        for (int i = 0; i < kH; ++i) b2[i] = p[kOffB2 + i];
        for (int i = 0; i < kH * kOut; ++i) W3[i] = M.r(p[kOffW3 + i]);
        for (int i = 0; i < kOut; ++i) b3[i] = p[kOffB3 + i];
    }
};

struct Act {
    double x[kIn], z1[kH], h1[kH], z2[kH], h2[kH], y[kOut];
};

void forward_row(const Net &N, const Mode &M, const float *X, Act &A) {
    for (int k = 0; k < kIn; ++k) A.x[k] = M.r(X[k]);
    for (int n = 0; n < kH; ++n) {
        double acc = 0.0;
        for (int k = 0; k < kIn; ++k) acc += A.x[k] * N.W1[k * kH + n];
        A.z1[n] = M.s(M.s(acc) + N.b1[n]);
        A.h1[n] = M.r(gelu(A.z1[n], M.exact));
    }
    for (int n = 0; n < kH; ++n) {
        double acc = 0.0;
        for (int k = 0; k < kH; ++k) acc += A.h1[k] * N.W2[k * kH + n];
        A.z2[n] = M.s(M.s(acc) + N.b2[n]);
        A.h2[n] = M.r(gelu(A.z2[n], M.exact));
    }
    for (int n = 0; n < kOut; ++n) {
        double acc = 0.0;
        for (int k = 0; k < kH; ++k) acc += A.h2[k] * N.W3[k * kOut + n];
        A.y[n] = M.s(M.s(acc) + N.b3[n]);
    }
}

}  // namespace

extern "C" {

int orc_mlp_nparams(void) { return kNParams; }

float orc_bf16_round(float x) { return bf16r(x); }

// Features of the listed nodes from the nodal mass and velocity
// (component-major, nn = (n0+1)(n1+1)(n2+1) nodes, i fastest):
// X[row][4 nb + c], nb = (dz+1)*9 + (dy+1)*3 + (dx+1); c = 0: m / m_ref,
// c = 1..3: m v_c / (m_ref v_ref).  FP32 values (rounded to bf16 by the
// network).  Columns 108..127 are zero.
void orc_mlp_features(const double *mass, const double *vel, const int *ncells, int64_t nn,
                      const int64_t *nodes, int64_t nrows, double m_ref, double v_ref,
                      float *X) {
    const int nx = ncells[0] + 1, ny = ncells[1] + 1, nz = ncells[2] + 1;
    const double im = 1.0 / m_ref, imv = 1.0 / (m_ref * v_ref);
    for (int64_t r = 0; r < nrows; ++r) {
        const int64_t q = nodes[r];
        const int i = (int)(q % nx), j = (int)((q / nx) % ny), k = (int)(q / ((int64_t)nx * ny));
        float *x = X + r * kInPad;
        for (int c = 0; c < kInPad; ++c) x[c] = 0.0f;
        for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    const int a = i + dx, b = j + dy, c = k + dz;
                    if (a < 0 || a >= nx || b < 0 || b >= ny || c < 0 || c >= nz) continue;
                    const int64_t p = a + (int64_t)nx * (b + (int64_t)ny * c);
                    const int nb = (dz + 1) * 9 + (dy + 1) * 3 + (dx + 1);
                    const double m = mass[p];
                    x[4 * nb] = (float)(m * im);
                    for (int d = 0; d < 3; ++d) x[4 * nb + 1 + d] = (float)(m * vel[d * nn + p] * imv);
                }
    }
    (void)nn;
}

// Network outputs Y[row][3] for rows X[row][128] (only the first 108 used).
void orc_mlp_forward(const float *params, int64_t nrows, const float *X, int exact, float *Y) {
    const Mode M{exact != 0};
    const Net N(params, M);
    Act A;
    for (int64_t r = 0; r < nrows; ++r) {
        forward_row(N, M, X + r * kInPad, A);
        for (int n = 0; n < kOut; ++n) Y[r * kOut + n] = (float)A.y[n];
    }
}

// Loss L = (1/nrows) sum_rows sum_c (Y - T)^2 and its gradient with respect
// to every parameter (same layout as params).  dY = 2 (Y - T) / nrows.
double orc_mlp_grad(const float *params, int64_t nrows, const float *X, const float *T,
                    int exact, double *grad) {
    const Mode M{exact != 0};
    const Net N(params, M);
    std::vector<double> g(kNParams, 0.0);
    double loss = 0.0;
    Act A;
    for (int64_t r = 0; r < nrows; ++r) {
        forward_row(N, M, X + r * kInPad, A);
        double dy[kOut], dz2[kH], dz1[kH];
        for (int n = 0; n < kOut; ++n) {
            const double e = A.y[n] - (double)T[r * kOut + n];
            loss += e * e;
            dy[n] = M.r(M.s(2.0 * e / (double)nrows));
        }
        for (int k = 0; k < kH; ++k)
            for (int n = 0; n < kOut; ++n) g[kOffW3 + k * kOut + n] += A.h2[k] * dy[n];
        for (int n = 0; n < kOut; ++n) g[kOffB3 + n] += dy[n];
        for (int k = 0; k < kH; ++k) {
            double acc = 0.0;
            for (int n = 0; n < kOut; ++n) acc += dy[n] * N.W3[k * kOut + n];
            dz2[k] = M.r(M.s(M.s(acc) * gelu_d(A.z2[k], M.exact)));
        }
        for (int k = 0; k < kH; ++k)
            for (int n = 0; n < kH; ++n) g[kOffW2 + k * kH + n] += A.h1[k] * dz2[n];
        for (int n = 0; n < kH; ++n) g[kOffB2 + n] += dz2[n];
        for (int k = 0; k < kH; ++k) {
            double acc = 0.0;
            for (int n = 0; n < kH; ++n) acc += dz2[n] * N.W2[k * kH + n];
            dz1[k] = M.r(M.s(M.s(acc) * gelu_d(A.z1[k], M.exact)));
        }
        for (int k = 0; k < kIn; ++k)
            for (int n = 0; n < kH; ++n) g[kOffW1 + k * kH + n] += A.x[k] * dz1[n];
        for (int n = 0; n < kH; ++n) g[kOffB1 + n] += dz1[n];
    }
    for (int i = 0; i < kNParams; ++i) grad[i] = g[i];
    return loss / (double)nrows;
}

// One Adam step on the FP32 master parameters (t = 1, 2, ...).
void orc_mlp_adam(float *params, float *m, float *v, const float *grad, int n, float lr,
                  float beta1, float beta2, float eps, int t) {
    const float c1 = 1.0f - std::pow(beta1, (float)t), c2 = 1.0f - std::pow(beta2, (float)t);
    for (int i = 0; i < n; ++i) {
        m[i] = beta1 * m[i] + (1.0f - beta1) * grad[i];
        v[i] = beta2 * v[i] + (1.0f - beta2) * grad[i] * grad[i];
        const float mh = m[i] / c1, vh = v[i] / c2;
        params[i] -= lr * mh / (std::sqrt(vh) + eps);
    }
}

}  // extern "C"

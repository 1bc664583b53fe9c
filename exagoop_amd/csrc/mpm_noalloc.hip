// mpm_noalloc.hip — the energy and momentum sums of mpm_energies over
// particle arrays the caller holds in device memory, without allocating.
//
// Calculate_Total_Energies / Calculate_MWA_VelocityComponents
// (mpm_diagnostics.cpp:23-127) reduce over every particle.  Here the sums run
// in two launches: each workgroup sums a fixed chunk of particles into one
// row of per-workgroup partial sums, then one workgroup sums the rows in a
// fixed order.  The partial sums live in `scratch`, a file-scope device
// array; the library allocates no device memory for this entry point, so it
// can be called where hipMalloc cannot (under HIP graph capture, in serving
// loops).  The result does not depend on scheduling: the same inputs give the
// same bits on every call.

#include <mutex>

#include "mpm_context.h"
#include "mpm_math.h"

namespace {

constexpr int kEnBlock = 256;
constexpr int kEnPerThread = 4;
constexpr int64_t kEnChunk = (int64_t)kEnBlock * kEnPerThread;  // particles per workgroup
constexpr int64_t kEnMaxRows = MPM_ENERGIES_NOALLOC_MAX_PARTICLES / kEnChunk;
constexpr int kEnSums = 6;
static_assert(MPM_ENERGIES_NOALLOC_MAX_PARTICLES % kEnChunk == 0, "whole chunks");

// Per-workgroup partial sums, row b = workgroup b: 16384 x 6 doubles
// (768 KiB) on each device the library runs on.  Every call writes the rows
// it reads, so nothing has to start at zero.
__device__ __align__(256) double scratch[kEnMaxRows * kEnSums];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_dev[kMaxDevices];  // device address of `scratch`, per device

// TKE, TSE (off-diagonal products counted twice), sum m v, sum m of one
// particle: the summands of mpm_diagnostics.cpp:31-68 (3-D) and :98-112.
__device__ inline void particle_sums(const double *R, int64_t stride, int64_t i, double e[6]) {
    const double m = R[MPM_R_MASS * stride + i];
    double v2 = 0.0;
    for (int d = 0; d < 3; ++d)
        v2 += R[(MPM_R_VEL + d) * stride + i] * R[(MPM_R_VEL + d) * stride + i];
    e[0] += 0.5 * m * v2;
    const double *S = R + MPM_R_STRESS * stride + i;
    const double *E = R + MPM_R_STRAIN * stride + i;
    using xmpm::XX, xmpm::YY, xmpm::ZZ, xmpm::XY, xmpm::YZ, xmpm::XZ;
    const double se = S[XX * stride] * E[XX * stride] + S[YY * stride] * E[YY * stride] +
                      S[ZZ * stride] * E[ZZ * stride] +
                      2.0 * (S[XY * stride] * E[XY * stride] + S[YZ * stride] * E[YZ * stride] +
                             S[XZ * stride] * E[XZ * stride]);
    e[1] += 0.5 * R[MPM_R_VOLUME * stride + i] * se;
    for (int d = 0; d < 3; ++d) e[2 + d] += m * R[(MPM_R_VEL + d) * stride + i];
    e[5] += m;
}

// Tree sum of one value per thread in LDS, fixed pairing; thread 0's result.
__device__ inline void block_sum6(double (*red)[kEnBlock], double e[6]) {
    for (int q = 0; q < kEnSums; ++q) red[q][threadIdx.x] = e[q];
    __syncthreads();
    for (int s = kEnBlock / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int q = 0; q < kEnSums; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + s];
        __syncthreads();
    }
}

// Workgroup b sums particles [b * kEnChunk, (b + 1) * kEnChunk) into row b.
__global__ __launch_bounds__(kEnBlock) void k_energy_rows(int64_t np, int64_t stride,
                                                          const double *__restrict__ R,
                                                          const int64_t *__restrict__ id,
                                                          double *__restrict__ rows) {
    __shared__ double red[kEnSums][kEnBlock];
    double e[6] = {0, 0, 0, 0, 0, 0};
    const int64_t base = (int64_t)blockIdx.x * kEnChunk;
    for (int k = 0; k < kEnPerThread; ++k) {
        const int64_t i = base + (int64_t)k * kEnBlock + threadIdx.x;  // coalesced per k
        if (i < np && (!id || id[i] >= 0)) particle_sums(R, stride, i, e);
    }
    block_sum6(red, e);
    if (threadIdx.x == 0)
        for (int q = 0; q < kEnSums; ++q) rows[(int64_t)blockIdx.x * kEnSums + q] = red[q][0];
}

// One workgroup: thread t sums rows t, t + 256, ... in order, then the tree.
__global__ __launch_bounds__(kEnBlock) void k_energy_final(int64_t nrows,
                                                           const double *__restrict__ rows,
                                                           double *__restrict__ out) {
    __shared__ double red[kEnSums][kEnBlock];
    double e[6] = {0, 0, 0, 0, 0, 0};
    for (int64_t b = threadIdx.x; b < nrows; b += kEnBlock)
        for (int q = 0; q < kEnSums; ++q) e[q] += rows[b * kEnSums + q];
    block_sum6(red, e);
    if (threadIdx.x == 0)
        for (int q = 0; q < kEnSums; ++q) out[q] = red[q][0];
}

// Device address of `scratch` on the current device, looked up once per
// device with that device current.
int scratch_address(double **out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return xmpm::hip_err(e, "hipGetDevice");
    if (dev < 0 || dev >= kMaxDevices) return xmpm::set_err(MPM_E_ARG, "device index too large");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return xmpm::hip_err(e, "hipGetSymbolAddress(scratch)");
        g_scratch_dev[dev] = static_cast<double *>(p);
    }
    *out = g_scratch_dev[dev];
    return MPM_OK;
}

}  // namespace

extern "C" int mpm_energies_noalloc(int64_t np, int64_t stride, const double *rdata,
                                    const int64_t *id, double *out, void *stream) {
    using namespace xmpm;
    if (np < 0 || !out) return set_err(MPM_E_ARG, "mpm_energies_noalloc: np < 0 or null out");
    if (np > MPM_ENERGIES_NOALLOC_MAX_PARTICLES)
        return set_err(MPM_E_ARG,
                       "mpm_energies_noalloc: more particles than its scratch holds "
                       "(MPM_ENERGIES_NOALLOC_MAX_PARTICLES)");
    if (np > 0 && (!rdata || stride < np))
        return set_err(MPM_E_ARG, "mpm_energies_noalloc: null rdata or stride < np");
    double *rows = nullptr;
    int rc = scratch_address(&rows);
    if (rc) return rc;
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int64_t nrows = (np + kEnChunk - 1) / kEnChunk;
    if (nrows > 0)
        hipLaunchKernelGGL(k_energy_rows, dim3((unsigned)nrows), dim3(kEnBlock), 0, st, np,
                           stride, rdata, id, rows);
    hipLaunchKernelGGL(k_energy_final, dim3(1), dim3(kEnBlock), 0, st, nrows, rows, out);
    return check_launch();
}

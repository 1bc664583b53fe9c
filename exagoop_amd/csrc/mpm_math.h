// mpm_math.h — device math for the MI355X explicit MPM step.
//
// B-spline tables and 1-D evaluators, wall-BC projection, stencil bounds and
// constitutive laws.  Arithmetic order follows the ExaGOOP sources cited per
// function (division vs. reciprocal, product and accumulation order), so the
// FP64 results agree with the reference CPU step to rounding of the atomic
// deposit order only.  The library is built with -ffp-contract=off.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/exagoop_amd.h"

namespace xmpm {

// constants.H:10-24 (truncated literals are part of the spline tables).
constexpr double kFourByThree = 1.3333333333;
constexpr double kOneBySix = 0.16666666666667;
constexpr double kOneByThree = 0.333333333333;
constexpr double kTwoByThree = 0.6666666666667;
constexpr double kTiny = 1e-20;

enum { XX = 0, XY = 1, XZ = 2, YY = 3, YZ = 4, ZZ = 5 };

// Spline coefficient tables (interpolants.H:32-170), flattened as
// [type][segment][coef]; type 0 is unused padding.
struct SplineTables {
    double cval[5][4][4];
    double cder[5][4][3];
    double qval[5][4][3];
    double qder[5][4][2];
    int qmap[5][6];
};

__constant__ __device__ static const SplineTables kTab = {
    // cubic values a z^3 + b z^2 + c z + d
    {{{0, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0}},
     {{kOneBySix, 1.0, 2.0, kFourByThree},
      {-kOneBySix, 0.0, 1.0, 1.0},
      {kOneBySix, 0.0, -1.0, 1.0},
      {-kOneBySix, 1.0, -2.0, kFourByThree}},
     {{0.0, 0.0, 0.0, 0.0},
      {-kOneByThree, -1.0, 0.0, kTwoByThree},
      {0.5, -1.0, 0.0, kTwoByThree},
      {-kOneBySix, 1.0, -2.0, kFourByThree}},
     {{kOneBySix, 1.0, 2.0, kFourByThree},
      {-0.5, -1.0, 0.0, kTwoByThree},
      {0.5, -1.0, 0.0, kTwoByThree},
      {-kOneBySix, 1.0, -2.0, kFourByThree}},
     {{kOneBySix, 1.0, 2.0, kFourByThree},
      {-0.5, -1.0, 0.0, kTwoByThree},
      {kOneByThree, -1.0, 0.0, kTwoByThree},
      {0.0, 0.0, 0.0, 0.0}}},
    // cubic derivatives a z^2 + b z + c
    {{{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}},
     {{0.5, 2.0, 2.0}, {-0.5, 0.0, 1.0}, {0.5, 0.0, -1.0}, {-0.5, 2.0, -2.0}},
     {{0, 0, 0}, {-1.0, -2.0, 0.0}, {1.5, -2.0, 0.0}, {-0.5, 2.0, -2.0}},
     {{0.5, 2.0, 2.0}, {-1.5, -2.0, 0.0}, {1.5, -2.0, 0.0}, {-0.5, 2.0, -2.0}},
     {{0.5, 2.0, 2.0}, {-1.5, -2.0, 0.0}, {1.0, -2.0, 0.0}, {0, 0, 0}}},
    // quadratic values a z^2 + b z + c
    {{{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}},
     {{kTwoByThree, 2.0, 1.5},
      {-kFourByThree, 0.0, 1.0},
      {-kFourByThree, 0.0, 1.0},
      {kTwoByThree, -2.0, 1.5}},
     {{4.0 / 3.0, 8.0 / 3.0, 4.0 / 3.0},
      {-7.0 / 6.0, 1.0 / 6.0, 17.0 / 24.0},
      {0.5, -1.5, 1.125},
      {0, 0, 0}},
     {{0.5, 1.5, 1.125}, {-1.0, 0.0, 0.75}, {0.5, -1.5, 1.125}, {0, 0, 0}},
     {{0.5, 1.5, 1.125},
      {-7.0 / 6.0, -1.0 / 6.0, 17.0 / 24.0},
      {4.0 / 3.0, -8.0 / 3.0, 4.0 / 3.0},
      {0, 0, 0}}},
    // quadratic derivatives a z + b
    {{{0, 0}, {0, 0}, {0, 0}, {0, 0}},
     {{4.0 / 3.0, 2.0}, {-8.0 / 3.0, 0.0}, {-8.0 / 3.0, 0.0}, {4.0 / 3.0, -2.0}},
     {{8.0 / 3.0, 8.0 / 3.0}, {-14.0 / 6.0, 1.0 / 6.0}, {1.0, -1.5}, {0, 0}},
     {{1.0, 1.5}, {-2.0, 0.0}, {1.0, -1.5}, {0, 0}},
     {{1.0, 1.5}, {-14.0 / 6.0, -1.0 / 6.0}, {8.0 / 3.0, -8.0 / 3.0}, {0, 0}}},
    // quadratic segment map by floor(2 (z + shift))
    {{-1, -1, -1, -1, -1, -1},
     {0, 0, 1, 2, 3, 3},
     {0, 1, 1, 2, 2, -1},
     {0, 0, 1, 1, 2, 2},
     {0, 0, 1, 1, 2, -1}}};

// interval_idx_quadratic (interpolants.H:187-199).  The evaluators take the
// coefficient table as an argument so a kernel can pass a copy staged in LDS.
__device__ __forceinline__ int quad_seg(const SplineTables &T, int t, double z) {
    double shift = (t == 2) ? 1.0 : 1.5;
    int raw = (int)floor((z + shift) * 2);
    if (raw < 0 || raw >= 6) return -1;
    return T.qmap[t][raw];
}

// quadraticspline_1d / _der (interpolants.H:266-304)
__device__ __forceinline__ double quad_val(const SplineTables &T, int t, double z) {
    double z2 = z * z;
    int s = quad_seg(T, t, z);
    if (s < 0) return 0.0;
    const double *c = T.qval[t][s];
    return c[0] * z2 + c[1] * z + c[2];
}
__device__ __forceinline__ double quad_der(const SplineTables &T, int t, double z) {
    int s = quad_seg(T, t, z);
    if (s < 0) return 0.0;
    const double *c = T.qder[t][s];
    return c[0] * z + c[1];
}

// cubicspline_1d / _der (interpolants.H:215-250); a segment index outside
// [0,3] (|z| >= 2 from rounding at the support edge) evaluates to 0.
__device__ __forceinline__ double cubic_val(const SplineTables &T, int t, double z) {
    double z3 = z * z * z;
    double z2 = z * z;
    int s = (int)floor(z) + 2;
    if (s < 0 || s > 3) return 0.0;
    const double *c = T.cval[t][s];
    return c[0] * z3 + c[1] * z2 + c[2] * z + c[3];
}
__device__ __forceinline__ double cubic_der(const SplineTables &T, int t, double z) {
    double z2 = z * z;
    int s = (int)floor(z) + 2;
    if (s < 0 || s > 3) return 0.0;
    const double *c = T.cder[t][s];
    return c[0] * z2 + c[1] * z + c[2];
}
__device__ __forceinline__ double quad_val(int t, double z) { return quad_val(kTab, t, z); }
__device__ __forceinline__ double quad_der(int t, double z) { return quad_der(kTab, t, z); }
__device__ __forceinline__ double cubic_val(int t, double z) { return cubic_val(kTab, t, z); }
__device__ __forceinline__ double cubic_der(int t, double z) { return cubic_der(kTab, t, z); }

// Node shape-function type, arithmetic form of basisval*
// (interpolants.H:1419-1427) and if-chain form of *_interp_two (:946-956).
__device__ __forceinline__ int type_arith(int idx, int lo, int hi) {
    return 3 + (idx == lo || idx == hi + 1) * (1 - 3) + (idx == lo + 1) * (2 - 3) +
           (idx == hi) * (4 - 3);
}
__device__ __forceinline__ int type_chain(int idx, int lo, int hi) {
    if (idx == lo || idx == hi + 1) return 1;
    if (idx == lo + 1) return 2;
    if (idx == hi) return 4;
    return 3;
}

// compute_bounds (mpm_particle_grid_ops.cpp:255-299)
__device__ __forceinline__ void bounds(int ivd, int lod, int hid, int scheme, int per,
                                       int &bmin, int &bmax) {
    if (scheme == 1) {
        bmin = 0;
        bmax = 2;
    } else if (per) {
        bmin = -1;
        bmax = 3;
    } else if (ivd == lod) {
        bmin = 0;
        bmax = 3;
    } else if (ivd == hid) {
        bmin = -1;
        bmax = 2;
    } else {
        bmin = -1;
        bmax = 3;
    }
}

// One direction of basisval_and_grad (interpolants.H:1592-1632): value and
// derivative of the 1-D factor for stencil offset l of cell i.
__device__ __forceinline__ void basis1d(int od, int per, int i, int l, int lo, int hi,
                                        double x, double plo, double dx, double &val,
                                        double &der) {
    const double dxinv = 1.0 / dx;
    if (od == 1) {
        double r = (x - (plo + i * dx)) / dx;
        val = (l == 0) ? (1.0 - r) : r;
        der = (l == 0) ? -dxinv : dxinv;
    } else {
        int t = 3;
        if (per == 0) t = type_arith(i + l, lo, hi);
        double r = (x - (plo + (i + l) * dx)) / dx;
        if (od == 2) {
            val = quad_val(t, r);
            der = quad_der(t, r) * dxinv;
        } else {
            val = cubic_val(t, r);
            der = cubic_der(t, r) * dxinv;
        }
    }
}

// One direction of quadratic_interp_two / cubic_interp_two
// (interpolants.H:943-1005, 1295-1365): quadratic divides by dx, cubic
// multiplies by inv_dx = 1/dx.
__device__ __forceinline__ double interp1d(int order, int per, int idx, int lo, int hi,
                                           double x, double plo, double dx) {
    int t = per ? 3 : type_chain(idx, lo, hi);
    if (order == 2) {
        double r = (x - (plo + idx * dx)) / dx;
        return quad_val(t, r);
    }
    const double inv = 1.0 / dx;
    double r = (x - (plo + idx * dx)) * inv;
    return cubic_val(t, r);
}

// applybc (mpm_kernels.H:27-113); returns -1 for an unknown code.
__device__ __forceinline__ int applybc(const double in[3], double out[3], double mu,
                                       const double n[3], int bc) {
    int modify = 0;
    double vt[3], vth[3];
    double vn = 0.0;
    for (int d = 0; d < 3; ++d) vn += in[d] * n[d];
    for (int d = 0; d < 3; ++d) vt[d] = in[d] - vn * n[d];
    double vtm = 0.0;
    for (int d = 0; d < 3; ++d) vtm += vt[d] * vt[d];
    vtm = sqrt(vtm);
    for (int d = 0; d < 3; ++d) vth[d] = vt[d] / (vtm + kTiny);
    if (bc == MPM_BC_NOSLIP) {
        modify = 1;
        for (int d = 0; d < 3; ++d) out[d] = 0.0;
    } else if (bc == MPM_BC_SLIP) {
        modify = 1;
        for (int d = 0; d < 3; ++d) out[d] = vt[d];
    } else if (bc == MPM_BC_PARTIALSLIP) {
        if (vn <= 0.0) {
            modify = 1;
            if (vtm <= -mu * vn) {
                for (int d = 0; d < 3; ++d) out[d] = 0.0;
            } else {
                for (int d = 0; d < 3; ++d) out[d] = (vtm + mu * vn) * vth[d];
            }
        }
    } else if (bc != MPM_BC_PERIODIC) {
        return -1;
    }
    return modify;
}

// det F by cofactor expansion (mpm_particle_timestep.cpp:169-171)
__device__ __forceinline__ double det3(const double F[9]) {
    return F[0] * (F[4] * F[8] - F[5] * F[7]) - F[1] * (F[3] * F[8] - F[5] * F[6]) +
           F[2] * (F[3] * F[7] - F[4] * F[6]);
}

// linear_elastic (constitutive_models.H:21-54), 3-D branch; also used for
// the delta form (:76-109), which is the same linear map.
__device__ __forceinline__ void linear_elastic(const double e[6], double s[6], double E,
                                               double v) {
    double c1 = E / (1.0 + v) / (1.0 - 2.0 * v);
    double c2 = E / (1.0 + v);
    s[XX] = c1 * ((1 - v) * e[XX] + v * (e[YY] + e[ZZ]));
    s[YY] = c1 * ((1 - v) * e[YY] + v * (e[XX] + e[ZZ]));
    s[ZZ] = c1 * ((1 - v) * e[ZZ] + v * (e[XX] + e[YY]));
    s[XY] = c2 * e[XY];
    s[XZ] = c2 * e[XZ];
    s[YZ] = c2 * e[YZ];
}

// Newtonian_Fluid (constitutive_models.H:117-157), 3-D branch.
__device__ __forceinline__ void newtonian(const double D[6], double s[6], double mu,
                                          double p) {
    double tr = D[XX] + D[YY] + D[ZZ];
    double o3 = 1.0 / 3.0;
    s[XX] = 2.0 * mu * (D[XX] - o3 * tr) - p;
    s[YY] = 2.0 * mu * (D[YY] - o3 * tr) - p;
    s[ZZ] = 2.0 * mu * (D[ZZ] - o3 * tr) - p;
    s[XY] = 2.0 * mu * D[XY];
    s[XZ] = 2.0 * mu * D[XZ];
    s[YZ] = 2.0 * mu * D[YZ];
}

// pow and log of the constitutive updates, called out of line: inlined,
// their expansions set the register count (and so the occupancy) of the
// kernels that contain them; once per particle the call costs little.
__device__ __attribute__((noinline)) double ool_pow(double x, double y) { return pow(x, y); }
__device__ __attribute__((noinline)) double ool_log(double x) { return log(x); }

// neo_hookean (constitutive_models.H:200-250), 3-D branch.
__device__ __forceinline__ void neo_hookean(const double F[9], double s[6], double E,
                                            double v) {
    double mu = E / (2.0 * (1.0 + v));
    double lam = E * v / ((1.0 + v) * (1.0 - 2.0 * v));
    double B[6];
    auto rd = [&](int a, int b) {
        double acc = 0.0;
        for (int k = 0; k < 3; ++k) acc += F[a * 3 + k] * F[b * 3 + k];
        return acc;
    };
    B[XX] = rd(0, 0);
    B[YY] = rd(1, 1);
    B[XY] = rd(0, 1);
    B[ZZ] = rd(2, 2);
    B[XZ] = rd(0, 2);
    B[YZ] = rd(1, 2);
    double det = det3(F);
    det = (det < 1.e-12) ? 1.e-12 : det;  // std::max(det, 1e-12)
    const double vol = lam * ool_log(det) / det;
    const double muJ = mu / det;
    s[XX] = muJ * (B[XX] - 1.0) + vol;
    s[YY] = muJ * (B[YY] - 1.0) + vol;
    s[ZZ] = muJ * (B[ZZ] - 1.0) + vol;
    s[XY] = muJ * B[XY];
    s[XZ] = muJ * B[XZ];
    s[YZ] = muJ * B[YZ];
}

}  // namespace xmpm

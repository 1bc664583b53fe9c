// mpm_geometry.cpp — level-set geometry (host code): the signed distance of
// the analytic bodies the reference builds with AMReX EB2 implicit functions
// (sphere, plane, cylinder, box; mpm_eb.cpp:212-388), sampled on the valid
// nodal box in the layout mpm_add_levelset_body takes.  EB2 and its signed-
// distance fill are AMReX code absent from the reference tree; these are the
// exact Euclidean distances of the same shapes, negative inside the solid.

#include <cmath>

#include "mpm_context.h"

using namespace xmpm;

namespace {

double norm3(double a, double b, double c) { return std::sqrt(a * a + b * b + c * c); }

// Signed distance to the surface, negative inside the shape.
double sdf(int type, const double *p, const double x[3], double nrm_len) {
    switch (type) {
    case MPM_LS_SPHERE:
        return norm3(x[0] - p[0], x[1] - p[1], x[2] - p[2]) - p[3];
    case MPM_LS_PLANE:
        // the solid lies behind the plane (against its normal)
        return ((x[0] - p[0]) * p[3] + (x[1] - p[1]) * p[4] + (x[2] - p[2]) * p[5]) / nrm_len;
    case MPM_LS_CYLINDER: {
        const int dir = (int)p[5];
        double r2 = 0.0;
        for (int d = 0; d < 3; ++d)
            if (d != dir) r2 += (x[d] - p[d]) * (x[d] - p[d]);
        const double dr = std::sqrt(r2) - p[3];
        if (!(p[4] > 0.0)) return dr;  // unbounded along its axis
        const double dh = std::fabs(x[dir] - p[dir]) - 0.5 * p[4];
        const double a = std::fmax(dr, 0.0), b = std::fmax(dh, 0.0);
        return std::sqrt(a * a + b * b) + std::fmin(std::fmax(dr, dh), 0.0);
    }
    default: {  // MPM_LS_BOX
        double q[3], out2 = 0.0, inmax = -INFINITY;
        for (int d = 0; d < 3; ++d) {
            const double c = 0.5 * (p[d] + p[3 + d]), h = 0.5 * (p[3 + d] - p[d]);
            q[d] = std::fabs(x[d] - c) - h;
            out2 += std::fmax(q[d], 0.0) * std::fmax(q[d], 0.0);
            inmax = std::fmax(inmax, q[d]);
        }
        return std::sqrt(out2) + std::fmin(inmax, 0.0);
    }
    }
}

}  // namespace

extern "C" {

int mpm_levelset_analytic(const mpm_specs *s, int type, const double *params, int fluid_inside,
                          double *phi) {
    if (!s || !params || !phi) return set_err(MPM_E_ARG, "null argument");
    if (type < MPM_LS_SPHERE || type > MPM_LS_BOX) return set_err(MPM_E_ARG, "unknown body type");
    double nrm_len = 1.0;
    if (type == MPM_LS_SPHERE && !(params[3] > 0.0))
        return set_err(MPM_E_ARG, "sphere radius must be positive");
    if (type == MPM_LS_PLANE) {
        nrm_len = norm3(params[3], params[4], params[5]);
        if (!(nrm_len > 0.0)) return set_err(MPM_E_ARG, "plane normal must be non-zero");
    }
    if (type == MPM_LS_CYLINDER) {
        const int dir = (int)params[5];
        if (!(params[3] > 0.0) || dir < 0 || dir > 2 || params[5] != (double)dir)
            return set_err(MPM_E_ARG, "cylinder needs radius > 0 and direction 0, 1 or 2");
    }
    if (type == MPM_LS_BOX)
        for (int d = 0; d < 3; ++d)
            if (!(params[3 + d] > params[d])) return set_err(MPM_E_ARG, "box needs lo < hi");
    int n[3];
    double dx[3];
    for (int d = 0; d < 3; ++d) {
        n[d] = s->ncells[d];
        if (n[d] < 1 || !(s->phi[d] > s->plo[d])) return set_err(MPM_E_ARG, "bad domain");
        dx[d] = (s->phi[d] - s->plo[d]) / n[d];
    }
    int64_t q = 0;
    for (int k = 0; k <= n[2]; ++k)
        for (int j = 0; j <= n[1]; ++j)
            for (int i = 0; i <= n[0]; ++i, ++q) {
                const double x[3] = {s->plo[0] + i * dx[0], s->plo[1] + j * dx[1],
                                     s->plo[2] + k * dx[2]};
                const double v = sdf(type, params, x, nrm_len);
                phi[q] = fluid_inside ? -v : v;
            }
    return MPM_OK;
}

}  // extern "C"

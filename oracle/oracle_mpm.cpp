// oracle_mpm.cpp — CPU restatement of ExaGOOP's explicit MPM step.
//
// TEST INFRASTRUCTURE ONLY: the parity checker for the HIP library and the
// timed CPU baseline.  Never linked into the product.
//
// Each function cites the reference file:line it restates (paths relative to
// the ExaGOOP tree).  The restatement keeps the reference's arithmetic order
// (division vs. reciprocal, product order, accumulation order) and its data
// model: AoS particle records, per-node re-evaluation of the basis, periodic
// ghost-particle images for P2G, ghost-node fill for G2P.
//
// Parity status of the AMReX-provided pieces (AMReX is absent from the
// reference tree, so these are restated from AMReX's documented behaviour
// and are "parity unpinned"): getParticleCell, Geometry dx/dxi,
// Redistribute's periodic wrap and out-of-domain removal, FillBoundary's
// ghost-node fill, NeighborParticleContainer image selection.

#include "oracle_mpm.h"

#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// constants.H:10-24 — truncated literals kept as the reference writes them.
constexpr double kFourByThree = 1.3333333333;
constexpr double kOneBySix = 0.16666666666667;
constexpr double kOneByThree = 0.333333333333;
constexpr double kTwoByThree = 0.6666666666667;
constexpr double kTiny = 1e-20;

// Symmetric-tensor packing in 3-D (constants.H:80-86).
enum { XX = 0, XY = 1, XZ = 2, YY = 3, YZ = 4, ZZ = 5 };

// Boundary-modified spline tables (interpolants.H:32-170).  Rows are the
// shape-function type (1 boundary node, 2 lo+1, 3 interior/periodic, 4 hi),
// columns the polynomial segment.
const double kCubicVal[5][4][4] = {
    {{0, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0}},
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
     {0.0, 0.0, 0.0, 0.0}}};

const double kCubicDer[5][4][3] = {
    {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}},
    {{0.5, 2.0, 2.0}, {-0.5, 0.0, 1.0}, {0.5, 0.0, -1.0}, {-0.5, 2.0, -2.0}},
    {{0, 0, 0}, {-1.0, -2.0, 0.0}, {1.5, -2.0, 0.0}, {-0.5, 2.0, -2.0}},
    {{0.5, 2.0, 2.0}, {-1.5, -2.0, 0.0}, {1.5, -2.0, 0.0}, {-0.5, 2.0, -2.0}},
    {{0.5, 2.0, 2.0}, {-1.5, -2.0, 0.0}, {1.0, -2.0, 0.0}, {0, 0, 0}}};

const double kQuadVal[5][4][3] = {
    {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}},
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
     {0, 0, 0}}};

const double kQuadDer[5][4][2] = {
    {{0, 0}, {0, 0}, {0, 0}, {0, 0}},
    {{4.0 / 3.0, 2.0}, {-8.0 / 3.0, 0.0}, {-8.0 / 3.0, 0.0}, {4.0 / 3.0, -2.0}},
    {{8.0 / 3.0, 8.0 / 3.0}, {-14.0 / 6.0, 1.0 / 6.0}, {1.0, -1.5}, {0, 0}},
    {{2.0 / 2.0, 3.0 / 2.0}, {-2.0, 0.0}, {2.0 / 2.0, -3.0 / 2.0}, {0, 0}},
    {{1.0, 1.5}, {-14.0 / 6.0, -1.0 / 6.0}, {8.0 / 3.0, -8.0 / 3.0}, {0, 0}}};

const int kQuadSegMap[5][6] = {{-1, -1, -1, -1, -1, -1},
                               {0, 0, 1, 2, 3, 3},
                               {0, 1, 1, 2, 2, -1},
                               {0, 0, 1, 1, 2, 2},
                               {0, 0, 1, 1, 2, -1}};

// interval_idx_quadratic (interpolants.H:187-199).
inline int quad_segment(int type, double z) {
    double shift = (type == 2) ? 1.0 : 1.5;
    int raw = static_cast<int>(std::floor((z + shift) * 2));
    if (raw < 0 || raw >= 6) return -1;
    return kQuadSegMap[type][raw];
}

// quadraticspline_1d (interpolants.H:266-278).
inline double quad_val(int type, double z) {
    double z2 = z * z;
    int s = quad_segment(type, z);
    if (s < 0) return 0.0;
    const double *c = kQuadVal[type][s];
    return c[0] * z2 + c[1] * z + c[2];
}

// quadraticspline_1d_der (interpolants.H:293-304).
inline double quad_der(int type, double z) {
    int s = quad_segment(type, z);
    if (s < 0) return 0.0;
    const double *c = kQuadDer[type][s];
    return c[0] * z + c[1];
}

// cubicspline_1d (interpolants.H:215-226).  The reference indexes the table
// with floor(z)+2 unchecked; an index outside [0,3] (|z| >= 2, reachable
// only through rounding at the support edge, where the spline is ~0) is
// restated here as 0.
inline double cubic_val(int type, double z) {
    double z3 = z * z * z;
    double z2 = z * z;
    int s = static_cast<int>(std::floor(z)) + 2;
    if (s < 0 || s > 3) return 0.0;
    const double *c = kCubicVal[type][s];
    return c[0] * z3 + c[1] * z2 + c[2] * z + c[3];
}

// cubicspline_1d_der (interpolants.H:240-250).
inline double cubic_der(int type, double z) {
    double z2 = z * z;
    int s = static_cast<int>(std::floor(z)) + 2;
    if (s < 0 || s > 3) return 0.0;
    const double *c = kCubicDer[type][s];
    return c[0] * z2 + c[1] * z + c[2];
}

// Shape-function type by node index (interpolants.H:1419-1427): the
// arithmetic form used by basisval*.
inline int node_type_arith(int idx, int lo, int hi) {
    return 3 + (idx == lo || idx == hi + 1) * (1 - 3) + (idx == lo + 1) * (2 - 3) +
           (idx == hi) * (4 - 3);
}

// The if-chain form used by quadratic_interp_two / cubic_interp_two
// (interpolants.H:946-956, 1298-1308).
inline int node_type_chain(int idx, int lo, int hi) {
    if (idx == lo || idx == hi + 1) return 1;
    if (idx == lo + 1) return 2;
    if (idx == hi) return 4;
    return 3;
}

// basisval_and_grad (interpolants.H:1581-1654).
double basisval_and_grad(const int st[3], const int cell[3], const double xp[3],
                         const double plo[3], const double dx[3], const int od[3],
                         const int per[3], const int lo[3], const int hi[3],
                         double grad[3]) {
    double v1[3], d1[3];
    for (int d = 0; d < 3; ++d) {
        const int i = cell[d];
        const int l = st[d];
        const double dxinv = 1.0 / dx[d];
        if (od[d] == 1) {
            double r = (xp[d] - (plo[d] + i * dx[d])) / dx[d];
            v1[d] = (l == 0) ? (1.0 - r) : r;
            d1[d] = (l == 0) ? -dxinv : dxinv;
        } else if (od[d] == 2) {
            int t = 3;
            int idx = i + l;
            if (per[d] == 0) t = node_type_arith(idx, lo[d], hi[d]);
            double r = (xp[d] - (plo[d] + (i + l) * dx[d])) / dx[d];
            v1[d] = quad_val(t, r);
            d1[d] = quad_der(t, r) * dxinv;
        } else {
            int t = 3;
            int idx = i + l;
            if (per[d] == 0) t = node_type_arith(idx, lo[d], hi[d]);
            double r = (xp[d] - (plo[d] + (i + l) * dx[d])) / dx[d];
            v1[d] = cubic_val(t, r);
            d1[d] = cubic_der(t, r) * dxinv;
        }
    }
    double N = 1.0;
    for (int d = 0; d < 3; ++d) N *= v1[d];
    for (int dir = 0; dir < 3; ++dir) {
        double g = 1.0;
        for (int d = 0; d < 3; ++d) g *= (d == dir) ? d1[d] : v1[d];
        grad[dir] = g;
    }
    return N;
}

// basisval (interpolants.H:1392-1456).
double basisval(const int st[3], const int cell[3], const double xp[3],
                const double plo[3], const double dx[3], const int od[3],
                const int per[3], const int lo[3], const int hi[3]) {
    double c[3];
    for (int d = 0; d < 3; ++d) {
        const int i = cell[d];
        const int l = st[d];
        double r = (xp[d] - (plo[d] + i * dx[d])) / dx[d];
        if (od[d] == 1) {
            c[d] = (l == 0) ? (1.0 - r) : r;
        } else {
            int t = 3;
            int idx = i + l;
            if (per[d] == 0) t = node_type_arith(idx, lo[d], hi[d]);
            r = (xp[d] - (plo[d] + (i + l) * dx[d])) / dx[d];
            c[d] = (od[d] == 2) ? quad_val(t, r) : cubic_val(t, r);
        }
    }
    double v = 1.0;
    for (int d = 0; d < 3; ++d) v *= c[d];
    return v;
}

// basisvalder (interpolants.H:1478-1553), linear branch only is reached
// (level-set gradient, mpm_eb.H:243-288).
double basisvalder_linear(int dir, const int st[3], const int cell[3],
                          const double xp[3], const double plo[3],
                          const double dx[3]) {
    double prod = 1.0;
    for (int d = 0; d < 3; ++d) {
        const int i = cell[d];
        const int l = st[d];
        const double dxinv = 1.0 / dx[d];
        double r = (xp[d] - (plo[d] + i * dx[d])) / dx[d];
        double val = (l == 0) ? (1.0 - r) : r;
        double der = (l == 0) ? -dxinv : dxinv;
        prod *= (d == dir) ? der : val;
    }
    return prod;
}

// applybc (mpm_kernels.H:27-113).  Unknown codes return -1 (the reference
// aborts).
int applybc(const double in[3], double out[3], double mu, const double n[3],
            int bc) {
    int modify = 0;
    double vt[3], vth[3];
    double vn = 0.0;
    for (int d = 0; d < 3; ++d) vn += in[d] * n[d];
    for (int d = 0; d < 3; ++d) vt[d] = in[d] - vn * n[d];
    double vtm = 0.0;
    for (int d = 0; d < 3; ++d) vtm += vt[d] * vt[d];
    vtm = std::sqrt(vtm);
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
    } else if (bc == MPM_BC_PERIODIC) {
    } else {
        return -1;
    }
    return modify;
}

// compute_bounds (mpm_particle_grid_ops.cpp:255-299).
inline void compute_bounds(int ivd, int lod, int hid, int scheme, int per,
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

struct Part {
    double pos[3];
    double r[MPM_NREAL];
    int idata[MPM_NINT];
    int64_t id;  // upload index; < 0 means removed
    double th[MPM_NTHERMAL];  // realData 46..52 of the USE_TEMP build
};

struct LsBody {
    std::vector<double> phi;  // nodal with 1 ghost layer
    int bc;
    double mu;
    double w[3];
    // temperature condition (LevelSetBody, mpm_eb.H:78-118)
    int temp_bc = MPM_LS_TEMP_ADIABATIC;
    double T_wall = 0.0, heat_flux = 0.0, h_conv = 0.0, T_inf = 0.0;
};

inline void atomic_add(double *p, double v, bool par) {
    if (par) {
#pragma omp atomic
        *p += v;
    } else {
        *p += v;
    }
}

}  // namespace

struct orc_ctx {
    mpm_specs s;
    int n[3];
    int lo[3] = {0, 0, 0};
    int hi[3];
    double dx[3], dxi[3];
    int od[3];
    int ng_cells, ng_nodal;
    int N[3];  // nodal extents incl. ghosts
    std::vector<double> nd;
    std::vector<Part> parts;
    std::vector<Part> ghosts;
    std::vector<LsBody> bodies;
    mpm_wall_udf_fn udf[3][2] = {{nullptr, nullptr}, {nullptr, nullptr}, {nullptr, nullptr}};
    std::vector<double> udf_tab[3][2];
    double time = 0.0;
    double last_dt = 0.0;
    int64_t steps = 0;
    int nthreads = 1;
    bool thermal = false;
    mpm_thermal_specs ts;
    mpm_wall_temp_udf_fn tudf[3][2] = {{nullptr, nullptr}, {nullptr, nullptr}, {nullptr, nullptr}};
    std::vector<double> tudf_tab[3][2];

    inline double &A(int i, int j, int k, int c) {
        size_t idx = (size_t)c * N[0] * N[1] * N[2] +
                     ((size_t)(k + ng_nodal) * N[1] + (j + ng_nodal)) * N[0] +
                     (i + ng_nodal);
        return nd[idx];
    }
    bool par() const { return nthreads > 1; }

    // getParticleCell (AMReX; called at mpm_particle_grid_ops.cpp:430,983):
    // floor((x - plo) * dxi) + domain.lo.
    void cell(const double x[3], int iv[3]) const {
        for (int d = 0; d < 3; ++d)
            iv[d] = (int)std::floor((x[d] - s.plo[d]) * dxi[d]) + lo[d];
    }
};

namespace {

// Level-set interpolation (mpm_eb.H:176-228, 243-288), lsref = 1.
double ls_value(orc_ctx *c, const LsBody &b, const double x[3]) {
    int idx[3];
    for (int d = 0; d < 3; ++d)
        idx[d] = (int)std::floor((x[d] - c->s.plo[d] + kTiny) / c->dx[d]);
    double v = 0.0;
    for (int corner = 0; corner < 8; ++corner) {
        int off[3] = {corner & 1, (corner >> 1) & 1, (corner >> 2) & 1};
        int ii[3];
        for (int d = 0; d < 3; ++d) {
            ii[d] = idx[d] + off[d];
            if (ii[d] < -1) ii[d] = -1;
            if (ii[d] > c->n[d] + 1) ii[d] = c->n[d] + 1;
        }
        size_t p = ((size_t)(ii[2] + 1) * (c->n[1] + 3) + (ii[1] + 1)) * (c->n[0] + 3) +
                   (ii[0] + 1);
        const int od1[3] = {1, 1, 1};
        const int per1[3] = {1, 1, 1};
        const int z3[3] = {0, 0, 0};
        v += b.phi[p] * basisval(off, idx, x, c->s.plo, c->dx, od1, per1, z3, z3);
    }
    return v;
}

void ls_grad(orc_ctx *c, const LsBody &b, const double x[3], double g[3]) {
    int idx[3];
    for (int d = 0; d < 3; ++d)
        idx[d] = (int)std::floor((x[d] - c->s.plo[d] + kTiny) / c->dx[d]);
    for (int d = 0; d < 3; ++d) g[d] = 0.0;
    for (int corner = 0; corner < 8; ++corner) {
        int off[3] = {corner & 1, (corner >> 1) & 1, (corner >> 2) & 1};
        int ii[3];
        for (int d = 0; d < 3; ++d) {
            ii[d] = idx[d] + off[d];
            if (ii[d] < -1) ii[d] = -1;
            if (ii[d] > c->n[d] + 1) ii[d] = c->n[d] + 1;
        }
        size_t p = ((size_t)(ii[2] + 1) * (c->n[1] + 3) + (ii[1] + 1)) * (c->n[0] + 3) +
                   (ii[0] + 1);
        for (int d = 0; d < 3; ++d)
            g[d] += b.phi[p] * basisvalder_linear(d, off, idx, x, c->s.plo, c->dx);
    }
}

// RedistributeLocal's effect on one particle (AMReX enforcePeriodic +
// out-of-domain removal).  Periodic: shift by the period while outside
// [plo, rhi], rhi the largest double whose cell is n-1.  Non-periodic: a
// particle whose cell lies outside the domain is removed.
void redistribute_one(orc_ctx *c, Part &p) {
    for (int d = 0; d < 3; ++d) {
        if (!c->s.periodic[d]) continue;
        double plo = c->s.plo[d], phi = c->s.phi[d];
        double rhi = phi;
        while ((int)std::floor((rhi - plo) * c->dxi[d]) >= c->n[d])
            rhi = std::nextafter(rhi, -INFINITY);
        double L = phi - plo;
        // bounded shifts (a non-finite position would never settle)
        if (p.pos[d] > rhi) {
            for (int k = 0; k < 64 && p.pos[d] > rhi; ++k) p.pos[d] -= L;
            if (p.pos[d] < plo) p.pos[d] = plo;
        } else if (p.pos[d] < plo) {
            for (int k = 0; k < 64 && p.pos[d] < plo; ++k) p.pos[d] += L;
            if (p.pos[d] > rhi) p.pos[d] = rhi;
        }
    }
    int iv[3];
    c->cell(p.pos, iv);
    for (int d = 0; d < 3; ++d) {
        if (c->s.periodic[d]) continue;
        if (iv[d] < c->lo[d] || iv[d] > c->hi[d]) {
            p.id = -1;
            return;
        }
    }
}

void redistribute(orc_ctx *c) {
    for (auto &p : c->parts)
        if (p.id >= 0) redistribute_one(c, p);
}

// fillNeighbors for one periodic box: images of particles within ng_cells of
// a periodic face, including edge/corner images.
void fill_neighbors(orc_ctx *c) {
    c->ghosts.clear();
    for (const auto &p : c->parts) {
        if (p.id < 0) continue;
        int iv[3];
        c->cell(p.pos, iv);
        for (int sz = -1; sz <= 1; ++sz)
            for (int sy = -1; sy <= 1; ++sy)
                for (int sx = -1; sx <= 1; ++sx) {
                    int sh[3] = {sx, sy, sz};
                    if (!sx && !sy && !sz) continue;
                    bool ok = true;
                    for (int d = 0; d < 3 && ok; ++d) {
                        if (!sh[d]) continue;
                        if (!c->s.periodic[d]) ok = false;
                        else if (sh[d] > 0) ok = (iv[d] - c->lo[d] < c->ng_cells);
                        else ok = (c->hi[d] - iv[d] < c->ng_cells);
                    }
                    if (!ok) continue;
                    Part g = p;
                    for (int d = 0; d < 3; ++d)
                        g.pos[d] = p.pos[d] + sh[d] * (c->s.phi[d] - c->s.plo[d]);
                    c->ghosts.push_back(g);
                }
    }
}

// FillBoundary(periodicity) on the nodal data: ghost nodes in periodic
// directions take the value of the valid node one period away; ghost nodes
// with an out-of-range index in a non-periodic direction are left as they
// are.
void fill_boundary(orc_ctx *c) {
    const int g = c->ng_nodal;
    for (int k = -g; k <= c->n[2] + g; ++k)
        for (int j = -g; j <= c->n[1] + g; ++j)
            for (int i = -g; i <= c->n[0] + g; ++i) {
                int I[3] = {i, j, k};
                bool ghost = false, ok = true;
                int S[3];
                for (int d = 0; d < 3; ++d) {
                    S[d] = I[d];
                    if (I[d] < 0 || I[d] > c->n[d]) {
                        ghost = true;
                        if (!c->s.periodic[d]) ok = false;
                        else S[d] = (I[d] < 0) ? I[d] + c->n[d] : I[d] - c->n[d];
                    }
                }
                if (!ghost || !ok) continue;
                for (int comp = 0; comp < MPM_N_NCOMP_TEMP; ++comp)
                    c->A(i, j, k, comp) = c->A(S[0], S[1], S[2], comp);
            }
}

void eval_udf_tables(orc_ctx *c, double t) {
    // compute_udf_wall_vel_at_nodes (nodal_data_ops.cpp:929-1059), 3-D.
    for (int d = 0; d < 3; ++d)
        for (int h = 0; h < 2; ++h) {
            if (!c->udf[d][h]) continue;
            int p0 = (d == 0) ? 1 : 0;
            int p1 = (d == 2) ? 1 : 2;
            auto &tab = c->udf_tab[d][h];
            tab.assign((size_t)(c->n[p0] + 1) * (c->n[p1] + 1) * 3, 0.0);
            for (int j = 0; j <= c->n[p0]; ++j)
                for (int k = 0; k <= c->n[p1]; ++k) {
                    double co[3] = {0, 0, 0};
                    co[d] = h ? c->s.phi[d] : c->s.plo[d];
                    co[p0] = c->s.plo[p0] + j * c->dx[p0];
                    co[p1] = c->s.plo[p1] + k * c->dx[p1];
                    double vel[3] = {0, 0, 0};
                    c->udf[d][h](co[0], co[1], co[2], t, vel);
                    size_t f = ((size_t)j * (c->n[p1] + 1) + k) * 3;
                    for (int q = 0; q < 3; ++q) tab[f + q] = vel[q];
                }
        }
}

// Initialise_Domain's per-direction order (mpm_init.cpp:180-233).
int order_directional(const mpm_specs *s, int od[3]) {
    for (int d = 0; d < 3; ++d) {
        if (s->order_scheme == 1) od[d] = 1;
        else if (s->order_scheme == 2 || s->order_scheme == 3) {
            const int nc = s->ncells[d];
            od[d] = (s->periodic[d] == 0) ? ((nc < 5) ? 1 : s->order_scheme)
                                          : ((nc < 3) ? 1 : s->order_scheme);
        } else
            return MPM_E_ORDER;
    }
    return MPM_OK;
}

}  // namespace

extern "C" {

int orc_create(const mpm_specs *specs, orc_ctx **out) {
    if (!specs || !out) return MPM_E_ARG;
    int od[3];
    int rc = order_directional(specs, od);
    if (rc) return rc;
    for (int d = 0; d < 3; ++d) {
        if (specs->ncells[d] < 1 || !(specs->phi[d] > specs->plo[d])) return MPM_E_ARG;
        int b[2] = {specs->bclo[d], specs->bchi[d]};
        for (int q = 0; q < 2; ++q)
            if (b[q] < 0 || b[q] > MPM_BC_OUTFLOW) return MPM_E_BC;
    }
    orc_ctx *c = new orc_ctx();
    c->s = *specs;
    for (int d = 0; d < 3; ++d) {
        c->n[d] = specs->ncells[d];
        c->hi[d] = specs->ncells[d] - 1;
        // Geometry cell size: prob length / ncells, inverse as 1/dx.
        c->dx[d] = (specs->phi[d] - specs->plo[d]) / (double)specs->ncells[d];
        c->dxi[d] = 1.0 / c->dx[d];
        c->od[d] = od[d];
    }
    // mpm_init.cpp:164-172, 175-224
    c->ng_cells = 1;
    if (specs->order_scheme == 3) c->ng_cells = 2;
    if (specs->order_scheme == 2) c->ng_cells = 3;
    c->ng_nodal = (specs->order_scheme == 1) ? 1 : 3;
    for (int d = 0; d < 3; ++d) c->N[d] = c->n[d] + 1 + 2 * c->ng_nodal;
    // NUM_STATES of the USE_TEMP build (constants.H:31): the thermal
    // components follow the 20 mechanical ones.
    c->nd.assign((size_t)MPM_N_NCOMP_TEMP * c->N[0] * c->N[1] * c->N[2], 0.0);
    *out = c;
    return MPM_OK;
}

void orc_destroy(orc_ctx *c) { delete c; }

void orc_set_threads(orc_ctx *c, int n) { c->nthreads = n < 1 ? 1 : n; }

int orc_upload_particles(orc_ctx *c, int64_t np, const double *pos,
                         const double *rdata, const int *idata) {
    if (!c || np < 0 || (np > 0 && (!pos || !rdata || !idata))) return MPM_E_ARG;
    c->parts.resize(np);
    for (int64_t i = 0; i < np; ++i) {
        Part &p = c->parts[i];
        for (int d = 0; d < 3; ++d) p.pos[d] = pos[d * np + i];
        for (int f = 0; f < MPM_NREAL; ++f) p.r[f] = rdata[f * np + i];
        for (int f = 0; f < MPM_NINT; ++f) p.idata[f] = idata[f * np + i];
        for (int f = 0; f < MPM_NTHERMAL; ++f) p.th[f] = 0.0;
        p.id = i;
        int m = p.idata[MPM_I_MODEL];
        if (m < 0 || m > 2) return MPM_E_MODEL;
    }
    c->ghosts.clear();
    return MPM_OK;
}

int orc_num_particles(orc_ctx *c, int64_t *np) {
    int64_t n = 0;
    for (auto &p : c->parts) n += (p.id >= 0);
    *np = n;
    return MPM_OK;
}

int orc_download_particles(orc_ctx *c, double *pos, double *rdata, int *idata,
                           int64_t *id) {
    int64_t np = 0;
    orc_num_particles(c, &np);
    int64_t q = 0;
    for (auto &p : c->parts) {
        if (p.id < 0) continue;
        if (pos)
            for (int d = 0; d < 3; ++d) pos[d * np + q] = p.pos[d];
        if (rdata)
            for (int f = 0; f < MPM_NREAL; ++f) rdata[f * np + q] = p.r[f];
        if (idata)
            for (int f = 0; f < MPM_NINT; ++f) idata[f * np + q] = p.idata[f];
        if (id) id[q] = p.id;
        ++q;
    }
    return MPM_OK;
}

int orc_get_nodal(orc_ctx *c, int comp, double *out) {
    if (comp < 0 || comp >= (c->thermal ? (int)MPM_N_NCOMP_TEMP : (int)MPM_N_NCOMP) || !out)
        return MPM_E_ARG;
    size_t q = 0;
    for (int k = 0; k <= c->n[2]; ++k)
        for (int j = 0; j <= c->n[1]; ++j)
            for (int i = 0; i <= c->n[0]; ++i) out[q++] = c->A(i, j, k, comp);
    return MPM_OK;
}

// Reset_Nodaldata_to_Zero (nodal_data_ops.cpp:924-927): all comps, ghosts too.
int orc_reset_nodal(orc_ctx *c) {
    std::fill(c->nd.begin(), c->nd.end(), 0.0);
    return MPM_OK;
}

int orc_fill_neighbors(orc_ctx *c) {
    fill_neighbors(c);
    return MPM_OK;
}

// deposit_onto_grid_momentum (mpm_particle_grid_ops.cpp:333-628).
int orc_p2g(orc_ctx *c, int update_mass, int update_vel, int update_forces) {
    const mpm_specs &s = c->s;
    // zero (:368-403)
    for (int k = 0; k <= c->n[2]; ++k)
        for (int j = 0; j <= c->n[1]; ++j)
            for (int i = 0; i <= c->n[0]; ++i) {
                if (update_mass) c->A(i, j, k, MPM_N_MASS) = 0.0;
                if (update_vel)
                    for (int d = 0; d < 3; ++d) c->A(i, j, k, MPM_N_VEL + d) = 0.0;
                if (update_forces)
                    for (int d = 0; d < 3; ++d) c->A(i, j, k, MPM_N_FRC + d) = 0.0;
                if (update_forces == 2) c->A(i, j, k, 15) = 0.0;
            }
    const int64_t nreal = (int64_t)c->parts.size();
    const int64_t nt = nreal + (int64_t)c->ghosts.size();
    const bool par = c->par();
#pragma omp parallel for schedule(static) num_threads(c->nthreads) if (par)
    for (int64_t ip = 0; ip < nt; ++ip) {
        Part &p = ip < nreal ? c->parts[ip] : c->ghosts[ip - nreal];
        if (p.id < 0) continue;
        if (p.idata[MPM_I_PHASE] != 0) continue;
        double xp[3] = {p.pos[0], p.pos[1], p.pos[2]};
        int iv[3];
        c->cell(xp, iv);
        int mn[3], mx[3];
        for (int d = 0; d < 3; ++d) {
            int b0, b1;
            compute_bounds(iv[d], c->lo[d], c->hi[d], c->od[d], s.periodic[d], b0, b1);
            mn[d] = std::max(b0, 0 - iv[d]);
            mx[d] = std::min(b1, c->n[d] - iv[d] + 1);
        }
        for (int n = mn[2]; n < mx[2]; n++)
            for (int m = mn[1]; m < mx[1]; m++)
                for (int l = mn[0]; l < mx[0]; l++) {
                    int I = iv[0] + l, J = iv[1] + m, K = iv[2] + n;
                    int st[3] = {l, m, n};
                    double N, g[3] = {0, 0, 0};
                    if (update_forces)
                        N = basisval_and_grad(st, iv, xp, s.plo, c->dx, c->od, s.periodic,
                                              c->lo, c->hi, g);
                    else
                        N = basisval(st, iv, xp, s.plo, c->dx, c->od, s.periodic, c->lo,
                                     c->hi);
                    const double mass = p.r[MPM_R_MASS];
                    if (update_mass) atomic_add(&c->A(I, J, K, MPM_N_MASS), mass * N, par);
                    if (update_vel) {
                        double pc[3] = {mass * p.r[MPM_R_VEL] * N,
                                        mass * p.r[MPM_R_VEL + 1] * N,
                                        mass * p.r[MPM_R_VEL + 2] * N};
                        for (int d = 0; d < 3; ++d)
                            atomic_add(&c->A(I, J, K, MPM_N_VEL + d), pc[d], par);
                    }
                    if (update_forces) {
                        // get_tensor (interpolants.H:1673-1690)
                        double T[9];
                        int ind = 0;
                        for (int a = 0; a < 3; a++)
                            for (int b = a; b < 3; b++) {
                                T[a * 3 + b] = p.r[MPM_R_STRESS + ind];
                                T[b * 3 + a] = p.r[MPM_R_STRESS + ind];
                                ind++;
                            }
                        double bf[3] = {mass * s.gravity[0] * N, mass * s.gravity[1] * N,
                                        mass * s.gravity[2] * N};
                        if (s.external_loads && xp[0] > s.force_slab_lo[0] &&
                            xp[0] < s.force_slab_hi[0] && xp[1] > s.force_slab_lo[1] &&
                            xp[1] < s.force_slab_hi[1] && xp[2] > s.force_slab_lo[2] &&
                            xp[2] < s.force_slab_hi[2]) {
                            for (int d = 0; d < 3; d++) bf[d] += s.extforce[d] * N;
                        }
                        // tensor_vector_pdt (interpolants.H:1803-1815)
                        double tv[3];
                        for (int a = 0; a < 3; a++) {
                            tv[a] = 0.0;
                            for (int b = 0; b < 3; b++) tv[a] += T[a * 3 + b] * g[b];
                        }
                        double fi[3] = {-p.r[MPM_R_VOLUME] * tv[0], -p.r[MPM_R_VOLUME] * tv[1],
                                        -p.r[MPM_R_VOLUME] * tv[2]};
                        for (int d = 0; d < 3; d++)
                            atomic_add(&c->A(I, J, K, MPM_N_FRC + d), bf[d] + fi[d], par);
                    }
                }
    }
    // normalise (:590-627)
    for (int k = 0; k <= c->n[2]; ++k)
        for (int j = 0; j <= c->n[1]; ++j)
            for (int i = 0; i <= c->n[0]; ++i) {
                double m = c->A(i, j, k, MPM_N_MASS);
                if (update_vel && m > 0.0) {
                    for (int d = 0; d < 3; ++d) {
                        if (c->A(i, j, k, MPM_N_MASS) >= s.mass_tolerance)
                            c->A(i, j, k, MPM_N_VEL + d) /= c->A(i, j, k, MPM_N_MASS);
                        else
                            c->A(i, j, k, MPM_N_VEL + d) = 0.0;
                    }
                }
                if (update_forces == 2 && m > 0.0) {
                    if (m >= s.mass_tolerance) c->A(i, j, k, 15) /= m;
                    else c->A(i, j, k, 15) = 0.0;
                }
            }
    return MPM_OK;
}

// backup_current_velocity (nodal_data_ops.cpp:57-81).
int orc_backup_velocity(orc_ctx *c) {
    for (int k = 0; k <= c->n[2]; ++k)
        for (int j = 0; j <= c->n[1]; ++j)
            for (int i = 0; i <= c->n[0]; ++i)
                if (c->A(i, j, k, MPM_N_MASS) > 0.0) {
                    c->A(i, j, k, MPM_N_MASS_OLD) = c->A(i, j, k, MPM_N_MASS);
                    for (int d = 0; d < 3; d++)
                        c->A(i, j, k, MPM_N_DVEL + d) = c->A(i, j, k, MPM_N_VEL + d);
                }
    return MPM_OK;
}

// Nodal_Time_Update_Momentum (nodal_data_ops.cpp:451-482).
int orc_nodal_update(orc_ctx *c, double dt) {
    for (int k = 0; k <= c->n[2]; ++k)
        for (int j = 0; j <= c->n[1]; ++j)
            for (int i = 0; i <= c->n[0]; ++i) {
                if (c->A(i, j, k, MPM_N_MASS) >= c->s.mass_tolerance) {
                    for (int d = 0; d < 3; d++)
                        c->A(i, j, k, MPM_N_VEL + d) +=
                            c->A(i, j, k, MPM_N_FRC + d) / c->A(i, j, k, MPM_N_MASS) * dt;
                } else {
                    for (int d = 0; d < 3; d++) c->A(i, j, k, MPM_N_VEL + d) = 0.0;
                }
            }
    return MPM_OK;
}

// Apply_Nodal_BCs (utilities.cpp:183-204).
int orc_apply_nodal_bcs(orc_ctx *c, double dt, double t) {
    (void)dt;
    const mpm_specs &s = c->s;
    // nodal_bcs (nodal_data_ops.cpp:663-766), including the corner quirk:
    // relvel_in accumulates wall-velocity subtractions across directions and
    // the last direction's write wins.
    for (int k = 0; k <= c->n[2]; ++k)
        for (int j = 0; j <= c->n[1]; ++j)
            for (int i = 0; i <= c->n[0]; ++i) {
                int id[3] = {i, j, k};
                double rin[3], rout[3], wv[3] = {0, 0, 0};
                for (int d = 0; d < 3; d++) {
                    rin[d] = c->A(i, j, k, MPM_N_VEL + d);
                    rout[d] = rin[d];
                }
                for (int dir = 0; dir < 3; ++dir) {
                    int side = -1;
                    if (id[dir] == c->lo[dir]) side = 0;
                    else if (id[dir] == c->hi[dir] + 1) side = 1;
                    if (side < 0) continue;
                    const double *wvel = side ? s.wall_vel_hi : s.wall_vel_lo;
                    for (int d = 0; d < 3; ++d) {
                        wv[d] = wvel[dir * 3 + d];
                        rin[d] -= wv[d];
                    }
                    double nrm[3] = {0, 0, 0};
                    nrm[dir] = side ? -1.0 : 1.0;
                    int rc = applybc(rin, rout, side ? s.wall_mu_hi[dir] : s.wall_mu_lo[dir],
                                     nrm, side ? s.bchi[dir] : s.bclo[dir]);
                    if (rc < 0) return MPM_E_BC;
                    for (int q = 0; q < 3; ++q) c->A(i, j, k, MPM_N_VEL + q) = rout[q] + wv[q];
                }
            }
    // compute_udf_wall_vel_at_nodes + apply_udf_nodal_bcs (:929-1175)
    eval_udf_tables(c, t);
    bool any = false;
    for (int d = 0; d < 3; ++d) any |= (c->udf[d][0] || c->udf[d][1]);
    if (any) {
        for (int k = 0; k <= c->n[2]; ++k)
            for (int j = 0; j <= c->n[1]; ++j)
                for (int i = 0; i <= c->n[0]; ++i) {
                    int id[3] = {i, j, k};
                    for (int dir = 0; dir < 3; ++dir) {
                        const double *tab = nullptr;
                        int bc = -1;
                        if (id[dir] == c->lo[dir] && c->udf[dir][0]) {
                            tab = c->udf_tab[dir][0].data();
                            bc = s.bclo[dir];
                        } else if (id[dir] == c->hi[dir] + 1 && c->udf[dir][1]) {
                            tab = c->udf_tab[dir][1].data();
                            bc = s.bchi[dir];
                        }
                        if (!tab) continue;
                        int p0 = (dir == 0) ? 1 : 0;
                        int p1 = (dir == 2) ? 1 : 2;
                        size_t f = ((size_t)id[p0] * (c->n[p1] + 1) + id[p1]) * 3;
                        if (bc == MPM_BC_NOSLIP) {
                            for (int q = 0; q < 3; ++q) c->A(i, j, k, MPM_N_VEL + q) = tab[f + q];
                        } else {
                            c->A(i, j, k, MPM_N_VEL + dir) = tab[f + dir];
                        }
                    }
                }
    }
    // nodal_levelset_bcs (nodal_data_ops.cpp:129-220)
    for (const auto &b : c->bodies) {
        for (int k = 0; k <= c->n[2]; ++k)
            for (int j = 0; j <= c->n[1]; ++j)
                for (int i = 0; i <= c->n[0]; ++i) {
                    double xp[3] = {s.plo[0] + i * c->dx[0], s.plo[1] + j * c->dx[1],
                                    s.plo[2] + k * c->dx[2]};
                    double ls = ls_value(c, b, xp);
                    if (ls >= 0.0 || c->A(i, j, k, MPM_N_MASS) <= 0.0) continue;
                    double nrm[3] = {1.0, 0.0, 0.0};
                    ls_grad(c, b, xp, nrm);
                    double gm = 0.0;
                    for (int d = 0; d < 3; d++) gm += nrm[d] * nrm[d];
                    gm = std::sqrt(gm);
                    if (!(gm >= 1.0e-10)) continue;
                    for (int d = 0; d < 3; d++) nrm[d] /= gm;
                    double rin[3], rout[3];
                    for (int d = 0; d < 3; d++) {
                        rin[d] = c->A(i, j, k, MPM_N_VEL + d) - b.w[d];
                        rout[d] = rin[d];
                    }
                    double vn = 0.0;
                    for (int d = 0; d < 3; d++) vn += rin[d] * nrm[d];
                    if (vn > 0.0) continue;
                    if (applybc(rin, rout, b.mu, nrm, b.bc) < 0) return MPM_E_BC;
                    for (int d = 0; d < 3; d++) c->A(i, j, k, MPM_N_VEL + d) = rout[d] + b.w[d];
                }
    }
    // store_delta_velocity (nodal_data_ops.cpp:368-392)
    for (int k = 0; k <= c->n[2]; ++k)
        for (int j = 0; j <= c->n[1]; ++j)
            for (int i = 0; i <= c->n[0]; ++i)
                if (c->A(i, j, k, MPM_N_MASS) > 0.0)
                    for (int d = 0; d < 3; d++)
                        c->A(i, j, k, MPM_N_DVEL + d) =
                            c->A(i, j, k, MPM_N_VEL + d) - c->A(i, j, k, MPM_N_DVEL + d);
    return MPM_OK;
}

// interpolate_from_grid (mpm_particle_grid_ops.cpp:924-1098) with the
// interpolators bilin/quadratic/cubic_interp_two (interpolants.H:583-656,
// 863-1006, 1225-1366) and get_deformation_gradient_tensor (:1732-1787).
int orc_g2p(orc_ctx *c, int update_vel, int update_strainrate, double dt) {
    const mpm_specs &s = c->s;
    fill_boundary(c);
    const double alpha = s.alpha_pic_flip;
    const int64_t np = (int64_t)c->parts.size();
#pragma omp parallel for schedule(static) num_threads(c->nthreads) if (c->par())
    for (int64_t ip = 0; ip < np; ++ip) {
        Part &p = c->parts[ip];
        if (p.id < 0) continue;
        if (p.idata[MPM_I_PHASE] != 0) continue;
        double xp[3] = {p.pos[0], p.pos[1], p.pos[2]};
        int iv[3];
        c->cell(xp, iv);
        int mn[3], mx[3];
        for (int d = 0; d < 3; ++d)
            compute_bounds(iv[d], c->lo[d], c->hi[d], c->od[d], s.periodic[d], mn[d], mx[d]);
        if (update_vel) {
            for (int dim = 0; dim < 3; dim++) {
                double vals[2] = {0.0, 0.0};
                const int c0 = MPM_N_VEL + dim, c1 = MPM_N_DVEL + dim;
                const int o = c->od[dim];
                if (o == 1) {
                    const int i = iv[0], j = iv[1], k = iv[2];
                    double a = (xp[0] - (s.plo[0] + i * c->dx[0])) / c->dx[0];
                    double b = (xp[1] - (s.plo[1] + j * c->dx[1])) / c->dx[1];
                    double g = (xp[2] - (s.plo[2] + k * c->dx[2])) / c->dx[2];
                    double w000 = (1.0 - a) * (1.0 - b) * (1.0 - g);
                    double w100 = a * (1.0 - b) * (1.0 - g);
                    double w010 = (1.0 - a) * b * (1.0 - g);
                    double w110 = a * b * (1.0 - g);
                    double w001 = (1.0 - a) * (1.0 - b) * g;
                    double w101 = a * (1.0 - b) * g;
                    double w011 = (1.0 - a) * b * g;
                    double w111 = a * b * g;
                    for (int q = 0; q < 2; ++q) {
                        int cc = q ? c1 : c0;
                        vals[q] = w000 * c->A(i, j, k, cc) + w100 * c->A(i + 1, j, k, cc) +
                                  w010 * c->A(i, j + 1, k, cc) +
                                  w110 * c->A(i + 1, j + 1, k, cc) +
                                  w001 * c->A(i, j, k + 1, cc) +
                                  w101 * c->A(i + 1, j, k + 1, cc) +
                                  w011 * c->A(i, j + 1, k + 1, cc) +
                                  w111 * c->A(i + 1, j + 1, k + 1, cc);
                    }
                } else {
                    const double idx_[3] = {1.0 / c->dx[0], 1.0 / c->dx[1], 1.0 / c->dx[2]};
                    for (int n = mn[2]; n < mx[2]; ++n) {
                        int tz = node_type_chain(iv[2] + n, c->lo[2], c->hi[2]);
                        double rz = (o == 2)
                                        ? (xp[2] - (s.plo[2] + (iv[2] + n) * c->dx[2])) / c->dx[2]
                                        : (xp[2] - (s.plo[2] + (iv[2] + n) * c->dx[2])) * idx_[2];
                        int t = s.periodic[2] ? 3 : tz;
                        double nv = (o == 2) ? quad_val(t, rz) : cubic_val(t, rz);
                        for (int m = mn[1]; m < mx[1]; ++m) {
                            int ty = node_type_chain(iv[1] + m, c->lo[1], c->hi[1]);
                            double ry = (o == 2) ? (xp[1] - (s.plo[1] + (iv[1] + m) * c->dx[1])) / c->dx[1]
                                                 : (xp[1] - (s.plo[1] + (iv[1] + m) * c->dx[1])) * idx_[1];
                            int tt = s.periodic[1] ? 3 : ty;
                            double mv = (o == 2) ? quad_val(tt, ry) : cubic_val(tt, ry);
                            for (int l = mn[0]; l < mx[0]; ++l) {
                                int tx = node_type_chain(iv[0] + l, c->lo[0], c->hi[0]);
                                double rx = (o == 2) ? (xp[0] - (s.plo[0] + (iv[0] + l) * c->dx[0])) / c->dx[0]
                                                     : (xp[0] - (s.plo[0] + (iv[0] + l) * c->dx[0])) * idx_[0];
                                int t3 = s.periodic[0] ? 3 : tx;
                                double lv = (o == 2) ? quad_val(t3, rx) : cubic_val(t3, rx);
                                double w = lv * mv * nv;
                                vals[0] += w * c->A(iv[0] + l, iv[1] + m, iv[2] + n, c0);
                                vals[1] += w * c->A(iv[0] + l, iv[1] + m, iv[2] + n, c1);
                            }
                        }
                    }
                }
                p.r[MPM_R_VELP + dim] = vals[0];
                p.r[MPM_R_VEL + dim] = alpha * p.r[MPM_R_VEL + dim] + alpha * vals[1] +
                                       (1 - alpha) * vals[0];
            }
        }
        if (update_strainrate) {
            double L[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
            for (int n = mn[2]; n < mx[2]; ++n)
                for (int m = mn[1]; m < mx[1]; ++m)
                    for (int l = mn[0]; l < mx[0]; ++l) {
                        int st[3] = {l, m, n};
                        double g[3];
                        basisval_and_grad(st, iv, xp, s.plo, c->dx, c->od, s.periodic, c->lo,
                                          c->hi, g);
                        for (int d1 = 0; d1 < 3; ++d1) {
                            double v1 = c->A(iv[0] + l, iv[1] + m, iv[2] + n, MPM_N_VEL + d1);
                            for (int d2 = 0; d2 < 3; ++d2) L[d1][d2] += v1 * g[d2];
                        }
                    }
            // get_deformation_gradient_tensor (interpolants.H:1732-1787)
            double nrm = 0.0;
            for (int a = 0; a < 3; a++)
                for (int b = 0; b < 3; b++) nrm += L[a][b] * L[a][b];
            if (!(nrm >= 1.0e10)) {
                double Lp[3][3], Fp[3][3];
                for (int a = 0; a < 3; a++)
                    for (int b = 0; b < 3; b++) {
                        Lp[a][b] = (a == b) ? 1.0 + L[a][b] * dt : L[a][b] * dt;
                        Fp[a][b] = p.r[MPM_R_DEFGRAD + 3 * a + b];
                    }
                for (int a = 0; a < 3; ++a)
                    for (int b = 0; b < 3; ++b) {
                        double acc = 0.0;
                        for (int q = 0; q < 3; ++q) acc += Lp[a][q] * Fp[q][b];
                        p.r[MPM_R_DEFGRAD + a * 3 + b] = acc;
                    }
            }
            int ind = 0;
            for (int d1 = 0; d1 < 3; ++d1)
                for (int d2 = d1; d2 < 3; ++d2) {
                    p.r[MPM_R_STRAINRATE + ind] = 0.5 * (L[d1][d2] + L[d2][d1]);
                    ind++;
                }
        }
    }
    return MPM_OK;
}

// moveParticles (mpm_particle_timestep.cpp:226-578), then the
// Redistribute_Fill_Update that follows it (main.cpp:142).
int orc_move(orc_ctx *c, double dt, int step) {
    const mpm_specs &s = c->s;
    const int64_t np = (int64_t)c->parts.size();
    int err = 0;
#pragma omp parallel for schedule(static) num_threads(c->nthreads) if (c->par())
    for (int64_t ip = 0; ip < np; ++ip) {
        Part &p = c->parts[ip];
        if (p.id < 0) continue;
        for (int d = 0; d < 3; ++d) p.pos[d] += p.r[MPM_R_VELP + d] * dt;
        double rin[3], rout[3];
        for (int d = 0; d < 3; ++d) {
            rin[d] = p.r[MPM_R_VEL + d];
            rout[d] = p.r[MPM_R_VEL + d];
        }
        if (!c->bodies.empty() && p.idata[MPM_I_PHASE] == 0) {
            int hit = -1;
            double minphi = kTiny;
            for (size_t b = 0; b < c->bodies.size(); ++b) {
                double ph = ls_value(c, c->bodies[b], p.pos);
                if (ph < minphi) {
                    minphi = ph;
                    hit = (int)b;
                }
            }
            if (hit >= 0) {
                const LsBody &B = c->bodies[hit];
                double nrm[3];
                ls_grad(c, B, p.pos, nrm);
                double gm = 0.0;
                for (int d = 0; d < 3; ++d) gm += nrm[d] * nrm[d];
                gm = std::sqrt(gm);
                for (int d = 0; d < 3; ++d) nrm[d] /= (gm + kTiny);
                int mod = applybc(rin, rout, B.mu, nrm, B.bc);
                if (mod < 0) err = MPM_E_BC;
                if (mod > 0)
                    for (int d = 0; d < 3; ++d) p.pos[d] += 2.0 * std::fabs(minphi) * nrm[d];
                for (int d = 0; d < 3; ++d) p.r[MPM_R_VEL + d] = rout[d];
            }
        }
        double wv[3] = {0, 0, 0};
        for (int dir = 0; dir < 3; ++dir) {
            int side = -1;
            if (p.pos[dir] < s.plo[dir]) side = 0;
            else if (p.pos[dir] > s.phi[dir]) side = 1;
            if (side < 0) continue;
            const double *wvel = side ? s.wall_vel_hi : s.wall_vel_lo;
            for (int d = 0; d < 3; ++d) wv[d] = wvel[dir * 3 + d];
            const std::vector<double> &tab = c->udf_tab[dir][side];
            if (c->udf[dir][side] && !tab.empty()) {
                int p0 = (dir == 0) ? 1 : 0;
                int p1 = (dir == 2) ? 1 : 2;
                double f0 = (p.pos[p0] - s.plo[p0]) / c->dx[p0];
                double f1 = (p.pos[p1] - s.plo[p1]) / c->dx[p1];
                f0 = std::max(0.0, std::min((double)c->n[p0], f0));
                f1 = std::max(0.0, std::min((double)c->n[p1], f1));
                int j0 = (int)f0, j1 = std::min(j0 + 1, c->n[p0]);
                int k0 = (int)f1, k1 = std::min(k0 + 1, c->n[p1]);
                double wx1 = f0 - j0, wx0 = 1.0 - wx1;
                double wy1 = f1 - k0, wy0 = 1.0 - wy1;
                int n1 = c->n[p1] + 1;
                for (int q = 0; q < 3; ++q)
                    wv[q] = wx0 * wy0 * tab[(j0 * n1 + k0) * 3 + q] +
                            wx0 * wy1 * tab[(j0 * n1 + k1) * 3 + q] +
                            wx1 * wy0 * tab[(j1 * n1 + k0) * 3 + q] +
                            wx1 * wy1 * tab[(j1 * n1 + k1) * 3 + q];
            }
            for (int d = 0; d < 3; ++d) rin[d] -= wv[d];
            double nrm[3] = {0, 0, 0};
            nrm[dir] = side ? -1.0 : 1.0;
            int mod = applybc(rin, rout, side ? s.wall_mu_hi[dir] : s.wall_mu_lo[dir], nrm,
                              side ? s.bchi[dir] : s.bclo[dir]);
            if (mod < 0) err = MPM_E_BC;
            if (mod > 0) p.pos[dir] = 2.0 * (side ? s.phi[dir] : s.plo[dir]) - p.pos[dir];
            for (int q = 0; q < 3; ++q) p.r[MPM_R_VEL + q] = rout[q] + wv[q];
        }
    }
    if (err) return err;
    int nr = s.num_redist > 0 ? s.num_redist : 1;
    if (step % nr == 0) redistribute(c);
    fill_neighbors(c);
    return MPM_OK;
}

// Redistribute_Fill_Update (utilities.cpp:420-434) on its own: the
// redistribution every num_redist steps, the images refreshed every call.
int orc_redistribute(orc_ctx *c, int step) {
    const int nr = c->s.num_redist > 0 ? c->s.num_redist : 1;
    if (step % nr == 0) redistribute(c);
    fill_neighbors(c);
    return MPM_OK;
}

// updateVolume (mpm_particle_timestep.cpp:131-182).
int orc_update_volume(orc_ctx *c) {
    const int64_t np = (int64_t)c->parts.size();
#pragma omp parallel for schedule(static) num_threads(c->nthreads) if (c->par())
    for (int64_t ip = 0; ip < np; ++ip) {
        Part &p = c->parts[ip];
        if (p.id < 0 || p.idata[MPM_I_PHASE] != 0) continue;
        const double *F = &p.r[MPM_R_DEFGRAD];
        double det = F[0] * (F[4] * F[8] - F[5] * F[7]) - F[1] * (F[3] * F[8] - F[5] * F[6]) +
                     F[2] * (F[3] * F[7] - F[4] * F[6]);
        p.r[MPM_R_JACOBIAN] = det;
        p.r[MPM_R_VOLUME] = p.r[MPM_R_VOLINIT] * det;
        p.r[MPM_R_DENSITY] = p.r[MPM_R_MASS] / p.r[MPM_R_VOLUME];
    }
    return MPM_OK;
}

// Calculate_MP_Stress_Strain (utilities.cpp:372-399) and the constitutive
// updates (mpm_particle_container.cpp:43-272, constitutive_models.H:21-250).
static int stress_impl(orc_ctx *c, double dt, double rate, bool delta) {
    const int64_t np = (int64_t)c->parts.size();
    int err = 0;
#pragma omp parallel for schedule(static) num_threads(c->nthreads) if (c->par())
    for (int64_t ip = 0; ip < np; ++ip) {
        Part &p = c->parts[ip];
        if (p.id < 0 || p.idata[MPM_I_PHASE] != 0) continue;
        double *eps = &p.r[MPM_R_STRAIN];
        const double *D = &p.r[MPM_R_STRAINRATE];
        double *sig = &p.r[MPM_R_STRESS];
        const int model = p.idata[MPM_I_MODEL];
        for (int q = 0; q < 6; ++q) eps[q] += dt * D[q];
        eps[XX] += dt * rate;
        eps[YY] += dt * rate;
        eps[ZZ] += dt * rate;
        const double E = p.r[MPM_R_E], v = p.r[MPM_R_NU];
        if (delta) {
            if (model != 0) {
                err = MPM_E_UNSUPPORTED;
                continue;
            }
            double de[6];
            for (int q = 0; q < 6; ++q) de[q] = dt * D[q];
            de[XX] += dt * rate;
            de[YY] += dt * rate;
            de[ZZ] += dt * rate;
            double c1 = E / (1.0 + v) / (1.0 - 2.0 * v);
            double c2 = E / (1.0 + v);
            double ds[6];
            ds[XX] = c1 * ((1 - v) * de[XX] + v * (de[YY] + de[ZZ]));
            ds[YY] = c1 * ((1 - v) * de[YY] + v * (de[XX] + de[ZZ]));
            ds[ZZ] = c1 * ((1 - v) * de[ZZ] + v * (de[XX] + de[YY]));
            ds[XY] = c2 * de[XY];
            ds[XZ] = c2 * de[XZ];
            ds[YZ] = c2 * de[YZ];
            for (int q = 0; q < 6; ++q) sig[q] += ds[q];
            continue;
        }
        double out[6];
        if (model == 0) {
            double c1 = E / (1.0 + v) / (1.0 - 2.0 * v);
            double c2 = E / (1.0 + v);
            out[XX] = c1 * ((1 - v) * eps[XX] + v * (eps[YY] + eps[ZZ]));
            out[YY] = c1 * ((1 - v) * eps[YY] + v * (eps[XX] + eps[ZZ]));
            out[ZZ] = c1 * ((1 - v) * eps[ZZ] + v * (eps[XX] + eps[YY]));
            out[XY] = c2 * eps[XY];
            out[XZ] = c2 * eps[XZ];
            out[YZ] = c2 * eps[YZ];
        } else if (model == 1) {
            const double pinf = 0.0;
            p.r[MPM_R_PRESSURE] =
                p.r[MPM_R_BULK] * (std::pow(1.0 / p.r[MPM_R_JACOBIAN], p.r[MPM_R_GAMMA]) - 1.0) +
                pinf;
            const double mu = p.r[MPM_R_VISC], pr = p.r[MPM_R_PRESSURE];
            double tr = D[XX] + D[YY] + D[ZZ];
            double o3 = 1.0 / 3.0;
            out[XX] = 2.0 * mu * (D[XX] - o3 * tr) - pr;
            out[YY] = 2.0 * mu * (D[YY] - o3 * tr) - pr;
            out[ZZ] = 2.0 * mu * (D[ZZ] - o3 * tr) - pr;
            out[XY] = 2.0 * mu * D[XY];
            out[XZ] = 2.0 * mu * D[XZ];
            out[YZ] = 2.0 * mu * D[YZ];
        } else {
            const double *F = &p.r[MPM_R_DEFGRAD];
            double mu = E / (2.0 * (1.0 + v));
            double lam = E * v / ((1.0 + v) * (1.0 - 2.0 * v));
            auto rd = [&](int a, int b) {
                double acc = 0.0;
                for (int k = 0; k < 3; ++k) acc += F[a * 3 + k] * F[b * 3 + k];
                return acc;
            };
            double B[6];
            B[XX] = rd(0, 0);
            B[YY] = rd(1, 1);
            B[XY] = rd(0, 1);
            B[ZZ] = rd(2, 2);
            B[XZ] = rd(0, 2);
            B[YZ] = rd(1, 2);
            double det = F[0] * (F[4] * F[8] - F[5] * F[7]) - F[1] * (F[3] * F[8] - F[5] * F[6]) +
                         F[2] * (F[3] * F[7] - F[4] * F[6]);
            det = std::max(det, 1.e-12);
            const double vol = lam * std::log(det) / det;
            const double muJ = mu / det;
            out[XX] = muJ * (B[XX] - 1.0) + vol;
            out[YY] = muJ * (B[YY] - 1.0) + vol;
            out[ZZ] = muJ * (B[ZZ] - 1.0) + vol;
            out[XY] = muJ * B[XY];
            out[XZ] = muJ * B[XZ];
            out[YZ] = muJ * B[YZ];
        }
        for (int q = 0; q < 6; ++q) sig[q] = out[q];
    }
    return err;
}

int orc_stress(orc_ctx *c, double dt, double time) {
    const mpm_specs &s = c->s;
    const double rate = (time < s.applied_strainrate_time) ? s.applied_strainrate : 0.0;
    return stress_impl(c, dt, rate, s.calculate_strain_based_on_delta == 1);
}

// Calculate_time_step (mpm_particle_timestep.cpp:38-111).
int orc_calculate_dt(orc_ctx *c, double *dt_out) {
    const mpm_specs &s = c->s;
    if (s.fixed_timestep == 1) {
        *dt_out = s.timestep;
        return MPM_OK;
    }
    double dt = std::numeric_limits<double>::max();
    double dxmin = c->dx[0];
    for (int d = 1; d < 3; ++d) dxmin = std::min(dxmin, c->dx[d]);
    for (const auto &p : c->parts) {
        if (p.id < 0 || p.idata[MPM_I_PHASE] != 0) continue;
        double Cs;
        int m = p.idata[MPM_I_MODEL];
        if (m == 1) {
            Cs = std::sqrt(p.r[MPM_R_BULK] / p.r[MPM_R_DENSITY]);
        } else if (m == 0 || m == 2) {
            double E = p.r[MPM_R_E], nu = p.r[MPM_R_NU];
            double lam = E * nu / ((1 + nu) * (1 - 2.0 * nu));
            double mu = E / (2.0 * (1 + nu));
            Cs = std::sqrt((lam + 2.0 * mu) / p.r[MPM_R_DENSITY]);
        } else {
            return MPM_E_MODEL;
        }
        double vm = 0.0;
        for (int d = 0; d < 3; ++d) vm += p.r[MPM_R_VEL + d] * p.r[MPM_R_VEL + d];
        vm = std::sqrt(vm);
        double dtp = dxmin / (Cs + vm);
        dt = std::min(dt, dtp);
    }
    dt = s.CFL * dt;
    dt = std::max(std::min(dt, s.dt_max_limit), s.dt_min_limit);
    *dt_out = dt;
    return MPM_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Thermal module (the USE_TEMP build): heat conduction on the same
// particles and grid, as the reference's time loop and initial pass run it.
// ---------------------------------------------------------------------------
namespace {

enum {
    kMsp = MPM_N_MASS_SPHEAT,
    kMspt = MPM_N_MASS_SPHEAT_TEMP,
    kTemp = MPM_N_TEMPERATURE,
    kSrc = MPM_N_SOURCE_TEMP,
    kDtemp = MPM_N_DELTA_TEMPERATURE,
    kMcond = MPM_N_MASS_CONDUCTIVITY
};

// FabArray::SumBoundary(scomp, ncomp, periodicity) on the one-box nodal data
// (AMReX; called at mpm_particle_grid_ops.cpp:848-853): every valid node
// becomes the sum of the values at all of its periodic images inside the
// valid box, itself first, so a node on a periodic face adds the value of the
// opposite face node (which the ghost-particle images have already filled
// with the same deposits).
void sum_boundary(orc_ctx *c, int comp0, int ncomp) {
    const int *per = c->s.periodic;
    if (!per[0] && !per[1] && !per[2]) return;
    const int n0 = c->n[0] + 1, n1 = c->n[1] + 1, n2 = c->n[2] + 1;
    std::vector<double> tmp((size_t)n0 * n1 * n2);
    auto at = [&](int i, int j, int k) { return ((size_t)k * n1 + j) * n0 + i; };
    for (int comp = comp0; comp < comp0 + ncomp; ++comp) {
        for (int k = 0; k < n2; ++k)
            for (int j = 0; j < n1; ++j)
                for (int i = 0; i < n0; ++i) tmp[at(i, j, k)] = c->A(i, j, k, comp);
        for (int k = 0; k < n2; ++k)
            for (int j = 0; j < n1; ++j)
                for (int i = 0; i < n0; ++i) {
                    const int I[3] = {i, j, k};
                    int im[3][2], ni[3];
                    for (int d = 0; d < 3; ++d) {
                        ni[d] = 0;
                        im[d][ni[d]++] = I[d];
                        if (per[d] && I[d] == 0) im[d][ni[d]++] = c->n[d];
                        else if (per[d] && I[d] == c->n[d]) im[d][ni[d]++] = 0;
                    }
                    double sum = 0.0;
                    for (int a = 0; a < ni[2]; ++a)
                        for (int b = 0; b < ni[1]; ++b)
                            for (int e = 0; e < ni[0]; ++e)
                                sum += tmp[at(im[0][e], im[1][b], im[2][a])];
                    c->A(i, j, k, comp) = sum;
                }
    }
}

// deposit_onto_grid_temperature (mpm_particle_grid_ops.cpp:656-884).
void p2g_thermal(orc_ctx *c, bool reset, bool mass_temp, bool source) {
    const mpm_specs &s = c->s;
    if (reset)  // :681-700
        for (int k = 0; k <= c->n[2]; ++k)
            for (int j = 0; j <= c->n[1]; ++j)
                for (int i = 0; i <= c->n[0]; ++i) {
                    c->A(i, j, k, kMsp) = 0.0;
                    c->A(i, j, k, kMspt) = 0.0;
                    c->A(i, j, k, kSrc) = 0.0;
                    if (mass_temp) c->A(i, j, k, kMcond) = 0.0;
                }
    const int64_t nreal = (int64_t)c->parts.size();
    const int64_t nt = nreal + (int64_t)c->ghosts.size();
    const bool par = c->par();
#pragma omp parallel for schedule(static) num_threads(c->nthreads) if (par)
    for (int64_t ip = 0; ip < nt; ++ip) {
        const Part &p = ip < nreal ? c->parts[ip] : c->ghosts[ip - nreal];
        if (p.id < 0) continue;
        if (p.idata[MPM_I_PHASE] != 0) continue;
        double xp[3] = {p.pos[0], p.pos[1], p.pos[2]};
        int iv[3];
        c->cell(xp, iv);
        int mn[3], mx[3];
        for (int d = 0; d < 3; ++d) {
            int b0, b1;
            compute_bounds(iv[d], c->lo[d], c->hi[d], c->od[d], s.periodic[d], b0, b1);
            mn[d] = std::max(b0, 0 - iv[d]);
            mx[d] = std::min(b1, c->n[d] - iv[d] + 1);
        }
        const double mass = p.r[MPM_R_MASS], vol = p.r[MPM_R_VOLUME];
        const double cp = p.th[MPM_TH_SPECIFIC_HEAT], T = p.th[MPM_TH_TEMPERATURE];
        const double kc = p.th[MPM_TH_CONDUCTIVITY], Q = p.th[MPM_TH_HEAT_SOURCE];
        for (int n = mn[2]; n < mx[2]; n++)
            for (int m = mn[1]; m < mx[1]; m++)
                for (int l = mn[0]; l < mx[0]; l++) {
                    const int I = iv[0] + l, J = iv[1] + m, K = iv[2] + n;
                    const int st[3] = {l, m, n};
                    const double N = basisval(st, iv, xp, s.plo, c->dx, c->od, s.periodic,
                                              c->lo, c->hi);
                    if (mass_temp) {
                        atomic_add(&c->A(I, J, K, kMsp), mass * cp * N, par);
                        atomic_add(&c->A(I, J, K, kMspt), mass * cp * T * N, par);
                        atomic_add(&c->A(I, J, K, kMcond), mass * kc * N, par);
                    }
                    if (source) {
                        // basisvalder per direction: the gradient of
                        // basisval_and_grad (the same products).
                        double g[3];
                        basisval_and_grad(st, iv, xp, s.plo, c->dx, c->od, s.periodic, c->lo,
                                          c->hi, g);
                        double net = 0.0;
                        for (int d = 0; d < 3; ++d) net += p.th[MPM_TH_HEAT_FLUX + d] * g[d];
                        const double int_src = net * vol;
                        const double ext_src = vol * Q * N;
                        atomic_add(&c->A(I, J, K, kSrc), int_src + ext_src, par);
                    }
                }
    }
    if (mass_temp) {
        sum_boundary(c, kMsp, 2);
        sum_boundary(c, kMcond, 1);
    }
    if (source) sum_boundary(c, kSrc, 1);
    // :857-881 — a node without thermal mass keeps its temperature.
    if (mass_temp)
        for (int k = 0; k <= c->n[2]; ++k)
            for (int j = 0; j <= c->n[1]; ++j)
                for (int i = 0; i <= c->n[0]; ++i)
                    if (c->A(i, j, k, kMsp) > 0.0) {
                        if (c->A(i, j, k, kMsp) >= s.mass_tolerance)
                            c->A(i, j, k, kTemp) = c->A(i, j, k, kMspt) / c->A(i, j, k, kMsp);
                        else
                            c->A(i, j, k, kTemp) = 0.0;
                    }
}

// backup_current_temperature (nodal_data_ops.cpp:95-114).
void backup_temperature(orc_ctx *c) {
    for (int k = 0; k <= c->n[2]; ++k)
        for (int j = 0; j <= c->n[1]; ++j)
            for (int i = 0; i <= c->n[0]; ++i)
                if (c->A(i, j, k, kMsp) > 0.0) c->A(i, j, k, kDtemp) = c->A(i, j, k, kTemp);
}

// Nodal_Time_Update_Temperature (nodal_data_ops.cpp:500-527).
void nodal_update_temperature(orc_ctx *c, double dt) {
    for (int k = 0; k <= c->n[2]; ++k)
        for (int j = 0; j <= c->n[1]; ++j)
            for (int i = 0; i <= c->n[0]; ++i) {
                if (c->A(i, j, k, kMsp) >= c->s.mass_tolerance)
                    c->A(i, j, k, kTemp) += c->A(i, j, k, kSrc) / c->A(i, j, k, kMsp) * dt;
                else
                    c->A(i, j, k, kTemp) = 0.0;
            }
}

// nodal_bcs_temperature (nodal_data_ops.cpp:786-910) in the order a serial
// ParallelFor visits the nodal box (k outermost, i fastest): a wall node
// reads its inward neighbour after that neighbour's own update when the
// neighbour comes earlier in the sweep.  Unknown codes change nothing.
void bcs_temperature(orc_ctx *c) {
    const mpm_thermal_specs &t = c->ts;
    for (int k = 0; k <= c->n[2]; ++k)
        for (int j = 0; j <= c->n[1]; ++j)
            for (int i = 0; i <= c->n[0]; ++i) {
                const int I[3] = {i, j, k};
                bool applied = false;
                for (int d = 0; d < 3; ++d) {
                    const bool is_lo = I[d] == c->lo[d];
                    const bool is_hi = I[d] == c->hi[d] + 1;
                    if (!is_lo && !is_hi) continue;
                    const int bc = is_lo ? t.bclo[d] : t.bchi[d];
                    int nb[3] = {i, j, k};
                    nb[d] += is_lo ? 1 : -1;
                    double &T = c->A(i, j, k, kTemp);
                    const double Tnb = c->A(nb[0], nb[1], nb[2], kTemp);
                    if (bc == MPM_BC_TEMP_PERIODIC) {
                        applied = true;
                    } else if (bc == MPM_BC_TEMP_ISOTHERMAL) {
                        T = is_lo ? t.T_wall_lo[d] : t.T_wall_hi[d];
                        applied = true;
                    } else if (bc == MPM_BC_TEMP_ADIABATIC) {
                        if (!applied) {
                            T = Tnb;
                            applied = true;
                        }
                    } else if (bc == MPM_BC_TEMP_USERFLUX) {
                        const double q = is_lo ? t.flux_lo[d] : t.flux_hi[d];
                        const double mk = c->A(i, j, k, kMcond), m = c->A(i, j, k, MPM_N_MASS);
                        const double k_node = (m > 0.0) ? mk / m : 1.0;
                        T = Tnb + q * c->dx[d] / k_node;
                        applied = true;
                    } else if (bc == MPM_BC_TEMP_CONVECTION) {
                        const double hc = is_lo ? t.h_lo[d] : t.h_hi[d];
                        const double Tinf = is_lo ? t.T_inf_lo[d] : t.T_inf_hi[d];
                        const double mk = c->A(i, j, k, kMcond), m = c->A(i, j, k, MPM_N_MASS);
                        const double k_node = (m > 0.0) ? mk / m : 1.0;
                        const double Bi = hc * c->dx[d] / k_node;
                        T = (Tnb + Bi * Tinf) / (1.0 + Bi);
                        applied = true;
                    }
                }
            }
}

bool any_temp_udf(const orc_ctx *c) {
    for (int d = 0; d < 3; ++d)
        for (int h = 0; h < 2; ++h)
            if (c->tudf[d][h]) return true;
    return false;
}

// compute_udf_temp_at_nodes (nodal_data_ops.cpp:1178-1306), 3-D: the face
// functions at the face nodes, two values per node.
void eval_temp_udf_tables(orc_ctx *c, double t) {
    for (int d = 0; d < 3; ++d)
        for (int h = 0; h < 2; ++h) {
            if (!c->tudf[d][h]) continue;
            const int p0 = (d == 0) ? 1 : 0;
            const int p1 = (d == 2) ? 1 : 2;
            auto &tab = c->tudf_tab[d][h];
            tab.assign((size_t)(c->n[p0] + 1) * (c->n[p1] + 1) * 2, 0.0);
            for (int j = 0; j <= c->n[p0]; ++j)
                for (int k = 0; k <= c->n[p1]; ++k) {
                    double co[3] = {0, 0, 0};
                    co[d] = h ? c->s.phi[d] : c->s.plo[d];
                    co[p0] = c->s.plo[p0] + j * c->dx[p0];
                    co[p1] = c->s.plo[p1] + k * c->dx[p1];
                    double out[2] = {0.0, 0.0};
                    c->tudf[d][h](co[0], co[1], co[2], t, out);
                    const size_t f = ((size_t)j * (c->n[p1] + 1) + k) * 2;
                    tab[f] = out[0];
                    tab[f + 1] = out[1];
                }
        }
}

// apply_udf_nodal_bcs_temperature (nodal_data_ops.cpp:1308-1446): a second
// serial sweep over the nodal box; on a face with a function, the face's
// temperature BC type applied with the function's values (other types
// change nothing).
void udf_bcs_temperature(orc_ctx *c) {
    if (!any_temp_udf(c)) return;
    const mpm_thermal_specs &t = c->ts;
    for (int k = 0; k <= c->n[2]; ++k)
        for (int j = 0; j <= c->n[1]; ++j)
            for (int i = 0; i <= c->n[0]; ++i) {
                const int I[3] = {i, j, k};
                for (int dir = 0; dir < 3; ++dir) {
                    const double *tab = nullptr;
                    int bc = -1, sign = 0;
                    if (I[dir] == c->lo[dir] && c->tudf[dir][0]) {
                        tab = c->tudf_tab[dir][0].data();
                        bc = t.bclo[dir];
                        sign = 1;
                    } else if (I[dir] == c->hi[dir] + 1 && c->tudf[dir][1]) {
                        tab = c->tudf_tab[dir][1].data();
                        bc = t.bchi[dir];
                        sign = -1;
                    }
                    if (!tab) continue;
                    const int p0 = (dir == 0) ? 1 : 0;
                    const int p1 = (dir == 2) ? 1 : 2;
                    const size_t flat = ((size_t)I[p0] * (c->n[p1] + 1) + I[p1]) * 2;
                    const double v0 = tab[flat], v1 = tab[flat + 1];
                    double &T = c->A(i, j, k, kTemp);
                    int nb[3] = {i, j, k};
                    nb[dir] += sign;
                    if (bc == MPM_BC_TEMP_ISOTHERMAL) {
                        T = v0;
                    } else if (bc == MPM_BC_TEMP_USERFLUX || bc == MPM_BC_TEMP_CONVECTION) {
                        const double Tnb = c->A(nb[0], nb[1], nb[2], kTemp);
                        const double mk = c->A(i, j, k, kMcond), m = c->A(i, j, k, MPM_N_MASS);
                        const double k_node = (m > 0.0) ? mk / m : 1.0;
                        if (bc == MPM_BC_TEMP_USERFLUX) {
                            T = Tnb + v0 * c->dx[dir] / k_node;
                        } else {
                            const double Bi = v0 * c->dx[dir] / k_node;
                            T = (Tnb + Bi * v1) / (1.0 + Bi);
                        }
                    }
                }
            }
}

// nodal_levelset_bcs_temperature (nodal_data_ops.cpp:235-350): per body, in
// order, a serial sweep over the nodal box; dirichlet_only applies the
// isothermal bodies only.  The neighbour may be a ghost node (the level-set
// box is the nodal box grown by one).
void ls_bcs_temperature(orc_ctx *c, bool dirichlet_only) {
    const mpm_specs &s = c->s;
    for (const auto &b : c->bodies) {
        if (b.temp_bc == MPM_LS_TEMP_ADIABATIC) continue;
        if (dirichlet_only && b.temp_bc != MPM_LS_TEMP_ISOTHERMAL) continue;
        for (int k = 0; k <= c->n[2]; ++k)
            for (int j = 0; j <= c->n[1]; ++j)
                for (int i = 0; i <= c->n[0]; ++i) {
                    const double xp[3] = {s.plo[0] + i * c->dx[0], s.plo[1] + j * c->dx[1],
                                          s.plo[2] + k * c->dx[2]};
                    const double lsval = ls_value(c, b, xp);
                    if (lsval >= 0.0 || c->A(i, j, k, kMsp) <= 0.0) continue;
                    double nrm[3] = {1.0, 0.0, 0.0};
                    ls_grad(c, b, xp, nrm);
                    double gm = 0.0;
                    for (int d = 0; d < 3; d++) gm += nrm[d] * nrm[d];
                    gm = std::sqrt(gm);
                    if (gm < 1.0e-10) continue;
                    for (int d = 0; d < 3; d++) nrm[d] /= gm;
                    if (b.temp_bc == MPM_LS_TEMP_ISOTHERMAL) {
                        c->A(i, j, k, kTemp) = b.T_wall;
                        continue;
                    }
                    int dom = 0;
                    for (int d = 1; d < 3; d++)
                        if (std::abs(nrm[d]) > std::abs(nrm[dom])) dom = d;
                    int nb[3] = {i, j, k};
                    nb[dom] += (nrm[dom] > 0.0) ? 1 : -1;
                    if (nb[dom] < -1 || nb[dom] > c->n[dom] + 1) continue;
                    const double mk = c->A(i, j, k, kMcond), m = c->A(i, j, k, MPM_N_MASS);
                    const double k_node = (m > 0.0) ? mk / m : 1.0;
                    const double T_nb = c->A(nb[0], nb[1], nb[2], kTemp);
                    if (b.temp_bc == MPM_LS_TEMP_HEATFLUX) {
                        c->A(i, j, k, kTemp) = T_nb + b.heat_flux * c->dx[dom] / k_node;
                    } else if (b.temp_bc == MPM_LS_TEMP_CONVECTIVE) {
                        const double Bi = b.h_conv * c->dx[dom] / k_node;
                        c->A(i, j, k, kTemp) = (T_nb + Bi * b.T_inf) / (1.0 + Bi);
                    }
                }
    }
}

// store_delta_temperature (nodal_data_ops.cpp:409-430).
void store_delta_temperature(orc_ctx *c) {
    for (int k = 0; k <= c->n[2]; ++k)
        for (int j = 0; j <= c->n[1]; ++j)
            for (int i = 0; i <= c->n[0]; ++i)
                if (c->A(i, j, k, kMsp) > 0.0)
                    c->A(i, j, k, kDtemp) = c->A(i, j, k, kTemp) - c->A(i, j, k, kDtemp);
}

// interpolate_from_grid_temperature (mpm_particle_grid_ops.cpp:1129-1270)
// with bilin_interp / quadratic_interp / cubic_interp (interpolants.H:513-565,
// 680-862, 1029-1210), chosen by the x-direction order for all directions.
void g2p_thermal(orc_ctx *c, bool update_T, bool update_q) {
    const mpm_specs &s = c->s;
    fill_boundary(c);
    const int64_t np = (int64_t)c->parts.size();
#pragma omp parallel for schedule(static) num_threads(c->nthreads) if (c->par())
    for (int64_t ip = 0; ip < np; ++ip) {
        Part &p = c->parts[ip];
        if (p.id < 0) continue;
        if (p.idata[MPM_I_PHASE] != 0) continue;
        double xp[3] = {p.pos[0], p.pos[1], p.pos[2]};
        int iv[3];
        c->cell(xp, iv);
        int mn[3], mx[3];
        for (int d = 0; d < 3; ++d)
            compute_bounds(iv[d], c->lo[d], c->hi[d], c->od[d], s.periodic[d], mn[d], mx[d]);
        if (update_T) {
            const int o = c->od[0];
            double value = 0.0;
            if (o == 1) {
                const int i = iv[0], j = iv[1], k = iv[2];
                const double a = (xp[0] - (s.plo[0] + i * c->dx[0])) / c->dx[0];
                const double b = (xp[1] - (s.plo[1] + j * c->dx[1])) / c->dx[1];
                const double g = (xp[2] - (s.plo[2] + k * c->dx[2])) / c->dx[2];
                value = (1.0 - a) * (1.0 - b) * (1.0 - g) * c->A(i, j, k, kDtemp);
                value += a * (1.0 - b) * (1.0 - g) * c->A(i + 1, j, k, kDtemp);
                value += (1.0 - a) * b * (1.0 - g) * c->A(i, j + 1, k, kDtemp);
                value += a * b * (1.0 - g) * c->A(i + 1, j + 1, k, kDtemp);
                value += (1.0 - a) * (1.0 - b) * g * c->A(i, j, k + 1, kDtemp);
                value += a * (1.0 - b) * g * c->A(i + 1, j, k + 1, kDtemp);
                value += (1.0 - a) * b * g * c->A(i, j + 1, k + 1, kDtemp);
                value += a * b * g * c->A(i + 1, j + 1, k + 1, kDtemp);
            } else {
                const double inv[3] = {1.0 / c->dx[0], 1.0 / c->dx[1], 1.0 / c->dx[2]};
                auto w1 = [&](int d, int idx) {
                    const int t = s.periodic[d] ? 3 : node_type_chain(idx, c->lo[d], c->hi[d]);
                    const double x0 = s.plo[d] + idx * c->dx[d];
                    return (o == 2) ? quad_val(t, (xp[d] - x0) / c->dx[d])
                                    : cubic_val(t, (xp[d] - x0) * inv[d]);
                };
                for (int n = mn[2]; n < mx[2]; ++n)
                    for (int m = mn[1]; m < mx[1]; ++m)
                        for (int l = mn[0]; l < mx[0]; ++l) {
                            const double lv = w1(0, iv[0] + l), mv = w1(1, iv[1] + m),
                                         nv = w1(2, iv[2] + n);
                            value += lv * mv * nv * c->A(iv[0] + l, iv[1] + m, iv[2] + n, kDtemp);
                        }
            }
            p.th[MPM_TH_TEMPERATURE] += value;
        }
        if (update_q) {
            double gT[3] = {0.0, 0.0, 0.0};
            for (int n = mn[2]; n < mx[2]; ++n)
                for (int m = mn[1]; m < mx[1]; ++m)
                    for (int l = mn[0]; l < mx[0]; ++l) {
                        const int st[3] = {l, m, n};
                        double g[3];
                        basisval_and_grad(st, iv, xp, s.plo, c->dx, c->od, s.periodic, c->lo,
                                          c->hi, g);
                        const double Tn = c->A(iv[0] + l, iv[1] + m, iv[2] + n, kTemp);
                        for (int d = 0; d < 3; ++d) gT[d] += Tn * g[d];
                    }
            for (int d = 0; d < 3; ++d)
                p.th[MPM_TH_HEAT_FLUX + d] = -p.th[MPM_TH_CONDUCTIVITY] * gT[d];
        }
    }
}

// The thermal block of the initial pass (mpm_init.cpp:417-441), on nodal
// data that the reference has just zeroed (mpm_init.cpp:280-281).
void init_thermal(orc_ctx *c) {
    const size_t nodes = (size_t)c->N[0] * c->N[1] * c->N[2];
    std::fill(c->nd.begin() + (size_t)MPM_N_MASS_SPHEAT * nodes,
              c->nd.begin() + (size_t)MPM_N_NCOMP_TEMP * nodes, 0.0);
    p2g_thermal(c, true, true, false);
    backup_temperature(c);
    bcs_temperature(c);
    store_delta_temperature(c);
    g2p_thermal(c, true, true);
}

int check_thermal_specs(const mpm_thermal_specs *t) {
    for (int d = 0; d < 3; ++d)
        if (t->bclo[d] < MPM_BC_TEMP_PERIODIC || t->bclo[d] > MPM_BC_TEMP_CONVECTION ||
            t->bchi[d] < MPM_BC_TEMP_PERIODIC || t->bchi[d] > MPM_BC_TEMP_CONVECTION)
            return MPM_E_BC;
    return MPM_OK;
}

}  // namespace

extern "C" {

int orc_enable_thermal(orc_ctx *c, const mpm_thermal_specs *t) {
    if (!c || !t) return MPM_E_ARG;
    int rc = check_thermal_specs(t);
    if (rc) return rc;
    c->thermal = true;
    c->ts = *t;
    for (auto &p : c->parts)
        for (int f = 0; f < MPM_NTHERMAL; ++f) p.th[f] = 0.0;
    return MPM_OK;
}

int orc_upload_thermal(orc_ctx *c, int64_t n, const double *th) {
    if (!c || !c->thermal || n < 0 || (n > 0 && !th)) return MPM_E_ARG;
    for (auto &p : c->parts) {
        if (p.id < 0) continue;
        if (p.id >= n) return MPM_E_ARG;
        for (int f = 0; f < MPM_NTHERMAL; ++f) p.th[f] = th[(size_t)f * n + p.id];
    }
    return MPM_OK;
}

int orc_set_wall_temp_udf(orc_ctx *c, int dir, int hi, mpm_wall_temp_udf_fn fn) {
    if (!c || dir < 0 || dir > 2 || hi < 0 || hi > 1) return MPM_E_ARG;
    c->tudf[dir][hi] = fn;
    if (!fn) c->tudf_tab[dir][hi].clear();
    return MPM_OK;
}

int orc_set_levelset_body_temp(orc_ctx *c, int body, int temp_bc, double T_wall,
                               double heat_flux, double h_conv, double T_inf) {
    if (!c || body < 0 || body >= (int)c->bodies.size()) return MPM_E_ARG;
    if (temp_bc != MPM_LS_TEMP_ADIABATIC && temp_bc != MPM_LS_TEMP_ISOTHERMAL &&
        temp_bc != MPM_LS_TEMP_HEATFLUX && temp_bc != MPM_LS_TEMP_CONVECTIVE)
        return MPM_E_BC;
    LsBody &b = c->bodies[body];
    b.temp_bc = temp_bc;
    b.T_wall = T_wall;
    b.heat_flux = heat_flux;
    b.h_conv = h_conv;
    b.T_inf = T_inf;
    return MPM_OK;
}

int orc_download_thermal(orc_ctx *c, int64_t n, double *th) {
    if (!c || !c->thermal || n < 0 || (n > 0 && !th)) return MPM_E_ARG;
    std::fill(th, th + (size_t)MPM_NTHERMAL * n, std::numeric_limits<double>::quiet_NaN());
    for (const auto &p : c->parts) {
        if (p.id < 0) continue;
        if (p.id >= n) return MPM_E_ARG;
        for (int f = 0; f < MPM_NTHERMAL; ++f) th[(size_t)f * n + p.id] = p.th[f];
    }
    return MPM_OK;
}

// Initialise_Internal_Forces (mpm_init.cpp:368-401).
int orc_init_internal_forces(orc_ctx *c) {
    fill_neighbors(c);
    int rc = orc_p2g(c, 1, 1, 1);
    if (rc) return rc;
    orc_backup_velocity(c);
    rc = orc_g2p(c, 0, 1, 0.0);
    if (rc) return rc;
    // apply_constitutive_model(dt = 0, applied_strainrate): always the
    // total-strain form, with the applied rate regardless of time.
    rc = stress_impl(c, 0.0, c->s.applied_strainrate, false);
    if (rc) return rc;
    if (c->thermal) init_thermal(c);
    return MPM_OK;
}

int orc_step(orc_ctx *c, int nsteps) {
    const mpm_specs &s = c->s;
    int nr = s.num_redist > 0 ? s.num_redist : 1;
    for (int it = 0; it < nsteps; ++it) {
        c->steps++;
        const int step = (int)c->steps;
        // Redistribute_Fill_Update (utilities.cpp:420-434)
        if (step % nr == 0) redistribute(c);
        fill_neighbors(c);
        double dt;
        int rc = orc_calculate_dt(c, &dt);
        if (rc) return rc;
        orc_reset_nodal(c);
        if ((rc = orc_p2g(c, 1, 1, 1))) return rc;
        orc_backup_velocity(c);
        orc_nodal_update(c, dt);
        if ((rc = orc_apply_nodal_bcs(c, dt, c->time))) return rc;
        if (c->thermal) {  // main.cpp:99-105
            p2g_thermal(c, true, true, true);
            backup_temperature(c);
            nodal_update_temperature(c, dt);
            // Apply_Nodal_BCs_Temperature (utilities.cpp:221-240)
            bcs_temperature(c);
            if (any_temp_udf(c)) {
                eval_temp_udf_tables(c, c->time);
                udf_bcs_temperature(c);
            }
            ls_bcs_temperature(c, s.stress_update_scheme == MPM_MUSL);
            store_delta_temperature(c);
        }
        if (s.stress_update_scheme == MPM_USL) {
            orc_g2p(c, 1, 1, dt);
            if (c->thermal) g2p_thermal(c, true, true);  // main.cpp:112-114
            if ((rc = orc_move(c, dt, step))) return rc;
        } else {
            orc_g2p(c, 1, 0, dt);
            // second P2G uses the images filled at the start of the step
            if ((rc = orc_p2g(c, 0, 1, 0))) return rc;
            if ((rc = orc_apply_nodal_bcs(c, dt, c->time))) return rc;
            orc_g2p(c, 0, 1, dt);
            if (c->thermal) {  // main.cpp:130-135
                g2p_thermal(c, true, false);
                p2g_thermal(c, true, false, true);
                bcs_temperature(c);
                udf_bcs_temperature(c);
                ls_bcs_temperature(c, false);
                g2p_thermal(c, false, true);
            }
            if ((rc = orc_move(c, dt, step))) return rc;
        }
        orc_update_volume(c);
        if ((rc = orc_stress(c, dt, c->time))) return rc;
        c->time += dt;
        c->last_dt = dt;
    }
    return MPM_OK;
}

int orc_get_time(orc_ctx *c, double *time, int64_t *steps, double *last_dt) {
    if (time) *time = c->time;
    if (steps) *steps = c->steps;
    if (last_dt) *last_dt = c->last_dt;
    return MPM_OK;
}

int orc_add_levelset_body(orc_ctx *c, const double *phi, int bc, double mu,
                          const double w[3]) {
    if (!phi || (int)c->bodies.size() >= MPM_MAX_LS_BODIES) return MPM_E_ARG;
    if (bc < 0 || bc > MPM_BC_OUTFLOW) return MPM_E_BC;
    LsBody b;
    const int G0 = c->n[0] + 3, G1 = c->n[1] + 3, G2 = c->n[2] + 3;
    b.phi.assign((size_t)G0 * G1 * G2, 1.0);
    for (int k = 0; k <= c->n[2]; ++k)
        for (int j = 0; j <= c->n[1]; ++j)
            for (int i = 0; i <= c->n[0]; ++i)
                b.phi[((size_t)(k + 1) * G1 + (j + 1)) * G0 + (i + 1)] =
                    phi[((size_t)k * (c->n[1] + 1) + j) * (c->n[0] + 1) + i];
    // FillBoundary(periodicity) of the one-layer ghost ring
    for (int k = -1; k <= c->n[2] + 1; ++k)
        for (int j = -1; j <= c->n[1] + 1; ++j)
            for (int i = -1; i <= c->n[0] + 1; ++i) {
                int I[3] = {i, j, k}, S[3];
                bool ghost = false, ok = true;
                for (int d = 0; d < 3; ++d) {
                    S[d] = I[d];
                    if (I[d] < 0 || I[d] > c->n[d]) {
                        ghost = true;
                        if (!c->s.periodic[d]) ok = false;
                        else S[d] = I[d] < 0 ? I[d] + c->n[d] : I[d] - c->n[d];
                    }
                }
                if (!ghost || !ok) continue;
                b.phi[((size_t)(k + 1) * G1 + (j + 1)) * G0 + (i + 1)] =
                    b.phi[((size_t)(S[2] + 1) * G1 + (S[1] + 1)) * G0 + (S[0] + 1)];
            }
    b.bc = bc;
    b.mu = mu;
    for (int d = 0; d < 3; ++d) b.w[d] = w ? w[d] : 0.0;
    c->bodies.push_back(b);
    return MPM_OK;
}

int orc_set_wall_udf(orc_ctx *c, int dir, int hi, mpm_wall_udf_fn fn) {
    if (dir < 0 || dir > 2 || hi < 0 || hi > 1) return MPM_E_ARG;
    c->udf[dir][hi] = fn;
    if (!fn) c->udf_tab[dir][hi].clear();
    return MPM_OK;
}

// Calculate_Total_Energies (mpm_diagnostics.cpp:23-71) and the mass-weighted
// velocity sums (:94-127), over all live particles.
int orc_energies(orc_ctx *c, double out[6]) {
    double tke = 0, tse = 0, mv[3] = {0, 0, 0}, m = 0;
    for (const auto &p : c->parts) {
        if (p.id < 0) continue;
        double v2 = 0.0;
        for (int d = 0; d < 3; ++d) v2 += p.r[MPM_R_VEL + d] * p.r[MPM_R_VEL + d];
        tke += 0.5 * p.r[MPM_R_MASS] * v2;
        const double *S = &p.r[MPM_R_STRESS], *e = &p.r[MPM_R_STRAIN];
        double se = S[XX] * e[XX] + S[YY] * e[YY] + S[ZZ] * e[ZZ] +
                    2.0 * (S[XY] * e[XY] + S[YZ] * e[YZ] + S[XZ] * e[XZ]);
        tse += 0.5 * p.r[MPM_R_VOLUME] * se;
        for (int d = 0; d < 3; ++d) mv[d] += p.r[MPM_R_MASS] * p.r[MPM_R_VEL + d];
        m += p.r[MPM_R_MASS];
    }
    out[0] = tke;
    out[1] = tse;
    out[2] = mv[0];
    out[3] = mv[1];
    out[4] = mv[2];
    out[5] = m;
    return MPM_OK;
}

double orc_quadraticspline_1d(int t, double z) { return quad_val(t, z); }
double orc_quadraticspline_1d_der(int t, double z) { return quad_der(t, z); }
double orc_cubicspline_1d(int t, double z) { return cubic_val(t, z); }
double orc_cubicspline_1d_der(int t, double z) { return cubic_der(t, z); }

double orc_basisval_and_grad(const int st[3], const int cell[3], const double xp[3],
                             const double plo[3], const double dx[3], const int od[3],
                             const int per[3], const int lo[3], const int hi[3],
                             double grad[3]) {
    return basisval_and_grad(st, cell, xp, plo, dx, od, per, lo, hi, grad);
}

int orc_applybc(const double in[3], double out[3], double mu, const double n[3], int bc) {
    return applybc(in, out, mu, n, bc);
}

}  // extern "C"

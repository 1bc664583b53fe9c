// mpm_io.cpp — particle input formats (host code): the ASCII particle file
// reader of InitParticles (mpm_init.cpp:942-1167), the in-built box
// generator InitParticles(mincoords, maxcoords, ...) with generate_particle
// (mpm_init.cpp:1199-1382, 1386-1473), and a writer of the particle file in
// the layout the reference's generator writes
// (Generate_MPs_Inputfile_Generic.py:1020-1127).  Arrays are component-major
// as everywhere in the C ABI; a NULL pos asks for the particle count only.

#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "mpm_context.h"

using namespace xmpm;

namespace {

// constants.H:10-23 (the reference's truncated literals)
constexpr double kFourByThree = 1.3333333333;
constexpr double kPi = 3.141592654;
constexpr double kThree = 3.0;
constexpr double kFourth = 0.25;
constexpr double kHalf = 0.5;

struct Out {
    int64_t np;
    double *pos, *r;
    int *idata;
};

// Fields shared by both initialisers: F = I, J = 1, pressure 0, strain,
// stress and strain rate zero, rigid_body_id = -1, v' = 0.
void base_record(Out &o, int64_t i) {
    const int64_t n = o.np;
    for (int f = 0; f < MPM_NREAL; ++f) o.r[f * n + i] = 0.0;
    for (int d = 0; d < 3; ++d) o.r[(MPM_R_DEFGRAD + 4 * d) * n + i] = 1.0;
    o.r[MPM_R_JACOBIAN * n + i] = 1.0;
    o.idata[MPM_I_RIGID_ID * n + i] = -1;
}

// Reads one value with the reference's failure rule (safe_read: a failed
// extraction aborts with the given message).
template <typename T>
bool read_tok(std::istream &is, T &v, const char *what) {
    if (!(is >> v)) {
        set_err(MPM_E_ARG, what);
        return false;
    }
    return true;
}

int read_header(std::ifstream &ifs, int64_t &np) {
    std::string label;
    int dim = -1;
    if (!read_tok(ifs, label, "Error reading 'dim:' label")) return MPM_E_ARG;
    if (!read_tok(ifs, dim, "Error reading dimension")) return MPM_E_ARG;
    if (label != "dim:") return set_err(MPM_E_ARG, "mpm_particles.dat: Expected 'dim:' at line 1");
    if (dim != 3) return set_err(MPM_E_ARG, "Dimension mismatch between particle file and build");
    std::string label2;
    if (!read_tok(ifs, label2, "Error reading 'number_of_material_points:'")) return MPM_E_ARG;
    long n = -1;
    if (!read_tok(ifs, n, "Error reading number_of_material_points")) return MPM_E_ARG;
    if (label2 != "number_of_material_points:")
        return set_err(MPM_E_ARG, "Expected 'number_of_material_points:' at line 2");
    if (n <= 0) return set_err(MPM_E_ARG, "Invalid number_of_material_points");
    std::string line;
    std::getline(ifs, line);  // rest of line 2
    std::getline(ifs, line);  // line 3
    if (line.empty() || line[0] != '#')
        return set_err(MPM_E_ARG, "Expected header line beginning with '#'");
    np = n;
    return MPM_OK;
}

// InitParticles(filename); th non-null: the USE_TEMP build's layout, four
// more columns per particle after the model's (mpm_init.cpp:1097-1108).
int read_file(const char *path, int64_t *np, double *pos, double *rdata, int *idata, double *th,
              double totals[2]) {
    if (!path || !np) return set_err(MPM_E_ARG, "null argument");
    std::ifstream ifs(path);
    if (!ifs.good()) return set_err(MPM_E_ARG, "cannot open particle file");
    int64_t n = 0;
    int rc = read_header(ifs, n);
    if (rc) return rc;
    if (!pos) {
        *np = n;
        return MPM_OK;
    }
    if (*np != n) return set_err(MPM_E_ARG, "array size differs from number_of_material_points");
    if (!rdata || !idata) return set_err(MPM_E_ARG, "null argument");
    Out o{n, pos, rdata, idata};
    double tmass = 0.0, tvol = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        if (ifs.peek() == EOF)
            return set_err(MPM_E_ARG, "Unexpected EOF while reading particle data");
        base_record(o, i);
        int phase = 0, model = 0;
        double x[3], rad, dens, v[3];
        if (!read_tok(ifs, phase, "Error reading phase")) return MPM_E_ARG;
        for (int d = 0; d < 3; ++d)
            if (!read_tok(ifs, x[d], "Error reading particle position")) return MPM_E_ARG;
        if (!read_tok(ifs, rad, "Error reading radius")) return MPM_E_ARG;
        if (!read_tok(ifs, dens, "Error reading density")) return MPM_E_ARG;
        for (int d = 0; d < 3; ++d)
            if (!read_tok(ifs, v[d], "Error reading velocity")) return MPM_E_ARG;
        if (!read_tok(ifs, model, "Error reading constitutive_model")) return MPM_E_ARG;
        double E = 0.0, nu = 0.0, K = 0.0, gam = 0.0, mu = 0.0;
        if (model == 0 || model == 2) {
            if (!read_tok(ifs, E, "Error reading E")) return MPM_E_ARG;
            if (!read_tok(ifs, nu, "Error reading nu")) return MPM_E_ARG;
        } else if (model == 1) {
            if (!read_tok(ifs, K, "Error reading Bulk_modulus")) return MPM_E_ARG;
            if (!read_tok(ifs, gam, "Error reading Gama_pressure")) return MPM_E_ARG;
            if (!read_tok(ifs, mu, "Error reading Dynamic_viscosity")) return MPM_E_ARG;
        } else {
            return set_err(MPM_E_MODEL, "Incorrect constitutive model");
        }
        if (th) {
            double T, cp, kc, Q;
            if (!read_tok(ifs, T, "Error reading temperature")) return MPM_E_ARG;
            if (!read_tok(ifs, cp, "Error reading specific_heat")) return MPM_E_ARG;
            if (!read_tok(ifs, kc, "Error reading thermal_conductivity")) return MPM_E_ARG;
            if (!read_tok(ifs, Q, "Error reading heat_source")) return MPM_E_ARG;
            th[MPM_TH_TEMPERATURE * n + i] = T;
            th[MPM_TH_SPECIFIC_HEAT * n + i] = cp;
            th[MPM_TH_CONDUCTIVITY * n + i] = kc;
            for (int d = 0; d < 3; ++d) th[(MPM_TH_HEAT_FLUX + d) * n + i] = 0.0;
            th[MPM_TH_HEAT_SOURCE * n + i] = Q;
        }
        for (int d = 0; d < 3; ++d) {
            pos[d * n + i] = x[d];
            rdata[(MPM_R_VEL + d) * n + i] = v[d];
        }
        idata[MPM_I_PHASE * n + i] = phase;
        idata[MPM_I_MODEL * n + i] = model;
        rdata[MPM_R_RADIUS * n + i] = rad;
        rdata[MPM_R_DENSITY * n + i] = dens;
        rdata[MPM_R_E * n + i] = E;
        rdata[MPM_R_NU * n + i] = nu;
        rdata[MPM_R_BULK * n + i] = K;
        rdata[MPM_R_GAMMA * n + i] = gam;
        rdata[MPM_R_VISC * n + i] = mu;
        const double vol = kFourByThree * kPi * std::pow(rad, kThree);
        rdata[MPM_R_VOLUME * n + i] = vol;
        rdata[MPM_R_MASS * n + i] = dens * vol;
        rdata[MPM_R_VOLINIT * n + i] = vol;
        if (phase == 0) {
            tmass += dens * vol;
            tvol += vol;
        }
    }
    if (totals) {
        totals[0] = tmass;
        totals[1] = tvol;
    }
    return MPM_OK;
}

int load_file(mpm_ctx *c, const char *path, bool thermal, double totals[2]) {
    if (!c || !path) return set_err(MPM_E_ARG, "null argument");
    if (thermal && !c->thermal) return set_err(MPM_E_ARG, "thermal module not enabled");
    int64_t n = 0;
    int rc = read_file(path, &n, nullptr, nullptr, nullptr, nullptr, nullptr);
    if (rc) return rc;
    std::vector<double> pos(3 * n), r((size_t)MPM_NREAL * n), th;
    std::vector<int> idata((size_t)MPM_NINT * n);
    if (thermal) th.resize((size_t)MPM_NTHERMAL * n);
    rc = read_file(path, &n, pos.data(), r.data(), idata.data(), thermal ? th.data() : nullptr,
                   totals);
    if (rc) return rc;
    rc = mpm_upload_particles(c, n, pos.data(), r.data(), idata.data());
    if (rc || !thermal) return rc;
    return mpm_upload_thermal(c, n, th.data());
}

int write_file(const char *path, int64_t np, const double *pos, const double *rdata,
               const int *idata, const double *th) {
    if (!path || np < 1 || !pos || !rdata || !idata) return set_err(MPM_E_ARG, "bad argument");
    FILE *f = std::fopen(path, "w");
    if (!f) return set_err(MPM_E_ARG, "cannot open particle file for writing");
    const int m0 = idata[MPM_I_MODEL * np];
    std::fprintf(f, "dim: 3\nnumber_of_material_points: %lld\n", (long long)np);
    std::fprintf(f, "# phase x y z radius density vx vy vz cm_id%s%s\n",
                 m0 == 1 ? " Bulk_modulus Gamma_pressure Dynamic_viscosity" : " E nu",
                 th ? " T spheat thermcond heatsrc" : "");
    for (int64_t i = 0; i < np; ++i) {
        const int model = idata[MPM_I_MODEL * np + i];
        std::fprintf(f, "%d", idata[MPM_I_PHASE * np + i]);
        for (int d = 0; d < 3; ++d) std::fprintf(f, "\t%.6e", pos[d * np + i]);
        std::fprintf(f, "\t%.6e\t%.6e", rdata[MPM_R_RADIUS * np + i], rdata[MPM_R_DENSITY * np + i]);
        for (int d = 0; d < 3; ++d) std::fprintf(f, "\t%.6e", rdata[(MPM_R_VEL + d) * np + i]);
        std::fprintf(f, "\t%d", model);
        if (model == 1)
            std::fprintf(f, "\t%.6e\t%.6e\t%.6e", rdata[MPM_R_BULK * np + i],
                         rdata[MPM_R_GAMMA * np + i], rdata[MPM_R_VISC * np + i]);
        else
            std::fprintf(f, "\t%.6e\t%.6e", rdata[MPM_R_E * np + i], rdata[MPM_R_NU * np + i]);
        if (th)
            std::fprintf(f, "\t%.6e\t%.6e\t%.6e\t%.6e", th[MPM_TH_TEMPERATURE * np + i],
                         th[MPM_TH_SPECIFIC_HEAT * np + i], th[MPM_TH_CONDUCTIVITY * np + i],
                         th[MPM_TH_HEAT_SOURCE * np + i]);
        std::fprintf(f, "\n");
    }
    const bool ok = std::ferror(f) == 0;
    std::fclose(f);
    return ok ? MPM_OK : set_err(MPM_E_ARG, "write failed");
}

}  // namespace

extern "C" {

int mpm_read_particle_file(const char *path, int64_t *np, double *pos, double *rdata, int *idata,
                           double totals[2]) {
    return read_file(path, np, pos, rdata, idata, nullptr, totals);
}

int mpm_read_particle_file_thermal(const char *path, int64_t *np, double *pos, double *rdata,
                                   int *idata, double *th, double totals[2]) {
    if (pos && !th) return set_err(MPM_E_ARG, "null argument");
    return read_file(path, np, pos, rdata, idata, pos ? th : nullptr, totals);
}

int mpm_load_particle_file(mpm_ctx *c, const char *path, double totals[2]) {
    return load_file(c, path, false, totals);
}

int mpm_load_particle_file_thermal(mpm_ctx *c, const char *path, double totals[2]) {
    return load_file(c, path, true, totals);
}

int mpm_autogen_particles(const mpm_specs *s, const double mincoords[3],
                          const double maxcoords[3], const double vel[3], const int ppc[3],
                          double dens, int model, double E, double nu, double K, double gamma,
                          double visc, int multi_per_cell, int64_t *np, double *pos,
                          double *rdata, int *idata, double totals[2]) {
    if (!s || !mincoords || !maxcoords || !vel || !ppc || !np)
        return set_err(MPM_E_ARG, "null argument");
    double dx[3], cell_vol = 1.0;
    for (int d = 0; d < 3; ++d) {
        if (s->ncells[d] < 1 || !(s->phi[d] > s->plo[d])) return set_err(MPM_E_ARG, "bad domain");
        dx[d] = (s->phi[d] - s->plo[d]) / s->ncells[d];
        cell_vol *= dx[d];
    }
    // mpm_init.cpp:1239-1251: in 3-D the chain of else-ifs stops at the
    // SPACEDIM >= 2 branch, so ppc_z keeps its initial 1 whatever ppc[2] is.
    const int ppc_x = ppc[0], ppc_y = ppc[1], ppc_z = 1;
    if (multi_per_cell && (ppc_x < 1 || ppc_y < 1))
        return set_err(MPM_E_ARG, "particles per cell must be >= 1");
    std::vector<double> offx(ppc_x > 0 ? ppc_x : 1), offy(ppc_y > 0 ? ppc_y : 1), offz(1);
    for (int i = 0; i < ppc_x; ++i) offx[i] = (i + 0.5) / ppc_x;  // make_ppc_offsets
    for (int i = 0; i < ppc_y; ++i) offy[i] = (i + 0.5) / ppc_y;
    offz[0] = (0 + 0.5) / ppc_z;
    const int np_cell = ppc_x * ppc_y * ppc_z;
    const double vol_particle = multi_per_cell ? cell_vol / np_cell : cell_vol;
    auto inside = [&](const double c[3]) {
        for (int d = 0; d < 3; ++d)
            if (!(c[d] >= mincoords[d] && c[d] <= maxcoords[d])) return false;
        return true;
    };
    // Cells in the order a single-box MFIter visits them (x fastest).
    auto visit = [&](auto &&emit) {
        for (int k = 0; k < s->ncells[2]; ++k)
            for (int j = 0; j < s->ncells[1]; ++j)
                for (int i = 0; i < s->ncells[0]; ++i) {
                    const int iv[3] = {i, j, k};
                    if (!multi_per_cell) {
                        double c[3];
                        for (int d = 0; d < 3; ++d) c[d] = s->plo[d] + (iv[d] + kHalf) * dx[d];
                        if (inside(c)) emit(c);
                        continue;
                    }
                    double base[3];
                    for (int d = 0; d < 3; ++d) base[d] = s->plo[d] + iv[d] * dx[d];
                    for (int ix = 0; ix < ppc_x; ++ix)
                        for (int iy = 0; iy < ppc_y; ++iy)
                            for (int iz = 0; iz < ppc_z; ++iz) {
                                const double c[3] = {base[0] + offx[ix] * dx[0],
                                                     base[1] + offy[iy] * dx[1],
                                                     base[2] + offz[iz] * dx[2]};
                                if (inside(c)) emit(c);
                            }
                }
    };
    int64_t count = 0;
    visit([&](const double *) { ++count; });
    if (!pos) {
        *np = count;
        return MPM_OK;
    }
    if (*np != count) return set_err(MPM_E_ARG, "array size differs from the generated count");
    if (!rdata || !idata) return set_err(MPM_E_ARG, "null argument");
    Out o{count, pos, rdata, idata};
    const int64_t n = count;
    // generate_particle (mpm_init.cpp:1386-1473)
    const double rad = std::pow(kThree * kFourth * vol_particle / kPi, 1.0 / 3.0);
    double tmass = 0.0, tvol = 0.0;
    int64_t i = 0;
    visit([&](const double *c) {
        base_record(o, i);
        for (int d = 0; d < 3; ++d) {
            pos[d * n + i] = c[d];
            rdata[(MPM_R_VEL + d) * n + i] = vel[d];
        }
        idata[MPM_I_PHASE * n + i] = 0;
        idata[MPM_I_MODEL * n + i] = model;
        rdata[MPM_R_RADIUS * n + i] = rad;
        rdata[MPM_R_DENSITY * n + i] = dens;
        rdata[MPM_R_E * n + i] = E;
        rdata[MPM_R_NU * n + i] = nu;
        rdata[MPM_R_BULK * n + i] = K;
        rdata[MPM_R_GAMMA * n + i] = gamma;
        rdata[MPM_R_VISC * n + i] = visc;
        rdata[MPM_R_VOLUME * n + i] = vol_particle;
        rdata[MPM_R_MASS * n + i] = dens * vol_particle;
        rdata[MPM_R_VOLINIT * n + i] = vol_particle;
        tmass += dens * vol_particle;
        tvol += vol_particle;
        ++i;
    });
    if (totals) {
        totals[0] = tmass;
        totals[1] = tvol;
    }
    return MPM_OK;
}

int mpm_write_particle_file(const char *path, int64_t np, const double *pos, const double *rdata,
                            const int *idata) {
    return write_file(path, np, pos, rdata, idata, nullptr);
}

int mpm_write_particle_file_thermal(const char *path, int64_t np, const double *pos,
                                    const double *rdata, const int *idata, const double *th) {
    if (!th) return set_err(MPM_E_ARG, "null argument");
    return write_file(path, np, pos, rdata, idata, th);
}

}  // extern "C"

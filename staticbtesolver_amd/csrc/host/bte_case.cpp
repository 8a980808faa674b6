// bte_case.cpp -- C entry points over the C++ facade (include/bte_case.h).
#include "bte_case.h"

#include <cstdio>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>

#include "StaticBTESolver/StaticBTESolver.h"

struct bte_case {
    std::unique_ptr<BTEMesh> mesh;
    std::unique_ptr<BTEBand> bands;
    std::unique_ptr<BTEBoundaryCondition> bcs;
    std::unique_ptr<StaticBTESolver> solver;
    std::string err;
    int device = -1;
};

extern "C" {

int bte_case_open(const char* geometry, double Lx, double Ly, double Lz, const char* band_file, const char* bc_file,
                  int DM, int ntheta, int nphi, double T_ref, int device, bte_case** out, char* err, int errlen) {
    auto set_err = [&](const std::string& m) {
        if (err && errlen > 0) std::snprintf(err, errlen, "%s", m.c_str());
    };
    if (!out || !geometry || !band_file || !bc_file) {
        set_err("null argument");
        return BTE_ERR_ARG;
    }
    *out = nullptr;
    auto c = std::make_unique<bte_case>();
    c->device = device;
    try {
        std::string g(geometry);
        const auto dot = g.find_last_of('.');
        if (dot == std::string::npos) {
            c->mesh = std::make_unique<BTEMesh>(std::stoi(g), Lx);
        } else if (g.substr(dot + 1) == "mphtxt") {
            std::ifstream f(g);
            c->mesh = std::make_unique<BTEMesh>(f, Lx, Ly, Lz);
        } else {
            set_err("file format not supported");
            return BTE_ERR_ARG;
        }
        std::ifstream bf(band_file);
        c->bands = std::make_unique<BTEBand>(bf);
        std::ifstream cf(bc_file);
        c->bcs = std::make_unique<BTEBoundaryCondition>(cf);
    } catch (const std::exception& e) {
        set_err(e.what());
        return BTE_ERR_ARG;
    }
    c->solver = std::make_unique<StaticBTESolver>(c->mesh.get(), c->bcs.get(), c->bands.get());
    c->solver->setVerbose(false);
    c->solver->setDevice(device);
    c->solver->setParam(DM, ntheta, nphi, T_ref);
    if (c->solver->status() != BTE_OK) {
        set_err(c->solver->error());
        return c->solver->status();
    }
    if (c->solver->preprocess() != BTE_OK) {
        set_err(c->solver->error());
        return c->solver->status();
    }
    *out = c.release();
    return BTE_OK;
}

void bte_case_close(bte_case* c) { delete c; }

void bte_case_sizes(const bte_case* c, int* o) {
    const StaticBTESolver& s = *c->solver;
    o[0] = s.N_cell();
    o[1] = s.N_dir();
    o[2] = s.N_band();
    o[3] = s.N_face();
    o[4] = (int)c->bcs->size();
    o[5] = c->mesh->dim;
}

void bte_case_params(const bte_case* c, double* o) {
    o[0] = c->solver->T_ref();
    o[1] = c->solver->solid_angle();
}

int bte_case_solve(bte_case* c, const bte_opts* opts, int verbose, bte_result* res) {
    if (!c || !res) return BTE_ERR_ARG;
    StaticBTESolver& s = *c->solver;
    bte_opts o;
    if (opts)
        o = *opts;
    else
        bte_default_opts(&o);
    s.setOptions(o);
    s.setVerbose(verbose != 0);
    s.solve(o.max_iter);
    if (s.status() != BTE_OK && s.status() != BTE_ERR_NOCONV) {
        c->err = s.error();
        return s.status();
    }
    const bte_result& r = s.result();
    const int N = s.N_cell(), B = s.N_band(), nbc = (int)c->bcs->size();
    if (res->T) std::copy(s.cell_temperature().begin(), s.cell_temperature().end(), res->T);
    if (res->T_band)
        for (int i = 0; i < N; i++)
            for (int b = 0; b < B; b++) res->T_band[(size_t)i * B + b] = s.cell_band_temperature()[i][b];
    if (res->bc_flux) std::copy(s.bc_heat_flux().begin(), s.bc_heat_flux().end(), res->bc_flux);
    if (res->bc_band_flux)
        for (int k = 0; k < nbc; k++)
            for (int b = 0; b < B; b++) res->bc_band_flux[(size_t)k * B + b] = s.bc_band_heat_flux()[k][b];
    if (res->margins)
        for (int i = 0; i < (int)s.margins().size() && i < res->margins_capacity; i++) res->margins[i] = s.margins()[i];
    res->iterations = r.iterations;
    res->converged = r.converged;
    res->inner_iters = r.inner_iters;
    res->inner_failures = r.inner_failures;
    res->solve_seconds = r.solve_seconds;
    res->setup_seconds = r.setup_seconds;
    res->learned_accepted = r.learned_accepted;
    res->learned_proposed = r.learned_proposed;
    if (s.status() == BTE_ERR_NOCONV) c->err = s.error();
    return s.status();
}

int bte_case_geometry(bte_case* c, double* volume, double* centre, double* area, double* normal, int* nbr) {
    StaticBTESolver& s = *c->solver;
    if (volume) std::copy(s.cell_volume().begin(), s.cell_volume().end(), volume);
    if (centre) std::copy(s.cell_centers().begin(), s.cell_centers().end(), centre);
    if (area) std::copy(s.cell_face_area().begin(), s.cell_face_area().end(), area);
    if (normal) std::copy(s.cell_face_normal().begin(), s.cell_face_normal().end(), normal);
    if (nbr) std::copy(s.cell_neighbor_indices().begin(), s.cell_neighbor_indices().end(), nbr);
    return BTE_OK;
}

int bte_case_quadrature(bte_case* c, double* omega, double* dir, double* S) {
    StaticBTESolver& s = *c->solver;
    if (omega) std::copy(s.control_angles().begin(), s.control_angles().end(), omega);
    if (dir) std::copy(s.direction_vectors().begin(), s.direction_vectors().end(), dir);
    if (S) std::copy(s.S().begin(), s.S().end(), S);
    return BTE_OK;
}

bte_ctx* bte_case_ctx(bte_case* c) {
    bte_ctx* ctx = c->solver->context();
    if (!ctx) c->err = c->solver->error();
    return ctx;
}

const char* bte_case_error(const bte_case* c) { return c->err.c_str(); }

}  // extern "C"

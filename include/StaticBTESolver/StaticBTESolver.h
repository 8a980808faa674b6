// StaticBTESolver.h -- the solver class of the StaticBTESolver C++ API, running on one MI355X.
//
// Same constructor, setParam() and solve() as the reference (includes/StaticBTESolver/
// StaticBTESolver.h:81-95) and the same stdout lines; solve() runs the source iteration through
// the C ABI in include/bte.h.  Additions: options (precision, preconditioner, ...), read accessors
// for the results the reference keeps private, and a status instead of exit(1) on bad input.
#ifndef STATICBTE_AMD_STATICBTESOLVER_H
#define STATICBTE_AMD_STATICBTESOLVER_H

#include <memory>
#include <string>
#include <vector>

#include "../bte.h"
#include "BTEBand.h"
#include "BTEBoundaryCondition.h"
#include "BTEMesh.h"
#include "utility.h"

class StaticBTESolver {
public:
    // The three pointers are not owned and must outlive solve() (as in the reference).
    StaticBTESolver(BTEMesh* mesh, BTEBoundaryCondition* bcs, BTEBand* bands);
    ~StaticBTESolver();
    StaticBTESolver(const StaticBTESolver&) = delete;
    StaticBTESolver& operator=(const StaticBTESolver&) = delete;

    // DM: material dimension (2 or 3); num_theta, num_phi: angular resolution; T_ref: reference
    // temperature (forced to 0 on 1D meshes, StaticBTESolver.cpp:55).
    void setParam(int DM, int num_theta, int num_phi, double T_ref);
    // Runs at most max_iter outer iterations; prints the reference's per-iteration lines.
    void solve(int max_iter);

    // ---- additions ----
    void setOptions(const bte_opts& opts);
    const bte_opts& options() const { return opts_; }
    void setVerbose(bool v) { verbose_ = v; }
    int status() const { return status_; }                // BTE_OK or a negative bte_status
    const std::string& error() const { return error_; }
    int device() const { return device_; }
    void setDevice(int d) { device_ = d; }

    int N_cell() const { return N_cell_; }
    int N_dir() const { return N_dir_; }
    int N_band() const { return N_band_; }
    int N_face() const { return N_face_; }
    double T_ref() const { return T_ref_; }
    double solid_angle() const { return solid_angle_; }
    const std::vector<double>& cell_temperature() const { return cell_temperature_; }
    const vector2D<double>& cell_band_temperature() const { return cell_band_temperature_; }
    const std::vector<double>& bc_heat_flux() const { return bc_heat_flux_; }
    const vector2D<double>& bc_band_heat_flux() const { return bc_band_heat_flux_; }
    const std::vector<double>& margins() const { return margins_; }
    int iterations() const { return result_.iterations; }
    const bte_result& result() const { return result_; }
    // Geometry / quadrature of _preprocess as flat arrays (also the C ABI inputs).
    const std::vector<double>& control_angles() const { return omega_; }
    const std::vector<double>& direction_vectors() const { return dir_; }  // [N_dir*3]
    const std::vector<double>& S() const { return S_; }                    // [N_dir*3]
    const std::vector<double>& cell_volume() const { return vol_; }
    const std::vector<double>& cell_centers() const { return centre_; }    // [N_cell*3]
    const std::vector<double>& cell_face_area() const { return area_; }    // [N_cell*N_face]
    const std::vector<double>& cell_face_normal() const { return normal_; }  // [N_cell*N_face*3]
    const std::vector<int>& cell_neighbor_indices() const { return nbr_; }   // [N_cell*N_face]
    // Final intensities e[b][d][c] after solve().
    std::vector<double> intensity();
    // Writes "x/Lx y/Ly [z/Lz] T" per cell (the commented-out Tempcell.dat writer,
    // StaticBTESolver.cpp:818-831).
    bool writeTemperature(const std::string& path) const;
    // Prepares geometry and quadrature without solving (idempotent).
    int preprocess();
    // The C ABI context holding this problem on the GPU (created and loaded on first use; owned
    // by the solver).  NULL on failure, see error().
    bte_ctx* context();
    int DM() const { return DM_; }
    int num_theta() const { return num_theta_; }
    int num_phi() const { return num_phi_; }

private:
    int fail(int code, const std::string& msg);
    int upload();

    BTEMesh* mesh_;
    BTEBoundaryCondition* bcs_;
    BTEBand* bands_;
    int DM_ = 3, num_theta_ = 4, num_phi_ = 4;
    double T_ref_ = 300, solid_angle_ = 4 * PI, WFACTOR_ = 0;
    int N_cell_ = 0, N_dir_ = 0, N_band_ = 0, N_face_ = 0;
    bool params_set_ = false, preprocessed_ = false;
    std::vector<double> omega_, dir_, S_, vol_, centre_, area_, normal_;
    std::vector<int> nbr_;
    std::vector<double> cell_temperature_, bc_heat_flux_, margins_;
    vector2D<double> cell_band_temperature_, bc_band_heat_flux_;
    bte_opts opts_;
    bte_result result_{};
    bte_ctx* ctx_ = nullptr;
    int device_ = -1;
    bool verbose_ = true;
    bool uploaded_ = false;  // the context holds this problem (cleared by setParam)
    int status_ = 0;
    std::string error_;
};

#endif

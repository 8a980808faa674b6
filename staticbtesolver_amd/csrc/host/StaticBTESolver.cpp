// StaticBTESolver.cpp -- C++ facade of the MI355X solver (reference class StaticBTESolver,
// src/StaticBTESolver.cpp).  Host work here is the once-per-solve preprocessing (angular
// quadrature, cell volumes, face areas and outward normals, face neighbours); the source iteration
// itself runs on the GPU through the C ABI (include/bte.h).
#include "StaticBTESolver/StaticBTESolver.h"

#include <algorithm>
#include <array>
#include <cstdint>
#include <chrono>
#include <cstdio>
#include <fstream>
#include <unordered_map>

namespace {

const double kPi = M_PI;

struct V3 {
    double x, y, z;
};
inline V3 pt(const std::shared_ptr<staticbtesolver::Point>& p) { return {p->x, p->y, p->z}; }
inline V3 sub(const V3& a, const V3& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline double dot(const V3& a, const V3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(const V3& a, const V3& b) { return {a.y * b.z - b.y * a.z, b.x * a.z - a.x * b.z, a.x * b.y - b.x * a.y}; }
inline double norm(const V3& a) { return std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }

// Sorted-vertex key of a face (an edge in 2D, a triangle in 3D).
struct FaceKey {
    std::array<int, 3> v;
    bool operator==(const FaceKey& o) const { return v == o.v; }
};
struct FaceKeyHash {
    size_t operator()(const FaceKey& k) const {
        uint64_t h = 1469598103934665603ull;
        for (int x : k.v) h = (h ^ (uint64_t)(uint32_t)x) * 1099511628211ull;
        return (size_t)h;
    }
};
FaceKey make_key(int a, int b, int c = -1) {
    FaceKey k{{a, b, c}};
    std::sort(k.v.begin(), k.v.end());
    return k;
}

}  // namespace

StaticBTESolver::StaticBTESolver(BTEMesh* mesh, BTEBoundaryCondition* bcs, BTEBand* bands)
    : mesh_(mesh), bcs_(bcs), bands_(bands) {
    bte_default_opts(&opts_);
}

StaticBTESolver::~StaticBTESolver() {
    if (ctx_) bte_ctx_destroy(ctx_);
}

int StaticBTESolver::fail(int code, const std::string& msg) {
    status_ = code;
    error_ = msg;
    if (verbose_) std::printf("%s\n", msg.c_str());
    return code;
}

void StaticBTESolver::setOptions(const bte_opts& opts) { opts_ = opts; }

// setParam, StaticBTESolver.cpp:48-94.
void StaticBTESolver::setParam(int DM, int num_theta, int num_phi, double T_ref) {
    status_ = BTE_OK;
    error_.clear();
    DM_ = DM;
    num_theta_ = num_theta;
    num_phi_ = num_phi;
    T_ref_ = T_ref;
    solid_angle_ = 4 * kPi;
    params_set_ = false;
    preprocessed_ = false;
    uploaded_ = false;
    if (!mesh_ || !bands_ || !bcs_) {
        fail(BTE_ERR_ARG, "StaticBTESolver: mesh, boundary conditions and bands must not be null");
        return;
    }
    if (DM != 2 && DM != 3) {
        fail(BTE_ERR_ARG, "DM must be 2 or 3");
        return;
    }
    if (num_theta <= 0 || num_phi <= 0) {
        fail(BTE_ERR_ARG, "num_theta and num_phi must be positive");
        return;
    }
    const int dim = mesh_->dim;
    if (dim == 1) {
        T_ref_ = 0;  // 1D intensities are absolute (:55)
        N_cell_ = (int)mesh_->elements1D.size();
        N_dir_ = (DM == 3) ? num_theta * num_phi : num_theta;
        if (DM == 2) solid_angle_ = 2 * kPi;
        N_face_ = 2;
    } else if (dim == 2) {
        N_cell_ = (int)mesh_->elements2D.size();
        WFACTOR_ = (DM == 3) ? 2.0 : 4 * kPi;
        N_dir_ = (DM == 3) ? 4 * num_theta * num_phi : 4 * num_phi;
        N_face_ = 3;
    } else if (dim == 3) {
        if (DM == 2) {
            // :84-87 would set 4*nt*np directions while _preprocess builds 4*np (out of bounds)
            fail(BTE_ERR_ARG, "DM=2 on a 3D mesh is not supported");
            return;
        }
        N_cell_ = (int)mesh_->elements3D.size();
        WFACTOR_ = 1.0;
        N_dir_ = 8 * num_theta * num_phi;
        N_face_ = 4;
    } else {
        fail(BTE_ERR_MESH, "DG is not set properly.");
        return;
    }
    N_band_ = (int)bands_->size();
    if (N_cell_ <= 0 || N_band_ <= 0) {
        fail(BTE_ERR_ARG, "empty mesh or band list");
        return;
    }
    params_set_ = true;
}

// _preprocess, StaticBTESolver.cpp:372-688.
int StaticBTESolver::preprocess() {
    if (!params_set_) return status_ ? status_ : fail(BTE_ERR_STATE, "setParam must be called before solve");
    if (preprocessed_) return BTE_OK;
    const int dim = mesh_->dim, N = N_cell_, F = N_face_, D = N_dir_;
    omega_.assign(D, 0);
    dir_.assign(3 * (size_t)D, 0);
    S_.assign(3 * (size_t)D, 0);
    // ---- angular quadrature (:374-492) ----
    if (DM_ == 2 && dim >= 2) {
        // 4*num_phi in-plane directions at the midpoints of equal phi intervals, S = omega * d
        const double dphi = 0.5 * kPi / num_phi_;
        for (int k = 0; k < D; k++) {
            const double phi = (k + 0.5) * dphi, w = WFACTOR_ / D;
            omega_[k] = w;
            dir_[3 * k] = std::sin(phi);
            dir_[3 * k + 1] = std::cos(phi);
            S_[3 * k] = w * std::sin(phi);
            S_[3 * k + 1] = w * std::cos(phi);
        }
    } else if (DM_ == 2) {
        // 1D: Gauss-Legendre in theta on [0, pi], weight doubled (the half-circle is symmetric)
        std::vector<double> g = staticbtesolver::GaussIntegrationPoints(0, kPi, D);
        WFACTOR_ = 0;
        for (int k = 0; k < D; k++) {
            const double ct = std::cos(g[2 * k]);
            omega_[k] = g[2 * k + 1] * 2;
            dir_[3 * k] = ct;
            dir_[3 * k + 1] = std::pow(1 - ct * ct, 0.5);
            WFACTOR_ += omega_[k];
        }
    } else if (dim >= 2) {
        // midpoint rule in theta and phi; S is the exact integral of s over each angular patch
        const double dth = 0.5 * kPi / num_theta_, dph = 0.5 * kPi / num_phi_;
        const int nth = (dim == 2) ? num_theta_ : 2 * num_theta_, nph = 4 * num_phi_;
        for (int it = 0; it < nth; it++) {
            const double th = (it + 0.5) * dth;
            for (int ip = 0; ip < nph; ip++) {
                const double ph = (ip + 0.5) * dph;
                const int k = ip + it * nph;
                omega_[k] = WFACTOR_ * 2 * std::sin(th) * std::sin(0.5 * dth) * dph;
                const double radial = std::sin(0.5 * dph) * (dth - std::cos(2 * th) * std::sin(dth));
                S_[3 * k] = WFACTOR_ * std::sin(ph) * radial;
                S_[3 * k + 1] = WFACTOR_ * std::cos(ph) * radial;
                S_[3 * k + 2] = WFACTOR_ * 0.5 * dph * std::sin(2 * th) * std::sin(dth);
                dir_[3 * k] = std::sin(th) * std::sin(ph);
                dir_[3 * k + 1] = std::sin(th) * std::cos(ph);
                dir_[3 * k + 2] = std::cos(th);
            }
        }
    } else {
        // 1D with DM=3: Gauss in cos(theta) on [-1, 1] and in phi on [-pi, pi]; the reference uses
        // node/pi as the cosine of phi (:476)
        std::vector<double> g = staticbtesolver::GaussIntegrationPoints(-1, 1, num_theta_);
        std::vector<double> gp = staticbtesolver::GaussIntegrationPoints(-kPi, kPi, num_phi_);
        WFACTOR_ = 0;
        for (int i = 0; i < num_theta_; i++) {
            const double ct = g[2 * i], st = std::pow(1 - ct * ct, 0.5);
            for (int j = 0; j < num_phi_; j++) {
                const double cp = gp[2 * j] / kPi, sp = std::pow(1 - cp * cp, 0.5);
                const int k = i * num_phi_ + j;
                omega_[k] = g[2 * i + 1] * gp[2 * j + 1];
                dir_[3 * k] = ct;
                dir_[3 * k + 1] = st * cp;
                dir_[3 * k + 2] = st * sp;
                WFACTOR_ += omega_[k];
            }
        }
    }
    // ---- cells: centres, volumes (:494-543), faces: areas, outward normals (:544-598) ----
    vol_.assign(N, 0);
    centre_.assign(3 * (size_t)N, 0);
    area_.assign((size_t)N * F, 0);
    normal_.assign(3 * (size_t)N * F, 0);
    nbr_.assign((size_t)N * F, 0);
    const auto& P = mesh_->meshPts;
    auto set_normal = [&](int c, int f, V3 n) {
        const double l = norm(n);
        normal_[3 * ((size_t)c * F + f)] = n.x / l;
        normal_[3 * ((size_t)c * F + f) + 1] = n.y / l;
        normal_[3 * ((size_t)c * F + f) + 2] = n.z / l;
    };
    if (dim == 1) {
        for (int c = 0; c < N; c++) {
            centre_[3 * c] = mesh_->L_x / N * (c + 0.5);
            vol_[c] = mesh_->L_x / N;
            area_[2 * c] = area_[2 * c + 1] = 1;
            normal_[6 * c] = -1;
            normal_[6 * c + 3] = 1;
            nbr_[2 * c] = c - 1;
            nbr_[2 * c + 1] = c + 1;
        }
        nbr_[0] = -1;               // left end: boundary 0 (:609)
        nbr_[2 * (N - 1) + 1] = -2;  // right end: boundary 1 (:610)
    } else if (dim == 2) {
        for (int c = 0; c < N; c++) {
            const int* ix = mesh_->elements2D[c]->index;
            const V3 a = pt(P[ix[0]]), b = pt(P[ix[1]]), cc = pt(P[ix[2]]);
            centre_[3 * c] = (a.x + b.x + cc.x) / 3;
            centre_[3 * c + 1] = (a.y + b.y + cc.y) / 3;
            vol_[c] = 0.5 * norm(cross(sub(cc, a), sub(b, a)));
            for (int e = 0; e < 3; e++) {
                // face e is the edge opposite vertex e; its normal points away from vertex e
                const V3 p0 = pt(P[ix[e]]), p1 = pt(P[ix[(e + 1) % 3]]), p2 = pt(P[ix[(e + 2) % 3]]);
                area_[3 * c + e] = norm(sub(p1, p2));
                V3 n = {p2.y - p1.y, p1.x - p2.x, 0};
                if (dot(n, sub(p0, p1)) > 0) n = {0 - n.x, 0 - n.y, 0 - n.z};
                set_normal(c, e, n);
            }
        }
    } else {
        for (int c = 0; c < N; c++) {
            const int* ix = mesh_->elements3D[c]->index;
            const V3 a = pt(P[ix[0]]), b = pt(P[ix[1]]), cc = pt(P[ix[2]]), d = pt(P[ix[3]]);
            centre_[3 * c] = (a.x + b.x + cc.x + d.x) / 4;
            centre_[3 * c + 1] = (a.y + b.y + cc.y + d.y) / 4;
            centre_[3 * c + 2] = (a.z + b.z + cc.z + d.z) / 4;
            vol_[c] = staticbtesolver::getVolume(*P[ix[0]], *P[ix[1]], *P[ix[2]], *P[ix[3]]);
            for (int f = 0; f < 4; f++) {
                const V3 p0 = pt(P[ix[f]]), p1 = pt(P[ix[(f + 1) % 4]]), p2 = pt(P[ix[(f + 2) % 4]]),
                         p3 = pt(P[ix[(f + 3) % 4]]);
                V3 n = cross(sub(p3, p1), sub(p3, p2));
                area_[4 * c + f] = 0.5 * norm(n);
                if (dot(n, sub(p0, p1)) > 0) n = {0 - n.x, 0 - n.y, 0 - n.z};
                set_normal(c, f, n);
            }
        }
    }
    // ---- neighbours (:600-687): the reference takes the lowest-numbered other cell sharing the
    // face, and if the face keeps neighbour 0 it takes the first boundary element (file order) on
    // it, encoded -entity-1.  Hash maps of sorted face keys give the same answer in O(N). ----
    if (dim >= 2) {
        const int nv = dim;  // vertices per face
        std::unordered_map<FaceKey, std::vector<int>, FaceKeyHash> cells_of;
        cells_of.reserve((size_t)N * F);
        auto face_key = [&](int c, int f) {
            const int* ix = (dim == 2) ? mesh_->elements2D[c]->index : mesh_->elements3D[c]->index;
            return nv == 2 ? make_key(ix[(f + 1) % F], ix[(f + 2) % F]) : make_key(ix[(f + 1) % F], ix[(f + 2) % F], ix[(f + 3) % F]);
        };
        for (int c = 0; c < N; c++)
            for (int f = 0; f < F; f++) cells_of[face_key(c, f)].push_back(c);  // ascending c
        std::unordered_map<FaceKey, int, FaceKeyHash> boundary;
        if (dim == 2) {
            for (const auto& s : mesh_->elements1D) boundary.emplace(make_key(s->index[0], s->index[1]), s->entity_index);
        } else {
            for (const auto& t : mesh_->elements2D)
                boundary.emplace(make_key(t->index[0], t->index[1], t->index[2]), t->entity_index);
        }
        for (int c = 0; c < N; c++)
            for (int f = 0; f < F; f++) {
                const FaceKey k = face_key(c, f);
                int nb = 0;
                for (int o : cells_of[k])
                    if (o != c) {
                        nb = o;
                        break;
                    }
                if (nb == 0) {
                    auto it = boundary.find(k);
                    if (it != boundary.end()) nb = -it->second - 1;
                }
                nbr_[(size_t)c * F + f] = nb;
            }
    }
    preprocessed_ = true;
    return BTE_OK;
}

int StaticBTESolver::upload() {
    int rc;
    if (!ctx_ && (rc = bte_ctx_create(device_, &ctx_)) != BTE_OK) return fail(rc, "no usable HIP device for the solver");
    if ((rc = bte_set_geometry(ctx_, mesh_->dim, N_cell_, N_face_, vol_.data(), area_.data(), normal_.data(),
                               nbr_.data())))
        return fail(rc, bte_last_error(ctx_));
    if ((rc = bte_set_quadrature(ctx_, DM_, N_dir_, omega_.data(), dir_.data(), S_.data(), num_theta_, num_phi_,
                                 solid_angle_, T_ref_)))
        return fail(rc, bte_last_error(ctx_));
    std::vector<double> v, tau, C, Lr;
    for (const auto& b : *bands_) {
        v.push_back(b.group_velocity);
        tau.push_back(b.relaxation_time);
        C.push_back(b.Ctot);
        Lr.push_back(b.Lr);
    }
    if ((rc = bte_set_bands(ctx_, N_band_, v.data(), tau.data(), C.data(), Lr.data())))
        return fail(rc, bte_last_error(ctx_));
    std::vector<int> idx, type;
    std::vector<double> T;
    for (const auto& bc : *bcs_) {
        idx.push_back(bc.index);
        type.push_back(bc.type);
        T.push_back(bc.temperature);
    }
    if ((rc = bte_set_bcs(ctx_, (int)idx.size(), idx.data(), type.data(), T.data())))
        return fail(rc, bte_last_error(ctx_));
    return BTE_OK;
}

// solve, StaticBTESolver.cpp:863-878 (prints as _iteration :691-797 and _get_heat_flux :365-368).
void StaticBTESolver::solve(int max_iter) {
    if (preprocess() != BTE_OK) return;
    if (upload() != BTE_OK) return;
    uploaded_ = true;
    const int nbc = (int)bcs_->size();
    cell_temperature_.assign(N_cell_, 0);
    std::vector<double> Tb((size_t)N_cell_ * N_band_), bflux(nbc), bbf((size_t)nbc * N_band_);
    margins_.assign(std::max(max_iter, 0), 0);
    bte_opts o = opts_;
    o.max_iter = max_iter;
    o.verbose = verbose_ ? 1 : 0;
    result_ = bte_result{};
    result_.T = cell_temperature_.data();
    result_.T_band = Tb.data();
    result_.bc_flux = bflux.data();
    result_.bc_band_flux = bbf.data();
    result_.margins = margins_.data();
    result_.margins_capacity = (int)margins_.size();
    const int rc = bte_solve(ctx_, &o, &result_);
    if (rc != BTE_OK && rc != BTE_ERR_NOCONV) {
        fail(rc, bte_last_error(ctx_));
        return;
    }
    status_ = rc;
    if (rc == BTE_ERR_NOCONV) error_ = bte_last_error(ctx_);
    margins_.resize(result_.iterations);
    cell_band_temperature_.assign(N_cell_, std::vector<double>(N_band_));
    for (int c = 0; c < N_cell_; c++)
        for (int b = 0; b < N_band_; b++) cell_band_temperature_[c][b] = Tb[(size_t)c * N_band_ + b];
    bc_heat_flux_ = bflux;
    bc_band_heat_flux_.assign(nbc, std::vector<double>(N_band_));
    for (int k = 0; k < nbc; k++)
        for (int b = 0; b < N_band_; b++) bc_band_heat_flux_[k][b] = bbf[(size_t)k * N_band_ + b];
    if (verbose_) std::printf("Time taken by iteration: %g milliseconds\n", result_.solve_seconds * 1e3);
    result_.T = result_.T_band = result_.bc_flux = result_.bc_band_flux = result_.margins = nullptr;
}

bte_ctx* StaticBTESolver::context() {
    if (preprocess() != BTE_OK) return nullptr;
    if (!uploaded_ && upload() != BTE_OK) return nullptr;
    uploaded_ = true;
    return ctx_;
}

std::vector<double> StaticBTESolver::intensity() {
    std::vector<double> ee;
    if (!ctx_) return ee;
    ee.resize((size_t)N_band_ * N_dir_ * N_cell_);
    if (bte_get_intensity(ctx_, ee.data()) != BTE_OK) ee.clear();
    return ee;
}

bool StaticBTESolver::writeTemperature(const std::string& path) const {
    std::ofstream out(path);
    if (!out) return false;
    for (int c = 0; c < N_cell_ && c < (int)cell_temperature_.size(); c++) {
        out << centre_[3 * c] / mesh_->L_x << " " << centre_[3 * c + 1] / mesh_->L_y << " ";
        if (mesh_->dim == 3) out << centre_[3 * c + 2] / mesh_->L_z << " ";
        out << cell_temperature_[c] << "\n";
    }
    return true;
}

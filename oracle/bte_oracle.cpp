// bte_oracle.cpp -- CPU restatement of StaticBTESolver's steady-state source iteration.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker for the MI355X library: it is loaded by
// tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and by nothing else.
// The product (staticbtesolver_amd/) never links, loads or calls it.
//
// Every function restates one piece of the reference (yuchengs/StaticBTESolver) and cites the
// file:line it follows.  Parity status of each piece (how it is pinned) is stated at the
// function.  Three inner solvers are provided:
//   ORC_INNER_EXACT   : the upwind matrix is triangular under a topological order of its
//                       dependency graph, so a downwind sweep solves it exactly; if the graph has
//                       a cycle, right-preconditioned GMRES(50) with the partial sweep as the
//                       preconditioner runs to a true relative residual of 1e-14.
//   ORC_INNER_PETSC   : restatement of PETSc's documented default KSP for a SeqAIJ matrix, as
//                       StaticBTESolver.cpp:270-308 configures it: GMRES(30), classical
//                       Gram-Schmidt, left ILU(0) in natural order, rtol 1e-9 on the
//                       preconditioned residual norm, maxits 1000, zero initial guess.
//   ORC_INNER_BICGSTAB: restatement of the reference GPU solver src/bicgstab.hpp:138-300
//                       (right-Jacobi BiCGSTAB, tol 1e-9, maxit 1000, breakdown restarts).
//
// Parity pins (tests/test_oracle_*.py):
//   - reader + outer loop + RHS (types 1/2/31/32/33) + T update + margin + heat flux: the five
//     golden logs tests/{2DM1DG,3DM1DG,2DM2DG,3DM2DG,3DM3DG}/*.sol of the reference (6 digits)
//     and the 2Dcase2 CPU log; also the equal-wall-temperature exact fixed point.
//   - quadrature: Gauss-Legendre against numpy.polynomial.legendre.leggauss; sum of weights =
//     solid angle; sum of omega*d = 0; S integrals against brute-force patch quadrature.
//   - geometry: sum_f A_f n_f = 0 per cell, total volume, outward normals, neighbour symmetry.
//   - assembly: dense brute-force assembly written independently in the test; interior row
//     sums equal omega*V.
//   - inner solvers: true residual of each solve, and EXACT vs GMRES vs BiCGSTAB agreement.

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <memory>
#include <sstream>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace {

const double kPi = 3.14159265358979323846;

struct V3 {
    double x = 0, y = 0, z = 0;
};
inline double dot3(const V3& a, const V3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 sub3(const V3& a, const V3& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
// utility.cpp:14-16 (cross_prod component formulas)
inline V3 cross3(const V3& p1, const V3& p2) {
    return {p1.y * p2.z - p2.y * p1.z, p2.x * p1.z - p1.x * p2.z, p1.x * p2.y - p2.x * p1.y};
}
inline double len3(const V3& a) { return std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }

struct Csr {
    std::vector<int> rowptr, col;
    std::vector<double> val;
};

struct Case {
    std::string err;
    // mesh (BTEMesh.h:35-42)
    int dim = 0;
    double Lx = 0, Ly = 0, Lz = 0;
    std::vector<V3> pts;
    std::vector<std::array<int, 2>> seg;
    std::vector<int> seg_ent;
    std::vector<std::array<int, 3>> tri;
    std::vector<int> tri_ent;
    std::vector<std::array<int, 4>> tet;
    // bands (utility.h:251-268)
    std::vector<double> vg, tau, Ctot, Lr;
    // boundary conditions (utility.h:232-246)
    std::vector<int> bc_index, bc_type;
    std::vector<double> bc_T;
    // parameters (StaticBTESolver.cpp:48-94)
    int DM = 3, num_theta = 4, num_phi = 4;
    double T_ref = 300, solid_angle = 4 * kPi, WFACTOR = 0;
    int N_cell = 0, N_dir = 0, N_band = 0, N_face = 0;
    bool params_set = false, preprocessed = false;
    // quadrature + geometry (StaticBTESolver.cpp:372-688)
    std::vector<double> omega;
    std::vector<V3> dirv, S;
    std::vector<double> vol;
    std::vector<V3> centre;
    std::vector<double> farea;  // [c*F+f]
    std::vector<V3> fnorm;      // [c*F+f]
    std::vector<int> nbr;       // [c*F+f]
    // state
    std::vector<double> ee_curr, ee_prev;  // [b][d][c]
    std::vector<double> Tband;             // [c][b]
    std::vector<double> dens;              // [c][b]
    std::vector<double> T;                 // [c]
    std::vector<double> margins;
    std::vector<double> flux, band_flux;  // [bc], [bc][b]
    int iters = 0;
    // inner solver settings
    int inner_mode = 0;
    double inner_tol = -1;  // <0: mode default
    int inner_maxit = -1;
    long long inner_iter_total = 0;
    double worst_true_resid = 0;
    double stop_margin = 1e-5;
    // exact mode: topological sweep order and lagged entries per direction.  The upwind pattern of
    // (b, d) depends on d only (the band enters as the scalar v_b tau_b), so the order found for
    // one band of d serves every band.  Geometry and quadrature are fixed when the case is created,
    // so the cache holds for the case's lifetime.
    std::vector<std::vector<int>> order_of_dir;
    std::vector<std::vector<char>> lagged_of_dir;
    // workers of an outer iteration's systems (directions split round-robin, as the reference
    // splits them over MPI ranks); guards err when a worker reports a bad boundary type
    int n_threads = 1;
    std::mutex err_mu;
};

// ---------------------------------------------------------------------------------------------
// Readers
// ---------------------------------------------------------------------------------------------

// BTEMesh.cpp:7-130 (COMSOL .mphtxt reader).  Parity: reader goldens + oracle/_ref dump.
bool read_mphtxt(Case& cs, const char* path, double Lx, double Ly, double Lz) {
    std::ifstream in(path);
    if (!in.is_open()) {
        cs.err = std::string("file not open: ") + path;
        return false;
    }
    cs.Lx = Lx;
    cs.Ly = Ly;
    cs.Lz = Lz;
    std::string line;
    bool found = false;
    while (std::getline(in, line))
        if (line == "4 Mesh # class") {
            found = true;
            break;
        }
    if (!found) {
        cs.err = "no '4 Mesh # class' line";
        return false;
    }
    std::getline(in, line);  // "4 # version"
    std::getline(in, line);
    cs.dim = std::atoi(line.c_str());
    std::getline(in, line);
    int npts = std::atoi(line.c_str());
    while (std::getline(in, line))
        if (line == "# Mesh point coordinates") break;
    cs.pts.resize(npts);
    for (int i = 0; i < npts; i++) {
        double x = 0, y = 0, z = 0;
        if (cs.dim == 3) {
            in >> x >> y >> z;
            cs.pts[i] = {x * Lx, y * Ly, z * Lz};
        } else if (cs.dim == 2) {
            in >> x >> y;
            cs.pts[i] = {x * Lx, y * Ly, 0};
        } else {
            in >> x;
            cs.pts[i] = {x * Lx, 0, 0};
        }
    }
    auto seek = [&](const char* tag) {
        while (std::getline(in, line))
            if (line == tag) return true;
        return false;
    };
    auto skip4 = [&]() {
        for (int i = 0; i < 4; i++) std::getline(in, line);
    };
    if (!seek("3 edg # type name")) {
        cs.err = "no edg section";
        return false;
    }
    skip4();
    int n1 = std::atoi(line.c_str());
    std::getline(in, line);
    cs.seg.resize(n1);
    cs.seg_ent.assign(n1, 0);
    for (int i = 0; i < n1; i++) in >> cs.seg[i][0] >> cs.seg[i][1];
    skip4();
    for (int i = 0; i < n1; i++) in >> cs.seg_ent[i];
    if (cs.dim == 1) return true;
    if (!seek("3 tri # type name")) {
        cs.err = "no tri section";
        return false;
    }
    skip4();
    int n2 = std::atoi(line.c_str());
    std::getline(in, line);
    cs.tri.resize(n2);
    cs.tri_ent.assign(n2, 0);
    for (int i = 0; i < n2; i++) in >> cs.tri[i][0] >> cs.tri[i][1] >> cs.tri[i][2];
    if (cs.dim == 2) return true;
    skip4();
    for (int i = 0; i < n2; i++) in >> cs.tri_ent[i];
    if (!seek("3 tet # type name")) {
        cs.err = "no tet section";
        return false;
    }
    skip4();
    int n3 = std::atoi(line.c_str());
    std::getline(in, line);
    cs.tet.resize(n3);
    for (int i = 0; i < n3; i++) in >> cs.tet[i][0] >> cs.tet[i][1] >> cs.tet[i][2] >> cs.tet[i][3];
    return true;
}

// BTEMesh.cpp:132-147 (uniform 1D mesh of N_cell segments on [0, L_x]).
void make_1d(Case& cs, int n, double Lx) {
    cs.dim = 1;
    cs.Lx = Lx;
    cs.Ly = cs.Lz = 0;
    cs.pts.resize(n + 1);
    for (int i = 0; i <= n; i++) cs.pts[i] = {Lx / n * i, 0, 0};
    cs.seg.resize(n);
    cs.seg_ent.assign(n, 0);
    for (int i = 0; i < n; i++) cs.seg[i] = {i, i + 1};
}

// BTEBand.cpp:8-26: one band per line "v tau Ctot Lr", stop at the first line of length <= 1.
bool read_bands(Case& cs, const char* path) {
    std::ifstream in(path);
    if (!in.is_open()) {
        cs.err = std::string("file not open: ") + path;
        return false;
    }
    std::string line;
    while (std::getline(in, line)) {
        if (line.length() <= 1) break;
        std::stringstream ss(line);
        double v = 0, t = 0, c = 0, l = 0;
        ss >> v >> t >> c >> l;
        cs.vg.push_back(v);
        cs.tau.push_back(t);
        cs.Ctot.push_back(c);
        cs.Lr.push_back(l);
    }
    return true;
}

// BTEBoundaryCondition.cpp:7-23: count, skip the rest of that line and two more, then rows
// "k type T" stored with index = -k-1.
bool read_bcs(Case& cs, const char* path) {
    std::ifstream in(path);
    if (!in.is_open()) {
        cs.err = std::string("file not open: ") + path;
        return false;
    }
    int num = 0;
    in >> num;
    std::string line;
    for (int i = 0; i < 3; i++) std::getline(in, line);
    for (int i = 0; i < num; i++) {
        int k = 0, type = 0;
        double T = 0;
        in >> k >> type >> T;
        cs.bc_index.push_back(-k - 1);
        cs.bc_type.push_back(type);
        cs.bc_T.push_back(T);
    }
    return true;
}

// ---------------------------------------------------------------------------------------------
// setParam (StaticBTESolver.cpp:48-94)
// ---------------------------------------------------------------------------------------------
bool set_param(Case& cs, int DM, int nt, int np, double Tref) {
    cs.num_theta = nt;
    cs.num_phi = np;
    cs.T_ref = Tref;
    cs.DM = DM;
    cs.solid_angle = 4 * kPi;
    if (cs.dim == 1) {
        cs.T_ref = 0;  // :55 (quirk Q2)
        cs.N_cell = (int)cs.seg.size();
        if (DM == 3)
            cs.N_dir = nt * np;
        else if (DM == 2) {
            cs.N_dir = nt;
            cs.solid_angle = 2 * kPi;
        }
        cs.N_face = 2;
    }
    if (cs.dim == 2) {
        cs.N_cell = (int)cs.tri.size();
        if (DM == 3) {
            cs.WFACTOR = 2.0;
            cs.N_dir = 4 * nt * np;
        } else if (DM == 2) {
            cs.WFACTOR = 4 * kPi;
            cs.N_dir = 4 * np;
        }
        cs.N_face = 3;
    } else if (cs.dim == 3) {
        cs.N_cell = (int)cs.tet.size();
        if (DM == 3) {
            cs.WFACTOR = 1.0;
            cs.N_dir = 8 * nt * np;
        } else if (DM == 2) {
            // :84-87 sets N_dir = 4*nt*np but _preprocess builds only 4*np directions (quirk Q9):
            // the reference indexes out of bounds.  Rejected here.
            cs.err = "DM=2 on a 3D mesh is unsupported";
            return false;
        }
        cs.N_face = 4;
    }
    if (DM != 2 && DM != 3) {
        cs.err = "DM must be 2 or 3";
        return false;
    }
    cs.N_band = (int)cs.vg.size();
    cs.Tband.assign((size_t)cs.N_cell * cs.N_band, 0.0);
    cs.dens.assign((size_t)cs.N_cell * cs.N_band, 0.0);
    cs.T.assign(cs.N_cell, 0.0);
    cs.params_set = true;
    return true;
}

// utility.cpp:69-113 GaussIntegrationPoints(a, b, n): Gauss-Legendre nodes (ascending) and
// weights on [a, b].  Restated as Newton iteration on P_n from the Chebyshev initial guess.
// Parity: numpy leggauss (test) and the reference's own utility.cpp built under oracle/_ref.
std::vector<double> gauss_legendre(double a, double b, int n) {
    std::vector<double> out;
    for (int i = 0; i < n; i++) {
        // ascending: i=0 is the node near -1
        double y = std::cos((2.0 * (n - 1 - i) + 1) * kPi / (2.0 * n));
        double dp = 1;
        for (int it = 0; it < 100; it++) {
            double p0 = 1, p1 = y;
            for (int k = 2; k <= n; k++) {
                double p2 = ((2 * k - 1) * y * p1 - (k - 1) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            double pn = (n == 1) ? y : p1;
            double pnm1 = (n == 1) ? 1 : p0;
            dp = n * (pnm1 - y * pn) / (1 - y * y);
            double dy = pn / dp;
            y -= dy;
            if (std::fabs(dy) <= 1e-16) break;
        }
        // final derivative at the converged node
        double p0 = 1, p1 = y;
        for (int k = 2; k <= n; k++) {
            double p2 = ((2 * k - 1) * y * p1 - (k - 1) * p0) / k;
            p0 = p1;
            p1 = p2;
        }
        double pn = (n == 1) ? y : p1;
        double pnm1 = (n == 1) ? 1 : p0;
        dp = n * (pnm1 - y * pn) / (1 - y * y);
        out.push_back((a * (1 - y) + b * (1 + y)) / 2);
        out.push_back((b - a) / ((1 - y * y) * dp * dp));
    }
    return out;
}

// _preprocess angular part, StaticBTESolver.cpp:374-492.
void make_quadrature(Case& cs) {
    cs.omega.clear();
    cs.dirv.clear();
    cs.S.clear();
    if (cs.DM == 2) {
        if (cs.dim >= 2) {  // :375-395
            double dphi = 0.5 * kPi / cs.num_phi;
            int n = 4 * cs.num_phi;
            for (int k = 0; k < n; k++) {
                double phi = (k + 0.5) * dphi;
                double w = cs.WFACTOR / n;
                cs.omega.push_back(w);
                cs.S.push_back({w * std::sin(phi), w * std::cos(phi), 0});
                cs.dirv.push_back({std::sin(phi), std::cos(phi), 0});
            }
        } else {  // :397-417, 1D: Gauss-Legendre in theta on [0, pi], weights doubled
            std::vector<double> g = gauss_legendre(0, kPi, cs.N_dir);
            cs.WFACTOR = 0;
            for (int d = 0; d < cs.N_dir; d++) {
                double ct = std::cos(g[2 * d]);
                double st = std::pow(1 - ct * ct, 0.5);
                cs.omega.push_back(g[2 * d + 1] * 2);
                cs.dirv.push_back({ct, st, 0});
                cs.WFACTOR += g[2 * d + 1] * 2;
            }
            cs.S.assign(cs.N_dir, V3{});
        }
    } else {
        if (cs.dim >= 2) {  // :420-461 midpoint rule in theta and phi, closed-form S
            double dth = 0.5 * kPi / cs.num_theta;
            double dph = 0.5 * kPi / cs.num_phi;
            int ntheta = (cs.dim == 2) ? cs.num_theta : 2 * cs.num_theta;
            int nphi = 4 * cs.num_phi;
            cs.omega.assign((size_t)ntheta * nphi, 0);
            cs.S.assign((size_t)ntheta * nphi, V3{});
            cs.dirv.assign((size_t)ntheta * nphi, V3{});
            for (int it = 0; it < ntheta; it++) {
                double th = (it + 0.5) * dth;
                for (int ip = 0; ip < nphi; ip++) {
                    double ph = (ip + 0.5) * dph;
                    int k = ip + it * nphi;
                    cs.omega[k] = cs.WFACTOR * 2 * std::sin(th) * std::sin(0.5 * dth) * dph;
                    double radial = std::sin(0.5 * dph) * (dth - std::cos(2 * th) * std::sin(dth));
                    cs.S[k] = {cs.WFACTOR * std::sin(ph) * radial, cs.WFACTOR * std::cos(ph) * radial,
                               cs.WFACTOR * 0.5 * dph * std::sin(2 * th) * std::sin(dth)};
                    cs.dirv[k] = {std::sin(th) * std::sin(ph), std::sin(th) * std::cos(ph), std::cos(th)};
                }
            }
        } else {  // :463-491 1D with DM=3: Gauss in cos(theta) on [-1,1], Gauss on [-pi,pi] / pi
            std::vector<double> g = gauss_legendre(-1, 1, cs.num_theta);
            std::vector<double> gp = gauss_legendre(-kPi, kPi, cs.num_phi);
            cs.omega.assign(cs.N_dir, 0);
            cs.WFACTOR = 0;
            for (int i = 0; i < cs.num_theta; i++) {
                double ct = g[2 * i];
                double st = std::pow(1 - ct * ct, 0.5);
                for (int j = 0; j < cs.num_phi; j++) {
                    double cp = gp[2 * j] / kPi;  // the reference divides the node by pi
                    double sp = std::pow(1 - cp * cp, 0.5);
                    cs.omega[i * cs.num_phi + j] = g[2 * i + 1] * gp[2 * j + 1];
                    cs.dirv.push_back({ct, st * cp, st * sp});
                }
            }
            for (int d = 0; d < cs.N_dir; d++) cs.WFACTOR += cs.omega[d];
            cs.S.assign(cs.N_dir, V3{});
        }
    }
}

// _preprocess geometry part, StaticBTESolver.cpp:494-687; getArea/getVolume utility.cpp:27-57.
// Parity: geometric invariants in tests; neighbour search follows the reference's "first match"
// order (vertex map in cell order, then boundary elements in file order) and keeps quirk Q4
// (an unmatched face keeps neighbour 0).
bool make_geometry(Case& cs) {
    const int N = cs.N_cell, F = cs.N_face;
    cs.vol.assign(N, 0);
    cs.centre.assign(N, V3{});
    cs.farea.assign((size_t)N * F, 0);
    cs.fnorm.assign((size_t)N * F, V3{});
    cs.nbr.assign((size_t)N * F, 0);
    std::vector<std::vector<int>> ptmap;
    if (cs.dim == 1) {
        for (int c = 0; c < N; c++) {
            cs.centre[c] = {cs.Lx / N * (c + 0.5), 0, 0};
            cs.vol[c] = cs.Lx / N;
            cs.farea[c * 2 + 0] = 1;
            cs.farea[c * 2 + 1] = 1;
            cs.fnorm[c * 2 + 0] = {-1, 0, 0};
            cs.fnorm[c * 2 + 1] = {1, 0, 0};
            cs.nbr[c * 2 + 0] = c - 1;
            cs.nbr[c * 2 + 1] = c + 1;
        }
        cs.nbr[0] = -0 - 1;               // :609 left end -> BC 0
        cs.nbr[(N - 1) * 2 + 1] = -1 - 1;  // :610 right end -> BC 1
        return true;
    }
    ptmap.resize(cs.pts.size());
    if (cs.dim == 2) {
        for (int c = 0; c < N; c++) {
            const V3 &p1 = cs.pts[cs.tri[c][0]], &p2 = cs.pts[cs.tri[c][1]], &p3 = cs.pts[cs.tri[c][2]];
            cs.centre[c] = {(p1.x + p2.x + p3.x) / 3, (p1.y + p2.y + p3.y) / 3, 0};
            cs.vol[c] = 0.5 * len3(cross3(sub3(p3, p1), sub3(p2, p1)));
            for (int v : cs.tri[c]) ptmap[v].push_back(c);
            for (int e = 0; e < 3; e++) {
                const V3& q0 = cs.pts[cs.tri[c][e % 3]];
                const V3& q1 = cs.pts[cs.tri[c][(e + 1) % 3]];
                const V3& q2 = cs.pts[cs.tri[c][(e + 2) % 3]];
                cs.farea[c * 3 + e] = len3(sub3(q1, q2));
                V3 p01 = sub3(q0, q1);
                V3 n = {q2.y - q1.y, q1.x - q2.x, 0};
                if (dot3(n, p01) > 0) n = {0 - n.x, 0 - n.y, 0 - n.z};
                double l = len3(n);
                cs.fnorm[c * 3 + e] = {n.x / l, n.y / l, n.z / l};
            }
        }
        for (int c = 0; c < N; c++) {
            for (int e = 0; e < 3; e++) {
                int v1 = cs.tri[c][(e + 1) % 3], v2 = cs.tri[c][(e + 2) % 3];
                for (int o : ptmap[v1]) {
                    if (o == c) continue;
                    int hit = 0;
                    for (int i = 0; i < 3; i++)
                        if (cs.tri[o][i] == v1 || cs.tri[o][i] == v2) hit++;
                    if (hit >= 2) {
                        cs.nbr[c * 3 + e] = o;
                        break;
                    }
                }
                if (cs.nbr[c * 3 + e] == 0) {
                    for (size_t s = 0; s < cs.seg.size(); s++) {
                        int hit = 0;
                        for (int i = 0; i < 2; i++)
                            if (cs.seg[s][i] == v1 || cs.seg[s][i] == v2) hit++;
                        if (hit >= 2) {
                            cs.nbr[c * 3 + e] = -cs.seg_ent[s] - 1;
                            break;
                        }
                    }
                }
            }
        }
        return true;
    }
    if (cs.dim == 3) {
        for (int c = 0; c < N; c++) {
            const V3 &p1 = cs.pts[cs.tet[c][0]], &p2 = cs.pts[cs.tet[c][1]], &p3 = cs.pts[cs.tet[c][2]],
                     &p4 = cs.pts[cs.tet[c][3]];
            cs.centre[c] = {(p1.x + p2.x + p3.x + p4.x) / 4, (p1.y + p2.y + p3.y + p4.y) / 4,
                            (p1.z + p2.z + p3.z + p4.z) / 4};
            V3 a = sub3(p1, p4), b = sub3(p2, p4), d = sub3(p3, p4);
            double det = a.x * (b.y * d.z - b.z * d.y) - b.x * (a.y * d.z - a.z * d.y) + d.x * (a.y * b.z - a.z * b.y);
            cs.vol[c] = 1.0 / 6 * std::fabs(det);
            for (int v : cs.tet[c]) ptmap[v].push_back(c);
            for (int f = 0; f < 4; f++) {
                const V3& q0 = cs.pts[cs.tet[c][f % 4]];
                const V3& q1 = cs.pts[cs.tet[c][(f + 1) % 4]];
                const V3& q2 = cs.pts[cs.tet[c][(f + 2) % 4]];
                const V3& q3 = cs.pts[cs.tet[c][(f + 3) % 4]];
                V3 p31 = sub3(q3, q1), p32 = sub3(q3, q2), p01 = sub3(q0, q1);
                cs.farea[c * 4 + f] = 0.5 * len3(cross3(p31, p32));
                V3 n = cross3(p31, p32);
                if (dot3(n, p01) > 0) n = {0 - n.x, 0 - n.y, 0 - n.z};
                double l = len3(n);
                cs.fnorm[c * 4 + f] = {n.x / l, n.y / l, n.z / l};
            }
        }
        for (int c = 0; c < N; c++) {
            for (int f = 0; f < 4; f++) {
                int v1 = cs.tet[c][(f + 1) % 4], v2 = cs.tet[c][(f + 2) % 4], v3 = cs.tet[c][(f + 3) % 4];
                for (int o : ptmap[v1]) {
                    if (o == c) continue;
                    int hit = 0;
                    for (int i = 0; i < 4; i++)
                        if (cs.tet[o][i] == v1 || cs.tet[o][i] == v2 || cs.tet[o][i] == v3) hit++;
                    if (hit >= 3) {
                        cs.nbr[c * 4 + f] = o;
                        break;
                    }
                }
                if (cs.nbr[c * 4 + f] == 0) {
                    for (size_t s = 0; s < cs.tri.size(); s++) {
                        int hit = 0;
                        for (int i = 0; i < 3; i++)
                            if (cs.tri[s][i] == v1 || cs.tri[s][i] == v2 || cs.tri[s][i] == v3) hit++;
                        if (hit >= 3) {
                            cs.nbr[c * 4 + f] = -cs.tri_ent[s] - 1;
                            break;
                        }
                    }
                }
            }
        }
        return true;
    }
    cs.err = "mesh dimension must be 1, 2 or 3";
    return false;
}

bool preprocess(Case& cs) {
    if (cs.preprocessed) return true;
    make_quadrature(cs);
    if (!make_geometry(cs)) return false;
    cs.preprocessed = true;
    return true;
}

// Coefficient used for face f of cell c in direction d: v*tau*A_f*(S_d . n_f), or with the unit
// direction in 1D (StaticBTESolver.cpp:130-136, 166-172).
inline double face_coef(const Case& cs, int b, int d, int c, int f) {
    double t = cs.vg[b] * cs.tau[b] * cs.farea[(size_t)c * cs.N_face + f];
    const V3& n = cs.fnorm[(size_t)c * cs.N_face + f];
    if (cs.dim > 1)
        t *= dot3(cs.S[d], n);
    else
        t *= dot3(cs.dirv[d], n);
    return t;
}

// _get_Ke, StaticBTESolver.cpp:122-160.  Sorted-column CSR of the upwind operator.
void assemble(const Case& cs, int b, int d, Csr& A) {
    const int N = cs.N_cell, F = cs.N_face;
    A.rowptr.assign(N + 1, 0);
    A.col.clear();
    A.val.clear();
    std::vector<std::pair<int, double>> row;
    for (int c = 0; c < N; c++) {
        row.clear();
        double diag = 0;
        for (int f = 0; f < F; f++) {
            double t = face_coef(cs, b, d, c, f);
            if (dot3(cs.dirv[d], cs.fnorm[(size_t)c * F + f]) >= 0)
                diag += t;
            else if (cs.nbr[(size_t)c * F + f] >= 0)
                row.emplace_back(cs.nbr[(size_t)c * F + f], t);
        }
        if (cs.dim > 1)
            diag += cs.vol[c] * cs.omega[d];
        else
            diag += cs.vol[c];
        row.emplace_back(c, diag);
        std::sort(row.begin(), row.end());
        for (auto& p : row) {
            A.col.push_back(p.first);
            A.val.push_back(p.second);
        }
        A.rowptr[c + 1] = (int)A.col.size();
    }
}

// Mirror direction for specular boundary types, StaticBTESolver.cpp:205-252.  Returns -1 for an
// unsupported combination (the reference exits there).
int mirror_dir(const Case& cs, int type, int d) {
    const int p = cs.num_phi;
    if (type == 31) {
        int it = d / (4 * p), ip = d % (4 * p);
        if (cs.dim == 2)
            ip = (ip >= 2 * p) ? 6 * p - ip - 1 : 2 * p - ip - 1;
        else if (cs.dim == 3)
            ip = 4 * p - ip - 1;
        else
            return -1;
        return it * 4 * p + ip;
    }
    if (type == 32) {
        int it = d / (4 * p), ip = d % (4 * p);
        if (cs.dim == 2) return it * 4 * p + (4 * p - ip - 1);
        if (cs.dim == 3) {
            int ix = ip / (2 * p);
            int isx = 2 * p - (ip % (2 * p)) - 1;
            return it * 4 * p + isx + ix * 2 * p;
        }
        return -1;
    }
    if (type == 33) {
        if (cs.dim < 3) return -1;
        int iz = 2 * p - (d / (4 * p)) - 1;  // uses num_phi for the theta index (quirk Q8)
        int isz = d % (4 * p);
        return isz + iz * 4 * p;
    }
    return -1;
}

inline double eprev(const Case& cs, int b, int d, int c) {
    return cs.ee_prev[((size_t)b * cs.N_dir + d) * cs.N_cell + c];
}

// _get_Re, StaticBTESolver.cpp:162-269.
bool rhs(Case& cs, int b, int d, std::vector<double>& Re) {
    const int N = cs.N_cell, F = cs.N_face;
    Re.assign(N, 0.0);
    for (int c = 0; c < N; c++) {
        for (int f = 0; f < F; f++) {
            const V3& n = cs.fnorm[(size_t)c * F + f];
            if (dot3(cs.dirv[d], n) >= 0) continue;
            double a = face_coef(cs, b, d, c, f);
            int nb = cs.nbr[(size_t)c * F + f];
            for (size_t k = 0; k < cs.bc_index.size(); k++) {
                if (nb != cs.bc_index[k]) continue;
                int type = cs.bc_type[k];
                if (type == 1) {
                    Re[c] -= cs.Ctot[b] / cs.solid_angle * (cs.bc_T[k] - cs.T_ref) * a;
                } else if (type == 2) {
                    double es = 0, tw = 0;
                    for (int d2 = 0; d2 < cs.N_dir; d2++) {
                        if (dot3(cs.dirv[d2], n) > 0) {
                            if (cs.dim == 3) {
                                es += eprev(cs, b, d2, c) * dot3(cs.S[d2], n) * cs.omega[d2];
                                tw += dot3(cs.S[d2], n) * cs.omega[d2];
                            } else if (cs.dim == 2) {
                                es += eprev(cs, b, d2, c) * dot3(cs.S[d2], n);
                                tw += dot3(cs.S[d2], n);
                            } else {
                                es += eprev(cs, b, d2, c) * dot3(cs.dirv[d2], n) * cs.omega[d2];
                            }
                        }
                    }
                    if (cs.dim != 1)
                        es = es / tw;
                    else
                        es = es / kPi;
                    Re[c] -= es * a;
                } else if (type == 31 || type == 32 || type == 33) {
                    int m = mirror_dir(cs, type, d);
                    if (m < 0 || m >= cs.N_dir) {
                        std::lock_guard<std::mutex> lk(cs.err_mu);
                        cs.err = "boundary condition type " + std::to_string(type) + " not supported here";
                        return false;
                    }
                    Re[c] -= eprev(cs, b, m, c) * a;
                } else {
                    std::lock_guard<std::mutex> lk(cs.err_mu);
                    cs.err = "boundary condition type " + std::to_string(type) + " not supported";
                    return false;
                }
            }
        }
        if (cs.dim != 1)
            Re[c] += cs.dens[(size_t)c * cs.N_band + b] * cs.vol[c] * cs.omega[d];
        else
            Re[c] += cs.dens[(size_t)c * cs.N_band + b] * cs.vol[c];
    }
    return true;
}

// ---------------------------------------------------------------------------------------------
// Inner solvers
// ---------------------------------------------------------------------------------------------
void spmv(const Csr& A, const double* x, double* y) {
    int n = (int)A.rowptr.size() - 1;
    for (int i = 0; i < n; i++) {
        double s = 0;
        for (int k = A.rowptr[i]; k < A.rowptr[i + 1]; k++) s += A.val[k] * x[A.col[k]];
        y[i] = s;
    }
}

double nrm2(const std::vector<double>& v) {
    double s = 0;
    for (double x : v) s += x * x;
    return std::sqrt(s);
}

double true_rel_resid(const Csr& A, const std::vector<double>& x, const std::vector<double>& b) {
    std::vector<double> r(b.size());
    spmv(A, x.data(), r.data());
    for (size_t i = 0; i < r.size(); i++) r[i] = b[i] - r[i];
    double nb = nrm2(b);
    return nb > 0 ? nrm2(r) / nb : nrm2(r);
}

// Topological order of the dependency graph (row i depends on column j != i).  Returns the order
// of the rows that could be placed; `lagged` marks entries dropped to break cycles.
void sweep_order(const Csr& A, std::vector<int>& order, std::vector<char>& lagged) {
    int n = (int)A.rowptr.size() - 1;
    std::vector<int> indeg(n, 0);
    std::vector<std::vector<int>> out(n);
    lagged.assign(A.col.size(), 0);
    for (int i = 0; i < n; i++)
        for (int k = A.rowptr[i]; k < A.rowptr[i + 1]; k++)
            if (A.col[k] != i) {
                indeg[i]++;
                out[A.col[k]].push_back(i);
            }
    std::vector<char> placed(n, 0);
    order.clear();
    std::vector<int> queue;
    for (int i = 0; i < n; i++)
        if (indeg[i] == 0) queue.push_back(i);
    size_t head = 0;
    while ((int)order.size() < n) {
        if (head == queue.size()) {
            // cycle: lag the remaining incoming edges of the lowest unplaced row
            int pick = -1;
            for (int i = 0; i < n; i++)
                if (!placed[i]) {
                    pick = i;
                    break;
                }
            for (int k = A.rowptr[pick]; k < A.rowptr[pick + 1]; k++) {
                int j = A.col[k];
                if (j != pick && !placed[j]) lagged[k] = 1;
            }
            indeg[pick] = 0;
            queue.push_back(pick);
        }
        int i = queue[head++];
        if (placed[i]) continue;
        placed[i] = 1;
        order.push_back(i);
        for (int o : out[i])
            if (!placed[o] && --indeg[o] == 0) queue.push_back(o);
    }
}

// y = M^-1 v for the sweep preconditioner (forward substitution in topological order, lagged
// entries ignored).
void sweep_apply(const Csr& A, const std::vector<int>& order, const std::vector<char>& lagged,
                 const double* v, double* y) {
    for (int i : order) {
        double s = v[i], dg = 1;
        for (int k = A.rowptr[i]; k < A.rowptr[i + 1]; k++) {
            if (A.col[k] == i)
                dg = A.val[k];
            else if (!lagged[k])
                s -= A.val[k] * y[A.col[k]];
        }
        y[i] = s / dg;
    }
}

// Restarted GMRES(m), classical Gram-Schmidt, Givens rotations.  left=true: preconditioned
// residual norm ||M^-1 r|| is monitored (PETSc default left preconditioning); left=false: right
// preconditioning, the monitored norm is the true residual norm.
template <typename Prec>
int gmres(const Csr& A, const std::vector<double>& b, std::vector<double>& x, int m, double rtol,
          double abstol, int maxits, bool left, Prec prec) {
    const int n = (int)b.size();
    std::vector<double> r(n), w(n), z(n);
    std::vector<std::vector<double>> Vb(m + 1, std::vector<double>(n));
    std::vector<double> H((size_t)(m + 1) * m), cs(m), sn(m), g(m + 1);
    auto residual = [&](std::vector<double>& out) {
        spmv(A, x.data(), w.data());
        for (int i = 0; i < n; i++) w[i] = b[i] - w[i];
        if (left)
            prec(w.data(), out.data());
        else
            out = w;
    };
    int its = 0;
    double rnorm0 = -1;
    std::vector<double> rr(n);
    while (its < maxits) {
        residual(rr);
        double beta = nrm2(rr);
        if (rnorm0 < 0) rnorm0 = beta;
        double ttol = std::max(rtol * rnorm0, abstol);
        if (beta <= ttol) break;
        for (int i = 0; i < n; i++) Vb[0][i] = rr[i] / beta;
        std::fill(g.begin(), g.end(), 0.0);
        g[0] = beta;
        int j = 0;
        bool done = false;
        for (; j < m && its < maxits; j++) {
            its++;
            if (left) {
                spmv(A, Vb[j].data(), w.data());
                prec(w.data(), z.data());
            } else {
                prec(Vb[j].data(), w.data());
                spmv(A, w.data(), z.data());
            }
            // classical Gram-Schmidt: all projections against the unmodified vector
            for (int i = 0; i <= j; i++) {
                double h = 0;
                for (int k = 0; k < n; k++) h += z[k] * Vb[i][k];
                H[(size_t)i * m + j] = h;
            }
            for (int i = 0; i <= j; i++) {
                double h = H[(size_t)i * m + j];
                for (int k = 0; k < n; k++) z[k] -= h * Vb[i][k];
            }
            double hn = nrm2(z);
            H[(size_t)(j + 1) * m + j] = hn;
            for (int i = 0; i < j; i++) {
                double a = H[(size_t)i * m + j], bb = H[(size_t)(i + 1) * m + j];
                H[(size_t)i * m + j] = cs[i] * a + sn[i] * bb;
                H[(size_t)(i + 1) * m + j] = -sn[i] * a + cs[i] * bb;
            }
            double a = H[(size_t)j * m + j], bb = H[(size_t)(j + 1) * m + j];
            double den = std::sqrt(a * a + bb * bb);
            cs[j] = den > 0 ? a / den : 1;
            sn[j] = den > 0 ? bb / den : 0;
            H[(size_t)j * m + j] = den;
            H[(size_t)(j + 1) * m + j] = 0;
            g[j + 1] = -sn[j] * g[j];
            g[j] = cs[j] * g[j];
            bool happy = hn <= 1e-30 * beta;
            if (!happy)
                for (int k = 0; k < n; k++) Vb[j + 1][k] = z[k] / hn;
            if (std::fabs(g[j + 1]) <= ttol || happy) {
                j++;
                done = true;
                break;
            }
        }
        // back substitution and update
        std::vector<double> yv(j, 0.0);
        for (int i = j - 1; i >= 0; i--) {
            double s = g[i];
            for (int k = i + 1; k < j; k++) s -= H[(size_t)i * m + k] * yv[k];
            yv[i] = s / H[(size_t)i * m + i];
        }
        std::fill(w.begin(), w.end(), 0.0);
        for (int i = 0; i < j; i++)
            for (int k = 0; k < n; k++) w[k] += yv[i] * Vb[i][k];
        if (left) {
            for (int k = 0; k < n; k++) x[k] += w[k];
        } else {
            prec(w.data(), z.data());
            for (int k = 0; k < n; k++) x[k] += z[k];
        }
        if (done) break;
    }
    return its;
}

// ILU(0) in natural order on the CSR pattern (PETSc PCILU default levels 0, natural ordering).
struct Ilu0 {
    const Csr* A;
    std::vector<double> lu;
    std::vector<int> diag;
    void factor(const Csr& M) {
        A = &M;
        int n = (int)M.rowptr.size() - 1;
        lu = M.val;
        diag.assign(n, -1);
        for (int i = 0; i < n; i++)
            for (int k = M.rowptr[i]; k < M.rowptr[i + 1]; k++)
                if (M.col[k] == i) diag[i] = k;
        std::vector<int> pos(n, -1);
        for (int i = 0; i < n; i++) {
            for (int k = M.rowptr[i]; k < M.rowptr[i + 1]; k++) pos[M.col[k]] = k;
            for (int k = M.rowptr[i]; k < M.rowptr[i + 1]; k++) {
                int j = M.col[k];
                if (j >= i) break;
                lu[k] /= lu[diag[j]];
                for (int kk = diag[j] + 1; kk < M.rowptr[j + 1]; kk++) {
                    int jj = M.col[kk];
                    if (pos[jj] >= 0) lu[pos[jj]] -= lu[k] * lu[kk];
                }
            }
            for (int k = M.rowptr[i]; k < M.rowptr[i + 1]; k++) pos[M.col[k]] = -1;
        }
    }
    void apply(const double* v, double* y) const {
        const Csr& M = *A;
        int n = (int)M.rowptr.size() - 1;
        for (int i = 0; i < n; i++) {
            double s = v[i];
            for (int k = M.rowptr[i]; k < diag[i]; k++) s -= lu[k] * y[M.col[k]];
            y[i] = s;
        }
        for (int i = n - 1; i >= 0; i--) {
            double s = y[i];
            for (int k = diag[i] + 1; k < M.rowptr[i + 1]; k++) s -= lu[k] * y[M.col[k]];
            y[i] = s / lu[diag[i]];
        }
    }
};

// src/bicgstab.hpp:138-300 restated on the CPU: right-Jacobi BiCGSTAB with the reference's
// first-step form p = r + beta(p - omega q) from rho_prev = alpha = omega = 1, breakdown
// restarts (r_hat = 6.8 b + 0.6), stop on the recursive residual ||r||/||r0|| < tol.
int bicgstab_ref(const Csr& A, const std::vector<double>& b, std::vector<double>& x, double tol, int maxit) {
    const int n = (int)b.size();
    std::vector<double> jac(n, 0.0);
    for (int i = 0; i < n; i++)
        for (int k = A.rowptr[i]; k < A.rowptr[i + 1]; k++)
            if (A.col[k] == i) jac[i] = 1.0 / A.val[k];
    std::vector<double> r(b), p(b), rw(b), q(n, 0.0), ph(n), z(n), t(n);
    x.assign(n, 0.0);
    auto dot = [&](const std::vector<double>& a, const std::vector<double>& c) {
        double s = 0;
        for (int i = 0; i < n; i++) s += a[i] * c[i];
        return s;
    };
    double nrmr0 = std::sqrt(dot(r, r)), nrmr = nrmr0;
    double rho = 1, rhop = 1, omega = 1, alpha = 1;
    int it = 0;
    auto restart = [&](bool copy_b) {
        rho = alpha = omega = 1;
        std::fill(x.begin(), x.end(), 0.0);
        std::fill(q.begin(), q.end(), 0.0);
        std::fill(p.begin(), p.end(), 0.0);
        r = b;
        if (copy_b) rw = b;
        for (int i = 0; i < n; i++) rw[i] = b[i] * 6.8 + 0.6;
    };
    for (int i = 1; i <= maxit; i++) {
        it = i;
        if (nrmr < 1e-50) break;
        rhop = rho;
        rho = dot(rw, r);
        if (rho < 1e-200 && rho > -1e-200) {
            restart(false);
            continue;
        }
        double beta = (rho / rhop) * (alpha / omega);
        for (int k = 0; k < n; k++) p[k] = (p[k] - omega * q[k]) * beta + r[k];
        for (int k = 0; k < n; k++) ph[k] = jac[k] * p[k];
        spmv(A, ph.data(), q.data());
        double tmp = dot(rw, q);
        if (tmp < 1e-200 && tmp > -1e-200) {
            restart(true);
            continue;
        }
        alpha = rho / tmp;
        for (int k = 0; k < n; k++) r[k] -= alpha * q[k];
        for (int k = 0; k < n; k++) z[k] = jac[k] * r[k];
        spmv(A, z.data(), t.data());
        double ts = dot(t, r), tt = dot(t, t);
        omega = ts / tt;
        for (int k = 0; k < n; k++) x[k] += alpha * ph[k];
        for (int k = 0; k < n; k++) x[k] += omega * z[k];
        for (int k = 0; k < n; k++) r[k] -= omega * t[k];
        nrmr = std::sqrt(dot(r, r));
        if (nrmr / nrmr0 < tol) break;
    }
    return it;
}

int inner_solve(Case& cs, const Csr& A, const std::vector<double>& b, std::vector<double>& x, int dir = -1) {
    const int n = (int)b.size();
    x.assign(n, 0.0);
    if (cs.inner_mode == 1) {  // PETSc default restatement
        Ilu0 ilu;
        ilu.factor(A);
        double tol = cs.inner_tol > 0 ? cs.inner_tol : 1e-9;
        int maxit = cs.inner_maxit > 0 ? cs.inner_maxit : 1000;
        return gmres(A, b, x, 30, tol, 1e-50, maxit, true,
                     [&](const double* v, double* y) { ilu.apply(v, y); });
    }
    if (cs.inner_mode == 2) {
        double tol = cs.inner_tol > 0 ? cs.inner_tol : 1e-9;
        int maxit = cs.inner_maxit > 0 ? cs.inner_maxit : 1000;
        return bicgstab_ref(A, b, x, tol, maxit);
    }
    // exact
    std::vector<int> order_tmp;
    std::vector<char> lagged_tmp;
    const std::vector<int>* op = &order_tmp;
    const std::vector<char>* lp = &lagged_tmp;
    if (dir >= 0) {
        if ((int)cs.order_of_dir.size() <= dir) {
            cs.order_of_dir.resize(dir + 1);
            cs.lagged_of_dir.resize(dir + 1);
        }
        if (cs.order_of_dir[dir].empty()) sweep_order(A, cs.order_of_dir[dir], cs.lagged_of_dir[dir]);
        op = &cs.order_of_dir[dir];
        lp = &cs.lagged_of_dir[dir];
    } else {
        sweep_order(A, order_tmp, lagged_tmp);
    }
    const std::vector<int>& order = *op;
    const std::vector<char>& lagged = *lp;
    bool cyclic = false;
    for (char l : lagged) cyclic |= (l != 0);
    if (!cyclic) {
        sweep_apply(A, order, lagged, b.data(), x.data());
        return 1;
    }
    double tol = cs.inner_tol > 0 ? cs.inner_tol : 1e-14;
    return gmres(A, b, x, 50, tol, 0.0, 100000, false,
                 [&](const double* v, double* y) { sweep_apply(A, order, lagged, v, y); });
}

// _get_cell_temperature (:96-104) for all bands, then _recover_temperature (:106-120).
void update_temperature(Case& cs) {
    const int N = cs.N_cell, B = cs.N_band, D = cs.N_dir;
    for (int b = 0; b < B; b++)
        for (int c = 0; c < N; c++) {
            double e0 = 0;
            for (int d = 0; d < D; d++) e0 += cs.ee_curr[((size_t)b * D + d) * N + c] * cs.omega[d];
            cs.Tband[(size_t)c * B + b] = e0 / cs.Ctot[b] + cs.T_ref;
        }
    for (int c = 0; c < N; c++) {
        double rrn = 0, rnn = 0;
        for (int b = 0; b < B; b++) {
            rrn += cs.Lr[b] * cs.Tband[(size_t)c * B + b];
            rnn += cs.Lr[b];
        }
        cs.T[c] = rrn / rnn;
    }
    for (int c = 0; c < N; c++)
        for (int b = 0; b < B; b++)
            cs.dens[(size_t)c * B + b] = cs.Ctot[b] / cs.solid_angle * (cs.T[c] - cs.T_ref);
}

// _get_margin, StaticBTESolver.cpp:311-321.
double margin(const Case& cs) {
    double m = 0;
    for (size_t i = 0; i < cs.ee_curr.size(); i++) m += std::pow(cs.ee_curr[i] - cs.ee_prev[i], 2);
    return m / cs.N_band / cs.N_cell;
}

// _get_heat_flux, StaticBTESolver.cpp:323-370.
void heat_flux(Case& cs) {
    const int N = cs.N_cell, B = cs.N_band, D = cs.N_dir, F = cs.N_face;
    int nbc = (int)cs.bc_index.size();
    cs.band_flux.assign((size_t)nbc * B, 0.0);
    cs.flux.assign(nbc, 0.0);
    for (int k = 0; k < nbc; k++) {
        for (int b = 0; b < B; b++) {
            double len = 0, heat = 0;
            for (int c = 0; c < N; c++)
                for (int f = 0; f < F; f++) {
                    if (cs.nbr[(size_t)c * F + f] != cs.bc_index[k]) continue;
                    const V3& n = cs.fnorm[(size_t)c * F + f];
                    double A = cs.farea[(size_t)c * F + f];
                    for (int d = 0; d < D; d++) {
                        double e = cs.ee_curr[((size_t)b * D + d) * N + c];
                        if (cs.DM == 3 && cs.dim == 3)
                            heat += e * dot3(cs.dirv[d], n) * A * cs.omega[d];
                        else if (cs.dim == 1)
                            heat += e * dot3(cs.dirv[d], n) * cs.omega[d];
                        else
                            heat += e * dot3(cs.S[d], n) * A;
                    }
                    len += A;
                }
            double q = (cs.dim != 1) ? heat / len * cs.vg[b] : heat * cs.vg[b];
            cs.band_flux[(size_t)k * B + b] = q;
            cs.flux[k] += q;
        }
    }
}

// The systems [sys_begin, sys_end) of one outer iteration (StaticBTESolver.cpp:724-780): each
// (band, direction) system assembled (_get_Ke), given its right-hand side (_get_Re) and solved into
// its slice of ee_curr.  With W = n_threads workers, worker w takes the directions d = w, w + W, ...
// of every band -- the reference's split of directions over MPI ranks (:725) -- so the workers
// write disjoint slices and each direction's cached sweep order belongs to one worker.  Every
// system's arithmetic is the serial loop's.  Returns false on a right-hand-side error.
bool solve_systems(Case& cs, int sys_begin, int sys_end, bool check_resid) {
    const int N = cs.N_cell, D = cs.N_dir;
    const int W = std::max(1, std::min(cs.n_threads, D));
    if ((int)cs.order_of_dir.size() < D) {
        cs.order_of_dir.resize(D);
        cs.lagged_of_dir.resize(D);
    }
    std::vector<long long> its(W, 0);
    std::vector<double> worst(W, 0.0);
    std::vector<char> ok(W, 1);
    auto work = [&](int w) {
        Csr A;
        std::vector<double> Re, sol;
        for (int s = sys_begin; s < sys_end; s++) {
            const int b = s / D, d = s % D;
            if (d % W != w) continue;
            assemble(cs, b, d, A);
            if (!rhs(cs, b, d, Re)) {
                ok[w] = 0;
                return;
            }
            its[w] += inner_solve(cs, A, Re, sol, d);
            if (check_resid) worst[w] = std::max(worst[w], true_rel_resid(A, sol, Re));
            std::copy(sol.begin(), sol.end(), cs.ee_curr.begin() + ((size_t)b * D + d) * N);
        }
    };
    if (W == 1) {
        work(0);
    } else {
        std::vector<std::thread> pool;
        for (int w = 0; w < W; w++) pool.emplace_back(work, w);
        for (auto& t : pool) t.join();
    }
    for (int w = 0; w < W; w++) {
        cs.inner_iter_total += its[w];
        cs.worst_true_resid = std::max(cs.worst_true_resid, worst[w]);
        if (!ok[w]) return false;
    }
    return true;
}

// _iteration (:690-808) + _postprocess heat flux (:810-835), all systems in order.
int solve(Case& cs, int max_iter, int verbose) {
    if (!preprocess(cs)) return -1;
    const int N = cs.N_cell, B = cs.N_band, D = cs.N_dir;
    // check BC types referenced by boundary faces up front
    cs.ee_curr.assign((size_t)B * D * N, 0.0);
    std::fill(cs.dens.begin(), cs.dens.end(), 0.0);
    cs.margins.clear();
    cs.iters = 0;
    cs.inner_iter_total = 0;
    cs.worst_true_resid = 0;
    if (verbose)
        std::printf("N_cell: %d\nN_dir: %d\nN_band: %d\nnum_theta: %d\nnum_phi: %d\n\n", N, D, B, cs.num_theta,
                    cs.num_phi);
    for (int it = 0; it < max_iter; it++) {
        cs.ee_prev = cs.ee_curr;
        std::fill(cs.ee_curr.begin(), cs.ee_curr.end(), 0.0);
        if (!solve_systems(cs, 0, B * D, true)) return -1;
        update_temperature(cs);
        double m = margin(cs);
        cs.margins.push_back(m);
        cs.iters = it + 1;
        if (verbose) std::printf("Iteration #%d\t Margin per band per cell: %g\n", it, m);
        if (m <= cs.stop_margin) break;
    }
    heat_flux(cs);
    if (verbose)
        for (size_t k = 0; k < cs.flux.size(); k++)
            std::printf("heat flux of Boundary Condition #%d: %g\n", -1 - cs.bc_index[k], cs.flux[k]);
    return cs.iters;
}

}  // namespace

// =============================================================================================
// C entry points for the test harness (ctypes) and the CPU-baseline timer.
// =============================================================================================
extern "C" {

typedef Case orc_case;

// geometry: an .mphtxt path, or a decimal integer N for the 1D uniform mesh (BTEMain.cpp:118-122)
orc_case* orc_create(const char* geometry, double Lx, double Ly, double Lz, const char* band_file,
                     const char* bc_file, int DM, int ntheta, int nphi, double T_ref, char* errbuf, int errlen) {
    Case* cs = new Case();
    std::string g(geometry);
    bool ok;
    if (g.find('.') == std::string::npos) {
        make_1d(*cs, std::atoi(g.c_str()), Lx);
        ok = true;
    } else {
        ok = read_mphtxt(*cs, geometry, Lx, Ly, Lz);
    }
    ok = ok && read_bands(*cs, band_file) && read_bcs(*cs, bc_file) && set_param(*cs, DM, ntheta, nphi, T_ref) &&
         preprocess(*cs);
    if (!ok) {
        if (errbuf && errlen > 0) std::snprintf(errbuf, errlen, "%s", cs->err.c_str());
        delete cs;
        return nullptr;
    }
    return cs;
}

void orc_destroy(orc_case* cs) { delete cs; }

const char* orc_error(orc_case* cs) { return cs->err.c_str(); }

// sizes: N_cell, N_dir, N_band, N_face, n_bc, dim, n_pts, n_boundary_elements
void orc_sizes(orc_case* cs, int* out) {
    out[0] = cs->N_cell;
    out[1] = cs->N_dir;
    out[2] = cs->N_band;
    out[3] = cs->N_face;
    out[4] = (int)cs->bc_index.size();
    out[5] = cs->dim;
    out[6] = (int)cs->pts.size();
    out[7] = cs->dim == 3 ? (int)cs->tri.size() : (int)cs->seg.size();
}

void orc_params(orc_case* cs, double* out) {
    out[0] = cs->T_ref;
    out[1] = cs->solid_angle;
    out[2] = cs->WFACTOR;
}

void orc_set_inner(orc_case* cs, int mode, double tol, int maxit) {
    cs->inner_mode = mode;
    cs->inner_tol = tol;
    cs->inner_maxit = maxit;
}

void orc_set_stop(orc_case* cs, double stop_margin) { cs->stop_margin = stop_margin; }

// omega[D], dir[D*3], S[D*3]
void orc_quadrature(orc_case* cs, double* omega, double* dir, double* S) {
    for (int d = 0; d < cs->N_dir; d++) {
        omega[d] = cs->omega[d];
        dir[3 * d] = cs->dirv[d].x;
        dir[3 * d + 1] = cs->dirv[d].y;
        dir[3 * d + 2] = cs->dirv[d].z;
        S[3 * d] = cs->S[d].x;
        S[3 * d + 1] = cs->S[d].y;
        S[3 * d + 2] = cs->S[d].z;
    }
}

// vol[N], centre[N*3], area[N*F], normal[N*F*3], nbr[N*F]
void orc_geometry(orc_case* cs, double* vol, double* centre, double* area, double* normal, int* nbr) {
    size_t NF = (size_t)cs->N_cell * cs->N_face;
    for (int c = 0; c < cs->N_cell; c++) {
        vol[c] = cs->vol[c];
        centre[3 * c] = cs->centre[c].x;
        centre[3 * c + 1] = cs->centre[c].y;
        centre[3 * c + 2] = cs->centre[c].z;
    }
    for (size_t i = 0; i < NF; i++) {
        area[i] = cs->farea[i];
        normal[3 * i] = cs->fnorm[i].x;
        normal[3 * i + 1] = cs->fnorm[i].y;
        normal[3 * i + 2] = cs->fnorm[i].z;
        nbr[i] = cs->nbr[i];
    }
}

// mesh points [n_pts*3] (scaled) and elements of the cell dimension [N_cell*(dim+1)]
void orc_mesh(orc_case* cs, double* pts, int* cells) {
    for (size_t i = 0; i < cs->pts.size(); i++) {
        pts[3 * i] = cs->pts[i].x;
        pts[3 * i + 1] = cs->pts[i].y;
        pts[3 * i + 2] = cs->pts[i].z;
    }
    for (int c = 0; c < cs->N_cell; c++) {
        if (cs->dim == 1)
            for (int k = 0; k < 2; k++) cells[2 * c + k] = cs->seg[c][k];
        else if (cs->dim == 2)
            for (int k = 0; k < 3; k++) cells[3 * c + k] = cs->tri[c][k];
        else
            for (int k = 0; k < 4; k++) cells[4 * c + k] = cs->tet[c][k];
    }
}

// CSR of system (b, d); returns nnz.  Buffers must hold N+1 and N*F+1 entries.
int orc_assemble(orc_case* cs, int b, int d, int* rowptr, int* col, double* val) {
    Csr A;
    assemble(*cs, b, d, A);
    std::copy(A.rowptr.begin(), A.rowptr.end(), rowptr);
    std::copy(A.col.begin(), A.col.end(), col);
    std::copy(A.val.begin(), A.val.end(), val);
    return (int)A.col.size();
}

// RHS of system (b, d) for a given ee_prev [B*D*N] and density [N*B]; returns 0 or -1.
int orc_rhs(orc_case* cs, int b, int d, const double* ee_prev, const double* dens, double* Re) {
    size_t tot = (size_t)cs->N_band * cs->N_dir * cs->N_cell;
    cs->ee_prev.assign(ee_prev, ee_prev + tot);
    cs->dens.assign(dens, dens + (size_t)cs->N_cell * cs->N_band);
    std::vector<double> r;
    if (!rhs(*cs, b, d, r)) return -1;
    std::copy(r.begin(), r.end(), Re);
    return 0;
}

// One inner solve of a CSR system with the selected mode; returns iterations.
int orc_inner_solve(int mode, double tol, int maxit, int n, const int* rowptr, const int* col, const double* val,
                    const double* b, double* x) {
    Case cs;
    cs.inner_mode = mode;
    cs.inner_tol = tol;
    cs.inner_maxit = maxit;
    Csr A;
    A.rowptr.assign(rowptr, rowptr + n + 1);
    A.col.assign(col, col + rowptr[n]);
    A.val.assign(val, val + rowptr[n]);
    std::vector<double> bb(b, b + n), xx;
    int it = inner_solve(cs, A, bb, xx);
    std::copy(xx.begin(), xx.end(), x);
    return it;
}

// Levels of the topological sweep order for a CSR (level[i]); returns number of lagged entries.
int orc_sweep_levels(int n, const int* rowptr, const int* col, int* level) {
    Csr A;
    A.rowptr.assign(rowptr, rowptr + n + 1);
    A.col.assign(col, col + rowptr[n]);
    A.val.assign(rowptr[n], 1.0);
    std::vector<int> order;
    std::vector<char> lagged;
    sweep_order(A, order, lagged);
    int nl = 0;
    for (char l : lagged) nl += l;
    for (int i : order) {
        int lv = 0;
        for (int k = A.rowptr[i]; k < A.rowptr[i + 1]; k++)
            if (A.col[k] != i && !lagged[k]) lv = std::max(lv, level[A.col[k]] + 1);
        level[i] = lv;
    }
    return nl;
}

int orc_solve(orc_case* cs, int max_iter, int verbose) { return solve(*cs, max_iter, verbose); }

// Workers of the outer iteration's systems (1: the serial loop, as one reference MPI rank).
void orc_set_threads(orc_case* cs, int n) { cs->n_threads = std::max(1, n); }

// Runs `n_outer` outer iterations on a subset of systems [sys_begin, sys_end) of the first
// iteration only (timing sample for the CPU baseline).  Returns elapsed seconds.
double orc_time_systems(orc_case* cs, int sys_begin, int sys_end) {
    const int N = cs->N_cell, B = cs->N_band, D = cs->N_dir;
    cs->ee_prev.assign((size_t)B * D * N, 0.0);
    cs->ee_curr.assign((size_t)B * D * N, 0.0);
    auto t0 = std::chrono::steady_clock::now();
    if (!solve_systems(*cs, sys_begin, sys_end, false)) return -1;
    auto t1 = std::chrono::steady_clock::now();
    return std::chrono::duration<double>(t1 - t0).count();
}

// results: T[N], Tband[N*B] (cell-major as the reference), flux[n_bc], band_flux[n_bc*B],
// margins[iters]
void orc_results(orc_case* cs, double* T, double* Tband, double* flux, double* band_flux, double* margins) {
    if (T) std::copy(cs->T.begin(), cs->T.end(), T);
    if (Tband) std::copy(cs->Tband.begin(), cs->Tband.end(), Tband);
    if (flux) std::copy(cs->flux.begin(), cs->flux.end(), flux);
    if (band_flux) std::copy(cs->band_flux.begin(), cs->band_flux.end(), band_flux);
    if (margins) std::copy(cs->margins.begin(), cs->margins.end(), margins);
}

// final intensities [B*D*N]
void orc_intensity(orc_case* cs, double* ee) { std::copy(cs->ee_curr.begin(), cs->ee_curr.end(), ee); }

// Stepping form of solve() for studies of outer-iteration corrections: orc_begin resets the state,
// orc_step runs n outer iterations (returns the last margin), orc_get_T / orc_set_T read / replace
// the cell temperature and the emission density derived from it (StaticBTESolver.cpp:115-119).
int orc_begin(orc_case* cs) {
    if (!preprocess(*cs)) return -1;
    cs->ee_curr.assign((size_t)cs->N_band * cs->N_dir * cs->N_cell, 0.0);
    std::fill(cs->dens.begin(), cs->dens.end(), 0.0);
    std::fill(cs->T.begin(), cs->T.end(), 0.0);
    cs->margins.clear();
    cs->iters = 0;
    return 0;
}

double orc_step(orc_case* cs, int n) {
    const int B = cs->N_band, D = cs->N_dir;
    double m = 0;
    for (int it = 0; it < n; it++) {
        cs->ee_prev = cs->ee_curr;
        if (!solve_systems(*cs, 0, B * D, false)) return -1;
        update_temperature(*cs);
        m = margin(*cs);
        cs->margins.push_back(m);
        cs->iters++;
    }
    return m;
}

void orc_get_T(orc_case* cs, double* T) { std::copy(cs->T.begin(), cs->T.end(), T); }

void orc_set_T(orc_case* cs, const double* T) {
    const int N = cs->N_cell, B = cs->N_band;
    for (int c = 0; c < N; c++) {
        cs->T[c] = T[c];
        for (int b = 0; b < B; b++) cs->dens[(size_t)c * B + b] = cs->Ctot[b] / cs->solid_angle * (T[c] - cs->T_ref);
    }
}

int orc_iterations(orc_case* cs) { return cs->iters; }
double orc_worst_resid(orc_case* cs) { return cs->worst_true_resid; }
long long orc_inner_iters(orc_case* cs) { return cs->inner_iter_total; }

// Gauss-Legendre restatement exposed for pinning: out[2*i] node, out[2*i+1] weight.
void orc_gauss(double a, double b, int n, double* out) {
    std::vector<double> g = gauss_legendre(a, b, n);
    std::copy(g.begin(), g.end(), out);
}

}  // extern "C"

#ifdef ORC_MAIN
// CLI with BTEcmd's flags for timing: bte_oracle_cli -g G -m band -b bc -d DM -t T -p P -T Tref
//   -I maxIter -x Lx -y Ly -z Lz [-M inner_mode] [-S stop]
int main(int argc, char** argv) {
    std::string g, m, bcf;
    double Lx = 0, Ly = 0, Lz = 0, Tref = 300, stop = 1e-5;
    int DM = 3, nt = 4, np = 4, maxIter = 10000, mode = 1;
    for (int i = 1; i + 1 < argc; i += 2) {
        std::string k = argv[i];
        const char* v = argv[i + 1];
        if (k == "-g") g = v;
        else if (k == "-m") m = v;
        else if (k == "-b") bcf = v;
        else if (k == "-d") DM = std::atoi(v);
        else if (k == "-t") nt = std::atoi(v);
        else if (k == "-p") np = std::atoi(v);
        else if (k == "-T") Tref = std::atof(v);
        else if (k == "-I") maxIter = std::atoi(v);
        else if (k == "-x") Lx = std::atof(v);
        else if (k == "-y") Ly = std::atof(v);
        else if (k == "-z") Lz = std::atof(v);
        else if (k == "-M") mode = std::atoi(v);
        else if (k == "-S") stop = std::atof(v);
    }
    char err[256] = {0};
    orc_case* cs = orc_create(g.c_str(), Lx, Ly, Lz, m.c_str(), bcf.c_str(), DM, nt, np, Tref, err, 256);
    if (!cs) {
        std::printf("error: %s\n", err);
        return 1;
    }
    orc_set_inner(cs, mode, -1, -1);
    orc_set_stop(cs, stop);
    auto t0 = std::chrono::steady_clock::now();
    int it = orc_solve(cs, maxIter, 1);
    auto t1 = std::chrono::steady_clock::now();
    if (it < 0) {
        std::printf("error: %s\n", cs->err.c_str());
        return 1;
    }
    std::printf("Time taken by iteration: %g milliseconds\n", std::chrono::duration<double>(t1 - t0).count() * 1e3);
    orc_destroy(cs);
    return 0;
}
#endif

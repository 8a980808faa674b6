// BTEMain.cpp -- BTEcmd, the command-line tool (flags and defaults of the reference's
// src/BTEMain.cpp:27-110): -g geometry (.mphtxt or N for 1D), -m band file, -b boundary file,
// -d DM, -t ntheta, -p nphi, -T T_ref, -I maxIter, -x/-y/-z length scales.  Extra flags of this
// build: --precision fp64|mixed, --precond sweep|jacobi, --stop <margin>, --device <n>,
// --temperature <file> (writes x/Lx y/Ly [z/Lz] T per cell).
#include <getopt.h>

#include <chrono>
#include <cstring>
#include <iostream>
#include <memory>
#include <stdexcept>
#include <string>

#include "StaticBTESolver/StaticBTESolver.h"

using namespace std;

int main(int argc, char** argv) {
    static struct option longopts[] = {{"geometry", required_argument, nullptr, 'g'},
                                       {"material", required_argument, nullptr, 'm'},
                                       {"boundary", required_argument, nullptr, 'b'},
                                       {"DM", required_argument, nullptr, 'd'},
                                       {"ntheta", required_argument, nullptr, 't'},
                                       {"nphi", required_argument, nullptr, 'p'},
                                       {"T_ref", required_argument, nullptr, 'T'},
                                       {"maxIter", required_argument, nullptr, 'I'},
                                       {"L_x", required_argument, nullptr, 'x'},
                                       {"L_y", required_argument, nullptr, 'y'},
                                       {"L_z", required_argument, nullptr, 'z'},
                                       {"precision", required_argument, nullptr, 1},
                                       {"precond", required_argument, nullptr, 2},
                                       {"stop", required_argument, nullptr, 3},
                                       {"device", required_argument, nullptr, 4},
                                       {"temperature", required_argument, nullptr, 5},
                                       {nullptr, 0, nullptr, 0}};
    string geo, band, bc, tfile;
    double Lx = 0, Ly = 0, Lz = 0, T_ref = 300;
    int DM = 3, ntheta = 4, nphi = 4, maxIter = 10000, device = -1;
    bool bPresent = false;
    bte_opts opts;
    bte_default_opts(&opts);
    int c, idx = 0;
    try {
        while ((c = getopt_long(argc, argv, "g:m:b:d:t:p:T:I:x:y:z:", longopts, &idx)) != -1) {
            switch (c) {
                case 'g': geo = optarg; break;
                case 'm': band = optarg; break;
                case 'b': bc = optarg; bPresent = true; break;
                case 'd': DM = stoi(optarg); break;
                case 't': ntheta = stoi(optarg); break;
                case 'p': nphi = stoi(optarg); break;
                case 'T': T_ref = stod(optarg); break;
                case 'I': maxIter = stoi(optarg); break;
                case 'x': Lx = stod(optarg); break;
                case 'y': Ly = stod(optarg); break;
                case 'z': Lz = stod(optarg); break;
                case 1: opts.precision = strcmp(optarg, "mixed") == 0 ? BTE_PREC_MIXED : BTE_PREC_FP64; break;
                case 2: opts.precond = strcmp(optarg, "jacobi") == 0 ? BTE_PC_JACOBI : BTE_PC_SWEEP; break;
                case 3: opts.stop_margin = stod(optarg); break;
                case 4: device = stoi(optarg); break;
                case 5: tfile = optarg; break;
                default: cerr << "Unknown option " << c << endl; return 1;
            }
        }
    } catch (const std::exception& e) {
        cerr << "bad flag value: " << e.what() << endl;
        return 1;
    }
    auto start = chrono::high_resolution_clock::now();
    std::unique_ptr<BTEMesh> mesh;
    std::unique_ptr<BTEBand> bands;
    std::unique_ptr<BTEBoundaryCondition> bcs;
    try {
        const auto dot = geo.find_last_of('.');
        if (dot == string::npos) {
            mesh = std::make_unique<BTEMesh>(stoi(geo), Lx);
        } else if (geo.substr(dot + 1) == "mphtxt") {
            ifstream f(geo);
            mesh = std::make_unique<BTEMesh>(f, Lx, Ly, Lz);
        } else if (geo.substr(dot + 1) == "geo") {
            cout << "gmsh is not supported" << endl;
            return 1;
        } else {
            cout << "file format not supported" << endl;
            return 1;
        }
        ifstream bfile(band);
        bands = std::make_unique<BTEBand>(bfile);
        if (!bPresent) {
            cerr << "Must specify boundary conditions" << endl;
            return 1;
        }
        ifstream cfile(bc);
        bcs = std::make_unique<BTEBoundaryCondition>(cfile);
    } catch (const std::exception& e) {
        cout << e.what() << endl;
        return 1;
    }
    StaticBTESolver solver(mesh.get(), bcs.get(), bands.get());
    solver.setDevice(device);
    solver.setOptions(opts);
    solver.setParam(DM, ntheta, nphi, T_ref);
    if (solver.status() != BTE_OK) return 1;
    solver.solve(maxIter);
    if (solver.status() != BTE_OK && solver.status() != BTE_ERR_NOCONV) return 1;
    if (!tfile.empty() && !solver.writeTemperature(tfile)) cerr << "cannot write " << tfile << endl;
    auto stop = chrono::high_resolution_clock::now();
    cout << "Time taken by process: " << chrono::duration<double, milli>(stop - start).count() << " milliseconds"
         << endl;
    return 0;
}

// ref_harness.cpp -- test-only C entry points over the reference's own reader and utility
// sources (BTEMesh.cpp, BTEBand.cpp, BTEBoundaryCondition.cpp, utility.cpp), compiled from
// /root/reference into oracle/_ref/.  Used only to cross-check the oracle's restatement of
// those pieces; never shipped, never loaded by the product.
#include <fstream>
#include "StaticBTESolver/BTEBand.h"
#include "StaticBTESolver/BTEBoundaryCondition.h"
#include "StaticBTESolver/BTEMesh.h"
#include "StaticBTESolver/utility.h"

extern "C" {

// GaussIntegrationPoints(a, b, n) -> out[2n]
void ref_gauss(double a, double b, int n, double* out) {
    std::vector<double> g = staticbtesolver::GaussIntegrationPoints(a, b, n);
    for (size_t i = 0; i < g.size(); i++) out[i] = g[i];
}

// counts[0..4] = dim, n_pts, n_seg, n_tri, n_tet; then fills the arrays when non-null.
int ref_read_mesh(const char* path, double Lx, double Ly, double Lz, int* counts, double* pts, int* seg,
                  int* seg_ent, int* tri, int* tri_ent, int* tet) {
    std::ifstream in(path);
    if (!in.is_open()) return -1;
    BTEMesh m(in, Lx, Ly, Lz);
    counts[0] = m.dim;
    counts[1] = (int)m.meshPts.size();
    counts[2] = (int)m.elements1D.size();
    counts[3] = (int)m.elements2D.size();
    counts[4] = (int)m.elements3D.size();
    if (pts)
        for (size_t i = 0; i < m.meshPts.size(); i++) {
            pts[3 * i] = m.meshPts[i]->x;
            pts[3 * i + 1] = m.meshPts[i]->y;
            pts[3 * i + 2] = m.meshPts[i]->z;
        }
    if (seg)
        for (size_t i = 0; i < m.elements1D.size(); i++) {
            seg[2 * i] = m.elements1D[i]->index[0];
            seg[2 * i + 1] = m.elements1D[i]->index[1];
            seg_ent[i] = m.elements1D[i]->entity_index;
        }
    if (tri)
        for (size_t i = 0; i < m.elements2D.size(); i++) {
            for (int k = 0; k < 3; k++) tri[3 * i + k] = m.elements2D[i]->index[k];
            tri_ent[i] = m.elements2D[i]->entity_index;
        }
    if (tet)
        for (size_t i = 0; i < m.elements3D.size(); i++)
            for (int k = 0; k < 4; k++) tet[4 * i + k] = m.elements3D[i]->index[k];
    return 0;
}

// bands -> out[4*nb]; returns nb
int ref_read_bands(const char* path, double* out) {
    std::ifstream in(path);
    if (!in.is_open()) return -1;
    BTEBand b(in);
    int n = (int)b.size();
    for (int i = 0; i < n; i++) {
        out[4 * i] = b[i].group_velocity;
        out[4 * i + 1] = b[i].relaxation_time;
        out[4 * i + 2] = b[i].Ctot;
        out[4 * i + 3] = b[i].Lr;
    }
    return n;
}

// bcs -> index[n], type[n], T[n]; returns n
int ref_read_bcs(const char* path, int* index, int* type, double* T) {
    std::ifstream in(path);
    if (!in.is_open()) return -1;
    BTEBoundaryCondition b(in);
    int n = (int)b.size();
    for (int i = 0; i < n; i++) {
        index[i] = b[i].index;
        type[i] = b[i].type;
        T[i] = b[i].temperature;
    }
    return n;
}

double ref_tri_area(const double* p) {
    staticbtesolver::Point a(p[0], p[1], p[2]), b(p[3], p[4], p[5]), c(p[6], p[7], p[8]);
    return staticbtesolver::getArea(a, b, c);
}

double ref_tet_volume(const double* p) {
    staticbtesolver::Point a(p[0], p[1], p[2]), b(p[3], p[4], p[5]), c(p[6], p[7], p[8]), d(p[9], p[10], p[11]);
    return staticbtesolver::getVolume(a, b, c, d);
}
}

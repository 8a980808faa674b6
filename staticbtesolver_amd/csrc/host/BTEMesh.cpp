// BTEMesh.cpp -- COMSOL .mphtxt reader and 1D mesh of the C++ API.
//
// Accepts the files the reference reader accepts (BTEMesh.cpp:7-130): the "4 Mesh # class"
// object, its sdim and point count, "# Mesh point coordinates", and the edg / tri / tet element
// types with their "# Elements" and "# Geometric entity indices" blocks.  Sections are located by
// their header comments rather than by fixed line offsets; trailing '\r' is tolerated.
#include "StaticBTESolver/BTEMesh.h"

#include <sstream>
#include <stdexcept>
#include <string>

using staticbtesolver::Point;

namespace {

struct LineReader {
    std::istream& in;
    std::string line;
    bool next() {
        if (!std::getline(in, line)) return false;
        if (!line.empty() && line.back() == '\r') line.pop_back();
        return true;
    }
    // content before the '#' comment, and the comment text
    static std::string comment(const std::string& l) {
        auto p = l.find('#');
        return p == std::string::npos ? std::string() : l.substr(p + 1);
    }
    static long long lead_int(const std::string& l) { return std::atoll(l.c_str()); }
};

bool contains(const std::string& s, const char* t) { return s.find(t) != std::string::npos; }

}  // namespace

BTEMesh::BTEMesh(std::ifstream& inFile, double Lx, double Ly, double Lz) {
    if (!inFile.is_open()) throw std::runtime_error("DEBUG: file not open");
    L_x = Lx;
    L_y = Ly;
    L_z = Lz;
    LineReader r{inFile, {}};
    bool found = false;
    while (r.next())
        if (r.line == "4 Mesh # class") {
            found = true;
            break;
        }
    if (!found) throw std::runtime_error("mphtxt: no '4 Mesh # class' object");
    long long npts = -1;
    dim = 0;
    while (r.next()) {
        const std::string c = LineReader::comment(r.line);
        if (contains(c, "sdim")) dim = (int)LineReader::lead_int(r.line);
        if (contains(c, "number of mesh points")) npts = LineReader::lead_int(r.line);
        if (contains(r.line, "# Mesh point coordinates")) break;
    }
    if (dim < 1 || dim > 3 || npts < 0) throw std::runtime_error("mphtxt: missing sdim / point count");
    meshPts.reserve(npts);
    for (long long i = 0; i < npts; i++) {
        double x = 0, y = 0, z = 0;
        inFile >> x;
        if (dim >= 2) inFile >> y;
        if (dim == 3) inFile >> z;
        meshPts.push_back(std::make_shared<Point>(x * Lx, dim >= 2 ? y * Ly : 0.0, dim == 3 ? z * Lz : 0.0));
    }
    if (!inFile) throw std::runtime_error("mphtxt: truncated point list");
    // element types
    std::string type;
    while (r.next()) {
        const std::string c = LineReader::comment(r.line);
        if (contains(c, "type name")) {
            std::istringstream ss(r.line);
            int len = 0;
            ss >> len >> type;
            continue;
        }
        if (!contains(c, "number of elements")) continue;
        const long long n = LineReader::lead_int(r.line);
        int nodes = type == "edg" ? 2 : type == "tri" ? 3 : type == "tet" ? 4 : 0;
        bool want = (type == "edg") || (type == "tri" && dim >= 2) || (type == "tet" && dim == 3);
        while (r.next() && !contains(r.line, "# Elements")) {
        }
        std::vector<int> idx((size_t)n * std::max(nodes, 1));
        if (nodes == 0) {
            for (long long i = 0; i < n; i++) {
                int v;
                inFile >> v;
            }
        } else {
            for (auto& v : idx) inFile >> v;
        }
        if (!inFile) throw std::runtime_error("mphtxt: truncated element list");
        // geometric entity indices follow for every type (not for tets in some writers)
        std::vector<int> ent;
        std::streampos pos = inFile.tellg();
        bool have_ent = false;
        while (r.next()) {
            const std::string c2 = LineReader::comment(r.line);
            if (contains(c2, "number of geometric entity indices")) {
                const long long ne = LineReader::lead_int(r.line);
                while (r.next() && !contains(r.line, "# Geometric entity indices")) {
                }
                ent.resize(ne);
                for (auto& e : ent) inFile >> e;
                have_ent = true;
                break;
            }
            if (contains(c2, "type name") || contains(c2, "number of elements")) {
                inFile.clear();
                inFile.seekg(pos);
                break;
            }
        }
        if (!want) continue;
        if (type == "edg") {
            for (long long i = 0; i < n; i++)
                elements1D.push_back(std::make_shared<staticbtesolver::Segment>(
                    idx[2 * i], idx[2 * i + 1], have_ent && i < (long long)ent.size() ? ent[i] : 0));
        } else if (type == "tri") {
            // the reference keeps triangle entity indices only for 3D meshes (BTEMesh.cpp:105-110)
            for (long long i = 0; i < n; i++)
                elements2D.push_back(std::make_shared<staticbtesolver::Triangle>(
                    idx[3 * i], idx[3 * i + 1], idx[3 * i + 2],
                    (dim == 3 && have_ent && i < (long long)ent.size()) ? ent[i] : 0));
        } else if (type == "tet") {
            for (long long i = 0; i < n; i++)
                elements3D.push_back(std::make_shared<staticbtesolver::Tetrahedron>(idx[4 * i], idx[4 * i + 1],
                                                                                    idx[4 * i + 2], idx[4 * i + 3]));
        }
    }
}

BTEMesh::BTEMesh(int N_cell, double Lx) {
    L_x = Lx;
    L_y = 0;
    L_z = 0;
    dim = 1;
    meshPts.reserve(N_cell + 1);
    for (int i = 0; i <= N_cell; i++) meshPts.push_back(std::make_shared<Point>(Lx / N_cell * i));
    elements1D.reserve(N_cell);
    for (int i = 0; i < N_cell; i++) elements1D.push_back(std::make_shared<staticbtesolver::Segment>(i, i + 1));
}

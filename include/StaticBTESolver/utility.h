// utility.h -- value types of the StaticBTESolver C++ API (same names and members as the
// reference's includes/StaticBTESolver/utility.h:17-336, so code written against the reference
// compiles unchanged).
#ifndef STATICBTE_AMD_UTILITY_H
#define STATICBTE_AMD_UTILITY_H

#include <cmath>
#include <cstddef>
#include <iostream>
#include <memory>
#include <vector>

#ifndef PI
#define PI M_PI
#endif

namespace staticbtesolver {

// Point with the reference's operators (utility.h:24-96).
struct Point {
    double x, y, z;
    explicit Point(double xx = 0, double yy = 0, double zz = 0) : x{xx}, y{yy}, z{zz} {}
    Point(const Point&) = default;
    Point& operator=(const Point&) = default;
    Point operator+(const Point& o) const { return Point(x + o.x, y + o.y, z + o.z); }
    Point operator-(const Point& o) const { return Point(x - o.x, y - o.y, z - o.z); }
    Point operator/(double s) const { return Point(x / s, y / s, z / s); }
    double operator*(const Point& o) const { return x * o.x + y * o.y + z * o.z; }  // dot product
    double length() const { return std::sqrt(x * x + y * y + z * z); }
    double distance(const Point& o) const { return (*this - o).length(); }
    friend std::ostream& operator<<(std::ostream& os, const Point& p);
};

double dot_prod(const Point& a, const Point& b);
double dot_prod(const std::shared_ptr<Point>& a, const std::shared_ptr<Point>& b);
Point cross_prod(const Point& a, const Point& b);
std::shared_ptr<Point> cross_prod(const std::shared_ptr<Point>& a, const std::shared_ptr<Point>& b);
double getLength(const std::shared_ptr<Point>& a, const std::shared_ptr<Point>& b);
double getArea(const Point& a, const Point& b, const Point& c);
double getArea(const std::shared_ptr<Point>& a, const std::shared_ptr<Point>& b, const std::shared_ptr<Point>& c);
double getVolume(const Point& a, const Point& b, const Point& c, const Point& d);
double getVolume(const std::shared_ptr<Point>& a, const std::shared_ptr<Point>& b, const std::shared_ptr<Point>& c,
                 const std::shared_ptr<Point>& d);

template <typename A>
using vector2D = std::vector<std::vector<A>>;
template <typename A>
using vector3D = std::vector<vector2D<A>>;
template <typename A>
using vector4D = std::vector<vector3D<A>>;

struct Segment {
    int index[2]{};
    int entity_index;
    explicit Segment(int a = 0, int b = 0, int e = 0) : index{a, b}, entity_index(e) {}
};

struct Triangle {
    int index[3]{};
    int entity_index;
    explicit Triangle(int a = 0, int b = 0, int c = 0, int e = 0) : index{a, b, c}, entity_index(e) {}
};

struct Tetrahedron {
    int index[4]{};
    explicit Tetrahedron(int a = 0, int b = 0, int c = 0, int d = 0) : index{a, b, c, d} {}
};

// index = -k-1 for boundary entity k; type 1 isothermal, 2 diffuse, 31/32/33 specular.
struct BoundaryCondition {
    int index;
    int type;
    double temperature;
    explicit BoundaryCondition(int index = 0, int type = 0, int temp = 0)
        : index(index), type(type), temperature(temp) {}
    friend std::ostream& operator<<(std::ostream& os, const BoundaryCondition& bc);
};

struct Band {
    double group_velocity;
    double relaxation_time;
    double Ctot;
    double Lr;
    explicit Band(double vg = 0, double rt = 0, double ct = 0, double lr = 0)
        : group_velocity(vg), relaxation_time(rt), Ctot(ct), Lr(lr) {}
    friend std::ostream& operator<<(std::ostream& os, const Band& b);
};

// Row-major dim1 x dim2 array of doubles (utility.h:270-331).
class ContinuousArray {
public:
    double* data{};
    int dim1{};
    int dim2{};
    ContinuousArray() = default;
    ContinuousArray(int d1, int d2) { init(d1, d2); }
    ContinuousArray(const ContinuousArray& o) { assign(o); }
    ContinuousArray& operator=(const ContinuousArray& o) {
        if (this != &o) assign(o);
        return *this;
    }
    void init(int d1, int d2) {
        delete[] data;
        dim1 = d1;
        dim2 = d2;
        data = new double[(size_t)d1 * d2]();
    }
    double* get_ptr(int i, int j) const { return data + ((size_t)i * dim2 + j); }
    double get(int i, int j) const { return *get_ptr(i, j); }
    void clear() const {
        for (size_t k = 0; k < (size_t)dim1 * dim2; k++) data[k] = 0.0;
    }
    void print() const;
    ~ContinuousArray() { delete[] data; }

private:
    void assign(const ContinuousArray& o) {
        init(o.dim1, o.dim2);
        for (size_t k = 0; k < (size_t)dim1 * dim2; k++) data[k] = o.data[k];
    }
};

// Gauss-Legendre nodes and weights on [a, b], interleaved {x0, w0, x1, w1, ...} ascending in x
// (utility.cpp:69-113).
std::vector<double> GaussIntegrationPoints(double a, double b, int N);

// Resident set size of this process in kB (utility.cpp:125-136).
size_t get_host_memory();

}  // namespace staticbtesolver

#endif

// utility.cpp -- geometry helpers and Gauss-Legendre rule of the C++ API
// (reference src/utility.cpp:10-136 provides the same functions).
#include "StaticBTESolver/utility.h"

#include <cstdio>
#include <cstring>
#include <iomanip>

namespace staticbtesolver {

std::ostream& operator<<(std::ostream& os, const Point& p) {
    os << std::fixed << std::setprecision(18) << "(" << p.x << ", " << p.y << " " << p.z << ")";
    return os;
}

std::ostream& operator<<(std::ostream& os, const BoundaryCondition& bc) {
    return os << "Index: " << bc.index << ", Condition: " << bc.type << ", Temperature: " << bc.temperature;
}

std::ostream& operator<<(std::ostream& os, const Band& b) {
    return os << "Group velocity: " << b.group_velocity << ", relaxation time: " << b.relaxation_time
              << ", Ctot: " << b.Ctot << ", Lr: " << b.Lr;
}

void ContinuousArray::print() const {
    for (int i = 0; i < dim1; i++) {
        for (int j = 0; j < dim2; j++) std::cout << get(i, j) << "  ";
        std::cout << std::endl;
    }
}

double dot_prod(const Point& a, const Point& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
double dot_prod(const std::shared_ptr<Point>& a, const std::shared_ptr<Point>& b) { return dot_prod(*a, *b); }

Point cross_prod(const Point& a, const Point& b) {
    return Point(a.y * b.z - b.y * a.z, b.x * a.z - a.x * b.z, a.x * b.y - b.x * a.y);
}
std::shared_ptr<Point> cross_prod(const std::shared_ptr<Point>& a, const std::shared_ptr<Point>& b) {
    return std::make_shared<Point>(cross_prod(*a, *b));
}

double getLength(const std::shared_ptr<Point>& a, const std::shared_ptr<Point>& b) { return (*a - *b).length(); }

// Half the norm of (c - a) x (b - a).
double getArea(const Point& a, const Point& b, const Point& c) { return 0.5 * cross_prod(c - a, b - a).length(); }
// Half the norm of (c - a) x (c - b).
double getArea(const std::shared_ptr<Point>& a, const std::shared_ptr<Point>& b, const std::shared_ptr<Point>& c) {
    return 0.5 * cross_prod(*c - *a, *c - *b).length();
}

// |det[a-d, b-d, c-d]| / 6, determinant expanded along the first column.
double getVolume(const Point& a, const Point& b, const Point& c, const Point& d) {
    const Point u = a - d, v = b - d, w = c - d;
    const double det = u.x * (v.y * w.z - v.z * w.y) - v.x * (u.y * w.z - u.z * w.y) + w.x * (u.y * v.z - u.z * v.y);
    return 1.0 / 6 * std::fabs(det);
}
double getVolume(const std::shared_ptr<Point>& a, const std::shared_ptr<Point>& b, const std::shared_ptr<Point>& c,
                 const std::shared_ptr<Point>& d) {
    return getVolume(*a, *b, *c, *d);
}

// Legendre P_n and P_{n-1} at x by the three-term recurrence.
static void legendre(int n, double x, double& pn, double& pn1) {
    double a = 1.0, b = x;
    if (n == 0) {
        pn = 1;
        pn1 = 0;
        return;
    }
    for (int k = 2; k <= n; k++) {
        const double c = ((2.0 * k - 1) * x * b - (k - 1.0) * a) / k;
        a = b;
        b = c;
    }
    pn = b;
    pn1 = a;
}

std::vector<double> GaussIntegrationPoints(double a, double b, int N) {
    std::vector<double> out(2 * (size_t)std::max(N, 0));
    for (int i = 0; i < N; i++) {
        // roots ascending; Chebyshev-like start for the i-th root from the left
        double x = -std::cos(M_PI * (i + 0.75) / (N + 0.5));
        double pn = 0, pn1 = 0, dp = 1;
        for (int it = 0; it < 64; it++) {
            legendre(N, x, pn, pn1);
            dp = N * (pn1 - x * pn) / (1 - x * x);
            const double dx = pn / dp;
            x -= dx;
            if (std::fabs(dx) < 1e-16) break;
        }
        legendre(N, x, pn, pn1);
        dp = N * (pn1 - x * pn) / (1 - x * x);
        out[2 * i] = 0.5 * (a * (1 - x) + b * (1 + x));
        out[2 * i + 1] = (b - a) / ((1 - x * x) * dp * dp);
    }
    return out;
}

size_t get_host_memory() {
    FILE* f = std::fopen("/proc/self/status", "r");
    if (!f) return 0;
    char line[256];
    size_t kb = 0;
    while (std::fgets(line, sizeof line, f))
        if (std::strncmp(line, "VmRSS:", 6) == 0) {
            kb = (size_t)std::strtoull(line + 6, nullptr, 10);
            break;
        }
    std::fclose(f);
    return kb;
}

}  // namespace staticbtesolver

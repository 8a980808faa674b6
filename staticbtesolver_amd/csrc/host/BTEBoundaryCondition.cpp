// BTEBoundaryCondition.cpp -- boundary condition file reader of the C++ API (format of the
// reference's BTEBoundaryCondition.cpp:7-23).
#include "StaticBTESolver/BTEBoundaryCondition.h"

#include <stdexcept>
#include <string>

BTEBoundaryCondition::BTEBoundaryCondition(std::ifstream& inFile) {
    if (!inFile.is_open()) throw std::runtime_error("DEBUG: file not open");
    int num = 0;
    inFile >> num;
    std::string skip;
    for (int i = 0; i < 3; i++) std::getline(inFile, skip);  // rest of the count line + 2 lines
    for (int i = 0; i < num; i++) {
        int k = 0;
        staticbtesolver::BoundaryCondition bc;
        inFile >> k >> bc.type >> bc.temperature;
        if (!inFile) throw std::runtime_error("boundary condition file: expected " + std::to_string(num) + " rows");
        bc.index = -k - 1;
        boundaryConditions.push_back(bc);
    }
}

BTEBoundaryCondition::const_iterator BTEBoundaryCondition::begin() const { return boundaryConditions.begin(); }
BTEBoundaryCondition::const_iterator BTEBoundaryCondition::end() const { return boundaryConditions.end(); }
std::size_t BTEBoundaryCondition::size() const { return boundaryConditions.size(); }
const staticbtesolver::BoundaryCondition& BTEBoundaryCondition::operator[](std::size_t i) const {
    return boundaryConditions[i];
}

std::ostream& operator<<(std::ostream& os, const BTEBoundaryCondition& bc) {
    for (const auto& b : bc) os << b << std::endl;
    return os;
}

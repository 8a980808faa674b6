// BTEBoundaryCondition.h -- boundary conditions of the StaticBTESolver C++ API (reference
// includes/StaticBTESolver/BTEBoundaryCondition.h:29-54).
#ifndef STATICBTE_AMD_BTEBOUNDARYCONDITION_H
#define STATICBTE_AMD_BTEBOUNDARYCONDITION_H

#include <fstream>
#include <iostream>
#include <vector>

#include "utility.h"

class BTEBoundaryCondition {
    using BoundaryConditionList = std::vector<staticbtesolver::BoundaryCondition>;

public:
    BoundaryConditionList boundaryConditions;
    BTEBoundaryCondition() = default;
    // First the number of boundaries, then two skipped lines, then one "k type T" row per boundary;
    // stored with index = -k-1 (BTEBoundaryCondition.cpp:7-23).  Throws std::runtime_error if the
    // stream is not open.
    explicit BTEBoundaryCondition(std::ifstream& inFile);
    BTEBoundaryCondition(const BTEBoundaryCondition&) = default;
    BTEBoundaryCondition& operator=(const BTEBoundaryCondition&) = default;

    using iterator = BoundaryConditionList::iterator;
    using const_iterator = BoundaryConditionList::const_iterator;
    const_iterator begin() const;
    const_iterator end() const;
    std::size_t size() const;
    const staticbtesolver::BoundaryCondition& operator[](std::size_t i) const;
    friend std::ostream& operator<<(std::ostream& os, const BTEBoundaryCondition& bc);
};

#endif

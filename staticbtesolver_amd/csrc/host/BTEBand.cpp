// BTEBand.cpp -- band file reader of the C++ API (format of the reference's BTEBand.cpp:8-26).
#include "StaticBTESolver/BTEBand.h"

#include <sstream>
#include <stdexcept>
#include <string>

BTEBand::BTEBand(std::ifstream& inFile) {
    if (!inFile.is_open()) throw std::runtime_error("DEBUG: file not open");
    std::string line;
    while (std::getline(inFile, line)) {
        if (!line.empty() && line.back() == '\r') line.pop_back();
        if (line.length() <= 1) break;  // the reference stops at the first (near-)empty line
        std::istringstream ss(line);
        Band b;
        ss >> b.group_velocity >> b.relaxation_time >> b.Ctot >> b.Lr;
        bands.push_back(b);
    }
}

BTEBand::const_iterator BTEBand::begin() const { return bands.begin(); }
BTEBand::const_iterator BTEBand::end() const { return bands.end(); }
std::size_t BTEBand::size() { return bands.size(); }
std::size_t BTEBand::size() const { return bands.size(); }
const Band& BTEBand::operator[](std::size_t i) const { return bands[i]; }

std::ostream& operator<<(std::ostream& os, const BTEBand& band) {
    for (const auto& b : band) os << b << std::endl;
    return os;
}

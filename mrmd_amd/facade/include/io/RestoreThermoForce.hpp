// reference include path io/RestoreThermoForce.hpp of the mrmd action API
// (RestoreThermoForce.cpp:24-90): a thermodynamic force built on `subdomain` with the bin
// width of the file's first two bin centres, holding the file's forces
#pragma once

#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "mrmd_facade/core.hpp"

namespace mrmd
{
namespace io
{
inline action::ThermodynamicForce restoreThermoForce(const std::string& filename,
                                                     const data::Subdomain& subdomain,
                                                     const std::vector<real_t>& targetDensities = {1_r},
                                                     const std::vector<real_t>& thermodynamicForceModulations = {1_r},
                                                     const bool enforceSymmetry = false,
                                                     const bool usePeriodicity = false,
                                                     const idx_t maxNumForces = 10)
{
    std::ifstream file(filename);
    if (!file) throw std::runtime_error("restoreThermoForce: cannot open " + filename);
    std::string line, word;
    std::getline(file, line);
    std::vector<real_t> grid;
    {
        std::stringstream ss(line);
        while (ss >> word) grid.push_back(std::stod(word));
    }
    if (grid.size() < 2) throw std::runtime_error("restoreThermoForce: the grid line needs two bin centres");
    std::vector<std::vector<real_t>> forces;
    while (std::getline(file, line))
    {
        std::stringstream ss(line);
        std::vector<real_t> row;
        while (ss >> word) row.push_back(std::stod(word));
        if (row.empty()) continue;
        forces.push_back(row);
        if (idx_c(forces.size()) > maxNumForces) throw std::runtime_error("restoreThermoForce: too many forces");
    }
    action::ThermodynamicForce tf(targetDensities, subdomain, grid[1] - grid[0], thermodynamicForceModulations,
                                  enforceSymmetry, usePeriodicity);
    auto h = tf.getForce();
    if (idx_c(forces.size()) != h.numHistograms)
        throw std::runtime_error("restoreThermoForce: the file's force count differs from the atom types");
    for (idx_t t = 0; t < h.numHistograms; ++t)
    {
        if (idx_c(forces[size_t(t)].size()) != h.numBins)
            throw std::runtime_error("restoreThermoForce: the file's bin count differs from the subdomain's");
        for (idx_t b = 0; b < h.numBins; ++b) h(b, t) = forces[size_t(t)][size_t(b)];
    }
    tf.setForce(h);
    return tf;
}
}  // namespace io
}  // namespace mrmd

// reference include path io/DumpThermoForce.hpp of the mrmd action API: the thermodynamic
// force table as text (DumpThermoForce.cpp:24-64, DumpProfile.cpp:22-48)
#pragma once

#include <fstream>
#include <string>
#include <vector>

#include "mrmd_facade/core.hpp"

namespace mrmd
{
namespace io
{
namespace impl
{
/// one line of values separated by single spaces, default stream formatting
inline void dumpLine(std::ofstream& f, const std::vector<real_t>& v)
{
    for (size_t i = 0; i < v.size(); ++i) f << v[i] << (i + 1 < v.size() ? " " : "");
    f << std::endl;
}
}  // namespace impl

/// the grid line, then the force of type `typeId`
inline void dumpThermoForce(const std::string& filename, const action::ThermodynamicForce& tf, const idx_t& typeId)
{
    std::ofstream f(filename);
    impl::dumpLine(f, data::createGrid(tf.getForce()));
    impl::dumpLine(f, tf.getForce(typeId));
}

/// the grid line, then one line per atom type
inline void dumpThermoForce(const std::string& filename, const action::ThermodynamicForce& tf)
{
    const auto force = tf.getForce();
    std::ofstream f(filename);
    impl::dumpLine(f, data::createGrid(force));
    for (idx_t t = 0; t < force.numHistograms; ++t) impl::dumpLine(f, tf.getForce(t));
}
}  // namespace io
}  // namespace mrmd

// Reference unit tests restated against the drop-in headers (same class names and calls):
// VelocityVerlet.test.cpp, VelocityVerletLangevinThermostat.test.cpp, LJ_IdealGas.test.cpp,
// ThermodynamicForce.test.cpp, Slab.test.cpp, CheckRegion.test.cpp, UpdateMolecules.test.cpp,
// ContributeMoleculeForceToAtoms.test.cpp, GhostExchange.test.cpp,
// MultiResPeriodicGhostExchange.test.cpp, io/ThermoForce.test.cpp, Spherical.test.cpp,
// BerendsenThermostat.test.cpp.
#include <cstdio>

#include "action/BerendsenThermostat.hpp"
#include "action/ContributeMoleculeForceToAtoms.hpp"
#include "action/LJ_IdealGas.hpp"
#include "action/ThermodynamicForce.hpp"
#include "action/UpdateMolecules.hpp"
#include "action/VelocityVerlet.hpp"
#include "action/VelocityVerletLangevinThermostat.hpp"
#include "check.hpp"
#include "communication/GhostLayer.hpp"
#include "communication/MultiResGhostLayer.hpp"
#include "data/MoleculesFromAtoms.hpp"
#include "io/DumpThermoForce.hpp"
#include "io/RestoreThermoForce.hpp"
#include "weighting_function/CheckRegion.hpp"
#include "weighting_function/Slab.hpp"
#include "weighting_function/Spherical.hpp"
#include "analysis/KineticEnergy.hpp"

using namespace mrmd;

static data::Atoms singleAtom()
{
    data::HostAtoms h(1);
    h.getPos()(0, 0) = 2_r, h.getPos()(0, 1) = 3_r, h.getPos()(0, 2) = 4_r;
    h.getVel()(0, 0) = 7_r, h.getVel()(0, 1) = 5_r, h.getVel()(0, 2) = 3_r;
    h.getForce()(0, 0) = 9_r, h.getForce()(0, 1) = 7_r, h.getForce()(0, 2) = 8_r;
    h.getMass()(0) = 1.5_r;
    h.getCharge()(0) = 0.5_r;
    h.numLocalAtoms = 1;
    return data::Atoms(h);
}

static void velocityVerlet()
{
    auto atoms = singleAtom();
    action::VelocityVerlet::preForceIntegrate(atoms, 4_r);
    data::HostAtoms h(atoms);
    EXPECT_FLOAT_EQ(h.getVel()(0, 0), 19_r);
    EXPECT_FLOAT_EQ(h.getVel()(0, 1), 14.333333_r);
    EXPECT_FLOAT_EQ(h.getVel()(0, 2), 13.666667_r);
    EXPECT_FLOAT_EQ(h.getPos()(0, 0), 78_r);
    EXPECT_FLOAT_EQ(h.getPos()(0, 1), 60.333332_r);
    EXPECT_FLOAT_EQ(h.getPos()(0, 2), 58.666668_r);
    auto atoms2 = singleAtom();
    action::VelocityVerlet::postForceIntegrate(atoms2, 4_r);
    data::HostAtoms h2(atoms2);
    EXPECT_FLOAT_EQ(h2.getVel()(0, 0), 19_r);
    EXPECT_FLOAT_EQ(h2.getPos()(0, 0), 2_r);
}

static void langevin()
{
    auto atoms = singleAtom();
    action::VelocityVerletLangevinThermostat integrator(0.5_r, 0.5_r);
    integrator.preForceIntegrate_apply_if(atoms, 4_r, pred::Never{});
    data::HostAtoms h(atoms);
    EXPECT_FLOAT_EQ(h.getVel()(0, 0), 19_r);
    EXPECT_FLOAT_EQ(h.getPos()(0, 1), 60.333332_r);
    auto atoms2 = singleAtom();
    integrator.preForceIntegrate(atoms2, 4_r);
    data::HostAtoms h2(atoms2);
    EXPECT_FALSE(std::abs(h2.getVel()(0, 0) - 19_r) < 1e-6);
    // local thermostat: x > 78 is false after the first step, true after the second
    auto atoms3 = singleAtom();
    integrator.preForceIntegrate_apply_if(atoms3, 4_r, pred::Greater{78_r});
    data::HostAtoms h3(atoms3);
    EXPECT_FLOAT_EQ(h3.getPos()(0, 0), 78_r);
}

static void ljIdealGas(real_t lambda, real_t scale)
{
    data::HostMolecules hm(2);
    hm.getPos()(0, 0) = -0.5_r;
    hm.getPos()(1, 0) = +0.5_r;
    hm.getAtomsOffset()(0) = 0, hm.getNumAtoms()(0) = 2;
    hm.getAtomsOffset()(1) = 2, hm.getNumAtoms()(1) = 2;
    hm.getModulatedLambda()(0) = lambda, hm.getModulatedLambda()(1) = lambda;
    hm.numLocalMolecules = 2;
    data::Molecules molecules(hm);
    data::HostAtoms ha(4);
    const real_t p[4][2] = {{-0.5, -0.5}, {-0.5, 0.5}, {0.5, -0.5}, {0.5, 0.5}};
    for (int i = 0; i < 4; ++i)
    {
        ha.getPos()(i, 0) = p[i][0];
        ha.getPos()(i, 1) = p[i][1];
        ha.getRelativeMass()(i) = 0.5_r;
        ha.getForce()(i, 0) = ha.getForce()(i, 1) = ha.getForce()(i, 2) = (lambda == 0 ? 2_r : 0_r);
    }
    ha.numLocalAtoms = 4;
    data::Atoms atoms(ha);
    HalfVerletList vl;
    real_t minGrid[3] = {-1_r, -1_r, -1_r}, maxGrid[3] = {1_r, 1_r, 1_r};
    vl.build(molecules.getPos(), 0, molecules.numLocalMolecules, 2_r, 1_r, minGrid, maxGrid, 4);
    constexpr real_t sigma = 0.9_r;
    action::LJ_IdealGas LJ(0_r, 2.5_r * sigma, sigma, 2_r, true);
    LJ.run(molecules, vl, atoms);
    data::HostAtoms h(atoms);
    if (lambda == 0)
    {
        for (int i = 0; i < 4; ++i)
            for (int d = 0; d < 3; ++d) EXPECT_FLOAT_EQ(h.getForce()(i, d), 2_r);
        return;
    }
    const real_t fx = 0.22156665_r * scale, fy = 1.3825009_r * scale;
    EXPECT_FLOAT_EQ(h.getForce()(0, 0), -fx);
    EXPECT_FLOAT_EQ(h.getForce()(0, 1), +fy);
    EXPECT_FLOAT_EQ(h.getForce()(1, 1), -fy);
    EXPECT_FLOAT_EQ(h.getForce()(2, 0), +fx);
    EXPECT_FLOAT_EQ(h.getForce()(3, 1), -fy);
}

static void thermodynamicForce()
{
    data::Subdomain subdomain({0_r, 0_r, 0_r}, {10_r, 1_r, 1_r}, 1_r);
    action::ThermodynamicForce tf({1_r}, subdomain, 1_r, {1_r});
    auto force = tf.getForce();
    for (idx_t i = 0; i < force.numBins; ++i) force(i, 0) = real_c(i);
    tf.setForce(force);
    data::HostAtoms ha(45);
    idx_t n = 0;
    for (idx_t i = 0; i < 10; ++i)
        for (idx_t j = 0; j < i; ++j) ha.getPos()(n++, 0) = real_c(i) + 0.5_r;
    ha.numLocalAtoms = 45;
    data::Atoms atoms(ha);
    tf.sample(atoms);
    auto density = tf.getDensityProfile();
    for (idx_t i = 0; i < density.numBins; ++i) EXPECT_FLOAT_EQ(density(i, 0), real_c(i));
    tf.update(1_r, 0_r);
    auto updated = tf.getForce();
    for (idx_t i = 0; i < updated.numBins; ++i) EXPECT_FLOAT_EQ(updated(i, 0), real_c(i) - 1_r);
    EXPECT_EQ(tf.getNumberOfDensityProfileSamples(), 0);
}

// io/ThermoForce.test.cpp:25-120: forces (bin + 1)(type + 1) on 100 bins of the x extent of
// subdomain (1, 2, 3)-(4, 6, 8), dumped and restored; single type and five types
static action::ThermodynamicForce createThermoForce(idx_t numBins, idx_t numForces, const data::Subdomain& subdomain)
{
    const real_t binWidth = (subdomain.maxCorner[0] - subdomain.minCorner[0]) / real_c(numBins);
    action::ThermodynamicForce tf(std::vector<real_t>(size_t(numForces), 1_r), subdomain, binWidth,
                                  std::vector<real_t>(size_t(numForces), 1_r));
    auto f = tf.getForce();
    for (idx_t b = 0; b < numBins; ++b)
        for (idx_t t = 0; t < numForces; ++t) f(b, t) = 1_r * (real_c(b) + 1) * (real_c(t) + 1);
    tf.setForce(f);
    return tf;
}

static void thermoForceDumpRestore()
{
    const data::Subdomain subdomain({1_r, 2_r, 3_r}, {4_r, 6_r, 8_r}, 0.5_r);
    {
        auto tf1 = createThermoForce(100, 1, subdomain);
        io::dumpThermoForce("dummySingleForce.txt", tf1, 0);
        auto tf2 = io::restoreThermoForce("dummySingleForce.txt", subdomain);
        const auto g1 = data::createGrid(tf1.getForce()), g2 = data::createGrid(tf2.getForce());
        const auto f1 = tf1.getForce(0), f2 = tf2.getForce(0);
        EXPECT_EQ(idx_c(g2.size()), 100);
        for (size_t b = 0; b < 100; ++b)
        {
            EXPECT_FLOAT_EQ(g1[b], g2[b]);
            EXPECT_FLOAT_EQ(f1[b], f2[b]);
        }
    }
    {
        auto tf1 = createThermoForce(100, 5, subdomain);
        io::dumpThermoForce("dummyMultipleForces.txt", tf1);
        const std::vector<real_t> ones(5, 1_r);
        auto tf2 = io::restoreThermoForce("dummyMultipleForces.txt", subdomain, ones, ones);
        const auto h1 = tf1.getForce(), h2 = tf2.getForce();
        const auto g1 = data::createGrid(h1), g2 = data::createGrid(h2);
        for (idx_t b = 0; b < 100; ++b)
        {
            EXPECT_FLOAT_EQ(g1[size_t(b)], g2[size_t(b)]);
            for (idx_t t = 0; t < 5; ++t) EXPECT_FLOAT_EQ(h1(b, t), h2(b, t));
        }
    }
}

// weighting_function/Spherical.test.cpp:22-85
static void spherical()
{
    const Point3D center = {2_r, 3_r, 4_r};
    const auto weight = weighting_function::Spherical(center, 2_r, 2_r, 2);
    Point3D pos = center;
    auto w = weight(pos[0], pos[1], pos[2]);
    for (auto i = 0; i < 60; ++i)
    {
        const auto old = w;
        w = weight(pos[0], pos[1], pos[2]);
        EXPECT_LE(w, old);
        for (int d = 0; d < 3; ++d) pos[d] += 0.1_r;
    }
    EXPECT_FLOAT_EQ(weight(2.1_r, 3.1_r, 4.1_r), 1_r);
    EXPECT_FLOAT_EQ(weight(6.1_r, 7.1_r, 8.1_r), 0_r);
    const auto eps = 1e-10_r;
    pos = center;
    for (auto i = 0; i < 60; ++i)
    {
        real_t l0, gx0, gy0, gz0, l1, gx1, gy1, gz1;
        weight(pos[0], 0_r, 0_r, l0, gx0, gy0, gz0);
        weight(pos[0] + eps, 0_r, 0_r, l1, gx1, gy1, gz1);
        EXPECT_FLOAT_EQ((l1 - l0) / eps, 0.5_r * (gx0 + gx1));
        pos[0] += 0.1_r;
    }
}

// action/BerendsenThermostat.test.cpp:24-45 on the single-atom fixture
static void berendsen()
{
    for (const real_t gamma : {0_r, 1_r})
    {
        auto atoms = singleAtom();
        auto T = analysis::getMeanKineticEnergy(atoms) * (2_r / 3_r);
        action::BerendsenThermostat::apply(atoms, T, 3.8_r, gamma);
        T = analysis::getMeanKineticEnergy(atoms) * (2_r / 3_r);
        EXPECT_FLOAT_EQ(T, gamma == 0_r ? 41.5_r : 3.8_r);
    }
}

static void slabAndRegions()
{
    auto weight = weighting_function::Slab({2_r, 3_r, 4_r}, 2_r, 2_r, 2);
    EXPECT_TRUE(weight.isInATRegion(2_r, 10_r, 10_r));
    EXPECT_TRUE(weight.isInHYRegion(3.1_r, 10_r, 10_r));
    EXPECT_TRUE(weight.isInCGRegion(5.2_r, 10_r, 10_r));
    real_t tmp, w;
    weight(2.9_r, 3.1_r, 4.1_r, tmp, w, tmp, tmp, tmp);
    EXPECT_FLOAT_EQ(w, 1_r);
    weight(5.1_r, 7.1_r, 8.1_r, tmp, w, tmp, tmp, tmp);
    EXPECT_FLOAT_EQ(w, 0_r);
    EXPECT_TRUE(weighting_function::isInATRegion(1_r));
    EXPECT_TRUE(weighting_function::isInHYRegion(0.5_r));
    EXPECT_TRUE(weighting_function::isInCGRegion(0_r));
}

static void moleculesUpdateAndContribute()
{
    data::HostAtoms ha(4);
    const real_t p[4][2] = {{1, 0}, {0, 1}, {-1, 0}, {0, -1}};
    const real_t rm[4] = {0.25, 0.75, 0.25, 0.75};
    for (int i = 0; i < 4; ++i)
    {
        ha.getPos()(i, 0) = p[i][0];
        ha.getPos()(i, 1) = p[i][1];
        ha.getRelativeMass()(i) = rm[i];
        ha.getMass()(i) = rm[i] * 4;
    }
    ha.numLocalAtoms = 4;
    data::Atoms atoms(ha);
    data::HostMolecules hm(2);
    hm.getAtomsOffset()(0) = 0, hm.getNumAtoms()(0) = 2;
    hm.getAtomsOffset()(1) = 2, hm.getNumAtoms()(1) = 2;
    hm.getForce()(0, 0) = 5_r, hm.getForce()(0, 1) = 6_r, hm.getForce()(0, 2) = 7_r;
    hm.numLocalMolecules = 2;
    data::Molecules molecules(hm);
    action::UpdateMolecules::update(molecules, atoms, weighting_function::Slab({100_r, 0_r, 0_r}, 1_r, 1_r, 1));
    data::HostMolecules h(molecules);
    EXPECT_FLOAT_EQ(h.getPos()(0, 0), 0.25_r);
    EXPECT_FLOAT_EQ(h.getPos()(0, 1), 0.75_r);
    EXPECT_FLOAT_EQ(h.getPos()(1, 0), -0.25_r);
    action::ContributeMoleculeForceToAtoms::update(molecules, atoms);
    data::HostAtoms hf(atoms);
    EXPECT_FLOAT_EQ(hf.getForce()(0, 0), 5_r * 0.25_r);
    EXPECT_FLOAT_EQ(hf.getForce()(1, 2), 7_r * 0.75_r);
}

static void ghostExchange()
{
    data::Subdomain subdomain({0_r, 0_r, 0_r}, {3_r, 3_r, 3_r}, 0.7_r);
    data::HostAtoms h(27);
    idx_t idx = 0;
    for (real_t x = 0.5_r; x < 3_r; x += 1_r)
        for (real_t y = 0.5_r; y < 3_r; y += 1_r)
            for (real_t z = 0.5_r; z < 3_r; z += 1_r)
            {
                h.getPos()(idx, 0) = x, h.getPos()(idx, 1) = y, h.getPos()(idx, 2) = z;
                ++idx;
            }
    h.numLocalAtoms = 27;
    data::Atoms atoms(h);
    communication::GhostLayer gl;
    gl.createGhostAtoms(atoms, subdomain);
    EXPECT_EQ(atoms.numGhostAtoms, 98);
    auto corr = gl.getCorrespondingRealAtom(atoms.numLocalAtoms + atoms.numGhostAtoms);
    for (idx_t i = 0; i < 27; ++i) EXPECT_EQ(corr[size_t(i)], -1);
    for (idx_t i = 27; i < 27 + 98; ++i) EXPECT_TRUE(corr[size_t(i)] >= 0 && corr[size_t(i)] < 27);

    // MultiResPeriodicGhostExchange.test.cpp: 2 atoms per molecule
    data::HostMolecules hm(27);
    data::HostAtoms ha(54);
    idx = 0;
    for (real_t x = 0.5_r; x < 3_r; x += 1_r)
        for (real_t y = 0.5_r; y < 3_r; y += 1_r)
            for (real_t z = 0.5_r; z < 3_r; z += 1_r)
            {
                hm.getPos()(idx, 0) = x, hm.getPos()(idx, 1) = y, hm.getPos()(idx, 2) = z;
                hm.getAtomsOffset()(idx) = 2 * idx, hm.getNumAtoms()(idx) = 2;
                for (int i = 0; i < 2; ++i)
                {
                    ha.getPos()(2 * idx + i, 0) = x + 0.1_r * i;
                    ha.getPos()(2 * idx + i, 1) = y + 0.2_r * i;
                    ha.getPos()(2 * idx + i, 2) = z + 0.3_r * i;
                }
                ++idx;
            }
    hm.numLocalMolecules = 27;
    ha.numLocalAtoms = 54;
    data::Molecules molecules(hm);
    data::Atoms atoms2(ha);
    communication::MultiResGhostLayer mgl;
    mgl.createGhostAtoms(molecules, atoms2, subdomain);
    EXPECT_EQ(molecules.numGhostMolecules, 5 * 5 * 5 - 3 * 3 * 3);
    EXPECT_EQ(atoms2.numGhostAtoms, (5 * 5 * 5 - 3 * 3 * 3) * 2);
}

int main(int argc, char** argv)
{
    ScopeGuard guard(argc, argv);
    velocityVerlet();
    langevin();
    ljIdealGas(1_r, 1_r);
    ljIdealGas(0.5_r, 0.5_r);
    ljIdealGas(0_r, 0_r);
    thermodynamicForce();
    thermoForceDumpRestore();
    spherical();
    berendsen();
    slabAndRegions();
    moleculesUpdateAndContribute();
    ghostExchange();
    CHECK_DONE();
}

// The reference's ESPResSo++ comparison (tests/LennardJones/LennardJones.cpp:82-143) written
// against the drop-in headers: same includes, same class names and call sequence.
#include <cstdio>

#include "action/LennardJones.hpp"
#include "action/VelocityVerlet.hpp"
#include "check.hpp"
#include "communication/GhostLayer.hpp"
#include "data/Subdomain.hpp"
#include "io/RestoreTXT.hpp"

using namespace mrmd;

int main(int argc, char** argv)
{
    if (argc < 2)
    {
        std::printf("usage: %s positions.txt\n", argv[0]);
        return 2;
    }
    ScopeGuard guard(argc, argv);
    {
        constexpr double rc = 2.5, skin = 0.3, dt = 0.005;
        auto subdomain = data::Subdomain({0_r, 0_r, 0_r}, {33.8585, 33.8585, 33.8585}, rc + skin);
        auto atoms = io::restoreAtoms(argv[1]);
        EXPECT_EQ(atoms.numLocalAtoms, 32768);

        auto ghostExchange = communication::GhostLayer();
        ghostExchange.createGhostAtoms(atoms, subdomain);
        fence();
        EXPECT_EQ(atoms.numLocalAtoms, 32768);
        EXPECT_EQ(atoms.numGhostAtoms, 22104);
        atoms.resize(size_t(atoms.numLocalAtoms + atoms.numGhostAtoms));

        double cellRatio = 1.0_r;
        HalfVerletList verletList(atoms.getPos(), 0, atoms.numLocalAtoms, rc + skin, cellRatio,
                                  subdomain.minGhostCorner.data(), subdomain.maxGhostCorner.data());
        EXPECT_EQ(verletList.numPairs(), 1310403);

        action::LennardJones LJ(rc, 1_r, 1_r);
        LJ.apply(atoms, verletList);
        real_t totalEnergy = LJ.getEnergy();
        EXPECT_FLOAT_EQ(totalEnergy, -94795.927_r);
        std::printf("ghosts %lld pairs %lld energy %.6f\n", (long long)atoms.numGhostAtoms,
                    (long long)verletList.numPairs(), totalEnergy);

        action::VelocityVerlet::preForceIntegrate(atoms, dt);
        LJ.apply(atoms, verletList);
        action::VelocityVerlet::postForceIntegrate(atoms, dt);
        fence();
    }
    CHECK_DONE();
}

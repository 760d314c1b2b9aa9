// The reference's Argon NVE example (examples/02_LennardJones_NVE/02_LennardJones_NVE.cpp:
// 87-216) written against the drop-in headers: the phase point restored from the GRO file,
// velocity Verlet with the half-skin rebuild rule, ghost exchange, the half Verlet list and
// the unshifted, capped 4-argument LennardJones.  Output is cut to what the parity test
// compares: per output step the potential and kinetic energy per atom and the rebuild count.
#include <cstdio>
#include <cstdlib>
#include <limits>

#include "action/LennardJones.hpp"
#include "action/VelocityVerlet.hpp"
#include "analysis/KineticEnergy.hpp"
#include "check.hpp"
#include "communication/GhostLayer.hpp"
#include "data/Atoms.hpp"
#include "data/Subdomain.hpp"
#include "io/RestoreGRO.hpp"

using namespace mrmd;

int main(int argc, char** argv)
{
    if (argc < 4)
    {
        std::printf("usage: %s file.gro nsteps output_interval\n", argv[0]);
        return 2;
    }
    ScopeGuard guard(argc, argv);
    {
        // Config of the example (02_LennardJones_NVE.cpp:49-85)
        const idx_t nsteps = std::atoll(argv[2]);
        const idx_t outputInterval = std::atoll(argv[3]);
        constexpr real_t dt = 0.002_r;
        constexpr real_t sigma = 1_r, epsilon = 1_r;
        constexpr real_t r_cut = 2.5_r * sigma, r_cap = 0.7_r * sigma;
        constexpr real_t skin = 0.1_r * sigma;
        constexpr real_t neighborCutoff = r_cut + skin;
        constexpr real_t cell_ratio = 1_r;
        constexpr idx_t estimatedMaxNeighbors = 60;

        auto subdomain = data::Subdomain({0_r, 0_r, 0_r}, {0_r, 0_r, 0_r}, neighborCutoff);
        auto atoms = data::Atoms(0);
        io::restoreGRO(argv[1], subdomain, atoms);
        EXPECT_EQ(atoms.numLocalAtoms, 4096);
        EXPECT_EQ(subdomain.diameter[0], 20_r);
        const auto volume = subdomain.diameter[0] * subdomain.diameter[1] * subdomain.diameter[2];
        std::printf("rho: %.6f\n", real_c(atoms.numLocalAtoms) / volume);

        communication::GhostLayer ghostLayer;
        HalfVerletList verletList;
        real_t maxAtomDisplacement = std::numeric_limits<real_t>::max();
        idx_t rebuildCounter = 0;
        action::LennardJones lennardJones(r_cut, sigma, epsilon, r_cap);

        for (idx_t step = 0; step < nsteps; ++step)
        {
            maxAtomDisplacement += action::VelocityVerlet::preForceIntegrate(atoms, dt);
            if (maxAtomDisplacement >= skin * 0.5_r)
            {
                maxAtomDisplacement = 0_r;
                ghostLayer.exchangeRealAtoms(atoms, subdomain);
                ghostLayer.createGhostAtoms(atoms, subdomain);
                verletList.build(atoms.getPos(), 0, atoms.numLocalAtoms, neighborCutoff, cell_ratio,
                                 subdomain.minGhostCorner.data(), subdomain.maxGhostCorner.data(),
                                 estimatedMaxNeighbors);
                ++rebuildCounter;
            }
            else
            {
                ghostLayer.updateGhostAtoms(atoms, subdomain);
            }
            atoms.setForce(0_r);
            lennardJones.apply(atoms, verletList);
            ghostLayer.contributeBackGhostToReal(atoms);
            action::VelocityVerlet::postForceIntegrate(atoms, dt);
            if (step % outputInterval == 0)
            {
                const auto E0 = lennardJones.getEnergy() / real_c(atoms.numLocalAtoms);
                const auto Ek = analysis::getMeanKineticEnergy(atoms);
                std::printf("step %lld E0 %.17g Ek %.17g rebuilds %lld Nghost %lld\n", (long long)step, E0, Ek,
                            (long long)rebuildCounter, (long long)atoms.numGhostAtoms);
            }
        }
        fence();
        std::printf("rebuilds %lld\n", (long long)rebuildCounter);
    }
    CHECK_DONE();
}

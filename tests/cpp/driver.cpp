// A reference-style driver over the C++ adapter (include/svm.hpp), in the
// shape of the reference's sphericalVertex.cpp / minimization.cpp drivers:
// build a model, stabilise with Brownian dynamics at T = 0 and (A0, P0) =
// (1, 1), run active dynamics at the target p0, then minimise with FIRE; then
// a few overdamped steps of the Voronoi model of the same cell points.
//
//   driver N seed p0 steps
//
// Prints "energy_after_stabilise", "energy_after_active", FIRE iterations,
// force and final energy, and a checksum of the final positions.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <memory>

#include "svm.hpp"

int main(int argc, char** argv) {
    const int N = argc > 1 ? std::atoi(argv[1]) : 512;
    const unsigned long long seed = argc > 2 ? std::strtoull(argv[2], nullptr, 10) : 1137;
    const double p0 = argc > 3 ? std::atof(argv[3]) : 3.8;
    const int steps = argc > 4 ? std::atoi(argv[4]) : 100;
    try {
        auto model = std::make_shared<svm::sphericalVertexModel>(N, seed, 1.0, p0);
        model->setPreferredParameters(1.0, 1.0);
        auto bd = std::make_shared<svm::brownianDynamics>(true);
        bd->setT(0.0);
        bd->setDeltaT(0.01);
        svm::Simulation sim;
        sim.setConfiguration(model);
        sim.addUpdater(bd, model);
        for (int i = 0; i < 200; ++i) sim.performTimestep();
        std::printf("energy_after_stabilise %.17g\n", sim.computePotentialEnergy());
        model->setPreferredParameters(1.0, p0);
        auto active = std::make_shared<svm::sphericalSelfPropelledParticle>();
        active->setV0(0.1);
        active->setTau(1.0);
        active->setMu(1.0);
        active->setDeltaT(0.01);
        sim.clearUpdaters();
        sim.addUpdater(active, model);
        for (int i = 0; i < steps; ++i) sim.performTimestep();
        std::printf("energy_after_active %.17g\n", sim.computePotentialEnergy());
        auto fire = std::make_shared<svm::energyMinimizerFIRE>(model);
        fire->setFIREParameters(0.001, 0.99, 0.05, 1.1, 0.95, 0.9, 4, 1e-10, 0.0);
        fire->setStopRms(true);
        fire->setMaximumIterations(20000);
        fire->minimize();
        std::printf("fire_iterations %d\nfire_force %.6g\nenergy_after_fire %.17g\n", fire->getCurrentIterations(),
                    fire->getMaxForce(), sim.computePotentialEnergy());
        double cs = 0.0;
        for (const auto& p : model->returnPositions()) cs += p[0] + 2.0 * p[1] + 3.0 * p[2];
        std::printf("position_checksum %.17g\n", cs);
        // the Voronoi model of the same cell points: overdamped Brownian steps
        // of the cell points, each followed by the retriangulation
        auto vor = std::make_shared<svm::sphericalVoronoi>(N, seed, 1.0, p0);
        auto vbd = std::make_shared<svm::brownianDynamics>(true);
        vbd->setT(0.0);
        vbd->setDeltaT(0.01);
        svm::Simulation vsim;
        vsim.setConfiguration(vor);
        vsim.addUpdater(vbd, vor);
        for (int i = 0; i < 5; ++i) vsim.performTimestep();
        std::printf("voronoi_energy %.17g\n", vsim.computePotentialEnergy());
        double vcs = 0.0;
        for (const auto& p : vor->returnPositions()) vcs += p[0] + 2.0 * p[1] + 3.0 * p[2];
        std::printf("voronoi_checksum %.17g\n", vcs);
        // metric neighbour lists of the vertex model's final positions, as
        // sphericalModel::setRadius configures them (range 1, sphereRadius
        // max|coordinate| + 2.5)
        const auto pos = model->returnPositions();
        double rmax = 0.0;
        for (const auto& p : pos)
            for (int k = 0; k < 3; ++k) rmax = std::fmax(rmax, std::fabs(p[k]));
        svm::neighborList nl(1.0, rmax + 2.5);
        nl.computeNeighborLists(pos);
        long long pairs = 0;
        double ncs = 0.0;
        for (size_t i = 0; i < pos.size(); ++i) {
            pairs += nl.neighborsPerParticle[i];
            for (int k = 0; k < nl.neighborsPerParticle[i]; ++k)
                ncs += nl.particleIndices[i * nl.Nmax + k] + nl.neighborDistances[i * nl.Nmax + k];
        }
        std::printf("neighbor_pairs %lld\nneighbor_nmax %d\nneighbor_checksum %.17g\n", pairs, nl.Nmax, ncs);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}

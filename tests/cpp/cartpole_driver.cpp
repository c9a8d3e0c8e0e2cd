// cartpole_driver.cpp -- the headless cart-pole run of Matthew_Sim_CBM
// (src/main.cpp:34-163 with --nogui -e cartpole; SimThread::SimThread,
// ::setupMFs and ::run, src/simthread.cpp:7-87,96-165), written against the
// reference's C++ API only: CRandomSFMT0, Cartpole, CBMState, CBMSimCore,
// PoissonRegenCells.
//
//   cartpole_driver <actParams> <conParams> <seed> <maxSteps> <trace.bin> [Cartpole options]
//
// Cartpole options are those of Cartpole::getOptions (--maxNumTrials,
// --maxTrialLen, --trackLen, --logfile, --simStateDir) and --closedLoop.
// Writes one row of 13 floats per step (the columns of cbm_env.h's
// cbm_cartpole_col) and prints the steps run.  The reference hands
// CBMSimCore &randSeed as the array of zone generator seeds; zone z gets
// randSeed + z here.
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <vector>

#include <CBMCoreInclude/interface/cbmsimcore.h>
#include <CBMEnvInclude/environments/cartpole.hpp>
#include <CBMStateInclude/interfaces/cbmstate.h>
#include <CBMToolsInclude/poissonregencells.h>
#include <CXXToolsInclude/randGenerators/sfmt.h>

int main(int argc, char **argv)
{
    if (argc < 6) {
        std::fprintf(stderr, "usage: cartpole_driver actParams conParams seed maxSteps trace.bin [options]\n");
        return 2;
    }
    const int randSeed = std::atoi(argv[3]);
    const long maxSteps = std::atol(argv[4]);
    try {
        std::srand(randSeed);
        CRandomSFMT0 randGen(randSeed);
        Cartpole env(&randGen, argc, argv);
        const int numMZ = env.numRequiredMZ();

        std::vector<int> crSeeds, arSeeds;
        for (int i = 0; i < numMZ; i++) {
            crSeeds.push_back(std::rand());
            arSeeds.push_back(std::rand());
        }
        std::fstream conP(argv[2], std::fstream::in), actP(argv[1], std::fstream::in);
        CBMState state(actP, conP, numMZ, randSeed, crSeeds.data(), arSeeds.data());
        std::vector<int> zoneSeeds(numMZ);
        for (int z = 0; z < numMZ; z++) zoneSeeds[z] = randSeed + z;
        CBMSimCore core(&state, zoneSeeds.data());

        env.setupMossyFibers(&state);
        const unsigned numMF = state.sizes().numMF, numNC = state.sizes().numNC;
        PoissonRegenCells mfs(numMF, randSeed, 4.0f, 1.0f);
        const bool closed = env.getClosedLoop();

        FILE *out = std::fopen(argv[5], "wb");
        if (!out) throw std::runtime_error("cannot open the trace file");
        long step = 0;
        for (; step < maxSteps && !env.terminated(); step++) {
            float *mfFreq = env.getState();
            const ct_uint8_t *apMF = mfs.calcActivity(mfFreq);
            core.updateMFInput(apMF);
            core.calcActivity();
            const unsigned long l0 = env.leftMZ().errorsDelivered(), r0 = env.rightMZ().errorsDelivered();
            env.step(&core);

            float ncL = 0, ncR = 0;
            if (closed) {
                ncL = env.leftMZ().lastNCSum();
                ncR = env.rightMZ().lastNCSum();
            } else {
                const ct_uint8_t *a = core.getMZoneList()[0]->exportAPNC();
                const ct_uint8_t *b = core.getMZoneList()[1]->exportAPNC();
                for (unsigned i = 0; i < numNC; i++) {
                    ncL += a[i];
                    ncR += b[i];
                }
            }
            int spikes = 0;
            for (unsigned i = 0; i < numMF; i++) spikes += apMF[i];
            const float row[13] = {env.getPoleAngle(),
                                   env.getPoleVelocity(),
                                   env.getCartPosition(),
                                   env.getCartVelocity(),
                                   env.getForceLeft(),
                                   env.getForceRight(),
                                   (float)(env.leftMZ().errorsDelivered() - l0),
                                   (float)(env.rightMZ().errorsDelivered() - r0),
                                   (float)env.getTimeAloft(),
                                   (float)env.getTrialNum(),
                                   ncL,
                                   ncR,
                                   (float)spikes};
            std::fwrite(row, sizeof(float), 13, out);
        }
        std::fclose(out);
        std::printf("%ld\n", step);
    } catch (const std::exception &e) {
        std::fprintf(stderr, "cartpole_driver: %s\n", e.what());
        return 1;
    }
    return 0;
}

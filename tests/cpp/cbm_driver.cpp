// cbm_driver.cpp -- a driver written against the reference's C++ API only
// (CBMState, CBMSimCore, InNetInterface, MZoneInterface), the way
// ECManagementDelay uses it (Eyelid_Sim_CBM_Dev/src/ectrial/ecmanagementdelay.cpp:201-273):
// per ms updateMFInput, an error drive at one step, a plasticity-step change,
// calcActivity, and export reads.
//
//   cbm_driver <actParams> <conParams> <steps> <seed> <out.bin>
//
// Writes the final exports as raw arrays (apBufGR, vmGR, apBufGO, vmGO,
// apBufSC, apBufPC, vmPC, apBufBC, apBufIO, vmIO, apBufNC, vmNC) and the eyelid
// trace of the EyelidIntegrator on zone 0's NC spikes (one float per step) for
// comparison.
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <vector>

#include <CBMCoreInclude/interface/cbmsimcore.h>
#include <CBMStateInclude/interfaces/cbmstate.h>
#include <CBMToolsInclude/eyelidintegrator.h>

int main(int argc, char **argv)
{
    if (argc != 6) {
        std::fprintf(stderr, "usage: cbm_driver actParams conParams steps seed out.bin\n");
        return 2;
    }
    std::fstream act(argv[1], std::ios::in), con(argv[2], std::ios::in);
    const int steps = std::atoi(argv[3]);
    const int seed = std::atoi(argv[4]);
    int crSeed = seed + 101, arSeed = seed + 1001, rSeed = seed + 1001;
    try {
        CBMState state(act, con, 1, seed, &crSeed, &arSeed);
        CBMSimCore sim(&state, &rSeed);
        const unsigned numMF = state.sizes().numMF;
        std::vector<ct_uint8_t> mf(numMF);
        // the delay protocol's readout (ecmanagementdelay.cpp:92-93)
        EyelidIntegrator eyelid(state.sizes().numNC, 1, 10.5, 0.012, -0.13, 0.13, 100);
        std::vector<float> lid;
        for (int t = 0; t < steps; t++) {
            for (unsigned i = 0; i < numMF; i++) mf[i] = ((i * 7u + (unsigned)t * 13u) % 97u) < 3u;
            if (t == steps / 2) sim.updateErrDrive(0, 1.0f);
            if (t == steps / 3) sim.getMZoneList()[0]->setGRPCPlastSteps(-0.0009f, 0.0005f);
            sim.updateMFInput(mf.data());
            sim.calcActivity();
            lid.push_back(eyelid.calcStep(sim.getMZoneList()[0]->exportAPNC()));
        }
        InNetInterface *in = sim.getInputNet();
        MZoneInterface *z = sim.getMZoneList()[0];
        const cbm_sizes &s = state.sizes();
        FILE *f = std::fopen(argv[5], "wb");
        if (!f) return 3;
        std::fwrite(in->exportAPBufGR(), 4, s.numGR, f);
        std::fwrite(in->exportVmGR(), 4, s.numGR, f);
        std::fwrite(in->exportAPBufGO(), 4, s.numGO, f);
        std::fwrite(in->exportVmGO(), 4, s.numGO, f);
        std::fwrite(in->exportAPBufSC(), 4, s.numSC, f);
        std::fwrite(z->exportAPBufPC(), 4, s.numPC, f);
        std::fwrite(z->exportVmPC(), 4, s.numPC, f);
        std::fwrite(z->exportAPBufBC(), 4, s.numBC, f);
        std::fwrite(z->exportAPBufIO(), 4, s.numIO, f);
        std::fwrite(z->exportVmIO(), 4, s.numIO, f);
        std::fwrite(z->exportAPBufNC(), 4, s.numNC, f);
        std::fwrite(z->exportVmNC(), 4, s.numNC, f);
        std::fwrite(lid.data(), 4, lid.size(), f);
        std::fclose(f);
        std::printf("cbm_driver: %d steps of %u GR done\n", steps, s.numGR);
    } catch (const std::exception &e) {
        std::fprintf(stderr, "cbm_driver: %s\n", e.what());
        return 1;
    }
    return 0;
}

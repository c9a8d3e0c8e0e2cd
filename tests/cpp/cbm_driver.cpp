// cbm_driver.cpp -- a driver written against the reference's C++ API only
// (CBMState, CBMSimCore, InNetInterface, MZoneInterface), the way
// ECManagementDelay uses it (Eyelid_Sim_CBM_Dev/src/ectrial/ecmanagementdelay.cpp:201-273):
// per ms updateMFInput, an error drive at one step, a plasticity-step change,
// a GR->GO increment set and reset (InNetInterface::setGIncGRtoGO), calcActivity,
// and export reads.
//
//   cbm_driver <actParams> <conParams> <steps> <seed> <out.bin> [<trials.dat> [<state.st>]]
//
// Writes the final exports as raw arrays (apBufGR, vmGR, apBufGO, vmGO,
// apBufSC, apBufPC, vmPC, apBufBC, apBufIO, vmIO, apBufNC, vmNC) and the eyelid
// trace of the EyelidIntegrator on zone 0's NC spikes (one float per step) for
// comparison.  With trials.dat it also records one trial of ECTrialsData over
// the whole run (steps/3 ms before, during and after the "CS"), with the
// histograms, rasters and eyelid trace of ECManagementDelay::calcSimActivity
// (ecmanagementdelay.cpp:233-271); with state.st it saves the final state
// with CBMSimCore::writeToState(fstream&).
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <vector>

#include <CBMCoreInclude/interface/cbmsimcore.h>
#include <CBMDataInclude/interfaces/ectrialsdata.h>
#include <CBMStateInclude/interfaces/cbmstate.h>
#include <CBMToolsInclude/eyelidintegrator.h>

int main(int argc, char **argv)
{
    if (argc < 6 || argc > 8) {
        std::fprintf(stderr, "usage: cbm_driver actParams conParams steps seed out.bin [trials.dat [state.st]]\n");
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
        const cbm_sizes &sz = state.sizes();
        ECTrialsData *data = nullptr;
        if (argc >= 7) {
            // the parameters of ECManagementDelay (ecmanagementdelay.cpp:115-160)
            std::map<std::string, PSHParams> psh;
            PSHParams pp;
            pp.numTimeStepsPerBin = 10;
            pp.numCells = sz.numGO;
            psh["go"] = psh["grInputGO"] = psh["goInputGO"] = psh["goOutSynScale"] = pp;
            pp.numCells = sz.numMF;
            psh["mf"] = pp;
            pp.numCells = sz.numGR / 10;
            psh["gr"] = pp;
            pp.numCells = sz.numSC;
            psh["sc"] = pp;
            pp.numCells = sz.numBC;
            psh["bc"] = pp;
            pp.numCells = sz.numPC;
            psh["pc"] = pp;
            pp.numCells = sz.numIO;
            psh["io"] = pp;
            std::map<std::string, RasterParams> ras;
            ras["go"].numCells = sz.numGO;
            ras["sc"].numCells = sz.numSC;
            ras["bc"].numCells = sz.numBC;
            ras["pc"].numCells = sz.numPC;
            ras["nc"].numCells = sz.numNC;
            ras["io"].numCells = sz.numIO;
            EyelidOutParams ep;
            ep.numTimeStepSmooth = 4;
            data = new ECTrialsData(steps / 3, steps / 3, steps - 2 * (steps / 3), 1, 1, psh, ras,
                                    std::map<std::string, RawUIntParams>(), ep);
        }
        for (int t = 0; t < steps; t++) {
            for (unsigned i = 0; i < numMF; i++) mf[i] = ((i * 7u + (unsigned)t * 13u) % 97u) < 3u;
            if (t == steps / 2) sim.updateErrDrive(0, 1.0f);
            if (t == steps / 3) sim.getMZoneList()[0]->setGRPCPlastSteps(-0.0009f, 0.0005f);
            if (t == steps / 4) sim.getInputNet()->setGIncGRtoGO(0.0003f);
            if (t == 3 * steps / 4) sim.getInputNet()->resetGIncGRtoGO();
            sim.updateMFInput(mf.data());
            sim.calcActivity();
            const float pos = eyelid.calcStep(sim.getMZoneList()[0]->exportAPNC());
            lid.push_back(pos);
            if (data) {
                InNetInterface *in = sim.getInputNet();
                MZoneInterface *z0 = sim.getMZoneList()[0];
                if (t > 0 && t % (int)data->getTSPerRasterUpdate() == 0) {
                    data->updateRaster("go", in->exportAPBufGO());
                    data->updateRaster("sc", in->exportAPBufSC());
                    data->updateRaster("bc", z0->exportAPBufBC());
                    data->updateRaster("pc", z0->exportAPBufPC());
                    data->updateRaster("nc", z0->exportAPBufNC());
                    data->updateRaster("io", z0->exportAPBufIO());
                }
                data->updatePSH("go", in->exportAPGO());
                data->updatePSH("grInputGO", in->exportSumGRInputGO());
                data->updatePSH("goInputGO", in->exportSumGOInputGO());
                data->updatePSH("goOutSynScale", in->exportGOOutSynScaleGOGO());
                data->updatePSH("mf", mf.data());
                data->updatePSH("gr", in->exportAPGR());
                data->updatePSH("sc", in->exportAPSC());
                data->updatePSH("bc", z0->exportAPBC());
                data->updatePSH("pc", z0->exportAPPC());
                data->updatePSH("io", z0->exportAPIO());
                data->updateEyelid(pos);
            }
        }
        if (data) {
            std::fstream df(argv[6], std::ios::out | std::ios::binary | std::ios::trunc);
            data->writeData(df);
            delete data;
        }
        if (argc >= 8) {
            std::fstream sf(argv[7], std::ios::out | std::ios::binary | std::ios::trunc);
            sim.writeToState(sf);
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

// eyelidintegrator.h -- eyelid position readout of the nucleus cells' spikes,
// with the constructor and calcStep of the reference's EyelidIntegrator
// (CBM_TOOLS_LIB_DEV/CBMToolsInclude/eyelidintegrator.h, implementation
// CBM_TOOLS_LIB_DEV/src/eyelidintegrator.cpp:13-68), so trial drivers that use
// it link unchanged.
#ifndef CBM_AMD_EYELIDINTEGRATOR_H
#define CBM_AMD_EYELIDINTEGRATOR_H

#include <vector>

#include <CXXToolsInclude/stdDefinitions/pstdint.h>

class EyelidIntegrator {
public:
    // numInputCells NC inputs; per-input conductance decays with gDecayTau and
    // rises by gIncrease per spike; their sum plus gShift (floored at 0) pulls
    // the lid towards maxAmplitude against a leak of gLeakRaw/(6-msPerTimeStep)
    EyelidIntegrator(unsigned int numInputCells, unsigned int msPerTimeStep, float gDecayTau, float gIncrease,
                     float gShift, float gLeakRaw, float maxAmplitude);

    // one time step from the NC spike bytes; returns the lid position
    float calcStep(const ct_uint8_t *apIn);

private:
    std::vector<float> g;
    float decay, inc, shift, leak, amp, pos;
};

#endif

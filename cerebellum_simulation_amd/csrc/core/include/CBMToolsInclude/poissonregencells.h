// poissonregencells.h -- mossy-fibre spike generator with a relative
// refractory threshold (CBM_TOOLS_LIB_DEV/CBMToolsInclude/poissonregencells.h:19-43,
// src/poissonregencells.cpp:10-66), as drivers use it: per step, cell i
// recovers its threshold towards 1 by 1 - exp(-msPerTS/threshDecayTau), spikes
// when a uniform draw falls below frequency[i] (Hz) * sPerTS * threshold, and
// a spike sets its threshold to 0.
#ifndef CBM_POISSONREGENCELLS_H
#define CBM_POISSONREGENCELLS_H

#include <vector>

#include <CXXToolsInclude/randGenerators/sfmt.h>
#include <CXXToolsInclude/stdDefinitions/pstdint.h>

class PoissonRegenCells
{
public:
    // The draws come from a generator seeded with IRandom(0, INT_MAX) of a
    // generator seeded with randSeed (one thread, as the reference runs).
    PoissonRegenCells(unsigned int numCells, int randSeed, float threshDecayTau, float msPerTimeStep);
    ~PoissonRegenCells();

    // One step: spikes (0/1) of every cell for the given rates; the array
    // stays valid until the next call.  A rate that is NaN never spikes.
    const ct_uint8_t *calcActivity(const float *frequencies);

    // Not in the reference: draw the next step's uniforms now (one per cell,
    // in cell order, the draws calcActivity would make), so that a driver can
    // do it while the network steps; the next calcActivity uses them.
    void prepare();

    // The relative thresholds after the last step (not in the reference).
    const float *thresholds() const { return threshs.data(); }

private:
    PoissonRegenCells();

    CRandomSFMT0 *gen;
    unsigned int nCells;
    float msPerTS, sPerTS, threshDecay;
    std::vector<ct_uint8_t> aps;
    std::vector<float> threshs;
    std::vector<double> draws;
    bool prepared = false;
};

#endif

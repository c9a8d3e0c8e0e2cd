// poissonregencells.cpp -- see CBMToolsInclude/poissonregencells.h.
//
// Arithmetic as the reference writes it (poissonregencells.cpp:23-33,49-66),
// binary32 with no contraction: threshDecay = 1 - exp(-msPerTS/tau) with the
// binary32 exp that <math.h>'s exp(float) overload selects; the threshold
// update t + (1 - t)*decay; the spike test draw < (freq*sPerTS)*t with the
// binary32 product widened to double for the comparison.
#include <CBMToolsInclude/poissonregencells.h>

#include <climits>
#include <cmath>

PoissonRegenCells::PoissonRegenCells(unsigned int numCells, int randSeed, float threshDecayTau,
                                     float msPerTimeStep)
    : nCells(numCells), msPerTS(msPerTimeStep), aps(numCells, 0), threshs(numCells, 1.0f)
{
    CRandomSFMT0 seeder(randSeed);
    gen = new CRandomSFMT0(seeder.IRandom(0, INT_MAX));
    sPerTS = msPerTimeStep / 1000;
    threshDecay = 1 - std::exp(-msPerTS / threshDecayTau);
}

PoissonRegenCells::~PoissonRegenCells() { delete gen; }

void PoissonRegenCells::prepare()
{
    draws.resize(nCells);
    for (unsigned int i = 0; i < nCells; i++) draws[i] = gen->Random();
    prepared = true;
}

const ct_uint8_t *PoissonRegenCells::calcActivity(const float *frequencies)
{
    if (!prepared) prepare();
    prepared = false;
    for (unsigned int i = 0; i < nCells; i++) {
        float t = threshs[i];
        t = t + (1 - t) * threshDecay;
        const float p = (frequencies[i] * sPerTS) * t;
        const bool spike = draws[i] < (double)p;
        aps[i] = spike ? 1 : 0;
        threshs[i] = spike ? 0.0f : t;
    }
    return aps.data();
}

// microzone.cpp -- see CBMEnvInclude/microzone.hpp.
//
// update() in binary32 as the reference writes it (microzone.cpp:31-50): the
// NC count is summed as a float, force += pow(count/numNC * forceScale,
// forcePow) with the float pow overload, force *= forceDecay; the moving sum
// adds the new count before it drops the one windowLength updates old, and the
// average is clamped to [0, 1].
#include <CBMEnvInclude/microzone.hpp>

#include <algorithm>
#include <cmath>

Microzone::Microzone()
    : mzNum(0), numNC(0), force(0), forceScale(0), forcePow(1), forceDecay(1), simCore(NULL),
      window(windowLength, 0.0f), windowPos(0), actSum(0), movingAvg(0), errTotal(0), errPending(0), lastSum(0)
{
}

Microzone::Microzone(std::string name, int mzNum, float forceScale, float forcePow, float forceDecay, int numNC,
                     CBMSimCore *simCore)
    : name(name), mzNum(mzNum), numNC(numNC), force(0), forceScale(forceScale), forcePow(forcePow),
      forceDecay(forceDecay), simCore(simCore), window(windowLength, 0.0f), windowPos(0), actSum(0),
      movingAvg(0), errTotal(0), errPending(0), lastSum(0)
{
}

bool Microzone::initialized() { return simCore != NULL; }

void Microzone::initialize(CBMSimCore *core, int n)
{
    simCore = core;
    numNC = n;
}

void Microzone::deliverError()
{
    errTotal++;
    errPending++;
    if (simCore) simCore->updateErrDrive((unsigned int)mzNum, 1.0f);
}

void Microzone::smartDeliverError()
{
    if (movingAvg < errSaturatedThreshold) deliverError();
}

void Microzone::update() { update(getApNC()); }

void Microzone::update(const ct_uint8_t *apNC)
{
    float sum = 0;
    for (int i = 0; i < numNC; i++) sum += apNC[i];
    lastSum = sum;
    force += std::pow((sum / float(numNC)) * forceScale, forcePow);
    force *= forceDecay;

    actSum += sum;
    const float oldest = window[windowPos];
    window[windowPos] = sum;
    windowPos = (windowPos + 1) % windowLength;
    actSum -= oldest;
    movingAvg = std::min(1.0f, std::max(0.0f, actSum / float(windowLength)));
}

const ct_uint8_t *Microzone::getApNC() { return simCore->getMZoneList()[mzNum]->exportAPNC(); }

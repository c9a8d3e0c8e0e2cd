// microzone.hpp -- one output microzone of an environment (the interface of
// Matthew_Sim_CBM/includes/microzone.hpp:14-86, behaviour of
// src/microzone.cpp:5-55): it turns its zone's NC spikes into a leaky force
// and a 100-step moving average of the NC count, and delivers error
// (climbing-fibre drive 1.0) to its zone.  The boost serialisation of the
// original is not kept.
#ifndef CBM_ENV_MICROZONE_HPP
#define CBM_ENV_MICROZONE_HPP

#include <string>
#include <vector>

#include <CBMCoreInclude/interface/cbmsimcore.h>
#include <CXXToolsInclude/stdDefinitions/pstdint.h>

class Microzone
{
public:
    Microzone();
    Microzone(std::string name, int mzNum, float forceScale, float forcePow, float forceDecay, int numNC = 0,
              CBMSimCore *simCore = NULL);

    // Attach the core (and the NC count) this microzone reads and drives.
    void initialize(CBMSimCore *core, int numNC);
    bool initialized();

    // updateErrDrive(mzNum, 1.0) on the core.
    void deliverError();
    // deliverError() unless the moving average has reached 0.8.
    void smartDeliverError();

    // Read the zone's NC spikes from the core and update force and average:
    //   force = (force + pow(nNC/numNC * forceScale, forcePow)) * forceDecay
    void update();
    // The same from spikes the caller supplies (numNC bytes).
    void update(const ct_uint8_t *apNC);

    float getForce() { return force; }
    float getMovingAverage() { return movingAvg; }
    const ct_uint8_t *getApNC();
    std::string getName() { return name; }
    void setName(std::string newName) { name = newName; }

    // Not in the reference: errors delivered in all and since the last
    // takeErrors() call, and the NC count of the last update().
    unsigned long errorsDelivered() const { return errTotal; }
    unsigned int takeErrors()
    {
        unsigned int e = errPending;
        errPending = 0;
        return e;
    }
    float lastNCSum() const { return lastSum; }

protected:
    std::string name;
    int mzNum, numNC;
    // The reference never initialises force; it starts at 0 here.
    float force, forceScale, forcePow, forceDecay;
    CBMSimCore *simCore;

    static const unsigned int windowLength = 100;
    // ring of the last windowLength NC counts (zeros before the first update)
    std::vector<float> window;
    unsigned int windowPos;
    float actSum, movingAvg;
    static constexpr float errSaturatedThreshold = .8f;

    unsigned long errTotal;
    unsigned int errPending;
    float lastSum;
};

#endif

// statevariable.hpp -- a variable of an environment encoded in the rates of
// a group of mossy fibres (the interface of Matthew_Sim_CBM/includes/
// statevariable.hpp:13-160).  Kinds:
//   HIGH_FREQ  the group fires at its excited rate at all times;
//   GAUSSIAN   the value, clamped to [min, max], excites each MF of the group by
//              a Gaussian tuning curve centred on its slot of the range (the
//              group's MFs tile the range in order; width gaussWidth slots,
//              peak normalised to 1): rate = g*excited + (1-g)*relaxed;
//   MANUAL     the environment supplies a 0..1 level per MF:
//              rate = 1.5*level*excited + (1-level)*relaxed.
// Arithmetic follows the reference's expression types: binary32 except the
// Gaussian's exp and square roots, which are binary64 (C's exp and
// std::sqrt of double operands), rounded to binary32 on assignment.
#ifndef CBM_ENV_STATEVARIABLE_HPP
#define CBM_ENV_STATEVARIABLE_HPP

#include <algorithm>
#include <cassert>
#include <cmath>
#include <iostream>
#include <math.h>
#include <string>
#include <vector>

#include <CXXToolsInclude/randGenerators/sfmt.h>

enum UpdateType { MANUAL, GAUSSIAN, HIGH_FREQ };

// What an environment needs of a state variable whatever its owner's type
// (the reference keeps StateVariable<Derived> objects in a vector of
// StateVariable<Environment> pointers by a cast; a common base does it here).
class StateVariableBase
{
public:
    StateVariableBase() {}
    StateVariableBase(std::string name, UpdateType type, float mfProportion)
        : name(name), numMF(0), mfProportion(mfProportion), mfFreq(NULL), mfFreqRelaxed(NULL),
          mfFreqExcited(NULL), type(type), minSVVal(0), maxSVVal(0), gaussWidth(6.0f)
    {
    }
    virtual ~StateVariableBase() {}

    // Write this variable's MF rates into the rate vector.
    virtual void update() = 0;

    // Take the last numMF entries of unassignedMFs, last first.
    void assignOrderedMFInds(std::vector<int> &unassignedMFs)
    {
        if (!mfInds.empty()) return;
        for (int i = 0; i < numMF; i++) {
            mfInds.push_back(unassignedMFs.back());
            unassignedMFs.pop_back();
        }
    }

    // Draw numMF entries without replacement: IRandom(0, size-1) per draw.
    void assignRandomMFInds(std::vector<int> &unassignedMFs, CRandomSFMT0 *randGen)
    {
        if (!mfInds.empty()) return;
        for (int i = 0; i < numMF; ++i) {
            const int k = randGen->IRandom(0, (int)unassignedMFs.size() - 1);
            mfInds.push_back(unassignedMFs[k]);
            unassignedMFs.erase(unassignedMFs.begin() + k);
        }
    }

    // The group takes totalMF * mfProportion MFs (binary32 product, truncated).
    void initialize(int totalMF, std::vector<float> *freq, std::vector<float> *relaxed, std::vector<float> *excited)
    {
        numMF = (int)(totalMF * mfProportion);
        mfFreq = freq;
        mfFreqRelaxed = relaxed;
        mfFreqExcited = excited;
    }

    void gaussMFAct(float svVal)
    {
        std::vector<float> &f = *mfFreq, &lo = *mfFreqRelaxed, &hi = *mfFreqExcited;
        const float cur = std::max(minSVVal, std::min(maxSVVal, svVal));
        const float range = maxSVVal - minSVVal;
        const float slot = range / (float)mfInds.size();
        float centre = minSVVal + slot / 2.0;
        const float sd = gaussWidth * slot;
        const float sd2 = sd * sd;
        // 1 / sqrt(2 pi sd^2): the density at the centre
        const float peak = 1.0 / std::sqrt(2 * M_PI * sd2);
        for (size_t j = 0; j < mfInds.size(); j++) {
            const float d = cur - centre;
            const float arg = -1 * (d * d) / (2 * sd2);
            const float dens = ::exp((double)arg) / std::sqrt(2 * M_PI * sd2);
            const float g = dens / peak;
            const int m = mfInds[j];
            f[m] = g * hi[m] + (1 - g) * lo[m];
            centre += slot;
        }
    }

    std::string getName() { return name; }
    int getNumMF() { return numMF; }

    std::string name;
    int numMF;
    float mfProportion;
    std::vector<int> mfInds;
    std::vector<float> *mfFreq, *mfFreqRelaxed, *mfFreqExcited;
    UpdateType type;
    float minSVVal, maxSVVal, gaussWidth;
};

template <class env>
class StateVariable : public StateVariableBase
{
public:
    StateVariable() {}
    StateVariable(std::string name, UpdateType type, float mfProportion)
        : StateVariableBase(name, type, mfProportion), environment(NULL), getValue(NULL), func_getMFFreq(NULL)
    {
    }

    void initializeGaussian(float minVal, float maxVal, env *e, float (env::*getSV)(), float width = 6.0)
    {
        assert(type == GAUSSIAN);
        minSVVal = minVal;
        maxSVVal = maxVal;
        environment = e;
        getValue = getSV;
        gaussWidth = width;
    }

    void initializeManual(env *e, float *(env::*getMFFreq)())
    {
        environment = e;
        func_getMFFreq = getMFFreq;
    }

    void update() override
    {
        assert(mfFreq && mfFreqRelaxed && mfFreqExcited);
        std::vector<float> &f = *mfFreq, &lo = *mfFreqRelaxed, &hi = *mfFreqExcited;
        if (type == MANUAL) {
            const float *lvl = (environment->*func_getMFFreq)();
            for (size_t j = 0; j < mfInds.size(); j++) {
                const float v = lvl[j];
                assert(v >= 0 && v <= 1);
                const int m = mfInds[j];
                f[m] = 1.5 * v * hi[m] + (1 - v) * lo[m];
            }
        } else if (type == GAUSSIAN) {
            gaussMFAct((environment->*getValue)());
        } else if (type == HIGH_FREQ) {
            for (size_t j = 0; j < mfInds.size(); j++) f[mfInds[j]] = hi[mfInds[j]];
        } else {
            std::cerr << "StateVariable: unknown update type" << std::endl;
            std::abort();
        }
    }

    env *environment;
    float (env::*getValue)();
    float *(env::*func_getMFFreq)();
};

#endif

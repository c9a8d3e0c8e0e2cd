// environment.cpp -- see CBMEnvInclude/environments/environment.hpp.
#include <CBMEnvInclude/environments/environment.hpp>

#include <cassert>
#include <sstream>

Environment::Environment(CRandomSFMT0 *randGen) : randGen(randGen), timestep(0), numMF(0), numNC(0) {}
Environment::~Environment() {}

int Environment::numRequiredMZ() { return 1; }

void Environment::setupMossyFibers(CBMState *simState)
{
    setupMossyFibers((int)simState->sizes().numMF, (int)simState->sizes().numNC);
}

void Environment::setupMossyFibers(int nMF, int nNC)
{
    numMF = nMF;
    numNC = nNC;
    mfFreq.assign(numMF, 0.0f);
    mfFreqRelaxed.assign(numMF, 0.0f);
    mfFreqExcited.assign(numMF, 0.0f);
    // background 1..10 Hz (environment.cpp:26-30)
    const float bgMin = 1, bgMax = 10;
    for (int i = 0; i < numMF; i++) mfFreqRelaxed[i] = randGen->Random() * (bgMax - bgMin) + bgMin;
    // the reference's excited range is (max - max) wide, so every MF gets its
    // minimum, 40 Hz; the draws are still taken (environment.cpp:32-37)
    const float exMin = 40, exMax = 50;
    for (int i = 0; i < numMF; i++) {
        mfFreqExcited[i] = randGen->Random() * (exMax - exMax) + exMin;
        mfExcited.push_back(false);
    }
}

void Environment::setupStateVariables(bool randomizeMFs, std::ofstream &logfile)
{
    assert(!mfFreq.empty());
    for (StateVariableBase *sv : stateVariables) sv->initialize(numMF, &mfFreq, &mfFreqRelaxed, &mfFreqExcited);
    std::vector<int> pool;
    for (int i = numMF - 1; i >= 0; i--) pool.push_back(i);
    for (StateVariableBase *sv : stateVariables) {
        if (randomizeMFs)
            sv->assignRandomMFInds(pool, randGen);
        else
            sv->assignOrderedMFInds(pool);
    }
    if (logfile.is_open()) {
        for (StateVariableBase *sv : stateVariables) {
            logfile << "StateVariable " << sv->name << " " << sv->mfInds.size();
            for (int m : sv->mfInds) logfile << " " << m;
            logfile << "\n";
        }
        for (Microzone *mz : microzones) logfile << "Microzone " << mz->getName() << "\n";
    }
}

float *Environment::getState()
{
    for (size_t i = 0; i < mfFreq.size(); i++) mfFreq[i] = mfExcited[i] ? mfFreqExcited[i] : mfFreqRelaxed[i];
    return mfFreq.data();
}

void Environment::step(CBMSimCore *simCore)
{
    assert((int)microzones.size() == numRequiredMZ());
    for (Microzone *mz : microzones) {
        if (!mz->initialized()) mz->initialize(simCore, numNC);
        mz->update();
    }
    timestep++;
}

bool Environment::terminated() { return false; }

std::vector<std::string> Environment::getMZNames()
{
    std::vector<std::string> names;
    if (!microzones.empty()) {
        for (Microzone *mz : microzones) names.push_back(mz->getName());
    } else {
        for (int i = 0; i < numRequiredMZ(); i++) {
            std::ostringstream o;
            o << "MZ" << i;
            names.push_back(o.str());
        }
    }
    return names;
}

std::vector<std::string> Environment::getStateVariableNames()
{
    std::vector<std::string> names;
    for (StateVariableBase *sv : stateVariables) names.push_back(sv->getName());
    return names;
}

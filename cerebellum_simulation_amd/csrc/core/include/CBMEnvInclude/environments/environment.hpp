// environment.hpp -- base class of the closed-loop environments that drive a
// network (the interface of Matthew_Sim_CBM/includes/environments/
// environment.hpp:15-80, behaviour of src/environments/environment.cpp:11-123):
// it owns the mossy-fibre rate vectors (relaxed 1-10 Hz and excited rates drawn
// once), the state variables that encode the environment into them, and the
// microzones that read the network's output.  The boost serialisation and the
// archive written to the log are not kept.
#ifndef CBM_ENV_ENVIRONMENT_HPP
#define CBM_ENV_ENVIRONMENT_HPP

#include <fstream>
#include <string>
#include <vector>

#include <CBMCoreInclude/interface/cbmsimcore.h>
#include <CBMStateInclude/interfaces/cbmstate.h>
#include <CXXToolsInclude/randGenerators/sfmt.h>
#include <CXXToolsInclude/stdDefinitions/pstdint.h>

#include "../microzone.hpp"
#include "../statevariable.hpp"

class Environment
{
public:
    explicit Environment(CRandomSFMT0 *randGen);
    virtual ~Environment();

    virtual int numRequiredMZ();

    // Size the MF vectors from the state and draw the relaxed and excited
    // rates: numMF draws of U[1,10) Hz, then numMF draws for the excited rates
    // (which the reference's formula makes 40 Hz for every MF).
    virtual void setupMossyFibers(CBMState *simState);
    // The same from the counts (no state object).
    void setupMossyFibers(int numMF, int numNC);

    // MF rates (Hz) for the next step.
    virtual float *getState();

    // Advance one step after the network's step; may deliver errors.
    virtual void step(CBMSimCore *simCore);

    virtual bool terminated();

    virtual std::vector<std::string> getMZNames();
    virtual std::vector<std::string> getStateVariableNames();
    std::vector<StateVariableBase *> getStateVariables() { return stateVariables; }
    std::vector<Microzone *> getMicrozones() { return microzones; }

    int getNumMF() const { return numMF; }
    int getNumNC() const { return numNC; }
    const std::vector<float> &getRelaxedFreqs() const { return mfFreqRelaxed; }
    const std::vector<float> &getExcitedFreqs() const { return mfFreqExcited; }

    std::vector<bool> mfExcited;

protected:
    std::vector<float> mfFreq, mfFreqRelaxed, mfFreqExcited;
    CRandomSFMT0 *randGen;
    int timestep;
    int numMF, numNC;
    std::vector<Microzone *> microzones;
    std::vector<StateVariableBase *> stateVariables;

    // Size each state variable's MF group and assign its MFs, randomly (in
    // state-variable order, from the pool numMF-1 .. 0) or in order; writes
    // one line per variable to the log when it is open.
    void setupStateVariables(bool randomizeMFs, std::ofstream &logfile);
};

#endif

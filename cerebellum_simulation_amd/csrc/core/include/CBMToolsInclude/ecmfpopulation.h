// ecmfpopulation.h -- MF firing rates of the eyelid-conditioning driver:
// background, CS-tonic and CS-phasic rates per MF
// (CBM_TOOLS_LIB_DEV/CBMToolsInclude/ecmfpopulation.h, used by
// Eyelid_Sim_CBM_Dev/src/ectrial/ecmanagementdelay.cpp:20,42,61,88-92).
#ifndef CBM_ECMFPOPULATION_H
#define CBM_ECMFPOPULATION_H

#include <fstream>
#include <vector>

#include <CXXToolsInclude/stdDefinitions/pstdint.h>
#include <CXXToolsInclude/randGenerators/randomc.h>
#include <CXXToolsInclude/randGenerators/sfmt.h>

#include <CBMDataInclude/interfaces/ectrialsdata.h>
#include <CBMDataInclude/peristimhist/peristimhistfloat.h>

class ECMFPopulation
{
public:
    // Rates estimated from the "mf" histogram of recorded trials: the mean
    // rate before the CS, in its first bins and in its last bins.
    ECMFPopulation(ECTrialsData *data);
    // Seeded rates: background MFs, a fraction of context MFs, and fractions
    // of CS-tonic and CS-phasic MFs, each with rates drawn uniformly in Hz
    // from the generator seeded with randSeed.
    ECMFPopulation(int numMF, int randSeed, float fracCSTMF, float fracCSPMF, float fracCtxtMF, float bgFreqMin,
                   float csBGFreqMin, float ctxtFreqMin, float csTFreqMin, float csPFreqMin, float bgFreqMax,
                   float csBGFreqMax, float ctxtFreqMax, float csTFreqMax, float csPFreqMax);
    // The binary file writeToFile writes.
    ECMFPopulation(std::fstream &infile);
    ~ECMFPopulation();

    // numMF (u32), then the background, CS-tonic and CS-phasic rates (f32 each).
    void writeToFile(std::fstream &outfile);

    float *getMFBG();
    float *getMFInCSTonic();
    float *getMFFreqInCSPhasic();

private:
    ct_uint32_t numMF;
    std::vector<float> bg, tonic, phasic;
};

#endif

// eyelidout.h -- eyelid traces per trial, smoothed as recorded
// (CBM_DATA_LIB_DEV/CBMDataInclude/outdata/eyelidout.h).
#ifndef CBM_EYELIDOUT_H
#define CBM_EYELIDOUT_H

#include <fstream>
#include <vector>

#include <CXXToolsInclude/stdDefinitions/pstdint.h>

class EyelidOut
{
public:
    EyelidOut(unsigned int numTrials, unsigned int msPerTimeStep, unsigned int numTimeStepSmooth,
              unsigned int msPerTrial);
    EyelidOut(std::fstream &infile);
    ~EyelidOut();

    void writeData(std::fstream &outfile);
    bool updateData(float data);

    ct_uint32_t getNumTrials();
    ct_uint32_t getNumPointsPerTrial();
    ct_uint32_t getMSPerPoint();
    std::vector<float> getEyelidData(unsigned int trialN);

private:
    ct_uint32_t nTrials, msPerTS, numTSSmooth, msPerTrial, tsPerTrial, curTrialN, curTS;
    std::vector<float> rawData;  // [trial][point]
};

#endif

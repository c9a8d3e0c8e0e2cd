// peristimhistfloat.h -- peri-stimulus histograms, FP32 bins
// (CBM_DATA_LIB_DEV/CBMDataInclude/peristimhist/peristimhistfloat.h).
#ifndef CBM_PERISTIMHISTFLOAT_H
#define CBM_PERISTIMHISTFLOAT_H

#include <fstream>
#include <vector>

#include <CXXToolsInclude/stdDefinitions/pstdint.h>

class PeriStimHistFloat
{
public:
    PeriStimHistFloat(int numSteps, int stepsPerBin, int numCells);
    PeriStimHistFloat(std::fstream &infile);
    ~PeriStimHistFloat();

    void writeData(std::fstream &outfile);
    void update(const ct_uint8_t *condition);  // spikes: adds 1 per nonzero byte
    void update(const ct_uint32_t *counts);
    void update(const float *input);

    std::vector<float> getCellPSH(int cellN);
    std::vector<std::vector<float> > getPopPSHs();
    ct_uint32_t getNumCells();
    ct_uint32_t getNumBins();
    ct_uint32_t getBinWidthInMS();
    ct_uint32_t getNumTrials();

private:
    float *bin(int b) { return &cells[(size_t)b * nCells]; }
    void advance();
    ct_uint32_t nCells, numTS, tsPerBin, numBins, curTS, curTrialN;
    std::vector<float> cells;  // [bin][cell]
};

#endif

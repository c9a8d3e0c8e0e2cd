// rawuintdata.h -- raw per-step uint32 rows per trial
// (CBM_DATA_LIB_DEV/CBMDataInclude/outdata/rawuintdata.h).
#ifndef CBM_RAWUINTDATA_H
#define CBM_RAWUINTDATA_H

#include <fstream>
#include <string>
#include <vector>

#include <CXXToolsInclude/stdDefinitions/pstdint.h>

class RawUIntData
{
public:
    RawUIntData(std::string label, unsigned int numRows, unsigned int numTrials, unsigned int numPointsPerTrial);
    RawUIntData(std::fstream &infile);
    ~RawUIntData();

    void writeData(std::fstream &outfile);
    bool updateData(const ct_uint32_t *input);

    unsigned int getNumTrials();
    unsigned int getNumDataRows();
    unsigned int getNumPointsPerTrial();
    const ct_uint32_t *getTrial(unsigned int trialN) { return &data[(size_t)trialN * numPointsPerTrial * numDataRows]; }

private:
    std::string dataLabel;
    unsigned int numDataRows, maxNumTrials, numPointsPerTrial, curTrialN, curPoint;
    std::vector<ct_uint32_t> data;  // [trial][point][row]
};

#endif

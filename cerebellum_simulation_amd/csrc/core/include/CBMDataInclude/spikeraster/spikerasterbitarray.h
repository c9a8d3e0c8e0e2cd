// spikerasterbitarray.h -- per-trial rasters of 32-step spike words
// (CBM_DATA_LIB_DEV/CBMDataInclude/spikeraster/spikerasterbitarray.h).
#ifndef CBM_SPIKERASTERBITARRAY_H
#define CBM_SPIKERASTERBITARRAY_H

#include <CXXToolsInclude/stdDefinitions/pstdint.h>

#include "../interfaces/ispikeraster.h"

class SpikeRasterBitArray : public ISpikeRaster
{
public:
    SpikeRasterBitArray(std::string cellType, unsigned int numCells, unsigned int numTrials, unsigned int msPerTrial,
                        unsigned int msPerTimeStep);
    SpikeRasterBitArray(std::fstream &infile);
    ~SpikeRasterBitArray();

    void writeData(std::fstream &outfile);
    // one word per cell: the cell's spike register after 32 more steps
    bool updateRaster(const ct_uint32_t *input);
    std::vector<std::vector<int> > getCellSpikeTimes(unsigned int cellN, int offset);
    std::vector<int> getCellSpikeTimes(unsigned int cellN, unsigned int trialN, int offset);
    std::vector<bool> getPopSpikesAtTS(unsigned int timeStep, unsigned int trialN);
    unsigned int getNumTrials();
    unsigned int getUpdateInterval();

private:
    void allocateMemory();
    unsigned int msPerTS, curTrialN, curTS, curBufN, tsPerTrial, bufsPerTrial;
    std::vector<std::vector<ct_uint32_t> > raster;  // [trial][buf * nCells + cell]
};

#endif

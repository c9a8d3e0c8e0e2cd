// ispikeraster.h -- spike raster interface of the trial data files
// (CBM_DATA_LIB_DEV/CBMDataInclude/interfaces/ispikeraster.h).
#ifndef CBM_ISPIKERASTER_H
#define CBM_ISPIKERASTER_H

#include <fstream>
#include <string>
#include <vector>

class ISpikeRaster
{
public:
    ISpikeRaster(std::string cellType, unsigned int numCells, unsigned int numTrials, unsigned int msPerTrial);
    ISpikeRaster(std::fstream &infile);
    virtual ~ISpikeRaster();

    virtual void writeData(std::fstream &outfile);
    virtual bool updateRaster(const unsigned int *input) = 0;
    // spike times (ms from `offset`) of one cell in every recorded trial / one trial
    virtual std::vector<std::vector<int> > getCellSpikeTimes(unsigned int cellN, int offset) = 0;
    virtual std::vector<int> getCellSpikeTimes(unsigned int cellN, unsigned int trialN, int offset) = 0;
    virtual std::vector<bool> getPopSpikesAtTS(unsigned int timeStep, unsigned int trialN) = 0;
    virtual unsigned int getUpdateInterval() = 0;

    std::string getCellType();
    unsigned int getNumCells();
    virtual unsigned int getNumTrials();
    unsigned int getMSPerTrial();

protected:
    std::string cellT;
    unsigned int nCells, nTrials, msPerTrial;
};

#endif

// ectrialsdata.cpp -- the trial data library of the eyelid drivers:
// ECTrialsData (CBM_DATA_LIB_DEV/src/interfaces/ectrialsdata.cpp:10-263),
// PeriStimHistFloat (src/peristimhist/peristimhistfloat.cpp), SpikeRasterBitArray
// and ISpikeRaster (src/spikeraster/spikerasterbitarray.cpp,
// src/interfaces/ispikeraster.cpp), EyelidOut (src/outdata/eyelidout.cpp) and
// RawUIntData (src/outdata/rawuintdata.cpp).
//
// The file layout is theirs: each object writes a line of space-separated
// unsigned integers (its sizes and write positions) and then its raw array;
// ECTrialsData writes its own line, the eyelid data, then the histograms,
// rasters and raw rows each preceded by a count line and one name line per
// object, in name order.  A reader skips the one character after each text
// line before the raw bytes.
#include <CBMDataInclude/interfaces/ectrialsdata.h>

#include <cmath>
#include <stdexcept>

namespace {

template <typename T>
void rawRW(bool read, std::fstream &f, std::vector<T> &v)
{
    if (v.empty()) return;
    if (read)
        f.read(reinterpret_cast<char *>(v.data()), (std::streamsize)(v.size() * sizeof(T)));
    else
        f.write(reinterpret_cast<const char *>(v.data()), (std::streamsize)(v.size() * sizeof(T)));
    if (!f) throw std::runtime_error("trial data: stream ended inside an array");
}

}  // namespace

// ---------------------------------------------------------------- ISpikeRaster
ISpikeRaster::ISpikeRaster(std::string cellType, unsigned int numCells, unsigned int numTrials,
                           unsigned int msPerTrial)
    : cellT(cellType), nCells(numCells), nTrials(numTrials), msPerTrial(msPerTrial)
{
}

ISpikeRaster::ISpikeRaster(std::fstream &infile)
{
    infile >> cellT >> nCells >> nTrials >> msPerTrial;
    infile.seekg(1, std::ios::cur);
}

ISpikeRaster::~ISpikeRaster() {}

void ISpikeRaster::writeData(std::fstream &outfile)
{
    outfile << cellT << " " << nCells << " " << nTrials << " " << msPerTrial << " ";
}

std::string ISpikeRaster::getCellType() { return cellT; }
unsigned int ISpikeRaster::getNumCells() { return nCells; }
unsigned int ISpikeRaster::getNumTrials() { return nTrials; }
unsigned int ISpikeRaster::getMSPerTrial() { return msPerTrial; }

// ---------------------------------------------------------------- SpikeRasterBitArray
SpikeRasterBitArray::SpikeRasterBitArray(std::string cellType, unsigned int numCells, unsigned int numTrials,
                                         unsigned int msPerTrial, unsigned int msPerTimeStep)
    : ISpikeRaster(cellType, numCells, numTrials, msPerTrial), msPerTS(msPerTimeStep), curTrialN(0), curTS(0),
      curBufN(0)
{
    tsPerTrial = msPerTrial / msPerTS;
    bufsPerTrial = tsPerTrial / 32;  // whole 32-step words only
    allocateMemory();
}

SpikeRasterBitArray::SpikeRasterBitArray(std::fstream &infile) : ISpikeRaster(infile)
{
    infile >> msPerTS >> curTrialN >> curTS >> curBufN >> tsPerTrial >> bufsPerTrial;
    infile.seekg(1, std::ios::cur);
    allocateMemory();
    for (auto &t : raster) rawRW(true, infile, t);
}

SpikeRasterBitArray::~SpikeRasterBitArray() {}

void SpikeRasterBitArray::allocateMemory()
{
    raster.assign(nTrials, std::vector<ct_uint32_t>((size_t)bufsPerTrial * nCells, 0));
}

void SpikeRasterBitArray::writeData(std::fstream &outfile)
{
    ISpikeRaster::writeData(outfile);
    outfile << msPerTS << " " << curTrialN << " " << curTS << " " << curBufN << " " << tsPerTrial << " "
            << bufsPerTrial << std::endl;
    for (auto &t : raster) rawRW(false, outfile, t);
}

bool SpikeRasterBitArray::updateRaster(const ct_uint32_t *input)
{
    if (curTrialN >= nTrials) return false;
    ct_uint32_t *row = &raster[curTrialN][(size_t)curBufN * nCells];
    for (unsigned int i = 0; i < nCells; i++) row[i] = input[i];
    curBufN++;
    curTS += 32;
    if (curBufN >= bufsPerTrial) {
        curTrialN++;
        curBufN = 0;
        curTS = 0;
    }
    return true;
}

std::vector<std::vector<int> > SpikeRasterBitArray::getCellSpikeTimes(unsigned int cellN, int offset)
{
    std::vector<std::vector<int> > out;
    for (unsigned int t = 0; t < curTrialN; t++) out.push_back(getCellSpikeTimes(cellN, t, offset));
    return out;
}

// bit 31 of a word is the earliest of its 32 steps
std::vector<int> SpikeRasterBitArray::getCellSpikeTimes(unsigned int cellN, unsigned int trialN, int offset)
{
    std::vector<int> out;
    int time = offset;
    for (unsigned int b = 0; b < bufsPerTrial; b++) {
        ct_uint32_t w = raster[trialN][(size_t)b * nCells + cellN];
        for (int j = 0; j < 32; j++) {
            if (w & 0x80000000u) out.push_back(time);
            time += (int)msPerTS;
            w <<= 1;
        }
    }
    return out;
}

std::vector<bool> SpikeRasterBitArray::getPopSpikesAtTS(unsigned int timeStep, unsigned int trialN)
{
    std::vector<bool> aps(nCells, false);
    const unsigned int b = timeStep / 32, bit = timeStep % 32;
    if (b >= bufsPerTrial) return aps;
    const ct_uint32_t mask = 0x80000000u >> bit;
    for (unsigned int i = 0; i < nCells; i++) aps[i] = (raster[trialN][(size_t)b * nCells + i] & mask) != 0;
    return aps;
}

unsigned int SpikeRasterBitArray::getNumTrials() { return curTrialN; }
unsigned int SpikeRasterBitArray::getUpdateInterval() { return 32; }

// ---------------------------------------------------------------- PeriStimHistFloat
PeriStimHistFloat::PeriStimHistFloat(int numSteps, int stepsPerBin, int numCells)
    : nCells((ct_uint32_t)numCells), numTS((ct_uint32_t)numSteps), tsPerBin((ct_uint32_t)stepsPerBin), curTS(0),
      curTrialN(0)
{
    numBins = (ct_uint32_t)std::ceil(numTS / (float)tsPerBin);
    cells.assign((size_t)numBins * nCells, 0.0f);
}

PeriStimHistFloat::PeriStimHistFloat(std::fstream &infile)
{
    infile >> nCells >> numTS >> tsPerBin >> numBins >> curTS >> curTrialN;
    infile.seekg(1, std::ios::cur);
    cells.assign((size_t)numBins * nCells, 0.0f);
    rawRW(true, infile, cells);
}

PeriStimHistFloat::~PeriStimHistFloat() {}

void PeriStimHistFloat::writeData(std::fstream &outfile)
{
    outfile << nCells << " " << numTS << " " << tsPerBin << " " << numBins << " " << curTS << " " << curTrialN
            << std::endl;
    rawRW(false, outfile, cells);
}

void PeriStimHistFloat::advance()
{
    curTS++;
    if (curTS >= numTS) {
        curTS = 0;
        curTrialN++;
    }
}

void PeriStimHistFloat::update(const ct_uint8_t *condition)
{
    float *b = bin((int)(curTS / tsPerBin));
    for (ct_uint32_t i = 0; i < nCells; i++) b[i] += (float)(condition[i] > 0);
    advance();
}

void PeriStimHistFloat::update(const ct_uint32_t *counts)
{
    float *b = bin((int)(curTS / tsPerBin));
    for (ct_uint32_t i = 0; i < nCells; i++) b[i] += (float)counts[i];
    advance();
}

void PeriStimHistFloat::update(const float *input)
{
    float *b = bin((int)(curTS / tsPerBin));
    for (ct_uint32_t i = 0; i < nCells; i++) b[i] += input[i];
    advance();
}

std::vector<float> PeriStimHistFloat::getCellPSH(int cellN)
{
    std::vector<float> v(numBins);
    for (ct_uint32_t b = 0; b < numBins; b++) v[b] = cells[(size_t)b * nCells + cellN];
    return v;
}

std::vector<std::vector<float> > PeriStimHistFloat::getPopPSHs()
{
    std::vector<std::vector<float> > v(nCells);
    for (ct_uint32_t i = 0; i < nCells; i++) v[i] = getCellPSH((int)i);
    return v;
}

ct_uint32_t PeriStimHistFloat::getNumCells() { return nCells; }
ct_uint32_t PeriStimHistFloat::getNumBins() { return numBins; }
ct_uint32_t PeriStimHistFloat::getBinWidthInMS() { return tsPerBin; }
ct_uint32_t PeriStimHistFloat::getNumTrials() { return curTrialN; }

// ---------------------------------------------------------------- EyelidOut
EyelidOut::EyelidOut(unsigned int numTrials, unsigned int msPerTimeStep, unsigned int numTimeStepSmooth,
                     unsigned int msPerTrial)
    : nTrials(numTrials), msPerTS(msPerTimeStep), numTSSmooth(numTimeStepSmooth), msPerTrial(msPerTrial),
      curTrialN(0), curTS(0)
{
    tsPerTrial = msPerTrial / msPerTS;
    rawData.assign((size_t)nTrials * tsPerTrial, 0.0f);
}

EyelidOut::EyelidOut(std::fstream &infile)
{
    infile >> nTrials >> msPerTS >> numTSSmooth >> msPerTrial >> tsPerTrial >> curTrialN >> curTS;
    rawData.assign((size_t)nTrials * tsPerTrial, 0.0f);
    infile.seekg(1, std::ios::cur);
    rawRW(true, infile, rawData);
}

EyelidOut::~EyelidOut() {}

void EyelidOut::writeData(std::fstream &outfile)
{
    outfile << nTrials << " " << msPerTS << " " << numTSSmooth << " " << msPerTrial << " " << tsPerTrial << " "
            << curTrialN << " " << curTS << std::endl;
    rawRW(false, outfile, rawData);
}

// The stored value is the mean of this input and the numTSSmooth-1 values
// stored before it in the trial -- which are themselves smoothed
// (eyelidout.cpp:62-93); the first numTSSmooth-1 points are stored raw.
bool EyelidOut::updateData(float data)
{
    if (curTrialN >= nTrials) return false;
    float *trial = &rawData[(size_t)curTrialN * tsPerTrial];
    if (curTS < numTSSmooth - 1) {
        trial[curTS] = data;
    } else {
        float avg = 0;
        for (int j = -((int)numTSSmooth - 1); j < 0; j++) avg += trial[(int)curTS + j];
        avg += data;
        avg = avg / numTSSmooth;
        trial[curTS] = avg;
    }
    curTS++;
    if (curTS >= tsPerTrial) {
        curTrialN++;
        curTS = 0;
    }
    return true;
}

ct_uint32_t EyelidOut::getNumTrials() { return curTrialN; }
ct_uint32_t EyelidOut::getNumPointsPerTrial() { return tsPerTrial; }
ct_uint32_t EyelidOut::getMSPerPoint() { return msPerTS; }

std::vector<float> EyelidOut::getEyelidData(unsigned int trialN)
{
    return std::vector<float>(rawData.begin() + (size_t)trialN * tsPerTrial,
                              rawData.begin() + (size_t)(trialN + 1) * tsPerTrial);
}

// ---------------------------------------------------------------- RawUIntData
RawUIntData::RawUIntData(std::string label, unsigned int numRows, unsigned int numTrials,
                         unsigned int numPointsPerTrial)
    : dataLabel(label), numDataRows(numRows), maxNumTrials(numTrials), numPointsPerTrial(numPointsPerTrial),
      curTrialN(0), curPoint(0)
{
    data.assign((size_t)maxNumTrials * numPointsPerTrial * numDataRows, 0u);
}

// The reference reads the arrays before allocating them (rawuintdata.cpp:21-32);
// here they are allocated from the header first.
RawUIntData::RawUIntData(std::fstream &infile)
{
    infile >> dataLabel >> numDataRows >> maxNumTrials >> numPointsPerTrial >> curTrialN >> curPoint;
    infile.seekg(1, std::ios::cur);
    data.assign((size_t)maxNumTrials * numPointsPerTrial * numDataRows, 0u);
    rawRW(true, infile, data);
}

RawUIntData::~RawUIntData() {}

void RawUIntData::writeData(std::fstream &outfile)
{
    outfile << dataLabel << " " << numDataRows << " " << maxNumTrials << " " << numPointsPerTrial << " "
            << curTrialN << " " << curPoint << std::endl;
    rawRW(false, outfile, data);
}

bool RawUIntData::updateData(const ct_uint32_t *input)
{
    if (curTrialN >= maxNumTrials) return false;
    ct_uint32_t *row = &data[((size_t)curTrialN * numPointsPerTrial + curPoint) * numDataRows];
    for (unsigned int i = 0; i < numDataRows; i++) row[i] = input[i];
    curPoint++;
    if (curPoint >= numPointsPerTrial) {
        curTrialN++;
        curPoint = 0;
    }
    return true;
}

unsigned int RawUIntData::getNumTrials() { return maxNumTrials; }
unsigned int RawUIntData::getNumDataRows() { return numDataRows; }
unsigned int RawUIntData::getNumPointsPerTrial() { return numPointsPerTrial; }

// ---------------------------------------------------------------- ECTrialsData
ECTrialsData::ECTrialsData(unsigned int msPreCS, unsigned int msCS, unsigned int msPostCS, unsigned int msPerTimeStep,
                           unsigned int numTrials, std::map<std::string, PSHParams> periStimH,
                           std::map<std::string, RasterParams> cellRasters,
                           std::map<std::string, RawUIntParams> rawUInt, EyelidOutParams eyelidout)
    : msPreCS(msPreCS), msCS(msCS), msPostCS(msPostCS), msPerTS(msPerTimeStep), numTrials(numTrials)
{
    msTotal = msPreCS + msCS + msPostCS;
    tsTotal = msTotal / msPerTS;
    eyelidOut = new EyelidOut(numTrials, msPerTS, eyelidout.numTimeStepSmooth, msTotal);
    for (auto &kv : rawUInt) uintData[kv.first] = new RawUIntData(kv.first, kv.second.numRows, numTrials, tsTotal);
    for (auto &kv : periStimH)
        pshs[kv.first] = new PeriStimHistFloat((int)tsTotal, (int)kv.second.numTimeStepsPerBin, (int)kv.second.numCells);
    for (auto &kv : cellRasters)
        rasters[kv.first] = new SpikeRasterBitArray(kv.first, kv.second.numCells, numTrials, msTotal, msPerTS);
}

ECTrialsData::ECTrialsData(std::fstream &infile)
{
    infile >> msPreCS >> msCS >> msPostCS >> msTotal >> msPerTS >> tsTotal >> numTrials;
    infile.seekg(1, std::ios::cur);
    eyelidOut = new EyelidOut(infile);
    ct_uint32_t n = 0;
    infile >> n;
    for (ct_uint32_t i = 0; i < n; i++) {
        std::string name;
        infile >> name;
        infile.seekg(1, std::ios::cur);
        pshs[name] = new PeriStimHistFloat(infile);
    }
    n = 0;
    infile >> n;
    for (ct_uint32_t i = 0; i < n; i++) {
        std::string name;
        infile >> name;
        infile.seekg(1, std::ios::cur);
        rasters[name] = new SpikeRasterBitArray(infile);
    }
    n = 0;
    infile >> n;
    for (ct_uint32_t i = 0; i < n; i++) {
        std::string name;
        infile >> name;
        infile.seekg(1, std::ios::cur);
        uintData[name] = new RawUIntData(infile);
    }
}

ECTrialsData::~ECTrialsData()
{
    for (auto &kv : pshs) delete kv.second;
    for (auto &kv : rasters) delete kv.second;
    for (auto &kv : uintData) delete kv.second;
    delete eyelidOut;
}

void ECTrialsData::writeData(std::fstream &outfile)
{
    outfile << msPreCS << " " << msCS << " " << msPostCS << " " << msTotal << " " << msPerTS << " " << tsTotal << " "
            << numTrials << std::endl;
    eyelidOut->writeData(outfile);
    outfile << pshs.size() << std::endl;
    for (auto &kv : pshs) {
        outfile << kv.first << std::endl;
        kv.second->writeData(outfile);
    }
    outfile << rasters.size() << std::endl;
    for (auto &kv : rasters) {
        outfile << kv.first << std::endl;
        kv.second->writeData(outfile);
    }
    outfile << uintData.size() << std::endl;
    for (auto &kv : uintData) {
        outfile << kv.first << std::endl;
        kv.second->writeData(outfile);
    }
}

void ECTrialsData::updatePSH(std::string cellT, const ct_uint8_t *aps) { pshs.at(cellT)->update(aps); }
void ECTrialsData::updatePSH(std::string cellT, const ct_uint32_t *counts) { pshs.at(cellT)->update(counts); }
void ECTrialsData::updatePSH(std::string cellT, const float *input) { pshs.at(cellT)->update(input); }
void ECTrialsData::updateRaster(std::string cellT, const ct_uint32_t *apBufs) { rasters.at(cellT)->updateRaster(apBufs); }
void ECTrialsData::updateRawUInt(std::string label, const ct_uint32_t *input) { uintData.at(label)->updateData(input); }
void ECTrialsData::updateEyelid(float data) { eyelidOut->updateData(data); }

unsigned int ECTrialsData::getTSPerRasterUpdate()
{
    return rasters.empty() ? 0u : rasters.begin()->second->getUpdateInterval();
}

ISpikeRaster *ECTrialsData::getRaster(std::string cellT) { return rasters.at(cellT); }
PeriStimHistFloat *ECTrialsData::getPSH(std::string cellT) { return pshs.at(cellT); }
RawUIntData *ECTrialsData::getUIntData(std::string label) { return uintData.at(label); }
EyelidOut *ECTrialsData::getEyelidData() { return eyelidOut; }

unsigned int ECTrialsData::getMSPreTrial() { return msPreCS; }
unsigned int ECTrialsData::getMSTrial() { return msCS; }
unsigned int ECTrialsData::getMSPostTrial() { return msPostCS; }
unsigned int ECTrialsData::getMSPerTimeStep() { return msPerTS; }
unsigned int ECTrialsData::getMSTotal() { return msTotal; }
unsigned int ECTrialsData::getTimeStepsTotal() { return tsTotal; }
unsigned int ECTrialsData::getMaxNumTrials() { return numTrials; }

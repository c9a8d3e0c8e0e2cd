// ecmfpopulation.cpp -- MF rates of the eyelid driver
// (CBM_TOOLS_LIB_DEV/src/ecmfpopulation.cpp:11-190).  The seeded constructor
// keeps the reference's order of generator draws, so the same seed gives the
// same rates: one Random() per MF for its background rate, then the CS-tonic,
// CS-phasic and context MFs picked by IRandom(0, numMF-1) with redraws on a
// taken MF, then the picked MFs' rates in MF order.
#include <CBMToolsInclude/ecmfpopulation.h>

#include <stdexcept>

ECMFPopulation::ECMFPopulation(ECTrialsData *data)
{
    PeriStimHistFloat *psh = data->getPSH("mf");
    if (!psh) throw std::runtime_error("ECMFPopulation: the trial data has no \"mf\" histogram");
    numMF = psh->getNumCells();
    const int msPerBin = (int)psh->getBinWidthInMS();
    const int numTrials = (int)psh->getNumTrials();
    // spikes in a bin summed over trials -> Hz
    const float scaling = numTrials * (msPerBin / 1000.0);
    const int pre = (int)data->getMSPreTrial() / msPerBin;
    const int trial = (int)data->getMSTrial() / msPerBin;
    bg.resize(numMF);
    tonic.resize(numMF);
    phasic.resize(numMF);
    auto sum = [](const std::vector<float> &h, int from, int to) {
        float s = 0;
        for (int j = from; j < to; j++) s += h[j];
        return s;
    };
    for (ct_uint32_t i = 0; i < numMF; i++) {
        const std::vector<float> h = psh->getCellPSH((int)i);
        bg[i] = sum(h, 2, pre - 2) / ((pre - 4) * scaling);                  // before the CS, edges left out
        phasic[i] = sum(h, pre + 1, pre + 3) / (2 * scaling);                 // bins 1-2 of the CS
        tonic[i] = sum(h, pre + trial - 6, pre + trial - 1) / (5 * scaling);  // the CS's last bins but one
    }
}

ECMFPopulation::ECMFPopulation(int numMF, int randSeed, float fracCSTMF, float fracCSPMF, float fracCtxtMF,
                               float bgFreqMin, float csBGFreqMin, float ctxtFreqMin, float csTFreqMin,
                               float csPFreqMin, float bgFreqMax, float csBGFreqMax, float ctxtFreqMax,
                               float csTFreqMax, float csPFreqMax)
    : numMF((ct_uint32_t)numMF), bg(numMF), tonic(numMF), phasic(numMF)
{
    CRandomSFMT0 gen(randSeed);
    enum Kind : unsigned char { kBG = 0, kCSTonic = 1, kCSPhasic = 2, kContext = 3 };
    std::vector<unsigned char> kind(numMF, kBG);
    for (int i = 0; i < numMF; i++) {
        bg[i] = gen.Random() * (bgFreqMax - bgFreqMin) + bgFreqMin;
        tonic[i] = phasic[i] = bg[i];
    }
    // pick `count` MFs of kind k among the MFs still free of the kinds before it
    auto pick = [&](int count, Kind k) {
        for (int c = 0; c < count; c++)
            for (;;) {
                const int m = gen.IRandom(0, numMF - 1);
                if (kind[m] != kBG) continue;
                kind[m] = k;
                break;
            }
    };
    pick((int)(fracCSTMF * numMF), kCSTonic);
    pick((int)(fracCSPMF * numMF), kCSPhasic);
    pick((int)(fracCtxtMF * numMF), kContext);
    for (int i = 0; i < numMF; i++) {
        switch (kind[i]) {
        case kContext:
            bg[i] = gen.Random() * (ctxtFreqMax - ctxtFreqMin) + ctxtFreqMin;
            tonic[i] = phasic[i] = bg[i];
            break;
        case kCSTonic:
            bg[i] = gen.Random() * (csBGFreqMax - csBGFreqMin) + csBGFreqMin;
            tonic[i] = gen.Random() * (csTFreqMax - csTFreqMin) + csTFreqMin;
            phasic[i] = tonic[i];
            break;
        case kCSPhasic:
            bg[i] = gen.Random() * (csBGFreqMax - csBGFreqMin) + csBGFreqMin;
            tonic[i] = bg[i];
            phasic[i] = gen.Random() * (csPFreqMax - csPFreqMin) + csPFreqMin;
            break;
        default:
            break;
        }
    }
}

ECMFPopulation::ECMFPopulation(std::fstream &infile)
{
    infile.read(reinterpret_cast<char *>(&numMF), sizeof(numMF));
    bg.resize(numMF);
    tonic.resize(numMF);
    phasic.resize(numMF);
    infile.read(reinterpret_cast<char *>(bg.data()), numMF * sizeof(float));
    infile.read(reinterpret_cast<char *>(tonic.data()), numMF * sizeof(float));
    infile.read(reinterpret_cast<char *>(phasic.data()), numMF * sizeof(float));
}

ECMFPopulation::~ECMFPopulation() {}

void ECMFPopulation::writeToFile(std::fstream &outfile)
{
    outfile.write(reinterpret_cast<const char *>(&numMF), sizeof(numMF));
    outfile.write(reinterpret_cast<const char *>(bg.data()), numMF * sizeof(float));
    outfile.write(reinterpret_cast<const char *>(tonic.data()), numMF * sizeof(float));
    outfile.write(reinterpret_cast<const char *>(phasic.data()), numMF * sizeof(float));
}

float *ECMFPopulation::getMFBG() { return bg.data(); }
float *ECMFPopulation::getMFInCSTonic() { return tonic.data(); }
float *ECMFPopulation::getMFFreqInCSPhasic() { return phasic.data(); }

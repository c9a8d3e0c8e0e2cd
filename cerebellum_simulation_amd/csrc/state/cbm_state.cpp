// cbm_state.cpp -- network-state library: parameter files, derived constants,
// connectivity generation and initial activity state (include/cbm_state.h).
//
// Everything here runs once, at construction, on the host.  Arrays are kept
// in the CBMState layout (cell-major [cell][slot], unused slots UINT32_MAX) and
// addressed with the field ids of cbm_hip.h.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "../../../include/cbm_state.h"
#include "sfmt.hpp"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

constexpr int kMaxZones = 8;

// ---------------------------------------------------------------- params
// ActivityParams::ActivityParams + updateParams
// (CBM_STATE_LIB_DEV/src/params/activityparams.cpp:12-42, 436-636).  The
// reference stores every value as float and derives in float; exp() of a float
// expression binds to the float overload.
struct ActMap {
    std::map<std::string, float> m;
    float operator[](const char *k) { return m[k]; }
};

int parseAct(const char *text, ActMap &am)
{
    std::istringstream in(text ? text : "");
    std::string key;
    float val;
    while (in >> key >> val) {
        if (key == "activityParamEnd") return CBM_OK;
        am.m[key] = val;
    }
    return fail(CBM_E_ARG, "actParams text has no activityParamEnd line");
}

void deriveAct(ActMap &m, cbm_act_params &p)
{
    using std::exp;
    std::memset(&p, 0, sizeof(p));
    p.msPerTimeStep = m["msPerTimeStep"];
    float msPerHistBinMF = m["msPerHistBinMF"];
    p.numTSinMFHist = (uint32_t)(msPerHistBinMF / p.msPerTimeStep);
    if (m.m.find("coupleRiRjRatioGO") == m.m.end()) m.m["coupleRiRjRatioGO"] = 0;
    if (m.m.find("goGABAGOGOSynRecTau") == m.m.end()) m.m["goGABAGOGOSynRecTau"] = 1;
    if (m.m.find("goGABAGOGOSynDepF") == m.m.end()) m.m["goGABAGOGOSynDepF"] = 1;
    const float ms = p.msPerTimeStep;

    p.eLeakGO = m["eLeakGO"];
    p.eGABAGO = m["eGABAGO"];
    p.threshMaxGO = m["threshMaxGO"];
    p.threshRestGO = m["threshBaseGO"];
    p.gIncMFtoGO = m["gMFIncGO"];
    p.gIncGRtoGO = m["gGRIncGO"];
    p.gGABAIncGOtoGO = m["gGOIncGO"];
    p.coupleRiRjRatioGO = m["coupleRiRjRatioGO"];
    p.gMGluRScaleGRtoGO = m["gMGluRScaleGO"];
    p.gMGluRIncScaleGO = m["gMGluRIncScaleGO"];
    p.mGluRScaleGO = m["mGluRScaleGO"];
    p.gluScaleGO = m["gluScaleGO"];
    p.gLeakGO = m["rawGLeakGO"] / (6 - ms);
    p.gDecMFtoGO = exp(-ms / m["gMFDecayTGO"]);
    p.gDecGRtoGO = exp(-ms / m["gGRDecayTGO"]);
    p.gGABADecGOtoGO = exp(-ms / m["gGODecayTGO"]);
    p.goGABAGOGOSynRec = 1 - exp(-ms / m["goGABAGOGOSynRecTau"]);
    p.goGABAGOGOSynDepF = m["goGABAGOGOSynDepF"];
    p.mGluRDecayGO = m["mGluRDecayGO"];
    p.gMGluRIncDecayGO = m["gMGluRIncDecayGO"];
    p.gMGluRDecGRtoGO = m["gMGluRDecayGO"];
    p.gluDecayGO = m["gluDecayGO"];
    p.threshDecGO = 1 - exp(-ms / m["threshDecayTGO"]);

    p.eLeakGR = m["eLeakGR"];
    p.eGOGR = m["eGOGR"];
    p.threshMaxGR = m["threshMaxGR"];
    p.threshRestGR = m["threshBaseGR"];
    p.gIncMFtoGR = m["gMFIncGR"];
    p.gIncGOtoGR = m["gGOIncGR"];
    p.gDecMFtoGR = exp(-ms / m["gMFDecayTGR"]);
    p.gDecGOtoGR = exp(-ms / m["gGODecayTGR"]);
    p.threshDecGR = 1 - exp(-ms / m["threshDecayTGR"]);
    p.gLeakGR = m["rawGLeakGR"] / (6 - ms);
    p.tsPerHistBinGR = (uint32_t)(m["msPerHistBinGR"] / ms);

    p.eLeakSC = m["eLeakSC"];
    p.gLeakSC = m["rawGLeakSC"] / (6 - ms);
    p.gDecGRtoSC = exp(-ms / m["gPFDecayTSC"]);
    p.threshMaxSC = m["threshMaxSC"];
    p.threshRestSC = m["threshBaseSC"];
    p.threshDecSC = 1 - exp(-ms / m["threshDecayTSC"]);
    p.gIncGRtoSC = m["pfIncSC"];

    p.eLeakBC = m["eLeakBC"];
    p.ePCtoBC = m["ePCBC"];
    p.gLeakBC = m["rawGLeakBC"] / (6 - ms);
    p.gDecGRtoBC = exp(-ms / m["gPFDecayTBC"]);
    p.gDecPCtoBC = exp(-ms / m["gPCDecayTBC"]);
    p.threshDecBC = 1 - exp(-ms / m["threshDecayTBC"]);
    p.threshRestBC = m["threshBaseBC"];
    p.threshMaxBC = m["threshMaxBC"];
    p.gIncGRtoBC = m["pfIncConstBC"];
    p.gIncPCtoBC = m["pcIncConstBC"];

    p.initSynWofGRtoPC = m["pfSynWInitPC"];
    p.eLeakPC = m["eLeakPC"];
    p.eBCtoPC = m["eBCPC"];
    p.eSCtoPC = m["eSCPC"];
    p.threshMaxPC = m["threshMaxPC"];
    p.threshRestPC = m["threshBasePC"];
    p.threshDecPC = 1 - exp(-ms / m["threshDecayTPC"]);
    p.gLeakPC = m["rawGLeakPC"] / (6 - ms);
    p.gDecGRtoPC = exp(-ms / m["gPFDecayTPC"]);
    p.gDecBCtoPC = exp(-ms / m["gBCDecayTPC"]);
    p.gDecSCtoPC = exp(-ms / m["gSCDecayTPC"]);
    p.gIncSCtoPC = m["gSCIncConstPC"];
    p.gIncGRtoPC = m["gPFScaleConstPC"];
    p.gIncBCtoPC = m["gBCScaleConstPC"];
    p.tsPopHistPC = (uint32_t)(40 / ms);
    p.tsPerPopHistBinPC = (uint32_t)(5 / ms);
    p.numPopHistBinsPC = p.tsPopHistPC / p.tsPerPopHistBinPC;

    p.coupleRiRjRatioIO = m["coupleScaleIO"];
    p.eLeakIO = m["eLeakIO"];
    p.eNCtoIO = m["eNCIO"];
    p.gLeakIO = m["rawGLeakIO"] / (6 - ms);
    p.gDecTSofNCtoIO = m["gNCDecTSIO"];
    p.gDecTTofNCtoIO = m["gNCDecTTIO"];
    p.gDecT0ofNCtoIO = m["gNCDecT0IO"];
    p.gIncNCtoIO = m["gNCIncScaleIO"];
    p.gIncTauNCtoIO = m["gNCIncTIO"];
    p.threshRestIO = m["threshBaseIO"];
    p.threshMaxIO = m["threshMaxIO"];
    p.threshDecIO = 1 - exp(-ms / m["threshDecayTIO"]);
    p.tsLTDDurationIO = (uint32_t)(m["msLTDDurationIO"] / ms);
    p.tsLTDStartAPIO = (int32_t)(m["msLTDStartAPIO"] / ms);
    p.tsLTPStartAPIO = (int32_t)(m["msLTPStartAPIO"] / ms);
    p.synLTPStepSizeGRtoPC = m["grPCLTPIncIO"];
    p.synLTDStepSizeGRtoPC = m["grPCLTDDecIO"];
    p.grPCHistCheckBinIO = (uint32_t)std::abs(p.tsLTDStartAPIO / ((int)p.tsPerHistBinGR));
    p.maxExtIncVIO = m["maxErrDriveIO"];

    p.eLeakNC = m["eLeakNC"];
    p.ePCtoNC = m["ePCNC"];
    p.gmaxNMDADecMFtoNC = exp(-ms / m["mfNMDADecayTNC"]);
    p.gmaxAMPADecMFtoNC = exp(-ms / m["mfAMPADecayTNC"]);
    p.gNMDAIncMFtoNC = 1 - exp(-ms / m["rawGMFNMDAIncNC"]);
    p.gAMPAIncMFtoNC = 1 - exp(-ms / m["rawGMFAMPAIncNC"]);
    p.gIncAvgPCtoNC = m["gPCScaleAvgNC"];
    p.gDecPCtoNC = exp(-ms / m["gPCDecayTNC"]);
    p.gLeakNC = m["rawGLeakNC"] / (6 - ms);
    p.threshDecNC = 1 - exp(-ms / m["threshDecayTNC"]);
    p.threshMaxNC = m["threshMaxNC"];
    p.threshRestNC = m["threshBaseNC"];
    p.relPDecTSofNCtoIO = m["outIORelPDecTSNC"];
    p.relPDecTTofNCtoIO = m["outIORelPDecTTNC"];
    p.relPDecT0ofNCtoIO = m["outIORelPDecT0NC"];
    p.relPIncNCtoIO = m["outIORelPIncScaleNC"];
    p.relPIncTauNCtoIO = m["outIORelPIncTNC"];
    p.initSynWofMFtoNC = m["mfSynWInitNC"];
    p.synLTDPCPopActThreshMFtoNC = m["mfNCLTDThreshNC"];
    p.synLTPPCPopActThreshMFtoNC = m["mfNCLTPThreshNC"];
    p.synLTDStepSizeMFtoNC = m["mfNCLTDDecNC"];
    p.synLTPStepSizeMFtoNC = m["mfNCLTPIncNC"];
}

// ConnectivityParams::ConnectivityParams (connectivityparams.cpp:12-184)
struct ConParams {
    std::map<std::string, uint32_t> m;
    std::vector<std::array<float, 3>> gogoGABA, gogoCoup;
    uint32_t maxGABAIn = 0, maxGABAOut = 0, maxCoupIn = 1, maxCoupOut = 1;
    uint32_t get(const char *k) { return m[k]; }
};

int parseCon(const char *text, ConParams &c)
{
    std::istringstream in(text ? text : "");
    std::string key;
    bool hasCoup = false;
    while (in >> key) {
        if (key == "connectivityParamEnd") {
            uint32_t v;
            in >> v;
            if (!hasCoup) c.gogoCoup.assign(1, {1.f, 1.f, 0.f});
            return CBM_OK;
        }
        if (key == "GOGOLocalMatrix" || key == "GOGOCoupLocalMatrix") {
            std::string k2;
            uint32_t nin, nout;
            in >> k2 >> nin >> k2 >> nout;
            std::vector<std::array<float, 3>> rows(nout);
            for (auto &r : rows) in >> r[0] >> r[1] >> r[2];
            if (key == "GOGOLocalMatrix") {
                c.maxGABAIn = nin;
                c.maxGABAOut = nout;
                c.gogoGABA = rows;
            } else {
                hasCoup = true;
                c.maxCoupIn = nin;
                c.maxCoupOut = nout;
                c.gogoCoup = rows;
            }
        } else {
            uint32_t v;
            in >> v;
            c.m[key] = v;
        }
        if (!in) break;
    }
    return fail(CBM_E_ARG, "conParams text has no connectivityParamEnd line");
}

}  // namespace

struct cbm_state {
    cbm_sizes s{};
    cbm_act_params p{};
    cbm_gen_spec gen{};
    uint32_t grPFVel = 1, grAFDelay = 1, msPerStep = 1;
    std::vector<uint8_t> f[CBM_F_COUNT][kMaxZones];

    template <typename T>
    T *F(int field, int z = 0) { return reinterpret_cast<T *>(f[field][z].data()); }
};

namespace {

int allocFields(cbm_state *st)
{
    for (int fid = 0; fid < CBM_F_COUNT; fid++) {
        size_t e, r, c;
        int zoned, kind;
        if (cbm_field_shape(fid, &st->s, &e, &r, &c, &zoned, &kind) != CBM_OK) continue;
        int nz = zoned ? (int)st->s.numZones : 1;
        for (int z = 0; z < nz; z++) st->f[fid][z].assign(e * r * c, 0);
    }
    return CBM_OK;
}

void fillU32(std::vector<uint8_t> &v, uint32_t val)
{
    uint32_t *p = reinterpret_cast<uint32_t *>(v.data());
    std::fill(p, p + v.size() / 4, val);
}

void fillF32(std::vector<uint8_t> &v, float val)
{
    float *p = reinterpret_cast<float *>(v.data());
    std::fill(p, p + v.size() / 4, val);
}

// draw k distinct values in [0, n) (rejection on repeats)
void drawDistinct(cbm::Sfmt &g, uint32_t n, uint32_t k, uint32_t *out)
{
    for (uint32_t a = 0; a < k; a++) {
        for (;;) {
            uint32_t v = (uint32_t)g.IRandom(0, (int)n - 1);
            bool dup = false;
            for (uint32_t b = 0; b < a; b++) dup = dup || out[b] == v;
            if (!dup) {
                out[a] = v;
                break;
            }
        }
    }
}

int wrap(int x, int n) { return ((x % n) + n) % n; }

// delay bit of a GR->target distance (innetconnectivitystate.cpp:932-934,957-959)
uint32_t delayMask(const cbm_state *st, int dist)
{
    uint32_t d = (uint32_t)dist / st->grPFVel + st->grAFDelay;
    return 0x1u << (int)(d / st->msPerStep);
}

}  // namespace

extern "C" {

const char *cbm_state_last_error(void) { return g_err.c_str(); }

int cbm_act_params_from_text(const char *text, cbm_act_params *out)
{
    if (!text || !out) return fail(CBM_E_ARG, "null argument");
    ActMap am;
    int rc = parseAct(text, am);
    if (rc) return rc;
    deriveAct(am, *out);
    return CBM_OK;
}

int cbm_state_create(const char *actText, const char *conText, const cbm_gen_spec *gen,
                     cbm_state **out)
{
    if (!actText || !conText || !gen || !out) return fail(CBM_E_ARG, "null argument");
    if (gen->kind != CBM_GEN_UNIFORM) return fail(CBM_E_UNSUPPORTED, "unknown generator kind");
    if (gen->numZones < 1 || gen->numZones > (uint32_t)kMaxZones)
        return fail(CBM_E_ARG, "numZones must be 1..8");
    ActMap am;
    ConParams cp;
    int rc = parseAct(actText, am);
    if (rc) return rc;
    rc = parseCon(conText, cp);
    if (rc) return rc;

    cbm_state *st = new cbm_state();
    st->gen = *gen;
    deriveAct(am, st->p);
    cbm_sizes &s = st->s;
    // ConnectivityParams derived sizes (connectivityparams.cpp:90-184)
    s.grX = 1u << cp.get("grXP2");
    s.grY = 1u << cp.get("grYP2");
    s.numGR = s.grX * s.grY;
    s.goX = 1u << cp.get("goXP2");
    s.goY = 1u << cp.get("goYP2");
    s.numGO = s.goX * s.goY;
    s.numMF = 1u << cp.get("numMFP2");
    s.numSC = 1u << cp.get("numSCP2");
    s.numPC = 1u << cp.get("numPCP2");
    s.numBC = s.numPC * 4;
    s.numNC = 1u << cp.get("numNCP2");
    s.numIO = 1u << cp.get("numIOP2");
    s.numZones = gen->numZones;
    s.numpPCfromGRtoPC = s.numGR / s.numPC;
    s.numpBCfromGRtoBC = s.numGR / s.numBC;
    s.numpSCfromGRtoSC = s.numGR / s.numSC;
    s.numpBCfromBCtoPC = 4;
    s.numpPCfromPCtoBC = 16;
    s.numpPCfromSCtoPC = s.numSC / s.numPC;
    s.numpNCfromPCtoNC = s.numPC / s.numNC * cp.get("numpPCfromPCtoNC");
    s.numpNCfromMFtoNC = s.numMF / s.numNC;
    s.numpIOfromNCtoIO = s.numNC;
    s.numpIOInIOIO = s.numIO - 1;
    s.numPopHistBinsPC = st->p.numPopHistBinsPC;
    s.maxnumpGRfromMFtoGR = gen->nMFperGR;
    s.maxnumpGRfromGOtoGR = gen->nGOperGR;
    s.maxnumpGRfromGRtoGO = gen->nGOoutPerGR;
    s.maxnumpGOGABAOutGOGO = cp.maxGABAOut;
    s.maxnumpGOCoupInGOGO = cp.maxCoupIn;
    st->grPFVel = cp.get("grPFVelInGRXPerTStep");
    st->grAFDelay = cp.get("grAFDelayInTStep");
    st->msPerStep = (uint32_t)st->p.msPerTimeStep;
    if (st->grPFVel == 0 || st->msPerStep == 0) {
        delete st;
        return fail(CBM_E_ARG, "grPFVelInGRXPerTStep and msPerTimeStep must be positive");
    }
    if (gen->nMFperGR > s.numMF || gen->nGOperGR > s.numGO || gen->nGOoutPerGR > s.numGO ||
        gen->nMFperGO > s.numMF) {
        delete st;
        return fail(CBM_E_ARG, "fan-in larger than the source population");
    }
    uint32_t pcnc = cp.get("numpPCfromPCtoNC");

    // --- input net: GR draws, then GO dendrites on MFs, then GO-GO wiring
    cbm::Sfmt rng(gen->innetSeed);
    const uint32_t nGR = s.numGR;
    std::vector<uint32_t> pMF((size_t)nGR * gen->nMFperGR), pGOin((size_t)nGR * gen->nGOperGR),
        pOut((size_t)nGR * gen->nGOoutPerGR);
    for (uint32_t i = 0; i < nGR; i++) {
        drawDistinct(rng, s.numMF, gen->nMFperGR, &pMF[(size_t)i * gen->nMFperGR]);
        drawDistinct(rng, s.numGO, gen->nGOperGR, &pGOin[(size_t)i * gen->nGOperGR]);
        uint32_t *o = &pOut[(size_t)i * gen->nGOoutPerGR];
        if (gen->grgoWindowX == 0) {
            drawDistinct(rng, s.numGO, gen->nGOoutPerGR, o);
        } else {
            int cx = (int)((i % s.grX) * ((float)s.goX / s.grX));
            int cy = (int)((i / s.grX) * ((float)s.goY / s.grY));
            for (uint32_t a = 0; a < gen->nGOoutPerGR; a++) {
                for (;;) {
                    int dx = rng.IRandom(0, (int)gen->grgoWindowX - 1) - (int)gen->grgoWindowX / 2;
                    int dy = rng.IRandom(0, (int)gen->grgoWindowY - 1) - (int)gen->grgoWindowY / 2;
                    uint32_t v = (uint32_t)(wrap(cy + dy, (int)s.goY) * (int)s.goX + wrap(cx + dx, (int)s.goX));
                    bool dup = false;
                    for (uint32_t b = 0; b < a; b++) dup = dup || o[b] == v;
                    if (!dup) {
                        o[a] = v;
                        break;
                    }
                }
            }
        }
    }
    std::vector<std::vector<uint32_t>> mfOut(s.numMF);
    {
        std::vector<uint32_t> tmp(gen->nMFperGO);
        for (uint32_t go = 0; go < s.numGO; go++) {
            drawDistinct(rng, s.numMF, gen->nMFperGO, tmp.data());
            for (uint32_t m : tmp) mfOut[m].push_back(go);
        }
    }
    uint32_t maxMFGO = 1;
    for (auto &v : mfOut) maxMFGO = std::max<uint32_t>(maxMFGO, (uint32_t)v.size());
    s.maxnumpMFfromMFtoGO = maxMFGO;

    allocFields(st);
    {
        int32_t *nm = st->F<int32_t>(CBM_F_NUMP_GR_FROM_MF_TO_GR);
        int32_t *ng = st->F<int32_t>(CBM_F_NUMP_GR_FROM_GO_TO_GR);
        int32_t *no = st->F<int32_t>(CBM_F_NUMP_GR_FROM_GR_TO_GO);
        std::memcpy(st->F<uint32_t>(CBM_F_P_GR_FROM_MF_TO_GR), pMF.data(), pMF.size() * 4);
        std::memcpy(st->F<uint32_t>(CBM_F_P_GR_FROM_GO_TO_GR), pGOin.data(), pGOin.size() * 4);
        std::memcpy(st->F<uint32_t>(CBM_F_P_GR_FROM_GR_TO_GO), pOut.data(), pOut.size() * 4);
        for (uint32_t i = 0; i < nGR; i++) {
            nm[i] = (int32_t)gen->nMFperGR;
            ng[i] = (int32_t)gen->nGOperGR;
            no[i] = (int32_t)gen->nGOoutPerGR;
        }
        // delays: InNetConnectivityState::assignGRDelays (innetconnectivitystate.cpp:920-962)
        uint32_t *dB = st->F<uint32_t>(CBM_F_P_GR_DELAYMASK_GR_TO_BSP);
        uint32_t *dG = st->F<uint32_t>(CBM_F_P_GR_DELAYMASK_GR_TO_GO);
        const uint32_t kOut = gen->nGOoutPerGR;
        for (uint32_t i = 0; i < nGR; i++) {
            int grPosX = (int)(i % s.grX);
            int distB = std::abs((int)(s.grX / 2 - grPosX));
            dB[i] = delayMask(st, distB);
            for (uint32_t j = 0; j < kOut; j++) {
                int goPosX = (int)((pOut[(size_t)i * kOut + j] % s.goX) * (((float)s.grX) / s.goX));
                int d = std::abs(goPosX - grPosX);
                if (d > (int)(s.grX / 2)) d = goPosX < grPosX ? goPosX + (int)s.grX - grPosX : grPosX + (int)s.grX - goPosX;
                dG[(size_t)i * kOut + j] = delayMask(st, d);
            }
        }
        int32_t *nmg = st->F<int32_t>(CBM_F_NUMP_MF_FROM_MF_TO_GO);
        uint32_t *pmg = st->F<uint32_t>(CBM_F_P_MF_FROM_MF_TO_GO);
        fillU32(st->f[CBM_F_P_MF_FROM_MF_TO_GO][0], UINT32_MAX);
        for (uint32_t m = 0; m < s.numMF; m++) {
            nmg[m] = (int32_t)mfOut[m].size();
            for (size_t j = 0; j < mfOut[m].size(); j++) pmg[(size_t)m * maxMFGO + j] = mfOut[m][j];
        }
        // GO->GO: connectGOGO (innetconnectivitystate.cpp:852-918)
        int32_t *nGab = st->F<int32_t>(CBM_F_NUMP_GO_GABA_OUT_GOGO);
        uint32_t *pGab = st->F<uint32_t>(CBM_F_P_GO_GABA_OUT_GOGO);
        int32_t *nCp = st->F<int32_t>(CBM_F_NUMP_GO_COUP_IN_GOGO);
        uint32_t *pCp = st->F<uint32_t>(CBM_F_P_GO_COUP_IN_GOGO);
        fillU32(st->f[CBM_F_P_GO_GABA_OUT_GOGO][0], UINT32_MAX);
        fillU32(st->f[CBM_F_P_GO_COUP_IN_GOGO][0], UINT32_MAX);
        for (uint32_t i = 0; i < s.numGO; i++) {
            int cx = (int)(i % s.goX), cy = (int)(i / s.goX);
            for (uint32_t j = 0; j < cp.maxGABAOut; j++) {
                if (rng.Random() >= (1 - cp.gogoGABA[j][2])) {
                    int dx = wrap(cx + (int)cp.gogoGABA[j][0], (int)s.goX);
                    int dy = wrap(cy + (int)cp.gogoGABA[j][1], (int)s.goY);
                    pGab[(size_t)i * cp.maxGABAOut + nGab[i]] = s.goX * dy + dx;
                    nGab[i]++;
                }
            }
            for (uint32_t j = 0; j < cp.maxCoupOut; j++) {
                if (rng.Random() >= (1 - cp.gogoCoup[j][2])) {
                    int dx = wrap(cx + (int)cp.gogoCoup[j][0], (int)s.goX);
                    int dy = wrap(cy + (int)cp.gogoCoup[j][1], (int)s.goY);
                    uint32_t dst = s.goX * dy + dx;
                    if ((uint32_t)nCp[dst] < cp.maxCoupIn) {
                        pCp[(size_t)dst * cp.maxCoupIn + nCp[dst]] = i;
                        nCp[dst]++;
                    }
                }
            }
        }
    }

    // --- microzones: MZoneConnectivityState (mzoneconnectivitystate.cpp:213-417)
    for (uint32_t z = 0; z < s.numZones; z++) {
        const uint32_t ratio = s.numBC / s.numPC;
        uint32_t *bcpc = st->F<uint32_t>(CBM_F_P_BC_FROM_BC_TO_PC, z);
        static const int off[4][4] = {{1, -1, 2, -2}, {1, -1, 3, -3}, {3, -3, 6, -6}, {4, -4, 9, -9}};
        for (uint32_t i = 0; i < s.numPC; i++)
            for (int r = 0; r < 4; r++)
                for (int k = 0; k < 4; k++)
                    bcpc[(size_t)(i * ratio + r) * 4 + k] = (uint32_t)wrap((int)i + off[r][k], (int)s.numPC);
        uint32_t *pcbc = st->F<uint32_t>(CBM_F_P_PC_FROM_PC_TO_BC, z);
        for (uint32_t i = 0; i < s.numPC; i++)
            for (uint32_t j = 0; j < s.numpPCfromPCtoBC; j++)
                pcbc[(size_t)i * s.numpPCfromPCtoBC + j] = (uint32_t)wrap((int)(i * ratio) - 6 + (int)j, (int)s.numBC);
        uint32_t *scpc = st->F<uint32_t>(CBM_F_P_PC_FROM_SC_TO_PC, z);
        for (uint32_t i = 0; i < s.numSC; i++)
            scpc[(size_t)(i / s.numpPCfromSCtoPC) * s.numpPCfromSCtoPC + i % s.numpPCfromSCtoPC] = i;
        // PC->NC: regular block, then random extra PCs drawn with IRandomX
        uint32_t *pcnc_ = st->F<uint32_t>(CBM_F_P_NC_FROM_PC_TO_NC, z);
        fillU32(st->f[CBM_F_P_NC_FROM_PC_TO_NC][z], UINT32_MAX);
        const uint32_t per = s.numPC / s.numNC, kN = s.numpNCfromPCtoNC;
        std::vector<std::vector<uint32_t>> pcTo(s.numPC);
        for (uint32_t i = 0; i < s.numNC; i++)
            for (uint32_t j = 0; j < per; j++) {
                pcnc_[(size_t)i * kN + j] = i * per + j;
                pcTo[i * per + j].push_back(i);
            }
        if (pcnc > 1) {
            cbm::Sfmt r2(gen->mzoneConSeed[z]);
            for (uint32_t i = 0; i + 1 < s.numNC; i++) {
                for (uint32_t j = kN / 3; j < kN; j++) {
                    bool any = false;
                    for (uint32_t q = 0; q < s.numPC && !any; q++)
                        any = pcTo[q].size() < pcnc &&
                              std::find(pcTo[q].begin(), pcTo[q].end(), i) == pcTo[q].end();
                    if (!any) {
                        delete st;
                        return fail(CBM_E_UNSUPPORTED,
                                    "PC->NC wiring has no eligible PC (the reference loop would not end); "
                                    "use numpPCfromPCtoNC 1 for these sizes");
                    }
                    uint32_t pc;
                    for (;;) {
                        pc = (uint32_t)r2.IRandomX(0, (int)s.numPC - 1);
                        if (pcTo[pc].size() >= pcnc) continue;
                        if (std::find(pcTo[pc].begin(), pcTo[pc].end(), i) != pcTo[pc].end()) continue;
                        break;
                    }
                    pcnc_[(size_t)i * kN + j] = pc;
                    pcTo[pc].push_back(i);
                }
            }
            uint32_t numSyn = per;
            for (uint32_t h = 1; h < pcnc; h++)
                for (uint32_t i = 0; i < s.numPC; i++)
                    if (pcTo[i].size() < pcnc && numSyn < kN) {
                        pcnc_[(size_t)(s.numNC - 1) * kN + numSyn] = i;
                        pcTo[i].push_back(s.numNC - 1);
                        numSyn++;
                    }
        }
        uint32_t *ncio = st->F<uint32_t>(CBM_F_P_IO_FROM_NC_TO_IO, z);
        for (uint32_t i = 0; i < s.numIO; i++)
            for (uint32_t j = 0; j < s.numNC; j++) ncio[(size_t)i * s.numNC + j] = j;
        uint32_t *mfnc = st->F<uint32_t>(CBM_F_P_NC_FROM_MF_TO_NC, z);
        for (uint32_t i = 0; i < s.numMF; i++) mfnc[i] = i;  // [i/k][i%k] = i
        uint32_t *ioio = st->F<uint32_t>(CBM_F_P_IO_IN_IOIO, z);
        for (uint32_t i = 0; i < s.numIO; i++) {
            uint32_t in = 0;
            for (uint32_t j = 0; j < s.numpIOInIOIO; j++) {
                if (in == i) in++;
                ioio[(size_t)i * s.numpIOInIOIO + j] = in;
                in++;
            }
        }
        st->F<int32_t>(CBM_F_ZONE_RNG_SEED, z)[0] = gen->mzoneRngSeed[z];
        st->F<uint64_t>(CBM_F_ZONE_RNG_SKIP, z)[0] = s.numGR;  // MZone::testReduction draws
    }

    // --- initial activity values (innetactivitystate.cpp:277-337, mzoneactivitystate.cpp:357-445)
    const cbm_act_params &p = st->p;
    fillF32(st->f[CBM_F_VM_GO][0], p.eLeakGO);
    fillF32(st->f[CBM_F_THRESH_GO][0], p.threshRestGO);
    fillF32(st->f[CBM_F_GO_GABA_OUT_SYNSCALE_GOGO][0], 1.0f);
    fillF32(st->f[CBM_F_THRESH_GR][0], p.threshRestGR);
    fillF32(st->f[CBM_F_VM_GR][0], p.eLeakGR);
    fillF32(st->f[CBM_F_THRESH_SC][0], p.threshRestSC);
    fillF32(st->f[CBM_F_VM_SC][0], p.eLeakSC);
    for (uint32_t z = 0; z < s.numZones; z++) {
        fillF32(st->f[CBM_F_THRESH_BC][z], p.threshRestBC);
        fillF32(st->f[CBM_F_VM_BC][z], p.eLeakBC);
        fillF32(st->f[CBM_F_VM_PC][z], p.eLeakPC);
        fillF32(st->f[CBM_F_THRESH_PC][z], p.threshRestPC);
        fillF32(st->f[CBM_F_PFPC_SYNW][z], p.initSynWofGRtoPC);
        fillF32(st->f[CBM_F_THRESH_IO][z], p.threshRestIO);
        fillF32(st->f[CBM_F_VM_IO][z], p.eLeakIO);
        fillF32(st->f[CBM_F_THRESH_NC][z], p.threshRestNC);
        fillF32(st->f[CBM_F_VM_NC][z], p.eLeakNC);
        fillF32(st->f[CBM_F_G_PC_SCALE_NC][z], p.gIncAvgPCtoNC);
        fillF32(st->f[CBM_F_MF_SYNW_NC][z], p.initSynWofMFtoNC);
        float steps[2] = {p.synLTDStepSizeGRtoPC, p.synLTPStepSizeGRtoPC};
        std::memcpy(st->f[CBM_F_GRPC_PLAST_STEPS][z].data(), steps, 8);
    }
    std::memcpy(st->f[CBM_F_GINC_GR_TO_GO][0].data(), &p.gIncGRtoGO, 4);
    *out = st;
    return CBM_OK;
}

int cbm_state_destroy(cbm_state *st)
{
    delete st;
    return CBM_OK;
}

int cbm_state_sizes(const cbm_state *st, cbm_sizes *out)
{
    if (!st || !out) return fail(CBM_E_ARG, "null argument");
    *out = st->s;
    return CBM_OK;
}

int cbm_state_params(const cbm_state *st, cbm_act_params *out)
{
    if (!st || !out) return fail(CBM_E_ARG, "null argument");
    *out = st->p;
    return CBM_OK;
}

int cbm_state_field(cbm_state *st, int field, int zone, void **ptr, size_t *bytes)
{
    if (!st || !ptr || !bytes || field < 0 || field >= CBM_F_COUNT) return fail(CBM_E_ARG, "bad argument");
    size_t e, r, c;
    int zoned, kind;
    if (cbm_field_shape(field, &st->s, &e, &r, &c, &zoned, &kind) != CBM_OK) return fail(CBM_E_ARG, "bad field");
    if (!zoned) zone = 0;
    if (zone < 0 || zone >= (int)st->s.numZones) return fail(CBM_E_ARG, "bad zone");
    *ptr = st->f[field][zone].data();
    *bytes = st->f[field][zone].size();
    return CBM_OK;
}

int cbm_sfmt_sample(int seed, uint64_t skip, int kind, int lo, int hi, int n, void *out)
{
    if (!out || n < 0 || kind < 0 || kind > 3) return fail(CBM_E_ARG, "bad argument");
    cbm::Sfmt g(seed);
    for (int i = 0; i < n; i++) {
        switch (kind) {
        case 0: static_cast<uint32_t *>(out)[i] = g.BRandom(); break;
        case 1:
            if (i == 0)
                for (uint64_t k = 0; k < skip; k++) (void)g.Random();
            static_cast<double *>(out)[i] = g.Random();
            break;
        case 2: static_cast<int32_t *>(out)[i] = g.IRandom(lo, hi); break;
        default: static_cast<int32_t *>(out)[i] = g.IRandomX(lo, hi); break;
        }
    }
    return CBM_OK;
}

int cbm_state_load_into(cbm_state *st, cbm_handle *h)
{
    if (!st || !h) return fail(CBM_E_ARG, "null argument");
    for (int f = 0; f < CBM_F_COUNT; f++) {
        size_t e, r, c;
        int zoned, kind;
        if (cbm_field_shape(f, &st->s, &e, &r, &c, &zoned, &kind) != CBM_OK) continue;
        if (f == CBM_F_ZONE_RNG_STATE || f == CBM_F_CUR_TIME || f == CBM_F_SUM_GR_INPUT_GO ||
            f == CBM_F_SUM_GO_INPUT_GO || f == CBM_F_PIPE_PFPC_SUM || f == CBM_F_PIPE_PFBC_SUM ||
            f == CBM_F_PIPE_PFSC_SUM || f == CBM_F_PIPE_GRGO_SUM || f == CBM_F_MF_INPUT)
            continue;
        int nz = zoned ? (int)st->s.numZones : 1;
        for (int z = 0; z < nz; z++) {
            if (st->f[f][z].empty()) continue;
            int rc = cbm_write(h, f, z, st->f[f][z].data(), st->f[f][z].size());
            if (rc) return fail(rc, std::string("cbm_write failed: ") + cbm_last_error());
        }
    }
    int rc = cbm_commit(h);
    if (rc) return fail(rc, std::string("cbm_commit failed: ") + cbm_last_error());
    return CBM_OK;
}

}  // extern "C"

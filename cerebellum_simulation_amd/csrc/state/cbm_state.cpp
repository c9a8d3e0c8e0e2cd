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
#include <fstream>
#include <iterator>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "../../../include/cbm_state.h"
#include "innet_gen.hpp"
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
    // the parameter files as parsed (the checkpoint header writes them back)
    ConParams con;
    std::map<std::string, float> act;
    // checkpoint sections the step does not use, kept byte for byte
    std::map<std::string, std::vector<uint8_t>> blob[kMaxZones];
    bool fromCheckpoint = false;  // its header is written back unchanged
    bool refGen = false;          // built by the glomerular generator: the header is the file's
    uint64_t droppedConnections = 0;

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

int cbm_state_param_at(const cbm_state *st, int which, size_t i, const char **name, double *value)
{
    if (!st || !name || !value || (which != 0 && which != 1)) return fail(CBM_E_ARG, "bad argument");
    if (which == 0) {
        if (i >= st->act.size()) return CBM_E_ARG;
        auto it = std::next(st->act.begin(), (std::ptrdiff_t)i);
        *name = it->first.c_str();
        *value = it->second;
    } else {
        if (i >= st->con.m.size()) return CBM_E_ARG;
        auto it = std::next(st->con.m.begin(), (std::ptrdiff_t)i);
        *name = it->first.c_str();
        *value = it->second;
    }
    return CBM_OK;
}

int cbm_state_set_act_param(cbm_state *st, const char *name, float value)
{
    if (!st || !name) return fail(CBM_E_ARG, "bad argument");
    auto it = st->act.find(name);
    if (it == st->act.end()) return fail(CBM_E_ARG, std::string("no activity parameter ") + name);
    it->second = value;
    ActMap am;
    am.m = st->act;
    deriveAct(am, st->p);
    return CBM_OK;
}

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
    if (gen->kind != CBM_GEN_UNIFORM && gen->kind != CBM_GEN_REFERENCE)
        return fail(CBM_E_UNSUPPORTED, "unknown generator kind");
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
    st->con = cp;
    st->act = am.m;
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
    const bool refGen = gen->kind == CBM_GEN_REFERENCE;
    st->refGen = refGen;
    s.maxnumpGRfromMFtoGR = refGen ? cp.get("maxnumpGRfromMFtoGR") : gen->nMFperGR;
    s.maxnumpGRfromGOtoGR = refGen ? cp.get("maxnumpGRfromGOtoGR") : gen->nGOperGR;
    s.maxnumpGRfromGRtoGO = refGen ? cp.get("maxnumpGRfromGRtoGO") : gen->nGOoutPerGR;
    s.maxnumpGOGABAOutGOGO = cp.maxGABAOut;
    s.maxnumpGOCoupInGOGO = cp.maxCoupIn;
    st->grPFVel = cp.get("grPFVelInGRXPerTStep");
    st->grAFDelay = cp.get("grAFDelayInTStep");
    st->msPerStep = (uint32_t)st->p.msPerTimeStep;
    if (st->grPFVel == 0 || st->msPerStep == 0) {
        delete st;
        return fail(CBM_E_ARG, "grPFVelInGRXPerTStep and msPerTimeStep must be positive");
    }
    if (!refGen && (gen->nMFperGR > s.numMF || gen->nGOperGR > s.numGO || gen->nGOoutPerGR > s.numGO ||
                    gen->nMFperGO > s.numMF)) {
        delete st;
        return fail(CBM_E_ARG, "fan-in larger than the source population");
    }
    uint32_t pcnc = cp.get("numpPCfromPCtoNC");

    cbm::Sfmt rng(gen->innetSeed);
    const uint32_t nGR = s.numGR;
    if (refGen) {
        // --- input net by the reference's glomerular generator, then GO-GO wiring
        static const char *keys[] = {"glXP2", "glYP2", "maxnumpGLfromGLtoGR", "lownumpGLfromGLtoGR",
                                     "maxnumpGLfromGLtoGO", "maxnumpGLfromGOtoGL", "maxnumpGOfromGRtoGO",
                                     "maxnumpGOfromGLtoGO", "maxnumpGOfromGOtoGL", "maxnumpGRfromGLtoGR",
                                     "maxnumpGRfromGOtoGR", "maxnumpGRfromMFtoGR", "maxnumpGRfromGRtoGO"};
        for (const char *k : keys)
            if (!cp.m.count(k)) {
                delete st;
                return fail(CBM_E_ARG, std::string("the glomerular generator needs conParams key ") + k);
            }
        cbm::InnetGenParams q{};
        q.glX = 1u << cp.get("glXP2");
        q.glY = 1u << cp.get("glYP2");
        q.grX = s.grX;
        q.grY = s.grY;
        q.goX = s.goX;
        q.goY = s.goY;
        q.numMF = s.numMF;
        q.maxGLtoGR = cp.get("maxnumpGLfromGLtoGR");
        q.lowGLtoGR = cp.get("lownumpGLfromGLtoGR");
        q.maxGLtoGO = cp.get("maxnumpGLfromGLtoGO");
        q.maxGLfromGO = cp.get("maxnumpGLfromGOtoGL");
        q.maxGRfromGL = cp.get("maxnumpGRfromGLtoGR");
        q.maxGRfromGO = s.maxnumpGRfromGOtoGR;
        q.maxGRfromMF = s.maxnumpGRfromMFtoGR;
        q.maxGRtoGO = s.maxnumpGRfromGRtoGO;
        q.maxGOfromGR = cp.get("maxnumpGOfromGRtoGO");
        q.maxGOfromGL = cp.get("maxnumpGOfromGLtoGO");
        q.maxGOtoGL = cp.get("maxnumpGOfromGOtoGL");
        q.spanGRDenX = cp.get("spanGRDenOnGLX");
        q.spanGRDenY = cp.get("spanGRDenOnGLY");
        q.spanGODecX = cp.get("spanGODecDenOnGLX");
        q.spanGODecY = cp.get("spanGODecDenOnGLY");
        q.spanGOAxonX = cp.get("spanGOAxonOnGLX");
        q.spanGOAxonY = cp.get("spanGOAxonOnGLY");
        q.spanGOAscX = cp.get("spanGOAscDenOnGRX");
        q.spanGOAscY = cp.get("spanGOAscDenOnGRY");
        cbm::InnetGen rg;
        const std::string why = cbm::generateInnet(q, rng, rg);
        if (!why.empty()) {
            delete st;
            return fail(CBM_E_UNSUPPORTED, "glomerular generator: " + why);
        }
        s.maxnumpMFfromMFtoGO = rg.mfToGO.cols;
        allocFields(st);
        auto put = [&](int field, const cbm::ConList &l, int numField) {
            uint32_t *dst = st->F<uint32_t>(field);
            const size_t w = st->f[field][0].size() / 4 / l.rows;
            for (uint32_t i = 0; i < l.rows; i++)
                for (size_t j = 0; j < w; j++) dst[(size_t)i * w + j] = j < l.cols ? l.p[(size_t)i * l.cols + j] : UINT32_MAX;
            std::memcpy(st->f[numField][0].data(), l.num.data(), (size_t)l.rows * 4);
        };
        put(CBM_F_P_GR_FROM_MF_TO_GR, rg.grFromMF, CBM_F_NUMP_GR_FROM_MF_TO_GR);
        put(CBM_F_P_GR_FROM_GO_TO_GR, rg.grFromGO, CBM_F_NUMP_GR_FROM_GO_TO_GR);
        put(CBM_F_P_GR_FROM_GR_TO_GO, rg.grToGO, CBM_F_NUMP_GR_FROM_GR_TO_GO);
        put(CBM_F_P_MF_FROM_MF_TO_GO, rg.mfToGO, CBM_F_NUMP_MF_FROM_MF_TO_GO);
        // the lists the step does not use, in the checkpoint's layout
        auto keep = [&](const char *name, const void *p, size_t bytes) {
            st->blob[0][name].assign(static_cast<const uint8_t *>(p), static_cast<const uint8_t *>(p) + bytes);
        };
        auto keepList = [&](const char *numName, const char *listName, const cbm::ConList &l) {
            keep(numName, l.num.data(), l.num.size() * 4);
            keep(listName, l.p.data(), l.p.size() * 4);
        };
        keep("haspGLfromMFtoGL", rg.hasGLfromMF.data(), rg.hasGLfromMF.size());
        keep("pGLfromMFtoGL", rg.glFromMF.data(), rg.glFromMF.size() * 4);
        keepList("numpGLfromGLtoGO", "pGLfromGLtoGO", rg.glToGO);
        keepList("numpGLfromGOtoGL", "pGLfromGOtoGL", rg.glFromGO);
        keepList("numpGLfromGLtoGR", "pGLfromGLtoGR", rg.glToGR);
        keepList("numpMFfromMFtoGL", "pMFfromMFtoGL", rg.mfToGL);
        keepList("numpMFfromMFtoGR", "pMFfromMFtoGR", rg.mfToGR);
        keepList("numpGOfromGLtoGO", "pGOfromGLtoGO", rg.goFromGL);
        keepList("numpGOfromGOtoGL", "pGOfromGOtoGL", rg.goToGL);
        keepList("numpGOfromMFtoGO", "pGOfromMFtoGO", rg.goFromMF);
        keepList("numpGOfromGOtoGR", "pGOfromGOtoGR", rg.goToGR);
        keepList("numpGOfromGRtoGO", "pGOfromGRtoGO", rg.goFromGR);
        keepList("numpGRfromGLtoGR", "pGRfromGLtoGR", rg.grFromGL);
        st->droppedConnections = rg.droppedConnections;
    } else {
    // --- input net: GR draws, then GO dendrites on MFs, then GO-GO wiring
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
        int32_t *nmg = st->F<int32_t>(CBM_F_NUMP_MF_FROM_MF_TO_GO);
        uint32_t *pmg = st->F<uint32_t>(CBM_F_P_MF_FROM_MF_TO_GO);
        fillU32(st->f[CBM_F_P_MF_FROM_MF_TO_GO][0], UINT32_MAX);
        for (uint32_t m = 0; m < s.numMF; m++) {
            nmg[m] = (int32_t)mfOut[m].size();
            for (size_t j = 0; j < mfOut[m].size(); j++) pmg[(size_t)m * maxMFGO + j] = mfOut[m][j];
        }
    }
    }  // synthetic generator
    {
        // delays: InNetConnectivityState::assignGRDelays (innetconnectivitystate.cpp:920-962);
        // slots past a GR's GO-target count keep UINT_MAX, as initializeVals leaves them
        uint32_t *dB = st->F<uint32_t>(CBM_F_P_GR_DELAYMASK_GR_TO_BSP);
        uint32_t *dG = st->F<uint32_t>(CBM_F_P_GR_DELAYMASK_GR_TO_GO);
        const int32_t *no = st->F<int32_t>(CBM_F_NUMP_GR_FROM_GR_TO_GO);
        const uint32_t *pOutF = st->F<uint32_t>(CBM_F_P_GR_FROM_GR_TO_GO);
        const uint32_t kOut = s.maxnumpGRfromGRtoGO;
        fillU32(st->f[CBM_F_P_GR_DELAYMASK_GR_TO_GO][0], UINT32_MAX);
        for (uint32_t i = 0; i < nGR; i++) {
            int grPosX = (int)(i % s.grX);
            int distB = std::abs((int)(s.grX / 2 - grPosX));
            dB[i] = delayMask(st, distB);
            for (uint32_t j = 0; j < (uint32_t)no[i] && j < kOut; j++) {
                int goPosX = (int)((pOutF[(size_t)i * kOut + j] % s.goX) * (((float)s.grX) / s.goX));
                int d = std::abs(goPosX - grPosX);
                if (d > (int)(s.grX / 2)) d = goPosX < grPosX ? goPosX + (int)s.grX - grPosX : grPosX + (int)s.grX - goPosX;
                dG[(size_t)i * kOut + j] = delayMask(st, d);
            }
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

// ---------------------------------------------------------------- checkpoint
// The CBMState file format: CBMState::writeState / CBMState(fstream&)
// (CBM_STATE_LIB_DEV/src/interfaces/cbmstate.cpp:12-33,86-103).  A text
// header -- the zone count, ConnectivityParams::writeParams
// (connectivityparams.cpp:192-218), ActivityParams::writeParams
// (activityparams.cpp:243-253) -- then the raw arrays of the four stateRW
// methods in their order: InNetConnectivityState (innetconnectivitystate.cpp:
// 530-606), MZoneConnectivityState per zone (mzoneconnectivitystate.cpp:
// 142-163), InNetActivityState (innetactivitystate.cpp:233-275),
// MZoneActivityState per zone (mzoneactivitystate.cpp:299-355).  Array sizes
// follow from the header's connectivity parameters as ConnectivityParams
// derives them (connectivityparams.cpp:90-184).
namespace {

struct RefSizes {
    uint32_t numGL, glToGR, glToGO, glFromGO;
    uint32_t numMF, mfToGL, mfToGO, mfToGR;
    uint32_t numGO, goFromGR, goFromGL, goFromMF, goToGL, goToGR;
    uint32_t numGR, grToGO, grFromGO, grFromMFkey;
    uint32_t numSC, numPC, numBC, numNC, numIO;
    uint32_t pcFromGR, pcFromBC, pcToBC, pcFromSC, pcToNC, bcToPC, bcFromPC, scToPC;
    uint32_t ncFromPC, ncToIO, ncFromMF, ioToPC, ioFromNC, ioIn, ioOut;
    uint32_t gabaIn, gabaOut, coupIn, coupOut;
};

RefSizes deriveRef(const ConParams &c)
{
    auto m = [&](const char *k) -> uint32_t {
        auto it = c.m.find(k);
        return it == c.m.end() ? 0u : it->second;
    };
    RefSizes r{};
    r.numGL = (1u << m("glXP2")) * (1u << m("glYP2"));
    r.glToGR = m("maxnumpGLfromGLtoGR");
    r.glToGO = m("maxnumpGLfromGLtoGO");
    r.glFromGO = m("maxnumpGLfromGOtoGL");
    r.numMF = 1u << m("numMFP2");
    r.mfToGL = r.numGL / r.numMF;
    r.mfToGO = r.mfToGL * r.glToGO;
    r.mfToGR = r.mfToGL * r.glToGR;
    r.numGO = (1u << m("goXP2")) * (1u << m("goYP2"));
    r.goFromGR = m("maxnumpGOfromGRtoGO");
    r.goFromGL = m("maxnumpGOfromGLtoGO");
    r.goFromMF = r.goFromGL;
    r.goToGL = m("maxnumpGOfromGOtoGL");
    r.goToGR = r.goToGL * r.glToGR;
    r.numGR = (1u << m("grXP2")) * (1u << m("grYP2"));
    r.grToGO = m("maxnumpGRfromGRtoGO");
    r.grFromGO = m("maxnumpGRfromGOtoGR");
    r.grFromMFkey = m("maxnumpGRfromMFtoGR");
    r.numSC = 1u << m("numSCP2");
    r.numPC = 1u << m("numPCP2");
    r.numBC = r.numPC * 4;
    r.numNC = 1u << m("numNCP2");
    r.numIO = 1u << m("numIOP2");
    r.pcFromGR = r.numGR / r.numPC;
    r.pcFromBC = 16;
    r.pcToBC = 16;
    r.pcFromSC = r.numSC / r.numPC;
    r.pcToNC = m("numpPCfromPCtoNC");
    r.bcToPC = 4;
    r.bcFromPC = 4;
    r.scToPC = 1;
    r.ncFromPC = r.numPC / r.numNC * r.pcToNC;
    r.ncToIO = r.numIO;
    r.ncFromMF = r.numMF / r.numNC;
    r.ioToPC = r.numPC / r.numIO;
    r.ioFromNC = r.numNC;
    r.ioIn = r.numIO - 1;
    r.ioOut = r.numIO - 1;
    r.gabaIn = c.maxGABAIn;
    r.gabaOut = c.maxGABAOut;
    r.coupIn = c.maxCoupIn;
    r.coupOut = c.maxCoupOut;
    return r;
}

// One array of the file: its name, its field here (-1: kept as a blob, or
// synthesised when writing a state built here), element size, rows and
// columns as the file lays them out, and whether each zone has its own.
struct Sec {
    const char *name;
    int field;
    uint32_t elem;
    uint64_t rows, cols;
    bool zoned;
};

std::vector<Sec> sections(const RefSizes &r, uint32_t popBins)
{
    const uint64_t gl = r.numGL, mf = r.numMF, go = r.numGO, gr = r.numGR;
    std::vector<Sec> v = {
        // InNetConnectivityState::stateRW
        {"haspGLfromMFtoGL", -1, 1, gl, 1, false},
        {"pGLfromMFtoGL", -1, 4, gl, 1, false},
        {"numpGLfromGLtoGO", -1, 4, gl, 1, false},
        {"pGLfromGLtoGO", -1, 4, gl, r.glToGO, false},
        {"numpGLfromGOtoGL", -1, 4, gl, 1, false},
        {"pGLfromGOtoGL", -1, 4, gl, r.glFromGO, false},
        {"numpGLfromGLtoGR", -1, 4, gl, 1, false},
        {"pGLfromGLtoGR", -1, 4, gl, r.glToGR, false},
        {"numpMFfromMFtoGL", -1, 4, mf, 1, false},
        {"pMFfromMFtoGL", -1, 4, mf, r.mfToGL, false},
        {"numpMFfromMFtoGR", -1, 4, mf, 1, false},
        {"pMFfromMFtoGR", -1, 4, mf, r.mfToGR, false},
        {"numpMFfromMFtoGO", CBM_F_NUMP_MF_FROM_MF_TO_GO, 4, mf, 1, false},
        {"pMFfromMFtoGO", CBM_F_P_MF_FROM_MF_TO_GO, 4, mf, r.mfToGO, false},
        {"numpGOfromGLtoGO", -1, 4, go, 1, false},
        {"pGOfromGLtoGO", -1, 4, go, r.goFromGL, false},
        {"numpGOfromGOtoGL", -1, 4, go, 1, false},
        {"pGOfromGOtoGL", -1, 4, go, r.goToGL, false},
        {"numpGOfromMFtoGO", -1, 4, go, 1, false},
        {"pGOfromMFtoGO", -1, 4, go, r.goFromMF, false},
        {"numpGOfromGOtoGR", -1, 4, go, 1, false},
        {"pGOfromGOtoGR", -1, 4, go, r.goToGR, false},
        {"numpGOfromGRtoGO", -1, 4, go, 1, false},
        {"pGOfromGRtoGO", -1, 4, go, r.goFromGR, false},
        {"numpGOGABAInGOGO", -1, 4, go, 1, false},
        {"pGOGABAInGOGO", -1, 4, go, r.gabaIn, false},
        {"numpGOGABAOutGOGO", CBM_F_NUMP_GO_GABA_OUT_GOGO, 4, go, 1, false},
        {"pGOGABAOutGOGO", CBM_F_P_GO_GABA_OUT_GOGO, 4, go, r.gabaOut, false},
        {"numpGOCoupInGOGO", CBM_F_NUMP_GO_COUP_IN_GOGO, 4, go, 1, false},
        {"pGOCoupInGOGO", CBM_F_P_GO_COUP_IN_GOGO, 4, go, r.coupIn, false},
        {"numpGOCoupOutGOGO", -1, 4, go, 1, false},
        {"pGOCoupOutGOGO", -1, 4, go, r.coupOut, false},
        {"pGRDelayMaskfromGRtoBSP", CBM_F_P_GR_DELAYMASK_GR_TO_BSP, 4, gr, 1, false},
        {"numpGRfromGLtoGR", -1, 4, gr, 1, false},
        // allocated and written with maxnumpGRfromGOtoGR columns (innetconnectivitystate.cpp:517,595)
        {"pGRfromGLtoGR", -1, 4, gr, r.grFromGO, false},
        {"numpGRfromGRtoGO", CBM_F_NUMP_GR_FROM_GR_TO_GO, 4, gr, 1, false},
        {"pGRfromGRtoGO", CBM_F_P_GR_FROM_GR_TO_GO, 4, gr, r.grToGO, false},
        {"pGRDelayMaskfromGRtoGO", CBM_F_P_GR_DELAYMASK_GR_TO_GO, 4, gr, r.grToGO, false},
        {"numpGRfromGOtoGR", CBM_F_NUMP_GR_FROM_GO_TO_GR, 4, gr, 1, false},
        {"pGRfromGOtoGR", CBM_F_P_GR_FROM_GO_TO_GR, 4, gr, r.grFromGO, false},
        {"numpGRfromMFtoGR", CBM_F_NUMP_GR_FROM_MF_TO_GR, 4, gr, 1, false},
        // also maxnumpGRfromGOtoGR columns (innetconnectivitystate.cpp:527,605)
        {"pGRfromMFtoGR", CBM_F_P_GR_FROM_MF_TO_GR, 4, gr, r.grFromGO, false},
        // MZoneConnectivityState::stateRW, per zone
        {"pBCfromBCtoPC", CBM_F_P_BC_FROM_BC_TO_PC, 4, r.numBC, r.bcToPC, true},
        {"pBCfromPCtoBC", -1, 4, r.numBC, r.bcFromPC, true},
        {"pSCfromSCtoPC", -1, 4, r.numSC, r.scToPC, true},
        {"pPCfromBCtoPC", -1, 4, r.numPC, r.pcFromBC, true},
        // written in the place of pPCfromPCtoBC, which the file never holds
        // (mzoneconnectivitystate.cpp:150)
        {"pBCfromPCtoBC#2", -1, 4, r.numPC, r.pcToBC, true},
        {"pPCfromSCtoPC", CBM_F_P_PC_FROM_SC_TO_PC, 4, r.numPC, r.pcFromSC, true},
        {"pPCfromPCtoNC", -1, 4, r.numPC, r.pcToNC, true},
        {"pPCfromIOtoPC", -1, 4, r.numPC, 1, true},
        {"pNCfromPCtoNC", CBM_F_P_NC_FROM_PC_TO_NC, 4, r.numNC, r.ncFromPC, true},
        {"pNCfromNCtoIO", -1, 4, r.numNC, r.ncToIO, true},
        {"pNCfromMFtoNC", CBM_F_P_NC_FROM_MF_TO_NC, 4, r.numNC, r.ncFromMF, true},
        {"pIOfromIOtoPC", -1, 4, r.numIO, r.ioToPC, true},
        {"pIOfromNCtoIO", CBM_F_P_IO_FROM_NC_TO_IO, 4, r.numIO, r.ioFromNC, true},
        {"pIOInIOIO", CBM_F_P_IO_IN_IOIO, 4, r.numIO, r.ioIn, true},
        {"pIOOutIOIO", -1, 4, r.numIO, r.ioOut, true},
        // InNetActivityState::stateRW
        {"histMF", CBM_F_HIST_MF, 1, mf, 1, false},
        {"apBufMF", CBM_F_APBUF_MF, 4, mf, 1, false},
        {"apGO", CBM_F_AP_GO, 1, go, 1, false},
        {"apBufGO", CBM_F_APBUF_GO, 4, go, 1, false},
        {"vGO", CBM_F_VM_GO, 4, go, 1, false},
        {"vCoupleGO", CBM_F_VCOUPLE_GO, 4, go, 1, false},
        {"threshCurGO", CBM_F_THRESH_GO, 4, go, 1, false},
        {"inputMFGO", CBM_F_INPUT_MF_GO, 4, go, 1, false},
        {"inputGOGO", CBM_F_INPUT_GO_GO, 4, go, 1, false},
        {"gMFGO", CBM_F_G_MF_GO, 4, go, 1, false},
        {"gGRGO", CBM_F_G_GR_GO, 4, go, 1, false},
        {"gGOGO", CBM_F_G_GO_GO, 4, go, 1, false},
        {"gMGluRGO", CBM_F_G_MGLUR_GO, 4, go, 1, false},
        {"gMGluRIncGO", CBM_F_G_MGLURINC_GO, 4, go, 1, false},
        {"mGluRGO", CBM_F_MGLUR_GO, 4, go, 1, false},
        {"gluGO", CBM_F_GLU_GO, 4, go, 1, false},
        {"apGR", CBM_F_AP_GR, 1, gr, 1, false},
        {"apBufGR", CBM_F_APBUF_GR, 4, gr, 1, false},
        {"gMFGR", -1, 4, gr, r.grFromMFkey, false},
        {"gMFSumGR", CBM_F_G_MF_SUM_GR, 4, gr, 1, false},
        {"gGOGR", -1, 4, gr, r.grFromGO, false},
        {"gGOSumGR", CBM_F_G_GO_SUM_GR, 4, gr, 1, false},
        {"threshGR", CBM_F_THRESH_GR, 4, gr, 1, false},
        {"vGR", CBM_F_VM_GR, 4, gr, 1, false},
        {"gKCaGR", CBM_F_GKCA_GR, 4, gr, 1, false},
        {"historyGR", CBM_F_HIST_GR, 8, gr, 1, false},
        {"apSC", CBM_F_AP_SC, 1, r.numSC, 1, false},
        {"apBufSC", CBM_F_APBUF_SC, 4, r.numSC, 1, false},
        {"gPFSC", CBM_F_G_PF_SC, 4, r.numSC, 1, false},
        {"threshSC", CBM_F_THRESH_SC, 4, r.numSC, 1, false},
        {"vSC", CBM_F_VM_SC, 4, r.numSC, 1, false},
        {"inputSumPFSC", CBM_F_INPUT_SUM_PF_SC, 4, r.numSC, 1, false},
        // MZoneActivityState::stateRW, per zone
        {"apBC", CBM_F_AP_BC, 1, r.numBC, 1, true},
        {"apBufBC", CBM_F_APBUF_BC, 4, r.numBC, 1, true},
        {"inputPCBC", CBM_F_INPUT_PC_BC, 4, r.numBC, 1, true},
        {"gPFBC", CBM_F_G_PF_BC, 4, r.numBC, 1, true},
        {"gPCBC", CBM_F_G_PC_BC, 4, r.numBC, 1, true},
        {"threshBC", CBM_F_THRESH_BC, 4, r.numBC, 1, true},
        {"vBC", CBM_F_VM_BC, 4, r.numBC, 1, true},
        {"apPC", CBM_F_AP_PC, 1, r.numPC, 1, true},
        {"apBufPC", CBM_F_APBUF_PC, 4, r.numPC, 1, true},
        {"inputBCPC", CBM_F_INPUT_BC_PC, 4, r.numPC, 1, true},
        {"inputSCPC", CBM_F_INPUT_SC_PC, 1, r.numPC, r.pcFromSC, true},
        {"pfSynWeightPC", CBM_F_PFPC_SYNW, 4, r.numPC, r.pcFromGR, true},
        {"inputSumPFPC", CBM_F_INPUT_SUM_PF_PC, 4, r.numPC, 1, true},
        {"gPFPC", CBM_F_G_PF_PC, 4, r.numPC, 1, true},
        {"gBCPC", CBM_F_G_BC_PC, 4, r.numPC, 1, true},
        {"gSCPC", CBM_F_G_SC_PC, 4, r.numPC, r.pcFromSC, true},
        {"vPC", CBM_F_VM_PC, 4, r.numPC, 1, true},
        {"threshPC", CBM_F_THRESH_PC, 4, r.numPC, 1, true},
        {"histPCPopAct", CBM_F_HIST_PC_POPACT, 4, popBins, 1, true},
        {"apIO", CBM_F_AP_IO, 1, r.numIO, 1, true},
        {"apBufIO", CBM_F_APBUF_IO, 4, r.numIO, 1, true},
        {"inputNCIO", CBM_F_INPUT_NC_IO, 1, r.numIO, r.ioFromNC, true},
        {"gNCIO", CBM_F_G_NC_IO, 4, r.numIO, r.ioFromNC, true},
        {"threshIO", CBM_F_THRESH_IO, 4, r.numIO, 1, true},
        {"vIO", CBM_F_VM_IO, 4, r.numIO, 1, true},
        {"vCoupleIO", CBM_F_VCOUPLE_IO, 4, r.numIO, 1, true},
        {"pfPCPlastTimerIO", CBM_F_PFPC_PLAST_TIMER_IO, 4, r.numIO, 1, true},
        {"apNC", CBM_F_AP_NC, 1, r.numNC, 1, true},
        {"apBufNC", CBM_F_APBUF_NC, 4, r.numNC, 1, true},
        {"inputPCNC", CBM_F_INPUT_PC_NC, 1, r.numNC, r.ncFromPC, true},
        {"inputMFNC", CBM_F_INPUT_MF_NC, 1, r.numNC, r.ncFromMF, true},
        {"gPCNC", CBM_F_G_PC_NC, 4, r.numNC, r.ncFromPC, true},
        {"gPCScaleNC", CBM_F_G_PC_SCALE_NC, 4, r.numNC, r.ncFromPC, true},
        {"mfSynWeightNC", CBM_F_MF_SYNW_NC, 4, r.numNC, r.ncFromMF, true},
        {"gmaxNMDAofMFtoNC", CBM_F_GMAX_NMDA_MF_NC, 4, r.numNC, r.ncFromMF, true},
        {"gmaxAMPAofMFtoNC", CBM_F_GMAX_AMPA_MF_NC, 4, r.numNC, r.ncFromMF, true},
        {"gMFNMDANC", CBM_F_G_MF_NMDA_NC, 4, r.numNC, r.ncFromMF, true},
        {"gMFAMPANC", CBM_F_G_MF_AMPA_NC, 4, r.numNC, r.ncFromMF, true},
        {"threshNC", CBM_F_THRESH_NC, 4, r.numNC, 1, true},
        {"vNC", CBM_F_VM_NC, 4, r.numNC, 1, true},
        {"synIOPReleaseNC", CBM_F_SYN_IO_PRELEASE_NC, 4, r.numNC, 1, true},
    };
    return v;
}

// the array groups, in file order: InNet con, MZone con (per zone), InNet act, MZone act (per zone)
enum { G_INNET_CON, G_MZONE_CON, G_INNET_ACT, G_MZONE_ACT };
int groupOf(const std::vector<Sec> &all, size_t i)
{
    const std::string n = all[i].name;
    size_t firstMZ = 0, firstIA = 0, firstMA = 0;
    for (size_t k = 0; k < all.size(); k++) {
        if (!firstMZ && std::string(all[k].name) == "pBCfromBCtoPC") firstMZ = k;
        if (!firstIA && std::string(all[k].name) == "histMF") firstIA = k;
        if (!firstMA && std::string(all[k].name) == "apBC") firstMA = k;
    }
    return i < firstMZ ? G_INNET_CON : i < firstIA ? G_MZONE_CON : i < firstMA ? G_INNET_ACT : G_MZONE_ACT;
}

// Copy rows x fromCols elements into rows x toCols (fill the extra columns;
// fails if a dropped column holds anything but the fill).
bool repack(const uint8_t *src, uint64_t rows, uint64_t fromCols, uint8_t *dst, uint64_t toCols, uint32_t elem,
            uint32_t fill)
{
    for (uint64_t rI = 0; rI < rows; rI++)
        for (uint64_t c = 0; c < std::max(fromCols, toCols); c++) {
            if (c < fromCols && c < toCols) {
                std::memcpy(dst + (rI * toCols + c) * elem, src + (rI * fromCols + c) * elem, elem);
            } else if (c < toCols) {
                std::memcpy(dst + (rI * toCols + c) * elem, &fill, elem);
            } else {
                uint32_t v = 0;
                std::memcpy(&v, src + (rI * fromCols + c) * elem, std::min<uint32_t>(elem, 4));
                if (v != fill) return false;
            }
        }
    return true;
}

bool isIndexList(const char *name) { return name[0] == 'p' && std::strncmp(name, "pfSyn", 5) != 0 && std::strncmp(name, "pfPC", 4) != 0; }

// Content of a section a state built here does not hold: what the
// reference's constructors leave in it (mzoneconnectivitystate.cpp:190-417:
// pBCfromPCtoBC, pPCfromBCtoPC and pIOOutIOIO are never filled), the reverse
// lists of the lists the step uses, or empty lists (count 0, entries
// UINT_MAX) and zero conductances.
std::vector<uint8_t> synth(cbm_state *st, const Sec &sc, const RefSizes &r, int z)
{
    const uint64_t n = sc.rows * sc.cols;
    std::vector<uint8_t> out(n * sc.elem, 0);
    uint32_t *u = reinterpret_cast<uint32_t *>(out.data());
    const std::string nm = sc.name;
    if (sc.elem == 4 && isIndexList(sc.name)) std::fill(u, u + n, UINT32_MAX);
    if (nm == "pSCfromSCtoPC") {
        for (uint64_t i = 0; i < r.numSC; i++) u[i] = (uint32_t)(i / r.pcFromSC);
    } else if (nm == "pPCfromIOtoPC") {
        for (uint64_t i = 0; i < r.numPC; i++) u[i] = (uint32_t)(i / r.ioToPC);
    } else if (nm == "pNCfromNCtoIO") {
        for (uint64_t i = 0; i < r.numIO; i++)
            for (uint64_t j = 0; j < r.numNC; j++) u[j * r.ncToIO + i] = (uint32_t)i;
    } else if (nm == "pIOfromIOtoPC") {
        for (uint64_t i = 0; i < r.numIO; i++)
            for (uint64_t j = 0; j < r.ioToPC; j++) u[i * r.ioToPC + j] = (uint32_t)(i * r.ioToPC + j);
    } else if (nm == "pPCfromPCtoNC") {
        // connectPCtoNC (mzoneconnectivitystate.cpp:276-358) appends NC i to
        // a PC's list: first the regular block, then the extra NCs in NC
        // order, then the last NC's fill
        const uint32_t *nc = st->F<uint32_t>(CBM_F_P_NC_FROM_PC_TO_NC, z);
        const uint64_t kN = r.ncFromPC, reg = kN / 3 ? kN / 3 : kN;
        std::vector<uint32_t> cnt(r.numPC, 0);
        auto add = [&](uint64_t i, uint64_t j) {
            const uint32_t pc = nc[i * kN + j];
            if (pc < r.numPC && cnt[pc] < r.pcToNC) u[(uint64_t)pc * r.pcToNC + cnt[pc]++] = (uint32_t)i;
        };
        for (uint64_t i = 0; i < r.numNC; i++)
            for (uint64_t j = 0; j < reg && j < kN; j++) add(i, j);
        for (uint64_t i = 0; i + 1 < r.numNC; i++)
            for (uint64_t j = reg; j < kN; j++) add(i, j);
        for (uint64_t j = reg; j < kN; j++) add(r.numNC - 1, j);
    } else if (nm == "numpGOGABAInGOGO" || nm == "pGOGABAInGOGO" || nm == "numpGOCoupOutGOGO" ||
               nm == "pGOCoupOutGOGO") {
        // reverse of the GO-GO out-list (GABA) or in-list (coupling), in source order
        const bool gaba = nm.find("GABA") != std::string::npos;
        const uint32_t *num = st->F<uint32_t>(gaba ? CBM_F_NUMP_GO_GABA_OUT_GOGO : CBM_F_NUMP_GO_COUP_IN_GOGO);
        const uint32_t *lst = st->F<uint32_t>(gaba ? CBM_F_P_GO_GABA_OUT_GOGO : CBM_F_P_GO_COUP_IN_GOGO);
        const uint64_t w = gaba ? st->s.maxnumpGOGABAOutGOGO : st->s.maxnumpGOCoupInGOGO;
        const uint64_t cols = gaba ? r.gabaIn : r.coupOut;
        std::vector<uint32_t> c(r.numGO, 0);
        std::vector<uint32_t> lists(r.numGO * cols, UINT32_MAX);
        for (uint64_t i = 0; i < r.numGO; i++)
            for (uint64_t k = 0; k < num[i] && k < w; k++) {
                const uint32_t t = lst[i * w + k];
                if (t < r.numGO && c[t] < cols) lists[(uint64_t)t * cols + c[t]++] = (uint32_t)i;
            }
        if (nm[0] == 'n')
            std::memcpy(out.data(), c.data(), r.numGO * 4);
        else
            std::memcpy(out.data(), lists.data(), lists.size() * 4);
    } else if (sc.elem == 4 && nm.compare(0, 4, "nump") == 0) {
        std::fill(u, u + n, 0u);
    }
    return out;
}

// Header of a state built here: its parameter files, with the sizes the file
// layout derives made to match the arrays held here (the glomerulus layer of
// the reference's generator does not exist here: one glomerulus per MF).
ConParams headerOf(cbm_state *st)
{
    if (st->fromCheckpoint || st->refGen) return st->con;
    ConParams c = st->con;
    RefSizes r = deriveRef(c);
    const cbm_sizes &s = st->s;
    uint32_t mfP2 = 0;
    while ((1u << mfP2) < s.numMF) mfP2++;
    if (r.mfToGO != s.maxnumpMFfromMFtoGO || r.mfToGL == 0) {
        c.m["glXP2"] = mfP2;
        c.m["glYP2"] = 0;
        c.m["maxnumpGLfromGLtoGO"] = s.maxnumpMFfromMFtoGO;
    }
    const uint32_t wGR = std::max(s.maxnumpGRfromMFtoGR, s.maxnumpGRfromGOtoGR);
    if (r.grFromGO != wGR) c.m["maxnumpGRfromGOtoGR"] = wGR;
    if (r.grFromMFkey != s.maxnumpGRfromMFtoGR) c.m["maxnumpGRfromMFtoGR"] = s.maxnumpGRfromMFtoGR;
    if (r.grToGO != s.maxnumpGRfromGRtoGO) c.m["maxnumpGRfromGRtoGO"] = s.maxnumpGRfromGRtoGO;
    return c;
}

int writeCheckpoint(cbm_state *st, std::ostream &f)
{
    const ConParams hc = headerOf(st);
    const RefSizes r = deriveRef(hc);
    const cbm_sizes &s = st->s;
    if (r.numGR != s.numGR || r.numGO != s.numGO || r.numMF != s.numMF || r.numPC != s.numPC || r.numNC != s.numNC ||
        r.numIO != s.numIO || r.numSC != s.numSC)
        return fail(CBM_E_STATE, "the parameter header does not describe this state's sizes");
    // InNetActivityState::stateRW sets these on write as on read (innetactivitystate.cpp:247-248)
    fillF32(st->f[CBM_F_INPUT_GOGABA_SYNDEP_GO][0], 1.0f);
    fillF32(st->f[CBM_F_GO_GABA_OUT_SYNSCALE_GOGO][0], 0.0f);
    f << s.numZones << std::endl;
    for (const auto &kv : hc.m) f << kv.first << " " << kv.second << std::endl;
    f << "GOGOLocalMatrix" << std::endl;
    f << "maxnumpGOInGOtoGO " << hc.maxGABAIn << std::endl;
    f << "maxnumpGOOutGOtoGO " << hc.maxGABAOut << std::endl;
    for (uint32_t i = 0; i < hc.maxGABAOut; i++)
        f << hc.gogoGABA[i][0] << " " << hc.gogoGABA[i][1] << " " << hc.gogoGABA[i][2] << std::endl;
    f << "GOGOCoupLocalMatrix" << std::endl;
    f << "maxnumpGOCoupInGOtoGO " << hc.maxCoupIn << std::endl;
    f << "maxnumpGOCoupOutGOtoGO " << hc.maxCoupOut << std::endl;
    for (uint32_t i = 0; i < hc.maxCoupOut; i++)
        f << hc.gogoCoup[i][0] << " " << hc.gogoCoup[i][1] << " " << hc.gogoCoup[i][2] << std::endl;
    f << "connectivityParamEnd 1" << std::endl;
    for (const auto &kv : st->act) f << kv.first << " " << kv.second << std::endl;
    f << "activityParamEnd 1" << std::endl;

    const std::vector<Sec> all = sections(r, s.numPopHistBinsPC);
    auto emit = [&](size_t i, int z) -> int {
        const Sec &sc = all[i];
        const uint64_t bytes = sc.rows * sc.cols * sc.elem;
        std::vector<uint8_t> buf;
        if (sc.field >= 0) {
            size_t e, rr, cc;
            int zoned, kind;
            cbm_field_shape(sc.field, &st->s, &e, &rr, &cc, &zoned, &kind);
            const std::vector<uint8_t> &src = st->f[sc.field][sc.zoned ? z : 0];
            if (rr * cc * e == bytes) {
                buf = src;
            } else {
                buf.assign(bytes, 0);
                if (rr != sc.rows || e != sc.elem ||
                    !repack(src.data(), rr, cc, buf.data(), sc.cols, sc.elem, isIndexList(sc.name) ? UINT32_MAX : 0))
                    return fail(CBM_E_STATE, std::string("array ") + sc.name + " does not fit the file layout");
            }
        } else {
            const std::string key = sc.name;
            auto it = st->blob[sc.zoned ? z : 0].find(key == "pBCfromPCtoBC#2" ? "pBCfromPCtoBC" : key);
            if (it != st->blob[sc.zoned ? z : 0].end() && it->second.size() == bytes)
                buf = it->second;
            else if (key == "pBCfromPCtoBC#2")
                buf = synth(st, all[i - 3], r, z);  // the same array as pBCfromPCtoBC
            else
                buf = synth(st, sc, r, z);
            buf.resize(bytes, 0);
        }
        f.write(reinterpret_cast<const char *>(buf.data()), (std::streamsize)bytes);
        return CBM_OK;
    };
    for (int g = G_INNET_CON; g <= G_MZONE_ACT; g++) {
        const bool perZone = g == G_MZONE_CON || g == G_MZONE_ACT;
        for (int z = 0; z < (perZone ? (int)s.numZones : 1); z++)
            for (size_t i = 0; i < all.size(); i++)
                if (groupOf(all, i) == g) {
                    int rc = emit(i, z);
                    if (rc) return rc;
                }
    }
    f.flush();
    if (!f) return fail(CBM_E_ARG, "checkpoint write failed");
    return CBM_OK;
}

int readCheckpoint(std::istream &f, cbm_state **out)
{
    *out = nullptr;
    int nz = 0;
    f >> nz;
    if (!f || nz < 1 || nz > kMaxZones) return fail(CBM_E_UNSUPPORTED, "zone count must be 1..8");
    // ConnectivityParams(fstream&) (connectivityparams.cpp:12-88)
    ConParams cp;
    std::string key;
    bool hasCoup = false, ended = false;
    while (f >> key) {
        uint32_t v;
        if (key == "connectivityParamEnd") {
            f >> v;
            ended = true;
            break;
        }
        if (key == "GOGOLocalMatrix" || key == "GOGOCoupLocalMatrix") {
            std::string k2;
            uint32_t nin, nout;
            f >> k2 >> nin >> k2 >> nout;
            std::vector<std::array<float, 3>> rows(nout);
            for (auto &rw : rows) f >> rw[0] >> rw[1] >> rw[2];
            if (key == "GOGOLocalMatrix") {
                cp.maxGABAIn = nin;
                cp.maxGABAOut = nout;
                cp.gogoGABA = rows;
            } else {
                hasCoup = true;
                cp.maxCoupIn = nin;
                cp.maxCoupOut = nout;
                cp.gogoCoup = rows;
            }
        } else {
            f >> v;
            cp.m[key] = v;
        }
    }
    if (!ended) return fail(CBM_E_ARG, "no connectivityParamEnd in the checkpoint header");
    if (!hasCoup) cp.gogoCoup.assign(1, {1.f, 1.f, 0.f});
    // ActivityParams(fstream&) (activityparams.cpp:12-42)
    ActMap am;
    ended = false;
    float fv;
    while (f >> key >> fv) {
        if (key == "activityParamEnd") {
            ended = true;
            break;
        }
        am.m[key] = fv;
    }
    if (!ended) return fail(CBM_E_ARG, "no activityParamEnd in the checkpoint header");
    f.seekg(1, std::ios::cur);  // CBMState(fstream&) skips the line end (cbmstate.cpp:19)

    cbm_state *st = new cbm_state();
    st->con = cp;
    st->act = am.m;
    st->fromCheckpoint = true;
    deriveAct(am, st->p);
    const RefSizes r = deriveRef(cp);
    cbm_sizes &s = st->s;
    s.grX = 1u << cp.get("grXP2");
    s.grY = 1u << cp.get("grYP2");
    s.numGR = r.numGR;
    s.goX = 1u << cp.get("goXP2");
    s.goY = 1u << cp.get("goYP2");
    s.numGO = r.numGO;
    s.numMF = r.numMF;
    s.numSC = r.numSC;
    s.numPC = r.numPC;
    s.numBC = r.numBC;
    s.numNC = r.numNC;
    s.numIO = r.numIO;
    s.numZones = (uint32_t)nz;
    s.maxnumpGRfromMFtoGR = r.grFromGO;  // the file's MF->GR rows have maxnumpGRfromGOtoGR columns
    s.maxnumpGRfromGOtoGR = r.grFromGO;
    s.maxnumpGRfromGRtoGO = r.grToGO;
    s.maxnumpMFfromMFtoGO = r.mfToGO;
    s.maxnumpGOGABAOutGOGO = r.gabaOut;
    s.maxnumpGOCoupInGOGO = r.coupIn;
    s.numpPCfromGRtoPC = r.pcFromGR;
    s.numpBCfromGRtoBC = r.numGR / r.numBC;
    s.numpSCfromGRtoSC = r.numGR / r.numSC;
    s.numpBCfromBCtoPC = r.bcToPC;
    s.numpPCfromPCtoBC = r.pcToBC;
    s.numpPCfromSCtoPC = r.pcFromSC;
    s.numpNCfromPCtoNC = r.ncFromPC;
    s.numpNCfromMFtoNC = r.ncFromMF;
    s.numpIOfromNCtoIO = r.ioFromNC;
    s.numpIOInIOIO = r.ioIn;
    s.numPopHistBinsPC = st->p.numPopHistBinsPC;
    st->grPFVel = std::max(1u, cp.get("grPFVelInGRXPerTStep"));
    st->grAFDelay = cp.get("grAFDelayInTStep");
    st->msPerStep = std::max(1u, (uint32_t)st->p.msPerTimeStep);
    allocFields(st);

    const std::vector<Sec> all = sections(r, s.numPopHistBinsPC);
    for (int g = G_INNET_CON; g <= G_MZONE_ACT; g++) {
        const bool perZone = g == G_MZONE_CON || g == G_MZONE_ACT;
        for (int z = 0; z < (perZone ? nz : 1); z++)
            for (size_t i = 0; i < all.size(); i++) {
                if (groupOf(all, i) != g) continue;
                const Sec &sc = all[i];
                const uint64_t bytes = sc.rows * sc.cols * sc.elem;
                std::vector<uint8_t> buf(bytes);
                f.read(reinterpret_cast<char *>(buf.data()), (std::streamsize)bytes);
                if (!f) {
                    delete st;
                    return fail(CBM_E_ARG, std::string("checkpoint ends inside ") + sc.name);
                }
                if (sc.field >= 0) {
                    std::vector<uint8_t> &dst = st->f[sc.field][sc.zoned ? z : 0];
                    if (dst.size() != bytes) {
                        delete st;
                        return fail(CBM_E_UNSUPPORTED, std::string("array ") + sc.name + " has an unexpected size");
                    }
                    dst = buf;
                } else {
                    // the second copy of pBCfromPCtoBC overwrites the first, as on a reference read
                    const std::string k = std::string(sc.name) == "pBCfromPCtoBC#2" ? "pBCfromPCtoBC" : sc.name;
                    std::vector<uint8_t> &b = st->blob[sc.zoned ? z : 0][k];
                    if (b.size() == buf.size() || b.empty())
                        b = buf;
                    else
                        std::memcpy(b.data(), buf.data(), std::min(b.size(), buf.size()));
                }
            }
    }
    // Not in the file.  pPCfromPCtoBC is rebuilt as connectPCtoBC makes it
    // (mzoneconnectivitystate.cpp:241-255); the reference leaves it
    // unset after a read.  Everything else starts as a new state does.
    for (int z = 0; z < nz; z++) {
        uint32_t *pcbc = st->F<uint32_t>(CBM_F_P_PC_FROM_PC_TO_BC, z);
        const uint32_t ratio = s.numBC / s.numPC;
        for (uint32_t i = 0; i < s.numPC; i++)
            for (uint32_t j = 0; j < s.numpPCfromPCtoBC; j++)
                pcbc[(size_t)i * s.numpPCfromPCtoBC + j] = (uint32_t)wrap((int)(i * ratio) - 6 + (int)j, (int)s.numBC);
        float steps[2] = {st->p.synLTDStepSizeGRtoPC, st->p.synLTPStepSizeGRtoPC};
        std::memcpy(st->f[CBM_F_GRPC_PLAST_STEPS][z].data(), steps, 8);
        st->F<uint64_t>(CBM_F_ZONE_RNG_SKIP, z)[0] = s.numGR;
    }
    fillF32(st->f[CBM_F_INPUT_GOGABA_SYNDEP_GO][0], 1.0f);
    fillF32(st->f[CBM_F_GO_GABA_OUT_SYNSCALE_GOGO][0], 0.0f);
    std::memcpy(st->f[CBM_F_GINC_GR_TO_GO][0].data(), &st->p.gIncGRtoGO, 4);
    *out = st;
    return CBM_OK;
}

}  // namespace

extern "C" {

int cbm_state_write_checkpoint(cbm_state *st, const char *path)
{
    if (!st || !path) return fail(CBM_E_ARG, "null argument");
    std::ofstream f(path, std::ios::out | std::ios::binary | std::ios::trunc);
    if (!f) return fail(CBM_E_ARG, std::string("cannot open ") + path);
    return writeCheckpoint(st, f);
}

int cbm_state_read_checkpoint(const char *path, cbm_state **out)
{
    if (!path || !out) return fail(CBM_E_ARG, "null argument");
    std::ifstream f(path, std::ios::in | std::ios::binary);
    if (!f) return fail(CBM_E_ARG, std::string("cannot open ") + path);
    return readCheckpoint(f, out);
}

int cbm_state_checkpoint_to_memory(cbm_state *st, void **data, size_t *bytes)
{
    if (!st || !data || !bytes) return fail(CBM_E_ARG, "null argument");
    std::ostringstream o(std::ios::out | std::ios::binary);
    int rc = writeCheckpoint(st, o);
    if (rc) return rc;
    const std::string b = o.str();
    *data = std::malloc(b.size() ? b.size() : 1);
    if (!*data) return fail(CBM_E_NOMEM, "out of host memory");
    std::memcpy(*data, b.data(), b.size());
    *bytes = b.size();
    return CBM_OK;
}

int cbm_state_checkpoint_from_memory(const void *data, size_t bytes, cbm_state **out, size_t *consumed)
{
    if (!data || !out) return fail(CBM_E_ARG, "null argument");
    std::istringstream in(std::string(static_cast<const char *>(data), bytes), std::ios::in | std::ios::binary);
    int rc = readCheckpoint(in, out);
    if (rc) return rc;
    if (consumed) {
        const std::streamoff pos = in.tellg();
        *consumed = pos < 0 ? bytes : (size_t)pos;
    }
    return CBM_OK;
}

void cbm_state_free_memory(void *p) { std::free(p); }

}  // extern "C"

// cbmsimcore.cpp -- the reference's C++ core API over the two C-ABIs.
//
//   CBMState     -> cbm_state_* (include/cbm_state.h)
//   CBMSimCore   -> cbm_* (include/cbm_hip.h): construct = cbm_create +
//                   cbm_state_load_into, calcActivity = cbm_step(1)
//   InNetInterface / MZoneInterface exports -> cbm_read into host buffers that
//                   stay valid until the next call of the same export
//                   (the reference returns host arrays it refreshes the same way,
//                   CBM_CORE_LIB_MGPU_DEV/src/innetmodules/innet.cpp:679-798).
//
// Errors are reported by throwing std::runtime_error with the library's message.
#include <CBMCoreInclude/interface/cbmsimcore.h>

#include <algorithm>
#include <climits>
#include <cstring>
#include <ctime>
#include <iterator>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include <cbm_hip.h>
#include "../state/sfmt.hpp"

namespace {

void hipCheck(int rc, const char *what)
{
    if (rc != CBM_OK) throw std::runtime_error(std::string(what) + ": " + cbm_last_error());
}

void stateCheck(int rc, const char *what)
{
    if (rc != CBM_OK) throw std::runtime_error(std::string(what) + ": " + cbm_state_last_error());
}

std::string slurp(std::fstream &f)
{
    std::ostringstream o;
    o << f.rdbuf();
    return o.str();
}

uint32_t conKey(const std::string &text, const char *key, uint32_t dflt)
{
    std::istringstream in(text);
    std::string k;
    while (in >> k) {
        if (k == key) {
            uint32_t v;
            if (in >> v) return v;
        }
    }
    return dflt;
}

}  // namespace

// ---------------------------------------------------------------- CBMState
CBMState::CBMState(std::fstream &actPFile, std::fstream &conPFile, unsigned int nZones) : st(nullptr)
{
    cbm::Sfmt g((int)time(0));
    int innet = g.IRandom(0, INT_MAX);
    std::vector<int> cr(nZones), ar(nZones);
    for (unsigned int i = 0; i < nZones; i++) {
        cr[i] = g.IRandom(0, INT_MAX);
        ar[i] = g.IRandom(0, INT_MAX);
    }
    build(slurp(actPFile), slurp(conPFile), nZones, innet, cr.data(), ar.data());
}

CBMState::CBMState(std::fstream &actPFile, std::fstream &conPFile, unsigned int nZones, int innetCRSeed,
                   int *mzoneCRSeed, int *mzoneARSeed)
    : st(nullptr)
{
    build(slurp(actPFile), slurp(conPFile), nZones, innetCRSeed, mzoneCRSeed, mzoneARSeed);
}

void CBMState::build(const std::string &act, const std::string &con, unsigned int nZones, int innetCRSeed,
                     const int *mzoneCRSeed, const int *mzoneARSeed)
{
    if (nZones < 1 || nZones > 8) throw std::runtime_error("CBMState: numZones must be 1..8");
    cbm_gen_spec g;
    std::memset(&g, 0, sizeof(g));
    // a parameter file with a glomerulus layer (CBM_Params/conParams.txt) gets
    // the reference's own generator; one without it the synthetic one
    g.kind = conKey(con, "glXP2", UINT32_MAX) != UINT32_MAX ? CBM_GEN_REFERENCE : CBM_GEN_UNIFORM;
    g.numZones = nZones;
    g.innetSeed = innetCRSeed;
    // fan-ins of the benchmark networks; the GR->GO out-degree is capped by the
    // file's maxnumpGRfromGRtoGO
    g.nMFperGR = 4;
    g.nGOperGR = 3;
    g.nGOoutPerGR = std::min<uint32_t>(2, conKey(con, "maxnumpGRfromGRtoGO", 2));
    g.nMFperGO = conKey(con, "maxnumpGOfromGLtoGO", 16);
    for (unsigned int z = 0; z < nZones; z++) {
        g.mzoneConSeed[z] = mzoneCRSeed ? mzoneCRSeed[z] : 0;
        g.mzoneRngSeed[z] = mzoneARSeed ? mzoneARSeed[z] : 0;
    }
    stateCheck(cbm_state_create(act.c_str(), con.c_str(), &g, &st), "cbm_state_create");
    cbm_state_sizes(st, &sz);
    cbm_state_params(st, &ap);
}

CBMState::CBMState(std::fstream &infile) : st(nullptr)
{
    const std::streampos start = infile.tellg();
    const std::string buf((std::istreambuf_iterator<char>(infile)), std::istreambuf_iterator<char>());
    size_t used = 0;
    stateCheck(cbm_state_checkpoint_from_memory(buf.data(), buf.size(), &st, &used), "CBMState(fstream&)");
    infile.clear();
    infile.seekg(start + (std::streamoff)used);
    cbm_state_sizes(st, &sz);
    cbm_state_params(st, &ap);
}

void CBMState::writeState(std::fstream &outfile)
{
    void *data = nullptr;
    size_t n = 0;
    stateCheck(cbm_state_checkpoint_to_memory(st, &data, &n), "CBMState::writeState");
    outfile.write(static_cast<const char *>(data), (std::streamsize)n);
    cbm_state_free_memory(data);
}

CBMState::~CBMState()
{
    deleteViews();
    if (st) cbm_state_destroy(st);
}

ct_uint32_t CBMState::getNumZones() { return sz.numZones; }

void *CBMState::field(int f, int zone, size_t *bytes)
{
    void *p = nullptr;
    size_t b = 0;
    stateCheck(cbm_state_field(st, f, zone, &p, &b), "cbm_state_field");
    if (bytes) *bytes = b;
    return p;
}

// ---------------------------------------------------------------- views
namespace {

template <typename T>
const T *readInto(cbm_handle *h, std::vector<T> &buf, int field, int zone, const cbm_sizes &s)
{
    buf.resize(cbm_field_bytes(field, &s) / sizeof(T));
    hipCheck(cbm_read(h, field, zone, buf.data(), buf.size() * sizeof(T)), "cbm_read");
    return buf.data();
}

}  // namespace

class CBMInNetView : public InNetInterface
{
public:
    CBMInNetView(cbm_handle *h, const cbm_sizes &s) : h(h), s(s) {}
    void setGIncGRtoGO(float inc) override { hipCheck(cbm_set_ginc_grgo(h, inc), "setGIncGRtoGO"); }
    void resetGIncGRtoGO() override { hipCheck(cbm_reset_ginc_grgo(h), "resetGIncGRtoGO"); }
    const ct_uint8_t *exportAPSC() override { return readInto(h, apSC, CBM_F_AP_SC, 0, s); }
    const ct_uint8_t *exportAPGO() override { return readInto(h, apGO, CBM_F_AP_GO, 0, s); }
    const ct_uint8_t *exportAPGR() override { return readInto(h, apGR, CBM_F_AP_GR, 0, s); }
    const ct_uint8_t *exportHistMF() override { return readInto(h, histMF, CBM_F_HIST_MF, 0, s); }
    const ct_uint32_t *exportAPBufMF() override { return readInto(h, bufMF, CBM_F_APBUF_MF, 0, s); }
    const ct_uint32_t *exportAPBufGR() override { return readInto(h, bufGR, CBM_F_APBUF_GR, 0, s); }
    const ct_uint32_t *exportAPBufGO() override { return readInto(h, bufGO, CBM_F_APBUF_GO, 0, s); }
    const ct_uint32_t *exportAPBufSC() override { return readInto(h, bufSC, CBM_F_APBUF_SC, 0, s); }
    const float *exportVmGR() override { return readInto(h, vGR, CBM_F_VM_GR, 0, s); }
    const float *exportVmGO() override { return readInto(h, vGO, CBM_F_VM_GO, 0, s); }
    const float *exportVmSC() override { return readInto(h, vSC, CBM_F_VM_SC, 0, s); }
    const float *exportGESumGR() override { return readInto(h, gE, CBM_F_G_MF_SUM_GR, 0, s); }
    const float *exportGISumGR() override { return readInto(h, gI, CBM_F_G_GO_SUM_GR, 0, s); }
    const ct_uint32_t *exportSumGRInputGO() override { return readInto(h, grIn, CBM_F_SUM_GR_INPUT_GO, 0, s); }
    const float *exportSumGOInputGO() override { return readInto(h, goIn, CBM_F_SUM_GO_INPUT_GO, 0, s); }
    const float *exportGOOutSynScaleGOGO() override
    {
        return readInto(h, scale, CBM_F_GO_GABA_OUT_SYNSCALE_GOGO, 0, s);
    }

private:
    cbm_handle *h;
    cbm_sizes s;
    std::vector<ct_uint8_t> apSC, apGO, apGR, histMF;
    std::vector<ct_uint32_t> bufMF, bufGR, bufGO, bufSC, grIn;
    std::vector<float> vGR, vGO, vSC, gE, gI, goIn, scale;
};

class CBMZoneView : public MZoneInterface
{
public:
    CBMZoneView(cbm_handle *h, const cbm_sizes &s, int z) : h(h), s(s), z(z) {}
    void setGRPCPlastSteps(float ltd, float ltp) override
    {
        hipCheck(cbm_set_grpc_steps(h, z, ltd, ltp), "setGRPCPlastSteps");
    }
    void resetGRPCPlastSteps() override { hipCheck(cbm_reset_grpc_steps(h, z), "resetGRPCPlastSteps"); }
    const ct_uint8_t *exportAPBC() override { return readInto(h, apBC, CBM_F_AP_BC, z, s); }
    const ct_uint8_t *exportAPPC() override { return readInto(h, apPC, CBM_F_AP_PC, z, s); }
    const ct_uint8_t *exportAPIO() override { return readInto(h, apIO, CBM_F_AP_IO, z, s); }
    const ct_uint8_t *exportAPNC() override { return readInto(h, apNC, CBM_F_AP_NC, z, s); }
    const ct_uint32_t *exportAPBufBC() override { return readInto(h, bufBC, CBM_F_APBUF_BC, z, s); }
    const ct_uint32_t *exportAPBufPC() override { return readInto(h, bufPC, CBM_F_APBUF_PC, z, s); }
    const ct_uint32_t *exportAPBufIO() override { return readInto(h, bufIO, CBM_F_APBUF_IO, z, s); }
    const ct_uint32_t *exportAPBufNC() override { return readInto(h, bufNC, CBM_F_APBUF_NC, z, s); }
    const float *exportVmBC() override { return readInto(h, vBC, CBM_F_VM_BC, z, s); }
    const float *exportVmPC() override { return readInto(h, vPC, CBM_F_VM_PC, z, s); }
    const float *exportVmNC() override { return readInto(h, vNC, CBM_F_VM_NC, z, s); }
    const float *exportVmIO() override { return readInto(h, vIO, CBM_F_VM_IO, z, s); }

private:
    cbm_handle *h;
    cbm_sizes s;
    int z;
    std::vector<ct_uint8_t> apBC, apPC, apIO, apNC;
    std::vector<ct_uint32_t> bufBC, bufPC, bufIO, bufNC;
    std::vector<float> vBC, vPC, vNC, vIO;
};

// ---------------------------------------------------------------- CBMSimCore
CBMSimCore::CBMSimCore(CBMState *state, int gpuIndStart, int numGPUP2)
{
    // CBMSimCore.cpp:19-35: zone seeds drawn from a time-seeded generator
    cbm::Sfmt g((int)time(0));
    std::vector<int> seeds(state->getNumZones());
    for (auto &s : seeds) s = g.IRandom(0, INT_MAX);
    construct(state, seeds.data(), gpuIndStart, numGPUP2);
}

CBMSimCore::CBMSimCore(CBMState *state, int *mzoneRSeed, int gpuIndStart, int numGPUP2)
{
    construct(state, mzoneRSeed, gpuIndStart, numGPUP2);
}

CBMSimCore::CBMSimCore(cbm_handle *handle, const cbm_sizes &sizes)
    : simState(nullptr), h(handle), ownsHandle(false), numZones(sizes.numZones)
{
    if (!handle) throw std::runtime_error("CBMSimCore: null handle");
    makeViews(sizes);
}

void CBMSimCore::makeViews(const cbm_sizes &s)
{
    inputNet = new CBMInNetView(h, s);
    zones = new CBMZoneView *[numZones];
    zoneList = new MZoneInterface *[numZones];
    for (unsigned int z = 0; z < numZones; z++) {
        zones[z] = new CBMZoneView(h, s, (int)z);
        zoneList[z] = zones[z];
    }
}

void CBMSimCore::construct(CBMState *state, const int *mzoneRSeed, int gpuIndStart, int numGPUP2)
{
    if (numGPUP2 > 0) throw std::runtime_error("CBMSimCore: one GPU per network (numGPUP2 must be -1 or 0)");
    simState = state;
    numZones = state->getNumZones();
    h = nullptr;
    ownsHandle = true;
    cbm_options o;
    std::memset(&o, 0, sizeof(o));
    o.grgoLag = 2;
    o.device = gpuIndStart < 0 ? 0 : gpuIndStart;
    hipCheck(cbm_create(&state->sizes(), &state->params(), &o, &h), "cbm_create");
    // MZone's generator seed (mzone.cpp:18) comes from the core, not the state
    for (unsigned int z = 0; z < numZones; z++) {
        size_t b = 0;
        int32_t *seed = static_cast<int32_t *>(state->field(CBM_F_ZONE_RNG_SEED, (int)z, &b));
        *seed = mzoneRSeed[z];
    }
    stateCheck(cbm_state_load_into(state->handle(), h), "cbm_state_load_into");
    makeViews(state->sizes());
}

CBMSimCore::~CBMSimCore()
{
    delete inputNet;
    for (unsigned int z = 0; z < numZones; z++) delete zones[z];
    delete[] zones;
    delete[] zoneList;
    if (h && ownsHandle) cbm_destroy(h);
}

void CBMSimCore::calcActivity() { hipCheck(cbm_step(h, 1), "calcActivity"); }

void CBMSimCore::updateMFInput(const ct_uint8_t *mfIn) { hipCheck(cbm_set_mf(h, mfIn), "updateMFInput"); }

void CBMSimCore::updateErrDrive(unsigned int zoneN, float errDriveRelative)
{
    hipCheck(cbm_set_err_drive(h, (int)zoneN, errDriveRelative), "updateErrDrive");
}

void CBMSimCore::writeToState()
{
    if (!simState) throw std::runtime_error("CBMSimCore::writeToState: this core has no CBMState");
    const cbm_sizes &s = simState->sizes();
    for (int f = 0; f < CBM_F_COUNT; f++) {
        size_t e, r, c;
        int zoned, kind;
        if (cbm_field_shape(f, &s, &e, &r, &c, &zoned, &kind) != CBM_OK || kind != CBM_KIND_STATE) continue;
        for (unsigned int z = 0; z < (zoned ? numZones : 1u); z++) {
            size_t b = 0;
            void *p = simState->field(f, (int)z, &b);
            if (b) hipCheck(cbm_read(h, f, (int)z, p, b), "writeToState");
        }
    }
}

void CBMSimCore::writeToState(std::fstream &outfile)
{
    writeToState();
    simState->writeState(outfile);
    for (int f : {CBM_F_INPUT_GOGABA_SYNDEP_GO, CBM_F_GO_GABA_OUT_SYNSCALE_GOGO}) {
        size_t b = 0;
        void *p = simState->field(f, 0, &b);
        hipCheck(cbm_write(h, f, 0, p, b), "writeToState(fstream&)");
    }
}

InNetInterface *CBMSimCore::getInputNet() { return inputNet; }

MZoneInterface **CBMSimCore::getMZoneList() { return zoneList; }

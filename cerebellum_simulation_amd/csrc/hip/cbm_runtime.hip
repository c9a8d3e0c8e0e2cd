// cbm_runtime.hip -- host side of the C-ABI in include/cbm_hip.h.
//
// Owns one network on one GPU: stages the CBMState-layout fields written by
// the caller, validates them, packs granule connectivity into the device
// layout of the GR kernel (four 16-bit indices per GR, 8-bit delay bits),
// builds the gather lists of the small-population kernel, and launches the
// two kernels of each step on the handle's stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../state/sfmt.hpp"
#include "cbm_dev.hpp"

namespace cbmdev {
template <int CPT>
__global__ void gr_step_kernel(const GrArgs a);
__global__ void small_step_kernel(const SmallArgs *args, const SmallStep st);
}  // namespace cbmdev

using namespace cbmdev;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &m)
{
    g_err = m;
    return code;
}

#define HIP_TRY(expr)                                                                        \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess)                                                                \
            return fail(CBM_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));     \
    } while (0)

bool isPow2(uint32_t x) { return x && !(x & (x - 1)); }

// fields that the small kernel reads from device storage besides the state
bool deviceField(int f)
{
    size_t e, r, c;
    int zoned, kind;
    cbm_sizes dummy{};
    if (cbm_field_shape(f, &dummy, &e, &r, &c, &zoned, &kind) != CBM_OK) return false;
    switch (f) {
    case CBM_F_P_NC_FROM_MF_TO_NC: case CBM_F_P_PC_FROM_SC_TO_PC: case CBM_F_P_NC_FROM_PC_TO_NC:
    case CBM_F_P_IO_FROM_NC_TO_IO: case CBM_F_P_IO_IN_IOIO: case CBM_F_NUMP_GO_COUP_IN_GOGO:
    case CBM_F_P_GO_COUP_IN_GOGO: case CBM_F_SUM_GR_INPUT_GO: case CBM_F_SUM_GO_INPUT_GO:
        return true;
    case CBM_F_AP_GR: case CBM_F_INPUT_SUM_PF_PC: case CBM_F_INPUT_SUM_PF_SC: case CBM_F_ERR_DRIVE:
        return false;
    default:
        return kind == CBM_KIND_STATE;
    }
}

}  // namespace

struct cbm_handle {
    cbm_sizes s{};
    cbm_act_params p{};
    cbm_options o{};
    hipStream_t stream = nullptr;
    bool committed = false;
    int cpt = 4;

    std::vector<uint8_t> host[CBM_F_COUNT];  // staged fields, zones stacked
    bool written[CBM_F_COUNT] = {};
    void *dfield[CBM_F_COUNT] = {};

    uint2 *mfIn = nullptr, *goIn = nullptr, *goOut = nullptr;
    uint32_t *goOutDly = nullptr;
    uint8_t *bspDly = nullptr;
    float *chunk = nullptr, *pfpcRing = nullptr;
    uint32_t *bcRing = nullptr, *scRing = nullptr, *goRing = nullptr, *rngState = nullptr;
    uint8_t *dmf = nullptr;
    uint8_t *pinnedMF = nullptr;
    Csr mfToGoIn{}, gabaIn{}, bcToPcIn{}, pcToBcIn{};
    uint32_t anyCoupling = 0;
    SmallArgs *smallArgs = nullptr;  // device copy of the constant small-kernel arguments
    size_t smallLds = 0;             // dynamic LDS of the small kernel (synapse partials)

    uint64_t curTime = 0;
    float errDrive[kMaxZones] = {};
    float ltd[kMaxZones] = {}, ltp[kMaxZones] = {};
    float gIncGRtoGO = 0;

    std::vector<void *> allocs;
    size_t devBytes = 0;
    hipEvent_t ev[2] = {};  // GR-kernel timing pair used by cbm_profile_steps

    size_t zoneBytes(int f) const { return cbm_field_bytes(f, &s); }
    int zonesOf(int f) const
    {
        int zoned = 0;
        cbm_field_shape(f, &s, nullptr, nullptr, nullptr, &zoned, nullptr);
        return zoned ? (int)s.numZones : 1;
    }
    template <typename T>
    T *H(int f, int z = 0) { return reinterpret_cast<T *>(host[f].data() + (size_t)z * zoneBytes(f)); }

    int dalloc(void **p, size_t bytes)
    {
        if (bytes == 0) bytes = 4;
        hipError_t e = hipMalloc(p, bytes);
        if (e != hipSuccess) return fail(CBM_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
        allocs.push_back(*p);
        devBytes += bytes;
        HIP_TRY(hipMemset(*p, 0, bytes));
        return CBM_OK;
    }
    template <typename T>
    int upload(T **dst, const std::vector<T> &v)
    {
        int rc = dalloc((void **)dst, v.size() * sizeof(T));
        if (rc) return rc;
        if (!v.empty()) HIP_TRY(hipMemcpy(*dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
        return CBM_OK;
    }
};

namespace {

int checkIndex(const char *what, uint32_t v, uint32_t n)
{
    if (v >= n) return fail(CBM_E_ARG, std::string(what) + ": index out of range");
    return CBM_OK;
}

// delay mask -> bit index (kNever for an empty mask)
int maskBit(uint32_t m, uint8_t *out)
{
    if (m == 0) {
        *out = kNever;
        return CBM_OK;
    }
    if (!isPow2(m)) return fail(CBM_E_UNSUPPORTED, "delay masks must have a single bit set");
    uint8_t b = 0;
    while ((1u << b) != m) b++;
    *out = b;
    return CBM_OK;
}

int buildCsr(cbm_handle *h, uint32_t rows, const std::vector<std::vector<uint32_t>> &lists, Csr *out,
             std::vector<uint32_t> *ptrAcc, std::vector<uint32_t> *idxAcc)
{
    (void)h;
    (void)out;
    uint32_t base = (uint32_t)idxAcc->size();
    for (uint32_t r = 0; r < rows; r++) {
        ptrAcc->push_back(base);
        for (uint32_t v : lists[r]) idxAcc->push_back(v);
        base = (uint32_t)idxAcc->size();
    }
    ptrAcc->push_back(base);
    return CBM_OK;
}

int validateSizes(const cbm_sizes &s)
{
    if (s.numGR == 0 || s.numGR % 1024) return fail(CBM_E_UNSUPPORTED, "numGR must be a positive multiple of 1024");
    if (!isPow2(s.numGR)) return fail(CBM_E_UNSUPPORTED, "numGR must be a power of two");
    if (s.numMF < 4 || s.numMF % 4 || s.numMF > 0xFFFF) return fail(CBM_E_UNSUPPORTED, "numMF must be a multiple of 4 below 65535");
    if (s.numGO < 4 || s.numGO % 4 || s.numGO > 0xFFFF) return fail(CBM_E_UNSUPPORTED, "numGO must be a multiple of 4 below 65535");
    if (s.numZones < 1 || s.numZones > (uint32_t)kMaxZones) return fail(CBM_E_UNSUPPORTED, "numZones must be 1..8");
    if (!isPow2(s.numPC) || !isPow2(s.numBC) || !isPow2(s.numSC) || !isPow2(s.numIO) || !isPow2(s.numNC))
        return fail(CBM_E_UNSUPPORTED, "PC, BC, SC, IO and NC counts must be powers of two");
    if (s.numIO > (uint32_t)kMaxIO || s.numIO > s.numPC) return fail(CBM_E_UNSUPPORTED, "numIO must be <= 64 and <= numPC");
    if (s.numpPCfromGRtoPC * s.numPC != s.numGR || s.numpBCfromGRtoBC * s.numBC != s.numGR ||
        s.numpSCfromGRtoSC * s.numSC != s.numGR)
        return fail(CBM_E_ARG, "PF block sizes must equal numGR / population");
    if (s.numpSCfromGRtoSC < 64 || s.numpBCfromGRtoBC < 64 || s.numpPCfromGRtoPC < 64)
        return fail(CBM_E_UNSUPPORTED, "PF blocks of SC, BC and PC must hold at least 64 GRs");
    if (s.maxnumpGRfromMFtoGR > 4 || s.maxnumpGRfromGOtoGR > 4 || s.maxnumpGRfromGRtoGO > 4)
        return fail(CBM_E_UNSUPPORTED, "GR fan-in and GR->GO fan-out above 4 are not supported");
    if (s.numpIOfromNCtoIO != s.numNC) return fail(CBM_E_ARG, "numpIOfromNCtoIO must equal numNC");
    return CBM_OK;
}

}  // namespace

extern "C" {

int cbm_abi_version(void) { return CBM_ABI_VERSION; }

const char *cbm_last_error(void) { return g_err.c_str(); }

int cbm_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int cbm_create(const cbm_sizes *sizes, const cbm_act_params *params, const cbm_options *opts,
               cbm_handle **out)
{
    if (!sizes || !params || !out) return fail(CBM_E_ARG, "null argument");
    int rc = validateSizes(*sizes);
    if (rc) return rc;
    if (params->tsPerHistBinGR == 0 || params->tsPerPopHistBinPC == 0 || params->numTSinMFHist == 0 ||
        params->numPopHistBinsPC == 0)
        return fail(CBM_E_ARG, "derived step counts must be positive");
    if (params->numPopHistBinsPC != sizes->numPopHistBinsPC)
        return fail(CBM_E_ARG, "numPopHistBinsPC differs between sizes and params");
    int ndev = cbm_device_count();
    if (ndev <= 0) return fail(CBM_E_HIP, "no HIP device available");
    cbm_handle *h = new cbm_handle();
    h->s = *sizes;
    h->p = *params;
    if (opts) h->o = *opts;
    if (h->o.grgoLag != 1) h->o.grgoLag = 2;
    if (h->o.device < 0 || h->o.device >= ndev) {
        delete h;
        return fail(CBM_E_ARG, "device ordinal out of range");
    }
    h->cpt = h->o.grCellsPerThread ? (int)h->o.grCellsPerThread : 4;
    if (h->cpt != 1 && h->cpt != 2 && h->cpt != 4) {
        delete h;
        return fail(CBM_E_ARG, "grCellsPerThread must be 1, 2 or 4");
    }
    hipError_t e = hipSetDevice(h->o.device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipHostMalloc((void **)&h->pinnedMF, h->s.numMF, hipHostMallocDefault);
    if (e != hipSuccess) {
        std::string m = hipGetErrorString(e);
        delete h;
        return fail(CBM_E_HIP, "stream/pinned setup: " + m);
    }
    for (int f = 0; f < CBM_F_COUNT; f++) {
        size_t b = cbm_field_bytes(f, &h->s);
        h->host[f].assign(b * h->zonesOf(f), 0);
    }
    for (uint32_t z = 0; z < h->s.numZones; z++) {
        h->ltd[z] = params->synLTDStepSizeGRtoPC;
        h->ltp[z] = params->synLTPStepSizeGRtoPC;
    }
    h->gIncGRtoGO = params->gIncGRtoGO;
    *out = h;
    return CBM_OK;
}

int cbm_destroy(cbm_handle *h)
{
    if (!h) return fail(CBM_E_ARG, "null handle");
    (void)hipSetDevice(h->o.device);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    for (void *p : h->allocs) (void)hipFree(p);
    if (h->pinnedMF) (void)hipHostFree(h->pinnedMF);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return CBM_OK;
}

static int fieldCheck(cbm_handle *h, int field, int zone, size_t bytes, int *zoned, int *kind)
{
    size_t e, r, c;
    if (field < 0 || field >= CBM_F_COUNT ||
        cbm_field_shape(field, &h->s, &e, &r, &c, zoned, kind) != CBM_OK)
        return fail(CBM_E_ARG, "unknown field");
    if (*zoned ? (zone < 0 || zone >= (int)h->s.numZones) : false) return fail(CBM_E_ARG, "zone out of range");
    if (bytes != e * r * c) return fail(CBM_E_ARG, "byte count does not match the field shape");
    return CBM_OK;
}

// ring slot of pipeline row r (0 = last step)
static uint32_t ringSlot(const cbm_handle *h, int r)
{
    return (uint32_t)((h->curTime + (uint64_t)kRing * 4 - (uint64_t)r) % kRing);
}

int cbm_write(cbm_handle *h, int field, int zone, const void *src, size_t bytes)
{
    if (!h || (!src && bytes)) return fail(CBM_E_ARG, "null argument");
    int zoned, kind;
    int rc = fieldCheck(h, field, zone, bytes, &zoned, &kind);
    if (rc) return rc;
    if (!zoned) zone = 0;
    const size_t off = (size_t)zone * bytes;
    (void)hipSetDevice(h->o.device);
    if (!h->committed) {
        if (bytes) std::memcpy(h->host[field].data() + off, src, bytes);
        h->written[field] = true;
        if (field == CBM_F_CUR_TIME) std::memcpy(&h->curTime, src, 8);
        if (field == CBM_F_GINC_GR_TO_GO) std::memcpy(&h->gIncGRtoGO, src, 4);
        if (field == CBM_F_ERR_DRIVE) std::memcpy(&h->errDrive[zone], src, 4);
        if (field == CBM_F_GRPC_PLAST_STEPS) {
            std::memcpy(&h->ltd[zone], src, 4);
            std::memcpy(&h->ltp[zone], (const uint8_t *)src + 4, 4);
        }
        return CBM_OK;
    }
    if (kind == CBM_KIND_CON && field != CBM_F_ZONE_RNG_STATE)
        return fail(CBM_E_STATE, "connectivity cannot change after cbm_commit");
    HIP_TRY(hipStreamSynchronize(h->stream));
    switch (field) {
    case CBM_F_AP_GR: case CBM_F_INPUT_SUM_PF_PC: case CBM_F_INPUT_SUM_PF_SC:
        return CBM_OK;  // derived from other state; accepted and ignored
    case CBM_F_CUR_TIME: {
        // the pipeline rows keep their meaning: re-seat them around the new time
        std::vector<uint8_t> pf(cbm_field_bytes(CBM_F_PIPE_PFPC_SUM, &h->s) * h->s.numZones);
        std::vector<uint8_t> bc(cbm_field_bytes(CBM_F_PIPE_PFBC_SUM, &h->s)), sc(cbm_field_bytes(CBM_F_PIPE_PFSC_SUM, &h->s)),
            go(cbm_field_bytes(CBM_F_PIPE_GRGO_SUM, &h->s));
        const size_t pfz = cbm_field_bytes(CBM_F_PIPE_PFPC_SUM, &h->s);
        for (uint32_t z = 0; z < h->s.numZones; z++)
            if ((rc = cbm_read(h, CBM_F_PIPE_PFPC_SUM, (int)z, pf.data() + z * pfz, pfz))) return rc;
        if ((rc = cbm_read(h, CBM_F_PIPE_PFBC_SUM, 0, bc.data(), bc.size()))) return rc;
        if ((rc = cbm_read(h, CBM_F_PIPE_PFSC_SUM, 0, sc.data(), sc.size()))) return rc;
        if ((rc = cbm_read(h, CBM_F_PIPE_GRGO_SUM, 0, go.data(), go.size()))) return rc;
        std::memcpy(&h->curTime, src, 8);
        for (uint32_t z = 0; z < h->s.numZones; z++)
            if ((rc = cbm_write(h, CBM_F_PIPE_PFPC_SUM, (int)z, pf.data() + z * pfz, pfz))) return rc;
        if ((rc = cbm_write(h, CBM_F_PIPE_PFBC_SUM, 0, bc.data(), bc.size()))) return rc;
        if ((rc = cbm_write(h, CBM_F_PIPE_PFSC_SUM, 0, sc.data(), sc.size()))) return rc;
        return cbm_write(h, CBM_F_PIPE_GRGO_SUM, 0, go.data(), go.size());
    }
    case CBM_F_GINC_GR_TO_GO: std::memcpy(&h->gIncGRtoGO, src, 4); return CBM_OK;
    case CBM_F_ERR_DRIVE: std::memcpy(&h->errDrive[zone], src, 4); return CBM_OK;
    case CBM_F_GRPC_PLAST_STEPS:
        std::memcpy(&h->ltd[zone], src, 4);
        std::memcpy(&h->ltp[zone], (const uint8_t *)src + 4, 4);
        return CBM_OK;
    case CBM_F_MF_INPUT: HIP_TRY(hipMemcpy(h->dmf, src, bytes, hipMemcpyHostToDevice)); return CBM_OK;
    case CBM_F_ZONE_RNG_STATE:
        if (((const uint32_t *)src)[352] > 352) return fail(CBM_E_ARG, "generator index above 352");
        HIP_TRY(hipMemcpy(h->rngState + (size_t)zone * 353, src, bytes, hipMemcpyHostToDevice));
        return CBM_OK;
    case CBM_F_PIPE_PFPC_SUM: {
        const size_t row = (size_t)h->s.numPC * 4;
        for (int r = 0; r < 2; r++)
            HIP_TRY(hipMemcpy(h->pfpcRing + ((size_t)ringSlot(h, r) * h->s.numZones + zone) * h->s.numPC,
                              (const uint8_t *)src + r * row, row, hipMemcpyHostToDevice));
        return CBM_OK;
    }
    case CBM_F_PIPE_PFBC_SUM: case CBM_F_PIPE_PFSC_SUM: case CBM_F_PIPE_GRGO_SUM: {
        uint32_t *ring = field == CBM_F_PIPE_PFBC_SUM ? h->bcRing : field == CBM_F_PIPE_PFSC_SUM ? h->scRing : h->goRing;
        const uint32_t n = field == CBM_F_PIPE_PFBC_SUM ? h->s.numBC : field == CBM_F_PIPE_PFSC_SUM ? h->s.numSC : h->s.numGO;
        for (int r = 0; r < 2; r++)
            HIP_TRY(hipMemcpy(ring + (size_t)ringSlot(h, r) * n, (const uint8_t *)src + (size_t)r * n * 4,
                              (size_t)n * 4, hipMemcpyHostToDevice));
        HIP_TRY(hipMemset(ring + (size_t)ringSlot(h, -1) * n, 0, (size_t)n * 4));
        return CBM_OK;
    }
    default:
        break;
    }
    if (!h->dfield[field]) return fail(CBM_E_ARG, "field is not writable");
    HIP_TRY(hipMemcpy((uint8_t *)h->dfield[field] + off, src, bytes, hipMemcpyHostToDevice));
    return CBM_OK;
}

int cbm_commit(cbm_handle *h)
{
    if (!h) return fail(CBM_E_ARG, "null handle");
    if (h->committed) return fail(CBM_E_STATE, "already committed");
    const cbm_sizes &s = h->s;
    HIP_TRY(hipSetDevice(h->o.device));
    int rc;

    // ---- granule connectivity: validate and pack
    const uint32_t nGR = s.numGR;
    std::vector<uint2> mfIn(nGR), goIn(nGR), goOut(nGR);
    std::vector<uint32_t> goOutDly(nGR);
    std::vector<uint8_t> bsp(nGR);
    {
        const int32_t *nM = h->H<int32_t>(CBM_F_NUMP_GR_FROM_MF_TO_GR);
        const uint32_t *pM = h->H<uint32_t>(CBM_F_P_GR_FROM_MF_TO_GR);
        const int32_t *nG = h->H<int32_t>(CBM_F_NUMP_GR_FROM_GO_TO_GR);
        const uint32_t *pG = h->H<uint32_t>(CBM_F_P_GR_FROM_GO_TO_GR);
        const int32_t *nO = h->H<int32_t>(CBM_F_NUMP_GR_FROM_GR_TO_GO);
        const uint32_t *pO = h->H<uint32_t>(CBM_F_P_GR_FROM_GR_TO_GO);
        const uint32_t *dO = h->H<uint32_t>(CBM_F_P_GR_DELAYMASK_GR_TO_GO);
        const uint32_t *dB = h->H<uint32_t>(CBM_F_P_GR_DELAYMASK_GR_TO_BSP);
        const uint32_t kM = s.maxnumpGRfromMFtoGR, kG = s.maxnumpGRfromGOtoGR, kO = s.maxnumpGRfromGRtoGO;
        for (uint32_t i = 0; i < nGR; i++) {
            uint16_t a[4] = {kNoCell, kNoCell, kNoCell, kNoCell}, b[4] = {kNoCell, kNoCell, kNoCell, kNoCell},
                     c[4] = {kNoCell, kNoCell, kNoCell, kNoCell};
            uint8_t dl[4] = {kNever, kNever, kNever, kNever};
            if (nM[i] < 0 || (uint32_t)nM[i] > kM || nG[i] < 0 || (uint32_t)nG[i] > kG || nO[i] < 0 ||
                (uint32_t)nO[i] > kO)
                return fail(CBM_E_ARG, "GR fan-in/out count above its maximum");
            for (int k = 0; k < nM[i]; k++) {
                if ((rc = checkIndex("P_GR_FROM_MF_TO_GR", pM[(size_t)i * kM + k], s.numMF))) return rc;
                a[k] = (uint16_t)pM[(size_t)i * kM + k];
            }
            for (int k = 0; k < nG[i]; k++) {
                if ((rc = checkIndex("P_GR_FROM_GO_TO_GR", pG[(size_t)i * kG + k], s.numGO))) return rc;
                b[k] = (uint16_t)pG[(size_t)i * kG + k];
            }
            for (int k = 0; k < nO[i]; k++) {
                if ((rc = checkIndex("P_GR_FROM_GR_TO_GO", pO[(size_t)i * kO + k], s.numGO))) return rc;
                c[k] = (uint16_t)pO[(size_t)i * kO + k];
                if ((rc = maskBit(dO[(size_t)i * kO + k], &dl[k]))) return rc;
            }
            if ((rc = maskBit(dB[i], &bsp[i]))) return rc;
            mfIn[i] = make_uint2(a[0] | ((uint32_t)a[1] << 16), a[2] | ((uint32_t)a[3] << 16));
            goIn[i] = make_uint2(b[0] | ((uint32_t)b[1] << 16), b[2] | ((uint32_t)b[3] << 16));
            goOut[i] = make_uint2(c[0] | ((uint32_t)c[1] << 16), c[2] | ((uint32_t)c[3] << 16));
            goOutDly[i] = dl[0] | ((uint32_t)dl[1] << 8) | ((uint32_t)dl[2] << 16) | ((uint32_t)dl[3] << 24);
        }
    }

    // ---- gather lists of the small kernel
    std::vector<uint32_t> mfgPtr, mfgIdx, gabPtr, gabIdx, bcpPtr, bcpIdx, pcbPtr, pcbIdx;
    {
        std::vector<std::vector<uint32_t>> in(s.numGO), gab(s.numGO);
        const int32_t *n = h->H<int32_t>(CBM_F_NUMP_MF_FROM_MF_TO_GO);
        const uint32_t *pp = h->H<uint32_t>(CBM_F_P_MF_FROM_MF_TO_GO);
        for (uint32_t m = 0; m < s.numMF; m++) {
            if (n[m] < 0 || (uint32_t)n[m] > s.maxnumpMFfromMFtoGO) return fail(CBM_E_ARG, "MF->GO count above maximum");
            for (int j = 0; j < n[m]; j++) {
                uint32_t g = pp[(size_t)m * s.maxnumpMFfromMFtoGO + j];
                if ((rc = checkIndex("P_MF_FROM_MF_TO_GO", g, s.numGO))) return rc;
                in[g].push_back(m);
            }
        }
        const int32_t *ng = h->H<int32_t>(CBM_F_NUMP_GO_GABA_OUT_GOGO);
        const uint32_t *pg = h->H<uint32_t>(CBM_F_P_GO_GABA_OUT_GOGO);
        for (uint32_t i = 0; i < s.numGO; i++) {
            if (ng[i] < 0 || (uint32_t)ng[i] > s.maxnumpGOGABAOutGOGO) return fail(CBM_E_ARG, "GO GABA count above maximum");
            for (int j = 0; j < ng[i]; j++) {
                uint32_t g = pg[(size_t)i * s.maxnumpGOGABAOutGOGO + j];
                if ((rc = checkIndex("P_GO_GABA_OUT_GOGO", g, s.numGO))) return rc;
                gab[g].push_back(i);  // i ascending, j ascending: the reference's order
            }
        }
        buildCsr(h, s.numGO, in, nullptr, &mfgPtr, &mfgIdx);
        buildCsr(h, s.numGO, gab, nullptr, &gabPtr, &gabIdx);
        const int32_t *nc = h->H<int32_t>(CBM_F_NUMP_GO_COUP_IN_GOGO);
        const uint32_t *pc = h->H<uint32_t>(CBM_F_P_GO_COUP_IN_GOGO);
        h->anyCoupling = 0;
        for (uint32_t i = 0; i < s.numGO; i++) {
            if (nc[i] < 0 || (uint32_t)nc[i] > s.maxnumpGOCoupInGOGO) return fail(CBM_E_ARG, "GO coupling count above maximum");
            for (int j = 0; j < nc[i]; j++)
                if ((rc = checkIndex("P_GO_COUP_IN_GOGO", pc[(size_t)i * s.maxnumpGOCoupInGOGO + j], s.numGO))) return rc;
            if (nc[i] > 0 && h->p.coupleRiRjRatioGO != 0.0f) h->anyCoupling = 1;
        }
    }
    for (uint32_t z = 0; z < s.numZones; z++) {
        std::vector<std::vector<uint32_t>> bp(s.numPC), pb(s.numBC);
        const uint32_t *b = h->H<uint32_t>(CBM_F_P_BC_FROM_BC_TO_PC, z);
        for (uint32_t i = 0; i < s.numBC; i++)
            for (uint32_t j = 0; j < s.numpBCfromBCtoPC; j++) {
                uint32_t pc = b[(size_t)i * s.numpBCfromBCtoPC + j];
                if ((rc = checkIndex("P_BC_FROM_BC_TO_PC", pc, s.numPC))) return rc;
                bp[pc].push_back(i);
            }
        const uint32_t *q = h->H<uint32_t>(CBM_F_P_PC_FROM_PC_TO_BC, z);
        for (uint32_t i = 0; i < s.numPC; i++)
            for (uint32_t j = 0; j < s.numpPCfromPCtoBC; j++) {
                uint32_t bc = q[(size_t)i * s.numpPCfromPCtoBC + j];
                if ((rc = checkIndex("P_PC_FROM_PC_TO_BC", bc, s.numBC))) return rc;
                pb[bc].push_back(i);
            }
        buildCsr(h, s.numPC, bp, nullptr, &bcpPtr, &bcpIdx);
        buildCsr(h, s.numBC, pb, nullptr, &pcbPtr, &pcbIdx);
        struct { int f; uint32_t n; } chk[] = {
            {CBM_F_P_PC_FROM_SC_TO_PC, s.numSC}, {CBM_F_P_NC_FROM_PC_TO_NC, s.numPC},
            {CBM_F_P_NC_FROM_MF_TO_NC, s.numMF}, {CBM_F_P_IO_FROM_NC_TO_IO, s.numNC},
            {CBM_F_P_IO_IN_IOIO, s.numIO}};
        for (auto &c : chk) {
            const uint32_t *v = h->H<uint32_t>(c.f, z);
            size_t cnt = h->zoneBytes(c.f) / 4;
            for (size_t k = 0; k < cnt; k++)
                if ((rc = checkIndex("microzone connectivity", v[k], c.n))) return rc;
        }
    }

    // ---- device storage
    if ((rc = h->upload(&h->mfIn, mfIn)) || (rc = h->upload(&h->goIn, goIn)) || (rc = h->upload(&h->goOut, goOut)) ||
        (rc = h->upload(&h->goOutDly, goOutDly)) || (rc = h->upload(&h->bspDly, bsp)))
        return rc;
    uint32_t *tmp;
    if ((rc = h->upload(&tmp, mfgPtr))) return rc;
    h->mfToGoIn.ptr = tmp;
    if ((rc = h->upload(&tmp, mfgIdx))) return rc;
    h->mfToGoIn.idx = tmp;
    if ((rc = h->upload(&tmp, gabPtr))) return rc;
    h->gabaIn.ptr = tmp;
    if ((rc = h->upload(&tmp, gabIdx))) return rc;
    h->gabaIn.idx = tmp;
    // per-zone CSR pointers are stored with a per-zone index base
    if ((rc = h->upload(&tmp, bcpPtr))) return rc;
    h->bcToPcIn.ptr = tmp;
    if ((rc = h->upload(&tmp, bcpIdx))) return rc;
    h->bcToPcIn.idx = tmp;
    if ((rc = h->upload(&tmp, pcbPtr))) return rc;
    h->pcToBcIn.ptr = tmp;
    if ((rc = h->upload(&tmp, pcbIdx))) return rc;
    h->pcToBcIn.idx = tmp;

    for (int f = 0; f < CBM_F_COUNT; f++) {
        if (!deviceField(f)) continue;
        const size_t b = h->host[f].size();
        if ((rc = h->dalloc(&h->dfield[f], b))) return rc;
        if (b) HIP_TRY(hipMemcpy(h->dfield[f], h->host[f].data(), b, hipMemcpyHostToDevice));
    }
    const uint32_t nZ = s.numZones;
    if ((rc = h->dalloc((void **)&h->chunk, (size_t)nZ * (nGR / 64) * 4))) return rc;
    if ((rc = h->dalloc((void **)&h->pfpcRing, (size_t)kRing * nZ * s.numPC * 4))) return rc;
    if ((rc = h->dalloc((void **)&h->bcRing, (size_t)kRing * s.numBC * 4))) return rc;
    if ((rc = h->dalloc((void **)&h->scRing, (size_t)kRing * s.numSC * 4))) return rc;
    if ((rc = h->dalloc((void **)&h->goRing, (size_t)kRing * s.numGO * 4))) return rc;
    if ((rc = h->dalloc((void **)&h->dmf, s.numMF))) return rc;
    HIP_TRY(hipMemcpy(h->dmf, h->host[CBM_F_MF_INPUT].data(), s.numMF, hipMemcpyHostToDevice));
    {
        std::vector<uint32_t> rs((size_t)nZ * 353);
        for (uint32_t z = 0; z < nZ; z++) {
            cbm::Sfmt g(h->H<int32_t>(CBM_F_ZONE_RNG_SEED, z)[0]);
            const uint64_t skip = h->H<uint64_t>(CBM_F_ZONE_RNG_SKIP, z)[0];
            for (uint64_t k = 0; k < skip; k++) (void)g.Random();
            g.save(&rs[(size_t)z * 353]);
        }
        if ((rc = h->upload(&h->rngState, rs))) return rc;
    }
    {
        SmallArgs a;
        for (int f = 0; f < CBM_F_COUNT; f++) a.f[f] = h->dfield[f];
        a.chunk = h->chunk;
        a.pfpcRing = h->pfpcRing;
        a.bcRing = h->bcRing;
        a.scRing = h->scRing;
        a.goRing = h->goRing;
        a.rngState = h->rngState;
        a.mfToGoIn = h->mfToGoIn;
        a.gabaIn = h->gabaIn;
        a.bcToPcIn = h->bcToPcIn;
        a.pcToBcIn = h->pcToBcIn;
        a.coupN = (const int32_t *)h->dfield[CBM_F_NUMP_GO_COUP_IN_GOGO];
        a.coupP = (const uint32_t *)h->dfield[CBM_F_P_GO_COUP_IN_GOGO];
        a.s = s;
        a.p = h->p;
        a.grgoLag = h->o.grgoLag;
        a.anyCoupling = h->anyCoupling;
        // synapse partials: SC->PC, MF->NC (NMDA, AMPA), PC->NC, NC->IO
        h->smallLds = 4 * (size_t)nZ *
                      (s.numSC + 2 * (size_t)s.numNC * s.numpNCfromMFtoNC + (size_t)s.numNC * s.numpNCfromPCtoNC +
                       (size_t)s.numIO * s.numpIOfromNCtoIO);
        if (h->smallLds > 150 * 1024)
            return fail(CBM_E_UNSUPPORTED, "small-population synapse partials exceed the LDS of one workgroup");
        if (h->smallLds > 64 * 1024)
            HIP_TRY(hipFuncSetAttribute((const void *)small_step_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                        (int)h->smallLds));
        if ((rc = h->dalloc((void **)&h->smallArgs, sizeof(SmallArgs)))) return rc;
        HIP_TRY(hipMemcpy(h->smallArgs, &a, sizeof(SmallArgs), hipMemcpyHostToDevice));
    }
    h->committed = true;
    // pipeline rows given before commit
    if (h->written[CBM_F_PIPE_PFPC_SUM]) {
        const size_t zb = h->zoneBytes(CBM_F_PIPE_PFPC_SUM);
        for (uint32_t z = 0; z < nZ; z++)
            if ((rc = cbm_write(h, CBM_F_PIPE_PFPC_SUM, (int)z, h->host[CBM_F_PIPE_PFPC_SUM].data() + z * zb, zb))) return rc;
    }
    for (int f : {CBM_F_PIPE_PFBC_SUM, CBM_F_PIPE_PFSC_SUM, CBM_F_PIPE_GRGO_SUM})
        if (h->written[f] && (rc = cbm_write(h, f, 0, h->host[f].data(), h->host[f].size()))) return rc;
    if (h->written[CBM_F_ZONE_RNG_STATE])
        for (uint32_t z = 0; z < nZ; z++)
            if ((rc = cbm_write(h, CBM_F_ZONE_RNG_STATE, (int)z, h->H<uint8_t>(CBM_F_ZONE_RNG_STATE, z), 353 * 4))) return rc;
    HIP_TRY(hipDeviceSynchronize());
    return CBM_OK;
}

int cbm_read(cbm_handle *h, int field, int zone, void *dst, size_t bytes)
{
    if (!h || (!dst && bytes)) return fail(CBM_E_ARG, "null argument");
    int zoned, kind;
    int rc = fieldCheck(h, field, zone, bytes, &zoned, &kind);
    if (rc) return rc;
    if (!zoned) zone = 0;
    const size_t off = (size_t)zone * bytes;
    if (!h->committed || kind == CBM_KIND_CON) {
        if (bytes) std::memcpy(dst, h->host[field].data() + off, bytes);
        return CBM_OK;
    }
    HIP_TRY(hipSetDevice(h->o.device));
    HIP_TRY(hipStreamSynchronize(h->stream));
    switch (field) {
    case CBM_F_CUR_TIME: std::memcpy(dst, &h->curTime, 8); return CBM_OK;
    case CBM_F_GINC_GR_TO_GO: std::memcpy(dst, &h->gIncGRtoGO, 4); return CBM_OK;
    case CBM_F_ERR_DRIVE: std::memcpy(dst, &h->errDrive[zone], 4); return CBM_OK;
    case CBM_F_GRPC_PLAST_STEPS:
        std::memcpy(dst, &h->ltd[zone], 4);
        std::memcpy((uint8_t *)dst + 4, &h->ltp[zone], 4);
        return CBM_OK;
    case CBM_F_MF_INPUT: HIP_TRY(hipMemcpy(dst, h->dmf, bytes, hipMemcpyDeviceToHost)); return CBM_OK;
    case CBM_F_ZONE_RNG_STATE:
        HIP_TRY(hipMemcpy(dst, h->rngState + (size_t)zone * 353, bytes, hipMemcpyDeviceToHost));
        return CBM_OK;
    case CBM_F_AP_GR: {
        std::vector<uint32_t> ab(h->s.numGR);
        HIP_TRY(hipMemcpy(ab.data(), h->dfield[CBM_F_APBUF_GR], ab.size() * 4, hipMemcpyDeviceToHost));
        uint8_t *d = (uint8_t *)dst;
        for (size_t i = 0; i < ab.size(); i++) d[i] = (uint8_t)(ab[i] & 1u);
        return CBM_OK;
    }
    case CBM_F_PIPE_PFPC_SUM: case CBM_F_INPUT_SUM_PF_PC: {
        const size_t row = (size_t)h->s.numPC * 4;
        const int r0 = field == CBM_F_PIPE_PFPC_SUM ? 0 : 1, nr = field == CBM_F_PIPE_PFPC_SUM ? 2 : 1;
        for (int r = 0; r < nr; r++)
            HIP_TRY(hipMemcpy((uint8_t *)dst + r * row,
                              h->pfpcRing + ((size_t)ringSlot(h, r0 + r) * h->s.numZones + zone) * h->s.numPC, row,
                              hipMemcpyDeviceToHost));
        return CBM_OK;
    }
    case CBM_F_PIPE_PFBC_SUM: case CBM_F_PIPE_PFSC_SUM: case CBM_F_PIPE_GRGO_SUM: case CBM_F_INPUT_SUM_PF_SC: {
        const bool bc = field == CBM_F_PIPE_PFBC_SUM, go = field == CBM_F_PIPE_GRGO_SUM;
        uint32_t *ring = bc ? h->bcRing : go ? h->goRing : h->scRing;
        const uint32_t n = bc ? h->s.numBC : go ? h->s.numGO : h->s.numSC;
        const int r0 = field == CBM_F_INPUT_SUM_PF_SC ? 1 : 0, nr = field == CBM_F_INPUT_SUM_PF_SC ? 1 : 2;
        for (int r = 0; r < nr; r++)
            HIP_TRY(hipMemcpy((uint8_t *)dst + (size_t)r * n * 4, ring + (size_t)ringSlot(h, r0 + r) * n, (size_t)n * 4,
                              hipMemcpyDeviceToHost));
        return CBM_OK;
    }
    default:
        break;
    }
    if (!h->dfield[field]) return fail(CBM_E_ARG, "field is not readable");
    HIP_TRY(hipMemcpy(dst, (uint8_t *)h->dfield[field] + off, bytes, hipMemcpyDeviceToHost));
    return CBM_OK;
}

int cbm_set_mf(cbm_handle *h, const uint8_t *mf)
{
    if (!h || !mf) return fail(CBM_E_ARG, "null argument");
    if (!h->committed) return cbm_write(h, CBM_F_MF_INPUT, 0, mf, h->s.numMF);
    HIP_TRY(hipSetDevice(h->o.device));
    HIP_TRY(hipStreamSynchronize(h->stream));  // the pinned stage may still be in flight
    std::memcpy(h->pinnedMF, mf, h->s.numMF);
    HIP_TRY(hipMemcpyAsync(h->dmf, h->pinnedMF, h->s.numMF, hipMemcpyHostToDevice, h->stream));
    return CBM_OK;
}

int cbm_set_err_drive(cbm_handle *h, int zone, float rel)
{
    if (!h || zone < 0 || zone >= (int)h->s.numZones) return fail(CBM_E_ARG, "bad zone");
    h->errDrive[zone] = rel * h->p.maxExtIncVIO;
    return CBM_OK;
}

int cbm_set_grpc_steps(cbm_handle *h, int zone, float ltd, float ltp)
{
    if (!h || zone < 0 || zone >= (int)h->s.numZones) return fail(CBM_E_ARG, "bad zone");
    h->ltd[zone] = ltd;
    h->ltp[zone] = ltp;
    return CBM_OK;
}

int cbm_reset_grpc_steps(cbm_handle *h, int zone)
{
    if (!h) return fail(CBM_E_ARG, "null handle");
    return cbm_set_grpc_steps(h, zone, h->p.synLTDStepSizeGRtoPC, h->p.synLTPStepSizeGRtoPC);
}

int cbm_set_ginc_grgo(cbm_handle *h, float inc)
{
    if (!h) return fail(CBM_E_ARG, "null handle");
    h->gIncGRtoGO = inc;
    return CBM_OK;
}

int cbm_reset_ginc_grgo(cbm_handle *h)
{
    if (!h) return fail(CBM_E_ARG, "null handle");
    h->gIncGRtoGO = h->p.gIncGRtoGO;
    return CBM_OK;
}

static int launchStep(cbm_handle *h, const uint8_t *mf, bool timeGR)
{
    const cbm_sizes &s = h->s;
    const cbm_act_params &p = h->p;
    const uint64_t t = ++h->curTime;
    const uint32_t slot = (uint32_t)(t % kRing);

    GrArgs g;
    g.v = (float *)h->dfield[CBM_F_VM_GR];
    g.thr = (float *)h->dfield[CBM_F_THRESH_GR];
    g.gKCa = (float *)h->dfield[CBM_F_GKCA_GR];
    g.gE = (float *)h->dfield[CBM_F_G_MF_SUM_GR];
    g.gI = (float *)h->dfield[CBM_F_G_GO_SUM_GR];
    g.apBuf = (uint32_t *)h->dfield[CBM_F_APBUF_GR];
    g.hist = (uint64_t *)h->dfield[CBM_F_HIST_GR];
    g.w = (float *)h->dfield[CBM_F_PFPC_SYNW];
    g.mfIn = h->mfIn;
    g.goIn = h->goIn;
    g.goOut = h->goOut;
    g.goOutDly = h->goOutDly;
    g.bspDly = h->bspDly;
    g.mf = mf;
    g.apGOprev = (const uint8_t *)h->dfield[CBM_F_AP_GO];
    g.chunk = h->chunk;
    g.bcCnt = h->bcRing + (size_t)slot * s.numBC;
    g.scCnt = h->scRing + (size_t)slot * s.numSC;
    g.goCnt = h->goRing + (size_t)slot * s.numGO;
    g.plastTimer = (const int32_t *)h->dfield[CBM_F_PFPC_PLAST_TIMER_IO];
    g.numGR = s.numGR;
    g.numMF = s.numMF;
    g.numGO = s.numGO;
    g.numZones = s.numZones;
    g.numIO = s.numIO;
    g.log2Lbc = ilog2(s.numpBCfromGRtoBC);
    g.log2Lsc = ilog2(s.numpSCfromGRtoSC);
    g.log2GRperIO = ilog2(s.numGR / s.numIO);
    g.doPlast = (t % p.tsPerHistBinGR) == 0;
    g.gLeak = p.gLeakGR;
    g.eLeak = p.eLeakGR;
    g.eGO = p.eGOGR;
    g.threshRest = p.threshRestGR;
    g.threshMax = p.threshMaxGR;
    g.threshDec = p.threshDecGR;
    g.decMF = p.gDecMFtoGR;
    g.incMF = p.gIncMFtoGR;
    g.decGO = p.gDecGOtoGR;
    g.incGO = p.gIncGOtoGR;
    g.ltdEnd = p.tsLTDStartAPIO + (int)p.tsLTDDurationIO;
    g.ltpStart = p.tsLTPStartAPIO;
    for (int z = 0; z < kMaxZones; z++) {
        g.ltd[z] = h->ltd[z];
        g.ltp[z] = h->ltp[z];
    }
    const int tileGR = 256 * h->cpt;
    const uint32_t blocks = s.numGR / tileGR;
    const size_t lds = (size_t)s.numGO * 4 + 2 * 4 * h->cpt * 4 + (size_t)s.numZones * s.numIO * 4 + s.numMF + s.numGO;
    if (timeGR) HIP_TRY(hipEventRecord(h->ev[0], h->stream));
    if (h->cpt == 4)
        hipLaunchKernelGGL(gr_step_kernel<4>, dim3(blocks), dim3(256), lds, h->stream, g);
    else if (h->cpt == 2)
        hipLaunchKernelGGL(gr_step_kernel<2>, dim3(blocks), dim3(256), lds, h->stream, g);
    else
        hipLaunchKernelGGL(gr_step_kernel<1>, dim3(blocks), dim3(256), lds, h->stream, g);
    HIP_TRY(hipGetLastError());
    if (timeGR) HIP_TRY(hipEventRecord(h->ev[1], h->stream));

    SmallStep st;
    st.mf = mf;
    st.t = t;
    for (int z = 0; z < kMaxZones; z++) st.errDrive[z] = h->errDrive[z];
    st.gIncGRtoGO = h->gIncGRtoGO;
    hipLaunchKernelGGL(small_step_kernel, dim3(1), dim3(kSmallThreads), h->smallLds, h->stream, h->smallArgs, st);
    HIP_TRY(hipGetLastError());
    // MZone::calcIOActivities clears errDrive after use (mzone.cpp:334)
    for (int z = 0; z < kMaxZones; z++) h->errDrive[z] = 0.0f;
    return CBM_OK;
}

int cbm_step(cbm_handle *h, int n)
{
    if (!h || n < 0) return fail(CBM_E_ARG, "bad argument");
    if (!h->committed) return fail(CBM_E_STATE, "cbm_commit has not been called");
    HIP_TRY(hipSetDevice(h->o.device));
    for (int k = 0; k < n; k++) {
        int rc = launchStep(h, h->dmf, false);
        if (rc) return rc;
    }
    return CBM_OK;
}

static int checkTrain(cbm_handle *h, const uint8_t *mf_dev, size_t pitch, int n)
{
    if (!h || !mf_dev || n < 0) return fail(CBM_E_ARG, "bad argument");
    if (!h->committed) return fail(CBM_E_STATE, "cbm_commit has not been called");
    if (((uintptr_t)mf_dev & 3) || (pitch & 3) || pitch < h->s.numMF)
        return fail(CBM_E_ARG, "MF train must be 4-byte aligned with pitch >= numMF, pitch % 4 == 0");
    hipPointerAttribute_t attr;
    if (hipPointerGetAttributes(&attr, mf_dev) != hipSuccess || attr.type != hipMemoryTypeDevice) {
        (void)hipGetLastError();
        return fail(CBM_E_ARG, "MF train must be device memory");
    }
    return CBM_OK;
}

int cbm_run_mf_train(cbm_handle *h, const uint8_t *mf_dev, size_t pitch, int n, const float *err)
{
    int rc = checkTrain(h, mf_dev, pitch, n);
    if (rc) return rc;
    HIP_TRY(hipSetDevice(h->o.device));
    for (int k = 0; k < n; k++) {
        if (err)
            for (uint32_t z = 0; z < h->s.numZones; z++) {
                const float e = err[(size_t)k * h->s.numZones + z];
                if (e != 0.0f) h->errDrive[z] = e * h->p.maxExtIncVIO;
            }
        if ((rc = launchStep(h, mf_dev + (size_t)k * pitch, false))) return rc;
    }
    if (n > 0)
        HIP_TRY(hipMemcpyAsync(h->dmf, mf_dev + (size_t)(n - 1) * pitch, h->s.numMF, hipMemcpyDeviceToDevice, h->stream));
    return CBM_OK;
}

int cbm_sync(cbm_handle *h)
{
    if (!h) return fail(CBM_E_ARG, "null handle");
    HIP_TRY(hipSetDevice(h->o.device));
    HIP_TRY(hipStreamSynchronize(h->stream));
    return CBM_OK;
}

void *cbm_stream(cbm_handle *h) { return h ? (void *)h->stream : nullptr; }

int cbm_profile_steps(cbm_handle *h, const uint8_t *mf_dev, size_t pitch, int n, double *grUs, double *stepUs)
{
    int rc = checkTrain(h, mf_dev, pitch, n);
    if (rc) return rc;
    if (n == 0) return fail(CBM_E_ARG, "n must be positive");
    HIP_TRY(hipSetDevice(h->o.device));
    HIP_TRY(hipStreamSynchronize(h->stream));
    // one event pair around every GR launch, read back after the whole run so
    // that timing does not insert host waits between steps
    std::vector<hipEvent_t> ev(2 * (size_t)n + 2);
    for (auto &e : ev) HIP_TRY(hipEventCreate(&e));
    HIP_TRY(hipEventRecord(ev[2 * n], h->stream));
    for (int k = 0; k < n; k++) {
        h->ev[0] = ev[2 * k];
        h->ev[1] = ev[2 * k + 1];
        if ((rc = launchStep(h, mf_dev + (size_t)k * pitch, true))) break;
    }
    h->ev[0] = h->ev[1] = nullptr;
    if (!rc) {
        HIP_TRY(hipEventRecord(ev[2 * n + 1], h->stream));
        HIP_TRY(hipEventSynchronize(ev[2 * n + 1]));
        double grTotal = 0;
        for (int k = 0; k < n; k++) {
            float ms = 0;
            HIP_TRY(hipEventElapsedTime(&ms, ev[2 * k], ev[2 * k + 1]));
            grTotal += ms;
        }
        float total = 0;
        HIP_TRY(hipEventElapsedTime(&total, ev[2 * n], ev[2 * n + 1]));
        if (grUs) *grUs = grTotal * 1000.0 / n;
        if (stepUs) *stepUs = (double)total * 1000.0 / n;
        HIP_TRY(hipMemcpyAsync(h->dmf, mf_dev + (size_t)(n - 1) * pitch, h->s.numMF, hipMemcpyDeviceToDevice,
                               h->stream));
    }
    HIP_TRY(hipStreamSynchronize(h->stream));
    for (auto &e : ev) (void)hipEventDestroy(e);
    return rc;
}

int cbm_field_info(int field, const cbm_sizes *s, size_t *elem, size_t *rows, size_t *cols, int *zoned,
                   int *kind, int *is_float)
{
    if (!s) return fail(CBM_E_ARG, "null sizes");
    int rc = cbm_field_shape(field, s, elem, rows, cols, zoned, kind);
    if (rc) return fail(rc, "unknown field");
    if (is_float) *is_float = cbm_field_is_float(field);
    return CBM_OK;
}

int cbm_device_bytes(cbm_handle *h, size_t *bytes)
{
    if (!h || !bytes) return fail(CBM_E_ARG, "null argument");
    *bytes = h->devBytes;
    return CBM_OK;
}

}  // extern "C"

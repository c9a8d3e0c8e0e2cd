// cbm_runtime.hip -- host side of the C-ABI in include/cbm_hip.h.
//
// Owns one network on one GPU: stages the CBMState-layout fields written by
// the caller, validates them, packs granule connectivity into the device
// layout of the GR kernel (four 16-bit indices per GR, 8-bit delay bits),
// builds the gather lists of the small-population kernel, and launches each
// step (one fused launch, or two kernels on two streams).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../state/sfmt.hpp"
#include "cbm_handle.hpp"

using namespace cbmdev;

namespace {
thread_local std::string g_err;
}  // namespace

int cbmdev::fail(int code, const std::string &m)
{
    g_err = m;
    return code;
}

namespace {

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
    case CBM_F_AP_GO: case CBM_F_PFPC_PLAST_TIMER_IO:  // double-buffered by step parity
        return false;
    default:
        return kind == CBM_KIND_STATE;
    }
}

}  // namespace

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

uint32_t roundUp(uint32_t x, uint32_t m) { return (x + m - 1) / m * m; }

uint32_t maxLen(const std::vector<std::vector<uint32_t>> &lists)
{
    size_t m = 0;
    for (const auto &l : lists) m = std::max(m, l.size());
    return (uint32_t)m;
}

// Lists of `rows` rows padded to `width` slots, slot-major (out[k*rows + r]),
// kNoCell in unused slots.
void appendEll(std::vector<uint16_t> *out, const std::vector<std::vector<uint32_t>> &lists, uint32_t width)
{
    const size_t rows = lists.size(), base = out->size();
    out->resize(base + (size_t)width * rows, kNoCell);
    for (size_t r = 0; r < rows; r++)
        for (size_t k = 0; k < lists[r].size(); k++) (*out)[base + k * rows + r] = (uint16_t)lists[r][k];
}

// Dynamic LDS layout of the small kernel.  Every region starts on 16 bytes;
// the u16 index blob holds, in order, P_NC_FROM_MF_TO_NC, P_PC_FROM_SC_TO_PC,
// P_NC_FROM_PC_TO_NC, P_IO_FROM_NC_TO_IO, P_IO_IN_IOIO and the BC->PC and
// PC->BC in-lists, each padded to 8 elements.
SmallLds smallLayout(const cbm_sizes &s, uint32_t slotsBCPC, uint32_t slotsPCBC, bool coupling)
{
    SmallLds L{};
    const uint32_t nZ = s.numZones;
    uint32_t o = 0;
    auto take = [&](uint32_t bytes) {
        const uint32_t at = o;
        o = roundUp(o + bytes, 16);
        return at;
    };
    L.mf = take(s.numMF);
    L.hmf = take(s.numMF);
    L.goMF = take(4 * s.numGO);
    L.apGO = take(s.numGO);
    L.scaleGO = take(4 * s.numGO);
    L.vGO = take(coupling ? 4 * s.numGO : 0);
    uint32_t f = 0;  // floats
    L.synPC = f; f = roundUp(f + nZ * s.numSC, 4);
    L.synN = f;  f = roundUp(f + nZ * s.numNC * s.numpNCfromMFtoNC, 4);
    L.synA = f;  f = roundUp(f + nZ * s.numNC * s.numpNCfromMFtoNC, 4);
    L.synP = f;  f = roundUp(f + nZ * s.numNC * s.numpNCfromPCtoNC, 4);
    L.synIO = f; f = roundUp(f + nZ * s.numIO * s.numpIOfromNCtoIO, 4);
    L.syn = take(4 * f);
    L.apPC = take(nZ * s.numPC);
    L.apBC = take(nZ * s.numBC);
    L.apSC = take(s.numSC);
    L.apIO = take(nZ * s.numIO);
    L.apNC = take(nZ * s.numNC);
    L.vIO = take(4 * nZ * s.numIO);
    L.rel = take(4 * nZ * s.numNC);
    L.ncSum = take(4 * 3 * nZ * s.numNC);
    L.inPF = take(4 * nZ * s.numPC);
    L.inBC = take(4 * s.numBC);
    L.inSC = take(4 * s.numSC);
    L.inGO = take(4 * s.numGO);
    L.inTm = take(4 * nZ * s.numIO);
    L.rng = take(4 * nZ * kRngZone);
    L.rngNext = take(4 * nZ * kRngNext);
    L.popStride = 3 + s.numPopHistBinsPC + 2;
    L.pop = take(4 * nZ * L.popStride);
    uint32_t e = 0;  // u16 elements
    L.iMFNC = e; e = roundUp(e + nZ * s.numNC * s.numpNCfromMFtoNC, 8);
    L.iSCPC = e; e = roundUp(e + nZ * s.numPC * s.numpPCfromSCtoPC, 8);
    L.iPCNC = e; e = roundUp(e + nZ * s.numNC * s.numpNCfromPCtoNC, 8);
    L.iNCIO = e; e = roundUp(e + nZ * s.numIO * s.numpIOfromNCtoIO, 8);
    L.iIOIO = e; e = roundUp(e + nZ * s.numIO * s.numpIOInIOIO, 8);
    L.iBCPC = e; e = roundUp(e + nZ * slotsBCPC * s.numPC, 8);
    L.iPCBC = e; e = roundUp(e + nZ * slotsPCBC * s.numBC, 8);
    L.slotsBCPC = slotsBCPC;
    L.slotsPCBC = slotsPCBC;
    L.idx16 = e / 8;
    L.idx = take(2 * e);
    L.total = o;
    // thread offsets: consecutive wave-aligned ranges, smallest groups first,
    // wrapping around the block (a group larger than the block covers it all)
    auto place = [](std::initializer_list<std::pair<uint32_t *, uint32_t>> groups) {
        uint32_t cur = 0;
        for (auto &g : groups) {
            *g.first = cur % kSmallThreads;
            cur += roundUp(std::max<uint32_t>(g.second, 1), 64);
        }
    };
    place({{&L.oSynIO, nZ * s.numIO * s.numpIOfromNCtoIO}, {&L.oSynPC, nZ * s.numNC * s.numpNCfromPCtoNC},
           {&L.oSynSC, nZ * s.numSC}, {&L.oSynMF, nZ * s.numNC * s.numpNCfromMFtoNC}});
    place({{&L.oIO, nZ * s.numIO}, {&L.oNC, nZ * s.numNC}, {&L.oPC, nZ * s.numPC}, {&L.oBC, nZ * s.numBC},
           {&L.oSC, s.numSC}, {&L.oNCS, 3 * 64}});
    place({{&L.oRIO, nZ * s.numIO}, {&L.oRBCPC, nZ * s.numPC}, {&L.oRPCBC, nZ * s.numBC},
           {&L.oRPCNC, nZ * s.numNC * s.numpNCfromPCtoNC}, {&L.oRSCPC, nZ * s.numPC * s.numpPCfromSCtoPC}});
    L.l2PC = ilog2(s.numPC);
    L.l2BC = ilog2(s.numBC);
    L.l2IO = ilog2(s.numIO);
    L.l2NC = ilog2(s.numNC);
    return L;
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
    h->cpt = h->o.grCellsPerThread ? (int)h->o.grCellsPerThread : 1;
    h->nt = h->o.grBlockThreads ? (int)h->o.grBlockThreads : 256;
    // fused GR tiles: 2 cells per thread unless 1 is asked for (or the net is one tile)
    h->fusedCpt = (h->o.grCellsPerThread == 1 || sizes->numGR < 2048) ? 1 : 2;
    const bool shapeOk = (h->nt == 256 && (h->cpt == 1 || h->cpt == 2 || h->cpt == 4)) ||
                         (h->nt == 1024 && (h->cpt == 1 || h->cpt == 2));
    if (!shapeOk || h->o.stepMode > 5) {
        delete h;
        return fail(CBM_E_ARG, "GR kernel shape must be 256 threads x 1, 2 or 4 cells or 1024 x 1 or 2; stepMode 0..5");
    }
    hipError_t e = hipSetDevice(h->o.device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&h->stream2, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&h->evG, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&h->evS, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventRecord(h->evG, h->stream);
    if (e == hipSuccess) e = hipEventRecord(h->evS, h->stream2);
    if (e == hipSuccess) e = hipHostMalloc((void **)&h->pinnedMF, h->s.numMF, hipHostMallocDefault);
    if (e == hipSuccess) {
        const cbm_sizes &z = h->s;
        const size_t io = std::max({(size_t)z.numZones * z.numPC, (size_t)z.numZones * z.numBC, (size_t)z.numSC,
                                    (size_t)z.numZones * z.numIO, (size_t)z.numZones * z.numNC, (size_t)z.numGO});
        e = hipHostMalloc((void **)&h->pinnedIO, io, hipHostMallocCoherent);
    }
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
    if (h->batch) return fail(CBM_E_STATE, "the network belongs to a batch: destroy the batch first");
    (void)hipSetDevice(h->o.device);
    if (h->looping) {  // stop the closed loop's launch before its buffers go
        __atomic_store_n(reinterpret_cast<uint32_t *>(h->loopHost), kHostStop | (uint32_t)h->loopK, __ATOMIC_RELEASE);
        h->looping = false;
    }
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    if (h->stream2) (void)hipStreamSynchronize(h->stream2);
    for (void *p : h->allocs) (void)hipFree(p);
    if (h->rec) (void)hipFree(h->rec);
    if (h->pinnedMF) (void)hipHostFree(h->pinnedMF);
    if (h->pinnedIO) (void)hipHostFree(h->pinnedIO);
    if (h->loopHost) (void)hipHostFree(h->loopHost);
    if (h->loopMF) (void)hipFree(h->loopMF);
    if (h->loopErr) (void)hipFree(h->loopErr);
    if (h->loopFlag) (void)hipFree(h->loopFlag);
    if (h->resStatusHost) (void)hipHostFree(h->resStatusHost);
    if (h->resErr) (void)hipFree(h->resErr);
    if (h->evG) (void)hipEventDestroy(h->evG);
    if (h->evS) (void)hipEventDestroy(h->evS);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    if (h->stream2) (void)hipStreamDestroy(h->stream2);
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

}  // extern "C"

// A closed loop's pinned block: the step count the host posts (offset 0), the
// step count the microzone block stores after a step's spikes (256), the two
// input rows (MF bytes then the zones' error drive), then the two output rows.
constexpr size_t kLoopOutSeqOff = 256, kLoopInOff = 512;

int cbmdev::syncAll(cbm_handle *h)
{
    if (h->looping) return fail(CBM_E_STATE, "a closed loop is running on the network: cbm_loop_end first");
    HIP_TRY(hipStreamSynchronize(h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream2));
    if (h->resCheck) {  // the status of the last resident run, copied after it
        h->resCheck = false;
        if (*h->resStatusHost) {
            h->resident = false;
            return fail(CBM_E_HIP, "resident run: a wait between steps timed out (the run's workgroups were not all "
                                   "resident at once); the network state is undefined");
        }
    }
    return CBM_OK;
}

// GR-kernel instantiations by launch shape and connectivity format (defined below)
static const void *grSpanKernel(int cpt, bool compact);
static const void *fusedKernel(const cbm_handle *h);

extern "C" {

// device address of a parity-buffered field as of the last completed step
static uint8_t *parityField(cbm_handle *h, int field, int zone)
{
    const uint32_t par = (uint32_t)(h->curTime & 1u);
    if (field == CBM_F_AP_GO) return h->apGO2 + (size_t)par * h->s.numGO;
    if (field == CBM_F_PFPC_PLAST_TIMER_IO)
        return (uint8_t *)(h->timer2 + ((size_t)par * h->s.numZones + zone) * h->s.numIO);
    return nullptr;
}

// PF->PC sums of the last step still sit as chunk sums until the next small
// kernel folds them; fold them into their ring slot now (same order as that fold)
static int foldLastStep(cbm_handle *h)
{
    if (h->skipFold) return CBM_OK;  // the slot holds host-written sums
    const cbm_sizes &s = h->s;
    const uint32_t nChunks = s.numGR / 64, tasks = s.numZones * s.numPC;
    const float *c = h->chunk2 + (size_t)(h->curTime & 1u) * s.numZones * nChunks;
    float *out = h->pfpcRing + (size_t)ringSlot(h, 0) * s.numZones * s.numPC;
    hipLaunchKernelGGL(pfpc_fold_kernel, dim3((tasks + 3) / 4), dim3(256), 0, h->stream, c, out, nChunks,
                       s.numpPCfromGRtoPC / 64, s.numPC, tasks);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(h->stream));
    return CBM_OK;
}

// ---- GR history: the reference's words <-> the device's bit-planes
// a[i] bit j <-> a[j] bit i (the masked-swap transpose of a 64 x 64 bit block)
static void transpose64(uint64_t a[64])
{
    uint64_t m = 0x00000000FFFFFFFFull;
    for (int j = 32; j != 0; j >>= 1, m ^= (m << j)) {
        for (int k = 0; k < 64; k = ((k | j) + 1) & ~j) {
            const uint64_t t = ((a[k] >> j) ^ a[k | j]) & m;
            a[k] ^= t << j;
            a[k | j] ^= t;
        }
    }
}

// the history update index at time T (its words' bit k is in plane histPlane(c, k))
static uint64_t histUpdate(const cbm_handle *h) { return h->curTime / h->p.tsPerHistBinGR; }

// words[numGR] (bit k of GR g's word) -> planes[64][numGR / 64]
static void histToPlanes(const uint64_t *words, uint32_t numGR, uint64_t c, uint64_t *planes)
{
    const uint32_t nChunks = numGR / 64;
    uint64_t blk[64];
    for (uint32_t q = 0; q < nChunks; q++) {
        std::memcpy(blk, words + (size_t)q * 64, sizeof(blk));
        transpose64(blk);  // blk[k] bit l = bit k of GR 64q + l
        for (int k = 0; k < 64; k++) planes[(size_t)histPlane(c, k) * nChunks + q] = blk[k];
    }
}

static void histFromPlanes(const uint64_t *planes, uint32_t numGR, uint64_t c, uint64_t *words)
{
    const uint32_t nChunks = numGR / 64;
    uint64_t blk[64];
    for (uint32_t q = 0; q < nChunks; q++) {
        for (int k = 0; k < 64; k++) blk[k] = planes[(size_t)histPlane(c, k) * nChunks + q];
        transpose64(blk);
        std::memcpy(words + (size_t)q * 64, blk, sizeof(blk));
    }
}

static int writeHist(cbm_handle *h, const void *src)
{
    std::vector<uint64_t> planes(h->s.numGR);
    histToPlanes((const uint64_t *)src, h->s.numGR, histUpdate(h), planes.data());
    HIP_TRY(hipMemcpy(h->dfield[CBM_F_HIST_GR], planes.data(), planes.size() * 8, hipMemcpyHostToDevice));
    return CBM_OK;
}

static int readHist(cbm_handle *h, void *dst)
{
    std::vector<uint64_t> planes(h->s.numGR);
    HIP_TRY(hipMemcpy(planes.data(), h->dfield[CBM_F_HIST_GR], planes.size() * 8, hipMemcpyDeviceToHost));
    histFromPlanes(planes.data(), h->s.numGR, histUpdate(h), (uint64_t *)dst);
    return CBM_OK;
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
    if ((rc = syncAll(h))) return rc;
    if (uint8_t *d = parityField(h, field, zone)) {
        HIP_TRY(hipMemcpy(d, src, bytes, hipMemcpyHostToDevice));
        return CBM_OK;
    }
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
        std::vector<uint8_t> apgo(h->s.numGO), tim(h->host[CBM_F_PFPC_PLAST_TIMER_IO].size());
        std::vector<uint64_t> hist(h->s.numGR);  // its planes are placed by the time
        std::vector<uint32_t> rng((size_t)h->s.numZones * 353);  // in the parity slot of the time
        const size_t tz = h->zoneBytes(CBM_F_PFPC_PLAST_TIMER_IO);
        if ((rc = readHist(h, hist.data()))) return rc;
        for (uint32_t z = 0; z < h->s.numZones; z++)
            if ((rc = cbm_read(h, CBM_F_ZONE_RNG_STATE, (int)z, &rng[(size_t)z * 353], 353 * 4))) return rc;
        if ((rc = cbm_read(h, CBM_F_AP_GO, 0, apgo.data(), apgo.size()))) return rc;
        for (uint32_t z = 0; z < h->s.numZones; z++)
            if ((rc = cbm_read(h, CBM_F_PFPC_PLAST_TIMER_IO, (int)z, tim.data() + z * tz, tz))) return rc;
        std::memcpy(&h->curTime, src, 8);
        if ((rc = writeHist(h, hist.data()))) return rc;
        for (uint32_t z = 0; z < h->s.numZones; z++)
            if ((rc = cbm_write(h, CBM_F_ZONE_RNG_STATE, (int)z, &rng[(size_t)z * 353], 353 * 4))) return rc;
        if ((rc = cbm_write(h, CBM_F_AP_GO, 0, apgo.data(), apgo.size()))) return rc;
        for (uint32_t z = 0; z < h->s.numZones; z++)
            if ((rc = cbm_write(h, CBM_F_PFPC_PLAST_TIMER_IO, (int)z, tim.data() + z * tz, tz))) return rc;
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
    case CBM_F_HIST_GR: return writeHist(h, src);
    case CBM_F_PC_POPACT_CTRS:
        if (((const uint32_t *)src)[1] >= h->s.numPopHistBinsPC)
            return fail(CBM_E_ARG, "PC_POPACT_CTRS[1] must be below numPopHistBinsPC");
        break;
    case CBM_F_ZONE_RNG_STATE: {
        if (((const uint32_t *)src)[352] > 352) return fail(CBM_E_ARG, "generator index above 352");
        // into this step's parity slot with generation 0; both prebuilt blocks
        // of the zone lose their tag
        const size_t nZ = h->s.numZones, par = h->curTime & 1u;
        uint32_t st[kRngZone];
        std::memcpy(st, src, 353 * 4);
        st[353] = 0;
        HIP_TRY(hipMemcpy(h->rngState + (par * nZ + (size_t)zone) * kRngZone, st, sizeof(st), hipMemcpyHostToDevice));
        const uint32_t none = 0;
        for (size_t q = 0; q < 2; q++)
            HIP_TRY(hipMemcpy(h->rngNext + (q * nZ + (size_t)zone) * kRngNext + 352, &none, 4, hipMemcpyHostToDevice));
        return CBM_OK;
    }
    case CBM_F_PIPE_PFPC_SUM: {
        const size_t row = (size_t)h->s.numPC * 4;
        if ((rc = foldLastStep(h))) return rc;  // other zones' row 0 keeps its value
        for (int r = 0; r < 2; r++)
            HIP_TRY(hipMemcpy(h->pfpcRing + ((size_t)ringSlot(h, r) * h->s.numZones + zone) * h->s.numPC,
                              (const uint8_t *)src + r * row, row, hipMemcpyHostToDevice));
        h->skipFold = true;
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

// the resident-run kernel of a span shape (spill: the small populations'
// fields that do not fit in LDS stay in global memory; built for the 5-chunk
// span only)
static const void *residentKernel(int cpt, bool loop, bool spill)
{
    if (spill)
        return loop ? (const void *)step_resident_kernel<5, true, true> : (const void *)step_resident_kernel<5, false, true>;
#define CBM_RK(C) \
    (loop ? (const void *)step_resident_kernel<C, true, false> : (const void *)step_resident_kernel<C, false, false>)
    return cpt == 2 ? CBM_RK(2) : cpt == 3 ? CBM_RK(3) : CBM_RK(5);
#undef CBM_RK
}
static const void *residentKernel(const cbm_handle *h, bool loop)
{
    return residentKernel(h->resCpt, loop, h->resSpill);
}

// the six MF->NC synapse arrays, the first fields a resident run leaves in
// global memory when not all fit in LDS (large, and each read once a step in
// order)
static bool mfNcField(int f)
{
    return f == CBM_F_INPUT_MF_NC || f == CBM_F_MF_SYNW_NC || f == CBM_F_GMAX_NMDA_MF_NC ||
           f == CBM_F_GMAX_AMPA_MF_NC || f == CBM_F_G_MF_NMDA_NC || f == CBM_F_G_MF_AMPA_NC;
}

// Decide at commit whether cbm_run_mf_train runs resident (stepMode 0 or 5;
// the even split and compact connectivity records; the input layer's state
// fields fit in LDS beside the small step's scratch (the microzones' may stay
// partly in global memory with the 5-chunk span); one 1024-thread workgroup
// of the kernel fits on every CU), and set it up: the LDS of the GR workgroups
// (their span's working LDS and the records of their cells), the fields' place
// in the small block's LDS (h->smallHost), the step counters and the status
// word.  The launch asks for more than half of a CU's LDS, so that no CU holds
// two of its workgroups and the grid, one workgroup per CU, is on the device
// at once when nothing else is.
static int setupResident(cbm_handle *h, int cus)
{
    const cbm_sizes &s = h->s;
    const int sm = (int)h->o.stepMode;
    h->resident = false;
    h->resSpill = false;
    // the GR cells split evenly over one block per CU but three (the two
    // parts of the small populations and the generator block)
    const uint32_t spans = cus > 4 ? (uint32_t)cus - 3 : 0u;
    const uint32_t per = spans ? (s.numGR / 64 + spans - 1) / spans * 64 : 0u;
    const int cpt = per == 0 ? 0 : per <= 2048 ? 2 : per <= 3072 ? 3 : per <= 5120 ? 5 : 0;
    if ((sm == 0 || sm == 5) && cpt && h->conn) {
        int ldsPerCU = 0;
        HIP_TRY(hipDeviceGetAttribute(&ldsPerCU, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, h->o.device));
        const uint32_t recOff = roundUp(grTileLds(1024, cpt, s.numGO, s.numMF, s.numZones, s.numIO), 16);
        // each part's fields after the step's scratch LDS in its own block.
        // When they do not all fit, the 5-chunk span's spilling kernel places
        // the input layer's first in the microzones' synapse scratch (which
        // the input-layer part does not use), and leaves the microzones'
        // MF->NC synapse arrays in global memory, then each microzone field
        // in turn that does not fit in what is left
        SmallArgs &sa = h->smallHost;
        const uint32_t base = roundUp((uint32_t)h->smallLds, 16);
        size_t lds = 0;
        bool big = false;
        for (int pass = 0; pass < (cpt == 5 ? 2 : 1); pass++) {
            big = pass == 1;
            for (int f = 0; f < CBM_F_COUNT; f++) sa.fLds[f] = kNotInLds;
            uint32_t off = base, top = base, n = 0, nIn = 0;
            uint32_t hole = sa.lds.syn;
            const uint32_t holeEnd = sa.lds.apPC;
            for (int e = 0; e < kSmallFieldCount; e++) {
                if (e == kSmallInFields) {
                    off = base;
                    nIn = n;
                }
                const int f = kSmallFields[e];
                const bool in = e < kSmallInFields;
                const uint32_t bytes = (uint32_t)(h->zoneBytes(f) * (size_t)h->zonesOf(f));
                if (big && !in && (mfNcField(f) || off + bytes > kSmallLdsMax)) continue;
                if (big && in && hole + bytes <= holeEnd) {
                    sa.fLds[f] = hole;
                    hole = roundUp(hole + bytes, 16);
                } else {
                    sa.fLds[f] = off;
                    off = roundUp(off + bytes, 16);
                    top = std::max(top, off);
                }
                sa.ldsField[n] = (uint32_t)f;
                sa.ldsBytes[n++] = bytes;
            }
            sa.nLdsFields = n;
            sa.nInFields = nIn;
            const size_t need = std::max<size_t>(top, recOff + (size_t)cpt * 1024 * 16);
            lds = std::max<size_t>(need, roundUp((uint32_t)ldsPerCU / 2 + 16, 16));
            if (lds <= kSmallLdsMax) break;
        }
        if (ldsPerCU > 0 && lds <= kSmallLdsMax) {
            const void *k = residentKernel(cpt, false, big), *kl = residentKernel(cpt, true, big);
            HIP_TRY(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
            HIP_TRY(hipFuncSetAttribute(kl, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
            int perCU = 0, perCUl = 0;
            HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, k, kSmallThreads, lds));
            HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCUl, kl, kSmallThreads, lds));
            if (perCU == 1 && perCUl == 1 && cus >= (int)spans + 3) {
                int rc;
                if ((rc = h->dalloc((void **)&h->resBar, (size_t)(kBarGroups + 3 * kFlagCopies) * kBarStride * 4)))
                    return rc;
                if ((rc = h->dalloc((void **)&h->resApSC3, 3 * (size_t)s.numSC))) return rc;
                if ((rc = h->dalloc((void **)&h->resMfFlag, (size_t)kFlagCopies * kBarStride * 4))) return rc;
                if ((rc = h->dalloc((void **)&h->resStatus, 4))) return rc;
                if (!h->resStatusHost) {
                    HIP_TRY(hipHostMalloc((void **)&h->resStatusHost, 4, hipHostMallocDefault));
                    *h->resStatusHost = 0;
                }
                h->resBarBase = h->resFlagBase = h->resFlagInBase = h->resPbBase = 0;
                h->resSpans = spans;
                h->resCpt = cpt;
                h->residentLds = lds;
                h->residentRecLds = recOff;
                h->resSpill = big;
                h->resident = true;
            }
        }
    }
    if (!h->resident) {
        for (int f = 0; f < CBM_F_COUNT; f++) h->smallHost.fLds[f] = kNotInLds;
        h->smallHost.nLdsFields = h->smallHost.nInFields = 0;
    }
    if (sm == 5 && !h->resident)
        return fail(CBM_E_UNSUPPORTED, "stepMode 5: resident runs need the even GR split, compact connectivity "
                                       "records, the input layer's state in LDS and one 1024-thread "
                                       "workgroup per CU");
    return CBM_OK;
}

int cbm_commit(cbm_handle *h)
{
    if (!h) return fail(CBM_E_ARG, "null handle");
    if (h->committed) return fail(CBM_E_STATE, "already committed");
    const cbm_sizes &s = h->s;
    HIP_TRY(hipSetDevice(h->o.device));
    int rc;

    if (s.numGR % (uint32_t)(h->nt * h->cpt) || s.numGR % (1024u * (uint32_t)h->fusedCpt))
        return fail(CBM_E_UNSUPPORTED, "numGR must be a multiple of the GR tile");
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

    // ---- connectivity of the small kernel
    const uint32_t nZ = s.numZones;
    std::vector<uint16_t> mfgoOut, gabaEll, blob;
    uint32_t mfgoPad = 8, gabaPad = 4, slotsBCPC = 0, slotsPCBC = 0;
    {
        // MF -> GO targets per MF, row-major, padded to a multiple of 8 slots
        const int32_t *n = h->H<int32_t>(CBM_F_NUMP_MF_FROM_MF_TO_GO);
        const uint32_t *pp = h->H<uint32_t>(CBM_F_P_MF_FROM_MF_TO_GO);
        mfgoPad = roundUp(std::max<uint32_t>(s.maxnumpMFfromMFtoGO, 1), 8);
        mfgoOut.assign((size_t)s.numMF * mfgoPad, kNoCell);
        for (uint32_t m = 0; m < s.numMF; m++) {
            if (n[m] < 0 || (uint32_t)n[m] > s.maxnumpMFfromMFtoGO) return fail(CBM_E_ARG, "MF->GO count above maximum");
            for (int j = 0; j < n[m]; j++) {
                uint32_t g = pp[(size_t)m * s.maxnumpMFfromMFtoGO + j];
                if ((rc = checkIndex("P_MF_FROM_MF_TO_GO", g, s.numGO))) return rc;
                mfgoOut[(size_t)m * mfgoPad + j] = (uint16_t)g;
            }
        }
        // GABA sources per GO, ascending source then slot: the reference's order
        std::vector<std::vector<uint32_t>> gab(s.numGO);
        const int32_t *ng = h->H<int32_t>(CBM_F_NUMP_GO_GABA_OUT_GOGO);
        const uint32_t *pg = h->H<uint32_t>(CBM_F_P_GO_GABA_OUT_GOGO);
        for (uint32_t i = 0; i < s.numGO; i++) {
            if (ng[i] < 0 || (uint32_t)ng[i] > s.maxnumpGOGABAOutGOGO) return fail(CBM_E_ARG, "GO GABA count above maximum");
            for (int j = 0; j < ng[i]; j++) {
                uint32_t g = pg[(size_t)i * s.maxnumpGOGABAOutGOGO + j];
                if ((rc = checkIndex("P_GO_GABA_OUT_GOGO", g, s.numGO))) return rc;
                gab[g].push_back(i);
            }
        }
        gabaPad = roundUp(std::max<uint32_t>(maxLen(gab), 1), 4);
        gabaEll.assign((size_t)s.numGO * gabaPad, kNoCell);
        for (uint32_t g = 0; g < s.numGO; g++)
            for (size_t k = 0; k < gab[g].size(); k++) gabaEll[(size_t)g * gabaPad + k] = (uint16_t)gab[g][k];
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
    {
        // microzone lists: validate, then build the u16 blob in the LDS order
        std::vector<std::vector<std::vector<uint32_t>>> bp(nZ), pb(nZ);
        for (uint32_t z = 0; z < nZ; z++) {
            bp[z].resize(s.numPC);
            pb[z].resize(s.numBC);
            const uint32_t *b = h->H<uint32_t>(CBM_F_P_BC_FROM_BC_TO_PC, z);
            for (uint32_t i = 0; i < s.numBC; i++)
                for (uint32_t j = 0; j < s.numpBCfromBCtoPC; j++) {
                    uint32_t pc = b[(size_t)i * s.numpBCfromBCtoPC + j];
                    if ((rc = checkIndex("P_BC_FROM_BC_TO_PC", pc, s.numPC))) return rc;
                    bp[z][pc].push_back(i);
                }
            const uint32_t *q = h->H<uint32_t>(CBM_F_P_PC_FROM_PC_TO_BC, z);
            for (uint32_t i = 0; i < s.numPC; i++)
                for (uint32_t j = 0; j < s.numpPCfromPCtoBC; j++) {
                    uint32_t bc = q[(size_t)i * s.numpPCfromPCtoBC + j];
                    if ((rc = checkIndex("P_PC_FROM_PC_TO_BC", bc, s.numBC))) return rc;
                    pb[z][bc].push_back(i);
                }
            slotsBCPC = std::max(slotsBCPC, maxLen(bp[z]));
            slotsPCBC = std::max(slotsPCBC, maxLen(pb[z]));
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
        const uint32_t *ctr = h->H<uint32_t>(CBM_F_PC_POPACT_CTRS);
        for (uint32_t z = 0; z < nZ; z++)
            if (ctr[z * 3 + 1] >= s.numPopHistBinsPC) return fail(CBM_E_ARG, "PC_POPACT_CTRS[1] must be below numPopHistBinsPC");
        const SmallLds L = smallLayout(s, slotsBCPC, slotsPCBC, h->anyCoupling != 0);
        blob.assign((size_t)L.idx16 * 8, kNoCell);
        auto put = [&](uint32_t off, int f) {
            const size_t per = h->zoneBytes(f) / 4;
            for (uint32_t z = 0; z < nZ; z++) {
                const uint32_t *v = h->H<uint32_t>(f, z);
                for (size_t k = 0; k < per; k++) blob[off + z * per + k] = (uint16_t)v[k];
            }
        };
        put(L.iMFNC, CBM_F_P_NC_FROM_MF_TO_NC);
        put(L.iSCPC, CBM_F_P_PC_FROM_SC_TO_PC);
        put(L.iPCNC, CBM_F_P_NC_FROM_PC_TO_NC);
        put(L.iNCIO, CBM_F_P_IO_FROM_NC_TO_IO);
        put(L.iIOIO, CBM_F_P_IO_IN_IOIO);
        for (uint32_t z = 0; z < nZ; z++) {
            std::vector<uint16_t> e1, e2;
            appendEll(&e1, bp[z], slotsBCPC);
            appendEll(&e2, pb[z], slotsPCBC);
            std::copy(e1.begin(), e1.end(), blob.begin() + L.iBCPC + (size_t)z * slotsBCPC * s.numPC);
            std::copy(e2.begin(), e2.end(), blob.begin() + L.iPCBC + (size_t)z * slotsPCBC * s.numBC);
        }
        if (L.total > kSmallLdsMax)
            return fail(CBM_E_UNSUPPORTED, "small-population state exceeds the LDS of one workgroup");
        h->smallLds = L.total;
    }

    // ---- device storage
    const bool compact = h->o.connFormat != 1 && s.numMF <= 4095 && s.numGO <= 2047 && s.maxnumpGRfromMFtoGR <= 4 &&
                         s.maxnumpGRfromGOtoGR <= 3 && s.maxnumpGRfromGRtoGO <= 2;
    if (compact) {
        std::vector<uint4> conn(nGR);
        for (uint32_t i = 0; i < nGR; i++) {
            const uint16_t m[4] = {(uint16_t)(mfIn[i].x & 0xFFFF), (uint16_t)(mfIn[i].x >> 16),
                                   (uint16_t)(mfIn[i].y & 0xFFFF), (uint16_t)(mfIn[i].y >> 16)};
            const uint16_t g[3] = {(uint16_t)(goIn[i].x & 0xFFFF), (uint16_t)(goIn[i].x >> 16),
                                   (uint16_t)(goIn[i].y & 0xFFFF)};
            const uint16_t o[2] = {(uint16_t)(goOut[i].x & 0xFFFF), (uint16_t)(goOut[i].x >> 16)};
            conn[i] = packGrConn(m, g, o, bsp[i], (uint8_t)(goOutDly[i] & 0xFF), (uint8_t)((goOutDly[i] >> 8) & 0xFF),
                                 s.numMF);
        }
        if ((rc = h->upload(&h->conn, conn))) return rc;
    } else if ((rc = h->upload(&h->mfIn, mfIn)) || (rc = h->upload(&h->goIn, goIn)) ||
               (rc = h->upload(&h->goOut, goOut)) || (rc = h->upload(&h->goOutDly, goOutDly)) ||
               (rc = h->upload(&h->bspDly, bsp))) {
        return rc;
    }
    uint16_t *dMfgo = nullptr, *dGaba = nullptr, *dBlob = nullptr;
    if ((rc = h->upload(&dMfgo, mfgoOut)) || (rc = h->upload(&dGaba, gabaEll)) || (rc = h->upload(&dBlob, blob)))
        return rc;

    for (int f = 0; f < CBM_F_COUNT; f++) {
        if (!deviceField(f)) continue;
        const size_t b = h->host[f].size();
        if ((rc = h->dalloc(&h->dfield[f], b))) return rc;
        if (f == CBM_F_HIST_GR) {
            if ((rc = writeHist(h, h->host[f].data()))) return rc;
        } else if (b) {
            HIP_TRY(hipMemcpy(h->dfield[f], h->host[f].data(), b, hipMemcpyHostToDevice));
        }
    }
    const uint32_t par = (uint32_t)(h->curTime & 1u);
    if ((rc = h->dalloc((void **)&h->apGO2, 2 * (size_t)s.numGO))) return rc;
    HIP_TRY(hipMemcpy(h->apGO2 + (size_t)par * s.numGO, h->host[CBM_F_AP_GO].data(), s.numGO, hipMemcpyHostToDevice));
    const size_t timerBytes = h->host[CBM_F_PFPC_PLAST_TIMER_IO].size();
    if ((rc = h->dalloc((void **)&h->timer2, 2 * timerBytes))) return rc;
    HIP_TRY(hipMemcpy((uint8_t *)h->timer2 + par * timerBytes, h->host[CBM_F_PFPC_PLAST_TIMER_IO].data(), timerBytes,
                      hipMemcpyHostToDevice));
    if ((rc = h->dalloc((void **)&h->chunk2, 2 * (size_t)nZ * (nGR / 64) * 4))) return rc;
    if ((rc = h->dalloc((void **)&h->pfpcRing, (size_t)kRing * nZ * s.numPC * 4))) return rc;
    if ((rc = h->dalloc((void **)&h->bcRing, (size_t)kRing * s.numBC * 4))) return rc;
    if ((rc = h->dalloc((void **)&h->scRing, (size_t)kRing * s.numSC * 4))) return rc;
    if ((rc = h->dalloc((void **)&h->goRing, (size_t)kRing * s.numGO * 4))) return rc;
    if ((rc = h->dalloc((void **)&h->dmf, s.numMF))) return rc;
    HIP_TRY(hipMemcpy(h->dmf, h->host[CBM_F_MF_INPUT].data(), s.numMF, hipMemcpyHostToDevice));
    {
        // this step's parity slot holds the state; generation 0, no prebuilt block
        std::vector<uint32_t> rs(2 * (size_t)nZ * kRngZone, 0u), rn(2 * (size_t)nZ * kRngNext, 0u);
        for (uint32_t z = 0; z < nZ; z++) {
            cbm::Sfmt g(h->H<int32_t>(CBM_F_ZONE_RNG_SEED, z)[0]);
            const uint64_t skip = h->H<uint64_t>(CBM_F_ZONE_RNG_SKIP, z)[0];
            for (uint64_t k = 0; k < skip; k++) (void)g.Random();
            g.save(&rs[((size_t)par * nZ + z) * kRngZone]);
        }
        if ((rc = h->upload(&h->rngState, rs)) || (rc = h->upload(&h->rngNext, rn))) return rc;
    }
    {
        SmallArgs a;
        for (int f = 0; f < CBM_F_COUNT; f++) a.f[f] = h->dfield[f];
        a.chunk2 = h->chunk2;
        a.pfpcRing = h->pfpcRing;
        a.bcRing = h->bcRing;
        a.scRing = h->scRing;
        a.goRing = h->goRing;
        a.rngState = h->rngState;
        a.rngNext = h->rngNext;
        a.apGO2 = h->apGO2;
        a.timer2 = h->timer2;
        a.idxBlob = reinterpret_cast<const uint4 *>(dBlob);
        a.mfgoOut = dMfgo;
        a.gabaEll = dGaba;
        a.coupN = (const int32_t *)h->dfield[CBM_F_NUMP_GO_COUP_IN_GOGO];
        a.coupP = (const uint32_t *)h->dfield[CBM_F_P_GO_COUP_IN_GOGO];
        a.mfgoPad = mfgoPad;
        a.gabaPad = gabaPad;
        a.lds = smallLayout(s, slotsBCPC, slotsPCBC, h->anyCoupling != 0);
        a.s = s;
        a.p = h->p;
        a.grgoLag = h->o.grgoLag;
        a.anyCoupling = h->anyCoupling;
        a.rec = nullptr;
        a.recRow = a.recCap = 0;
        for (int k = 0; k < 6; k++) a.recOff[k] = 0xFFFFFFFFu;
        for (int f = 0; f < CBM_F_COUNT; f++) a.fLds[f] = kNotInLds;
        a.nLdsFields = 0;
        h->smallHost = a;
        if (h->smallLds > 64 * 1024)
            HIP_TRY(hipFuncSetAttribute((const void *)small_step_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                        (int)h->smallLds));
        // fused step: every block gets max(small, GR tile) LDS (one 1024-thread
        // block per CU either way, the small step's registers decide that)
        h->mode = h->o.stepMode ? (int)h->o.stepMode : 1;
        int cus = 0;
        HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, h->o.device));
        // GR work of the fused step: split evenly over one block per CU beside
        // the small step's when every block gets 1,024 to 5,120 cells (mode 1;
        // mode 4 whenever it fits), else tiles of 1024 x fusedCpt cells (mode 3
        // always).  The even split leaves no last round of leftover tiles: on
        // the 1M-GR network it takes 16.8 us per step against 18.7 for tiles.
        if (h->mode == 1 || h->mode == 4 || h->mode == 5) {
            // one CU each for the small step and the generator block
            const uint32_t spans = cus > 2 ? (uint32_t)cus - 2 : 1u;
            const uint32_t per = (s.numGR / 64 + spans - 1) / spans * 64;  // most cells of a span
            const int cpt = per <= 2048 ? 2 : per <= 3072 ? 3 : per <= 5120 ? 5 : 0;
            if (cpt && (h->mode >= 4 || per >= 1024)) {
                h->spans = spans;
                h->spanCpt = cpt;
            } else if (h->mode >= 4) {
                return fail(CBM_E_UNSUPPORTED, "stepMode 4/5: more than 5,120 GR per block");
            }
        }
        h->fusedLds = std::max<size_t>(
            h->smallLds, grTileLds(1024, h->spans ? h->spanCpt : h->fusedCpt, s.numGO, s.numMF, nZ, s.numIO));
        if (h->fused() && h->fusedLds > 64 * 1024)
            HIP_TRY(hipFuncSetAttribute(fusedKernel(h), hipFuncAttributeMaxDynamicSharedMemorySize, (int)h->fusedLds));
        const size_t grLds = grTileLds(1024, h->spans ? h->spanCpt : h->fusedCpt, s.numGO, s.numMF, nZ, s.numIO);
        if (h->spans && grLds > 64 * 1024)
            HIP_TRY(hipFuncSetAttribute(grSpanKernel(h->spanCpt, h->conn != nullptr),
                                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)grLds));
        if ((rc = setupResident(h, cus))) return rc;  // (may lay out the fields in LDS: h->smallHost)
        if ((rc = h->dalloc((void **)&h->smallArgs, sizeof(SmallArgs)))) return rc;
        HIP_TRY(hipMemcpy(h->smallArgs, &h->smallHost, sizeof(SmallArgs), hipMemcpyHostToDevice));
    }
    h->skipFold = true;
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
    if ((rc = syncAll(h))) return rc;
    if (const uint8_t *d = parityField(h, field, zone)) {
        HIP_TRY(hipMemcpy(dst, d, bytes, hipMemcpyDeviceToHost));
        return CBM_OK;
    }
    switch (field) {
    case CBM_F_CUR_TIME: std::memcpy(dst, &h->curTime, 8); return CBM_OK;
    case CBM_F_GINC_GR_TO_GO: std::memcpy(dst, &h->gIncGRtoGO, 4); return CBM_OK;
    case CBM_F_ERR_DRIVE: std::memcpy(dst, &h->errDrive[zone], 4); return CBM_OK;
    case CBM_F_GRPC_PLAST_STEPS:
        std::memcpy(dst, &h->ltd[zone], 4);
        std::memcpy((uint8_t *)dst + 4, &h->ltp[zone], 4);
        return CBM_OK;
    case CBM_F_MF_INPUT: HIP_TRY(hipMemcpy(dst, h->dmf, bytes, hipMemcpyDeviceToHost)); return CBM_OK;
    case CBM_F_HIST_GR: return readHist(h, dst);
    case CBM_F_ZONE_RNG_STATE:
        HIP_TRY(hipMemcpy(dst, h->rngState + ((size_t)(h->curTime & 1u) * h->s.numZones + (size_t)zone) * kRngZone,
                          bytes, hipMemcpyDeviceToHost));
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
        if (r0 == 0 && (rc = foldLastStep(h))) return rc;
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

// cells of recorder population k, zones stacked
static uint32_t recCells(const cbm_sizes &s, int k)
{
    const uint32_t nZ = s.numZones;
    switch (k) {
    case CBM_REC_PC: return nZ * s.numPC;
    case CBM_REC_BC: return nZ * s.numBC;
    case CBM_REC_SC: return s.numSC;
    case CBM_REC_IO: return nZ * s.numIO;
    case CBM_REC_NC: return nZ * s.numNC;
    case CBM_REC_GO: return s.numGO;
    case CBM_REC_PF: return s.numGR / 64;
    default: return 0;
    }
}

int cbm_set_recorder(cbm_handle *h, uint32_t mask, uint32_t capacity)
{
    if (!h) return fail(CBM_E_ARG, "null handle");
    if (!h->committed) return fail(CBM_E_STATE, "cbm_commit has not been called");
    if (mask >> 7) return fail(CBM_E_ARG, "unknown recorder population");
    HIP_TRY(hipSetDevice(h->o.device));
    int rc = syncAll(h);
    if (rc) return rc;
    SmallArgs &a = h->smallHost;
    if (h->rec) (void)hipFree(h->rec);
    h->rec = nullptr;
    a.rec = nullptr;
    a.recRow = a.recCap = 0;
    for (int k = 0; k < 6; k++) a.recOff[k] = 0xFFFFFFFFu;
    h->recOffPF = 0xFFFFFFFFu;
    if (mask && capacity) {
        uint32_t row = 0;
        for (int k = 0; k < 7; k++)
            if (mask & (1u << k)) {
                (k < 6 ? a.recOff[k] : h->recOffPF) = row;
                row += recCells(h->s, k);
            }
        HIP_TRY(hipMalloc((void **)&h->rec, (size_t)row * capacity));
        HIP_TRY(hipMemset(h->rec, 0, (size_t)row * capacity));
        a.rec = h->rec;
        a.recRow = row;
        a.recCap = capacity;
        h->recFrom = h->curTime + 1;
    }
    HIP_TRY(hipMemcpy(h->smallArgs, &a, sizeof(SmallArgs), hipMemcpyHostToDevice));
    return CBM_OK;
}

int cbm_read_recorded(cbm_handle *h, int population, uint64_t t0, int n, uint8_t *dst)
{
    if (!h || (n > 0 && !dst) || n < 0) return fail(CBM_E_ARG, "bad argument");
    const SmallArgs &a = h->smallHost;
    if (population < 0 || population > CBM_REC_PF || !a.rec)
        return fail(CBM_E_ARG, "population is not recorded");
    const uint32_t off = population == CBM_REC_PF ? h->recOffPF : a.recOff[population];
    if (off == 0xFFFFFFFFu) return fail(CBM_E_ARG, "population is not recorded");
    if (n == 0) return CBM_OK;
    if (t0 < h->recFrom || t0 + (uint64_t)n - 1 > h->curTime || h->curTime - t0 >= a.recCap)
        return fail(CBM_E_ARG, "steps outside the recorded window");
    HIP_TRY(hipSetDevice(h->o.device));
    int rc = syncAll(h);
    if (rc) return rc;
    const uint32_t cells = recCells(h->s, population);
    // a device destination takes one strided copy per ring segment; a host
    // one takes the segment's whole rows in one copy and the columns are
    // picked on the host (a strided copy into host memory moves row by row)
    hipPointerAttribute_t attr;
    const bool toDevice = hipPointerGetAttributes(&attr, dst) == hipSuccess && attr.type == hipMemoryTypeDevice;
    if (!toDevice) (void)hipGetLastError();
    int done = 0;
    while (done < n) {
        const uint64_t t = t0 + (uint64_t)done;
        const uint32_t r = (uint32_t)(t % a.recCap);
        const int run = (int)std::min<uint64_t>((uint64_t)(n - done), (uint64_t)(a.recCap - r));
        const uint8_t *src = h->rec + (size_t)r * a.recRow;
        if (toDevice) {
            HIP_TRY(hipMemcpy2D(dst + (size_t)done * cells, cells, src + off, a.recRow, cells, run, hipMemcpyDefault));
        } else {
            h->recStage.resize((size_t)run * a.recRow);
            HIP_TRY(hipMemcpy(h->recStage.data(), src, h->recStage.size(), hipMemcpyDeviceToHost));
            for (int i = 0; i < run; i++)
                std::memcpy(dst + (size_t)(done + i) * cells, h->recStage.data() + (size_t)i * a.recRow + off, cells);
        }
        done += run;
    }
    return CBM_OK;
}

int cbm_set_mf(cbm_handle *h, const uint8_t *mf)
{
    if (!h || !mf) return fail(CBM_E_ARG, "null argument");
    if (!h->committed) return cbm_write(h, CBM_F_MF_INPUT, 0, mf, h->s.numMF);
    HIP_TRY(hipSetDevice(h->o.device));
    int rc = syncAll(h);  // the pinned stage and the previous step may still be in flight
    if (rc) return rc;
    std::memcpy(h->pinnedMF, mf, h->s.numMF);
    HIP_TRY(hipMemcpyAsync(h->dmf, h->pinnedMF, h->s.numMF, hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipEventRecord(h->evG, h->stream));  // the next small kernel waits for the copy too
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
    if (h->looping) return fail(CBM_E_STATE, "a closed loop is running: the PF->PC steps change between loops");
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
    if (h->looping) return fail(CBM_E_STATE, "a closed loop is running: gIncGRtoGO changes between loops");
    h->gIncGRtoGO = inc;
    return CBM_OK;
}

int cbm_reset_ginc_grgo(cbm_handle *h)
{
    if (!h) return fail(CBM_E_ARG, "null handle");
    if (h->looping) return fail(CBM_E_STATE, "a closed loop is running: gIncGRtoGO changes between loops");
    h->gIncGRtoGO = h->p.gIncGRtoGO;
    return CBM_OK;
}

// The GR-kernel instantiations by launch shape and connectivity format
// (compact records or the five arrays).
static const void *grTileKernel(int nt, int cpt, bool compact)
{
#define CBM_K(C, N) (compact ? (const void *)gr_step_kernel<C, N, 1> : (const void *)gr_step_kernel<C, N, 2>)
    if (nt == 1024) return cpt == 2 ? CBM_K(2, 1024) : CBM_K(1, 1024);
    return cpt == 4 ? CBM_K(4, 256) : cpt == 2 ? CBM_K(2, 256) : CBM_K(1, 256);
#undef CBM_K
}

static const void *grSpanKernel(int cpt, bool compact)
{
#define CBM_K(C) (compact ? (const void *)gr_span_kernel<C, 1> : (const void *)gr_span_kernel<C, 2>)
    return cpt == 2 ? CBM_K(2) : cpt == 3 ? CBM_K(3) : CBM_K(5);
#undef CBM_K
}

// the fused step kernel of the handle's launch shape
static const void *fusedKernel(const cbm_handle *h)
{
    const bool compact = h->conn != nullptr;
    if (h->spans) {
#define CBM_K(C) (compact ? (const void *)step_span_kernel<C, 1> : (const void *)step_span_kernel<C, 2>)
        return h->spanCpt == 2 ? CBM_K(2) : h->spanCpt == 3 ? CBM_K(3) : CBM_K(5);
#undef CBM_K
    }
#define CBM_K(C) (compact ? (const void *)step_kernel<C, 1> : (const void *)step_kernel<C, 2>)
    return h->fusedCpt == 2 ? CBM_K(2) : CBM_K(1);
#undef CBM_K
}

// Launch the GR tiles alone (nt x cpt shape) on `st`.
static hipError_t launchGr(cbm_handle *h, const GrArgs &g, const GrStep &gs, int nt, int cpt, hipStream_t st)
{
    const cbm_sizes &s = h->s;
    const uint32_t blocks = s.numGR / (uint32_t)(nt * cpt);
    const size_t lds = grTileLds(nt, cpt, s.numGO, s.numMF, s.numZones, s.numIO);
    GrArgs ga = g;
    GrStep gsa = gs;
    void *args[] = {&ga, &gsa};
    return hipLaunchKernel(grTileKernel(nt, cpt, h->conn != nullptr), dim3(blocks), dim3(nt), args, lds, st);
}

}  // extern "C"

// the step-independent part of the GR kernel arguments
void cbmdev::grConstArgs(const cbm_handle *h, GrArgs *gp_)
{
    const cbm_sizes &s = h->s;
    const cbm_act_params &p = h->p;
    GrArgs &g = *gp_;
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
    g.conn = h->conn;
    g.numGR = s.numGR;
    g.numMF = s.numMF;
    g.numGO = s.numGO;
    g.numZones = s.numZones;
    g.numIO = s.numIO;
    g.numPC = s.numPC;
    g.perPC = s.numpPCfromGRtoPC / 64;
    g.log2Lbc = ilog2(s.numpBCfromGRtoBC);
    g.log2Lsc = ilog2(s.numpSCfromGRtoSC);
    g.log2GRperIO = ilog2(s.numGR / s.numIO);
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
}

extern "C" {

enum Timing { kNoTiming = 0, kTimeKernel = 1, kSplit = 2 };

// One step t.  The GR tiles of t need the GO spikes and plasticity timers of
// t-1 (small populations of t-1); the small populations of t need the chunk
// sums and ring slots of t-1 (GR tiles of t-1).  So the two halves of step t
// are independent and run concurrently:
//  fused (mode 1): one launch; block 0 is the small step, the rest GR tiles;
//    consecutive launches on one stream are ordered by the stream
//  two streams (mode 2):
//    stream : wait(evS = small t-1) -> GR t -> record evG
//    stream2: wait(evG = GR t-1)    -> small t -> record evS
//  split timing: GR t then small t, in order on `stream`
static int launchStep(cbm_handle *h, const uint8_t *mf, int timing)
{
    const cbm_sizes &s = h->s;
    const cbm_act_params &p = h->p;
    const uint64_t t = ++h->curTime;
    const uint32_t slot = (uint32_t)(t % kRing);
    const uint32_t par = (uint32_t)(t & 1u), parPrev = par ^ 1u;

    GrArgs g;
    grConstArgs(h, &g);
    GrStep gs;
    gs.mf = mf;
    gs.apGOprev = h->apGO2 + (size_t)parPrev * s.numGO;
    gs.chunk = h->chunk2 + (size_t)par * s.numZones * (s.numGR / 64);
    // the GR launch of step t folds the chunk sums of t-1 into their ring slot,
    // unless the host wrote that slot
    gs.foldSrc = h->skipFold ? nullptr : h->chunk2 + (size_t)parPrev * s.numZones * (s.numGR / 64);
    gs.foldDst = h->pfpcRing + (size_t)((t + kRing - 1) % kRing) * s.numZones * s.numPC;
    gs.bcCnt = h->bcRing + (size_t)slot * s.numBC;
    gs.scCnt = h->scRing + (size_t)slot * s.numSC;
    gs.goCnt = h->goRing + (size_t)slot * s.numGO;
    gs.plastTimer = h->timer2 + (size_t)parPrev * s.numZones * s.numIO;
    gs.doPlast = (t % p.tsPerHistBinGR) == 0;
    gs.histOld = g.hist + (size_t)histPlane(t / p.tsPerHistBinGR, 40) * (s.numGR / 64);
    gs.histNew = g.hist + (size_t)histPlane(t / p.tsPerHistBinGR, 0) * (s.numGR / 64);
    gs.pfCnt = (h->rec && h->recOffPF != 0xFFFFFFFFu)
                   ? h->rec + (size_t)(t % h->smallHost.recCap) * h->smallHost.recRow + h->recOffPF
                   : nullptr;
    SmallStep st;
    st.mf = mf;
    st.t = t;
    for (int z = 0; z < kMaxZones; z++) st.errDrive[z] = h->errDrive[z];
    st.gIncGRtoGO = h->gIncGRtoGO;
    st.par = par;
    st.slotPrev = (uint32_t)((t + kRing - 1) % kRing);
    st.slotT2 = (uint32_t)((t + kRing - 2) % kRing);
    st.slotG = (uint32_t)((t + kRing - h->o.grgoLag) % kRing);
    st.slotNext = (uint32_t)((t + 1) % kRing);
    st.popHist = (t % p.tsPerPopHistBinPC) == 0;
    st.resetMFHist = (t % p.numTSinMFHist) == 0;
    st.ioOut = h->ioOut;
    st.ioPop = h->ioPop;
    st.apSCout = nullptr;  // (the whole step on one block)
    st.apSCin = nullptr;
    st.scFromPrev = 0;
    h->skipFold = false;

    if (timing == kSplit) {
        HIP_TRY(hipEventRecord(h->ev[0], h->stream));
        if (h->spans) {
            const size_t lds = grTileLds(1024, h->spanCpt, s.numGO, s.numMF, s.numZones, s.numIO);
            uint32_t spans = h->spans;
            void *args[] = {&g, &gs, &spans};
            HIP_TRY(hipLaunchKernel(grSpanKernel(h->spanCpt, h->conn != nullptr), dim3(spans), dim3(1024), args, lds,
                                    h->stream));
        } else {
            HIP_TRY(launchGr(h, g, gs, h->fused() ? 1024 : h->nt, h->fused() ? h->fusedCpt : h->cpt, h->stream));
        }
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipEventRecord(h->ev[1], h->stream));
        HIP_TRY(hipEventRecord(h->ev[2], h->stream));
        hipLaunchKernelGGL(small_step_kernel, dim3(1), dim3(kSmallThreads), h->smallLds, h->stream, h->smallArgs, st);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipEventRecord(h->ev[3], h->stream));
    } else if (h->fused()) {
        const uint32_t tiles = s.numGR / (1024u * (uint32_t)h->fusedCpt);
        if (timing == kTimeKernel) HIP_TRY(hipEventRecord(h->ev[0], h->stream));
        const SmallArgs *sa = h->smallArgs;
        uint32_t spans = h->spans;
        void *args[] = {&sa, &st, &g, &gs, &spans};  // the tile kernel takes the first four
        HIP_TRY(hipLaunchKernel(fusedKernel(h), dim3(2 + (spans ? spans : tiles)), dim3(kSmallThreads), args,
                                h->fusedLds, h->stream));
        if (timing == kTimeKernel) HIP_TRY(hipEventRecord(h->ev[1], h->stream));
    } else {
        HIP_TRY(hipStreamWaitEvent(h->stream, h->evS, 0));
        if (timing == kTimeKernel) HIP_TRY(hipEventRecord(h->ev[0], h->stream));
        HIP_TRY(launchGr(h, g, gs, h->nt, h->cpt, h->stream));
        if (timing == kTimeKernel) HIP_TRY(hipEventRecord(h->ev[1], h->stream));
        HIP_TRY(hipStreamWaitEvent(h->stream2, h->evG, 0));
        hipLaunchKernelGGL(small_step_kernel, dim3(1), dim3(kSmallThreads), h->smallLds, h->stream2, h->smallArgs, st);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipEventRecord(h->evG, h->stream));
        HIP_TRY(hipEventRecord(h->evS, h->stream2));
    }
    // MZone::calcIOActivities clears errDrive after use (mzone.cpp:334)
    for (int z = 0; z < kMaxZones; z++) h->errDrive[z] = 0.0f;
    return CBM_OK;
}

// make `stream` wait for the last small kernel, so that work queued on it
// afterwards (and cbm_stream() users) sees the whole step
static int join(cbm_handle *h)
{
    HIP_TRY(hipStreamWaitEvent(h->stream, h->evS, 0));
    return CBM_OK;
}

int cbm_step_io_begin(cbm_handle *h, const uint8_t *mf, int population)
{
    if (!h || !mf) return fail(CBM_E_ARG, "null argument");
    if (!h->committed) return fail(CBM_E_STATE, "cbm_commit has not been called");
    if (h->ioPending) return fail(CBM_E_STATE, "cbm_step_io_end has not been called for the last step");
    if (population < CBM_REC_PC || population > CBM_REC_GO)
        return fail(CBM_E_ARG, "cbm_step_io_begin: population must be PC, BC, SC, IO, NC or GO");
    if (h->looping) {
        // a closed loop: the step's MF row and error drive into the input row
        // of its parity, then the step count the generator block polls
        if (population != h->loopPop) return fail(CBM_E_ARG, "cbm_step_io_begin: not the closed loop's population");
        if (h->loopK >= h->loopMax) return fail(CBM_E_STATE, "cbm_step_io_begin: the closed loop's steps are used up");
        uint8_t *in = h->loopHost + kLoopInOff + (size_t)(h->loopK & 1) * h->loopInPitch;
        std::memcpy(in, mf, h->s.numMF);
        std::memcpy(in + h->loopInErr, h->errDrive, 4 * (size_t)h->s.numZones);
        for (int z = 0; z < kMaxZones; z++) h->errDrive[z] = 0.0f;
        __atomic_store_n(reinterpret_cast<uint32_t *>(h->loopHost), (uint32_t)(h->loopK + 1), __ATOMIC_RELEASE);
        h->ioPending = true;
        return CBM_OK;
    }
    HIP_TRY(hipSetDevice(h->o.device));
    int rc = syncAll(h);  // the pinned stages of the last step are free again
    if (rc) return rc;
    std::memcpy(h->pinnedMF, mf, h->s.numMF);
    HIP_TRY(hipMemcpyAsync(h->dmf, h->pinnedMF, h->s.numMF, hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipEventRecord(h->evG, h->stream));
    // the step's small-population kernel stores the population's spikes (the
    // zones one after another) into pinnedIO: no copy after the kernel
    h->ioOut = h->pinnedIO;
    h->ioPop = (uint32_t)population;
    rc = launchStep(h, h->dmf, kNoTiming);
    h->ioOut = nullptr;
    if (rc) return rc;
    if ((rc = join(h))) return rc;
    h->ioBytes = recCells(h->s, population);
    h->ioPending = true;
    return CBM_OK;
}

int cbm_step_io_end(cbm_handle *h, uint8_t *out)
{
    if (!h || !out) return fail(CBM_E_ARG, "null argument");
    if (!h->ioPending) return fail(CBM_E_STATE, "cbm_step_io_begin has not been called");
    if (h->looping) {
        // a closed loop: wait for the step count the microzone block stores
        // after the step's spikes (bounded: the kernel bounds its waits too)
        const volatile uint32_t *seq = reinterpret_cast<const uint32_t *>(h->loopHost + kLoopOutSeqOff);
        const uint32_t want = (uint32_t)(h->loopK + 1);
        const auto t0 = std::chrono::steady_clock::now();
        uint32_t spin = 0;
        while ((int32_t)(__atomic_load_n(seq, __ATOMIC_ACQUIRE) - want) < 0) {
            if (++spin % 1024 == 0 && std::chrono::steady_clock::now() - t0 > std::chrono::seconds(5))
                return fail(CBM_E_HIP, "closed loop: the step's output did not arrive (the loop's launch ended; "
                                       "cbm_loop_end reports why)");
        }
        std::memcpy(out, h->loopHost + h->loopOutOff + (size_t)(h->loopK & 1) * 4 * h->loopOutWords, h->loopOutBytes);
        h->loopK++;
        h->ioPending = false;
        return CBM_OK;
    }
    HIP_TRY(hipSetDevice(h->o.device));
    h->ioPending = false;
    HIP_TRY(hipStreamSynchronize(h->stream));
    std::memcpy(out, h->pinnedIO, h->ioBytes);
    return CBM_OK;
}

int cbm_step(cbm_handle *h, int n)
{
    if (!h || n < 0) return fail(CBM_E_ARG, "bad argument");
    if (!h->committed) return fail(CBM_E_STATE, "cbm_commit has not been called");
    HIP_TRY(hipSetDevice(h->o.device));
    for (int k = 0; k < n; k++) {
        int rc = launchStep(h, h->dmf, kNoTiming);
        if (rc) return rc;
    }
    return join(h);
}

}  // extern "C"

int cbmdev::checkTrain(cbm_handle *h, const uint8_t *mf_dev, size_t pitch, int n)
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

// Resident launches of the process go to a device one after another: each
// waits for the previous one, whichever handle made it, so two runs never
// hold part of the CUs each while their workgroups wait for the rest.  Other
// kernels may run beside a resident launch: they only delay the workgroups
// they keep off the CUs, and a wait that lasts past its bound ends the run
// with an error.  (A cooperative launch would give the same guarantee, but on
// this runtime every kernel the process launched after one, on any stream,
// ran slower: the 256-network batch step by 8 %.)
static constexpr int kResDevices = 64;
static std::mutex g_resMu;
static hipEvent_t g_resLast[kResDevices];

static int launchResident(cbm_handle *h, const void *kern, uint32_t grid, void **args)
{
    const int dev = h->o.device;
    if (dev < 0 || dev >= kResDevices) return fail(CBM_E_UNSUPPORTED, "device ordinal above 63");
    std::lock_guard<std::mutex> lock(g_resMu);
    hipEvent_t &last = g_resLast[dev];
    if (!last) HIP_TRY(hipEventCreateWithFlags(&last, hipEventDisableTiming));
    else HIP_TRY(hipStreamWaitEvent(h->stream, last, 0));
    HIP_TRY(hipLaunchKernel(kern, dim3(grid), dim3(kSmallThreads), args, h->residentLds, h->stream));
    HIP_TRY(hipEventRecord(last, h->stream));
    return CBM_OK;
}

// Steps of one resident launch at most: the step counters grow by the grid
// size per step and are compared modulo 2^32.
constexpr int kResMaxSteps = 1 << 20;

// cbm_run_mf_train as resident launches (step_resident_kernel): the error
// drive goes to the device first, each launch derives its steps' values as
// launchStep does, and the launches' barrier status is copied to pinned host
// memory for the next synchronisation to read.
static void runFields(cbm_handle *h, ResRun *rp);
static void resAdvance(cbm_handle *h, uint32_t m);

static int runResident(cbm_handle *h, const uint8_t *mf_dev, size_t pitch, int n, const float *err)
{
    const uint32_t nZ = h->s.numZones;
    if (err) {
        const size_t need = (size_t)n * nZ;
        if (need > h->resErrCap) {
            if (h->resErr) {
                HIP_TRY(hipStreamSynchronize(h->stream));  // an earlier run may still read it
                (void)hipFree(h->resErr);
                h->resErr = nullptr;
                h->resErrCap = 0;
            }
            const hipError_t e = hipMalloc((void **)&h->resErr, need * sizeof(float));
            if (e != hipSuccess) return fail(CBM_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
            h->resErrCap = need;
        }
        HIP_TRY(hipMemcpyAsync(h->resErr, err, need * sizeof(float), hipMemcpyHostToDevice, h->stream));
    }
    ResRun r;
    runFields(h, &r);
    r.pitch = pitch;
    const uint32_t grid = h->resSpans + 3;
    const void *kern = residentKernel(h, false);
    const SmallArgs *sa = h->smallArgs;
    for (int k0 = 0; k0 < n;) {
        const int m = std::min(n - k0, kResMaxSteps);
        r.mf = mf_dev + (size_t)k0 * pitch;
        r.err = err ? h->resErr + (size_t)k0 * nZ : nullptr;
        r.t0 = h->curTime;
        for (int z = 0; z < kMaxZones; z++) r.errDrive0[z] = h->errDrive[z];
        r.skipFold0 = h->skipFold ? 1u : 0u;
        r.n = (uint32_t)m;
        r.barBase = h->resBarBase;
        r.flagBase = h->resFlagBase;
        r.flagInBase = h->resFlagInBase;
        r.pbBase = h->resPbBase;
        void *args[] = {&sa, &r};
        const int rc = launchResident(h, kern, grid, args);
        if (rc) return rc;
        resAdvance(h, (uint32_t)m);
        k0 += m;
    }
    HIP_TRY(hipMemcpyAsync(h->resStatusHost, h->resStatus, 4, hipMemcpyDeviceToHost, h->stream));
    h->resCheck = true;
    return CBM_OK;
}

// The fields of a resident run that do not depend on its steps.
static void runFields(cbm_handle *h, ResRun *rp)
{
    ResRun &r = *rp;
    grConstArgs(h, &r.g);
    r.maxExtIncVIO = h->p.maxExtIncVIO;
    r.gIncGRtoGO = h->gIncGRtoGO;
    r.spans = h->resSpans;
    r.recOffPF = (h->rec && h->recOffPF != 0xFFFFFFFFu) ? h->recOffPF : 0xFFFFFFFFu;
    r.bar = h->resBar;
    r.flag = h->resBar + kBarGroups * kBarStride;
    r.flagIn = h->resBar + (kBarGroups + kFlagCopies) * kBarStride;
    r.pb = h->resBar + (kBarGroups + 2 * kFlagCopies) * kBarStride;
    r.apSC3 = h->resApSC3;
    r.waitTicks = kResWaitTicks;
    r.hostSeq = nullptr;
    r.hostIn = nullptr;
    r.inPitch = r.inErr = 0;
    r.hostOutSeq = nullptr;
    r.hostOut = nullptr;
    r.outWords = r.ioPop = 0;
    r.errDev = nullptr;
    r.mfFlag = h->resMfFlag;
    r.status = h->resStatus;
    r.recLds = h->residentRecLds;
}

// The host's records after m steps of a resident launch.
static void resAdvance(cbm_handle *h, uint32_t m)
{
    h->resBarBase += m * h->resSpans;  // each GR block counts each step
    h->resFlagBase += m;
    h->resFlagInBase += m;
    h->resPbBase += m;
    h->curTime += (uint64_t)m;
    for (int z = 0; z < kMaxZones; z++) h->errDrive[z] = 0.0f;  // used by the first step
    h->skipFold = false;
}

extern "C" {

int cbm_loop_begin(cbm_handle *h, int max_steps, int population)
{
    if (!h) return fail(CBM_E_ARG, "null handle");
    if (!h->committed) return fail(CBM_E_STATE, "cbm_commit has not been called");
    if (h->looping) return fail(CBM_E_STATE, "cbm_loop_begin: a closed loop is running already");
    if (h->ioPending) return fail(CBM_E_STATE, "cbm_step_io_end has not been called for the last step");
    if (!h->resident) return fail(CBM_E_UNSUPPORTED, "cbm_loop_begin: the network does not run resident (cbm_resident)");
    if (population != CBM_REC_PC && population != CBM_REC_BC && population != CBM_REC_IO && population != CBM_REC_NC)
        return fail(CBM_E_ARG, "cbm_loop_begin: population must be PC, BC, IO or NC");
    if (max_steps < 1 || max_steps > kResMaxSteps) return fail(CBM_E_ARG, "cbm_loop_begin: max_steps out of range");
    HIP_TRY(hipSetDevice(h->o.device));
    int rc = syncAll(h);
    if (rc) return rc;
    const cbm_sizes &s = h->s;
    const uint32_t nZ = s.numZones;
    h->loopInErr = roundUp(s.numMF, 16);
    h->loopInPitch = roundUp(h->loopInErr + 4 * nZ, 256);
    h->loopOutBytes = (uint32_t)recCells(s, population);
    h->loopOutWords = (h->loopOutBytes + 3) / 4;
    h->loopOutOff = (uint32_t)(kLoopInOff + 2 * (size_t)h->loopInPitch);
    const size_t hostBytes = h->loopOutOff + 2 * roundUp(4 * h->loopOutWords, 256);
    if (hostBytes > h->loopHostCap) {
        if (h->loopHost) (void)hipHostFree(h->loopHost);
        h->loopHost = nullptr;
        h->loopHostCap = 0;
        HIP_TRY(hipHostMalloc((void **)&h->loopHost, hostBytes, hipHostMallocCoherent));
        h->loopHostCap = hostBytes;
    }
    std::memset(h->loopHost, 0, hostBytes);
    h->loopPitch = roundUp(s.numMF, 256);
    if ((size_t)max_steps > h->loopCap) {
        if (h->loopMF) (void)hipFree(h->loopMF);
        if (h->loopErr) (void)hipFree(h->loopErr);
        h->loopMF = nullptr;
        h->loopErr = nullptr;
        h->loopCap = 0;
        hipError_t e = hipMalloc((void **)&h->loopMF, (size_t)max_steps * h->loopPitch);
        if (e == hipSuccess) e = hipMalloc((void **)&h->loopErr, (size_t)max_steps * nZ * sizeof(float));
        if (e != hipSuccess) return fail(CBM_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
        h->loopCap = (size_t)max_steps;
    }
    if (!h->loopFlag) {
        const hipError_t e = hipMalloc((void **)&h->loopFlag, (size_t)kFlagCopies * kBarStride * 4);
        if (e != hipSuccess) return fail(CBM_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    }
    HIP_TRY(hipMemsetAsync(h->loopFlag, 0, (size_t)kFlagCopies * kBarStride * 4, h->stream));
    ResRun r;
    runFields(h, &r);
    r.mf = h->loopMF;
    r.pitch = h->loopPitch;
    r.err = nullptr;
    r.t0 = h->curTime;
    for (int z = 0; z < kMaxZones; z++) r.errDrive0[z] = 0.0f;  // (each step's drive comes with its MF row)
    r.skipFold0 = h->skipFold ? 1u : 0u;
    r.n = (uint32_t)max_steps;
    r.barBase = h->resBarBase;
    r.flagBase = h->resFlagBase;
    r.flagInBase = h->resFlagInBase;
    r.pbBase = h->resPbBase;
    r.waitTicks = kLoopWaitTicks;
    r.hostSeq = reinterpret_cast<const uint32_t *>(h->loopHost);
    r.hostIn = h->loopHost + kLoopInOff;
    r.inPitch = h->loopInPitch;
    r.inErr = h->loopInErr;
    r.hostOutSeq = reinterpret_cast<uint32_t *>(h->loopHost + kLoopOutSeqOff);
    r.hostOut = reinterpret_cast<uint32_t *>(h->loopHost + h->loopOutOff);
    r.outWords = h->loopOutWords;
    r.ioPop = (uint32_t)population;
    r.errDev = h->loopErr;
    r.mfFlag = h->loopFlag;
    const SmallArgs *sa = h->smallArgs;
    void *args[] = {&sa, &r};
    if ((rc = launchResident(h, residentKernel(h, true), h->resSpans + 3, args))) return rc;
    h->looping = true;
    h->loopK = 0;
    h->loopMax = max_steps;
    h->loopPop = population;
    return CBM_OK;
}

int cbm_loop_end(cbm_handle *h, int *steps)
{
    if (!h) return fail(CBM_E_ARG, "null handle");
    if (!h->looping) return fail(CBM_E_STATE, "cbm_loop_end: no closed loop is running");
    HIP_TRY(hipSetDevice(h->o.device));
    if (h->ioPending) {  // a step posted and not collected: collect it, so that it counts
        std::vector<uint8_t> out(h->loopOutBytes);
        const int rc = cbm_step_io_end(h, out.data());
        if (rc) {
            h->looping = false;
            h->ioPending = false;
            return rc;
        }
    }
    // the stop before the next step; the launch ends after the steps posted
    const uint32_t m = (uint32_t)h->loopK;
    __atomic_store_n(reinterpret_cast<uint32_t *>(h->loopHost), kHostStop | m, __ATOMIC_RELEASE);
    h->looping = false;
    h->ioPending = false;
    if (m > 0)  // the MF input field holds the last step's row, as after cbm_step_io
        HIP_TRY(hipMemcpyAsync(h->dmf, h->loopMF + (size_t)(m - 1) * h->loopPitch, h->s.numMF,
                               hipMemcpyDeviceToDevice, h->stream));
    HIP_TRY(hipMemcpyAsync(h->resStatusHost, h->resStatus, 4, hipMemcpyDeviceToHost, h->stream));
    h->resCheck = true;
    resAdvance(h, m);
    if (steps) *steps = (int)m;
    return syncAll(h);
}

int cbm_run_mf_train(cbm_handle *h, const uint8_t *mf_dev, size_t pitch, int n, const float *err)
{
    int rc = checkTrain(h, mf_dev, pitch, n);
    if (rc) return rc;
    HIP_TRY(hipSetDevice(h->o.device));
    if (h->resident && n >= 2) {
        if ((rc = runResident(h, mf_dev, pitch, n, err))) return rc;
    } else {
        for (int k = 0; k < n; k++) {
            if (err)
                for (uint32_t z = 0; z < h->s.numZones; z++) {
                    const float e = err[(size_t)k * h->s.numZones + z];
                    if (e != 0.0f) h->errDrive[z] = e * h->p.maxExtIncVIO;
                }
            if ((rc = launchStep(h, mf_dev + (size_t)k * pitch, kNoTiming))) return rc;
        }
    }
    if ((rc = join(h))) return rc;
    if (n > 0) {
        HIP_TRY(hipMemcpyAsync(h->dmf, mf_dev + (size_t)(n - 1) * pitch, h->s.numMF, hipMemcpyDeviceToDevice, h->stream));
        HIP_TRY(hipEventRecord(h->evG, h->stream));
    }
    return CBM_OK;
}

int cbm_sync(cbm_handle *h)
{
    if (!h) return fail(CBM_E_ARG, "null handle");
    HIP_TRY(hipSetDevice(h->o.device));
    return syncAll(h);
}

void *cbm_stream(cbm_handle *h) { return h ? (void *)h->stream : nullptr; }

int cbm_stream_ceiling(int device, uint64_t cells, int read_words, int write_words, int launches,
                       double *us_per_launch, uint64_t *bytes_per_launch)
{
    if (!us_per_launch || cells == 0 || launches <= 0 || read_words < 1 || read_words > kMaxStreamWords ||
        write_words < 0 || write_words > read_words)
        return fail(CBM_E_ARG, "bad argument");
    if (cells > (uint64_t)UINT32_MAX * 4096u) return fail(CBM_E_ARG, "too many cells");
    HIP_TRY(hipSetDevice(device));
    StreamArgs a{};
    a.cells = cells;
    a.nRead = read_words;
    a.nWrite = write_words;
    std::vector<void *> bufs;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    auto done = [&](int rc) {
        for (void *p : bufs) (void)hipFree(p);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (st) (void)hipStreamDestroy(st);
        return rc;
    };
    hipError_t e = hipSuccess;
    for (int w = 0; w < read_words && e == hipSuccess; w++) {
        void *p = nullptr;
        e = hipMalloc(&p, cells * 4);
        if (e == hipSuccess) {
            bufs.push_back(p);
            a.word[w] = (uint32_t *)p;
            e = hipMemset(p, 0, cells * 4);
        }
    }
    if (e == hipSuccess) {
        void *p = nullptr;
        e = hipMalloc(&p, cells * 16);
        if (e == hipSuccess) {
            bufs.push_back(p);
            a.rec = (const uint4 *)p;
            e = hipMemset(p, 0, cells * 16);
        }
    }
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e != hipSuccess) return done(fail(CBM_E_HIP, std::string("stream ceiling setup: ") + hipGetErrorString(e)));
    const dim3 grid((uint32_t)((cells + 4095) / 4096)), block(1024);
    for (int i = 0; i < 10; i++) hipLaunchKernelGGL(stream_ceiling_kernel, grid, block, 0, st, a);
    e = hipEventRecord(e0, st);
    for (int i = 0; i < launches && e == hipSuccess; i++) {
        hipLaunchKernelGGL(stream_ceiling_kernel, grid, block, 0, st, a);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(e1, st);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e != hipSuccess) return done(fail(CBM_E_HIP, std::string("stream ceiling: ") + hipGetErrorString(e)));
    *us_per_launch = ms * 1000.0 / launches;
    if (bytes_per_launch) *bytes_per_launch = cells * (uint64_t)(4 * read_words + 16 + 4 * write_words);
    return done(CBM_OK);
}

// n steps with event pairs around the timed launches (one pair per step, or
// two in split mode), read back after the whole run so that timing inserts no
// host waits between steps.  us[0], us[1]: mean durations of the pairs;
// *stepUs: mean step time over the run.
static int profileRun(cbm_handle *h, const uint8_t *mf_dev, size_t pitch, int n, int timing, double us[2],
                      double *stepUs)
{
    int rc = checkTrain(h, mf_dev, pitch, n);
    if (rc) return rc;
    if (n == 0) return fail(CBM_E_ARG, "n must be positive");
    HIP_TRY(hipSetDevice(h->o.device));
    if ((rc = syncAll(h))) return rc;
    const int per = timing == kSplit ? 4 : 2;
    std::vector<hipEvent_t> ev((size_t)per * n + 2);
    for (auto &e : ev) HIP_TRY(hipEventCreate(&e));
    // a fused step is one kernel: it is timed over the whole region of n
    // back-to-back launches (per-launch events would add their own overhead)
    const bool region = timing == kTimeKernel && h->fused();
    HIP_TRY(hipStreamWaitEvent(h->stream, h->evS, 0));
    HIP_TRY(hipEventRecord(ev[(size_t)per * n], h->stream));
    for (int k = 0; k < n; k++) {
        for (int e = 0; e < per; e++) h->ev[e] = ev[(size_t)per * k + e];
        if ((rc = launchStep(h, mf_dev + (size_t)k * pitch, region ? kNoTiming : timing))) break;
    }
    for (auto &e : h->ev) e = nullptr;
    if (!rc) rc = join(h);
    if (!rc) {
        HIP_TRY(hipEventRecord(ev[(size_t)per * n + 1], h->stream));
        HIP_TRY(hipEventSynchronize(ev[(size_t)per * n + 1]));
        double tot[2] = {0, 0};
        for (int k = 0; k < n && !region; k++)
            for (int j = 0; j < per / 2; j++) {
                float ms = 0;
                HIP_TRY(hipEventElapsedTime(&ms, ev[(size_t)per * k + 2 * j], ev[(size_t)per * k + 2 * j + 1]));
                tot[j] += ms;
            }
        float total = 0;
        HIP_TRY(hipEventElapsedTime(&total, ev[(size_t)per * n], ev[(size_t)per * n + 1]));
        us[0] = (region ? (double)total : tot[0]) * 1000.0 / n;
        us[1] = tot[1] * 1000.0 / n;
        if (stepUs) *stepUs = (double)total * 1000.0 / n;
        HIP_TRY(hipMemcpyAsync(h->dmf, mf_dev + (size_t)(n - 1) * pitch, h->s.numMF, hipMemcpyDeviceToDevice,
                               h->stream));
        HIP_TRY(hipEventRecord(h->evG, h->stream));
    }
    HIP_TRY(hipStreamSynchronize(h->stream));
    for (auto &e : ev) (void)hipEventDestroy(e);
    return rc;
}

int cbm_profile_steps(cbm_handle *h, const uint8_t *mf_dev, size_t pitch, int n, double *kernelUs, double *stepUs)
{
    double us[2] = {0, 0};
    const int rc = profileRun(h, mf_dev, pitch, n, kTimeKernel, us, stepUs);
    if (!rc && kernelUs) *kernelUs = us[0];
    return rc;
}

int cbm_profile_split(cbm_handle *h, const uint8_t *mf_dev, size_t pitch, int n, double *grUs, double *smallUs)
{
    double us[2] = {0, 0};
    const int rc = profileRun(h, mf_dev, pitch, n, kSplit, us, nullptr);
    if (!rc) {
        if (grUs) *grUs = us[0];
        if (smallUs) *smallUs = us[1];
    }
    return rc;
}

int cbm_step_mode(cbm_handle *h)
{
    if (!h) return fail(CBM_E_ARG, "null handle");
    if (!h->committed) return fail(CBM_E_STATE, "cbm_commit has not been called");
    return h->mode;
}

int cbm_resident(cbm_handle *h)
{
    if (!h) return fail(CBM_E_ARG, "null handle");
    if (!h->committed) return fail(CBM_E_STATE, "cbm_commit has not been called");
    return h->resident ? 1 : 0;
}

int cbm_gr_launch(cbm_handle *h, uint32_t *blocks, uint32_t *threads, uint32_t *cellsPerThread, uint32_t *evenSplit)
{
    if (!h || !blocks || !threads || !cellsPerThread || !evenSplit) return fail(CBM_E_ARG, "null argument");
    if (!h->committed) return fail(CBM_E_STATE, "cbm_commit has not been called");
    const int nt = h->fused() ? 1024 : h->nt;
    const int cpt = h->spans ? h->spanCpt : (h->fused() ? h->fusedCpt : h->cpt);
    *blocks = h->spans ? h->spans : h->s.numGR / (uint32_t)(nt * cpt);
    *threads = (uint32_t)nt;
    *cellsPerThread = (uint32_t)cpt;
    *evenSplit = h->spans ? 1u : 0u;
    return CBM_OK;
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

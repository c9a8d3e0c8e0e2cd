// cbm_batch.hip -- host side of the cbm_batch_* C-ABI (include/cbm_hip.h):
// a ragged pack of committed networks stepped together, one launch per step.
//
// Block n < count of the launch steps network n's small populations; the
// other blocks are GR tiles of every network, listed in a tile table built
// largest network first.  Each network keeps its own sizes, parameters and
// state buffers (those of its handle); the batch only adds a device array of
// per-network descriptors (BatchNet) and the tile table, so a batched step
// computes exactly what each network's own step computes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "cbm_handle.hpp"

namespace cbmdev {
// FMT: 1 = compact connectivity records, 2 = the five arrays, 0 = decided per tile
template <int CPT, int FMT>
__global__ void batch_step_kernel(const BatchNet *nets, const uint2 *tiles, uint32_t nNets, const uint32_t *kbase,
                                  uint32_t k);
__global__ void batch_advance_kernel(uint32_t *kbase, uint32_t steps);
__global__ void sweep_step_kernel(const BatchNet *nets, const SweepNet *sw, uint32_t nNets, const SweepRun *run,
                                  const uint32_t *kbase, uint32_t k);
__global__ void sweep_q_kernel(const float *prob, uint64_t pitch, uint32_t n, uint32_t nCl, float *Q);
__global__ void cluster_mean_kernel(const float *w, uint32_t numGR, uint32_t nZ, uint32_t log2Cells, float *W);
}  // namespace cbmdev

using namespace cbmdev;

struct cbm_batch {
    std::vector<cbm_handle *> nets;
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;       // end of the last run, waited on by the networks' streams
    int cpt = 2;                     // GR cells per thread of the batched tiles
    int fmt = 0;                     // connectivity format shared by the pack (0: mixed)
    uint32_t graphSteps = 0;         // steps per captured graph (0: plain launches)
    std::vector<BatchNet> hnet;
    BatchNet *dnet = nullptr;
    uint2 *dtiles = nullptr;
    uint32_t ntiles = 0;
    size_t lds = 0;
    uint32_t *kbase = nullptr;       // run-step counter of a captured graph
    float *derr = nullptr;           // error drive trains of the current run
    size_t derrCap = 0;
    hipGraphExec_t graph = nullptr;  // captured window of graphSteps launches
    int graphCpt = 0;
    // surrogate sweep: cluster weights per network (cbm_batch_sweep)
    std::vector<SweepNet> hsw;
    SweepNet *dsw = nullptr;
    size_t swLds = 0;
    uint32_t swCells = 0;
    float *swQ = nullptr;            // per-step bin probabilities of the plasticity rule
    size_t swQCap = 0;
    SweepRun hswRun{}, *dswRun = nullptr;  // probability rows of the current sweep run
    hipGraphExec_t swGraph = nullptr;      // kSweepGraphSteps sweep steps reading the counter kbase
};

constexpr uint32_t kSweepGraphSteps = 64;

// Staging of the error-drive rows of cbm_batch_run_mf_train_noalloc (1 MiB,
// CBM_NOALLOC_ERR_FLOATS floats): a file-scope array that every GPU the
// library runs on holds until the process ends.  Host code reaches it only
// through its device address, resolved once per device with that device
// current (the array's name is a host-side stand-in, not a device address).
static __device__ __align__(256) float scratch[CBM_NOALLOC_ERR_FLOATS];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratchMu;
float *g_scratch[kMaxDevices];  // device address of `scratch` on each device (nullptr: not resolved)

// The calling thread's current device must be `device`.
int resolveScratch(int device)
{
    if (device < 0 || device >= kMaxDevices) return fail(CBM_E_UNSUPPORTED, "device ordinal above 63");
    std::lock_guard<std::mutex> lock(g_scratchMu);
    if (g_scratch[device]) return CBM_OK;
    float *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
        return fail(CBM_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch[device] = ptr;
    return CBM_OK;
}

float *scratchOf(int device)
{
    std::lock_guard<std::mutex> lock(g_scratchMu);
    return device >= 0 && device < kMaxDevices ? g_scratch[device] : nullptr;
}
}  // namespace

namespace {

const void *batchKernel(int cpt, int fmt)
{
#define CBM_K(C) (fmt == 1   ? (const void *)batch_step_kernel<C, 1> \
                  : fmt == 2 ? (const void *)batch_step_kernel<C, 2> \
                             : (const void *)batch_step_kernel<C, 0>)
    return cpt == 2 ? CBM_K(2) : CBM_K(1);
#undef CBM_K
}

int buildTiles(cbm_batch *b)
{
    std::vector<int> order(b->nets.size());
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](int x, int y) { return b->nets[x]->s.numGR > b->nets[y]->s.numGR; });
    std::vector<uint2> tiles;
    for (int n : order) {
        const uint32_t nt = b->nets[n]->s.numGR / (1024u * (uint32_t)b->cpt);
        for (uint32_t t = 0; t < nt; t++) tiles.push_back(make_uint2((uint32_t)n, t));
    }
    if (b->dtiles) (void)hipFree(b->dtiles);
    b->dtiles = nullptr;
    HIP_TRY(hipMalloc((void **)&b->dtiles, tiles.size() * sizeof(uint2)));
    HIP_TRY(hipMemcpy(b->dtiles, tiles.data(), tiles.size() * sizeof(uint2), hipMemcpyHostToDevice));
    b->ntiles = (uint32_t)tiles.size();
    size_t lds = 0;
    for (cbm_handle *h : b->nets)
        lds = std::max(lds, std::max<size_t>(h->smallLds, grTileLds(1024, b->cpt, h->s.numGO, h->s.numMF,
                                                                         h->s.numZones, h->s.numIO)));
    b->lds = lds;
    bool allCompact = true, allArrays = true;
    for (cbm_handle *h : b->nets) {
        allCompact = allCompact && h->conn != nullptr;
        allArrays = allArrays && h->conn == nullptr;
    }
    b->fmt = allCompact ? 1 : allArrays ? 2 : 0;
    if (lds > 64 * 1024)
        HIP_TRY(hipFuncSetAttribute(batchKernel(b->cpt, b->fmt), hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    if (b->graph) (void)hipGraphExecDestroy(b->graph);
    b->graph = nullptr;
    return CBM_OK;
}

int launch(cbm_batch *b, const uint32_t *kbase, uint32_t k)
{
    const BatchNet *nets = b->dnet;
    const uint2 *tiles = b->dtiles;
    uint32_t nNets = (uint32_t)b->nets.size();
    void *args[] = {&nets, &tiles, &nNets, &kbase, &k};
    HIP_TRY(hipLaunchKernel(batchKernel(b->cpt, b->fmt), dim3(nNets + b->ntiles), dim3(kSmallThreads), args, b->lds,
                            b->stream));
    return CBM_OK;
}

// Capture graphSteps launches reading the counter kbase, plus its advance.
int captureGraph(cbm_batch *b)
{
    if (b->graph && b->graphCpt == b->cpt) return CBM_OK;
    if (b->graph) (void)hipGraphExecDestroy(b->graph);
    b->graph = nullptr;
    hipGraph_t g = nullptr;
    HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    int rc = CBM_OK;
    for (uint32_t k = 0; k < b->graphSteps && !rc; k++) rc = launch(b, b->kbase, k);
    if (!rc) hipLaunchKernelGGL(batch_advance_kernel, dim3(1), dim3(64), 0, b->stream, b->kbase, b->graphSteps);
    const hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (rc) return rc;
    if (e != hipSuccess) return fail(CBM_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    const hipError_t e2 = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e2 != hipSuccess) return fail(CBM_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e2));
    b->graphCpt = b->cpt;
    return CBM_OK;
}

// Every network's device and host state that a step changes, copied aside
// and back (cbm_batch_autotune runs trial steps on the live networks).
struct Snapshot {
    std::vector<std::pair<void *, size_t>> bufs;  // (device pointer, bytes)
    std::vector<void *> copies;
    struct Host {
        uint64_t curTime;
        float errDrive[kMaxZones];
        bool skipFold;
    };
    std::vector<Host> host;
};

void listState(cbm_handle *h, std::vector<std::pair<void *, size_t>> *out)
{
    const cbm_sizes &s = h->s;
    for (int f = 0; f < CBM_F_COUNT; f++) {
        size_t e, r, c;
        int zoned, kind;
        if (!h->dfield[f] || cbm_field_shape(f, &s, &e, &r, &c, &zoned, &kind) != CBM_OK) continue;
        if (kind == CBM_KIND_STATE || kind == CBM_KIND_RUNTIME) out->push_back({h->dfield[f], h->host[f].size()});
    }
    const size_t nZ = s.numZones;
    out->push_back({h->apGO2, 2 * (size_t)s.numGO});
    out->push_back({h->timer2, 2 * nZ * s.numIO * 4});
    out->push_back({h->chunk2, 2 * nZ * (s.numGR / 64) * 4});
    out->push_back({h->pfpcRing, (size_t)kRing * nZ * s.numPC * 4});
    out->push_back({h->bcRing, (size_t)kRing * s.numBC * 4});
    out->push_back({h->scRing, (size_t)kRing * s.numSC * 4});
    out->push_back({h->goRing, (size_t)kRing * s.numGO * 4});
    out->push_back({h->rngState, 2 * nZ * kRngZone * 4});
    out->push_back({h->rngNext, 2 * nZ * kRngNext * 4});
    out->push_back({h->dmf, s.numMF});
}

int takeSnapshot(cbm_batch *b, Snapshot *sn)
{
    sn->bufs.clear();
    sn->host.clear();
    for (cbm_handle *h : b->nets) {
        listState(h, &sn->bufs);
        Snapshot::Host hs;
        hs.curTime = h->curTime;
        std::memcpy(hs.errDrive, h->errDrive, sizeof(hs.errDrive));
        hs.skipFold = h->skipFold;
        sn->host.push_back(hs);
    }
    for (auto &bf : sn->bufs) {
        void *p = nullptr;
        HIP_TRY(hipMalloc(&p, bf.second));
        sn->copies.push_back(p);
        HIP_TRY(hipMemcpyAsync(p, bf.first, bf.second, hipMemcpyDeviceToDevice, b->stream));
    }
    HIP_TRY(hipStreamSynchronize(b->stream));
    return CBM_OK;
}

int restoreSnapshot(cbm_batch *b, Snapshot *sn)
{
    for (size_t i = 0; i < sn->bufs.size(); i++)
        HIP_TRY(hipMemcpyAsync(sn->bufs[i].first, sn->copies[i], sn->bufs[i].second, hipMemcpyDeviceToDevice,
                               b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    for (void *p : sn->copies) (void)hipFree(p);
    sn->copies.clear();
    for (size_t n = 0; n < b->nets.size(); n++) {
        cbm_handle *h = b->nets[n];
        h->curTime = sn->host[n].curTime;
        std::memcpy(h->errDrive, sn->host[n].errDrive, sizeof(h->errDrive));
        h->skipFold = sn->host[n].skipFold;
    }
    return CBM_OK;
}

// Run n batched steps; time the launches when evs is given (one pair).
// noAlloc: stage the error drive in the device's `scratch` array and launch
// step by step (no buffer growth, no graph capture).
int run(cbm_batch *b, const uint8_t *const *mf, const size_t *pitch, int n, const float *const *err,
        hipEvent_t *evs, const float *prob = nullptr, size_t probPitch = 0, bool noAlloc = false)
{
    if (!mf || !pitch || n < 0) return fail(CBM_E_ARG, "bad argument");
    HIP_TRY(hipSetDevice(b->device));
    HIP_TRY(hipStreamSynchronize(b->stream));  // the previous run still reads the descriptors
    const size_t N = b->nets.size();
    int rc;
    for (size_t i = 0; i < N; i++) {
        if ((rc = checkTrain(b->nets[i], mf[i], pitch[i], n))) return rc;
        if ((rc = syncAll(b->nets[i]))) return rc;
    }
    if (n == 0) return CBM_OK;
    // error drive trains of this run, packed into one device buffer
    size_t errFloats = 0;
    for (size_t i = 0; i < N; i++)
        if (err && err[i]) errFloats += (size_t)n * b->nets[i]->s.numZones;
    float *errStage = b->derr;
    if (noAlloc) {
        if (errFloats > CBM_NOALLOC_ERR_FLOATS)
            return fail(CBM_E_ARG, "cbm_batch_run_mf_train_noalloc: n x (numZones of the networks with an error "
                                   "drive) exceeds CBM_NOALLOC_ERR_FLOATS");
        errStage = scratchOf(b->device);
        if (!errStage) return fail(CBM_E_STATE, "scratch array not resolved on this device");
    } else if (errFloats > b->derrCap) {
        if (b->derr) (void)hipFree(b->derr);
        b->derr = nullptr;
        b->derrCap = 0;
        HIP_TRY(hipMalloc((void **)&b->derr, errFloats * 4));
        b->derrCap = errFloats;
        errStage = b->derr;
    }
    size_t off = 0;
    for (size_t i = 0; i < N; i++) {
        cbm_handle *h = b->nets[i];
        BatchNet &d = b->hnet[i];
        grConstArgs(h, &d.gr);  // current LTD/LTP steps
        d.mf = mf[i];
        d.pitch = pitch[i];
        d.err = nullptr;
        if (err && err[i]) {
            const size_t cnt = (size_t)n * h->s.numZones;
            HIP_TRY(hipMemcpy(errStage + off, err[i], cnt * 4, hipMemcpyHostToDevice));
            d.err = errStage + off;
            off += cnt;
        }
        d.t0 = h->curTime;
        std::memcpy(d.errDrive0, h->errDrive, sizeof(d.errDrive0));
        d.gIncGRtoGO = h->gIncGRtoGO;
        d.skipFold0 = h->skipFold ? 1u : 0u;
    }
    HIP_TRY(hipMemcpy(b->dnet, b->hnet.data(), N * sizeof(BatchNet), hipMemcpyHostToDevice));

    if (evs) HIP_TRY(hipEventRecord(evs[0], b->stream));
    int k = 0;
    if (prob) {
        uint32_t maxCl = 0;
        for (const SweepNet &w : b->hsw) maxCl = std::max(maxCl, w.nCl);
        const size_t need = (size_t)n * probPitch;
        if (need > b->swQCap) {
            if (b->swQ) (void)hipFree(b->swQ);
            b->swQ = nullptr;
            b->swQCap = 0;
            HIP_TRY(hipMalloc((void **)&b->swQ, need * 4));
            b->swQCap = need;
        }
        const size_t cnt = (size_t)n * maxCl;
        hipLaunchKernelGGL(sweep_q_kernel, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, b->stream, prob,
                           (uint64_t)probPitch, (uint32_t)n, maxCl, b->swQ);
        HIP_TRY(hipGetLastError());
        b->hswRun = SweepRun{prob, b->swQ, (uint64_t)probPitch};
        HIP_TRY(hipMemcpyAsync(b->dswRun, &b->hswRun, sizeof(SweepRun), hipMemcpyHostToDevice, b->stream));
        const dim3 grid(2 * (uint32_t)N), block(kSmallThreads);
        if ((uint32_t)n >= kSweepGraphSteps) {
            if (!b->swGraph) {
                hipGraph_t g = nullptr;
                HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
                for (uint32_t j = 0; j < kSweepGraphSteps; j++)
                    hipLaunchKernelGGL(sweep_step_kernel, grid, block, b->swLds, b->stream, b->dnet, b->dsw,
                                       (uint32_t)N, b->dswRun, b->kbase, j);
                hipLaunchKernelGGL(batch_advance_kernel, dim3(1), dim3(64), 0, b->stream, b->kbase, kSweepGraphSteps);
                const hipError_t e = hipStreamEndCapture(b->stream, &g);
                if (e != hipSuccess) return fail(CBM_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
                const hipError_t e2 = hipGraphInstantiate(&b->swGraph, g, nullptr, nullptr, 0);
                (void)hipGraphDestroy(g);
                if (e2 != hipSuccess) return fail(CBM_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e2));
            }
            const uint32_t zero = 0;
            HIP_TRY(hipMemcpyAsync(b->kbase, &zero, 4, hipMemcpyHostToDevice, b->stream));
            HIP_TRY(hipStreamSynchronize(b->stream));  // the host values above must be copied before they go
            for (; k + (int)kSweepGraphSteps <= n; k += (int)kSweepGraphSteps) HIP_TRY(hipGraphLaunch(b->swGraph, b->stream));
        }
        for (; k < n; k++) {
            hipLaunchKernelGGL(sweep_step_kernel, grid, block, b->swLds, b->stream, b->dnet, b->dsw, (uint32_t)N,
                               b->dswRun, nullptr, (uint32_t)k);
            HIP_TRY(hipGetLastError());
        }
    }
    if (b->graphSteps && (uint32_t)n >= b->graphSteps && !prob && !noAlloc) {
        if ((rc = captureGraph(b))) return rc;
        const uint32_t zero = 0;
        HIP_TRY(hipMemcpyAsync(b->kbase, &zero, 4, hipMemcpyHostToDevice, b->stream));
        HIP_TRY(hipStreamSynchronize(b->stream));  // the host value above must be copied before it goes
        for (; k + (int)b->graphSteps <= n; k += (int)b->graphSteps) HIP_TRY(hipGraphLaunch(b->graph, b->stream));
    }
    for (; k < n; k++)
        if ((rc = launch(b, nullptr, (uint32_t)k))) return rc;
    if (evs) HIP_TRY(hipEventRecord(evs[1], b->stream));

    // each network now stands after step t0 + n: its host shadows follow, its
    // streams wait for the batch, and its MF buffer holds the last row
    for (size_t i = 0; i < N; i++) {
        cbm_handle *h = b->nets[i];
        h->curTime += (uint64_t)n;
        for (int z = 0; z < kMaxZones; z++) h->errDrive[z] = 0.0f;
        // after a sweep the ring slot of the last step holds the cluster
        // tile's PF->PC sums and the chunk sums are stale: nothing to fold
        h->skipFold = prob != nullptr;
        HIP_TRY(hipMemcpyAsync(h->dmf, mf[i] + (size_t)(n - 1) * pitch[i], h->s.numMF, hipMemcpyDeviceToDevice,
                               b->stream));
    }
    HIP_TRY(hipEventRecord(b->done, b->stream));
    for (cbm_handle *h : b->nets) {
        HIP_TRY(hipStreamWaitEvent(h->stream, b->done, 0));
        HIP_TRY(hipStreamWaitEvent(h->stream2, b->done, 0));
        HIP_TRY(hipEventRecord(h->evG, h->stream));
        HIP_TRY(hipEventRecord(h->evS, h->stream2));
    }
    return CBM_OK;
}

// Cluster weights of every network from its per-GR PF->PC weights.
int sweepSetup(cbm_batch *b, uint32_t cells)
{
    if (!cells || (cells & (cells - 1)) || cells < 64) return fail(CBM_E_ARG, "cellsPerCluster must be a power of two >= 64");
    uint32_t log2 = 0;
    while ((1u << log2) < cells) log2++;
    size_t lds = b->lds;
    for (cbm_handle *h : b->nets) {
        const cbm_sizes &sz = h->s;
        const uint32_t nCl = sz.numGR / cells;
        if (sz.numGR % cells || nCl % sz.numPC || nCl % sz.numBC || nCl % sz.numSC || nCl % sz.numIO || nCl == 0)
            return fail(CBM_E_UNSUPPORTED, "clusters must tile every PC, BC, SC and IO block of GRs");
        if (nCl > 8 * (uint32_t)kSmallThreads) return fail(CBM_E_UNSUPPORTED, "at most 8192 clusters per network");
        lds = std::max(lds, (size_t)4 * (sz.numZones * sz.numIO + (size_t)nCl * (1 + sz.numZones)));
    }
    HIP_TRY(hipSetDevice(b->device));
    HIP_TRY(hipStreamSynchronize(b->stream));
    for (SweepNet &w : b->hsw)
        if (w.W) (void)hipFree(w.W);
    b->hsw.assign(b->nets.size(), SweepNet{nullptr, 0, 0});
    for (size_t i = 0; i < b->nets.size(); i++) {
        cbm_handle *h = b->nets[i];
        SweepNet &w = b->hsw[i];
        w.nCl = h->s.numGR / cells;
        w.log2Cells = log2;
        const size_t n = (size_t)h->s.numZones * w.nCl;
        HIP_TRY(hipMalloc((void **)&w.W, n * 4));
        int rc = syncAll(h);
        if (rc) return rc;
        hipLaunchKernelGGL(cluster_mean_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, b->stream,
                           b->hnet[i].gr.w, h->s.numGR, h->s.numZones, log2, w.W);
        HIP_TRY(hipGetLastError());
    }
    if (!b->dsw) HIP_TRY(hipMalloc((void **)&b->dsw, b->nets.size() * sizeof(SweepNet)));
    if (!b->dswRun) HIP_TRY(hipMalloc((void **)&b->dswRun, sizeof(SweepRun)));
    if (b->swGraph) (void)hipGraphExecDestroy(b->swGraph);  // captured with the old LDS size
    b->swGraph = nullptr;
    HIP_TRY(hipMemcpy(b->dsw, b->hsw.data(), b->hsw.size() * sizeof(SweepNet), hipMemcpyHostToDevice));
    if (lds > 64 * 1024)
        HIP_TRY(hipFuncSetAttribute((const void *)sweep_step_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)lds));
    b->swLds = lds;
    b->swCells = cells;
    HIP_TRY(hipStreamSynchronize(b->stream));
    return CBM_OK;
}

}  // namespace

extern "C" {

int cbm_batch_create(cbm_handle *const *nets, int count, const cbm_batch_options *opts, cbm_batch **out)
{
    if (!nets || count <= 0 || !out) return fail(CBM_E_ARG, "bad argument");
    for (int i = 0; i < count; i++) {
        if (!nets[i]) return fail(CBM_E_ARG, "null handle");
        if (!nets[i]->committed) return fail(CBM_E_STATE, "every network must be committed");
        if (nets[i]->batch) return fail(CBM_E_STATE, "a network can belong to one batch only");
        if (nets[i]->o.device != nets[0]->o.device) return fail(CBM_E_ARG, "the networks must share one device");
        for (int j = 0; j < i; j++)
            if (nets[j] == nets[i]) return fail(CBM_E_ARG, "a network appears twice");
        if (nets[i]->s.numGR % 2048u) return fail(CBM_E_UNSUPPORTED, "batched networks need numGR % 2048 == 0");
    }
    cbm_batch *b = new cbm_batch();
    b->nets.assign(nets, nets + count);
    b->device = nets[0]->o.device;
    if (opts && opts->cellsPerThread) b->cpt = (int)opts->cellsPerThread;
    if (opts) b->graphSteps = opts->graphSteps;
    auto bail = [&](int rc) {
        cbm_batch_destroy(b);
        return rc;
    };
    if (b->cpt != 1 && b->cpt != 2) return bail(fail(CBM_E_ARG, "cellsPerThread must be 1 or 2"));
    hipError_t e = hipSetDevice(b->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->done, hipEventDisableTiming);
    if (e == hipSuccess) e = hipMalloc((void **)&b->dnet, (size_t)count * sizeof(BatchNet));
    if (e == hipSuccess) e = hipMalloc((void **)&b->kbase, 4);
    if (e != hipSuccess) return bail(fail(CBM_E_HIP, std::string("batch setup: ") + hipGetErrorString(e)));
    b->hnet.resize(count);
    for (int i = 0; i < count; i++) {
        cbm_handle *h = nets[i];
        BatchNet &d = b->hnet[i];
        std::memset(&d, 0, sizeof(d));
        d.small = h->smallArgs;
        grConstArgs(h, &d.gr);
        d.chunk2 = h->chunk2;
        d.pfpcRing = h->pfpcRing;
        d.bcRing = h->bcRing;
        d.scRing = h->scRing;
        d.goRing = h->goRing;
        d.apGO2 = h->apGO2;
        d.timer2 = h->timer2;
        d.maxExtIncVIO = h->p.maxExtIncVIO;
        d.grgoLag = h->o.grgoLag;
        d.tsPerHistBinGR = h->p.tsPerHistBinGR;
        d.tsPerPopHistBinPC = h->p.tsPerPopHistBinPC;
        d.numTSinMFHist = h->p.numTSinMFHist;
    }
    int rc = buildTiles(b);
    if (rc) return bail(rc);
    if ((rc = resolveScratch(b->device))) return bail(rc);  // for cbm_batch_run_mf_train_noalloc
    for (cbm_handle *h : b->nets) h->batch = b;
    *out = b;
    return CBM_OK;
}

int cbm_batch_destroy(cbm_batch *b)
{
    if (!b) return fail(CBM_E_ARG, "null batch");
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    for (cbm_handle *h : b->nets)
        if (h->batch == b) h->batch = nullptr;
    if (b->graph) (void)hipGraphExecDestroy(b->graph);
    if (b->dnet) (void)hipFree(b->dnet);
    if (b->dtiles) (void)hipFree(b->dtiles);
    if (b->kbase) (void)hipFree(b->kbase);
    if (b->derr) (void)hipFree(b->derr);
    for (SweepNet &w : b->hsw)
        if (w.W) (void)hipFree(w.W);
    if (b->dsw) (void)hipFree(b->dsw);
    if (b->swQ) (void)hipFree(b->swQ);
    if (b->dswRun) (void)hipFree(b->dswRun);
    if (b->swGraph) (void)hipGraphExecDestroy(b->swGraph);
    if (b->done) (void)hipEventDestroy(b->done);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return CBM_OK;
}

int cbm_batch_run_mf_train(cbm_batch *b, const uint8_t *const *mf_dev, const size_t *pitch, int n,
                           const float *const *err_host)
{
    if (!b) return fail(CBM_E_ARG, "null batch");
    return run(b, mf_dev, pitch, n, err_host, nullptr);
}

int cbm_batch_run_mf_train_noalloc(cbm_batch *b, const uint8_t *const *mf_dev, const size_t *pitch, int n,
                                   const float *const *err_host)
{
    if (!b) return fail(CBM_E_ARG, "null batch");
    return run(b, mf_dev, pitch, n, err_host, nullptr, nullptr, 0, true);
}

int cbm_batch_sync(cbm_batch *b)
{
    if (!b) return fail(CBM_E_ARG, "null batch");
    HIP_TRY(hipSetDevice(b->device));
    HIP_TRY(hipStreamSynchronize(b->stream));
    return CBM_OK;
}

int cbm_batch_profile(cbm_batch *b, const uint8_t *const *mf_dev, const size_t *pitch, int n, double *kernel_us)
{
    if (!b || n <= 0) return fail(CBM_E_ARG, "bad argument");
    hipEvent_t ev[2];
    HIP_TRY(hipSetDevice(b->device));
    HIP_TRY(hipEventCreate(&ev[0]));
    HIP_TRY(hipEventCreate(&ev[1]));
    int rc = run(b, mf_dev, pitch, n, nullptr, ev);
    if (!rc) {
        float ms = 0;
        HIP_TRY(hipEventSynchronize(ev[1]));
        HIP_TRY(hipEventElapsedTime(&ms, ev[0], ev[1]));
        if (kernel_us) *kernel_us = ms * 1000.0 / n;
    }
    (void)hipEventDestroy(ev[0]);
    (void)hipEventDestroy(ev[1]);
    return rc;
}

int cbm_batch_autotune(cbm_batch *b, const uint8_t *const *mf_dev, const size_t *pitch, int steps, int *chosen)
{
    if (!b || steps <= 0) return fail(CBM_E_ARG, "bad argument");
    HIP_TRY(hipSetDevice(b->device));
    for (cbm_handle *h : b->nets) {
        int rc = syncAll(h);
        if (rc) return rc;
    }
    Snapshot sn;
    int rc = takeSnapshot(b, &sn);
    if (rc) {
        for (void *p : sn.copies) (void)hipFree(p);
        return rc;
    }
    const uint32_t graphSteps = b->graphSteps;
    b->graphSteps = 0;  // time the plain launches
    double best = 1e300;
    int bestCpt = b->cpt;
    for (int cpt : {1, 2}) {
        b->cpt = cpt;
        if ((rc = buildTiles(b))) break;
        double us = 0;
        // one untimed step first, then the timed ones, each from the snapshot
        if ((rc = restoreSnapshot(b, &sn)) || (rc = takeSnapshot(b, &sn))) break;
        if ((rc = cbm_batch_profile(b, mf_dev, pitch, steps, &us))) break;
        if (us < best) {
            best = us;
            bestCpt = cpt;
        }
        if ((rc = restoreSnapshot(b, &sn)) || (rc = takeSnapshot(b, &sn))) break;
    }
    b->graphSteps = graphSteps;
    const int rc2 = restoreSnapshot(b, &sn);
    if (rc) return rc;
    if (rc2) return rc2;
    b->cpt = bestCpt;
    if ((rc = buildTiles(b))) return rc;
    if (chosen) *chosen = bestCpt;
    return CBM_OK;
}

int cbm_batch_info(cbm_batch *b, int *cells_per_thread, int *blocks, int *graph_steps)
{
    if (!b) return fail(CBM_E_ARG, "null batch");
    if (cells_per_thread) *cells_per_thread = b->cpt;
    if (blocks) *blocks = (int)(b->nets.size() + b->ntiles);
    if (graph_steps) *graph_steps = (int)b->graphSteps;
    return CBM_OK;
}

}  // extern "C"

extern "C" {

int cbm_batch_sweep(cbm_batch *b, const float *prob, size_t probPitch, int n, const uint8_t *mf_dev, size_t mfPitch,
                    const float *const *err_host, uint32_t cellsPerCluster)
{
    if (!b || !prob || !mf_dev || n < 0) return fail(CBM_E_ARG, "bad argument");
    const size_t N = b->nets.size();
    uint32_t maxCl = 0;
    for (cbm_handle *h : b->nets) maxCl = std::max(maxCl, h->s.numGR / std::max(cellsPerCluster, 1u));
    if (probPitch < maxCl || ((uintptr_t)prob & 3)) return fail(CBM_E_ARG, "probability rows shorter than the cluster count");
    hipPointerAttribute_t attr;
    if (hipPointerGetAttributes(&attr, prob) != hipSuccess || attr.type != hipMemoryTypeDevice) {
        (void)hipGetLastError();
        return fail(CBM_E_ARG, "probabilities must be device memory");
    }
    int rc;
    if (b->swCells != cellsPerCluster && (rc = sweepSetup(b, cellsPerCluster))) return rc;
    std::vector<const uint8_t *> mf(N, mf_dev);
    std::vector<size_t> pitch(N, mfPitch);
    return run(b, mf.data(), pitch.data(), n, err_host, nullptr, prob, probPitch);
}

int cbm_batch_sweep_weights(cbm_batch *b, int net, float *host)
{
    if (!b || !host || net < 0 || net >= (int)b->nets.size()) return fail(CBM_E_ARG, "bad argument");
    if (!b->swCells) return fail(CBM_E_STATE, "no sweep has run");
    HIP_TRY(hipSetDevice(b->device));
    HIP_TRY(hipStreamSynchronize(b->stream));
    const SweepNet &w = b->hsw[net];
    HIP_TRY(hipMemcpy(host, w.W, (size_t)b->nets[net]->s.numZones * w.nCl * 4, hipMemcpyDeviceToHost));
    return CBM_OK;
}

int cbm_batch_sweep_reset(cbm_batch *b)
{
    if (!b) return fail(CBM_E_ARG, "null batch");
    if (!b->swCells) return CBM_OK;
    return sweepSetup(b, b->swCells);
}

}  // extern "C"

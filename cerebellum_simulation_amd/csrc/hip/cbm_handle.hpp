// cbm_handle.hpp -- the network handle behind the cbm_* C-ABI, shared by the
// single-network runtime (cbm_runtime.hip) and the batch (cbm_batch.hip).
// Not part of the public interface.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "cbm_dev.hpp"

namespace cbmdev {
// FMT: 1 = compact 16-byte connectivity records, 2 = the five arrays
template <int CPT, int NT, int FMT>
__global__ void gr_step_kernel(const GrArgs a, const GrStep st);
template <int CPT, int FMT>
__global__ void step_kernel(const SmallArgs *args, const SmallStep st, const GrArgs g, const GrStep gs);
template <int CPT, int FMT>
__global__ void step_span_kernel(const SmallArgs *args, const SmallStep st, const GrArgs g, const GrStep gs,
                                 const uint32_t spans);
template <int CPT, int FMT>
__global__ void gr_span_kernel(const GrArgs a, const GrStep st, const uint32_t spans);
__global__ void small_step_kernel(const SmallArgs *args, const SmallStep st);
template <int CPT, bool LOOP, bool BIGNC>  // LOOP: a closed loop (cbm_loop_begin); SPILL: some small-population fields left in global memory
__global__ void step_resident_kernel(const SmallArgs *args, const ResRun r);
__global__ void stream_ceiling_kernel(StreamArgs a);
__global__ void pfpc_fold_kernel(const float *chunk, float *out, uint32_t nChunks, uint32_t perPC, uint32_t numPC,
                                 uint32_t tasks);

// records the message returned by cbm_last_error() and returns code
int fail(int code, const std::string &m);
}  // namespace cbmdev

#define HIP_TRY(expr)                                                                        \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess)                                                                \
            return cbmdev::fail(CBM_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

using namespace cbmdev;  // private header: the runtime's own names

struct cbm_handle {
    cbm_batch *batch = nullptr;      // the batch the network belongs to, if any
    cbm_sizes s{};
    cbm_act_params p{};
    cbm_options o{};
    hipStream_t stream = nullptr;   // GR kernel, host copies; joined with stream2 after every call
    hipStream_t stream2 = nullptr;  // small-population kernel
    hipEvent_t evG = nullptr;       // recorded on stream after each GR kernel (and host copies)
    hipEvent_t evS = nullptr;       // recorded on stream2 after each small kernel
    bool committed = false;
    int cpt = 4, nt = 256;   // GR kernel shape of the two-stream step
    int fusedCpt = 1;        // GR cells per thread of the fused step
    int mode = 0;            // 1, 3, 4 fused, 2 two streams (decided at commit)
    uint32_t spans = 0;      // fused step: GR blocks of the even split (0: power-of-two tiles)
    int spanCpt = 0;         // GR cells per thread of those blocks
    size_t fusedLds = 0;     // dynamic LDS of the fused step kernel
    bool fused() const { return mode != 2; }
    // resident runs (cbm_run_mf_train of two or more steps): one launch of
    // spans + 3 workgroups, one per CU, GR state in registers, step counters
    // and flags between steps
    bool resident = false;
    bool resSpill = false;            // some small-population fields stay in global memory
    size_t residentLds = 0;           // its dynamic LDS (the small step's, or a span's + its records)
    uint32_t residentRecLds = 0;      // byte offset of the GR records in that LDS
    uint32_t *resBar = nullptr;       // GR step counters [kBarGroups * kBarStride], then the step-count
                                      // copies [kFlagCopies * kBarStride] of the microzone, input-layer and
                                      // generator blocks
    uint32_t resBarBase = 0;          // the counters' sum after the last resident launch
    uint32_t resFlagBase = 0;         // the small block's step count after it
    uint32_t resPbBase = 0;           // the generator block's step count after it
    uint32_t resFlagInBase = 0;       // the input-layer block's step count after it
    uint32_t resSpans = 0;            // GR blocks of a resident launch
    int resCpt = 0;                   // their cells per thread
    uint8_t *resApSC3 = nullptr;      // SC spikes by step mod 3, input-layer block -> microzone block
    uint32_t *resMfFlag = nullptr;    // the MF-row step count the blocks poll: zero outside closed loops
    // a closed loop (cbm_loop_begin .. cbm_loop_end): one resident launch
    // that takes each step's MF row and error drive from pinned memory and
    // returns the step's spikes of loopPop there
    bool looping = false;
    int loopK = 0, loopMax = 0, loopPop = 0;
    uint8_t *loopHost = nullptr;      // pinned, coherent: step counts, [2] input rows, [2] output rows
    size_t loopHostCap = 0;
    uint32_t loopInPitch = 0, loopInErr = 0, loopOutOff = 0, loopOutWords = 0, loopOutBytes = 0;
    uint8_t *loopMF = nullptr;        // device: [loopCap][loopPitch] MF rows
    float *loopErr = nullptr;         // device: [loopCap][numZones] error drive
    uint32_t *loopFlag = nullptr;     // device: the MF-row step count of the loop's launch
    size_t loopCap = 0, loopPitch = 0;
    uint32_t *resStatus = nullptr;    // device: 1 when a barrier wait timed out
    uint32_t *resStatusHost = nullptr;  // pinned copy, read at the next synchronisation
    bool resCheck = false;            // a resident run's status is still to be read
    float *resErr = nullptr;          // device copy of a run's error drive
    size_t resErrCap = 0;             // its capacity in floats

    std::vector<uint8_t> host[CBM_F_COUNT];  // staged fields, zones stacked
    bool written[CBM_F_COUNT] = {};
    void *dfield[CBM_F_COUNT] = {};

    uint2 *mfIn = nullptr, *goIn = nullptr, *goOut = nullptr;
    uint4 *conn = nullptr;     // compact GR connectivity (nullptr: the five arrays)
    uint32_t *goOutDly = nullptr;
    uint8_t *bspDly = nullptr;
    float *chunk2 = nullptr, *pfpcRing = nullptr;  // chunk sums [2][nZ][numGR/64] by step parity
    uint32_t *bcRing = nullptr, *scRing = nullptr, *goRing = nullptr, *rngState = nullptr, *rngNext = nullptr;
    uint8_t *apGO2 = nullptr;   // [2][numGO]
    int32_t *timer2 = nullptr;  // [2][nZ][numIO]
    bool skipFold = true;       // the ring slot of the last step holds host-written PF->PC sums
    uint8_t *dmf = nullptr;
    uint8_t *pinnedMF = nullptr;
    // cbm_step_io_begin/_end: the small step stores one population's spikes
    // into pinnedIO (coherent pinned memory); ioOut is set for that one launch
    uint8_t *pinnedIO = nullptr;
    uint8_t *ioOut = nullptr;
    uint32_t ioPop = 0;
    uint32_t ioBytes = 0;
    bool ioPending = false;
    uint32_t anyCoupling = 0;
    SmallArgs *smallArgs = nullptr;  // device copy of the constant small-kernel arguments
    SmallArgs smallHost{};           // host copy of the same (re-uploaded when the recorder changes)
    uint8_t *rec = nullptr;          // spike recorder ring (owned; freed on change and destroy)
    uint64_t recFrom = 0;            // first step written to the current recorder
    uint32_t recOffPF = 0xFFFFFFFFu; // offset of the PF chunk counts in a recorder row
    std::vector<uint8_t> recStage;   // host copy of recorder rows being read
    size_t smallLds = 0;             // dynamic LDS of the small kernel

    uint64_t curTime = 0;
    float errDrive[kMaxZones] = {};
    float ltd[kMaxZones] = {}, ltp[kMaxZones] = {};
    float gIncGRtoGO = 0;

    std::vector<void *> allocs;
    size_t devBytes = 0;
    hipEvent_t ev[4] = {};  // timing events of cbm_profile_steps / cbm_profile_split

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


namespace cbmdev {
// wait for both streams of the handle
int syncAll(cbm_handle *h);
// the step-independent part of the GR kernel arguments
void grConstArgs(const cbm_handle *h, GrArgs *g);
// device MF train checks shared by cbm_run_mf_train and the batch
int checkTrain(cbm_handle *h, const uint8_t *mf_dev, size_t pitch, int n);
}  // namespace cbmdev

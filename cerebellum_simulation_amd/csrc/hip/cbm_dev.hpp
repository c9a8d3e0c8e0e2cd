// cbm_dev.hpp -- device-side data layout and kernel argument blocks of the
// MI355X network step.  Shared by the kernels (cbm_kernels.hip) and the host
// runtime (cbm_runtime.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#include "../../../include/cbm_hip.h"

namespace cbmdev {

constexpr int kMaxZones = 8;
constexpr int kMaxIO = 64;
constexpr uint16_t kNoCell = 0xFFFFu;   // unused connectivity slot
constexpr uint8_t kNever = 0xFFu;       // delay mask 0: never delivers
constexpr int kRing = 3;                // input-sum ring slots (t, t-1, t-2)
constexpr int kSmallThreads = 1024;     // threads of the small-population kernel

// Granule-cell kernel arguments.  Per-GR arrays are structure-of-arrays of
// numGR elements; connectivity is packed four 16-bit cell indices per GR
// (kNoCell = empty slot) and one 8-bit delay (bit index of the spike register)
// per output slot.
struct GrArgs {
    float *v, *thr, *gKCa, *gE, *gI;
    uint32_t *apBuf;
    uint64_t *hist;
    float *w;                  // [numZones][numGR] PF->PC weights
    const uint2 *mfIn;         // 4 x u16 MF inputs
    const uint2 *goIn;         // 4 x u16 GO inputs
    const uint2 *goOut;        // 4 x u16 GO targets
    const uint32_t *goOutDly;  // 4 x u8 delay bit per target
    const uint8_t *bspDly;     // delay bit towards BC/PC/SC
    const uint8_t *mf;         // raw MF bytes of this step
    const uint8_t *apGOprev;   // GO spikes of the previous step
    float *chunk;              // [numZones][numGR/64] PF->PC sums of 64-GR chunks
    uint32_t *bcCnt, *scCnt, *goCnt;  // this step's ring slot (zeroed beforehand)
    const int32_t *plastTimer;        // [numZones][numIO]
    uint32_t numGR, numMF, numGO, numZones, numIO;
    uint32_t log2Lbc, log2Lsc, log2GRperIO;
    uint32_t doPlast;          // t % tsPerHistBinGR == 0: plasticity + history this step
    float gLeak, eLeak, eGO, threshRest, threshMax, threshDec;
    float decMF, incMF, decGO, incGO;
    int32_t ltdEnd, ltpStart;  // tsLTDStartAPIO + tsLTDDurationIO, tsLTPStartAPIO
    float ltd[kMaxZones], ltp[kMaxZones];
};

// CSR list: row r covers idx[ptr[r] .. ptr[r+1])
struct Csr {
    const uint32_t *ptr;
    const uint32_t *idx;
};

// Small-population kernel arguments.  Field storage is the CBMState layout;
// per-zone fields hold numZones consecutive copies.
struct SmallArgs {
    void *f[CBM_F_COUNT];      // device storage of every state field
    const float *chunk;        // [numZones][numGR/64]
    float *pfpcRing;           // [kRing][numZones][numPC]
    uint32_t *bcRing, *scRing, *goRing;  // [kRing][n]
    uint32_t *rngState;        // [numZones][353]
    Csr mfToGoIn;              // per GO: MF sources (any order)
    Csr gabaIn;                // per GO: GABA sources in the reference's accumulation order
    Csr bcToPcIn;              // per PC: BC sources
    Csr pcToBcIn;              // per BC: PC sources (with multiplicity)
    const int32_t *coupN;      // numpGOCoupInGOGO
    const uint32_t *coupP;     // pGOCoupInGOGO [numGO][maxCoup]
    cbm_sizes s;
    cbm_act_params p;
    uint32_t grgoLag;
    uint32_t anyCoupling;      // some GO has a coupling input and ratio != 0
};

// Per-step values of the small-population kernel (passed by value; the
// constant SmallArgs block lives in device memory).
struct SmallStep {
    const uint8_t *mf;         // raw MF bytes of this step
    uint64_t t;                // this step (curTime after the increment)
    float errDrive[kMaxZones];
    float gIncGRtoGO;
};

__host__ __device__ inline uint32_t ilog2(uint32_t x)
{
    uint32_t r = 0;
    while ((1u << r) < x) r++;
    return r;
}

}  // namespace cbmdev

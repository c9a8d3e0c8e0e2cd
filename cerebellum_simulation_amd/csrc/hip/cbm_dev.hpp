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
// Zone generator state on the device, per zone: the 352 words of the current
// SFMT block, the index of the next unused word and the block's generation
// (blocks regenerated since the host last set the state); and the prebuilt
// next block, 352 words and the generation + 1 it holds (a tag that matches
// nothing after a host write).  Both are double-buffered by step parity.
constexpr int kRngZone = 354, kRngNext = 353;
constexpr int kSmallThreads = 1024;     // threads of the small-population kernel
// The state fields the small-population step reads and writes, which a
// resident run keeps in LDS from the first step to the last (when they fit
// beside the step's scratch LDS): first the input layer's (GO, SC, the MF
// spike buffer), held by the block that steps them, then the microzones'.
constexpr int kSmallFieldCount = 64, kSmallInFields = 23;
constexpr int kSmallFields[kSmallFieldCount] = {
    // input layer
    CBM_F_APBUF_GO, CBM_F_APBUF_MF, CBM_F_APBUF_SC, CBM_F_AP_SC,
    CBM_F_GLU_GO, CBM_F_GO_GABA_OUT_SYNSCALE_GOGO, CBM_F_G_GO_GO, CBM_F_G_GR_GO,
    CBM_F_G_MF_GO, CBM_F_G_MGLURINC_GO, CBM_F_G_MGLUR_GO, CBM_F_G_PF_SC,
    CBM_F_INPUT_GOGABA_SYNDEP_GO, CBM_F_INPUT_GO_GO, CBM_F_INPUT_MF_GO, CBM_F_MGLUR_GO,
    CBM_F_SUM_GO_INPUT_GO, CBM_F_SUM_GR_INPUT_GO, CBM_F_THRESH_GO, CBM_F_THRESH_SC,
    CBM_F_VCOUPLE_GO, CBM_F_VM_GO, CBM_F_VM_SC,
    // microzones
    CBM_F_APBUF_BC, CBM_F_APBUF_IO, CBM_F_APBUF_NC, CBM_F_APBUF_PC,
    CBM_F_AP_BC, CBM_F_AP_IO, CBM_F_AP_NC, CBM_F_AP_PC,
    CBM_F_GMAX_AMPA_MF_NC, CBM_F_GMAX_NMDA_MF_NC, CBM_F_G_BC_PC, CBM_F_G_MF_AMPA_NC,
    CBM_F_G_MF_NMDA_NC, CBM_F_G_NC_IO, CBM_F_G_PC_BC, CBM_F_G_PC_NC,
    CBM_F_G_PC_SCALE_NC, CBM_F_G_PF_BC, CBM_F_G_PF_PC, CBM_F_G_SC_PC,
    CBM_F_HIST_MF, CBM_F_HIST_PC_POPACT, CBM_F_INPUT_BC_PC, CBM_F_INPUT_MF_NC,
    CBM_F_INPUT_NC_IO, CBM_F_INPUT_PC_BC, CBM_F_INPUT_PC_NC, CBM_F_INPUT_SC_PC,
    CBM_F_MFNC_NOLTP_NOLTD, CBM_F_MF_SYNW_NC, CBM_F_PC_POPACT_CTRS, CBM_F_SYN_IO_PRELEASE_NC,
    CBM_F_THRESH_BC, CBM_F_THRESH_IO, CBM_F_THRESH_NC, CBM_F_THRESH_PC,
    CBM_F_VCOUPLE_IO, CBM_F_VM_BC, CBM_F_VM_IO, CBM_F_VM_NC,
    CBM_F_VM_PC};
constexpr uint32_t kNotInLds = 0xFFFFFFFFu;
constexpr uint32_t kSmallLdsMax = 152 * 1024;  // dynamic LDS cap of the small kernel

// Granule-cell kernel arguments that do not change from step to step.
// Per-GR arrays are structure-of-arrays of numGR elements; connectivity is
// packed four 16-bit cell indices per GR (kNoCell = empty slot) and one 8-bit
// delay (bit index of the spike register) per output slot.
// The GR history (CBM_F_HIST_GR, one uint64_t per GR in the reference's
// layout: bit 0 the latest 5-step bin, shifted left at every update) lives on
// the device as a ring of 64 bit-planes, plane-major: plane q holds one bit
// per GR, word (q * numGR / 64 + GR / 64), bit GR % 64.  Update c (at step
// t = c * tsPerHistBinGR) reads bit 39 of the old word from plane (c - 40)
// mod 64 and writes its bits into plane c mod 64, the plane of the bit the
// shift drops, so an update moves two bits per GR instead of the 16 bytes of
// reading and writing the word.  At time T the word's bit k is in plane
// (T / tsPerHistBinGR - k) mod 64.
__host__ __device__ inline uint32_t histPlane(uint64_t update, int back)
{
    return (uint32_t)((update - (uint64_t)back) & 63u);
}

struct GrArgs {
    float *v, *thr, *gKCa, *gE, *gI;
    uint32_t *apBuf;
    uint64_t *hist;            // history bit-planes [64][numGR / 64]
    float *w;                  // [numZones][numGR] PF->PC weights
    const uint2 *mfIn;         // 4 x u16 MF inputs
    const uint2 *goIn;         // 4 x u16 GO inputs
    const uint2 *goOut;        // 4 x u16 GO targets
    const uint32_t *goOutDly;  // 4 x u8 delay bit per target
    const uint8_t *bspDly;     // delay bit towards BC/PC/SC
    // compact format (nullptr: the arrays above): one 16-byte record per GR,
    // see packGrConn
    const uint4 *conn;
    uint32_t numGR, numMF, numGO, numZones, numIO, numPC, perPC;  // perPC: 64-GR chunks per PC
    uint32_t log2Lbc, log2Lsc, log2GRperIO;
    float gLeak, eLeak, eGO, threshRest, threshMax, threshDec;
    float decMF, incMF, decGO, incGO;
    int32_t ltdEnd, ltpStart;  // tsLTDStartAPIO + tsLTDDurationIO, tsLTPStartAPIO
    float ltd[kMaxZones], ltp[kMaxZones];
};

// Granule-cell arguments of one step.
struct GrStep {
    const uint8_t *mf;         // raw MF bytes of this step
    const uint8_t *apGOprev;   // GO spikes of the previous step
    float *chunk;              // [numZones][numGR/64] PF->PC sums of 64-GR chunks, this step
    const float *foldSrc;      // chunk sums of the previous step (nullptr: nothing to fold)
    float *foldDst;            // their ring slot, [numZones][numPC]
    uint32_t *bcCnt, *scCnt, *goCnt;  // this step's ring slot (zeroed beforehand)
    const int32_t *plastTimer;        // [numZones][numIO] as of the previous step
    uint32_t doPlast;          // t % tsPerHistBinGR == 0: plasticity + history this step
    // GR history planes of this update (doPlast): bit 39 of each GR's history
    // word is read from histOld and this update's bits go to histNew
    const uint64_t *histOld;
    uint64_t *histNew;
    uint8_t *pfCnt;            // [numGR/64] delayed PF spikes per 64-GR chunk (nullptr: not recorded)
};

// Byte offsets of the small kernel's dynamic LDS regions (16-byte aligned),
// and element offsets of the index lists inside the staged u16 index blob.
struct SmallLds {
    uint32_t mf, hmf;              // u8 [numMF]: raw MF bytes, MF history after this step
    uint32_t goMF;                 // u32 [numGO]: MF->GO counts being scattered
    uint32_t apGO, scaleGO, vGO;   // u8 / f32 / f32 [numGO]
    uint32_t syn;                  // f32 synapse partials (offsets below, in floats)
    uint32_t synPC, synN, synA, synP, synIO;
    uint32_t apPC, apBC, apSC, apIO, apNC;  // u8 spikes of this step (zones stacked)
    uint32_t vIO, rel;             // f32 [nZ*numIO], [nZ*numNC]
    uint32_t ncSum;                // f32 [3][nZ*numNC]: the NC input sums (MF NMDA, MF AMPA, PC) of this step
    uint32_t inPF, inBC, inSC, inGO, inTm;  // this step's inputs from the other half of the step: PF->PC
                                   // sums [nZ*numPC], PF->BC/SC and GR->GO counts [numBC/numSC/numGO],
                                   // plasticity timers [nZ*numIO] (staged in phase 0)
    uint32_t rng;                  // u32 [nZ][kRngZone]: generator words, index, generation
    uint32_t rngNext;              // u32 [nZ][kRngNext]: the prebuilt next block and its tag
    uint32_t pop;                  // u32 [nZ][popStride]: 3 counters, bins, 2 flags
    uint32_t popStride;
    uint32_t idx;                  // u16 blob (16-byte copy units below)
    uint32_t iMFNC, iSCPC, iPCNC, iNCIO, iIOIO, iBCPC, iPCBC;  // u16 element offsets
    uint32_t slotsBCPC, slotsPCBC; // ELL widths of the BC->PC and PC->BC in-lists
    uint32_t idx16;                // blob length in 16-byte units
    uint32_t total;                // bytes
    // thread offsets: element i of a population (or synapse group) is handled
    // by thread (i + offset) mod kSmallThreads, so that the small populations
    // land on different waves instead of all on wave 0
    uint32_t oSynSC, oSynMF, oSynPC, oSynIO;      // SC->PC, MF->NC, PC->NC, NC->IO synapses
    uint32_t oPC, oBC, oSC, oIO, oNC;             // cell updates
    uint32_t oNCS;                                // NC input sums: three waves (MF NMDA, MF AMPA, PC)
    uint32_t oRBCPC, oRSCPC, oRPCBC, oRPCNC, oRIO; // output routing
    uint32_t l2PC, l2BC, l2IO, l2NC;              // log2 of the power-of-two populations
};

// Small-population kernel arguments.  Field storage is the CBMState layout;
// per-zone fields hold numZones consecutive copies.  Fields that both kernels
// touch within one step are double-buffered by step parity.
struct SmallArgs {
    void *f[CBM_F_COUNT];      // device storage of every state field (nullptr: see below)
    const float *chunk2;       // [2][numZones][numGR/64] chunk sums by step parity
    float *pfpcRing;           // [kRing][numZones][numPC]
    uint32_t *bcRing, *scRing, *goRing;  // [kRing][n]
    uint32_t *rngState;        // [2][numZones][kRngZone], by step parity
    uint32_t *rngNext;         // [2][numZones][kRngNext], by step parity
    uint8_t *apGO2;            // [2][numGO] GO spikes by step parity
    int32_t *timer2;           // [2][numZones][numIO] PF->PC plasticity timers by step parity
    const uint4 *idxBlob;      // u16 index lists of the microzone outputs, staged into LDS
    const uint16_t *mfgoOut;   // [numMF][mfgoPad] GO targets of each MF (kNoCell-padded)
    const uint16_t *gabaEll;   // [numGO][gabaPad] GABA sources in increasing order (kNoCell-padded)
    const int32_t *coupN;      // numpGOCoupInGOGO
    const uint32_t *coupP;     // pGOCoupInGOGO [numGO][maxCoup]
    uint32_t mfgoPad, gabaPad;  // row widths of mfgoOut and gabaEll (multiples of 8 and 4)
    SmallLds lds;
    cbm_sizes s;
    cbm_act_params p;
    uint32_t grgoLag;
    uint32_t anyCoupling;      // some GO has a coupling input and ratio != 0
    // spike recorder: row (t mod recCap) of rec holds the recorded populations'
    // spikes of step t; recOff[k] is population k's offset in the row
    // (CBM_REC_* order: PC, BC, SC, IO, NC, GO; UINT32_MAX = not recorded)
    uint8_t *rec;
    uint32_t recRow, recCap;
    uint32_t recOff[6];
    // resident run with the fields in LDS: fLds[f] is field f's byte offset
    // in the small block's dynamic LDS (kNotInLds: not resident); the copy
    // list gives each resident field's offset and bytes (all zones)
    uint32_t fLds[CBM_F_COUNT];
    uint32_t nLdsFields, nInFields;  // entries [0, nInFields) the input-layer part's, the rest the microzones'
    uint32_t ldsField[kSmallFieldCount], ldsBytes[kSmallFieldCount];
};

// Per-step values of the small-population kernel (passed by value; the
// constant SmallArgs block lives in device memory).  Slots and flags are the
// host's 64-bit step arithmetic, done once per step.
struct SmallStep {
    const uint8_t *mf;         // raw MF bytes of this step
    uint64_t t;                // this step (curTime after the increment)
    float errDrive[kMaxZones];
    float gIncGRtoGO;
    uint32_t par;              // t & 1: parity buffer written this step
    uint32_t slotPrev, slotT2, slotG, slotNext;  // ring slots of t-1, t-2, t-grgoLag, t+1
    uint32_t popHist;          // t % tsPerPopHistBinPC == 0: PC population histogram + MF->NC weights
    uint32_t resetMFHist;      // t % numTSinMFHist == 0
    // this step's spikes of population ioPop (CBM_REC_* order) are stored
    // into ioOut, pinned host memory (nullptr: none)
    uint8_t *ioOut;
    uint32_t ioPop;
    // a resident run that splits the step over two blocks: the input-layer
    // part writes this step's SC spikes to apSCout, the microzone part routes
    // the previous step's from apSCin to PC (scFromPrev; 0 in a launch's
    // first step, whose INPUT_SC_PC the step before routed)
    uint8_t *apSCout;
    const uint8_t *apSCin;
    uint32_t scFromPrev;
};

// Arguments of the streaming-ceiling kernel: nRead 4-byte arrays (the first
// nWrite written back) and one 16-byte record array, `cells` long each.
constexpr int kMaxStreamWords = 8;
struct StreamArgs {
    uint32_t *word[kMaxStreamWords];
    const uint4 *rec;
    uint64_t cells;
    int nRead, nWrite;
};

// One network of a batch: its constant arguments, the buffers the per-step
// arguments point into, and the run the batch is executing.  Kernels derive
// step t = t0 + k + 1 of run step k from these.
struct BatchNet {
    const SmallArgs *small;    // the network's device argument block
    GrArgs gr;                 // step-independent GR arguments
    float *chunk2, *pfpcRing;
    uint32_t *bcRing, *scRing, *goRing;
    uint8_t *apGO2;
    int32_t *timer2;
    const uint8_t *mf;         // MF row k at mf + k * pitch (device memory)
    uint64_t pitch;
    const float *err;          // [n][numZones] relative error drive, or nullptr
    uint64_t t0;               // curTime before the run
    float errDrive0[kMaxZones];  // drive pending from the host for the first step
    float maxExtIncVIO;
    float gIncGRtoGO;
    uint32_t skipFold0;        // the host wrote the PF->PC sums of step t0
    uint32_t grgoLag, tsPerHistBinGR, tsPerPopHistBinPC, numTSinMFHist;
};

// A resident run: n steps of one network in one launch (one block per CU), the GR cells' state held in the registers of their blocks from the
// first step to the last, step counters and flags between steps.  Step t = t0 + k + 1
// of run step k takes MF row k; its per-step values are derived on the
// device as launchStep derives them on the host.
constexpr int kBarGroups = 8;      // step counters of the GR blocks (one per group of blocks)
constexpr int kFlagCopies = 16;    // copies of a small-population or generator block's step count
constexpr int kBarStride = 1024;   // words between counters and between flag copies (4 KB: different memory
                                   // channels, so that the pollers of one do not queue behind another's)
struct ResRun {
    GrArgs g;
    const uint8_t *mf;          // MF row k at mf + k * pitch (device memory)
    uint64_t pitch;
    const float *err;           // [n][numZones] relative error drive (device), or nullptr
    uint64_t t0;                // curTime before the run
    float errDrive0[kMaxZones]; // drive pending from the host for the first step
    float maxExtIncVIO, gIncGRtoGO;
    uint32_t n;                 // steps
    uint32_t spans;             // GR blocks (1 .. spans); block spans + 1 steps the microzones, spans + 2
                                // folds the PF->PC sums and prebuilds generator blocks
    uint32_t skipFold0;         // the host wrote the PF->PC sums of step t0
    uint32_t recOffPF;          // offset of the PF chunk counts in a recorder row (UINT32_MAX: none)
    uint32_t *bar;              // [kBarGroups * kBarStride] step counters of the GR blocks
    uint32_t barBase;           // their sum before this launch
    uint32_t *flag;             // [kFlagCopies * kBarStride] the microzone part's step count, copies
    uint32_t flagBase;          // its value before this launch
    uint32_t *flagIn;           // [kFlagCopies * kBarStride] the input-layer part's step count, copies
    uint32_t flagInBase;        // its value before this launch
    uint8_t *apSC3;             // [3][numSC] SC spikes by step mod 3 (input-layer part -> microzone part)
    uint32_t *pb;               // [kFlagCopies * kBarStride] the generator block's step count, copies
    uint32_t pbBase;            // its value before this launch
    uint32_t *status;           // set to 1 when a barrier wait timed out
    uint32_t recLds;            // byte offset in dynamic LDS of the GR blocks' records [CPT][1024]
    uint64_t waitTicks;         // bound of a wait (s_memrealtime ticks)
    // a closed loop (cbm_loop_begin; hostSeq nullptr: none): the host posts
    // step k's MF row and error drive in hostIn[k & 1] and then k + 1 in
    // *hostSeq (kHostStop | k: stop before step k); the generator block copies
    // them to mf + k * pitch and errDev[k] and publishes k + 1 in mfFlag
    // (+ kStopBit to stop); the microzone part stores the step's spikes of
    // population ioPop in hostOut[k & 1] and then k + 1 in *hostOutSeq
    const uint32_t *hostSeq;
    const uint8_t *hostIn;      // [2][inPitch]: MF bytes, then numZones floats at inErr
    uint32_t inPitch, inErr;
    uint32_t *hostOutSeq;
    uint32_t *hostOut;          // [2][outWords]: the spikes, 4 to a word
    uint32_t outWords, ioPop;
    float *errDev;              // [n][numZones] error drive (device)
    uint32_t *mfFlag;           // [kFlagCopies * kBarStride] steps whose MF row is in place (0 at launch)
};
constexpr uint32_t kHostStop = 1u << 31;  // in hostSeq: stop
constexpr uint64_t kResWaitTicks = 20000000ull;    // bound of a wait between blocks: 0.2 s of s_memrealtime (100 MHz)
constexpr uint64_t kLoopWaitTicks = 200000000ull;  // in a closed loop, where the host posts each step: 2 s
constexpr uint32_t kStopBit = 1u << 30;   // in mfFlag: stop

// Per network of a surrogate sweep: cluster weights [numZones][nCl] and the
// cluster size (2^log2Cells consecutive GRs).
struct SweepNet {
    float *W;
    uint32_t nCl, log2Cells;
};

// The probability rows of the current sweep run (device memory): row k of
// prob drives run step k; Q row k is its plasticity bin probability.
struct SweepRun {
    const float *prob;
    const float *Q;
    uint64_t pitch;
};

// Compact GR connectivity: up to 4 MF inputs as 12-bit cell indices (numMF =
// empty), 3 GO inputs and 2 GO targets as 11-bit cell indices (kNoCell11 =
// empty) and the three delays as 6-bit bit indices (kNever6 = never), in two
// 64-bit halves:
//   lo: m0 m1 m2 m3 (12 bits each, from bit 0) | g0 (48..58)
//   hi: g1 g2 o0 o1 (11 bits each, from bit 0) | delay o0 (44..49) | delay o1 (50..55) | PF delay (56..61)
// Usable when numMF is at most 4095 and numGO at most 2047.  An empty input
// reads a bit of the spike bitmaps that is always zero: bit numMF of the MF
// bitmap (it holds numMF + 1 bits at least), bit 2047 of the GO bitmap (at
// least 2,048 bits).
constexpr uint32_t kNoMF12 = 0xFFFu;  // (the field mask)
constexpr uint32_t kNoCell11 = 0x7FFu;
constexpr uint32_t kNever6 = 0x3Fu;
constexpr uint32_t kMinMFBitmapWords = 64, kMinGOBitmapWords = 64;

__host__ __device__ inline uint4 packGrConn(const uint16_t m[4], const uint16_t g[3], const uint16_t o[2],
                                            uint8_t dBsp, uint8_t dO0, uint8_t dO1, uint32_t numMF)
{
    auto mx = [numMF](uint16_t v) -> uint64_t { return v == kNoCell ? numMF : v; };
    auto ix = [](uint16_t v) -> uint64_t { return v == kNoCell ? kNoCell11 : v; };
    auto dl = [](uint8_t v) -> uint64_t { return v == kNever ? kNever6 : v; };
    const uint64_t lo = mx(m[0]) | mx(m[1]) << 12 | mx(m[2]) << 24 | mx(m[3]) << 36 | ix(g[0]) << 48;
    const uint64_t hi = ix(g[1]) | ix(g[2]) << 11 | ix(o[0]) << 22 | ix(o[1]) << 33 | dl(dO0) << 44 | dl(dO1) << 50 |
                        dl(dBsp) << 56;
    return make_uint4((uint32_t)lo, (uint32_t)(lo >> 32), (uint32_t)hi, (uint32_t)(hi >> 32));
}

// Bytes of dynamic LDS a GR tile or span of nt threads x cpt cells uses.
__host__ __device__ constexpr uint32_t grTileLds(uint32_t nt, uint32_t cpt, uint32_t numGO, uint32_t numMF,
                                                 uint32_t numZones, uint32_t numIO)
{
    return numGO * 4 + 2 * (nt * cpt / 64 + 1) * 4 + numZones * numIO * 4 +
           ((numMF >> 5) + 1 > kMinMFBitmapWords ? (numMF >> 5) + 1 : kMinMFBitmapWords) * 4 +
           ((numGO >> 5) + 1 > kMinGOBitmapWords ? (numGO >> 5) + 1 : kMinGOBitmapWords) * 4;
}

__host__ __device__ inline uint32_t ilog2(uint32_t x)
{
    uint32_t r = 0;
    while ((1u << r) < x) r++;
    return r;
}

}  // namespace cbmdev

// cbm_kernels.hip -- CDNA4 (gfx950) kernels of one 1 ms cerebellar network step.
//
// Two kernels per step, launched back to back on one stream:
//
//  gr_step_kernel     every granule cell (GR), fused: the membrane/threshold/AHP
//                     update (kernels.cu:79-110), the MF and GO input sums
//                     (kernels.cu:169-199), PF->PC plasticity gated by history
//                     bit 39 (kernels.cu:282-291, mzone.cpp:643-724), the history
//                     shift (kernels.cu:241-246), and this step's PF->BC/SC/PC and
//                     GR->GO output sums (kernels.cu:112-167,247-376).  One pass
//                     over the per-GR arrays; spike vectors staged in LDS; PF sums
//                     are wave ballots/popcounts and a fixed-order 64-lane
//                     butterfly; the GR->GO histogram lives in LDS per block.
//
//  small_step_kernel  one workgroup: the small populations in the reference's
//                     order -- PC/BC/SC (mzone.cpp:187-256, innet.cpp:892-911),
//                     GO (innet.cpp:816-980), IO/NC and their outputs, the
//                     NC->IO release draws from the zone's SFMT stream and
//                     MF->NC plasticity (mzone.cpp:258-605) -- plus the second
//                     level of the PF->PC sum.
//
// Floating-point expressions keep the reference's operand order and types and
// the file is built with -ffp-contract=off, so binary32 results round where the
// reference's source rounds.
#include "cbm_dev.hpp"

namespace cbmdev {

// ---------------------------------------------------------------- helpers
__device__ __forceinline__ float wave_sum_butterfly(float x)
{
    // xor butterfly over 64 lanes; every lane ends with the same value, equal
    // to the halving tree x[i]+=x[i+k] for k=32,16,...,1 evaluated at lane 0
#pragma unroll
    for (int k = 32; k >= 1; k >>= 1) x = x + __shfl_xor(x, k, 64);
    return x;
}

__device__ __forceinline__ uint32_t lo16(uint32_t w) { return w & 0xFFFFu; }
__device__ __forceinline__ uint32_t hi16(uint32_t w) { return w >> 16; }

// stage a byte vector as 0/1 flags in LDS (n multiple of 4)
__device__ __forceinline__ void stage_flags(uint8_t *dst, const uint8_t *src, uint32_t n, int tid,
                                            int nthr)
{
    const uint32_t *s32 = reinterpret_cast<const uint32_t *>(src);
    uint32_t *d32 = reinterpret_cast<uint32_t *>(dst);
    for (uint32_t i = tid; i < n / 4; i += nthr) {
        uint32_t w = s32[i];
        uint32_t o = ((w & 0x000000FFu) != 0) | (((w & 0x0000FF00u) != 0) << 8) |
                     (((w & 0x00FF0000u) != 0) << 16) | (((w & 0xFF000000u) != 0) << 24);
        d32[i] = o;
    }
}

// ---------------------------------------------------------------- GR kernel
constexpr int kGrThreads = 256;

template <int CPT>
__global__ void __launch_bounds__(kGrThreads) gr_step_kernel(const GrArgs a)
{
    extern __shared__ uint32_t smem[];
    constexpr int kSeg = 4 * CPT;  // at most this many 64-GR segments per tile
    uint32_t *s_goh = smem;                          // [numGO] GR->GO counts of the tile
    uint32_t *s_sc = s_goh + a.numGO;                // [kSeg]
    uint32_t *s_bc = s_sc + kSeg;                    // [kSeg]
    float *s_step = reinterpret_cast<float *>(s_bc + kSeg);  // [numZones*numIO]
    uint8_t *s_mf = reinterpret_cast<uint8_t *>(s_step + a.numZones * a.numIO);  // [numMF]
    uint8_t *s_go = s_mf + a.numMF;                  // [numGO]

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    for (uint32_t i = tid; i < a.numGO; i += kGrThreads) s_goh[i] = 0;
    if (tid < 2 * kSeg) s_sc[tid] = 0;  // s_sc and s_bc are contiguous
    if (a.doPlast) {
        for (uint32_t i = tid; i < a.numZones * a.numIO; i += kGrThreads) {
            const uint32_t z = i / a.numIO;
            const int32_t tm = a.plastTimer[i];
            // MZone::runPFPCPlastCUDA step choice (mzone.cpp:660-674)
            s_step[i] = tm < a.ltdEnd ? a.ltd[z] : (tm > a.ltpStart ? a.ltp[z] : 0.0f);
        }
    }
    stage_flags(s_mf, a.mf, a.numMF, tid, kGrThreads);
    stage_flags(s_go, a.apGOprev, a.numGO, tid, kGrThreads);
    __syncthreads();

    const uint32_t tileBase = blockIdx.x * (kGrThreads * CPT);
    const uint32_t firstSc = tileBase >> a.log2Lsc, firstBc = tileBase >> a.log2Lbc;
    const uint32_t nChunks = a.numGR >> 6;

#pragma unroll
    for (int j = 0; j < CPT; j++) {
        const uint32_t gr = tileBase + j * kGrThreads + tid;
        float v = a.v[gr];
        float gk = a.gKCa[gr];
        float th = a.thr[gr];
        float gE = a.gE[gr];
        float gI = a.gI[gr];
        uint32_t ab = a.apBuf[gr];
        const uint2 mi = a.mfIn[gr];
        const uint2 gi = a.goIn[gr];
        const uint2 go = a.goOut[gr];
        const uint32_t gd = a.goOutDly[gr];
        const uint32_t bsp = a.bspDly[gr];

        // calcActivityGRGPU (kernels.cu:92-109)
        v = v + (a.gLeak + gk * gk * gk * gk) * (a.eLeak - v) - (gE)*v + (gI) * (a.eGO - v);
        th = th + (a.threshRest - th) * a.threshDec;
        const uint32_t ap = v > th;
        const float thN = ap * a.threshMax + (!ap) * th;
        gk = gk * 0.9999f;
        const float gkN = ap * (gk + 0.003f) + (!ap) * gk;
        ab = (ab << 1) | (ap);

        // updateGRInOPGPU for MF then GO inputs (kernels.cu:185-198)
        int cm = 0, cg = 0;
        {
            const uint32_t m0 = lo16(mi.x), m1 = hi16(mi.x), m2 = lo16(mi.y), m3 = hi16(mi.y);
            if (m0 != kNoCell) cm += s_mf[m0];
            if (m1 != kNoCell) cm += s_mf[m1];
            if (m2 != kNoCell) cm += s_mf[m2];
            if (m3 != kNoCell) cm += s_mf[m3];
            const uint32_t g0 = lo16(gi.x), g1 = hi16(gi.x), g2 = lo16(gi.y), g3 = hi16(gi.y);
            if (g0 != kNoCell) cg += s_go[g0];
            if (g1 != kNoCell) cg += s_go[g1];
            if (g2 != kNoCell) cg += s_go[g2];
            if (g3 != kNoCell) cg += s_go[g3];
        }
        gE = gE * a.decMF + a.incMF * (cm);
        gI = gI * a.decGO + a.incGO * (cg);

        a.v[gr] = v;
        a.thr[gr] = thN;
        a.gKCa[gr] = gkN;
        a.apBuf[gr] = ab;
        a.gE[gr] = gE;
        a.gI[gr] = gI;

        uint32_t bit39 = 0;
        if (a.doPlast) {
            const uint64_t h = a.hist[gr];
            bit39 = (uint32_t)(h >> 39) & 1u;
            // updateGRHistory (kernels.cu:244-245): mask 0x1F is hard-coded there
            a.hist[gr] = (h << 1) | (uint64_t)(((ab & 0x0000001Fu) > 0) * 0x00000001u);
        }

        // delayed spike towards BC/PC/SC (kernels.cu:256,269)
        const uint32_t d = (bsp == kNever) ? 0u : ((ab >> bsp) & 1u);
        const uint64_t bal = __ballot(d);

        // PF->PC: plasticity, then this step's 64-GR chunk sum, per zone
        const uint32_t io = gr >> a.log2GRperIO;
        for (uint32_t z = 0; z < a.numZones; z++) {
            float* wz = a.w + (size_t)z * a.numGR;
            float wv = 0.0f;
            if (a.doPlast || bal != 0) wv = wz[gr];
            if (a.doPlast) {
                const float plastStep = s_step[z * a.numIO + io];
                wv = wv + (bit39 > 0) * plastStep;
                wv = (wv > 0) * wv;
                wv = (wv > 1) + (wv <= 1) * wv;
                wz[gr] = wv;
            }
            float x = 0.0f;
            if (bal != 0) x = wave_sum_butterfly(wv * d);
            if (lane == 0) a.chunk[(size_t)z * nChunks + (gr >> 6)] = x;
        }

        // PF->BC and PF->SC spike counts of this wave's 64 GRs
        if (lane == 0 && bal != 0) {
            const uint32_t c = (uint32_t)__popcll(bal);
            atomicAdd(&s_sc[(gr >> a.log2Lsc) - firstSc], c);
            atomicAdd(&s_bc[(gr >> a.log2Lbc) - firstBc], c);
        }

        // GR->GO outputs (kernels.cu:133-144)
        if (ab != 0) {
            const uint32_t t0 = lo16(go.x), t1 = hi16(go.x), t2 = lo16(go.y), t3 = hi16(go.y);
            const uint32_t d0 = gd & 0xFFu, d1 = (gd >> 8) & 0xFFu, d2 = (gd >> 16) & 0xFFu, d3 = gd >> 24;
            if (t0 != kNoCell && d0 != kNever && ((ab >> d0) & 1u)) atomicAdd(&s_goh[t0], 1u);
            if (t1 != kNoCell && d1 != kNever && ((ab >> d1) & 1u)) atomicAdd(&s_goh[t1], 1u);
            if (t2 != kNoCell && d2 != kNever && ((ab >> d2) & 1u)) atomicAdd(&s_goh[t2], 1u);
            if (t3 != kNoCell && d3 != kNever && ((ab >> d3) & 1u)) atomicAdd(&s_goh[t3], 1u);
        }
    }
    __syncthreads();

    for (uint32_t i = tid; i < a.numGO; i += kGrThreads) {
        const uint32_t c = s_goh[i];
        if (c) atomicAdd(&a.goCnt[i], c);
    }
    const uint32_t tileGR = kGrThreads * CPT;
    const uint32_t nSc = (tileGR >> a.log2Lsc) ? (tileGR >> a.log2Lsc) : 1;
    const uint32_t nBc = (tileGR >> a.log2Lbc) ? (tileGR >> a.log2Lbc) : 1;
    if (tid < (int)nSc && s_sc[tid]) atomicAdd(&a.scCnt[firstSc + tid], s_sc[tid]);
    if (tid < (int)nBc && s_bc[tid]) atomicAdd(&a.bcCnt[firstBc + tid], s_bc[tid]);
}

template __global__ void gr_step_kernel<1>(const GrArgs);
template __global__ void gr_step_kernel<2>(const GrArgs);
template __global__ void gr_step_kernel<4>(const GrArgs);

// ---------------------------------------------------------------- SFMT on device
// The zone generator of MZone (CRandomSFMT0, sfmt.cpp:210-246,337-355),
// regenerated by four lanes of one wave: lane k owns 32-bit lane k of every
// 128-bit state word, and the two byte shifts of the recursion take the
// neighbouring lane's word with a shuffle.
__device__ void sfmt_regen4(uint32_t *st, int k)
{
    const uint32_t msk = k == 0 ? 0xeffff7fbu : k == 1 ? 0xffffffefu : k == 2 ? 0xdfdfbfffu : 0x7fffdbfdu;
    uint32_t r1 = st[4 * 86 + k], r2 = st[4 * 87 + k];
#pragma unroll 1
    for (int i = 0; i < 88; i++) {
        int jj = i + 68;
        if (jj >= 88) jj -= 88;
        const uint32_t av = st[4 * i + k], bv = st[4 * jj + k];
        const uint32_t aPrev = __shfl(av, k == 0 ? 0 : k - 1, 4);
        const uint32_t r1Next = __shfl(r1, k == 3 ? 3 : k + 1, 4);
        const uint32_t sa = (av << 24) | (k == 0 ? 0u : (aPrev >> 8));
        const uint32_t sc = (r1 >> 24) | (k == 3 ? 0u : (r1Next << 8));
        const uint32_t r = av ^ sa ^ ((bv >> 7) & msk) ^ sc ^ (r2 << 14);
        st[4 * i + k] = r;
        r1 = r2;
        r2 = r;
    }
}

// ---------------------------------------------------------------- small kernel
template <typename T>
__device__ __forceinline__ T *F(const SmallArgs &a, int f)
{
    return reinterpret_cast<T *>(a.f[f]);
}

__global__ void __launch_bounds__(kSmallThreads) small_step_kernel(const SmallArgs *__restrict__ args,
                                                                   const SmallStep st)
{
    const SmallArgs &a = *args;
    const cbm_sizes &s = a.s;
    const cbm_act_params &p = a.p;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    constexpr int nThr = kSmallThreads, nWave = kSmallThreads / 64;
    const uint64_t t = st.t;
    const uint32_t slotT = (uint32_t)(t % kRing), slotT2 = (uint32_t)((t + kRing - 2) % kRing);
    const uint32_t slotG = (uint32_t)((t + kRing - a.grgoLag) % kRing);
    const uint32_t nZ = s.numZones;

    __shared__ uint32_t s_rng[352];
    __shared__ int s_doLTD[kMaxZones], s_doLTP[kMaxZones];

    // LDS partials of the synapse-parallel sub-phase (offsets in floats)
    extern __shared__ float s_syn[];
    const uint32_t kSC = s.numpPCfromSCtoPC, kMF = s.numpNCfromMFtoNC, kPC = s.numpNCfromPCtoNC;
    const uint32_t kNCIO = s.numpIOfromNCtoIO;
    const uint32_t offPC = 0;
    const uint32_t offN = offPC + nZ * s.numSC;
    const uint32_t offA = offN + nZ * s.numNC * kMF;
    const uint32_t offP = offA + nZ * s.numNC * kMF;
    const uint32_t offIO = offP + nZ * s.numNC * kPC;

    // ======================= phase 1a: MF history, PF->PC fold, synapses
    // InNet::updateMFActivties (innet.cpp:800-815)
    {
        uint8_t *__restrict__ histMF = F<uint8_t>(a, CBM_F_HIST_MF);
        uint32_t *__restrict__ apBufMF = F<uint32_t>(a, CBM_F_APBUF_MF);
        for (uint32_t i = tid; i < s.numMF; i += nThr) {
            const uint8_t ap = st.mf[i] > 0;
            const uint8_t h = histMF[i];
            const uint32_t b = apBufMF[i];
            histMF[i] = h || ap;
            apBufMF[i] = (b << 1) | (ap * 0x00000001u);
        }
    }
    // second level of this step's PF->PC sums: one wave per (zone, PC)
    {
        const uint32_t nChunks = s.numGR >> 6, perPC = s.numpPCfromGRtoPC >> 6;
        for (uint32_t task = wave; task < nZ * s.numPC; task += nWave) {
            const uint32_t z = task / s.numPC, pc = task % s.numPC;
            const float *__restrict__ c = a.chunk + (size_t)z * nChunks + (size_t)pc * perPC;
            float acc = 0.0f;
#pragma unroll 8
            for (uint32_t q = lane; q < perPC; q += 64) acc = acc + c[q];
            acc = wave_sum_butterfly(acc);
            if (lane == 0) a.pfpcRing[((size_t)slotT * nZ + z) * s.numPC + pc] = acc;
        }
    }
    // SC->PC synapses of calcPCActivities (mzone.cpp:201-207)
    {
        float *__restrict__ gSC = F<float>(a, CBM_F_G_SC_PC);
        const uint8_t *__restrict__ inSC = F<uint8_t>(a, CBM_F_INPUT_SC_PC);
        for (uint32_t q = tid; q < nZ * s.numSC; q += nThr) {
            float g = gSC[q];
            const uint8_t in = inSC[q];
            g = g + p.gIncSCtoPC * (1 - g) * in;
            g = g * p.gDecSCtoPC;
            gSC[q] = g;
            s_syn[offPC + q] = g;
        }
    }
    // MF->NC and PC->NC synapses of calcNCActivities (mzone.cpp:358-383)
    {
        const uint8_t *__restrict__ inMF = F<uint8_t>(a, CBM_F_INPUT_MF_NC);
        const float *__restrict__ wMF = F<float>(a, CBM_F_MF_SYNW_NC);
        float *__restrict__ gmN = F<float>(a, CBM_F_GMAX_NMDA_MF_NC);
        float *__restrict__ gmA = F<float>(a, CBM_F_GMAX_AMPA_MF_NC);
        float *__restrict__ gN = F<float>(a, CBM_F_G_MF_NMDA_NC);
        float *__restrict__ gA = F<float>(a, CBM_F_G_MF_AMPA_NC);
        for (uint32_t q = tid; q < nZ * s.numNC * kMF; q += nThr) {
            const uint8_t in = inMF[q];
            const float w = wMF[q];
            float mN = gmN[q], mA = gmA[q], n = gN[q], aa = gA[q];
            mN = mN * p.gmaxNMDADecMFtoNC + in * w * (1 - mN);
            n = n + p.gNMDAIncMFtoNC * (mN - n);
            mA = mA * p.gmaxAMPADecMFtoNC + in * w * (1 - mA);
            aa = aa + p.gAMPAIncMFtoNC * (mA - aa);
            gmN[q] = mN;
            gN[q] = n;
            gmA[q] = mA;
            gA[q] = aa;
            s_syn[offN + q] = n;
            s_syn[offA + q] = aa;
        }
        const uint8_t *__restrict__ inPC = F<uint8_t>(a, CBM_F_INPUT_PC_NC);
        const float *__restrict__ scPC = F<float>(a, CBM_F_G_PC_SCALE_NC);
        float *__restrict__ gPC = F<float>(a, CBM_F_G_PC_NC);
        for (uint32_t q = tid; q < nZ * s.numNC * kPC; q += nThr) {
            const uint8_t in = inPC[q];
            const float sc = scPC[q];
            float g = gPC[q];
            g = g * p.gDecPCtoNC + in * sc * (1 - g);
            gPC[q] = g;
            s_syn[offP + q] = g;
        }
    }
    // NC->IO synapses of calcIOActivities (mzone.cpp:272-280)
    {
        float *__restrict__ gNC = F<float>(a, CBM_F_G_NC_IO);
        uint8_t *__restrict__ inNC = F<uint8_t>(a, CBM_F_INPUT_NC_IO);
        const float ms = p.msPerTimeStep;
        for (uint32_t q = tid; q < nZ * s.numIO * kNCIO; q += nThr) {
            float g = gNC[q];
            const uint8_t in = inNC[q];
            g = g * expf(-ms / (-p.gDecTSofNCtoIO * expf(-g / p.gDecTTofNCtoIO) + p.gDecT0ofNCtoIO));
            g = g + in * p.gIncNCtoIO * expf(-g / p.gIncTauNCtoIO);
            gNC[q] = g;
            inNC[q] = 0;
            s_syn[offIO + q] = g;
        }
    }
    __syncthreads();

    // ======================= phase 1b: cell updates
    // MZone::calcPCActivities (mzone.cpp:187-226)
    for (uint32_t idx = tid; idx < nZ * s.numPC; idx += nThr) {
        const uint32_t z = idx / s.numPC, i = idx % s.numPC;
        float *__restrict__ gPFp = F<float>(a, CBM_F_G_PF_PC);
        float *__restrict__ gBCp = F<float>(a, CBM_F_G_BC_PC);
        float *__restrict__ vPC = F<float>(a, CBM_F_VM_PC);
        float *__restrict__ thPC = F<float>(a, CBM_F_THRESH_PC);
        uint32_t *__restrict__ apBuf = F<uint32_t>(a, CBM_F_APBUF_PC);
        const float pfSum = a.pfpcRing[((size_t)slotT2 * nZ + z) * s.numPC + i];
        const uint32_t inBC = F<uint32_t>(a, CBM_F_INPUT_BC_PC)[idx];
        float gPF = gPFp[idx], gBC = gBCp[idx], v = vPC[idx], th = thPC[idx];
        const uint32_t ab = apBuf[idx];
        float gSCPCSum = 0;
        for (uint32_t j = 0; j < kSC; j++) gSCPCSum += s_syn[offPC + idx * kSC + j];
        gPF = gPF + pfSum * p.gIncGRtoPC;
        gPF = gPF * p.gDecGRtoPC;
        gBC = gBC + inBC * p.gIncBCtoPC;
        gBC = gBC * p.gDecBCtoPC;
        v = v + (p.gLeakPC * (p.eLeakPC - v)) - (gPF * v) + (gBC * (p.eBCtoPC - v)) +
            (gSCPCSum * (p.eSCtoPC - v));
        th = th + (p.threshDecPC * (p.threshRestPC - th));
        const uint8_t ap = v > th;
        gPFp[idx] = gPF;
        gBCp[idx] = gBC;
        vPC[idx] = v;
        F<uint8_t>(a, CBM_F_AP_PC)[idx] = ap;
        apBuf[idx] = (ab << 1) | (ap * 0x00000001u);
        thPC[idx] = ap * p.threshMaxPC + (!ap) * th;
        if (ap) atomicAdd(&F<uint32_t>(a, CBM_F_PC_POPACT_CTRS)[z * 3 + 2], 1u);
    }
    // MZone::calcBCActivities (mzone.cpp:228-256)
    for (uint32_t idx = tid; idx < nZ * s.numBC; idx += nThr) {
        const uint32_t i = idx % s.numBC;
        float *__restrict__ gPFp = F<float>(a, CBM_F_G_PF_BC);
        float *__restrict__ gPCp = F<float>(a, CBM_F_G_PC_BC);
        float *__restrict__ vBC = F<float>(a, CBM_F_VM_BC);
        float *__restrict__ thBC = F<float>(a, CBM_F_THRESH_BC);
        uint32_t *__restrict__ apBuf = F<uint32_t>(a, CBM_F_APBUF_BC);
        const uint32_t pf = a.bcRing[(size_t)slotT2 * s.numBC + i];
        const uint32_t inPC = F<uint32_t>(a, CBM_F_INPUT_PC_BC)[idx];
        float gPF = gPFp[idx], gPC = gPCp[idx], v = vBC[idx], th = thBC[idx];
        const uint32_t ab = apBuf[idx];
        gPF = gPF + (pf * p.gIncGRtoBC);
        gPF = gPF * p.gDecGRtoBC;
        gPC = gPC + (inPC * p.gIncPCtoBC);
        gPC = gPC * p.gDecPCtoBC;
        v = v + (p.gLeakBC * (p.eLeakBC - v)) - (gPF * v) + (gPC * (p.ePCtoBC - v));
        th = th + p.threshDecBC * (p.threshRestBC - th);
        const uint8_t ap = v > th;
        gPFp[idx] = gPF;
        gPCp[idx] = gPC;
        vBC[idx] = v;
        F<uint8_t>(a, CBM_F_AP_BC)[idx] = ap;
        apBuf[idx] = (ab << 1) | (ap * 0x00000001u);
        thBC[idx] = ap * p.threshMaxBC + (!ap) * (th);
    }
    // InNet::calcSCActivities (innet.cpp:892-911)
    for (uint32_t i = tid; i < s.numSC; i += nThr) {
        float *__restrict__ gPFp = F<float>(a, CBM_F_G_PF_SC);
        float *__restrict__ vSC = F<float>(a, CBM_F_VM_SC);
        float *__restrict__ thSC = F<float>(a, CBM_F_THRESH_SC);
        uint32_t *__restrict__ apBuf = F<uint32_t>(a, CBM_F_APBUF_SC);
        const uint32_t pf = a.scRing[(size_t)slotT2 * s.numSC + i];
        float gPF = gPFp[i], v = vSC[i], th = thSC[i];
        const uint32_t ab = apBuf[i];
        gPF = gPF + (pf * p.gIncGRtoSC);
        gPF = gPF * p.gDecGRtoSC;
        v = v + (p.gLeakSC * (p.eLeakSC - v)) - gPF * v;
        const uint8_t ap = v > th;
        th = th + p.threshDecSC * (p.threshRestSC - th);
        gPFp[i] = gPF;
        vSC[i] = v;
        F<uint8_t>(a, CBM_F_AP_SC)[i] = ap;
        apBuf[i] = (ab << 1) | (ap * 0x00000001u);
        thSC[i] = ap * p.threshMaxSC + (!ap) * (th);
    }
    // InNet::calcGOActivities (innet.cpp:827-876)
    for (uint32_t i = tid; i < s.numGO; i += nThr) {
        uint32_t *__restrict__ inMFp = F<uint32_t>(a, CBM_F_INPUT_MF_GO);
        uint32_t *__restrict__ inGOp = F<uint32_t>(a, CBM_F_INPUT_GO_GO);
        float *__restrict__ inDepp = F<float>(a, CBM_F_INPUT_GOGABA_SYNDEP_GO);
        float *__restrict__ gMFp = F<float>(a, CBM_F_G_MF_GO);
        float *__restrict__ gGRp = F<float>(a, CBM_F_G_GR_GO);
        float *__restrict__ gGOp = F<float>(a, CBM_F_G_GO_GO);
        float *__restrict__ glup = F<float>(a, CBM_F_GLU_GO);
        float *__restrict__ mGp = F<float>(a, CBM_F_MGLUR_GO);
        float *__restrict__ gIncp = F<float>(a, CBM_F_G_MGLURINC_GO);
        float *__restrict__ gMGp = F<float>(a, CBM_F_G_MGLUR_GO);
        float *__restrict__ thp = F<float>(a, CBM_F_THRESH_GO);
        float *__restrict__ vGOp = F<float>(a, CBM_F_VM_GO);
        const float *__restrict__ vCp = F<float>(a, CBM_F_VCOUPLE_GO);
        uint32_t *__restrict__ apBuf = F<uint32_t>(a, CBM_F_APBUF_GO);
        const uint32_t totalGRIn = a.goRing[(size_t)slotG * s.numGO + i];
        const uint32_t inMF = inMFp[i];
        const float inDep = inDepp[i];
        float gMF = gMFp[i], gGR = gGRp[i], gGO = gGOp[i], glu = glup[i], mG = mGp[i], gInc = gIncp[i],
              gMG = gMGp[i], thv = thp[i], v = vGOp[i];
        const float vC = vCp[i];
        const uint32_t ab = apBuf[i];
        gMF = inMF * p.gIncMFtoGO + gMF * p.gDecMFtoGO;
        gGR = totalGRIn * st.gIncGRtoGO + gGR * p.gDecGRtoGO;
        gGO = inDep * p.gGABAIncGOtoGO + gGO * p.gGABADecGOtoGO;
        glu = glu * p.gluDecayGO + totalGRIn * p.gluScaleGO * exp(-1.5 * glu);
        mG = mG * p.mGluRDecayGO + glu * p.mGluRScaleGO * expf(-mG);
        gInc = gInc * p.gMGluRIncDecayGO + mG * p.gMGluRIncScaleGO * expf(-gInc);
        gMG = gMG * p.gMGluRDecGRtoGO + gInc * p.gMGluRScaleGRtoGO;
        thv = thv + (p.threshRestGO - thv) * p.threshDecGO;
        v = v + (p.gLeakGO * (p.eLeakGO - v)) + (gGO * (p.eGABAGO - v)) - (gMF + gGR) * v + vC;
        const uint8_t ap = v > thv;
        F<uint32_t>(a, CBM_F_SUM_GR_INPUT_GO)[i] = totalGRIn;
        F<float>(a, CBM_F_SUM_GO_INPUT_GO)[i] = inDep;
        gMFp[i] = gMF;
        gGRp[i] = gGR;
        gGOp[i] = gGO;
        glup[i] = glu;
        mGp[i] = mG;
        gIncp[i] = gInc;
        gMGp[i] = gMG;
        vGOp[i] = v;
        F<uint8_t>(a, CBM_F_AP_GO)[i] = ap;
        apBuf[i] = (ab << 1) | (ap * 0x00000001u);
        thp[i] = ap * p.threshMaxGO + (!ap) * thv;
        inMFp[i] = 0;
        inGOp[i] = 0;
        inDepp[i] = 0;
    }
    // MZone::calcIOActivities (mzone.cpp:282-334)
    for (uint32_t idx = tid; idx < nZ * s.numIO; idx += nThr) {
        const uint32_t z = idx / s.numIO;
        float *__restrict__ vIO = F<float>(a, CBM_F_VM_IO);
        float *__restrict__ thIO = F<float>(a, CBM_F_THRESH_IO);
        uint32_t *__restrict__ apBuf = F<uint32_t>(a, CBM_F_APBUF_IO);
        float v = vIO[idx], th = thIO[idx];
        const float vC = F<float>(a, CBM_F_VCOUPLE_IO)[idx];
        const uint32_t ab = apBuf[idx];
        float gNCSum = 0;
        for (uint32_t j = 0; j < kNCIO; j++) gNCSum = gNCSum + s_syn[offIO + idx * kNCIO + j];
        gNCSum = gNCSum / 3.1;
        v = v + p.gLeakIO * (p.eLeakIO - v) + gNCSum * 1.5 * (p.eNCtoIO - v) + vC + st.errDrive[z];
        const uint8_t ap = v > th;
        vIO[idx] = v;
        F<uint8_t>(a, CBM_F_AP_IO)[idx] = ap;
        apBuf[idx] = (ab << 1) | (ap * 0x00000001u);
        thIO[idx] = p.threshMaxIO * ap + (!ap) * (th + p.threshDecIO * (p.threshRestIO - th));
    }
    // MZone::calcNCActivities (mzone.cpp:371-396)
    for (uint32_t idx = tid; idx < nZ * s.numNC; idx += nThr) {
        float *__restrict__ vNC = F<float>(a, CBM_F_VM_NC);
        float *__restrict__ thNC = F<float>(a, CBM_F_THRESH_NC);
        uint32_t *__restrict__ apBuf = F<uint32_t>(a, CBM_F_APBUF_NC);
        float v = vNC[idx], th = thNC[idx];
        const uint32_t ab = apBuf[idx];
        const float ms = p.msPerTimeStep;
        float gMFNMDASum = 0, gMFAMPASum = 0, gPCNCSum = 0;
        for (uint32_t j = 0; j < kMF; j++) {
            gMFNMDASum = gMFNMDASum + s_syn[offN + idx * kMF + j];
            gMFAMPASum = gMFAMPASum + s_syn[offA + idx * kMF + j];
        }
        gMFNMDASum = gMFNMDASum * ms / ((float)kMF);
        gMFAMPASum = gMFAMPASum * ms / ((float)kMF);
        gMFNMDASum = gMFNMDASum * v / (-80.0f);
        for (uint32_t j = 0; j < kPC; j++) gPCNCSum = gPCNCSum + s_syn[offP + idx * kPC + j];
        gPCNCSum = gPCNCSum * ms / ((float)kPC);
        v = v + p.gLeakNC * (p.eLeakNC - v) - (gMFNMDASum + gMFAMPASum) * v + gPCNCSum * (p.ePCtoNC - v);
        th = th + p.threshDecNC * (p.threshRestNC - th);
        const uint8_t ap = v > th;
        vNC[idx] = v;
        F<uint8_t>(a, CBM_F_AP_NC)[idx] = ap;
        apBuf[idx] = (ab << 1) | (ap * 0x00000001u);
        thNC[idx] = ap * p.threshMaxNC + (!ap) * th;
    }
    __syncthreads();

    // ======================= phase 2: output routing (reads phase-1 spikes)
    // InNet::updateMFtoGOOut (innet.cpp:912-931), as a gather over each GO's MFs
    for (uint32_t g = tid; g < s.numGO; g += nThr) {
        uint32_t c = 0;
        for (uint32_t q = a.mfToGoIn.ptr[g]; q < a.mfToGoIn.ptr[g + 1]; q++) c += st.mf[a.mfToGoIn.idx[q]] > 0;
        F<uint32_t>(a, CBM_F_INPUT_MF_GO)[g] = c;
    }
    // InNet::updateGOtoGOOut (innet.cpp:937-950): each GO gathers its GABA
    // inputs in increasing source order, which is the order the reference adds them
    {
        const uint8_t *apGO = F<uint8_t>(a, CBM_F_AP_GO);
        const float *scale = F<float>(a, CBM_F_GO_GABA_OUT_SYNSCALE_GOGO);
        for (uint32_t g = tid; g < s.numGO; g += nThr) {
            uint32_t c = 0;
            float dep = 0;
            for (uint32_t q = a.gabaIn.ptr[g]; q < a.gabaIn.ptr[g + 1]; q++) {
                const uint32_t src = a.gabaIn.idx[q];
                if (apGO[src]) {
                    c++;
                    dep += scale[src];
                }
            }
            F<uint32_t>(a, CBM_F_INPUT_GO_GO)[g] = c;
            F<float>(a, CBM_F_INPUT_GOGABA_SYNDEP_GO)[g] = dep;
        }
    }
    // GO electrical coupling (innet.cpp:961-978): vCoupleGO is a Jacobi sweep;
    // the threshold update is in place in cell order, so it runs on one lane
    {
        const float *vGO = F<float>(a, CBM_F_VM_GO);
        float *vC = F<float>(a, CBM_F_VCOUPLE_GO);
        const uint32_t kc = s.maxnumpGOCoupInGOGO;
        const float ratio = p.coupleRiRjRatioGO;
        for (uint32_t i = tid; i < s.numGO; i += nThr) {
            float acc = 0;
            for (int j = 0; j < a.coupN[i]; j++) acc = acc + ratio * (vGO[a.coupP[(size_t)i * kc + j]] - vGO[i]);
            vC[i] = acc;
        }
        if (a.anyCoupling && tid == 0) {
            float *th = F<float>(a, CBM_F_THRESH_GO);
            for (uint32_t i = 0; i < s.numGO; i++) {
                float tc = 0;
                for (int j = 0; j < a.coupN[i]; j++) tc = tc + ratio * (th[a.coupP[(size_t)i * kc + j]] - th[i]);
                th[i] = th[i] + tc;
            }
        }
    }
    for (uint32_t z = 0; z < nZ; z++) {
        // MZone::updateMFNCOut (mzone.cpp:514-524): the caller's raw MF bytes
        {
            const size_t n = (size_t)s.numNC * s.numpNCfromMFtoNC;
            const uint32_t *pMF = F<uint32_t>(a, CBM_F_P_NC_FROM_MF_TO_NC) + z * n;
            uint8_t *inMF = F<uint8_t>(a, CBM_F_INPUT_MF_NC) + z * n;
            for (size_t q = tid; q < n; q += nThr) inMF[q] = st.mf[pMF[q]];
        }
        // MZone::updateBCPCOut (mzone.cpp:449-465)
        {
            const uint8_t *apBC = F<uint8_t>(a, CBM_F_AP_BC) + (size_t)z * s.numBC;
            uint32_t *inBCPC = F<uint32_t>(a, CBM_F_INPUT_BC_PC) + (size_t)z * s.numPC;
            const uint32_t *ptr = a.bcToPcIn.ptr + (size_t)z * (s.numPC + 1);
            for (uint32_t i = tid; i < s.numPC; i += nThr) {
                uint32_t c = 0;
                for (uint32_t q = ptr[i]; q < ptr[i + 1]; q++) c += apBC[a.bcToPcIn.idx[q]] != 0;
                inBCPC[i] = c;
            }
        }
        // MZone::updateSCPCOut (mzone.cpp:467-477)
        {
            const size_t n = (size_t)s.numPC * s.numpPCfromSCtoPC;
            const uint32_t *pSC = F<uint32_t>(a, CBM_F_P_PC_FROM_SC_TO_PC) + z * n;
            uint8_t *inSC = F<uint8_t>(a, CBM_F_INPUT_SC_PC) + z * n;
            const uint8_t *apSC = F<uint8_t>(a, CBM_F_AP_SC);
            for (size_t q = tid; q < n; q += nThr) inSC[q] = apSC[pSC[q]];
        }
        // MZone::updatePCOut (mzone.cpp:409-447)
        {
            const uint8_t *apPC = F<uint8_t>(a, CBM_F_AP_PC) + (size_t)z * s.numPC;
            uint32_t *inPCBC = F<uint32_t>(a, CBM_F_INPUT_PC_BC) + (size_t)z * s.numBC;
            const uint32_t *ptr = a.pcToBcIn.ptr + (size_t)z * (s.numBC + 1);
            for (uint32_t i = tid; i < s.numBC; i += nThr) {
                uint32_t c = 0;
                for (uint32_t q = ptr[i]; q < ptr[i + 1]; q++) c += apPC[a.pcToBcIn.idx[q]];
                inPCBC[i] = c;
            }
            const size_t n = (size_t)s.numNC * s.numpNCfromPCtoNC;
            const uint32_t *pNC = F<uint32_t>(a, CBM_F_P_NC_FROM_PC_TO_NC) + z * n;
            uint8_t *inPCNC = F<uint8_t>(a, CBM_F_INPUT_PC_NC) + z * n;
            for (size_t q = tid; q < n; q += nThr) inPCNC[q] = apPC[pNC[q]];
        }
        // MZone::updateIOOut (mzone.cpp:479-493)
        {
            int32_t *timer = F<int32_t>(a, CBM_F_PFPC_PLAST_TIMER_IO) + (size_t)z * s.numIO;
            const uint8_t *apIO = F<uint8_t>(a, CBM_F_AP_IO) + (size_t)z * s.numIO;
            const float *vIO = F<float>(a, CBM_F_VM_IO) + (size_t)z * s.numIO;
            float *vCIO = F<float>(a, CBM_F_VCOUPLE_IO) + (size_t)z * s.numIO;
            const uint32_t *pIO = F<uint32_t>(a, CBM_F_P_IO_IN_IOIO) + (size_t)z * s.numIO * s.numpIOInIOIO;
            for (uint32_t i = tid; i < s.numIO; i += nThr) {
                timer[i] = (!apIO[i]) * (timer[i] + 1) + apIO[i] * p.tsLTDStartAPIO;
                float vc = 0;
                for (uint32_t j = 0; j < s.numpIOInIOIO; j++)
                    vc = vc + p.coupleRiRjRatioIO * (vIO[pIO[i * s.numpIOInIOIO + j]] - vIO[i]);
                vCIO[i] = vc;
            }
        }
        // MZone::updateNCOut release probability (mzone.cpp:497-503)
        {
            float *rel = F<float>(a, CBM_F_SYN_IO_PRELEASE_NC) + (size_t)z * s.numNC;
            const uint8_t *apNC = F<uint8_t>(a, CBM_F_AP_NC) + (size_t)z * s.numNC;
            const float ms = p.msPerTimeStep;
            for (uint32_t i = tid; i < s.numNC; i += nThr) {
                float r = rel[i];
                r = r * expf(-ms / (p.relPDecTSofNCtoIO * expf(-r / p.relPDecTTofNCtoIO) + p.relPDecT0ofNCtoIO));
                r = r + apNC[i] * p.relPIncNCtoIO * expf(-r / p.relPIncTauNCtoIO);
                rel[i] = r;
            }
        }
        // MZone::updateMFNCSyn population histogram (mzone.cpp:536-580)
        if (tid == 0) {
            s_doLTD[z] = 0;
            s_doLTP[z] = 0;
            if (t % p.tsPerPopHistBinPC == 0) {
                uint32_t *ctrs = F<uint32_t>(a, CBM_F_PC_POPACT_CTRS) + z * 3;
                uint32_t *hist = F<uint32_t>(a, CBM_F_HIST_PC_POPACT) + (size_t)z * s.numPopHistBinsPC;
                uint8_t *flags = F<uint8_t>(a, CBM_F_MFNC_NOLTP_NOLTD) + z * 2;
                ctrs[0] = ctrs[0] - hist[ctrs[1]] + ctrs[2];
                hist[ctrs[1]] = ctrs[2];
                ctrs[2] = 0;
                ctrs[1]++;
                ctrs[1] = ctrs[1] % p.numPopHistBinsPC;
                const float avg = ((float)ctrs[0]) / p.numPopHistBinsPC;
                if (avg >= p.synLTDPCPopActThreshMFtoNC && !flags[1]) {
                    s_doLTD[z] = 1;
                    flags[1] = 1;
                } else if (avg < p.synLTDPCPopActThreshMFtoNC) {
                    flags[1] = 0;
                }
                if (avg <= p.synLTPPCPopActThreshMFtoNC && !flags[0]) {
                    s_doLTP[z] = 1;
                    flags[0] = 1;
                } else if (avg > p.synLTPPCPopActThreshMFtoNC) {
                    flags[0] = 0;
                }
            }
        }
    }
    __syncthreads();

    // ======================= phase 3: scale update, draws, MF->NC weights
    {
        const uint8_t *apGO = F<uint8_t>(a, CBM_F_AP_GO);
        float *scale = F<float>(a, CBM_F_GO_GABA_OUT_SYNSCALE_GOGO);
        const float rec = p.goGABAGOGOSynRec, depF = p.goGABAGOGOSynDepF;
        for (uint32_t i = tid; i < s.numGO; i += nThr)
            scale[i] = scale[i] + (!apGO[i]) * (1 - scale[i]) * rec - apGO[i] * (scale[i] * (1 - depF));
    }
    // MZone::updateNCOut draws (mzone.cpp:505-511): one Random() per (IO, NC) in
    // row order from the zone's stream
    if (wave == 0) {
        for (uint32_t z = 0; z < nZ; z++) {
            uint32_t *g = a.rngState + (size_t)z * 353;
            for (int i = lane; i < 352; i += 64) s_rng[i] = g[i];
            uint32_t ix = g[352];
            __threadfence_block();
            if (lane < 4) {
                const float *rel = F<float>(a, CBM_F_SYN_IO_PRELEASE_NC) + (size_t)z * s.numNC;
                const uint32_t *pNC = F<uint32_t>(a, CBM_F_P_IO_FROM_NC_TO_IO) + (size_t)z * s.numIO * s.numNC;
                uint8_t *inNC = F<uint8_t>(a, CBM_F_INPUT_NC_IO) + (size_t)z * s.numIO * s.numNC;
                const uint32_t n = s.numIO * s.numpIOfromNCtoIO;
#pragma unroll 1
                for (uint32_t q = 0; q < n; q++) {
                    if (ix >= 351) {
                        __threadfence_block();
                        sfmt_regen4(s_rng, lane);
                        __threadfence_block();
                        ix = 0;
                    }
                    const uint64_t r = (uint64_t)s_rng[ix] | ((uint64_t)s_rng[ix + 1] << 32);
                    ix += 2;
                    const double u = (double)(int64_t)(r >> 12) * (1.0 / 4503599627370496.0);
                    if (lane == 0) inNC[q] = (u < rel[pNC[q]]);
                }
            }
            ix = __shfl(ix, 0, 64);
            __threadfence_block();
            for (int i = lane; i < 352; i += 64) g[i] = s_rng[i];
            if (lane == 0) g[352] = ix;
            __threadfence_block();
        }
    }
    // MZone::updateMFNCSyn weights (mzone.cpp:586-600)
    for (uint32_t z = 0; z < nZ; z++) {
        if (t % p.tsPerPopHistBinPC != 0) continue;
        const int doLTD = s_doLTD[z], doLTP = s_doLTP[z];
        const size_t n = (size_t)s.numNC * s.numpNCfromMFtoNC;
        const uint32_t *pMF = F<uint32_t>(a, CBM_F_P_NC_FROM_MF_TO_NC) + z * n;
        float *w = F<float>(a, CBM_F_MF_SYNW_NC) + z * n;
        const uint8_t *histMF = F<uint8_t>(a, CBM_F_HIST_MF);
        for (size_t q = tid; q < n; q += nThr) {
            const float delta = histMF[pMF[q]] * (doLTD * p.synLTDStepSizeMFtoNC + doLTP * p.synLTPStepSizeMFtoNC);
            float wv = w[q];
            wv = wv + delta;
            wv = (wv > 0) * wv;
            wv = (wv <= 1) * wv + (wv > 1);
            w[q] = wv;
        }
    }
    __syncthreads();

    // ======================= phase 4: MF history reset, next ring slot
    if (t % p.numTSinMFHist == 0) {
        uint8_t *histMF = F<uint8_t>(a, CBM_F_HIST_MF);
        for (uint32_t i = tid; i < s.numMF; i += nThr) histMF[i] = 0;
    }
    {
        const uint32_t nx = (uint32_t)((t + 1) % kRing);
        for (uint32_t i = tid; i < s.numBC; i += nThr) a.bcRing[(size_t)nx * s.numBC + i] = 0;
        for (uint32_t i = tid; i < s.numSC; i += nThr) a.scRing[(size_t)nx * s.numSC + i] = 0;
        for (uint32_t i = tid; i < s.numGO; i += nThr) a.goRing[(size_t)nx * s.numGO + i] = 0;
    }
}

}  // namespace cbmdev

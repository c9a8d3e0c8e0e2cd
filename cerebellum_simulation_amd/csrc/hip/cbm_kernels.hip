// cbm_kernels.hip -- CDNA4 (gfx950) kernels of one 1 ms cerebellar network step.
//
// Two kernels per step.  The small kernel of step t reads only what step t-1
// produced, so it runs on a second stream concurrently with the GR kernel of
// step t; fields both touch within one step are double-buffered by parity.
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
//                     The first tiles also fold the previous step's chunk
//                     sums into per-PC sums.
//
//  small_step_kernel  one workgroup: the small populations in the reference's
//                     order -- PC/BC/SC (mzone.cpp:187-256, innet.cpp:892-911),
//                     GO (innet.cpp:816-980), IO/NC and their outputs, the
//                     NC->IO release draws from the zone's SFMT stream and
//                     MF->NC plasticity (mzone.cpp:258-605).
//
//  step_kernel        both in one launch: block 0 is the small step, the other
//                     blocks are GR tiles (step_span_kernel: the GR cells split
//                     evenly over one block per remaining CU).

// Floating-point expressions keep the reference's operand order and types and
// the file is built with -ffp-contract=off, so binary32 results round where the
// reference's source rounds.
#include "cbm_dev.hpp"

namespace cbmdev {

// ---------------------------------------------------------------- helpers
// Device pointers read from an argument block in memory are generic; casting
// them to the global address space where they are used gives
// global_load/global_store/global_atomic, which do not share the LDS wait
// counter the way flat accesses do.
template <typename T>
using G = __attribute__((address_space(1))) T;
template <typename T>
__device__ __forceinline__ G<T> *ga(T *p)
{
    return (G<T> *)p;
}
// atomicAdd (relaxed, agent scope) through a global pointer
__device__ __forceinline__ void gadd(uint32_t *p, uint32_t v)
{
    __hip_atomic_fetch_add(ga(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Sum of a value over the 64 lanes of a wave, valid in lane 0: the halving
// tree x[i] += x[i+k] for k = 32, 16, ..., 1 (the pairs and order of an xor
// butterfly, read at lane 0), with lane i < k reading lane i + k through the
// cross-lane swaps and row shifts of the VALU instead of LDS permutes.
__device__ __forceinline__ float wave_sum_lane0(float x)
{
    x = x + __uint_as_float(__builtin_amdgcn_permlane32_swap(__float_as_uint(x), __float_as_uint(x), false, false)[1]);
    x = x + __uint_as_float(__builtin_amdgcn_permlane16_swap(__float_as_uint(x), __float_as_uint(x), false, false)[1]);
    x = x + __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x108, 0xF, 0xF, true));  // row_shl:8
    x = x + __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x104, 0xF, 0xF, true));  // row_shl:4
    x = x + __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x102, 0xF, 0xF, true));  // row_shl:2
    x = x + __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x101, 0xF, 0xF, true));  // row_shl:1
    return x;
}

__device__ __forceinline__ uint32_t lo16(uint32_t w) { return w & 0xFFFFu; }
__device__ __forceinline__ uint32_t hi16(uint32_t w) { return w >> 16; }

// Block-uniform values derived in vector arithmetic (64-bit remainders,
// pointers read from a record in memory), moved to scalar registers so that
// they do not occupy vector registers across a step.
__device__ __forceinline__ uint32_t sreg(uint32_t x) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)x); }
__device__ __forceinline__ uint64_t sreg64(uint64_t x)
{
    return ((uint64_t)sreg((uint32_t)(x >> 32)) << 32) | sreg((uint32_t)x);
}
template <typename T>
__device__ __forceinline__ T *sptr(T *p)
{
    return reinterpret_cast<T *>(sreg64(reinterpret_cast<uint64_t>(p)));
}

// ---------------------------------------------------------------- cross-block data
// Arrays that one block writes and another reads within one launch of the
// resident run (X = true): the GO spikes, plasticity timers, chunk sums, ring
// slots and generator blocks.  gfx950 keeps one L2 per XCD and these L2s are
// not coherent with each other inside a kernel, so such arrays are accessed
// with relaxed agent-scope atomics: loads and stores (sc1) that go through to
// the device's point of coherence.  Step counters and flags (res_arrive,
// res_publish, res_wait) order them between steps.  Data a block alone
// touches during the run keeps plain
// accesses and stays in its XCD's caches.  X = false: plain accesses.
template <bool X, typename T>
__device__ __forceinline__ T xld(const T *p)
{
    if constexpr (X) return __hip_atomic_load(ga(const_cast<T *>(p)), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else return *ga(p);
}
template <bool X, typename T>
__device__ __forceinline__ void xst(T *p, T v)
{
    if constexpr (X) __hip_atomic_store(ga(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else *ga(p) = v;
}

// ---------------------------------------------------------------- PF->PC fold
// Second level of the PF->PC sum of one (zone, PC), run by one full wave: lane
// l adds the 64-GR chunk sums l, l+64, l+128, ... in order, then the lanes are
// combined as by the xor butterfly (the result is valid in lane 0).
template <bool X = false>
__device__ __forceinline__ float fold_pc(const float *__restrict__ c, uint32_t perPC, int lane)
{
    float acc = 0.0f;
    if constexpr (X) {
        // eight loads in flight at clamped indices, then the adds in the same
        // order under the same bounds (an atomic load under a branch is not
        // issued ahead of it)
        for (uint32_t q0 = (uint32_t)lane; q0 < perPC; q0 += 8 * 64) {
            float x[8];
#pragma unroll
            for (int j = 0; j < 8; j++) x[j] = xld<X>(c + min(q0 + (uint32_t)j * 64u, perPC - 1u));
#pragma unroll
            for (int j = 0; j < 8; j++)
                if (q0 + (uint32_t)j * 64u < perPC) acc = acc + x[j];
        }
    } else {
#pragma unroll 8
        for (uint32_t q = lane; q < perPC; q += 64) acc = acc + c[q];
    }
    return wave_sum_lane0(acc);
}

// ---------------------------------------------------------------- GR span
// Second level of the previous step's PF->PC sums, one wave per (zone, PC),
// spread over the waves of the GR blocks (block `rank` of `ranks`): the small
// populations read them one step later.
template <int NT, bool X = false>
__device__ __forceinline__ void gr_fold(const GrArgs &a, const GrStep &st, const uint32_t rank, const uint32_t ranks)
{
    if (!st.foldSrc) return;
    uint32_t tid = threadIdx.x;
    if constexpr (X) asm volatile("" : "+v"(tid));  // as in gr_span_impl
    const int lane = tid & 63;
    const uint32_t nChunks = a.numGR >> 6, waves = ranks * (uint32_t)(NT / 64);
    for (uint32_t task = rank * (uint32_t)(NT / 64) + (tid >> 6); task < a.numZones * a.numPC;
         task += waves) {  // wave-uniform
        const uint32_t z = task / a.numPC, pc = task % a.numPC;
        const float acc = fold_pc<X>(st.foldSrc + (size_t)z * nChunks + (size_t)pc * a.perPC, a.perPC, lane);
        if (lane == 0) xst<X>(st.foldDst + task, acc);
    }
}

// The per-cell state of a block's GR cells held in registers across the steps
// of a resident run (cell j of thread tid: first + j * 1024 + tid): membrane,
// threshold, K-Ca conductance, MF and GO conductances, spike register and
// zone-0 PF->PC weight; the cells' compact connectivity records are held in
// LDS (rec[j * 1024 + tid]).
template <int CPT>
struct GrRegs {
    float v[CPT], th[CPT], gk[CPT], gE[CPT], gI[CPT], w0[CPT];
    uint32_t ab[CPT];
    const uint4 *rec;
};

// Streaming accesses of the per-GR arrays, non-temporal: the tiles' tens of
// MB pass through the L2 without evicting what the small-population block and
// the staging reads keep there (the small step's round trips then hit the
// cache instead of waiting behind the tiles' HBM traffic).
template <typename T>
__device__ __forceinline__ T ntl(const T *p)
{
    return __builtin_nontemporal_load(ga(p));
}
__device__ __forceinline__ uint2 ntl(const uint2 *p)
{
    typedef uint32_t v2u __attribute__((ext_vector_type(2)));
    const v2u r = __builtin_nontemporal_load(ga(reinterpret_cast<const v2u *>(p)));
    return make_uint2(r.x, r.y);
}
__device__ __forceinline__ uint4 ntl(const uint4 *p)
{
    typedef uint32_t v4u __attribute__((ext_vector_type(4)));
    const v4u r = __builtin_nontemporal_load(ga(reinterpret_cast<const v4u *>(p)));
    return make_uint4(r.x, r.y, r.z, r.w);
}
template <typename T>
__device__ __forceinline__ void nts(T *p, T v)
{
    __builtin_nontemporal_store(v, ga(p));
}
// NT: non-temporal or ordinary.  The compact-record build streams
// non-temporally (networks of up to ~2M GR, whose GR pass is short enough for
// the small-population block's latency to matter); the five-array build, used
// for the largest networks, keeps the ordinary policy, under which the MALL
// keeps more of their streams (4M GR: about 6 % faster per step).
template <bool NT, typename T>
__device__ __forceinline__ T gld(const T *p)
{
    if constexpr (NT) return ntl(p);
    else if constexpr (std::is_scalar<T>::value) return *ga(p);
    else return *p;  // (the vector types have no constructor from a global-space object)
}
template <bool NT, typename T>
__device__ __forceinline__ void gst(T *p, T v)
{
    if constexpr (NT) nts(p, v);
    else if constexpr (std::is_scalar<T>::value) *ga(p) = v;
    else *p = v;
}

// Element i of a per-GR array as its base plus a 32-bit byte offset: the
// addressing form with a scalar base and one vector offset, so a cell's loads
// and stores share one offset register instead of a 64-bit address each
// (byte offsets stay below 2^32 for every supported network size).
template <typename T>
__device__ __forceinline__ T *el(T *base, uint32_t i)
{
    typedef typename std::conditional<std::is_const<T>::value, const char, char>::type B;
    return reinterpret_cast<T *>(reinterpret_cast<B *>(base) + (uint32_t)(i * (uint32_t)sizeof(T)));
}

// 1 when bit idx of the LDS bitmap bm is set
__device__ __forceinline__ uint32_t bit_at(const uint32_t *bm, uint32_t idx) { return (bm[idx >> 5] >> (idx & 31u)) & 1u; }

// `count` consecutive granule cells from `first`, both multiples of 64 and
// count <= NT*CPT: thread tid owns cells first + j*NT + tid below first +
// count (whole waves are in or out).  The block is `rank` of `ranks` for the
// fold tasks.  The MF and GO spikes are staged in LDS as bitmaps (one word per
// 32 cells, one bank each, so the cells' random reads never conflict); each
// bitmap holds at least 2,048 bits and one past its cells, all zero beyond
// them, so a compact record's absent input (0x7FF) reads a zero bit as it is
// and the five-array layout's kNoCell is clamped onto bit numMF / numGO.
// Order of the loads: the spike
// bytes and the plasticity timers first, then every load of the span's cells
// in straight-line code (the connectivity format is a template argument and
// the records are unpacked after the barrier), so no load waits for an
// earlier one, the staging waits only for the spike bytes, and each cell's
// update starts as soon as its own loads have arrived.  FULL: count ==
// NT*CPT; COMPACT: one 16-byte connectivity record per GR (a.conn).
template <int CPT, int NT, bool FULL, bool COMPACT, bool PLAST, bool RES = false>
__device__ __forceinline__ void gr_span_impl(const GrArgs &a, const GrStep &st, const uint32_t first,
                                             const uint32_t count, const uint32_t rank, const uint32_t ranks,
                                             uint32_t *smem, GrRegs<CPT> *R = nullptr)
{
    constexpr int kSeg = NT * CPT / 64 + 1;  // PF blocks a span can touch (blocks hold >= 64 GRs)
    constexpr int kStage = 2048 / NT;        // spike bytes of each vector per thread held in registers
    const uint32_t nStep = a.numZones * a.numIO;
    // bitmap words: one past the cells, and at least 64 so that every 11-bit
    // index of a compact record (0x7FF when absent) reads a zero bit
    const uint32_t nMFb = max((a.numMF >> 5) + 1, kMinMFBitmapWords), nGOb = max((a.numGO >> 5) + 1, kMinGOBitmapWords);
    uint32_t *s_goh = smem;                          // [numGO] GR->GO counts of the span
    uint32_t *s_sc = s_goh + a.numGO;                // [kSeg]
    uint32_t *s_bc = s_sc + kSeg;                    // [kSeg]
    float *s_step = reinterpret_cast<float *>(s_bc + kSeg);  // [numZones*numIO]
    uint32_t *s_mfb = reinterpret_cast<uint32_t *>(s_step + nStep);  // [nMFb] MF spike bitmap
    uint32_t *s_gob = s_mfb + nMFb;                  // [nGOb] GO spike bitmap (previous step)

    int tid = threadIdx.x;
    // resident: the thread index through an opaque copy, so that the values
    // derived from it (LDS addresses, cell offsets) are computed again in each
    // step instead of being held in registers across the run's step loop
    if constexpr (RES) asm volatile("" : "+v"(tid));
    const int lane = tid & 63;
    const uint32_t wbase = (uint32_t)(tid & ~63);
    // the plasticity step: a variant of its own (PLAST), or in a resident run
    // a uniform branch of one variant (two inlined variants would hold more
    // registers than the cells' state leaves)
    const bool plast = RES ? (PLAST && st.doPlast != 0) : PLAST;
    // spike bytes, from clamped addresses (a load under a branch makes the
    // compiler wait for every load in flight before the value is used)
    uint32_t smf[kStage], sgo[kStage];
#pragma unroll
    for (int k = 0; k < kStage; k++) {
        const uint32_t i = tid + k * NT;
        smf[k] = *ga(st.mf + min(i, a.numMF - 1));
        sgo[k] = xld<RES>(st.apGOprev + min(i, a.numGO - 1));
    }
    const int32_t tm0 = xld<RES>(st.plastTimer + min((uint32_t)tid, nStep - 1));
    // MZone::runPFPCPlastCUDA step choice (mzone.cpp:660-674)
    // (the zone's step sizes picked by compares among register copies: an
    // indexed read of the argument block would be one more memory load to
    // wait for, and so would a select of its fields, which the compiler turns
    // into one)
    auto reg = [](float x) { return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(x))); };
    auto stepOf = [&](uint32_t i, int32_t tm) {
        const uint32_t z = i / a.numIO;
        float ltd = reg(a.ltd[0]), ltp = reg(a.ltp[0]);
#pragma unroll
        for (int zz = 1; zz < kMaxZones; zz++)
            if (z == (uint32_t)zz) {
                ltd = reg(a.ltd[zz]);
                ltp = reg(a.ltp[zz]);
            }
        return tm < a.ltdEnd ? ltd : (tm > a.ltpStart ? ltp : 0.0f);
    };
    // 64 spike flags of one wave into two bitmap words
    auto stage64 = [&](uint32_t *bm, uint32_t nWords, uint32_t base, bool on) {
        const uint64_t m = __ballot(on);
        if (lane == 0 && (base >> 5) < nWords) bm[base >> 5] = (uint32_t)m;
        if (lane == 0 && (base >> 5) + 1 < nWords) bm[(base >> 5) + 1] = (uint32_t)(m >> 32);
    };
    // what the registers do not hold (large networks only), before the cells'
    // loads: a loop that loads makes the compiler wait for every load in
    // flight before it
    if (plast && nStep > (uint32_t)NT)
        for (uint32_t i = tid + NT; i < nStep; i += NT) s_step[i] = stepOf(i, xld<RES>(st.plastTimer + i));
    for (uint32_t base = kStage * NT + wbase; base < 32 * nMFb; base += NT)  // wave-uniform
        stage64(s_mfb, nMFb, base, base + lane < a.numMF && *ga(st.mf + min(base + lane, a.numMF - 1)) != 0);
    for (uint32_t base = kStage * NT + wbase; base < 32 * nGOb; base += NT)
        stage64(s_gob, nGOb, base, base + lane < a.numGO && xld<RES>(st.apGOprev + min(base + lane, a.numGO - 1)) != 0);
    __builtin_amdgcn_sched_barrier(0);  // keep the loads above ahead of the cells' loads

    float v[CPT], gk[CPT], th[CPT], gE[CPT], gI[CPT], w0[CPT];
    uint32_t ab[CPT];
    uint4 rec[CPT];                       // COMPACT
    uint2 mi_[CPT], gi_[CPT], go_[CPT];   // !COMPACT
    uint32_t gd_[CPT], bsp_[CPT];         // !COMPACT
    uint64_t hist[CPT] = {};              // PLAST: the wave's 64 history words' bit 39, one plane word
    auto cellOf = [&](int j) -> uint32_t {
        const uint32_t o = (uint32_t)(j * NT + tid);
        return first + (FULL ? o : min(o, count - 1u));
    };
    // (unconditional loads: cells past the span read its last cell, and the
    // update below skips them; a load under a branch lets the compiler move
    // work on its value into the branch and wait for it there).  Cell j's
    // loads, history words included, are all issued before cell j+1's: loads
    // return in issue order, so cell j's update then waits only for its own
    // loads while the later cells' are still in flight.
#pragma unroll
    for (int j = 0; j < CPT; j++) {
        const uint32_t gr = cellOf(j);
        __builtin_amdgcn_sched_barrier(0);
        if constexpr (RES) {  // the block's cells are in registers (R)
            if (plast) hist[j] = *ga(st.histOld + (gr >> 6));
            continue;
        }
        v[j] = gld<COMPACT>(el(a.v, gr));
        gk[j] = gld<COMPACT>(el(a.gKCa, gr));
        th[j] = gld<COMPACT>(el(a.thr, gr));
        gE[j] = gld<COMPACT>(el(a.gE, gr));
        gI[j] = gld<COMPACT>(el(a.gI, gr));
        ab[j] = gld<COMPACT>(el(a.apBuf, gr));
        if constexpr (COMPACT) {
            rec[j] = gld<COMPACT>(el(a.conn, gr));
        } else {
            mi_[j] = gld<COMPACT>(el(a.mfIn, gr));
            gi_[j] = gld<COMPACT>(el(a.goIn, gr));
            go_[j] = gld<COMPACT>(el(a.goOut, gr));
            gd_[j] = gld<COMPACT>(el(a.goOutDly, gr));
            bsp_[j] = gld<COMPACT>(el(a.bspDly, gr));
        }
        w0[j] = gld<COMPACT>(el(a.w, gr));
        if constexpr (PLAST) hist[j] = *ga(st.histOld + (gr >> 6));  // one address per wave
    }
    __builtin_amdgcn_sched_barrier(0);

    // LDS writes of what was loaded first, before any branch or loop that
    // loads (after one, the compiler waits for every load in flight)
#pragma unroll
    for (int k = 0; k < kStage; k++) {
        const uint32_t base = k * NT + wbase, i = base + lane;
        stage64(s_mfb, nMFb, base, i < a.numMF && smf[k] != 0);
        stage64(s_gob, nGOb, base, i < a.numGO && sgo[k] != 0);
    }
    if (plast && (uint32_t)tid < nStep) s_step[tid] = stepOf(tid, tm0);
    for (uint32_t i = tid; i < a.numGO; i += NT) s_goh[i] = 0;
    if (tid < 2 * kSeg) s_sc[tid] = 0;  // s_sc and s_bc are contiguous
    __syncthreads();
    __builtin_amdgcn_sched_barrier(0);  // no cell update ahead of the barrier: it would wait for the cell's loads

    const uint32_t firstSc = first >> a.log2Lsc, firstBc = first >> a.log2Lbc;
    const uint32_t nChunks = a.numGR >> 6;

#pragma unroll
    for (int j = 0; j < CPT; j++) {
        if (!FULL && j * NT + (tid & ~63) >= count) continue;  // wave-uniform
        // resident: one cell at a time (interleaving the cells' updates would
        // hold more temporaries than the registers the cells' state leaves)
        if constexpr (RES) __builtin_amdgcn_sched_barrier(0);
        const uint32_t gr = first + j * NT + tid;
        float vv, g, t, e, i;
        uint32_t b;
        uint4 rj;
        if constexpr (RES) {
            vv = R->v[j], g = R->gk[j], t = R->th[j], e = R->gE[j], i = R->gI[j], b = R->ab[j], rj = R->rec[j * NT + tid];
        } else {
            vv = v[j], g = gk[j], t = th[j], e = gE[j], i = gI[j], b = ab[j];
            if constexpr (COMPACT) rj = rec[j];
        }
        // connectivity fields: MF and GO inputs clamped onto the zero bit of
        // the bitmaps when absent, the delay of the spike towards BC/PC/SC
        // (>= 32: never), and up to four GO targets with their delays
        uint32_t in[8], dBsp, tg[4], dg[4];
        if constexpr (COMPACT) {
            // packGrConn: 12-bit MF and 11-bit GO indices (numMF and 0x7FF
            // absent: a zero bit), 6-bit delays (0x3F never)
            const uint64_t lo = (uint64_t)rj.x | (uint64_t)rj.y << 32;
            const uint64_t hi = (uint64_t)rj.z | (uint64_t)rj.w << 32;
            auto f12 = [](uint64_t w, int sh) { return (uint32_t)(w >> sh) & kNoMF12; };
            auto f11 = [](uint64_t w, int sh) { return (uint32_t)(w >> sh) & kNoCell11; };
            auto f6 = [](uint64_t w, int sh) { return (uint32_t)(w >> sh) & kNever6; };
            in[0] = f12(lo, 0);
            in[1] = f12(lo, 12);
            in[2] = f12(lo, 24);
            in[3] = f12(lo, 36);
            in[4] = f11(lo, 48);
            in[5] = f11(hi, 0);
            in[6] = f11(hi, 11);
            in[7] = a.numGO;
            dBsp = f6(hi, 56);
            tg[0] = f11(hi, 22);
            tg[1] = f11(hi, 33);
            tg[2] = tg[3] = kNoCell11;
            dg[0] = f6(hi, 44);
            dg[1] = f6(hi, 50);
            dg[2] = dg[3] = kNever6;
        } else {
            const uint2 mi = mi_[j], gi = gi_[j], go = go_[j];
            in[0] = min(lo16(mi.x), a.numMF);
            in[1] = min(hi16(mi.x), a.numMF);
            in[2] = min(lo16(mi.y), a.numMF);
            in[3] = min(hi16(mi.y), a.numMF);
            in[4] = min(lo16(gi.x), a.numGO);
            in[5] = min(hi16(gi.x), a.numGO);
            in[6] = min(lo16(gi.y), a.numGO);
            in[7] = min(hi16(gi.y), a.numGO);
            dBsp = bsp_[j] == kNever ? 63u : bsp_[j];
            tg[0] = lo16(go.x);
            tg[1] = hi16(go.x);
            tg[2] = lo16(go.y);
            tg[3] = hi16(go.y);
            const uint32_t dl = gd_[j];
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const uint32_t x = (dl >> (8 * q)) & 0xFFu;
                dg[q] = x == kNever ? 63u : x;
            }
        }

        // calcActivityGRGPU (kernels.cu:92-109)
        vv = vv + (a.gLeak + g * g * g * g) * (a.eLeak - vv) - (e)*vv + (i) * (a.eGO - vv);
        t = t + (a.threshRest - t) * a.threshDec;
        const uint32_t ap = vv > t;
        const float thN = ap * a.threshMax + (!ap) * t;
        g = g * 0.9999f;
        const float gkN = ap * (g + 0.003f) + (!ap) * g;
        b = (b << 1) | (ap);

        // updateGRInOPGPU for MF then GO inputs (kernels.cu:185-198)
        int cm = 0, cg = 0;
        cm = (int)(bit_at(s_mfb, in[0]) + bit_at(s_mfb, in[1]) + bit_at(s_mfb, in[2]) + bit_at(s_mfb, in[3]));
        cg = (int)(bit_at(s_gob, in[4]) + bit_at(s_gob, in[5]) + bit_at(s_gob, in[6]));
        if constexpr (!COMPACT) cg += (int)bit_at(s_gob, in[7]);
        e = e * a.decMF + a.incMF * (cm);
        i = i * a.decGO + a.incGO * (cg);

        if constexpr (RES) {
            R->v[j] = vv, R->th[j] = thN, R->gk[j] = gkN, R->ab[j] = b, R->gE[j] = e, R->gI[j] = i;
        } else {
            gst<COMPACT>(el(a.v, gr), vv);
            gst<COMPACT>(el(a.thr, gr), thN);
            gst<COMPACT>(el(a.gKCa, gr), gkN);
            gst<COMPACT>(el(a.apBuf, gr), b);
            gst<COMPACT>(el(a.gE, gr), e);
            gst<COMPACT>(el(a.gI, gr), i);
        }

        uint32_t bit39 = 0;
        if (plast) {
            bit39 = (uint32_t)(hist[j] >> lane) & 1u;
            // updateGRHistory (kernels.cu:244-245): mask 0x1F is hard-coded
            // there; the new bit of every GR of the wave in one plane word
            const uint64_t nb = __ballot((b & 0x0000001Fu) > 0);
            if (lane == 0) gst<COMPACT>(st.histNew + (gr >> 6), nb);
        }

        // delayed spike towards BC/PC/SC (kernels.cu:256,269)
        const uint32_t d = (uint32_t)((uint64_t)b >> dBsp) & 1u;
        const uint64_t bal = __ballot(d);

        // PF->PC: plasticity, then this step's 64-GR chunk sum, per zone
        const uint32_t io = gr >> a.log2GRperIO;
        for (uint32_t z = 0; z < a.numZones; z++) {  // (numZones is 1 in most networks)
            float *wz = a.w + (size_t)z * a.numGR;
            float wv = z == 0 ? (RES ? R->w0[j] : w0[j]) : gld<COMPACT>(el(wz, gr));
            if (plast) {
                const float plastStep = s_step[z * a.numIO + io];
                wv = wv + (bit39 > 0) * plastStep;
                wv = (wv > 0) * wv;
                wv = (wv > 1) + (wv <= 1) * wv;
                if (RES && z == 0) R->w0[j] = wv;
                else gst<COMPACT>(el(wz, gr), wv);
            }
            float x = 0.0f;
            if (bal != 0) x = wave_sum_lane0(wv * d);
            if (lane == 0) xst<RES>(st.chunk + (size_t)z * nChunks + (gr >> 6), x);
        }

        // PF->BC and PF->SC spike counts of this wave's 64 GRs
        if (st.pfCnt && lane == 0) *ga(st.pfCnt + (gr >> 6)) = (uint8_t)__popcll(bal);
        if (lane == 0 && bal != 0) {
            const uint32_t c = (uint32_t)__popcll(bal);
            atomicAdd(&s_sc[(gr >> a.log2Lsc) - firstSc], c);
            atomicAdd(&s_bc[(gr >> a.log2Lbc) - firstBc], c);
        }

        // GR->GO outputs (kernels.cu:133-144)
        // (a target is absent when its index is numGO or more; a delay of 32 or
        // more never sees a spike)
        if (b != 0) {
#pragma unroll
            for (int q = 0; q < (COMPACT ? 2 : 4); q++)
                if (tg[q] < a.numGO && (((uint64_t)b >> dg[q]) & 1u)) atomicAdd(&s_goh[tg[q]], 1u);
        }
    }
    __syncthreads();

    for (uint32_t i = tid; i < a.numGO; i += NT) {
        const uint32_t c = s_goh[i];
        if (c) gadd(st.goCnt + i, c);
    }
    {
        // PF blocks the span touches; blocks at either end may be shared with
        // the neighbouring span, whose counts add in by the same atomics
        const uint32_t nSc = ((first + count - 1) >> a.log2Lsc) - firstSc + 1;
        const uint32_t nBc = ((first + count - 1) >> a.log2Lbc) - firstBc + 1;
        if ((uint32_t)tid < nSc && s_sc[tid]) gadd(st.scCnt + firstSc + tid, s_sc[tid]);
        if ((uint32_t)tid < nBc && s_bc[tid]) gadd(st.bcCnt + firstBc + tid, s_bc[tid]);
    }

    gr_fold<NT, RES>(a, st, rank, ranks);
}

// Connectivity format of a kernel instantiation: kFmtAny decides per launch
// from a.conn (the batch, whose networks may differ); a kernel built for one
// format holds only that variant, which keeps its scalar registers from
// spilling.
constexpr int kFmtAny = 0, kFmtCompact = 1, kFmtArrays = 2;

template <int CPT, int NT, bool FULL, int FMT>
__device__ __forceinline__ void gr_span(const GrArgs &a, const GrStep &st, const uint32_t first, const uint32_t count,
                                        const uint32_t rank, const uint32_t ranks, uint32_t *smem)
{
    if (!FULL && count == 0) {  // an empty span (more blocks than 64-GR chunks) only folds
        gr_fold<NT>(a, st, rank, ranks);
        return;
    }
    // the plasticity step (every tsPerHistBinGR-th) is a variant of its own, so
    // that its history words load in order with each cell's other arrays
    constexpr bool kCompact = FMT == kFmtCompact;
    if constexpr (FMT != kFmtAny) {
        if (st.doPlast)
            gr_span_impl<CPT, NT, FULL, kCompact, true>(a, st, first, count, rank, ranks, smem);
        else
            gr_span_impl<CPT, NT, FULL, kCompact, false>(a, st, first, count, rank, ranks, smem);
    } else if (a.conn) {
        if (st.doPlast)
            gr_span_impl<CPT, NT, FULL, true, true>(a, st, first, count, rank, ranks, smem);
        else
            gr_span_impl<CPT, NT, FULL, true, false>(a, st, first, count, rank, ranks, smem);
    } else {
        if (st.doPlast)
            gr_span_impl<CPT, NT, FULL, false, true>(a, st, first, count, rank, ranks, smem);
        else
            gr_span_impl<CPT, NT, FULL, false, false>(a, st, first, count, rank, ranks, smem);
    }
}

// One tile of NT*CPT consecutive granule cells (tile-aligned).
template <int CPT, int NT, int FMT>
__device__ __forceinline__ void gr_tile(const GrArgs &a, const GrStep &st, const uint32_t tile, uint32_t *smem)
{
    gr_span<CPT, NT, true, FMT>(a, st, tile * (uint32_t)(NT * CPT), (uint32_t)(NT * CPT), tile,
                                a.numGR / (uint32_t)(NT * CPT), smem);
}

// Span r of `spans`: chunks [r*C/spans, (r+1)*C/spans) of the C = numGR/64
// 64-GR chunks, so that every block gets the same work within one chunk.
template <int CPT, int NT, int FMT>
__device__ __forceinline__ void gr_span_of(const GrArgs &a, const GrStep &st, const uint32_t r, const uint32_t spans,
                                           uint32_t *smem)
{
    const uint64_t nC = a.numGR >> 6;
    const uint32_t c0 = (uint32_t)(r * nC / spans), c1 = (uint32_t)((r + 1) * nC / spans);
    gr_span<CPT, NT, false, FMT>(a, st, c0 << 6, (c1 - c0) << 6, r, spans, smem);
}

// GR tiles alone (the two-stream step, and GR-only timing)
template <int CPT, int NT, int FMT>
__global__ void __launch_bounds__(NT) gr_step_kernel(const GrArgs a, const GrStep st)
{
    extern __shared__ uint32_t smem[];
    gr_tile<CPT, NT, FMT>(a, st, blockIdx.x, smem);
}

#define CBM_GR_STEP(CPT, NT)                                                                    \
    template __global__ void gr_step_kernel<CPT, NT, kFmtCompact>(const GrArgs, const GrStep); \
    template __global__ void gr_step_kernel<CPT, NT, kFmtArrays>(const GrArgs, const GrStep);
CBM_GR_STEP(1, 256)
CBM_GR_STEP(2, 256)
CBM_GR_STEP(4, 256)
CBM_GR_STEP(1, 1024)
CBM_GR_STEP(2, 1024)
#undef CBM_GR_STEP

// GR spans alone (GR-only timing of the span-split fused step)
template <int CPT, int FMT>
__global__ void __launch_bounds__(1024) gr_span_kernel(const GrArgs a, const GrStep st, const uint32_t spans)
{
    extern __shared__ uint32_t smem[];
    gr_span_of<CPT, 1024, FMT>(a, st, blockIdx.x, spans, smem);
}

#define CBM_GR_SPAN(CPT)                                                                                       \
    template __global__ void gr_span_kernel<CPT, kFmtCompact>(const GrArgs, const GrStep, const uint32_t); \
    template __global__ void gr_span_kernel<CPT, kFmtArrays>(const GrArgs, const GrStep, const uint32_t);
CBM_GR_SPAN(2)
CBM_GR_SPAN(3)
CBM_GR_SPAN(5)
#undef CBM_GR_SPAN

// ---------------------------------------------------------------- SFMT on device
// The zone generator of MZone (CRandomSFMT0, sfmt.cpp:210-246,337-355),
// regenerated by the four lanes of one quad: lane k owns 32-bit lane k of
// every 128-bit state word, and the two byte shifts of the recursion take the
// neighbouring lane's word by a quad permutation of the VALU (no LDS round
// trip on the recursion's dependent chain).
__device__ void sfmt_regen4(uint32_t *st, int k)
{
    const uint32_t msk = k == 0 ? 0xeffff7fbu : k == 1 ? 0xffffffefu : k == 2 ? 0xdfdfbfffu : 0x7fffdbfdu;
    uint32_t r1 = st[4 * 86 + k], r2 = st[4 * 87 + k];
    // Eight words per batch, their inputs read before any of them is written:
    // word i reads words i and (i+68) mod 88, and for i >= 20 the latter was
    // regenerated at least 13 words earlier, i.e. in a previous batch.
#pragma unroll 1
    for (int i0 = 0; i0 < 88; i0 += 8) {
        uint32_t av[8], bv[8];
#pragma unroll
        for (int j = 0; j < 8; j++) {
            int jj = i0 + j + 68;
            if (jj >= 88) jj -= 88;
            av[j] = st[4 * (i0 + j) + k];
            bv[j] = st[4 * jj + k];
        }
#pragma unroll
        for (int j = 0; j < 8; j++) {
            // quad_perm [0,0,1,2]: lane k - 1; quad_perm [1,2,3,3]: lane k + 1
            const uint32_t aPrev = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)av[j], 0x90, 0xF, 0xF, false);
            const uint32_t r1Next = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)r1, 0xF9, 0xF, 0xF, false);
            const uint32_t sa = (av[j] << 24) | (k == 0 ? 0u : (aPrev >> 8));
            const uint32_t sc = (r1 >> 24) | (k == 3 ? 0u : (r1Next << 8));
            const uint32_t r = av[j] ^ sa ^ ((bv[j] >> 7) & msk) ^ sc ^ (r2 << 14);
            st[4 * (i0 + j) + k] = r;
            r1 = r2;
            r2 = r;
        }
    }
}

// The same fold as a stand-alone launch (four waves per block, one task per
// wave), for reading the last step's PF->PC sums between steps.
__global__ void __launch_bounds__(256) pfpc_fold_kernel(const float *__restrict__ chunk, float *__restrict__ out,
                                                        uint32_t nChunks, uint32_t perPC, uint32_t numPC,
                                                        uint32_t tasks)
{
    const uint32_t task = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (task >= tasks) return;  // wave-uniform
    const uint32_t z = task / numPC, pc = task % numPC;
    const float acc = fold_pc(chunk + (size_t)z * nChunks + (size_t)pc * perPC, perPC, lane);
    if (lane == 0) out[task] = acc;
}

// ---------------------------------------------------------------- small kernel

template <typename T>
__device__ __forceinline__ G<T> *gp(T *p)
{
    return (G<T> *)(p);
}

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));

// xld / xst for global-space pointers
template <bool X, typename T>
__device__ __forceinline__ T xldg(const G<T> *p)
{
    if constexpr (X) return __hip_atomic_load(const_cast<G<T> *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else return *p;
}
template <bool X, typename T>
__device__ __forceinline__ void xstg(G<T> *p, T v)
{
    if constexpr (X) __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else *p = v;
}

template <typename T>
__device__ __forceinline__ G<T> *F(const SmallArgs &a, int f)
{
    return (G<T> *)(a.f[f]);
}

// Field f of the small populations: in the small block's LDS during a
// resident run that keeps them there (LF 1), else in global memory (LF 0);
// LF 2: in LDS unless the run left it in global memory (fLds[f] kNotInLds)
template <int LF, typename T>
__device__ __forceinline__ auto SF(const SmallArgs &a, int f, uint8_t *lds)
{
    if constexpr (LF == 1) return reinterpret_cast<T *>(lds + a.fLds[f]);
    else if constexpr (LF == 2) {
        const uint32_t o = a.fLds[f];
        return o != kNotInLds ? reinterpret_cast<T *>(lds + o) : reinterpret_cast<T *>(a.f[f]);
    } else return F<T>(a, f);
}

template <typename T>
__device__ __forceinline__ T *at(uint8_t *base, uint32_t off)
{
    return reinterpret_cast<T *>(base + off);
}

// Sequential sum of n LDS floats in index order (16-byte reads when aligned).
__device__ __forceinline__ float lds_sum_seq(const float *x, uint32_t n, float acc)
{
    if ((n & 3u) == 0 && ((uintptr_t)x & 15u) == 0) {
        const float4 *x4 = reinterpret_cast<const float4 *>(x);
        uint32_t j = 0;
        // chunks of 32 terms: the chunk's 8 loads are in flight together and
        // each add waits only for the load of its own term
        for (; j + 8 <= n / 4; j += 8) {
            float4 v[8];
#pragma unroll
            for (int q = 0; q < 8; q++) v[q] = x4[j + q];
#pragma unroll
            for (int q = 0; q < 8; q++) {
                acc = acc + v[q].x;
                acc = acc + v[q].y;
                acc = acc + v[q].z;
                acc = acc + v[q].w;
            }
        }
#pragma unroll 4
        for (; j < n / 4; j++) {
            const float4 v = x4[j];
            acc = acc + v.x;
            acc = acc + v.y;
            acc = acc + v.z;
            acc = acc + v.w;
        }
        return acc;
    }
    for (uint32_t j = 0; j < n; j++) acc = acc + x[j];
    return acc;
}

// One workgroup steps every population other than GR.  The phases and what
// each reads:
//   0  stage MF bytes, output index lists, generator words and the PC
//      population histogram into LDS; MF history; MF->GO scatter of this
//      step's spiking MFs; per-synapse conductances (SC->PC, MF->NC, PC->NC,
//      NC->IO) into LDS
//   1  MF->NC routing; PC, BC, SC, GO, IO and NC updates (spikes to LDS);
//      PC/BC/SC read the PF sums of t-2, GO the GR->GO count of t-grgoLag
//   2  output routing from the LDS spikes; population histogram; clear the
//      ring slot of step t+1
//   3  GO->GO depression scale; NC->IO release draws (wave 0); MF->NC weights
// Within a phase every loop iteration issues all of its global loads before
// its first store, so each phase costs about one memory round trip, and each
// population or synapse group starts on its own wave, so that no wave runs
// the updates of every small population in turn.  X: inside a resident run
// (the arrays other blocks share go through xldg / xstg); LF: with the state
// fields in LDS (SF).  PART: the whole step, or one of its two parts on a
// block of its own (a resident run): the input layer (GO, SC, MF->GO and
// GO->GO) or the microzones (PC, BC, IO, NC, their synapses, outputs and
// plasticity).  The microzone part routes the SC spikes to PC at the start of
// the next step (from the input-layer part's apSCout) instead of at the end
// of this one.  SPILL: with the fields in LDS, a field the run could not fit
// there stays in global memory (fLds[f] kNotInLds; checked per field).
constexpr int kPartAll = 0, kPartIn = 1, kPartMz = 2;
template <bool X = false, bool LF = false, int PART = kPartAll, bool SPILL = false>
__device__ __forceinline__ void small_step(const SmallArgs &a, const SmallStep &st, const float *errDrive,
                                           uint8_t *lds)
{
    const cbm_sizes &s = a.s;
    const cbm_act_params &p = a.p;
    const SmallLds &L = a.lds;
    constexpr int LFM = LF ? (SPILL ? 2 : 1) : 0;  // where SF finds the fields
    int tid = threadIdx.x;
    // resident: as in gr_span_impl, values derived from the thread index are
    // computed again in each step instead of being held across the step loop
    if constexpr (X) asm volatile("" : "+v"(tid));
    const int lane = tid & 63, wave = tid >> 6;
    constexpr int nThr = kSmallThreads, nWave = kSmallThreads / 64;
    const uint32_t par = st.par, parPrev = par ^ 1u;
    const uint32_t nZ = s.numZones;
    const uint32_t kSC = s.numpPCfromSCtoPC, kMF = s.numpNCfromMFtoNC, kPC = s.numpNCfromPCtoNC;
    const uint32_t kNCIO = s.numpIOfromNCtoIO, kIOIO = s.numpIOInIOIO;
    const uint32_t bins = s.numPopHistBinsPC;
    constexpr bool IN = PART != kPartMz, MZ = PART != kPartIn;

    uint8_t *s_mf = lds + L.mf;
    uint8_t *s_hmf = lds + L.hmf;
    uint32_t *s_goMF = at<uint32_t>(lds, L.goMF);
    uint8_t *s_apGO = lds + L.apGO;
    float *s_scale = at<float>(lds, L.scaleGO);
    float *s_vGO = at<float>(lds, L.vGO);
    float *s_syn = at<float>(lds, L.syn);
    uint8_t *s_apPC = lds + L.apPC, *s_apBC = lds + L.apBC, *s_apSC = lds + L.apSC;
    uint8_t *s_apIO = lds + L.apIO, *s_apNC = lds + L.apNC;
    float *s_vIO = at<float>(lds, L.vIO), *s_rel = at<float>(lds, L.rel);
    uint32_t *s_rng = at<uint32_t>(lds, L.rng);
    uint32_t *s_pop = at<uint32_t>(lds, L.pop);
    const uint16_t *s_idx = at<uint16_t>(lds, L.idx);
    __shared__ int s_doLTD[kMaxZones], s_doLTP[kMaxZones];

    // ======================= phase 0
    if constexpr (IN)
        for (uint32_t i = tid; i < s.numGO; i += nThr) s_goMF[i] = 0;
    __syncthreads();
    // staging, InNet::updateMFActivties (innet.cpp:800-815) and
    // InNet::updateMFtoGOOut (innet.cpp:912-931): each spiking MF adds one to
    // every GO it targets; the counts are integers, so the order is free
    {
        const uint32_t nMFw = s.numMF / 4, nRng = nZ * kRngZone, nNext = nZ * kRngNext, nPop = nZ * L.popStride;
        const uint32_t nPCc = nZ * s.numPC, nIOc = nZ * s.numIO;
        const uint32_t n = max(max(max(max(nMFw, L.idx16), max(max(nRng, nNext), nPop)), s.numMF),
                               max(max(nPCc, s.numBC), max(max(s.numSC, s.numGO), nIOc)));
        const G<uint32_t> *rngIn = gp(a.rngState) + (size_t)parPrev * nRng;
        const G<uint32_t> *nextIn = gp(a.rngNext) + (size_t)parPrev * nNext;
        uint32_t *s_next = at<uint32_t>(lds, L.rngNext);
        const G<uint32_t> *mfw = (const G<uint32_t> *)st.mf;
        uint32_t *mfl = at<uint32_t>(lds, L.mf);
        u32x4 *bd = at<u32x4>(lds, L.idx);
        auto histMF = SF<LFM, uint8_t>(a, CBM_F_HIST_MF, lds);
        auto apBufMF = SF<LFM, uint32_t>(a, CBM_F_APBUF_MF, lds);
        for (uint32_t i = tid; i < n; i += nThr) {
            // every array is read at a clamped index (in bounds, no branch),
            // so all of the iteration's loads are issued before the first wait
            const uint32_t im = min(i, s.numMF - 1), ip = min(i, nPop - 1);
            const uint32_t w = mfw[min(i, nMFw - 1)];
            const u32x4 bl = ((const G<u32x4> *)a.idxBlob)[min(i, L.idx16 - 1)];
            const uint32_t r = MZ ? xldg<X>(rngIn + min(i, nRng - 1)) : 0u;
            const uint32_t rn = MZ ? xldg<X>(nextIn + min(i, nNext - 1)) : 0u;
            // the inputs phase 1 reads from the other half of the step (PC/BC/SC
            // the PF sums of t-2, GO the GR->GO count of t-grgoLag, IO the
            // plasticity timer of t-1): nothing writes them before phase 2
            const float inPF = MZ ? xldg<X>(gp(a.pfpcRing) + (size_t)st.slotT2 * nPCc + min(i, nPCc - 1)) : 0.0f;
            const uint32_t inBC = MZ ? xldg<X>(gp(a.bcRing) + (size_t)st.slotT2 * s.numBC + min(i, s.numBC - 1)) : 0u;
            const uint32_t inSC = IN ? xldg<X>(gp(a.scRing) + (size_t)st.slotT2 * s.numSC + min(i, s.numSC - 1)) : 0u;
            const uint32_t inGO = IN ? xldg<X>(gp(a.goRing) + (size_t)st.slotG * s.numGO + min(i, s.numGO - 1)) : 0u;
            const int32_t inTm = MZ ? xldg<X>(gp(a.timer2) + (size_t)parPrev * nIOc + min(i, nIOc - 1)) : 0;
            // (microzone part) the previous step's SC spikes
            const uint8_t scPrev = PART == kPartMz ? xldg<X>(gp(st.apSCin) + min(i, s.numSC - 1)) : (uint8_t)0;
            const uint32_t pz = ip / L.popStride, pk = ip % L.popStride;
            const auto pw = pk < 3 ? SF<LFM, uint32_t>(a, CBM_F_PC_POPACT_CTRS, lds) + pz * 3 + pk
                                           : SF<LFM, uint32_t>(a, CBM_F_HIST_PC_POPACT, lds) + (size_t)pz * bins +
                                                 min(pk - 3, bins - 1);
            const uint32_t pv32 = *pw;
            const uint8_t pv8 = SF<LFM, uint8_t>(a, CBM_F_MFNC_NOLTP_NOLTD, lds)[pz * 2 + (pk >= 3 + bins ? pk - 3 - bins : 0)];
            const uint8_t m = gp(st.mf)[im];
            const uint8_t h = histMF[im];
            const uint32_t b = apBufMF[im];
            const uint32_t pv = pk < 3 + bins ? pv32 : pv8;
            if (i < nMFw) mfl[i] = w;
            if (MZ && i < L.idx16) bd[i] = bl;
            if (MZ && i < nPCc) at<float>(lds, L.inPF)[i] = inPF;
            if (MZ && i < s.numBC) at<uint32_t>(lds, L.inBC)[i] = inBC;
            if (IN && i < s.numSC) at<uint32_t>(lds, L.inSC)[i] = inSC;
            if (IN && i < s.numGO) at<uint32_t>(lds, L.inGO)[i] = inGO;
            if (MZ && i < nIOc) at<int32_t>(lds, L.inTm)[i] = inTm;
            if (PART == kPartMz && i < s.numSC) s_apSC[i] = scPrev;
            if (MZ && i < nRng) s_rng[i] = r;
            if (MZ && i < nNext) s_next[i] = rn;
            if (MZ && i < nPop) s_pop[i] = pv;
            if (i < s.numMF) {
                const uint8_t ap = m > 0;
                if constexpr (MZ) s_hmf[i] = h || ap;
                if constexpr (IN) apBufMF[i] = (b << 1) | (ap * 0x00000001u);
                if (IN && ap) {
                    const G<u32x4> *row = (const G<u32x4> *)(a.mfgoOut + (size_t)i * a.mfgoPad);
                    const uint32_t nv = a.mfgoPad / 8;
                    for (uint32_t c0 = 0; c0 < nv; c0 += 4) {
                        u32x4 wv[4];
#pragma unroll
                        for (int k = 0; k < 4; k++) wv[k] = row[min(c0 + k, nv - 1)];
#pragma unroll
                        for (int k = 0; k < 4; k++) {
                            if (c0 + k >= nv) break;
                            const uint32_t v4[4] = {wv[k].x, wv[k].y, wv[k].z, wv[k].w};
#pragma unroll
                            for (int e = 0; e < 4; e++) {
                                const uint32_t lo = v4[e] & 0xFFFFu, hi = v4[e] >> 16;
                                if (lo != kNoCell) atomicAdd(&s_goMF[lo], 1u);
                                if (hi != kNoCell) atomicAdd(&s_goMF[hi], 1u);
                            }
                        }
                    }
                }
            }
        }
    }
    // (microzone part) the SC->PC routing of the previous step reads the
    // staged index lists and SC spikes
    if constexpr (PART == kPartMz) __syncthreads();
    // synapses: SC->PC of calcPCActivities (mzone.cpp:201-207), MF->NC and
    // PC->NC of calcNCActivities (mzone.cpp:358-383), NC->IO of
    // calcIOActivities (mzone.cpp:272-280)
    if constexpr (MZ) {
        auto gSC = SF<LFM, float>(a, CBM_F_G_SC_PC, lds);
        auto inSC = SF<LFM, uint8_t>(a, CBM_F_INPUT_SC_PC, lds);
        auto inMF = SF<LFM, uint8_t>(a, CBM_F_INPUT_MF_NC, lds);
        auto wMF = SF<LFM, float>(a, CBM_F_MF_SYNW_NC, lds);
        auto gmN = SF<LFM, float>(a, CBM_F_GMAX_NMDA_MF_NC, lds);
        auto gmA = SF<LFM, float>(a, CBM_F_GMAX_AMPA_MF_NC, lds);
        auto gN = SF<LFM, float>(a, CBM_F_G_MF_NMDA_NC, lds);
        auto gA = SF<LFM, float>(a, CBM_F_G_MF_AMPA_NC, lds);
        auto inPC = SF<LFM, uint8_t>(a, CBM_F_INPUT_PC_NC, lds);
        auto scPC = SF<LFM, float>(a, CBM_F_G_PC_SCALE_NC, lds);
        auto gPC = SF<LFM, float>(a, CBM_F_G_PC_NC, lds);
        auto gNC = SF<LFM, float>(a, CBM_F_G_NC_IO, lds);
        auto inNC = SF<LFM, uint8_t>(a, CBM_F_INPUT_NC_IO, lds);
        const uint32_t nS = nZ * s.numSC, nM = nZ * s.numNC * kMF, nP = nZ * s.numNC * kPC,
                       nI = nZ * s.numIO * kNCIO;
        const uint32_t rounds = (max(max(nS, nM), max(nP, nI)) + nThr - 1) / nThr;
        const float ms = p.msPerTimeStep;
        for (uint32_t rd = 0; rd < rounds; rd++) {
            // group-local indices (each group starts on its own wave), clamped for the loads
            const uint32_t jS = rd * nThr + ((tid - L.oSynSC) & (nThr - 1)), jM = rd * nThr + ((tid - L.oSynMF) & (nThr - 1));
            const uint32_t jP = rd * nThr + ((tid - L.oSynPC) & (nThr - 1)), jI = rd * nThr + ((tid - L.oSynIO) & (nThr - 1));
            const uint32_t qs = min(jS, nS - 1), qm = min(jM, nM - 1), qp = min(jP, nP - 1), qi = min(jI, nI - 1);
            float g1 = gSC[qs];
            // (microzone part: MZone::updateSCPCOut (mzone.cpp:467-477) of the
            // previous step, here)
            const uint8_t i1 = (PART == kPartMz && st.scFromPrev) ? s_apSC[s_idx[L.iSCPC + qs]] : inSC[qs];
            const uint8_t i2 = inMF[qm];
            const float w = wMF[qm];
            float mN = gmN[qm], mA = gmA[qm], nn = gN[qm], aa = gA[qm];
            const uint8_t i3 = inPC[qp];
            const float sc = scPC[qp];
            float g3 = gPC[qp];
            float g4 = gNC[qi];
            const uint8_t i4 = inNC[qi];
            if (jS < nS) {
                const uint32_t q = jS;
                g1 = g1 + p.gIncSCtoPC * (1 - g1) * i1;
                g1 = g1 * p.gDecSCtoPC;
                if (PART == kPartMz && st.scFromPrev) inSC[q] = i1;
                gSC[q] = g1;
                s_syn[L.synPC + q] = g1;
            }
            if (jM < nM) {
                const uint32_t q = jM;
                mN = mN * p.gmaxNMDADecMFtoNC + i2 * w * (1 - mN);
                nn = nn + p.gNMDAIncMFtoNC * (mN - nn);
                mA = mA * p.gmaxAMPADecMFtoNC + i2 * w * (1 - mA);
                aa = aa + p.gAMPAIncMFtoNC * (mA - aa);
                gmN[q] = mN;
                gN[q] = nn;
                gmA[q] = mA;
                gA[q] = aa;
                s_syn[L.synN + q] = nn;
                s_syn[L.synA + q] = aa;
            }
            if (jP < nP) {
                const uint32_t q = jP;
                g3 = g3 * p.gDecPCtoNC + i3 * sc * (1 - g3);
                gPC[q] = g3;
                s_syn[L.synP + q] = g3;
            }
            if (jI < nI) {
                const uint32_t q = jI;
                g4 = g4 * expf(-ms / (-p.gDecTSofNCtoIO * expf(-g4 / p.gDecTTofNCtoIO) + p.gDecT0ofNCtoIO));
                g4 = g4 + i4 * p.gIncNCtoIO * expf(-g4 / p.gIncTauNCtoIO);
                gNC[q] = g4;
                s_syn[L.synIO + q] = g4;
            }
        }
    }
    __syncthreads();

    // ======================= phase 1
    // MZone::updateMFNCOut (mzone.cpp:514-524): the caller's raw MF bytes
    if constexpr (MZ) {
        auto inMF = SF<LFM, uint8_t>(a, CBM_F_INPUT_MF_NC, lds);
        for (uint32_t q = tid; q < nZ * s.numNC * kMF; q += nThr) inMF[q] = s_mf[s_idx[L.iMFNC + q]];
    }
    // cell updates: MZone::calcPCActivities (mzone.cpp:187-226),
    // calcBCActivities (mzone.cpp:228-256), InNet::calcSCActivities
    // (innet.cpp:892-911), InNet::calcGOActivities (innet.cpp:827-876),
    // MZone::calcIOActivities (mzone.cpp:282-334) with the plasticity timer of
    // MZone::updateIOOut (mzone.cpp:484), and MZone::calcNCActivities
    // (mzone.cpp:371-396) with the release probability of MZone::updateNCOut
    // (mzone.cpp:497-503); the last two depend only on the cell's own spike.
    // Cell i of every population but GO is handled by the same thread; GO
    // follows in its own pass to bound the registers held at once.
    {
        const uint32_t nPCc = nZ * s.numPC, nBCc = nZ * s.numBC, nSCc = s.numSC, nGOc = s.numGO;
        const uint32_t nIOc = nZ * s.numIO, nNCc = nZ * s.numNC;
        // the NC input sums of calcNCActivities (mzone.cpp:357-384), each in
        // the reference's order: the MF NMDA, MF AMPA and PC sums of every NC
        // on three waves of their own, beside the other populations' updates
        // (at 1M GR the MF sums are 128-term chains, the longest of the
        // phase); the NC update reads them after a barrier
        float *ncSum = at<float>(lds, L.ncSum);
        if constexpr (MZ) {
            const uint32_t w = (tid - L.oNCS) & (nThr - 1), c = w >> 6;
            if (c < 3) {
                const uint32_t n = c == 2 ? kPC : kMF, base = c == 0 ? L.synN : c == 1 ? L.synA : L.synP;
                for (uint32_t i = w & 63u; i < nNCc; i += 64) ncSum[c * nNCc + i] = lds_sum_seq(s_syn + base + i * n, n, 0.0f);
            }
        }
        const uint32_t rounds = (max(max(nPCc, nBCc), max(nSCc, nIOc)) + nThr - 1) / nThr;
        for (uint32_t rd = 0; rd < rounds; rd++) {
            // population-local cell indices: each population starts on its own wave
            const uint32_t cP = rd * nThr + ((tid - L.oPC) & (nThr - 1)), cB = rd * nThr + ((tid - L.oBC) & (nThr - 1));
            const uint32_t cS = rd * nThr + ((tid - L.oSC) & (nThr - 1)), cI = rd * nThr + ((tid - L.oIO) & (nThr - 1));
            const bool hPC = cP < nPCc, hBC = cB < nBCc, hSC = cS < nSCc, hIO = cI < nIOc;
            // ---- loads, at clamped indices (in bounds, no branch)
            const uint32_t iP = min(cP, nPCc - 1), iB = min(cB, nBCc - 1), iS = min(cS, nSCc - 1);
            const uint32_t iI = min(cI, nIOc - 1);
            const float pcPF = at<float>(lds, L.inPF)[iP];
            const uint32_t pcIn = SF<LFM, uint32_t>(a, CBM_F_INPUT_BC_PC, lds)[iP];
            const float pcGPF = SF<LFM, float>(a, CBM_F_G_PF_PC, lds)[iP], pcGBC = SF<LFM, float>(a, CBM_F_G_BC_PC, lds)[iP];
            const float pcV = SF<LFM, float>(a, CBM_F_VM_PC, lds)[iP], pcTh = SF<LFM, float>(a, CBM_F_THRESH_PC, lds)[iP];
            const uint32_t pcAb = SF<LFM, uint32_t>(a, CBM_F_APBUF_PC, lds)[iP];
            const uint32_t bcPF = at<uint32_t>(lds, L.inBC)[iB & (s.numBC - 1)];
            const uint32_t bcIn = SF<LFM, uint32_t>(a, CBM_F_INPUT_PC_BC, lds)[iB];
            const float bcGPF = SF<LFM, float>(a, CBM_F_G_PF_BC, lds)[iB], bcGPC = SF<LFM, float>(a, CBM_F_G_PC_BC, lds)[iB];
            const float bcV = SF<LFM, float>(a, CBM_F_VM_BC, lds)[iB], bcTh = SF<LFM, float>(a, CBM_F_THRESH_BC, lds)[iB];
            const uint32_t bcAb = SF<LFM, uint32_t>(a, CBM_F_APBUF_BC, lds)[iB];
            const uint32_t scPF = at<uint32_t>(lds, L.inSC)[iS];
            const float scGPF = SF<LFM, float>(a, CBM_F_G_PF_SC, lds)[iS], scV = SF<LFM, float>(a, CBM_F_VM_SC, lds)[iS];
            const float scTh = SF<LFM, float>(a, CBM_F_THRESH_SC, lds)[iS];
            const uint32_t scAb = SF<LFM, uint32_t>(a, CBM_F_APBUF_SC, lds)[iS];
            const float ioV = SF<LFM, float>(a, CBM_F_VM_IO, lds)[iI], ioTh = SF<LFM, float>(a, CBM_F_THRESH_IO, lds)[iI];
            const float ioVC = SF<LFM, float>(a, CBM_F_VCOUPLE_IO, lds)[iI];
            const uint32_t ioAb = SF<LFM, uint32_t>(a, CBM_F_APBUF_IO, lds)[iI];
            const int32_t ioTm = at<int32_t>(lds, L.inTm)[iI];
            // ---- PC
            if (MZ && hPC) {
                const uint32_t i = iP, z = i >> L.l2PC;
                float gPF = pcGPF, gBC = pcGBC, vv = pcV, th = pcTh;
                const float gSCPCSum = lds_sum_seq(s_syn + L.synPC + i * kSC, kSC, 0.0f);
                gPF = gPF + pcPF * p.gIncGRtoPC;
                gPF = gPF * p.gDecGRtoPC;
                gBC = gBC + pcIn * p.gIncBCtoPC;
                gBC = gBC * p.gDecBCtoPC;
                vv = vv + (p.gLeakPC * (p.eLeakPC - vv)) - (gPF * vv) + (gBC * (p.eBCtoPC - vv)) +
                     (gSCPCSum * (p.eSCtoPC - vv));
                th = th + (p.threshDecPC * (p.threshRestPC - th));
                const uint8_t ap = vv > th;
                SF<LFM, float>(a, CBM_F_G_PF_PC, lds)[i] = gPF;
                SF<LFM, float>(a, CBM_F_G_BC_PC, lds)[i] = gBC;
                SF<LFM, float>(a, CBM_F_VM_PC, lds)[i] = vv;
                SF<LFM, uint8_t>(a, CBM_F_AP_PC, lds)[i] = ap;
                s_apPC[i] = ap;
                SF<LFM, uint32_t>(a, CBM_F_APBUF_PC, lds)[i] = (pcAb << 1) | (ap * 0x00000001u);
                SF<LFM, float>(a, CBM_F_THRESH_PC, lds)[i] = ap * p.threshMaxPC + (!ap) * th;
                if (ap) atomicAdd(&s_pop[z * L.popStride + 2], 1u);
            }
            // ---- BC
            if (MZ && hBC) {
                const uint32_t i = iB;
                float gPF = bcGPF, gPC = bcGPC, vv = bcV, th = bcTh;
                gPF = gPF + (bcPF * p.gIncGRtoBC);
                gPF = gPF * p.gDecGRtoBC;
                gPC = gPC + (bcIn * p.gIncPCtoBC);
                gPC = gPC * p.gDecPCtoBC;
                vv = vv + (p.gLeakBC * (p.eLeakBC - vv)) - (gPF * vv) + (gPC * (p.ePCtoBC - vv));
                th = th + p.threshDecBC * (p.threshRestBC - th);
                const uint8_t ap = vv > th;
                SF<LFM, float>(a, CBM_F_G_PF_BC, lds)[i] = gPF;
                SF<LFM, float>(a, CBM_F_G_PC_BC, lds)[i] = gPC;
                SF<LFM, float>(a, CBM_F_VM_BC, lds)[i] = vv;
                SF<LFM, uint8_t>(a, CBM_F_AP_BC, lds)[i] = ap;
                s_apBC[i] = ap;
                SF<LFM, uint32_t>(a, CBM_F_APBUF_BC, lds)[i] = (bcAb << 1) | (ap * 0x00000001u);
                SF<LFM, float>(a, CBM_F_THRESH_BC, lds)[i] = ap * p.threshMaxBC + (!ap) * (th);
            }
            // ---- SC
            if (IN && hSC) {
                const uint32_t i = iS;
                float gPF = scGPF, vv = scV, th = scTh;
                gPF = gPF + (scPF * p.gIncGRtoSC);
                gPF = gPF * p.gDecGRtoSC;
                vv = vv + (p.gLeakSC * (p.eLeakSC - vv)) - gPF * vv;
                const uint8_t ap = vv > th;
                th = th + p.threshDecSC * (p.threshRestSC - th);
                SF<LFM, float>(a, CBM_F_G_PF_SC, lds)[i] = gPF;
                SF<LFM, float>(a, CBM_F_VM_SC, lds)[i] = vv;
                SF<LFM, uint8_t>(a, CBM_F_AP_SC, lds)[i] = ap;
                s_apSC[i] = ap;
                SF<LFM, uint32_t>(a, CBM_F_APBUF_SC, lds)[i] = (scAb << 1) | (ap * 0x00000001u);
                SF<LFM, float>(a, CBM_F_THRESH_SC, lds)[i] = ap * p.threshMaxSC + (!ap) * (th);
            }
            // ---- IO
            if (MZ && hIO) {
                const uint32_t i = iI, z = i >> L.l2IO;
                float vv = ioV;
                float gNCSum = lds_sum_seq(s_syn + L.synIO + i * kNCIO, kNCIO, 0.0f);
                gNCSum = gNCSum / 3.1;
                vv = vv + p.gLeakIO * (p.eLeakIO - vv) + gNCSum * 1.5 * (p.eNCtoIO - vv) + ioVC + errDrive[z];
                const uint8_t ap = vv > ioTh;
                SF<LFM, float>(a, CBM_F_VM_IO, lds)[i] = vv;
                s_vIO[i] = vv;
                SF<LFM, uint8_t>(a, CBM_F_AP_IO, lds)[i] = ap;
                s_apIO[i] = ap;
                SF<LFM, uint32_t>(a, CBM_F_APBUF_IO, lds)[i] = (ioAb << 1) | (ap * 0x00000001u);
                SF<LFM, float>(a, CBM_F_THRESH_IO, lds)[i] =
                    p.threshMaxIO * ap + (!ap) * (ioTh + p.threshDecIO * (p.threshRestIO - ioTh));
                xstg<X>(gp(a.timer2) + (size_t)par * nIOc + i, (int32_t)((!ap) * (ioTm + 1) + ap * p.tsLTDStartAPIO));
            }
        }
        for (uint32_t i = tid; IN && i < nGOc; i += nThr) {
            float gMF = 0, gGR = 0, gGO = 0, glu = 0, mG = 0, gInc = 0, gMG = 0, thv = 0, v = 0, vC = 0, sc = 0, inDep = 0;
            uint32_t totalGRIn = 0, inMF = 0, goAb = 0;
            {
                totalGRIn = at<uint32_t>(lds, L.inGO)[i];
                inMF = SF<LFM, uint32_t>(a, CBM_F_INPUT_MF_GO, lds)[i];
                inDep = SF<LFM, float>(a, CBM_F_INPUT_GOGABA_SYNDEP_GO, lds)[i];
                gMF = SF<LFM, float>(a, CBM_F_G_MF_GO, lds)[i];
                gGR = SF<LFM, float>(a, CBM_F_G_GR_GO, lds)[i];
                gGO = SF<LFM, float>(a, CBM_F_G_GO_GO, lds)[i];
                glu = SF<LFM, float>(a, CBM_F_GLU_GO, lds)[i];
                mG = SF<LFM, float>(a, CBM_F_MGLUR_GO, lds)[i];
                gInc = SF<LFM, float>(a, CBM_F_G_MGLURINC_GO, lds)[i];
                gMG = SF<LFM, float>(a, CBM_F_G_MGLUR_GO, lds)[i];
                thv = SF<LFM, float>(a, CBM_F_THRESH_GO, lds)[i];
                v = SF<LFM, float>(a, CBM_F_VM_GO, lds)[i];
                vC = SF<LFM, float>(a, CBM_F_VCOUPLE_GO, lds)[i];
                sc = SF<LFM, float>(a, CBM_F_GO_GABA_OUT_SYNSCALE_GOGO, lds)[i];
                goAb = SF<LFM, uint32_t>(a, CBM_F_APBUF_GO, lds)[i];
            }
            {
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
                SF<LFM, uint32_t>(a, CBM_F_SUM_GR_INPUT_GO, lds)[i] = totalGRIn;
                SF<LFM, float>(a, CBM_F_SUM_GO_INPUT_GO, lds)[i] = inDep;
                SF<LFM, float>(a, CBM_F_G_MF_GO, lds)[i] = gMF;
                SF<LFM, float>(a, CBM_F_G_GR_GO, lds)[i] = gGR;
                SF<LFM, float>(a, CBM_F_G_GO_GO, lds)[i] = gGO;
                SF<LFM, float>(a, CBM_F_GLU_GO, lds)[i] = glu;
                SF<LFM, float>(a, CBM_F_MGLUR_GO, lds)[i] = mG;
                SF<LFM, float>(a, CBM_F_G_MGLURINC_GO, lds)[i] = gInc;
                SF<LFM, float>(a, CBM_F_G_MGLUR_GO, lds)[i] = gMG;
                SF<LFM, float>(a, CBM_F_VM_GO, lds)[i] = v;
                xstg<X>(gp(a.apGO2) + (size_t)par * s.numGO + i, ap);
                s_apGO[i] = ap;
                s_scale[i] = sc;
                if (a.anyCoupling) s_vGO[i] = v;
                SF<LFM, uint32_t>(a, CBM_F_APBUF_GO, lds)[i] = (goAb << 1) | (ap * 0x00000001u);
                SF<LFM, float>(a, CBM_F_THRESH_GO, lds)[i] = ap * p.threshMaxGO + (!ap) * thv;
            }
        }
        // ---- NC, once its input sums are in
        if constexpr (MZ) {
            __syncthreads();
            for (uint32_t i = (tid - L.oNC) & (nThr - 1); i < nNCc; i += nThr) {
                float vv = SF<LFM, float>(a, CBM_F_VM_NC, lds)[i], th = SF<LFM, float>(a, CBM_F_THRESH_NC, lds)[i];
                float r = SF<LFM, float>(a, CBM_F_SYN_IO_PRELEASE_NC, lds)[i];
                const uint32_t ncAb = SF<LFM, uint32_t>(a, CBM_F_APBUF_NC, lds)[i];
                const float ms = p.msPerTimeStep;
                float gMFNMDASum = ncSum[i], gMFAMPASum = ncSum[nNCc + i], gPCNCSum = ncSum[2 * nNCc + i];
                gMFNMDASum = gMFNMDASum * ms / ((float)kMF);
                gMFAMPASum = gMFAMPASum * ms / ((float)kMF);
                gMFNMDASum = gMFNMDASum * vv / (-80.0f);
                gPCNCSum = gPCNCSum * ms / ((float)kPC);
                vv = vv + p.gLeakNC * (p.eLeakNC - vv) - (gMFNMDASum + gMFAMPASum) * vv + gPCNCSum * (p.ePCtoNC - vv);
                th = th + p.threshDecNC * (p.threshRestNC - th);
                const uint8_t ap = vv > th;
                SF<LFM, float>(a, CBM_F_VM_NC, lds)[i] = vv;
                SF<LFM, uint8_t>(a, CBM_F_AP_NC, lds)[i] = ap;
                s_apNC[i] = ap;
                SF<LFM, uint32_t>(a, CBM_F_APBUF_NC, lds)[i] = (ncAb << 1) | (ap * 0x00000001u);
                SF<LFM, float>(a, CBM_F_THRESH_NC, lds)[i] = ap * p.threshMaxNC + (!ap) * th;
                r = r * expf(-ms / (p.relPDecTSofNCtoIO * expf(-r / p.relPDecTTofNCtoIO) + p.relPDecT0ofNCtoIO));
                r = r + ap * p.relPIncNCtoIO * expf(-r / p.relPIncTauNCtoIO);
                SF<LFM, float>(a, CBM_F_SYN_IO_PRELEASE_NC, lds)[i] = r;
                s_rel[i] = r;
            }
        }
    }
    __syncthreads();

    // ======================= phase 2
    // MF->GO counts; InNet::updateGOtoGOOut (innet.cpp:937-950): each GO
    // gathers its GABA inputs in increasing source order, which is the order
    // the reference adds them
    if constexpr (IN) {
        auto inMFGO = SF<LFM, uint32_t>(a, CBM_F_INPUT_MF_GO, lds);
        auto inGO = SF<LFM, uint32_t>(a, CBM_F_INPUT_GO_GO, lds);
        auto inDep = SF<LFM, float>(a, CBM_F_INPUT_GOGABA_SYNDEP_GO, lds);
        for (uint32_t g = tid; g < s.numGO; g += nThr) {
            uint32_t c = 0;
            float dep = 0;
            const G<u32x2> *row = (const G<u32x2> *)(a.gabaEll + (size_t)g * a.gabaPad);
            for (uint32_t k0 = 0; k0 < a.gabaPad / 4; k0++) {
                const u32x2 w = row[k0];
                const uint32_t src[4] = {w.x & 0xFFFFu, w.x >> 16, w.y & 0xFFFFu, w.y >> 16};
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    if (src[k] != kNoCell && s_apGO[src[k]]) {
                        c++;
                        dep += s_scale[src[k]];
                    }
                }
            }
            inMFGO[g] = s_goMF[g];
            inGO[g] = c;
            inDep[g] = dep;
        }
    }
    // GO electrical coupling (innet.cpp:961-978): vCoupleGO is a Jacobi sweep;
    // the threshold update is in place in cell order, so it runs on one lane.
    // Without coupling inputs (or with a zero ratio) both terms are zero.
    if constexpr (IN) {
        auto vC = SF<LFM, float>(a, CBM_F_VCOUPLE_GO, lds);
        if (a.anyCoupling) {
            const uint32_t kc = s.maxnumpGOCoupInGOGO;
            const float ratio = p.coupleRiRjRatioGO;
            for (uint32_t i = tid; i < s.numGO; i += nThr) {
                float acc = 0;
                for (int j = 0; j < gp(a.coupN)[i]; j++) acc = acc + ratio * (s_vGO[gp(a.coupP)[(size_t)i * kc + j]] - s_vGO[i]);
                vC[i] = acc;
            }
            if (tid == 0) {
                auto th = SF<LFM, float>(a, CBM_F_THRESH_GO, lds);
                for (uint32_t i = 0; i < s.numGO; i++) {
                    float tc = 0;
                    for (int j = 0; j < gp(a.coupN)[i]; j++) tc = tc + ratio * (th[gp(a.coupP)[(size_t)i * kc + j]] - th[i]);
                    th[i] = th[i] + tc;
                }
            }
        } else {
            for (uint32_t i = tid; i < s.numGO; i += nThr) vC[i] = 0.0f;
        }
    }
    // (input-layer part) this step's SC spikes for the microzone part
    if constexpr (PART == kPartIn)
        for (uint32_t i = tid; i < s.numSC; i += nThr) xstg<X>(gp(st.apSCout) + i, s_apSC[i]);
    // microzone outputs: MZone::updateBCPCOut (mzone.cpp:449-465),
    // updateSCPCOut (mzone.cpp:467-477), updatePCOut (mzone.cpp:409-447) and
    // the IO coupling of updateIOOut (mzone.cpp:486-489), from LDS
    if constexpr (MZ) {
        auto inBCPC = SF<LFM, uint32_t>(a, CBM_F_INPUT_BC_PC, lds);
        for (uint32_t i = (tid - L.oRBCPC) & (nThr - 1); i < nZ * s.numPC; i += nThr) {
            const uint32_t z = i >> L.l2PC, pc = i & (s.numPC - 1);
            const uint16_t *ell = s_idx + L.iBCPC + (size_t)z * L.slotsBCPC * s.numPC + pc;
            uint32_t c = 0;
            // (every LDS read at a clamped index, then selects: no read
            // waits for the one before)
            for (uint32_t k0 = 0; k0 < L.slotsBCPC; k0 += 8) {
                uint32_t b[8], v[8];
#pragma unroll
                for (int k = 0; k < 8; k++) {
                    const uint32_t e = ell[(size_t)min(k0 + k, L.slotsBCPC - 1) * s.numPC];
                    b[k] = k0 + k < L.slotsBCPC ? e : kNoCell;
                }
#pragma unroll
                for (int k = 0; k < 8; k++) v[k] = s_apBC[z * s.numBC + min(b[k], s.numBC - 1)];
#pragma unroll
                for (int k = 0; k < 8; k++) c += (b[k] != kNoCell) & (v[k] != 0);
            }
            inBCPC[i] = c;
        }
        if constexpr (PART == kPartAll) {  // (microzone part: at the next step's start)
            auto inSC = SF<LFM, uint8_t>(a, CBM_F_INPUT_SC_PC, lds);
            for (uint32_t q = (tid - L.oRSCPC) & (nThr - 1); q < nZ * s.numPC * kSC; q += nThr)
                inSC[q] = s_apSC[s_idx[L.iSCPC + q]];
        }
        auto inPCBC = SF<LFM, uint32_t>(a, CBM_F_INPUT_PC_BC, lds);
        for (uint32_t i = (tid - L.oRPCBC) & (nThr - 1); i < nZ * s.numBC; i += nThr) {
            const uint32_t z = i >> L.l2BC, bc = i & (s.numBC - 1);
            const uint16_t *ell = s_idx + L.iPCBC + (size_t)z * L.slotsPCBC * s.numBC + bc;
            uint32_t c = 0;
            for (uint32_t k0 = 0; k0 < L.slotsPCBC; k0 += 8) {
                uint32_t pc[8], v[8];
#pragma unroll
                for (int k = 0; k < 8; k++) {
                    const uint32_t e = ell[(size_t)min(k0 + k, L.slotsPCBC - 1) * s.numBC];
                    pc[k] = k0 + k < L.slotsPCBC ? e : kNoCell;
                }
#pragma unroll
                for (int k = 0; k < 8; k++) v[k] = s_apPC[z * s.numPC + min(pc[k], s.numPC - 1)];
#pragma unroll
                for (int k = 0; k < 8; k++) c += (pc[k] != kNoCell) * v[k];
            }
            inPCBC[i] = c;
        }
        auto inPCNC = SF<LFM, uint8_t>(a, CBM_F_INPUT_PC_NC, lds);
        const uint32_t per = s.numNC * kPC;
        for (uint32_t q = (tid - L.oRPCNC) & (nThr - 1); q < nZ * per; q += nThr)
            inPCNC[q] = s_apPC[(q / per) * s.numPC + s_idx[L.iPCNC + q]];
        auto vCIO = SF<LFM, float>(a, CBM_F_VCOUPLE_IO, lds);
        for (uint32_t i = (tid - L.oRIO) & (nThr - 1); i < nZ * s.numIO; i += nThr) {
            const uint32_t z = i >> L.l2IO;
            float vc = 0;
            for (uint32_t j = 0; j < kIOIO; j++)
                vc = vc + p.coupleRiRjRatioIO * (s_vIO[z * s.numIO + s_idx[L.iIOIO + i * kIOIO + j]] - s_vIO[i]);
            vCIO[i] = vc;
        }
    }
    // MZone::updateMFNCSyn population histogram (mzone.cpp:536-580), one lane per zone
    if (MZ && tid < (int)nZ) {
        const uint32_t z = tid;
        uint32_t *ctrs = s_pop + z * L.popStride, *hist = ctrs + 3, *flags = ctrs + 3 + bins;
        int doLTD = 0, doLTP = 0;
        if (st.popHist) {
            const uint32_t b = ctrs[1] < bins ? ctrs[1] : 0;  // commit keeps ctrs[1] < bins
            ctrs[0] = ctrs[0] - hist[b] + ctrs[2];
            hist[b] = ctrs[2];
            ctrs[2] = 0;
            ctrs[1]++;
            ctrs[1] = ctrs[1] % p.numPopHistBinsPC;
            const float avg = ((float)ctrs[0]) / p.numPopHistBinsPC;
            if (avg >= p.synLTDPCPopActThreshMFtoNC && !flags[1]) {
                doLTD = 1;
                flags[1] = 1;
            } else if (avg < p.synLTDPCPopActThreshMFtoNC) {
                flags[1] = 0;
            }
            if (avg <= p.synLTPPCPopActThreshMFtoNC && !flags[0]) {
                doLTP = 1;
                flags[0] = 1;
            } else if (avg > p.synLTPPCPopActThreshMFtoNC) {
                flags[0] = 0;
            }
        }
        s_doLTD[z] = doLTD;
        s_doLTP[z] = doLTP;
        for (uint32_t k = 0; k < 3; k++) SF<LFM, uint32_t>(a, CBM_F_PC_POPACT_CTRS, lds)[z * 3 + k] = ctrs[k];
        for (uint32_t k = 0; k < bins; k++) SF<LFM, uint32_t>(a, CBM_F_HIST_PC_POPACT, lds)[(size_t)z * bins + k] = hist[k];
        for (uint32_t k = 0; k < 2; k++) SF<LFM, uint8_t>(a, CBM_F_MFNC_NOLTP_NOLTD, lds)[z * 2 + k] = (uint8_t)flags[k];
    }
    // spike recorder (row t mod recCap), and one population's spikes for the
    // host's step-and-read call straight into its pinned buffer
    {
        const uint8_t *src[6] = {s_apPC, s_apBC, s_apSC, s_apIO, s_apNC, s_apGO};
        const uint32_t cnt[6] = {nZ * s.numPC, nZ * s.numBC, s.numSC, nZ * s.numIO, nZ * s.numNC, s.numGO};
        if (a.rec) {
            uint8_t *row = a.rec + (size_t)(st.t % a.recCap) * a.recRow;
            for (int k = 0; k < 6; k++) {
                if (a.recOff[k] == 0xFFFFFFFFu) continue;
                // (a part records its own populations: SC and GO the input layer)
                if (PART != kPartAll && (PART == kPartIn) != (k == CBM_REC_SC || k == CBM_REC_GO)) continue;
                for (uint32_t i = tid; i < cnt[k]; i += nThr) gp(row)[a.recOff[k] + i] = src[k][i];
            }
        }
        if (st.ioOut) {
#pragma unroll
            for (int k = 0; k < 6; k++)
                if (st.ioPop == (uint32_t)k)
                    for (uint32_t i = tid; i < cnt[k]; i += nThr) st.ioOut[i] = src[k][i];
        }
    }
    // the ring slot of step t+1 is the slot PC/BC/SC read above (t-2): clear it now
    for (uint32_t i = tid; MZ && i < s.numBC; i += nThr) xstg<X>(gp(a.bcRing) + (size_t)st.slotNext * s.numBC + i, 0u);
    for (uint32_t i = tid; IN && i < s.numSC; i += nThr) xstg<X>(gp(a.scRing) + (size_t)st.slotNext * s.numSC + i, 0u);
    for (uint32_t i = tid; IN && i < s.numGO; i += nThr) xstg<X>(gp(a.goRing) + (size_t)st.slotNext * s.numGO + i, 0u);
    __syncthreads();

    // ======================= phase 3
    // MZone::updateNCOut draws (mzone.cpp:505-511): one Random() per (IO, NC)
    // in row order from the zone's stream.  Draw k of a run uses generator
    // words ix+2k, ix+2k+1, so up to 64 draws that fit before the next
    // regeneration are taken at once, one per lane.  A regeneration takes the
    // block prebuilt by an earlier launch when its tag says it follows the
    // current one, and computes it here otherwise.
    if (MZ && wave == 0) {
        auto inNC = SF<LFM, uint8_t>(a, CBM_F_INPUT_NC_IO, lds);
        const uint32_t n = s.numIO * kNCIO;
        const uint32_t *s_next = at<uint32_t>(lds, L.rngNext);
        for (uint32_t z = 0; z < nZ; z++) {
            uint32_t *g = s_rng + (size_t)z * kRngZone;
            const uint32_t *nx = s_next + (size_t)z * kRngNext;
            uint32_t ix = g[352], gen = g[353];
            const uint16_t *pn = s_idx + L.iNCIO + (size_t)z * n;
            const float *rel = s_rel + (size_t)z * s.numNC;
#pragma unroll 1
            for (uint32_t done = 0; done < n;) {
                if (ix >= 351) {
                    __threadfence_block();
                    if (nx[352] == gen + 1u) {  // wave-uniform
                        for (uint32_t w = lane; w < 352; w += 64) g[w] = nx[w];
                    } else if (lane < 4) {
                        sfmt_regen4(g, lane);
                    }
                    __threadfence_block();
                    gen += 1u;
                    ix = 0;
                }
                uint32_t k = (352u - ix) >> 1;
                k = k < 64u ? k : 64u;
                k = k < n - done ? k : n - done;
                if ((uint32_t)lane < k) {
                    const uint32_t q = done + lane, w = ix + 2 * lane;
                    const uint64_t r = (uint64_t)g[w] | ((uint64_t)g[w + 1] << 32);
                    const double u = (double)(int64_t)(r >> 12) * (1.0 / 4503599627370496.0);
                    inNC[(size_t)z * n + q] = (u < rel[pn[q]]);
                }
                ix += 2 * k;
                done += k;
            }
            if (lane == 0) {
                g[352] = ix;
                g[353] = gen;
            }
        }
        __threadfence_block();
        for (uint32_t i = lane; i < nZ * kRngZone; i += 64) xstg<X>(gp(a.rngState) + (size_t)par * nZ * kRngZone + i, s_rng[i]);
    }
    // GO->GO synaptic depression scale (innet.cpp:952-954), after the GABA
    // gather above has used the previous scale
    if constexpr (IN) {
        auto scale = SF<LFM, float>(a, CBM_F_GO_GABA_OUT_SYNSCALE_GOGO, lds);
        const float rec = p.goGABAGOGOSynRec, depF = p.goGABAGOGOSynDepF;
        for (uint32_t i = tid; i < s.numGO; i += nThr) {
            const float sc = s_scale[i];
            const uint8_t ap = s_apGO[i];
            scale[i] = sc + (!ap) * (1 - sc) * rec - ap * (sc * (1 - depF));
        }
    }
    // MZone::updateMFNCSyn weights (mzone.cpp:586-600)
    if (MZ && st.popHist) {
        const uint32_t per = s.numNC * kMF;
        auto w = SF<LFM, float>(a, CBM_F_MF_SYNW_NC, lds);
        for (uint32_t q = tid; q < nZ * per; q += nThr) {
            const uint32_t z = q / per;
            const int doLTD = s_doLTD[z], doLTP = s_doLTP[z];
            const float delta = s_hmf[s_idx[L.iMFNC + q]] * (doLTD * p.synLTDStepSizeMFtoNC + doLTP * p.synLTPStepSizeMFtoNC);
            float wv = w[q];
            wv = wv + delta;
            wv = (wv > 0) * wv;
            wv = (wv <= 1) * wv + (wv > 1);
            w[q] = wv;
        }
    }
    // MF history, cleared every numTSinMFHist steps (innet.cpp:982-995), kept
    // by the microzone part (MF->NC plasticity reads it)
    if constexpr (MZ) {
        auto histMF = SF<LFM, uint8_t>(a, CBM_F_HIST_MF, lds);
        for (uint32_t i = tid; i < s.numMF; i += nThr) histMF[i] = st.resetMFHist ? 0 : s_hmf[i];
    }
}

// The zone generators' next blocks, built ahead of use by the last block of
// a fused launch: from the state the previous step left, block gen + 1 into
// this step's slot, or a copy of the previous slot when that already holds
// it.  The whole block moves the words (one load round trip for all of them),
// and lanes 4z..4z+3 of wave 0 run zone z's recursion in LDS as sfmt_regen4
// does.  The small step takes a prebuilt block only from a slot an earlier
// launch wrote.
template <bool X = false>
__device__ void sfmt_prebuild(const SmallArgs &a, const SmallStep &st, uint32_t *lds)
{
    __shared__ uint32_t s_gen[kMaxZones], s_copy[kMaxZones];
    const uint32_t nZ = a.s.numZones, tid = threadIdx.x;
    const uint32_t par = st.par, parPrev = par ^ 1u;
    const G<uint32_t> *cur = gp(a.rngState) + (size_t)parPrev * nZ * kRngZone;
    const G<uint32_t> *nin = gp(a.rngNext) + (size_t)parPrev * nZ * kRngNext;
    G<uint32_t> *nout = gp(a.rngNext) + (size_t)par * nZ * kRngNext;
    if (tid < nZ) {
        const uint32_t gen = xldg<X>(cur + (size_t)tid * kRngZone + 353), tag = xldg<X>(nin + (size_t)tid * kRngNext + 352);
        s_gen[tid] = gen;
        s_copy[tid] = tag == gen + 1u;
    }
    __syncthreads();
    // the previous slot's block (copy) or the current state (regenerate), at
    // most three words per thread for eight zones, all loads issued first
    constexpr uint32_t kPer = (kMaxZones * 352 + kSmallThreads - 1) / kSmallThreads;
    const uint32_t n = nZ * 352;
    uint32_t v[kPer];
#pragma unroll
    for (int j = 0; j < (int)kPer; j++) {
        const uint32_t i = min(tid + (uint32_t)j * kSmallThreads, n - 1u), z = i / 352, w = i % 352;
        v[j] = s_copy[z] ? xldg<X>(nin + (size_t)z * kRngNext + w) : xldg<X>(cur + (size_t)z * kRngZone + w);
    }
#pragma unroll
    for (int j = 0; j < (int)kPer; j++) {
        const uint32_t i = tid + (uint32_t)j * kSmallThreads;
        if (i < n) lds[i] = v[j];
    }
    __syncthreads();
    {
        const uint32_t z = tid >> 2, k = tid & 3u;
        if (tid < 64 && z < nZ && !s_copy[z]) sfmt_regen4(lds + (size_t)z * 352, (int)k);  // whole quads
    }
    __syncthreads();
    for (uint32_t i = tid; i < n; i += kSmallThreads) xstg<X>(nout + (size_t)(i / 352) * kRngNext + i % 352, lds[i]);
    if (tid < nZ) xstg<X>(nout + (size_t)tid * kRngNext + 352, s_gen[tid] + 1u);
}

// the small populations alone (the two-stream step, and per-kernel timing)
__global__ void __launch_bounds__(kSmallThreads) small_step_kernel(const SmallArgs *__restrict__ args,
                                                                   const SmallStep st)
{
    extern __shared__ __align__(16) uint8_t lds[];
    small_step(*args, st, st.errDrive, lds);
}

// One launch per step: block 0 steps the small populations, blocks 1.. are GR
// tiles of 1024 threads (one block per CU: the small step needs up to 128
// VGPRs per thread), and the last block prebuilds the generators' next blocks.
template <int CPT, int FMT>
__global__ void __launch_bounds__(kSmallThreads) step_kernel(const SmallArgs *__restrict__ args, const SmallStep st, const GrArgs g, const GrStep gs)
{
    extern __shared__ __align__(16) uint8_t lds[];
    if (blockIdx.x == 0)
        small_step(*args, st, st.errDrive, lds);
    else if (blockIdx.x == gridDim.x - 1)
        sfmt_prebuild(*args, st, reinterpret_cast<uint32_t *>(lds));
    else
        gr_tile<CPT, kSmallThreads, FMT>(g, gs, blockIdx.x - 1, reinterpret_cast<uint32_t *>(lds));
}

template __global__ void step_kernel<1, kFmtCompact>(const SmallArgs *, const SmallStep, const GrArgs, const GrStep);
template __global__ void step_kernel<2, kFmtCompact>(const SmallArgs *, const SmallStep, const GrArgs, const GrStep);
template __global__ void step_kernel<1, kFmtArrays>(const SmallArgs *, const SmallStep, const GrArgs, const GrStep);
template __global__ void step_kernel<2, kFmtArrays>(const SmallArgs *, const SmallStep, const GrArgs, const GrStep);

// The same with the GR cells split evenly over `spans` blocks (one per CU
// but the small step's and the generator block's): all of a step's GR loads
// are issued in one wave of blocks that runs beside the small step.
template <int CPT, int FMT>
__global__ void __launch_bounds__(kSmallThreads) step_span_kernel(const SmallArgs *__restrict__ args, const SmallStep st,
                                                                  const GrArgs g, const GrStep gs, const uint32_t spans)
{
    extern __shared__ __align__(16) uint8_t lds[];
    if (blockIdx.x == 0)
        small_step(*args, st, st.errDrive, lds);
    else if (blockIdx.x > spans)
        sfmt_prebuild(*args, st, reinterpret_cast<uint32_t *>(lds));
    else
        gr_span_of<CPT, kSmallThreads, FMT>(g, gs, blockIdx.x - 1, spans, reinterpret_cast<uint32_t *>(lds));
}

#define CBM_STEP_SPAN(CPT, FMT)                                                                              \
    template __global__ void step_span_kernel<CPT, FMT>(const SmallArgs *, const SmallStep, const GrArgs, \
                                                        const GrStep, const uint32_t);
CBM_STEP_SPAN(2, kFmtCompact)
CBM_STEP_SPAN(3, kFmtCompact)
CBM_STEP_SPAN(5, kFmtCompact)
CBM_STEP_SPAN(2, kFmtArrays)
CBM_STEP_SPAN(3, kFmtArrays)
CBM_STEP_SPAN(5, kFmtArrays)
#undef CBM_STEP_SPAN

// ---------------------------------------------------------------- resident run
// Step-to-step synchronisation of a resident run, point to point instead of a
// grid barrier.  Step t of
//  * the small populations needs every GR block to have finished step t-1
//    (their ring-slot counts) and the generator block step t-1 (the
//    prebuilt generator block, the fold of the chunk sums of t-2);
//  * a GR block needs the small populations to have finished step t-1 (GO
//    spikes, plasticity timers, cleared ring slot) and with it the generator
//    block (the fold of the chunk sums of t-2, whose buffer step t
//    overwrites);
//  * the generator block needs the small populations to have finished step
//    t-1 (the generator state) and every GR block step t-1 (the chunk sums
//    it folds).
// So a GR block that finishes a step adds one to the counter of its group
// (blocks b and b + 8 share one; eight counters 4 KB apart), the generator
// block to a counter of its own, and the small block, once it has finished a
// step and the generator block has too, writes its step count into
// kFlagCopies flag words 4 KB apart; each other block polls one of them, so
// that no word has more than ~16 pollers, with a growing pause between polls
// (up to ~0.4 us) so that their loads do not crowd the memory channels the
// small block uses meanwhile.
// A signalling block first waits until its own stores and atomics are
// acknowledged, so what it wrote for the other blocks has reached the point
// of coherence.  Counters and flags only grow (compared modulo 2^32).  Every
// wait is bounded: a block that waits ~0.2 s records the failure in *status
// and leaves the run, so a launch whose blocks were not all resident ends
// instead of hanging.
constexpr uint64_t kWaitTicks = kResWaitTicks;

// every thread's stores and atomics acknowledged, then the block joins up
__device__ __forceinline__ void res_drain()
{
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
}

// a GR or generator block: one more step done (on counter ctr)
__device__ __forceinline__ void res_arrive(uint32_t *ctr)
{
    res_drain();
    if (threadIdx.x == 0) __hip_atomic_fetch_add(ctr, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// the small block: steps done so far
__device__ __forceinline__ void res_publish(uint32_t *flag, uint32_t v)
{
    res_drain();
    if (threadIdx.x < (uint32_t)kFlagCopies)
        __hip_atomic_store(flag + threadIdx.x * kBarStride, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Block-wide wait until the sum of the N0 words w0[q * kBarStride] reaches t0
// and, when N1 / N2 / N3 is 1, the word w1 reaches t1, w2 reaches t2 and w3
// reaches t3: thread 0 loads every word of a poll at once, so the conditions
// cost one memory round trip together.  BACKOFF: the pause between polls
// doubles from 64 to 1,024 cycles.  0 (and the status word set) when the wait
// took longer than `ticks`; 2 when w3 carries kStopBit above t3; else 1.
template <int N0, int N1, int N2, int N3, bool BACKOFF>
__device__ __forceinline__ int res_waitn(const uint32_t *w0, uint32_t t0, const uint32_t *w1, uint32_t t1,
                                         const uint32_t *w2, uint32_t t2, const uint32_t *w3, uint32_t t3,
                                         uint32_t *status, uint64_t ticks = kWaitTicks)
{
    __shared__ int s_ok;
    if (threadIdx.x == 0) {
        const uint64_t start = __builtin_amdgcn_s_memrealtime();
        uint32_t pause = 0;
        int ok = 1;
        auto ld = [](const uint32_t *p) {
            return __hip_atomic_load(const_cast<uint32_t *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        };
        while (true) {
            uint32_t c[N0 ? N0 : 1], c1 = 0, c2 = 0, c3 = 0;
#pragma unroll
            for (int q = 0; q < N0; q++) c[q] = ld(w0 + q * kBarStride);
            if (N1) c1 = ld(w1);
            if (N2) c2 = ld(w2);
            if (N3) c3 = ld(w3);
            uint32_t sum = 0;
#pragma unroll
            for (int q = 0; q < N0; q++) sum += c[q];
            if ((!N0 || (int32_t)(sum - t0) >= 0) && (!N1 || (int32_t)(c1 - t1) >= 0) &&
                (!N2 || (int32_t)(c2 - t2) >= 0) && (!N3 || (int32_t)(c3 - t3) >= 0)) {
                if (N3 && c3 - t3 >= kStopBit) ok = 2;
                break;
            }
            if (__builtin_amdgcn_s_memrealtime() - start > ticks) {
                __hip_atomic_store(status, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                ok = 0;
                break;
            }
            if (!BACKOFF) {
                __builtin_amdgcn_s_sleep(1);
            } else {
                // s_sleep takes an immediate: 1, 2, 4, 8, 16 x 64 cycles
                switch (pause) {
                case 0: __builtin_amdgcn_s_sleep(1); break;
                case 1: __builtin_amdgcn_s_sleep(2); break;
                case 2: __builtin_amdgcn_s_sleep(4); break;
                case 3: __builtin_amdgcn_s_sleep(8); break;
                default: __builtin_amdgcn_s_sleep(16); break;
                }
                pause++;
            }
        }
        s_ok = ok;
    }
    __syncthreads();
    return s_ok;
}

template <int N0, int N1, int N2, bool BACKOFF>
__device__ __forceinline__ bool res_wait3(const uint32_t *w0, uint32_t t0, const uint32_t *w1, uint32_t t1,
                                          const uint32_t *w2, uint32_t t2, uint32_t *status)
{
    return res_waitn<N0, N1, N2, 0, BACKOFF>(w0, t0, w1, t1, w2, t2, nullptr, 0, status) != 0;
}

// block-wide wait until the sum of the n words w[q * kBarStride] reaches target
template <int N, bool BACKOFF>
__device__ __forceinline__ bool res_wait(const uint32_t *w, uint32_t target, uint32_t *status)
{
    return res_wait3<N, 0, 0, BACKOFF>(w, target, nullptr, 0, nullptr, 0, status);
}

// The PF->PC fold of the chunk sums of one step (as gr_fold), by the 16 waves
// of the generator block: two (zone, PC) tasks per wave at a time, the loads
// of both in flight together, each task summed in gr_fold's order.
__device__ __forceinline__ void res_fold(const float *src, float *dst, uint32_t nTasks, uint32_t numPC, uint32_t nC,
                                         uint32_t perPC, uint32_t tid)
{
    const uint32_t lane = tid & 63u, wave = tid >> 6;
    constexpr uint32_t kWaves = kSmallThreads / 64;
    for (uint32_t t0 = wave; t0 < nTasks; t0 += 2 * kWaves) {  // wave-uniform
        const uint32_t t1 = min(t0 + kWaves, nTasks - 1u);
        const float *c0 = src + (size_t)(t0 / numPC) * nC + (size_t)(t0 % numPC) * perPC;
        const float *c1 = src + (size_t)(t1 / numPC) * nC + (size_t)(t1 % numPC) * perPC;
        float a0 = 0.0f, a1 = 0.0f;
        for (uint32_t q0 = lane; q0 < perPC; q0 += 8 * 64) {
            float x0[8], x1[8];
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const uint32_t q = min(q0 + (uint32_t)j * 64u, perPC - 1u);
                x0[j] = xld<true>(c0 + q);
                x1[j] = xld<true>(c1 + q);
            }
#pragma unroll
            for (int j = 0; j < 8; j++)
                if (q0 + (uint32_t)j * 64u < perPC) {
                    a0 = a0 + x0[j];
                    a1 = a1 + x1[j];
                }
        }
        a0 = wave_sum_lane0(a0);
        a1 = wave_sum_lane0(a1);
        if (lane == 0) {
            xst<true>(dst + t0, a0);
            if (t0 + kWaves < nTasks) xst<true>(dst + t1, a1);
        }
    }
}

// The small populations' state fields of copy-list entries [e0, e1) between
// global memory and the block's LDS (in: into LDS), 16 bytes at a time and
// the tail byte by byte.
__device__ __forceinline__ void res_fields(const SmallArgs &a, uint8_t *lds, bool in, uint32_t tid, uint32_t e0,
                                           uint32_t e1)
{
    for (uint32_t e = e0; e < e1; e++) {
        const int f = (int)a.ldsField[e];
        uint8_t *g = reinterpret_cast<uint8_t *>(a.f[f]);
        uint8_t *l = lds + a.fLds[f];
        const uint32_t bytes = a.ldsBytes[e], n16 = bytes / 16;
        for (uint32_t i = tid; i < n16; i += kSmallThreads) {
            if (in) reinterpret_cast<u32x4 *>(l)[i] = *ga(reinterpret_cast<const u32x4 *>(g) + i);
            else *ga(reinterpret_cast<u32x4 *>(g) + i) = reinterpret_cast<const u32x4 *>(l)[i];
        }
        for (uint32_t i = n16 * 16 + tid; i < bytes; i += kSmallThreads) {
            if (in) l[i] = *ga(g + i);
            else *ga(g + i) = l[i];
        }
    }
}

// One part of the small populations' step in a resident run (PART: the
// input layer on block 0, the microzones on block spans + 1), its state
// fields held in the block's LDS from the first step to the last.
template <int PART, bool LOOP, bool SPILL>
__device__ __forceinline__ void res_part(const SmallArgs &a, const ResRun &r, float *s_err, uint8_t *lds)
{
    constexpr bool IN = PART == kPartIn;
    const uint32_t nZ = a.s.numZones, nSC = a.s.numSC;
    const uint32_t e0 = IN ? 0u : a.nInFields, e1 = IN ? a.nInFields : a.nLdsFields;
    res_fields(a, lds, true, threadIdx.x, e0, e1);
    __syncthreads();
    // the relative error drive of the next step, loaded a step ahead (thread
    // z: zone z; the microzone part's IO reads it)
    auto errOf = [&](uint32_t k) {
        return (!IN && r.err && threadIdx.x < nZ && k < r.n) ? r.err[(size_t)k * nZ + threadIdx.x] : 0.0f;
    };
    float errNext = errOf(0);
    uint32_t kDone = r.n;
    for (uint32_t k = 0; k < r.n; k++) {
        // every GR block done with step k - 1, and for the microzone part the
        // input-layer part (whose SC spikes of step k - 1 it routes) and the
        // generator block (whose fold of step k - 1 it reads); in a closed
        // loop, the step's MF row in place (or the stop)
        const uint32_t *mfw = r.mfFlag + (blockIdx.x % kFlagCopies) * kBarStride;
        const uint64_t ticks = LOOP ? r.waitTicks : kWaitTicks;
        const int w = IN ? res_waitn<kBarGroups, 0, 0, LOOP, false>(r.bar, r.barBase + k * r.spans, nullptr, 0,
                                                                    nullptr, 0, mfw, k + 1, r.status, ticks)
                         : res_waitn<kBarGroups, 1, 1, LOOP, false>(r.bar, r.barBase + k * r.spans, r.flagIn,
                                                                    r.flagInBase + k, r.pb, r.pbBase + k, mfw, k + 1,
                                                                    r.status, ticks);
        if (!w) return;
        if constexpr (LOOP) {
            if (w == 2) {
                kDone = k;
                break;
            }
        }
        const uint64_t t = r.t0 + k + 1;
        SmallStep st;
        st.mf = r.mf + (size_t)k * r.pitch;
        st.t = t;
        st.ioOut = nullptr;
        st.ioPop = 0;
        uint32_t tid = threadIdx.x;
        asm volatile("" : "+v"(tid));  // as in small_step
        if (tid < (uint32_t)kMaxZones) {
            const uint32_t z = tid;
            float e = k == 0 ? r.errDrive0[z] : 0.0f;
            if (z < nZ && errNext != 0.0f) e = errNext * r.maxExtIncVIO;
            if (!IN && LOOP && z < nZ) e = xldg<true>(gp(r.errDev) + (size_t)k * nZ + z);  // (closed loop)
            s_err[z] = e;
        }
        st.gIncGRtoGO = r.gIncGRtoGO;
        st.par = (uint32_t)(t & 1u);
        st.slotPrev = sreg((uint32_t)((t + kRing - 1) % kRing));
        st.slotT2 = sreg((uint32_t)((t + kRing - 2) % kRing));
        st.slotG = sreg((uint32_t)((t + kRing - a.grgoLag) % kRing));
        st.slotNext = sreg((uint32_t)((t + 1) % kRing));
        st.popHist = sreg((t % a.p.tsPerPopHistBinPC) == 0);
        st.resetMFHist = sreg((t % a.p.numTSinMFHist) == 0);
        // the SC spikes of step t (written) and t-1 (read), in three slots:
        // the input-layer part can be a step ahead of the microzone part
        st.apSCout = r.apSC3 + (size_t)sreg((uint32_t)(t % 3)) * nSC;
        st.apSCin = r.apSC3 + (size_t)sreg((uint32_t)((t + 2) % 3)) * nSC;
        st.scFromPrev = k > 0;
        __syncthreads();
        // the argument block through a pointer the compiler cannot follow
        // across steps (its fields are loaded again in each step instead of
        // being held, and spilled, across the loop), in the constant address
        // space: the block does not change during the run, so its fields are
        // read with scalar loads that no store of the step has to wait for
        uint64_t pa = reinterpret_cast<uint64_t>(&a);
        asm volatile("" : "+s"(pa));
        const SmallArgs *ap = (const SmallArgs *)(const __attribute__((address_space(4))) SmallArgs *)pa;
        small_step<true, true, PART, SPILL>(*ap, st, s_err, lds);
        if constexpr (IN) {
            res_publish(r.flagIn, r.flagInBase + k + 1);
        } else {
            errNext = errOf(k + 1);
            // (closed loop) the step's spikes of the requested population to
            // the host, four to a word, through to system memory; the drain of
            // the publish orders them before the step count the host polls
            if constexpr (LOOP) {
                const uint32_t pops[4] = {a.lds.apPC, a.lds.apBC, a.lds.apIO, a.lds.apNC};
                const uint32_t pop = r.ioPop == CBM_REC_PC ? 0u : r.ioPop == CBM_REC_BC ? 1u : r.ioPop == CBM_REC_IO ? 2u : 3u;
                const uint8_t *src = lds + pops[pop];
                uint32_t *dst = r.hostOut + (size_t)(k & 1u) * r.outWords;
                for (uint32_t i = threadIdx.x; i < r.outWords; i += kSmallThreads) {
                    const uint32_t v = (uint32_t)src[4 * i] | (uint32_t)src[4 * i + 1] << 8 |
                                       (uint32_t)src[4 * i + 2] << 16 | (uint32_t)src[4 * i + 3] << 24;
                    __hip_atomic_store(dst + i, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                }
            }
            res_publish(r.flag, r.flagBase + k + 1);
            if (LOOP && threadIdx.x == 0)
                __hip_atomic_store(r.hostOutSeq, k + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        }
    }
    if constexpr (!IN) {
        // the SC->PC routing of the last step (MZone::updateSCPCOut), which
        // the microzone part does at the start of the next step
        if (kDone > 0) {
            if (!res_wait<1, false>(r.flagIn, r.flagInBase + kDone, r.status)) return;
            const G<uint8_t> *sc = gp(r.apSC3) + (size_t)((r.t0 + kDone) % 3) * nSC;
            uint8_t *s_apSC = lds + a.lds.apSC;
            for (uint32_t i = threadIdx.x; i < nSC; i += kSmallThreads) s_apSC[i] = xldg<true>(sc + i);
            __syncthreads();
            auto inSC = SF<SPILL ? 2 : 1, uint8_t>(a, CBM_F_INPUT_SC_PC, lds);
            const uint16_t *s_idx = at<uint16_t>(lds, a.lds.idx);
            for (uint32_t q = threadIdx.x; q < nZ * a.s.numPC * a.s.numpPCfromSCtoPC; q += kSmallThreads)
                inSC[q] = s_apSC[s_idx[a.lds.iSCPC + q]];
        }
    }
    __syncthreads();
    res_fields(a, lds, false, threadIdx.x, e0, e1);
}

// The generators' next blocks and the PF->PC fold in a resident run (the
// last block).
template <bool LOOP>
__device__ __forceinline__ void res_prebuild(const SmallArgs &a, const ResRun &r, uint8_t *lds)
{
    const uint32_t nZ = a.s.numZones, nC = r.g.numGR >> 6;
    for (uint32_t k = 0; k < r.n; k++) {
        if constexpr (LOOP) {
            // a closed loop: the host's step k (or its stop), copied from
            // pinned memory to the device and published to the other blocks
            __shared__ uint32_t s_v;
            if (threadIdx.x == 0) {
                const uint64_t start = __builtin_amdgcn_s_memrealtime();
                uint32_t v;
                while (true) {
                    v = __hip_atomic_load(const_cast<uint32_t *>(r.hostSeq), __ATOMIC_RELAXED,
                                          __HIP_MEMORY_SCOPE_SYSTEM);
                    if ((v & kHostStop) || (int32_t)(v - (k + 1)) >= 0) break;
                    if (__builtin_amdgcn_s_memrealtime() - start > r.waitTicks) {
                        __hip_atomic_store(r.status, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        v = 0xFFFFFFFFu;
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
                s_v = v;
            }
            __syncthreads();
            const uint32_t v = s_v;
            if (v == 0xFFFFFFFFu) return;
            if (v & kHostStop) {
                res_publish(r.mfFlag, k + 1 + kStopBit);
                break;
            }
            const uint32_t *in = reinterpret_cast<const uint32_t *>(r.hostIn + (size_t)(k & 1u) * r.inPitch);
            uint32_t *row = reinterpret_cast<uint32_t *>(const_cast<uint8_t *>(r.mf) + (size_t)k * r.pitch);
            const uint32_t nW = a.s.numMF / 4;
            for (uint32_t i = threadIdx.x; i < nW + nZ; i += kSmallThreads) {
                const uint32_t x = __hip_atomic_load(const_cast<uint32_t *>(i < nW ? in + i : in + r.inErr / 4 + (i - nW)),
                                                     __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                if (i < nW) xstg<true>(gp(row) + i, x);
                else xstg<true>(gp(reinterpret_cast<uint32_t *>(r.errDev)) + (size_t)k * nZ + (i - nW), x);
            }
            res_publish(r.mfFlag, k + 1);
        }
        // the microzone part done with step k - 1 (the generator state), every
        // GR block done with step k - 1 (the chunk sums)
        if (!res_wait3<kBarGroups, 1, 0, true>(r.bar, r.barBase + k * r.spans,
                                                r.flag + (blockIdx.x % kFlagCopies) * kBarStride, r.flagBase + k,
                                                nullptr, 0, r.status))
            return;
        const uint64_t t = r.t0 + k + 1;
        uint32_t tid = threadIdx.x;
        asm volatile("" : "+v"(tid));  // as in small_step
        // the chunk sums of step t-1 into their ring slot (read at step t+1)
        if (!(k == 0 && r.skipFold0))
            res_fold(a.chunk2 + (size_t)((t - 1) & 1u) * nZ * nC,
                     a.pfpcRing + (size_t)sreg((uint32_t)((t + kRing - 1) % kRing)) * nZ * a.s.numPC, nZ * a.s.numPC,
                     a.s.numPC, nC, r.g.perPC, tid);
        SmallStep st;
        st.par = (uint32_t)(t & 1u);
        sfmt_prebuild<true>(a, st, reinterpret_cast<uint32_t *>(lds));
        res_publish(r.pb, r.pbBase + k + 1);
    }
}

// The GR cells of span `rank` in a resident run (blocks 1 .. spans).
template <int CPT, bool LOOP>
__device__ __forceinline__ void res_gr(const SmallArgs &a, const ResRun &r, const uint32_t rank, uint8_t *lds)
{
    const GrArgs &g = r.g;
    const uint32_t nZ = g.numZones, nC = g.numGR >> 6, tid = threadIdx.x;
    const uint32_t perHist = a.p.tsPerHistBinGR;
    const uint32_t c0 = (uint32_t)((uint64_t)rank * nC / r.spans), c1 = (uint32_t)((uint64_t)(rank + 1) * nC / r.spans);
    const uint32_t first = c0 << 6, count = (c1 - c0) << 6;
    uint32_t *smem = reinterpret_cast<uint32_t *>(lds);
    GrRegs<CPT> R;
    uint4 *recL = reinterpret_cast<uint4 *>(lds + r.recLds);
    R.rec = recL;
    if (count) {
#pragma unroll
        for (int j = 0; j < CPT; j++) {
            const uint32_t gr = first + min((uint32_t)j * 1024u + tid, count - 1u);
            R.v[j] = g.v[gr];
            R.th[j] = g.thr[gr];
            R.gk[j] = g.gKCa[gr];
            R.gE[j] = g.gE[gr];
            R.gI[j] = g.gI[gr];
            R.ab[j] = g.apBuf[gr];
            R.w0[j] = g.w[gr];
            recL[j * 1024 + tid] = g.conn[gr];
        }
    }
    for (uint32_t k = 0; k < r.n; k++) {
        // both parts of the small populations done with step k - 1
        const uint32_t cp = (blockIdx.x % kFlagCopies) * kBarStride;
        // (and the generator block done with step k - 1: its fold read the
        // chunk sums this step overwrites; in a closed loop, the step's MF
        // row in place, or the stop)
        const int w = res_waitn<1, 1, 1, LOOP, true>(r.flagIn + cp, r.flagInBase + k, r.flag + cp, r.flagBase + k,
                                                     r.pb + cp, r.pbBase + k, r.mfFlag + cp, k + 1, r.status,
                                                     LOOP ? r.waitTicks : kWaitTicks);
        if (!w) return;
        if constexpr (LOOP)
            if (w == 2) break;
        const uint64_t t = r.t0 + k + 1;
        const uint32_t par = (uint32_t)(t & 1u), parPrev = par ^ 1u;
        const uint32_t slot = (uint32_t)(t % kRing);
        float *chunk2 = const_cast<float *>(a.chunk2);
        // (uniform values in scalar registers: see sptr)
        GrStep gs;
        gs.mf = (sptr(r.mf + (size_t)k * r.pitch));
        gs.apGOprev = (sptr(a.apGO2 + (size_t)parPrev * g.numGO));
        gs.chunk = (sptr(chunk2 + (size_t)par * nZ * nC));
        gs.foldSrc = nullptr;  // the generator block folds (res_prebuild)
        gs.foldDst = nullptr;
        gs.bcCnt = (sptr(a.bcRing + (size_t)slot * a.s.numBC));
        gs.scCnt = (sptr(a.scRing + (size_t)slot * a.s.numSC));
        gs.goCnt = (sptr(a.goRing + (size_t)slot * g.numGO));
        gs.plastTimer = (sptr(a.timer2 + (size_t)parPrev * nZ * g.numIO));
        gs.doPlast = sreg((t % perHist) == 0);
        const uint64_t upd = t / perHist;
        gs.histOld = (sptr(g.hist + (size_t)histPlane(upd, 40) * nC));
        gs.histNew = (sptr(g.hist + (size_t)histPlane(upd, 0) * nC));
        gs.pfCnt = (sptr((a.rec && r.recOffPF != 0xFFFFFFFFu)
                               ? a.rec + (size_t)(t % a.recCap) * a.recRow + r.recOffPF
                               : nullptr));
        if (count)
            gr_span_impl<CPT, kSmallThreads, false, true, true, true>(g, gs, first, count, rank, r.spans, smem, &R);
        res_arrive(r.bar + (blockIdx.x & (kBarGroups - 1)) * kBarStride);
    }
    // the cells' state back into the per-GR arrays
#pragma unroll
    for (int j = 0; j < CPT; j++) {
        if ((uint32_t)j * 1024u + (tid & ~63u) >= count) continue;  // wave-uniform
        const uint32_t gr = first + (uint32_t)j * 1024u + tid;
        g.v[gr] = R.v[j];
        g.thr[gr] = R.th[j];
        g.gKCa[gr] = R.gk[j];
        g.gE[gr] = R.gE[j];
        g.gI[gr] = R.gI[j];
        g.apBuf[gr] = R.ab[j];
        g.w[gr] = R.w0[j];
    }
}

// n steps of one network in one launch of spans + 3 blocks of 1024 threads,
// one per CU, all on the device at once (the host launches resident runs of a
// device one after another and asks for LDS that no two of them fit in): block 0 steps the input-layer small populations
// (GO, SC), block spans + 1 the microzones (PC, BC, IO, NC), each holding its
// populations' state in LDS, blocks 1 .. spans the GR cells of their even
// split (chunks [r*C/spans, (r+1)*C/spans) of the C 64-GR chunks, as
// step_span_kernel), whose state stays in registers (GrRegs) from the first
// step to the last and is stored back at the end, and block spans + 2 folds
// the PF->PC chunk sums and prebuilds the zone generators' next blocks.
// Each step does the work of one step_span_kernel launch, and the waits on
// step counters and flags stand in for the launch boundary.  CPT: cells per
// thread (every span holds at most 1024 * CPT cells); compact connectivity
// only.
template <int CPT, bool LOOP, bool SPILL>
__global__ void __launch_bounds__(kSmallThreads) step_resident_kernel(const SmallArgs *__restrict__ args, const ResRun r)
{
    extern __shared__ __align__(16) uint8_t lds[];
    __shared__ float s_err[kMaxZones];  // in LDS: the small step indexes it by zone
    if (blockIdx.x == 0)
        res_part<kPartIn, LOOP, false>(*args, r, s_err, lds);
    else if (blockIdx.x == r.spans + 1)
        res_part<kPartMz, LOOP, SPILL>(*args, r, s_err, lds);
    else if (blockIdx.x > r.spans)
        res_prebuild<LOOP>(*args, r, lds);
    else
        res_gr<CPT, LOOP>(*args, r, blockIdx.x - 1, lds);
}

#define CBM_RESIDENT(CPT, SPILL)                                                                     \
    template __global__ void step_resident_kernel<CPT, false, SPILL>(const SmallArgs *, const ResRun); \
    template __global__ void step_resident_kernel<CPT, true, SPILL>(const SmallArgs *, const ResRun);
CBM_RESIDENT(2, false)
CBM_RESIDENT(3, false)
CBM_RESIDENT(5, false)
CBM_RESIDENT(5, true)  // (1M-GR networks of many MFs: the cart-pole network)

// ---------------------------------------------------------------- streaming ceiling
// A per-cell pass with the GR step's streams and no arithmetic to speak of:
// every cell reads nRead 4-byte values and a 16-byte record, and writes the
// first nWrite values back (the 4-byte streams non-temporal).  Block b covers cells
// [b * 4096, b * 4096 + 4096), four per thread, 1024 cells apart.
__global__ void __launch_bounds__(1024) stream_ceiling_kernel(StreamArgs a)
{
    constexpr int kCpt = 4;
    uint32_t x[kCpt][kMaxStreamWords];
    uint4 r[kCpt];
    const uint64_t first = (uint64_t)blockIdx.x * (1024u * kCpt) + threadIdx.x;
#pragma unroll
    for (int j = 0; j < kCpt; j++) {
        const uint64_t g = first + (uint64_t)j * 1024u;
        const bool in = g < a.cells;
#pragma unroll
        for (int w = 0; w < kMaxStreamWords; w++)
            x[j][w] = in && w < a.nRead ? __builtin_nontemporal_load(a.word[w] + g) : 0u;
        r[j] = in ? a.rec[g] : make_uint4(0, 0, 0, 0);
    }
#pragma unroll
    for (int j = 0; j < kCpt; j++) {
        const uint64_t g = first + (uint64_t)j * 1024u;
        if (g >= a.cells) continue;
        const uint32_t k = r[j].x ^ r[j].y ^ r[j].z ^ r[j].w;
#pragma unroll
        for (int w = 0; w < kMaxStreamWords; w++)
            if (w < a.nWrite) __builtin_nontemporal_store(x[j][w] + x[j][(w + 1) % kMaxStreamWords] + k, a.word[w] + g);
    }
}

// ---------------------------------------------------------------- batch

// The small populations of network d at run step kk (block-wide).
__device__ __forceinline__ void batch_small(const BatchNet &d, uint32_t kk, float *s_err, uint8_t *lds)
{
    const uint64_t t = sreg64(d.t0 + kk + 1);
    const uint32_t par = (uint32_t)(t & 1u);
    const uint32_t nZ = d.gr.numZones;
    SmallStep st;
    st.mf = d.mf + (size_t)kk * d.pitch;
    st.t = t;
    st.ioOut = nullptr;
    st.ioPop = 0;
    st.apSCout = nullptr;  // (the whole step on one block)
    st.apSCin = nullptr;
    st.scFromPrev = 0;
    if (threadIdx.x < (uint32_t)kMaxZones) {
        const uint32_t z = threadIdx.x;
        float e = kk == 0 ? d.errDrive0[z] : 0.0f;
        if (d.err && z < nZ) {
            const float r = d.err[(size_t)kk * nZ + z];
            if (r != 0.0f) e = r * d.maxExtIncVIO;
        }
        s_err[z] = e;
    }
    st.gIncGRtoGO = d.gIncGRtoGO;
    st.par = par;
    st.slotPrev = sreg((uint32_t)((t + kRing - 1) % kRing));
    st.slotT2 = sreg((uint32_t)((t + kRing - 2) % kRing));
    st.slotG = sreg((uint32_t)((t + kRing - d.grgoLag) % kRing));
    st.slotNext = sreg((uint32_t)((t + 1) % kRing));
    st.popHist = sreg((t % d.tsPerPopHistBinPC) == 0);
    st.resetMFHist = sreg((t % d.numTSinMFHist) == 0);
    __syncthreads();
    // d.small points into device global memory: taken through a scalar
    // register as a global pointer, the argument block is read with global
    // and scalar loads as a kernel argument is, not with flat loads whose
    // results held 22 vector registers in scratch memory
    const SmallArgs *sa = (const SmallArgs *)(const G<SmallArgs> *)sreg64((uint64_t)d.small);
    small_step(*sa, st, s_err, lds);
}

// One launch per step for a pack of networks: block n < nNets steps network
// n's small populations, the other blocks are GR tiles (network, tile) from
// the tile table.  Run step kk = *kbase + k (kbase: device counter advanced by
// a captured graph, or nullptr).
template <int CPT, int FMT>
__global__ void __launch_bounds__(kSmallThreads) batch_step_kernel(const BatchNet *__restrict__ nets,
                                                                   const uint2 *__restrict__ tiles, uint32_t nNets,
                                                                   const uint32_t *kbase, uint32_t k)
{
    extern __shared__ __align__(16) uint8_t lds[];
    __shared__ float s_err[kMaxZones];  // in LDS: the small step indexes it by zone
    const uint32_t kk = (kbase ? *kbase : 0u) + k;
    const bool small = blockIdx.x < nNets;
    const uint2 tt = small ? make_uint2(blockIdx.x, 0u) : tiles[blockIdx.x - nNets];
    const BatchNet &d = nets[tt.x];
    const uint64_t t = d.t0 + kk + 1;
    const uint32_t par = (uint32_t)(t & 1u), parPrev = par ^ 1u;
    const uint8_t *mf = d.mf + (size_t)kk * d.pitch;
    const uint32_t nZ = d.gr.numZones;
    if (small) {
        batch_small(d, kk, s_err, lds);
    } else {
        const GrArgs &g = d.gr;
        const uint32_t slot = (uint32_t)(t % kRing);
        const size_t nChunks = g.numGR / 64;
        GrStep gs;
        gs.mf = mf;
        gs.apGOprev = d.apGO2 + (size_t)parPrev * g.numGO;
        gs.chunk = d.chunk2 + (size_t)par * nZ * nChunks;
        gs.foldSrc = (kk == 0 && d.skipFold0) ? nullptr : d.chunk2 + (size_t)parPrev * nZ * nChunks;
        gs.foldDst = d.pfpcRing + (size_t)((t + kRing - 1) % kRing) * nZ * g.numPC;
        gs.bcCnt = d.bcRing + (size_t)slot * d.small->s.numBC;
        gs.scCnt = d.scRing + (size_t)slot * d.small->s.numSC;
        gs.goCnt = d.goRing + (size_t)slot * g.numGO;
        gs.plastTimer = d.timer2 + (size_t)parPrev * nZ * g.numIO;
        gs.doPlast = (t % d.tsPerHistBinGR) == 0;
        const uint64_t upd = t / d.tsPerHistBinGR;
        gs.histOld = g.hist + (size_t)histPlane(upd, 40) * nChunks;
        gs.histNew = g.hist + (size_t)histPlane(upd, 0) * nChunks;
        gs.pfCnt = nullptr;
        gr_tile<CPT, kSmallThreads, FMT>(g, gs, tt.y, reinterpret_cast<uint32_t *>(lds));
    }
}

// FMT: every network of the pack with compact records, every one with the
// five arrays, or a mix (decided per tile)
#define CBM_BATCH(CPT, FMT) \
    template __global__ void batch_step_kernel<CPT, FMT>(const BatchNet *, const uint2 *, uint32_t, const uint32_t *, uint32_t);
CBM_BATCH(1, kFmtCompact)
CBM_BATCH(2, kFmtCompact)
CBM_BATCH(1, kFmtArrays)
CBM_BATCH(2, kFmtArrays)
CBM_BATCH(1, kFmtAny)
CBM_BATCH(2, kFmtAny)
#undef CBM_BATCH

// advances the run-step counter of a captured batch graph
__global__ void batch_advance_kernel(uint32_t *kbase, uint32_t steps)
{
    if (threadIdx.x == 0) *kbase += steps;
}

// ---------------------------------------------------------------- surrogate sweep
// The input layer of a network replaced by cluster probabilities: the
// cluster tile below takes the place of the GR tiles, and the network's small
// populations step exactly as above.  prob row kk holds, for every cluster of
// `cells` consecutive GRs, the probability that one of its cells emits a
// delayed PF spike at run step kk; rows kk-199 .. kk-1 must be readable.
// The tile writes what the GR tiles would write into this step's ring
// slots, in expectation:
//   PF->PC sum of PC j      = sum over its clusters of cells * W[z][c] * p[c]
//   PF->BC / PF->SC counts  = nearest integer to sum of cells * p[c]
//   GR->GO counts           = 0 (Golgi cells only feed the replaced layer)
// and applies the PF->PC rule of the GR tiles (history bit 39: a spike in
// the 5-step bin that ends 195 steps before) to the cluster weight with the
// probability of such a spike, 1 - prod_{j=195..199} (1 - p[kk-j][c]):
//   W = clamp(W + q * step(z, IO of c), 0, 1)   every tsPerHistBinGR steps.
// Sums run in a fixed order, so a run is deterministic.
// out(i, sum of src[i*len .. i*len+len-1]) for i < nSeg, len a power of two;
// every sum in the same fixed order (lanes stride over a segment, then an
// xor butterfly), so results do not depend on timing.
template <typename Out>
__device__ __forceinline__ void segment_sums(const float *src, uint32_t nSeg, uint32_t len, Out out)
{
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6, waves = blockDim.x >> 6;
    if (len >= 64) {
        for (uint32_t i = wave; i < nSeg; i += waves) {
            float a = 0.0f;
            for (uint32_t e = lane; e < len; e += 64) a += src[(size_t)i * len + e];
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) a += __shfl_xor(a, o);
            if (lane == 0) out(i, a);
        }
    } else {
        const uint32_t per = 64 / len;  // segments per wave
        for (uint32_t i0 = wave * per; i0 < nSeg; i0 += waves * per) {
            const uint32_t i = i0 + lane / len;
            float a = i < nSeg ? src[(size_t)i * len + (lane % len)] : 0.0f;
            for (uint32_t o = len >> 1; o > 0; o >>= 1) a += __shfl_xor(a, (int)o);
            if (lane % len == 0 && i < nSeg) out(i, a);
        }
    }
}

__device__ __forceinline__ void cluster_tile(const BatchNet &d, const SweepNet &sw, const float *prob,
                                          const float *Q, uint64_t pitch, uint32_t kk, uint8_t *lds)
{
    constexpr int kPer = 4;  // clusters per thread and group
    const GrArgs &g = d.gr;
    const uint64_t t = d.t0 + kk + 1;
    const uint32_t slot = (uint32_t)(t % kRing), parPrev = (uint32_t)(t & 1u) ^ 1u;
    const uint32_t nZ = g.numZones, nCl = sw.nCl;
    const float cells = (float)(1u << sw.log2Cells);
    const bool doPlast = (t % d.tsPerHistBinGR) == 0;
    float *s_step = reinterpret_cast<float *>(lds);  // [nZ * numIO]
    float *s_cnt = s_step + nZ * g.numIO;             // [nCl]
    float *s_pc = s_cnt + nCl;                        // [nZ][nCl]
    const uint32_t tid = threadIdx.x, nt = blockDim.x;
    if (doPlast)
        for (uint32_t i = tid; i < nZ * g.numIO; i += nt) {
            const uint32_t z = i / g.numIO;
            const int32_t tm = d.timer2[(size_t)parPrev * nZ * g.numIO + i];
            s_step[i] = tm < g.ltdEnd ? g.ltd[z] : (tm > g.ltpStart ? g.ltp[z] : 0.0f);
        }
    __syncthreads();
    // this thread's clusters c = c0 + tid + i*nt, four at a time: every load
    // of a group before any store
    const float *p = prob + (size_t)kk * pitch, *q = Q + (size_t)kk * pitch;
    for (uint32_t c0 = 0; c0 < nCl; c0 += kPer * nt) {
        float pv[kPer], qv[kPer];
#pragma unroll
        for (int i = 0; i < kPer; i++) {
            const uint32_t c = c0 + tid + i * nt;
            const bool in = c < nCl;
            pv[i] = in ? p[c] : 0.0f;
            qv[i] = in && doPlast ? q[c] : 0.0f;
        }
        for (uint32_t z = 0; z < nZ; z++) {
            float wv[kPer];
#pragma unroll
            for (int i = 0; i < kPer; i++) {
                const uint32_t c = c0 + tid + i * nt;
                wv[i] = c < nCl ? sw.W[(size_t)z * nCl + c] : 0.0f;
            }
#pragma unroll
            for (int i = 0; i < kPer; i++) {
                const uint32_t c = c0 + tid + i * nt;
                if (c >= nCl) continue;
                const float cnt = pv[i] * cells;
                float w = wv[i];
                if (doPlast) {
                    const uint32_t io = (c << sw.log2Cells) >> g.log2GRperIO;
                    w = w + qv[i] * s_step[z * g.numIO + io];
                    w = (w > 0) * w;
                    w = (w > 1) + (w <= 1) * w;
                    sw.W[(size_t)z * nCl + c] = w;
                }
                if (z == 0) s_cnt[c] = cnt;
                s_pc[(size_t)z * nCl + c] = w * cnt;
            }
        }
    }
    __syncthreads();
    const uint32_t numBC = d.small->s.numBC, numSC = d.small->s.numSC;
    float *pfpc = d.pfpcRing + (size_t)slot * nZ * g.numPC;
    segment_sums(s_pc, nZ * g.numPC, nCl / g.numPC, [&](uint32_t i, float a) { pfpc[i] = a; });
    uint32_t *bc = d.bcRing + (size_t)slot * numBC;
    segment_sums(s_cnt, numBC, nCl / numBC, [&](uint32_t i, float a) { bc[i] = (uint32_t)rintf(a); });
    uint32_t *sc = d.scRing + (size_t)slot * numSC;
    segment_sums(s_cnt, numSC, nCl / numSC, [&](uint32_t i, float a) { sc[i] = (uint32_t)rintf(a); });
}

// Q[k][c] = 1 - prod_{j=195..199} (1 - prob[k-j][c]): the probability of a
// PF spike in the 5-step history bin the plasticity rule of step k reads,
// shared by every network of a sweep (they share prob).
__global__ void sweep_q_kernel(const float *prob, uint64_t pitch, uint32_t n, uint32_t nCl, float *Q)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)n * nCl) return;
    const uint32_t k = (uint32_t)(i / nCl), c = (uint32_t)(i % nCl);
    const float *p = prob + (size_t)k * pitch + c;
    float keep = 1.0f;
#pragma unroll
    for (int j = 195; j <= 199; j++) keep *= 1.0f - p[-(int64_t)j * (int64_t)pitch];
    Q[(size_t)k * pitch + c] = 1.0f - keep;
}

// One step of every sweep network: block n < nNets steps network n's small
// populations, block nNets + n is its cluster tile (the two are independent
// within a step, as the GR tiles and the small step of the exact step are).
// Run step kk = *kbase + k (kbase: counter advanced by a captured graph, or
// nullptr); the run's probability rows come from the device block `run`, so
// one captured graph serves every run.
__global__ void __launch_bounds__(kSmallThreads) sweep_step_kernel(const BatchNet *__restrict__ nets,
                                                                   const SweepNet *__restrict__ sw, uint32_t nNets,
                                                                   const SweepRun *run, const uint32_t *kbase,
                                                                   uint32_t k)
{
    extern __shared__ __align__(16) uint8_t lds[];
    __shared__ float s_err[kMaxZones];
    const uint32_t kk = (kbase ? *kbase : 0u) + k;
    if (blockIdx.x < nNets)
        batch_small(nets[blockIdx.x], kk, s_err, lds);
    else
        cluster_tile(nets[blockIdx.x - nNets], sw[blockIdx.x - nNets], run->prob, run->Q, run->pitch, kk, lds);
}

// W[z][c] = mean of the PF->PC weights of cluster c's cells, in cell order
__global__ void cluster_mean_kernel(const float *w, uint32_t numGR, uint32_t nZ, uint32_t log2Cells, float *W)
{
    const uint32_t nCl = numGR >> log2Cells;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nZ * nCl) return;
    const uint32_t z = i / nCl, c = i % nCl, cells = 1u << log2Cells;
    const float *src = w + (size_t)z * numGR + (size_t)c * cells;
    float acc = 0.0f;
    for (uint32_t j = 0; j < cells; j++) acc += src[j];
    W[i] = acc / (float)cells;
}

}  // namespace cbmdev

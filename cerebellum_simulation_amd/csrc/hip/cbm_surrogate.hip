// Learned surrogate of the granule-Golgi input layer: causal dilated temporal
// convolutions on the CDNA4 matrix cores (v_mfma_f32_16x16x32_bf16), BF16
// operands, FP32 accumulation, fused epilogues.  The public interface and the
// model's definition are in include/cbm_surrogate.h.
//
// Layout: every activation is time-major, one row per ms and one column per
// channel, so a convolution tap is a row offset.  A batch is a list of
// segments; segment s of the batch occupies rows s*P .. s*P+P-1 (P = lead +
// outputs), and the MF rows it reads start at segStart[s] of the caller's MF
// stream.  A row offset that leaves the segment reads zero.
//
// Kernels
//   conv_kernel<ASRC, EPI>  C[r][n] = sum_tap sum_c A[r + tap*shift][c] * Bt[n][tap*cin + c]
//       forward (shift = -dil, Bt = W[o][tap][c]) and data gradient
//       (shift = +dil, A = dZ, Bt = W^T[c][tap][o]); 128x128 tile per
//       256-thread block, 64-deep K steps through double-buffered LDS,
//       each wave 64x64 = 4x4 MFMA tiles.  Epilogues: bias+ReLU -> BF16,
//       bias+sigmoid -> FP32 probabilities, bias+sigmoid+cross-entropy ->
//       BF16 dZ and the loss, ReLU mask -> BF16.
//   wgrad_kernel<SSRC>  partial[split][o][tap*cin + c] = sum_{r in split} G[r][o] * S[r - tap*dil][c]
//       the reduction runs over time rows, which are strided in both
//       operands: tiles are transposed on their way into LDS.
//   colsum_kernel, reduce_kernel, adam_kernel, targets_kernel.

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "cbm_handle.hpp"
#include "cbm_surrogate.h"

namespace sg {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef uint16_t u16;

#ifndef SG_BK
#define SG_BK 32
#endif
// 128 x 128 output tile per 256-thread block, K steps of BK through
// double-buffered LDS (row stride BK + 8 elements)
constexpr int BM = 128, BN = 128, BK = SG_BK, LDK = BK + 8;
constexpr int kThreads = 256;
static_assert(BK == 32 || BK == 64, "K step");
// conv_kernel's operand tiles: unpadded rows of BK elements (64 B at BK = 32)
// whose 16-byte chunks are XOR-swizzled by row, so that neither the loader's
// stores (16 lanes: 4 rows x 4 chunks) nor the MFMA fragment reads (16 lanes:
// 16 consecutive rows, one chunk) meet in a bank, for 32 or 64 LDS banks:
// within an aligned group of 16 rows, the 4 rows of equal r mod 4 get 4
// distinct chunk permutations, and so do the 4 rows of equal parity within 8.
// The pattern repeats every 16 rows, so a wave's fragments 16 rows apart
// share their swizzle.
__device__ __forceinline__ int swz(int r, int c)
{
    if constexpr (BK == 32)
        return c ^ ((((r & 15) >> 1) ^ ((r & 15) >> 3)) & 3);
    else
        return c;
}
constexpr int LDC = BK == 32 ? BK : LDK;  // conv_kernel's operand row stride
// operand loader of conv_kernel: 16-byte vectors, kVec per tile row; a pass of
// the block covers kRowStep rows, kPass passes cover a 128-row tile
constexpr int kVec = BK / 8, kRowStep = kThreads / kVec, kPass = BM / kRowStep;

enum { SRC_BF16 = 0, SRC_SPIKE = 1 };
enum { EPI_RELU = 0, EPI_SIGMOID = 1, EPI_BCE = 2, EPI_MASK = 3 };

__host__ __device__ inline u16 f2bf(float x)
{
    uint32_t u;
    memcpy(&u, &x, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) return (u16)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
    return (u16)(u >> 16);
}
__host__ __device__ inline float bf2f(u16 h)
{
    const uint32_t u = (uint32_t)h << 16;
    float x;
    memcpy(&x, &u, 4);
    return x;
}

// 8 spike bytes -> 8 BF16 values (1.0 for a nonzero byte)
__device__ inline uint4 spikes_to_bf16(uint2 v)
{
    uint32_t w[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const uint32_t src = i < 2 ? v.x : v.y;
        const uint32_t b0 = (src >> (16 * (i & 1))) & 0xFFu, b1 = (src >> (16 * (i & 1) + 8)) & 0xFFu;
        w[i] = (b0 ? 0x3F80u : 0u) | ((b1 ? 0x3F80u : 0u) << 16);
    }
    return make_uint4(w[0], w[1], w[2], w[3]);
}

struct ConvArgs {
    const void *A;          // time-major source
    size_t lda;             // elements (BF16) or bytes (spikes) per source row
    const int64_t *seg;     // first source row of each segment (nullptr: s*P)
    int P;                  // batch rows per segment
    int shift;              // source row offset per tap
    int cin;                // channels per tap; multiple of BK
    int K;                  // taps*cin
    const u16 *Bt;          // [N][K]
    int M, N;               // batch rows, output channels
    const float *bias;      // [N]
    u16 *out;               // BF16 [M][ldo] (RELU, BCE, MASK)
    size_t ldo;
    float *prob;            // FP32 outputs of valid rows only (SIGMOID): [(s*(P-lead) + pos-lead)][ldp]
    size_t ldp;
    const u16 *mask;        // forward activation whose sign masks the gradient (MASK)
    size_t ldm;
    const u16 *Y;           // targets, rows indexed like the MF stream (BCE)
    size_t ldy;
    const int64_t *segY;    // first target row of each segment (BCE)
    int lead;               // rows of a segment without a full receptive field
    float scale;            // 1 / number of outputs (BCE)
    float *loss;
    float *part;            // split-K: FP32 partial sums [split][M][N] instead of the epilogue
    int kPerSplit;          // K steps (of BK) per split
    const u16 *zeros;       // 16 zero bytes: the source of rows outside their segment (conv_dma_kernel)
};

// The epilogue of a 128 x 128 tile held as each wave's 4 x 4 MFMA
// accumulators (wave rows wm.., columns wn..): split-K partial sums, or the
// fused epilogue EPI through the block's shared buffer (SMEM_BYTES), which
// the caller has finished reading.
template <int EPI, int SMEM_BYTES>
__device__ __forceinline__ void conv_epilogue(const ConvArgs &a, const f32x4 (&acc)[4][4], u16 *smem, int m0, int n0,
                                              int wm, int wn)
{
    const int tid = threadIdx.x, lane = tid & 63;
    if (a.part) {  // split-K: this split's partial sums
        float *dst = a.part + (size_t)blockIdx.z * a.M * a.N;
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int r = m0 + wm + i * 16 + (lane >> 4) * 4 + v;
                if (r >= a.M) continue;
#pragma unroll
                for (int j = 0; j < 4; j++) dst[(size_t)r * a.N + n0 + wn + j * 16 + (lane & 15)] = acc[i][j][v];
            }
        return;
    }
    // Through LDS: a lane holds column (lane & 15) and rows
    // 4*(lane >> 4) + v of each 16x16 tile; the tile is staged and then
    // written out row by row with the lanes on consecutive columns, so every
    // global access of the epilogue is a contiguous run per wave.  When the
    // operand buffers cannot hold the whole 128-row tile (BK = 32), it goes
    // through in two halves of 64 rows (the waves with wm == 64 * half stage
    // theirs).
    constexpr int LDT = BN + 4;  // floats per staged row
    constexpr int HM = BM * LDT * 4 + BM * 8 <= SMEM_BYTES ? BM : BM / 2;
    static_assert(HM * LDT * 4 + BM * 8 <= SMEM_BYTES, "staged tile must fit the shared buffer");
    float *tile = reinterpret_cast<float *>(smem);
    int *rowSeg = reinterpret_cast<int *>(tile + HM * LDT);  // segment of each tile row
    int *rowPos = rowSeg + BM;                               // position in it, -1 past M
    if (tid < BM) {
        const int r = m0 + tid;
        const int sg = r < a.M ? r / a.P : 0;
        rowSeg[tid] = sg;
        rowPos[tid] = r < a.M ? r - sg * a.P : -1;
    }
    // a thread takes 4 consecutive columns of every 8th row: a wave covers two
    // whole tile rows per instruction
    const int c4 = (tid & 31) * 4, n = n0 + c4;
    float bn[4] = {0.f, 0.f, 0.f, 0.f};
    if (EPI != EPI_MASK) *(float4 *)bn = *(const float4 *)&a.bias[n];
    float lsum = 0.f;
#pragma unroll
    for (int half = 0; half < BM / HM; half++) {
        // BCE: the target words of this part's rows, requested before any of
        // them is used (one round trip for the part instead of one per row;
        // rows without a target load nothing)
        uint2 ypre[EPI == EPI_BCE ? HM / 8 : 1];
        if constexpr (EPI == EPI_BCE) {
#pragma unroll
            for (int it = 0; it < HM / 8; it++) {
                const int r = m0 + half * HM + it * 8 + (tid >> 5);
                const int sg = r < a.M ? r / a.P : 0, pos = r < a.M ? r - sg * a.P : -1;
                const bool use = pos >= a.lead;
                const int64_t yrow = use ? a.segY[sg] + pos : 0;
                ypre[it] = use ? *(const uint2 *)&a.Y[(size_t)yrow * a.ldy + n] : make_uint2(0u, 0u);
            }
        }
        if (HM == BM || wm == half * HM) {
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int j = 0; j < 4; j++)
#pragma unroll
                    for (int v = 0; v < 4; v++)
                        tile[(wm - half * HM + i * 16 + (lane >> 4) * 4 + v) * LDT + wn + j * 16 + (lane & 15)] =
                            acc[i][j][v];
        }
        __syncthreads();
#pragma unroll
        for (int it = 0; it < HM / 8; it++) {
            const int hrow = it * 8 + (tid >> 5), row = half * HM + hrow, pos = rowPos[row];
            if (pos < 0) continue;
            const int sg = rowSeg[row], r = m0 + row;
            float x[4];
            *(float4 *)x = *(const float4 *)&tile[hrow * LDT + c4];
            if (EPI == EPI_RELU) {
                u16 o[4];
#pragma unroll
                for (int c = 0; c < 4; c++) {
                    const float z = x[c] + bn[c];
                    o[c] = f2bf(z > 0.f ? z : 0.f);
                }
                *(uint2 *)&a.out[(size_t)r * a.ldo + n] = *(const uint2 *)o;
            } else if (EPI == EPI_MASK) {
                u16 m[4], o[4];
                *(uint2 *)m = *(const uint2 *)&a.mask[(size_t)r * a.ldm + n];
#pragma unroll
                for (int c = 0; c < 4; c++) o[c] = (m[c] & 0x8000u) || !(m[c] & 0x7FFFu) ? (u16)0 : f2bf(x[c]);
                *(uint2 *)&a.out[(size_t)r * a.ldo + n] = *(const uint2 *)o;
            } else if (EPI == EPI_SIGMOID) {
                if (pos >= a.lead) {
                    float pr[4];
#pragma unroll
                    for (int c = 0; c < 4; c++) pr[c] = __builtin_amdgcn_rcpf(1.f + __expf(-(x[c] + bn[c])));
                    *(float4 *)&a.prob[((size_t)sg * (a.P - a.lead) + (pos - a.lead)) * a.ldp + n] = *(const float4 *)pr;
                }
            } else {  // EPI_BCE
                u16 o[4] = {0, 0, 0, 0};
                if (pos >= a.lead) {
                    u16 yv[4];
                    *(uint2 *)yv = ypre[it];  // row half * HM + it * 8 + (tid >> 5)
#pragma unroll
                    for (int c = 0; c < 4; c++) {
                        // one exponential serves both: e = exp(-|z|), u = 1 + e,
                        // sigmoid(z) = 1/u (z >= 0) or e/u, and the loss term
                        // softplus(z) - y z = max(z, 0) + log1p(e) - y z with
                        // log1p(e) = log(u) * e / (u - 1) (Goldberg's correction
                        // for the rounding of u; e itself when u rounds to 1)
                        const float z = x[c] + bn[c], y = bf2f(yv[c]);
                        const float e = expf(-fabsf(z)), u = 1.f + e;
                        const float rr = __builtin_amdgcn_rcpf(u);
                        const float pp = z >= 0.f ? rr : e * rr;
                        o[c] = f2bf((pp - y) * a.scale);
                        const float lp = u == 1.f ? e : __logf(u) * e * __builtin_amdgcn_rcpf(u - 1.f);
                        lsum += fmaxf(z, 0.f) + lp - y * z;
                    }
                }
                *(uint2 *)&a.out[(size_t)r * a.ldo + n] = *(const uint2 *)o;
            }
        }
        __syncthreads();
    }
    if (EPI == EPI_BCE) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) lsum += __shfl_xor(lsum, o);
        if (lane == 0 && lsum != 0.f) atomicAdd(a.loss, lsum * a.scale);
    }
}

template <int ASRC, int EPI>
__global__ void __launch_bounds__(kThreads) conv_kernel(const ConvArgs a)
{
    // the operand tiles, double-buffered; after the last K step the same bytes
    // hold the output tile for the epilogue (128 x 128 floats, or its halves)
    constexpr int kOperands = 2 * (BM + BN) * LDC;
    constexpr int kStaged = ((BM * (BN + 4) * 4 + BM * 8 <= kOperands * 2 ? BM : BM / 2) * (BN + 4) * 4 + BM * 8) / 2;
    __shared__ __attribute__((aligned(16))) u16 smem[kOperands > kStaged ? kOperands : kStaged];
    u16(*sA)[BM * LDC] = reinterpret_cast<u16(*)[BM * LDC]>(smem);
    u16(*sB)[BN * LDC] = reinterpret_cast<u16(*)[BN * LDC]>(smem + 2 * BM * LDC);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;

    // loader: kPass rows (lr + kRowStep*i) x 8 consecutive k per thread, for
    // A and B.  Per row: its position in the segment and the source row of
    // position 0.
    const int lr = tid / kVec, lk = (tid % kVec) * 8;
    // (scalars, not arrays: staging arrays were left in scratch memory)
    int64_t ab0 = 0, ab1 = 0, ab2 = 0, ab3 = 0;
    int ap0 = 0, ap1 = 0, ap2 = 0, ap3 = 0;
#define SG_ROW(i, AB, AP)                                                       \
    if (i < kPass) {                                                            \
        const int r = m0 + lr + kRowStep * i;                                   \
        const int s = r < a.M ? r / a.P : 0;                                    \
        AP = r < a.M ? r - s * a.P : -(1 << 30);                                \
        AB = r < a.M ? (a.seg ? a.seg[s] : (int64_t)s * a.P) : 0;               \
    }
    SG_ROW(0, ab0, ap0) SG_ROW(1, ab1, ap1) SG_ROW(2, ab2, ap2) SG_ROW(3, ab3, ap3)
#undef SG_ROW

    uint4 ra0, ra1, ra2, ra3, rb0, rb1, rb2, rb3;
#define SG_LOADA(i, AB, AP, RA)                                                                         \
    if (i < kPass) {                                                                                    \
        const int sp = AP + tap * a.shift;                                                              \
        const bool ok = sp >= 0 && sp < a.P;                                                            \
        const int64_t row = AB + (ok ? sp : 0);                                                         \
        if (ASRC == SRC_BF16)                                                                           \
            RA = ok ? *(const uint4 *)((const u16 *)a.A + (size_t)row * a.lda + c) : make_uint4(0, 0, 0, 0); \
        else                                                                                            \
            RA = ok ? spikes_to_bf16(*(const uint2 *)((const uint8_t *)a.A + (size_t)row * a.lda + c))  \
                    : make_uint4(0, 0, 0, 0);                                                           \
    }
#define SG_LOADB(i, RB) \
    if (i < kPass) RB = *(const uint4 *)(bp + (size_t)(kRowStep * i) * a.K);
#define SG_LOAD(kb)                                                                                     \
    {                                                                                                   \
        const int k0 = (kb) * BK;                                                                       \
        const int tap = k0 / a.cin, c = k0 - tap * a.cin + lk;                                          \
        SG_LOADA(0, ab0, ap0, ra0) SG_LOADA(1, ab1, ap1, ra1) SG_LOADA(2, ab2, ap2, ra2) SG_LOADA(3, ab3, ap3, ra3) \
        const u16 *bp = a.Bt + (size_t)(n0 + lr) * a.K + k0 + lk;                                       \
        SG_LOADB(0, rb0) SG_LOADB(1, rb1) SG_LOADB(2, rb2) SG_LOADB(3, rb3)                             \
    }
#define SG_STORE1(i, RA, RB)                                                              \
    if (i < kPass) {                                                                      \
        const int row = lr + kRowStep * i, at = row * LDC + swz(row, lk >> 3) * 8;        \
        *(uint4 *)&sA[buf_][at] = RA;                                                     \
        *(uint4 *)&sB[buf_][at] = RB;                                                     \
    }
#define SG_STORE(buf)                                                                          \
    {                                                                                          \
        const int buf_ = (buf);                                                                \
        SG_STORE1(0, ra0, rb0) SG_STORE1(1, ra1, rb1) SG_STORE1(2, ra2, rb2) SG_STORE1(3, ra3, rb3) \
    }

    f32x4 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

    const int wm = (wave >> 1) * 64, wn = (wave & 1) * 64;
    const int fr = lane & 15, fk = (lane >> 4) * 8;
    // this block's K steps: all of them, or its split's
    const int kb0 = a.part ? (int)blockIdx.z * a.kPerSplit : 0;
    const int nk = a.part ? min(a.K / BK, kb0 + a.kPerSplit) : a.K / BK;
    SG_LOAD(kb0)
    SG_STORE(0)
    __syncthreads();
    for (int kb = kb0; kb < nk; kb++) {
        const int buf = (kb - kb0) & 1;
        // the next K step's loads stay ahead of this step's MFMAs (left to
        // itself the scheduler moves them down to their LDS stores)
        if (kb + 1 < nk) SG_LOAD(kb + 1)
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int kk = 0; kk < BK; kk += 32) {
            bf16x8 af[4], bfr[4];
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int row = wm + i * 16 + fr;
                af[i] = *(const bf16x8 *)&sA[buf][row * LDC + swz(row, (kk + fk) >> 3) * 8];
            }
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int row = wn + j * 16 + fr;
                bfr[j] = *(const bf16x8 *)&sB[buf][row * LDC + swz(row, (kk + fk) >> 3) * 8];
            }
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int j = 0; j < 4; j++)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i], bfr[j], acc[i][j], 0, 0, 0);
        }
        if (kb + 1 < nk) SG_STORE(buf ^ 1)
        __syncthreads();
    }
#undef SG_LOADA
#undef SG_LOADB
#undef SG_LOAD
#undef SG_STORE1
#undef SG_STORE

    conv_epilogue<EPI, (int)sizeof(smem)>(a, acc, smem, m0, n0, wm, wn);
}

// conv_kernel for BF16 sources with the operand tiles streamed straight from
// global memory into LDS (16-byte global_load_lds per lane, no staging
// registers) through a ring of kStages stages, two K steps ahead of the MFMAs.
// Wave w's DMA instruction q fills rows (2w + q) * 16 .. + 15 of the stage's
// A tile (and then of its B tile); lane l writes bytes 16 l .. 16 l + 15 of
// that 1 KB, i.e. row (2w + q) * 16 + (l >> 2) at physical chunk l & 3, so it
// loads the logical chunk swz(row, l & 3) (the swizzle is an XOR, its own
// inverse).  Rows outside their segment load 16 zero bytes.  Each wave waits
// for its own DMA with an explicit vmcnt and the block meets at a bare
// s_barrier (a __syncthreads release fence would wait for every DMA in
// flight); compiler barriers keep LDS accesses on their side of it.
constexpr int kStages = 3;
template <int EPI>
__global__ void __launch_bounds__(kThreads) conv_dma_kernel(const ConvArgs a)
{
    static_assert(BK == 32 && BM == 128 && BN == 128 && kThreads == 256, "DMA map of conv_dma_kernel");
    typedef __attribute__((address_space(3))) void lds_t;
    typedef __attribute__((address_space(1))) void gbl_t;
    constexpr int kStageU16 = (BM + BN) * BK;  // 16 KB: A rows 0..127, then B rows 0..127
    constexpr int kOps = kStages * kStageU16;
    constexpr int kStaged = ((BM / 2) * (BN + 4) * 4 + BM * 8) / 2;
    __shared__ __attribute__((aligned(16))) u16 smem[kOps > kStaged ? kOps : kStaged];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
    const int lrow = lane >> 2, lch = swz(lrow, lane & 3);

    // the lane's two A rows (q = 0, 1): position in the segment, source row of position 0
    int64_t ab0 = 0, ab1 = 0;
    int ap0 = 0, ap1 = 0;
#define SG_ROW(q, AB, AP)                                                   \
    {                                                                       \
        const int r = m0 + (2 * wave + q) * 16 + lrow;                      \
        const int s = r < a.M ? r / a.P : 0;                                \
        AP = r < a.M ? r - s * a.P : -(1 << 30);                            \
        AB = r < a.M ? (a.seg ? a.seg[s] : (int64_t)s * a.P) : 0;           \
    }
    SG_ROW(0, ab0, ap0) SG_ROW(1, ab1, ap1)
#undef SG_ROW
    const u16 *bq0 = a.Bt + (size_t)(n0 + (2 * wave + 0) * 16 + lrow) * a.K + lch * 8;
    const u16 *bq1 = a.Bt + (size_t)(n0 + (2 * wave + 1) * 16 + lrow) * a.K + lch * 8;
    const u16 *A = (const u16 *)a.A;
#define SG_DMA_A(q, AB, AP)                                                                              \
    {                                                                                                    \
        const int sp = AP + tap * a.shift;                                                               \
        const bool ok = sp >= 0 && sp < a.P;                                                             \
        const u16 *src = ok ? A + (size_t)(AB + sp) * a.lda + c : a.zeros;                               \
        __builtin_amdgcn_global_load_lds((gbl_t *)src, (lds_t *)(stg + (2 * wave + q) * 16 * BK), 16, 0, 0); \
    }
#define SG_ISSUE(kb, st)                                                                                 \
    {                                                                                                    \
        const int k0 = (kb) * BK;                                                                        \
        const int tap = k0 / a.cin, c = k0 - tap * a.cin + lch * 8;                                      \
        u16 *stg = smem + (st) * kStageU16;                                                              \
        SG_DMA_A(0, ab0, ap0) SG_DMA_A(1, ab1, ap1)                                                      \
        __builtin_amdgcn_global_load_lds((gbl_t *)(bq0 + k0), (lds_t *)(stg + BM * BK + (2 * wave + 0) * 16 * BK), 16, 0, 0); \
        __builtin_amdgcn_global_load_lds((gbl_t *)(bq1 + k0), (lds_t *)(stg + BM * BK + (2 * wave + 1) * 16 * BK), 16, 0, 0); \
    }
    // s_waitcnt with vmcnt = n (n < 16), expcnt and lgkmcnt not waited for
#define SG_WAIT_VM(n) __builtin_amdgcn_s_waitcnt(((n) & 0xF) | (7 << 4) | (0xF << 8))
#define SG_BARRIER()                    \
    {                                   \
        asm volatile("" ::: "memory");  \
        __builtin_amdgcn_s_barrier();   \
        asm volatile("" ::: "memory");  \
    }

    f32x4 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int wm = (wave >> 1) * 64, wn = (wave & 1) * 64;
    const int fr = lane & 15, fk = (lane >> 4) * 8;
    const int kb0 = a.part ? (int)blockIdx.z * a.kPerSplit : 0;
    const int nk = a.part ? min(a.K / BK, kb0 + a.kPerSplit) : a.K / BK;
    const int cnt = nk - kb0;  // >= 1
    SG_ISSUE(kb0, 0)
    if (cnt > 1) SG_ISSUE(kb0 + 1, 1)
    for (int it = 0; it < cnt; it++) {
        const int st = it % kStages;
        // this step's stage has landed (each wave's own four DMAs), the next
        // step's may still be in flight
        if (it + 1 < cnt)
            SG_WAIT_VM(4);
        else
            SG_WAIT_VM(0);
        SG_BARRIER()
        // every wave is done with the stage read in the step before: refill it
        if (it + 2 < cnt) SG_ISSUE(kb0 + it + 2, (it + 2) % kStages)
        const u16 *sA = smem + st * kStageU16, *sB = sA + BM * BK;
        bf16x8 af[4], bfr[4];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int row = wm + i * 16 + fr;
            af[i] = *(const bf16x8 *)&sA[row * BK + swz(row, fk >> 3) * 8];
        }
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int row = wn + j * 16 + fr;
            bfr[j] = *(const bf16x8 *)&sB[row * BK + swz(row, fk >> 3) * 8];
        }
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int j = 0; j < 4; j++)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i], bfr[j], acc[i][j], 0, 0, 0);
    }
#undef SG_DMA_A
#undef SG_ISSUE
#undef SG_WAIT_VM
#undef SG_BARRIER
    __syncthreads();  // every wave is done with the last stage (no DMA is in flight)
    conv_epilogue<EPI, (int)sizeof(smem)>(a, acc, smem, m0, n0, wm, wn);
}

struct WgradArgs {
    const u16 *G;        // [R][ldg]: gradient of the layer's pre-activation
    size_t ldg;
    const void *S;       // layer input, time-major (BF16 or spikes)
    size_t lds;
    const int64_t *seg;  // first source row of each segment of S (nullptr: s*P)
    int P, dil, cin, taps;
    int R;               // batch rows
    int rowsPerSplit;    // multiple of BK
    int Nw;              // taps*cin
    float *part;         // [split][cout][Nw]
    int cout;
};

// C[o][tap*cin + c] over this block's split of rows; both operands have the
// reduction index (time) as their row, so each thread loads 4 rows x 8
// columns and writes them transposed: 8 LDS words of 4 rows each.  The 16
// lanes that share an LDS cycle take 16 different row quads of the same 8
// columns, so their 8-byte writes cover all 32 banks once.
template <int SSRC>
__global__ void __launch_bounds__(kThreads) wgrad_kernel(const WgradArgs a)
{
    __shared__ __attribute__((aligned(16))) u16 sG[2][BM * LDK];
    __shared__ __attribute__((aligned(16))) u16 sS[2][BN * LDK];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int n0 = blockIdx.x * BN, m0 = blockIdx.y * BM, split = blockIdx.z;
    const int tap = n0 / a.cin, c0 = n0 - tap * a.cin;
    const int r0 = split * a.rowsPerSplit;
    const int r1 = min(a.R, r0 + a.rowsPerSplit);
    constexpr int kQ = BK / 16;  // rows per thread of each 8-column group
    const int cc = (tid >> 4) * 8, rg = (tid & 15) * kQ;  // 8 columns x rows rg..rg+kQ-1

    uint4 rgv[kQ], rsv[kQ];
#define SG_WLOAD(kb)                                                                                          \
    {                                                                                                         \
        const int rb = r0 + (kb) * BK + rg;                                                                   \
        _Pragma("unroll") for (int q = 0; q < kQ; q++)                                                        \
        {                                                                                                     \
            const int r = rb + q;                                                                             \
            const bool in = r < r1;                                                                           \
            rgv[q] = in ? *(const uint4 *)(a.G + (size_t)r * a.ldg + m0 + cc) : make_uint4(0, 0, 0, 0);       \
            const int s = in ? r / a.P : 0;                                                                   \
            const int sp = in ? r - s * a.P - tap * a.dil : -1;                                               \
            const bool ok = sp >= 0;                                                                          \
            const int64_t row = ok ? (a.seg ? a.seg[s] : (int64_t)s * a.P) + sp : 0;                          \
            if (SSRC == SRC_BF16)                                                                             \
                rsv[q] = ok ? *(const uint4 *)((const u16 *)a.S + (size_t)row * a.lds + c0 + cc) : make_uint4(0, 0, 0, 0); \
            else                                                                                              \
                rsv[q] = ok ? spikes_to_bf16(*(const uint2 *)((const uint8_t *)a.S + (size_t)row * a.lds + c0 + cc)) \
                            : make_uint4(0, 0, 0, 0);                                                         \
        }                                                                                                     \
    }
    // written transposed: column cc + e gets this thread's kQ rows in one
    // 2 * kQ-byte word; the 16 lanes of a column group cover 32 * kQ bytes
    // of one LDS row
#define SG_WSTORE(buf)                                                                       \
    {                                                                                        \
        _Pragma("unroll") for (int e = 0; e < 8; e++)                                        \
        {                                                                                    \
            uint32_t wg[kQ / 2], ws[kQ / 2];                                                 \
            _Pragma("unroll") for (int q = 0; q < kQ; q += 2)                                \
            {                                                                                \
                const u16 *g0 = (const u16 *)&rgv[q], *g1 = (const u16 *)&rgv[q + 1];       \
                const u16 *s0 = (const u16 *)&rsv[q], *s1 = (const u16 *)&rsv[q + 1];       \
                wg[q / 2] = (uint32_t)g0[e] | ((uint32_t)g1[e] << 16);                       \
                ws[q / 2] = (uint32_t)s0[e] | ((uint32_t)s1[e] << 16);                       \
            }                                                                                \
            if constexpr (kQ == 4) {                                                         \
                *(uint2 *)&sG[buf][(cc + e) * LDK + rg] = make_uint2(wg[0], wg[1]);          \
                *(uint2 *)&sS[buf][(cc + e) * LDK + rg] = make_uint2(ws[0], ws[1]);          \
            } else {                                                                         \
                *(uint32_t *)&sG[buf][(cc + e) * LDK + rg] = wg[0];                          \
                *(uint32_t *)&sS[buf][(cc + e) * LDK + rg] = ws[0];                          \
            }                                                                                \
        }                                                                                    \
    }

    f32x4 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int wm = (wave >> 1) * 64, wn = (wave & 1) * 64;
    const int fr = lane & 15, fk = (lane >> 4) * 8;
    const int nk = r1 > r0 ? (r1 - r0 + BK - 1) / BK : 0;
    if (nk > 0) {
        SG_WLOAD(0)
        SG_WSTORE(0)
    }
    __syncthreads();
    for (int kb = 0; kb < nk; kb++) {
        const int buf = kb & 1;
        if (kb + 1 < nk) SG_WLOAD(kb + 1)
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int kk = 0; kk < BK; kk += 32) {
            bf16x8 af[4], bfr[4];
#pragma unroll
            for (int i = 0; i < 4; i++) af[i] = *(const bf16x8 *)&sG[buf][(wm + i * 16 + fr) * LDK + kk + fk];
#pragma unroll
            for (int j = 0; j < 4; j++) bfr[j] = *(const bf16x8 *)&sS[buf][(wn + j * 16 + fr) * LDK + kk + fk];
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int j = 0; j < 4; j++)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i], bfr[j], acc[i][j], 0, 0, 0);
        }
        if (kb + 1 < nk) SG_WSTORE(buf ^ 1)
        __syncthreads();
    }
#undef SG_WLOAD
#undef SG_WSTORE
    float *out = a.part + (size_t)split * a.cout * a.Nw;
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int o = m0 + wm + i * 16 + (lane >> 4) * 4 + v;
#pragma unroll
            for (int j = 0; j < 4; j++) out[(size_t)o * a.Nw + n0 + wn + j * 16 + (lane & 15)] = acc[i][j][v];
        }
}

// partial column sums of G over row chunks: part[chunk][c]
__global__ void __launch_bounds__(256) colsum_kernel(const u16 *G, size_t ldg, int R, int C, int rowsPerChunk,
                                                     float *part)
{
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= C) return;
    const int r0 = blockIdx.y * rowsPerChunk, r1 = min(R, r0 + rowsPerChunk);
    float s = 0.f;
#pragma unroll 8
    for (int r = r0; r < r1; r++) s += bf2f(G[(size_t)r * ldg + c]);
    part[(size_t)blockIdx.y * C + c] = s;
}

// The masked data gradient from split-K partials: out[r][n] = BF16(sum_s
// part[s][r][n], in split order), or 0 where the forward activation mask[r][n]
// is not positive (the EPI_MASK epilogue).
__global__ void __launch_bounds__(256) mask_reduce_kernel(const float *part, int splits, int M, int N, const u16 *mask,
                                                          size_t ldm, u16 *out, size_t ldo)
{
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= (size_t)M * N) return;
    const size_t r = i / (size_t)N, n = i - r * (size_t)N;
    float x = 0.f;
    for (int k = 0; k < splits; k++) x += part[(size_t)k * M * N + i];
    const u16 m = mask[r * ldm + n];
    out[r * ldo + n] = (m & 0x8000u) || !(m & 0x7FFFu) ? (u16)0 : f2bf(x);
}

// A hidden layer's activations from split-K partials: out[r][n] =
// BF16(max(sum_s part[s][r][n] + bias[n], 0)), the sum in split order (the
// EPI_RELU epilogue).
__global__ void __launch_bounds__(256) relu_reduce_kernel(const float *part, int splits, int M, int N,
                                                          const float *bias, u16 *out, size_t ldo)
{
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= (size_t)M * N) return;
    const size_t r = i / (size_t)N, n = i - r * (size_t)N;
    float x = 0.f;
    for (int k = 0; k < splits; k++) x += part[(size_t)k * M * N + i];
    const float z = x + bias[n];
    out[r * ldo + n] = f2bf(z > 0.f ? z : 0.f);
}

// out[i] = sum_s part[s][i], in split order
__global__ void __launch_bounds__(256) reduce_kernel(const float *part, int splits, size_t n, float *out)
{
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    float s = 0.f;
    // in split order; unrolled so that eight loads are in flight at a time
#pragma unroll 8
    for (int k = 0; k < splits; k++) s += part[(size_t)k * n + i];
    out[i] = s;
}

// Adam on one parameter array; writes the BF16 copy in the forward layout and,
// for a weight W[o][tap][c], the transposed copy Wt[c][tap][o].
__global__ void __launch_bounds__(256) adam_kernel(float *theta, float *m, float *v, const float *g, size_t n,
                                                   float lr, float b1, float b2, float eps, float bc1, float bc2,
                                                   u16 *wf, u16 *wt, int cout, int taps, int cin)
{
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float th = theta[i] - lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
    theta[i] = th;
    if (wf) {
        const u16 h = f2bf(th);
        wf[i] = h;
        const int c = (int)(i % cin), tap = (int)((i / cin) % taps), o = (int)(i / ((size_t)cin * taps));
        wt[((size_t)c * taps + tap) * cout + o] = h;
    }
}

__global__ void __launch_bounds__(256) bf16_copies_kernel(const float *theta, size_t n, u16 *wf, u16 *wt, int cout,
                                                          int taps, int cin)
{
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const u16 h = f2bf(theta[i]);
    wf[i] = h;
    const int c = (int)(i % cin), tap = (int)((i / cin) % taps), o = (int)(i / ((size_t)cin * taps));
    wt[((size_t)c * taps + tap) * cout + o] = h;
}

__global__ void __launch_bounds__(256) targets_kernel(const uint8_t *pf, size_t pitch, int n, uint32_t clusters,
                                                      uint32_t per, u16 *y, size_t ypitch)
{
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= (size_t)n * clusters) return;
    const size_t r = i / clusters;
    const uint32_t k = (uint32_t)(i % clusters);
    uint32_t s = 0;
    for (uint32_t q = 0; q < per; q++) s += pf[r * pitch + (size_t)k * per + q];
    y[r * ypitch + k] = f2bf((float)s / (float)(64u * per));
}

__global__ void zero_kernel(float *p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.f;
}

__global__ void copy_loss_kernel(const float *src, float *dst) { *dst = *src; }

}  // namespace sg

using namespace sg;

struct cbm_sg {
    cbm_sg_desc d{};
    int lead = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double lastFlops = 0;
    // parameters: master FP32, Adam moments, gradient; per layer W then b
    size_t off[6]{}, cnt[6]{}, nParams = 0;
    int cout[3]{}, cin[3]{};
    float *theta = nullptr, *mom = nullptr, *vel = nullptr, *grad = nullptr;
    u16 *wf[3]{}, *wt[3]{};
    uint64_t adamStep = 0;
    // activations and gradients of the current batch
    u16 *h1 = nullptr, *h2 = nullptr, *dz3 = nullptr, *dz2 = nullptr, *dz1 = nullptr;
    float *part = nullptr;
    size_t partFloats = 0;
    float *loss = nullptr;
    u16 *zeros = nullptr;  // 64 zero bytes
    int lastRows = 0;
};

namespace {

int check_batch(const cbm_sg *sg, const cbm_sg_batch *b, bool training, int *P, int *R)
{
    if (!b || !b->mf || !b->segStart || b->numSeg < 1 || b->outputs < 1) return fail(CBM_E_ARG, "bad batch");
    if (((uintptr_t)b->mf & 15) || (b->mfPitch & 15) || b->mfPitch < sg->d.numIn)
        return fail(CBM_E_ARG, "MF rows must be 16-byte aligned with a pitch that is a multiple of 16 and >= numIn");
    if (training && (!b->target || ((uintptr_t)b->target & 15) || (b->targetPitch & 7) || b->targetPitch < sg->d.numOut))
        return fail(CBM_E_ARG, "targets must be 16-byte aligned with a pitch (elements) that is a multiple of 8 and >= numOut");
    const int64_t rows = (int64_t)b->numSeg * (sg->lead + b->outputs);
    if (rows > (int64_t)sg->d.maxRows) return fail(CBM_E_ARG, "batch has more rows than maxRows");
    *P = sg->lead + b->outputs;
    *R = (int)rows;
    return CBM_OK;
}

dim3 grid_for(size_t n) { return dim3((unsigned)((n + 255) / 256)); }

// conv_dma_kernel for the data gradient (EPI_MASK: N = 128 columns, split
// over K, few blocks per CU, where the deeper pipeline pays: 154 -> 135 us per
// pass at the full shape); the wide forward passes ran 3-7 % slower with it
// than with register staging and four blocks per CU
template <int E>
constexpr bool conv_dma(int src)
{
    return src == SRC_BF16 && E == EPI_MASK;
}

int conv(cbm_sg *sg, int src, int epi, const ConvArgs &a0)
{
    ConvArgs a = a0;
    a.part = nullptr;
    a.kPerSplit = 0;
    int splits = 1;
    const int tiles = (a.N / BN) * ((a.M + BM - 1) / BM), nk = a.K / BK;
    if ((epi == EPI_MASK || epi == EPI_RELU) && tiles < 512 && nk >= 16) {
        // too few tiles to fill the chip over a long K: split K, sum in order
        splits = std::min(8, std::min(nk / 8, (1024 + tiles - 1) / tiles));
        while (splits > 1 && (size_t)splits * a.M * a.N > sg->partFloats) splits--;
        if (splits > 1) {
            a.kPerSplit = (nk + splits - 1) / splits;
            splits = (nk + a.kPerSplit - 1) / a.kPerSplit;
            a.part = sg->part;
        }
    }
    const dim3 grid(a.N / BN, (a.M + BM - 1) / BM, splits);
    a.zeros = sg->zeros;
    // (spike bytes always go through registers: they are widened to BF16 on the way)
#define CONV_LAUNCH(S, E)                                                                         \
    if (conv_dma<E>(S))                                                                           \
        hipLaunchKernelGGL((conv_dma_kernel<E>), grid, dim3(kThreads), 0, sg->stream, a);         \
    else                                                                                          \
        hipLaunchKernelGGL((conv_kernel<S, E>), grid, dim3(kThreads), 0, sg->stream, a);
#define CONV_CASE(S, E)                                                                         \
    if (src == S && epi == E) {                                                                 \
        CONV_LAUNCH(S, E)                                                                       \
        sg->lastFlops += 2.0 * a.M * (double)a.N * a.K;                                         \
        HIP_TRY(hipGetLastError());                                                             \
        if (a.part) {                                                                           \
            hipLaunchKernelGGL(relu_reduce_kernel, grid_for((size_t)a.M * a.N), dim3(256), 0,   \
                               sg->stream, a.part, splits, a.M, a.N, a.bias, a.out, a.ldo);     \
            HIP_TRY(hipGetLastError());                                                         \
        }                                                                                       \
        return CBM_OK;                                                                          \
    }
    CONV_CASE(SRC_SPIKE, EPI_RELU)
    CONV_CASE(SRC_BF16, EPI_RELU)
    CONV_CASE(SRC_BF16, EPI_SIGMOID)
    CONV_CASE(SRC_BF16, EPI_BCE)
    if (src == SRC_BF16 && epi == EPI_MASK) {
        CONV_LAUNCH(SRC_BF16, EPI_MASK)
        sg->lastFlops += 2.0 * a.M * (double)a.N * a.K;
        HIP_TRY(hipGetLastError());
        if (a.part) {
            hipLaunchKernelGGL(mask_reduce_kernel, grid_for((size_t)a.M * a.N), dim3(256), 0, sg->stream, a.part,
                               splits, a.M, a.N, a.mask, a.ldm, a.out, a.ldo);
            HIP_TRY(hipGetLastError());
        }
        return CBM_OK;
    }
#undef CONV_CASE
#undef CONV_LAUNCH
    return fail(CBM_E_ARG, "unsupported convolution variant");
}

// layer l's forward; X = its input
ConvArgs fwd_args(const cbm_sg *sg, int l, const void *X, size_t ldx, const int64_t *seg, int P, int R)
{
    ConvArgs a{};
    a.A = X;
    a.lda = ldx;
    a.seg = seg;
    a.P = P;
    a.shift = -(int)sg->d.dil[l];
    a.cin = sg->cin[l];
    a.K = (int)sg->d.taps * sg->cin[l];
    a.Bt = sg->wf[l];
    a.M = R;
    a.N = sg->cout[l];
    a.bias = sg->theta + sg->off[2 * l + 1];
    a.lead = sg->lead;
    return a;
}

int forward_hidden(cbm_sg *sg, const cbm_sg_batch *b, int P, int R)
{
    ConvArgs a = fwd_args(sg, 0, b->mf, b->mfPitch, b->segStart, P, R);
    a.out = sg->h1;
    a.ldo = sg->d.hidden;
    int rc = conv(sg, SRC_SPIKE, EPI_RELU, a);
    if (rc) return rc;
    a = fwd_args(sg, 1, sg->h1, sg->d.hidden, nullptr, P, R);
    a.out = sg->h2;
    a.ldo = sg->d.hidden;
    return conv(sg, SRC_BF16, EPI_RELU, a);
}

int wgrad(cbm_sg *sg, int l, int src, const u16 *G, const void *S, size_t lds, const int64_t *seg, int P, int R)
{
    WgradArgs w{};
    w.G = G;
    w.ldg = sg->cout[l];
    w.S = S;
    w.lds = lds;
    w.seg = seg;
    w.P = P;
    w.dil = (int)sg->d.dil[l];
    w.cin = sg->cin[l];
    w.taps = (int)sg->d.taps;
    w.R = R;
    w.Nw = w.taps * w.cin;
    w.cout = sg->cout[l];
    const int tiles = (w.Nw / BN) * (w.cout / BM);
    const int kSteps = (R + BK - 1) / BK;
    int splits = (1024 + tiles - 1) / tiles;
    splits = std::max(1, std::min(splits, kSteps));
    const size_t n = (size_t)w.cout * w.Nw;
    while (splits > 1 && (size_t)splits * n > sg->partFloats) splits--;
    w.rowsPerSplit = ((kSteps + splits - 1) / splits) * BK;
    splits = (R + w.rowsPerSplit - 1) / w.rowsPerSplit;
    w.part = sg->part;
    const dim3 grid(w.Nw / BN, w.cout / BM, splits);
    if (src == SRC_SPIKE)
        hipLaunchKernelGGL(wgrad_kernel<SRC_SPIKE>, grid, dim3(kThreads), 0, sg->stream, w);
    else
        hipLaunchKernelGGL(wgrad_kernel<SRC_BF16>, grid, dim3(kThreads), 0, sg->stream, w);
    HIP_TRY(hipGetLastError());
    sg->lastFlops += 2.0 * R * (double)n;
    hipLaunchKernelGGL(reduce_kernel, grid_for(n), dim3(256), 0, sg->stream, sg->part, splits, n,
                       sg->grad + sg->off[2 * l]);
    HIP_TRY(hipGetLastError());
    // bias gradient: column sums of G
    const int C = sg->cout[l];
    const int chunks = std::max(1, std::min(256, R / 64));
    const int per = (R + chunks - 1) / chunks;
    hipLaunchKernelGGL(colsum_kernel, dim3((C + 255) / 256, chunks), dim3(256), 0, sg->stream, G, (size_t)C, R, C,
                       per, sg->part);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(reduce_kernel, grid_for(C), dim3(256), 0, sg->stream, sg->part, chunks, (size_t)C,
                       sg->grad + sg->off[2 * l + 1]);
    HIP_TRY(hipGetLastError());
    return CBM_OK;
}

// forward with the loss epilogue, then every gradient into sg->grad
int forward_backward(cbm_sg *sg, const cbm_sg_batch *b, int P, int R)
{
    hipLaunchKernelGGL(zero_kernel, dim3(1), dim3(64), 0, sg->stream, sg->loss, 1);
    HIP_TRY(hipGetLastError());
    int rc = forward_hidden(sg, b, P, R);
    if (rc) return rc;
    const size_t H = sg->d.hidden, O = sg->d.numOut;
    ConvArgs a = fwd_args(sg, 2, sg->h2, H, nullptr, P, R);
    a.out = sg->dz3;
    a.ldo = O;
    a.Y = b->target;
    a.ldy = b->targetPitch;
    a.segY = b->segStart;
    a.scale = 1.0f / ((float)b->numSeg * (float)b->outputs * (float)O);
    a.loss = sg->loss;
    if ((rc = conv(sg, SRC_BF16, EPI_BCE, a))) return rc;
    // layer 3: weight gradient, then the data gradient masked by H2 > 0
    if ((rc = wgrad(sg, 2, SRC_BF16, sg->dz3, sg->h2, H, nullptr, P, R))) return rc;
    ConvArgs d{};
    d.A = sg->dz3;
    d.lda = O;
    d.P = P;
    d.shift = (int)sg->d.dil[2];
    d.cin = (int)O;
    d.K = (int)(sg->d.taps * O);
    d.Bt = sg->wt[2];
    d.M = R;
    d.N = (int)H;
    d.out = sg->dz2;
    d.ldo = H;
    d.mask = sg->h2;
    d.ldm = H;
    if ((rc = conv(sg, SRC_BF16, EPI_MASK, d))) return rc;
    // layer 2
    if ((rc = wgrad(sg, 1, SRC_BF16, sg->dz2, sg->h1, H, nullptr, P, R))) return rc;
    d.A = sg->dz2;
    d.lda = H;
    d.shift = (int)sg->d.dil[1];
    d.cin = (int)H;
    d.K = (int)(sg->d.taps * H);
    d.Bt = sg->wt[1];
    d.out = sg->dz1;
    d.mask = sg->h1;
    if ((rc = conv(sg, SRC_BF16, EPI_MASK, d))) return rc;
    // layer 1 (no gradient of the spikes)
    return wgrad(sg, 0, SRC_SPIKE, sg->dz1, b->mf, b->mfPitch, b->segStart, P, R);
}

int refresh_copies(cbm_sg *sg)
{
    for (int l = 0; l < 3; l++) {
        const size_t n = sg->cnt[2 * l];
        hipLaunchKernelGGL(bf16_copies_kernel, grid_for(n), dim3(256), 0, sg->stream, sg->theta + sg->off[2 * l], n,
                           sg->wf[l], sg->wt[l], sg->cout[l], (int)sg->d.taps, sg->cin[l]);
        HIP_TRY(hipGetLastError());
    }
    return CBM_OK;
}

int clear_adam(cbm_sg *sg)
{
    HIP_TRY(hipMemsetAsync(sg->mom, 0, sg->nParams * sizeof(float), sg->stream));
    HIP_TRY(hipMemsetAsync(sg->vel, 0, sg->nParams * sizeof(float), sg->stream));
    sg->adamStep = 0;
    return CBM_OK;
}

uint64_t splitmix(uint64_t &x)
{
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

}  // namespace

extern "C" {

int cbm_sg_create(const cbm_sg_desc *desc, cbm_sg **out)
{
    if (!desc || !out) return fail(CBM_E_ARG, "null argument");
    *out = nullptr;
    const cbm_sg_desc &d = *desc;
    if (!d.numIn || d.numIn % 128 || !d.hidden || d.hidden % 128 || !d.numOut || d.numOut % 128)
        return fail(CBM_E_UNSUPPORTED, "numIn, hidden and numOut must be positive multiples of 128");
    if (d.taps < 1 || d.taps > 16) return fail(CBM_E_UNSUPPORTED, "taps must be 1..16");
    if (!d.dil[0] || !d.dil[1] || !d.dil[2]) return fail(CBM_E_UNSUPPORTED, "dilations must be >= 1");
    if (d.maxRows < 1 || d.maxRows > (1u << 26)) return fail(CBM_E_UNSUPPORTED, "maxRows must be 1..2^26");
    HIP_TRY(hipSetDevice(d.device));
    cbm_sg *sg = new cbm_sg;
    sg->d = d;
    sg->lead = (int)((d.taps - 1) * (d.dil[0] + d.dil[1] + d.dil[2]));
    sg->cin[0] = (int)d.numIn;
    sg->cin[1] = sg->cin[2] = (int)d.hidden;
    sg->cout[0] = sg->cout[1] = (int)d.hidden;
    sg->cout[2] = (int)d.numOut;
    size_t o = 0;
    for (int l = 0; l < 3; l++) {
        sg->off[2 * l] = o;
        sg->cnt[2 * l] = (size_t)sg->cout[l] * d.taps * sg->cin[l];
        o += sg->cnt[2 * l];
        sg->off[2 * l + 1] = o;
        sg->cnt[2 * l + 1] = (size_t)sg->cout[l];
        o += sg->cnt[2 * l + 1];
    }
    sg->nParams = o;
    const size_t R = d.maxRows;
    size_t part = 0;
    // room for 8 row splits of the largest weight gradient
    for (int l = 0; l < 3; l++) part = std::max(part, (size_t)sg->cout[l] * d.taps * sg->cin[l] * 8);
    part = std::max(part, (size_t)256 * d.numOut);
    sg->partFloats = part;
    auto alloc = [&](void **p, size_t bytes) -> int {
        hipError_t e = hipMalloc(p, bytes);
        if (e != hipSuccess) return fail(CBM_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
        return CBM_OK;
    };
    int rc = CBM_OK;
    rc = rc ? rc : alloc((void **)&sg->theta, o * 4);
    rc = rc ? rc : alloc((void **)&sg->mom, o * 4);
    rc = rc ? rc : alloc((void **)&sg->vel, o * 4);
    rc = rc ? rc : alloc((void **)&sg->grad, o * 4);
    for (int l = 0; l < 3 && !rc; l++) {
        rc = alloc((void **)&sg->wf[l], sg->cnt[2 * l] * 2);
        rc = rc ? rc : alloc((void **)&sg->wt[l], sg->cnt[2 * l] * 2);
    }
    rc = rc ? rc : alloc((void **)&sg->h1, R * d.hidden * 2);
    rc = rc ? rc : alloc((void **)&sg->h2, R * d.hidden * 2);
    rc = rc ? rc : alloc((void **)&sg->dz2, R * d.hidden * 2);
    rc = rc ? rc : alloc((void **)&sg->dz1, R * d.hidden * 2);
    rc = rc ? rc : alloc((void **)&sg->dz3, R * d.numOut * 2);
    rc = rc ? rc : alloc((void **)&sg->part, part * 4);
    rc = rc ? rc : alloc((void **)&sg->loss, 4);
    rc = rc ? rc : alloc((void **)&sg->zeros, 64);
    if (!rc && hipMemset(sg->zeros, 0, 64) != hipSuccess) rc = fail(CBM_E_HIP, "hipMemset failed");
    if (!rc && hipStreamCreateWithFlags(&sg->stream, hipStreamNonBlocking) != hipSuccess)
        rc = fail(CBM_E_HIP, "hipStreamCreate failed");
    if (!rc && (hipEventCreate(&sg->ev0) != hipSuccess || hipEventCreate(&sg->ev1) != hipSuccess))
        rc = fail(CBM_E_HIP, "hipEventCreate failed");
    if (!rc) {
        HIP_TRY(hipMemset(sg->theta, 0, o * 4));
        HIP_TRY(hipMemset(sg->mom, 0, o * 4));
        HIP_TRY(hipMemset(sg->vel, 0, o * 4));
        rc = refresh_copies(sg);
        if (!rc) HIP_TRY(hipStreamSynchronize(sg->stream));
    }
    if (rc) {
        cbm_sg_destroy(sg);
        return rc;
    }
    *out = sg;
    return CBM_OK;
}

int cbm_sg_destroy(cbm_sg *sg)
{
    if (!sg) return CBM_OK;
    (void)hipSetDevice(sg->d.device);
    if (sg->stream) (void)hipStreamSynchronize(sg->stream);
    void *ptrs[] = {sg->theta, sg->mom, sg->vel, sg->grad, sg->wf[0], sg->wf[1], sg->wf[2], sg->wt[0], sg->wt[1],
                    sg->wt[2], sg->h1, sg->h2, sg->dz1, sg->dz2, sg->dz3, sg->part, sg->loss, sg->zeros};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    if (sg->ev0) (void)hipEventDestroy(sg->ev0);
    if (sg->ev1) (void)hipEventDestroy(sg->ev1);
    if (sg->stream) (void)hipStreamDestroy(sg->stream);
    delete sg;
    return CBM_OK;
}

int cbm_sg_lead(const cbm_sg *sg) { return sg ? sg->lead : fail(CBM_E_ARG, "null surrogate"); }

int cbm_sg_num_params(const cbm_sg *sg, size_t *n)
{
    if (!sg || !n) return fail(CBM_E_ARG, "null argument");
    *n = sg->nParams;
    return CBM_OK;
}

int cbm_sg_get_params(cbm_sg *sg, float *host)
{
    if (!sg || !host) return fail(CBM_E_ARG, "null argument");
    HIP_TRY(hipSetDevice(sg->d.device));
    HIP_TRY(hipStreamSynchronize(sg->stream));
    HIP_TRY(hipMemcpy(host, sg->theta, sg->nParams * 4, hipMemcpyDeviceToHost));
    return CBM_OK;
}

int cbm_sg_set_params(cbm_sg *sg, const float *host)
{
    if (!sg || !host) return fail(CBM_E_ARG, "null argument");
    HIP_TRY(hipSetDevice(sg->d.device));
    HIP_TRY(hipStreamSynchronize(sg->stream));
    HIP_TRY(hipMemcpy(sg->theta, host, sg->nParams * 4, hipMemcpyHostToDevice));
    int rc = refresh_copies(sg);
    if (!rc) rc = clear_adam(sg);
    if (!rc) HIP_TRY(hipStreamSynchronize(sg->stream));
    return rc;
}

int cbm_sg_init_params(cbm_sg *sg, uint64_t seed, float outBias)
{
    if (!sg) return fail(CBM_E_ARG, "null surrogate");
    std::vector<float> h(sg->nParams, 0.0f);
    uint64_t x = seed;
    for (int l = 0; l < 3; l++) {
        const double fanIn = (double)sg->d.taps * sg->cin[l];
        // uniform(-a, a): variance 2/fanIn for the ReLU layers, 1/fanIn for the output layer
        const double a = std::sqrt((l < 2 ? 6.0 : 3.0) / fanIn);
        for (size_t i = 0; i < sg->cnt[2 * l]; i++) {
            const double u = (double)(splitmix(x) >> 11) * (1.0 / 9007199254740992.0);
            h[sg->off[2 * l] + i] = (float)((2.0 * u - 1.0) * a);
        }
    }
    for (size_t i = 0; i < sg->cnt[5]; i++) h[sg->off[5] + i] = outBias;
    return cbm_sg_set_params(sg, h.data());
}

int cbm_sg_forward(cbm_sg *sg, const cbm_sg_batch *b, float *prob, size_t probPitch)
{
    if (!sg || !prob || probPitch < sg->d.numOut) return fail(CBM_E_ARG, "bad argument");
    if (((uintptr_t)prob & 15) || (probPitch & 3))
        return fail(CBM_E_ARG, "prob must be 16-byte aligned with a pitch (elements) that is a multiple of 4");
    int P, R;
    int rc = check_batch(sg, b, false, &P, &R);
    if (rc) return rc;
    HIP_TRY(hipSetDevice(sg->d.device));
    sg->lastFlops = 0;
    HIP_TRY(hipEventRecord(sg->ev0, sg->stream));
    if ((rc = forward_hidden(sg, b, P, R))) return rc;
    ConvArgs a = fwd_args(sg, 2, sg->h2, sg->d.hidden, nullptr, P, R);
    a.prob = prob;
    a.ldp = probPitch;
    if ((rc = conv(sg, SRC_BF16, EPI_SIGMOID, a))) return rc;
    HIP_TRY(hipEventRecord(sg->ev1, sg->stream));
    sg->lastRows = R;
    return CBM_OK;
}

int cbm_sg_loss_grad(cbm_sg *sg, const cbm_sg_batch *b, float *loss, float *grad)
{
    if (!sg) return fail(CBM_E_ARG, "null surrogate");
    int P, R;
    int rc = check_batch(sg, b, true, &P, &R);
    if (rc) return rc;
    HIP_TRY(hipSetDevice(sg->d.device));
    sg->lastFlops = 0;
    HIP_TRY(hipEventRecord(sg->ev0, sg->stream));
    if ((rc = forward_backward(sg, b, P, R))) return rc;
    HIP_TRY(hipEventRecord(sg->ev1, sg->stream));
    sg->lastRows = R;
    HIP_TRY(hipStreamSynchronize(sg->stream));
    if (loss) HIP_TRY(hipMemcpy(loss, sg->loss, 4, hipMemcpyDeviceToHost));
    if (grad) HIP_TRY(hipMemcpy(grad, sg->grad, sg->nParams * 4, hipMemcpyDeviceToHost));
    return CBM_OK;
}

int cbm_sg_train_step(cbm_sg *sg, const cbm_sg_batch *b, const cbm_sg_adam *opt, float *lossOut)
{
    if (!sg || !opt) return fail(CBM_E_ARG, "null argument");
    if (!(opt->lr >= 0) || !(opt->beta1 >= 0 && opt->beta1 < 1) || !(opt->beta2 >= 0 && opt->beta2 < 1) ||
        !(opt->eps > 0))
        return fail(CBM_E_ARG, "bad Adam settings");
    int P, R;
    int rc = check_batch(sg, b, true, &P, &R);
    if (rc) return rc;
    HIP_TRY(hipSetDevice(sg->d.device));
    sg->lastFlops = 0;
    HIP_TRY(hipEventRecord(sg->ev0, sg->stream));
    if ((rc = forward_backward(sg, b, P, R))) return rc;
    sg->adamStep++;
    const float bc1 = 1.0f - (float)std::pow((double)opt->beta1, (double)sg->adamStep);
    const float bc2 = 1.0f - (float)std::pow((double)opt->beta2, (double)sg->adamStep);
    for (int k = 0; k < 6; k++) {
        const size_t n = sg->cnt[k], o = sg->off[k];
        const int l = k / 2;
        const bool w = (k % 2) == 0;
        hipLaunchKernelGGL(adam_kernel, grid_for(n), dim3(256), 0, sg->stream, sg->theta + o, sg->mom + o,
                           sg->vel + o, sg->grad + o, n, opt->lr, opt->beta1, opt->beta2, opt->eps, bc1, bc2,
                           w ? sg->wf[l] : nullptr, w ? sg->wt[l] : nullptr, sg->cout[l], (int)sg->d.taps, sg->cin[l]);
        HIP_TRY(hipGetLastError());
    }
    if (lossOut) {
        hipLaunchKernelGGL(copy_loss_kernel, dim3(1), dim3(1), 0, sg->stream, sg->loss, lossOut);
        HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipEventRecord(sg->ev1, sg->stream));
    sg->lastRows = R;
    return CBM_OK;
}

int cbm_sg_read_hidden(cbm_sg *sg, int layer, uint16_t *host, size_t rows)
{
    if (!sg || !host || (layer != 1 && layer != 2)) return fail(CBM_E_ARG, "bad argument");
    if (rows > (size_t)sg->lastRows) return fail(CBM_E_ARG, "more rows than the last batch");
    HIP_TRY(hipSetDevice(sg->d.device));
    HIP_TRY(hipStreamSynchronize(sg->stream));
    HIP_TRY(hipMemcpy(host, layer == 1 ? sg->h1 : sg->h2, rows * sg->d.hidden * 2, hipMemcpyDeviceToHost));
    return CBM_OK;
}

int cbm_sg_targets_from_pf(cbm_sg *sg, const uint8_t *pf, size_t pfPitch, int n, uint32_t numChunks,
                           uint32_t chunksPerCluster, uint16_t *y, size_t yPitch)
{
    if (!sg || !pf || !y || n < 0 || !chunksPerCluster || numChunks % chunksPerCluster || pfPitch < numChunks)
        return fail(CBM_E_ARG, "bad argument");
    const uint32_t clusters = numChunks / chunksPerCluster;
    if (yPitch < clusters) return fail(CBM_E_ARG, "target pitch below the cluster count");
    if (n == 0) return CBM_OK;
    HIP_TRY(hipSetDevice(sg->d.device));
    hipLaunchKernelGGL(targets_kernel, grid_for((size_t)n * clusters), dim3(256), 0, sg->stream, pf, pfPitch, n,
                       clusters, chunksPerCluster, y, yPitch);
    HIP_TRY(hipGetLastError());
    return CBM_OK;
}

int cbm_sg_sync(cbm_sg *sg)
{
    if (!sg) return fail(CBM_E_ARG, "null surrogate");
    HIP_TRY(hipSetDevice(sg->d.device));
    HIP_TRY(hipStreamSynchronize(sg->stream));
    return CBM_OK;
}

int cbm_sg_last_timing(cbm_sg *sg, double *us, double *flops)
{
    if (!sg || !us || !flops) return fail(CBM_E_ARG, "null argument");
    HIP_TRY(hipSetDevice(sg->d.device));
    HIP_TRY(hipEventSynchronize(sg->ev1));
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, sg->ev0, sg->ev1));
    *us = ms * 1000.0;
    *flops = sg->lastFlops;
    return CBM_OK;
}

}  // extern "C"

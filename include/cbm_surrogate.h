/* cbm_surrogate.h -- learned surrogate of the granule-Golgi input layer.
 *
 * A three-layer causal temporal-convolution network maps the MF spikes of the
 * last 50 ms (one row of numIn bytes per ms) to the probability that a cell of
 * each GR cluster emits a PF spike in the next ms.  Layer l is
 *     Z_l[t][o] = b_l[o] + sum_{tap < taps} sum_c X_l[t - tap*dil_l][c] * W_l[o][tap][c]
 * with ReLU after layers 1 and 2 and a logistic sigmoid after layer 3; rows
 * before the start of a segment read as zero.  With taps 8 and dilations
 * 1, 2, 4 the receptive field is 1 + 7*(1+2+4) = 50 rows.
 *
 * Arithmetic: BF16 weights and activations, FP32 accumulation on the matrix
 * cores (hand-written MFMA kernels), FP32 master weights and Adam moments.
 * Training minimises the binary cross-entropy between the predicted
 * probability and the observed fraction of the cluster's cells that spiked
 * (soft targets), averaged over every output of the batch.
 *
 * No reference counterpart: the reference simulates every GR.  The surrogate
 * is a separate path; the exact step (cbm_hip.h) is unchanged by it and
 * serves as its judge.  All pointers marked "device" are device memory of
 * the surrogate's device; all calls are asynchronous on the surrogate's own
 * stream unless they return host data.  Status codes are those of cbm_hip.h.
 */
#ifndef CBM_SURROGATE_H
#define CBM_SURROGATE_H

#include <stddef.h>
#include <stdint.h>

#include "cbm_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct cbm_sg cbm_sg;

typedef struct cbm_sg_desc {
    uint32_t numIn;    /* MF channels per row; multiple of 128 */
    uint32_t hidden;   /* channels of layers 1 and 2; multiple of 128 */
    uint32_t numOut;   /* GR clusters; multiple of 128 */
    uint32_t taps;     /* kernel length, 1..16 */
    uint32_t dil[3];   /* dilation of each layer, >= 1 */
    uint32_t maxRows;  /* largest batch: segments x (lead + outputs per segment) */
    int device;        /* HIP device ordinal */
} cbm_sg_desc;

/* A batch of segments cut from one stream of MF rows.  Segment s covers
 * rows segStart[s] .. segStart[s] + lead + outputs - 1, where lead =
 * (taps-1)*(dil[0]+dil[1]+dil[2]); its j-th output (j < outputs) is the
 * prediction for the window that ends at row segStart[s] + lead + j, and its
 * training target is row segStart[s] + lead + j of `target`. */
typedef struct cbm_sg_batch {
    const uint8_t *mf;        /* device: MF rows, nonzero byte = spike; 16-byte aligned */
    size_t mfPitch;           /* bytes between rows, multiple of 16, >= numIn */
    const int64_t *segStart;  /* device: first row of each segment */
    int numSeg;
    int outputs;              /* outputs per segment, >= 1 */
    const uint16_t *target;   /* device: BF16 bit patterns, [rows][numOut] (training only) */
    size_t targetPitch;       /* elements between target rows, multiple of 8 */
} cbm_sg_batch;

typedef struct cbm_sg_adam {
    float lr, beta1, beta2, eps;
} cbm_sg_adam;

int cbm_sg_create(const cbm_sg_desc *desc, cbm_sg **out);
int cbm_sg_destroy(cbm_sg *sg);

/* Rows of history each output needs before it (the receptive field minus 1). */
int cbm_sg_lead(const cbm_sg *sg);

/* Number of FP32 parameters, in the order W1, b1, W2, b2, W3, b3 with
 * W_l[o][tap][c] (output channel major, then tap, then input channel). */
int cbm_sg_num_params(const cbm_sg *sg, size_t *n);
int cbm_sg_get_params(cbm_sg *sg, float *host);
/* Set the master weights (and their BF16 copies); clears the Adam state. */
int cbm_sg_set_params(cbm_sg *sg, const float *host);
/* Seeded uniform initialisation scaled by each layer's fan-in; the output
 * biases are set to outBias (the logit of the mean target is a good start).
 * Clears the Adam state. */
int cbm_sg_init_params(cbm_sg *sg, uint64_t seed, float outBias);

/* Inference: probabilities, device FP32 [numSeg * outputs][probPitch]; prob
 * 16-byte aligned and probPitch a multiple of 4. */
int cbm_sg_forward(cbm_sg *sg, const cbm_sg_batch *b, float *prob, size_t probPitch);

/* Forward + backward without an update.  loss (host, nullable) receives the
 * mean cross-entropy; grad (host, nullable) the FP32 gradient of every
 * parameter in the cbm_sg_get_params order.  Synchronous. */
int cbm_sg_loss_grad(cbm_sg *sg, const cbm_sg_batch *b, float *loss, float *grad);

/* One Adam step on the batch.  lossOut (device, nullable) receives this
 * step's mean cross-entropy. */
int cbm_sg_train_step(cbm_sg *sg, const cbm_sg_batch *b, const cbm_sg_adam *opt, float *lossOut);

/* Copy layer 1 or 2's BF16 activations of the last forward, rows
 * [numSeg * (lead + outputs)][hidden], to host memory (tests). */
int cbm_sg_read_hidden(cbm_sg *sg, int layer, uint16_t *host, size_t rows);

/* Training targets from recorded PF counts (cbm_read_recorded, CBM_REC_PF):
 * y[r][k] = BF16( sum of `chunksPerCluster` consecutive chunk counts of row r
 * / (64 * chunksPerCluster) ), for n rows and numChunks/chunksPerCluster
 * clusters.  pf and y are device memory. */
int cbm_sg_targets_from_pf(cbm_sg *sg, const uint8_t *pf, size_t pfPitch, int n, uint32_t numChunks,
                           uint32_t chunksPerCluster, uint16_t *y, size_t yPitch);

int cbm_sg_sync(cbm_sg *sg);
/* Kernel time of the last cbm_sg_train_step / cbm_sg_forward in microseconds
 * (HIP events on the surrogate's stream), and the matrix-core flops it did. */
int cbm_sg_last_timing(cbm_sg *sg, double *us, double *flops);

/* ---------------------------------------------------------------- sweeps */
/* Step the networks of a batch with their input layer replaced by cluster
 * probabilities (e.g. from cbm_sg_forward): fast sweeps of the microzone side
 * (PC, BC, SC, IO, NC and PF->PC plasticity) over many parameter sets that
 * share one MF train.  Each network's small populations step exactly as in
 * cbm_batch_run_mf_train; its GR layer is not stepped.  A cluster is
 * `cellsPerCluster` consecutive GRs (a power of two >= 64 that tiles every
 * PC, BC, SC and IO block).  Per step, in expectation over the cluster's
 * cells:
 *   PF->PC input of a PC  = sum over its clusters of cells * W[z][c] * p[c]
 *   PF->BC, PF->SC counts = nearest integer to sum of cells * p[c]
 *   GR->GO counts         = 0 (Golgi cells drive only the replaced layer)
 *   every tsPerHistBinGR steps: W[z][c] = clamp(W + q * step, 0, 1), q = the
 *     probability of a PF spike in the 5-step bin 195 steps back (the history
 *     bit 39 of the exact rule, kernels.cu:282-291, mzone.cpp:643-724).
 * prob: device FP32 rows of probPitch floats; row k is run step k (the
 * probabilities of the PF spikes of step t0 + k + 1), and rows -199 .. -1
 * before it must be readable (the plasticity look-back).  The cluster
 * weights W start as the cluster means of each network's per-GR weights on
 * the first sweep (or after cbm_batch_sweep_reset) and carry over between
 * sweeps; the per-GR weights are not changed.  No reference counterpart. */
int cbm_batch_sweep(cbm_batch *b, const float *prob, size_t probPitch, int n, const uint8_t *mf_dev, size_t mfPitch,
                    const float *const *err_host, uint32_t cellsPerCluster);
/* Cluster weights of network `net`, [numZones][numGR / cellsPerCluster]. */
int cbm_batch_sweep_weights(cbm_batch *b, int net, float *host);
/* Re-derive every cluster weight from the per-GR weights. */
int cbm_batch_sweep_reset(cbm_batch *b);

#ifdef __cplusplus
}
#endif
#endif

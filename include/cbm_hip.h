/*
 * cbm_hip.h -- C-ABI of the MI355X cerebellar network step.
 *
 * One handle = one network (an input net plus numZones microzones) resident
 * on one GPU.  The functions below replace the reference's device FFI and the
 * host-side small-population loops of one 1 ms step:
 *
 *   cbm_create / cbm_commit   <- InNet::initCUDA + MZone::initCUDA
 *                                (CBM_CORE_LIB_MGPU_DEV/src/innetmodules/innet.cpp:168-606,
 *                                 CBM_CORE_LIB_MGPU_DEV/src/mzonemodules/mzone.cpp:71-137)
 *   cbm_set_mf                <- CBMSimCore::updateMFInput
 *                                (CBM_CORE_LIB_MGPU_DEV/src/interface/cbmsimcore.cpp:312-320)
 *   cbm_set_err_drive         <- CBMSimCore::updateErrDrive / MZone::setErrDrive
 *                                (cbmsimcore.cpp:322-325, mzone.cpp:167-173)
 *   cbm_set_grpc_steps        <- MZone::setGRPCPlastSteps / resetGRPCPlastSteps (mzone.cpp:726-736)
 *   cbm_set_ginc_grgo         <- InNet::setGIncGRtoGO / resetGIncGRtoGO (innet.cpp:669-677)
 *   cbm_step                  <- CBMSimCore::calcActivity (cbmsimcore.cpp:139-310), i.e. the nine
 *                                live kernels of CBM_CORE_LIB_MGPU_DEV/src/cuda/kernels.cu:79-376 and
 *                                the host loops of innet.cpp:800-995 and mzone.cpp:187-605
 *   cbm_run_mf_train          <- a driver loop of updateMFInput + calcActivity over a
 *                                device-resident spike train (no counterpart; one launch sequence
 *                                per ms without host round trips)
 *   cbm_read                  <- InNet/MZone export* (innet.cpp:679-798, mzone.cpp:739-788)
 *                                and writeToState (innet.cpp:608-667, mzone.cpp:139-165)
 *   cbm_write                 <- state upload of the constructors (innet.cpp:406-481)
 *
 * All functions return 0 on success and a negative CBM_E_* code on failure;
 * cbm_last_error() gives a message for the calling thread.  A handle is not
 * thread safe: one host thread drives it, as in the reference.
 *
 * Arrays cross the boundary in the CBMState layout of the reference
 * (cell-major 2-D arrays [cell][slot], unused slots = UINT32_MAX), so a
 * CBMState can be handed over field by field.
 */
#ifndef CBM_HIP_H
#define CBM_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CBM_ABI_VERSION 5

/* ---------------------------------------------------------------- errors */
#define CBM_OK 0
#define CBM_E_ARG (-1)      /* bad argument (null, wrong size, bad field)      */
#define CBM_E_STATE (-2)    /* call not valid in this state (e.g. not committed) */
#define CBM_E_HIP (-3)      /* HIP runtime error                                */
#define CBM_E_UNSUPPORTED (-4) /* network outside what the kernels support      */
#define CBM_E_NOMEM (-5)

/* ---------------------------------------------------------------- sizes */
/* Population sizes and fan-ins: the fields of ConnectivityParams that the step
 * uses (CBM_STATE_LIB_DEV/src/params/connectivityparams.cpp:90-184). */
typedef struct cbm_sizes {
    uint32_t numGR, grX, grY;
    uint32_t numGO, goX, goY;
    uint32_t numMF;
    uint32_t numSC, numBC, numPC, numNC, numIO;
    uint32_t numZones;

    uint32_t maxnumpGRfromMFtoGR;   /* slots per GR in P_GR_FROM_MF_TO_GR       */
    uint32_t maxnumpGRfromGOtoGR;   /* slots per GR in P_GR_FROM_GO_TO_GR       */
    uint32_t maxnumpGRfromGRtoGO;   /* slots per GR in P_GR_FROM_GR_TO_GO       */
    uint32_t maxnumpMFfromMFtoGO;   /* slots per MF in P_MF_FROM_MF_TO_GO       */
    uint32_t maxnumpGOGABAOutGOGO;  /* slots per GO in P_GO_GABA_OUT_GOGO       */
    uint32_t maxnumpGOCoupInGOGO;   /* slots per GO in P_GO_COUP_IN_GOGO        */

    uint32_t numpPCfromGRtoPC;      /* numGR/numPC: contiguous PF block per PC  */
    uint32_t numpBCfromGRtoBC;      /* numGR/numBC                              */
    uint32_t numpSCfromGRtoSC;      /* numGR/numSC                              */
    uint32_t numpBCfromBCtoPC;      /* 4                                        */
    uint32_t numpPCfromPCtoBC;      /* 16                                       */
    uint32_t numpPCfromSCtoPC;      /* numSC/numPC                              */
    uint32_t numpNCfromPCtoNC;      /* numPC/numNC*numpPCfromPCtoNC             */
    uint32_t numpNCfromMFtoNC;      /* numMF/numNC                              */
    uint32_t numpIOfromNCtoIO;      /* numNC (all to all)                       */
    uint32_t numpIOInIOIO;          /* numIO-1                                  */
    uint32_t numPopHistBinsPC;      /* tsPopHistPC/tsPerPopHistBinPC (8)        */
} cbm_sizes;

/* ---------------------------------------------------------------- params */
/* Derived activity constants, exactly as ActivityParams::updateParams computes
 * them (CBM_STATE_LIB_DEV/src/params/activityparams.cpp:436-636).  Computed
 * once on the host and shared by every consumer. */
typedef struct cbm_act_params {
    float msPerTimeStep;
    uint32_t numTSinMFHist;
    /* golgi */
    float eLeakGO, eGABAGO, threshMaxGO, threshRestGO;
    float gIncMFtoGO, gIncGRtoGO, gGABAIncGOtoGO, coupleRiRjRatioGO;
    float gMGluRScaleGRtoGO, gMGluRIncScaleGO, mGluRScaleGO, gluScaleGO, gLeakGO;
    float goGABAGOGOSynRec, goGABAGOGOSynDepF;
    float gDecMFtoGO, gDecGRtoGO, gGABADecGOtoGO;
    float mGluRDecayGO, gMGluRIncDecayGO, gMGluRDecGRtoGO, gluDecayGO, threshDecGO;
    /* granule */
    float eLeakGR, eGOGR, threshMaxGR, threshRestGR, gIncMFtoGR, gIncGOtoGR;
    float gDecMFtoGR, gDecGOtoGR, threshDecGR, gLeakGR;
    uint32_t tsPerHistBinGR;
    /* stellate */
    float eLeakSC, gLeakSC, gDecGRtoSC, threshMaxSC, threshRestSC, threshDecSC, gIncGRtoSC;
    /* basket */
    float eLeakBC, ePCtoBC, gLeakBC, gDecGRtoBC, gDecPCtoBC, threshDecBC, threshRestBC,
          threshMaxBC, gIncGRtoBC, gIncPCtoBC;
    /* purkinje */
    float initSynWofGRtoPC, eLeakPC, eBCtoPC, eSCtoPC, threshMaxPC, threshRestPC, threshDecPC,
          gLeakPC, gDecGRtoPC, gDecBCtoPC, gDecSCtoPC, gIncSCtoPC, gIncGRtoPC, gIncBCtoPC;
    uint32_t tsPopHistPC, tsPerPopHistBinPC, numPopHistBinsPC;
    float synLTPStepSizeGRtoPC, synLTDStepSizeGRtoPC;
    /* inferior olive */
    float coupleRiRjRatioIO, eLeakIO, eNCtoIO, gLeakIO, gDecTSofNCtoIO, gDecTTofNCtoIO,
          gDecT0ofNCtoIO, gIncNCtoIO, gIncTauNCtoIO, threshRestIO, threshMaxIO, threshDecIO;
    uint32_t tsLTDDurationIO;
    int32_t tsLTDStartAPIO, tsLTPStartAPIO;
    uint32_t grPCHistCheckBinIO;
    float maxExtIncVIO;
    /* nucleus */
    float eLeakNC, ePCtoNC, gmaxNMDADecMFtoNC, gmaxAMPADecMFtoNC, gNMDAIncMFtoNC, gAMPAIncMFtoNC,
          gIncAvgPCtoNC, gDecPCtoNC, gLeakNC, threshDecNC, threshMaxNC, threshRestNC,
          relPDecTSofNCtoIO, relPDecTTofNCtoIO, relPDecT0ofNCtoIO, relPIncNCtoIO, relPIncTauNCtoIO,
          initSynWofMFtoNC, synLTDPCPopActThreshMFtoNC, synLTPPCPopActThreshMFtoNC,
          synLTDStepSizeMFtoNC, synLTPStepSizeMFtoNC;
} cbm_act_params;

/* Options of this implementation (no reference counterpart). */
typedef struct cbm_options {
    /* Age, in steps, of the GR->GO input count that the Golgi update of step t
     * reads: 1 = the count of apBuf after step t-1 ("fresh"), 2 = after step
     * t-2 ("stale").  The reference reads a host buffer while an async copy
     * into it is in flight (innet.cpp:1236-1237 vs innet.cpp:835-838), so both
     * occur there; 2 is the default. */
    uint32_t grgoLag;
    /* HIP device ordinal to run on. */
    int32_t device;
    /* Granule cells per thread and threads per block of the GR kernel in the
     * two-stream step; 0 = pick (256 x 1).  The fused step uses 1024-thread
     * blocks with 2 cells per thread, or 1 when grCellsPerThread is 1. */
    uint32_t grCellsPerThread, grBlockThreads;
    /* How a step is launched: 0 = pick (1, with resident runs below where the
     * network allows them), 1 = one fused launch per step (block
     * 0 steps the small populations; the GR cells are split evenly over one
     * block per remaining CU when that gives each block 1,024 to 5,120 cells,
     * else into tiles of 1024 x 2 cells), 2 = the GR and small-population
     * kernels on two streams, ordered by events, 3 = fused with GR tiles
     * always, 4 = fused with the even split whenever it gives each block at
     * most 5,120 cells (else an error), 5 = as 4, and cbm_run_mf_train runs
     * of two or more steps are resident, or cbm_commit fails.  Resident:
     * the whole run is one launch of one workgroup per CU whose GR
     * workgroups hold their cells' state in registers (and connectivity in
     * LDS), and whose small-population workgroups hold theirs in LDS, from
     * the first step to the last, synchronised point to point between steps
     * (networks with the even split, compact connectivity records and an
     * input layer whose state fits in LDS; of the microzones' state, what
     * does not fit stays in global memory).  The resident launches of a process run
     * on a device one after another; a kernel of another stream or process
     * that keeps CUs busy for longer than 0.2 s while a resident run waits
     * for them makes that run fail (CBM_E_HIP at the next cbm_sync or call;
     * the network's state is then undefined).  All give the same bits. */
    uint32_t stepMode;
    /* Device layout of the GR connectivity: 0 = pick (one 16-byte record per
     * GR when numMF is at most 4095, numGO at most 2047 and the fan-ins at most 4
     * MF, 3 GO and 2 GO targets; else 29 bytes in five arrays), 1 = the
     * five arrays always.  Both give the same bits. */
    uint32_t connFormat;
} cbm_options;

/* ---------------------------------------------------------------- fields */
/* Every array that crosses the boundary.  Shapes are given by
 * cbm_field_shape() below.  Z = one array per microzone (zone argument). */
enum cbm_field {
    /* input-net connectivity (InNetConnectivityState, innetconnectivitystate.h:130-143) */
    CBM_F_NUMP_GR_FROM_MF_TO_GR = 0, CBM_F_P_GR_FROM_MF_TO_GR,
    CBM_F_NUMP_GR_FROM_GO_TO_GR, CBM_F_P_GR_FROM_GO_TO_GR,
    CBM_F_NUMP_GR_FROM_GR_TO_GO, CBM_F_P_GR_FROM_GR_TO_GO,
    CBM_F_P_GR_DELAYMASK_GR_TO_GO, CBM_F_P_GR_DELAYMASK_GR_TO_BSP,
    CBM_F_NUMP_MF_FROM_MF_TO_GO, CBM_F_P_MF_FROM_MF_TO_GO,
    CBM_F_NUMP_GO_GABA_OUT_GOGO, CBM_F_P_GO_GABA_OUT_GOGO,
    CBM_F_NUMP_GO_COUP_IN_GOGO, CBM_F_P_GO_COUP_IN_GOGO,
    /* microzone connectivity (Z) (MZoneConnectivityState) */
    CBM_F_P_BC_FROM_BC_TO_PC, CBM_F_P_PC_FROM_PC_TO_BC, CBM_F_P_PC_FROM_SC_TO_PC,
    CBM_F_P_NC_FROM_PC_TO_NC, CBM_F_P_NC_FROM_MF_TO_NC, CBM_F_P_IO_FROM_NC_TO_IO,
    CBM_F_P_IO_IN_IOIO,
    /* zone random generator (Z): seed, number of Random() draws discarded after
     * seeding (MZone::testReduction draws numGR, mzone.cpp:846-849), and the
     * full generator state (352 words + index) for snapshots */
    CBM_F_ZONE_RNG_SEED, CBM_F_ZONE_RNG_SKIP, CBM_F_ZONE_RNG_STATE,
    /* input-net activity (InNetActivityState, innetactivitystate.h:41-89) */
    CBM_F_HIST_MF, CBM_F_APBUF_MF,
    CBM_F_AP_GO, CBM_F_APBUF_GO, CBM_F_VM_GO, CBM_F_VCOUPLE_GO, CBM_F_THRESH_GO,
    CBM_F_INPUT_MF_GO, CBM_F_INPUT_GO_GO, CBM_F_INPUT_GOGABA_SYNDEP_GO,
    CBM_F_GO_GABA_OUT_SYNSCALE_GOGO,
    CBM_F_G_MF_GO, CBM_F_G_GR_GO, CBM_F_G_GO_GO, CBM_F_G_MGLUR_GO, CBM_F_G_MGLURINC_GO,
    CBM_F_MGLUR_GO, CBM_F_GLU_GO,
    CBM_F_AP_GR, CBM_F_APBUF_GR, CBM_F_G_MF_SUM_GR, CBM_F_G_GO_SUM_GR, CBM_F_THRESH_GR,
    CBM_F_VM_GR, CBM_F_GKCA_GR, CBM_F_HIST_GR,
    CBM_F_AP_SC, CBM_F_APBUF_SC, CBM_F_G_PF_SC, CBM_F_THRESH_SC, CBM_F_VM_SC,
    CBM_F_INPUT_SUM_PF_SC,
    /* microzone activity (Z) (MZoneActivityState, mzoneactivitystate.h:47-100) */
    CBM_F_AP_BC, CBM_F_APBUF_BC, CBM_F_INPUT_PC_BC, CBM_F_G_PF_BC, CBM_F_G_PC_BC,
    CBM_F_THRESH_BC, CBM_F_VM_BC,
    CBM_F_AP_PC, CBM_F_APBUF_PC, CBM_F_INPUT_BC_PC, CBM_F_INPUT_SC_PC, CBM_F_PFPC_SYNW,
    CBM_F_INPUT_SUM_PF_PC, CBM_F_G_PF_PC, CBM_F_G_BC_PC, CBM_F_G_SC_PC, CBM_F_VM_PC,
    CBM_F_THRESH_PC, CBM_F_HIST_PC_POPACT, CBM_F_PC_POPACT_CTRS,
    CBM_F_AP_IO, CBM_F_APBUF_IO, CBM_F_INPUT_NC_IO, CBM_F_ERR_DRIVE, CBM_F_G_NC_IO,
    CBM_F_THRESH_IO, CBM_F_VM_IO, CBM_F_VCOUPLE_IO, CBM_F_PFPC_PLAST_TIMER_IO,
    CBM_F_MFNC_NOLTP_NOLTD,
    CBM_F_AP_NC, CBM_F_APBUF_NC, CBM_F_INPUT_PC_NC, CBM_F_INPUT_MF_NC, CBM_F_G_PC_NC,
    CBM_F_G_PC_SCALE_NC, CBM_F_MF_SYNW_NC, CBM_F_GMAX_NMDA_MF_NC, CBM_F_GMAX_AMPA_MF_NC,
    CBM_F_G_MF_NMDA_NC, CBM_F_G_MF_AMPA_NC, CBM_F_THRESH_NC, CBM_F_VM_NC,
    CBM_F_SYN_IO_PRELEASE_NC,
    /* run-time values that live between steps but not in CBMState */
    CBM_F_CUR_TIME,          /* u64: steps taken (CBMSimCore::curTime)                  */
    CBM_F_MF_INPUT,          /* u8[numMF]: MF spikes used by the next step               */
    CBM_F_GINC_GR_TO_GO,     /* f32: InNet::tempGIncGRtoGO                               */
    CBM_F_GRPC_PLAST_STEPS,  /* f32[2] (Z): MZone::tempGRPCLTDStep, tempGRPCLTPStep      */
    CBM_F_SUM_GR_INPUT_GO,   /* u32[numGO]: exportSumGRInputGO                           */
    CBM_F_SUM_GO_INPUT_GO,   /* f32[numGO]: exportSumGOInputGO                           */
    /* pipeline of population input sums; row 0 = sum over apBuf after the last
     * step, row 1 = the step before (the two-step PF lag, cbmsimcore.cpp:139-310) */
    CBM_F_PIPE_PFPC_SUM,     /* f32[2][numPC] (Z)                                        */
    CBM_F_PIPE_PFBC_SUM,     /* u32[2][numBC]                                            */
    CBM_F_PIPE_PFSC_SUM,     /* u32[2][numSC]                                            */
    CBM_F_PIPE_GRGO_SUM,     /* u32[2][numGO]                                            */
    CBM_F_COUNT
};

enum cbm_field_kind { CBM_KIND_CON = 0, CBM_KIND_STATE = 1, CBM_KIND_RUNTIME = 2 };

/* Shape of a field: element size in bytes, rows x cols elements, whether it is
 * per zone, and its kind.  Returns 0, or CBM_E_ARG for an unknown field. */
static inline int cbm_field_shape(int f, const cbm_sizes *s, size_t *elem, size_t *rows,
                                  size_t *cols, int *zoned, int *kind)
{
    size_t e = 4, r = 1, c = 1;
    int z = 0, k = CBM_KIND_STATE;
    switch (f) {
    case CBM_F_NUMP_GR_FROM_MF_TO_GR: case CBM_F_NUMP_GR_FROM_GO_TO_GR:
    case CBM_F_NUMP_GR_FROM_GR_TO_GO: case CBM_F_P_GR_DELAYMASK_GR_TO_BSP:
        r = s->numGR; k = CBM_KIND_CON; break;
    case CBM_F_P_GR_FROM_MF_TO_GR: r = s->numGR; c = s->maxnumpGRfromMFtoGR; k = CBM_KIND_CON; break;
    case CBM_F_P_GR_FROM_GO_TO_GR: r = s->numGR; c = s->maxnumpGRfromGOtoGR; k = CBM_KIND_CON; break;
    case CBM_F_P_GR_FROM_GR_TO_GO: case CBM_F_P_GR_DELAYMASK_GR_TO_GO:
        r = s->numGR; c = s->maxnumpGRfromGRtoGO; k = CBM_KIND_CON; break;
    case CBM_F_NUMP_MF_FROM_MF_TO_GO: r = s->numMF; k = CBM_KIND_CON; break;
    case CBM_F_P_MF_FROM_MF_TO_GO: r = s->numMF; c = s->maxnumpMFfromMFtoGO; k = CBM_KIND_CON; break;
    case CBM_F_NUMP_GO_GABA_OUT_GOGO: case CBM_F_NUMP_GO_COUP_IN_GOGO:
        r = s->numGO; k = CBM_KIND_CON; break;
    case CBM_F_P_GO_GABA_OUT_GOGO: r = s->numGO; c = s->maxnumpGOGABAOutGOGO; k = CBM_KIND_CON; break;
    case CBM_F_P_GO_COUP_IN_GOGO: r = s->numGO; c = s->maxnumpGOCoupInGOGO; k = CBM_KIND_CON; break;
    case CBM_F_P_BC_FROM_BC_TO_PC: r = s->numBC; c = s->numpBCfromBCtoPC; z = 1; k = CBM_KIND_CON; break;
    case CBM_F_P_PC_FROM_PC_TO_BC: r = s->numPC; c = s->numpPCfromPCtoBC; z = 1; k = CBM_KIND_CON; break;
    case CBM_F_P_PC_FROM_SC_TO_PC: r = s->numPC; c = s->numpPCfromSCtoPC; z = 1; k = CBM_KIND_CON; break;
    case CBM_F_P_NC_FROM_PC_TO_NC: r = s->numNC; c = s->numpNCfromPCtoNC; z = 1; k = CBM_KIND_CON; break;
    case CBM_F_P_NC_FROM_MF_TO_NC: r = s->numNC; c = s->numpNCfromMFtoNC; z = 1; k = CBM_KIND_CON; break;
    case CBM_F_P_IO_FROM_NC_TO_IO: r = s->numIO; c = s->numpIOfromNCtoIO; z = 1; k = CBM_KIND_CON; break;
    case CBM_F_P_IO_IN_IOIO: r = s->numIO; c = s->numpIOInIOIO; z = 1; k = CBM_KIND_CON; break;
    case CBM_F_ZONE_RNG_SEED: z = 1; k = CBM_KIND_CON; break;
    case CBM_F_ZONE_RNG_SKIP: e = 8; z = 1; k = CBM_KIND_CON; break;
    case CBM_F_ZONE_RNG_STATE: r = 353; z = 1; k = CBM_KIND_RUNTIME; break;

    case CBM_F_HIST_MF: e = 1; r = s->numMF; break;
    case CBM_F_APBUF_MF: r = s->numMF; break;
    case CBM_F_AP_GO: e = 1; r = s->numGO; break;
    case CBM_F_APBUF_GO: case CBM_F_VM_GO: case CBM_F_VCOUPLE_GO: case CBM_F_THRESH_GO:
    case CBM_F_INPUT_MF_GO: case CBM_F_INPUT_GO_GO: case CBM_F_INPUT_GOGABA_SYNDEP_GO:
    case CBM_F_GO_GABA_OUT_SYNSCALE_GOGO: case CBM_F_G_MF_GO: case CBM_F_G_GR_GO:
    case CBM_F_G_GO_GO: case CBM_F_G_MGLUR_GO: case CBM_F_G_MGLURINC_GO: case CBM_F_MGLUR_GO:
    case CBM_F_GLU_GO:
        r = s->numGO; break;
    case CBM_F_AP_GR: e = 1; r = s->numGR; break;
    case CBM_F_APBUF_GR: case CBM_F_G_MF_SUM_GR: case CBM_F_G_GO_SUM_GR: case CBM_F_THRESH_GR:
    case CBM_F_VM_GR: case CBM_F_GKCA_GR:
        r = s->numGR; break;
    case CBM_F_HIST_GR: e = 8; r = s->numGR; break;
    case CBM_F_AP_SC: e = 1; r = s->numSC; break;
    case CBM_F_APBUF_SC: case CBM_F_G_PF_SC: case CBM_F_THRESH_SC: case CBM_F_VM_SC:
    case CBM_F_INPUT_SUM_PF_SC:
        r = s->numSC; break;

    case CBM_F_AP_BC: e = 1; r = s->numBC; z = 1; break;
    case CBM_F_APBUF_BC: case CBM_F_INPUT_PC_BC: case CBM_F_G_PF_BC: case CBM_F_G_PC_BC:
    case CBM_F_THRESH_BC: case CBM_F_VM_BC:
        r = s->numBC; z = 1; break;
    case CBM_F_AP_PC: e = 1; r = s->numPC; z = 1; break;
    case CBM_F_APBUF_PC: case CBM_F_INPUT_BC_PC: case CBM_F_INPUT_SUM_PF_PC: case CBM_F_G_PF_PC:
    case CBM_F_G_BC_PC: case CBM_F_VM_PC: case CBM_F_THRESH_PC:
        r = s->numPC; z = 1; break;
    case CBM_F_INPUT_SC_PC: e = 1; r = s->numPC; c = s->numpPCfromSCtoPC; z = 1; break;
    case CBM_F_G_SC_PC: r = s->numPC; c = s->numpPCfromSCtoPC; z = 1; break;
    case CBM_F_PFPC_SYNW: r = s->numPC; c = s->numpPCfromGRtoPC; z = 1; break;
    case CBM_F_HIST_PC_POPACT: r = s->numPopHistBinsPC; z = 1; break;
    case CBM_F_PC_POPACT_CTRS: r = 3; z = 1; break;
    case CBM_F_AP_IO: e = 1; r = s->numIO; z = 1; break;
    case CBM_F_APBUF_IO: case CBM_F_THRESH_IO: case CBM_F_VM_IO: case CBM_F_VCOUPLE_IO:
    case CBM_F_PFPC_PLAST_TIMER_IO:
        r = s->numIO; z = 1; break;
    case CBM_F_INPUT_NC_IO: e = 1; r = s->numIO; c = s->numpIOfromNCtoIO; z = 1; break;
    case CBM_F_G_NC_IO: r = s->numIO; c = s->numpIOfromNCtoIO; z = 1; break;
    case CBM_F_ERR_DRIVE: z = 1; break;
    case CBM_F_MFNC_NOLTP_NOLTD: e = 1; r = 2; z = 1; break;
    case CBM_F_AP_NC: e = 1; r = s->numNC; z = 1; break;
    case CBM_F_APBUF_NC: case CBM_F_THRESH_NC: case CBM_F_VM_NC: case CBM_F_SYN_IO_PRELEASE_NC:
        r = s->numNC; z = 1; break;
    case CBM_F_INPUT_PC_NC: e = 1; r = s->numNC; c = s->numpNCfromPCtoNC; z = 1; break;
    case CBM_F_G_PC_NC: case CBM_F_G_PC_SCALE_NC: r = s->numNC; c = s->numpNCfromPCtoNC; z = 1; break;
    case CBM_F_INPUT_MF_NC: e = 1; r = s->numNC; c = s->numpNCfromMFtoNC; z = 1; break;
    case CBM_F_MF_SYNW_NC: case CBM_F_GMAX_NMDA_MF_NC: case CBM_F_GMAX_AMPA_MF_NC:
    case CBM_F_G_MF_NMDA_NC: case CBM_F_G_MF_AMPA_NC:
        r = s->numNC; c = s->numpNCfromMFtoNC; z = 1; break;

    case CBM_F_CUR_TIME: e = 8; k = CBM_KIND_RUNTIME; break;
    case CBM_F_MF_INPUT: e = 1; r = s->numMF; k = CBM_KIND_RUNTIME; break;
    case CBM_F_GINC_GR_TO_GO: k = CBM_KIND_RUNTIME; break;
    case CBM_F_GRPC_PLAST_STEPS: r = 2; z = 1; k = CBM_KIND_RUNTIME; break;
    case CBM_F_SUM_GR_INPUT_GO: case CBM_F_SUM_GO_INPUT_GO: r = s->numGO; k = CBM_KIND_RUNTIME; break;
    case CBM_F_PIPE_PFPC_SUM: r = 2; c = s->numPC; z = 1; k = CBM_KIND_RUNTIME; break;
    case CBM_F_PIPE_PFBC_SUM: r = 2; c = s->numBC; k = CBM_KIND_RUNTIME; break;
    case CBM_F_PIPE_PFSC_SUM: r = 2; c = s->numSC; k = CBM_KIND_RUNTIME; break;
    case CBM_F_PIPE_GRGO_SUM: r = 2; c = s->numGO; k = CBM_KIND_RUNTIME; break;
    default: return CBM_E_ARG;
    }
    if (elem) *elem = e;
    if (rows) *rows = r;
    if (cols) *cols = c;
    if (zoned) *zoned = z;
    if (kind) *kind = k;
    return CBM_OK;
}

/* 1 if the field holds IEEE binary32 values, 0 if integers. */
static inline int cbm_field_is_float(int f)
{
    switch (f) {
    case CBM_F_VM_GO: case CBM_F_VCOUPLE_GO: case CBM_F_THRESH_GO:
    case CBM_F_INPUT_GOGABA_SYNDEP_GO: case CBM_F_GO_GABA_OUT_SYNSCALE_GOGO:
    case CBM_F_G_MF_GO: case CBM_F_G_GR_GO: case CBM_F_G_GO_GO: case CBM_F_G_MGLUR_GO:
    case CBM_F_G_MGLURINC_GO: case CBM_F_MGLUR_GO: case CBM_F_GLU_GO:
    case CBM_F_G_MF_SUM_GR: case CBM_F_G_GO_SUM_GR: case CBM_F_THRESH_GR: case CBM_F_VM_GR:
    case CBM_F_GKCA_GR: case CBM_F_G_PF_SC: case CBM_F_THRESH_SC: case CBM_F_VM_SC:
    case CBM_F_G_PF_BC: case CBM_F_G_PC_BC: case CBM_F_THRESH_BC: case CBM_F_VM_BC:
    case CBM_F_PFPC_SYNW: case CBM_F_INPUT_SUM_PF_PC: case CBM_F_G_PF_PC: case CBM_F_G_BC_PC:
    case CBM_F_G_SC_PC: case CBM_F_VM_PC: case CBM_F_THRESH_PC: case CBM_F_ERR_DRIVE:
    case CBM_F_G_NC_IO: case CBM_F_THRESH_IO: case CBM_F_VM_IO: case CBM_F_VCOUPLE_IO:
    case CBM_F_G_PC_NC: case CBM_F_G_PC_SCALE_NC: case CBM_F_MF_SYNW_NC:
    case CBM_F_GMAX_NMDA_MF_NC: case CBM_F_GMAX_AMPA_MF_NC: case CBM_F_G_MF_NMDA_NC:
    case CBM_F_G_MF_AMPA_NC: case CBM_F_THRESH_NC: case CBM_F_VM_NC:
    case CBM_F_SYN_IO_PRELEASE_NC: case CBM_F_GINC_GR_TO_GO: case CBM_F_GRPC_PLAST_STEPS:
    case CBM_F_SUM_GO_INPUT_GO: case CBM_F_PIPE_PFPC_SUM:
        return 1;
    default:
        return 0;
    }
}

static inline size_t cbm_field_bytes(int f, const cbm_sizes *s)
{
    size_t e = 0, r = 0, c = 0;
    if (cbm_field_shape(f, s, &e, &r, &c, 0, 0) != CBM_OK) return 0;
    return e * r * c;
}

/* ---------------------------------------------------------------- handle */
typedef struct cbm_handle cbm_handle;

/* Version of the ABI the library was built with (CBM_ABI_VERSION). */
int cbm_abi_version(void);

/* Message for the last failing call on this thread ("" if none). */
const char *cbm_last_error(void);

/* Number of HIP devices visible (0 if none). */
int cbm_device_count(void);

/* Create a network of the given sizes and derived constants.  opts may be
 * NULL (defaults).  Connectivity and initial state are then given with
 * cbm_write() and the network is made runnable with cbm_commit(). */
int cbm_create(const cbm_sizes *sizes, const cbm_act_params *params, const cbm_options *opts,
               cbm_handle **out);
int cbm_destroy(cbm_handle *h);

/* Copy `bytes` bytes from host memory into a field (bytes must equal
 * cbm_field_bytes).  Connectivity fields may only be written before
 * cbm_commit(); state and run-time fields at any time. */
int cbm_write(cbm_handle *h, int field, int zone, const void *src, size_t bytes);

/* Validate connectivity, build the device layouts and upload everything. */
int cbm_commit(cbm_handle *h);

/* Synchronous copy of a field into host memory. */
int cbm_read(cbm_handle *h, int field, int zone, void *dst, size_t bytes);

/* updateMFInput: numMF bytes, nonzero = spike, used by the next step. */
int cbm_set_mf(cbm_handle *h, const uint8_t *mf);
/* updateErrDrive: errDrive of the zone = rel * maxExtIncVIO for the next step. */
int cbm_set_err_drive(cbm_handle *h, int zone, float rel);
/* setGRPCPlastSteps; cbm_reset_grpc_steps restores the parameter values. */
int cbm_set_grpc_steps(cbm_handle *h, int zone, float ltd, float ltp);
int cbm_reset_grpc_steps(cbm_handle *h, int zone);
/* setGIncGRtoGO; cbm_reset_ginc_grgo restores gIncGRtoGO. */
int cbm_set_ginc_grgo(cbm_handle *h, float inc);
int cbm_reset_ginc_grgo(cbm_handle *h);

/* n 1-ms steps with the MF input last set (calcActivity n times).  Returns
 * after the work is queued; cbm_sync() waits for it. */
int cbm_step(cbm_handle *h, int n);

/* The per-ms round trip of a closed-loop driver in one call pair: the MF
 * input from host memory, one step, and the spikes of one population of
 * every zone back to the host -- what SimThread::run and Microzone::update do
 * with updateMFInput, calcActivity and exportAPNC per zone
 * (Matthew_Sim_CBM/src/simthread.cpp:105-112, src/microzone.cpp:31-33), with
 * one host synchronisation instead of one per call.  _begin queues the copy
 * in and the step, whose small-population kernel stores the population's
 * spikes into pinned host memory (population: CBM_REC_PC, _BC, _SC, _IO, _NC
 * or _GO, not _PF), and returns; the host may work meanwhile (the next step's
 * draws, say).  _end waits and writes the spikes, zones stacked
 * ([numZones][cells] for the microzone populations), to out.  Every _begin
 * needs its _end before the next _begin. */
int cbm_step_io_begin(cbm_handle *h, const uint8_t *mf, int population);
int cbm_step_io_end(cbm_handle *h, uint8_t *out);

/* A closed loop on a resident network (cbm_resident): cbm_loop_begin
 * launches one resident run of up to max_steps steps that takes each step's
 * input from pinned host memory, and the cbm_step_io_begin / _end pairs that
 * follow (population: the one given here, CBM_REC_PC, _BC, _IO or _NC) post
 * the step's MF bytes and the error drive set by cbm_set_err_drive, and
 * collect the step's spikes, with no launch and no stream synchronisation per
 * step.  The results are those of the same pairs without the loop.
 * cbm_loop_end stops the run after the steps posted (*steps: their number)
 * and reports a wait that timed out as CBM_E_HIP.  Between the two only
 * cbm_step_io_begin / _end and cbm_set_err_drive may be called on the handle
 * (other calls fail with CBM_E_STATE); the host must post each step within
 * 2 s of the last one's output, or the run ends with an error.  The loop's
 * launch holds every CU until cbm_loop_end: kernels of other handles or
 * streams of the process wait for it.  No reference
 * counterpart: it serves the closed-loop drivers of SimThread::run
 * (Matthew_Sim_CBM/src/simthread.cpp:60-165), whose per-ms round trip it
 * shortens. */
int cbm_loop_begin(cbm_handle *h, int max_steps, int population);
int cbm_loop_end(cbm_handle *h, int *steps);

/* n steps where step k reads MF spikes from mf_dev + k*pitch (a DEVICE
 * pointer to n rows of numMF bytes, e.g. a pre-generated spike train).
 * err_drive_host, if not NULL, gives [n][numZones] relative error drives
 * (0 = none) applied before each step.  A run of two or more steps is
 * resident with cbm_options.stepMode 0 or 5 when the network allows it
 * (cbm_resident); a resident run
 * whose workgroups were not all resident at once ends after a bounded wait
 * with CBM_E_HIP at the next call that waits for the handle's work, and the
 * network's state is then undefined. */
int cbm_run_mf_train(cbm_handle *h, const uint8_t *mf_dev, size_t pitch, int n,
                     const float *err_drive_host);

/* Wait for all queued work of the handle. */
int cbm_sync(cbm_handle *h);

/* The HIP stream (hipStream_t) the handle launches on, as an opaque pointer. */
void *cbm_stream(cbm_handle *h);

/* Time the dominant kernel over the next n steps with HIP events on the
 * stream it is launched on: runs them with the MF train as in
 * cbm_run_mf_train and reports the mean duration of the step kernel (fused
 * modes: one event pair around the n back-to-back launches, divided by n) or
 * of the GR kernel (two-stream mode: an event pair around each launch), and
 * the mean whole-step duration, in microseconds. */
int cbm_profile_steps(cbm_handle *h, const uint8_t *mf_dev, size_t pitch, int n,
                      double *kernel_us, double *step_us);

/* Run the next n steps as a GR-kernel launch followed by a small-population
 * launch, in order on one stream, and report each kernel's mean duration in
 * microseconds (HIP events).  The results equal those of the normal step. */
int cbm_profile_split(cbm_handle *h, const uint8_t *mf_dev, size_t pitch, int n,
                      double *gr_kernel_us, double *small_kernel_us);

/* Streaming ceiling of a per-cell pass on `device` (a measurement aid; no
 * counterpart in the reference): `cells` cells, each reading `read_words`
 * 4-byte values (one per array) and one 16-byte record and writing the first
 * `write_words` of the values back in place -- the GR step's per-cell streams
 * (read_words 7, write_words 6: five state values, the spike register, the
 * PF->PC weight; the connectivity record).  Non-temporal loads and stores,
 * 1024-thread blocks of 4,096 contiguous cells.  Allocates the arrays,
 * launches `launches` times after 10 untimed launches and reports the mean
 * launch duration in microseconds (HIP events) and the bytes one launch
 * moves.  read_words 1..8, write_words <= read_words. */
int cbm_stream_ceiling(int device, uint64_t cells, int read_words, int write_words, int launches,
                       double *us_per_launch, uint64_t *bytes_per_launch);

/* The launch mode the handle uses (1, 3, 4 or 5 = fused, 2 = two streams; see
 * cbm_options.stepMode), or a negative CBM_E_* code. */
int cbm_step_mode(cbm_handle *h);

/* 1 when cbm_run_mf_train runs of two or more steps are resident (see
 * cbm_options.stepMode), 0 when they launch step by step, or a negative
 * CBM_E_* code.  No reference counterpart. */
int cbm_resident(cbm_handle *h);

/* How the step's GR work is launched: *blocks GR workgroups of *threads
 * threads with *cellsPerThread cells each at most, *evenSplit = 1 when the
 * cells are split evenly over the blocks (one per CU beside the small step)
 * rather than into power-of-two tiles. */
int cbm_gr_launch(cbm_handle *h, uint32_t *blocks, uint32_t *threads, uint32_t *cellsPerThread,
                  uint32_t *evenSplit);

/* ---------------------------------------------------------------- recorder */
/* Populations the spike recorder can keep (bit masks for cbm_set_recorder, and
 * the population argument of cbm_read_recorded). */
enum cbm_rec_population {
    CBM_REC_PC = 0, CBM_REC_BC = 1, CBM_REC_SC = 2, CBM_REC_IO = 3, CBM_REC_NC = 4, CBM_REC_GO = 5,
    /* PF spikes as the PCs, BCs and SCs see them (after each GR's PF delay),
     * counted per 64-GR chunk: numGR/64 bytes per step, values 0..64.  Only
     * single-network steps write them (not batched runs). */
    CBM_REC_PF = 6
};

/* Spike recorder for trial readouts (the per-ms exportAP* calls of the
 * reference's trial drivers, e.g. Eyelid_Sim_CBM_Dev/src/ectrial/
 * ecmanagementdelay.cpp:233-266, without a host round trip per step): from
 * the next step on, step t writes the spike bytes of every population whose
 * bit (1 << CBM_REC_x) is in `mask` into row t mod `capacity` of a device
 * ring.  capacity 0 (or mask 0) turns recording off. */
int cbm_set_recorder(cbm_handle *h, uint32_t mask, uint32_t capacity);

/* Copy the recorded spikes of one population for steps t0 .. t0+n-1 (all
 * still in the ring) into dst, [n][cells] bytes with zones stacked.  dst may
 * be host or device memory. */
int cbm_read_recorded(cbm_handle *h, int population, uint64_t t0, int n, uint8_t *dst);

/* ---------------------------------------------------------------- batches */
/* A ragged pack of committed networks stepped together, one launch per step
 * for all of them: block n < count steps network n's small populations, the
 * other blocks are the GR tiles of every network from a tile table built
 * largest network first.  Each network keeps its own sizes, parameters, seeds
 * and state, and a batched run leaves every network exactly as its own
 * cbm_run_mf_train would.  No reference counterpart: the reference steps one
 * network per process. */
typedef struct cbm_batch cbm_batch;

typedef struct cbm_batch_options {
    /* GR cells per thread of the batched tiles: 0 = 2, or 1 or 2 (see
     * cbm_batch_autotune). */
    uint32_t cellsPerThread;
    /* Steps per captured HIP graph (0 = plain launches): runs of at least this
     * many steps replay one graph per window. */
    uint32_t graphSteps;
} cbm_batch_options;

/* The networks must be committed, share a device, have numGR a multiple of
 * 2048 and belong to no other batch.  They stay owned by the caller and must
 * outlive the batch (cbm_destroy refuses a network still in a batch).  Between
 * batched runs each network can be read, written and stepped on its own. */
int cbm_batch_create(cbm_handle *const *nets, int count, const cbm_batch_options *opts, cbm_batch **out);
int cbm_batch_destroy(cbm_batch *b);

/* n steps of every network.  Network i reads MF row k at mf_dev[i] + k *
 * pitch[i] (device memory, as cbm_run_mf_train) and, when err_host and
 * err_host[i] are non-null, the relative error drive err_host[i][k *
 * numZones + z] (0 = none). */
int cbm_batch_run_mf_train(cbm_batch *b, const uint8_t *const *mf_dev, const size_t *pitch, int n,
                           const float *const *err_host);

/* cbm_batch_run_mf_train without allocating: no counterpart in the reference
 * (which runs one network per process and has no batch).  It is for callers
 * that cannot allocate device memory -- code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc: it calls no hipMalloc,
 * hipFree or other allocating routine (cbm_batch_run_mf_train may grow its
 * error-drive staging buffer and instantiate graphs).  The error-drive rows
 * are staged in a fixed device array of CBM_NOALLOC_ERR_FLOATS floats, one
 * per GPU, held until the process ends; a run whose n x (sum of numZones over
 * the networks given an error drive) exceeds it fails with CBM_E_ARG.  Steps
 * are launched one by one (no graph).  Every call on one GPU shares that
 * array: calls that overlap on one GPU (other threads, or another batch's run
 * still in flight on another stream) race. */
#define CBM_NOALLOC_ERR_FLOATS (1u << 18)
int cbm_batch_run_mf_train_noalloc(cbm_batch *b, const uint8_t *const *mf_dev, const size_t *pitch, int n,
                                   const float *const *err_host);

/* Wait for the batch's queued work. */
int cbm_batch_sync(cbm_batch *b);

/* n batched steps (as cbm_batch_run_mf_train without error drive), reporting
 * the mean time per batched step in microseconds (HIP events around the
 * launches on the batch's stream). */
int cbm_batch_profile(cbm_batch *b, const uint8_t *const *mf_dev, const size_t *pitch, int n, double *step_us);

/* Autotuner of the GR tile shape: times `steps` batched steps with 1 and with
 * 2 GR cells per thread on the batch's own networks (each trial starts from a
 * device snapshot of every network), keeps the faster shape, and restores
 * every network to its state before the call. */
int cbm_batch_autotune(cbm_batch *b, const uint8_t *const *mf_dev, const size_t *pitch, int steps,
                       int *chosen_cells_per_thread);

/* Current GR cells per thread, launch blocks per step and graph window. */
int cbm_batch_info(cbm_batch *b, int *cells_per_thread, int *blocks, int *graph_steps);

/* cbm_field_shape / cbm_field_is_float as exported functions, for hosts that
 * cannot use the inline versions (FFI bindings). */
int cbm_field_info(int field, const cbm_sizes *s, size_t *elem, size_t *rows, size_t *cols,
                   int *zoned, int *kind, int *is_float);

/* Bytes of device memory held by the handle. */
int cbm_device_bytes(cbm_handle *h, size_t *bytes);

#ifdef __cplusplus
}
#endif

#endif /* CBM_HIP_H */

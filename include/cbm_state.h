/*
 * cbm_state.h -- C-ABI of the network-state library (CBMState side).
 *
 * Builds what CBMSimCore consumes: ConnectivityParams / ActivityParams parsed
 * from the reference's key/value text files, connectivity, and initial
 * activity state, all in the CBMState layout and with the field vocabulary of
 * cbm_hip.h.  It replaces, for non-C++ hosts, the constructors of
 *   CBMState                 CBM_STATE_LIB_DEV/src/interfaces/cbmstate.cpp:12-204
 *   ConnectivityParams       CBM_STATE_LIB_DEV/src/params/connectivityparams.cpp:12-184
 *   ActivityParams           CBM_STATE_LIB_DEV/src/params/activityparams.cpp:12-42,436-636
 *   InNet/MZone activity     CBM_STATE_LIB_DEV/src/state/innetactivitystate.cpp:277-337,
 *                            CBM_STATE_LIB_DEV/src/state/mzoneactivitystate.cpp:357-445
 *   MZoneConnectivityState   CBM_STATE_LIB_DEV/src/state/mzoneconnectivitystate.cpp:190-417
 * and adds the seeded synthetic input-net generator the benchmark networks use.
 */
#ifndef CBM_STATE_H
#define CBM_STATE_H

#include <stddef.h>
#include <stdint.h>
#include "cbm_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct cbm_state cbm_state;

/* Input-net generators. */
enum cbm_gen_kind {
    /* Each GR draws nMFperGR distinct MFs and nGOperGR distinct GOs uniformly,
     * and projects to nGOoutPerGR distinct GOs (uniformly, or within a window
     * of the GO grid when grgoWindowX > 0); each GO draws nMFperGO distinct MFs.
     * Delay masks follow InNetConnectivityState::assignGRDelays
     * (innetconnectivitystate.cpp:920-962); GO-GO wiring follows connectGOGO
     * (innetconnectivitystate.cpp:852-918). */
    CBM_GEN_UNIFORM = 0,
    /* The reference's own generator (InNetConnectivityState, CBM_STATE_LIB_DEV/
     * src/state/innetconnectivitystate.cpp:13-44,716-1122): GR dendrites, GO
     * dendrites and GO axons placed on the glomerulus grid by the spatial
     * connector, MFs given glomeruli at random, the MF->GR, MF->GO and GO->GR
     * lists relayed through the glomeruli, GO ascending dendrites onto GRs,
     * then GO-GO wiring and delays, all from one SFMT stream seeded with
     * innetSeed in the reference's order.  Sizes and spans come from the
     * conParams keys (glXP2, glYP2, maxnumpGLfromGLtoGR, lownumpGLfromGLtoGR,
     * spanGRDenOnGLX, ...; CBM_Params/conParams.txt holds them); the nMFperGR
     * .. grgoWindow fields are not used.  A checkpoint of such a state holds
     * the glomerulus layer and every list as the reference's does. */
    CBM_GEN_REFERENCE = 1
};

typedef struct cbm_gen_spec {
    int32_t kind;            /* cbm_gen_kind                                    */
    uint32_t numZones;
    int32_t innetSeed;       /* seed of the input-net generator                 */
    uint32_t nMFperGR, nGOperGR, nGOoutPerGR, nMFperGO;
    uint32_t grgoWindowX, grgoWindowY; /* 0 = uniform GR->GO targets; else a
                                          window in GO-grid cells            */
    int32_t mzoneConSeed[8]; /* MZoneConnectivityState seeds (PC->NC wiring)    */
    int32_t mzoneRngSeed[8]; /* MZone generator seeds (NC->IO release draws)    */
} cbm_gen_spec;

/* Parse the text of an actParams file and derive the constants.  Unknown keys
 * are kept; missing keys read as 0 (std::map semantics of the reference). */
int cbm_act_params_from_text(const char *text, cbm_act_params *out);

/* Build a complete network state from the texts of an actParams and a
 * conParams file and a generator spec. */
int cbm_state_create(const char *actText, const char *conText, const cbm_gen_spec *gen,
                     cbm_state **out);
int cbm_state_destroy(cbm_state *st);

int cbm_state_sizes(const cbm_state *st, cbm_sizes *out);
int cbm_state_params(const cbm_state *st, cbm_act_params *out);

/* Entry i, in key order, of a parameter file as parsed: which = 0 the actParams
 * file (ActivityParams::paramMap; ActivityParams::getParamCopy / getParam,
 * CBM_STATE_LIB_DEV/src/params/activityparams.cpp:404-421), 1 the conParams
 * file (ConnectivityParams::paramMap; ConnectivityParams::getParamCopy,
 * connectivityparams.cpp:372-390).  *name stays valid until the state is
 * destroyed or an activity parameter is set.  CBM_E_ARG past the last entry. */
int cbm_state_param_at(const cbm_state *st, int which, size_t i, const char **name, double *value);

/* ActivityParams::setParam (activityparams.cpp:423-434): change an activity
 * parameter the file gave and re-derive the step's constants
 * (cbm_state_params).  CBM_E_ARG for a name the file did not give.  A
 * simulation handle loaded from the state before keeps the constants it was
 * given (the reference's modules read the parameter object every step). */
int cbm_state_set_act_param(cbm_state *st, const char *name, float value);

/* Pointer to the state's storage of a field (CBMState layout) and its size in
 * bytes.  The pointer stays valid until cbm_state_destroy. */
int cbm_state_field(cbm_state *st, int field, int zone, void **ptr, size_t *bytes);

/* Hand every connectivity, state and generator-seed field of the state to a
 * simulation handle (cbm_write for each) and commit it. */
int cbm_state_load_into(cbm_state *st, cbm_handle *h);

/* Outputs of the library's own SFMT generator (the CRandomSFMT0 contract the
 * state and core code rely on), for checking against the reference's:
 * kind 0 = BRandom() as uint32, 1 = Random() as double after `skip` Random()
 * calls, 2 = IRandom(lo,hi) and 3 = IRandomX(lo,hi) as int32. */
int cbm_sfmt_sample(int seed, uint64_t skip, int kind, int lo, int hi, int n, void *out);

/* ---------------------------------------------------------------- checkpoints
 * The CBMState file format (CBMState::writeState / CBMState(fstream&),
 * CBM_STATE_LIB_DEV/src/interfaces/cbmstate.cpp:12-33,86-103): the zone count,
 * the connectivity and activity parameters as text
 * (connectivityparams.cpp:192-218, activityparams.cpp:243-253), then the raw
 * arrays of InNetConnectivityState, each zone's MZoneConnectivityState,
 * InNetActivityState and each zone's MZoneActivityState, in the order and
 * sizes of their stateRW methods (innetconnectivitystate.cpp:530-606,
 * mzoneconnectivitystate.cpp:142-163, innetactivitystate.cpp:233-275,
 * mzoneactivitystate.cpp:299-355).  As there:
 *  - writing sets the GO synaptic-depression fields INPUT_GOGABA_SYNDEP_GO to 1
 *    and GO_GABA_OUT_SYNSCALE_GOGO to 0 in the state (innetactivitystate.cpp:247-248),
 *    and reading sets them the same way;
 *  - the MZone section holds pBCfromPCtoBC twice and never pPCfromPCtoBC
 *    (mzoneconnectivitystate.cpp:145,150).  A read rebuilds P_PC_FROM_PC_TO_BC
 *    as connectPCtoBC makes it (the reference leaves it unset).
 * Arrays the step does not use (the glomerulus layer, reverse lists,
 * per-synapse GR conductances) are kept as read and written back unchanged,
 * so a read followed by a write reproduces the file.  A state built by
 * cbm_state_create has no glomerulus layer: its header gets one glomerulus per
 * MF and the column counts of its own lists; the lists it does not hold are
 * written as the reference's constructors leave them (mzoneconnectivitystate.cpp:
 * 190-417) or as reverse lists of the ones it holds.  Not in the format: the
 * step counter, the pipeline rows, the zone generators (seeded by CBMSimCore),
 * and the MZoneActivityState members its stateRW skips
 * (mzoneactivitystate.h:69-86 against mzoneactivitystate.cpp:299-355):
 * PC_POPACT_CTRS, ERR_DRIVE and MFNC_NOLTP_NOLTD.  The reference leaves those
 * uninitialised after a read (mzoneactivitystate.cpp:21-29); here they read
 * as zero. */
int cbm_state_write_checkpoint(cbm_state *st, const char *path);
int cbm_state_read_checkpoint(const char *path, cbm_state **out);
/* The same to and from memory: *data is allocated (free it with
 * cbm_state_free_memory); *consumed receives the bytes the state took. */
int cbm_state_checkpoint_to_memory(cbm_state *st, void **data, size_t *bytes);
int cbm_state_checkpoint_from_memory(const void *data, size_t bytes, cbm_state **out, size_t *consumed);
void cbm_state_free_memory(void *p);

/* Message of the last failing call on this thread. */
const char *cbm_state_last_error(void);

#ifdef __cplusplus
}
#endif

#endif

/*
 * orc_innet.h -- oracle restatement of the reference's input-net
 * connectivity generator.
 *
 * TEST INFRASTRUCTURE: only tests/ use it, as the checker of the library's
 * glomerular generator (csrc/state/innet_gen.cpp, CBM_GEN_REFERENCE).
 *
 * Restated from CBM_STATE_LIB_DEV/src/state/innetconnectivitystate.cpp:
 *   InNetConnectivityState(params, msPerStep, seed)  13-44 (call order)
 *   initializeVals   607-714 (empty lists: count 0, slots UINT_MAX)
 *   connectGRGL      716-732     connectGOGL   734-753
 *   connectMFGL      755-796     translateMFGL 798-829
 *   translateGOGL    831-838     connectGRGO   840-850
 *   connectCommon    964-1083    translateCommon 1085-1122
 * with the list widths of allocateMemory (462-527) and the derived sizes of
 * ConnectivityParams (CBM_STATE_LIB_DEV/src/params/connectivityparams.cpp:90-122).
 * Draws come from the oracle's SFMT (orc_sfmt.c).  Stops after connectGRGO:
 * GO-GO wiring and delays are checked elsewhere.
 */
#ifndef ORC_INNET_H
#define ORC_INNET_H
#include <stdint.h>

typedef struct orc_innet_params {
    uint32_t glXP2, glYP2, grXP2, grYP2, goXP2, goYP2, numMFP2;
    uint32_t maxnumpGLfromGLtoGR, lownumpGLfromGLtoGR, maxnumpGLfromGLtoGO, maxnumpGLfromGOtoGL;
    uint32_t maxnumpGRfromGLtoGR, maxnumpGRfromGOtoGR, maxnumpGRfromGRtoGO;
    uint32_t maxnumpGOfromGRtoGO, maxnumpGOfromGLtoGO, maxnumpGOfromGOtoGL;
    uint32_t spanGRDenOnGLX, spanGRDenOnGLY, spanGODecDenOnGLX, spanGODecDenOnGLY;
    uint32_t spanGOAxonOnGLX, spanGOAxonOnGLY, spanGOAscDenOnGRX, spanGOAscDenOnGRY;
} orc_innet_params;

typedef struct orc_innet orc_innet;

/* 0 on success; -1 when a list would overflow the widths the reference
 * allocates (the reference would write past them). */
int orc_innet_generate(const orc_innet_params *p, int seed, orc_innet **out);
void orc_innet_free(orc_innet *g);
/* Rows and columns of a named list ("pGLfromGLtoGR", "pGRfromMFtoGR", ...,
 * the reference's array names; "haspGLfromMFtoGL" and "pGLfromMFtoGL" have one
 * column and no count); 0 when the name is unknown. */
int orc_innet_shape(orc_innet *g, const char *name, uint32_t *rows, uint32_t *cols);
/* Copy a list's counts (rows int32, nullable) and slots (rows*cols uint32, or
 * rows bytes for haspGLfromMFtoGL). */
int orc_innet_copy(orc_innet *g, const char *name, int32_t *num, void *slots);
/* Connections the connector gave up after its attempt limit. */
uint64_t orc_innet_dropped(orc_innet *g);

#endif

/*
 * cbm_oracle.h -- CPU oracle of the cerebellar 1 ms network step.
 *
 * TEST INFRASTRUCTURE.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * or as the timed CPU baseline -- never as the product path.
 *
 * The oracle restates CBMSimCore::calcActivity (CBM_CORE_LIB_MGPU_DEV/src/
 * interface/cbmsimcore.cpp:139-310) as one sequential C loop nest in the
 * reference's order, with the reference's two-phase pipeline lags made
 * explicit.  It keeps the network in the CBMState layout and speaks the same
 * field vocabulary as the product C-ABI (include/cbm_hip.h), so a test can
 * drive both with the same calls.
 *
 * Two builds of the same source:
 *   precision 32  "mimic":  binary32 state and the reference's C/C++ type
 *                           promotions, no FMA contraction; the PF->PC sum in
 *                           the canonical order documented in cbm_oracle.c
 *   precision 64  "replay": the same recurrence with every real in binary64
 *                           (constants are the binary32 values, widened)
 *
 * How each part is pinned (against checks its formulas do not restate):
 *   GR update, MF/GO conductances, plasticity and history, PF/GO sums,
 *   pipeline lags, NC->IO release draws, MF history  tests/test_oracle.py
 *   Golgi glutamate/mGluR chain, NC conductances and membrane, IO
 *   conductances and coupling, MF->NC plasticity      tests/test_oracle_pins.py
 *   PC, BC, SC, GO and IO membranes, thresholds and
 *   spikes, SC->PC saturation, MF->GO and GO->GO
 *   counts, GO depression and coupling, GO->GR input,
 *   IO timer and error drive, microzone routing        tests/test_oracle_pins_small.py
 *   (each pin is also run against a planted mistake, build/liborc_plant.so)
 *   SFMT (orc_sfmt.c)                                  tests/test_sfmt.py (the
 *                                                       reference's own generator)
 *   derived constants                                  tests/test_params.py
 *   input-net generator (orc_innet.c)                  tests/test_innet_gen.py
 * Parity unpinned: the generator's stream equality with the reference's own
 * build of InNetConnectivityState, which cannot be compiled here.
 */
#ifndef CBM_ORACLE_H
#define CBM_ORACLE_H

#include <stddef.h>
#include <stdint.h>
#include "../include/cbm_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_net orc_net;

int orc_create(const cbm_sizes *s, const cbm_act_params *p, const cbm_options *o, int precision,
               orc_net **out);
int orc_destroy(orc_net *n);
/* Same byte layout as cbm_write/cbm_read (binary32 for real fields; the
 * replay build converts on the way in and out). */
int orc_write(orc_net *n, int field, int zone, const void *src, size_t bytes);
int orc_read(orc_net *n, int field, int zone, void *dst, size_t bytes);
/* Real fields in binary64 (the replay's own values; the mimic widens). */
int orc_read_f64(orc_net *n, int field, int zone, double *dst, size_t count);
int orc_commit(orc_net *n);
int orc_set_mf(orc_net *n, const uint8_t *mf);
int orc_set_err_drive(orc_net *n, int zone, float rel);
int orc_set_grpc_steps(orc_net *n, int zone, float ltd, float ltp);
int orc_set_ginc_grgo(orc_net *n, float inc);
/* n steps with the MF input last set */
int orc_step(orc_net *n, int nsteps);
/* n steps reading MF rows mf + k*pitch (host memory) */
int orc_run_mf_train(orc_net *n, const uint8_t *mf, size_t pitch, int nsteps,
                     const float *err_drive);
/* Number of OpenMP threads the GR loops use (1 unless built with OpenMP). */
int orc_threads(void);
/* Planted mistake k of the build with ORC_PLANTS (build/liborc_plant.so; 0 =
 * none): tests show with it that a pin fails on a plausible mistake.  Other
 * builds accept only 0. */
int orc_set_plant(int k);

/* Canonical PF->PC block sum used by the mimic (exposed for tests):
 * x has len elements, len a multiple of 64. */
float orc_canonical_sum_f32(const float *x, size_t len);

#ifdef __cplusplus
}
#endif

#endif

/*
 * cbm_env.h -- closed-loop environments on a simulation handle: the cart-pole
 * of Matthew_Sim_CBM driven by two microzones.
 *
 * It replaces, for hosts that hold a cbm_handle rather than the C++ objects,
 * the loop of the headless driver:
 *   main                 Matthew_Sim_CBM/src/main.cpp:34-163 (--nogui -e cartpole)
 *   SimThread::run       Matthew_Sim_CBM/src/simthread.cpp:96-165 (one step:
 *                        env->getState, mfs->calcActivity, updateMFInput,
 *                        calcActivity, env->step)
 *   SimThread::setupMFs  simthread.cpp:60-87 (PoissonRegenCells(numMF, seed, 4, 1))
 *   Cartpole, Microzone  src/environments/cartpole.cpp:24-285, src/microzone.cpp:5-55
 * The C++ classes themselves (Cartpole, Microzone, Environment,
 * StateVariable, PoissonRegenCells, CRandomSFMT0) are in libcbm_core under
 * CBMEnvInclude/ and CBMToolsInclude/ for C++ drivers.
 *
 * Host-side code: every call runs on the calling thread; the network steps
 * on its own device.  Status codes are those of cbm_hip.h; the message of
 * the last failure is cbm_env_last_error().
 */
#ifndef CBM_ENV_H
#define CBM_ENV_H

#include <stddef.h>
#include <stdint.h>

#include "cbm_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct cbm_cartpole cbm_cartpole;

typedef struct cbm_cartpole_opts {
    int32_t maxNumTrials;    /* trials before the run ends (reference default 20)       */
    int32_t closedLoop;      /* 0: forces stay 0, as the reference's Cartpole::step never
                                updates its microzones; 1: both microzones read their NC
                                spikes every step and push the cart                      */
    int64_t maxTrialLen;     /* steps after which a trial ends (default 1000000)          */
    float trackLen;          /* <= 0: infinite track (default)                           */
    int32_t verbose;         /* print one line per trial, as the reference does          */
    const char *logfile;     /* NULL or "": no log file                                  */
    const char *simStateDir; /* NULL or "": no milStepStart/milStepEnd/failureState files */
} cbm_cartpole_opts;

/* Columns of a trace row, one row per step, after the step. */
enum cbm_cartpole_col {
    CBM_CP_THETA = 0,   /* pole angle (rad)                                   */
    CBM_CP_THETA_DOT,   /* pole angular velocity                              */
    CBM_CP_X,           /* cart position                                      */
    CBM_CP_X_DOT,       /* cart velocity                                      */
    CBM_CP_FORCE_LEFT,  /* force of microzone 0                               */
    CBM_CP_FORCE_RIGHT, /* force of microzone 1                               */
    CBM_CP_ERR_LEFT,    /* errors delivered to microzone 0 in this step       */
    CBM_CP_ERR_RIGHT,   /* errors delivered to microzone 1 in this step       */
    CBM_CP_TIME_ALOFT,  /* steps of the current trial                         */
    CBM_CP_TRIAL,       /* trial number                                       */
    CBM_CP_NC_LEFT,     /* NC spikes of zone 0 this step (traced runs only)    */
    CBM_CP_NC_RIGHT,    /* NC spikes of zone 1 this step (traced runs only)    */
    CBM_CP_MF_SPIKES,   /* MF spikes fed to the network this step             */
    CBM_CP_COLS
};

/* A cart-pole on network h (numZones >= 2; zones 0 and 1 are ForceLeft and
 * ForceRight).  envSeed seeds the environment's generator (main.cpp:87:
 * CRandomSFMT0 randGen(randSeed)), which the constructor draws the first pole tilt
 * from and setupMossyFibers the MF rates and groups from; mfSeed seeds the
 * MF spike generator (PoissonRegenCells(numMF, mfSeed, 4, 1)).  h may be NULL:
 * then only the detached calls below work, with s giving numMF and numNC. */
int cbm_cartpole_create(cbm_handle *h, const cbm_sizes *s, int envSeed, int mfSeed, const cbm_cartpole_opts *opts,
                        cbm_cartpole **out);
int cbm_cartpole_destroy(cbm_cartpole *c);

/* Up to n steps of the driver loop, ending early when the environment has
 * run its trials; *done receives the steps run.  trace (nullable) receives
 * n rows of CBM_CP_COLS floats (rows after *done are left alone). */
int cbm_cartpole_run(cbm_cartpole *c, int n, float *trace, int *done);

/* 1 when the environment has run its trials, else 0. */
int cbm_cartpole_terminated(cbm_cartpole *c);

/* Detached steps, for hosts that step the network themselves:
 * the MF rates (numMF floats, Hz) of Cartpole::getState ... */
int cbm_cartpole_get_state(cbm_cartpole *c, float *mfFreq);
/* ... their spikes from the MF generator (numMF bytes) ... */
int cbm_cartpole_mf_spikes(cbm_cartpole *c, const float *mfFreq, uint8_t *apMF);
/* ... and the environment's step after the network's, with the NC spikes of
 * zones 0 and 1 (numNC bytes each; read only in closed loop, may be NULL
 * otherwise).  row (nullable) receives a trace row; err receives the errors
 * the step delivered to microzones 0 and 1 (the host applies them as an
 * error drive of 1.0 to those zones before the next network step). */
int cbm_cartpole_step(cbm_cartpole *c, const uint8_t *ncLeft, const uint8_t *ncRight, float *row, int32_t err[2]);

/* Which MF group each MF belongs to: 0 always-on, 1 pole velocity, 2 pole
 * angle, 3 cart velocity, 4 cart position, -1 none (numMF int32). */
int cbm_cartpole_mf_groups(cbm_cartpole *c, int32_t *group);
/* The relaxed and excited MF rates (numMF floats each, nullable). */
int cbm_cartpole_mf_rates(cbm_cartpole *c, float *relaxed, float *excited);

const char *cbm_env_last_error(void);

#ifdef __cplusplus
}
#endif
#endif

// cbm_env.cpp -- the C-ABI of include/cbm_env.h over the C++ environment
// classes (CBMEnvInclude/) and PoissonRegenCells.
#include <cbm_env.h>

#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include <CBMCoreInclude/interface/cbmsimcore.h>
#include <CBMEnvInclude/environments/cartpole.hpp>
#include <CBMToolsInclude/poissonregencells.h>

namespace {
thread_local std::string g_err;

int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

template <class F>
int guarded(F &&f)
{
    try {
        return f();
    } catch (const std::exception &e) {
        return fail(CBM_E_ARG, e.what());
    }
}
}  // namespace

struct cbm_cartpole {
    std::unique_ptr<CRandomSFMT0> rng;
    std::unique_ptr<CBMSimCore> core;  // non-owning view of the caller's handle
    std::unique_ptr<Cartpole> env;
    std::unique_ptr<PoissonRegenCells> mfs;
    cbm_sizes sz;
    bool closedLoop;

    void fillRow(float *row, int errL, int errR, float ncL, float ncR, int mfSpikes)
    {
        Cartpole &e = *env;
        row[CBM_CP_THETA] = e.getPoleAngle();
        row[CBM_CP_THETA_DOT] = e.getPoleVelocity();
        row[CBM_CP_X] = e.getCartPosition();
        row[CBM_CP_X_DOT] = e.getCartVelocity();
        row[CBM_CP_FORCE_LEFT] = e.getForceLeft();
        row[CBM_CP_FORCE_RIGHT] = e.getForceRight();
        row[CBM_CP_ERR_LEFT] = (float)errL;
        row[CBM_CP_ERR_RIGHT] = (float)errR;
        row[CBM_CP_TIME_ALOFT] = (float)e.getTimeAloft();
        row[CBM_CP_TRIAL] = (float)e.getTrialNum();
        row[CBM_CP_NC_LEFT] = ncL;
        row[CBM_CP_NC_RIGHT] = ncR;
        row[CBM_CP_MF_SPIKES] = (float)mfSpikes;
    }
};

extern "C" {

const char *cbm_env_last_error(void) { return g_err.c_str(); }

int cbm_cartpole_create(cbm_handle *h, const cbm_sizes *s, int envSeed, int mfSeed, const cbm_cartpole_opts *o,
                        cbm_cartpole **out)
{
    if (!s || !out) return fail(CBM_E_ARG, "cbm_cartpole_create: null sizes or out");
    if (s->numZones < 2) return fail(CBM_E_ARG, "cbm_cartpole_create: the cart-pole needs two microzones");
    if (s->numMF == 0 || s->numNC == 0) return fail(CBM_E_ARG, "cbm_cartpole_create: empty network");
    return guarded([&] {
        std::unique_ptr<cbm_cartpole> c(new cbm_cartpole());
        c->sz = *s;
        CartpoleOptions opts;
        if (o) {
            opts.maxNumTrials = o->maxNumTrials;
            opts.maxTrialLen = (long)o->maxTrialLen;
            opts.trackLen = o->trackLen;
            opts.closedLoop = o->closedLoop != 0;
            opts.verbose = o->verbose != 0;
            if (o->logfile) opts.logfile = o->logfile;
            if (o->simStateDir) opts.simStateDir = o->simStateDir;
        }
        c->closedLoop = opts.closedLoop;
        if (h && !opts.simStateDir.empty())
            return fail(CBM_E_UNSUPPORTED,
                        "cbm_cartpole_create: state files need a CBMState (use the C++ Cartpole with a CBMSimCore)");
        c->rng.reset(new CRandomSFMT0(envSeed));
        if (h) c->core.reset(new CBMSimCore(h, *s));
        c->env.reset(new Cartpole(c->rng.get(), opts));
        c->env->setupMossyFibers((int)s->numMF, (int)s->numNC);
        c->mfs.reset(new PoissonRegenCells(s->numMF, mfSeed, 4.0f, 1.0f));
        *out = c.release();
        return CBM_OK;
    });
}

int cbm_cartpole_destroy(cbm_cartpole *c)
{
    delete c;
    return CBM_OK;
}

int cbm_cartpole_terminated(cbm_cartpole *c)
{
    if (!c) return fail(CBM_E_ARG, "null cart-pole");
    return c->env->terminated() ? 1 : 0;
}

int cbm_cartpole_run(cbm_cartpole *c, int n, float *trace, int *done)
{
    if (!c || n < 0) return fail(CBM_E_ARG, "cbm_cartpole_run: bad arguments");
    if (!c->core) return fail(CBM_E_ARG, "cbm_cartpole_run: created without a simulation handle");
    return guarded([&] {
        // The loop of SimThread::run with one host round trip per step: the
        // MF spikes go in, the step runs and every zone's NC spikes come back
        // in one cbm_step_io pair, and the MF generator draws the next step's
        // uniforms while the network steps.  On a resident network the pairs
        // go to one closed-loop launch (cbm_loop_begin), unless steps write
        // the state to files.
        int k = 0;
        const unsigned nNC = c->sz.numNC, nMF = c->sz.numMF;
        cbm_handle *h = c->core->handle();
        std::vector<uint8_t> nc((size_t)c->sz.numZones * nNC);
        struct Loop {  // ends the closed loop on every way out
            cbm_handle *h = nullptr;
            int end()
            {
                cbm_handle *x = h;
                h = nullptr;
                return x ? cbm_loop_end(x, nullptr) : CBM_OK;
            }
            ~Loop() { (void)end(); }
        } loop;
        if (n > 0 && cbm_resident(h) == 1 && !c->env->savesState()) {
            const int rc = cbm_loop_begin(h, n, CBM_REC_NC);
            if (rc) return fail(rc, std::string("cbm_loop_begin: ") + cbm_last_error());
            loop.h = h;
        }
        for (; k < n && !c->env->terminated(); k++) {
            float *freq = c->env->getState();
            const ct_uint8_t *ap = c->mfs->calcActivity(freq);
            int rc = cbm_step_io_begin(h, ap, CBM_REC_NC);
            if (rc) return fail(rc, std::string("cbm_step_io_begin: ") + cbm_last_error());
            c->mfs->prepare();
            rc = cbm_step_io_end(h, nc.data());
            if (rc) return fail(rc, std::string("cbm_step_io_end: ") + cbm_last_error());
            const unsigned long l0 = c->env->leftMZ().errorsDelivered(), r0 = c->env->rightMZ().errorsDelivered();
            c->env->stepWithSpikes(c->core.get(), nc.data(), nc.data() + nNC);
            if (trace) {
                float ncL = 0, ncR = 0;
                for (unsigned i = 0; i < nNC; i++) {
                    ncL += nc[i];
                    ncR += nc[nNC + i];
                }
                int spikes = 0;
                for (unsigned i = 0; i < nMF; i++) spikes += ap[i];
                c->fillRow(trace + (size_t)k * CBM_CP_COLS, (int)(c->env->leftMZ().errorsDelivered() - l0),
                           (int)(c->env->rightMZ().errorsDelivered() - r0), ncL, ncR, spikes);
            }
        }
        const int rc = loop.end();
        if (rc) return fail(rc, std::string("cbm_loop_end: ") + cbm_last_error());
        if (done) *done = k;
        return CBM_OK;
    });
}

int cbm_cartpole_get_state(cbm_cartpole *c, float *mfFreq)
{
    if (!c || !mfFreq) return fail(CBM_E_ARG, "cbm_cartpole_get_state: null argument");
    return guarded([&] {
        const float *f = c->env->getState();
        std::memcpy(mfFreq, f, sizeof(float) * c->sz.numMF);
        return CBM_OK;
    });
}

int cbm_cartpole_mf_spikes(cbm_cartpole *c, const float *mfFreq, uint8_t *apMF)
{
    if (!c || !mfFreq || !apMF) return fail(CBM_E_ARG, "cbm_cartpole_mf_spikes: null argument");
    const ct_uint8_t *ap = c->mfs->calcActivity(mfFreq);
    std::memcpy(apMF, ap, c->sz.numMF);
    return CBM_OK;
}

int cbm_cartpole_step(cbm_cartpole *c, const uint8_t *ncLeft, const uint8_t *ncRight, float *row, int32_t err[2])
{
    if (!c) return fail(CBM_E_ARG, "cbm_cartpole_step: null cart-pole");
    if (c->closedLoop && (!ncLeft || !ncRight)) return fail(CBM_E_ARG, "cbm_cartpole_step: closed loop needs NC spikes");
    return guarded([&] {
        Cartpole &e = *c->env;
        const unsigned long l0 = e.leftMZ().errorsDelivered(), r0 = e.rightMZ().errorsDelivered();
        e.stepDetached(ncLeft, ncRight);
        const int el = (int)(e.leftMZ().errorsDelivered() - l0), er = (int)(e.rightMZ().errorsDelivered() - r0);
        if (err) {
            err[0] = el;
            err[1] = er;
        }
        if (row) {
            float ncL = 0, ncR = 0;
            if (ncLeft && ncRight)
                for (unsigned i = 0; i < c->sz.numNC; i++) {
                    ncL += ncLeft[i];
                    ncR += ncRight[i];
                }
            c->fillRow(row, el, er, ncL, ncR, 0);
        }
        return CBM_OK;
    });
}

int cbm_cartpole_mf_groups(cbm_cartpole *c, int32_t *group)
{
    if (!c || !group) return fail(CBM_E_ARG, "cbm_cartpole_mf_groups: null argument");
    for (unsigned i = 0; i < c->sz.numMF; i++) group[i] = -1;
    std::vector<StateVariableBase *> svs = c->env->getStateVariables();
    for (size_t g = 0; g < svs.size(); g++)
        for (int m : svs[g]->mfInds) group[m] = (int32_t)g;
    return CBM_OK;
}

int cbm_cartpole_mf_rates(cbm_cartpole *c, float *relaxed, float *excited)
{
    if (!c) return fail(CBM_E_ARG, "cbm_cartpole_mf_rates: null cart-pole");
    const size_t n = c->sz.numMF;
    if (relaxed) std::memcpy(relaxed, c->env->getRelaxedFreqs().data(), n * sizeof(float));
    if (excited) std::memcpy(excited, c->env->getExcitedFreqs().data(), n * sizeof(float));
    return CBM_OK;
}

}  // extern "C"

/*
 * orc_api.c -- public entry points of the CPU oracle (cbm_oracle.h).
 * TEST INFRASTRUCTURE ONLY: the checker and the timed CPU baseline.
 */
#include <stdlib.h>
#include <string.h>
#include "cbm_oracle.h"

int orc_f32_create(const cbm_sizes *, const cbm_act_params *, const cbm_options *, void **);
void orc_f32_destroy(void *);
int orc_f32_io(void *, int, int, void *, size_t, int);
int orc_f32_read_f64(void *, int, int, double *, size_t);
int orc_f32_commit(void *);
void orc_f32_step(void *);
int orc_f64_create(const cbm_sizes *, const cbm_act_params *, const cbm_options *, void **);
void orc_f64_destroy(void *);
int orc_f64_io(void *, int, int, void *, size_t, int);
int orc_f64_read_f64(void *, int, int, double *, size_t);
int orc_f64_commit(void *);
void orc_f64_step(void *);

struct orc_net {
    int prec;
    void *impl;
    cbm_sizes s;
    float maxExtIncVIO;
};

#define DISPATCH(n, fn, ...) ((n)->prec == 32 ? orc_f32_##fn(__VA_ARGS__) : orc_f64_##fn(__VA_ARGS__))

int orc_create(const cbm_sizes *s, const cbm_act_params *p, const cbm_options *o, int precision,
               orc_net **out)
{
    if (!s || !p || !out || (precision != 32 && precision != 64)) return CBM_E_ARG;
    orc_net *n = (orc_net *)calloc(1, sizeof(orc_net));
    if (!n) return CBM_E_NOMEM;
    n->prec = precision;
    n->s = *s;
    n->maxExtIncVIO = p->maxExtIncVIO;
    int rc = precision == 32 ? orc_f32_create(s, p, o, &n->impl) : orc_f64_create(s, p, o, &n->impl);
    if (rc) { free(n); return rc; }
    *out = n;
    return CBM_OK;
}

int orc_destroy(orc_net *n)
{
    if (!n) return CBM_E_ARG;
    if (n->prec == 32) orc_f32_destroy(n->impl); else orc_f64_destroy(n->impl);
    free(n);
    return CBM_OK;
}

int orc_write(orc_net *n, int f, int z, const void *src, size_t bytes)
{
    if (!n || !src) return CBM_E_ARG;
    return DISPATCH(n, io, n->impl, f, z, (void *)src, bytes, 1);
}

int orc_read(orc_net *n, int f, int z, void *dst, size_t bytes)
{
    if (!n || !dst) return CBM_E_ARG;
    return DISPATCH(n, io, n->impl, f, z, dst, bytes, 0);
}

int orc_read_f64(orc_net *n, int f, int z, double *dst, size_t count)
{
    if (!n || !dst) return CBM_E_ARG;
    return DISPATCH(n, read_f64, n->impl, f, z, dst, count);
}

int orc_commit(orc_net *n)
{
    if (!n) return CBM_E_ARG;
    return DISPATCH(n, commit, n->impl);
}

int orc_set_mf(orc_net *n, const uint8_t *mf)
{
    return orc_write(n, CBM_F_MF_INPUT, 0, mf, n->s.numMF);
}

/* MZone::setErrDrive (mzone.cpp:167-173): errDrive = rel * maxExtIncVIO */
int orc_set_err_drive(orc_net *n, int zone, float rel)
{
    float v = rel * n->maxExtIncVIO;
    return orc_write(n, CBM_F_ERR_DRIVE, zone, &v, 4);
}

int orc_set_grpc_steps(orc_net *n, int zone, float ltd, float ltp)
{
    float v[2] = {ltd, ltp};
    return orc_write(n, CBM_F_GRPC_PLAST_STEPS, zone, v, 8);
}

int orc_set_ginc_grgo(orc_net *n, float inc)
{
    return orc_write(n, CBM_F_GINC_GR_TO_GO, 0, &inc, 4);
}

int orc_step(orc_net *n, int nsteps)
{
    if (!n || nsteps < 0) return CBM_E_ARG;
    for (int k = 0; k < nsteps; k++) {
        if (n->prec == 32) orc_f32_step(n->impl); else orc_f64_step(n->impl);
    }
    return CBM_OK;
}

int orc_run_mf_train(orc_net *n, const uint8_t *mf, size_t pitch, int nsteps, const float *err)
{
    if (!n || !mf || nsteps < 0) return CBM_E_ARG;
    for (int k = 0; k < nsteps; k++) {
        int rc = orc_set_mf(n, mf + (size_t)k * pitch);
        if (rc) return rc;
        if (err)
            for (uint32_t z = 0; z < n->s.numZones; z++) {
                float e = err[(size_t)k * n->s.numZones + z];
                if (e != 0.0f) orc_set_err_drive(n, (int)z, e);
            }
        if (n->prec == 32) orc_f32_step(n->impl); else orc_f64_step(n->impl);
    }
    return CBM_OK;
}

#ifdef ORC_PLANTS
int orc_plant = 0;
#endif

int orc_set_plant(int k)
{
#ifdef ORC_PLANTS
    orc_plant = k;
    return CBM_OK;
#else
    return k ? CBM_E_UNSUPPORTED : CBM_OK;
#endif
}

int orc_threads(void)
{
#ifdef _OPENMP
    extern int omp_get_max_threads(void);
    return omp_get_max_threads();
#else
    return 1;
#endif
}

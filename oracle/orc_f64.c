/* orc_f64.c -- binary64 "replay" build of the oracle (TEST INFRASTRUCTURE ONLY). */
#define REAL double
#define EXPR exp
#define SFX _f64
#define ORC_PREC 64
#include "orc_impl.inc"

int orc_f64_create(const cbm_sizes *s, const cbm_act_params *p, const cbm_options *o, void **out)
{
    return create_f64(s, p, o, (orc_net_f64 **)out);
}
void orc_f64_destroy(void *n) { free_all_f64((orc_net_f64 *)n); }
int orc_f64_io(void *n, int f, int z, void *b, size_t bytes, int w) { return io_f64((orc_net_f64 *)n, f, z, b, bytes, w); }
int orc_f64_read_f64(void *n, int f, int z, double *d, size_t c) { return read_f64_f64((orc_net_f64 *)n, f, z, d, c); }
int orc_f64_commit(void *n) { return commit_f64((orc_net_f64 *)n); }
void orc_f64_step(void *n) { step1_f64((orc_net_f64 *)n); }

/* orc_f32.c -- binary32 "mimic" build of the oracle (TEST INFRASTRUCTURE ONLY). */
#define REAL float
#define EXPR expf
#define SFX _f32
#define ORC_PREC 32
#include "orc_impl.inc"

int orc_f32_create(const cbm_sizes *s, const cbm_act_params *p, const cbm_options *o, void **out)
{
    return create_f32(s, p, o, (orc_net_f32 **)out);
}
void orc_f32_destroy(void *n) { free_all_f32((orc_net_f32 *)n); }
int orc_f32_io(void *n, int f, int z, void *b, size_t bytes, int w) { return io_f32((orc_net_f32 *)n, f, z, b, bytes, w); }
int orc_f32_read_f64(void *n, int f, int z, double *d, size_t c) { return read_f64_f32((orc_net_f32 *)n, f, z, d, c); }
int orc_f32_commit(void *n) { return commit_f32((orc_net_f32 *)n); }
void orc_f32_step(void *n) { step1_f32((orc_net_f32 *)n); }

float orc_canonical_sum_f32(const float *x, size_t len)
{
    float lanes[64], chunk[64];
    for (int l = 0; l < 64; l++) lanes[l] = 0;
    for (size_t c = 0; c < len / 64; c++) {
        for (int l = 0; l < 64; l++) chunk[l] = x[c * 64 + l];
        lanes[c % 64] = lanes[c % 64] + halving64_f32(chunk);
    }
    return halving64_f32(lanes);
}

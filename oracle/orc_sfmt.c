/*
 * orc_sfmt.c -- oracle SFMT-11213 (see orc_sfmt.h for the contract and the
 * reference lines it follows).  TEST INFRASTRUCTURE ONLY.
 *
 * The 128-bit words of the published algorithm are held as four 32-bit
 * little-endian lanes; the two byte shifts of the recursion (by SL2 and SR2
 * bytes) are written out lane by lane.
 */
#include "orc_sfmt.h"

#define POS1 68
#define SL1 14
#define SL2 3  /* bytes */
#define SR1 7
#define SR2 3  /* bytes */

static const uint32_t MSK[4] = {0xeffff7fbu, 0xffffffefu, 0xdfdfbfffu, 0x7fffdbfdu};
static const uint32_t PARITY[4] = {0x00000001u, 0x00000000u, 0xe8148000u, 0xd0c7afa3u};

static void rec(uint32_t *r, const uint32_t *a, const uint32_t *b, const uint32_t *c,
                const uint32_t *d)
{
    uint32_t al[4], cr[4];
    /* a shifted left by SL2 bytes as one 128-bit number */
    al[0] = a[0] << (8 * SL2);
    al[1] = (a[1] << (8 * SL2)) | (a[0] >> (32 - 8 * SL2));
    al[2] = (a[2] << (8 * SL2)) | (a[1] >> (32 - 8 * SL2));
    al[3] = (a[3] << (8 * SL2)) | (a[2] >> (32 - 8 * SL2));
    /* c shifted right by SR2 bytes as one 128-bit number */
    cr[3] = c[3] >> (8 * SR2);
    cr[2] = (c[2] >> (8 * SR2)) | (c[3] << (32 - 8 * SR2));
    cr[1] = (c[1] >> (8 * SR2)) | (c[2] << (32 - 8 * SR2));
    cr[0] = (c[0] >> (8 * SR2)) | (c[1] << (32 - 8 * SR2));
    for (int k = 0; k < 4; k++)
        r[k] = a[k] ^ al[k] ^ ((b[k] >> SR1) & MSK[k]) ^ cr[k] ^ (d[k] << SL1);
}

void orc_sfmt_generate(orc_sfmt *g)
{
    uint32_t r1[4], r2[4], r[4];
    uint32_t *s = g->s;
    for (int k = 0; k < 4; k++) {
        r1[k] = s[4 * (ORC_SFMT_N - 2) + k];
        r2[k] = s[4 * (ORC_SFMT_N - 1) + k];
    }
    for (int i = 0; i < ORC_SFMT_N; i++) {
        int j = i + POS1;
        if (j >= ORC_SFMT_N) j -= ORC_SFMT_N;
        rec(r, &s[4 * i], &s[4 * j], r1, r2);
        for (int k = 0; k < 4; k++) {
            s[4 * i + k] = r[k];
            r1[k] = r2[k];
            r2[k] = r[k];
        }
    }
    g->ix = 0;
}

void orc_sfmt_init(orc_sfmt *g, int seed)
{
    uint32_t y = (uint32_t)seed;
    g->s[0] = y;
    for (uint32_t i = 1; i < ORC_SFMT_N32; i++) {
        y = 1812433253u * (y ^ (y >> 30)) + i;
        g->s[i] = y;
    }
    /* period certification: parity of the first 128 bits against PARITY */
    uint32_t t = 0;
    for (int i = 0; i < 4; i++) t ^= PARITY[i] & g->s[i];
    for (int i = 16; i > 0; i >>= 1) t ^= t >> i;
    if (!(t & 1)) {
        int done = 0;
        for (int i = 0; i < 4 && !done; i++)
            for (uint32_t bit = 1; bit && !done; bit <<= 1)
                if (PARITY[i] & bit) {
                    g->s[i] ^= bit;
                    done = 1;
                }
    }
    g->last_interval = 0;
    g->rlimit = 0;
    orc_sfmt_generate(g);
}

uint32_t orc_sfmt_brandom(orc_sfmt *g)
{
    if (g->ix >= ORC_SFMT_N32) orc_sfmt_generate(g);
    return g->s[g->ix++];
}

double orc_sfmt_random(orc_sfmt *g)
{
    if (g->ix >= ORC_SFMT_N32 - 1) orc_sfmt_generate(g);
    uint64_t r = (uint64_t)g->s[g->ix] | ((uint64_t)g->s[g->ix + 1] << 32);
    g->ix += 2;
    return (double)(int64_t)(r >> 12) * (1.0 / 4503599627370496.0); /* 2^-52 */
}

int orc_sfmt_irandom(orc_sfmt *g, int min, int max)
{
    if (max <= min) return max == min ? min : (int)0x80000000u;
    uint32_t interval = (uint32_t)max - (uint32_t)min + 1u; /* wraps modulo 2^32 */
    uint64_t lr = (uint64_t)orc_sfmt_brandom(g) * interval;
    return (int32_t)(uint32_t)(lr >> 32) + min;
}

int orc_sfmt_irandomx(orc_sfmt *g, int min, int max)
{
    if (max <= min) return max == min ? min : (int)0x80000000u;
    uint32_t interval = (uint32_t)max - (uint32_t)min + 1u; /* wraps modulo 2^32 */
    if (interval != g->last_interval) {
        g->rlimit = (uint32_t)(((uint64_t)1 << 32) / interval) * interval - 1;
        g->last_interval = interval;
    }
    uint64_t lr;
    do {
        lr = (uint64_t)orc_sfmt_brandom(g) * interval;
    } while ((uint32_t)lr > g->rlimit);
    return (int32_t)(uint32_t)(lr >> 32) + min;
}

void orc_sfmt_random_fill(orc_sfmt *g, double *out, int n)
{
    for (int i = 0; i < n; i++) out[i] = orc_sfmt_random(g);
}

int orc_sfmt_size(void) { return (int)sizeof(orc_sfmt); }

/*
 * orc_sfmt.h -- oracle restatement of the reference's random generator.
 *
 * TEST INFRASTRUCTURE: this file belongs to the CPU oracle.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may use it.
 *
 * The reference draws from CRandomSFMT0 = SIMD-oriented Fast Mersenne Twister
 * (Saito & Matsumoto) with Mersenne exponent 11213 and no Mother-of-All mixing
 * (CBM_Cart_On_Cart/includes/sfmt.h:136-159,226-252; the library copy lives in
 * the absent CXX_TOOLS_LIB).  The contract restated here:
 *   seeding       sfmt.cpp:78-96   s[0]=seed, s[i]=1812433253*(s[i-1]^(s[i-1]>>30))+i
 *   certification sfmt.cpp:177-206 parity of s[0..3] & PARITY must be odd
 *   generation    sfmt.cpp:210-246 the SFMT recursion over 88 128-bit words
 *   BRandom       sfmt.cpp:249-258 next 32-bit word
 *   Random        sfmt.cpp:337-355 next two words as a 64-bit value, >>12, *2^-52
 *   IRandom       sfmt.cpp:280-298 min + (BRandom()*interval)>>32
 *   IRandomX      sfmt.cpp:300-335 same with rejection of the top remainders
 */
#ifndef ORC_SFMT_H
#define ORC_SFMT_H

#include <stdint.h>

#define ORC_SFMT_N 88
#define ORC_SFMT_N32 (ORC_SFMT_N * 4)

typedef struct orc_sfmt {
    uint32_t s[ORC_SFMT_N32];
    uint32_t ix;
    uint32_t last_interval;
    uint32_t rlimit;
} orc_sfmt;

void orc_sfmt_init(orc_sfmt *g, int seed);
void orc_sfmt_generate(orc_sfmt *g);
uint32_t orc_sfmt_brandom(orc_sfmt *g);
double orc_sfmt_random(orc_sfmt *g);
int orc_sfmt_irandom(orc_sfmt *g, int min, int max);
int orc_sfmt_irandomx(orc_sfmt *g, int min, int max);
/* n consecutive Random() outputs */
void orc_sfmt_random_fill(orc_sfmt *g, double *out, int n);
/* sizeof(orc_sfmt), for hosts that allocate it */
int orc_sfmt_size(void);

#endif

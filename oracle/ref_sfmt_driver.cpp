// ref_sfmt_driver.cpp -- C entry points around the reference's own SFMT class.
//
// TEST INFRASTRUCTURE.  Built by oracle/Makefile together with the vendored
// reference source CBM_Cart_On_Cart/src/sfmt.cpp (compiled where it lies under
// the reference tree, never copied) into oracle/_ref/libref_sfmt.so.  Used only
// to write tests/golden/sfmt_vectors.json and to check the oracle's SFMT.
#include <stdint.h>
#include "sfmt.h"  // CBM_Cart_On_Cart/includes/sfmt.h, found through -I

extern "C" {

// n successive BRandom() outputs of CRandomSFMT0(seed).
void ref_sfmt_brandom(int seed, int n, uint32_t *out)
{
    CRandomSFMT0 g(seed);
    for (int i = 0; i < n; i++) out[i] = g.BRandom();
}

// n successive Random() outputs, after `skip` BRandom() calls.
void ref_sfmt_random(int seed, int skip, int n, double *out)
{
    CRandomSFMT0 g(seed);
    for (int i = 0; i < skip; i++) g.BRandom();
    for (int i = 0; i < n; i++) out[i] = g.Random();
}

// n successive Random() outputs after `skip` Random() calls (the draws
// MZone::testReduction makes before the first NC->IO draw).
void ref_sfmt_random_after(int seed, long skip, int n, double *out)
{
    CRandomSFMT0 g(seed);
    for (long i = 0; i < skip; i++) g.Random();
    for (int i = 0; i < n; i++) out[i] = g.Random();
}

// n successive IRandom(lo, hi) outputs.
void ref_sfmt_irandom(int seed, int lo, int hi, int n, int *out)
{
    CRandomSFMT0 g(seed);
    for (int i = 0; i < n; i++) out[i] = g.IRandom(lo, hi);
}

// n successive IRandomX(lo, hi) outputs.
void ref_sfmt_irandomx(int seed, int lo, int hi, int n, int *out)
{
    CRandomSFMT0 g(seed);
    for (int i = 0; i < n; i++) out[i] = g.IRandomX(lo, hi);
}

}  // extern "C"

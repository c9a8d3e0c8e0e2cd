// sfmt.h -- CRandomSFMT0, the generator class that drivers, environments and
// tools draw from (interface: CBM_Cart_On_Cart/includes/sfmt.h:213-243, the
// class without the Mother-of-All combination), on the library's own
// SFMT-11213 (csrc/state/sfmt.hpp).
#ifndef CBM_CXXTOOLS_SFMT_H
#define CBM_CXXTOOLS_SFMT_H

#include <stdint.h>

#include "../../../../state/sfmt.hpp"

class CRandomSFMT0
{
public:
    explicit CRandomSFMT0(int seed) : g(seed) {}
    void RandomInit(int seed) { g.seed_(seed); }
    // integer in [min, max]; (int)0x80000000 when max < min
    int IRandom(int min, int max) { return g.IRandom(min, max); }
    // the same without the multiply-and-shift bias (rejection)
    int IRandomX(int min, int max) { return g.IRandomX(min, max); }
    // double in [0, 1) with 52 random bits
    double Random() { return g.Random(); }
    uint32_t BRandom() { return g.BRandom(); }

private:
    cbm::Sfmt g;
};

#endif

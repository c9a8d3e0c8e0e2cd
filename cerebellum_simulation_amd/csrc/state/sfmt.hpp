// sfmt.hpp -- SFMT-11213 random generator with the interface the reference's
// state and core code draw from (CRandomSFMT0: RandomInit, BRandom, Random,
// IRandom, IRandomX; CBM_Cart_On_Cart/includes/sfmt.h:226-252).
//
// Written from the published SFMT algorithm (Saito & Matsumoto, "SIMD-oriented
// Fast Mersenne Twister", 2008) with the parameter set for exponent 11213 and
// the seeding / period-certification / output conventions of the reference's
// generator:
//   * seed:   w[0]=seed, w[i]=1812433253*(w[i-1]^(w[i-1]>>30))+i over 352 words
//   * certify: odd parity of (w[0..3] & PARITY), else flip the lowest PARITY bit
//   * the state is regenerated once at seeding, then whenever it is used up
//   * Random(): if fewer than two words remain, regenerate; take two words as a
//     little-endian 64-bit value v and return (v>>12) * 2^-52
//   * IRandom(a,b) = a + (BRandom()*(b-a+1) >> 32); IRandomX rejects draws whose
//     low 32 bits exceed floor(2^32/n)*n-1.
#pragma once
#include <cstdint>

namespace cbm {

class Sfmt {
public:
    static constexpr int N = 88;       // 128-bit words
    static constexpr int N32 = N * 4;  // 32-bit words

    explicit Sfmt(int seed = 0) { seed_(seed); }
    void seed_(int seed)
    {
        uint32_t y = (uint32_t)seed;
        w_[0] = y;
        for (uint32_t i = 1; i < (uint32_t)N32; i++) {
            y = 1812433253u * (y ^ (y >> 30)) + i;
            w_[i] = y;
        }
        static const uint32_t parity[4] = {0x00000001u, 0x00000000u, 0xe8148000u, 0xd0c7afa3u};
        uint32_t p = 0;
        for (int i = 0; i < 4; i++) p ^= parity[i] & w_[i];
        for (int sh = 16; sh > 0; sh >>= 1) p ^= p >> sh;
        if ((p & 1u) == 0) {
            bool fixed = false;
            for (int i = 0; i < 4 && !fixed; i++)
                for (uint32_t bit = 1; bit != 0 && !fixed; bit <<= 1)
                    if (parity[i] & bit) {
                        w_[i] ^= bit;
                        fixed = true;
                    }
        }
        lastInterval_ = 0;
        rlimit_ = 0;
        refill();
    }

    uint32_t BRandom()
    {
        if (ix_ >= (uint32_t)N32) refill();
        return w_[ix_++];
    }

    double Random()
    {
        if (ix_ >= (uint32_t)N32 - 1) refill();
        uint64_t v = (uint64_t)w_[ix_] | ((uint64_t)w_[ix_ + 1] << 32);
        ix_ += 2;
        return (double)(int64_t)(v >> 12) * (1.0 / 4503599627370496.0);
    }

    // The interval length hi - lo + 1 wraps modulo 2^32 (IRandom(0, INT_MAX)
    // draws from 2^31 values); it is formed in unsigned arithmetic.
    int IRandom(int lo, int hi)
    {
        if (hi <= lo) return hi == lo ? lo : (int)0x80000000u;
        uint32_t n = (uint32_t)hi - (uint32_t)lo + 1u;
        return (int32_t)(uint32_t)(((uint64_t)BRandom() * n) >> 32) + lo;
    }

    int IRandomX(int lo, int hi)
    {
        if (hi <= lo) return hi == lo ? lo : (int)0x80000000u;
        uint32_t n = (uint32_t)hi - (uint32_t)lo + 1u;
        if (n != lastInterval_) {
            rlimit_ = (uint32_t)((((uint64_t)1) << 32) / n) * n - 1;
            lastInterval_ = n;
        }
        uint64_t v;
        do {
            v = (uint64_t)BRandom() * n;
        } while ((uint32_t)v > rlimit_);
        return (int32_t)(uint32_t)(v >> 32) + lo;
    }

    // 352 state words followed by the read index (the layout of the
    // CBM_F_ZONE_RNG_STATE field).
    void save(uint32_t *out) const
    {
        for (int i = 0; i < N32; i++) out[i] = w_[i];
        out[N32] = ix_;
    }
    void load(const uint32_t *in)
    {
        for (int i = 0; i < N32; i++) w_[i] = in[i];
        ix_ = in[N32];
    }

private:
    // one pass of the recursion over the whole state (128-bit words held as
    // four little-endian 32-bit lanes)
    void refill()
    {
        static const uint32_t msk[4] = {0xeffff7fbu, 0xffffffefu, 0xdfdfbfffu, 0x7fffdbfdu};
        uint32_t r1[4], r2[4];
        for (int k = 0; k < 4; k++) {
            r1[k] = w_[4 * (N - 2) + k];
            r2[k] = w_[4 * (N - 1) + k];
        }
        for (int i = 0; i < N; i++) {
            const uint32_t *a = &w_[4 * i];
            const uint32_t *b = &w_[4 * ((i + 68) % N)];
            uint32_t sa[4] = {a[0] << 24, (a[1] << 24) | (a[0] >> 8), (a[2] << 24) | (a[1] >> 8),
                              (a[3] << 24) | (a[2] >> 8)};
            uint32_t sc[4] = {(r1[0] >> 24) | (r1[1] << 8), (r1[1] >> 24) | (r1[2] << 8),
                              (r1[2] >> 24) | (r1[3] << 8), r1[3] >> 24};
            uint32_t r[4];
            for (int k = 0; k < 4; k++) r[k] = a[k] ^ sa[k] ^ ((b[k] >> 7) & msk[k]) ^ sc[k] ^ (r2[k] << 14);
            for (int k = 0; k < 4; k++) {
                w_[4 * i + k] = r[k];
                r1[k] = r2[k];
                r2[k] = r[k];
            }
        }
        ix_ = 0;
    }

    uint32_t w_[N32];
    uint32_t ix_ = 0;
    uint32_t lastInterval_ = 0, rlimit_ = 0;
};

}  // namespace cbm

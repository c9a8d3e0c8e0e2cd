"""Write tests/golden/sfmt_vectors.json from the reference's own SFMT class.

Run in the build container (where the reference tree exists):
    make -C oracle ref && python tests/golden/make_sfmt_vectors.py
The library oracle/_ref/libref_sfmt.so is CBM_Cart_On_Cart/src/sfmt.cpp compiled
where it lies, with the C entry points of oracle/ref_sfmt_driver.cpp.  The
vectors are data: generator outputs for fixed seeds.
"""
import ctypes as C
import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
SEEDS = [0, 1, 12345, -7, 2147483647]


def main():
    L = C.CDLL(os.path.join(ROOT, "oracle", "_ref", "libref_sfmt.so"))
    out = {"generator": "CRandomSFMT0 (CBM_Cart_On_Cart/src/sfmt.cpp, MEXP 11213)", "seeds": {}}
    for s in SEEDS:
        b = (C.c_uint32 * 1000)()
        L.ref_sfmt_brandom(s, 1000, b)
        r0 = (C.c_double * 700)()
        L.ref_sfmt_random(s, 0, 700, r0)
        r3 = (C.c_double * 400)()
        L.ref_sfmt_random(s, 3, 400, r3)
        ra = (C.c_double * 300)()
        L.ref_sfmt_random_after.argtypes = [C.c_int, C.c_long, C.c_int, C.c_void_p]
        L.ref_sfmt_random_after(s, 16384, 300, ra)
        ir = (C.c_int * 500)()
        L.ref_sfmt_irandom(s, 0, 99, 500, ir)
        ix = (C.c_int * 500)()
        L.ref_sfmt_irandomx(s, 0, 30, 500, ix)
        out["seeds"][str(s)] = {
            "brandom": list(b), "random": [x.hex() for x in r0], "random_skip3_words": [x.hex() for x in r3],
            "random_after_16384_random": [x.hex() for x in ra], "irandom_0_99": list(ir),
            "irandomx_0_30": list(ix),
        }
    with open(os.path.join(HERE, "sfmt_vectors.json"), "w") as f:
        json.dump(out, f, indent=0)
    print("wrote", os.path.join(HERE, "sfmt_vectors.json"))


if __name__ == "__main__":
    main()

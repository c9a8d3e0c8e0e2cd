"""Restatement of ECMFPopulation (CBM_TOOLS_LIB_DEV/src/ecmfpopulation.cpp),
the eyelid driver's MF rates, on the oracle's SFMT generator.

TEST INFRASTRUCTURE: only tests/ use this module, as the checker of the
library's C++ ECMFPopulation (csrc/core/ecmfpopulation.cpp)."""
from __future__ import annotations

import numpy as np

from cartpole_ref import Gen

f32 = np.float32


def seeded(num_mf, seed, frac_cst, frac_csp, frac_ctx, bg_min, csbg_min, ctx_min, cst_min, csp_min,
           bg_max, csbg_max, ctx_max, cst_max, csp_max):
    """(bg, tonic, phasic) of ECMFPopulation's seeded constructor
    (ecmfpopulation.cpp:67-180): Random() per MF for the background rate
    (float arguments, binary64 product and sum, stored in binary32), then CS-
    tonic, CS-phasic and context MFs picked with IRandom(0, numMF-1) and
    redrawn when taken, then the picked MFs' rates in MF order."""
    g = Gen(seed)
    a = [f32(x) for x in (bg_min, csbg_min, ctx_min, cst_min, csp_min, bg_max, csbg_max, ctx_max, cst_max, csp_max)]
    bg_min, csbg_min, ctx_min, cst_min, csp_min, bg_max, csbg_max, ctx_max, cst_max, csp_max = a

    def rate(lo, hi):  # Random()*(max-min)+min with float max, min
        return f32(g.random() * float(f32(hi - lo)) + float(lo))

    bg = np.zeros(num_mf, np.float32)
    for i in range(num_mf):
        bg[i] = rate(bg_min, bg_max)
    tonic, phasic = bg.copy(), bg.copy()
    cst = np.zeros(num_mf, bool)
    csp = np.zeros(num_mf, bool)
    ctx = np.zeros(num_mf, bool)
    for _ in range(int(f32(frac_cst) * f32(num_mf))):
        while True:
            m = g.irandom(0, num_mf - 1)
            if not cst[m]:
                cst[m] = True
                break
    for _ in range(int(f32(frac_csp) * f32(num_mf))):
        while True:
            m = g.irandom(0, num_mf - 1)
            if not (csp[m] or cst[m]):
                csp[m] = True
                break
    for _ in range(int(f32(frac_ctx) * f32(num_mf))):
        while True:
            m = g.irandom(0, num_mf - 1)
            if not (ctx[m] or csp[m] or cst[m]):
                ctx[m] = True
                break
    for i in range(num_mf):
        if ctx[i]:
            bg[i] = rate(ctx_min, ctx_max)
            tonic[i] = phasic[i] = bg[i]
        if cst[i]:
            bg[i] = rate(csbg_min, csbg_max)
            tonic[i] = rate(cst_min, cst_max)
            phasic[i] = tonic[i]
        if csp[i]:
            bg[i] = rate(csbg_min, csbg_max)
            tonic[i] = bg[i]
            phasic[i] = rate(csp_min, csp_max)
    return bg, tonic, phasic, (cst, csp, ctx)


def from_psh(cells: np.ndarray, ms_per_bin: int, num_trials: int, ms_pre: int, ms_trial: int):
    """(bg, tonic, phasic) of the constructor from recorded trials
    (ecmfpopulation.cpp:11-62): cells [numMF][bins] spike counts summed over
    trials; binary32 sums in bin order."""
    scaling = f32(num_trials * (ms_per_bin / 1000.0))
    pre, trial = ms_pre // ms_per_bin, ms_trial // ms_per_bin
    n = cells.shape[0]
    out = [np.zeros(n, np.float32) for _ in range(3)]

    def s(row, lo, hi):
        acc = f32(0)
        for j in range(lo, hi):
            acc = f32(acc + f32(row[j]))
        return acc

    for i in range(n):
        row = cells[i]
        out[0][i] = f32(s(row, 2, pre - 2) / f32(f32(pre - 4) * scaling))
        out[2][i] = f32(s(row, pre + 1, pre + 3) / f32(f32(2) * scaling))
        out[1][i] = f32(s(row, pre + trial - 6, pre + trial - 1) / f32(f32(5) * scaling))
    return out[0], out[1], out[2]


def file_bytes(bg, tonic, phasic) -> bytes:
    """ECMFPopulation::writeToFile (ecmfpopulation.cpp:157-177)."""
    return (np.uint32(len(bg)).tobytes() + np.asarray(bg, np.float32).tobytes() +
            np.asarray(tonic, np.float32).tobytes() + np.asarray(phasic, np.float32).tobytes())

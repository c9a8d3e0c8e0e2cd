"""ECMFPopulation, the eyelid driver's MF rates
(CBM_TOOLS_LIB_DEV/src/ecmfpopulation.cpp, used at
Eyelid_Sim_CBM_Dev/src/ectrial/ecmanagementdelay.cpp:20,42,61,88-92): the
library's C++ class (tests/cpp/ecmfpop_driver) writes the same file as the
oracle restatement (oracle/ecmfpop_ref.py) from the same seed; the oracle is
pinned by the category counts and rate ranges the constructor's arguments fix,
by the reference generator's stream (tests/test_sfmt.py), and, for the
constructor from recorded trials, by MFs firing with fixed periods, whose
rates come out as 1000 / period Hz."""
import os
import subprocess

import numpy as np
import pytest

import ecmfpop_ref as R
from cerebellum_simulation_amd.ectrials import ECTrialsData

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "cpp", "build", "ecmfpop_driver")
# fractions and Hz ranges of the delay protocol's MF populations
ARGS = dict(num_mf=1024, seed=7, frac_cst=0.03, frac_csp=0.02, frac_ctx=0.1, bg_min=1, csbg_min=1, ctx_min=30,
            cst_min=90, csp_min=200, bg_max=10, csbg_max=5, ctx_max=60, cst_max=110, csp_max=250)


def _cpp_seeded(tmp_path, a):
    out = tmp_path / "pop.bin"
    order = ("num_mf", "seed", "frac_cst", "frac_csp", "frac_ctx", "bg_min", "csbg_min", "ctx_min", "cst_min",
             "csp_min", "bg_max", "csbg_max", "ctx_max", "cst_max", "csp_max")
    r = subprocess.run([DRIVER, "seeded", str(out)] + [repr(a[k]) for k in order], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return out.read_bytes()


@pytest.mark.parametrize("seed", [7, 12345])
def test_seeded_file_equals_oracle(tmp_path, seed):
    a = {**ARGS, "seed": seed}
    bg, ton, pha, _ = R.seeded(**a)
    assert _cpp_seeded(tmp_path, a) == R.file_bytes(bg, ton, pha)


def test_oracle_categories_and_ranges():
    bg, ton, pha, (cst, csp, ctx) = R.seeded(**ARGS)
    n = ARGS["num_mf"]
    assert cst.sum() == int(np.float32(0.03) * n) and csp.sum() == int(np.float32(0.02) * n)
    assert ctx.sum() == int(np.float32(0.1) * n)
    assert not (cst & csp).any() and not (ctx & (cst | csp)).any()
    plain = ~(cst | csp | ctx)
    assert ((bg[plain] >= 1) & (bg[plain] < 10)).all()
    assert ((bg[ctx] >= 30) & (bg[ctx] < 60)).all()
    assert ((bg[cst | csp] >= 1) & (bg[cst | csp] < 5)).all()
    assert ((ton[cst] >= 90) & (ton[cst] < 110)).all() and np.array_equal(pha[cst], ton[cst])
    assert ((pha[csp] >= 200) & (pha[csp] < 250)).all() and np.array_equal(ton[csp], bg[csp])
    assert np.array_equal(ton[~cst], bg[~cst]) and np.array_equal(pha[~(csp | cst)], bg[~(csp | cst)])


def test_file_round_trip(tmp_path):
    bg, ton, pha, _ = R.seeded(**ARGS)
    src = tmp_path / "in.bin"
    src.write_bytes(R.file_bytes(bg, ton, pha))
    out = tmp_path / "out.bin"
    r = subprocess.run([DRIVER, "copy", str(src), str(out)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    assert out.read_bytes() == src.read_bytes()


def test_rates_from_recorded_trials(tmp_path):
    """MF i fires every p_pre ms before the CS, every p_cs ms in the CS (and
    every p_post after it); the rates estimated from three recorded trials are
    1000 / period Hz."""
    n, trials, pre, cs, post = 24, 3, 500, 500, 500
    rng = np.random.default_rng(3)
    # periods that divide each window the estimate reads (460 ms before the
    # CS, 20 ms at its start, 50 ms at its end) and its start
    p_pre = rng.choice([2, 4, 5, 10, 20], n)
    p_cs = rng.choice([2, 5, 10], n)
    data = ECTrialsData(pre, cs, post, 1, trials, psh={"mf": (10, n)})
    for _ in range(trials):
        for t in range(pre + cs + post):
            p = p_pre if t < pre or t >= pre + cs else p_cs
            data.psh["mf"].update(((t % p) == 0).astype(np.uint8))
    src = tmp_path / "trials.dat"
    src.write_bytes(data.write())
    out = tmp_path / "pop.bin"
    r = subprocess.run([DRIVER, "trials", str(src), str(out)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    cells = data.psh["mf"].cells.T  # [numMF][bins]
    bg, ton, pha = R.from_psh(cells, 10, trials, pre, cs)
    assert out.read_bytes() == R.file_bytes(bg, ton, pha)
    np.testing.assert_allclose(bg, 1000.0 / p_pre, rtol=1e-6)
    np.testing.assert_allclose(ton, 1000.0 / p_cs, rtol=1e-6)
    np.testing.assert_allclose(pha, 1000.0 / p_cs, rtol=1e-6)


def test_oracle_order_matters():
    """The pick order is part of the contract: drawing the CS-phasic MFs
    before the CS-tonic ones (a plausible restatement mistake) changes the
    population."""
    a = dict(ARGS)
    bg, ton, pha, (cst, csp, _) = R.seeded(**a)
    b = dict(a, frac_cst=a["frac_csp"], frac_csp=a["frac_cst"], cst_min=a["csp_min"], cst_max=a["csp_max"],
             csp_min=a["cst_min"], csp_max=a["cst_max"])
    bg2, ton2, pha2, (cst2, csp2, _) = R.seeded(**b)
    assert not np.array_equal(cst, csp2)

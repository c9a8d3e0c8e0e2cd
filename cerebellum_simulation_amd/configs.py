"""Network configurations of the benchmark and the parity tests.

* ``ACT_PARAMS_DEV`` -- the activity parameter set the configurations use: the
  key/value pairs of CBM_Params/actParams_dev.txt, in file order.
* ``conparams_text`` -- a conParams file for given population sizes, in the
  reference's key/value format (CBM_STATE_LIB_DEV/src/params/connectivityparams.cpp:12-87),
  with the GO-GO local matrices of CBM_Params/conParams_dev.txt.
* ``CONFIGS`` -- the workloads: "c0" 16,384 GR (the CPU-runnable case), "c1" the
  1M-GR eyelid network, "c2" the 4M-GR dense network, and small test networks.
* ``mf_train`` -- seeded Bernoulli mossy-fibre spike trains (background 1-10 Hz,
  context 30-60 Hz, a CS subset at 100 Hz during 500-1000 ms of each 2,000 ms trial).
"""
from __future__ import annotations

import numpy as np

ACT_PARAMS_DEV = [
    ("msPerTimeStep", "1"),
    ("msPerHistBinMF", "5"),
    ("eLeakGO", "-70.0"),
    ("eMGluRGO", "-96.0"),
    ("eGABAGO", "-80.0"),
    ("threshMaxGO", "-10.0"),
    ("threshBaseGO", "-33.0"),
    ("gMFIncGO", "0.0075"),
    ("gGRIncGO", "0.0001"),
    ("gGOIncGO", "0.00684"),
    ("coupleRiRjRatioGO", "0.00"),
    ("gMGluRScaleGO", "0"),
    ("gMGluRIncScaleGO", "0.7"),
    ("mGluRScaleGO", "0.1"),
    ("gluScaleGO", "0.01"),
    ("rawGLeakGO", "0.02"),
    ("gMFDecayTGO", "4.5"),
    ("gGRDecayTGO", "4.5"),
    ("gGODecayTGO", "25.0"),
    ("mGluRDecayGO", "0.98"),
    ("gMGluRIncDecayGO", "0.98"),
    ("gMGluRDecayGO", "0.98"),
    ("gluDecayGO", "0.98"),
    ("threshDecayTGO", "20.0"),
    ("eLeakGR", "-70.0"),
    ("eGOGR", "-80.0"),
    ("eMFGR", "0.0"),
    ("threshMaxGR", "-20.0"),
    ("threshBaseGR", "-40.0"),
    ("gMFIncGR", "0.00374"),
    ("gGOIncGR", "0.022"),
    ("gMFDecayTGR", "55.0"),
    ("gGODecayTGR", "25.0"),
    ("threshDecayTGR", "3.0"),
    ("rawGLeakGR", "0.1"),
    ("msPerHistBinGR", "5"),
    ("eLeakSC", "-60.0"),
    ("rawGLeakSC", "0.2"),
    ("gPFDecayTSC", "4.15"),
    ("threshMaxSC", "0.0"),
    ("threshBaseSC", "-50.0"),
    ("threshDecayTSC", "22.0"),
    ("pfIncSC", "0.00132"),
    ("eLeakBC", "-70.0"),
    ("ePCBC", "-70.0"),
    ("rawGLeakBC", "0.2"),
    ("gPFDecayTBC", "4.15"),
    ("gPCDecayTBC", "5.0"),
    ("threshDecayTBC", "10.0"),
    ("threshBaseBC", "-50.0"),
    ("threshMaxBC", "0"),
    ("pfIncConstBC", "0.00054"),
    ("pcIncConstBC", "0.002"),
    ("pfSynWInitPC", "0.5"),
    ("eLeakPC", "-60.0"),
    ("eBCPC", "-80.0"),
    ("eSCPC", "-80.0"),
    ("threshMaxPC", "-48"),
    ("threshBasePC", "-60"),
    ("threshDecayTPC", "5.0"),
    ("rawGLeakPC", "0.2"),
    ("gPFDecayTPC", "4.15"),
    ("gBCDecayTPC", "5.0"),
    ("gSCDecayTPC", "4.15"),
    ("gSCIncConstPC", "0.00005"),
    ("gPFScaleConstPC", "0.00001365"),
    ("gBCScaleConstPC", "0.0007"),
    ("coupleScaleIO", "0.05"),
    ("eLeakIO", "-60.0"),
    ("eNCIO", "-80.0"),
    ("rawGLeakIO", "0.15"),
    ("gNCDecTSIO", "0.5"),
    ("gNCDecTTIO", "70.0"),
    ("gNCDecT0IO", "0.56"),
    ("gNCIncScaleIO", "0.003"),
    ("gNCIncTIO", "300.0"),
    ("threshBaseIO", "-61.0"),
    ("threshMaxIO", "10.0"),
    ("threshDecayTIO", "122"),
    ("msLTDDurationIO", "100"),
    ("msLTDStartAPIO", "-200"),
    ("msLTPStartAPIO", "0"),
    ("grPCLTPIncIO", "0.0002"),
    ("grPCLTDDecIO", "-0.0018"),
    ("maxErrDriveIO", "30"),
    ("eLeakNC", "-65.0"),
    ("ePCNC", "-80.0"),
    ("mfNMDADecayTNC", "50.0"),
    ("mfAMPADecayTNC", "6.0"),
    ("rawGMFNMDAIncNC", "3.0"),
    ("rawGMFAMPAIncNC", "3.0"),
    ("gPCScaleAvgNC", "0.115"),
    ("gPCDecayTNC", "4.15"),
    ("rawGLeakNC", "0.1"),
    ("threshDecayTNC", "5.0"),
    ("threshMaxNC", "-40.0"),
    ("threshBaseNC", "-72.0"),
    ("outIORelPDecTSNC", "40.0"),
    ("outIORelPDecTTNC", "1.0"),
    ("outIORelPDecT0NC", "78.0"),
    ("outIORelPIncScaleNC", "0.25"),
    ("outIORelPIncTNC", "0.8"),
    ("mfSynWInitNC", "0.005"),
    ("mfNCLTDThreshNC", "15"),
    ("mfNCLTPThreshNC", "1"),
    ("mfNCLTDDecNC", "-0.0000125"),
    ("mfNCLTPIncNC", "0.00003"),
]


def act_params_text(overrides: dict | None = None) -> str:
    """actParams file text: ACT_PARAMS_DEV with optional overrides."""
    kv = dict(ACT_PARAMS_DEV)
    for k, v in (overrides or {}).items():
        kv[k] = str(v)
    return "".join(f"{k} {v}\n" for k, v in kv.items()) + "activityParamEnd 1\n"


# GO-GO local matrices of CBM_Params/conParams_dev.txt: GABA to four
# neighbours with probability 1, coupling matrix with probability 0
_GOGO_GABA = [(0, -1, 1.0), (-1, -1, 1.0), (-1, 0, 1.0), (-1, 1, 1.0)]
_GOGO_COUP = [(1, 0, 0), (0, 1, 0), (-1, 0, 0), (0, -1, 0)]


def conparams_text(*, grXP2, grYP2, goXP2, goYP2, numMFP2, numSCP2, numPCP2, numNCP2, numIOP2,
                   numpPCfromPCtoNC=3, grPFVelInGRXPerTStep=147, grAFDelayInTStep=1,
                   gogo_gaba=None, gogo_coup=None) -> str:
    gaba = _GOGO_GABA if gogo_gaba is None else gogo_gaba
    coup = _GOGO_COUP if gogo_coup is None else gogo_coup
    lines = [
        f"numMFP2 {numMFP2}", f"goXP2 {goXP2}", f"goYP2 {goYP2}",
        "GOGOLocalMatrix", f"maxnumpGOInfromGOtoGO {len(gaba)}", f"maxnumpGOOutfromGOtoGO {len(gaba)}",
        *[f"{a} {b} {c}" for a, b, c in gaba],
        "GOGOCoupLocalMatrix", f"maxnumpGOCoupInfromGOtoGO {len(coup)}", f"maxnumpGOCoupOutfromGOtoGO {len(coup)}",
        *[f"{a} {b} {c}" for a, b, c in coup],
        f"grXP2 {grXP2}", f"grYP2 {grYP2}", f"grPFVelInGRXPerTStep {grPFVelInGRXPerTStep}",
        f"grAFDelayInTStep {grAFDelayInTStep}",
        f"numSCP2 {numSCP2}", f"numPCP2 {numPCP2}", f"numpPCfromPCtoNC {numpPCfromPCtoNC}",
        f"numNCP2 {numNCP2}", f"numIOP2 {numIOP2}", "connectivityParamEnd 1",
    ]
    return "\n".join(lines) + "\n"


# The glomerulus-layer parameters of the reference's CBM_Params/conParams.txt
# (the file drivers load by default), for its glomerular generator
# (NetState(..., generator="reference")).
GLOMERULAR_1M = dict(
    glXP2=9, glYP2=7, maxnumpGLfromGLtoGR=80, lownumpGLfromGLtoGR=64, maxnumpGLfromGLtoGO=1,
    maxnumpGLfromGOtoGL=1, numMFP2=11, goXP2=6, goYP2=4, maxnumpGOfromGRtoGO=4096, maxnumpGOfromGLtoGO=16,
    maxnumpGOfromGOtoGL=48, spanGODecDenOnGLX=12, spanGODecDenOnGLY=12, spanGOAscDenOnGRX=2048,
    spanGOAscDenOnGRY=48, spanGOAxonOnGLX=12, spanGOAxonOnGLY=12, grXP2=11, grYP2=9, grPFVelInGRXPerTStep=147,
    grAFDelayInTStep=1, maxnumpGRfromGRtoGO=4, maxnumpGRfromGLtoGR=4, maxnumpGRfromGOtoGR=4, maxnumpGRfromMFtoGR=4,
    spanGRDenOnGLX=4, spanGRDenOnGLY=4, numSCP2=9, numPCP2=5, numpPCfromPCtoNC=3, numNCP2=3, numIOP2=2)
# the GO-GO matrices of that file: GABA to the 8 neighbours with probability
# 0.7 (and a self entry with probability 0), coupling entries all 0
GOGO_GABA_1M = [(0, 1, 0.7), (0, -1, 0.7), (1, 0, 0.7), (-1, 0, 0.7), (1, 1, 0.7), (-1, 1, 0.7), (-1, -1, 0.7),
                (1, -1, 0.7), (0, 0, 0)]
GOGO_COUP_1M = [(0, 1, 0), (0, -1, 0), (1, 0, 0), (-1, 0, 0), (1, 1, 0), (-1, 1, 0), (1, -1, 0), (-1, -1, 0)]

GLOMERULAR = {
    "g1m": GLOMERULAR_1M,
    # the same grid ratios at a quarter of the GRs: 262,144 GR, 16,384 glomeruli, 1,024 MF, 256 GO
    "g262k": {**GLOMERULAR_1M, "glXP2": 8, "glYP2": 6, "numMFP2": 10, "goXP2": 5, "goYP2": 3, "grXP2": 10,
              "grYP2": 8, "spanGOAscDenOnGRX": 1024, "grPFVelInGRXPerTStep": 100},
    # a test network: 2,048 GR on 128 glomeruli, 16 MF, 32 GO; 4 PC, 2 NC, 1 IO
    "g2k": {**GLOMERULAR_1M, "glXP2": 4, "glYP2": 3, "numMFP2": 4, "goXP2": 3, "goYP2": 2, "grXP2": 6,
            "grYP2": 5, "maxnumpGOfromGRtoGO": 224, "maxnumpGOfromGLtoGO": 2, "maxnumpGOfromGOtoGL": 3,
            "spanGODecDenOnGLX": 4, "spanGODecDenOnGLY": 4, "spanGOAxonOnGLX": 4, "spanGOAxonOnGLY": 4,
            "spanGOAscDenOnGRX": 64, "spanGOAscDenOnGRY": 8, "grPFVelInGRXPerTStep": 10, "numSCP2": 4,
            "numPCP2": 2, "numpPCfromPCtoNC": 1, "numNCP2": 1, "numIOP2": 0},
}


def glomerular_conparams_text(name: str = "g1m", **over) -> str:
    """conParams text with the glomerulus layer, in the layout of the
    reference's CBM_Params/conParams.txt."""
    kv = {**GLOMERULAR[name], **over}
    gogo = ["GOGOLocalMatrix", f"maxnumpGOInfromGOtoGO {len(GOGO_GABA_1M)}",
            f"maxnumpGOOutfromGOtoGO {len(GOGO_GABA_1M)}", *[f"{a} {b} {c}" for a, b, c in GOGO_GABA_1M],
            "GOGOCoupLocalMatrix", f"maxnumpGOCoupInfromGOtoGO {len(GOGO_COUP_1M)}",
            f"maxnumpGOCoupOutfromGOtoGO {len(GOGO_COUP_1M)}", *[f"{a} {b} {c}" for a, b, c in GOGO_COUP_1M]]
    lines = [f"{k} {v}" for k, v in kv.items()]
    at = lines.index(f"spanGOAxonOnGLY {kv['spanGOAxonOnGLY']}") + 1
    return "\n".join(lines[:at] + gogo + lines[at:] + ["connectivityParamEnd 1"]) + "\n"


# name -> (conParams keys, generator keys)
CONFIGS = {
    # 16,384 GR; 256 GO; 256 MF; 8 PC; 32 BC; 128 SC; 1 IO; 2 NC
    "c0": (dict(grXP2=8, grYP2=6, goXP2=4, goYP2=4, numMFP2=8, numSCP2=7, numPCP2=3, numNCP2=1,
                numIOP2=0, numpPCfromPCtoNC=1, grPFVelInGRXPerTStep=40),
           dict(n_mf_per_gr=4, n_go_per_gr=3, n_go_out_per_gr=2, n_mf_per_go=16)),
    # 1,048,576 GR; 1,024 GO; 1,024 MF; 32 PC; 128 BC; 512 SC; 4 IO; 8 NC
    "c1": (dict(grXP2=11, grYP2=9, goXP2=6, goYP2=4, numMFP2=10, numSCP2=9, numPCP2=5, numNCP2=3,
                numIOP2=2, numpPCfromPCtoNC=3, grPFVelInGRXPerTStep=147),
           dict(n_mf_per_gr=4, n_go_per_gr=3, n_go_out_per_gr=2, n_mf_per_go=16)),
    # 4,194,304 GR; 4,096 GO; 4,096 MF; 32 PC; 128 BC; 512 SC; 4 IO; 8 NC; 4 GO targets per GR
    # drawn in a window of the 128 x 32 GO grid
    "c2": (dict(grXP2=12, grYP2=10, goXP2=7, goYP2=5, numMFP2=12, numSCP2=9, numPCP2=5, numNCP2=3,
                numIOP2=2, numpPCfromPCtoNC=3, grPFVelInGRXPerTStep=294),
           dict(n_mf_per_gr=4, n_go_per_gr=3, n_go_out_per_gr=4, n_mf_per_go=16, grgo_window=(16, 4))),
    # the cart-pole network: the sizes of the reference's CBM_Params/conParams.txt
    # (1,048,576 GR; 1,024 GO; 2,048 MF; 32 PC; 128 BC; 512 SC; 4 IO; 8 NC) with
    # two microzones
    "cp1m": (dict(grXP2=11, grYP2=9, goXP2=6, goYP2=4, numMFP2=11, numSCP2=9, numPCP2=5, numNCP2=3,
                  numIOP2=2, numpPCfromPCtoNC=3, grPFVelInGRXPerTStep=147),
             dict(n_mf_per_gr=4, n_go_per_gr=3, n_go_out_per_gr=2, n_mf_per_go=16, num_zones=2)),
    # test networks
    "t4k": (dict(grXP2=6, grYP2=6, goXP2=3, goYP2=3, numMFP2=6, numSCP2=4, numPCP2=2, numNCP2=1,
                 numIOP2=0, numpPCfromPCtoNC=1, grPFVelInGRXPerTStep=10),
            dict(n_mf_per_gr=4, n_go_per_gr=3, n_go_out_per_gr=2, n_mf_per_go=8)),
    "t8k2z": (dict(grXP2=7, grYP2=6, goXP2=4, goYP2=3, numMFP2=7, numSCP2=5, numPCP2=3, numNCP2=2,
                   numIOP2=1, numpPCfromPCtoNC=3, grPFVelInGRXPerTStep=20),
              dict(n_mf_per_gr=4, n_go_per_gr=3, n_go_out_per_gr=2, n_mf_per_go=8, num_zones=2)),
}


# Size classes of the ragged batch: 16K, 65K, 262K and 1M GR, with GO, MF
# and PC growing from the 16K-GR network (c0) towards the 1M-GR one (c1).
BATCH_CLASSES = {
    "b16k": CONFIGS["c0"],
    # 65,536 GR; 512 GO; 512 MF; 16 PC; 64 BC; 256 SC; 2 IO; 4 NC
    "b65k": (dict(grXP2=9, grYP2=7, goXP2=5, goYP2=4, numMFP2=9, numSCP2=8, numPCP2=4, numNCP2=2,
                  numIOP2=1, numpPCfromPCtoNC=3, grPFVelInGRXPerTStep=70),
             dict(n_mf_per_gr=4, n_go_per_gr=3, n_go_out_per_gr=2, n_mf_per_go=16)),
    # 262,144 GR; 1,024 GO; 1,024 MF; 32 PC; 128 BC; 512 SC; 4 IO; 8 NC
    "b262k": (dict(grXP2=10, grYP2=8, goXP2=5, goYP2=5, numMFP2=10, numSCP2=9, numPCP2=5, numNCP2=3,
                   numIOP2=2, numpPCfromPCtoNC=3, grPFVelInGRXPerTStep=100),
              dict(n_mf_per_gr=4, n_go_per_gr=3, n_go_out_per_gr=2, n_mf_per_go=16)),
    "b1m": CONFIGS["c1"],
}
# the 256-network ragged batch: 64 x 16K + 96 x 65K + 64 x 262K + 32 x 1M GR
BATCH_MIX = [("b16k", 64), ("b65k", 96), ("b262k", 64), ("b1m", 32)]


def build_state(name: str, seed: int = 0, act_overrides: dict | None = None, *, gogo_inhibition: bool = True,
                gogo_coup=None, **gen_overrides):
    """NetState of a named configuration or batch size class (GR inputs
    drawn with `seed`); gogo_inhibition=False leaves the GO-GO GABA matrix
    empty; gogo_coup replaces the GO-GO coupling matrix ([(dx, dy, p)])."""
    from .network import NetState
    con, gen = CONFIGS[name] if name in CONFIGS else BATCH_CLASSES[name]
    gen = {**gen, **gen_overrides}
    con_text = conparams_text(**con, gogo_gaba=None if gogo_inhibition else [], gogo_coup=gogo_coup)
    return NetState(act_params_text(act_overrides), con_text, innet_seed=seed, **gen)


def batch_plan(mix=None, base_seed: int = 1000):
    """[(size class, seed, gogo_inhibition)] of a ragged batch: every network
    has its own seed (which also picks its CS MF subset) and GO-GO inhibition
    on or off, half and half within each class."""
    plan = []
    for cls, count in (BATCH_MIX if mix is None else mix):
        for i in range(count):
            seed = base_seed + len(plan) * 7919
            plan.append((cls, seed, i % 2 == 0))
    return plan


def cs_mfs(num_mf: int, seed: int = 0, cs_frac: float = 0.03, ctx_frac: float = 0.10) -> np.ndarray:
    """Indices of the CS MFs that mf_train(num_mf, ..., seed=seed) drives."""
    rng = np.random.default_rng(seed)
    n_ctx = int(round(ctx_frac * num_mf))
    rng.permutation(num_mf)
    rng.uniform(1.0, 10.0, num_mf)
    rng.uniform(30.0, 60.0, n_ctx)
    return np.sort(rng.choice(num_mf, max(1, int(cs_frac * num_mf)), replace=False))


def trained_pfpc_weights(state, cs: np.ndarray, w_cs: float = 0.0, w_other: float = 0.5) -> np.ndarray:
    """PF->PC weights of a network that has learned the CS: GRs with at least
    one CS MF among their inputs at w_cs, the others at w_other."""
    s = state.sizes
    pm = state.field("P_GR_FROM_MF_TO_GR")
    nm = state.field("NUMP_GR_FROM_MF_TO_GR")
    valid = np.arange(pm.shape[1])[None, :] < nm[:, None]
    hit = (np.isin(pm, cs) & valid).any(axis=1)
    return np.where(hit, np.float32(w_cs), np.float32(w_other)).astype(np.float32)


def mf_train(num_mf: int, steps: int, seed: int = 0, *, t0: int = 0, trial_ms: int = 2000,
             cs: bool = True, cs_frac: float = 0.03, cs_on: int = 500, cs_off: int = 1000,
             cs_hz: float = 100.0, ctx_frac: float = 0.10, spike_seed: int | None = None) -> np.ndarray:
    """uint8 [steps, num_mf] Bernoulli spike train.  Rates per MF are drawn
    once per seed: background MFs U[1,10] Hz, context MFs U[30,60] Hz; when
    cs is set a seeded subset of cs_frac of the MFs fires at cs_hz during
    [cs_on, cs_off) of every trial.  Step k is trial time (t0 + k) % trial_ms.
    spike_seed, when given, draws the spikes from a generator of its own, so
    trials of one network (same rates and CS subset) get fresh spikes."""
    rng = np.random.default_rng(seed)
    n_ctx = int(round(ctx_frac * num_mf))
    perm = rng.permutation(num_mf)
    rate = rng.uniform(1.0, 10.0, num_mf)
    rate[perm[:n_ctx]] = rng.uniform(30.0, 60.0, n_ctx)
    p = np.broadcast_to((rate * 1e-3).astype(np.float64), (steps, num_mf)).copy()
    if cs:
        n_cs = max(1, int(cs_frac * num_mf))
        cs_idx = rng.choice(num_mf, n_cs, replace=False)
        tt = (t0 + np.arange(steps)) % trial_ms
        in_cs = (tt >= cs_on) & (tt < cs_off)
        p[np.ix_(in_cs, cs_idx)] = cs_hz * 1e-3
    draw = rng if spike_seed is None else np.random.default_rng(spike_seed)
    return (draw.random((steps, num_mf)) < p).astype(np.uint8)


def mf_rates(num_mf: int, seed: int = 0, *, cs_frac: float = 0.03, ctx_frac: float = 0.10):
    """(rate in Hz per MF, CS MF indices) that mf_train(num_mf, ..., seed=seed) uses."""
    rng = np.random.default_rng(seed)
    n_ctx = int(round(ctx_frac * num_mf))
    perm = rng.permutation(num_mf)
    rate = rng.uniform(1.0, 10.0, num_mf)
    rate[perm[:n_ctx]] = rng.uniform(30.0, 60.0, n_ctx)
    cs_idx = rng.choice(num_mf, max(1, int(cs_frac * num_mf)), replace=False)
    return rate, np.sort(cs_idx)


def order_gr_by_inputs(state, key_mfs) -> np.ndarray:
    """Relabel the GRs of `state` in place so that, within every SC block of
    numGR/numSC consecutive GRs, the GRs with the most inputs from `key_mfs`
    come first (stable otherwise).  Every PF target block (SC, BC, PC, IO) is
    a union of SC blocks and every GR field is GR-major, so this is an exact
    relabelling of the network; it makes consecutive GRs alike, which is what
    a summary per cluster of consecutive GRs (the surrogate's) needs.
    Returns the permutation (new GR g was old GR perm[g])."""
    s = state.sizes
    pm = state.field("P_GR_FROM_MF_TO_GR")
    nm = state.field("NUMP_GR_FROM_MF_TO_GR")
    valid = np.arange(pm.shape[1])[None, :] < nm[:, None]
    hits = (np.isin(pm, np.asarray(key_mfs)) & valid).sum(axis=1)
    blk = s.numGR // s.numSC
    perm = np.empty(s.numGR, np.int64)
    for b in range(s.numSC):
        lo = b * blk
        perm[lo:lo + blk] = lo + np.argsort(-hits[lo:lo + blk], kind="stable")
    from ._abi import FIELDS, field_info
    for name in FIELDS:
        info = field_info(FIELDS[name], s)
        if info["kind"] == 2:
            continue
        if name == "PFPC_SYNW":
            for z in range(state.num_zones):
                a = state.field(name, z)
                a[...] = a.reshape(-1)[perm].reshape(a.shape)
        elif not info["zoned"] and info["rows"] == s.numGR:
            a = state.field(name)
            a[...] = a[perm]
    return perm


def us_err_drive(steps: int, num_zones: int = 1, *, t0: int = 0, trial_ms: int = 2000,
                 us_ms: int = 999) -> np.ndarray:
    """float32 [steps, num_zones] relative error drive: 1.0 at trial time
    us_ms (the step before CS offset, as ECManagementDelay::calcSimActivity
    delivers it, Eyelid_Sim_CBM_Dev/src/ectrial/ecmanagementdelay.cpp:207-210)."""
    e = np.zeros((steps, num_zones), np.float32)
    tt = (t0 + np.arange(steps)) % trial_ms
    e[tt == us_ms, :] = 1.0
    return e

"""Trial-level parity: full 2,000-ms eyelid delay trials on the GPU against the
oracle (binary32 mimic and binary64 replay), with the trial driver of
cerebellum_simulation_amd/eyelid.py and the device spike recorder.

The networks are the 262K-GR size class (8 NC, 4 IO; two trials) and the
1M-GR eyelid network (32 PC, 8 NC, 4 IO; one trial), each with PF->PC weights
of a network that has learned the CS (GRs with a CS MF input at 0), so that the
eyelid closes during the CS and the conditioned-response onset is defined.
Bars, against the binary32 mimic (the reference's own arithmetic), on every
trial: PF->PC weights within 1e-4 relative (L2) -- in practice identical --,
identical PC and NC spikes and eyelid trace, CR onsets within 1 ms (or neither
has one).  Against the binary64 replay, on the first trial: PC and NC spike
counts within 1 %, CR onset within 1 ms, weights within W64_BAR relative
(1e-3 at 262K GR; 5e-3 at 1M GR, where the distance after one trial is
2.8e-3: each of the 4 IO cells gates the plasticity of a quarter of the GRs,
and one IO spike that binary64 rounding moves by a step shifts the LTD/LTP
window of 262,144 weights at once).  Later
trials are printed, not asserted, against binary64: a binary64 network rounds
differently from the binary32 one the reference computes, its IO/PC/NC spikes
drift apart by a few per trial and the plasticity they gate follows (weights
3e-4 apart after one trial here, 8e-3 after two), a property of the two
arithmetics rather than of the kernels.
"""
import numpy as np
import pytest

from cerebellum_simulation_amd import Network, configs
from cerebellum_simulation_amd.eyelid import DelayTrials, cr_onset

pytestmark = pytest.mark.gpu

TRIAL_MS = 2000
W64_BAR = {"b262k": 1e-3, "c1": 5e-3}


def _rates_close(a, b, tol=0.01):
    a, b = int(a.sum()), int(b.sum())
    return abs(a - b) <= tol * max(a, b, 1)


@pytest.mark.parametrize("cfg,trials", [("b262k", 2), ("c1", 1)])
def test_full_trials_against_oracle(gpu, cfg, trials):
    import torch
    import orc as O
    print(f"[case] {trials} full eyelid trial(s), {cfg}, GPU vs oracle", flush=True)
    st = configs.build_state(cfg, seed=0)
    s = st.sizes
    mseed = 100
    cs = configs.cs_mfs(s.numMF, seed=mseed)
    w = configs.trained_pfpc_weights(st, cs)
    net = Network.from_state(st)
    net.write("PFPC_SYNW", w)
    net.set_recorder(["PC", "NC"], TRIAL_MS)
    o32 = O.Oracle.from_state(st, precision=32, openmp=True)
    o64 = O.Oracle.from_state(st, precision=64, openmp=True)
    for o in (o32, o64):
        o.write("PFPC_SYNW", w)
    drv = {k: DelayTrials(st.params, s.numNC) for k in ("gpu", "o32", "o64")}
    for trial in range(trials):
        train = configs.mf_train(s.numMF, TRIAL_MS, seed=mseed + trial)
        dtr = torch.from_numpy(train).cuda()
        sp = {"gpu": drv["gpu"].run_trial_device(net, dtr, st.num_zones),
              "o32": drv["o32"].run_trial_stepwise(o32, train, st.num_zones, s.numPC),
              "o64": drv["o64"].run_trial_stepwise(o64, train, st.num_zones, s.numPC)}
        eye = {k: drv[k].eyelid.trace(sp[k]["NC"]) for k in sp}
        onset = {k: cr_onset(eye[k], 500, 1000) for k in eye}
        wg = net.read("PFPC_SYNW").astype(np.float64)
        w32 = o32.read("PFPC_SYNW").astype(np.float64)
        w64 = o64.read_f64("PFPC_SYNW")
        rel64 = float(np.linalg.norm(wg - w64) / np.linalg.norm(w64))
        print(f"{cfg} trial {trial}: NC {[int(sp[k]['NC'].sum()) for k in sp]}, PC {[int(sp[k]['PC'].sum()) for k in sp]}, "
              f"onset {onset}, eyelid max {[float(eye[k].max()) for k in eye]}, w rel vs f64 {rel64:.2e}, "
              f"w equal to f32 {np.mean(wg == w32):.6f}", flush=True)
        rel32 = float(np.linalg.norm(wg - w32) / np.linalg.norm(w32))
        assert rel32 <= 1e-4
        assert np.array_equal(sp["gpu"]["NC"], sp["o32"]["NC"]) and np.array_equal(sp["gpu"]["PC"], sp["o32"]["PC"])
        assert float(np.max(np.abs(eye["gpu"] - eye["o32"]))) == 0.0
        assert (onset["gpu"] is None) == (onset["o32"] is None)
        if onset["gpu"] is not None:
            assert abs(onset["gpu"] - onset["o32"]) <= 1
        if trial == 0:
            # the pre-trained network closes the lid during the CS
            assert onset["gpu"] is not None and onset["o64"] is not None
            assert abs(onset["gpu"] - onset["o64"]) <= 1
            for pop in ("NC", "PC"):
                assert _rates_close(sp["gpu"][pop], sp["o64"][pop]), pop
            assert rel64 <= W64_BAR[cfg]

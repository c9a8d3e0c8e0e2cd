"""Closed loops (cbm_loop_begin / cbm_loop_end): one resident run that takes
each step's MF bytes and error drive from pinned host memory and hands the
step's spikes back there must give the bits of the same cbm_step_io pairs run
without the loop (one launch and one synchronisation per step), which
test_gpu_parity holds to the oracle: every returned spike, every state and
run-time field afterwards, and the run that follows the loop.  The per-ms call
pattern is that of a closed-loop driver (Matthew_Sim_CBM/src/simthread.cpp:
60-165: updateMFInput, calcActivity, exportAPNC per zone), with error drives
set between steps as Microzone::update does (src/microzone.cpp:31-55)."""
import numpy as np
import pytest

from cerebellum_simulation_amd import Network, configs

from test_gpu_parity import compare

pytestmark = pytest.mark.gpu

# 4K-GR and 8K-GR two-zone networks (one 64-GR chunk per span or none), the
# 1M-GR network of the headline metric, and the 1M-GR two-zone network of
# 2,048 MFs of the cart-pole task (whose MF->NC synapse arrays do not fit in
# the microzone block's LDS beside its other fields and stay in global memory)
CASES = [("t4k", 70, "NC"), ("t8k2z", 90, "NC"), ("c1", 45, "PC"), ("cp1m", 40, "NC")]


@pytest.mark.parametrize("name,steps,pop", CASES, ids=[c[0] for c in CASES])
def test_loop_equals_step_io(gpu, name, steps, pop):
    import torch
    print(f"[case] loop {name} steps={steps} pop={pop}", flush=True)
    st = configs.build_state(name, seed=3)
    zones = st.num_zones
    tr = configs.mf_train(st.sizes.numMF, steps + 20, seed=5, t0=480, trial_ms=2000, cs_on=500, cs_off=1000)
    ref = Network.from_state(st, step_mode=4)
    res = Network.from_state(st, step_mode=5)
    assert not ref.resident and res.resident

    # (a loop's launch holds every CU until cbm_loop_end: the reference
    # network steps first)
    def drive(net, k):
        if k % 11 == 4:
            net.update_err_drive(k % zones, 0.5)
        return net.step_io(tr[k], pop)

    want = [drive(ref, k) for k in range(steps)]
    ref.sync()
    res.loop_begin(steps + 8, pop)
    with pytest.raises(RuntimeError):
        res.sync()  # only step_io and the error drive while the loop runs
    with pytest.raises(RuntimeError):
        res.set_ginc_gr_to_go(1.0)
    got = [drive(res, k) for k in range(steps)]
    assert res.loop_end() == steps
    for k in range(steps):
        assert np.array_equal(got[k], want[k]), f"{name}: step {k}: the loop's {pop} spikes differ"
    ref.sync()
    res.sync()
    assert int(res.read("CUR_TIME")[0]) == steps
    bad = compare(res, ref, zones)
    assert not bad, f"{name}: state after the loop differs from step-by-step calls: {bad}"
    # the network goes on from the loop's state: a resident run with a knob
    # changed, then a second loop
    dtr = torch.from_numpy(tr[steps:steps + 12]).cuda()
    for net in (ref, res):
        net.set_ginc_gr_to_go(net.params.gIncGRtoGO * 1.25)
        net.run_mf_train(dtr)
    ref.sync()
    res.sync()
    want = [ref.step_io(tr[k], pop) for k in range(steps + 12, steps + 20)]
    ref.sync()
    res.loop_begin(8, pop)
    got = [res.step_io(tr[k], pop) for k in range(steps + 12, steps + 20)]
    assert res.loop_end() == 8
    for i in range(8):
        assert np.array_equal(got[i], want[i]), f"{name}: second loop, step {steps + 12 + i}"
    ref.sync()
    res.sync()
    bad = compare(res, ref, zones)
    assert not bad, f"{name}: state after the second loop differs: {bad}"
    # a loop ended before any step leaves the network as it was
    res.loop_begin(4, pop)
    assert res.loop_end() == 0
    assert not compare(res, ref, zones)
    ref.close()
    res.close()

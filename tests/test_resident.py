"""Resident runs of cbm_run_mf_train: the whole run is one launch of one
workgroup per CU whose GR workgroups keep their cells' state in registers
(connectivity records in LDS) and whose two small-population workgroups keep
theirs in LDS from the first step to the last, with step counters and flags
between steps.  They must give the bits of the same steps
launched one by one (the fused even-split step, which test_gpu_parity,
test_trial and test_batch_c3 hold to the oracle): every state and run-time
field, and every recorded spike, over runs of several lengths (a run of one
step launches step by step), with the CS, the US drive, plasticity steps and
driver knobs changed between runs."""
import numpy as np
import pytest

from cerebellum_simulation_amd import Network, configs

from test_gpu_parity import compare

pytestmark = pytest.mark.gpu

# (network, run lengths): 1M GR (5 cells per thread), 262K GR (2 cells per
# thread), the 16K-GR network (one or two 64-GR chunks per span) and the 8K-GR
# two-zone network (spans without cells, which only fold)
CASES = [("c1", [150, 1, 249, 2, 198]), ("b262k", [300, 1, 299]), ("c0", [97, 203, 1, 99]),
         ("t8k2z", [64, 1, 135, 100])]
POPS = ["PC", "BC", "SC", "IO", "NC", "GO", "PF"]


@pytest.mark.parametrize("name,runs", CASES, ids=[c[0] for c in CASES])
def test_resident_run_equals_step_by_step(gpu, name, runs):
    import torch
    print(f"[case] resident {name} runs={runs}", flush=True)
    st = configs.build_state(name, seed=7)
    zones = st.num_zones
    steps = sum(runs)
    tr = configs.mf_train(st.sizes.numMF, steps, seed=11, t0=430, trial_ms=2000, cs_on=500, cs_off=1000)
    err = configs.us_err_drive(steps, zones, t0=430, trial_ms=2000, us_ms=520)
    dtr = torch.from_numpy(tr).cuda()
    ref = Network.from_state(st, step_mode=4)
    res = Network.from_state(st, step_mode=5)
    assert not ref.resident and res.resident
    for net in (ref, res):
        net.set_recorder(POPS, steps)
        net.update_err_drive(0, 0.25)  # pending before the first run
    k0 = 0
    for i, n in enumerate(runs):
        for net in (ref, res):
            if i == 1:
                net.set_ginc_gr_to_go(net.params.gIncGRtoGO * 1.5)
                net.set_grpc_plast_steps(zones - 1, -0.01, 0.002)
            if i == 2:
                net.reset_ginc_gr_to_go()
            net.run_mf_train(dtr[k0:k0 + n], err_drive=err[k0:k0 + n])
        k0 += n
    ref.sync()
    res.sync()
    assert int(res.read("CUR_TIME")[0]) == steps
    bad = compare(res, ref, zones)
    assert not bad, f"{name}: resident run differs from step-by-step launches: {bad}"
    for pop in POPS:
        a, b = res.read_recorded(pop, 1, steps), ref.read_recorded(pop, 1, steps)
        assert np.array_equal(a, b), f"{name}: recorded {pop} spikes differ"
    assert res.read_recorded("PC", 1, steps).any() and (res.read("APBUF_GR") != 0).any()
    ref.close()
    res.close()

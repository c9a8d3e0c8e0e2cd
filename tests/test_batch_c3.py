"""The ragged batch at the sizes of the 256-network workload: one network of
each size class (16K, 65K, 262K and 1M GR), GO-GO inhibition on and off,
stepped by the batch kernel with the steps captured in HIP graphs.  Every
member equals its own run bit for bit, and the 262K- and 1M-GR members equal
the binary32 oracle (the reference's step restated on the CPU) in every
field, as the single-network 1M-GR test does."""
import numpy as np
import pytest

from cerebellum_simulation_amd import Batch, Network, configs

from test_batch import COMPARE, _assert_same, _snapshot
from test_gpu_parity import assert_parity, compare

pytestmark = pytest.mark.gpu

PLAN = [("b16k", 1101, True), ("b65k", 1102, False), ("b262k", 1103, True), ("b1m", 1104, False),
        ("b262k", 1105, False), ("b1m", 1106, True)]
STEPS = 64


@pytest.mark.parametrize("graph_steps", [16])
def test_c3_classes_equal_own_runs_and_oracle(gpu, graph_steps):
    import torch

    import orc as O
    print(f"[case] config-3 classes {[p[0] for p in PLAN]} graph_steps={graph_steps}", flush=True)
    states = [configs.build_state(c, seed=s, gogo_inhibition=inh) for c, s, inh in PLAN]
    assert {st.sizes.numGR for st in states} == {16384, 65536, 262144, 1048576}
    trains = [configs.mf_train(st.sizes.numMF, STEPS, seed=s, t0=460, trial_ms=2000, cs_on=500, cs_off=1000)
              for st, (_, s, _) in zip(states, PLAN)]
    errs = [configs.us_err_drive(STEPS, st.num_zones, t0=460, trial_ms=2000, us_ms=499) for st in states]
    dtr = [torch.from_numpy(t).cuda() for t in trains]
    solo = [Network.from_state(st) for st in states]
    for n, t, e in zip(solo, dtr, errs):
        n.run_mf_train(t, err_drive=e)
    grouped = [Network.from_state(st) for st in states]
    b = Batch(grouped, graph_steps=graph_steps)
    assert b.info()["graph_steps"] == graph_steps
    b.run_mf_train(dtr, err_drive=errs)
    b.sync()
    for i, (x, y) in enumerate(zip(solo, grouped)):
        _assert_same(_snapshot(x), _snapshot(y), f"member {i} {PLAN[i]}")
    for i, (st, net) in enumerate(zip(states, grouped)):
        if st.sizes.numGR < 262144:
            continue
        print(f"oracle for member {i} {PLAN[i]} ({st.sizes.numGR} GR)", flush=True)
        o32 = O.Oracle.from_state(st, precision=32)
        o32.run_mf_train(trains[i], err_drive=errs[i])
        bad = compare(net, o32, st.num_zones)
        print(f"member {i}: {bad}", flush=True)
        assert_parity(bad, where=f"batched member {i} {PLAN[i]}")
        assert int(net.read("CUR_TIME")[0]) == STEPS
        assert (net.read("APBUF_GR") != 0).any()
        o32.close()
    b.close()

"""The ragged batch (cbm_batch_*): one launch per step for many networks.

A batched run must leave every network exactly as its own run leaves it
(same kernels, same per-network state and order of operations), with plain
launches and with captured graphs; the autotuner must restore every network.
The single-network path it is compared with is itself checked against the
oracle in test_gpu_parity.py.
"""
import numpy as np
import pytest

from cerebellum_simulation_amd import Batch, Network, configs

pytestmark = pytest.mark.gpu

COMPARE = ("APBUF_GR", "VM_GR", "THRESH_GR", "GKCA_GR", "HIST_GR", "PFPC_SYNW", "VM_GO", "APBUF_GO",
           "GO_GABA_OUT_SYNSCALE_GOGO", "VM_PC", "APBUF_NC", "VM_IO", "MF_SYNW_NC", "ZONE_RNG_STATE",
           "PIPE_PFPC_SUM", "PIPE_GRGO_SUM", "PIPE_PFBC_SUM", "CUR_TIME", "AP_GO", "PFPC_PLAST_TIMER_IO")

MIX = [("b16k", 2), ("t4k", 1), ("t8k2z", 1), ("b65k", 1)]


def _plan():
    return configs.batch_plan(MIX, base_seed=77)


def _build(plan):
    states = [configs.build_state(cls, seed=seed, gogo_inhibition=inh) for cls, seed, inh in plan]
    return states, [Network.from_state(st) for st in states]


def _trains(states, steps, seed0=0):
    import torch
    trs = [torch.from_numpy(configs.mf_train(st.sizes.numMF, steps, seed=seed0 + i)).cuda()
           for i, st in enumerate(states)]
    errs = [configs.us_err_drive(steps, st.num_zones, t0=0, trial_ms=steps, us_ms=steps // 2) for st in states]
    return trs, errs


def _snapshot(net):
    out = {}
    for f in COMPARE:
        for zz in range(net.sizes.numZones if net.info(f)["zoned"] else 1):
            out[(f, zz)] = net.read(f, zz)
    return out


def _assert_same(a, b, where):
    for k in a:
        assert np.array_equal(a[k], b[k]), (where, k)


@pytest.mark.parametrize("graph_steps", [0, 16])
def test_batch_equals_individual_runs(gpu, graph_steps):
    print(f"[case] batch vs individual, graph_steps={graph_steps}", flush=True)
    plan = _plan()
    states, solo = _build(plan)
    _, grouped = _build(plan)
    steps = 120
    trs, errs = _trains(states, steps)
    for net, tr, e in zip(solo, trs, errs):
        net.run_mf_train(tr, err_drive=e)
    b = Batch(grouped, graph_steps=graph_steps)
    b.run_mf_train(trs[:], n=60, err_drive=[e[:60] for e in errs])
    b.run_mf_train([t[60:] for t in trs], err_drive=[e[60:] for e in errs])
    b.sync()
    for i, (x, y) in enumerate(zip(solo, grouped)):
        _assert_same(_snapshot(x), _snapshot(y), f"net {i} {plan[i]}")
    # the networks stay usable alone after a batched run
    for x, y in zip(solo, grouped):
        x.calc_activity(3)
        y.calc_activity(3)
        _assert_same(_snapshot(x), _snapshot(y), "after solo steps")
    b.close()


def test_batch_autotune_restores_state(gpu):
    print("[case] batch autotune", flush=True)
    plan = _plan()
    states, nets = _build(plan)
    trs, errs = _trains(states, 40, seed0=5)
    b = Batch(nets)
    b.run_mf_train([t[:10] for t in trs])
    before = [_snapshot(n) for n in nets]
    cpt = b.autotune([t[10:] for t in trs], steps=8)
    assert cpt in (1, 2) and b.info()["cells_per_thread"] == cpt
    for n, s0 in zip(nets, before):
        _assert_same(s0, _snapshot(n), "after autotune")
    # and the batch still runs correctly with the chosen shape
    _, solo = _build(plan)
    for n, t in zip(solo, trs):
        n.run_mf_train(t[:10])
        n.run_mf_train(t[10:30])
    b.run_mf_train([t[10:30] for t in trs])
    b.sync()
    for x, y in zip(solo, nets):
        _assert_same(_snapshot(x), _snapshot(y), "after tuned run")
    b.close()


def test_batch_errors(gpu):
    from cerebellum_simulation_amd import CbmError
    st = configs.build_state("t4k", seed=1)
    n1 = Network.from_state(st)
    b = Batch([n1])
    with pytest.raises(CbmError, match="one batch"):
        Batch([n1])
    from cerebellum_simulation_amd import hip_lib
    assert hip_lib().cbm_destroy(n1._h) != 0  # still in a batch
    b.close()


@pytest.mark.parametrize("pack", ["arrays", "mixed"])
def test_batch_connectivity_formats(gpu, pack):
    """Packs whose networks all keep the five connectivity arrays, or mix
    them with compact records, take their own kernel builds; each network
    still ends as its own run leaves it."""
    print(f"[case] batch pack format {pack}", flush=True)
    plan = _plan()
    fmts = [1 if pack == "arrays" or i % 2 else 0 for i in range(len(plan))]
    states = [configs.build_state(cls, seed=seed, gogo_inhibition=inh) for cls, seed, inh in plan]
    solo = [Network.from_state(st, conn_format=f) for st, f in zip(states, fmts)]
    grouped = [Network.from_state(st, conn_format=f) for st, f in zip(states, fmts)]
    steps = 60
    trs, errs = _trains(states, steps, seed0=3)
    for net, tr, e in zip(solo, trs, errs):
        net.run_mf_train(tr, err_drive=e)
    b = Batch(grouped)
    b.run_mf_train(trs, err_drive=errs)
    b.sync()
    for i, (x, y) in enumerate(zip(solo, grouped)):
        _assert_same(_snapshot(x), _snapshot(y), f"net {i} {plan[i]} format {fmts[i]}")
    b.close()

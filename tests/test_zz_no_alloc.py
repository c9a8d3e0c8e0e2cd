"""cbm_batch_run_mf_train_noalloc: the batch run that allocates nothing
(error-drive rows staged in the library's fixed device array, plain
launches) leaves every network exactly as cbm_batch_run_mf_train does, over
two calls in a row, the second smaller; a run too large for the array is
refused with CBM_E_ARG before anything changes."""
import ctypes as C

import numpy as np
import pytest

from cerebellum_simulation_amd import Batch, Network, _abi, configs

from test_batch import _assert_same, _snapshot

pytestmark = pytest.mark.gpu

PLAN = [("b16k", 501, True), ("t8k2z", 502, False), ("b16k", 503, False), ("t4k", 504, True)]


def _nets():
    states = [configs.build_state(c, seed=s, gogo_inhibition=inh) for c, s, inh in PLAN]
    return states, [Network.from_state(st) for st in states]


def _args(trains, errs, k0, n):
    rows = [t[k0:k0 + n] for t in trains]
    ptrs = (C.c_void_p * len(rows))(*[r.data_ptr() for r in rows])
    pitch = (C.c_size_t * len(rows))(*[r.stride(0) for r in rows])
    keep = [None if e is None else np.ascontiguousarray(e[k0:k0 + n], np.float32) for e in errs]
    err = (C.c_void_p * len(keep))(*[None if e is None else e.ctypes.data for e in keep])
    return ptrs, pitch, err, keep


def test_noalloc_batch_run_equals_allocating_run(gpu):
    import torch
    print("[case] no-alloc batch run vs cbm_batch_run_mf_train", flush=True)
    L = _abi.hip_lib()
    states, ref = _nets()
    _, got = _nets()
    steps = 60
    trains = [torch.from_numpy(configs.mf_train(st.sizes.numMF, steps, seed=s, t0=470, trial_ms=600,
                                                cs_on=500, cs_off=560)).cuda() for st, (_, s, _) in zip(states, PLAN)]
    errs = [configs.us_err_drive(steps, st.num_zones, t0=470, trial_ms=600, us_ms=499) for st in states]
    bref, bgot = Batch(ref), Batch(got)
    # first call: 44 steps, every network with an error drive
    p, q, e, keep = _args(trains, errs, 0, 44)
    assert L.cbm_batch_run_mf_train(bref._b, p, q, 44, e) == 0
    assert L.cbm_batch_run_mf_train_noalloc(bgot._b, p, q, 44, e) == 0
    bref.sync()
    bgot.sync()
    for i, (x, y) in enumerate(zip(ref, got)):
        _assert_same(_snapshot(x), _snapshot(y), f"first call, net {i} {PLAN[i]}")
    # second call, smaller: 16 steps, drive for two networks only
    errs2 = [errs[0], None, errs[2], None]
    p, q, e, keep = _args(trains, errs2, 44, 16)
    assert L.cbm_batch_run_mf_train(bref._b, p, q, 16, e) == 0
    assert L.cbm_batch_run_mf_train_noalloc(bgot._b, p, q, 16, e) == 0
    bref.sync()
    bgot.sync()
    for i, (x, y) in enumerate(zip(ref, got)):
        _assert_same(_snapshot(x), _snapshot(y), f"second call, net {i} {PLAN[i]}")
    assert int(got[0].read("CUR_TIME")[0]) == steps
    assert any((n.read("APBUF_PC") != 0).any() for n in got)
    bref.close()
    bgot.close()


def test_noalloc_batch_run_refuses_oversized_drive(gpu):
    import torch
    L = _abi.hip_lib()
    states, nets = _nets()
    zones = sum(st.num_zones for st in states)
    n = _abi.NOALLOC_ERR_FLOATS // zones + 1
    print(f"[case] no-alloc refusal n={n} zones={zones}", flush=True)
    trains = [torch.zeros((n, st.sizes.numMF), dtype=torch.uint8, device="cuda") for st in states]
    errs = [np.zeros((n, st.num_zones), np.float32) for st in states]
    b = Batch(nets)
    before = [_snapshot(x) for x in nets]
    p, q, e, keep = _args(trains, errs, 0, n)
    rc = L.cbm_batch_run_mf_train_noalloc(b._b, p, q, n, e)
    assert rc == _abi.ERRORS["CBM_E_ARG"], rc
    L.cbm_last_error.restype = C.c_char_p
    assert b"CBM_NOALLOC_ERR_FLOATS" in L.cbm_last_error()
    b.sync()
    for x, s0 in zip(nets, before):
        _assert_same(s0, _snapshot(x), "after the refused call")
    b.close()

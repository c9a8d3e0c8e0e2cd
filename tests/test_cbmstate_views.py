"""The CBMState views the drivers call (CBM_STATE_LIB_DEV/CBMStateInclude/
interfaces/cbmstate.h:47-53: getConnectivityParams, getActivityParams,
getInnetConState, getMZoneActState), printed by tests/cpp/views_driver, equal
the state's sizes, parameter files and arrays (the same state built from the
same files and seeds in Python)."""
import os
import subprocess

import numpy as np

from cerebellum_simulation_amd import configs
from cerebellum_simulation_amd.network import NetState

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "cpp", "build", "views_driver")


def _lists(st, num, lst, rows):
    n = st.field(num).ravel()
    p = st.field(lst).reshape(rows, -1)
    return [list(p[i, :n[i]]) for i in range(rows)]


def test_views_equal_the_state(tmp_path):
    act, con = tmp_path / "act.txt", tmp_path / "con.txt"
    act.write_text(configs.act_params_text())
    con_text = configs.conparams_text(**configs.CONFIGS["c0"][0])
    con.write_text(con_text)
    seed = 3
    r = subprocess.run([DRIVER, str(act), str(con), str(seed)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    out = {}
    for line in r.stdout.splitlines():
        k, *v = line.split()
        out[k] = v
    st = NetState(configs.act_params_text(), con_text, num_zones=1, innet_seed=seed, n_mf_per_go=16)
    s = st.sizes
    assert [int(x) for x in out["sizes"]] == [s.numGR, s.numGO, s.numMF, s.numSC, s.numBC, s.numPC, s.numNC,
                                              s.numIO, s.grX, s.grY, s.goX, s.goY, 0]
    assert out["con"] == ["grXP2", str(int(round(np.log2(s.grX))))]
    acts = dict(configs.ACT_PARAMS_DEV)
    assert out["act"][:4] == ["msPerTimeStep", "1", "msLTDStartAPIO", str(int(float(acts["msLTDStartAPIO"])))]
    assert int(out["act"][5]) == len(acts)
    assert out["setParam"] == ["unknown", "0", "known", "1", "now", "-150"]
    grMF = _lists(st, "NUMP_GR_FROM_MF_TO_GR", "P_GR_FROM_MF_TO_GR", s.numGR)
    grGO = _lists(st, "NUMP_GR_FROM_GO_TO_GR", "P_GR_FROM_GO_TO_GR", s.numGR)
    grOut = _lists(st, "NUMP_GR_FROM_GR_TO_GO", "P_GR_FROM_GR_TO_GO", s.numGR)
    mfGO = _lists(st, "NUMP_MF_FROM_MF_TO_GO", "P_MF_FROM_MF_TO_GO", s.numMF)
    gaba = _lists(st, "NUMP_GO_GABA_OUT_GOGO", "P_GO_GABA_OUT_GOGO", s.numGO)
    ints = lambda k: [int(x) for x in out[k]]  # noqa: E731
    assert ints("grMF5") == grMF[5] and ints("grGO5") == grGO[5]
    assert ints("goGR3") == [g for g in range(s.numGR) for t in grOut[g] if t == 3]
    assert ints("goGOGR3") == [g for g in range(s.numGR) for t in grGO[g] if t == 3]
    assert ints("mfGO2") == mfGO[2]
    assert ints("goMF4") == [m for m in range(s.numMF) for t in mfGO[m] if t == 4]
    assert ints("goOut1") == gaba[1]
    assert ints("goIn1") == [g for g in range(s.numGO) for t in gaba[g] if t == 1]
    # delete, delete again (absent), add, add again (present); the pair moves to the end
    assert ints("gogo") == [1, 0, 1, 0]
    assert ints("goOut1after") == gaba[1][1:] + gaba[1][:1]
    assert ints("pop") == [s.numGR, s.numGR, s.numGO]
    w0 = st.field("PFPC_SYNW")
    assert int(out["w"][0]) == s.numGR and float(out["w"][1]) == float(w0.ravel()[0])
    init = float(acts["pfSynWInitPC"])
    assert [float(x) for x in out["wreset"]] == [init, init]

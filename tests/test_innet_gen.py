"""The reference's glomerular input-net generator (CBM_GEN_REFERENCE,
csrc/state/innet_gen.cpp), with parameter files in the layout of
CBM_Params/conParams.txt.

On the CPU:
* the library's lists equal those of the oracle's independent restatement
  (oracle/orc_innet.c) slot for slot, for a 2K-GR network at two seeds, a
  variant whose grids have different X and Y ratios, and a 262K-GR network with
  the grid ratios of conParams.txt;
* invariants the reference's code fixes, checked on the library's lists:
  every glomerulus has one MF owner and the last MF keeps its leftovers
  uncounted, so it reaches no GR or GO (reference quirk: connectMFGL never
  counts them); each list and its reverse hold the same pairs; the MF->GR,
  MF->GO and GO->GR lists are the glomerulus relays of the dendrite and axon
  lists; fan-ins stay within their limits with empty slots UINT_MAX; a GR's
  dendrites are distinct; every dendrite lies within its span of the GR's
  position scaled onto the glomerulus grid -- the Y position scaled by the X
  ratio, as the reference does;
* a checkpoint of such a state writes its parameter file's header and the
  glomerulus layer, and reads back to the same bytes.
On the GPU: a reference-wired network (GRs with fewer than 4 MF inputs and
0..4 GO inputs and targets) steps like the binary32 oracle.
"""
import numpy as np
import pytest

import orc as O
from cerebellum_simulation_amd import configs
from cerebellum_simulation_amd.network import NetState
from test_checkpoint import _offsets

U32MAX = np.uint32(0xFFFFFFFF)
FIELD_LISTS = {"pGRfromMFtoGR": ("NUMP_GR_FROM_MF_TO_GR", "P_GR_FROM_MF_TO_GR"),
               "pGRfromGOtoGR": ("NUMP_GR_FROM_GO_TO_GR", "P_GR_FROM_GO_TO_GR"),
               "pGRfromGRtoGO": ("NUMP_GR_FROM_GR_TO_GO", "P_GR_FROM_GR_TO_GO"),
               "pMFfromMFtoGO": ("NUMP_MF_FROM_MF_TO_GO", "P_MF_FROM_MF_TO_GO")}


def _con(name="g2k", **over):
    return {**configs.GLOMERULAR[name], **over}


def _state(name="g2k", seed=1, zones=1, **over):
    return NetState(configs.act_params_text(), configs.glomerular_conparams_text(name, **over), num_zones=zones,
                    innet_seed=seed, generator="reference")


def _lists(st):
    """{reference list name: (counts, slots [rows, cols])} of a state, from its
    fields and the glomerulus layer of its checkpoint."""
    data = st.checkpoint_bytes()
    secs, _ = _offsets(data)
    raw = {name: data[off:off + n] for name, z, off, n in secs if z == 0}
    out = {}
    for name in O.INNET_LISTS:
        if name in FIELD_LISTS:
            nf, pf = FIELD_LISTS[name]
            num, slots = st.field(nf).reshape(-1).copy(), st.field(pf).copy()
        else:
            num = np.frombuffer(raw["num" + name], np.int32).copy()
            slots = np.frombuffer(raw[name], np.uint32).reshape(num.size, -1).copy()
        out[name] = (num, slots.reshape(num.size, -1))
    out["haspGLfromMFtoGL"] = np.frombuffer(raw["haspGLfromMFtoGL"], np.uint8).copy()
    out["pGLfromMFtoGL"] = np.frombuffer(raw["pGLfromMFtoGL"], np.uint32).copy()
    return out


def _pairs(num, slots, reverse=False):
    """Sorted (row, entry) pairs of a list (empty slots skipped)."""
    rows = np.repeat(np.arange(slots.shape[0]), slots.shape[1])
    vals = slots.reshape(-1)
    keep = vals != U32MAX
    a, b = rows[keep].astype(np.int64), vals[keep].astype(np.int64)
    p = np.stack([b, a] if reverse else [a, b], 1)
    return p[np.lexsort((p[:, 1], p[:, 0]))]


@pytest.mark.parametrize("name,seed,over", [("g2k", 1, {}), ("g2k", 2, {}), ("g2k", 3, {"glYP2": 4}),
                                            ("g262k", 5, {})],
                         ids=["2k-seed1", "2k-seed2", "2k-unequal-ratios", "262k"])
def test_library_equals_oracle(name, seed, over):
    st = _state(name, seed, **over)
    mine = _lists(st)
    ref = O.innet_generate(_con(name, **over), seed)
    for key in O.INNET_LISTS:
        num, slots = mine[key]
        rnum, rslots = ref[key]
        assert np.array_equal(num, rnum), key
        # the library stores its GR lists as wide as the step needs; the file's are as wide as the reference's
        w = min(slots.shape[1], rslots.shape[1])
        assert np.array_equal(slots[:, :w], rslots[:, :w]), key
        assert np.all(slots[:, w:] == U32MAX) and np.all(rslots[:, w:] == U32MAX), key
    assert np.array_equal(mine["haspGLfromMFtoGL"], ref["haspGLfromMFtoGL"])
    assert np.array_equal(mine["pGLfromMFtoGL"], ref["pGLfromMFtoGL"])


@pytest.mark.parametrize("over", [{}, {"glYP2": 4}], ids=["equal-ratios", "unequal-ratios"])
def test_invariants(over):
    c = _con(**over)
    st = _state(seed=7, **over)
    L = _lists(st)
    glX, glY = 1 << c["glXP2"], 1 << c["glYP2"]
    grX, grY = 1 << c["grXP2"], 1 << c["grYP2"]
    numGL, numMF, numGR = glX * glY, 1 << c["numMFP2"], grX * grY
    mfGL = numGL // numMF
    last = numMF - 1
    own, has = L["pGLfromMFtoGL"], L["haspGLfromMFtoGL"]
    # one MF owner per glomerulus; MFs 0..numMF-2 own mfGL each and count them,
    # the last owns the rest in increasing order without counting them
    assert np.all(has == 1) and np.all(own < numMF)
    assert np.array_equal(np.bincount(own, minlength=numMF), np.full(numMF, mfGL))
    mnum, mgl = L["pMFfromMFtoGL"]
    assert np.all(mnum[:last] == mfGL) and mnum[last] == 0
    assert np.array_equal(mgl[last], np.nonzero(own == last)[0])
    for m in range(last):
        assert np.all(own[mgl[m]] == m)
    # the last MF reaches nothing
    assert L["pMFfromMFtoGR"][0][last] == 0 and L["pMFfromMFtoGO"][0][last] == 0
    assert not np.any(L["pGRfromMFtoGR"][1] == last)
    # lists and their reverses hold the same pairs
    for fwd, rev in (("pGRfromGLtoGR", "pGLfromGLtoGR"), ("pGOfromGLtoGO", "pGLfromGLtoGO"),
                     ("pGOfromGOtoGL", "pGLfromGOtoGL"), ("pGOfromGRtoGO", "pGRfromGRtoGO"),
                     ("pMFfromMFtoGR", "pGRfromMFtoGR"), ("pMFfromMFtoGO", "pGOfromMFtoGO"),
                     ("pGOfromGOtoGR", "pGRfromGOtoGR")):
        assert np.array_equal(_pairs(*L[fwd]), _pairs(*L[rev], reverse=True)), fwd
    # counts match the filled slots; dest lists are compact with UINT_MAX after the count
    for key in O.INNET_LISTS:
        num, slots = L[key]
        filled = (slots != U32MAX).sum(1)
        if key == "pMFfromMFtoGL":  # the last MF's row is filled but not counted
            filled[last] = 0
        assert np.array_equal(num, filled), key
    for key in ("pGLfromGLtoGR", "pGLfromGLtoGO", "pGLfromGOtoGL", "pGRfromGRtoGO", "pGRfromMFtoGR",
                "pGRfromGOtoGR", "pMFfromMFtoGR", "pMFfromMFtoGO", "pGOfromGOtoGR", "pGOfromMFtoGO"):
        num, slots = L[key]
        cols = np.arange(slots.shape[1])[None, :]
        assert np.all((slots != U32MAX) == (cols < num[:, None])), key
    assert L["pGLfromGLtoGR"][0].max() <= c["maxnumpGLfromGLtoGR"]
    assert L["pGRfromGRtoGO"][0].max() <= c["maxnumpGRfromGRtoGO"]
    assert L["pGLfromGOtoGL"][0].max() <= c["maxnumpGLfromGOtoGL"]
    # relays through the glomeruli
    gnum, ggl = L["pGRfromGLtoGR"]
    fnum, fgo = L["pGLfromGOtoGL"]
    for gr in range(0, numGR, 7):
        gls = ggl[gr][ggl[gr] != U32MAX]
        want_mf = sorted(int(own[g]) for g in gls if own[g] != last)
        got = L["pGRfromMFtoGR"][1][gr]
        assert sorted(int(v) for v in got if v != U32MAX) == want_mf, gr
        want_go = sorted(int(go) for g in gls for go in fgo[g][:fnum[g]])
        got = L["pGRfromGOtoGR"][1][gr]
        assert sorted(int(v) for v in got if v != U32MAX) == want_go, gr
        assert len(set(gls.tolist())) == len(gls)  # distinct dendrites
    # every dendrite within +-span/2 of the GR's position scaled by the X ratio
    scale = grX / glX
    half = c["spanGRDenOnGLX"] // 2
    gr_ids = np.repeat(np.arange(numGR), ggl.shape[1])
    gl_ids = ggl.reshape(-1)
    keep = gl_ids != U32MAX
    gr_ids, gl_ids = gr_ids[keep], gl_ids[keep].astype(np.int64)
    cx = np.floor(gr_ids % grX / scale + 0.5).astype(np.int64)
    cy = np.floor(gr_ids // grX / scale + 0.5).astype(np.int64)
    dx = (gl_ids % glX - cx + glX // 2) % glX - glX // 2
    dy = (gl_ids // glX - cy + glY // 2) % glY - glY // 2
    assert np.abs(dx).max() <= half and np.abs(dy).max() <= half
    if grY / glY != scale:
        # scaled by its own ratio the Y offsets would be far larger
        cy2 = np.floor(gr_ids // grX / (grY / glY) + 0.5).astype(np.int64)
        dy2 = (gl_ids // glX - cy2 + glY // 2) % glY - glY // 2
        assert np.abs(dy2).max() > half


def test_checkpoint_of_a_reference_wired_state():
    st = _state(seed=4, zones=2)
    data = st.checkpoint_bytes()
    from test_checkpoint import _header
    zones, con, mats, act, _ = _header(data)
    want = _con()
    for k, v in want.items():
        assert con[k] == v, k
    assert mats["GOGOLocalMatrix"] == (9, 9)
    secs, end = _offsets(data)
    assert end == len(data)
    back = NetState.from_checkpoint_bytes(data)
    assert back.checkpoint_bytes() == data


def test_full_size_sizes_and_quirks():
    """The 1M-GR network of conParams.txt: the fan-ins the reference's file
    gives, and the GRs on the last MF's glomeruli one MF input short."""
    st = _state("g1m", seed=3)
    s = st.sizes
    assert (s.numGR, s.numGO, s.numMF) == (1 << 20, 1024, 2048)
    assert (s.maxnumpGRfromMFtoGR, s.maxnumpGRfromGOtoGR, s.maxnumpGRfromGRtoGO, s.maxnumpMFfromMFtoGO) == (4, 4, 4, 32)
    nm = st.field("NUMP_GR_FROM_MF_TO_GR")
    counts = np.bincount(nm, minlength=5)
    # 32 glomeruli x about 64 GRs each lose the last MF
    assert counts[4] > 0.99 * s.numGR and 1000 < counts[3] < 4000 and counts[:3].sum() < 50
    assert st.field("NUMP_MF_FROM_MF_TO_GO")[-1] == 0


@pytest.mark.gpu
def test_reference_wired_network_steps_like_oracle(gpu):
    from cerebellum_simulation_amd import Network
    from test_gpu_parity import assert_parity, compare
    print("[case] reference-wired 2K-GR network, 400 steps vs binary32 oracle", flush=True)
    st = _state(seed=9, zones=1)
    nm = st.field("NUMP_GR_FROM_MF_TO_GR")
    assert (nm < 4).any() and (st.field("NUMP_GR_FROM_GO_TO_GR") < 4).any()
    net = Network.from_state(st)
    o32 = O.Oracle.from_state(st, precision=32)
    steps = 400
    tr = configs.mf_train(st.sizes.numMF, steps, seed=6, trial_ms=400, cs_on=100, cs_off=300)
    err = configs.us_err_drive(steps, 1, trial_ms=400, us_ms=299)
    import torch
    dtr = torch.from_numpy(tr).cuda()
    for k0 in range(0, steps, 100):
        net.run_mf_train(dtr[k0:k0 + 100], err_drive=err[k0:k0 + 100])
        o32.run_mf_train(tr[k0:k0 + 100], err_drive=err[k0:k0 + 100])
        bad = compare(net, o32, 1)
        print(f"step {k0 + 100}: {bad}", flush=True)
        assert_parity(bad, where=f"reference-wired step {k0 + 100}")
    assert (net.read("APBUF_GR") != 0).mean() > 0.01 and net.read("APBUF_GO").any()
    net.close()


@pytest.mark.gpu
def test_full_size_reference_network_runs(gpu):
    from cerebellum_simulation_amd import Network
    print("[case] reference-wired 1M-GR network of conParams.txt, 200 steps", flush=True)
    st = _state("g1m", seed=3, zones=2)
    net = Network.from_state(st)
    tr = configs.mf_train(st.sizes.numMF, 200, seed=2)
    import torch
    net.run_mf_train(torch.from_numpy(tr).cuda())
    net.sync()
    assert (net.read("APBUF_GR") != 0).mean() > 0.001 and net.read("APBUF_GO").any()
    net.close()

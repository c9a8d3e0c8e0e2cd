"""Pins of the CPU oracle, each against something the oracle's formulas do not
restate: closed forms of the GR recurrences, which pipeline row each
population reads (the two-phase lags of cbmsimcore.cpp:139-310), direct counts
over the spike registers, the reference generator's own draws for the NC->IO
release (tests/golden/sfmt_vectors.json), and the binary64 replay."""
import math

import numpy as np
import pytest

import orc
from cerebellum_simulation_amd import configs


def _quiet(st, prec=32):
    o = orc.Oracle.from_state(st, precision=prec)
    return o


@pytest.fixture
def t4k():
    return configs.build_state("t4k", seed=0)


def test_gr_closed_forms_without_input(t4k):
    """No input: v stays at eLeak exactly; the threshold relaxes as
    rest + (th0-rest)(1-d)^n; gKCa decays as g0 * 0.9999^n (kernels.cu:95-104)."""
    o = _quiet(t4k)
    p, s = t4k.params, t4k.sizes
    o.write("THRESH_GR", np.full(s.numGR, -30.0, np.float32))
    o.write("GKCA_GR", np.full(s.numGR, 0.5, np.float32))
    zero = np.zeros(s.numMF, np.uint8)
    n = 40
    for _ in range(n):
        o.update_mf_input(zero)
        o.calc_activity(1)
    assert (o.read("VM_GR") == np.float32(p.eLeakGR)).all()
    assert (o.read("APBUF_GR") == 0).all()
    th = o.read("THRESH_GR").astype(np.float64)
    d = float(np.float32(p.threshDecGR))
    exp_th = p.threshRestGR + (-30.0 - p.threshRestGR) * (1 - d) ** n
    assert np.allclose(th, exp_th, rtol=1e-5)
    assert np.allclose(o.read("GKCA_GR"), 0.5 * float(np.float32(0.9999)) ** n, rtol=1e-5)


def test_mf_conductance_geometric_series(t4k):
    """Every MF active every step: each GR counts its 4 inputs, so
    gE_n = 4 inc (1 - dec^n) / (1 - dec) from gE_0 = 0 (kernels.cu:198)."""
    o = _quiet(t4k)
    p, s = t4k.params, t4k.sizes
    ones = np.ones(s.numMF, np.uint8)
    n = 50
    for _ in range(n):
        o.update_mf_input(ones)
        o.calc_activity(1)
    dec, inc = float(np.float32(p.gDecMFtoGR)), float(np.float32(p.gIncMFtoGR))
    assert np.allclose(o.read("G_MF_SUM_GR"), 4 * inc * (1 - dec ** n) / (1 - dec), rtol=1e-5)


def test_spike_resets_threshold_and_raises_gkca(t4k):
    o = _quiet(t4k)
    s = t4k.sizes
    v = o.read("VM_GR")
    v[::3] = 0.0  # far above threshold
    o.write("VM_GR", v)
    o.write("GKCA_GR", np.full(s.numGR, 0.25, np.float32))
    o.calc_activity(1)
    ap = (o.read("APBUF_GR") & 1).astype(bool)
    assert ap[::3].all() and not ap[1::3].any()
    assert np.array_equal(o.read("AP_GR").astype(bool), ap)
    assert (o.read("THRESH_GR")[ap] == np.float32(t4k.params.threshMaxGR)).all()
    g = o.read("GKCA_GR").astype(np.float64)
    assert np.allclose(g[ap], 0.25 * 0.9999 + 0.003, rtol=1e-6)
    assert np.allclose(g[~ap], 0.25 * 0.9999, rtol=1e-6)


def test_plasticity_reads_history_bit_39(t4k):
    """Every 5th step, weights of GRs whose history bit 39 is set move by the
    IO's LTD step (timer inside the LTD window), clamped to [0, 1]; bit 40 --
    the bin the host computes -- is not the one the kernel tests
    (kernels.cu:287, activityparams.cpp:599)."""
    o = _quiet(t4k)
    s, p = t4k.sizes, t4k.params
    hist = np.zeros(s.numGR, np.uint64)
    hist[0::4] = np.uint64(1) << np.uint64(39)
    hist[1::4] = np.uint64(1) << np.uint64(40)
    o.write("HIST_GR", hist)
    w = np.full(s.numGR, 0.5, np.float32)
    w[8] = 0.001  # clamps at 0
    o.write("PFPC_SYNW", w)
    o.write("PFPC_PLAST_TIMER_IO", np.full(s.numIO, -150, np.int32))
    o.calc_activity(4)  # t = 1..4: no plasticity
    assert np.array_equal(o.read("PFPC_SYNW").ravel(), w)
    o.write("PFPC_PLAST_TIMER_IO", np.full(s.numIO, -150, np.int32))
    o.calc_activity(1)  # t = 5
    w5 = o.read("PFPC_SYNW").ravel()
    ltd = np.float32(p.synLTDStepSizeGRtoPC)
    assert w5[0] == np.float32(0.5) + ltd and w5[8] == 0.0
    assert (w5[1::4] == 0.5).all() and (w5[2::4] == 0.5).all()
    # history shifted once, with bit 0 = a spike in the last 5 steps
    h5 = o.read("HIST_GR")
    assert h5[0] == np.uint64(1) << np.uint64(40) and h5[1] == np.uint64(1) << np.uint64(41)


def test_ltp_outside_window_and_upper_clamp(t4k):
    o = _quiet(t4k)
    s = t4k.sizes
    o.write("HIST_GR", np.full(s.numGR, np.uint64(1) << np.uint64(39), np.uint64))
    w = np.full(s.numGR, 0.9999, np.float32)
    o.write("PFPC_SYNW", w)
    o.write("PFPC_PLAST_TIMER_IO", np.full(s.numIO, 50, np.int32))  # > tsLTPStartAPIO
    o.write("CUR_TIME", np.array([4], np.uint64))
    o.calc_activity(1)
    assert (o.read("PFPC_SYNW") == 1.0).all()


def test_population_inputs_lag_two_steps(t4k):
    """PC, BC and SC read the PF sum of two steps back (pipeline row 1); the
    row written by this step is row 0 (cbmsimcore.cpp:146-236)."""
    o = _quiet(t4k)
    s, p = t4k.sizes, t4k.params
    A = np.full((2, s.numBC), 0, np.uint32)
    A[0], A[1] = 7, 40
    o.write("PIPE_PFBC_SUM", A)
    S = np.zeros((2, s.numSC), np.uint32)
    S[0], S[1] = 3, 20
    o.write("PIPE_PFSC_SUM", S)
    P = np.zeros((2, s.numPC), np.float32)
    P[0], P[1] = 10.0, 900.0
    o.write("PIPE_PFPC_SUM", P)
    o.calc_activity(1)
    gbc = o.read("G_PF_BC").astype(np.float64)
    dec, inc = float(np.float32(p.gDecGRtoBC)), float(np.float32(p.gIncGRtoBC))
    assert np.allclose(gbc, 40 * inc * dec, rtol=1e-6)
    gsc = o.read("G_PF_SC").astype(np.float64)
    assert np.allclose(gsc, 20 * float(np.float32(p.gIncGRtoSC)) * float(np.float32(p.gDecGRtoSC)), rtol=1e-6)
    gpc = o.read("G_PF_PC").astype(np.float64)
    assert np.allclose(gpc, 900 * float(np.float32(p.gIncGRtoPC)) * float(np.float32(p.gDecGRtoPC)), rtol=1e-6)
    # the pipeline moved: old row 0 is now row 1
    assert (o.read("PIPE_PFBC_SUM")[1] == 7).all() and (o.read("PIPE_PFSC_SUM")[1] == 3).all()
    assert (o.read("PIPE_PFPC_SUM")[1] == 10.0).all()
    # MZone::writeToState / InNet::writeToState hold the row the next step reads
    assert (o.read("INPUT_SUM_PF_PC") == 10.0).all() and (o.read("INPUT_SUM_PF_SC") == 3).all()


@pytest.mark.parametrize("lag", [1, 2])
def test_golgi_grgo_lag(t4k, lag):
    """The GO update reads the GR->GO count of one step back (lag 1, "fresh")
    or two (lag 2, "stale") -- the two outcomes of the reference's racy read
    (innet.cpp:835-838 vs 1236-1237)."""
    o = orc.Oracle.from_state(t4k, precision=32, grgo_lag=lag)
    s = t4k.sizes
    G = np.zeros((2, s.numGO), np.uint32)
    G[0], G[1] = 11, 29
    o.write("PIPE_GRGO_SUM", G)
    o.calc_activity(1)
    assert (o.read("SUM_GR_INPUT_GO") == (11 if lag == 1 else 29)).all()


def _direct_counts(o, st):
    s = st.sizes
    ab = o.read("APBUF_GR")
    d = (ab & st.field("P_GR_DELAYMASK_GR_TO_BSP")) != 0
    hit = (ab[:, None] & st.field("P_GR_DELAYMASK_GR_TO_GO")) != 0
    go = np.bincount(st.field("P_GR_FROM_GR_TO_GO")[hit], minlength=s.numGO)
    w = o.read("PFPC_SYNW").reshape(s.numPC, -1).astype(np.float64)
    return (d.reshape(s.numBC, -1).sum(1), d.reshape(s.numSC, -1).sum(1), go,
            (w * d.reshape(s.numPC, -1)).sum(1), d, hit)


def test_output_sums_match_direct_counts():
    st = configs.build_state("c0", seed=0)
    o = orc.Oracle.from_state(st, precision=32)
    v = o.read("VM_GR")
    v[::11] = -30.0
    o.write("VM_GR", v)
    o.run_mf_train(configs.mf_train(st.sizes.numMF, 30, seed=3))
    bc, sc, go, pc, d, hit = _direct_counts(o, st)
    assert np.array_equal(o.read("PIPE_PFBC_SUM")[0], bc)
    assert np.array_equal(o.read("PIPE_PFSC_SUM")[0], sc)
    assert np.array_equal(o.read("PIPE_GRGO_SUM")[0], go)
    # conservation: every delivered GR->GO spike lands in exactly one GO bin
    assert go.sum() == hit.sum() > 0
    assert bc.sum() == sc.sum() == d.sum()
    assert np.allclose(o.read("PIPE_PFPC_SUM")[0], pc, rtol=1e-6)


def test_canonical_sum():
    rng = np.random.default_rng(0)
    ints = rng.integers(0, 100, 32768).astype(np.float32)
    assert orc.canonical_sum_f32(ints) == float(ints.astype(np.int64).sum())  # exact below 2^24
    x = rng.random(32768).astype(np.float32)
    assert abs(orc.canonical_sum_f32(x) - math.fsum(x.astype(np.float64))) <= 1e-6 * x.sum()


def test_release_draws_follow_reference_stream(golden_sfmt):
    """NC->IO inputs are Random() < p in (IO, NC) order from the zone's
    generator, after numGR Random() draws (mzone.cpp:505-511, 846-849)."""
    st = configs.build_state("c0", seed=0, mzone_rng_seeds=[12345] * 8)
    o = orc.Oracle.from_state(st, precision=32)
    s = st.sizes
    ref = [float.fromhex(x) for x in golden_sfmt["seeds"]["12345"]["random_after_16384_random"]]
    k = 0
    for step in range(20):
        o.write("SYN_IO_PRELEASE_NC", np.linspace(0.1, 0.9, s.numNC).astype(np.float32))
        o.calc_activity(1)
        rel = o.read("SYN_IO_PRELEASE_NC").astype(np.float64)  # after this step's update
        got = o.read("INPUT_NC_IO").reshape(s.numIO, s.numNC)
        for i in range(s.numIO):
            for j in range(s.numNC):
                assert got[i, j] == (ref[k] < np.float32(rel[j])), (step, i, j)
                k += 1


def test_mf_history_or_and_reset(t4k):
    o = _quiet(t4k)
    s = t4k.sizes
    a = np.zeros(s.numMF, np.uint8)
    a[3] = 1
    b = np.zeros(s.numMF, np.uint8)
    b[5] = 2  # any nonzero byte is a spike
    o.update_mf_input(a)
    o.calc_activity(1)
    o.update_mf_input(b)
    o.calc_activity(1)
    h = o.read("HIST_MF")
    assert h[3] == 1 and h[5] == 1 and h.sum() == 2
    assert o.read("APBUF_MF")[3] == 2 and o.read("APBUF_MF")[5] == 1
    o.update_mf_input(np.zeros(s.numMF, np.uint8))
    o.calc_activity(3)  # t = 5: history cleared after use
    assert o.read("HIST_MF").sum() == 0


def test_binary32_mimic_tracks_binary64_replay():
    st = configs.build_state("c0", seed=0)
    o32 = orc.Oracle.from_state(st, precision=32)
    o64 = orc.Oracle.from_state(st, precision=64)
    tr = configs.mf_train(st.sizes.numMF, 100, seed=9)
    o32.run_mf_train(tr)
    o64.run_mf_train(tr)
    v32, v64 = o32.read("VM_GR").astype(np.float64), o64.read_f64("VM_GR")
    spk = (o32.read("APBUF_GR") != o64.read("APBUF_GR")).mean()
    assert spk < 1e-3
    same = o32.read("APBUF_GR") == o64.read("APBUF_GR")
    assert np.allclose(v32[same], v64[same], rtol=1e-4, atol=1e-4)


def test_eyelid_integrator_closed_forms():
    """EyelidIntegrator (eyelidintegrator.cpp:13-68): no input keeps the lid
    open; a constant spike train drives gIn to gInc/(1-gDecay) per cell and the
    lid to gSum*evMax/(gLeak+gSum); binary32 stays within 1e-5 of a binary64
    run of the same recurrence."""
    import numpy as np
    from cerebellum_simulation_amd.eyelid import EyelidIntegrator, cr_onset
    e = EyelidIntegrator(8)
    assert all(e.step(np.zeros(8, np.uint8)) == 0 for _ in range(50))
    e = EyelidIntegrator(8)
    ones = np.ones(8, np.uint8)
    for _ in range(3000):
        v = e.step(ones)
    gd, gi, gs, gl, em = (float(x) for x in (e.gDecay, e.gInc, e.gShift, e.gLeak, e.evMax))
    g_ss = gi / (1 - gd)
    assert abs(float(e.gIn[0]) - g_ss) <= 1e-5 * g_ss
    gsum = 8 * g_ss + gs
    v_ss = gsum * em / (gl + gsum)
    assert abs(float(v) - v_ss) <= 1e-4 * v_ss
    # binary64 run of the same recurrence on a random raster
    rng = np.random.default_rng(3)
    raster = (rng.random((400, 8)) < 0.4).astype(np.uint8)
    e = EyelidIntegrator(8)
    tr32 = e.trace(raster).astype(np.float64)
    g = np.zeros(8)
    vm = 0.0
    tr64 = []
    for row in raster:
        g = g * gd + row * gi
        s = max(g.sum() + gs, 0.0)
        vm = vm + gl * (0 - vm) + s * (em - vm)
        tr64.append(vm)
    assert np.max(np.abs(tr32 - np.array(tr64))) <= 1e-3
    # onset: first CS-window ms above baseline + 5
    tr = np.zeros(2000, np.float32)
    tr[640:900] = 20.0
    assert cr_onset(tr, 500, 1000) == 640
    assert cr_onset(np.zeros(2000, np.float32), 500, 1000) is None

"""The learned input-layer surrogate on the GPU, through the C-ABI
(include/cbm_surrogate.h), against the CPU restatement in
oracle/surrogate_ref.py, and the surrogate-driven sweep against the exact step.

* Forward pass with integer-valued weights (every activation an integer below
  256, so BF16 and FP32 are exact): layer-1 and layer-2 activations bit-exact,
  probabilities within 2e-7 of the binary64 logistic.
* Loss and every parameter gradient against the oracle (binary64 sums on the
  same BF16 operands): loss within 1e-5 relative, each gradient array within
  5e-3 relative (L2) -- the device sums in FP32 and rounds its BF16
  activations and gradients where its sums land, which moves a few of them by
  one BF16 step.  Real sizes (1,024 MFs, 128 channels, 4,096 clusters) and a
  small one; batches whose row count is not a multiple of the 128-row tile.
* One Adam step equals the closed form of a first Adam step on the device's
  own gradient.
* The PF recorder and the training targets: per-64-GR counts equal a count
  from the spike registers and PF delay masks; targets equal their sums.
* The sweep engine driven by the exact network's own cluster PF fractions,
  with PF->PC weights held fixed, reproduces the exact network's PC, BC, SC,
  IO and NC spikes bit for bit (every cluster sum is then exact), and its
  cluster-weight rule matches its closed form step by step.
* Training on data from the simulator predicts held-out PF rates better than
  each cluster's mean rate and than its mean over the training trials at the
  same trial time (the cross-entropy itself is dominated by the Bernoulli
  noise of 256 cells and falls only a few per cent).
"""
import numpy as np
import pytest

import surrogate_ref as R
from cerebellum_simulation_amd import Batch, Network, configs
from cerebellum_simulation_amd import surrogate as SG

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    return torch


def _stream(T, num_in, rate, seed):
    rng = np.random.default_rng(seed)
    return (rng.random((T, num_in)) < rate).astype(np.uint8)


def _make(dims, max_rows=4096):
    return SG.Surrogate(num_in=dims.num_in, hidden=dims.hidden, num_out=dims.num_out, taps=dims.taps,
                        dil=dims.dil, max_rows=max_rows)


def test_forward_integer_weights_exact(gpu):
    torch = _torch()
    print("[case] surrogate forward, integer weights, 256-128-256", flush=True)
    dims = R.Dims(256, 128, 256)
    sg = _make(dims)
    rng = np.random.default_rng(0)
    arrays = []
    for l, dens in zip(range(3), (0.02, 0.04, 0.04)):
        W = rng.integers(-1, 2, dims.shapes()[2 * l]) * (rng.random(dims.shapes()[2 * l]) < dens)
        b = rng.integers(-1, 3, dims.shapes()[2 * l + 1])
        arrays += [W, b]
    flat = dims.join(arrays).astype(np.float32)
    sg.set_params(flat)
    stream = _stream(500, dims.num_in, 0.08, 1)
    starts = np.array([0, 100, 57], np.int64)
    L = 37
    P = dims.lead + L
    H1, H2, Z = R.forward(dims, flat.astype(np.float64), R.segments(stream, starts, P))
    for a in (H1, H2, Z):  # integers within BF16's exact range: no rounding anywhere
        assert np.all(a == np.round(a)) and np.abs(a).max() <= 256
    assert H1.max() >= 2 and H2.max() >= 2 and np.abs(Z).max() >= 2
    mf = torch.from_numpy(stream).cuda()
    st = torch.from_numpy(starts).cuda()
    prob = sg.forward(mf, st, L).cpu().numpy()
    R_ = len(starts) * P
    h1 = sg.read_hidden(1, R_)
    h2 = sg.read_hidden(2, R_)
    assert np.array_equal(h1, R.bf16_bits(H1.reshape(R_, -1)))
    assert np.array_equal(h2, R.bf16_bits(H2.reshape(R_, -1)))
    want = R.sigmoid(Z[:, dims.lead:, :]).reshape(-1, dims.num_out)
    assert np.abs(prob - want).max() <= 2e-7
    sg.close()


@pytest.mark.parametrize("dims,segs,L", [(R.Dims(256, 128, 256), 5, 30), (R.Dims(1024, 128, 4096), 2, 30)],
                         ids=["small", "full"])
def test_loss_and_gradients_match_oracle(gpu, dims, segs, L):
    torch = _torch()
    print(f"[case] surrogate gradients {dims.num_in}-{dims.hidden}-{dims.num_out}, {segs} x {L}", flush=True)
    sg = _make(dims)
    sg.init_params(seed=3, out_bias=-2.5)
    flat = sg.get_params().astype(np.float64)
    T = 400
    stream = _stream(T, dims.num_in, 0.06, 2)
    rng = np.random.default_rng(4)
    ybits = R.bf16_bits(rng.integers(0, 41, (T, dims.num_out)) / 256.0)
    P = dims.lead + L
    starts = np.sort(rng.integers(0, T - P, segs)).astype(np.int64)
    mf = torch.from_numpy(stream).cuda()
    st = torch.from_numpy(starts).cuda()
    y = torch.from_numpy(ybits.view(np.int16)).cuda()
    loss, grad = sg.loss_grad(mf, st, L, y)
    X = R.segments(stream, starts, P)
    Y = R.segments(R.bits_to_f64(ybits), starts, P)
    lo, go = R.loss_grad(dims, flat, X, Y)
    print(f"loss gpu {loss:.7f} oracle {lo:.7f}", flush=True)
    assert abs(loss - lo) <= 1e-5 * abs(lo)
    for name, a, b in zip(("W1", "b1", "W2", "b2", "W3", "b3"), dims.split(grad), dims.split(go)):
        rel = float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))
        print(f"{name}: rel {rel:.2e} |g| {np.linalg.norm(b):.3e}", flush=True)
        assert rel <= 5e-3, name
    sg.close()


def test_adam_step_closed_form(gpu):
    torch = _torch()
    print("[case] surrogate Adam first step", flush=True)
    dims = R.Dims(256, 128, 256)
    sg = _make(dims)
    sg.init_params(seed=5, out_bias=-2.0)
    th0 = sg.get_params()
    T, L = 300, 40
    stream = _stream(T, dims.num_in, 0.06, 6)
    rng = np.random.default_rng(7)
    ybits = R.bf16_bits(rng.integers(0, 30, (T, dims.num_out)) / 256.0)
    starts = np.array([3, 150, 200], np.int64)
    mf, st = torch.from_numpy(stream).cuda(), torch.from_numpy(starts).cuda()
    y = torch.from_numpy(ybits.view(np.int16)).cuda()
    _, g = sg.loss_grad(mf, st, L, y)
    lr, b1, b2, eps = np.float32(1e-3), np.float32(0.9), np.float32(0.999), np.float32(1e-8)
    sg.train_step(mf, st, L, y, float(lr), float(b1), float(b2), float(eps))
    th1 = sg.get_params()
    m = (np.float32(1) - b1) * g
    v = (np.float32(1) - b2) * g * g
    bc1, bc2 = np.float32(1) - b1, np.float32(1) - b2
    want = th0 - lr * (m / bc1) / (np.sqrt(v / bc2) + eps)
    assert np.abs(th1 - want).max() <= 1e-6 * max(1.0, np.abs(th0).max())
    assert not np.array_equal(th1, th0)
    sg.close()


def test_pf_recorder_and_targets(gpu):
    torch = _torch()
    print("[case] PF recorder and targets, 65K GR", flush=True)
    st = configs.build_state("b65k", seed=1)
    s = st.sizes
    net = Network.from_state(st)
    net.set_recorder(["PF", "NC"], 64)
    train = configs.mf_train(s.numMF, 212, seed=2)
    # granule cells start firing some 30-60 steps into a run from the initial state
    net.run_mf_train(torch.from_numpy(train[:200]).cuda())
    dly = st.field("P_GR_DELAYMASK_GR_TO_BSP").astype(np.uint64)
    counts = []
    for k in range(200, 212):
        net.update_mf_input(train[k])
        net.calc_activity(1)
        ap = net.read("APBUF_GR").astype(np.uint64)
        pf = ((ap & dly) != 0).reshape(-1, 64).sum(axis=1)
        counts.append(pf)
    t0 = int(net.read("CUR_TIME")[0]) - 11
    rec = net.read_recorded("PF", t0, 12)
    assert np.array_equal(rec.astype(np.int64), np.array(counts))
    assert rec.sum() > 0
    dev = torch.empty((12, s.numGR // 64), dtype=torch.uint8, device="cuda")
    net.read_recorded("PF", t0, 12, out=dev)
    assert np.array_equal(dev.cpu().numpy(), rec)
    sg = SG.Surrogate(num_in=128, hidden=128, num_out=128, max_rows=256)
    y = torch.zeros((12, s.numGR // 256), dtype=torch.int16, device="cuda")
    sg.targets_from_pf(dev, 4, y)
    sg.sync()
    want = R.bf16_bits(rec.reshape(12, -1, 4).sum(axis=2) / 256.0)
    assert np.array_equal(y.cpu().numpy().view(np.uint16), want)
    sg.close()


def _sweep_pair(name, seed, steps, mf_seed):
    """An exact network run for `steps` with PF recording, and the cluster PF
    fractions of its steps as FP32 rows behind 199 rows of zeros."""
    torch = _torch()
    st = configs.build_state(name, seed=seed)
    s = st.sizes
    exact = Network.from_state(st)
    for z in range(st.num_zones):
        exact.set_grpc_plast_steps(z, 0.0, 0.0)
    exact.set_recorder(["PF", "PC", "BC", "SC", "IO", "NC"], steps)
    train = torch.from_numpy(configs.mf_train(s.numMF, steps, seed=mf_seed)).cuda()
    err = configs.us_err_drive(steps, st.num_zones, us_ms=350)
    exact.run_mf_train(train, err_drive=err)
    rec = {p: exact.read_recorded(p, 1, steps) for p in ("PF", "PC", "BC", "SC", "IO", "NC")}
    frac = rec["PF"].reshape(steps, -1, 4).sum(axis=2).astype(np.float32) / np.float32(256)
    prob = torch.zeros((199 + steps, frac.shape[1]), dtype=torch.float32, device="cuda")
    prob[199:] = torch.from_numpy(frac).cuda()
    return st, train, err, rec, prob


def test_sweep_with_exact_inputs_reproduces_microzone(gpu):
    print("[case] sweep driven by exact cluster PF fractions, 262K GR, fixed weights", flush=True)
    steps = 600
    st, train, err, rec, prob = _sweep_pair("b262k", 3, steps, 9)
    members = []
    for _ in range(2):
        n = Network.from_state(st)
        for z in range(st.num_zones):
            n.set_grpc_plast_steps(z, 0.0, 0.0)
        n.set_recorder(["PC", "BC", "SC", "IO", "NC"], steps)
        members.append(n)
    b = Batch(members)
    b.sweep(prob, 199, steps, train, err_drive=[err, err], cells_per_cluster=256)
    b.sync()
    for i, n in enumerate(members):
        for p in ("PC", "BC", "SC", "IO", "NC"):
            got = n.read_recorded(p, 1, steps)
            assert np.array_equal(got, rec[p]), (i, p, int(got.sum()), int(rec[p].sum()))
        w = b.sweep_weights(i)
        assert np.all(w == np.float32(0.5))
    assert rec["NC"].sum() > 0 and rec["PC"].sum() > 0 and rec["IO"].sum() > 0
    b.close()


def test_sweep_weight_rule_closed_form(gpu):
    torch = _torch()
    print("[case] sweep cluster-weight rule, 65K GR", flush=True)
    st = configs.build_state("b65k", seed=4)
    s = st.sizes
    net = Network.from_state(st)
    ltd, ltp = np.float32(-0.03), np.float32(0.02)
    net.set_grpc_plast_steps(0, float(ltd), float(ltp))
    b = Batch([net])
    n_cl = s.numGR // 256
    rng = np.random.default_rng(5)
    steps = 40
    p = rng.random((199 + steps, n_cl)).astype(np.float32) * np.float32(0.05)
    prob = torch.from_numpy(p).cuda()
    train = torch.from_numpy(configs.mf_train(s.numMF, steps, seed=6)).cuda()
    err = configs.us_err_drive(steps, 1, us_ms=10)
    prm = st.params
    ltd_end = int(prm.tsLTDStartAPIO + prm.tsLTDDurationIO)
    ltp_start = int(prm.tsLTPStartAPIO)
    io_of = (np.arange(n_cl) * 256) >> int(np.log2(s.numGR // s.numIO))
    W = None
    for k in range(steps):
        t = int(net.read("CUR_TIME")[0]) + 1
        timers = net.read("PFPC_PLAST_TIMER_IO").astype(np.int64)
        b.sweep(prob, 199 + k, 1, train[k:k + 1], err_drive=[err[k:k + 1]], cells_per_cluster=256)
        Wn = b.sweep_weights(0)[0]
        if W is None:
            W = np.full(n_cl, np.float32(0.5))
        if t % prm.tsPerHistBinGR == 0:
            keep = np.ones(n_cl, np.float32)
            for j in range(195, 200):
                keep = keep * (np.float32(1) - p[199 + k - j])
            q = np.float32(1) - keep
            tm = timers[io_of]
            step = np.where(tm < ltd_end, ltd, np.where(tm > ltp_start, ltp, np.float32(0))).astype(np.float32)
            w = W + q * step
            w = (w > 0) * w
            w = np.float32(1) * (w > 1) + (w <= 1) * w
            W = w.astype(np.float32)
        assert np.array_equal(Wn, W), k
    assert not np.all(W == np.float32(0.5))
    b.close()


def test_training_beats_cluster_and_phase_means(gpu):
    torch = _torch()
    print("[case] surrogate training on 262K-GR simulator data", flush=True)
    st = configs.build_state("b262k", seed=0)
    s = st.sizes
    net = Network.from_state(st)
    rates, cs = configs.mf_rates(s.numMF, seed=11)
    sg = SG.Surrogate(num_in=s.numMF, hidden=128, num_out=s.numGR // 256, max_rows=16384)
    trials, ntr = 14, 12
    data = SG.InputLayerData.record(net, sg, trials, rates, cs, seed=1)
    y = SG.bf16_bits_to_float(data.y)
    train_rows = (0, ntr * 2000)
    mean = y[:train_rows[1] - 1].mean(dim=0)
    phase = y[:train_rows[1]].view(ntr, 2000, -1).mean(dim=0)  # mean over training trials per trial time
    m = float(mean.mean())
    sg.init_params(seed=0, out_bias=float(np.log(m / (1 - m))))
    losses = SG.train(sg, data, 300, segments=32, outputs=256, lr=2e-3, lr_end=2e-4, rows=train_rows)
    print(f"training loss first {losses[:10].mean():.5f} last {losses[-10:].mean():.5f}", flush=True)
    assert losses[-10:].mean() < losses[:10].mean()
    lo, n = ntr * 2000 + 100, 2 * 2000 - 200
    tgt = data.y[lo:lo + n]
    rmse = SG.pf_rmse(sg.predict(data.mf, lo, n), tgt)
    base = SG.pf_rmse(mean.unsqueeze(0).expand(n, -1).contiguous(), tgt)
    ph = SG.pf_rmse(phase[torch.arange(lo, lo + n, device="cuda") % 2000].contiguous(), tgt)
    print(f"held-out PF rmse (spikes per cluster per ms): surrogate {rmse:.4f}, cluster mean {base:.4f}, "
          f"trial-phase mean {ph:.4f}", flush=True)
    assert rmse < 0.95 * base and rmse < ph
    sg.close()


def test_sweep_experiment_small(gpu):
    print("[case] surrogate-driven sweep judged by the exact step, 262K GR, 4 parameter sets", flush=True)
    # 64-cell clusters: an SC block of this network holds 512 GRs, about 59 of
    # them with a CS input, so 256-cell clusters would mix them with the others
    res = SG.sweep_experiment("b262k", train_trials=10, train_steps=250, segments=32, sweep_trials=2,
                              ltd_scales=(1, 8), ltp_scales=(1, 8), judged=(0, 3), cells=64,
                              log=lambda m: print(m, flush=True))
    print({k: v for k, v in res.items() if k not in ("exact_cr", "surrogate_cr")}, flush=True)
    print("exact CR", res["exact_cr"], "surrogate CR", res["surrogate_cr"], flush=True)
    assert res["pf_rmse"] < res["pf_rmse_cluster_mean"]
    # the pre-trained network closes the lid on every judged trial in the exact step
    assert all(o is not None for e in res["exact_cr"] for o, _ in e)
    # the surrogate-driven sweep does too, within 40 ms of the exact onset; the
    # sweep driven by each cluster's mean rate (no trained network) never does
    assert res["surrogate_vs_exact"]["onset_presence_mismatch"] == 0
    assert res["surrogate_vs_exact"]["onset_abs_err_ms"] <= 40
    assert res["cluster_mean_vs_exact"]["onset_presence_mismatch"] == res["cluster_mean_vs_exact"]["trials"]

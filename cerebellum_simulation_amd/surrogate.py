"""Learned surrogate of the granule-Golgi input layer, its training data, its
training loop, and the fast plasticity sweeps it drives.

The network and its kernels are defined in include/cbm_surrogate.h (HIP
library, BF16 MFMA kernels with FP32 accumulation); this module is the host
side: device memory comes from torch tensors, every computation runs in the
library.  No reference counterpart: the reference simulates every GR, and the
exact step (Network) stays the judge of the surrogate's PF rates and of the
conditioned responses of the sweeps it drives.

Data alignment.  MF row r of a stream is the MF input of one step; the
surrogate's output for the window that ends at row r predicts the PF spikes
of the step after it (a GR's spike at step t depends on MF input up to step
t - 1).  Targets: y[r][c] = fraction of cluster c's cells whose delayed PF
spike (the one PCs, BCs and SCs see) fired at the step after row r.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import _abi
from ._abi import check


def _torch():
    import torch
    return torch


class Surrogate:
    """A causal temporal-convolution network (numIn -> hidden -> hidden ->
    numOut channels, `taps` taps, dilations `dil`) on one GPU."""

    def __init__(self, num_in: int = 1024, hidden: int = 128, num_out: int = 4096, taps: int = 8,
                 dil=(1, 2, 4), max_rows: int = 1 << 15, device: int = 0):
        L = _abi.hip_lib()
        d = _abi.SgDesc()
        d.numIn, d.hidden, d.numOut, d.taps = num_in, hidden, num_out, taps
        for i in range(3):
            d.dil[i] = dil[i]
        d.maxRows, d.device = max_rows, device
        h = C.c_void_p()
        check(L.cbm_sg_create(C.byref(d), C.byref(h)), "cbm_sg_create")
        self._s = h
        self.num_in, self.hidden, self.num_out, self.taps, self.dil = num_in, hidden, num_out, taps, tuple(dil)
        self.max_rows = max_rows
        self.lead = L.cbm_sg_lead(h)
        n = C.c_size_t()
        check(L.cbm_sg_num_params(h, C.byref(n)), "cbm_sg_num_params")
        self.num_params = n.value

    # ---------------------------------------------------------------- parameters
    def init_params(self, seed: int = 0, out_bias: float = -3.0) -> None:
        check(_abi.hip_lib().cbm_sg_init_params(self._s, seed, out_bias), "cbm_sg_init_params")

    def get_params(self) -> np.ndarray:
        out = np.empty(self.num_params, np.float32)
        check(_abi.hip_lib().cbm_sg_get_params(self._s, out.ctypes.data), "cbm_sg_get_params")
        return out

    def set_params(self, flat) -> None:
        a = np.ascontiguousarray(flat, dtype=np.float32)
        if a.size != self.num_params:
            raise ValueError(f"expected {self.num_params} parameters")
        check(_abi.hip_lib().cbm_sg_set_params(self._s, a.ctypes.data), "cbm_sg_set_params")

    # ---------------------------------------------------------------- batches
    def _batch(self, mf, starts, outputs: int, target=None) -> _abi.SgBatch:
        torch = _torch()
        if not (mf.is_cuda and mf.dtype == torch.uint8 and mf.dim() == 2 and mf.stride(1) == 1):
            raise TypeError("mf must be a uint8 [rows, numIn] device tensor with contiguous rows")
        if not (starts.is_cuda and starts.dtype == torch.int64 and starts.is_contiguous()):
            raise TypeError("starts must be a contiguous int64 device tensor")
        b = _abi.SgBatch()
        b.mf, b.mfPitch = mf.data_ptr(), mf.stride(0)
        b.segStart, b.numSeg, b.outputs = starts.data_ptr(), starts.numel(), outputs
        if target is not None:
            if not (target.is_cuda and target.element_size() == 2 and target.dim() == 2 and target.stride(1) == 1):
                raise TypeError("target must be a 2-byte [rows, numOut] device tensor (BF16 bits)")
            b.target, b.targetPitch = target.data_ptr(), target.stride(0)
        # work queued by torch on its stream (inputs) must be done first
        torch.cuda.current_stream().synchronize()
        return b

    def forward(self, mf, starts, outputs: int, out=None):
        """Probabilities [numSeg * outputs, numOut] (FP32 device tensor)."""
        torch = _torch()
        if out is None:
            out = torch.empty((starts.numel() * outputs, self.num_out), dtype=torch.float32, device=mf.device)
        b = self._batch(mf, starts, outputs)
        check(_abi.hip_lib().cbm_sg_forward(self._s, C.byref(b), out.data_ptr(), out.stride(0)), "cbm_sg_forward")
        self.sync()
        return out

    def loss_grad(self, mf, starts, outputs: int, target):
        b = self._batch(mf, starts, outputs, target)
        loss = C.c_float()
        grad = np.empty(self.num_params, np.float32)
        check(_abi.hip_lib().cbm_sg_loss_grad(self._s, C.byref(b), C.byref(loss), grad.ctypes.data),
              "cbm_sg_loss_grad")
        return float(loss.value), grad

    def train_step(self, mf, starts, outputs: int, target, lr: float, beta1: float = 0.9, beta2: float = 0.999,
                   eps: float = 1e-8, loss_out=None, sync_inputs: bool = True) -> None:
        """One Adam step (asynchronous on the surrogate's stream)."""
        if sync_inputs:
            b = self._batch(mf, starts, outputs, target)
        else:
            b = _abi.SgBatch(mf.data_ptr(), mf.stride(0), starts.data_ptr(), starts.numel(), outputs,
                             target.data_ptr(), target.stride(0))
        o = _abi.SgAdam(lr, beta1, beta2, eps)
        check(_abi.hip_lib().cbm_sg_train_step(self._s, C.byref(b), C.byref(o),
                                                None if loss_out is None else loss_out.data_ptr()),
              "cbm_sg_train_step")

    def read_hidden(self, layer: int, rows: int) -> np.ndarray:
        out = np.empty((rows, self.hidden), np.uint16)
        check(_abi.hip_lib().cbm_sg_read_hidden(self._s, layer, out.ctypes.data, rows), "cbm_sg_read_hidden")
        return out

    def targets_from_pf(self, pf, chunks_per_cluster: int, out) -> None:
        """out[r][k] = BF16 fraction of cluster k's cells that fired, from
        recorded PF counts per 64-GR chunk (uint8 [rows, chunks] device)."""
        n, chunks = pf.shape
        check(_abi.hip_lib().cbm_sg_targets_from_pf(self._s, pf.data_ptr(), pf.stride(0), n, chunks,
                                                    chunks_per_cluster, out.data_ptr(), out.stride(0)),
              "cbm_sg_targets_from_pf")

    def last_timing(self) -> tuple[float, float]:
        us, fl = C.c_double(), C.c_double()
        check(_abi.hip_lib().cbm_sg_last_timing(self._s, C.byref(us), C.byref(fl)), "cbm_sg_last_timing")
        return us.value, fl.value

    def sync(self) -> None:
        check(_abi.hip_lib().cbm_sg_sync(self._s), "cbm_sg_sync")

    def close(self) -> None:
        if getattr(self, "_s", None) is not None and self._s.value:
            _abi.hip_lib().cbm_sg_destroy(self._s)
            self._s = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---------------------------------------------------------------- streams
    def predict(self, mf, first: int, n: int, outputs: int = 1024, out=None):
        """Probabilities of the windows ending at MF rows first .. first+n-1
        (first >= lead): [n, numOut] FP32 device tensor."""
        torch = _torch()
        if first < self.lead:
            raise ValueError("the first window needs lead rows before it")
        if out is None:
            out = torch.empty((n, self.num_out), dtype=torch.float32, device=mf.device)
        per = max(1, min(outputs, self.max_rows - self.lead))
        done = 0
        while done < n:
            segs = min(math.ceil((n - done) / per), max(1, self.max_rows // (self.lead + per)))
            starts = first - self.lead + done + per * torch.arange(segs, dtype=torch.int64, device=mf.device)
            prob = self.forward(mf, starts, per)
            take = min(segs * per, n - done)
            out[done:done + take] = prob[:take]
            done += take
        return out


# ---------------------------------------------------------------- data
def mf_train_device(rates_hz, cs_idx, steps: int, *, t0: int = 0, trial_ms: int = 2000, cs_on: int = 500,
                    cs_off: int = 1000, cs_hz: float = 100.0, generator=None, device="cuda"):
    """uint8 [steps, numMF] Bernoulli MF train on the GPU with the rates and CS
    subset of configs.mf_rates: the input statistics of configs.mf_train, drawn
    on the device for long training streams."""
    torch = _torch()
    p = torch.as_tensor(np.asarray(rates_hz) * 1e-3, dtype=torch.float32, device=device)
    p = p.unsqueeze(0).expand(steps, -1).clone()
    tt = (t0 + torch.arange(steps, device=device)) % trial_ms
    in_cs = (tt >= cs_on) & (tt < cs_off)
    cs = torch.as_tensor(np.asarray(cs_idx), dtype=torch.int64, device=device)
    rows = torch.nonzero(in_cs).squeeze(1)
    if rows.numel() and cs.numel():
        p[rows.unsqueeze(1), cs.unsqueeze(0)] = cs_hz * 1e-3
    r = torch.rand((steps, p.shape[1]), generator=generator, device=device)
    return (r < p).to(torch.uint8)


class InputLayerData:
    """MF rows and the exact network's cluster PF fractions, both on the GPU:
    mf uint8 [rows, numMF]; y BF16 bits (int16) [rows, clusters] with y[r] the
    target of the window ending at MF row r (the last row has none)."""

    def __init__(self, mf, y, cells_per_cluster: int):
        self.mf, self.y, self.cells = mf, y, cells_per_cluster

    @property
    def rows(self) -> int:
        return self.mf.shape[0]

    @classmethod
    def record(cls, net, sg: Surrogate, trials: int, rates_hz, cs_idx, *, trial_ms: int = 2000,
               cells_per_cluster: int = 256, seed: int = 0, us: bool = True, progress=None):
        """Run `trials` trials of the delay protocol's MF input on the exact
        network `net` (continuing its state) and record (MF rows, targets)."""
        torch = _torch()
        s = net.sizes
        chunks = s.numGR // 64
        per = cells_per_cluster // 64
        n_cl = chunks // per
        T = trials * trial_ms
        mf = torch.empty((T, s.numMF), dtype=torch.uint8, device="cuda")
        y = torch.zeros((T, n_cl), dtype=torch.int16, device="cuda")
        pf = torch.empty((trial_ms, chunks), dtype=torch.uint8, device="cuda")
        net.set_recorder(["PF"], trial_ms)
        gen = torch.Generator(device="cuda")
        gen.manual_seed(seed)
        err = None
        if us:
            from . import configs
            err = configs.us_err_drive(trial_ms, s.numZones)
        for i in range(trials):
            train = mf_train_device(rates_hz, cs_idx, trial_ms, trial_ms=trial_ms, generator=gen)
            mf[i * trial_ms:(i + 1) * trial_ms] = train
            torch.cuda.synchronize()
            t_first = int(net.read("CUR_TIME")[0]) + 1
            net.run_mf_train(train, err_drive=err)
            net.read_recorded("PF", t_first, trial_ms, out=pf)
            torch.cuda.synchronize()
            # step j of this trial is the target of MF row i*T + j - 1
            if i == 0:
                sg.targets_from_pf(pf[1:], per, y[0:trial_ms - 1])
            else:
                sg.targets_from_pf(pf, per, y[i * trial_ms - 1:(i + 1) * trial_ms - 1])
            sg.sync()
            if progress:
                progress(i)
        net.set_recorder([], 0)
        return cls(mf, y, cells_per_cluster)


def bf16_bits_to_float(y):
    torch = _torch()
    return (y.to(torch.int32) << 16).view(torch.float32)


def train(sg: Surrogate, data: InputLayerData, steps: int, *, segments: int = 64, outputs: int = 256,
          lr: float = 2e-3, lr_end: float = 2e-4, seed: int = 0, rows: tuple[int, int] | None = None):
    """Adam on random segments of rows [lo, hi) of the data (default: all);
    the learning rate decays linearly from lr to lr_end.  Returns the loss of
    every step (numpy)."""
    torch = _torch()
    lo, hi = rows if rows is not None else (0, data.rows)
    P = sg.lead + outputs
    last = hi - 1 - P  # the last row of a segment must have a target
    if last < lo:
        raise ValueError("not enough rows for one segment")
    gen = torch.Generator(device="cuda")
    gen.manual_seed(seed)
    starts = torch.randint(lo, last + 1, (steps, segments), generator=gen, device="cuda", dtype=torch.int64)
    losses = torch.zeros(steps, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    for i in range(steps):
        f = i / max(1, steps - 1)
        sg.train_step(data.mf, starts[i], outputs, data.y, lr + (lr_end - lr) * f, loss_out=losses[i:i + 1],
                      sync_inputs=False)
    sg.sync()
    return losses.cpu().numpy()


def pf_rmse(pred, target_bits, cells: int = 256) -> float:
    """RMS error, in PF spikes per cluster per ms, between predicted
    probabilities [n, clusters] and BF16 target fractions (device tensors)."""
    torch = _torch()
    y = bf16_bits_to_float(target_bits)
    d = (pred.to(torch.float64) - y.to(torch.float64)) * cells
    return float(torch.sqrt(torch.mean(d * d)))


# ---------------------------------------------------------------- sweeps
def _segments(dt, trial: int):
    """(first, last+1, LTD/LTP window on) of the delay protocol's trial
    `trial`, as row offsets of a stream of back-to-back trials."""
    base = trial * dt.trial_ms
    return [(base + a, base + b, what == "set") for a, b, what in dt.segments()]


def sweep_experiment(net_name: str = "c1", *, seed: int = 0, mf_seed: int = 11, train_trials: int = 1000,
                     train_steps: int = 600, segments: int = 64, outputs: int = 256, lr: float = 2e-3,
                     sweep_trials: int = 10, ltd_scales=(0.25, 0.5, 1, 2, 4, 8, 16, 32),
                     ltp_scales=(0.25, 0.5, 1, 2, 4, 8, 16, 32), judged=(0, 18, 36, 63), cells: int = 256,
                     hidden: int = 128, log=print) -> dict:
    """The learned input layer driving a plasticity sweep, judged by the exact step.

    1. The network `net_name` (GRs ordered within SC blocks by their number of
       CS MF inputs) runs `train_trials` delay trials; its MF rows and cluster
       PF fractions train the surrogate.
    2. A fresh MF stream of 1 + sweep_trials trials: the surrogate predicts the
       PF probabilities of the last sweep_trials (the first gives context).
    3. The sweep: one network per (LTD scale, LTP scale) pair, from PF->PC
       weights of a network that has learned the CS, runs the sweep trials
       with its input layer replaced by the predictions (cbm_batch_sweep); the
       protocol's LTD/LTP steps and the reset steps are scaled by the pair.
    4. The judge: the networks `judged` run the same trials with the exact
       step; PF rates (RMSE per cluster and ms), CR onset and CR amplitude per
       trial are compared.  The same judged pairs are also swept with each
       cluster's mean training rate in place of the predictions (the sweep
       without the trained network).
    Returns the figures as a dict."""
    import time
    torch = _torch()
    from . import configs
    from .eyelid import DelayTrials, EyelidBank, cr_amplitude, cr_onset
    from .network import Batch, Network
    res = {"network": net_name, "cells_per_cluster": cells}
    st = configs.build_state(net_name, seed=seed)
    s = st.sizes
    rates, cs = configs.mf_rates(s.numMF, seed=mf_seed)
    configs.order_gr_by_inputs(st, cs)
    w0 = configs.trained_pfpc_weights(st, cs)
    n_cl = s.numGR // cells
    Z = st.num_zones
    dt = DelayTrials(st.params, s.numNC)
    TM = dt.trial_ms

    # 1. data and training
    gen_net = Network.from_state(st)
    sg = Surrogate(num_in=s.numMF, hidden=hidden, num_out=n_cl, max_rows=max(16384, segments * (49 + outputs)))
    t = time.time()
    step = max(1, train_trials // 10)
    data = InputLayerData.record(gen_net, sg, train_trials, rates, cs, cells_per_cluster=cells, seed=seed + 1,
                                 progress=lambda i: (i + 1) % step == 0 and log(f"[c4] recorded trial {i + 1}"))
    torch.cuda.synchronize()
    res["record_s"] = time.time() - t
    res["windows"] = data.rows - 1
    gen_net.close()
    y = bf16_bits_to_float(data.y)
    mean = y[:-1].mean(dim=0)
    m = float(mean.mean())
    sg.init_params(seed=seed, out_bias=float(np.log(m / (1 - m))))
    t = time.time()
    losses = train(sg, data, train_steps, segments=segments, outputs=outputs, lr=lr, lr_end=lr / 10, seed=seed)
    res["train_s"] = time.time() - t
    us, fl = sg.last_timing()
    res["train_step_us"], res["train_step_tflops"] = us, fl / us / 1e6
    res["loss_first"], res["loss_last"] = float(losses[:10].mean()), float(losses[-10:].mean())
    log(f"[c4] recorded {res['windows']} windows in {res['record_s']:.1f}s, trained {train_steps} steps in "
        f"{res['train_s']:.1f}s ({res['train_step_tflops']:.0f} TFLOP/s per step); loss {res['loss_first']:.5f} -> "
        f"{res['loss_last']:.5f}")
    del data, y

    # 2. the sweep stream and its predictions
    K = sweep_trials
    gen = torch.Generator(device="cuda")
    gen.manual_seed(seed + 7)
    mf = mf_train_device(rates, cs, (1 + K) * TM, generator=gen)
    R0 = TM
    t = time.time()
    prob = sg.predict(mf, R0 - 200, 199 + K * TM)
    torch.cuda.synchronize()
    res["predict_s"] = time.time() - t
    us, fl = sg.last_timing()
    res["forward_tflops"] = fl / us / 1e6
    train_mf = mf[R0:]
    flat_prob = torch.zeros_like(prob)
    flat_prob[:] = mean.unsqueeze(0)
    err = configs.us_err_drive(K * TM, Z)
    pairs = [(a, b) for a in ltd_scales for b in ltp_scales]
    ltd_def, ltp_def = st.params.synLTDStepSizeGRtoPC, st.params.synLTPStepSizeGRtoPC

    def make(pair):
        n = Network.from_state(st)
        for z in range(Z):
            n.write("PFPC_SYNW", w0, z)
        n.set_recorder(["NC", "PC"], K * TM)
        return n

    def steps_for(pair, on):
        a, b = pair
        return (float(dt.ltd) * a, float(dt.ltp) * b) if on else (ltd_def * a, ltp_def * b)

    spikes = {}

    def readout(nets, tag):
        nc = np.stack([n.read_recorded("NC", 1, K * TM)[:, :s.numNC] for n in nets], axis=1)
        pc = np.stack([n.read_recorded("PC", 1, K * TM)[:, :s.numPC] for n in nets], axis=1)
        spikes[tag] = {"NC": nc.reshape(K, TM, len(nets), -1).sum(axis=(1, 3)).T.tolist(),
                       "PC": pc.reshape(K, TM, len(nets), -1).sum(axis=(1, 3)).T.tolist()}
        eye = EyelidBank(len(nets), s.numNC).trace(nc)
        out = []
        for i in range(len(nets)):
            per = []
            for k in range(K):
                tr = eye[k * TM:(k + 1) * TM, i]
                per.append((cr_onset(tr, dt.cs_on, dt.cs_off), cr_amplitude(tr, dt.cs_on, dt.cs_off)))
            out.append(per)
        return out

    def run_sweep(pairs_, p_, tag):
        nets = [make(p) for p in pairs_]
        b = Batch(nets)
        torch.cuda.synchronize()
        t0 = time.time()
        for k in range(K):
            for lo, hi, on in _segments(dt, k):
                for n, p in zip(nets, pairs_):
                    for z in range(Z):
                        n.set_grpc_plast_steps(z, *steps_for(p, on))
                b.sweep(p_, 199 + lo, hi - lo, train_mf[lo:hi], err_drive=[err[lo:hi]] * len(nets),
                        cells_per_cluster=cells)
        b.sync()
        el = time.time() - t0
        r = readout(nets, tag)
        b.close()
        for n in nets:
            n.close()
        return r, el

    # 3. the sweep
    sweep, t_sweep = run_sweep(pairs, prob, "surrogate")
    res["members"] = len(pairs)
    res["sweep_s"] = t_sweep
    res["sweep_network_ms_per_s"] = len(pairs) * K * TM / t_sweep
    log(f"[c4] sweep of {len(pairs)} parameter sets x {K} trials: {t_sweep:.2f}s, "
        f"{res['sweep_network_ms_per_s']:.0f} network-ms/s")

    # 4. the judge: exact runs of the judged pairs, and the sweep without the trained network
    jp = [pairs[j] for j in judged]
    exact_nets = [make(p) for p in jp]
    exact_nets[0].set_recorder(["NC", "PC", "PF"], K * TM)
    torch.cuda.synchronize()
    t0 = time.time()
    for k in range(K):
        for lo, hi, on in _segments(dt, k):
            for n, p in zip(exact_nets, jp):
                for z in range(Z):
                    n.set_grpc_plast_steps(z, *steps_for(p, on))
                n.run_mf_train(train_mf[lo:hi], err_drive=err[lo:hi])
    for n in exact_nets:
        n.sync()
    t_exact = time.time() - t0
    res["exact_network_ms_per_s"] = len(jp) * K * TM / t_exact
    exact = readout(exact_nets, "exact")
    pf = torch.empty((K * TM, s.numGR // 64), dtype=torch.uint8, device="cuda")
    exact_nets[0].read_recorded("PF", 1, K * TM, out=pf)
    torch.cuda.synchronize()
    yk = torch.zeros((K * TM, n_cl), dtype=torch.int16, device="cuda")
    sg.targets_from_pf(pf, cells // 64, yk)
    sg.sync()
    for n in exact_nets:
        n.close()
    res["pf_rmse"] = pf_rmse(prob[199:199 + K * TM], yk, cells)
    res["pf_rmse_cluster_mean"] = pf_rmse(flat_prob[199:199 + K * TM].contiguous(), yk, cells)
    flat, _ = run_sweep(jp, flat_prob, "cluster_mean")
    # the sweep driven by the exact network's own cluster PF fractions: what
    # summarising each cluster by its mean costs, without the surrogate's error
    own = torch.zeros_like(prob)
    own[199:199 + K * TM] = bf16_bits_to_float(yk)
    own_cr, _ = run_sweep(jp, own, "exact_fractions")

    def compare(got):
        d_on, d_amp, miss = [], [], 0
        for g, e in zip(got, exact):
            for (go, ga), (eo, ea) in zip(g, e):
                d_amp.append(abs(ga - ea))
                if (go is None) != (eo is None):
                    miss += 1
                elif go is not None:
                    d_on.append(abs(go - eo))
        return {"onset_abs_err_ms": float(np.mean(d_on)) if d_on else None, "onset_presence_mismatch": miss,
                "amplitude_abs_err": float(np.mean(d_amp)), "trials": len(d_amp)}

    res["judged"] = [list(p) for p in jp]
    res["surrogate_vs_exact"] = compare([sweep[j] for j in judged])
    res["cluster_mean_vs_exact"] = compare(flat)
    res["exact_fractions_vs_exact"] = compare(own_cr)
    res["spikes_per_trial"] = {k: {p: v[p][:2] for p in v} for k, v in spikes.items()}
    res["exact_fractions_cr"] = [[(o, round(a, 3)) for o, a in e] for e in own_cr]
    res["exact_cr"] = [[(o, round(a, 3)) for o, a in e] for e in exact]
    res["surrogate_cr"] = [[(o, round(a, 3)) for o, a in sweep[j]] for j in judged]
    res["speedup_vs_exact"] = res["sweep_network_ms_per_s"] / res["exact_network_ms_per_s"]
    log(f"[c4] exact: {res['exact_network_ms_per_s']:.0f} network-ms/s; PF rmse surrogate {res['pf_rmse']:.4f} vs "
        f"cluster mean {res['pf_rmse_cluster_mean']:.4f}; CR vs exact: surrogate {res['surrogate_vs_exact']}, "
        f"cluster mean {res['cluster_mean_vs_exact']}, exact fractions {res['exact_fractions_vs_exact']}")
    sg.close()
    return res

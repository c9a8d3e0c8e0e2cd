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

"""Eyelid-conditioning delay trials and their readouts.

* EyelidIntegrator: the eyelid position readout of the NC spikes,
  CBM_TOOLS_LIB_DEV/src/eyelidintegrator.cpp:13-68, in binary32 with the
  reference's operand order (expf from the C library, as exp(float) there).
* DelayTrials: the per-ms protocol of ECManagementDelay::calcSimActivity
  (Eyelid_Sim_CBM_Dev/src/ectrial/ecmanagementdelay.cpp:201-273): the US as a
  relative error drive of 1.0 on zone 0 at trial time csOff-1, PF->PC LTD/LTP
  steps of (-0.0005 * ratio, 0.0005) from csOn+200 until csOff+200 with
  ratio = ((csOff-csOn) + msLTDStartAPIO) / msLTDDurationIO floored at 0.5
  (ecmanagementdelay.cpp:85-90), and the eyelid integrator on zone 0's NC
  spikes after every step.  Trial time of step k is k mod the trial length
  (ecmanagementbase.cpp:55-56,88-95).  The MF trains are the caller's (the
  reference draws them with PoissonRegenCells; the trains here are seeded
  Bernoulli trains, configs.mf_train).
* cr_onset: the first ms of the CS window at which the eyelid has closed by
  more than a threshold over its pre-CS baseline.  The reference tree holds no
  onset definition (its analysis tools are not part of it); this one is ours.
"""
from __future__ import annotations

import ctypes

import numpy as np

_libm = ctypes.CDLL("libm.so.6")
_libm.expf.restype = ctypes.c_float
_libm.expf.argtypes = [ctypes.c_float]

f32 = np.float32


class EyelidIntegrator:
    """EyelidIntegrator(numInputCells, msPerTimeStep, gDecayTau, gIncrease,
    gShift, gLeakRaw, maxAmplitude) with the defaults the delay protocol uses
    (ecmanagementdelay.cpp:92-93)."""

    def __init__(self, num_in: int, ms_per_ts: int = 1, g_decay_tau: float = 10.5, g_increase: float = 0.012,
                 g_shift: float = -0.13, g_leak_raw: float = 0.13, max_amplitude: float = 100.0):
        self.gDecay = f32(_libm.expf(float(f32(-int(ms_per_ts)) / f32(g_decay_tau))))
        self.gInc = f32(g_increase)
        self.gShift = f32(g_shift)
        self.evMax = f32(max_amplitude)
        self.gLeak = f32(g_leak_raw) / f32(6 - int(ms_per_ts))
        self.vm = f32(0.0)
        self.gIn = np.zeros(num_in, f32)

    def step(self, ap_in) -> np.float32:
        """calcStep(apIn): one ms of the integrator; returns the eyelid position."""
        self.gIn = self.gIn * self.gDecay
        self.gIn = self.gIn + np.asarray(ap_in).astype(f32) * self.gInc
        g_sum = f32(0.0)
        for g in self.gIn:
            g_sum = f32(g_sum + g)
        g_sum = f32(g_sum + self.gShift)
        g_sum = f32(g_sum > 0) * g_sum
        self.vm = f32(f32(self.vm + f32(self.gLeak * f32(f32(0.0) - self.vm))) +
                      f32(g_sum * f32(self.evMax - self.vm)))
        return self.vm

    def trace(self, ap_nc: np.ndarray) -> np.ndarray:
        """Eyelid positions of a [steps, numNC] spike raster."""
        return np.array([self.step(row) for row in ap_nc], f32)


class EyelidBank:
    """Many EyelidIntegrators stepped together (one per row), with the
    arithmetic of EyelidIntegrator.step applied element-wise, so row m gives
    exactly what its own integrator would."""

    def __init__(self, count: int, num_in: int, ms_per_ts: int = 1, **kw):
        one = EyelidIntegrator(num_in, ms_per_ts, **kw)
        self.gDecay, self.gInc, self.gShift, self.evMax, self.gLeak = (one.gDecay, one.gInc, one.gShift,
                                                                         one.evMax, one.gLeak)
        self.vm = np.zeros(count, f32)
        self.gIn = np.zeros((count, num_in), f32)

    def trace(self, ap: np.ndarray) -> np.ndarray:
        """ap: [steps, count, num_in] spikes -> eyelid positions [steps, count]."""
        out = np.empty(ap.shape[:2], f32)
        zero = f32(0.0)
        for k in range(ap.shape[0]):
            self.gIn = self.gIn * self.gDecay
            self.gIn = self.gIn + ap[k].astype(f32) * self.gInc
            g = np.zeros(self.vm.shape, f32)
            for j in range(self.gIn.shape[1]):
                g = (g + self.gIn[:, j]).astype(f32)
            g = (g + self.gShift).astype(f32)
            g = ((g > 0).astype(f32) * g).astype(f32)
            leak = (self.gLeak * (zero - self.vm)).astype(f32)
            self.vm = ((self.vm + leak).astype(f32) + (g * (self.evMax - self.vm)).astype(f32)).astype(f32)
            out[k] = self.vm
        return out


def cr_amplitude(trace: np.ndarray, cs_on: int, cs_off: int, baseline_ms: int = 200) -> float:
    """Largest closure in [cs_on, cs_off) over the mean of the baseline_ms before cs_on."""
    base = float(np.mean(trace[cs_on - baseline_ms:cs_on], dtype=np.float64))
    return float(np.max(trace[cs_on:cs_off].astype(np.float64)) - base)


def cr_onset(trace: np.ndarray, cs_on: int, cs_off: int, threshold: float = 5.0, baseline_ms: int = 200):
    """First trial time in [cs_on, cs_off) at which trace - mean(trace over the
    baseline_ms before cs_on) exceeds threshold (eyelid units, 100 = closed);
    None when it never does."""
    base = float(np.mean(trace[cs_on - baseline_ms:cs_on], dtype=np.float64))
    above = np.nonzero(trace[cs_on:cs_off].astype(np.float64) - base > threshold)[0]
    return cs_on + int(above[0]) if above.size else None


class DelayTrials:
    """The delay-conditioning protocol on one network (a Network, or anything
    with the same update_mf_input / calc_activity / run_mf_train / read /
    set_grpc_plast_steps / reset_grpc_plast_steps / update_err_drive calls)."""

    def __init__(self, params, num_nc: int, *, trial_ms: int = 2000, cs_on: int = 500, cs_off: int = 1000,
                 zone: int = 0):
        self.trial_ms, self.cs_on, self.cs_off, self.zone = trial_ms, cs_on, cs_off, zone
        ms = int(round(params.msPerTimeStep))
        ltd_start = f32(params.tsLTDStartAPIO * ms)
        ltd_dur = f32(params.tsLTDDurationIO * ms)
        ratio = f32(f32(cs_off - cs_on) + ltd_start) / ltd_dur
        self.ratio = max(ratio, f32(0.5))
        self.ltd = f32(-0.0005) * self.ratio
        self.ltp = f32(0.0005)
        self.num_nc = num_nc
        self.eyelid = EyelidIntegrator(num_nc, ms)

    def segments(self):
        """(first, last+1, action before the first step) of one trial."""
        a, b = self.cs_on + 200, self.cs_off + 200
        return [(0, a, None), (a, b, "set"), (b, self.trial_ms, "reset")]

    def us_drive(self, num_zones: int) -> np.ndarray:
        e = np.zeros((self.trial_ms, num_zones), np.float32)
        e[self.cs_off - 1, self.zone] = 1.0
        return e

    def _action(self, model, what):
        if what == "set":
            model.set_grpc_plast_steps(self.zone, float(self.ltd), float(self.ltp))
        elif what == "reset":
            model.reset_grpc_plast_steps(self.zone)

    def run_trial_device(self, net, train_dev, num_zones: int) -> dict:
        """One trial of a GPU network from a device-resident MF train
        ([trial_ms, numMF]): the segments run as device-resident trains and the
        spikes come from the device recorder, which the caller has set to keep
        at least PC and NC for trial_ms steps.  Returns zone `zone`'s
        {"NC": [trial_ms, numNC], "PC": [trial_ms, numPC]} spikes."""
        err = self.us_drive(num_zones)
        t_start = int(net.read("CUR_TIME")[0]) + 1
        for a, b, what in self.segments():
            self._action(net, what)
            net.run_mf_train(train_dev[a:b], err_drive=err[a:b])
        out = {}
        for pop, n in (("NC", self.num_nc), ("PC", net.sizes.numPC)):
            x = net.read_recorded(pop, t_start, self.trial_ms)
            out[pop] = x[:, self.zone * n:(self.zone + 1) * n]
        return out

    def run_trial_stepwise(self, model, train: np.ndarray, num_zones: int, num_pc: int) -> dict:
        """The same trial one ms at a time, reading the spikes after every step
        (as the reference's driver does)."""
        nc = np.zeros((self.trial_ms, self.num_nc), np.uint8)
        pc = np.zeros((self.trial_ms, num_pc), np.uint8)
        for a, b, what in self.segments():
            self._action(model, what)
            for k in range(a, b):
                if k == self.cs_off - 1:
                    model.update_err_drive(self.zone, 1.0)
                model.update_mf_input(train[k])
                model.calc_activity(1)
                nc[k] = model.read("AP_NC", self.zone)
                pc[k] = model.read("AP_PC", self.zone)
        return {"NC": nc, "PC": pc}

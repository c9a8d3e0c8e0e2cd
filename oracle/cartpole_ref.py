"""Oracle restatement of the cart-pole environment, its two microzones and the
mossy-fibre spike generator.

TEST INFRASTRUCTURE: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use this module, as the checker.  It restates the reference
directly, in numpy binary32 scalars and arrays with the reference's binary64
promotions, drawing from the oracle's SFMT (orc_sfmt.c, pinned by the
reference's own generator):

  Environment.setupMossyFibers  Matthew_Sim_CBM/src/environments/environment.cpp:19-38
  Environment.setupStateVariables  environment.cpp:40-66 (random MF assignment)
  StateVariable.update / gaussMFAct  Matthew_Sim_CBM/includes/statevariable.hpp:92-143
  Cartpole (constructor, reset, getState, step, computePhysics, setMZErr,
    inFailure, calcForce, terminated)  src/environments/cartpole.cpp:24-285,
    constants includes/environments/cartpole.hpp:42-58,141-152
  Microzone.update / deliverError  src/microzone.cpp:23-50
  PoissonRegenCells  CBM_TOOLS_LIB_DEV/src/poissonregencells.cpp:10-66

Choices where the reference's behaviour is undefined or outside this path,
the same as the library's: the microzones' force starts at 0 (never
initialised there) and, with closed_loop, both microzones are updated from
their NC spikes at the start of every step (the reference's Cartpole::step
never updates them, which closed_loop=False keeps); the pole's exp in the
Gaussian tuning is C's binary64 exp; PoissonRegenCells' threshold decay uses
the binary32 expf; cos/sin of the angle are binary32 (cosf, sinf).

Parity pins (tests/test_cartpole.py): the equations of motion against the
point-mass cart-pole written in another algebraic form in binary64; the
Gaussian tuning against its closed form exp(-k^2/72) at slot centres; the
MF generator's mean rate against the renewal-theory rate of its recovering
threshold; the order and number of generator draws against the oracle SFMT's
stream.
"""
from __future__ import annotations

import ctypes as C
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
f32 = np.float32
FLT_MAX = float(np.finfo(np.float32).max)

_libm = C.CDLL("libm.so.6")
for _n in ("cosf", "sinf", "expf"):
    getattr(_libm, _n).restype = C.c_float
    getattr(_libm, _n).argtypes = [C.c_float]
_orc = None


def _lib():
    global _orc
    if _orc is None:
        path = os.path.join(HERE, "build", "liborc.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} missing: run `make -C oracle`")
        L = C.CDLL(path)
        L.orc_sfmt_init.argtypes = [C.c_void_p, C.c_int]
        L.orc_sfmt_random.argtypes = [C.c_void_p]
        L.orc_sfmt_random.restype = C.c_double
        L.orc_sfmt_irandom.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.orc_sfmt_irandom.restype = C.c_int
        L.orc_sfmt_random_fill.argtypes = [C.c_void_p, C.c_void_p, C.c_int]
        L.orc_sfmt_size.restype = C.c_int
        _orc = L
    return _orc


class Gen:
    """CRandomSFMT0 on the oracle's SFMT."""

    def __init__(self, seed: int):
        L = _lib()
        self._buf = C.create_string_buffer(L.orc_sfmt_size() + 64)
        self._p = (C.addressof(self._buf) + 63) & ~63
        L.orc_sfmt_init(self._p, int(seed))
        self.draws = 0  # Random() and IRandom() calls made

    def random(self) -> float:
        self.draws += 1
        return _lib().orc_sfmt_random(self._p)

    def irandom(self, lo: int, hi: int) -> int:
        self.draws += 1
        return _lib().orc_sfmt_irandom(self._p, lo, hi)

    def random_n(self, n: int) -> np.ndarray:
        out = np.empty(n, np.float64)
        _lib().orc_sfmt_random_fill(self._p, out.ctypes.data, n)
        self.draws += n
        return out


class PoissonRef:
    """PoissonRegenCells(numCells, seed, threshDecayTau, msPerTimeStep)."""

    def __init__(self, num_cells: int, seed: int, tau: float = 4.0, ms_per_ts: float = 1.0):
        seeder = Gen(seed)
        self.gen = Gen(seeder.irandom(0, 2**31 - 1))
        self.n = num_cells
        self.sPerTS = f32(f32(ms_per_ts) / f32(1000))
        self.decay = f32(f32(1) - f32(_libm.expf(float(f32(-f32(ms_per_ts)) / f32(tau)))))
        self.thresh = np.ones(num_cells, np.float32)

    def step(self, freq: np.ndarray) -> np.ndarray:
        t = self.thresh + (f32(1) - self.thresh) * self.decay
        with np.errstate(invalid="ignore"):
            p = (np.asarray(freq, np.float32) * self.sPerTS) * t
            spike = self.gen.random_n(self.n) < p.astype(np.float64)
        self.thresh = np.where(spike, f32(0), t).astype(np.float32)
        return spike.astype(np.uint8)


class MicrozoneRef:
    WINDOW = 100

    def __init__(self, num_nc: int, scale=0.1, power=1.0, decay=0.99):
        self.num_nc = num_nc
        self.scale, self.power, self.decay = f32(scale), f32(power), f32(decay)
        self.force = f32(0)
        self.queue = [f32(0)] * self.WINDOW
        self.act_sum = f32(0)
        self.avg = f32(0)
        self.errors = 0

    def update(self, ap_nc) -> None:
        s = f32(0)
        for v in np.asarray(ap_nc, np.uint8):
            s = f32(s + f32(v))
        drive = f32(f32(s / f32(self.num_nc)) * self.scale)
        # pow(x, 1) is x; other powers through binary32 pow
        inc = drive if self.power == f32(1) else f32(np.power(drive, self.power, dtype=np.float32))
        self.force = f32(f32(self.force + inc) * self.decay)
        self.act_sum = f32(self.act_sum + s)
        self.queue.append(s)
        self.act_sum = f32(self.act_sum - self.queue.pop(0))
        self.avg = f32(min(f32(1), max(f32(0), f32(self.act_sum / f32(self.WINDOW)))))


class CartpoleRef:
    GRAVITY, CART_MASS, POLE_MASS, LENGTH = f32(9.8), f32(1.0), f32(0.1), f32(7.5)
    FORCE_SCALE, FORCE_DECAY, TAU = f32(0.1), f32(0.99), f32(0.001)
    TIMEOUT = 1500
    LEFT_ANGLE, RIGHT_ANGLE = f32(-0.1745), f32(0.1745)
    MIN_POLE_VEL, MAX_POLE_VEL = f32(-0.5), f32(0.5)
    MIN_POLE_ACC, MAX_POLE_ACC = f32(-5.0), f32(5.0)
    PROPORTIONS = (f32(0.03), f32(0.06), f32(0.06), f32(0.06), f32(0.06))  # high, pole vel, pole ang, cart vel, pos

    def __init__(self, num_mf: int, num_nc: int, env_seed: int, mf_seed: int | None = None, *,
                 max_num_trials: int = 20, max_trial_len: int = 1_000_000, track_len: float = 0.0,
                 closed_loop: bool = False):
        self.gen = Gen(env_seed)
        self.num_mf, self.num_nc = num_mf, num_nc
        self.max_num_trials, self.max_trial_len = max_num_trials, max_trial_len
        self.closed_loop = closed_loop
        tl = f32(track_len)
        self.track_len = f32(FLT_MAX) if tl <= 0 else tl
        self.left_track = f32(-float(self.track_len) / 2.0)
        self.right_track = f32(float(self.track_len) / 2.0)
        self.trial, self.time_aloft, self.cycle = -1, 0, 0
        self.x = self.x_dot = self.theta = self.theta_dot = self.theta_dd = f32(0)
        self.force_left = self.force_right = f32(0)
        self.fallen, self.timeout, self.mil = False, 0, 0
        self.mz = [MicrozoneRef(num_nc), MicrozoneRef(num_nc)]
        self._reset()
        # MF rates: relaxed U[1, 10), then the excited draws (all 40 Hz)
        self.relaxed = np.array([f32(self.gen.random() * float(f32(9)) + 1.0) for _ in range(num_mf)], np.float32)
        self.excited = np.array([f32(self.gen.random() * 0.0 + 40.0) for _ in range(num_mf)], np.float32)
        # MF groups drawn without replacement from the pool numMF-1 .. 0
        pool = list(range(num_mf - 1, -1, -1))
        self.groups = []
        for p in self.PROPORTIONS:
            k = int(f32(num_mf) * p)
            g = []
            for _ in range(k):
                j = self.gen.irandom(0, len(pool) - 1)
                g.append(pool.pop(j))
            self.groups.append(np.array(g, np.int64))
        self.ranges = [None, (self.MIN_POLE_VEL, self.MAX_POLE_VEL), (self.LEFT_ANGLE, self.RIGHT_ANGLE),
                       (f32(-3.0), f32(3.0)), (self.left_track, self.right_track)]
        self.mfs = PoissonRef(num_mf, env_seed if mf_seed is None else mf_seed)

    # ------------------------------------------------------------ pieces
    def _reset(self) -> None:
        self.fallen = False
        self.timeout = 0
        self.x = self.x_dot = f32(0)
        self.theta = f32((self.gen.random() - 0.5) * 0.00001)
        self.theta_dot = f32((self.gen.random() - 0.5) * 0.00001)
        self.theta_dd = f32(0)
        self.trial += 1
        self.time_aloft = 0

    def _values(self):
        return [None, self.theta_dot, self.theta, self.x_dot, self.x]

    def gaussian(self, members: np.ndarray, lo, hi, value, freq: np.ndarray) -> None:
        np.seterr(over="ignore", invalid="ignore", divide="ignore")
        cur = max(lo, min(hi, value))
        rng = f32(hi - lo)
        slot = f32(rng / f32(len(members)))
        centre = f32(float(lo) + float(slot) / 2.0)
        sd = f32(f32(6.0) * slot)
        sd2 = f32(sd * sd)
        dens_norm = math.sqrt(2 * math.pi * float(sd2))  # binary64; inf when sd2 is
        peak = f32(1.0 / dens_norm)
        with np.errstate(all="ignore"):
            for m in members:
                d = f32(cur - centre)
                arg = f32(f32(f32(-1) * f32(d * d)) / f32(f32(2) * sd2))
                dens = f32(math.exp(float(arg)) / dens_norm)
                g = f32(dens / peak)
                freq[m] = f32(f32(g * self.excited[m]) + f32(f32(f32(1) - g) * self.relaxed[m]))
                centre = f32(centre + slot)

    def get_state(self) -> np.ndarray:
        freq = self.relaxed.copy()
        freq[self.groups[0]] = self.excited[self.groups[0]]
        if not self.timeout < self.TIMEOUT:
            vals = self._values()
            for g in range(1, 5):
                lo, hi = self.ranges[g]
                self.gaussian(self.groups[g], lo, hi, vals[g], freq)
        return freq

    def physics(self, force) -> None:
        g, M, m, L, tau = self.GRAVITY, self.CART_MASS, self.POLE_MASS, self.LENGTH, self.TAU
        c = f32(_libm.cosf(float(self.theta)))
        s = f32(_libm.sinf(float(self.theta)))
        td = self.theta_dot
        n1 = f32(f32(f32(-1) * force) * c)
        n1 = f32(n1 - f32(f32(g * m) * s))
        n1 = f32(n1 - f32(f32(g * M) * s))
        n1 = f32(n1 + f32(f32(f32(f32(f32(L * m) * td) * td) * c) * s))
        d1 = f32(L * f32(f32(m + M) - f32(f32(m * c) * c)))
        thetaacc = f32(f32(-n1) / d1)
        n2 = f32(f32(f32(f32(f32(-g) * L) * m) * c) * s)
        n2 = f32(n2 + f32(L * f32(f32(f32(-1) * force) + f32(f32(f32(f32(L * m) * td) * td) * s))))
        d2 = f32(f32(L * f32(M + m)) - f32(f32(L * m) * c))
        xacc = f32(f32(-n2) / d2)
        self.x = f32(self.x + f32(tau * self.x_dot))
        self.x_dot = f32(self.x_dot + f32(tau * xacc))
        self.theta = f32(self.theta + f32(tau * self.theta_dot))
        self.theta_dot = f32(self.theta_dot + f32(tau * thetaacc))
        self.theta_dot = min(max(self.MIN_POLE_VEL, self.theta_dot), self.MAX_POLE_VEL)
        self.theta_dd = min(max(self.MIN_POLE_ACC, thetaacc), self.MAX_POLE_ACC)
        while float(self.theta) >= math.pi:
            self.theta = f32(float(self.theta) - 2.0 * math.pi)
        while float(self.theta) < -math.pi:
            self.theta = f32(float(self.theta) + 2.0 * math.pi)
        return thetaacc, xacc

    def errors(self) -> tuple[int, int]:
        left = right = 0
        pmax = f32(0.01)
        p = min(f32(abs(self.theta)), pmax)
        hit = self.gen.random() < float(p)
        if self.theta >= 0 and self.theta_dot >= 0 and hit:
            left += 1
        if self.theta < 0 and self.theta_dot < 0 and hit:
            right += 1
        if self.track_len < f32(FLT_MAX):
            p = min(f32(f32(f32(0.005) * f32(abs(self.x))) / self.right_track), pmax)
            hit = self.gen.random() < float(p)
            if self.x < 0 and hit:
                left += 1
            if self.x > 0 and hit:
                right += 1
            p = min(f32(f32(0.005) * f32(abs(self.x_dot))), pmax)
            hit = self.gen.random() < float(p)
            if self.x_dot > 0 and self.x < 0 and hit:
                left += 1
            if self.x_dot < 0 and self.x > 0 and hit:
                left += 1
        return left, right

    def failure(self) -> str:
        if self.theta <= self.LEFT_ANGLE:
            return "PoleFallRight"
        if self.theta >= self.RIGHT_ANGLE:
            return "PoleFallLeft"
        if self.x <= self.left_track:
            return "CartFallLeft"
        if self.x >= self.right_track:
            return "CartFallRight"
        if self.time_aloft >= self.max_trial_len:
            return "MaxTrialLength"
        return "NoFailure"

    @property
    def terminated(self) -> bool:
        return self.trial >= self.max_num_trials

    def step(self, nc_left=None, nc_right=None) -> tuple[np.ndarray, tuple[int, int]]:
        """The environment's step after the network's; returns the trace row
        (cbm_cartpole_col order) and the errors delivered to zones 0 and 1."""
        self.cycle += 1
        self.timeout += 1
        if self.closed_loop:
            self.mz[0].update(nc_left)
            self.mz[1].update(nc_right)
        if self.fallen:
            mode = self.failure()
            if mode == "MaxTrialLength":
                self.mil += 1
            else:
                if self.mil >= 5:
                    self.max_num_trials = self.trial
                self.mil = 0
            self._reset()
        err = (0, 0)
        if self.timeout > self.TIMEOUT:
            self.force_right = self.mz[1].force
            self.force_left = self.mz[0].force
            self.physics(f32(self.force_right - self.force_left))
            err = self.errors()
            self.fallen = (self.theta <= self.LEFT_ANGLE or self.theta >= self.RIGHT_ANGLE or
                           self.x <= self.left_track or self.x >= self.right_track or
                           self.time_aloft >= self.max_trial_len)
            self.time_aloft += 1
        ncl = 0.0 if nc_left is None else float(np.asarray(nc_left, np.uint8).sum())
        ncr = 0.0 if nc_right is None else float(np.asarray(nc_right, np.uint8).sum())
        row = np.array([self.theta, self.theta_dot, self.x, self.x_dot, self.force_left, self.force_right,
                        err[0], err[1], self.time_aloft, self.trial, ncl, ncr, 0], np.float32)
        return row, err

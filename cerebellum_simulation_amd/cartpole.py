"""The cart-pole environment driven by two microzones (include/cbm_env.h).

The environment, its microzones and the MF spike generator are the C++
classes of libcbm_core (CBMEnvInclude/, CBMToolsInclude/poissonregencells.h);
this module is their host-side handle.  ``Cartpole.run`` is the loop of the
reference's headless driver (Matthew_Sim_CBM/src/simthread.cpp:96-165): per
step the environment's MF rates, MF spikes from PoissonRegenCells, one network
step, and the environment's step, which reads the two microzones' NC spikes
(closed loop) and delivers error to them.

``simthread_seeds`` and ``build_state`` reproduce how the reference's
driver seeds a run (main.cpp:73-87, simthread.cpp:7-43): the C library's
srand(seed) / rand() pairs for each zone's connectivity and activity seeds,
``CRandomSFMT0(seed)`` for the environment and ``PoissonRegenCells(numMF,
seed, 4, 1)`` for the MFs.  The reference hands ``&randSeed`` to CBMSimCore as
the array of zone generator seeds, which only holds zone 0's; zone z gets
seed + z here.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _abi
from ._abi import CartpoleOpts

COLS = ("theta", "theta_dot", "x", "x_dot", "force_left", "force_right", "err_left", "err_right", "time_aloft",
        "trial", "nc_left", "nc_right", "mf_spikes")
COL = {k: i for i, k in enumerate(COLS)}
GROUPS = ("high_freq", "pole_vel", "pole_ang", "cart_vel", "cart_pos")


def _check(rc: int, where: str) -> None:
    if rc:
        msg = _abi.core_lib().cbm_env_last_error().decode(errors="replace")
        raise _abi.CbmError(f"{where} failed ({rc}): {msg}")


def simthread_seeds(seed: int, num_zones: int = 2) -> tuple[list[int], list[int]]:
    """(connectivity seeds, activity seeds) per zone: srand(seed), then one
    rand() for each, zone by zone (simthread.cpp:13-18 after main.cpp:77-81)."""
    libc = C.CDLL("libc.so.6")
    libc.rand.restype = C.c_int
    libc.srand(C.c_uint(seed))
    cr, ar = [], []
    for _ in range(num_zones):
        cr.append(int(libc.rand()))
        ar.append(int(libc.rand()))
    return cr, ar


def _con_key(text: str, key: str, default: int) -> int:
    toks = text.split()
    for i, t in enumerate(toks[:-1]):
        if t == key:
            return int(toks[i + 1])
    return default


def build_state(act_text: str, con_text: str, seed: int, num_zones: int = 2):
    """The network a driver run with `seed` builds: CBMState(actPF, conPF,
    numMZ, seed, crSeeds, arSeeds) -- the reference's glomerular generator when
    the conParams text has a glomerulus layer, else the synthetic one with the
    fan-ins the C++ CBMState takes from the file: 4 MF and 3 GO inputs per GR,
    min(2, maxnumpGRfromGRtoGO) GO targets, maxnumpGOfromGLtoGO MF inputs per
    GO (16 when absent) -- and CBMSimCore's zone generators seeded seed,
    seed + 1, ..."""
    from .network import NetState
    cr, _ar = simthread_seeds(seed, num_zones)
    glomerular = _con_key(con_text, "glXP2", -1) >= 0
    return NetState(act_text, con_text, num_zones=num_zones, innet_seed=seed, mzone_con_seeds=cr,
                    mzone_rng_seeds=[seed + z for z in range(num_zones)], n_mf_per_gr=4, n_go_per_gr=3,
                    n_go_out_per_gr=min(2, _con_key(con_text, "maxnumpGRfromGRtoGO", 2)),
                    n_mf_per_go=_con_key(con_text, "maxnumpGOfromGLtoGO", 16),
                    generator="reference" if glomerular else "uniform")


class Cartpole:
    """A cart-pole on a network (net: Network with at least two zones), or
    detached (net None, sizes given) for hosts that step the network
    themselves."""

    def __init__(self, net=None, *, sizes=None, env_seed: int = 0, mf_seed: int | None = None,
                 max_num_trials: int = 20, max_trial_len: int = 1_000_000, track_len: float = 0.0,
                 closed_loop: bool = False, logfile: str | None = None, sim_state_dir: str | None = None,
                 verbose: bool = False):
        L = _abi.core_lib()
        s = net.sizes if net is not None else sizes
        if s is None:
            raise ValueError("give a network or its sizes")
        o = CartpoleOpts()
        o.maxNumTrials, o.closedLoop, o.maxTrialLen = max_num_trials, int(closed_loop), max_trial_len
        o.trackLen, o.verbose = track_len, int(verbose)
        self._logfile = (logfile or "").encode()
        self._statedir = (sim_state_dir or "").encode()
        o.logfile, o.simStateDir = self._logfile, self._statedir
        h = C.c_void_p()
        _check(L.cbm_cartpole_create(net._h if net is not None else None, C.byref(s), env_seed,
                                     env_seed if mf_seed is None else mf_seed, C.byref(o), C.byref(h)),
               "cbm_cartpole_create")
        self._c = h
        self.net = net
        self.num_mf, self.num_nc = s.numMF, s.numNC
        self.closed_loop = closed_loop

    def close(self) -> None:
        if self._c:
            _abi.core_lib().cbm_cartpole_destroy(self._c)
            self._c = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, n: int, trace: bool = True) -> np.ndarray | int:
        """Up to n driver steps; the trace rows of the steps run, or their count."""
        L = _abi.core_lib()
        done = C.c_int()
        buf = np.zeros((n, len(COLS)), np.float32) if trace else None
        _check(L.cbm_cartpole_run(self._c, n, buf.ctypes.data if trace else None, C.byref(done)),
               "cbm_cartpole_run")
        return buf[:done.value] if trace else done.value

    @property
    def terminated(self) -> bool:
        return bool(_abi.core_lib().cbm_cartpole_terminated(self._c))

    # ------------------------------------------------------------ detached
    def get_state(self) -> np.ndarray:
        out = np.empty(self.num_mf, np.float32)
        _check(_abi.core_lib().cbm_cartpole_get_state(self._c, out.ctypes.data), "cbm_cartpole_get_state")
        return out

    def mf_spikes(self, freq: np.ndarray) -> np.ndarray:
        f = np.ascontiguousarray(freq, np.float32)
        out = np.empty(self.num_mf, np.uint8)
        _check(_abi.core_lib().cbm_cartpole_mf_spikes(self._c, f.ctypes.data, out.ctypes.data),
               "cbm_cartpole_mf_spikes")
        return out

    def step(self, nc_left=None, nc_right=None) -> tuple[np.ndarray, tuple[int, int]]:
        row = np.zeros(len(COLS), np.float32)
        err = (C.c_int32 * 2)()
        a = None if nc_left is None else np.ascontiguousarray(nc_left, np.uint8)
        b = None if nc_right is None else np.ascontiguousarray(nc_right, np.uint8)
        _check(_abi.core_lib().cbm_cartpole_step(self._c, None if a is None else a.ctypes.data,
                                                 None if b is None else b.ctypes.data, row.ctypes.data, err),
               "cbm_cartpole_step")
        return row, (err[0], err[1])

    def mf_groups(self) -> np.ndarray:
        out = np.empty(self.num_mf, np.int32)
        _check(_abi.core_lib().cbm_cartpole_mf_groups(self._c, out.ctypes.data), "cbm_cartpole_mf_groups")
        return out

    def mf_rates(self) -> tuple[np.ndarray, np.ndarray]:
        lo, hi = np.empty(self.num_mf, np.float32), np.empty(self.num_mf, np.float32)
        _check(_abi.core_lib().cbm_cartpole_mf_rates(self._c, lo.ctypes.data, hi.ctypes.data), "cbm_cartpole_mf_rates")
        return lo, hi

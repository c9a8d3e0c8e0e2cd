"""ctypes wrapper of the CPU oracle (oracle/build/liborc.so).

TEST INFRASTRUCTURE: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import this module, as the checker or the timed CPU baseline.
It mirrors cerebellum_simulation_amd.Network so that a test drives the oracle
and the GPU path with the same calls.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from cerebellum_simulation_amd import _abi
from cerebellum_simulation_amd._abi import FIELDS, Options, field_info
from cerebellum_simulation_amd.network import field_dtype, field_shape

HERE = os.path.dirname(os.path.abspath(__file__))
_libs: dict[str, C.CDLL] = {}

# run-time fields that a fresh simulation starts from zero (never copied from a state)
SKIP_ON_LOAD = {"ZONE_RNG_STATE", "CUR_TIME", "SUM_GR_INPUT_GO", "SUM_GO_INPUT_GO", "PIPE_PFPC_SUM",
                "PIPE_PFBC_SUM", "PIPE_PFSC_SUM", "PIPE_GRGO_SUM", "MF_INPUT"}


def lib(openmp: bool = False, plant: bool = False) -> C.CDLL:
    """The oracle library: the single-threaded checker, the OpenMP build, or
    the checker with planted mistakes (orc_set_plant)."""
    name = "liborc_plant.so" if plant else ("liborc_omp.so" if openmp else "liborc.so")
    if name not in _libs:
        path = os.path.join(HERE, "build", name)
        if not os.path.exists(path):
            raise RuntimeError(f"{path} missing: run `make -C oracle`")
        L = C.CDLL(path)
        P, I, S, V = C.c_void_p, C.c_int, C.c_size_t, C.c_void_p
        L.orc_create.argtypes = [C.POINTER(_abi.Sizes), C.POINTER(_abi.ActParams), C.POINTER(Options), I,
                                 C.POINTER(P)]
        L.orc_destroy.argtypes = [P]
        L.orc_write.argtypes = [P, I, I, V, S]
        L.orc_read.argtypes = [P, I, I, V, S]
        L.orc_read_f64.argtypes = [P, I, I, V, S]
        L.orc_commit.argtypes = [P]
        L.orc_set_mf.argtypes = [P, V]
        L.orc_set_err_drive.argtypes = [P, I, C.c_float]
        L.orc_set_grpc_steps.argtypes = [P, I, C.c_float, C.c_float]
        L.orc_set_ginc_grgo.argtypes = [P, C.c_float]
        L.orc_step.argtypes = [P, I]
        L.orc_run_mf_train.argtypes = [P, V, S, I, V]
        L.orc_threads.restype = I
        L.orc_set_plant.argtypes = [I]
        L.orc_canonical_sum_f32.argtypes = [V, S]
        L.orc_canonical_sum_f32.restype = C.c_float
        _libs[name] = L
    return _libs[name]


def _check(rc: int, what: str) -> None:
    if rc:
        raise RuntimeError(f"oracle {what} failed ({rc})")


class Oracle:
    """The reference step restated on the CPU; precision 32 (mimic) or 64 (replay)."""

    def __init__(self, sizes, params, *, precision: int = 32, grgo_lag: int = 2, openmp: bool = False,
                 plant: int = 0):
        """plant=k > 0: the build with planted mistakes, with mistake k in place
        (a process-wide setting of that library: tests only)."""
        self.L = lib(openmp, plant=plant > 0)
        if plant:
            _check(self.L.orc_set_plant(plant), "set_plant")
        o = Options()
        o.grgoLag = grgo_lag
        h = C.c_void_p()
        _check(self.L.orc_create(C.byref(sizes), C.byref(params), C.byref(o), precision, C.byref(h)), "create")
        self._h = h
        self.sizes, self.params, self.precision = sizes, params, precision
        self._info = {}

    @classmethod
    def from_state(cls, state, **kw) -> "Oracle":
        orc = cls(state.sizes, state.params, **kw)
        for name, fid in FIELDS.items():
            if name in SKIP_ON_LOAD:
                continue
            info = field_info(fid, state.sizes)
            if info["kind"] == _abi.KIND_RUNTIME and name not in ("GINC_GR_TO_GO", "GRPC_PLAST_STEPS"):
                continue
            for z in range(state.num_zones if info["zoned"] else 1):
                arr = state.field(name, z)
                if arr.nbytes:
                    orc.write(name, arr, z)
        orc.commit()
        return orc

    def info(self, name):
        if name not in self._info:
            self._info[name] = field_info(FIELDS[name], self.sizes)
        return self._info[name]

    def read(self, name, zone=0):
        info = self.info(name)
        out = np.empty(field_shape(info), field_dtype(name, info))
        _check(self.L.orc_read(self._h, FIELDS[name], zone, out.ctypes.data, out.nbytes), f"read {name}")
        return out

    def read_f64(self, name, zone=0):
        info = self.info(name)
        out = np.empty(field_shape(info), np.float64)
        _check(self.L.orc_read_f64(self._h, FIELDS[name], zone, out.ctypes.data, out.size), f"read_f64 {name}")
        return out

    def write(self, name, value, zone=0):
        info = self.info(name)
        arr = np.ascontiguousarray(np.asarray(value, dtype=field_dtype(name, info)).reshape(field_shape(info)))
        _check(self.L.orc_write(self._h, FIELDS[name], zone, arr.ctypes.data, arr.nbytes), f"write {name}")

    def commit(self):
        _check(self.L.orc_commit(self._h), "commit")

    def update_mf_input(self, mf):
        arr = np.ascontiguousarray(mf, dtype=np.uint8)
        _check(self.L.orc_set_mf(self._h, arr.ctypes.data), "set_mf")

    def update_err_drive(self, zone, rel):
        _check(self.L.orc_set_err_drive(self._h, zone, rel), "set_err_drive")

    def set_grpc_plast_steps(self, zone, ltd, ltp):
        _check(self.L.orc_set_grpc_steps(self._h, zone, ltd, ltp), "set_grpc_steps")

    def reset_grpc_plast_steps(self, zone):
        self.set_grpc_plast_steps(zone, self.params.synLTDStepSizeGRtoPC, self.params.synLTPStepSizeGRtoPC)

    def set_ginc_gr_to_go(self, inc):
        _check(self.L.orc_set_ginc_grgo(self._h, inc), "set_ginc_grgo")

    def calc_activity(self, n=1):
        _check(self.L.orc_step(self._h, n), "step")

    def run_mf_train(self, train: np.ndarray, n=None, err_drive=None):
        train = np.ascontiguousarray(train, dtype=np.uint8)
        n = train.shape[0] if n is None else n
        err = None
        if err_drive is not None:
            ea = np.ascontiguousarray(err_drive, dtype=np.float32)
            err = ea.ctypes.data
        _check(self.L.orc_run_mf_train(self._h, train.ctypes.data, train.shape[1], n, err), "run_mf_train")

    def close(self):
        if getattr(self, "_h", None):
            self.L.orc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def canonical_sum_f32(x: np.ndarray) -> float:
    x = np.ascontiguousarray(x, dtype=np.float32)
    return float(lib().orc_canonical_sum_f32(x.ctypes.data, x.size))


# ---------------------------------------------------------------- input-net generator
_INNET_KEYS = ("glXP2", "glYP2", "grXP2", "grYP2", "goXP2", "goYP2", "numMFP2", "maxnumpGLfromGLtoGR",
               "lownumpGLfromGLtoGR", "maxnumpGLfromGLtoGO", "maxnumpGLfromGOtoGL", "maxnumpGRfromGLtoGR",
               "maxnumpGRfromGOtoGR", "maxnumpGRfromGRtoGO", "maxnumpGOfromGRtoGO", "maxnumpGOfromGLtoGO",
               "maxnumpGOfromGOtoGL", "spanGRDenOnGLX", "spanGRDenOnGLY", "spanGODecDenOnGLX", "spanGODecDenOnGLY",
               "spanGOAxonOnGLX", "spanGOAxonOnGLY", "spanGOAscDenOnGRX", "spanGOAscDenOnGRY")
INNET_LISTS = ("pGLfromGLtoGO", "pGLfromGOtoGL", "pGLfromGLtoGR", "pMFfromMFtoGL", "pMFfromMFtoGR", "pMFfromMFtoGO",
               "pGOfromGLtoGO", "pGOfromGOtoGL", "pGOfromMFtoGO", "pGOfromGOtoGR", "pGOfromGRtoGO", "pGRfromGLtoGR",
               "pGRfromGRtoGO", "pGRfromGOtoGR", "pGRfromMFtoGR")


class _InnetParams(C.Structure):
    _fields_ = [(k, C.c_uint32) for k in _INNET_KEYS]


def innet_generate(con: dict, seed: int) -> dict:
    """The oracle's input-net generator (orc_innet.c) on the conParams keys in
    `con`: {list name: (counts int32 [rows], slots uint32 [rows, cols])},
    plus "haspGLfromMFtoGL", "pGLfromMFtoGL" and "dropped"."""
    L = lib()
    L.orc_innet_generate.argtypes = [C.POINTER(_InnetParams), C.c_int, C.POINTER(C.c_void_p)]
    L.orc_innet_free.argtypes = [C.c_void_p]
    L.orc_innet_shape.argtypes = [C.c_void_p, C.c_char_p, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
    L.orc_innet_copy.argtypes = [C.c_void_p, C.c_char_p, C.c_void_p, C.c_void_p]
    L.orc_innet_dropped.argtypes = [C.c_void_p]
    L.orc_innet_dropped.restype = C.c_uint64
    p = _InnetParams(*[int(con[k]) for k in _INNET_KEYS])
    h = C.c_void_p()
    rc = L.orc_innet_generate(C.byref(p), int(seed), C.byref(h))
    try:
        _check(rc, "orc_innet_generate")
        out = {}
        r, c = C.c_uint32(), C.c_uint32()
        for name in INNET_LISTS:
            L.orc_innet_shape(h, name.encode(), C.byref(r), C.byref(c))
            num = np.empty(r.value, np.int32)
            slots = np.empty((r.value, c.value), np.uint32)
            L.orc_innet_copy(h, name.encode(), num.ctypes.data, slots.ctypes.data)
            out[name] = (num, slots)
        L.orc_innet_shape(h, b"pGLfromMFtoGL", C.byref(r), C.byref(c))
        has = np.empty(r.value, np.uint8)
        own = np.empty(r.value, np.uint32)
        L.orc_innet_copy(h, b"haspGLfromMFtoGL", None, has.ctypes.data)
        L.orc_innet_copy(h, b"pGLfromMFtoGL", None, own.ctypes.data)
        out["haspGLfromMFtoGL"], out["pGLfromMFtoGL"] = has, own
        out["dropped"] = int(L.orc_innet_dropped(h))
        return out
    finally:
        L.orc_innet_free(h)

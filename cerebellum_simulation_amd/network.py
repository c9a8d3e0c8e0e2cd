"""Python host mirror of the network step: ``NetState`` (CBMState side) and
``Network`` (the CBMSimCore-like handle over the HIP C-ABI).

The method names follow the reference's CBMSimCore / InNetInterface /
MZoneInterface (CBM_CORE_LIB_MGPU_DEV/CBMCoreInclude/interface/*.h):
``calc_activity`` is ``calcActivity``, ``update_mf_input`` is
``updateMFInput``, ``export_*`` are the export accessors.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _abi
from ._abi import FIELDS, ActParams, GenSpec, Options, Sizes, check, field_info

_SIGNED = {"NUMP_GR_FROM_MF_TO_GR", "NUMP_GR_FROM_GO_TO_GR", "NUMP_GR_FROM_GR_TO_GO",
           "NUMP_MF_FROM_MF_TO_GO", "NUMP_GO_GABA_OUT_GOGO", "NUMP_GO_COUP_IN_GOGO",
           "PFPC_PLAST_TIMER_IO", "ZONE_RNG_SEED"}


def field_dtype(name: str, info: dict) -> np.dtype:
    if info["elem"] == 1:
        return np.dtype(np.uint8)
    if info["elem"] == 8:
        return np.dtype(np.uint64)
    if info["is_float"]:
        return np.dtype(np.float32)
    return np.dtype(np.int32 if name in _SIGNED else np.uint32)


def field_shape(info: dict) -> tuple:
    return (info["rows"], info["cols"]) if info["cols"] != 1 else (info["rows"],)


class NetState:
    """A complete network description built by libcbm_state from actParams /
    conParams texts and a generator spec (CBMState's role)."""

    def __init__(self, act_text: str, con_text: str, *, num_zones: int = 1, innet_seed: int = 0,
                 n_mf_per_gr: int = 4, n_go_per_gr: int = 3, n_go_out_per_gr: int = 2,
                 n_mf_per_go: int = 16, grgo_window: tuple[int, int] = (0, 0),
                 mzone_con_seeds=None, mzone_rng_seeds=None, generator: str = "uniform"):
        """generator "uniform": the synthetic input net of the n_* fan-ins;
        "reference": the reference's glomerular generator (CBM_GEN_REFERENCE),
        sized by the conParams text's glomerulus keys."""
        L = _abi.state_lib()
        g = GenSpec()
        if generator not in ("uniform", "reference"):
            raise ValueError("generator must be 'uniform' or 'reference'")
        g.kind = 0 if generator == "uniform" else 1
        g.numZones = num_zones
        g.innetSeed = innet_seed
        g.nMFperGR, g.nGOperGR, g.nGOoutPerGR, g.nMFperGO = n_mf_per_gr, n_go_per_gr, n_go_out_per_gr, n_mf_per_go
        g.grgoWindowX, g.grgoWindowY = grgo_window
        for z in range(8):
            g.mzoneConSeed[z] = mzone_con_seeds[z] if mzone_con_seeds and z < len(mzone_con_seeds) \
                else innet_seed + 101 + z
            g.mzoneRngSeed[z] = mzone_rng_seeds[z] if mzone_rng_seeds and z < len(mzone_rng_seeds) \
                else innet_seed + 1001 + z
        h = C.c_void_p()
        rc = L.cbm_state_create(act_text.encode(), con_text.encode(), C.byref(g), C.byref(h))
        if rc:
            raise _abi.CbmError(f"cbm_state_create failed ({rc}): {L.cbm_state_last_error().decode()}")
        self._h = h
        self.sizes = Sizes()
        self.params = ActParams()
        L.cbm_state_sizes(h, C.byref(self.sizes))
        L.cbm_state_params(h, C.byref(self.params))
        self.num_zones = num_zones

    # ------------------------------------------------------------ checkpoints
    def write_checkpoint(self, path: str) -> None:
        """Save in the CBMState file format (CBMState::writeState layout)."""
        L = _abi.state_lib()
        rc = L.cbm_state_write_checkpoint(self._h, str(path).encode())
        if rc:
            raise _abi.CbmError(f"cbm_state_write_checkpoint failed ({rc}): {L.cbm_state_last_error().decode()}")

    def checkpoint_bytes(self) -> bytes:
        L = _abi.state_lib()
        d, n = C.c_void_p(), C.c_size_t()
        rc = L.cbm_state_checkpoint_to_memory(self._h, C.byref(d), C.byref(n))
        if rc:
            raise _abi.CbmError(f"cbm_state_checkpoint_to_memory failed ({rc}): {L.cbm_state_last_error().decode()}")
        try:
            return C.string_at(d, n.value)
        finally:
            L.cbm_state_free_memory(d)

    @classmethod
    def read_checkpoint(cls, path: str) -> "NetState":
        """A state saved in the CBMState file format (CBMState(fstream&))."""
        L = _abi.state_lib()
        h = C.c_void_p()
        rc = L.cbm_state_read_checkpoint(str(path).encode(), C.byref(h))
        if rc:
            raise _abi.CbmError(f"cbm_state_read_checkpoint failed ({rc}): {L.cbm_state_last_error().decode()}")
        return cls._wrap(h)

    @classmethod
    def from_checkpoint_bytes(cls, data: bytes) -> "NetState":
        L = _abi.state_lib()
        h = C.c_void_p()
        used = C.c_size_t()
        rc = L.cbm_state_checkpoint_from_memory(data, len(data), C.byref(h), C.byref(used))
        if rc:
            raise _abi.CbmError(f"cbm_state_checkpoint_from_memory failed ({rc}): {L.cbm_state_last_error().decode()}")
        return cls._wrap(h)

    @classmethod
    def _wrap(cls, h) -> "NetState":
        L = _abi.state_lib()
        self = cls.__new__(cls)
        self._h = h
        self.sizes = Sizes()
        self.params = ActParams()
        L.cbm_state_sizes(h, C.byref(self.sizes))
        L.cbm_state_params(h, C.byref(self.params))
        self.num_zones = self.sizes.numZones
        return self

    def field(self, name: str, zone: int = 0) -> np.ndarray:
        """numpy view (no copy) of a field in the CBMState layout."""
        info = field_info(FIELDS[name], self.sizes)
        ptr, nb = C.c_void_p(), C.c_size_t()
        L = _abi.state_lib()
        rc = L.cbm_state_field(self._h, FIELDS[name], zone, C.byref(ptr), C.byref(nb))
        if rc:
            raise _abi.CbmError(L.cbm_state_last_error().decode())
        dt = field_dtype(name, info)
        n = nb.value // dt.itemsize
        if n == 0 or not ptr.value:
            return np.zeros(field_shape(info), dt)
        buf = (C.c_uint8 * nb.value).from_address(ptr.value)
        buf._owner = self  # the view keeps the state alive
        return np.frombuffer(buf, dtype=dt, count=n).reshape(field_shape(info))

    def load_into(self, net: "Network") -> None:
        L = _abi.state_lib()
        rc = L.cbm_state_load_into(self._h, net._h)
        if rc:
            raise _abi.CbmError(f"cbm_state_load_into failed: {L.cbm_state_last_error().decode()}")

    def close(self) -> None:
        if getattr(self, "_h", None):
            _abi.state_lib().cbm_state_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def act_params_from_text(text: str) -> ActParams:
    p = ActParams()
    L = _abi.state_lib()
    rc = L.cbm_act_params_from_text(text.encode(), C.byref(p))
    if rc:
        raise _abi.CbmError(L.cbm_state_last_error().decode())
    return p


class Network:
    """One network on one GPU behind the cbm_* C-ABI (CBMSimCore's role)."""

    def __init__(self, sizes: Sizes, params: ActParams, *, grgo_lag: int = 2, device: int = 0,
                 cells_per_thread: int = 0, block_threads: int = 0, step_mode: int = 0, conn_format: int = 0):
        """step_mode: 0 = pick (1, with resident run_mf_train runs where the
        network allows them), 1 = one fused launch per step, 2 = GR and
        small-population kernels on two streams, 3 = fused with GR tiles of
        1024 x 2 cells, 4 = fused with the GR cells split evenly over the CUs
        (mode 1 picks the even split when each block gets 1,024 to 5,120
        cells), 5 = as 4, with run_mf_train runs of two or more steps
        resident: one launch per run, GR state in registers, see
        `resident` (cbm_options.stepMode)."""
        L = _abi.hip_lib()
        o = Options()
        o.grgoLag, o.device, o.grCellsPerThread = grgo_lag, device, cells_per_thread
        o.grBlockThreads, o.stepMode, o.connFormat = block_threads, step_mode, conn_format
        h = C.c_void_p()
        check(L.cbm_create(C.byref(sizes), C.byref(params), C.byref(o), C.byref(h)), "cbm_create")
        self._h = h
        self.sizes = sizes
        self.params = params
        self._info = {}

    @classmethod
    def from_state(cls, state: NetState, **kw) -> "Network":
        net = cls(state.sizes, state.params, **kw)
        state.load_into(net)
        return net

    # ------------------------------------------------------------ fields
    def info(self, name: str) -> dict:
        if name not in self._info:
            self._info[name] = field_info(FIELDS[name], self.sizes)
        return self._info[name]

    def read(self, name: str, zone: int = 0) -> np.ndarray:
        info = self.info(name)
        out = np.empty(field_shape(info), field_dtype(name, info))
        check(_abi.hip_lib().cbm_read(self._h, FIELDS[name], zone, out.ctypes.data, out.nbytes), f"cbm_read {name}")
        return out

    def write(self, name: str, value, zone: int = 0) -> None:
        info = self.info(name)
        arr = np.ascontiguousarray(np.asarray(value, dtype=field_dtype(name, info)).reshape(field_shape(info)))
        check(_abi.hip_lib().cbm_write(self._h, FIELDS[name], zone, arr.ctypes.data, arr.nbytes), f"cbm_write {name}")

    def commit(self) -> None:
        check(_abi.hip_lib().cbm_commit(self._h), "cbm_commit")

    # ------------------------------------------------------------ step API
    def update_mf_input(self, mf) -> None:
        arr = np.ascontiguousarray(mf, dtype=np.uint8)
        assert arr.size == self.sizes.numMF
        check(_abi.hip_lib().cbm_set_mf(self._h, arr.ctypes.data), "cbm_set_mf")

    def update_err_drive(self, zone: int, rel: float) -> None:
        check(_abi.hip_lib().cbm_set_err_drive(self._h, zone, rel), "cbm_set_err_drive")

    def set_grpc_plast_steps(self, zone: int, ltd: float, ltp: float) -> None:
        check(_abi.hip_lib().cbm_set_grpc_steps(self._h, zone, ltd, ltp), "cbm_set_grpc_steps")

    def reset_grpc_plast_steps(self, zone: int) -> None:
        check(_abi.hip_lib().cbm_reset_grpc_steps(self._h, zone), "cbm_reset_grpc_steps")

    def set_ginc_gr_to_go(self, inc: float) -> None:
        check(_abi.hip_lib().cbm_set_ginc_grgo(self._h, inc), "cbm_set_ginc_grgo")

    def reset_ginc_gr_to_go(self) -> None:
        check(_abi.hip_lib().cbm_reset_ginc_grgo(self._h), "cbm_reset_ginc_grgo")

    def calc_activity(self, n: int = 1) -> None:
        check(_abi.hip_lib().cbm_step(self._h, n), "cbm_step")

    def run_mf_train(self, train, n: int | None = None, err_drive=None) -> None:
        """n steps over a device-resident uint8 [steps, numMF] tensor."""
        ptr, pitch, rows = _device_rows(train)
        n = rows if n is None else n
        err = None
        if err_drive is not None:
            err_arr = np.ascontiguousarray(err_drive, dtype=np.float32)
            err = err_arr.ctypes.data
        check(_abi.hip_lib().cbm_run_mf_train(self._h, ptr, pitch, n, err), "cbm_run_mf_train")

    def profile_steps(self, train, n: int | None = None) -> tuple[float, float]:
        """(mean duration of the step's dominant kernel, mean step time) in us."""
        ptr, pitch, rows = _device_rows(train)
        n = rows if n is None else n
        gr, st = C.c_double(), C.c_double()
        check(_abi.hip_lib().cbm_profile_steps(self._h, ptr, pitch, n, C.byref(gr), C.byref(st)),
              "cbm_profile_steps")
        return gr.value, st.value

    def profile_split(self, train, n: int | None = None) -> tuple[float, float]:
        """(GR kernel, small-population kernel) mean durations in us, run one
        after the other on one stream."""
        ptr, pitch, rows = _device_rows(train)
        n = rows if n is None else n
        gr, sm = C.c_double(), C.c_double()
        check(_abi.hip_lib().cbm_profile_split(self._h, ptr, pitch, n, C.byref(gr), C.byref(sm)),
              "cbm_profile_split")
        return gr.value, sm.value

    REC = {"PC": 0, "BC": 1, "SC": 2, "IO": 3, "NC": 4, "GO": 5, "PF": 6}

    def set_recorder(self, populations, capacity: int) -> None:
        """Record the spikes of the named populations ("PC", "NC", ...) of
        every following step into a device ring of `capacity` steps."""
        mask = 0
        for p in populations:
            mask |= 1 << self.REC[p]
        check(_abi.hip_lib().cbm_set_recorder(self._h, mask, capacity), "cbm_set_recorder")

    def recorded_cells(self, population: str) -> int:
        s, z = self.sizes, self.sizes.numZones
        return {"PC": z * s.numPC, "BC": z * s.numBC, "SC": s.numSC, "IO": z * s.numIO, "NC": z * s.numNC,
                "GO": s.numGO, "PF": s.numGR // 64}[population]

    def read_recorded(self, population: str, t0: int, n: int, out=None):
        """uint8 [n, cells] spikes of steps t0 .. t0+n-1 (zones stacked); "PF"
        gives the delayed PF spikes counted per 64-GR chunk.  `out` may be a
        contiguous uint8 device tensor of that shape, filled in place."""
        cells = self.recorded_cells(population)
        if out is None:
            out = np.empty((n, cells), np.uint8)
            ptr = out.ctypes.data
        else:
            if tuple(out.shape) != (n, cells) or not out.is_contiguous() or out.element_size() != 1:
                raise ValueError(f"out must be a contiguous byte tensor of shape {(n, cells)}")
            ptr = out.data_ptr()
        check(_abi.hip_lib().cbm_read_recorded(self._h, self.REC[population], t0, n, ptr), "cbm_read_recorded")
        return out

    @property
    def step_mode(self) -> int:
        m = _abi.hip_lib().cbm_step_mode(self._h)
        check(min(m, 0), "cbm_step_mode")
        return m

    @property
    def resident(self) -> bool:
        """run_mf_train runs of two or more steps are resident (cbm_resident)."""
        m = _abi.hip_lib().cbm_resident(self._h)
        check(min(m, 0), "cbm_resident")
        return m == 1

    @property
    def gr_launch(self) -> dict:
        """GR workgroups of the step launch: blocks, threads, cells_per_thread
        and even_split (cells split evenly over one block per CU)."""
        v = [C.c_uint32() for _ in range(4)]
        check(_abi.hip_lib().cbm_gr_launch(self._h, *(C.byref(x) for x in v)), "cbm_gr_launch")
        return dict(zip(("blocks", "threads", "cells_per_thread", "even_split"), (x.value for x in v)))

    def sync(self) -> None:
        check(_abi.hip_lib().cbm_sync(self._h), "cbm_sync")

    def stream_ptr(self) -> int:
        """The hipStream_t the network's step kernels run on (for events or
        torch.cuda.ExternalStream)."""
        p = _abi.hip_lib().cbm_stream(self._h)
        if not p:
            raise RuntimeError("cbm_stream: no stream (the network is not committed)")
        return p

    STEP_IO_POPULATIONS = ("PC", "BC", "SC", "IO", "NC", "GO")

    def step_io(self, mf: np.ndarray, population: str = "NC") -> np.ndarray:
        """One step on the MF bytes `mf` (host, numMF) and this step's spikes
        of `population` (zones one after another), through
        cbm_step_io_begin/_end: the call pattern of a closed loop, replacing
        updateMFInput, calcActivity and exportAPNC per zone
        (CBM_CORE_LIB_MGPU_DEV/src/interface/cbmsimcore.cpp:312, :139;
        Matthew_Sim_CBM/src/microzone.cpp:31-33)."""
        L = _abi.hip_lib()
        mf = np.ascontiguousarray(mf, dtype=np.uint8)
        if mf.size < self.sizes.numMF:
            raise ValueError(f"step_io: {mf.size} MF bytes, the network has {self.sizes.numMF}")
        k = self.STEP_IO_POPULATIONS.index(population)
        s = self.sizes
        n = {0: s.numZones * s.numPC, 1: s.numZones * s.numBC, 2: s.numSC, 3: s.numZones * s.numIO,
             4: s.numZones * s.numNC, 5: s.numGO}[k]
        out = np.empty(n, dtype=np.uint8)
        check(L.cbm_step_io_begin(self._h, mf.ctypes.data, k), "cbm_step_io_begin")
        check(L.cbm_step_io_end(self._h, out.ctypes.data), "cbm_step_io_end")
        return out

    def loop_begin(self, max_steps: int, population: str = "NC") -> None:
        """Start a closed loop (cbm_loop_begin): the step_io calls that follow
        go to one resident run of up to max_steps steps, with no launch or
        stream synchronisation per step.  population: PC, BC, IO or NC."""
        k = self.STEP_IO_POPULATIONS.index(population)
        check(_abi.hip_lib().cbm_loop_begin(self._h, max_steps, k), "cbm_loop_begin")

    def loop_end(self) -> int:
        """Stop the closed loop after the steps posted; their number."""
        n = C.c_int()
        check(_abi.hip_lib().cbm_loop_end(self._h, C.byref(n)), "cbm_loop_end")
        return n.value

    def device_bytes(self) -> int:
        b = C.c_size_t()
        check(_abi.hip_lib().cbm_device_bytes(self._h, C.byref(b)), "cbm_device_bytes")
        return b.value

    # ------------------------------------------------------------ exports
    def export_ap_gr(self):
        return self.read("AP_GR")

    def export_ap_nc(self, zone: int = 0):
        return self.read("AP_NC", zone)

    def export_ap_pc(self, zone: int = 0):
        return self.read("AP_PC", zone)

    def export_ap_io(self, zone: int = 0):
        return self.read("AP_IO", zone)

    def close(self) -> None:
        if getattr(self, "_h", None):
            _abi.hip_lib().cbm_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def stream_ceiling(cells: int, read_words: int = 7, write_words: int = 6, launches: int = 200,
                   device: int = 0) -> tuple[float, int]:
    """(mean launch duration in us, bytes per launch) of a hand-written pass
    that only streams the GR step's per-cell arrays over `cells` cells: the
    bandwidth a GR pass of that size can reach on this GPU."""
    us, nbytes = C.c_double(), C.c_uint64()
    check(_abi.hip_lib().cbm_stream_ceiling(device, cells, read_words, write_words, launches, C.byref(us),
                                            C.byref(nbytes)), "cbm_stream_ceiling")
    return us.value, nbytes.value


def _device_rows(train):
    """(pointer, pitch, rows) of a 2-D uint8 device tensor."""
    if hasattr(train, "data_ptr"):
        assert train.dtype.itemsize == 1 and train.dim() == 2 and train.is_cuda
        return train.data_ptr(), train.stride(0), train.shape[0]
    raise TypeError("expected a torch uint8 tensor on the GPU")


class Batch:
    """A ragged pack of networks stepped together, one launch per step
    (the cbm_batch_* C-ABI).  The networks stay usable on their own between
    batched runs and must outlive the batch."""

    def __init__(self, nets: list, *, cells_per_thread: int = 0, graph_steps: int = 0):
        L = _abi.hip_lib()
        o = _abi.BatchOptions()
        o.cellsPerThread, o.graphSteps = cells_per_thread, graph_steps
        arr = (C.c_void_p * len(nets))(*[n._h.value for n in nets])
        b = C.c_void_p()
        check(L.cbm_batch_create(arr, len(nets), C.byref(o), C.byref(b)), "cbm_batch_create")
        self._b = b
        self.nets = list(nets)

    def _trains(self, trains, n):
        if len(trains) != len(self.nets):
            raise ValueError("one MF train per network")
        rows = [_device_rows(t) for t in trains]
        n = min(r[2] for r in rows) if n is None else n
        ptrs = (C.c_void_p * len(rows))(*[r[0] for r in rows])
        pitch = (C.c_size_t * len(rows))(*[r[1] for r in rows])
        return ptrs, pitch, n

    def run_mf_train(self, trains, n: int | None = None, err_drive=None) -> None:
        """trains[i]: uint8 [steps, numMF_i] device tensor; err_drive[i]: None
        or float32 [steps, numZones_i] relative drive."""
        ptrs, pitch, n = self._trains(trains, n)
        err = None
        if err_drive is not None:
            keep = [None if e is None else np.ascontiguousarray(e, dtype=np.float32) for e in err_drive]
            err = (C.c_void_p * len(keep))(*[None if e is None else e.ctypes.data for e in keep])
        check(_abi.hip_lib().cbm_batch_run_mf_train(self._b, ptrs, pitch, n, err), "cbm_batch_run_mf_train")

    def profile(self, trains, n: int | None = None) -> float:
        """Mean time of one batched step in us over n steps."""
        ptrs, pitch, n = self._trains(trains, n)
        us = C.c_double()
        check(_abi.hip_lib().cbm_batch_profile(self._b, ptrs, pitch, n, C.byref(us)), "cbm_batch_profile")
        return us.value

    def autotune(self, trains, steps: int = 20) -> int:
        ptrs, pitch, _ = self._trains(trains, steps)
        c = C.c_int()
        check(_abi.hip_lib().cbm_batch_autotune(self._b, ptrs, pitch, steps, C.byref(c)), "cbm_batch_autotune")
        return c.value

    def sweep(self, prob, first: int, n: int, train, err_drive=None, cells_per_cluster: int = 256) -> None:
        """n steps of every network with its input layer replaced by cluster
        probabilities (cbm_batch_sweep): prob FP32 [rows, clusters] device
        tensor whose row `first` + k drives run step k (first >= 199 rows of
        look-back); train uint8 [>= n, numMF] device tensor shared by all
        networks; err_drive[i]: None or float32 [n, numZones_i]."""
        torch = __import__("torch")
        if first < 199 or first + n > prob.shape[0]:
            raise ValueError("prob needs 199 rows before `first` and n rows from it")
        if prob.dtype != torch.float32 or prob.stride(1) != 1 or not prob.is_cuda:
            raise TypeError("prob must be an FP32 device tensor with contiguous rows")
        ptr, pitch, rows = _device_rows(train)
        if rows < n:
            raise ValueError("MF train shorter than n")
        err = None
        if err_drive is not None:
            keep = [None if e is None else np.ascontiguousarray(e, dtype=np.float32) for e in err_drive]
            err = (C.c_void_p * len(keep))(*[None if e is None else e.ctypes.data for e in keep])
        torch.cuda.current_stream().synchronize()
        check(_abi.hip_lib().cbm_batch_sweep(self._b, prob[first].data_ptr(), prob.stride(0), n, ptr, pitch, err,
                                             cells_per_cluster), "cbm_batch_sweep")
        self._cells = cells_per_cluster

    def sweep_weights(self, net: int) -> np.ndarray:
        """Cluster PF->PC weights of network `net`, [numZones, clusters]."""
        h = self.nets[net]
        out = np.empty(h.sizes.numZones * self._clusters(net), np.float32)
        check(_abi.hip_lib().cbm_batch_sweep_weights(self._b, net, out.ctypes.data), "cbm_batch_sweep_weights")
        return out.reshape(h.sizes.numZones, -1)

    def _clusters(self, net: int) -> int:
        return self.nets[net].sizes.numGR // self._cells if getattr(self, "_cells", 0) else 0

    def sweep_reset(self) -> None:
        check(_abi.hip_lib().cbm_batch_sweep_reset(self._b), "cbm_batch_sweep_reset")

    def info(self) -> dict:
        c, blocks, g = C.c_int(), C.c_int(), C.c_int()
        check(_abi.hip_lib().cbm_batch_info(self._b, C.byref(c), C.byref(blocks), C.byref(g)), "cbm_batch_info")
        return {"cells_per_thread": c.value, "blocks": blocks.value, "graph_steps": g.value}

    def sync(self) -> None:
        check(_abi.hip_lib().cbm_batch_sync(self._b), "cbm_batch_sync")

    def close(self) -> None:
        if getattr(self, "_b", None) is not None and self._b.value:
            _abi.hip_lib().cbm_batch_destroy(self._b)
            self._b = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

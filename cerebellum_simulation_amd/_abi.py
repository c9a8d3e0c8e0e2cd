"""ctypes view of the two C-ABIs (include/cbm_hip.h, include/cbm_state.h).

The struct layouts and the field enum are read from the C header itself, so
the Python mirror cannot drift from the library it binds.
"""
from __future__ import annotations

import ctypes as C
import os
import re

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
HEADER = os.path.join(ROOT, "include", "cbm_hip.h")
LIBDIR = os.path.join(PKG, "lib")

_CTYPES = {"float": C.c_float, "uint32_t": C.c_uint32, "int32_t": C.c_int32}


def _strip_comments(text: str) -> str:
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return re.sub(r"//[^\n]*", "", text)


def _parse_struct(text: str, name: str) -> list[tuple[str, type]]:
    m = re.search(r"typedef struct %s \{(.*?)\} %s;" % (name, name), text, flags=re.S)
    if not m:
        raise RuntimeError(f"struct {name} not found in {HEADER}")
    fields = []
    for decl in m.group(1).split(";"):
        decl = " ".join(decl.split())
        if not decl:
            continue
        typ, rest = decl.split(" ", 1)
        for member in rest.split(","):
            fields.append((member.strip(), _CTYPES[typ]))
    return fields


def _parse_enum(text: str, name: str) -> dict[str, int]:
    m = re.search(r"enum %s \{(.*?)\};" % name, text, flags=re.S)
    if not m:
        raise RuntimeError(f"enum {name} not found in {HEADER}")
    out, nxt = {}, 0
    for item in m.group(1).split(","):
        item = item.strip()
        if not item:
            continue
        if "=" in item:
            k, v = (x.strip() for x in item.split("="))
            nxt = int(v, 0)
        else:
            k = item
        out[k] = nxt
        nxt += 1
    return out


_TEXT = _strip_comments(open(HEADER).read())


class Sizes(C.Structure):
    _fields_ = _parse_struct(_TEXT, "cbm_sizes")


class ActParams(C.Structure):
    _fields_ = _parse_struct(_TEXT, "cbm_act_params")


class Options(C.Structure):
    _fields_ = _parse_struct(_TEXT, "cbm_options")


class BatchOptions(C.Structure):
    _fields_ = _parse_struct(_TEXT, "cbm_batch_options")


class GenSpec(C.Structure):
    _fields_ = [("kind", C.c_int32), ("numZones", C.c_uint32), ("innetSeed", C.c_int32),
                ("nMFperGR", C.c_uint32), ("nGOperGR", C.c_uint32), ("nGOoutPerGR", C.c_uint32),
                ("nMFperGO", C.c_uint32), ("grgoWindowX", C.c_uint32), ("grgoWindowY", C.c_uint32),
                ("mzoneConSeed", C.c_int32 * 8), ("mzoneRngSeed", C.c_int32 * 8)]


FIELDS: dict[str, int] = {k[len("CBM_F_"):]: v for k, v in _parse_enum(_TEXT, "cbm_field").items()
                          if k != "CBM_F_COUNT"}
FIELD_COUNT = len(FIELDS)
FIELD_NAMES = {v: k for k, v in FIELDS.items()}
KIND_CON, KIND_STATE, KIND_RUNTIME = 0, 1, 2
# CBM_E_* status codes and the no-allocation batch run's staging limit
ERRORS: dict[str, int] = {m.group(1): int(m.group(2)) for m in
                          re.finditer(r"#define (CBM_E_\w+) \((-?\d+)\)", open(HEADER).read())}
NOALLOC_ERR_FLOATS = 1 << int(re.search(r"#define CBM_NOALLOC_ERR_FLOATS \(1u << (\d+)\)",
                                        open(HEADER).read()).group(1))

_hip = None
_state = None


class SgDesc(C.Structure):
    """cbm_sg_desc (include/cbm_surrogate.h)."""
    _fields_ = [("numIn", C.c_uint32), ("hidden", C.c_uint32), ("numOut", C.c_uint32), ("taps", C.c_uint32),
                ("dil", C.c_uint32 * 3), ("maxRows", C.c_uint32), ("device", C.c_int)]


class SgBatch(C.Structure):
    """cbm_sg_batch (include/cbm_surrogate.h)."""
    _fields_ = [("mf", C.c_void_p), ("mfPitch", C.c_size_t), ("segStart", C.c_void_p), ("numSeg", C.c_int),
                ("outputs", C.c_int), ("target", C.c_void_p), ("targetPitch", C.c_size_t)]


class SgAdam(C.Structure):
    _fields_ = [("lr", C.c_float), ("beta1", C.c_float), ("beta2", C.c_float), ("eps", C.c_float)]


def _load(name: str) -> C.CDLL:
    path = os.path.join(LIBDIR, name)
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with `python -m cerebellum_simulation_amd.build`")
    return C.CDLL(path)


def hip_lib() -> C.CDLL:
    """libcbm_hip.so with argument types declared.

    When PyTorch is installed it is imported first: its wheel carries its own
    HIP runtime under the same soname as /opt/rocm's, and the process must use
    one runtime for both (the one loaded first wins; torch only initialises
    on its own)."""
    global _hip
    if _hip is None:
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = _load("libcbm_hip.so")
        P, I, S, V = C.c_void_p, C.c_int, C.c_size_t, C.c_void_p
        L.cbm_abi_version.restype = I
        L.cbm_last_error.restype = C.c_char_p
        L.cbm_device_count.restype = I
        L.cbm_create.argtypes = [C.POINTER(Sizes), C.POINTER(ActParams), C.POINTER(Options), C.POINTER(P)]
        L.cbm_destroy.argtypes = [P]
        L.cbm_write.argtypes = [P, I, I, V, S]
        L.cbm_read.argtypes = [P, I, I, V, S]
        L.cbm_commit.argtypes = [P]
        L.cbm_set_mf.argtypes = [P, V]
        L.cbm_set_err_drive.argtypes = [P, I, C.c_float]
        L.cbm_set_grpc_steps.argtypes = [P, I, C.c_float, C.c_float]
        L.cbm_reset_grpc_steps.argtypes = [P, I]
        L.cbm_set_ginc_grgo.argtypes = [P, C.c_float]
        L.cbm_reset_ginc_grgo.argtypes = [P]
        L.cbm_step.argtypes = [P, I]
        L.cbm_run_mf_train.argtypes = [P, V, S, I, V]
        L.cbm_sync.argtypes = [P]
        L.cbm_step_io_begin.argtypes = [P, V, I]
        L.cbm_step_io_end.argtypes = [P, V]
        L.cbm_loop_begin.argtypes = [P, I, I]
        L.cbm_loop_end.argtypes = [P, C.POINTER(C.c_int)]
        L.cbm_stream_ceiling.argtypes = [I, C.c_uint64, I, I, I, C.POINTER(C.c_double), C.POINTER(C.c_uint64)]
        L.cbm_stream.argtypes = [P]
        L.cbm_stream.restype = P
        L.cbm_profile_steps.argtypes = [P, V, S, I, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.cbm_profile_split.argtypes = [P, V, S, I, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.cbm_step_mode.argtypes = [P]
        L.cbm_resident.argtypes = [P]
        U32P = C.POINTER(C.c_uint32)
        L.cbm_gr_launch.argtypes = [P, U32P, U32P, U32P, U32P]
        L.cbm_set_recorder.argtypes = [P, C.c_uint32, C.c_uint32]
        L.cbm_read_recorded.argtypes = [P, I, C.c_uint64, I, V]
        PP = C.POINTER(C.c_void_p)
        SP = C.POINTER(C.c_size_t)
        L.cbm_batch_create.argtypes = [PP, I, C.POINTER(BatchOptions), C.POINTER(P)]
        L.cbm_batch_destroy.argtypes = [P]
        L.cbm_batch_run_mf_train.argtypes = [P, PP, SP, I, PP]
        L.cbm_batch_run_mf_train_noalloc.argtypes = [P, PP, SP, I, PP]
        L.cbm_batch_sync.argtypes = [P]
        L.cbm_batch_profile.argtypes = [P, PP, SP, I, C.POINTER(C.c_double)]
        L.cbm_batch_autotune.argtypes = [P, PP, SP, I, C.POINTER(I)]
        L.cbm_batch_info.argtypes = [P, C.POINTER(I), C.POINTER(I), C.POINTER(I)]
        L.cbm_device_bytes.argtypes = [P, C.POINTER(C.c_size_t)]
        L.cbm_field_info.argtypes = [I, C.POINTER(Sizes), C.POINTER(S), C.POINTER(S), C.POINTER(S),
                                     C.POINTER(I), C.POINTER(I), C.POINTER(I)]
        # learned surrogate (cbm_surrogate.h)
        F = C.POINTER(C.c_float)
        L.cbm_sg_create.argtypes = [C.POINTER(SgDesc), C.POINTER(P)]
        L.cbm_sg_destroy.argtypes = [P]
        L.cbm_sg_lead.argtypes = [P]
        L.cbm_sg_num_params.argtypes = [P, C.POINTER(S)]
        L.cbm_sg_get_params.argtypes = [P, V]
        L.cbm_sg_set_params.argtypes = [P, V]
        L.cbm_sg_init_params.argtypes = [P, C.c_uint64, C.c_float]
        L.cbm_sg_forward.argtypes = [P, C.POINTER(SgBatch), V, S]
        L.cbm_sg_loss_grad.argtypes = [P, C.POINTER(SgBatch), F, V]
        L.cbm_sg_train_step.argtypes = [P, C.POINTER(SgBatch), C.POINTER(SgAdam), V]
        L.cbm_sg_read_hidden.argtypes = [P, I, V, S]
        L.cbm_sg_targets_from_pf.argtypes = [P, V, S, I, C.c_uint32, C.c_uint32, V, S]
        L.cbm_sg_sync.argtypes = [P]
        L.cbm_sg_last_timing.argtypes = [P, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.cbm_batch_sweep.argtypes = [P, V, S, I, V, S, PP, C.c_uint32]
        L.cbm_batch_sweep_weights.argtypes = [P, I, V]
        L.cbm_batch_sweep_reset.argtypes = [P]
        want = int(re.search(r"#define CBM_ABI_VERSION (\d+)", open(HEADER).read()).group(1))
        if L.cbm_abi_version() != want:
            raise RuntimeError(f"libcbm_hip.so ABI version {L.cbm_abi_version()} != header's {want}: rebuild")
        _hip = L
    return _hip


def state_lib() -> C.CDLL:
    """libcbm_state.so with argument types declared."""
    global _state
    if _state is None:
        L = _load("libcbm_state.so")
        P, I, V = C.c_void_p, C.c_int, C.c_void_p
        L.cbm_state_last_error.restype = C.c_char_p
        L.cbm_act_params_from_text.argtypes = [C.c_char_p, C.POINTER(ActParams)]
        L.cbm_state_create.argtypes = [C.c_char_p, C.c_char_p, C.POINTER(GenSpec), C.POINTER(P)]
        L.cbm_state_destroy.argtypes = [P]
        L.cbm_state_sizes.argtypes = [P, C.POINTER(Sizes)]
        L.cbm_state_params.argtypes = [P, C.POINTER(ActParams)]
        L.cbm_state_field.argtypes = [P, I, I, C.POINTER(V), C.POINTER(C.c_size_t)]
        L.cbm_state_load_into.argtypes = [P, P]
        L.cbm_sfmt_sample.argtypes = [I, C.c_uint64, I, I, I, I, V]
        L.cbm_state_write_checkpoint.argtypes = [P, C.c_char_p]
        L.cbm_state_read_checkpoint.argtypes = [C.c_char_p, C.POINTER(P)]
        L.cbm_state_checkpoint_to_memory.argtypes = [P, C.POINTER(V), C.POINTER(C.c_size_t)]
        L.cbm_state_checkpoint_from_memory.argtypes = [V, C.c_size_t, C.POINTER(P), C.POINTER(C.c_size_t)]
        L.cbm_state_free_memory.argtypes = [V]
        L.cbm_state_free_memory.restype = None
        _state = L
    return _state


class CartpoleOpts(C.Structure):
    """cbm_cartpole_opts (include/cbm_env.h)."""
    _fields_ = [("maxNumTrials", C.c_int32), ("closedLoop", C.c_int32), ("maxTrialLen", C.c_int64),
                ("trackLen", C.c_float), ("verbose", C.c_int32), ("logfile", C.c_char_p),
                ("simStateDir", C.c_char_p)]


_core = None


def core_lib() -> C.CDLL:
    """libcbm_core.so (the C++ API and the environments of include/cbm_env.h)
    with the cbm_env.h argument types declared."""
    global _core
    if _core is None:
        hip_lib()
        state_lib()
        L = _load("libcbm_core.so")
        P, I, V = C.c_void_p, C.c_int, C.c_void_p
        L.cbm_env_last_error.restype = C.c_char_p
        L.cbm_cartpole_create.argtypes = [P, C.POINTER(Sizes), I, I, C.POINTER(CartpoleOpts), C.POINTER(P)]
        L.cbm_cartpole_destroy.argtypes = [P]
        L.cbm_cartpole_run.argtypes = [P, I, V, C.POINTER(I)]
        L.cbm_cartpole_terminated.argtypes = [P]
        L.cbm_cartpole_get_state.argtypes = [P, V]
        L.cbm_cartpole_mf_spikes.argtypes = [P, V, V]
        L.cbm_cartpole_step.argtypes = [P, V, V, V, V]
        L.cbm_cartpole_mf_groups.argtypes = [P, V]
        L.cbm_cartpole_mf_rates.argtypes = [P, V, V]
        _core = L
    return _core


class CbmError(RuntimeError):
    """A cbm_* call returned an error code."""


def check(rc: int, where: str, lib: C.CDLL | None = None) -> None:
    if rc != 0:
        L = lib or hip_lib()
        if L is _core:
            getter = L.cbm_env_last_error
        else:
            getter = L.cbm_state_last_error if hasattr(L, "cbm_state_last_error") and L is _state else L.cbm_last_error
        msg = getter().decode(errors="replace")
        raise CbmError(f"{where} failed ({rc}): {msg}")


def field_info(field: int, sizes: Sizes) -> dict:
    """Shape of a field: elem bytes, rows, cols, zoned, kind, is_float."""
    e, r, c = C.c_size_t(), C.c_size_t(), C.c_size_t()
    z, k, f = C.c_int(), C.c_int(), C.c_int()
    check(hip_lib().cbm_field_info(field, C.byref(sizes), C.byref(e), C.byref(r), C.byref(c),
                                   C.byref(z), C.byref(k), C.byref(f)), "cbm_field_info")
    return dict(elem=e.value, rows=r.value, cols=c.value, zoned=bool(z.value), kind=k.value,
                is_float=bool(f.value))

"""The C-ABI libraries load and export every declared entry point; the
Python struct mirrors match the C layouts; the product path refuses to run
without a GPU instead of falling back to the CPU."""
import ctypes as C
import os
import re
import subprocess

import pytest

from cerebellum_simulation_amd import _abi, configs, hip_lib, state_lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    text = _abi._strip_comments(open(header).read())
    return sorted(set(re.findall(r"^\s*(?:const char \*|void \*|int|float)\s*(\w+)\s*\(", text, flags=re.M)))


@pytest.mark.parametrize("header,libname", [("include/cbm_hip.h", "libcbm_hip.so"),
                                            ("include/cbm_surrogate.h", "libcbm_hip.so"),
                                            ("include/cbm_state.h", "libcbm_state.so"),
                                            ("include/cbm_env.h", "libcbm_core.so")])
def test_exports(header, libname):
    names = [n for n in _declared(os.path.join(ROOT, header)) if n.startswith("cbm_")]
    assert len(names) > 5
    L = C.CDLL(os.path.join(ROOT, "cerebellum_simulation_amd", "lib", libname))
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_oracle_exports():
    import orc
    L = orc.lib()
    for n in _declared(os.path.join(ROOT, "oracle", "cbm_oracle.h")):
        assert hasattr(L, n), n


def test_struct_layouts(tmp_path):
    src = tmp_path / "sz.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "cbm_hip.h"\n'
                   'int main(){printf("%zu %zu %zu %zu %zu\\n", sizeof(cbm_sizes), sizeof(cbm_act_params),'
                   ' sizeof(cbm_options), offsetof(cbm_act_params, maxExtIncVIO),'
                   ' offsetof(cbm_act_params, synLTPStepSizeMFtoNC));return 0;}\n')
    exe = tmp_path / "sz"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), "-o", str(exe), str(src)], check=True)
    got = list(map(int, subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()))
    assert got == [C.sizeof(_abi.Sizes), C.sizeof(_abi.ActParams), C.sizeof(_abi.Options),
                   _abi.ActParams.maxExtIncVIO.offset, _abi.ActParams.synLTPStepSizeMFtoNC.offset]


def test_field_table():
    st = configs.build_state("t4k")
    assert len(_abi.FIELDS) == max(_abi.FIELDS.values()) + 1
    for name, fid in _abi.FIELDS.items():
        info = _abi.field_info(fid, st.sizes)
        arr = st.field(name, 0)
        assert arr.nbytes == info["elem"] * info["rows"] * info["cols"], name
    with pytest.raises(_abi.CbmError):
        _abi.field_info(len(_abi.FIELDS), st.sizes)


def test_create_without_gpu_fails_loudly():
    if hip_lib().cbm_device_count() > 0:
        pytest.skip("a GPU is visible")
    from cerebellum_simulation_amd import CbmError, Network
    st = configs.build_state("t4k")
    with pytest.raises(CbmError, match="no HIP device"):
        Network(st.sizes, st.params)


def test_surrogate_struct_layouts(tmp_path):
    src = tmp_path / "sg.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "cbm_surrogate.h"\n'
                   'int main(){printf("%zu %zu %zu %zu %zu\\n", sizeof(cbm_sg_desc), sizeof(cbm_sg_batch),'
                   ' sizeof(cbm_sg_adam), offsetof(cbm_sg_desc, device), offsetof(cbm_sg_batch, targetPitch));'
                   'return 0;}\n')
    exe = tmp_path / "sg"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), "-o", str(exe), str(src)], check=True)
    got = list(map(int, subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()))
    assert got == [C.sizeof(_abi.SgDesc), C.sizeof(_abi.SgBatch), C.sizeof(_abi.SgAdam),
                   _abi.SgDesc.device.offset, _abi.SgBatch.targetPitch.offset]


def test_cartpole_struct_layout(tmp_path):
    src = tmp_path / "cp.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "cbm_env.h"\n'
                   'int main(){printf("%zu %zu %zu %zu %d\\n", sizeof(cbm_cartpole_opts),'
                   ' offsetof(cbm_cartpole_opts, maxTrialLen), offsetof(cbm_cartpole_opts, logfile),'
                   ' offsetof(cbm_cartpole_opts, simStateDir), (int)CBM_CP_COLS);return 0;}\n')
    exe = tmp_path / "cp"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), "-o", str(exe), str(src)], check=True)
    got = list(map(int, subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()))
    from cerebellum_simulation_amd import cartpole
    assert got == [C.sizeof(_abi.CartpoleOpts), _abi.CartpoleOpts.maxTrialLen.offset,
                   _abi.CartpoleOpts.logfile.offset, _abi.CartpoleOpts.simStateDir.offset, len(cartpole.COLS)]


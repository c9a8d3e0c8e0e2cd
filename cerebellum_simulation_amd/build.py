"""Build recipe of the native libraries (no cmake: three compiler lines).

    python -m cerebellum_simulation_amd.build          # product libraries
    python -m cerebellum_simulation_amd.build --all    # + the CPU oracle

Outputs, in-tree under ``cerebellum_simulation_amd/lib``:

* ``libcbm_hip.so``   -- HIP kernels for gfx950 + the ``cbm_*`` C-ABI
* ``libcbm_state.so`` -- parameter files, connectivity, initial state (``cbm_state_*``)
* ``libcbm_core.so``  -- the C++ ``CBMSimCore`` / ``InNet`` / ``MZone`` / ``CBMState`` API
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "lib")
INC = os.path.join(ROOT, "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")
ARCH = os.environ.get("CBM_OFFLOAD_ARCH", "gfx950")

# -ffp-contract=off: no fused multiply-add, so binary32 results round where the
# reference's source rounds.  The packed-fp32 feature is switched off for all
# device code.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", f"-I{INC}",
]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", f"-I{INC}"]


def _run(cmd: list[str]) -> None:
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_hip() -> str:
    out = os.path.join(LIB, "libcbm_hip.so")
    srcs = [os.path.join(CSRC, "hip", f) for f in ("cbm_runtime.hip", "cbm_batch.hip", "cbm_kernels.hip",
                                                      "cbm_surrogate.hip")]
    _run([HIPCC, *HIP_FLAGS, "-shared", "-o", out, *srcs])
    return out


def build_state() -> str:
    out = os.path.join(LIB, "libcbm_state.so")
    srcs = [os.path.join(CSRC, "state", f) for f in ("cbm_state.cpp", "innet_gen.cpp")]
    _run([CXX, *CXX_FLAGS, "-shared", "-o", out, *srcs, f"-L{LIB}", "-lcbm_hip",
          "-Wl,-rpath,$ORIGIN"])
    return out


def build_core() -> str | None:
    core = os.path.join(CSRC, "core")
    srcs = sorted(os.path.join(core, f) for f in os.listdir(core) if f.endswith(".cpp")) \
        if os.path.isdir(core) else []
    if not srcs:
        return None
    out = os.path.join(LIB, "libcbm_core.so")
    _run([CXX, *CXX_FLAGS, f"-I{os.path.join(CSRC, 'core', 'include')}", "-shared", "-o", out,
          *srcs, f"-L{LIB}", "-lcbm_state", "-lcbm_hip", "-Wl,-rpath,$ORIGIN"])
    return out


def build_driver() -> list[str]:
    """The C++ driver test programs (tests/cpp/*.cpp: the eyelid-style driver,
    the cart-pole driver and the ECMFPopulation file writer), written against
    the reference's CBMSimCore / CBMState / environment / tools API and linked
    to libcbm_core."""
    out_dir = os.path.join(ROOT, "tests", "cpp", "build")
    os.makedirs(out_dir, exist_ok=True)
    outs = []
    for name in ("cbm_driver", "cartpole_driver", "ecmfpop_driver", "views_driver"):
        out = os.path.join(out_dir, name)
        _run([CXX, *CXX_FLAGS, f"-I{os.path.join(CSRC, 'core', 'include')}", "-o", out,
              os.path.join(ROOT, "tests", "cpp", name + ".cpp"), f"-L{LIB}", "-lcbm_core", "-lcbm_state",
              "-lcbm_hip", "-Wl,-rpath,$ORIGIN/../../../cerebellum_simulation_amd/lib"])
        outs.append(out)
    return outs


def build_oracle() -> None:
    """The CPU oracle (test infrastructure) and, where the reference tree is
    present, oracle/_ref."""
    odir = os.path.join(ROOT, "oracle")
    _run(["make", "-C", odir, "-j8", "all"])
    if os.path.isdir("/root/reference") and shutil.which(CXX):
        _run(["make", "-C", odir, "ref"])


def build(all_parts: bool = True) -> None:
    os.makedirs(LIB, exist_ok=True)
    build_hip()
    build_state()
    build_core()
    if all_parts:
        build_driver()
        build_oracle()


if __name__ == "__main__":
    build(all_parts="--all" in sys.argv)

"""A C++ driver written against the reference's CBMSimCore / CBMState API
(tests/cpp/cbm_driver.cpp) links against libcbm_core and, on the GPU, gives
the same exports as the Python mirror of the same network; the ECTrialsData
file it records (C++ data library) equals the one the mirror records with the
numpy data library, and the state it saves with writeToState(fstream&) reads
back as the mirror's final state."""
import os
import subprocess

import numpy as np
import pytest

from cerebellum_simulation_amd import Network, NetState, configs, hip_lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "cpp", "build", "cbm_driver")
CON = configs.conparams_text(**configs.CONFIGS["t8k2z"][0])


def _files(tmp_path):
    a, c = tmp_path / "act.txt", tmp_path / "con.txt"
    a.write_text(configs.act_params_text())
    c.write_text(CON)
    return str(a), str(c)


def test_driver_is_built_and_linked():
    assert os.path.exists(DRIVER), "build the driver with __graft_entry__.build()"
    out = subprocess.run(["ldd", DRIVER], capture_output=True, text=True).stdout
    assert "libcbm_core.so" in out and "not found" not in out


def test_driver_without_gpu_reports_error(tmp_path):
    if hip_lib().cbm_device_count() > 0:
        pytest.skip("a GPU is visible")
    a, c = _files(tmp_path)
    r = subprocess.run([DRIVER, a, c, "5", "3", str(tmp_path / "o.bin")], capture_output=True, text=True)
    assert r.returncode == 1 and "no HIP device" in r.stderr


@pytest.mark.gpu
def test_driver_matches_python_mirror(gpu, tmp_path):
    print("[case] C++ driver vs Python mirror", flush=True)
    a, c = _files(tmp_path)
    # 100 steps: three 32-step raster words; the fresh network is quiet in the first
    steps, seed = 100, 3
    out = tmp_path / "o.bin"
    dat, ck = tmp_path / "trials.dat", tmp_path / "state.st"
    r = subprocess.run([DRIVER, a, c, str(steps), str(seed), str(out), str(dat), str(ck)], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    st = NetState(configs.act_params_text(), CON, num_zones=1, innet_seed=seed, n_mf_per_go=16)
    net = Network.from_state(st)
    s = st.sizes
    from cerebellum_simulation_amd.eyelid import EyelidIntegrator
    from cerebellum_simulation_amd.ectrials import ECTrialsData
    eye = EyelidIntegrator(s.numNC)
    lid = []
    third = steps // 3
    psh = {k: (10, n) for k, n in (("go", s.numGO), ("grInputGO", s.numGO), ("goInputGO", s.numGO),
                                   ("goOutSynScale", s.numGO), ("mf", s.numMF), ("gr", s.numGR // 10),
                                   ("sc", s.numSC), ("bc", s.numBC), ("pc", s.numPC), ("io", s.numIO))}
    rasters = {"go": s.numGO, "sc": s.numSC, "bc": s.numBC, "pc": s.numPC, "nc": s.numNC, "io": s.numIO}
    data = ECTrialsData(third, third, steps - 2 * third, 1, 1, psh=psh, rasters=rasters, eyelid_smooth=4)
    idx = np.arange(s.numMF, dtype=np.uint64)
    for t in range(steps):
        mf = (((idx * 7 + t * 13) % 97) < 3).astype(np.uint8)
        if t == steps // 2:
            net.update_err_drive(0, 1.0)
        if t == steps // 3:
            net.set_grpc_plast_steps(0, -0.0009, 0.0005)
        if t == steps // 4:
            net.set_ginc_gr_to_go(0.0003)
        if t == 3 * steps // 4:
            net.reset_ginc_gr_to_go()
        net.update_mf_input(mf)
        net.calc_activity(1)
        lid.append(eye.step(net.read("AP_NC", 0)))
        if t > 0 and t % 32 == 0:
            for k, f in (("go", "APBUF_GO"), ("sc", "APBUF_SC"), ("bc", "APBUF_BC"), ("pc", "APBUF_PC"),
                         ("nc", "APBUF_NC"), ("io", "APBUF_IO")):
                data.rasters[k].update(net.read(f).ravel())
        for k, f in (("go", "AP_GO"), ("grInputGO", "SUM_GR_INPUT_GO"), ("goInputGO", "SUM_GO_INPUT_GO"),
                     ("goOutSynScale", "GO_GABA_OUT_SYNSCALE_GOGO"), ("gr", "AP_GR"), ("sc", "AP_SC"),
                     ("bc", "AP_BC"), ("pc", "AP_PC"), ("io", "AP_IO")):
            data.psh[k].update(net.read(f).ravel())
        data.psh["mf"].update(mf)
        data.eyelid.update(lid[-1])
    raw = np.fromfile(out, dtype=np.uint32)
    parts = [("APBUF_GR", s.numGR), ("VM_GR", s.numGR), ("APBUF_GO", s.numGO), ("VM_GO", s.numGO),
             ("APBUF_SC", s.numSC), ("APBUF_PC", s.numPC), ("VM_PC", s.numPC), ("APBUF_BC", s.numBC),
             ("APBUF_IO", s.numIO), ("VM_IO", s.numIO), ("APBUF_NC", s.numNC), ("VM_NC", s.numNC)]
    off = 0
    for name, n in parts:
        got = raw[off:off + n]
        off += n
        want = net.read(name).ravel().view(np.uint32)
        assert np.array_equal(got, want), name
    assert np.array_equal(raw[off:].view(np.float32), np.array(lid, np.float32)), "eyelid trace"
    assert raw.size == off + steps
    assert dat.read_bytes() == data.write(), "ECTrialsData file"
    assert data.psh["pc"].cells.sum() > 0 and any(r.raster.any() for r in data.rasters.values())
    saved = NetState.read_checkpoint(str(ck))
    for name in ("VM_GR", "APBUF_GR", "HIST_GR", "VM_GO", "PFPC_SYNW", "VM_PC", "G_NC_IO", "SYN_IO_PRELEASE_NC"):
        assert np.array_equal(saved.field(name), net.read(name)), name

"""Trial data files (ECTrialsData) on the CPU: the C++ drop-in library
(csrc/core/ectrialsdata.cpp) and the numpy one (ectrials.py) restate the same
reference code independently, and must write byte-identical files for the same
updates and read each other's files.  Closed forms pin what neither can:
the eyelid smoothing (eyelidout.cpp:62-93: the mean of the input and the
numTSSmooth-1 points already stored, which are themselves smoothed), the
raster bit order (spikerasterbitarray.cpp:124-151: bit 31 of a word is its
earliest step), and histogram bins as sums over bins of the step inputs."""
import os
import subprocess

import numpy as np
import pytest

from cerebellum_simulation_amd import build
from cerebellum_simulation_amd.ectrials import (ECTrialsData, EyelidOut, SpikeRasterBitArray, TrialDataRecorder,
                                                ap_buf_words)

CFG = dict(ms_pre=100, ms_cs=60, ms_post=100, ms_per_ts=1, num_trials=3)
PSH = {"pc": (10, 8), "grInputGO": (10, 5), "goOutSynScale": (20, 5)}
RASTERS = {"pc": 8, "nc": 4}
RAW = {"x": 3}

CPP = r'''
#include <cstdio>
#include <fstream>
#include <vector>
#include <CBMDataInclude/interfaces/ectrialsdata.h>
int main(int argc, char **argv) {
    if (argc != 4) return 4;
    if (std::string(argv[1]) == "copy") {  // read a file and write it again
        std::fstream in(argv[2], std::ios::in | std::ios::binary);
        ECTrialsData d(in);
        std::fstream out(argv[3], std::ios::out | std::ios::binary | std::ios::trunc);
        d.writeData(out);
        return 0;
    }
    std::map<std::string, PSHParams> psh;
    psh["pc"] = PSHParams{10, 8}; psh["grInputGO"] = PSHParams{10, 5}; psh["goOutSynScale"] = PSHParams{20, 5};
    std::map<std::string, RasterParams> ras;
    ras["pc"] = RasterParams{8}; ras["nc"] = RasterParams{4};
    std::map<std::string, RawUIntParams> raw;
    raw["x"] = RawUIntParams{3};
    ECTrialsData d(100, 60, 100, 1, 3, psh, ras, raw, EyelidOutParams{4});
    FILE *f = std::fopen(argv[2], "rb");
    unsigned char pc[8]; unsigned int gi[5]; float go[5]; float eye; unsigned int x[3];
    unsigned int wpc[8], wnc[4];
    for (int t = 0; t < 3; t++)
        for (int ct = 0; ct < 260; ct++) {
            if (std::fread(pc, 1, 8, f) != 8 || std::fread(gi, 4, 5, f) != 5 || std::fread(go, 4, 5, f) != 5 ||
                std::fread(&eye, 4, 1, f) != 1 || std::fread(x, 4, 3, f) != 3) return 2;
            if (ct > 0 && ct % (int)d.getTSPerRasterUpdate() == 0) {
                if (std::fread(wpc, 4, 8, f) != 8 || std::fread(wnc, 4, 4, f) != 4) return 3;
                d.updateRaster("pc", wpc);
                d.updateRaster("nc", wnc);
            }
            d.updatePSH("pc", pc);
            d.updatePSH("grInputGO", gi);
            d.updatePSH("goOutSynScale", go);
            d.updateEyelid(eye);
            d.updateRawUInt("x", x);
        }
    std::fclose(f);
    std::fstream out(argv[3], std::ios::out | std::ios::binary | std::ios::trunc);
    d.writeData(out);
    return 0;
}
'''


@pytest.fixture(scope="module")
def prog(tmp_path_factory):
    d = tmp_path_factory.mktemp("ect")
    src = d / "ect.cpp"
    src.write_text(CPP)
    exe = d / "ect"
    inc = os.path.join(build.CSRC, "core", "include")
    subprocess.run(["g++", "-std=c++17", "-O1", f"-I{inc}", str(src), "-o", str(exe), f"-L{build.LIB}", "-lcbm_core",
                    "-lcbm_state", "-lcbm_hip", f"-Wl,-rpath,{build.LIB}"], check=True)
    return exe


def _trials(seed=0):
    rng = np.random.default_rng(seed)
    T = 260
    out = []
    for _ in range(3):
        pc = (rng.random((T, 8)) < 0.2).astype(np.uint8) * rng.integers(1, 3, (T, 8)).astype(np.uint8)
        nc = (rng.random((T, 4)) < 0.3).astype(np.uint8)
        gi = rng.integers(0, 50, (T, 5)).astype(np.uint32)
        go = rng.random((T, 5)).astype(np.float32)
        eye = (rng.random(T) * 80).astype(np.float32)
        x = rng.integers(0, 2**32, (T, 3), dtype=np.uint64).astype(np.uint32)
        out.append(dict(pc=pc, nc=nc, gi=gi, go=go, eye=eye, x=x))
    return out


def _script(trials, path):
    with open(path, "wb") as f:
        for tr in trials:
            steps = [ct for ct in range(1, 260) if ct % 32 == 0]
            wpc, wnc = ap_buf_words(tr["pc"], steps), ap_buf_words(tr["nc"], steps)
            w = 0
            for ct in range(260):
                f.write(tr["pc"][ct].tobytes() + tr["gi"][ct].tobytes() + tr["go"][ct].tobytes() +
                        tr["eye"][ct:ct + 1].tobytes() + tr["x"][ct].tobytes())
                if ct > 0 and ct % 32 == 0:
                    f.write(wpc[w].tobytes() + wnc[w].tobytes())
                    w += 1


def _python(trials, stepwise):
    d = ECTrialsData(**CFG, psh=PSH, rasters=RASTERS, raw_uint=RAW, eyelid_smooth=4)
    if stepwise:
        for tr in trials:
            steps = [ct for ct in range(1, 260) if ct % 32 == 0]
            wpc, wnc = ap_buf_words(tr["pc"], steps), ap_buf_words(tr["nc"], steps)
            w = 0
            for ct in range(260):
                if ct > 0 and ct % 32 == 0:
                    d.rasters["pc"].update(wpc[w])
                    d.rasters["nc"].update(wnc[w])
                    w += 1
                d.psh["pc"].update(tr["pc"][ct])
                d.psh["grInputGO"].update(tr["gi"][ct])
                d.psh["goOutSynScale"].update(tr["go"][ct])
                d.eyelid.update(tr["eye"][ct])
                d.uint["x"].update(tr["x"][ct])
    else:
        rec = TrialDataRecorder(d)
        for tr in trials:
            rec.add_trial(psh={"pc": tr["pc"], "grInputGO": tr["gi"], "goOutSynScale": tr["go"]},
                          rasters={"pc": tr["pc"], "nc": tr["nc"]}, eyelid=tr["eye"], raw_uint={"x": tr["x"]})
    return d


def test_cpp_and_numpy_write_the_same_file(prog, tmp_path):
    trials = _trials()
    _script(trials, tmp_path / "script.bin")
    subprocess.run([str(prog), "script", str(tmp_path / "script.bin"), str(tmp_path / "cpp.dat")], check=True)
    cpp = (tmp_path / "cpp.dat").read_bytes()
    assert _python(trials, stepwise=True).write() == cpp
    assert _python(trials, stepwise=False).write() == cpp


def test_each_reads_the_others_file(prog, tmp_path):
    trials = _trials(1)
    data = _python(trials, stepwise=False).write()
    (tmp_path / "py.dat").write_bytes(data)
    subprocess.run([str(prog), "copy", str(tmp_path / "py.dat"), str(tmp_path / "back.dat")], check=True)
    assert (tmp_path / "back.dat").read_bytes() == data
    d = ECTrialsData.read(data)
    assert d.write() == data
    assert d.tsTotal == 260 and sorted(d.psh) == sorted(PSH) and d.rasters["pc"].curTrialN == 3


def test_histogram_bins_are_sums_of_steps():
    trials = _trials(2)
    d = _python(trials, stepwise=False)
    pc = np.stack([t["pc"] for t in trials])  # [trial, T, cells]
    want = (pc > 0).sum(axis=0).reshape(26, 10, 8).sum(axis=1)
    assert np.array_equal(d.psh["pc"].cells, want.astype(np.float32))
    gi = np.stack([t["gi"] for t in trials]).sum(axis=0).reshape(26, 10, 5).sum(axis=1)
    assert np.array_equal(d.psh["grInputGO"].cells, gi.astype(np.float32))
    assert d.psh["pc"].curTrialN == 3 and d.psh["pc"].curTS == 0


def test_raster_bit_order_gives_spike_times():
    rng = np.random.default_rng(3)
    sp = (rng.random((260, 4)) < 0.2).astype(np.uint8)
    r = SpikeRasterBitArray("nc", 4, 1, 260, 1)
    for w in ap_buf_words(sp, [ct for ct in range(1, 260) if ct % 32 == 0]):
        r.update(w)
    assert r.bufsPerTrial == 8
    for c in range(4):
        # the word of step 32(u+1) covers steps 32u+1 .. 32u+32, earliest in bit 31
        want = [int(s) for s in np.nonzero(sp[:, c])[0] if 1 <= s <= 32 * r.bufsPerTrial]
        assert r.spike_times(c, 0, offset=1) == want


def test_eyelid_smoothing_closed_forms():
    e = EyelidOut(1, 1, 4, 12)
    for _ in range(12):
        e.update(10.0)
    assert np.all(e.rawData == np.float32(10.0))
    e = EyelidOut(1, 1, 4, 8)
    xs = [0, 0, 0, 8, 8, 8, 8, 8]
    for x in xs:
        e.update(x)
    # stored s_t = (s_{t-3} + s_{t-2} + s_{t-1} + x_t) / 4 for t >= 3, s_t = x_t before
    s = []
    for t, x in enumerate(xs):
        s.append(float(x) if t < 3 else (s[t - 3] + s[t - 2] + s[t - 1] + x) / 4)
    assert np.allclose(e.rawData[0], s, rtol=0, atol=1e-6)
    assert e.rawData[0][3] == np.float32(2.0) and e.rawData[0][4] == np.float32(2.5)

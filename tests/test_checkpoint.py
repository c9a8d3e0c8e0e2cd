"""The CBMState checkpoint format (include/cbm_state.h, cbm_state_*_checkpoint).

Pins, on the CPU:
* a layout restated here, independently of the library, from the reference's
  writers -- the header of cbmstate.cpp:86-103, connectivityparams.cpp:192-218
  and activityparams.cpp:243-253, then the arrays of the four stateRW methods
  (innetconnectivitystate.cpp:530-606, mzoneconnectivitystate.cpp:142-163,
  innetactivitystate.cpp:233-275, mzoneactivitystate.cpp:299-355) with the
  sizes ConnectivityParams derives from the header (connectivityparams.cpp:
  90-184): every array the state holds sits at its offset, and the file ends
  where the last one does;
* write -> read -> write reproduces the file byte for byte, and the state read
  back equals the state written;
* the reference's quirks: writing resets the GO synaptic-depression arrays in
  the state; the MZone section holds pBCfromPCtoBC twice and no pPCfromPCtoBC;
  a read rebuilds P_PC_FROM_PC_TO_BC as the reference's generator makes it;
* the C++ CBMState(fstream&) / writeState(fstream&) of the drop-in library
  read and write the same bytes.
On the GPU: a network checkpointed mid-run and restored into a new handle
continues bit for bit like the original, once what the format does not hold
is carried over: the step counter, the pipeline rows, the zone generator and
the MZoneActivityState members stateRW skips (MF->NC plasticity flags, PC
population-activity counters).
"""
import ctypes as C
import os

import numpy as np
import pytest

from cerebellum_simulation_amd import Network, _abi, configs
from cerebellum_simulation_amd.network import NetState

U32MAX = np.uint32(0xFFFFFFFF)


def _header(data: bytes):
    """(zones, connectivity key -> int, GO-GO matrix sizes, activity key -> text, header length)."""
    end = data.index(b"activityParamEnd 1\n") + len(b"activityParamEnd 1\n")
    toks = data[:end].decode().split()
    zones, i = int(toks[0]), 1
    con, mats = {}, {}
    while toks[i] != "connectivityParamEnd":
        if toks[i] in ("GOGOLocalMatrix", "GOGOCoupLocalMatrix"):
            nin, nout = int(toks[i + 2]), int(toks[i + 4])
            mats[toks[i]] = (nin, nout)
            i += 5 + 3 * nout
        else:
            con[toks[i]] = int(toks[i + 1])
            i += 2
    i += 2
    act = {}
    while toks[i] != "activityParamEnd":
        act[toks[i]] = toks[i + 1]
        i += 2
    return zones, con, mats, act, end


def _layout(con, mats, pop_bins):
    """[(name, group, bytes)] in file order, sizes as ConnectivityParams derives them."""
    g = lambda k: con.get(k, 0)  # noqa: E731
    numGL = (1 << g("glXP2")) * (1 << g("glYP2"))
    numMF = 1 << g("numMFP2")
    mfGL = numGL // numMF
    numGO = (1 << g("goXP2")) * (1 << g("goYP2"))
    numGR = (1 << g("grXP2")) * (1 << g("grYP2"))
    grGO = g("maxnumpGRfromGOtoGR")
    numSC, numPC, numNC, numIO = (1 << g("numSCP2")), (1 << g("numPCP2")), (1 << g("numNCP2")), (1 << g("numIOP2"))
    numBC = 4 * numPC
    gabaIn, gabaOut = mats["GOGOLocalMatrix"]
    coupIn, coupOut = mats.get("GOGOCoupLocalMatrix", (1, 1))
    ncPC = numPC // numNC * g("numpPCfromPCtoNC")
    innet_con = [
        ("haspGLfromMFtoGL", numGL), ("pGLfromMFtoGL", 4 * numGL),
        ("numpGLfromGLtoGO", 4 * numGL), ("pGLfromGLtoGO", 4 * numGL * g("maxnumpGLfromGLtoGO")),
        ("numpGLfromGOtoGL", 4 * numGL), ("pGLfromGOtoGL", 4 * numGL * g("maxnumpGLfromGOtoGL")),
        ("numpGLfromGLtoGR", 4 * numGL), ("pGLfromGLtoGR", 4 * numGL * g("maxnumpGLfromGLtoGR")),
        ("numpMFfromMFtoGL", 4 * numMF), ("pMFfromMFtoGL", 4 * numMF * mfGL),
        ("numpMFfromMFtoGR", 4 * numMF), ("pMFfromMFtoGR", 4 * numMF * mfGL * g("maxnumpGLfromGLtoGR")),
        ("NUMP_MF_FROM_MF_TO_GO", 4 * numMF), ("P_MF_FROM_MF_TO_GO", 4 * numMF * mfGL * g("maxnumpGLfromGLtoGO")),
        ("numpGOfromGLtoGO", 4 * numGO), ("pGOfromGLtoGO", 4 * numGO * g("maxnumpGOfromGLtoGO")),
        ("numpGOfromGOtoGL", 4 * numGO), ("pGOfromGOtoGL", 4 * numGO * g("maxnumpGOfromGOtoGL")),
        ("numpGOfromMFtoGO", 4 * numGO), ("pGOfromMFtoGO", 4 * numGO * g("maxnumpGOfromGLtoGO")),
        ("numpGOfromGOtoGR", 4 * numGO),
        ("pGOfromGOtoGR", 4 * numGO * g("maxnumpGOfromGOtoGL") * g("maxnumpGLfromGLtoGR")),
        ("numpGOfromGRtoGO", 4 * numGO), ("pGOfromGRtoGO", 4 * numGO * g("maxnumpGOfromGRtoGO")),
        ("numpGOGABAInGOGO", 4 * numGO), ("pGOGABAInGOGO", 4 * numGO * gabaIn),
        ("NUMP_GO_GABA_OUT_GOGO", 4 * numGO), ("P_GO_GABA_OUT_GOGO", 4 * numGO * gabaOut),
        ("NUMP_GO_COUP_IN_GOGO", 4 * numGO), ("P_GO_COUP_IN_GOGO", 4 * numGO * coupIn),
        ("numpGOCoupOutGOGO", 4 * numGO), ("pGOCoupOutGOGO", 4 * numGO * coupOut),
        ("P_GR_DELAYMASK_GR_TO_BSP", 4 * numGR),
        ("numpGRfromGLtoGR", 4 * numGR), ("pGRfromGLtoGR", 4 * numGR * grGO),
        ("NUMP_GR_FROM_GR_TO_GO", 4 * numGR), ("P_GR_FROM_GR_TO_GO", 4 * numGR * g("maxnumpGRfromGRtoGO")),
        ("P_GR_DELAYMASK_GR_TO_GO", 4 * numGR * g("maxnumpGRfromGRtoGO")),
        ("NUMP_GR_FROM_GO_TO_GR", 4 * numGR), ("P_GR_FROM_GO_TO_GR", 4 * numGR * grGO),
        ("NUMP_GR_FROM_MF_TO_GR", 4 * numGR), ("P_GR_FROM_MF_TO_GR", 4 * numGR * grGO),
    ]
    mzone_con = [
        ("P_BC_FROM_BC_TO_PC", 4 * numBC * 4), ("pBCfromPCtoBC", 4 * numBC * 4), ("pSCfromSCtoPC", 4 * numSC),
        ("pPCfromBCtoPC", 4 * numPC * 16), ("pBCfromPCtoBC#2", 4 * numPC * 16),
        ("P_PC_FROM_SC_TO_PC", 4 * numPC * (numSC // numPC)), ("pPCfromPCtoNC", 4 * numPC * g("numpPCfromPCtoNC")),
        ("pPCfromIOtoPC", 4 * numPC), ("P_NC_FROM_PC_TO_NC", 4 * numNC * ncPC), ("pNCfromNCtoIO", 4 * numNC * numIO),
        ("P_NC_FROM_MF_TO_NC", 4 * numNC * (numMF // numNC)), ("pIOfromIOtoPC", 4 * numIO * (numPC // numIO)),
        ("P_IO_FROM_NC_TO_IO", 4 * numIO * numNC), ("P_IO_IN_IOIO", 4 * numIO * (numIO - 1)),
        ("pIOOutIOIO", 4 * numIO * (numIO - 1)),
    ]
    f4 = lambda n: 4 * n  # noqa: E731
    innet_act = [
        ("HIST_MF", numMF), ("APBUF_MF", f4(numMF)), ("AP_GO", numGO), ("APBUF_GO", f4(numGO)), ("VM_GO", f4(numGO)),
        ("VCOUPLE_GO", f4(numGO)), ("THRESH_GO", f4(numGO)), ("INPUT_MF_GO", f4(numGO)), ("INPUT_GO_GO", f4(numGO)),
        ("G_MF_GO", f4(numGO)), ("G_GR_GO", f4(numGO)), ("G_GO_GO", f4(numGO)), ("G_MGLUR_GO", f4(numGO)),
        ("G_MGLURINC_GO", f4(numGO)), ("MGLUR_GO", f4(numGO)), ("GLU_GO", f4(numGO)),
        ("AP_GR", numGR), ("APBUF_GR", f4(numGR)), ("gMFGR", f4(numGR * g("maxnumpGRfromMFtoGR"))),
        ("G_MF_SUM_GR", f4(numGR)), ("gGOGR", f4(numGR * grGO)), ("G_GO_SUM_GR", f4(numGR)),
        ("THRESH_GR", f4(numGR)), ("VM_GR", f4(numGR)), ("GKCA_GR", f4(numGR)), ("HIST_GR", 8 * numGR),
        ("AP_SC", numSC), ("APBUF_SC", f4(numSC)), ("G_PF_SC", f4(numSC)), ("THRESH_SC", f4(numSC)),
        ("VM_SC", f4(numSC)), ("INPUT_SUM_PF_SC", f4(numSC)),
    ]
    pcSC, ncMF = numSC // numPC, numMF // numNC
    mzone_act = [
        ("AP_BC", numBC), ("APBUF_BC", f4(numBC)), ("INPUT_PC_BC", f4(numBC)), ("G_PF_BC", f4(numBC)),
        ("G_PC_BC", f4(numBC)), ("THRESH_BC", f4(numBC)), ("VM_BC", f4(numBC)),
        ("AP_PC", numPC), ("APBUF_PC", f4(numPC)), ("INPUT_BC_PC", f4(numPC)), ("INPUT_SC_PC", numPC * pcSC),
        ("PFPC_SYNW", f4(numGR)), ("INPUT_SUM_PF_PC", f4(numPC)), ("G_PF_PC", f4(numPC)), ("G_BC_PC", f4(numPC)),
        ("G_SC_PC", f4(numPC * pcSC)), ("VM_PC", f4(numPC)), ("THRESH_PC", f4(numPC)),
        ("HIST_PC_POPACT", f4(pop_bins)),
        ("AP_IO", numIO), ("APBUF_IO", f4(numIO)), ("INPUT_NC_IO", numIO * numNC), ("G_NC_IO", f4(numIO * numNC)),
        ("THRESH_IO", f4(numIO)), ("VM_IO", f4(numIO)), ("VCOUPLE_IO", f4(numIO)), ("PFPC_PLAST_TIMER_IO", f4(numIO)),
        ("AP_NC", numNC), ("APBUF_NC", f4(numNC)), ("INPUT_PC_NC", numNC * ncPC), ("INPUT_MF_NC", numNC * ncMF),
        ("G_PC_NC", f4(numNC * ncPC)), ("G_PC_SCALE_NC", f4(numNC * ncPC)), ("MF_SYNW_NC", f4(numNC * ncMF)),
        ("GMAX_NMDA_MF_NC", f4(numNC * ncMF)), ("GMAX_AMPA_MF_NC", f4(numNC * ncMF)),
        ("G_MF_NMDA_NC", f4(numNC * ncMF)), ("G_MF_AMPA_NC", f4(numNC * ncMF)), ("THRESH_NC", f4(numNC)),
        ("VM_NC", f4(numNC)), ("SYN_IO_PRELEASE_NC", f4(numNC)),
    ]
    return innet_con, mzone_con, innet_act, mzone_act


def _offsets(data: bytes):
    zones, con, mats, act, start = _header(data)
    pop_bins = 8
    groups = _layout(con, mats, pop_bins)
    out, off = [], start
    for gi, grp in enumerate(groups):
        for z in range(zones if gi in (1, 3) else 1):
            for name, n in grp:
                out.append((name, z, off, n))
                off += n
    return out, off


def _state(name="t8k2z", seed=3):
    st = configs.build_state(name, seed=seed)
    rng = np.random.default_rng(seed)
    # activity values that differ per cell, so misplaced arrays show
    for f in ("VM_GR", "THRESH_GR", "GKCA_GR", "G_MF_SUM_GR", "VM_GO", "G_GR_GO"):
        a = st.field(f)
        a[...] = rng.random(a.shape).astype(a.dtype)
    st.field("HIST_GR")[...] = rng.integers(0, 2**63, st.field("HIST_GR").shape, dtype=np.int64).astype(np.uint64)
    st.field("APBUF_GR")[...] = rng.integers(0, 2**32, st.field("APBUF_GR").shape, dtype=np.uint64).astype(np.uint32)
    for z in range(st.num_zones):
        for f in ("PFPC_SYNW", "VM_PC", "SYN_IO_PRELEASE_NC", "G_NC_IO"):
            a = st.field(f, z)
            a[...] = rng.random(a.shape).astype(a.dtype)
    return st


def test_layout_matches_reference_writers():
    st = _state()
    data = st.checkpoint_bytes()
    secs, end = _offsets(data)
    assert end == len(data)
    for name, z, off, n in secs:
        if name not in _abi.FIELDS:
            continue
        want = st.field(name, z)
        got = np.frombuffer(data[off:off + n], dtype=want.dtype)
        if name == "P_GR_FROM_GO_TO_GR":  # padded to the MF-input width of the file
            got = got.reshape(want.shape[0], -1)
            assert np.array_equal(got[:, :want.shape[1]], want) and np.all(got[:, want.shape[1]:] == U32MAX)
        else:
            assert got.size == want.size and np.array_equal(got, want.reshape(-1)), (name, z)


def test_header_text_format():
    st = _state()
    data = st.checkpoint_bytes()
    zones, con, mats, act, _ = _header(data)
    assert zones == st.num_zones
    text = data[:data.index(b"activityParamEnd 1\n")].decode().splitlines()
    keys = [ln.split()[0] for ln in text[1:text.index("GOGOLocalMatrix")]]
    assert keys == sorted(keys)  # std::map order
    acts = [ln.split()[0] for ln in text[text.index("connectivityParamEnd 1") + 1:]]
    assert acts == sorted(acts)
    # floats as std::ostream prints them: %g with 6 significant digits
    for k, v in act.items():
        assert v == ("%g" % float(np.float32(float(v)))), (k, v)
    assert mats["GOGOLocalMatrix"] == (st.sizes.maxnumpGOGABAOutGOGO, st.sizes.maxnumpGOGABAOutGOGO)


def test_round_trip_is_exact():
    st = _state()
    b1 = st.checkpoint_bytes()
    st2 = NetState.from_checkpoint_bytes(b1)
    assert st2.checkpoint_bytes() == b1
    for name, fid in _abi.FIELDS.items():
        info = _abi.field_info(fid, st.sizes)
        if info["kind"] == 2 or name == "ZONE_RNG_SEED":  # not in the format (CBMSimCore seeds it)
            continue
        for z in range(st.num_zones if info["zoned"] else 1):
            a, b = st.field(name, z), st2.field(name, z)
            if name == "P_GR_FROM_GO_TO_GR":
                assert np.array_equal(b[:, :a.shape[1]], a) and np.all(b[:, a.shape[1]:] == U32MAX)
            else:
                assert np.array_equal(a, b), (name, z)
    # derived activity constants survive the 6-digit header
    assert bytes(st.params) == bytes(st2.params)


def test_reference_quirks(tmp_path):
    st = _state()
    st.field("INPUT_GOGABA_SYNDEP_GO")[...] = 0.25
    st.field("GO_GABA_OUT_SYNSCALE_GOGO")[...] = 0.75
    path = tmp_path / "sim.st"
    st.write_checkpoint(str(path))
    # writing resets GO-GO synaptic depression in the state (innetactivitystate.cpp:247-248)
    assert np.all(st.field("INPUT_GOGABA_SYNDEP_GO") == 1) and np.all(st.field("GO_GABA_OUT_SYNSCALE_GOGO") == 0)
    data = path.read_bytes()
    secs, _ = _offsets(data)
    byname = {(n, z): (o, k) for n, z, o, k in secs}
    for z in range(st.num_zones):
        o1, k1 = byname[("pBCfromPCtoBC", z)]
        o2, k2 = byname[("pBCfromPCtoBC#2", z)]
        assert k1 == k2 and data[o1:o1 + k1] == data[o2:o2 + k2]
    st2 = NetState.read_checkpoint(str(path))
    for z in range(st.num_zones):
        assert np.array_equal(st2.field("P_PC_FROM_PC_TO_BC", z), st.field("P_PC_FROM_PC_TO_BC", z))
    assert np.all(st2.field("INPUT_GOGABA_SYNDEP_GO") == 1) and np.all(st2.field("GO_GABA_OUT_SYNSCALE_GOGO") == 0)


def test_truncated_file_fails(tmp_path):
    st = _state()
    data = st.checkpoint_bytes()
    p = tmp_path / "short.st"
    p.write_bytes(data[:-5])
    with pytest.raises(_abi.CbmError, match="ends inside"):
        NetState.read_checkpoint(str(p))


def test_cpp_cbmstate_reads_and_writes_the_same_bytes(tmp_path):
    """CBMState(fstream&) + writeState(fstream&) of the drop-in C++ library."""
    import subprocess
    from cerebellum_simulation_amd import build
    st = _state()
    src = tmp_path / "ck.st"
    st.write_checkpoint(str(src))
    prog = tmp_path / "rt.cpp"
    prog.write_text(
        '#include <fstream>\n#include <CBMStateInclude/interfaces/cbmstate.h>\n'
        'int main(int c, char **v) {\n'
        '  std::fstream in(v[1], std::ios::in | std::ios::binary);\n'
        '  CBMState s(in);\n'
        '  std::fstream out(v[2], std::ios::out | std::ios::binary | std::ios::trunc);\n'
        '  s.writeState(out);\n  return (int)s.getNumZones() == 2 ? 0 : 1;\n}\n')
    exe = tmp_path / "rt"
    inc = os.path.join(build.CSRC, "core", "include")
    subprocess.run(["g++", "-std=c++17", f"-I{inc}", f"-I{build.INC}", str(prog), "-o", str(exe), f"-L{build.LIB}",
                    "-lcbm_core", "-lcbm_state", "-lcbm_hip", f"-Wl,-rpath,{build.LIB}"], check=True)
    out = tmp_path / "ck2.st"
    subprocess.run([str(exe), str(src), str(out)], check=True)
    assert out.read_bytes() == src.read_bytes()


@pytest.mark.gpu
def test_restored_network_continues_bit_for_bit(gpu, tmp_path):
    import torch
    print("[case] checkpoint mid-run and continue, 8K-GR 2-zone network", flush=True)
    st = configs.build_state("t8k2z", seed=5)
    s = st.sizes
    a = Network.from_state(st)
    train = torch.from_numpy(configs.mf_train(s.numMF, 400, seed=6)).cuda()
    err = configs.us_err_drive(400, st.num_zones, us_ms=150)
    a.run_mf_train(train[:200], err_drive=err[:200])
    # CBMSimCore::writeToState(fstream&): the state takes the device arrays, is saved,
    # and the running network takes the GO synaptic-depression reset of the write
    for name, fid in _abi.FIELDS.items():
        info = _abi.field_info(fid, s)
        if info["kind"] != 1:
            continue
        for z in range(st.num_zones if info["zoned"] else 1):
            st.field(name, z)[...] = a.read(name, z)
    path = tmp_path / "mid.st"
    st.write_checkpoint(str(path))
    for name in ("INPUT_GOGABA_SYNDEP_GO", "GO_GABA_OUT_SYNSCALE_GOGO"):
        a.write(name, st.field(name))
    b = Network.from_state(NetState.read_checkpoint(str(path)))
    # not in the format: step counter, pipeline rows, MF input, zone generators
    for name in ("CUR_TIME", "MF_INPUT", "SUM_GR_INPUT_GO", "SUM_GO_INPUT_GO", "PIPE_PFBC_SUM", "PIPE_PFSC_SUM",
                 "PIPE_GRGO_SUM", "GINC_GR_TO_GO"):
        b.write(name, a.read(name))
    for z in range(st.num_zones):
        # not in the file: the zone generator, and the MZoneActivityState members
        # stateRW skips (mzoneactivitystate.cpp:299-355), the MF->NC plasticity
        # flags and the PC population-activity counters
        for name in ("ZONE_RNG_STATE", "PIPE_PFPC_SUM", "GRPC_PLAST_STEPS", "MFNC_NOLTP_NOLTD", "PC_POPACT_CTRS"):
            b.write(name, a.read(name, z), z)
    a.set_recorder(["PC", "NC", "IO", "GO"], 200)
    b.set_recorder(["PC", "NC", "IO", "GO"], 200)
    a.run_mf_train(train[200:], err_drive=err[200:])
    b.run_mf_train(train[200:], err_drive=err[200:])
    for pop in ("PC", "NC", "IO", "GO"):
        assert np.array_equal(a.read_recorded(pop, 201, 200), b.read_recorded(pop, 201, 200)), pop
    for name in ("VM_GR", "PFPC_SYNW", "HIST_GR", "MF_SYNW_NC"):
        for z in range(st.num_zones if _abi.field_info(_abi.FIELDS[name], s)["zoned"] else 1):
            assert np.array_equal(a.read(name, z), b.read(name, z)), name
    assert a.read_recorded("PC", 201, 200).sum() > 0

"""The cart-pole environment with two microzones (include/cbm_env.h,
CBMEnvInclude/, PoissonRegenCells).

On the CPU:
* the library's environment (C++, run detached with scripted NC spikes) and
  the oracle's restatement (oracle/cartpole_ref.py) give the same MF rates,
  MF spikes, trajectories, forces and errors step for step, open and closed
  loop, on an infinite and a finite track, across falls, time-outs, trials
  cut at their maximum length and the end of the run;
* the oracle is pinned: its accelerations against the point-mass cart-pole
  written in another algebraic form, its Gaussian tuning against the closed
  form exp(-k^2/72) at slot centres, its MF generator against the renewal-
  theory rate of the recovering threshold, its error rule's signs, and the
  number of generator draws per step.
On the GPU: the library's closed loop on a two-zone network equals the
oracle environment driving the same network step by step from Python, and
the C++ driver written against the reference's API (tests/cpp/cartpole_driver.cpp)
equals the library's loop.
"""
import math
import os
import subprocess

import numpy as np
import pytest

import cartpole_ref as R
from cerebellum_simulation_amd import _abi, configs
from cerebellum_simulation_amd import cartpole as CP

f32 = np.float32
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "cpp", "build", "cartpole_driver")


def _sizes(num_mf, num_nc):
    s = _abi.Sizes()
    s.numMF, s.numNC, s.numZones = num_mf, num_nc, 2
    return s


def _same(a, b) -> bool:
    return np.array_equal(np.asarray(a, np.float32), np.asarray(b, np.float32), equal_nan=True)


# open loop: the pole starts within 1e-5 rad and, with no force, takes about
# 9 s to fall, so those trials run long; closed loop: the scripted NC spikes
# push it over within a few hundred steps, and trials are cut at 1,200 steps
@pytest.mark.parametrize("closed,track,seed,max_len,steps",
                         [(False, 0.0, 3, 20000, 12500), (True, 0.0, 4, 1200, 9000), (True, 2.0, 5, 1200, 9000),
                          (False, 1.0, 6, 20000, 12500)],
                         ids=["open-infinite", "closed-infinite", "closed-track2", "open-track1"])
def test_library_and_oracle_agree_step_for_step(closed, track, seed, max_len, steps):
    num_mf, num_nc = 160, 4
    lib = CP.Cartpole(sizes=_sizes(num_mf, num_nc), env_seed=seed, mf_seed=seed + 50, closed_loop=closed,
                      track_len=track, max_trial_len=max_len, max_num_trials=4)
    ref = R.CartpoleRef(num_mf, num_nc, seed, seed + 50, closed_loop=closed, track_len=track, max_trial_len=max_len,
                        max_num_trials=4)
    lo, hi = lib.mf_rates()
    assert _same(lo, ref.relaxed) and _same(hi, ref.excited)
    groups = lib.mf_groups()
    for g, members in enumerate(ref.groups):
        assert np.array_equal(np.nonzero(groups == g)[0], np.sort(members)), g
    rng = np.random.default_rng(seed)
    seen = {"aloft": 0, "errors": 0, "trials": set(), "nan": False}
    for t in range(steps):
        if lib.terminated:
            assert ref.terminated
            break
        f_lib, f_ref = lib.get_state(), ref.get_state()
        assert _same(f_lib, f_ref), t
        seen["nan"] |= bool(np.isnan(f_ref).any())
        ap_lib, ap_ref = lib.mf_spikes(f_lib), ref.mfs.step(f_ref)
        assert np.array_equal(ap_lib, ap_ref), t
        ncl = (rng.random(num_nc) < 0.3).astype(np.uint8)
        ncr = (rng.random(num_nc) < 0.1).astype(np.uint8)
        row_lib, err_lib = lib.step(ncl, ncr)
        row_ref, err_ref = ref.step(ncl, ncr)
        assert _same(row_lib, row_ref), (t, row_lib, row_ref)
        assert tuple(err_lib) == tuple(err_ref), t
        seen["aloft"] += row_ref[R_COL("time_aloft")] > 0
        seen["errors"] += sum(err_ref)
        seen["trials"].add(int(row_ref[R_COL("trial")]))
    assert seen["aloft"] > 1000 and seen["errors"] > 0 and len(seen["trials"]) >= 2
    # an infinite track leaves the cart-position group's rates undefined
    # (the Gaussian's width overflows), as in the reference
    assert seen["nan"] == (track == 0.0)


def R_COL(name):
    return CP.COL[name]


def test_open_loop_keeps_forces_at_zero():
    lib = CP.Cartpole(sizes=_sizes(128, 4), env_seed=1, closed_loop=False)
    for _ in range(2000):
        lib.mf_spikes(lib.get_state())
        row, _ = lib.step(np.ones(4, np.uint8), np.ones(4, np.uint8))
        assert row[CP.COL["force_left"]] == 0 and row[CP.COL["force_right"]] == 0


def test_closed_loop_force_closed_form():
    # one zone fires every NC every step: force_n = sum_k 0.1 * 0.99^k, times 0.99
    ref = R.MicrozoneRef(4)
    for n in range(1, 200):
        ref.update(np.ones(4, np.uint8))
        want = 0.1 * 0.99 * (1 - 0.99 ** n) / (1 - 0.99)
        assert abs(float(ref.force) - want) <= 1e-5 * want
    assert ref.avg == f32(1)


def test_equations_of_motion_match_point_mass_cartpole():
    """theta'' = ((M+m) g sin + F cos - m L w^2 sin cos) / (L (M + m sin^2)),
    x'' = (F + m sin (g cos - L w^2)) / (M + m (1 - cos)) (the reference's
    cart equation, whose denominator has cos where the textbook has cos^2)."""
    ref = R.CartpoleRef(64, 2, 0)
    g, M, m, L = 9.8, 1.0, 0.1, 7.5
    rng = np.random.default_rng(0)
    for _ in range(500):
        th, w, F = rng.uniform(-0.5, 0.5), rng.uniform(-0.5, 0.5), rng.uniform(-2, 2)
        ref.theta, ref.theta_dot, ref.x, ref.x_dot = f32(th), f32(w), f32(0), f32(0)
        tacc, xacc = ref.physics(f32(F))
        th, w, F = float(f32(th)), float(f32(w)), float(f32(F))
        s, c = math.sin(th), math.cos(th)
        want_t = ((M + m) * g * s + F * c - m * L * w * w * s * c) / (L * (M + m * s * s))
        want_x = (F + m * s * (g * c - L * w * w)) / (M + m * (1 - c))
        assert abs(float(tacc) - want_t) <= 2e-5 * max(1.0, abs(want_t))
        assert abs(float(xacc) - want_x) <= 2e-5 * max(1.0, abs(want_x))
        # the Euler step moves the pole by tau * the old angular velocity
        assert ref.theta == f32(f32(th) + f32(f32(0.001) * f32(w)))


def test_gaussian_tuning_closed_form():
    ref = R.CartpoleRef(64, 2, 0)
    n, lo, hi = 10, f32(-0.5), f32(0.5)
    members = np.arange(n)
    freq = np.zeros(64, np.float32)
    relaxed, excited = ref.relaxed, ref.excited
    slot = f32(f32(hi - lo) / f32(n))
    for j in (0, 3, 9):
        centre = f32(float(lo) + float(slot) / 2.0)
        for _ in range(j):
            centre = f32(centre + slot)
        ref.gaussian(members, lo, hi, centre, freq)
        for k in range(n):
            gk = math.exp(-((k - j) ** 2) / 72.0)
            want = gk * float(excited[k]) + (1 - gk) * float(relaxed[k])
            assert abs(float(freq[k]) - want) <= 2e-5 * want, (j, k)
    # values beyond the range clamp to its ends
    a, b = freq.copy(), freq.copy()
    ref.gaussian(members, lo, hi, f32(7.0), a)
    ref.gaussian(members, lo, hi, hi, b)
    assert np.array_equal(a, b)


def _renewal_rate(p: float, d: float, kmax: int = 4000) -> float:
    """Spikes per step of a cell whose spike probability is p * (1 - (1-d)^k)
    k steps after its last spike."""
    surv, mean = 1.0, 0.0
    for k in range(1, kmax):
        h = p * (1 - (1 - d) ** k)
        mean += k * surv * h
        surv *= 1 - h
    return 1.0 / mean


def test_mf_generator_rate_matches_renewal_theory():
    n, steps = 4096, 1500
    ref = R.PoissonRef(n, 11)
    freq = np.full(n, 100.0, np.float32)
    total = 0
    for _ in range(steps):
        total += int(ref.step(freq).sum())
    d = 1 - math.exp(-0.25)
    want = _renewal_rate(0.1, d)
    got = total / (n * steps)
    assert abs(got - want) <= 0.01 * want, (got, want)
    # a wrong recovery (no refractoriness, or tau 2 instead of 4) is far off
    assert abs(_renewal_rate(0.1, 1.0) - want) > 0.05 * want
    assert abs(_renewal_rate(0.1, 1 - math.exp(-0.5)) - want) > 0.03 * want
    # a NaN rate never spikes
    ref2 = R.PoissonRef(8, 1)
    assert ref2.step(np.full(8, np.nan, np.float32)).sum() == 0


def test_library_mf_generator_equals_oracle():
    lib = CP.Cartpole(sizes=_sizes(512, 4), env_seed=2, mf_seed=77)
    ref = R.PoissonRef(512, 77)
    rng = np.random.default_rng(1)
    for _ in range(300):
        f = (rng.random(512) * 120).astype(np.float32)
        f[rng.random(512) < 0.02] = np.nan
        assert np.array_equal(lib.mf_spikes(f), ref.step(f))


def test_error_rule_signs_and_draws():
    ref = R.CartpoleRef(64, 2, 9, track_len=0.0)
    counts = {}
    for th, w in ((0.05, 0.1), (-0.05, -0.1), (0.05, -0.1), (-0.05, 0.1)):
        l = r = 0
        for _ in range(20000):
            ref.theta, ref.theta_dot = f32(th), f32(w)
            a, b = ref.errors()
            l, r = l + a, r + b
        counts[(th, w)] = (l, r)
    # falling right (theta, theta' > 0): the left microzone, at 1% per step
    assert counts[(0.05, 0.1)][1] == 0 and 150 < counts[(0.05, 0.1)][0] < 250
    assert counts[(-0.05, -0.1)][0] == 0 and 150 < counts[(-0.05, -0.1)][1] < 250
    assert counts[(0.05, -0.1)] == (0, 0) and counts[(-0.05, 0.1)] == (0, 0)
    # one draw per aloft step on an infinite track, three on a finite one
    for track, per in ((0.0, 1), (3.0, 3)):
        ref = R.CartpoleRef(64, 2, 9, track_len=track)
        d0 = ref.gen.draws
        assert d0 == 2 + 2 * 64 + sum(len(g) for g in ref.groups)
        ref.timeout = R.CartpoleRef.TIMEOUT
        ref.step()
        assert ref.gen.draws - d0 == per


@pytest.mark.gpu
def test_closed_loop_run_matches_oracle_environment(gpu):
    from cerebellum_simulation_amd import Network
    print("[case] cart-pole closed loop, 8K-GR two-zone network, library loop vs oracle env", flush=True)
    con = configs.conparams_text(**configs.CONFIGS["t8k2z"][0])
    act = configs.act_params_text()
    # long trials: the pole falls under the network's own force and the error
    # rule fires as it tilts
    seed, steps = 12, 12000
    st = CP.build_state(act, con, seed)
    a, b = Network.from_state(st), Network.from_state(st)
    s = st.sizes
    opts = dict(closed_loop=True, max_trial_len=20000, max_num_trials=6, track_len=2.0)
    lib = CP.Cartpole(a, env_seed=seed, mf_seed=seed, **opts)
    trace = lib.run(steps)
    ref = R.CartpoleRef(s.numMF, s.numNC, seed, seed, **opts)
    for t in range(trace.shape[0]):
        f = ref.get_state()
        ap = ref.mfs.step(f)
        b.update_mf_input(ap)
        b.calc_activity(1)
        ncl, ncr = b.read("AP_NC", 0).ravel(), b.read("AP_NC", 1).ravel()
        row, err = ref.step(ncl, ncr)
        row[CP.COL["mf_spikes"]] = ap.sum()
        assert _same(trace[t], row), (t, trace[t], row)
        for z in range(2):
            if err[z]:
                b.update_err_drive(z, 1.0)
    errs = trace[:, CP.COL["err_left"]].sum() + trace[:, CP.COL["err_right"]].sum()
    print(f"steps {trace.shape[0]}, trials {int(trace[-1, CP.COL['trial']]) + 1}, errors {errs}", flush=True)
    assert trace[:, CP.COL["nc_left"]].sum() > 0 and errs > 0
    assert (trace[:, CP.COL["force_left"]] != 0).any()
    lib.close()
    a.close()
    b.close()


@pytest.mark.gpu
@pytest.mark.parametrize("closed,glomerular", [(False, False), (True, False), (True, True)],
                         ids=["open", "closed", "closed-glomerular"])
def test_cpp_driver_matches_library_loop(gpu, tmp_path, closed, glomerular):
    """The glomerular case gives the driver a parameter file with a glomerulus
    layer: CBMState wires it with the reference's generator, as build_state does."""
    from cerebellum_simulation_amd import Network
    print(f"[case] cart-pole C++ driver vs library loop, closed={closed} glomerular={glomerular}", flush=True)
    con = (configs.glomerular_conparams_text("g2k", numNCP2=2, numIOP2=1) if glomerular else
           configs.conparams_text(**configs.CONFIGS["t8k2z"][0]))
    act = configs.act_params_text()
    (tmp_path / "act.txt").write_text(act)
    (tmp_path / "con.txt").write_text(con)
    seed, steps = 21, 3000
    args = ["--maxTrialLen", "700", "--maxNumTrials", "5", "--logfile", str(tmp_path / "cp.log"),
            "--simStateDir", ""]
    if closed:
        args.append("--closedLoop")
    r = subprocess.run([DRIVER, str(tmp_path / "act.txt"), str(tmp_path / "con.txt"), str(seed), str(steps),
                        str(tmp_path / "trace.bin"), *args], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    got = np.fromfile(tmp_path / "trace.bin", np.float32).reshape(-1, len(CP.COLS))
    st = CP.build_state(act, con, seed)
    net = Network.from_state(st)
    lib = CP.Cartpole(net, env_seed=seed, mf_seed=seed, closed_loop=closed, max_trial_len=700, max_num_trials=5)
    want = lib.run(steps)
    assert got.shape == want.shape and int(r.stdout.split()[-1]) == want.shape[0]
    assert _same(got, want)
    assert "EndTrial" in (tmp_path / "cp.log").read_text()
    lib.close()
    net.close()

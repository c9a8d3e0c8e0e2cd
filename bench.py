"""Benchmark: simulated ms per wall-second of the 1M-GR eyelid network on one GPU.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c1]

A step is one 1 ms network step (all GR, GO, SC, BC, PC, IO, NC) over one
row of a synthetic MF spike train that is resident in HBM before the timed
region (CS on a seeded 3% of MFs at 100 Hz in 500-1000 ms of every 2,000 ms
trial, the US as error drive at 999 ms).  Prints one JSON line.

The timed region also carries the outputs a driver reads each ms: the
device spike recorder stores every step's PC, NC and IO spikes, and the
recorded rows are copied to the host before the region ends.

roofline: the step's dominant kernel -- the resident kernel when the timed
run is resident (the whole run in one launch: GR state in registers, small
populations in LDS), else the fused step kernel (one launch per step: the GR
tiles plus one block for the small populations), or the GR kernel when the
two-stream step is used -- with the GR algorithmic bytes per step (numGR x
104 B at fan-ins (4,3,2), 120 B at (4,3,4); the compulsory per-GR bytes with
the reference's data types) divided by its mean duration per step, timed
with HIP events on the network's own stream around the launches of the timed
region, against MI355X's 8 TB/s.  A resident run does not stream those bytes
each step, so its frac can pass 1; per_step_launch gives the fused step
launched once per step on the same roofline.  traffic is the HBM bytes per
launch (per step inside a resident run) from the FETCH_SIZE / WRITE_SIZE
counters (profiles/traffic_<config>[_resident].json), and traffic_frac the
bandwidth those bytes give over the same duration.  The GR and
small-population kernels launched one after the other are timed too
(gr_only_us, small_only_us).

cpu_baseline: the CPU oracle (a C restatement of the same step), timed on
this host over a bounded sample of the same workload with OpenMP over the
host's cores (value) and with one thread (single_thread).

batched: the 256-network ragged pack (64 x 16K + 96 x 65K + 64 x 262K + 32 x
1M GR, each with its own seed, CS subset and GO-GO inhibition on or off)
stepped by one launch per ms, in network-ms per wall-second, against the same
networks run one at a time on the same GPU ("plain").

surrogate: config 4 -- the learned input layer trained on 2,000,000 windows of
the 1M-GR network, the sweep of 64 plasticity parameter sets it drives
(network-ms/s against the exact step), and the exact step's verdict on PF
rates and conditioned responses.
"""
import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "simulated ms per wall-second, 1M-GR net; batched network-ms/s vs plain"
HBM_PEAK_GBS = 8000.0
WORKLOADS = {
    "c1": "1M-GR eyelid network: 1,048,576 GR; 1,024 GO; 1,024 MF; 32 PC (32,768 PF each); 128 BC; "
          "512 SC; 4 IO; 8 NC; 4 MF + 3 GO inputs and 2 GO targets per GR, seed 0; 2,000 ms trials",
    "c2": "4M-GR dense network: 4,194,304 GR; 4,096 GO; 4,096 MF; 32 PC; 128 BC; 512 SC; 4 IO; 8 NC; "
          "4 GO targets per GR in a GO-grid window",
    "c0": "16K-GR network: 16,384 GR; 256 GO; 256 MF; 8 PC; 32 BC; 128 SC; 1 IO; 2 NC",
}


def bytes_per_gr(n_out: int) -> int:
    # 5 x FP32 state r+w (40) + apBuf r+w (8) + 4*(nMF+nGOin) indices (28)
    # + 8*nGOout target+mask (16 or 32) + PF delay mask (4) + weight read (4)
    # + (history r+w + weight write)/5 (4)
    return 40 + 8 + 28 + 8 * n_out + 4 + 4 + 4


def host_cpu() -> dict:
    """CPU model, logical CPUs, the ones this process may run on, and
    physical cores (distinct physical id / core id pairs of /proc/cpuinfo)."""
    model, cores, phys, core = platform.processor() or platform.machine(), set(), None, None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                k, _, v = line.partition(":")
                k, v = k.strip(), v.strip()
                if k == "model name":
                    model = v
                elif k == "physical id":
                    phys = v
                elif k == "core id":
                    core = v
                elif not k and phys is not None:
                    cores.add((phys, core))
        if phys is not None:
            cores.add((phys, core))
    except OSError:
        pass
    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return {"model": model, "logical_cpus": os.cpu_count(), "usable_cpus": usable,
            "physical_cores": len(cores) or None}


def cpu_baseline(cfg: str, sample_steps: int, omp_steps: int):
    """The binary32 oracle on the host: OpenMP over the host's cores and one
    thread, each over a bounded sample of the workload's steps."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import orc
    from cerebellum_simulation_amd import configs

    st = configs.build_state(cfg, seed=0)
    host = host_cpu()
    out = {}
    for openmp, steps in ((True, omp_steps), (False, sample_steps)):
        ref = orc.Oracle.from_state(st, precision=32, openmp=openmp)
        tr = configs.mf_train(st.sizes.numMF, steps, seed=11)
        t0 = time.perf_counter()
        ref.run_mf_train(tr)
        dt = time.perf_counter() - t0
        threads = ref.L.orc_threads() if openmp else 1
        out[openmp] = (steps / dt, dt, steps, threads)
        ref.close()
    v, dt, n, thr = out[True]
    v1, dt1, n1, _ = out[False]
    return {"value": round(v, 3), "unit": "sim-ms/s", "cores": thr, "kind": "port",
            "sample": f"{n} steps of the {cfg} network from its initial state, oracle binary32 build with OpenMP "
                      f"(oracle/build/liborc_omp.so, {thr} threads), {dt:.2f} s",
            "single_thread": {"value": round(v1, 3), "unit": "sim-ms/s", "cores": 1,
                              "sample": f"{n1} steps, oracle/build/liborc.so, {dt1:.2f} s"},
            "host": host}


def batch_bench(which: str, steps: int) -> dict:
    """Network-ms/s of the ragged batch (one launch per step for all networks)
    and of the same networks stepped one at a time."""
    import torch

    from cerebellum_simulation_amd import Batch, Network, configs, stream_ceiling

    mix = configs.BATCH_MIX if which == "full" else [(c, max(1, n // 8)) for c, n in configs.BATCH_MIX]
    plan = configs.batch_plan(mix)
    t_build = time.perf_counter()
    nets, trains = [], []
    warm, plain_steps = 10, max(20, steps // 4)
    graph_steps = 20
    total = warm + 2 * 8 + graph_steps + steps + plain_steps
    for i, (cls, seed, inh) in enumerate(plan):
        st = configs.build_state(cls, seed=seed, gogo_inhibition=inh)
        nets.append(Network.from_state(st))
        trains.append(torch.from_numpy(configs.mf_train(st.sizes.numMF, total, seed=seed)).to("cuda:0"))
    gr_total = sum(n.sizes.numGR for n in nets)
    print(f"[bench] batch of {len(nets)} networks ({gr_total} GR) built in {time.perf_counter() - t_build:.1f} s",
          file=sys.stderr, flush=True)
    b = Batch(nets, graph_steps=graph_steps)
    b.run_mf_train([t[:warm] for t in trains])
    cpt = b.autotune([t[warm:warm + 8] for t in trains], steps=8)
    # one window of warm-up steps captures the step graph for the chosen shape
    b.run_mf_train([t[warm + 8:warm + 8 + graph_steps] for t in trains])
    k0 = warm + 8 + graph_steps
    b.sync()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    b.run_mf_train([t[k0:k0 + steps] for t in trains])
    b.sync()
    el = time.perf_counter() - t0
    batched = len(nets) * steps / el
    k1 = k0 + steps
    # the same networks, one at a time (each network's own fused launch per step)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for n, t in zip(nets, trains):
        n.run_mf_train(t[k1:k1 + plain_steps])
        n.sync()
    el_plain = time.perf_counter() - t0
    plain = len(nets) * plain_steps / el_plain
    alg = sum(n.sizes.numGR * bytes_per_gr(n.sizes.maxnumpGRfromGRtoGO) for n in nets)
    step_us = el / steps * 1e6
    b.close()
    ceil_us, ceil_bytes = stream_ceiling(gr_total, launches=20)
    return {"networks": len(nets), "gr_total": gr_total, "mix": [list(m) for m in mix],
            "value": round(batched, 1), "unit": "network-ms/s", "plain": round(plain, 1),
            "speedup_vs_plain": round(batched / plain, 2), "steps": steps, "plain_steps": plain_steps,
            "step_us": round(step_us, 2), "cells_per_thread": cpt, "graph_steps": graph_steps,
            "stream_ceiling": {"us": round(ceil_us, 2), "bytes": ceil_bytes,
                               "GB/s": round(ceil_bytes / (ceil_us * 1e-6) / 1e9, 1),
                               "step_vs_ceiling": round(ceil_us / step_us, 4)},
            "roofline_frac": round(alg / (step_us * 1e-6) / 1e9 / HBM_PEAK_GBS, 4),
            "algorithmic_bytes_per_step": alg}


MFMA_BF16_PEAK_TFLOPS = 2500.0


def surrogate_bench(which: str) -> dict:
    """Config 4: the learned input layer (3-layer causal temporal conv, kernel
    8, 128 channels, BF16 MFMA with FP32 accumulation) trained on windows the
    1M-GR simulator generates, driving a sweep of 64 LTD/LTP parameter sets,
    judged by the exact step (surrogate.sweep_experiment)."""
    from cerebellum_simulation_amd import surrogate as SG
    log = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731
    if which == "kernels":
        return surrogate_kernels()
    if which == "full":
        res = SG.sweep_experiment("c1", train_trials=1000, train_steps=600, sweep_trials=10, log=log)
        workload = ("1M-GR network (config 1 sizes); 2,000,000 training windows of 50 ms x 1,024 MFs -> "
                    "4,096 clusters of 256 GRs; sweep of 64 (LTD, LTP) scale pairs x 10 trials")
    else:
        res = SG.sweep_experiment("b262k", train_trials=10, train_steps=250, segments=32, sweep_trials=2,
                                  ltd_scales=(1, 8), ltp_scales=(1, 8), judged=(0, 3), cells=64, log=log)
        workload = "262K-GR network; 20,000 windows -> 4,096 clusters of 64 GRs; 4 parameter sets x 2 trials"
    tf = res["train_step_tflops"]
    keep = ("windows", "record_s", "train_s", "train_step_us", "loss_first", "loss_last", "forward_tflops",
            "members", "sweep_s", "sweep_network_ms_per_s", "exact_network_ms_per_s", "speedup_vs_exact",
            "pf_rmse", "pf_rmse_cluster_mean", "judged", "surrogate_vs_exact", "cluster_mean_vs_exact",
            "exact_fractions_vs_exact", "cells_per_cluster")
    out = {"workload": workload, **{k: res[k] for k in keep}}
    out["roofline"] = {"bound": "mfma", "achieved": round(tf, 1), "peak": MFMA_BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                       "frac": round(tf / MFMA_BF16_PEAK_TFLOPS, 4), "kernel": "training step (all kernels)"}
    for k in ("record_s", "train_s", "train_step_us", "sweep_s", "sweep_network_ms_per_s", "exact_network_ms_per_s",
              "speedup_vs_exact", "pf_rmse", "pf_rmse_cluster_mean", "forward_tflops"):
        out[k] = round(out[k], 4)
    return out


def surrogate_kernels() -> dict:
    """The surrogate's kernels alone at the full shape, on synthetic windows
    (Bernoulli MF rows, random soft targets): 20 training steps of 64 x 256
    windows, one forward pass over 64 x 256 windows, and a sweep of 16
    members of the 1M-GR network over 2,000 steps.  For kernel profiles."""
    import torch
    from cerebellum_simulation_amd import Batch, Network, configs
    from cerebellum_simulation_amd import surrogate as SG
    gen = torch.Generator(device="cuda").manual_seed(0)
    rows = 40000
    mf = (torch.rand((rows, 1024), device="cuda", generator=gen) < 0.05).to(torch.uint8)
    y = (torch.randint(0, 20, (rows, 4096), device="cuda", generator=gen).float() / 256).to(torch.bfloat16)
    sg = SG.Surrogate(num_in=1024, hidden=128, num_out=4096, max_rows=20000)
    sg.init_params(0, -4.0)
    data = SG.InputLayerData(mf, y.view(torch.int16), 256)
    SG.train(sg, data, 20, segments=64, outputs=256)
    tr_us, tr_fl = sg.last_timing()
    sg.predict(mf, 64, 16384, outputs=256)
    fw_us, fw_fl = sg.last_timing()
    st = configs.build_state("c1", seed=0)
    nets = [Network.from_state(st) for _ in range(16)]
    b = Batch(nets)
    prob = torch.full((199 + 2000, 4096), 0.004, device="cuda")
    train = torch.from_numpy(configs.mf_train(st.sizes.numMF, 2000, seed=1)).cuda()
    t0 = time.perf_counter()
    b.sweep(prob, 199, 2000, train)
    b.sync()
    sw = time.perf_counter() - t0
    b.close()
    for n in nets:
        n.close()
    sg.close()
    return {"workload": "surrogate kernels, synthetic windows: training 64 x 256 windows (1,024 MF -> 128 -> 128 "
                        "-> 4,096), forward 64 x 256 windows, sweep of 16 x 1M-GR members x 2,000 steps",
            "train_step_us": round(tr_us, 1), "train_tflops": round(tr_fl / tr_us / 1e6, 1),
            "forward_us": round(fw_us, 1), "forward_tflops": round(fw_fl / fw_us / 1e6, 1),
            "sweep_step_us": round(sw / 2000 * 1e6, 2)}


def cartpole_bench(which: str) -> dict:
    """The cart-pole with two microzones in closed loop (include/cbm_env.h):
    per step the environment's MF rates, MF spikes from PoissonRegenCells,
    one network step, the NC spikes of both zones read back and the
    environment's step, all through the library's native loop."""
    from cerebellum_simulation_amd import Network, configs
    from cerebellum_simulation_amd import cartpole as CP
    name = "cp1m" if which == "full" else "t8k2z"
    st = CP.build_state(configs.act_params_text(), configs.conparams_text(**configs.CONFIGS[name][0]), seed=1)
    net = Network.from_state(st)
    env = CP.Cartpole(net, env_seed=1, mf_seed=1, closed_loop=True, max_num_trials=10**6)
    env.run(1600, trace=False)  # past the first time-out: the pole is aloft from here on
    net.sync()
    steps = 4000
    t0 = time.perf_counter()
    done = env.run(steps, trace=False)
    net.sync()
    el = time.perf_counter() - t0
    t0 = time.perf_counter()
    net.calc_activity(steps)
    net.sync()
    step_only = time.perf_counter() - t0
    res_only = None
    if net.resident:
        # the same network's step inside one resident run, over a synthetic MF train
        import torch
        dtr = torch.from_numpy(configs.mf_train(st.sizes.numMF, steps, seed=2)).to("cuda")
        net.sync()
        t0 = time.perf_counter()
        net.run_mf_train(dtr)
        net.sync()
        res_only = round((time.perf_counter() - t0) / steps * 1e6, 2)
    s = st.sizes
    out = {"workload": (f"cart-pole, two microzones, closed loop; {s.numGR:,} GR, {s.numMF:,} MF, {s.numNC} NC "
                        "per zone (the sizes of CBM_Params/conParams.txt)" if which == "full" else
                        f"cart-pole, two microzones, closed loop; {s.numGR:,} GR"),
           "steps": done, "sim_ms_per_s": round(done / el, 1), "us_per_step": round(el / done * 1e6, 2),
           "network_step_only_us": round(step_only / steps * 1e6, 2),
           "network_resident_step_us": res_only, "resident_loop": bool(net.resident)}
    env.close()
    net.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--config", default="c1")
    ap.add_argument("--profile-steps", type=int, default=400)
    ap.add_argument("--cpu-steps", type=int, default=0, help="one-thread CPU sample in steps (0 = pick ~10 s)")
    ap.add_argument("--cpu-omp-steps", type=int, default=0, help="OpenMP CPU sample in steps (0 = pick)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--step-mode", type=int, default=0, help="0 pick, 1 fused launch (even GR split when it fits, else tiles), 2 two streams, 3 fused with GR tiles, 4 fused with the even split")
    ap.add_argument("--cells-per-thread", type=int, default=0)
    ap.add_argument("--block-threads", type=int, default=0)
    ap.add_argument("--batch", default="full", choices=["full", "small", "none"],
                    help="ragged batch to measure beside the single network (small: 1/8 of each class)")
    ap.add_argument("--batch-steps", type=int, default=200)
    ap.add_argument("--surrogate", default="full", choices=["full", "small", "kernels", "none"],
                    help="config 4: the learned input layer and the plasticity sweep it drives")
    ap.add_argument("--cartpole", default="full", choices=["full", "small", "none"],
                    help="the cart-pole closed loop on two microzones (full: 1M-GR network, 2,048 MF)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB, 5 times (the known-size kernel that scales "
                         "the FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        n = args.calibrate_copy * 2**20 // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda:0")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes_read": n * 4, "calibrate_copy_bytes_written": n * 4}))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the network step runs "
              "on a single MI355X; a pack of independent networks would spread over GPUs by GR count, "
              "one process per GPU)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from cerebellum_simulation_amd import Network, configs, stream_ceiling

    t_build = time.perf_counter()
    st = configs.build_state(args.config, seed=0)
    net = Network.from_state(st, step_mode=args.step_mode, cells_per_thread=args.cells_per_thread,
                             block_threads=args.block_threads)
    s = st.sizes
    n = args.warmup + args.steps
    tr = configs.mf_train(s.numMF, n + args.profile_steps, seed=1)
    err = configs.us_err_drive(n + args.profile_steps, st.num_zones)
    dtr = torch.from_numpy(tr).to("cuda:0")
    torch.cuda.synchronize()
    print(f"[bench] {args.config}: built in {time.perf_counter() - t_build:.1f} s, "
          f"{net.device_bytes() / 2**20:.0f} MiB on device", file=sys.stderr, flush=True)

    rec = ("PC", "NC", "IO")
    net.set_recorder(rec, args.steps)
    net.run_mf_train(dtr[: args.warmup], err_drive=err[: args.warmup])
    net.sync()
    torch.cuda.synchronize()
    # HIP events on the network's own stream, around the timed launches
    stream = torch.cuda.ExternalStream(net.stream_ptr())
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t_first = int(net.read("CUR_TIME")[0]) + 1
    t0 = time.perf_counter()
    ev0.record(stream)
    net.run_mf_train(dtr[args.warmup:n], err_drive=err[args.warmup:n])
    ev1.record(stream)
    out_rows = {p: net.read_recorded(p, t_first, args.steps) for p in rec}  # the outputs, on the host
    net.sync()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    k_us = ev0.elapsed_time(ev1) * 1000.0 / args.steps
    sim_ms_per_s = args.steps / elapsed
    spikes = {p: int(v.sum()) for p, v in out_rows.items()}
    print(f"[bench] timed {args.steps} steps in {elapsed:.3f} s ({k_us:.3f} us per step on the GPU); "
          f"recorded spikes {spikes}", file=sys.stderr, flush=True)

    # the same launches alone, and the two halves one after the other
    half = args.profile_steps // 2
    k_us_alone, step_us = net.profile_steps(dtr[n:n + half])
    gr_only, small_only = net.profile_split(dtr[n + half:n + args.profile_steps])
    n_out = s.maxnumpGRfromGRtoGO
    alg_bytes = s.numGR * bytes_per_gr(n_out)
    achieved = alg_bytes / (k_us * 1e-6) / 1e9
    fused = net.step_mode != 2
    resident = net.resident
    launch = net.gr_launch
    step_kernel = ("step_span_kernel" if launch["even_split"] else "step_kernel") if fused else "gr_step_kernel"
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "kernel": "step_resident_kernel" if resident else step_kernel, "gr_launch": launch,
                "kernel_us": round(k_us, 3),
                "kernel_us_note": "HIP event pair on the network's stream around the timed region's launches, "
                                  "divided by the steps",
                "algorithmic_bytes_per_step": alg_bytes, "gr_only_us": round(gr_only, 3),
                "small_only_us": round(small_only, 3),
                "gr_only_frac": round(alg_bytes / (gr_only * 1e-6) / 1e9 / HBM_PEAK_GBS, 4)}
    # the step launched once per step (what cbm_step and the per-ms drivers
    # run): HBM-bound, its own roofline
    roofline["per_step_launch"] = {"kernel": step_kernel, "us": round(k_us_alone, 3),
                                   "achieved": round(alg_bytes / (k_us_alone * 1e-6) / 1e9, 1),
                                   "frac": round(alg_bytes / (k_us_alone * 1e-6) / 1e9 / HBM_PEAK_GBS, 4)}
    if resident:
        roofline["note"] = ("the timed steps run as one launch whose GR blocks keep their cells' state "
                            "in registers and LDS and whose two small-population blocks keep their populations in "
                            "LDS: the per-GR arrays are not streamed from HBM each step, so the algorithmic-byte "
                            "rate is not bounded by the HBM peak (traffic: the measured fabric bytes of one step "
                            "inside such a run); the step is bound by the chain of dependent memory round trips of "
                            "the small-population blocks and the GR blocks and the signalling between them")
    ceil_us, ceil_bytes = stream_ceiling(s.numGR)
    roofline["stream_ceiling"] = {"us": round(ceil_us, 3), "bytes": ceil_bytes,
                                  "GB/s": round(ceil_bytes / (ceil_us * 1e-6) / 1e9, 1),
                                  "kernel_vs_ceiling": round(ceil_us / k_us, 4)}
    traffic_file = os.path.join(ROOT, "profiles", f"traffic_{args.config}{'_resident' if resident else ''}.json")
    if os.path.exists(traffic_file):
        with open(traffic_file) as f:
            tr = json.load(f)
        # a resident run: the bytes of one step inside the run
        tr_bytes = tr.get("hbm_bytes_per_step") if resident else tr.get("hbm_bytes_per_launch")
        roofline["traffic"] = tr_bytes
        if tr_bytes:
            roofline["traffic_achieved"] = round(tr_bytes / (k_us * 1e-6) / 1e9, 1)
            roofline["traffic_frac"] = round(tr_bytes / (k_us * 1e-6) / 1e9 / HBM_PEAK_GBS, 4)

    batched = None
    if args.batch != "none":
        del dtr
        batched = batch_bench(args.batch, args.batch_steps)

    surrogate = None
    if args.surrogate != "none":
        try:
            surrogate = surrogate_bench(args.surrogate)
        except Exception as e:  # the headline line must still be printed
            surrogate = {"error": f"{type(e).__name__}: {e}"}

    cartpole = None
    if args.cartpole != "none":
        try:
            cartpole = cartpole_bench(args.cartpole)
        except Exception as e:  # the headline line must still be printed
            cartpole = {"error": f"{type(e).__name__}: {e}"}

    cpu = None
    if not args.no_cpu:
        cpu_steps = args.cpu_steps or {"c0": 5000, "c1": 1200, "c2": 300}.get(args.config, 100)
        omp_steps = args.cpu_omp_steps or {"c0": 5000, "c1": 4000, "c2": 1000}.get(args.config, 200)
        cpu = cpu_baseline(args.config, cpu_steps, omp_steps)

    out = {
        "metric": METRIC, "value": round(sim_ms_per_s, 2), "unit": "sim-ms/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(1000.0 * elapsed / args.steps, 5),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic: seeded Bernoulli MF trains, seeded synthetic connectivity; PC, NC and IO spikes of "
                "every step recorded and copied to the host in the timed region",
        "config": {"workload": WORKLOADS.get(args.config, args.config), "name": args.config,
                   "numGR": s.numGR, "numGO": s.numGO, "numMF": s.numMF, "numPC": s.numPC,
                   "parallelism": "single GPU",
                   "step_mode": "resident" if resident else "fused" if fused else "two streams"},
        "roofline": roofline, "cpu_baseline": cpu, "batched": batched, "surrogate": surrogate,
        "cartpole": cartpole,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

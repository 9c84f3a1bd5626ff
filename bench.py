#!/usr/bin/env python3
"""Benchmark of the MI355X DPD step engine.

Default workload: the polymer condensate box (64^3 cells at rho = 3, 786,432 beads;
16-mer (S P P P)x4 chains at 20% of the beads, a_SS = 10, other a = 25, gamma = 4.5,
Hookean bonds k = 128, l0 = 0.5, dt = 0.02): the single-box throughput run.  The state is
synthetic (uniform water, random-walk chains, Maxwell velocities) and is relaxed before
the warm-up steps (steps with periodic velocity rescaling to kT, then free steps).  One "step" is one full CSimBox::Evolve equivalent (sort, pair forces,
bonds, integration) of the whole box; inputs stay in HBM.

Prints one JSON line.  `value` is bead-steps/s on the GPU; `roofline` is the pair kernel
(the dominant kernel, timed as back-to-back launches between two HIP events) against HBM
bandwidth; `batched` is 256 replicas advanced together
against the same boxes one at a time; `other_configs` the water 20^3, binary 40^3 and
bilayer 48x48x32 boxes; `cpu_baseline` times the compiled reference (oracle/_ref/dpd) on
one host core (and, under `all_cores`, one copy per core of the job's CPU share at once) on
the condensate configuration, by differencing two run lengths so that set-up cancels.
"""
import argparse
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "osprey-dpd_amd"))

import numpy as np  # noqa: E402

import dpd_amd as D  # noqa: E402
from dpd_amd import tune as T  # noqa: E402
from dpd_amd import workloads as W  # noqa: E402

METRIC = "bead-steps/s at rho=3 (1 box and batched replicas) vs host CPU; force rel err"
UNIT = "bead-steps/s"
BYTES_PER_BEAD_STEP = 76     # FP32 x, v, f_old read + type; x, v, f written
BYTES_PER_BOND_STEP = 8
BYTES_PER_ANGLE_STEP = 12
HBM_PEAK_GBS = 8000.0
# useful FP32 flops per bead-step, reported beside the byte roofline: ~52 per interacting
# pair x 6.1 pairs + ~9 per rejected candidate x 34 candidates
FLOPS_PER_BEAD_STEP = 52 * 6.1 + 9 * 34

CONDENSATE_DMPCI = """dpd

Title	" condensate "
Date    02/12/26
Comment	" 16-mer (S P P P)x4 chains at 20% of the beads in water "

State	random

Bead  W
      0.5
      25
      4.5

Bead  S
      0.5
      25 10
      4.5 4.5

Bead  P
      0.5
      25 25 25
      4.5 4.5 4.5

Bond    S P  128  0.5
Bond    P P  128  0.5

Polymer	Water    0.984615385   " (W) "
Polymer	Chain    0.015384615   " (S (3 (P P P S)) P P P) "

Box         64 64 64       1  1  1
Density		3
Temp        1
RNGSeed		-20261202
Lambda		0.5
Step		0.02
Time		{T}
SamplePeriod     {T}
AnalysisPeriod	 {T}
DensityPeriod    {T}
DisplayPeriod    {T}
RestartPeriod    {T}
Grid		1  1  1
"""


def build_engine(wl, kernel):
    e = D.Engine(wl.cells, wl.a, wl.gamma, dt=wl.dt, seed=wl.seed, pair_kernel=kernel)
    e.upload(wl.ids, wl.types, wl.x, wl.v, unpbc=wl.u)
    if wl.bonds is not None or wl.pairs is not None:
        h, t, k, l0 = wl.bonds if wl.bonds is not None else ((), (), (), ())
        fi, se, mo, ph = wl.pairs if wl.pairs is not None else ((), (), (), ())
        e.upload_topology(h, t, k, l0, fi, se, mo, ph)
    if kernel == D.PAIR_AUTO:
        # the schedule the autotuner recorded for this shape (osprey-dpd_amd/tune.json)
        T.apply(e)
    return e


def _warm_graphs(e, warmup, steps):
    """Untimed: `warmup` steps, then one pass of the timed call itself, so that the graphs it
    replays (chunks of steps, the last one observing) are captured before the clock starts."""
    e.step(warmup, graph=True)
    e.step(steps, observe=True, graph=True)


def bench_batched(nbox=256, steps=100, warmup=20, nsingle=16, seed=20261202, pair_kernel=D.PAIR_AUTO):
    """Batched replicas: `nbox` water/lipid boxes with side lengths L ~ U{8..24}, a_ij ~
    U[20,90], k ~ U[64,256], advanced together (one graph replay per 50 steps), against
    the same kind of box advanced alone, one box per launch sequence."""
    from dpd_amd import surrogate as S
    rng = np.random.default_rng(seed)
    params = S.sample_params(nbox, rng)
    sizes = rng.integers(8, 25, size=nbox)
    seeds = rng.integers(1, 2 ** 31, size=nbox)
    e, meta = S.build_batch(params, sizes, seeds, pair_kernel=pair_kernel)
    if pair_kernel == D.PAIR_AUTO:
        T.apply(e)
    _warm_graphs(e, warmup, steps)
    t0 = time.perf_counter()
    e.step(steps, observe=True, graph=True)
    wall = time.perf_counter() - t0
    batched = e.n * steps / wall
    nb_total = e.n
    del e
    # the same kind of boxes one at a time (a sample of nsingle boxes)
    tot_bs, tot_t = 0.0, 0.0
    for i in range(nsingle):
        es, _ = S.build_batch(params[i:i + 1], sizes[i:i + 1], seeds[i:i + 1], pair_kernel=pair_kernel)
        if pair_kernel == D.PAIR_AUTO:
            T.apply(es)
        _warm_graphs(es, warmup, steps)
        t0 = time.perf_counter()
        es.step(steps, observe=True, graph=True)
        tot_t += time.perf_counter() - t0
        tot_bs += es.n * steps
        del es
    single = tot_bs / tot_t
    return {"boxes": nbox, "beads": int(nb_total), "steps": steps, "value": batched, "unit": UNIT,
            "ms_per_step": wall / steps * 1e3, "single_box_value": single, "single_box_sample": nsingle,
            "speedup_vs_single_box": batched / single,
            "workload": "256 water/lipid replicas, L~U{8..24}, rho=3, a_ij~U[20,90], k~U[64,256]"}


def bench_configs(steps=100, warmup=20):
    """The other single-GPU configurations (water 20^3, binary 40^3, bilayer 48x48x32):
    bead-steps/s with the automatic pair-kernel choice, graph replay, inputs resident."""
    out = {}
    for wl in (W.water(), W.binary(), W.bilayer()):
        e = build_engine(wl, D.PAIR_AUTO)
        _warm_graphs(e, warmup, steps)
        t0 = time.perf_counter()
        e.step(steps, observe=True, graph=True)
        wall = time.perf_counter() - t0
        o = e.observables()
        out[wl.name] = {"beads": wl.n, "bonds": wl.nbonds, "angles": wl.nangles, "steps": steps,
                        "value": wl.n * steps / wall, "ms_per_step": wall / steps * 1e3,
                        "pair_kernel": e.pair_kernel, "schedule": D.schedule_name(e.schedule), "kT": o["kT"],
                        "pressure": o["pressure"]}
        e.close()
    return out


def bench_mlp(nrows=1638, iters=1000, warm=50, seed=3):
    """The surrogate network's training iteration (7 -> 256 -> 256 -> 2: BF16 matrix-core
    forward and backward, FP32 Adam) on a synthetic full batch the size of the replica
    dataset's training split: ms per iteration and the rate of its eight matrix products
    (FLOPs of the padded shapes: inputs and outputs padded to 16)."""
    from dpd_amd.mlp import Surrogate
    rng = np.random.default_rng(seed)
    nin, H, nout = 7, 256, 2
    X = rng.normal(size=(nrows, nin))
    Y = rng.normal(size=(nrows, nout))
    net = Surrogate(nin, H, nout, seed=1)
    net.set_scaling(np.zeros(nin), np.ones(nin), np.zeros(nout), np.ones(nout))
    net.train(X, Y, iters=warm, lr=1e-3)
    t0 = time.perf_counter()
    net.train(X, Y, iters=iters, lr=1e-3)
    wall = time.perf_counter() - t0
    net.close()
    mp, ki, ko = (nrows + 15) // 16 * 16, 16, 16
    flops = 2 * mp * (2 * ki * H + 3 * H * H + 3 * H * ko)
    ms = wall / iters * 1e3
    return {"rows": nrows, "iters": iters, "ms_per_iter": ms, "gemm_flops_per_iter": flops,
            "gemm_tflops": flops / (ms * 1e-3) / 1e12, "net": f"{nin}-{H}-{H}-{nout}"}


def cpu_baseline_reference(t1=2, t2=12):
    """Time the reference binary on the condensate configuration: run lengths t1 and t2
    concurrently (one core each); per-step time = (wall2 - wall1)/(t2 - t1)."""
    exe = os.path.join(ROOT, "oracle", "_ref", "dpd")
    if not os.path.exists(exe):
        return None
    work = tempfile.mkdtemp(prefix="dpd_cpu_")
    try:
        procs = []
        for tag, T in (("a", t1), ("b", t2)):
            with open(os.path.join(work, "dmpci." + tag), "w") as fh:
                fh.write(CONDENSATE_DMPCI.format(T=T))
            procs.append((tag, T, time.perf_counter(),
                          subprocess.Popen([exe, tag], cwd=work, stdout=subprocess.DEVNULL,
                                           stderr=subprocess.DEVNULL)))
        walls = {}
        for tag, T, t0, p in procs:
            rc = p.wait(timeout=600)
            if rc != 0:
                return None
            walls[tag] = time.perf_counter() - t0
        with open(os.path.join(work, "dmpcis.a")) as fh:
            txt = fh.read()
        nbeads = int(txt.split("Number of beads created")[1].split()[0])
        per_step = (walls["b"] - walls["a"]) / (t2 - t1)
        if per_step <= 0:
            return None
        return {"value": nbeads / per_step, "unit": UNIT, "cores": 1, "kind": "reference",
                "sample": f"reference dpd binary (oracle/_ref) on the condensate box, {nbeads} beads; "
                          f"runs of {t1} and {t2} steps in parallel, differenced "
                          f"({walls['a']:.1f} s, {walls['b']:.1f} s wall)"}
    except Exception:
        return None
    finally:
        shutil.rmtree(work, ignore_errors=True)


def cpu_all_cores_reference(t1=2, t2=8, copies=None):
    """Whole-host throughput of the reference binary: `copies` independent condensate runs
    at once (one per core), first all of length t1, then all of length t2; per-step time
    under the same contention = (mean wall t2 - mean wall t1)/(t2 - t1)."""
    exe = os.path.join(ROOT, "oracle", "_ref", "dpd")
    if not os.path.exists(exe):
        return None
    if copies is None:
        copies = cpu_share()
    work = tempfile.mkdtemp(prefix="dpd_cpuall_")
    try:
        mean_wall = {}
        for T in (t1, t2):
            procs = []
            for c in range(copies):
                tag = f"c{c}t{T}"
                with open(os.path.join(work, "dmpci." + tag), "w") as fh:
                    fh.write(CONDENSATE_DMPCI.format(T=T))
                procs.append((time.perf_counter(), subprocess.Popen([exe, tag], cwd=work, stdout=subprocess.DEVNULL,
                                                                    stderr=subprocess.DEVNULL)))
            walls = []
            for t0, p in procs:
                if p.wait(timeout=900) != 0:
                    return None
                walls.append(time.perf_counter() - t0)
            mean_wall[T] = float(np.mean(walls))
            for name in os.listdir(work):
                if not name.startswith(("dmpci", "dmpcis")):
                    os.remove(os.path.join(work, name))
        with open(os.path.join(work, f"dmpcis.c0t{t1}")) as fh:
            nbeads = int(fh.read().split("Number of beads created")[1].split()[0])
        per_step = (mean_wall[t2] - mean_wall[t1]) / (t2 - t1)
        if per_step <= 0:
            return None
        return {"value": copies * nbeads / per_step, "unit": UNIT, "cores": copies,
                "sample": f"{copies} concurrent reference runs of the condensate box, {t1} then {t2} steps, "
                          f"differenced (mean wall {mean_wall[t1]:.1f} s, {mean_wall[t2]:.1f} s)"}
    except Exception:
        return None
    finally:
        shutil.rmtree(work, ignore_errors=True)


def cpu_baseline_port(wl, nsteps=3):
    """Fallback: the FP64 oracle's serial step on the same workload."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    box = O.Box(wl.cells, wl.a, wl.gamma, dt=wl.dt, seed=wl.seed)
    topo = O.Topo(*wl.bonds) if wl.bonds is not None else None
    s = O.System(box, wl.ids, wl.types, wl.x, wl.v, unpbc=wl.u, topo=topo)
    s.run(1)
    t0 = time.perf_counter()
    s.run(nsteps, step0=1)
    dt = time.perf_counter() - t0
    return {"value": wl.n * nsteps / dt, "unit": UNIT, "cores": 1, "kind": "port",
            "sample": f"FP64 oracle step, {wl.n} beads, {nsteps} steps after 1 warm-up step"}


def _loaded_hip_runtime():
    """The HIP runtime library this process already mapped (the engine links it)."""
    import ctypes as C
    with open("/proc/self/maps") as fh:
        for line in fh:
            path = line.split()[-1]
            if "libamdhip64.so" in path:
                return C.CDLL(path)
    return C.CDLL("libamdhip64.so")


def measured_copy_gbs(nbytes=1 << 30, reps=10):
    """HBM bandwidth of a plain device-to-device copy (read + write bytes per second),
    through the engine's own HIP runtime (a second runtime in the process, such as
    torch's, would not see the GPU)."""
    import ctypes as C
    hip = _loaded_hip_runtime()
    a, b, e0, e1 = C.c_void_p(), C.c_void_p(), C.c_void_p(), C.c_void_p()
    n = C.c_size_t(nbytes)
    d2d = 3   # hipMemcpyDeviceToDevice
    ok = (hip.hipMalloc(C.byref(a), n) == 0 and hip.hipMalloc(C.byref(b), n) == 0 and
          hip.hipMemset(a, 1, n) == 0 and hip.hipEventCreate(C.byref(e0)) == 0 and
          hip.hipEventCreate(C.byref(e1)) == 0 and hip.hipMemcpy(b, a, n, d2d) == 0)
    ms = C.c_float(0.0)
    if ok:
        hip.hipEventRecord(e0, None)
        for _ in range(reps):
            hip.hipMemcpyAsync(b, a, n, d2d, None)
        hip.hipEventRecord(e1, None)
        ok = hip.hipEventSynchronize(e1) == 0 and hip.hipEventElapsedTime(C.byref(ms), e0, e1) == 0
    for p in (a, b):
        if p.value:
            hip.hipFree(p)
    for ev in (e0, e1):
        if ev.value:
            hip.hipEventDestroy(ev)
    return 2.0 * nbytes * reps / (ms.value * 1e-3) / 1e9 if ok and ms.value > 0 else None


def copy_probe(nbytes=256 << 20):
    """One plain device copy of known size: what FETCH_SIZE / WRITE_SIZE report for it
    scales the counters of the pair kernel to bytes."""
    import torch
    a = torch.ones(nbytes // 4, device="cuda", dtype=torch.float32)
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": nbytes}))


def relax(e, rounds=8, steps=50, free=100):
    """Untimed relaxation of the synthetic start (uniform positions, random-walk chains):
    `rounds` blocks of `steps` steps, each followed by a velocity rescale to kT
    (RenormaliseMomenta), then `free` steps without rescaling, so that the timed steps run
    at the thermostat's kT rather than through the start-up transient."""
    for _ in range(rounds):
        e.step(steps, graph=True)
        e.renormalise_momenta()
    e.step(free, graph=True)


def host_cpu():
    """CPU model and core counts of this host."""
    model = None
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"model": model, "logical_cpus": os.cpu_count(), "affinity": len(os.sched_getaffinity(0)),
            "share": cpu_share()}


def cpu_share():
    """Cores this process may keep busy: the affinity mask, capped by OMP_NUM_THREADS when the
    host sets it to the job's CPU share."""
    n = len(os.sched_getaffinity(0))
    try:
        omp = int(os.environ.get("OMP_NUM_THREADS", "0"))
    except ValueError:
        omp = 0
    return min(n, omp) if omp > 0 else n


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--pair-kernel", type=int, default=D.PAIR_AUTO)
    ap.add_argument("--pair-launches", type=int, default=20,
                    help="back-to-back pair launches timed for the roofline figure")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-configs", action="store_true", help="skip the water, binary and bilayer boxes")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run one 256 MiB device-to-device copy (calibrates the PMC byte counters)")
    ap.add_argument("--configs-only", action="store_true",
                    help="only the water, binary and bilayer boxes (other_configs), one JSON line")
    ap.add_argument("--mlp-only", action="store_true",
                    help="only the surrogate network's training iteration, one JSON line")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe()
        return
    if args.configs_only:
        print(json.dumps({"other_configs": bench_configs()}), flush=True)
        return
    if args.mlp_only:
        print(json.dumps({"mlp": bench_mlp()}), flush=True)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "the engine runs on one GPU: --gpus must be 1 (sharding across "
                                   "GPUs is designed, not built)"}))
        sys.exit(2)

    wl = W.condensate()
    e = build_engine(wl, args.pair_kernel)
    relax(e)
    # warm-up steps, then one untimed pass of the timed call so that its graphs are captured
    _warm_graphs(e, args.warmup, args.steps)
    kT_start = e.observables()["kT"]
    t0 = time.perf_counter()
    e.step(args.steps, observe=True, graph=True)   # synchronises on return
    wall = time.perf_counter() - t0
    obs = e.observables()
    # the dominant kernel's launch time on the state the timed steps ended at: back-to-back
    # launches between two events on the engine's stream (no markers between launches)
    t_pair = e.time_pair(args.pair_launches) * 1e-3

    bead_steps = wl.n * args.steps
    value = bead_steps / wall
    ms_per_step = wall / args.steps * 1e3
    alg_bytes = BYTES_PER_BEAD_STEP * wl.n + BYTES_PER_BOND_STEP * wl.nbonds + BYTES_PER_ANGLE_STEP * wl.nangles
    achieved = alg_bytes / t_pair / 1e9
    kname = D.PAIR_KERNEL_NAMES[e.pair_kernel]
    if e.pair_kernel == D.PAIR_HITLIST and wl.nbonds + wl.nangles > 0:
        kname = "k_pair_hitlist_bonded"   # the bonded terms run in the same launch
    consistent = t_pair * 1e3 <= ms_per_step
    if not consistent:
        print(f"bench: pair launch {t_pair * 1e3:.4f} ms exceeds the step {ms_per_step:.4f} ms",
              file=sys.stderr, flush=True)
    # HBM bytes per launch and the issue limiter from the counter runs summarised under
    # profiles/, if they are of this kernel on this workload
    traffic, limiter = None, None
    tpath = os.path.join(ROOT, "profiles", "pair_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as fh:
            tj = json.load(fh)
        if tj.get("kernel") == kname and tj.get("workload") == wl.name:
            traffic = tj.get("bytes_per_launch")
            limiter = tj.get("limiter")

    copy_gbs = measured_copy_gbs()
    batched = None if args.no_batched else bench_batched(pair_kernel=args.pair_kernel)
    configs = None if (args.no_batched or args.no_configs) else bench_configs()

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline_reference()
        if cpu is None:
            print("bench: reference binary oracle/_ref/dpd unavailable; timing the oracle port instead",
                  file=sys.stderr, flush=True)
            cpu = cpu_baseline_port(wl)
        cpu["host"] = host_cpu()
        if cpu["kind"] == "reference":
            cpu["all_cores"] = cpu_all_cores_reference()

    out = {
        "metric": METRIC, "value": value, "unit": UNIT, "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
        "config": {"workload": wl.name, "beads": wl.n, "bonds": wl.nbonds, "rho": 3, "dt": wl.dt,
                   "types": "W,S,P a_SS=10 a=25 gamma=4.5", "parallelism": "single GPU",
                   "start": "synthetic, relaxed untimed (400 steps with kT rescaling + 100 free)",
                   "kT_start": kT_start},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "limiter": limiter,
                     "measured_copy_gbs": copy_gbs,
                     "flops_per_bead_step": FLOPS_PER_BEAD_STEP,
                     "step_tflops": FLOPS_PER_BEAD_STEP * value / 1e12,
                     "kernel": kname, "pair_kernel": e.pair_kernel, "schedule": D.schedule_name(e.schedule),
                     "pair_ms_per_launch": t_pair * 1e3,
                     "pair_launches_timed": args.pair_launches,
                     "pair_within_step": consistent,
                     "alg_bytes_per_launch": alg_bytes,
                     "step_gbs": alg_bytes / (wall / args.steps) / 1e9},
        "batched": batched,
        "other_configs": configs,
        "cpu_baseline": cpu,
        "speedup_vs_cpu": (value / cpu["value"]) if cpu else None,
        "speedup_vs_cpu_all_cores": (value / cpu["all_cores"]["value"]) if cpu and cpu.get("all_cores") else None,
        "observables": {"kT": obs["kT"], "pressure": obs["pressure"], "jumps": obs["jumps"]},
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()

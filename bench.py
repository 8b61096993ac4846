"""Benchmark: FWI gradient throughput (Gpt-updates/s, forward + adjoint).

Workload (one MI355X): 2D ring full-matrix capture on a 512 x 512 grid,
dx = 0.2 mm, dt = 50 ns, nt = 1000, 128-element ring of radius 22 mm, all 128
elements transmit, Ricker 500 kHz, true model = 1500 m/s water with three
seeded inclusions (c in [1400, 1800] m/s), initial model 1500 m/s.  Observed
data is synthesised on the device.  A step is one misfit-and-gradient
evaluation over all 128 shots: 2 * N * (nt - 1) * 128 = 6.70e10 grid-point
updates.

Prints one JSON line with the batched throughput as ``value``, the plain
per-shot path (one launch per shot per time step) beside it, the roofline of
the dominant kernel (per-launch HIP-event timing) and the reference CPU core
timed on a bounded sample of the same kind of workload.
"""

from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BYTES_FWD = 16.0   # algorithmic bytes of one forward update (FP32)
BYTES_ADJ = 28.0   # adjoint + imaging update
BYTES_PER_UPDATE = 22.0


WORKLOADS = {
    "C2": "2D ring FMC 512x512, dx=0.2mm, dt=50ns, nt=1000, 128-element ring R=22mm, "
          "128 shots, Ricker 500kHz, FP32 fields; one step = one misfit+gradient over all shots",
    "C2-256": "2D ring FMC 512x512, dx=0.2mm, dt=50ns, nt=1000, 256-element ring R=22mm, "
              "256 shots, Ricker 500kHz, FP32 fields; one step = one misfit+gradient",
    "C3": "2D ring FMC 1024x1024, dx=0.1mm, dt=25ns, nt=4000, 256-element ring R=44mm, "
          "256 shots, Ricker 500kHz, FP32 fields, checkpointed; one step = one misfit+gradient",
    "C4": "3D cylinder 256^3, dx=0.25mm, dt=30ns, nt=2000, 8 rings x 64 elements R=28mm, "
          "64 shots, Ricker 500kHz, FP32 fields, no stored forward levels (the adjoint sweep "
          "runs the forward backwards; --reverse off: checkpoint every 50 steps); one step = one "
          "misfit+gradient",
    "C5": "serving mix: Poisson-arriving FWI-gradient and forward requests (50/50), grids 2D "
          "{256^2, 512^2, 1024^2} and 3D {128^3, 192^3, 256^3}, 16-256 shots, seeded phantoms, "
          "admitted by HBM footprint; value = aggregate Gpt-updates/s over the run",
}


def serving(n_requests, nt_scale, rate=150.0, seed=5, with_3d=True):
    """Run the serving mix (make_mix: seed 5, arrivals for ~80 % utilisation at
    `rate` Gpt-updates/s; the mix of 2D / 3D gradient and forward requests at
    time axes x0.1 sustains about 160 on one MI355X when all requests are
    queued at once) through the continuous-batching server; observed data of
    gradient requests is synthesised on the device first (not timed)."""
    from openusct_amd.serve import Server, fill_observed, make_mix, summarize

    reqs = make_mix(n_requests, seed=seed, nt_scale=nt_scale, rate=rate, with_3d=with_3d)
    t0 = time.perf_counter()
    fill_observed(reqs)
    print(f"[bench] serving: {n_requests} requests, observed data ready in "
          f"{time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    step = max(1, n_requests // 10)

    def progress(k, n):
        if k % step == 0 or k == n:
            print(f"[bench] serving: {k}/{n} done", file=sys.stderr, flush=True)

    res = Server().run(reqs, progress=progress)
    s = summarize(res)
    s["gradient_requests"] = sum(1 for r in reqs if r.kind == "gradient")
    s["sample"] = (f"{n_requests} requests of the serving mix (seed 5), time axes x{nt_scale}, "
                   f"Poisson arrivals for ~80 % load at {rate:g} Gpt-updates/s")
    return {k: (round(v, 4) if isinstance(v, float) else v) for k, v in s.items()}


def dataset_io(traces, prob):
    """Acquisition-file leg: the workload's own traces written as a dataset file
    (float32, gzip chunks) and read back -- into host memory by Dataset.load and
    into HBM by read_fmc -- by the native library, beside h5py (the reference's
    loader, libhdf5) reading the same file where an h5py-enabled Python exists."""
    import shutil
    import subprocess
    import tempfile

    import numpy as np
    import torch

    from openusct_amd.dataset import ArrayGeometry, Dataset, read_fmc

    threads = len(os.sched_getaffinity(0))
    geom = ArrayGeometry(np.asarray(prob.geom.element_positions), prob.geom.radius_m, 5e5,
                         getattr(prob.geom, "array_type", "ring"))
    ds = Dataset(geom, traces, 1.0 / prob.dt, prob.wavelet, 5e5, 1500.0,
                 ground_truth={"c": 1.0 / np.sqrt(prob.m_true), "h_m": prob.h})
    raw = traces.size * 4
    d = tempfile.mkdtemp(prefix="usct_io_")
    try:
        path = os.path.join(d, "acq.h5")
        t0 = time.perf_counter()
        ds.save(path, threads=threads)
        t_save = time.perf_counter() - t0
        Dataset.load(path, threads=threads)  # page cache warm, as for every reader below
        t0 = time.perf_counter()
        back = Dataset.load(path, threads=threads)
        t_load = time.perf_counter() - t0
        exact = bool(np.array_equal(back.data, traces.astype(np.float32)))
        del back
        read_fmc(path, "cuda", threads=threads)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        dev = read_fmc(path, "cuda", threads=threads)
        torch.cuda.synchronize()
        t_dev = time.perf_counter() - t0
        exact = exact and bool(torch.equal(dev.cpu(), torch.from_numpy(traces.astype(np.float32)).double()))
        del dev
        h5py_gbs = None
        py39 = "/opt/conda/bin/python3.9"
        if os.path.exists(py39):
            code = ("import sys, time, h5py\n"
                    "f = h5py.File(sys.argv[1], 'r'); f['acquisition']['data'][()]\n"
                    "t0 = time.perf_counter(); x = f['acquisition']['data'][()]\n"
                    "print(time.perf_counter() - t0)\n")
            r = subprocess.run([py39, "-c", code, path], capture_output=True, text=True, timeout=300)
            if r.returncode == 0:
                h5py_gbs = raw / float(r.stdout.split()[-1]) / 1e9
        out = {"file_bytes": os.path.getsize(path), "raw_bytes": raw, "threads": threads,
               "save_s": round(t_save, 3), "save_GBps": round(raw / t_save / 1e9, 3),
               "load_GBps": round(raw / t_load / 1e9, 3), "to_device_GBps": round(raw / t_dev / 1e9, 3),
               "h5py_load_GBps": None if h5py_gbs is None else round(h5py_gbs, 3),
               "load_vs_h5py": None if h5py_gbs is None else round(raw / t_load / 1e9 / h5py_gbs, 2),
               "bit_exact": exact,
               "what": "float32 traces of this workload; load = Dataset.load into host memory, "
                       "to_device = read_fmc into HBM as float64 (pinned double buffer)"}
    finally:
        shutil.rmtree(d, ignore_errors=True)
    return out


def workload(name="C2"):
    from openusct_amd import configs
    return configs.BUILDERS[name]()


def cpu_baseline(budget_s=25.0):
    """Time the reference's own C++ core (oracle/_ref, built from its sources) on a
    bounded sample of the benchmark workload: the 512^2 / 128-element ring problem
    (nt = 1000) with its first k shots, k sized so the timed gradient takes about
    budget_s.  FP64 with OpenMP over all cores of this process's affinity mask;
    falls back to the oracle restatement when the reference core is not built."""
    from openusct_amd import configs
    from openusct_amd.plan import linear_indices

    cores = len(os.sched_getaffinity(0))
    prob = configs.c2()
    h, dt, nt, w = prob.h, prob.dt, prob.nt, prob.wavelet
    n = prob.shape[0]
    rec = linear_indices(prob.geom.idx, prob.shape).astype(np.int32)
    kind = "reference"
    try:
        sys.path.insert(0, os.path.join(ROOT, "oracle", "_ref"))
        import _uap  # the reference's pybind11 module, compiled from its own sources

        def fwd(m, els):
            return _uap.forward_fmc(m, h, dt, nt, rec[els].copy(), rec, w)

        def grad(m, els, d):
            return _uap.misfit_and_gradient(m, h, dt, nt, rec[els].copy(), rec, w, d, 0)
    except ImportError:
        from oracle import oracle as O
        kind = "port"

        def fwd(m, els):
            return O.forward_fmc(m, prob.geom, w, dt, h, nt, None, els)

        def grad(m, els, d):
            return O.misfit_and_gradient(m, prob.geom, w, dt, h, nt, d, None, els)
    # one shot first (also the warm-up) to size the sample to the time budget
    d1 = fwd(prob.m_true, [0])
    t0 = time.perf_counter()
    grad(prob.m0, [0], d1)
    t1 = time.perf_counter() - t0
    shots = int(max(1, min(len(prob.src_list), budget_s // max(t1, 1e-3))))
    els = list(prob.src_list[:shots])
    d = fwd(prob.m_true, els)
    t0 = time.perf_counter()
    grad(prob.m0, els, d)
    t = time.perf_counter() - t0
    upd = 2.0 * n * n * (nt - 1) * shots
    out = {"value": upd / t / 1e9, "unit": "Gpt-updates/s", "cores": cores, "kind": kind,
           "sample": f"512x512 grid, nt=1000, first {shots} of the 128 shots of the benchmark "
                     f"workload, FP64 misfit+gradient, {t:.1f} s"}
    if kind == "reference":
        # the same core on one thread: a child process, since this process's OpenMP
        # runtime already runs with every core
        try:
            r = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-one-thread"],
                               env=dict(os.environ, OMP_NUM_THREADS="1"), capture_output=True,
                               text=True, timeout=300)
            out["one_thread"] = json.loads(r.stdout.strip().splitlines()[-1])
        except Exception as e:  # the main line must still print
            out["one_thread"] = {"error": repr(e)}
    return out


def cpu_one_thread():
    """One shot of the reference core's gradient on the 512^2 / 128-element problem
    on one OpenMP thread (run in a child process)."""
    from openusct_amd import configs
    from openusct_amd.plan import linear_indices
    sys.path.insert(0, os.path.join(ROOT, "oracle", "_ref"))
    import _uap

    prob = configs.c2()
    h, dt, nt, w, n = prob.h, prob.dt, prob.nt, prob.wavelet, prob.shape[0]
    rec = linear_indices(prob.geom.idx, prob.shape).astype(np.int32)
    tx = rec[[prob.src_list[0]]].copy()
    d = _uap.forward_fmc(prob.m_true, h, dt, nt, tx, rec, w)
    from threadpoolctl import threadpool_info, threadpool_limits
    # one OpenMP thread set through the runtime (a host may fix OMP_NUM_THREADS)
    with threadpool_limits(limits=1, user_api="openmp"):
        omp = [i["num_threads"] for i in threadpool_info() if i.get("internal_api") == "openmp"]
        t0 = time.perf_counter()
        _uap.misfit_and_gradient(prob.m0, h, dt, nt, tx, rec, w, d, 0)
        t = time.perf_counter() - t0
    return {"value": 2.0 * n * n * (nt - 1) / t / 1e9, "unit": "Gpt-updates/s", "cores": 1,
            "omp_threads": omp, "sample": f"shot 0 of the benchmark workload, FP64 "
                                          f"misfit+gradient, {t:.2f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="C2", choices=sorted(WORKLOADS))
    ap.add_argument("--shots", type=int, default=0, help="use only the first N shots")
    ap.add_argument("--plain-shots", type=int, default=0,
                    help="shots of the plain per-shot path timed beside the batched one "
                         "(default: 128 in 2D, 2 for C4, whose plain plan keeps 134 GB of history)")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--ckpt-every", type=int, default=0)
    ap.add_argument("--reverse", default="auto", choices=["auto", "on", "off"],
                    help="3D gradients: run the forward backwards alongside the adjoint (on), "
                         "keep levels (off), or let the plan choose (auto)")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-profile", action="store_true",
                    help="skip the per-launch event timing pass (for counter runs)")
    ap.add_argument("--calib-copy-mb", type=int, default=0,
                    help="first run an elementwise copy of this many MiB (counter calibration)")
    ap.add_argument("--serve-requests", type=int, default=24,
                    help="requests of the serving sample reported beside the main line "
                         "(C5: of the serving run itself); 0 = skip")
    ap.add_argument("--gd-steps", type=int, default=-1,
                    help="steps of gradient descent through the sound-speed autograd Function "
                         "after the timed gradients (default: 5 for C3, else 0)")
    ap.add_argument("--serve-nt-scale", type=float, default=0.05,
                    help="time-axis scale of the serving requests")
    ap.add_argument("--no-io", action="store_true",
                    help="skip the acquisition-file leg (C2 only: write and read its traces)")
    ap.add_argument("--gn-products", type=int, default=-1,
                    help="Gauss-Newton Hessian-vector products timed on the same plan "
                         "(default: 3 for C2, else 0)")
    ap.add_argument("--cpu-one-thread", action="store_true", help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.cpu_one_thread:
        print(json.dumps(cpu_one_thread()), flush=True)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(shots would be sharded across GPUs with one all-reduce of g; not built).",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from openusct_amd.plan import Plan, linear_indices

    dev = torch.device("cuda", 0)
    if args.calib_copy_mb:
        # y = x * 1: reads and writes exactly calib_copy_mb MiB each
        x = torch.ones(args.calib_copy_mb << 18, dtype=torch.float32, device=dev)
        for _ in range(3):
            y = x * 1.0
        torch.cuda.synchronize()
        del x, y
    if args.workload == "C5":
        sv = serving(args.serve_requests or 100, args.serve_nt_scale)
        print(json.dumps({
            "metric": "FWI gradient Gpt-updates/s (fwd+adj) batched vs plain; p99 solve latency",
            "value": sv["gpt_updates_per_s"], "unit": "Gpt-updates/s", "n_gpus": 1, "steps": 1,
            "warmup": 0, "ms_per_step": round(sv["makespan_s"] * 1e3, 3),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
            "data": "synthetic (seeded phantoms, observed data simulated on the device)",
            "config": {"workload": WORKLOADS["C5"]}, "serving": sv,
            "roofline": None, "cpu_baseline": None}), flush=True)
        return
    prob = workload(args.workload)
    shape, h, dt, nt = prob.shape, prob.h, prob.dt, prob.nt
    N = int(np.prod(shape))
    rec = linear_indices(prob.geom.idx, shape)
    src = prob.src_list[: args.shots] if args.shots else prob.src_list
    tx = rec[src].copy()
    wd = torch.tensor(prob.wavelet, device=dev)
    mt = torch.tensor(prob.m_true, device=dev)
    md = torch.tensor(prob.m0, device=dev)
    reverse = {"auto": None, "on": True, "off": False}[args.reverse]
    # a 3D plan left to itself chooses between checkpoints and running the
    # forward backwards; an explicit --ckpt-every or --reverse off keeps levels
    ckpt = args.ckpt_every or (0 if reverse is not False and len(shape) == 3 else prob.ckpt_every)

    print(f"[bench] {args.workload}: building batched plan ({len(tx)} shots, {shape}, nt={nt})",
          file=sys.stderr, flush=True)
    plan = Plan(shape, h, dt, nt, tx, rec, method="batched", ckpt_every=ckpt,
                autotune=not args.no_autotune, reverse=reverse if len(shape) == 3 else None)
    print(f"[bench] plan {plan.info}", file=sys.stderr, flush=True)
    info = plan.info
    dobs = plan.forward_fmc(mt, wd)
    torch.cuda.synchronize()
    for _ in range(args.warmup):
        J, g = plan.misfit_and_gradient(md, wd, dobs)
    torch.cuda.synchronize()
    print("[bench] warmup done", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        J, g = plan.misfit_and_gradient(md, wd, dobs)
    torch.cuda.synchronize()
    dt_run = time.perf_counter() - t0
    ms = dt_run / args.steps * 1e3
    upd = 2.0 * N * (nt - 1) * len(tx)
    value = upd / (ms / 1e3) / 1e9
    print(f"[bench] batched {ms:.1f} ms/gradient, {value:.1f} Gpt-updates/s", file=sys.stderr,
          flush=True)

    gd = None
    gd_steps = args.gd_steps if args.gd_steps >= 0 else (5 if args.workload == "C3" else 0)
    if gd_steps:
        from openusct_amd.autograd import gradient_descent
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        c0 = torch.sqrt(1.0 / md)
        _, hist = gradient_descent(c0, plan, wd, dobs, steps=gd_steps, step_frac=0.005,
                                   c_bounds=(1400.0, 1800.0))
        torch.cuda.synchronize()
        gd = {"steps": gd_steps, "misfit_history": hist,
              "seconds": round(time.perf_counter() - t0, 2),
              "through": "openusct_amd.autograd.SoundSpeedMisfit (c -> J, backward = device adjoint)"}
        print(f"[bench] gradient descent {gd_steps} steps in {gd['seconds']} s: {hist}",
              file=sys.stderr, flush=True)

    gn = None
    gn_k = args.gn_products if args.gn_products >= 0 else (3 if args.workload == "C2" else 0)
    if gn_k:
        # Gauss-Newton products (J^T J) v: background forward, Born forward and
        # adjoint sweep per call; v a smooth direction of model size
        v = (g / g.abs().max()) * md.mean()
        plan.gn_hessian_vector(md, wd, v)  # first call allocates the Born state
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(gn_k):
            hv = plan.gn_hessian_vector(md, wd, v)
        torch.cuda.synchronize()
        ms_hv = (time.perf_counter() - t0) / gn_k * 1e3
        upd_hv = 3.0 * N * (nt - 1) * len(tx)
        gn = {"products": gn_k, "ms_per_product": round(ms_hv, 3),
              "vs_gradient": round(ms_hv / ms, 3),
              "gpt_updates_per_s": round(upd_hv / (ms_hv / 1e3) / 1e9, 3),
              "counted": "3 sweeps of N (nt-1) shots: background forward, Born forward, adjoint",
              "finite": bool(torch.isfinite(hv).all().item())}
        print(f"[bench] Gauss-Newton product {ms_hv:.1f} ms ({gn['vs_gradient']}x a gradient), "
              f"{gn['gpt_updates_per_s']} Gpt-updates/s", file=sys.stderr, flush=True)
        del hv, v

    if args.no_profile:
        prof = {"fwd_ms": float("nan"), "adj_ms": float("nan"), "n_fwd": 0, "n_adj": 0}
    else:
        prof = plan.profile_gradient(md, wd, dobs)
    per_pass = len(tx) / max(info["shot_passes"], 1)
    lv = max(int(prof.get("adj_levels", 1)), 1)    # time levels per adjoint launch
    lvf = max(int(prof.get("fwd_levels", 1)), 1)   # and per forward launch
    adj_bytes = BYTES_ADJ * N * per_pass * lv
    fwd_bytes = BYTES_FWD * N * per_pass * lvf
    adj_gbs = adj_bytes / (prof["adj_ms"] / 1e3) / 1e9
    fwd_gbs = fwd_bytes / (prof["fwd_ms"] / 1e3) / 1e9
    print(f"[bench] per launch ({lvf}/{lv} level(s)): fwd {prof['fwd_ms']:.3f} ms ({fwd_gbs:.0f} "
          f"GB/s alg), adj {prof['adj_ms']:.3f} ms ({adj_gbs:.0f} GB/s alg)", file=sys.stderr,
          flush=True)
    suffix = "_x2" if lv == 2 else ""
    rev = bool(info.get("reverse"))
    # minimum bytes of the adjoint kernel per launch: per shot-cell it reads
    # lam^{k+1}, lam^{k+2} and one stored level per time level, writes lam per level;
    # the reversing kernel instead reads forward levels k-1 and k (FP32 hi + bf16 lo)
    # and writes level k-2: 12 + 18 bytes
    kmin_bytes = (30.0 if rev else 4.0 * (2 + lv) + 4.0 * lv) * N * per_pass
    adj_kernel = ("k_adj2d" if len(shape) == 2 else ("k_radj3d" if rev else "k_adj3d")) + suffix
    fwd_kernel = ("k_fwd2d" if len(shape) == 2 else "k_fwd3d") + ("_x2" if lvf == 2 else "")
    if len(shape) == 3 and info.get("fwd_tile_kind") == 2:
        fwd_kernel = "k_fwd3d_x2_lds"
    Jv = float(J.item())
    traces = None
    if args.workload == "C2" and not args.no_io:
        traces = dobs.cpu().numpy()
    plan.close()
    del dobs
    torch.cuda.empty_cache()

    plain = None
    if not args.no_plain and args.workload in ("C2", "C2-256", "C4"):
        ps = min(args.plain_shots or (2 if len(shape) == 3 else 128), len(tx))
        pplan = Plan(shape, h, dt, nt, tx[:ps], rec, method="plain")
        pd = pplan.forward_fmc(mt, wd)
        pplan.misfit_and_gradient(md, wd, pd)  # warm-up
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        pplan.misfit_and_gradient(md, wd, pd)
        torch.cuda.synchronize()
        tp = time.perf_counter() - t0
        plain = 2.0 * N * (nt - 1) * ps / tp / 1e9
        print(f"[bench] plain {tp * 1e3:.1f} ms/gradient over {ps} shots, {plain:.1f} "
              f"Gpt-updates/s", file=sys.stderr, flush=True)
        pplan.close()

    io = None
    if traces is not None:
        try:
            io = dataset_io(traces, prob)
        except Exception as e:  # the main line must still print
            io = {"error": repr(e)}
            print(f"[bench] dataset file leg failed: {e!r}", file=sys.stderr, flush=True)
        if "error" not in io:
            print(f"[bench] dataset file {io['file_bytes'] / 1e6:.1f} MB: save {io['save_GBps']} GB/s, "
                  f"load {io['load_GBps']} GB/s, to HBM {io['to_device_GBps']} GB/s, h5py "
                  f"{io['h5py_load_GBps']} GB/s, exact {io['bit_exact']}", file=sys.stderr, flush=True)
        del traces

    sv = None
    if args.serve_requests:
        sv = serving(args.serve_requests, args.serve_nt_scale)
        print(f"[bench] serving p50 {sv['p50_s']:.3f} s, p99 {sv['p99_s']:.3f} s, "
              f"{sv['gpt_updates_per_s']:.1f} Gpt-updates/s", file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu:
        print("[bench] timing the reference CPU core", file=sys.stderr, flush=True)
        cpu = cpu_baseline()

    traffic = None
    tfile = os.path.join(ROOT, "profiles", {"C2": "pmc_traffic_ring512_128shots.json",
                                            "C4": "pmc_traffic_cylinder256_64shots.json"}.get(
                                                args.workload, "none"))
    if len(tx) == {"C2": 128, "C4": 64}.get(args.workload) and os.path.exists(tfile):
        # HBM bytes per launch of the adjoint kernel from the committed rocprofv3 --pmc
        # FETCH_SIZE / WRITE_SIZE runs (calibrated against a copy of known size)
        with open(tfile) as fh:
            k = json.load(fh)["kernels"].get(adj_kernel)
        if k is not None:
            traffic = round(k["read_bytes"] + k["write_bytes"])

    line = {
        "metric": "FWI gradient Gpt-updates/s (fwd+adj) batched vs plain; p99 solve latency",
        "value": round(value, 3),
        "unit": "Gpt-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded inclusion phantom, observed data simulated on the device)",
        "config": {
            "workload": WORKLOADS[args.workload] if not args.shots else
            WORKLOADS[args.workload] + f" (first {len(tx)} shots)",
            "grid": list(shape), "nt": nt, "shots": len(tx), "receivers": len(rec),
            "method": "batched", "plan": info, "peak_hbm_bytes": torch.cuda.max_memory_allocated() + info["workspace_bytes"],
            "recompute_level_fraction": round(max(prof["n_fwd"] - (nt - 1) * info["shot_passes"], 0) / max(prof["n_fwd"] + prof["n_adj"], 1), 4),
        },
        "plain_value": None if plain is None else round(plain, 3),
        "batched_vs_plain": None if plain is None else round(value / plain, 3),
        "misfit": Jv,
        "roofline": {
            "bound": "hbm",
            "kernel": adj_kernel + f" (adjoint + imaging{', forward level re-created' if rev else ''}"
                                   f", {lv} time level(s) per launch, all shots of a pass)",
            "achieved": round(adj_gbs, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(adj_gbs / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "per_launch_ms": round(prof["adj_ms"], 4),
            "alg_bytes_per_launch": adj_bytes,
            "levels_per_launch": lv,
            # HBM bytes the kernel itself cannot avoid per launch (its operands once):
            # two-level kernels read two levels and write two; against the time above
            "kernel_min_bytes_per_launch": kmin_bytes,
            "kernel_min_frac": round(kmin_bytes / (prof["adj_ms"] / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
            "fwd_kernel": {"name": fwd_kernel, "achieved": round(fwd_gbs, 1),
                           "per_launch_ms": round(prof["fwd_ms"], 4),
                           "alg_bytes_per_launch": fwd_bytes, "levels_per_launch": lvf},
            "gradient_level": round(value * BYTES_PER_UPDATE, 1),
        },
        "cpu_baseline": cpu,
        "serving": sv,
        "gradient_descent": gd,
        "gauss_newton": gn,
        "dataset_io": io,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

"""Benchmark: FWI gradient throughput (Gpt-updates/s, forward + adjoint).

Headline workload (one MI355X): the 256-shot 2D ring full-matrix capture at
512 x 512 -- dx = 0.2 mm, dt = 50 ns, nt = 1000, a 256-element ring of radius
22 mm, every element transmits, Ricker 500 kHz, true model 1500 m/s water with
three seeded inclusions (c in [1400, 1800] m/s), starting model 1500 m/s.
Observed data is synthesised on the device.  A step is one misfit-and-gradient
over all 256 shots: 2 * N * (nt - 1) * 256 = 1.34e11 grid-point updates.

The JSON line carries, besides the headline value:
  plain_value      the plain per-shot path (one launch per shot per time level)
                   on the same inputs;
  roofline         the dominant kernel of the gradient against the HBM peak,
                   on the bytes that kernel cannot avoid per launch (its
                   operands once), timed per launch with HIP events on the
                   plan's stream; HBM traffic per launch from committed
                   rocprofv3 --pmc runs; the survey's per-update byte basis
                   beside it under its own name;
  c4               the 3D cylinder at 256^3 with 64 shots (nt = 2000): the
                   same figures for one gradient;
  stream_ceiling   read, write and copy rates of streaming kernels
                   (libusct_bw.so) on the same device;
  cpu_baseline     the reference's own C++ core (FP64, OpenMP) on the 256^2,
                   8-shot problem, median of 3 after one warm-up, plus one thread;
  serving, gauss_newton, dataset_io   the other legs, on samples.
"""

from __future__ import annotations

import argparse
import ctypes
import json
import os
import statistics
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BYTES_FWD = 16.0   # the survey's algorithmic bytes of one forward update (FP32)
BYTES_ADJ = 28.0   # adjoint + imaging update
BYTES_PER_UPDATE = 22.0
METRIC = "FWI gradient Gpt-updates/s (fwd+adj) batched vs plain; p99 solve latency"

WORKLOADS = {
    "C1-16": "2D ring FMC 256x256, dx=0.2mm, dt=50ns, nt=1000, 64-element ring R=22mm, 16 shots "
             "(every 4th element), Ricker 500kHz, FP32 fields; one step = one misfit+gradient "
             "(a small request: launch overhead at its largest)",
    "C2-256": "2D ring FMC 512x512, dx=0.2mm, dt=50ns, nt=1000, 256-element ring R=22mm, "
              "256 shots, Ricker 500kHz, FP32 fields; one step = one misfit+gradient over all "
              "256 shots",
    "C2": "2D ring FMC 512x512, dx=0.2mm, dt=50ns, nt=1000, 128-element ring R=22mm, "
          "128 shots, Ricker 500kHz, FP32 fields; one step = one misfit+gradient over all shots",
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

# rocprofv3 --pmc summaries (HBM bytes per launch) committed under profiles/
TRAFFIC_FILES = {"C2-256": "pmc_traffic_ring512_256shots.json",
                 "C2": "pmc_traffic_ring512_128shots.json",
                 "C4": "pmc_traffic_cylinder256_64shots.json"}
FULL_SHOTS = {"C1-16": 16, "C2-256": 256, "C2": 128, "C3": 256, "C4": 64}


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
    t_serve = time.perf_counter()

    def progress(k, n):
        print(f"[bench] serving: {k}/{n} done at {time.perf_counter() - t_serve:.1f} s",
              file=sys.stderr, flush=True)

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


def workload(name):
    from openusct_amd import configs
    return configs.BUILDERS[name]()


# ---------------------------------------------------------------------------
# streaming ceiling (libusct_bw.so)
# ---------------------------------------------------------------------------
def _bw_lib():
    lib = ctypes.CDLL(os.path.join(ROOT, "openusct_amd", "libusct_bw.so"))
    lib.usct_bw_run.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                ctypes.c_int64, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                ctypes.POINTER(ctypes.c_float), ctypes.c_void_p]
    lib.usct_bw_run_pass.argtypes = lib.usct_bw_run.argtypes
    lib.usct_bw_pass_groups.argtypes = [ctypes.c_int64, ctypes.c_int]
    lib.usct_bw_pass_groups.restype = ctypes.c_int64
    lib.usct_bw_device_cus.restype = ctypes.c_int
    return lib


def stream_ceiling(gib=4, reps=5):
    """Best read / write / copy rates of the streaming kernels over `gib` GiB
    buffers (far beyond L2 and Infinity Cache): grid-stride loops over {2, 4, 8,
    16} work-groups per CU and single-pass grids with {1, 2, 4, 8} accesses per
    lane, each with ordinary or non-temporal accesses.  Copy counts read + write
    bytes.  Every configuration is kept in `all` (GB/s)."""
    import torch
    lib = _bw_lib()
    nbytes = gib << 30
    a = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    b = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda")
    cus = lib.usct_bw_device_cus()
    groups = max(cus * 64, lib.usct_bw_pass_groups(nbytes, 1))
    part = torch.empty(groups, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    out = {"buffer_bytes": nbytes, "cus": cus, "reps": reps}
    for kind, name, mult in ((0, "read", 1), (1, "write", 1), (2, "copy", 2)):
        best, every = None, {}
        for layout, params in (("grid-stride", (2, 4, 8, 16)), ("single-pass", (1, 2, 4, 8))):
            fn = lib.usct_bw_run if layout == "grid-stride" else lib.usct_bw_run_pass
            for prm in params:
                for nt in (0, 1):
                    ms = ctypes.c_float()
                    rc = fn(kind, a.data_ptr(), b.data_ptr(), part.data_ptr(), nbytes, prm, nt, reps,
                            ctypes.byref(ms), None)
                    if rc:
                        raise RuntimeError(f"usct_bw ({layout}, {name}) failed: {rc}")
                    gbs = mult * nbytes / (ms.value / 1e3) / 1e9
                    key = (f"{layout} {'blocks_per_cu' if layout == 'grid-stride' else 'per_lane'} "
                           f"{prm}{' nt' if nt else ''}")
                    every[key] = round(gbs, 1)
                    if best is None or gbs > best[0]:
                        best = (gbs, layout, prm, nt, ms.value)
        out[name] = {"GBps": round(best[0], 1), "layout": best[1],
                     ("blocks_per_cu" if best[1] == "grid-stride" else "per_lane"): best[2],
                     "nontemporal": bool(best[3]), "ms": round(best[4], 4), "all": every}
    del a, b, part
    torch.cuda.empty_cache()
    return out


def calib_copy(mib):
    """Counter calibration: three launches of the streaming copy kernel over `mib`
    MiB (each reads and writes exactly mib MiB)."""
    import torch
    lib = _bw_lib()
    nbytes = mib << 20
    a = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    ms = ctypes.c_float()
    rc = lib.usct_bw_run(2, a.data_ptr(), b.data_ptr(), None, nbytes, 8, 0, 2, ctypes.byref(ms),
                         None)
    torch.cuda.synchronize()
    if rc:
        raise RuntimeError("calibration copy failed")
    del a, b


# ---------------------------------------------------------------------------
# kernel byte model
# ---------------------------------------------------------------------------
def kernel_min_bytes(kernel, N, shots, groups):
    """(read, write) bytes a launch of `kernel` cannot avoid: every operand once.
    Per shot-cell: fields (FP32 4 B; forward state FP32 hi + bf16 lo = 6 B);
    per cell: the S' pair (8 B) and, for adjoints, the gradient accumulators of
    each lane group (hi + lo, read and written: 8 + 8 B per group)."""
    per_sc = {
        # read p^{n-1}, p^{n-2}; write p^n, p^{n+1}
        "k_fwd2d_x2": (12.0, 12.0),
        # read p^{n-1}, p^{n-2}; write p^n
        "k_fwd2d": (12.0, 6.0), "k_fwd3d": (12.0, 6.0),
        # read lam^{k+1}, lam^{k+2}, p^{k-1}, p^{k-2}; write lam^k, lam^{k-1}
        "k_adj2d_x2": (16.0, 8.0),
        # read lam^{k+1}, lam^{k+2}, one stored level; write lam^k
        "k_adj2d": (12.0, 4.0), "k_adj3d": (12.0, 4.0),
        # read lam^{k+1}, lam^{k+2}, forward levels k-1 and k (hi + lo); write lam^k, level k-2
        "k_radj3d": (20.0, 10.0),
    }[kernel]
    adj = "adj" in kernel
    rd = per_sc[0] * N * shots + 8.0 * N + (8.0 * N * groups if adj else 0.0)
    wr = per_sc[1] * N * shots + (8.0 * N * groups if adj else 0.0)
    return rd, wr


def load_traffic(name, kernel, shots, tile_kind=0):
    """Counted (read, write) bytes per launch of `kernel` from the workload's traffic
    file: the entry for the plan's tile kind where it is not 0 (none: no figure)."""
    f = TRAFFIC_FILES.get(name)
    path = os.path.join(ROOT, "profiles", f) if f else None
    if not path or shots != FULL_SHOTS.get(name) or not os.path.exists(path):
        return None
    with open(path) as fh:
        d = json.load(fh)
    k = (d.get("kernels_tile_kind", {}).get(str(tile_kind), {}) if tile_kind else d["kernels"]).get(kernel)
    return None if k is None else (k["read_bytes"], k["write_bytes"])


def touch_device_memory(gib=140):
    """Allocate, fill and free `gib` GiB of device memory (at most half of what is
    free) once, before the first plan.  Measured on MI355X boxes fresh from boot: the
    first plan a process creates there ran its 512^2 x 128-shot gradient at 178-179 ms
    for as long as it lived (40 s of back-to-back gradients), whatever its tiles, while
    a second plan with the same tiles in the same process, or a first plan created after
    such a fill, ran at 167-170 ms.  A long-running process has long since touched its
    device memory; this makes the benchmark's plans see that state.  Untimed."""
    import torch
    free, _ = torch.cuda.mem_get_info()
    n = min(gib << 30, free // 2) & ~((2 << 20) - 1)
    t0 = time.perf_counter()
    x = torch.empty(n, dtype=torch.uint8, device="cuda")
    x.fill_(0)
    torch.cuda.synchronize()
    del x
    torch.cuda.empty_cache()
    print(f"[bench] device memory touched: {n / 2**30:.0f} GiB in {time.perf_counter() - t0:.2f} s",
          flush=True)
    return n


def reduce_pmc(fetch_csv, write_csv, out_json, workload_text, calib_bytes, tile_kind=0,
               only=None):
    """Per-kernel HBM traffic from two rocprofv3 --pmc runs (FETCH_SIZE, then
    WRITE_SIZE, KiB per dispatch) of the same command started with
    --calib-copy-mb: each counter is divided by what it reads for the streaming
    copy kernel (k_bw_copy) against the bytes that kernel moves, and each kernel
    keeps the median over its dispatches of the most frequent grid size (the
    tuned tiling; the autotuner's timing launches run on other data).  With
    tile_kind != 0 (runs whose plan was restricted to that tile kind) the kernels
    named in `only` go into out_json's "kernels_tile_kind" entry for that kind, and
    the rest of out_json is kept."""
    import collections
    import csv

    def per_kernel(path, counter):
        d = collections.defaultdict(list)
        with open(path) as fh:
            for r in csv.DictReader(fh):
                if r["Counter_Name"] != counter:
                    continue
                name = r["Kernel_Name"].replace("(anonymous namespace)::", "")
                name = name.split("(")[0].split("<")[0].replace("void ", "").strip()
                name = name.split("::")[-1]
                d[name].append((int(r["Grid_Size"]), float(r["Counter_Value"]) * 1024.0))
        return d

    fetch, write = per_kernel(fetch_csv, "FETCH_SIZE"), per_kernel(write_csv, "WRITE_SIZE")
    calib = fetch.get("k_bw_copy"), write.get("k_bw_copy")
    if not calib[0] or not calib[1]:
        raise SystemExit("no k_bw_copy dispatches: start the profiled command with --calib-copy-mb")
    moved_bytes = float(calib_bytes)  # the copy reads and writes exactly this much
    f_scale = float(np.median([v for _, v in calib[0]])) / moved_bytes
    w_scale = float(np.median([v for _, v in calib[1]])) / moved_bytes
    out = {"workload": workload_text,
           "calibration": f"rocprofv3 --pmc on the streaming copy kernel (k_bw_copy, libusct_bw.so) of "
                          f"{moved_bytes / 2**20:.0f} MiB: FETCH_SIZE reads "
                          f"{f_scale * moved_bytes / 1024:.1f} KiB and WRITE_SIZE "
                          f"{w_scale * moved_bytes / 1024:.1f} KiB per dispatch against "
                          f"{moved_bytes / 1024:.0f} KiB moved each way",
           "fetch_scale": round(f_scale, 6), "write_scale": round(w_scale, 6), "kernels": {},
           "note": "bytes per dispatch: the median over the dispatches with the most frequent grid size "
                   "of each kernel (the tuned tiling; the median leaves out the autotuner's timing "
                   "launches, which run on other data), each counter divided by its calibration scale; "
                   "FETCH_SIZE counts the L2's requests to the fabric, so lines served by the Infinity "
                   "Cache count as fetched"}
    # The two runs tune their tiles separately and may pick different ones: each
    # counter takes its own run's most frequent grid for the kernel.
    for name in sorted(set(fetch) & set(write)):
        if name == "k_bw_copy" or not name.startswith("k_"):  # runtime copies and fills
            continue
        gf = collections.Counter(g for g, _ in fetch[name]).most_common(1)[0][0]
        gw = collections.Counter(g for g, _ in write[name]).most_common(1)[0][0]
        fv = [v for g, v in fetch[name] if g == gf]
        wv = [v for g, v in write[name] if g == gw]
        out["kernels"][name] = {"grid_size": gf, "dispatches": len(fetch[name]),
                                "read_bytes": int(round(float(np.median(fv)) / f_scale)),
                                "write_bytes": int(round(float(np.median(wv)) / w_scale))}
        if gw != gf:
            out["kernels"][name]["grid_size_write_run"] = gw
    if tile_kind:
        with open(out_json) as fh:
            base = json.load(fh)
        base.setdefault("kernels_tile_kind", {})[str(tile_kind)] = {
            k: v for k, v in out["kernels"].items() if only is None or k in only}
        base["note_tile_kind"] = ("kernels_tile_kind: the same counters for plans whose tuner was "
                                  "restricted to that tile kind (USCT_TUNE_KIND), same calibration")
        out = base
    with open(out_json, "w") as fh:
        json.dump(out, fh, indent=1)
        fh.write("\n")
    return out


# ---------------------------------------------------------------------------
# one gradient workload
# ---------------------------------------------------------------------------
def gradient_leg(name, steps, warmup, args, keep=False):
    """Batched plan of workload `name`: observed data, `warmup` untimed and `steps`
    timed gradients, then an eager pass with per-launch HIP events.  Returns the
    figures, and the plan, model tensors and data when ``keep``."""
    import torch
    from openusct_amd.plan import Plan, linear_indices

    dev = torch.device("cuda", 0)
    prob = workload(name)
    shape, h, dt, nt = prob.shape, prob.h, prob.dt, prob.nt
    N = int(np.prod(shape))
    rec = linear_indices(prob.geom.idx, shape)
    src = prob.src_list[: args.shots] if args.shots and name == args.workload else prob.src_list
    tx = rec[src].copy()
    wd = torch.tensor(prob.wavelet, device=dev)
    mt = torch.tensor(prob.m_true, device=dev)
    md = torch.tensor(prob.m0, device=dev)
    reverse = {"auto": None, "on": True, "off": False}[args.reverse]
    ckpt = args.ckpt_every or (0 if reverse is not False and len(shape) == 3 else prob.ckpt_every)
    torch.cuda.reset_peak_memory_stats()
    print(f"[bench] {name}: building batched plan ({len(tx)} shots, {shape}, nt={nt})",
          file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    plan = Plan(shape, h, dt, nt, tx, rec, method="batched", ckpt_every=ckpt,
                autotune=not args.no_autotune, reverse=reverse if len(shape) == 3 else None)
    t_plan = time.perf_counter() - t0
    info = plan.info
    print(f"[bench] {name}: plan in {t_plan:.1f} s: {info}", file=sys.stderr, flush=True)
    dobs = plan.forward_fmc(mt, wd)
    torch.cuda.synchronize()
    for _ in range(warmup):
        J, g = plan.misfit_and_gradient(md, wd, dobs)
    torch.cuda.synchronize()
    print(f"[bench] {name}: warmup done", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    for _ in range(steps):
        J, g = plan.misfit_and_gradient(md, wd, dobs)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) / steps * 1e3
    upd = 2.0 * N * (nt - 1) * len(tx)
    value = upd / (ms / 1e3) / 1e9
    print(f"[bench] {name}: batched {ms:.1f} ms/gradient, {value:.1f} Gpt-updates/s",
          file=sys.stderr, flush=True)
    peak = torch.cuda.max_memory_allocated() + info["workspace_bytes"]
    out = {"name": name, "workload": WORKLOADS[name] if len(tx) == len(prob.src_list) else
           WORKLOADS[name] + f" (first {len(tx)} shots)", "grid": list(shape), "nt": nt,
           "shots": len(tx), "receivers": len(rec), "plan": info, "plan_seconds": round(t_plan, 2),
           "ms_per_gradient": round(ms, 3), "value": value, "misfit": float(J.item()),
           "peak_hbm_bytes": int(peak), "updates": upd, "steps": steps, "warmup": warmup}
    if not args.no_profile:
        prof = plan.profile_gradient(md, wd, dobs)
        out["roofline"] = roofline(name, prof, info, N, len(tx), shape, value)
        out["recompute_level_fraction"] = round(
            max(prof["n_fwd"] - (nt - 1) * info["shot_passes"], 0) /
            max(prof["n_fwd"] + prof["n_adj"], 1), 4)
        r = out["roofline"]
        print(f"[bench] {name}: {r['kernel']} {r['per_launch_ms']:.3f} ms per launch, "
              f"{r['achieved']:.0f} GB/s on its minimum bytes (frac {r['frac']}); "
              f"{r['other_kernel']['name']} {r['other_kernel']['per_launch_ms']:.3f} ms",
              file=sys.stderr, flush=True)
    if keep:
        return out, (prob, plan, mt, md, wd, dobs, g, rec, tx)
    plan.close()
    del dobs, g
    torch.cuda.empty_cache()
    return out, None


def roofline(name, prof, info, N, shots, shape, value):
    """Roofline of the gradient's dominant kernel (the one with the larger share of
    the launch time), on the bytes it cannot avoid per launch."""
    rev = bool(info.get("reverse"))
    lva = max(int(prof.get("adj_levels", 1)), 1)
    lvf = max(int(prof.get("fwd_levels", 1)), 1)
    d2 = len(shape) == 2
    adj_k = ("k_adj2d" if d2 else ("k_radj3d" if rev else "k_adj3d")) + ("_x2" if lva == 2 else "")
    fwd_k = ("k_fwd2d" if d2 else "k_fwd3d") + ("_x2" if lvf == 2 else "")
    passes = max(info["shot_passes"], 1)
    per_pass = shots / passes
    groups = max(info["shots_per_pass"] // max(info["lanes_per_group"], 1), 1)
    t_adj = prof["adj_ms"] * prof["n_adj"] / lva
    t_fwd = prof["fwd_ms"] * prof["n_fwd"] / lvf

    def one(kernel, ms, lv, survey_bytes, tile_kind):
        rd, wr = kernel_min_bytes(kernel, N, per_pass, groups)
        ach = (rd + wr) / (ms / 1e3) / 1e9
        tr = load_traffic(name, kernel, shots, tile_kind)
        d = {"name": kernel, "per_launch_ms": round(ms, 4), "levels_per_launch": lv,
             "min_read_bytes": rd, "min_write_bytes": wr, "achieved": round(ach, 1),
             "frac": round(ach / HBM_PEAK_GBS, 4),
             "traffic": None if tr is None else round(tr[0] + tr[1]),
             "traffic_GBps": None if tr is None else round((tr[0] + tr[1]) / (ms / 1e3) / 1e9, 1),
             # the survey's per-update basis (28 B per adjoint update, 16 B per forward
             # update) counts every level's operands once per level; a two-level launch
             # reuses them from registers, so this figure may exceed the peak
             "survey_basis": {"bytes": survey_bytes * N * per_pass * lv,
                              "GBps": round(survey_bytes * N * per_pass * lv / (ms / 1e3) / 1e9, 1)}}
        return d

    a = one(adj_k, prof["adj_ms"], lva, BYTES_ADJ, int(info.get("tile_kind", 0)))
    f = one(fwd_k, prof["fwd_ms"], lvf, BYTES_FWD, int(info.get("fwd_tile_kind", 0)))
    dom, other = (a, f) if t_adj >= t_fwd else (f, a)
    what = {"k_adj2d_x2": "adjoint + imaging, two time levels per launch",
            "k_adj2d": "adjoint + imaging, one time level per launch",
            "k_adj3d": "3D adjoint + imaging from stored levels, one level per launch",
            "k_radj3d": "3D adjoint + imaging with the forward level re-created, one level per launch",
            "k_fwd2d_x2": "forward, two time levels per launch",
            "k_fwd2d": "forward, one level per launch", "k_fwd3d": "3D forward, one level per launch"}
    return {"bound": "hbm", "kernel": dom["name"], "what": what.get(dom["name"], ""),
            "achieved": dom["achieved"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": dom["frac"], "traffic": dom["traffic"], "traffic_GBps": dom["traffic_GBps"],
            "per_launch_ms": dom["per_launch_ms"], "levels_per_launch": dom["levels_per_launch"],
            "shots_per_launch": per_pass, "min_read_bytes": dom["min_read_bytes"],
            "min_write_bytes": dom["min_write_bytes"], "survey_basis": dom["survey_basis"],
            "time_share": round(max(t_adj, t_fwd) / max(t_adj + t_fwd, 1e-30), 3),
            "other_kernel": other,
            "gradient_survey_basis": {"bytes_per_update": BYTES_PER_UPDATE,
                                      "GBps": round(value * BYTES_PER_UPDATE, 1),
                                      "frac": round(value * BYTES_PER_UPDATE / HBM_PEAK_GBS, 4)}}


def plain_leg(prob, tx, rec, shots):
    """The plain per-shot path (one launch per shot per level, full history) on the
    first `shots` shots: one warm-up and one timed gradient."""
    import torch
    from openusct_amd.plan import Plan
    dev = torch.device("cuda", 0)
    shape, h, dt, nt = prob.shape, prob.h, prob.dt, prob.nt
    N = int(np.prod(shape))
    wd = torch.tensor(prob.wavelet, device=dev)
    mt = torch.tensor(prob.m_true, device=dev)
    md = torch.tensor(prob.m0, device=dev)
    ps = min(shots, len(tx))
    with Plan(shape, h, dt, nt, tx[:ps], rec, method="plain") as pplan:
        pd = pplan.forward_fmc(mt, wd)
        pplan.misfit_and_gradient(md, wd, pd)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        pplan.misfit_and_gradient(md, wd, pd)
        torch.cuda.synchronize()
        tp = time.perf_counter() - t0
    del pd
    torch.cuda.empty_cache()
    v = 2.0 * N * (nt - 1) * ps / tp / 1e9
    print(f"[bench] plain {tp * 1e3:.1f} ms/gradient over {ps} shots, {v:.1f} Gpt-updates/s",
          file=sys.stderr, flush=True)
    return v, ps


# ---------------------------------------------------------------------------
# CPU baseline: the reference's own core (oracle/_ref), in child processes
# ---------------------------------------------------------------------------
def cpu_child(which, reps, threads):
    """Time the reference's C++ core (FP64) in this process, which has not touched
    the device and whose OpenMP runtime starts with the environment it was given.
    which = "c1": the 256^2, 8-shot problem, one warm-up then `reps` timed
    gradients; "c2": the first shots of the 512^2, 128-element problem."""
    from openusct_amd import configs
    from openusct_amd.plan import linear_indices
    sys.path.insert(0, os.path.join(ROOT, "oracle", "_ref"))
    import _uap  # the reference's pybind11 module, compiled from its own sources

    prob = configs.c1() if which == "c1" else configs.c2()
    h, dt, nt, w = prob.h, prob.dt, prob.nt, prob.wavelet
    rec = linear_indices(prob.geom.idx, prob.shape).astype(np.int32)
    els = list(prob.src_list) if which == "c1" else list(prob.src_list[:reps])
    tx = rec[els].copy()
    d = _uap.forward_fmc(prob.m_true, h, dt, nt, tx, rec, w)
    times = []
    n_runs = reps + 1 if which == "c1" else 1
    from threadpoolctl import threadpool_info, threadpool_limits
    # the thread count is also set through the OpenMP runtime: a host may fix
    # OMP_NUM_THREADS for every process it starts
    with threadpool_limits(limits=threads, user_api="openmp"):
        omp = [i["num_threads"] for i in threadpool_info() if i.get("internal_api") == "openmp"]
        for _ in range(n_runs):
            t0 = time.perf_counter()
            _uap.misfit_and_gradient(prob.m0, h, dt, nt, tx, rec, w, d, 0)
            times.append(time.perf_counter() - t0)
    if which == "c1":
        times = times[1:]  # the first run is the warm-up
    upd = 2.0 * int(np.prod(prob.shape)) * (nt - 1) * len(els)
    return {"times_s": [round(t, 4) for t in times], "median_s": statistics.median(times),
            "updates": upd, "shots": len(els), "omp_threads": omp,
            "env": {k: os.environ.get(k) for k in ("OMP_NUM_THREADS", "OMP_PROC_BIND")}}


def _run_child(which, reps, threads):
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close",
               HIP_VISIBLE_DEVICES="")
    r = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-child", which,
                        "--cpu-reps", str(reps), "--cpu-threads", str(threads)], env=env,
                       capture_output=True, text=True, timeout=600)
    if r.returncode != 0:
        raise RuntimeError(r.stderr[-2000:])
    return json.loads(r.stdout.strip().splitlines()[-1])


def cpu_baseline():
    """The reference's own C++ core (oracle/_ref, built from its sources) on the
    256^2 / 64-element / 8-shot problem (nt = 1000, 1.048e9 updates): OpenMP on every
    core of this process's affinity mask with OMP_PROC_BIND=close, one warm-up and the
    median of 3 gradients; the same on one thread; and, as a second sample, the first
    shots of the 512^2 / 128-element problem."""
    if not os.path.isdir(os.path.join(ROOT, "oracle", "_ref")):
        return {"error": "oracle/_ref (the reference's core) is not built"}
    cores = len(os.sched_getaffinity(0))
    c1 = _run_child("c1", 3, cores)
    v = c1["updates"] / c1["median_s"] / 1e9
    print(f"[bench] CPU reference core, {cores} threads: {c1['median_s']:.3f} s per C1 gradient "
          f"({v:.3f} Gpt-updates/s)", file=sys.stderr, flush=True)
    out = {"value": round(v, 4), "unit": "Gpt-updates/s", "cores": cores, "kind": "reference",
           "sample": f"256x256 grid, 64-element ring, 8 shots, nt=1000 (1.048e9 updates), FP64 "
                     f"misfit+gradient by the reference's C++ core, OMP_PROC_BIND=close, one "
                     f"warm-up then the median of 3: {c1['times_s']} s",
           "c1": c1}
    try:
        one = _run_child("c1", 1, 1)
        out["one_thread"] = {"value": round(one["updates"] / one["median_s"] / 1e9, 4),
                             "cores": 1, "seconds": one["times_s"], "omp_threads": one["omp_threads"]}
    except Exception as e:  # the main line must still print
        out["one_thread"] = {"error": repr(e)}
    try:
        c2 = _run_child("c2", 4, cores)
        out["c2_sample"] = {"value": round(c2["updates"] / c2["median_s"] / 1e9, 4),
                            "sample": f"first {c2['shots']} shots of the 512^2, 128-element problem",
                            "seconds": c2["times_s"]}
    except Exception as e:
        out["c2_sample"] = {"error": repr(e)}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="C2-256", choices=sorted(WORKLOADS))
    ap.add_argument("--shots", type=int, default=0, help="use only the first N shots")
    ap.add_argument("--plain-shots", type=int, default=0,
                    help="shots of the plain per-shot path timed beside the batched one "
                         "(default: every shot in 2D, 2 in 3D, whose plain plan keeps a 134 GB history)")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-c4", action="store_true", help="skip the 3D (C4) leg of the 2D headline")
    ap.add_argument("--c4-steps", type=int, default=1)
    ap.add_argument("--no-ceiling", action="store_true", help="skip the streaming-kernel ceiling")
    ap.add_argument("--no-touch", action="store_true",
                    help="skip the one-time fill of device memory before the first plan")
    ap.add_argument("--only-ceiling", action="store_true",
                    help="measure only the streaming-kernel ceiling and print it as JSON")
    ap.add_argument("--ckpt-every", type=int, default=0)
    ap.add_argument("--reverse", default="auto", choices=["auto", "on", "off"],
                    help="3D gradients: run the forward backwards alongside the adjoint (on), "
                         "keep levels (off), or let the plan choose (auto)")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-profile", action="store_true",
                    help="skip the per-launch event timing pass (for counter runs)")
    ap.add_argument("--calib-copy-mb", type=int, default=0,
                    help="first run the streaming copy kernel over this many MiB (counter calibration)")
    ap.add_argument("--serve-requests", type=int, default=24,
                    help="requests of the serving sample reported beside the main line "
                         "(C5: of the serving run itself); 0 = skip")
    ap.add_argument("--gd-steps", type=int, default=-1,
                    help="steps of gradient descent through the sound-speed autograd Function "
                         "after the timed gradients (default: 5 for C3, else 0)")
    ap.add_argument("--serve-nt-scale", type=float, default=0.05,
                    help="time-axis scale of the serving requests")
    ap.add_argument("--no-io", action="store_true",
                    help="skip the acquisition-file leg (write and read the workload's traces)")
    ap.add_argument("--gn-products", type=int, default=-1,
                    help="Gauss-Newton Hessian-vector products timed on the same plan "
                         "(default: 3 for the 512^2 workloads, else 0)")
    ap.add_argument("--reduce-pmc", nargs=3, metavar=("FETCH_CSV", "WRITE_CSV", "OUT_JSON"),
                    help="no GPU: reduce two rocprofv3 --pmc counter files (FETCH_SIZE, WRITE_SIZE) "
                         "of a run started with --calib-copy-mb to per-kernel HBM bytes per launch")
    ap.add_argument("--pmc-workload", default="", help="workload text stored with --reduce-pmc")
    ap.add_argument("--pmc-tile-kind", type=int, default=0,
                    help="--reduce-pmc of runs restricted to this tile kind: add the kernels named "
                         "by --pmc-kernels to OUT_JSON's entry for the kind")
    ap.add_argument("--pmc-kernels", default="", help="comma-separated kernel names")
    ap.add_argument("--cpu-child", default="", help=argparse.SUPPRESS)
    ap.add_argument("--cpu-reps", type=int, default=3, help=argparse.SUPPRESS)
    ap.add_argument("--cpu-threads", type=int, default=1, help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.cpu_child:
        print(json.dumps(cpu_child(args.cpu_child, args.cpu_reps, args.cpu_threads)), flush=True)
        return
    if args.reduce_pmc:
        if not args.calib_copy_mb:
            raise SystemExit("--reduce-pmc needs --calib-copy-mb (the calibration copy's MiB)")
        out = reduce_pmc(*args.reduce_pmc, args.pmc_workload, args.calib_copy_mb << 20,
                         args.pmc_tile_kind, args.pmc_kernels.split(",") if args.pmc_kernels else None)
        print(json.dumps({k: v for k, v in out["kernels"].items()}), flush=True)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(shots would be sharded across GPUs with one all-reduce of g; not built).",
              file=sys.stderr)
        sys.exit(2)

    import torch

    if args.only_ceiling:
        print(json.dumps(stream_ceiling()), flush=True)
        return
    if args.calib_copy_mb:
        calib_copy(args.calib_copy_mb)
    if args.workload == "C5":
        sv = serving(args.serve_requests or 100, args.serve_nt_scale)
        print(json.dumps({
            "metric": METRIC, "value": sv["gpt_updates_per_s"], "unit": "Gpt-updates/s",
            "n_gpus": 1, "steps": 1, "warmup": 0, "ms_per_step": round(sv["makespan_s"] * 1e3, 3),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
            "data": "synthetic (seeded phantoms, observed data simulated on the device)",
            "config": {"workload": WORKLOADS["C5"]}, "serving": sv,
            "roofline": None, "cpu_baseline": None}), flush=True)
        return

    touched = 0 if args.no_touch else touch_device_memory()
    head, kept = gradient_leg(args.workload, args.steps, args.warmup, args, keep=True)
    prob, plan, mt, md, wd, dobs, g, rec, tx = kept

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
    gn_k = args.gn_products if args.gn_products >= 0 else \
        (3 if args.workload in ("C2", "C2-256") else 0)
    if gn_k:
        # Gauss-Newton products (J^T J) v: background forward, Born forward and
        # adjoint sweep per call; v a smooth direction of model size
        N = int(np.prod(prob.shape))
        v = (g / g.abs().max()) * md.mean()
        plan.gn_hessian_vector(md, wd, v)  # first call allocates the Born state
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(gn_k):
            hv = plan.gn_hessian_vector(md, wd, v)
        torch.cuda.synchronize()
        ms_hv = (time.perf_counter() - t0) / gn_k * 1e3
        upd_hv = 3.0 * N * (prob.nt - 1) * len(tx)
        gn = {"products": gn_k, "ms_per_product": round(ms_hv, 3),
              "vs_gradient": round(ms_hv / head["ms_per_gradient"], 3),
              "gpt_updates_per_s": round(upd_hv / (ms_hv / 1e3) / 1e9, 3),
              "counted": "3 sweeps of N (nt-1) shots: background forward, Born forward, adjoint",
              "finite": bool(torch.isfinite(hv).all().item())}
        print(f"[bench] Gauss-Newton product {ms_hv:.1f} ms ({gn['vs_gradient']}x a gradient), "
              f"{gn['gpt_updates_per_s']} Gpt-updates/s", file=sys.stderr, flush=True)
        del hv, v
    traces = None
    if args.workload in ("C2", "C2-256") and not args.no_io:
        traces = dobs.cpu().numpy()
    plan.close()
    del dobs, g, kept
    torch.cuda.empty_cache()

    plain, plain_shots = None, 0
    if not args.no_plain and args.workload in ("C2", "C2-256", "C4"):
        plain, plain_shots = plain_leg(prob, tx, rec, args.plain_shots or
                                       (2 if len(prob.shape) == 3 else len(tx)))

    c4 = None
    if not args.no_c4 and args.workload in ("C2", "C2-256"):
        c4, _ = gradient_leg("C4", args.c4_steps, 1, args)
        c4["value"] = round(c4["value"], 3)
        if not args.no_plain:
            cp = workload("C4")
            from openusct_amd.plan import linear_indices
            rec4 = linear_indices(cp.geom.idx, cp.shape)
            pv, ps = plain_leg(cp, rec4[cp.src_list].copy(), rec4, 2)
            c4["plain_value"] = round(pv, 3)
            c4["plain_shots"] = ps
            c4["batched_vs_plain"] = round(c4["value"] / pv, 3)

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

    ceiling = None
    if not args.no_ceiling:
        ceiling = stream_ceiling()
        print(f"[bench] streaming ceiling: read {ceiling['read']['GBps']} GB/s, write "
              f"{ceiling['write']['GBps']} GB/s, copy {ceiling['copy']['GBps']} GB/s",
              file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu:
        print("[bench] timing the reference CPU core", file=sys.stderr, flush=True)
        try:
            cpu = cpu_baseline()
        except Exception as e:  # the main line must still print
            cpu = {"error": repr(e)}

    value = head["value"]
    line = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Gpt-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_gradient"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded inclusion phantom, observed data simulated on the device)",
        "config": {
            "workload": head["workload"], "grid": head["grid"], "nt": head["nt"],
            "shots": head["shots"], "receivers": head["receivers"], "method": "batched",
            "plan": head["plan"], "peak_hbm_bytes": head["peak_hbm_bytes"],
            "recompute_level_fraction": head.get("recompute_level_fraction"),
        },
        "plain_value": None if plain is None else round(plain, 3),
        "plain_shots": plain_shots or None,
        "batched_vs_plain": None if plain is None else round(value / plain, 3),
        "misfit": head["misfit"],
        "roofline": head.get("roofline"),
        "c4": c4,
        "stream_ceiling": ceiling,
        "device_memory_touched_bytes": touched,
        "cpu_baseline": cpu,
        "serving": sv,
        "gradient_descent": gd,
        "gauss_newton": gn,
        "dataset_io": io,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

"""Continuous-batching request loop for FWI-gradient and forward-acquisition requests.

Requests of any grid size (2D or 3D) and shot count arrive over time.  The
loop admits a waiting request as soon as its device footprint fits the memory
budget and an execution slot is free (first fit in arrival order, with a
starvation guard that stops overtaking once the oldest request has waited
``max_wait_s``), so new work starts while earlier work is still running.
Every admitted request gets its own plan (device-resident model, wavefields,
checkpoints and traces) and its own HIP stream, so the device overlaps the
kernels of all requests in flight.  Each request is batched across its own
shots.  A request's plan set-up, input copies and kernel launches run on a
worker thread of its own: a long request enqueues thousands of launches, and
the HIP runtime blocks the enqueuing thread while the device's queues are
full, which in the loop's own thread held back admission and completion of
every other request.

Results do not depend on what else is in flight: a request's arithmetic is
fixed by its own inputs (tile shapes are chosen by rule, not by timing), and
its reductions run in a fixed order inside its own buffers.

The device memory of a finished request's plan is kept for the next plans
(``device_cache``, capped at the loop's budget) instead of being freed: hipFree
waits for the whole device, and a first hipMalloc of many GB maps fresh pages,
each of which took longer than a small request's solve.  What is kept is freed
when the loop ends.
"""

from __future__ import annotations

import collections
import concurrent.futures
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from .plan import Plan, device_cache, footprint_bytes


@dataclass
class Request:
    """One solve.  ``kind`` is "gradient" (needs ``dobs``) or "forward"."""
    rid: int
    kind: str
    shape: tuple
    h: float
    dt: float
    nt: int
    tx: np.ndarray            # linear source cells, one per shot
    rec: np.ndarray           # linear receiver cells
    m: np.ndarray             # squared slowness (float64)
    wavelet: np.ndarray
    dobs: np.ndarray | None = None
    arrival: float = 0.0      # seconds after the loop starts

    @property
    def updates(self):
        n = int(np.prod(self.shape)) * (self.nt - 1) * len(self.tx)
        return 2 * n if self.kind == "gradient" else n

    def footprint(self, cap=None):
        """Device bytes to reserve: the one-pass plan, or ``cap`` (the plan then runs
        its shots in several passes within it)."""
        b = footprint_bytes(self.shape, self.nt, len(self.tx), len(self.rec),
                            gradient=self.kind == "gradient")
        return b if cap is None else min(b, cap)


@dataclass
class Result:
    rid: int
    kind: str
    J: float | None
    grad: torch.Tensor | None
    data: torch.Tensor | None
    arrival: float
    admitted: float
    done: float
    updates: int
    workspace_bytes: int = 0
    enqueue_s: float = 0.0    # host seconds of plan set-up, copies and launches

    @property
    def latency(self):
        return self.done - self.arrival


@dataclass
class _Flight:
    req: Request
    held: int                 # bytes reserved at admission
    admitted: float
    future: concurrent.futures.Future | None = None
    plan: Plan | None = None
    event: torch.cuda.Event | None = None
    J: torch.Tensor | None = None
    grad: torch.Tensor | None = None
    data: torch.Tensor | None = None
    tensors: list = field(default_factory=list)
    enqueue_s: float = 0.0


class Server:
    """Single-device continuous-batching loop (one process, one GPU)."""

    def __init__(self, mem_budget=None, max_inflight=8, max_wait_s=5.0, use_graph=False):
        torch.cuda.init()
        free, total = torch.cuda.mem_get_info()
        self.budget = int(mem_budget if mem_budget is not None else 0.8 * free)
        self.max_inflight = int(max_inflight)
        self.max_wait_s = float(max_wait_s)
        self.use_graph = use_graph
        self.device = torch.device("cuda", torch.cuda.current_device())
        self.cap = self.budget // 2  # largest reservation of one request
        self._held = 0            # bytes of plans in flight

    # -- memory -----------------------------------------------------------------
    def _need(self, req):
        # kept blocks are 2 MiB multiples, up to 1.25x the buffer they serve
        return int(req.footprint(self.cap) * 1.25) + (64 << 20)

    def _fits(self, req, in_flight):
        need = self._need(req)
        if self._held + need <= self.budget:
            return True
        if need > self.budget:
            raise MemoryError(f"request {req.rid} needs {need} bytes, budget {self.budget}")
        return False

    # -- execution ----------------------------------------------------------------
    def _launch(self, req, now, pool):
        """Reserve the request's bytes and hand its set-up and launches to a worker."""
        if req.kind not in ("gradient", "forward"):
            raise ValueError(f"unknown request kind {req.kind!r}")
        fl = _Flight(req, self._need(req), now)
        self._held += fl.held
        fl.future = pool.submit(self._enqueue, fl)
        return fl

    def _enqueue(self, fl):
        """Worker thread: plan, input copies and the request's launches on its own
        stream, then an event that completes with them."""
        t = time.perf_counter()
        req = fl.req
        torch.cuda.set_device(self.device)
        stream = torch.cuda.Stream(device=self.device)
        est = req.footprint(self.cap)
        # the plan's own planner stays within what admission reserved
        fl.plan = Plan(req.shape, req.h, req.dt, req.nt, req.tx, req.rec, method="batched",
                       mem_budget=int(est * 1.05) + (64 << 20), autotune=False,
                       graph=self.use_graph, forward_only=req.kind == "forward")
        with torch.cuda.stream(stream):
            m = torch.as_tensor(req.m, dtype=torch.float64).to(self.device, non_blocking=False)
            w = torch.as_tensor(req.wavelet, dtype=torch.float64).to(self.device)
            if req.kind == "gradient":
                d = torch.as_tensor(req.dobs, dtype=torch.float64).to(self.device)
                fl.J, fl.grad = fl.plan.misfit_and_gradient(m, w, d, stream=stream)
                fl.tensors += [m, w, d]
            else:
                fl.data = fl.plan.forward_fmc(m, w, stream=stream)
                fl.tensors += [m, w]
            ev = torch.cuda.Event()
            ev.record(stream)
        fl.tensors.append(stream)
        fl.event = ev
        fl.enqueue_s = time.perf_counter() - t

    def run(self, requests, on_result=None, progress=None):
        """Serve ``requests`` (arrival times relative to now); returns results by rid."""
        device_cache(self.budget)
        try:
            with concurrent.futures.ThreadPoolExecutor(max_workers=self.max_inflight) as pool:
                return self._run(requests, on_result, progress, pool)
        finally:
            device_cache(0)

    @staticmethod
    def _finished(fl):
        """True once the worker has enqueued everything and the device has run it;
        a worker's exception is raised here."""
        if not fl.future.done():
            return False
        fl.future.result()
        return fl.event.query()

    def _run(self, requests, on_result, progress, pool):
        pending = collections.deque(sorted(requests, key=lambda r: r.arrival))
        queue, flights, results = [], [], {}
        t0 = time.perf_counter()
        n_total = len(pending)
        try:
            while pending or queue or flights:
                now = time.perf_counter() - t0
                while pending and pending[0].arrival <= now:
                    queue.append(pending.popleft())
                # admission: first fit in arrival order; no overtaking a request that
                # has waited longer than max_wait_s
                i = 0
                while i < len(queue) and len(flights) < self.max_inflight:
                    r = queue[i]
                    if self._fits(r, flights):
                        queue.pop(i)
                        flights.append(self._launch(r, time.perf_counter() - t0, pool))
                        continue
                    if now - queue[0].arrival > self.max_wait_s:
                        break
                    i += 1
                # completion
                still = []
                for fl in flights:
                    if self._finished(fl):
                        done = time.perf_counter() - t0
                        r = fl.req
                        res = Result(r.rid, r.kind, None if fl.J is None else float(fl.J.item()),
                                     fl.grad, fl.data, r.arrival, fl.admitted, done, r.updates,
                                     int(fl.plan.info["workspace_bytes"]), fl.enqueue_s)
                        results[r.rid] = res
                        fl.plan.close()  # its blocks are kept for the next plans
                        self._held -= fl.held
                        if on_result:
                            on_result(res)
                        if progress:
                            progress(len(results), n_total)
                    else:
                        still.append(fl)
                flights = still
                if not flights and not queue:
                    if pending:
                        time.sleep(max(0.0, min(pending[0].arrival - (time.perf_counter() - t0),
                                                0.01)))
                elif not any(f.future.done() and f.event.query() for f in flights):
                    time.sleep(2e-4)
        finally:
            # a failed request leaves the others' workers to finish before the pool
            # closes; their plans are closed here
            for fl in flights:
                try:
                    fl.future.result()
                except Exception:
                    pass
                if fl.plan is not None:
                    torch.cuda.synchronize(self.device)
                    fl.plan.close()
        return results


def summarize(results):
    """p50 / p99 latency and aggregate Gpt-updates/s of a finished run."""
    res = list(results.values())
    lat = np.array([r.latency for r in res])
    t_first = min(r.arrival for r in res)
    t_last = max(r.done for r in res)
    upd = sum(r.updates for r in res)
    enq = np.array([r.enqueue_s for r in res])
    return {"requests": len(res), "p50_s": float(np.percentile(lat, 50)),
            "p99_s": float(np.percentile(lat, 99)), "mean_s": float(lat.mean()),
            "makespan_s": t_last - t_first, "gpt_updates_per_s": upd / (t_last - t_first) / 1e9,
            "updates": upd, "enqueue_s_mean": float(enq.mean()), "enqueue_s_max": float(enq.max())}


def make_mix(n_requests=500, seed=5, nt_scale=1.0, shots=(16, 256), grids=None, rate=None,
             with_3d=True):
    """The serving mix: Poisson arrivals from default_rng(seed), gradient or forward with
    equal odds, grids from the C1/C3/C4 families, shots ~ U{shots}, phantoms seeded by
    request id.  Observed data of gradient requests is a forward model of the seeded
    phantom computed later by the caller (``fill_observed``)."""
    from . import phantom
    from .geometry import CylinderArray, RingArray
    from .plan import linear_indices
    from .sources import ricker

    rng = np.random.default_rng(seed)
    if grids is None:
        grids = [("2d", 256), ("2d", 512), ("2d", 1024)]
        if with_3d:
            grids += [("3d", 128), ("3d", 192), ("3d", 256)]
    reqs = []
    t = 0.0
    for rid in range(n_requests):
        kind = "gradient" if rng.random() < 0.5 else "forward"
        dim, n = grids[int(rng.integers(len(grids)))]
        ns = int(rng.integers(shots[0], shots[1] + 1))
        if dim == "2d":
            h, dt, nt = (2e-4, 5e-8, 1000) if n <= 512 else (1e-4, 2.5e-8, 4000)
            domain = (n - 1) * h
            geom = RingArray(max(ns, 16), 0.43 * domain, domain, h)
            shape = (n, n)
            c = phantom.seeded_discs(n, h, 1000 + rid, 3, inside_mm=0.31 * domain * 1e3)
        else:
            h, dt, nt = 2.5e-4, 3e-8, 2000
            domain = (n - 1) * h
            per = max(8, (ns + 7) // 8)
            geom = CylinderArray(8, per, 0.44 * domain, 0.44 * domain, domain, h)
            shape = (n, n, n)
            f = domain / 0.06375
            c = phantom.seeded_ellipsoids(n, h, 1000 + rid, 8, a_mm=(3 * f, 8 * f),
                                          inside_mm=20 * f)
        nt = max(8, int(round(nt * nt_scale)))
        rec = linear_indices(geom.idx, shape)
        tx = rec[np.linspace(0, len(rec) - 1, ns).astype(int)]
        m_true = phantom.velocity_to_m(c)
        m0 = np.full(shape, 1.0 / 1500.0 ** 2)
        req = Request(rid, kind, shape, h, dt, nt, tx, rec, m0 if kind == "gradient" else m_true,
                      ricker(nt, dt, 5e5), None, 0.0)
        req.m_true = m_true
        reqs.append(req)
    # arrival rate: about 80 % utilisation at `rate` Gpt-updates/s, if given
    if rate:
        mean_gap = np.mean([r.updates for r in reqs]) / (rate * 1e9) / 0.8
    else:
        mean_gap = 0.0
    for r in reqs:
        t += rng.exponential(mean_gap) if mean_gap > 0 else 0.0
        r.arrival = t
    return reqs


def fill_observed(reqs):
    """Synthesise observed data (device forward of each request's true model)."""
    for r in reqs:
        if r.kind != "gradient":
            continue
        with Plan(r.shape, r.h, r.dt, r.nt, r.tx, r.rec, autotune=False, graph=False,
                  forward_only=True) as p:
            d = p.forward_fmc(torch.as_tensor(r.m_true, device="cuda"),
                              torch.as_tensor(r.wavelet, device="cuda"))
            r.dobs = d.cpu().numpy()
    return reqs

"""Load generator for the request server (unc_server_*): reconstructions of
different sizes arriving as a Poisson process.

Request scenes: cameras log-uniform in [20, 800], points per camera uniform in
[25, 400], random-walk trajectories with track length mean 4 (uncertaintyte_amd
.scene.walk_scene) and a co-visibility window short enough that every camera
is pinned (the scene's null space is the 7-dim gauge, nothing more).  `n_distinct` seeded scenes are generated and the requests
cycle through them in a seeded random order.
"""
from __future__ import annotations

import threading
import time
from concurrent.futures import ProcessPoolExecutor
from dataclasses import dataclass, field
from typing import List

import numpy as np

from . import UncError, scene


def request_sizes(n: int, seed: int = 11):
    rng = np.random.default_rng(seed)
    cams = np.exp(rng.uniform(np.log(20), np.log(800), n)).astype(int)
    ppc = rng.integers(25, 401, n)
    return cams, ppc


def request_scene(n_cams: int, ppc: int, seed: int):
    # every camera must see enough points to pin its 9 parameters: the last
    # cameras of a walk are only in the windows of the last anchors, so the
    # window shrinks with the number of points per camera
    window = int(max(4, min(ppc // 5, n_cams // 4, 40)))
    sc = scene.walk_scene(int(n_cams), int(n_cams) * int(ppc), 4.0, seed=int(seed),
                          window=window)
    return scene.scene_jacobian(sc)


def _gen(args):
    return request_scene(*args)


def make_requests(n_requests: int = 512, n_distinct: int = 64, seed: int = 11,
                  workers: int = 8, max_cams: int = 800, pinned: bool = False):
    """(scenes, order): n_distinct seeded request scenes and the request order.
    pinned: the scenes' arrays in page-locked host memory (a server's receive
    buffers), so that their uploads run at the DMA engines' rate."""
    cams, ppc = request_sizes(n_distinct, seed)
    cams = np.minimum(cams, max_cams)
    args = [(int(c), int(p), seed * 1000 + i) for i, (c, p) in enumerate(zip(cams, ppc))]
    if workers > 1:
        with ProcessPoolExecutor(max_workers=workers) as ex:
            scenes = list(ex.map(_gen, args))
    else:
        scenes = [_gen(a) for a in args]
    if pinned:
        scenes = [s.pinned() for s in scenes]
    order = np.random.default_rng(seed + 1).integers(0, n_distinct, n_requests)
    return scenes, order


@dataclass
class LoadResult:
    latency_ms: List[float] = field(default_factory=list)
    queue_ms: List[float] = field(default_factory=list)
    makespan_s: float = 0.0
    n: int = 0
    failed: int = 0

    def summary(self):
        lat = np.array(self.latency_ms)
        q = np.array(self.queue_ms)
        return dict(requests=self.n, failed=self.failed,
                    solves_per_s=(self.n - self.failed) / self.makespan_s,
                    p50_ms=float(np.percentile(lat, 50)), p99_ms=float(np.percentile(lat, 99)),
                    mean_ms=float(lat.mean()), queue_p50_ms=float(np.percentile(q, 50)),
                    queue_p99_ms=float(np.percentile(q, 99)), makespan_s=self.makespan_s)


def run_load(server, scenes, order, rate: float, precision: int, seed: int = 5, check=None):
    """Submit scenes[order[i]] at Poisson arrival times of the given rate (1/s);
    returns LoadResult with per-request latency (submit -> done) and queueing.
    check(i, result), when given, is called with each request's result as it
    completes; a False return counts the request as failed."""
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, len(order)))
    tickets: List[int] = []
    lock = threading.Condition()
    res = LoadResult(n=len(order))
    done_submitting = [False]

    def waiter():
        k = 0
        while True:
            with lock:
                while k >= len(tickets) and not done_submitting[0]:
                    lock.wait()
                if k >= len(tickets):
                    return
                t = tickets[k]
            try:
                out, lat, q = server.wait(t)
                res.latency_ms.append(lat)
                res.queue_ms.append(q)
                if check is not None and not check(k, out):
                    res.failed += 1
            except UncError:
                res.failed += 1
            k += 1

    th = threading.Thread(target=waiter)
    th.start()
    t0 = time.perf_counter()
    for i, idx in enumerate(order):
        delay = t0 + arrivals[i] - time.perf_counter()
        if delay > 0:
            time.sleep(delay)
        t = server.submit(scenes[idx], precision=precision)
        with lock:
            tickets.append(t)
            lock.notify()
    with lock:
        done_submitting[0] = True
        lock.notify()
    th.join()
    res.makespan_s = time.perf_counter() - t0
    return res

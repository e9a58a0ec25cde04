"""Open-loop serving benchmark over the native serving loop (cvm_server_*).

Requests arrive as a Poisson stream; each is a residual + tangent evaluation on
one of the prepared sheets at its own displacement. A submitting thread issues
requests at their arrival times; a collecting thread waits for them in order.
Latency is the server's submit-to-completion time (x copied in, kernels, F
copied out; K stays in HBM in the slot or a caller buffer).
"""
from __future__ import annotations

import threading
import time

import numpy as np

import curvami_gpu as G
import workloads as W


def prepare_sheets(sides, kernel=0):
    """Device-resident sheets (neighbours, shape functions, assembly plan) by
    side. Returns (sheet descriptions, plates, seconds)."""
    sheets, plates, t0 = {}, {}, time.perf_counter()
    for n in sorted(set(sides)):
        sheets[n] = W.serving_sheet(n, kernel)
        plates[n] = G.plate_from_sheet(sheets[n])
    return sheets, plates, time.perf_counter() - t0


def displacements(reqs, sheets):
    """Rigid fold of the + side by the request's angle plus N(0, 1e-3) noise."""
    return [W.rigid_fold_displacement(sheets[n], ang, seed=seed, sigma=1e-3) for (n, ang, seed) in reqs]


def run_rate(server, sheet_ids, reqs, xs, rate, seed, flags):
    """One open-loop run at `rate` requests/s. Returns latency percentiles (ms)
    and the achieved throughput (requests/s)."""
    n = len(reqs)
    arrivals = W.poisson_arrivals(rate, n, seed)
    outs = [np.empty(x.size) for x in xs]
    ids = [None] * n
    lat = np.zeros(n); queue = np.zeros(n); dev = np.zeros(n)
    ready = threading.Semaphore(0)
    done_t = [0.0]

    def collect():
        for i in range(n):
            ready.acquire()
            q, d, t = server.wait(ids[i])
            queue[i], dev[i], lat[i] = q, d, t
        done_t[0] = time.perf_counter()

    th = threading.Thread(target=collect)
    b0 = server.batching_stats()
    th.start()
    t0 = time.perf_counter()
    for i in range(n):
        now = time.perf_counter() - t0
        if arrivals[i] > now:
            time.sleep(arrivals[i] - now)
        ids[i] = server.submit(sheet_ids[reqs[i][0]], xs[i], outs[i], flags)
        ready.release()
    th.join()
    span = done_t[0] - t0
    b1 = server.batching_stats()
    return {"rate": rate, "requests": n, "p50_ms": float(np.percentile(lat, 50)),
            "p99_ms": float(np.percentile(lat, 99)), "mean_ms": float(lat.mean()),
            "queue_p50_ms": float(np.percentile(queue, 50)), "device_p50_ms": float(np.percentile(dev, 50)),
            "throughput_rps": n / span, "offered_rps": n / arrivals[-1],
            "shared_tile_launches": b1["shared_tile_launches"] - b0["shared_tile_launches"],
            "requests_in_shared_launches": b1["requests_in_shared_launches"] - b0["requests_in_shared_launches"]}


def size_class_latency(server, sheet_ids, plates, n, x, flags, reps=20):
    """Latency of a request of side n alone (no other traffic): median and p99 ms."""
    out = np.empty(x.size)
    lat = []
    for _ in range(reps):
        rid = server.submit(sheet_ids[n], x, out, flags)
        lat.append(server.wait(rid)[2])
    lat = np.array(lat)
    return float(np.percentile(lat, 50)), float(np.percentile(lat, 99))

"""Serving loop over the request server (rtm_server_*, rtmuq2d_amd.Server).

A stream of migration requests arrives over wall-clock time; each is
submitted when its arrival time has passed, the server admits queued
requests by HBM footprint and advances everything in flight one block of
time steps per tick (continuous batching), and each completion is stamped
when its image reaches the host.  The request inputs are made on the host
at submission time (``make`` callbacks), overlapping the device's work on the
requests already in flight.
"""
from __future__ import annotations

import dataclasses
import time
from typing import Callable, Dict, List, Optional

import numpy as np


@dataclasses.dataclass
class Request:
    """One migration: ``make()`` returns (params, padded velocity, wavelet,
    seismogram, shot_x, shot_z); ``arrival`` is in seconds from the start."""
    arrival: float
    make: Callable
    label: str = ""


def poisson_arrivals(n: int, rate: float, seed: int) -> np.ndarray:
    """n arrival times of a Poisson process of the given rate (per second)."""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, n))


def percentiles(lat) -> Dict[str, float]:
    a = np.asarray(lat, np.float64)
    return {"p50": float(np.percentile(a, 50)), "p95": float(np.percentile(a, 95)),
            "p99": float(np.percentile(a, 99)), "mean": float(a.mean()), "max": float(a.max())}


def run(server, requests: List[Request], on_done: Optional[Callable] = None, log=None,
        progress: Optional[Callable] = None, progress_every: int = 100) -> Dict:
    """Drive ``server`` through ``requests`` (sorted by arrival).  Returns
    per-request latency (completion - arrival, seconds) of the completed
    requests, the wall time, the image throughput and the failed requests
    (index -> reason: the server retires a failed request and carries on).
    ``on_done(index, image)`` sees every result (image None for a failed
    request); ``progress(done, seconds)`` is called after every
    ``progress_every`` retirements."""
    order = sorted(range(len(requests)), key=lambda i: requests[i].arrival)
    t0 = time.perf_counter()
    now = lambda: time.perf_counter() - t0  # noqa: E731
    nxt = 0
    pending = 0
    id2idx: Dict[int, int] = {}
    done_at = [None] * len(requests)
    failed: Dict[int, str] = {}
    max_active = max_queued = 0
    while nxt < len(order) or pending:
        while nxt < len(order) and requests[order[nxt]].arrival <= now():
            i = order[nxt]
            p, vpad, wav, seis, sx, sz = requests[i].make()
            id2idx[server.submit(p, vpad, wav, seis, sx, sz)] = i
            pending += 1
            nxt += 1
        if pending:
            server.tick()
            st = server.stats()
            max_active = max(max_active, st["active"])
            max_queued = max(max_queued, st["queued"])
            for rid, img in server.done():
                i = id2idx.pop(rid)
                done_at[i] = now()
                if img is None:
                    failed[i] = getattr(server, "failed", {}).get(rid, "failed")
                pending -= 1
                ndone = len(requests) - pending - (len(order) - nxt)
                if progress and ndone % progress_every == 0:
                    progress(ndone, now())
                if on_done:
                    on_done(i, img)
                if log:
                    log(f"[serve] request {i} ({requests[i].label}) done at {done_at[i]:.2f} s, "
                        f"latency {done_at[i] - requests[i].arrival:.3f} s; active {st['active']}, queued {st['queued']}")
        elif nxt < len(order):
            time.sleep(max(0.0, min(requests[order[nxt]].arrival - now(), 0.05)))
    wall = now()
    ok = [i for i in range(len(requests)) if i not in failed]
    lat = [done_at[i] - requests[i].arrival for i in ok]
    return {"latency_s": percentiles(lat) if lat else {}, "wall_seconds": wall,
            "images_per_hour": len(ok) * 3600.0 / wall,
            "max_active": max_active, "max_queued": max_queued, "latencies": lat,
            "failed": failed}


def sequential_latencies(arrivals, service) -> List[float]:
    """Latencies of one-at-a-time service in arrival order (a single FIFO
    server whose service times are ``service``)."""
    t = 0.0
    out = []
    for a, s in sorted(zip(arrivals, service)):
        t = max(t, a) + s
        out.append(t - a)
    return out

"""The serving loop (rtmuq2d_amd/serve.py) on the CPU against a stand-in
server: every request is submitted once and collected once, latencies are
completion minus arrival, and progress is reported every N completions."""
from __future__ import annotations

import numpy as np

from rtmuq2d_amd import serve


class _FakeServer:
    """Completes each request two ticks after its submission."""

    def __init__(self):
        self.next_id = 0
        self.active = {}
        self.finished = []
        self.ticks = 0

    def submit(self, p, vpad, wav, seis, sx, sz):
        rid = self.next_id
        self.next_id += 1
        self.active[rid] = self.ticks
        return rid

    def tick(self):
        self.ticks += 1
        for rid, t0 in list(self.active.items()):
            if self.ticks - t0 >= 2:
                del self.active[rid]
                self.finished.append(rid)
        return len(self.active)

    def stats(self):
        return {"active": len(self.active), "queued": 0}

    def done(self):
        out = [(rid, np.zeros(1, np.float32)) for rid in self.finished]
        self.finished = []
        return out


def test_serve_loop_collects_every_request_and_reports_progress():
    n = 250
    reqs = [serve.Request(0.0, (lambda: (None, None, None, None, 0, 0)), "x") for _ in range(n)]
    seen, marks = [], []
    srv = _FakeServer()
    res = serve.run(srv, reqs, on_done=lambda i, img: seen.append(i),
                    progress=lambda k, t: marks.append(k), progress_every=100)
    assert sorted(seen) == list(range(n))
    assert srv.next_id == n
    assert marks == [100, 200]
    lat = res["latencies"]
    assert len(lat) == n and min(lat) >= 0.0 and res["wall_seconds"] >= max(lat)


class _FailingServer(_FakeServer):
    """Fails every third request (as the request server retires a request
    whose launch group failed: image None, the reason kept by the server)."""

    def __init__(self):
        super().__init__()
        self.failed = {}

    def done(self):
        out = []
        for rid in self.finished:
            if rid % 3 == 0:
                self.failed[rid] = f"request {rid} failed: launch group"
                out.append((rid, None))
            else:
                out.append((rid, np.zeros(1, np.float32)))
        self.finished = []
        return out


def test_serve_loop_reports_failed_requests_and_serves_the_rest():
    n = 30
    reqs = [serve.Request(0.0, (lambda: (None, None, None, None, 0, 0)), "x") for _ in range(n)]
    images = {}
    srv = _FailingServer()
    res = serve.run(srv, reqs, on_done=lambda i, img: images.__setitem__(i, img))
    assert sorted(images) == list(range(n))
    assert sorted(res["failed"]) == [i for i in range(n) if i % 3 == 0]
    assert all(images[i] is None for i in res["failed"])
    assert all("launch group" in why for why in res["failed"].values())
    assert len(res["latencies"]) == n - len(res["failed"])

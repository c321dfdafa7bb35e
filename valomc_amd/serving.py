"""Serving harness over vmc_server: a device-resident mesh cache and a
Poisson request stream driver that reports photons/s and p50/p99 latency.

The cache keeps Problems (meshes with their media and sources) resident on
the device, keyed by a hash of their content, and evicts the least recently
used ones that no running request holds when a new one would exceed its
byte budget.  ``drive`` submits each request at its arrival time and
collects completion latencies (submit to outputs written, timed inside the
library).
"""
from __future__ import annotations

import hashlib
import time
from collections import OrderedDict
from typing import Optional, Dict, List

import numpy as np

from . import FP32, FP64, Problem, Server


def content_key(p) -> str:
    h = hashlib.sha1()
    for name in ("H", "BH", "r", "mua", "mus", "g", "n", "BCType", "BCn", "BCLNormal", "BCLDirType",
                 "BCIntensity", "GaussianSigma"):
        a = getattr(p, name, None)
        if a is None:
            continue
        a = np.ascontiguousarray(a)
        h.update(name.encode())
        h.update(str(a.dtype).encode())
        h.update(str(a.shape).encode())
        h.update(a.tobytes())
    return h.hexdigest()


def device_bytes(p, slots: Optional[int] = None) -> int:
    """Resident element records of a problem (FP32 and FP64 copies), one per
    device slot: `slots` when the problem exists (Problem.slots), else the
    estimate for a grid mesh (3D: 4 slots per 3 tetrahedra; 2D: whole
    sectors of two-triangle cells, no padding)."""
    if slots is None:
        slots = -(-int(p.Ne) * 4 // 3) if p.dim == 3 else int(p.Ne)
        slots = -(-slots // 8) * 8
    per = 128 + 256 if p.dim == 3 else 128 + 192
    return int(slots) * per + int(p.Nb) * 16


class MeshCache:
    def __init__(self, budget_bytes: int = 16 << 30):
        self.budget = int(budget_bytes)
        self.used = 0
        self.entries: "OrderedDict[str, list]" = OrderedDict()   # key -> [Problem, bytes, holders]
        self.keys: Dict[int, str] = {}

    def get(self, p) -> Problem:
        key = self.keys.get(id(p)) or content_key(p)
        self.keys[id(p)] = key
        if key in self.entries:
            self.entries.move_to_end(key)
            e = self.entries[key]
            e[2] += 1
            return e[0]
        need = device_bytes(p)
        for k in list(self.entries):
            if self.used + need <= self.budget:
                break
            prob, b, holders = self.entries[k]
            if holders == 0:
                prob.close()
                self.used -= b
                del self.entries[k]
        prob = Problem(p)
        need = device_bytes(p, prob.slots)
        self.entries[key] = [prob, need, 1]
        self.used += need
        return prob

    def release(self, prob: Problem):
        for e in self.entries.values():
            if e[0] is prob:
                e[2] -= 1
                return

    def close(self):
        for prob, _, _ in self.entries.values():
            prob.close()
        self.entries.clear()
        self.used = 0


def drive(server: Server, cache: MeshCache, pool: List, reqs: List[dict], rate: float, seed: int = 7,
          keep=()):
    """Submit ``reqs`` at Poisson arrival times of the given rate (requests/s);
    a second thread waits for completions in submission order so output
    buffers are freed as requests finish.  Returns per-request records with
    the latency (and the element fluence for the indices in ``keep``) and the
    wall time from the first arrival to the last completion."""
    import queue
    import threading
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, len(reqs))
    by_kind = {"2d": [p for p in pool if p.notes.get("kind") == "2d"],
               "3d": [p for p in pool if p.notes.get("kind") == "3d"]}
    pending: "queue.Queue" = queue.Queue()
    out: List[dict] = []
    keep = set(keep)

    def waiter():
        while True:
            item = pending.get()
            if item is None:
                return
            i, q, prob, tk, t_sub = item
            r = server.wait(tk)
            rec = dict(q, submit_s=t_sub, latency_ms=r.total_ms, kernel_ms=r.kernel_ms)
            if i in keep:
                rec["ER"] = r.ER
            out.append(rec)
            cache.release(prob)

    th = threading.Thread(target=waiter, daemon=True)
    th.start()
    t0 = time.perf_counter()
    t_arr = t0
    for i, (q, gap) in enumerate(zip(reqs, gaps)):
        t_arr += gap
        now = time.perf_counter()
        if t_arr > now:
            time.sleep(t_arr - now)
        p = by_kind[q["kind"]][q["index"]]
        prob = cache.get(p)
        tk = server.submit(prob, q["photons"], seed=q["seed"], f=q["f"])
        pending.put((i, q, prob, tk, time.perf_counter() - t0))
    pending.put(None)
    th.join()
    wall = max(o["submit_s"] + o["latency_ms"] / 1e3 for o in out)
    return out, wall

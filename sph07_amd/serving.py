"""Continuous-batching request loop over the batched step (sph_batch_*).

Requests (independent 2D scenes with a step count) arrive on a seeded Poisson
process.  The loop:
  * admits arrived requests in arrival order while the device-memory budget
    (sph_scene_footprint) and the live-slot limit allow -- a request that
    arrives while others run is initialised on its own stream while the batch
    steps, and joins the batch at the next chunk boundary;
  * advances every live scene together, `chunk` steps at a time, in one
    batched launch sequence per step (state stays device-resident);
  * retires finished scenes, frees their memory and slot, and records the
    time to solution (completion - arrival).
The one-scene-at-a-time baseline runs the same requests in arrival order,
each alone through sph_solution.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass, field
from typing import Callable, List, Optional

import numpy as np

from . import scenes
from .deck import Deck


class SplitMix64:
    """SplitMix64 (Steele, Lea, Flood 2014): a seeded 64-bit generator."""

    def __init__(self, seed: int):
        self.s = seed & 0xFFFFFFFFFFFFFFFF

    def next_u64(self) -> int:
        self.s = (self.s + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        return z ^ (z >> 31)

    def uniform(self) -> float:
        return (self.next_u64() >> 11) * (1.0 / (1 << 53))

    def log_uniform(self, lo: float, hi: float) -> float:
        return math.exp(math.log(lo) + self.uniform() * (math.log(hi) - math.log(lo)))

    def randint(self, lo: int, hi: int) -> int:
        return lo + int(self.uniform() * (hi - lo + 1))


def scene_of_size(kind: str, n_target: int, nu: Optional[float] = None) -> Deck:
    """A 2D scene of the given kind with about n_target fluid particles."""
    if kind == "poiseuille":
        n = max(8, int(round(math.sqrt(n_target))))
        kw = {} if nu is None else {"nu": nu}
        return scenes.poiseuille(nx=n, ny=n, dx=1e-3 / n, **kw)
    if kind == "cavity":
        n = max(8, int(round(math.sqrt(n_target))))
        kw = {} if nu is None else {"nu": nu}
        return scenes.cavity(n=n, **kw)
    if kind == "droplet":
        n = max(16, int(round(math.sqrt(n_target))))
        return scenes.droplet(n_box=n)
    if kind == "cylinder":
        nx = max(24, int(round(math.sqrt(n_target / 2.0))))
        dx = 0.02 / nx
        return scenes.cylinder(nx=nx, ny=2 * nx, dx=dx, radius=max(1.25e-3, 3.0 * dx))
    raise ValueError(kind)


@dataclass
class Request:
    rid: int
    kind: str
    deck: Deck
    steps: int
    arrival: float
    t_admit: float = -1.0
    t_done: float = -1.0
    slot: int = -1
    result: Optional[dict] = None     # field id -> array, with keep_results

    @property
    def updates(self) -> int:
        return self.deck.np * self.steps


def draw_stream(seed: int = 0x5EED0004, n: int = 32, rate: float = 4.0,
                size_range=(4_000, 500_000), steps_range=(500, 5000),
                mix=(("poiseuille", 0.4), ("cavity", 0.3), ("droplet", 0.2), ("cylinder", 0.1)),
                steps_scale: float = 1.0):
    """(kind, size, steps, arrival) of each request: Poisson arrivals (rate per
    second), scene kinds by weight, log-uniform sizes, uniform step counts."""
    g = SplitMix64(seed)
    t = 0.0
    out = []
    cum = np.cumsum([w for _, w in mix])
    for _ in range(n):
        t += -math.log(1.0 - g.uniform()) / rate
        u = g.uniform() * cum[-1]
        kind = mix[int(np.searchsorted(cum, u, side="right"))][0]
        size = int(g.log_uniform(*size_range))
        steps = max(1, int(g.randint(*steps_range) * steps_scale))
        out.append((kind, size, steps, t))
    return out


def serving_stream(seed: int = 0x5EED0004, n: int = 32, rate: float = 4.0, **kw) -> List[Request]:
    return [Request(k, kind, scene_of_size(kind, size), steps, t)
            for k, (kind, size, steps, t) in enumerate(draw_stream(seed, n, rate, **kw))]


def ensemble(seed: int = 0x5EED0003, n: int = 512, size_range=(2_000, 64_000),
             nu_range=(1e-6, 1e-4), steps: int = 2000) -> List[Request]:
    """Independent scenes, alternately Poiseuille and lid-driven cavity, with
    log-uniform particle counts and viscosities; all arrive at t = 0."""
    g = SplitMix64(seed)
    out = []
    for k in range(n):
        kind = "poiseuille" if k % 2 == 0 else "cavity"
        size = int(g.log_uniform(*size_range))
        nu = g.log_uniform(*nu_range)
        out.append(Request(k, kind, scene_of_size(kind, size, nu=nu), steps, 0.0))
    return out


def percentile(xs, q):
    return float(np.percentile(np.asarray(xs, dtype=np.float64), q)) if len(xs) else float("nan")


@dataclass
class ServeStats:
    mode: str
    requests: int
    makespan_s: float
    particle_updates: int
    tts: List[float] = field(default_factory=list)

    @property
    def throughput(self) -> float:
        return self.particle_updates / self.makespan_s if self.makespan_s > 0 else 0.0

    def summary(self) -> dict:
        return {"mode": self.mode, "requests": self.requests, "makespan_s": round(self.makespan_s, 4),
                "particle_updates_per_s": self.throughput,
                "p50_time_to_solution_s": percentile(self.tts, 50),
                "p99_time_to_solution_s": percentile(self.tts, 99)}


def serve_batched(requests: List[Request], arm: int, hbm_budget: float = 200e9,
                  max_live: int = 512, chunk: int = 25, keep_results: bool = False,
                  clock: Callable[[], float] = time.perf_counter) -> ServeStats:
    """Continuous batching: admit on arrival (memory and slots permitting),
    advance all live scenes together, retire as they finish.  The host works
    while the device steps: each chunk is queued (sph_batch_launch) before the
    finished scenes are freed and the arrived requests are initialised on
    their own streams; those join the live set at the next chunk boundary."""
    from .native import Batch, F_RHO, F_V, F_X, footprint
    batch = Batch(2, arm)
    pending = sorted(requests, key=lambda r: r.arrival)
    qi = 0
    resident = {}
    used = 0
    t0 = clock()

    def admit() -> int:
        nonlocal qi, used
        now = clock() - t0
        n = 0
        while qi < len(pending) and pending[qi].arrival <= now and len(resident) < max_live:
            r = pending[qi]
            fp = footprint(r.deck)
            if used + fp > hbm_budget and resident:
                break
            r.slot = batch.add(r.deck, r.steps)
            r.t_admit = clock() - t0
            r.fp = fp
            used += fp
            resident[r.slot] = r
            qi += 1
            n += 1
        return n

    inflight = False
    while qi < len(pending) or resident:
        finished = []
        if inflight:
            batch.sync()
            inflight = False
            now = clock() - t0
            for s in batch.finished():
                r = resident[s]
                r.t_done = now
                st = batch.status(s)
                if st["error"]:
                    raise RuntimeError(f"request {r.rid} failed on the device (code {st['error']})")
                if keep_results:
                    r.result = {f: batch.field(s, f) for f in (F_X, F_V, F_RHO)}
                finished.append(s)
        if batch.live:
            batch.launch(chunk)
            inflight = True
        for s in finished:                 # freed at the next sync while steps are queued
            r = resident.pop(s)
            batch.remove(s)
            used -= r.fp
        admitted = admit()
        if not inflight and not admitted and not batch.live and qi < len(pending):
            time.sleep(max(0.0, pending[qi].arrival - (clock() - t0)))
    end = clock() - t0
    batch.close()
    first = min(r.arrival for r in requests)
    return ServeStats("batched", len(requests), end - first, sum(r.updates for r in requests),
                      [r.t_done - r.arrival for r in requests])


def serve_one_at_a_time(requests: List[Request], arm: int, keep_results: bool = False,
                        clock: Callable[[], float] = time.perf_counter) -> ServeStats:
    """Baseline: requests in arrival order, each alone to completion.  Each
    scene runs at the library's default launch shapes: sph_solution would
    otherwise time every shape on the first call of a scene of 65,536
    particles or more, inside the request's time to solution (the batched
    loop tunes on its own steps, at no extra cost)."""
    from .native import Scene, F_RHO, F_V, F_X
    pending = sorted(requests, key=lambda r: r.arrival)
    t0 = clock()
    for r in pending:
        now = clock() - t0
        if r.arrival > now:
            time.sleep(r.arrival - now)
        r.t_admit = clock() - t0
        sc = Scene(r.deck, arm)
        sc.autotune(0)
        sc.Solution(r.steps)
        sc.synchronize()
        r.t_done = clock() - t0
        if keep_results:
            r.result = {f: sc.field(f) for f in (F_X, F_V, F_RHO)}
        sc.close()
    end = clock() - t0
    first = min(r.arrival for r in requests)
    return ServeStats("one_at_a_time", len(requests), end - first, sum(r.updates for r in requests),
                      [r.t_done - r.arrival for r in requests])

"""Serving loop: simulation requests arriving over time share one GPU.

Requests are admitted into one resident batch (pp_batch_*) by device
footprint: a request needs an agent reservation, a contiguous run of grid
cells and a replica slot of the batch, all sized when the batch is created
from an HBM budget. The loop runs the batch in chunks of `chunk` iterations
(continuous batching): between chunks it retires finished requests, whose
slots, cells and reservations return to the pool, and admits the requests
that arrived meanwhile, first come first served; a request that does not fit
yet lets later, smaller ones in until it has waited `backfill_s` seconds,
after which admission holds for it. Each request's latency is its completion
(observed at the end of the chunk in which it finished) minus its arrival.

A request's results do not depend on what else runs: every replica of a
batch computes exactly the statistics of the same simulation run alone.
"""
from __future__ import annotations

import time
from collections import deque
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from .sim import Batch, Params, DONE, FAILED
from .workloads import footprint_bytes, reserve_for


@dataclass
class Request:
    rid: int
    arrival: float
    params: dict
    seed: int
    kind: str = ""
    L: int = 0
    admitted: Optional[float] = None
    done: Optional[float] = None
    state: int = 0
    error: int = 0
    steps: int = 0
    agent_steps: int = 0
    rows: Optional[np.ndarray] = field(default=None, repr=False)


def batch_limits(hbm_bytes: int, max_replicas: int = 256,
                 max_grass_restart: int = 15):
    """Agent and cell capacities of a batch filling hbm_bytes: cells at one
    per 1.5 reserved agents (the reservation is at least one agent per cell,
    v2s needs 0.96), 33 bytes per agent and 1 per cell."""
    agents = int(hbm_bytes / (33 + 1 / 1.5))
    agents = min(agents, (1 << 29) - 1)
    cells = min(int(agents / 1.5), 0xFFFFFFFE)
    while Batch.footprint(agents, cells, max_replicas,
                          max_grass_restart) > hbm_bytes:
        agents = int(agents * 0.98)
        cells = int(cells * 0.98)
    return agents, cells


class Server:
    def __init__(self, hbm_bytes: int, max_replicas: int = 256,
                 chunk: int = 32, backfill_s: float = 2.0,
                 keep_rows: bool = False, max_grass_restart: int = 15,
                 device: int = 0):
        self.agents, self.cells = batch_limits(hbm_bytes, max_replicas,
                                               max_grass_restart)
        self.batch = Batch(self.agents, self.cells, max_replicas,
                           max_grass_restart, device=device)
        self.hbm_bytes = Batch.footprint(self.agents, self.cells, max_replicas,
                                         max_grass_restart)
        self.chunk = chunk
        self.backfill_s = backfill_s
        self.keep_rows = keep_rows
        self.requests: list[Request] = []

    def close(self):
        self.batch.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def fits_ever(self, d: dict) -> bool:
        return (reserve_for(d) <= self.agents
                and d["grid_x"] * d["grid_y"] <= self.cells)

    def _try_admit(self, req: Request, now: float) -> bool:
        rid = self.batch.add(Params.from_values(**req.params), req.seed,
                             reserve_for(req.params))
        if rid is None:
            return False
        req.admitted = now
        req.state = 1
        self._running[rid] = req
        return True

    def _admit(self, queue: deque, now: float):
        blocked_since = None
        for req in list(queue):
            if self._try_admit(req, now):
                queue.remove(req)
                continue
            if blocked_since is None:
                blocked_since = req.arrival
                if now - blocked_since > self.backfill_s:
                    break           # hold admission for the oldest request

    def _retire(self, now: float):
        for brid, req in list(self._running.items()):
            st = self.batch.replica(brid)
            if st["state"] not in (DONE, FAILED):
                continue
            req.done = now
            req.state = st["state"]
            req.error = st["error"]
            req.steps = st["steps"]
            req.agent_steps = st["agent_steps"]
            if self.keep_rows:
                req.rows = self.batch.rows_array(brid)
            self.batch.release(brid)
            del self._running[brid]

    def run(self, stream, clock=time.perf_counter, log=None):
        """Serves a list of request dicts (arrival in seconds from the start,
        params, seed, ...) in real time. Returns the Request records and the
        summary."""
        self._running: dict[int, Request] = {}
        reqs = [Request(rid=i, arrival=s["arrival"], params=s["params"],
                        seed=s["seed"], kind=s.get("kind", ""),
                        L=s.get("L", 0)) for i, s in enumerate(stream)]
        for r in reqs:
            if not self.fits_ever(r.params):
                raise ValueError(f"request {r.rid} can never fit the batch")
        pending = deque(sorted(reqs, key=lambda r: r.arrival))
        queue: deque[Request] = deque()
        t0 = clock()
        ticks = 0
        chunks = 0
        max_running = 0
        last_log = 0.0
        while pending or queue or self._running:
            now = clock() - t0
            while pending and pending[0].arrival <= now:
                queue.append(pending.popleft())
            if queue:
                self._admit(queue, now)
            if self._running:
                max_running = max(max_running, len(self._running))
                self.batch.run(self.chunk)
                ticks += self.chunk
                chunks += 1
                self._retire(clock() - t0)
            elif pending:
                time.sleep(max(0.0, min(pending[0].arrival - now, 0.05)))
            if log and clock() - t0 - last_log > 10.0:
                last_log = clock() - t0
                done = sum(r.done is not None for r in reqs)
                log(f"serve t={last_log:.0f}s done {done}/{len(reqs)} "
                    f"running {len(self._running)} queued {len(queue)}")
        end = clock() - t0
        self.requests = reqs
        summary = summarize(reqs, end, ticks, chunks, self.hbm_bytes)
        summary["max_concurrent"] = max_running
        return reqs, summary


def summarize(reqs, wall, ticks, chunks, hbm_bytes):
    ok = [r for r in reqs if r.state == DONE]
    lat = np.array([r.done - r.arrival for r in ok]) if ok else np.zeros(0)
    wait = np.array([r.admitted - r.arrival for r in ok]) if ok else lat
    agent_steps = sum(r.agent_steps for r in reqs)
    return {
        "requests": len(reqs),
        "completed": len(ok),
        "failed": sum(r.state == FAILED for r in reqs),
        "wall_s": round(wall, 3),
        "p50_latency_s": round(float(np.percentile(lat, 50)), 3)
        if len(lat) else None,
        "p99_latency_s": round(float(np.percentile(lat, 99)), 3)
        if len(lat) else None,
        "p50_queue_s": round(float(np.percentile(wait, 50)), 3)
        if len(wait) else None,
        "agent_steps": int(agent_steps),
        "agent_steps_per_s": agent_steps / wall if wall > 0 else 0.0,
        "ticks": ticks,
        "chunks": chunks,
        "batch_hbm_bytes": int(hbm_bytes),
        "footprint_bytes_total": int(sum(footprint_bytes(r.params)
                                         for r in reqs)),
    }

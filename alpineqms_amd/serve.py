"""Many designs on one GPU: batched stepping (config 4) and a serving loop
with continuous admission (config 5).

The reference runs one design per process (QMS_3D.f90:117-121 reads one
inputs file; master.py:84 launches it).  Here every design is one library
handle with its own HIP stream, device-resident fields and captured
iteration graphs, so the kernels of different designs overlap on the GPU.

* ``run_batch``: set every design up (coefficients, steady state), then step
  all of them at once, each on a library thread (``Solver.run_async``), and
  time the stepping phase as a whole against the same steps one design at a
  time.
* ``serve``: requests arrive at given times; a request is admitted when a
  worker slot is free and its device footprint (``native.estimate_bytes``)
  fits the memory budget next to the requests in flight, in arrival order;
  each admitted request is set up, solved and stepped on its own worker
  thread while others run, and frees its memory when it ends.  Reported:
  per-request latency (arrival to last step), queueing delay, p50/p99, and
  aggregate DoF.steps/s.
"""
from __future__ import annotations

import threading
import time
from dataclasses import dataclass, field
from typing import Callable, List, Optional, Sequence

import numpy as np

from . import native, program
from .qms_io import material_slots

GPU_BYTES = 288 * 10**9       # HBM of one MI355X


def make_solver(design, precision: int, **kw) -> native.Solver:
    inp = design.inputs
    slots, props = material_slots(inp, design.geom)
    return native.Solver(inp.cells, inp.h, inp.dt, inp.bc_vector(), slots, props, precision=precision, **kw)


def velocities(design, first: int, count: int) -> np.ndarray:
    """Frame velocities of steps first .. first+count-1 (QMS_3D.f90:1431-1449)."""
    inp = design.inputs
    return np.array([program.step_velocity(inp.traj, inp.dt, n) for n in range(first, first + count)])


def footprint(design, precision: int, restart: int = 30) -> int:
    return native.estimate_bytes(design.inputs.cells, precision, restart)


@dataclass
class BatchResult:
    dof_steps: int
    batched_s: float
    sequential_s: float
    iters: List[int]

    @property
    def speedup(self) -> float:
        return self.sequential_s / self.batched_s


def run_batch(designs: Sequence, precision: int, steps: int, tol: float = 1e-12,
              sequential: bool = True) -> BatchResult:
    """Steps `steps` time steps of every design, all at once, after setting
    them up; with `sequential`, also the same steps one design at a time (on
    a second set of handles in the same state) for comparison."""
    def setup():
        out = []
        for d in designs:
            s = make_solver(d, precision, tol=tol)
            s.steady(program.steady_velocity(d.inputs.traj, d.inputs.dt))
            out.append(s)
        return out

    vs = [velocities(d, 1, steps) for d in designs]
    dof_steps = sum(d.dof for d in designs) * steps
    ss = setup()
    try:
        i0 = [s.stats().total_iters for s in ss]
        t0 = time.perf_counter()
        for s, v in zip(ss, vs):
            s.run_async(v)
        for s in ss:
            s.wait()
        t_batch = time.perf_counter() - t0
        iters = [int(s.stats().total_iters - a) for s, a in zip(ss, i0)]
    finally:
        for s in ss:
            s.close()
    t_seq = float("nan")
    if sequential:
        ss = setup()
        try:
            t0 = time.perf_counter()
            for s, v in zip(ss, vs):
                s.run(v)
            t_seq = time.perf_counter() - t0
        finally:
            for s in ss:
                s.close()
    return BatchResult(dof_steps, t_batch, t_seq, iters)


@dataclass
class RequestRecord:
    name: str
    dof: int
    steps: int
    bytes: int
    arrival: float
    start: float = 0.0
    end: float = 0.0
    iters: int = 0
    error: Optional[str] = None

    @property
    def latency(self) -> float:
        return self.end - self.arrival

    @property
    def wait(self) -> float:
        return self.start - self.arrival


@dataclass
class ServeResult:
    records: List[RequestRecord] = field(default_factory=list)
    wall_s: float = 0.0
    peak_bytes: int = 0
    peak_active: int = 0

    def percentile(self, q: float) -> float:
        lat = sorted(r.latency for r in self.records if r.error is None)
        if not lat:
            return float("nan")
        return float(np.percentile(lat, q))

    @property
    def dof_steps_per_s(self) -> float:
        return sum(r.dof * r.steps for r in self.records if r.error is None) / self.wall_s


def serve(requests: Sequence, precision: int = native.QMS_MIXED, budget_bytes: int = GPU_BYTES,
          max_active: int = 16, tol: float = 1e-12, on_result: Optional[Callable] = None) -> ServeResult:
    """Serve `requests` (objects with .arrival in seconds from the start and
    .design), admitting each, in arrival order, when a slot is free and its
    footprint fits the budget beside the requests in flight.  `on_result(req,
    solver)` runs on the worker before the handle is released (e.g. to read
    the final fields)."""
    res = ServeResult()
    lock = threading.Condition()
    state = {"bytes": 0, "active": 0}
    threads = []
    t_start = time.perf_counter()

    def now():
        return time.perf_counter() - t_start

    def work(req, rec: RequestRecord):
        s = None
        try:
            d = req.design
            s = make_solver(d, precision, tol=tol)
            s.steady(program.steady_velocity(d.inputs.traj, d.inputs.dt))
            i0 = s.stats().total_iters
            s.run(velocities(d, 1, d.inputs.steps))
            rec.iters = int(s.stats().total_iters - i0)
            if on_result is not None:
                on_result(req, s)
        except Exception as e:          # the request fails, the server goes on
            rec.error = str(e)
        finally:
            if s is not None:
                s.close()
            rec.end = now()
            with lock:
                state["bytes"] -= rec.bytes
                state["active"] -= 1
                lock.notify_all()

    for req in sorted(requests, key=lambda r: r.arrival):
        d = req.design
        rec = RequestRecord(d.name, d.dof, d.inputs.steps, footprint(d, precision), req.arrival)
        res.records.append(rec)
        if rec.bytes > budget_bytes:
            rec.error = f"footprint {rec.bytes} B exceeds the device budget {budget_bytes} B"
            rec.start = rec.end = max(now(), rec.arrival)
            continue
        delay = req.arrival - now()
        if delay > 0:
            time.sleep(delay)
        with lock:
            while state["active"] >= max_active or state["bytes"] + rec.bytes > budget_bytes:
                lock.wait()
            state["active"] += 1
            state["bytes"] += rec.bytes
            res.peak_bytes = max(res.peak_bytes, state["bytes"])
            res.peak_active = max(res.peak_active, state["active"])
        rec.start = now()
        t = threading.Thread(target=work, args=(req, rec), daemon=True)
        t.start()
        threads.append(t)
    for t in threads:
        t.join()
    res.wall_s = now()
    return res

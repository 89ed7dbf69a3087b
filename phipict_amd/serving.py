"""Serving: a request loop for PISO + MHD solves of mixed sizes on one GPU.

Requests (2D or 3D MHD ducts of different sizes and physics, each for a number of PISO steps)
arrive over time. A request is admitted when a slot is free and its estimated device-memory
footprint fits in what the HBM budget has left; it then builds its domain, fields and AMG
hierarchies on the device, where they stay resident while it runs, and steps on its own HIP
stream from its own host thread. Requests that arrive while others run are admitted as slots
and memory free up (continuous batching): their kernels overlap on the device, since the
library's solver workspaces belong to the calling thread and solves on different streams run
concurrently. The loop records arrival, admission and completion times, from which throughput
and latency percentiles follow.

A request's result does not depend on what else runs: every kernel's reductions are
deterministic and no buffer is shared between requests, so a request gives the same fields
alone and in the mix (bit for bit when the AMG lanes are chosen by rule rather than timing).
"""
from __future__ import annotations

import heapq
import math
import threading
import time
from dataclasses import dataclass, field
from typing import Callable, Optional, Sequence

import numpy as np
import torch

from . import grid

__all__ = ["DuctRequest", "RequestRecord", "ServingLoop", "build_duct", "estimate_footprint", "poisson_arrivals",
           "latency_percentiles"]


@dataclass
class DuctRequest:
    """One solve request: an MHD duct (3D, periodic in x, insulating walls in y and z) or
    channel (2D, periodic in x, walls in y) with ``cells`` = (nx, ny[, nz]), driven by a
    uniform pressure-gradient source, stepped ``steps`` times."""
    rid: int
    cells: tuple
    re: float
    ha: float
    steps: int = 50
    dt: float = 5e-4
    tol: float = 1e-8
    dtype: torch.dtype = torch.float32
    seed: int = 0
    arrival: float = 0.0  # seconds after the loop starts

    @property
    def dims(self) -> int:
        return len(self.cells)

    @property
    def ncells(self) -> int:
        return int(np.prod(self.cells))


@dataclass
class RequestRecord:
    rid: int
    arrival: float
    admitted: float = math.nan
    finished: float = math.nan
    footprint: int = 0
    error: Optional[str] = None
    fields: Optional[dict] = None  # host copies of velocity, pressure and potential (keep_fields)

    @property
    def latency(self) -> float:
        return self.finished - self.arrival

    @property
    def wait(self) -> float:
        return self.admitted - self.arrival


def build_duct(req: DuctRequest, device: torch.device):
    """The request's domain and MHD simulation: tanh wall refinement 1.5 on [-1, 1], length 4,
    nu = 1/Re, N = Ha^2/Re, source N (core velocity of order one), e_b = (0, 1, 0), seeded
    N(0, 1e-3) initial velocity; AMG-preconditioned CG for pressure and potential at the
    request's absolute tolerance, the pressure's raised to 64 eps ||b|| / sqrt(n) of the
    request's dtype, the potential solved in float64; BiCGStab momentum. A float32 pressure solve
    can stall above 1e-8 (its rounding floor follows the pressure, not the right-hand side, whose
    scale is small in a periodic duct): float32 requests keep the best iterate there, with the
    reference's warning, instead of failing (pressure_return_best_result)."""
    import phipict_amd as pa

    dt = req.dtype
    d = req.dims
    n = list(req.cells)
    axes = [np.linspace(0.0, 4.0, n[0] + 1)]
    for k in range(1, d):
        axes.append(np.array(grid.axis_coords(-1.0, 1.0, grid.make_weights_tanh(n[k], 1.5, "BOTH"))))
    mesh = np.meshgrid(*reversed(axes), indexing="ij")
    v = torch.as_tensor(np.stack(list(reversed(mesh))), dtype=dt).unsqueeze(0)
    gen = torch.Generator("cpu").manual_seed(req.seed)
    vel = (1e-3 * torch.randn(1, d, *reversed(n), generator=gen, dtype=torch.float64)).to(dt)
    N = req.ha ** 2 / req.re
    dom = pa.Domain(d, torch.tensor([1.0 / req.re], dtype=dt), dtype=dt, device=device)
    blk = dom.CreateBlock(velocity=vel, vertexCoordinates=v)
    for f in range(2, 2 * d):
        blk.CloseBoundary(f)
    src = torch.zeros(1, d, dtype=dt, device=device)
    src[0, 0] = N
    blk.setVelocitySource(src)
    dom.PrepareSolve()

    def forcing(domain, **kw):
        domain.getBlocks()[0].setVelocitySource(src)
        domain.UpdateDomainData()

    sim = pa.MHDSimulation(domain=dom, dt=req.dt, stuart_number=torch.tensor(N, dtype=dt),
                           e_b=torch.tensor([0.0, 1.0, 0.0], dtype=dt), non_orthogonal=False,
                           pressure_use_amg=True, potential_use_preconditioner=True,
                           pressure_tol=pa.SolverTolerance(rtol=64.0 * torch.finfo(dt).eps, atol=req.tol),
                           pressure_return_best_result=dt == torch.float32,
                           potential_tol=req.tol, potential_solve_dtype=torch.float64,
                           hooks={"PRE_VELOCITY_SETUP": forcing})
    return dom, sim


def estimate_footprint(req: DuctRequest) -> int:
    """Device bytes a request holds while it runs: fields, the three operators with their ELL
    copies, solver work vectors and the AMG hierarchies (levels, sliced copies, the patterns of
    A P), per cell for the dtype and dimension; a conservative model (measured peaks of the
    duct stay below it), used for admission."""
    s = 4 if req.dtype == torch.float32 else 8
    d = req.dims
    per_cell = (
        (4 * d + 8) * s             # fields, flat solve vectors, right-hand sides
        + 3 * (2 * d + 1) * (2 * s + 8)  # C, P, Epot: CSR values, ELL values and columns
        + 14 * s + 8 * 8            # solver work vectors (momentum, pressure, potential)
        + 2 * (900 if d == 3 else 500)   # two AMG hierarchies: levels, sliced copies, A P patterns
    )
    return int(per_cell * req.ncells + (64 << 20))


def poisson_arrivals(rate: float, count: int, seed: int = 6) -> list:
    """Arrival times of a Poisson process of the given rate (exponential gaps)."""
    gaps = np.random.default_rng(seed).exponential(1.0 / rate, count)
    return list(np.cumsum(gaps) - gaps[0])


def latency_percentiles(records: Sequence[RequestRecord], qs=(50, 99)) -> dict:
    lat = np.array([r.latency for r in records if r.error is None])
    return {f"p{q}": float(np.percentile(lat, q)) for q in qs} if lat.size else {}


@dataclass(order=True)
class _Pending:
    arrival: float
    rid: int
    req: DuctRequest = field(compare=False)


class ServingLoop:
    """Admit requests into at most ``max_slots`` concurrent runs whose estimated footprints sum
    to at most ``hbm_budget`` bytes; run each on its own stream and host thread; requests are
    admitted in arrival order (a request that does not fit yet waits, and so do later ones)."""

    def __init__(self, hbm_budget: int, max_slots: int = 4, device=None, keep_fields: bool = False,
                 builder: Callable = build_duct, footprint: Callable = estimate_footprint):
        self.hbm_budget = int(hbm_budget)
        self.max_slots = int(max_slots)
        self.device = device or torch.device("cuda", 0)
        self.keep_fields = keep_fields
        self.builder = builder
        self.footprint = footprint
        self.in_use = 0
        self.peak_in_use = 0
        self.peak_running = 0
        self._shapes: dict = {}
        self._shape_lock = threading.Lock()

    def _release(self, req: DuctRequest, dom) -> None:
        """Drop the cached AMG interpolations of the request's patterns unless a running
        request of the same grid still shares them: the budget accounts for running requests
        only."""
        from .solvers import amg

        with self._shape_lock:
            self._shapes[req.cells] -= 1
            if self._shapes[req.cells] > 0:
                return
        for mat in (dom.P, dom.Epot):
            if mat is not None:
                amg.release_pattern(mat)

    def _run_one(self, req: DuctRequest, rec: RequestRecord, t0: float, done: Callable) -> None:
        try:
            stream = torch.cuda.Stream(device=self.device)
            with torch.cuda.device(self.device), torch.cuda.stream(stream):
                dom, sim = self.builder(req, self.device)
                for _ in range(req.steps):
                    if not sim.single_step():
                        raise RuntimeError("step failed")
                if self.keep_fields:
                    blk = dom.getBlocks()[0]
                    rec.fields = {"velocity": blk.velocity.detach().cpu(), "pressure": blk.pressure.detach().cpu(),
                                  "potential": dom.epotResult.detach().cpu()}
                stream.synchronize()
            self._release(req, dom)
            del dom, sim
        except Exception as exc:  # recorded, the loop goes on
            rec.error = f"{type(exc).__name__}: {exc}"
        rec.finished = time.perf_counter() - t0
        done(req, rec)

    def run(self, requests: Sequence[DuctRequest]) -> list:
        """Serve the requests (arrival times relative to the start of the call); returns one
        record per request, in request order."""
        records = {r.rid: RequestRecord(r.rid, r.arrival) for r in requests}
        for r in requests:
            if self.footprint(r) > self.hbm_budget:
                records[r.rid].error = "footprint exceeds the HBM budget"
        queue = [_Pending(r.arrival, r.rid, r) for r in requests if records[r.rid].error is None]
        heapq.heapify(queue)
        lock = threading.Condition()
        running = {}
        t0 = time.perf_counter()

        def done(req, rec):
            with lock:
                self.in_use -= rec.footprint
                running.pop(req.rid, None)
                lock.notify_all()

        with lock:
            while queue or running:
                now = time.perf_counter() - t0
                # admit in arrival order while the head fits
                while queue and queue[0].arrival <= now and len(running) < self.max_slots:
                    need = self.footprint(queue[0].req)
                    if self.in_use + need > self.hbm_budget:
                        break
                    item = heapq.heappop(queue)
                    rec = records[item.rid]
                    rec.admitted, rec.footprint = now, need
                    self.in_use += need
                    with self._shape_lock:
                        self._shapes[item.req.cells] = self._shapes.get(item.req.cells, 0) + 1
                    self.peak_in_use = max(self.peak_in_use, self.in_use)
                    th = threading.Thread(target=self._run_one, args=(item.req, rec, t0, done), daemon=True)
                    running[item.rid] = th
                    self.peak_running = max(self.peak_running, len(running))
                    th.start()
                wait = None
                if queue and queue[0].arrival > now:
                    wait = queue[0].arrival - now
                lock.wait(timeout=wait if wait is not None else 0.5)
        return [records[r.rid] for r in requests]

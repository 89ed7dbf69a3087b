#!/usr/bin/env python3
"""PISO + inductionless-MHD step throughput on one MI355X.

Workload (default): the 3D rectangular MHD duct, one block of 192 x 96 x 96 cells (x streamwise,
periodic, driven by a uniform pressure-gradient source; y and z wall-refined tanh grids with
factor 1.5 on [-1, 1]; insulating no-slip walls), Re = 500, Ha = 100 (N = Ha^2/Re = 20),
e_b = (0, 1, 0), dt = 5e-4, two pressure correctors, BiCGStab momentum solves, AMG-preconditioned
CG pressure and potential solves to the absolute tolerance 1e-8; initial velocity N(0, 1e-3)
(seed 1). A step is one full PISO + magnetic step of the library (``MHDSimulation.single_step``)
with every field resident in HBM.

``--skew F`` jitters the duct's interior vertices by F times the smallest spacing (seeded) and
steps in the reference's default non-orthogonal mode (corner terms in the advection and pressure
matrices and on their right-hand sides); the CPU baseline is skipped then.
``--workload tjunction-adjoint`` times a checkpointed adjoint rollout instead.
``--workload uneven-batch`` times 64 independent 2D Hartmann channels of uneven sizes
(Nx, Ny log-uniform in [32, 256], Ha ~ U[5, 50], Re ~ U[50, 500], stretching ~ U[1, 2]; inflow
U = 1 with an advected outflow, insulating walls, dt = 1e-3, CG pressure and potential and
BiCGStab momentum to the absolute tolerance 1e-10) stepped as one domain of independent blocks
with segmented solves, against looping the plain path over the 64 instances one at a time.
``--workload serving`` serves Poisson arrivals of 2D and 3D MHD duct requests (cells
log-uniform in [16k, 2M], Ha ~ U[5, 200], Re ~ U[50, 1000], 50 PISO steps each, AMG-CG to 1e-8 in
FP32) admitted against an HBM budget into concurrent slots, and reports throughput and p50/p99
latency, against serving the same requests one at a time.

Printed: one JSON line with steps/s of the uninstrumented timed steps, the roofline of the
kernel with the largest share of the kernel time (algorithmic bytes per launch / average launch
time, from HIP events recorded on the kernel's own stream for a pseudo-random quarter of the
launches during as many further steps, so that the timer's event records stay out of the
throughput; kernels with the same work on every launch only) and the oracle's CPU time for a
bounded sample of the same physics.
"""
from __future__ import annotations

import argparse
import gc
import json
import math
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
MARK_BYTES = 256  # k_copy_bytes launches that bracket the timed steps (one workgroup)
RE, HA = 500.0, 100.0
DT = 5e-4
TOL = 1e-8
STRETCH = 1.5
LENGTH = 4.0


def parse_args(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", choices=["duct", "hartmann-c1", "tjunction-adjoint", "uneven-batch", "serving"],
                    default="duct",
                    help="duct: the PISO + MHD step throughput (default); tjunction-adjoint: the checkpointed "
                         "backward through a 3-block T-junction rollout and its finite-difference check")
    ap.add_argument("--dtype", choices=["f64", "f32"], default="f64")
    ap.add_argument("--size", default="192x96x96", help="cells nx x ny x nz")
    ap.add_argument("--inflow", action="store_true",
                    help="uniform inflow and advected outflow instead of the periodic, source-driven duct")
    ap.add_argument("--skew", type=float, default=0.0,
                    help="jitter the duct's interior vertices by this fraction of the smallest spacing and step in "
                         "the non-orthogonal mode (corner terms; CPU baseline off)")
    ap.add_argument("--timer-every", type=int, default=4, help="time a pseudo-random 1/k of the launches of each kernel")
    ap.add_argument("--timer-in-timed-region", action="store_true",
                    help="run the kernel timer during the timed steps (default: during as many further steps)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=2, help="timed oracle steps of the CPU baseline")
    ap.add_argument("--details", action="store_true", help="also print the per-kernel table to stderr")
    ap.add_argument("--host-profile", action="store_true",
                    help="profile the host (Python) side of the timed steps with cProfile; the result goes to stderr")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="before the warm-up, copy this many MiB three times with k_copy_bytes (the reference "
                         "point for memory counters)")
    ap.add_argument("--traffic-summary", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="no GPU run: summarize two rocprofv3 counter_collection.csv files (FETCH_SIZE, WRITE_SIZE) "
                         "of runs made with --copy-probe-mib into a markdown table on stdout")
    ap.add_argument("--trace-summary", metavar="KERNEL_TRACE_CSV",
                    help="no GPU run: per-kernel table of the timed steps (between the two marker launches) of a "
                         "rocprofv3 --kernel-trace csv of this script, as markdown on stdout")
    ap.add_argument("--traffic-file", default=None,
                    help="measured HBM bytes per launch (a --traffic-summary output) for roofline.traffic; default: "
                         "profiles/hbm_traffic_duct<size>_<dtype>.md when present")
    return ap.parse_args(argv)


def duct_vertices(nx, ny, nz):
    import numpy as np

    from phipict_amd import grid

    xs = np.linspace(0.0, LENGTH, nx + 1)
    ys = np.array(grid.axis_coords(-1.0, 1.0, grid.make_weights_tanh(ny, STRETCH, "BOTH")))
    zs = np.array(grid.axis_coords(-1.0, 1.0, grid.make_weights_tanh(nz, STRETCH, "BOTH")))
    Z, Y, X = np.meshgrid(zs, ys, xs, indexing="ij")
    return np.stack([X, Y, Z])


def make_sim(nx, ny, nz, dtype, device, skew=0.0, inflow=False):
    """The duct: periodic in x and driven by a pressure-gradient source, or (inflow) with a
    uniform inflow U = 1 at -x and an advected outflow at +x, both with an open potential
    (the pressure then carries the flow's physical pressure drop and its solve iterates)."""
    import numpy as np
    import torch

    import phipict_amd as pa

    if inflow:
        return make_inflow_sim(nx, ny, nz, dtype, device)
    N = HA ** 2 / RE
    G = N  # core velocity of order one
    v = duct_vertices(nx, ny, nz)
    if skew > 0:
        # seeded jitter of every vertex off the block faces: a non-orthogonal grid
        rng = np.random.default_rng(3)
        spacing = min(float(np.min(np.abs(np.diff(v[a], axis=2 - a)))) for a in range(3))
        v[:, 1:-1, 1:-1, 1:-1] += skew * spacing * rng.uniform(-1, 1, v[:, 1:-1, 1:-1, 1:-1].shape)
    gen = torch.Generator("cpu").manual_seed(1)
    vel = 1e-3 * torch.randn(1, 3, nz, ny, nx, generator=gen, dtype=torch.float64)
    dom = pa.Domain(3, torch.tensor([1.0 / RE], dtype=dtype), dtype=dtype, device=device)
    blk = dom.CreateBlock(velocity=vel.to(dtype), vertexCoordinates=torch.as_tensor(v, dtype=dtype).unsqueeze(0))
    for f in (2, 3, 4, 5):
        blk.CloseBoundary(f)
    source = torch.tensor([[G, 0.0, 0.0]], dtype=dtype, device=device)
    blk.setVelocitySource(source)
    dom.PrepareSolve()

    def forcing(domain, **kw):
        domain.getBlocks()[0].setVelocitySource(source)
        domain.UpdateDomainData()

    sim = pa.MHDSimulation(domain=dom, dt=DT, stuart_number=torch.tensor(N),
                           e_b=torch.tensor([0.0, 1.0, 0.0], dtype=dtype), non_orthogonal=skew > 0,
                           pressure_use_amg=True, potential_use_preconditioner=True, pressure_tol=TOL,
                           potential_tol=TOL, hooks={"PRE_VELOCITY_SETUP": forcing})
    sim.record_solver_stats = True
    return dom, sim


def make_inflow_sim(nx, ny, nz, dtype, device, U=1.0):
    import torch

    import phipict_amd as pa

    N = HA ** 2 / RE
    v = duct_vertices(nx, ny, nz)
    gen = torch.Generator("cpu").manual_seed(1)
    vel = 1e-3 * torch.randn(1, 3, nz, ny, nx, generator=gen, dtype=torch.float64)
    vel[:, 0] += U
    dom = pa.Domain(3, torch.tensor([1.0 / RE], dtype=dtype), dtype=dtype, device=device)
    blk = dom.CreateBlock(velocity=vel.to(dtype), vertexCoordinates=torch.as_tensor(v, dtype=dtype).unsqueeze(0))
    plug = torch.tensor([U, 0.0, 0.0], dtype=dtype, device=device)
    inlet = blk.CloseBoundary(0, velocity=plug.view(1, 3))
    inlet.setPotentialBC(pa.PotentialBC.OPEN)
    out = blk.CloseBoundary(1, velocity=torch.ones(1, 3, nz, ny, 1, dtype=dtype, device=device) * plug.view(1, 3, 1, 1, 1))
    out.setPotentialBC(pa.PotentialBC.OPEN)
    for f in (2, 3, 4, 5):
        blk.CloseBoundary(f)
    dom.PrepareSolve()
    pa.simulation.balance_boundary_fluxes(dom, [out])
    dom.UpdateDomainData()
    velm = torch.tensor([[U, 0.0, 0.0]], dtype=dtype, device=device)

    def advect(domain, **kw):
        pa.simulation.update_advective_boundaries(domain, [out], velm, DT)

    def no_source(domain, **kw):
        domain.getBlocks()[0].clearVelocitySource()
        domain.UpdateDomainData()

    sim = pa.MHDSimulation(domain=dom, dt=DT, stuart_number=torch.tensor(N),
                           e_b=torch.tensor([0.0, 1.0, 0.0], dtype=dtype), non_orthogonal=False,
                           pressure_use_amg=True, potential_use_preconditioner=True, pressure_tol=TOL,
                           potential_tol=TOL, hooks={"PRE": advect, "PRE_VELOCITY_SETUP": no_source})
    sim.record_solver_stats = True
    return dom, sim


# ----------------------------------------------------------------------------- adjoint workload
TJ_RE, TJ_HA, TJ_DT = 200.0, 20.0, 5e-3


def make_tjunction(n, device, B0, inflow, seed=4):
    """3-block T-junction duct: main arm 2n x n x n (x in [0, 2], inflow at -x), junction cube n^3
    (x in [2, 3], outlet at -y) and branch cube n^3 (y in [1, 2], outlet at +y), insulating
    no-slip walls, Re = 200, Ha = 20 (N = 2), e_b = B0 (0, 1, 0); outlets advected and
    flux-balanced every step. ``inflow``: the inflow face velocity [1, 3, n, n, 1]."""
    import numpy as np
    import torch

    import phipict_amd as pa

    dt = torch.float64

    def verts(x0, x1, nx, y0, y1):
        xs, ys, zs = np.linspace(x0, x1, nx + 1), np.linspace(y0, y1, n + 1), np.linspace(0.0, 1.0, n + 1)
        Z, Y, X = np.meshgrid(zs, ys, xs, indexing="ij")
        return torch.as_tensor(np.stack([X, Y, Z]), dtype=dt).unsqueeze(0)

    gen = torch.Generator("cpu").manual_seed(seed)
    dom = pa.Domain(3, torch.tensor([1.0 / TJ_RE], dtype=dt), dtype=dt, device=device)
    shapes = [(2 * n, 0.0, 2.0, 0.0, 1.0), (n, 2.0, 3.0, 0.0, 1.0), (n, 2.0, 3.0, 1.0, 2.0)]
    blocks = []
    for nx, x0, x1, y0, y1 in shapes:
        vel = 1e-2 * torch.randn(1, 3, n, n, nx, generator=gen, dtype=dt)
        blocks.append(dom.CreateBlock(velocity=vel, vertexCoordinates=verts(x0, x1, nx, y0, y1)))
    b0, b1, b2 = blocks
    b0.ConnectBlock(1, b1, 0)
    b1.ConnectBlock(3, b2, 2)
    for blk, faces in ((b0, (0, 2, 3, 4, 5)), (b1, (1, 2, 4, 5)), (b2, (0, 1, 3, 4, 5))):
        for f in faces:
            blk.CloseBoundary(f)
    inlet = b0.getBoundary(0)
    inlet.setVelocity(inflow)
    inlet.setPotentialBC(pa.PotentialBC.OPEN)
    flux = float(inflow[0, 0].detach().sum()) / (n * n)
    outs = []
    for blk, f, sign in ((b1, 2, -1.0), (b2, 3, 1.0)):
        bd = blk.getBoundary(f)
        v = torch.zeros(1, 3, n, 1, n, dtype=dt, device=device)
        # 45% of the inflow each: the balance below rescales them, a smooth function of the inflow
        v[0, 1] = 0.45 * sign * flux
        bd.setVelocity(v)
        bd.setPotentialBC(pa.PotentialBC.OPEN)
        outs.append(bd)
    dom.PrepareSolve()
    pa.simulation.balance_boundary_fluxes(dom, outs)
    velm = torch.tensor([[0.0, 0.5, 0.0]], dtype=dt, device=device)

    def advect(domain, **kw):
        pa.simulation.update_advective_boundaries(domain, outs, velm, TJ_DT)

    def no_source(domain, **kw):
        for b in domain.getBlocks():
            b.clearVelocitySource()
        domain.UpdateDomainData()

    tol = pa.SolverTolerance(rtol=1e-12)
    e_b = B0 * torch.tensor([0.0, 1.0, 0.0], dtype=dt)
    sim = pa.MHDSimulation(domain=dom, dt=TJ_DT, stuart_number=torch.tensor(TJ_HA ** 2 / TJ_RE, dtype=dt), e_b=e_b,
                           non_orthogonal=False, advection_tol=tol, pressure_tol=tol, potential_tol=tol,
                           hooks={"PRE": advect, "PRE_VELOCITY_SETUP": no_source}, flux_balance_tol=1e-6)
    return dom, sim


def tjunction_adjoint(args) -> int:
    """Gradients of the outlet MSE after ``steps`` steps w.r.t. the inflow field and B0 by the
    checkpointed backward; timed, and checked against central differences of the forward."""
    import numpy as np
    import torch

    from phipict_amd import adjoint as adj
    from phipict_amd.simulation import _face_slice

    device = torch.device("cuda", 0)
    n, steps, every = 32, args.steps if args.steps != 10 else 20, 5
    rng = np.random.default_rng(4)
    inflow = torch.zeros(1, 3, n, n, 1, dtype=torch.float64, device=device)
    inflow[0, 0] = torch.as_tensor(rng.normal(1.0, 0.1, (n, n, 1)), device=device)
    target = torch.as_tensor(np.random.default_rng(5).standard_normal((1, 3, n, 1, n)), device=device)

    def loss_fn(domain):
        return ((_face_slice(domain.getBlocks()[2], 3) - target) ** 2).mean()

    vin_leaf = [None]

    def rollout(b0, vin, grad):
        B0 = torch.tensor(b0, dtype=torch.float64, requires_grad=grad)
        v = vin.clone().requires_grad_(grad)
        vin_leaf[0] = v
        dom, sim = make_tjunction(n, device, B0, v)
        if grad:
            return adj.checkpointed_gradient(sim, steps, every, loss_fn), B0
        for _ in range(steps):
            if not sim.single_step():
                raise RuntimeError("forward step failed")
        return float(loss_fn(dom)), B0

    print("[bench] T-junction adjoint warm-up", file=sys.stderr, flush=True)
    rollout(1.0, inflow, True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    (loss, g0), B0 = rollout(1.0, inflow, True)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    g_in, g_b0 = vin_leaf[0].grad, float(B0.grad)
    print(f"[bench] forward + checkpointed backward of {steps} steps: {elapsed:.3f} s", file=sys.stderr, flush=True)
    # central differences: B0, and a random inflow direction (its net flux change rescales the
    # outlets, through the balance before and during the rollout)
    d = torch.as_tensor(np.random.default_rng(12).standard_normal((n, n, 1)), device=device)
    d_in = torch.zeros_like(inflow)
    d_in[0, 0] = d
    eps, eps_in = 1e-6, 1e-6
    t1 = time.perf_counter()
    fd_b0 = (rollout(1.0 + eps, inflow, False)[0] - rollout(1.0 - eps, inflow, False)[0]) / (2 * eps)
    fd_in = (rollout(1.0, inflow + eps_in * d_in, False)[0]
             - rollout(1.0, inflow - eps_in * d_in, False)[0]) / (2 * eps_in)
    fwd_time = (time.perf_counter() - t1) / 4
    an_in = float((g_in * d_in).sum())
    line = {
        "metric": "adjoint PISO steps/s (forward + checkpointed backward)",
        "value": round(steps / elapsed, 4),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": 1,
        "ms_per_step": round(1e3 * elapsed / steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"3-block MHD T-junction 2x{n}^3 + 2x{n}^3 cells ({4 * n ** 3} cells), Re=200, Ha=20, "
                               f"dt=5e-3, {steps} steps, checkpoint every {every}, CG/BiCGStab rtol 1e-12, "
                               "loss = outlet MSE vs a seeded target; gradients w.r.t. the inflow field and B0",
                   "cells": 4 * n ** 3},
        "forward_only_s": round(fwd_time, 3),
        "loss": loss,
        "gradcheck": {"b0": {"adjoint": g_b0, "fd": fd_b0, "rel_err": abs(g_b0 - fd_b0) / max(abs(fd_b0), 1e-300)},
                      "inflow": {"adjoint": an_in, "fd": fd_in,
                                 "rel_err": abs(an_in - fd_in) / max(abs(fd_in), 1e-300), "eps": eps_in},
                      "eps": eps},
    }
    print(json.dumps(line), flush=True)
    return 0


# ----------------------------------------------------------------------------- C1 in full
C1_STEPS, C1_DT, C1_TOL, C1_RE = 200, 1e-3, 1e-10, 100.0


def hartmann_c1(args) -> int:
    """The 2D Hartmann channel of the CPU-reference workload as stated: 64 x 64 cells, tanh
    stretching 1.5 toward the walls, Re = 100, Ha = 10, e_b = (0, 1, 0), uniform inflow U = 1 and
    an advected (zero-gradient) outflow, insulating no-slip walls, initial N(0, 1e-3) noise,
    dt = 1e-3, 200 steps with two correctors, BiCGStab momentum and CG pressure / potential to
    the absolute tolerance 1e-10, FP64. The whole 200 steps on the GPU, and in full in the
    oracle (one thread and the host's threads; the faster is the baseline), with the relative
    L2 differences of the final fields."""
    import numpy as np
    import torch

    import phipict_amd as pa

    sys.path.insert(0, str(ROOT / "tests"))
    import cases
    from oracle import oracle as O
    from oracle.piso_oracle import MHDSpec, OraclePISO, balance_boundary_fluxes, update_advective_boundaries

    device = torch.device("cuda", 0)
    specs = cases.hartmann_c1()
    N = (10.0 ** 2) / C1_RE

    def gpu_run(steps):
        dom = cases.build_domain(specs, 1.0 / C1_RE, C1_DT, torch.float64)
        outs = [dom.getBlocks()[0].getBoundary(1)]
        velm = torch.tensor([[1.0, 0.0]], dtype=torch.float64, device=device)

        def advect(domain, **kw):
            pa.simulation.update_advective_boundaries(domain, outs, velm, C1_DT)

        def no_source(domain, **kw):
            domain.getBlocks()[0].clearVelocitySource()
            domain.UpdateDomainData()

        sim = pa.MHDSimulation(domain=dom, dt=C1_DT, stuart_number=torch.tensor(N),
                               e_b=torch.tensor([0.0, 1.0, 0.0], dtype=torch.float64), non_orthogonal=False,
                               advection_tol=C1_TOL, pressure_tol=C1_TOL, potential_tol=C1_TOL,
                               hooks={"PRE": advect, "PRE_VELOCITY_SETUP": no_source})
        pa.simulation.balance_boundary_fluxes(dom, outs)
        dom.UpdateDomainData()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for i in range(steps):
            if not sim.single_step():
                raise RuntimeError(f"C1 step {i} failed")
        torch.cuda.synchronize()
        return dom, time.perf_counter() - t0

    gpu_run(args.warmup)
    print(f"[bench] C1: {C1_STEPS} GPU steps", file=sys.stderr, flush=True)
    dom, gpu_s = gpu_run(C1_STEPS)

    def oracle_run(threaded):
        O.use_threads(threaded)
        try:
            od = O.OracleDomain(specs, 1.0 / C1_RE, C1_DT, np.float64)
            osim = OraclePISO(od, C1_DT, advection_tol=C1_TOL, pressure_tol=C1_TOL,
                              mhd=MHDSpec(N, (0.0, 1.0, 0.0), potential_tol=C1_TOL),
                              pre_velocity_setup=lambda o: (update_advective_boundaries(o.od, [(0, 1)], np.array([1.0, 0.0]),
                                                                                        C1_DT), o.od.set_source(0, None)))
            balance_boundary_fluxes(od, [(0, 1)])
            t0 = time.perf_counter()
            for _ in range(C1_STEPS):
                osim.step()
            return od, osim, time.perf_counter() - t0
        finally:
            O.use_threads(False)

    threads = cpu_threads()
    os.environ["OMP_NUM_THREADS"] = str(threads)
    print(f"[bench] C1: {C1_STEPS} oracle steps, 1 and {threads} threads", file=sys.stderr, flush=True)
    od, osim, cpu1 = oracle_run(False)
    _, _, cpun = oracle_run(True)
    blk = dom.getBlocks()[0]
    errs = {"velocity": cases.rel_l2(cases.to_np(blk.velocity)[0], od.velocity[0]),
            "pressure": cases.rel_l2(cases.to_np(blk.pressure)[0], od.pressure[0]),
            "potential": cases.rel_l2(cases.to_np(dom.epotResult), osim.epot_result)}
    best_cores, best_s = (1, cpu1) if cpu1 <= cpun else (threads, cpun)
    cells = int(dom.getTotalSize())
    line = {
        "metric": "PISO steps/s",
        "value": round(C1_STEPS / gpu_s, 2),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": C1_STEPS,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * gpu_s / C1_STEPS, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "2D Hartmann MHD channel 64x64 cells, tanh 1.5 in y, Re=100, Ha=10, inflow U=1 + "
                               "advected outflow, dt=1e-3, 200 steps, 2 correctors, BiCGStab momentum + CG pressure "
                               "and potential to abs tol 1e-10", "cells": cells},
        "cell_updates_per_s": round(cells * C1_STEPS / gpu_s, 1),
        "rel_l2_vs_oracle": {k: float(f"{v:.3e}") for k, v in errs.items()},
        "cpu_baseline": {"value": round(C1_STEPS / best_s, 3), "unit": "steps/s", "cores": best_cores, "kind": "port",
                         "sample": (f"the oracle on the same case, all {C1_STEPS} steps: {cpu1:.2f} s on 1 thread, "
                                    f"{cpun:.2f} s on {threads} threads (OpenMP build)")},
    }
    print(json.dumps(line), flush=True)
    return 0


# ----------------------------------------------------------------------------- uneven batch
UB_COUNT, UB_DT, UB_TOL = 64, 1e-3, 1e-8


def uneven_instances(count=UB_COUNT):
    """(nx, ny, stretch, Re, Ha) of the uneven-batch workload: sizes log-uniform in [32, 256]
    (seed 2), Ha ~ U[5, 50], Re ~ U[50, 500], stretching ~ U[1, 2] (seed 3)."""
    import numpy as np

    rs = np.random.default_rng(2)
    sizes = np.rint(np.exp(rs.uniform(np.log(32), np.log(256), (count, 2)))).astype(int)
    rp = np.random.default_rng(3)
    ha, re, st = rp.uniform(5, 50, count), rp.uniform(50, 500, count), rp.uniform(1, 2, count)
    return [(int(sizes[i, 0]), int(sizes[i, 1]), float(st[i]), float(re[i]), float(ha[i])) for i in range(count)]


def make_uneven(instances, dtype, device, independent):
    """One domain holding the instances as independent blocks (independent=True), or one domain
    per instance (a list of (domain, simulation) pairs)."""
    import numpy as np
    import torch

    import phipict_amd as pa
    from phipict_amd import grid

    def add_block(dom, i, nx, ny, stretch, re):
        xs = np.linspace(0.0, 4.0, nx + 1)
        ys = np.array(grid.axis_coords(-1.0, 1.0, grid.make_weights_tanh(ny, stretch, "BOTH")))
        Y, X = np.meshgrid(ys, xs, indexing="ij")
        v = torch.as_tensor(np.stack([X, Y]), dtype=dtype).unsqueeze(0)
        gen = torch.Generator("cpu").manual_seed(100 + i)
        vel = (1e-3 * torch.randn(1, 2, ny, nx, generator=gen, dtype=torch.float64)).to(dtype)
        blk = dom.CreateBlock(velocity=vel, vertexCoordinates=v)
        blk.setViscosity(torch.tensor([1.0 / re], dtype=dtype, device=device))
        inlet = blk.CloseBoundary(0, velocity=torch.tensor([[1.0, 0.0]], dtype=dtype, device=device))
        inlet.setPotentialBC(pa.PotentialBC.OPEN)
        out = blk.CloseBoundary(1, velocity=torch.ones(1, 2, ny, 1, dtype=dtype, device=device)
                                * torch.tensor([1.0, 0.0], dtype=dtype, device=device).view(1, 2, 1, 1))
        out.setPotentialBC(pa.PotentialBC.OPEN)
        blk.CloseBoundary(2)
        blk.CloseBoundary(3)
        return out

    def finish(dom, outs, stuart):
        dom.PrepareSolve()
        pa.simulation.balance_boundary_fluxes(dom, outs)
        dom.UpdateDomainData()
        velm = torch.tensor([[1.0, 0.0]], dtype=dtype, device=device)

        def advect(domain, **kw):
            pa.simulation.update_advective_boundaries(domain, outs, velm, UB_DT)

        def no_source(domain, **kw):
            for b in domain.getBlocks():
                b.clearVelocitySource()
            domain.UpdateDomainData()

        return pa.MHDSimulation(domain=dom, dt=UB_DT, stuart_number=torch.tensor(stuart, dtype=dtype),
                                e_b=torch.tensor([0.0, 1.0, 0.0], dtype=dtype), non_orthogonal=False,
                                advection_tol=UB_TOL, pressure_tol=UB_TOL, potential_tol=UB_TOL,
                                hooks={"PRE": advect, "PRE_VELOCITY_SETUP": no_source})

    if independent:
        dom = pa.Domain(2, torch.tensor([0.01], dtype=dtype), dtype=dtype, device=device)
        outs = [add_block(dom, i, nx, ny, st, re) for i, (nx, ny, st, re, _ha) in enumerate(instances)]
        dom.setIndependentBlocks()
        sim = finish(dom, outs, [ha * ha / re for (_nx, _ny, _st, re, ha) in instances])
        return dom, sim
    pairs = []
    for i, (nx, ny, st, re, ha) in enumerate(instances):
        dom = pa.Domain(2, torch.tensor([0.01], dtype=dtype), dtype=dtype, device=device)
        out = add_block(dom, i, nx, ny, st, re)
        pairs.append((dom, finish(dom, [out], [ha * ha / re])))
    return pairs


def uneven_batch(args) -> int:
    import torch

    device = torch.device("cuda", 0)
    dtype = torch.float64 if args.dtype == "f64" else torch.float32
    inst = uneven_instances()
    cells = sum(nx * ny for nx, ny, *_ in inst)
    print(f"[bench] uneven batch: {len(inst)} instances, {cells} cells", file=sys.stderr, flush=True)
    import phipict_amd as pa

    runs = {}
    for graphs in (False, True):
        # the segmented solves' iteration chunks replayed from captured HIP graphs (useGraphs),
        # and launched directly for comparison
        pa.SetKrylovSettings(useGraphs=graphs)
        try:
            dom, sim = make_uneven(inst, dtype, device, True)
            for _ in range(args.warmup):
                if not sim.single_step():
                    raise RuntimeError("batched step failed")
            torch.cuda.synchronize()
            gc.collect()
            t0 = time.perf_counter()
            for _ in range(args.steps):
                if not sim.single_step():
                    raise RuntimeError("batched step failed")
            torch.cuda.synchronize()
            runs[graphs] = time.perf_counter() - t0
            tile = sim.__dict__.get("segment_tile_rows")
            del dom, sim
        finally:
            pa.SetKrylovSettings()
        print(f"[bench] batched (graphs {graphs}): {args.steps} steps in {runs[graphs]:.3f} s, tile rows {tile}",
              file=sys.stderr, flush=True)
    batched = runs[True]
    pairs = make_uneven(inst, dtype, device, False)
    for _ in range(max(1, args.warmup // 2)):
        for _d, s_ in pairs:
            s_.single_step()
    torch.cuda.synchronize()
    loop_steps = max(1, args.steps // 4)
    t1 = time.perf_counter()
    for _ in range(loop_steps):
        for _d, s_ in pairs:
            if not s_.single_step():
                raise RuntimeError("solo step failed")
    torch.cuda.synchronize()
    looped = time.perf_counter() - t1
    print(f"[bench] looped: {loop_steps} sweeps over {len(pairs)} instances in {looped:.3f} s", file=sys.stderr,
          flush=True)
    env_batched = len(inst) * args.steps / batched
    env_looped = len(inst) * loop_steps / looped
    line = {
        "metric": "uneven-batch env steps/s (batched, segmented solves)",
        "value": round(env_batched, 2),
        "unit": "env steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * batched / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.dtype,
        "data": "synthetic",
        "config": {"workload": f"{len(inst)} independent 2D Hartmann channels, Nx,Ny log-uniform [32,256], "
                               "Ha~U[5,50], Re~U[50,500], stretch~U[1,2], inflow + advected outflow, dt=1e-3, "
                               "CG pressure/potential + BiCGStab momentum, abs tol 1e-8, segmented solves from "
                               "captured HIP graphs",
                   "cells": cells},
        "graphs": True,
        "direct_launch_env_steps_per_s": round(len(inst) * args.steps / runs[False], 2),
        "segment_tile_rows": tile,
        "looped_env_steps_per_s": round(env_looped, 2),
        "batched_over_looped": round(env_batched / env_looped, 2),
        "cell_updates_per_s": round(cells * args.steps / batched, 1),
    }
    print(json.dumps(line), flush=True)
    return 0


# ----------------------------------------------------------------------------- serving
SV_COUNT, SV_RATE, SV_SLOTS, SV_BUDGET_GIB = 16, 2.0, 4, 96


def serving_requests(count=SV_COUNT, rate=SV_RATE, steps=50):
    """Mixed 2D / 3D duct requests: cells log-uniform in [16k, 2M] (seed 6), 3D ducts with
    nx = 2 ny = 2 nz, 2D channels with nx = 2 ny; Ha ~ U[5, 200], Re ~ U[50, 1000] (seed 7);
    Poisson arrivals of the given rate (seed 8)."""
    import numpy as np
    import torch

    from phipict_amd.serving import DuctRequest, poisson_arrivals

    rg = np.random.default_rng(6)
    cells = np.exp(rg.uniform(np.log(16e3), np.log(2e6), count))
    three = rg.random(count) < 0.5
    rp = np.random.default_rng(7)
    ha, re = rp.uniform(5, 200, count), rp.uniform(50, 1000, count)
    arrivals = poisson_arrivals(rate, count, seed=8)
    reqs = []
    for i in range(count):
        if three[i]:
            n = max(8, int(round((cells[i] / 2) ** (1 / 3))))
            shape = (2 * n, n, n)
        else:
            n = max(16, int(round((cells[i] / 2) ** 0.5)))
            shape = (2 * n, n)
        reqs.append(DuctRequest(rid=i, cells=shape, re=float(re[i]), ha=float(ha[i]), steps=steps, dt=5e-4, tol=1e-8,
                                dtype=torch.float32, seed=30 + i, arrival=float(arrivals[i])))
    return reqs


def serving_bench(args) -> int:
    from phipict_amd import serving

    reqs = serving_requests(steps=args.steps if args.steps != 10 else 50)
    cells = sum(r.ncells for r in reqs)
    budget = SV_BUDGET_GIB << 30
    print(f"[bench] serving {len(reqs)} requests ({cells} cells in total), {SV_SLOTS} slots, budget {SV_BUDGET_GIB} GiB",
          file=sys.stderr, flush=True)
    # warm-up: one small request of each dimension (library loading, first compilations)
    serving.ServingLoop(budget, SV_SLOTS).run([serving.DuctRequest(**{**reqs[0].__dict__, "cells": c, "steps": 2,
                                                                      "arrival": 0.0, "rid": 900 + k})
                                               for k, c in enumerate([(64, 32), (16, 8, 8)])])
    loop = serving.ServingLoop(budget, SV_SLOTS)
    t0 = time.perf_counter()
    recs = loop.run(reqs)
    served = time.perf_counter() - t0
    errors = [r.error for r in recs if r.error]
    print(f"[bench] served in {served:.2f} s, errors: {errors}", file=sys.stderr, flush=True)
    # the same requests one at a time, all available at once: the plain path looped
    one = serving.ServingLoop(budget, 1)
    t1 = time.perf_counter()
    solo = one.run([serving.DuctRequest(**{**r.__dict__, "arrival": 0.0}) for r in reqs])
    looped = time.perf_counter() - t1
    print(f"[bench] one at a time in {looped:.2f} s", file=sys.stderr, flush=True)
    lat = serving.latency_percentiles(recs)
    steps = sum(r.steps for r in reqs)
    line = {
        "metric": "served MHD solve requests: throughput and p50/p99 latency",
        "value": round(len(reqs) / served, 4),
        "unit": "requests/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": 2,
        "ms_per_step": None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{len(reqs)} Poisson arrivals (rate {SV_RATE}/s) of 2D/3D MHD duct requests, cells "
                               "log-uniform [16k, 2M], Ha~U[5,200], Re~U[50,1000], 50 PISO steps each, AMG-CG tol 1e-8 "
                               f"FP32 (potential solved in FP64), {SV_SLOTS} slots, HBM budget {SV_BUDGET_GIB} GiB",
                   "cells": cells},
        "latency_s": {k: round(v, 3) for k, v in lat.items()},
        "cell_updates_per_s": round(sum(r.ncells * r.steps for r in reqs) / served, 1),
        "makespan_s": round(served, 3),
        "one_at_a_time_s": round(looped, 3),
        "served_over_one_at_a_time": round(looped / served, 3),
        "peak_running": loop.peak_running,
        "peak_budget_in_use_gib": round(loop.peak_in_use / 2 ** 30, 2),
        "errors": errors + [r.error for r in solo if r.error],
    }
    print(json.dumps(line), flush=True)
    return 0


# ----------------------------------------------------------------------------- algorithmic bytes
def kernel_bytes(name, rows, s, sh, fine, levels, m=3, diag=False):
    """Algorithmic HBM bytes of one launch (neighbour reads counted once): ``fine`` = (n, nnz)
    of the system matrices, ``levels`` maps level rows -> (nnz_A, n_coarse, nnz_R, nnz_P).
    s / sh: bytes per value of the system / the AMG hierarchy; m: right-hand sides of the
    BiCGStab launches (the momentum solve: one per velocity component, sharing the matrix);
    diag: the grid's metrics are diagonal, so the assembly needs d + 1 metric values per cell
    (the diagonal of Minv and det; 2d + 1 where M is read too) instead of the full matrices."""
    n, nnz = fine
    d = 3
    minv = (d + 1) if diag else (d * d + 1)      # Minv rows and det
    full = (2 * d + 1) if diag else (2 * d * d + 1)  # M, Minv and det

    def M(nz, rws, val):
        return nz * (val + 4) + (rws + 1) * 4

    if name in ("cg_direction", "pcg_direction"):
        return M(nnz, n, s) + 4 * n * s
    if name in ("cg_update", "pcg_update"):
        return 6 * n * s
    if name == "pcg_dots":
        return 2 * n * s
    if name == "bicg_p":
        return 4 * m * n * s
    if name in ("bicg_v", "bicg_t"):
        return M(nnz, n, s) + 3 * m * n * s
    if name == "bicg_update":
        return 8 * m * n * s
    if name in ("mean",):
        return n * s
    if name == "subtract_mean":
        return 2 * n * s
    lv = levels.get(rows)
    if name == "amg_presmooth" and lv:
        vb = s if rows == n else sh
        return M(lv[0], rows, sh) + rows * (sh + vb + 2 * sh)
    if name in ("amg_jacobi", "amg_post_jacobi") and lv:
        vb = s if rows == n else sh
        out = s if (rows == n and name == "amg_post_jacobi") else sh
        return M(lv[0], rows, sh) + rows * (sh + vb + sh + out)
    if name == "amg_restrict":
        for nf, (na, nc, nr, npp, *_lanes) in levels.items():
            if nc == rows:
                return M(nr, nc, sh) + nf * sh + nc * sh
    if name == "amg_prolong" and lv:
        return M(lv[3], rows, sh) + lv[1] * sh + 2 * rows * sh
    if name == "advection_matrix":
        return n * (d * s + minv * s + (2 * d + 1) * s + s)
    if name == "advection_rhs":
        return n * ((2 * d + 2) * s)
    if name == "pressure_matrix":
        return n * (s + minv * s + (2 * d + 1) * s)
    if name == "pressure_rhs":
        return n * ((2 * d + 1) * (s + 4) + 4 + 3 * d * s + s)
    if name == "pressure_div":
        return n * (d * s + minv * s + s)
    if name == "correct_velocity":
        return n * (2 * s + 2 * d * s)
    if name == "epot_rhs":
        return n * (3 * s + minv * s + s)
    if name == "current_density":
        return n * (s + 3 * s + full * s + 3 * s)
    return None


def hierarchy_levels(sim):
    """Level rows -> (nnz_A, n_coarse, nnz_R, nnz_P, lanes_A, lanes_R, lanes_P) of the pressure
    and potential hierarchies (lanes of the sliced copies; 1 without one)."""
    out = {}
    for h in (getattr(sim, "_pressure_amg_hierarchy", None), getattr(sim, "_epot_amg_hierarchy", None)):
        if h is None:
            continue
        for i, lvl in enumerate(h.levels[:-1]):
            nxt = h.levels[i + 1]
            lanes = [lvl.shared[k].lanes if k in lvl.shared else 1 for k in ("A", "R", "P")]
            out[lvl.n] = (lvl.A.getNnz(), nxt.n, int(lvl.R_col.numel()), int(lvl.P_col.numel()), *lanes)
    return out


# ----------------------------------------------------------------------------- measured traffic
TRAFFIC_BEGIN, TRAFFIC_END = "<!-- traffic-json", "traffic-json -->"


def traffic_summary(fetch_csv, write_csv, copy_mib, command_note):
    """Markdown summary of the two counter runs; the figures also go into a JSON comment block
    that ``load_traffic`` reads back."""
    from phipict_amd import profiling

    tab = profiling.traffic_table(fetch_csv, write_csv, copy_mib << 20)
    entries = []
    for (base, targs, grid), r in sorted(tab["kernels"].items(), key=lambda kv: -kv[1]["fetch_bytes"] * kv[1]["dispatches"]):
        entries.append({"kernel": base, "targs": list(targs), "grid": grid, **r})
    lines = [command_note, "",
             f"Counter scale from `k_copy_bytes` moving {copy_mib} MiB: FETCH_SIZE reads {tab['fetch_scale']:.4f} x the "
             f"bytes read, WRITE_SIZE {tab['write_scale']:.4f} x the bytes written; the figures below are divided by "
             "these factors. Early exits: dispatches below 2% of the group's largest fetch (launches after the solve "
             "had converged); the means are over the other dispatches.", "",
             "| Kernel | Template | Grid | Dispatches | Early exits | Fetch MB | Write MB |",
             "|---|---|---:|---:|---:|---:|---:|"]
    for e in entries[:40]:
        lines.append(f"| `{e['kernel']}` | `{', '.join(e['targs'])}` | {e['grid']} | {e['dispatches']} | "
                     f"{e['early_exits']} | {e['fetch_bytes'] / 1e6:.3f} | {e['write_bytes'] / 1e6:.3f} |")
    lines += ["", TRAFFIC_BEGIN, json.dumps({"fetch_scale": tab["fetch_scale"], "write_scale": tab["write_scale"],
                                             "kernels": entries}), TRAFFIC_END]
    return "\n".join(lines) + "\n"


def trace_summary(path, steps):
    """Kernels of the timed region of a rocprofv3 kernel trace of this script: the dispatches
    between the two MARK_BYTES k_copy_bytes markers."""
    import csv

    import numpy as np

    from phipict_amd import profiling

    with open(path, newline="") as f:
        rows = sorted(csv.DictReader(f), key=lambda r: int(r["Start_Timestamp"]))
    mark_grid = 256
    marks = [i for i, r in enumerate(rows) if profiling.split_kernel_name(r["Kernel_Name"])[0] == "k_copy_bytes"
             and int(r.get("Grid_Size_X") or r.get("Grid_Size") or 0) == mark_grid]
    if len(marks) < 2:
        raise ValueError("the trace holds no pair of timed-region markers")
    a, b = marks[-2], marks[-1]
    win = rows[a + 1:b]
    t_lo, t_hi = int(rows[a]["End_Timestamp"]), int(rows[b]["Start_Timestamp"])
    stats = {}
    for r in win:
        base, targs = profiling.split_kernel_name(r["Kernel_Name"])
        key = base + (f"<{', '.join(targs)}>" if targs else "")
        d = int(r["End_Timestamp"]) - int(r["Start_Timestamp"])
        s = stats.setdefault(key, [0, 0, None, 0])
        s[0] += 1
        s[1] += d
        s[2] = d if s[2] is None else min(s[2], d)
        s[3] = max(s[3], d)
    busy = sum(s[1] for s in stats.values())
    span = t_hi - t_lo
    # idle time between consecutive dispatches (start of one minus the latest end before it)
    gaps, last_end = [], t_lo
    for r in win:
        st, en = int(r["Start_Timestamp"]), int(r["End_Timestamp"])
        gaps.append(max(0, st - last_end))
        last_end = max(last_end, en)
    gaps = np.asarray(gaps, dtype=np.float64) / 1e3 if gaps else np.zeros(1)
    lines = [f"Timed region: {steps} steps, {span / 1e6:.2f} ms from the first marker's end to the second marker's "
             f"start ({span / 1e6 / steps:.3f} ms per step); {len(win)} dispatches ({len(win) / steps:.0f} per step); "
             f"summed kernel time {busy / 1e6:.2f} ms ({busy / max(span, 1):.1%} of the span). Idle gap before a "
             f"dispatch: median {np.median(gaps):.2f} us, 90th percentile {np.percentile(gaps, 90):.2f} us, total "
             f"{gaps.sum() / 1e3:.2f} ms.", "",
             "| Kernel | Calls | Per step | Total ms | Avg us | Min us | Max us | % of kernel time |",
             "|---|---:|---:|---:|---:|---:|---:|---:|"]
    for key, (c, tot, lo, hi) in sorted(stats.items(), key=lambda kv: -kv[1][1])[:45]:
        lines.append(f"| `{key}` | {c} | {c / steps:.1f} | {tot / 1e6:.3f} | {tot / c / 1e3:.2f} | {lo / 1e3:.2f} | "
                     f"{hi / 1e3:.2f} | {tot / max(busy, 1):.2%} |")
    return "\n".join(lines) + "\n"


def load_traffic(path):
    try:
        text = Path(path).read_text()
    except OSError:
        return None
    if TRAFFIC_BEGIN not in text:
        return None
    body = text.split(TRAFFIC_BEGIN, 1)[1].split(TRAFFIC_END, 1)[0]
    return json.loads(body)


def level_grid_threads(rows, lanes, cus, ell=False):
    """Threads of a V-cycle level launch: 64 / lanes rows per wavefront (64 on the ELL finest
    level), four wavefronts per workgroup, at most 8 workgroups per CU."""
    per = 64 if ell else 64 // lanes
    waves = -(-rows // per)
    return min(max(1, -(-waves // 4)), 8 * cus) * 256


def traffic_for(table, name, rows, n, dtype_name, levels, cus):
    """Measured bytes per launch of the timer's kernel ``name`` at ``rows`` rows, or None.
    Level passes are told apart by their pass (k_level's first template argument), and their
    grid, which follows from the rows and the lanes per row of the matrix they read (``levels``
    maps level rows to (nnz_A, n_coarse, nnz_R, nnz_P, lanes_A, lanes_R, lanes_P))."""
    from phipict_amd import profiling

    if table is None:
        return None
    base = profiling.TIMER_TO_KERNEL.get(name)
    if name.startswith("amg_"):
        op = profiling.AMG_TIMER_OP.get(name)
        if op is None:
            return None
        if rows == n and name == "amg_presmooth":
            base, grid, targ0 = "k_pcg_decide", level_grid_threads(rows, 1, cus, ell=True), None
        elif rows == n and name in ("amg_jacobi", "amg_post_jacobi"):
            grid, targ0 = level_grid_threads(rows, 1, cus, ell=True), op
        elif name == "amg_restrict":
            fine = next((nf for nf, lv in levels.items() if lv[1] == rows), None)
            if fine is None:
                return None
            grid, targ0 = level_grid_threads(rows, levels[fine][5], cus), op
        elif rows in levels:
            lanes = levels[rows][6] if name == "amg_prolong" else levels[rows][4]
            grid, targ0 = level_grid_threads(rows, lanes, cus), op
        else:
            return None
        cands = [e for e in table["kernels"] if e["kernel"] == base and e["grid"] == grid
                 and (targ0 is None or (e["targs"] and e["targs"][0] == targ0))]
    else:
        if rows != n:
            return None
        cands = [e for e in table["kernels"] if e["kernel"] == base]
    if not cands:
        return None
    e = max(cands, key=lambda e: e["fetch_bytes"])
    return {"bytes": int(e["fetch_bytes"] + e["write_bytes"]), "fetch": int(e["fetch_bytes"]),
            "write": int(e["write_bytes"]), "dispatches": e["dispatches"], "early_exits": e["early_exits"],
            "kernel": f"{e['kernel']}<{', '.join(e['targs'])}>", "grid": e["grid"]}


# ----------------------------------------------------------------------------- CPU baseline
def cpu_threads() -> int:
    """Threads of the CPU baseline: OMP_NUM_THREADS when set (the host share given to this
    process), else every host core."""
    try:
        return max(1, int(os.environ.get("OMP_NUM_THREADS", "")))
    except ValueError:
        return os.cpu_count() or 1


def cpu_baseline(size, steps, dtype_name):
    """The oracle (the C restatement, built with OpenMP: cell and row loops shared among the
    host threads) on the same duct, grid and solver settings as the GPU run, for a reduced
    number of steps after one untimed step (AMG setup and first assembly); steps/s of the
    timed steps, no extrapolation."""
    import numpy as np

    from oracle import oracle as O
    from oracle.piso_oracle import MHDSpec, OraclePISO

    nx, ny, nz = size
    dtype = np.float64 if dtype_name == "f64" else np.float32
    threads = cpu_threads()
    os.environ["OMP_NUM_THREADS"] = str(threads)
    O.use_threads(True)
    try:
        v = duct_vertices(nx, ny, nz)
        rng = np.random.default_rng(1)
        cz = [nz, ny, nx]
        faces = [O.FaceSpec("periodic"), O.FaceSpec("periodic")] + [O.FaceSpec("fixed") for _ in range(4)]
        N = HA ** 2 / RE
        src = np.array([N, 0.0, 0.0])
        spec = O.BlockSpec(v, faces, velocity=1e-3 * rng.standard_normal([3] + cz), pressure=np.zeros(cz),
                           vsource=src)
        od = O.OracleDomain([spec], 1.0 / RE, DT, dtype)
        atol = 1e-8 if dtype == np.float64 else 1e-5
        sim = OraclePISO(od, DT, advection_tol=atol, pressure_tol=TOL, pressure_use_amg=True,
                         mhd=MHDSpec(N, potential_tol=TOL, use_amg=True),
                         pre_velocity_setup=lambda o: o.od.set_source(0, src.astype(dtype)))
        t0 = time.perf_counter()
        sim.step()  # AMG setup and first assembly
        setup = time.perf_counter() - t0
        t0 = time.perf_counter()
        for _ in range(steps):
            sim.step()
        el = time.perf_counter() - t0
    finally:
        O.use_threads(False)
    sps = steps / el
    cells = nx * ny * nz
    return {
        "value": round(sps, 5),
        "unit": "steps/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"oracle (C restatement, OpenMP, {threads} threads) on the same {nx}x{ny}x{nz} duct "
                   f"({cells} cells), {steps} timed steps after one untimed step with the AMG setup "
                   f"({setup:.1f} s); {el:.2f} s timed, {sps * cells:.3e} cell-updates/s"),
        "cell_updates_per_s": round(sps * cells, 1),
    }


def main(argv=None) -> int:
    args = parse_args(argv)
    if args.trace_summary:
        sys.stdout.write(trace_summary(args.trace_summary, args.steps))
        return 0
    if args.traffic_summary:
        note = (f"Counter runs: `rocprofv3 --pmc FETCH_SIZE` and `rocprofv3 --pmc WRITE_SIZE` over `python3 bench.py "
                f"... --copy-probe-mib {args.copy_probe_mib}`; summary by `python3 bench.py --copy-probe-mib "
                f"{args.copy_probe_mib} --traffic-summary <fetch csv> <write csv>`.")
        sys.stdout.write(traffic_summary(args.traffic_summary[0], args.traffic_summary[1], args.copy_probe_mib, note))
        return 0
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported (the PISO path would "
              "shard by blocks across GPUs, which is not built)", file=sys.stderr)
        return 2
    if args.workload == "tjunction-adjoint":
        return tjunction_adjoint(args)
    if args.workload == "uneven-batch":
        return uneven_batch(args)
    if args.workload == "hartmann-c1":
        return hartmann_c1(args)
    if args.workload == "serving":
        return serving_bench(args)
    import numpy as np
    import torch

    import phipict_amd as pa
    from phipict_amd import profiling

    nx, ny, nz = (int(v) for v in args.size.lower().split("x"))
    dtype = torch.float64 if args.dtype == "f64" else torch.float32
    device = torch.device("cuda", 0)
    print(f"[bench] building the {nx}x{ny}x{nz} duct ({args.dtype})", file=sys.stderr, flush=True)
    dom, sim = make_sim(nx, ny, nz, dtype, device, skew=args.skew, inflow=args.inflow)
    import ctypes

    from phipict_amd import _lib

    if args.copy_probe_mib > 0:
        nb = args.copy_probe_mib << 20
        src = torch.ones(nb // 8, dtype=torch.float64, device=device)
        dst = torch.empty_like(src)
        for _ in range(3):
            _lib.call("pp_copy_bytes", _lib.ptr(src), _lib.ptr(dst), ctypes.c_int64(nb), _lib.stream_ptr(device))
        torch.cuda.synchronize()
        if not torch.equal(src, dst):
            raise RuntimeError("k_copy_bytes copy mismatch")
        del src, dst
    for i in range(args.warmup):
        if not sim.single_step():
            raise RuntimeError(f"warm-up step {i} failed")
    torch.cuda.synchronize()
    print(f"[bench] warm-up done ({args.warmup} steps)", file=sys.stderr, flush=True)
    # A long-running driver keeps its setup objects (torch, the AMG hierarchy) out of the cyclic
    # collector's generations; otherwise a full collection (about 50 ms with torch loaded)
    # lands in the timed steps whenever the allocation counters say so.
    gc.collect()
    gc.freeze()
    log_start = len(sim.solver_log)
    if args.timer_in_timed_region:
        profiling.start(args.timer_every)
    prof = None
    if args.host_profile:
        import cProfile

        prof = cProfile.Profile()
    # two 256-byte k_copy_bytes launches bracket the timed steps in kernel traces
    mark = torch.zeros(64, dtype=torch.float64, device=device)

    def marker():
        _lib.call("pp_copy_bytes", _lib.ptr(mark), _lib.ptr(mark[32:]), ctypes.c_int64(MARK_BYTES),
                  _lib.stream_ptr(device))

    marker()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    if prof is not None:
        prof.enable()
    for i in range(args.steps):
        if not sim.single_step():
            raise RuntimeError(f"step {i} failed")
    if prof is not None:
        prof.disable()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    marker()
    if not args.timer_in_timed_region:
        # the per-kernel timings for the roofline come from as many further steps: the timer's
        # event records and bookkeeping stay out of the throughput figure
        torch.cuda.synchronize()
        profiling.start(args.timer_every)
        for i in range(args.steps):
            if not sim.single_step():
                raise RuntimeError(f"instrumented step {i} failed")
        torch.cuda.synchronize()
    profiling.stop()
    if prof is not None:
        import io
        import pstats

        buf = io.StringIO()
        st = pstats.Stats(prof, stream=buf)
        st.sort_stats("tottime").print_stats(30)
        st.sort_stats("cumulative").print_stats(30)
        print(buf.getvalue(), file=sys.stderr)
    timings = profiling.read()
    print(f"[bench] timed {args.steps} steps in {elapsed:.3f} s", file=sys.stderr, flush=True)
    n = dom.getTotalSize()
    s = 8 if dtype == torch.float64 else 4
    sh = 4  # hierarchy values are float32
    fine = (n, dom.C.getNnz())
    levels = hierarchy_levels(sim)
    rows = []
    diag = all(b.transform is None or b.isTransformDiagonal() for b in dom.getBlocks())
    for t in timings:
        b = kernel_bytes(t.name, t.rows, s, sh, fine, levels, diag=diag)
        rows.append((t, b))
    rows.sort(key=lambda tb: -tb[0].est_total_ms)
    if args.details:
        from phipict_amd.solvers import amg as amg_mod

        for entry in amg_mod.RAP_TUNING_LOG:
            print(f"[bench] refresh kernel tuning: rows {entry[0]} {entry[1]}: {entry[2]} -> {entry[3]}",
                  file=sys.stderr)
        tot = sum(t.est_total_ms for t, _ in rows)
        print(f"{'kernel':22s} {'rows':>9s} {'launches':>9s} {'avg us':>9s} {'share':>7s} {'GB/s':>8s}", file=sys.stderr)
        for t, b in rows:
            gbs = b / (t.avg_ms * 1e-3) / 1e9 if (b and t.sampled) else float("nan")
            print(f"{t.name:22s} {t.rows:9d} {t.launches:9d} {t.avg_ms * 1e3:9.1f} {t.est_total_ms / max(tot, 1e-12):7.1%} "
                  f"{gbs:8.1f}", file=sys.stderr)
    # the roofline kernel: the largest share among kernels that do the same work on every launch
    # (the BiCGStab phases run the three momentum right-hand sides together and skip the ones
    # that have converged, so their bytes per launch vary)
    dom_k = next(((t, b) for t, b in rows if b is not None and t.sampled > 0 and not t.name.startswith("bicg")), None)
    roofline = None
    if dom_k is not None:
        t, b = dom_k
        achieved = b / (t.avg_ms * 1e-3) / 1e9
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "kernel": t.name, "rows": t.rows,
                    "bytes_per_launch": int(b), "avg_us": round(t.avg_ms * 1e3, 2), "launches": t.launches,
                    "share_of_timed_kernels": round(t.est_total_ms / max(sum(x.est_total_ms for x, _ in rows), 1e-12), 3)}
        tfile = args.traffic_file or str(ROOT / "profiles" / f"hbm_traffic_duct{nx}x{ny}x{nz}_{args.dtype}.md")
        cus = torch.cuda.get_device_properties(device).multi_processor_count
        tr = traffic_for(load_traffic(tfile), t.name, t.rows, n, args.dtype, levels, cus)
        if tr is not None:
            roofline["traffic"] = tr["bytes"]
            roofline["traffic_detail"] = {**tr, "source": os.path.relpath(tfile, ROOT)}
    # the whole step against the roofline: the algorithmic bytes of every timed kernel with a
    # byte model (launches over the instrumented steps) per step, over the un-instrumented step
    # time; the BiCGStab phases are counted with all right-hand sides active (an upper bound)
    if roofline is not None:
        covered = [(t, b) for t, b in rows if b is not None and t.sampled > 0]
        step_bytes = sum(b * t.launches for t, b in covered) / args.steps
        step_s = elapsed / args.steps
        roofline["step"] = {
            "algorithmic_bytes": int(step_bytes), "ms": round(1e3 * step_s, 3),
            "achieved": round(step_bytes / step_s / 1e9, 1),
            "frac": round(step_bytes / step_s / 1e9 / HBM_PEAK_GBS, 4),
            "kernel_time_covered": round(sum(t.est_total_ms for t, _ in covered)
                                         / max(sum(t.est_total_ms for t, _ in rows), 1e-12), 3)}
    iters = {}
    for _, tag, infos in sim.solver_log[log_start:log_start + len(sim.solver_log[log_start:]) // (1 if args.timer_in_timed_region else 2)]:
        iters.setdefault(tag, []).append(max(i[0] for i in infos))
    cpu = None
    if not args.no_cpu_baseline and args.skew == 0 and not args.inflow:
        print(f"[bench] CPU baseline: the oracle on the {nx}x{ny}x{nz} duct, {args.cpu_steps} steps, "
              f"{cpu_threads()} threads", file=sys.stderr, flush=True)
        cpu = cpu_baseline((nx, ny, nz), args.cpu_steps, args.dtype)
    sps = args.steps / elapsed
    line = {
        "metric": "PISO steps/s",
        "value": round(sps, 4),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * elapsed / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.dtype,
        "data": "synthetic",
        "config": {"workload": f"3D MHD duct {nx}x{ny}x{nz} cells ("
                               + ("uniform inflow U=1, advected outflow, open potential at both ends" if args.inflow
                                  else "periodic x, pressure-gradient source") + "), Re=500, "
                               f"Ha=100, dt=5e-4, AMG-PCG pressure+potential tol 1e-8, BiCGStab momentum"
                               + (f", skewed grid (interior vertices jittered by {args.skew} of the spacing), "
                                  f"non-orthogonal corner terms" if args.skew > 0 else ""),
                   "cells": n, "nnz": dom.C.getNnz(), "correctors": 2,
                   "host": "Python driver; gc.collect() + gc.freeze() after the warm-up"},
        "cell_updates_per_s": round(sps * n, 1),
        "solver_iterations_mean": {k: round(float(np.mean(v)), 2) for k, v in iters.items()},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())

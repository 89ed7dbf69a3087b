"""PISO time step driver.

Restates the control flow of the reference's ``Simulation._PISO_split_step``
(src/phipict/core/piso_simulation.py:3525-4169), ``single_step``
(src/phipict/simulation/simulation.py:262-333) and the adaptive CFL substeps
(piso_simulation.py:149-215, 4171-4340): hooks, momentum prediction (BiCGStab),
pressure correctors (CG or AMG-PCG), velocity correction, mean projection of the pressure,
tolerance resolution and the pressure AMG refresh policy (:2222-2392). Every kernel and solve
runs in the gfx950 library through :mod:`phipict_amd.ops`.
"""
from __future__ import annotations

import logging
import numbers
from typing import Any, Callable, Optional

import numpy as np
import torch

from . import ops
from .domain import CSRmatrix, Domain, transform_size
from .hooks import Hook, Hooks
from .solvers import amg as amg_mod
from .solvers.tolerance import SolverTolerance, default_tolerance, resolve_tolerance

_LOG = logging.getLogger("phipict_amd.piso")


class LinsolveError(RuntimeError):
    """A linear solve did not converge or produced a non-finite residual
    (src/phipict/core/piso_diff.py:655-791)."""


def subtract_env_mean(flat: torch.Tensor, batch: int) -> torch.Tensor:
    """Subtract each batched environment's mean (piso_simulation.py:68-89)."""
    if batch == 1:
        return flat - torch.mean(flat)
    per_env = flat.reshape(batch, -1)
    return (per_env - per_env.mean(dim=1, keepdim=True)).reshape(flat.shape)


def subtract_env_mean_(flat: torch.Tensor, batch: int) -> torch.Tensor:
    """In-place :func:`subtract_env_mean` of a contiguous vector (the storage the domain's
    descriptors point at stays the same)."""
    per_env = flat.view(batch, -1)
    per_env.sub_(per_env.mean(dim=1, keepdim=True))
    return flat


def center_solution(domain: Domain, flat: torch.Tensor, inplace: bool = False) -> torch.Tensor:
    """Subtract the mean of every environment (piso_simulation.py:68-89), or of every instance
    of a domain of independent blocks (each instance's singular system fixes its own
    constant)."""
    if not domain.hasIndependentBlocks():
        return subtract_env_mean_(flat, domain.getBatchSize()) if inplace else subtract_env_mean(flat, domain.getBatchSize())
    ids, lengths = domain.segmentIds(), domain.segmentLengths()
    parts = flat.numel() // domain.getTotalSize()
    view = flat.view(parts, -1) if inplace else flat.reshape(parts, -1)
    means = torch.zeros(parts, lengths.numel(), dtype=flat.dtype, device=flat.device)
    means.index_add_(1, ids, view)
    means /= lengths
    if inplace:
        view.sub_(means[:, ids])
        return flat
    return (view - means[:, ids]).reshape(flat.shape)


def segmented_tolerances(tol, rhs: torch.Tensor, domain: Domain) -> list:
    """Absolute tolerance of every (component, instance) piece of a segmented solve: a
    SolverTolerance resolves against each piece's own right-hand side (||b||/sqrt(n) over the
    instance's rows), a number applies to all."""
    segs = domain.getSegments()
    S, N = len(segs) - 1, domain.getTotalSize()
    parts = rhs.numel() // N
    if not isinstance(tol, SolverTolerance):
        t = tol if tol is not None else default_tolerance(rhs.dtype)
        vals = [float(v) for v in t] if isinstance(t, (list, tuple)) else [float(t)]
        return vals if len(vals) in (1, parts * S) else [vals[0]]
    ids, lengths = domain.segmentIds(), domain.segmentLengths()
    sq = torch.zeros(parts, S, dtype=torch.float64, device=rhs.device)
    sq.index_add_(1, ids, rhs.detach().reshape(parts, N).to(torch.float64) ** 2)
    scales = (sq.sqrt() / lengths.to(torch.float64).sqrt()).reshape(-1).cpu().numpy()
    return tol._bound(scales, rhs.dtype)


#: rows per workgroup tile the segmented engine's autotuner tries
SEGMENT_TILE_CANDIDATES = (256, 512, 1024, 2048, 4096)
#: autotuner decisions: (instances, rows, {tile rows: microseconds per iteration}, chosen)
SEGMENT_TUNING_LOG: list = []


def tune_segment_tiles(A: CSRmatrix, segments, rhs: torch.Tensor, use_BiCG: bool = False, iterations: int = 24) -> int:
    """Rows per workgroup tile of the segmented engine for this set of instances: times a fixed
    number of iterations of the call's solver (tolerance 0, zero start) on the first right-hand
    side component for each
    candidate tile size with HIP events on the current stream and keeps the fastest. Larger
    tiles mean fewer workgroups and shorter folds, smaller ones more parallelism on small
    instances."""
    n = A.getRows()
    b = rhs.reshape(-1)[:n].contiguous()
    if not b.is_cuda:
        return 1024
    x = torch.zeros_like(b)
    timings = {}
    for tile in SEGMENT_TILE_CANDIDATES:
        if tile > max(256, n):
            break
        ops.SolveSegmented(A, segments, b, x.zero_(), maxIterations=2, tolerance=0.0, useBiCG=use_BiCG,
                           tileRows=tile)  # warm-up
        start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        start.record()
        ops.SolveSegmented(A, segments, b, x.zero_(), maxIterations=iterations, tolerance=0.0, useBiCG=use_BiCG,
                           tileRows=tile)
        stop.record()
        stop.synchronize()
        timings[tile] = 1e3 * start.elapsed_time(stop) / iterations
    best = min(timings, key=timings.get)
    SEGMENT_TUNING_LOG.append((len(segments) - 1, n, timings, best))
    return best


def check_solver_return_infos(infos, tol, max_iter, return_best_result: bool, tag: str = "") -> None:
    """Raise LinsolveError like the reference (piso_diff.py:659-791): always for a non-finite
    residual, for non-convergence unless the best iterate is returned."""
    if any(not i.isFiniteResidual for i in infos):
        bad = [k for k, i in enumerate(infos) if not i.isFiniteResidual]
        raise LinsolveError(f"Linear solve ({tag}, tol:{tol:.03e}) reported non-finite residual for RHS {bad} of "
                            f"{len(infos)}: {[str(i) for i in infos]}")
    if any(not i.converged for i in infos):
        bad = [k for k, i in enumerate(infos) if not i.converged]
        msg = f"Linear solve ({tag}, tol:{tol:.03e}) RHS {bad} of {len(infos)} did not converge after {max_iter} iterations."
        if return_best_result:
            _LOG.warning(msg)
        else:
            raise LinsolveError(msg + f" Solver infos: {[str(i) for i in infos]}")


# ---- boundary driving (piso_simulation.py:308-838)
# The reference computes these with tensor operations on the (small) face tensors; the same is
# done here, on the device, and without host synchronisation: the balance's "already balanced"
# test selects a scale of exactly 1 instead of branching on the host.

def _face_slice(block, face: int) -> torch.Tensor:
    """Block velocity in the cell layer next to ``face`` (piso_simulation.py:735-775)."""
    dims = block.getSpatialDims()
    axis, upper = face >> 1, face & 1
    idx = [slice(None), slice(None)] + [slice(None)] * dims
    idx[2 + dims - 1 - axis] = slice(-1, None) if upper else slice(0, 1)
    return block.velocity[tuple(idx)]


def _per_env_sum(flux: torch.Tensor) -> torch.Tensor:
    """Sum over the face cells, one value per environment ([N])."""
    return flux.reshape(flux.size(0), -1).sum(dim=1)


def get_fixed_boundary_flux(bound, bound_idx: int) -> torch.Tensor:
    """Total boundary-normal flux through a Dirichlet velocity face, per environment
    (piso_simulation.py:308-341)."""
    from .domain import BoundaryConditionType, FixedBoundary

    if not isinstance(bound, FixedBoundary):
        raise TypeError("boundary type not supported.")
    if bound.velocityType != BoundaryConditionType.DIRICHLET:
        raise ValueError("Only DIRICHLET boundaries are supported")
    return _per_env_sum(bound.GetFluxes())


def get_fixed_boundary_fluxes(list_idx_bound) -> torch.Tensor:
    """Net outward flux through (face index, boundary) pairs: [1], or [B] when a boundary has
    per-environment velocities (piso_simulation.py:344-382)."""
    if len(list_idx_bound) == 0:
        return None
    dom = list_idx_bound[0][1].getParentDomain()
    total = torch.zeros([1], dtype=dom.getDtype(), device=dom.getDevice())
    for face, bound in list_idx_bound:
        flux = get_fixed_boundary_flux(bound, face)
        total = total - flux if face % 2 == 0 else total + flux
    return total


def balance_boundary_fluxes(domain: Domain, free_bounds, tol=None) -> None:
    """Rescale the free boundaries so the net boundary flux vanishes, per environment
    (piso_simulation.py:533-610); a free boundary shared by all environments becomes
    per-environment when the scales differ."""
    if domain.hasIndependentBlocks() and domain.getNumBlocks() > 1:
        # every instance balances its own outflow against its own inflow
        for block in domain.getBlocks():
            own = [fb for fb in free_bounds if any(fb is b for _f, b in block.getFixedBoundaries())]
            if own:
                _balance(domain, block.getFixedBoundaries(), own, tol)
        return
    pairs = []
    for block in domain.getBlocks():
        pairs.extend(block.getFixedBoundaries())
    _balance(domain, pairs, free_bounds, tol)


def _balance(domain: Domain, pairs, free_bounds, tol) -> None:
    fixed = [pb for pb in pairs if not any(pb[1] is fb for fb in free_bounds)]
    free = [pb for pb in pairs if any(pb[1] is fb for fb in free_bounds)]
    if not free:
        return
    var_flux = get_fixed_boundary_fluxes(free)
    fix_flux = get_fixed_boundary_fluxes(fixed)
    if fix_flux is None:
        fix_flux = torch.zeros_like(var_flux)
    # an explicit tolerance passes through, otherwise the dtype default (piso_diff.py:397-421)
    balance_tol = tol if tol is not None else default_tolerance(domain.getDtype())
    # torch.allclose(fixed + variable, 0, atol=0.01 tol) decides on the host in the reference;
    # a scale of exactly 1 where balanced gives the same velocities without a synchronisation
    balanced = (fix_flux + var_flux).abs() <= float(balance_tol) * 0.01
    scale = torch.where(balanced.all(), torch.ones_like(var_flux), -fix_flux / var_flux)
    for face, bound in free:
        v = bound.velocity
        env_scale = scale.reshape([-1] + [1] * (v.dim() - 1)) if scale.numel() > 1 else scale
        bound.setVelocity(v * env_scale)


def get_advective_velocity(velms, velm_idx: int, bound, bound_idx: int) -> torch.Tensor:
    """Contravariant boundary-normal advection velocity of an outflow face, N1(D)HW
    (piso_simulation.py:455-530): the face's Minv row applied to ``velm``."""
    velm = velms if isinstance(velms, torch.Tensor) else velms[velm_idx]
    dom = bound.getParentDomain()
    dims = dom.getSpatialDims()
    axis = bound_idx >> 1
    shape = bound.face_shape()
    static = velm.dim() == 2
    velm = velm.to(device=dom.getDevice(), dtype=dom.getDtype())
    if not static and list(velm.shape[2:]) != shape:
        raise ValueError("varying velm size must match boundary")
    if bound.transform is not None:
        n_env = velm.size(0)
        if static:
            velm = velm.reshape([n_env, dims] + [1] * dims).expand(n_env, dims, *shape)
        t = bound.transform.reshape(-1, transform_size(dims))
        t_inv_row = t[:, dims * dims + axis * dims: dims * dims + (axis + 1) * dims]  # [fc, d]
        v = velm.reshape(n_env, dims, -1)  # [N, d, fc]
        adv = (t_inv_row.t()[None] * v).sum(dim=1)
        return adv.reshape([n_env, 1] + shape)
    if static:
        return velm[:, axis:axis + 1].reshape([velm.size(0), 1] + [1] * dims)
    return velm[:, axis:axis + 1]


def update_advective_boundaries(domain: Domain, bounds, velms, dt, tol=None) -> None:
    """Advect the velocity of outflow faces with ``velms`` over ``dt`` and balance the boundary
    fluxes afterwards (piso_simulation.py:611-838):
    ``u_b <- u_b - t (u_b - u_cell)``, ``t = 1 - 1/(1 + 2 dt v_n)``."""
    from .domain import FixedBoundary

    if not isinstance(bounds, list) or not bounds:
        raise ValueError("bounds must be a non-empty list of boundaries")
    dims, B = domain.getSpatialDims(), domain.getBatchSize()
    if isinstance(velms, torch.Tensor):
        if velms.dim() != 2 or velms.size(0) not in (1, B) or velms.size(1) != dims:
            raise ValueError("velms tensors have wrong shape")
    elif not (isinstance(velms, list) and len(velms) == len(bounds)):
        raise ValueError("velms tensors have wrong shape")
    if isinstance(dt, torch.Tensor) and dt.numel() > 1:
        dt = dt.to(device=domain.getDevice(), dtype=domain.getDtype()).view(-1, *([1] * (dims + 1)))
    found = []
    for block in domain.getBlocks():
        for face in range(2 * dims):
            bound = block.getBoundary(face)
            if isinstance(bound, FixedBoundary) and any(bound is b for b in bounds):
                found.append((block, face, bound))
    if len(found) != len(bounds):
        raise ValueError("Bounds passed as advective outflow are not (all) part of the domain.")
    for block, face, bound in found:
        if getattr(block, "passiveScalar", None) is not None:
            raise NotImplementedError("passive scalars are not implemented")
        if bound.isVelocityStatic():
            raise ValueError("outflow boundaries must be varying. Use FixedBoundary.makeVelocityVarying() to make "
                             "the velocity varying.")
        idx = next(i for i, b in enumerate(bounds) if b is bound)
        alpha = dt * 2 * get_advective_velocity(velms, idx, bound, face)
        t = 1 - 1 / (1 + alpha)
        v = bound.velocity
        bound.setVelocity(v - t * (v - _face_slice(block, face)))
    balance_boundary_fluxes(domain, bounds, tol=tol)


def get_max_time_step(domain: Domain, time_step_target: float, CFL_cond: float = 0.8,
                      with_transformations: bool = True) -> tuple:
    """Split ``time_step_target`` into equal substeps whose size meets ``CFL_cond`` at the
    current maximum velocity: ``(substep size, substeps)`` (piso_simulation.py:149-183)."""
    max_vel = domain.getMaxVelocity(True, with_transformations).cpu().numpy()
    max_time_step = CFL_cond / max_vel
    if max_time_step >= time_step_target:
        return time_step_target, 1
    substeps = int(np.ceil(time_step_target / max_time_step))
    return time_step_target / substeps, substeps


def _cfl_substep(time_step_target, max_vel, CFL_cond: float) -> tuple:
    """Size of the next substep of the remaining time ``time_step_target`` and the number of
    substeps it splits into at the maximum velocity ``max_vel`` (piso_simulation.py:186-215)."""
    if np.isclose(max_vel, 0):
        max_time_step = time_step_target
    else:
        max_time_step = CFL_cond / max_vel
    if max_time_step >= time_step_target:
        return time_step_target, 1
    substeps = int(np.ceil(time_step_target / max_time_step))
    return time_step_target / substeps, substeps


class Simulation:
    """PISO simulation of an incompressible flow on a multi-block domain (the reference's
    ``phipict.Simulation``; constructor arguments as simulation.py:166-248)."""

    NON_ORTHO_DIRECT_MATRIX = 1
    NON_ORTHO_DIRECT_RHS = 2
    NON_ORTHO_DIAGONAL_MATRIX = 4
    NON_ORTHO_DIAGONAL_RHS = 8
    NON_ORTHO_CENTER_MATRIX = 16
    PRESSURE_AMG_AUTO_MIN_ROWS = 200_000
    PRESSURE_AMG_AUTO_MIN_ROWS_BATCHED = 400_000
    _NON_ORTHO_MODE = NON_ORTHO_CENTER_MATRIX | NON_ORTHO_DIRECT_MATRIX | NON_ORTHO_DIAGONAL_RHS

    def __init__(self, domain: Optional[Domain] = None, dt: float = 1.0, substeps: int = 1, corrector_steps: int = 2,
                 advection_use_BiCG: bool = True, pressure_use_BiCG: bool = False, advection_tol=None,
                 pressure_tol=None, pressure_tol_intermediate=None, pressure_warm_start: bool = False,
                 pressure_use_amg: Optional[bool] = None, pressure_amg_options: Optional[dict] = None,
                 flux_balance_tol: float = 1e-5, advect_non_ortho_steps: int = 1, pressure_non_ortho_steps: int = 1,
                 normalize_pressure_result: bool = True, pressure_return_best_result: bool = False,
                 pressure_time_step_normalized: bool = False, velocity_corrector: str = "FD",
                 non_orthogonal: bool = True, differentiable: bool = False, hooks=None, prep_fn=None,
                 time_step: Optional[float] = None, adaptive_CFL: float = 0.8, adaptive_CFL_per_env: bool = True,
                 stop_fn: Optional[Callable[[], bool]] = None, advect_passive_scalar: bool = True,
                 **unsupported: Any) -> None:
        for k, v in unsupported.items():
            if k in ("density_viscosity", "scipy_solve_advection",
                     "scipy_solve_pressure", "preconditionBiCG", "BiCG_precondition_fallback", "convergence_tol",
                     "solver_double_fallback", "exclude_advection_solve_gradients",
                     "exclude_pressure_solve_gradients", "output_resampling_shape", "output_resampling_fill_max_steps"):
                if v not in (None, False, 0, 0.8, True) or k in ("scipy_solve_advection", "scipy_solve_pressure") and v:
                    raise NotImplementedError(f"option {k}={v!r} is not implemented")
                continue
            raise TypeError(f"unexpected argument {k!r}")
        if hooks is not None and prep_fn is not None:
            raise ValueError("Pass hooks or its former name prep_fn, not both.")
        if differentiable and velocity_corrector != "FD":
            raise NotImplementedError("the differentiable step uses the FD velocity corrector")
        self.domain = domain
        self.hooks = Hooks.coerce(hooks if hooks is not None else prep_fn)
        self.time_step = float(time_step if time_step is not None else dt)
        self.substeps = substeps
        self.adaptive_CFL = float(adaptive_CFL)
        self.adaptive_CFL_per_env = bool(adaptive_CFL_per_env)
        self.print_adaptive_step_info = False
        self.stop_fn = stop_fn
        self.advect_passive_scalar = bool(advect_passive_scalar)
        self.differentiable = bool(differentiable)
        # relative tolerance of the adjoint (transposed) solves, against the upstream gradient
        self.adjoint_rtol = 1e-12
        self.corrector_steps = int(corrector_steps)
        self.advection_use_BiCG = advection_use_BiCG
        self.pressure_use_BiCG = pressure_use_BiCG
        self.advection_tol = advection_tol
        self.pressure_tol = pressure_tol
        self.pressure_tol_intermediate = pressure_tol_intermediate
        self.pressure_warm_start = pressure_warm_start
        self.pressure_use_amg = pressure_use_amg
        self._pressure_amg_options = dict(pressure_amg_options or {})
        self.flux_balance_tol = flux_balance_tol
        self.advect_non_ortho_steps = advect_non_ortho_steps
        self.pressure_non_ortho_steps = pressure_non_ortho_steps
        self.normalize_pressure_result = normalize_pressure_result
        self.pressure_return_best_result = pressure_return_best_result
        self.pressure_time_step_normalized = pressure_time_step_normalized
        self._velocity_corrector_versions = {"FD": 1, "FVM_CENTER": 5, "FVM_FACE": 6}
        if velocity_corrector not in self._velocity_corrector_versions:
            raise ValueError(f"unknown velocity corrector {velocity_corrector!r}")
        self._velocity_corrector_version = self._velocity_corrector_versions[velocity_corrector]
        self.non_orthogonal = non_orthogonal
        self._non_ortho_flags = self._NON_ORTHO_MODE if non_orthogonal else 0
        self.linear_solve_max_iterations = 5000
        self.total_step = 0
        self.total_time = 0.0
        self._pressure_guess: dict = {}
        self._pressure_amg_hierarchy = None
        self._pressure_amg_version = None
        self._pressure_amg_refreshes = 0
        self._pressure_amg_baseline_iters: Optional[int] = None
        self._pressure_amg_stale = False
        self._pressure_rank_deficient_cache = None
        self.solver_log: list = []
        self.record_solver_stats = False

    @property
    def substeps(self) -> int:
        """Substeps per step; -1 for adaptive CFL substeps."""
        return self._substeps

    @substeps.setter
    def substeps(self, substeps) -> None:
        """An integer > 0, or ``"ADAPTIVE"`` (stored as -1) (piso_simulation.py:1596-1619)."""
        if isinstance(substeps, str) and substeps.upper() == "ADAPTIVE":
            substeps = -1
        else:
            if not isinstance(substeps, numbers.Integral):
                raise TypeError("substeps must be integral type or 'ADAPTIVE'.")
            if not substeps > 0:
                raise ValueError("substeps must be positive.")
        self._substeps = int(substeps)

    def _check_stop(self) -> bool:
        return self.stop_fn is not None and bool(self.stop_fn())

    # ---- helpers
    def _time_step_tensor(self) -> torch.Tensor:
        return torch.tensor([self.time_step], dtype=self.domain.getDtype())

    def _run_hook(self, hook: Hook, **kw) -> None:
        self.hooks.run(hook, domain=self.domain, total_step=self.total_step, total_time=self.total_time, **kw)

    def end_step(self, time_step) -> None:
        self.total_step += 1
        self.total_time += float(time_step.reshape(-1)[0]) if isinstance(time_step, torch.Tensor) else float(time_step)

    def _pressure_tol_for(self, cstep: int, pstep: int, corrector_steps: int):
        """Final corrector: pressure_tol, earlier ones: the intermediate tolerance (:2078-2105)."""
        if self.pressure_tol_intermediate is None:
            return self.pressure_tol
        final = cstep == corrector_steps - 1 and pstep == self.pressure_non_ortho_steps - 1
        return self.pressure_tol if final else self.pressure_tol_intermediate

    def _setup_pressure_matrix(self, domain, time_step, flags: int) -> None:
        """The pressure matrix of this corrector (pressure.cu:947-975). It depends on the
        advection diagonal, the metrics and the flags only, so a corrector after the first of a
        step finds it assembled already and skips the launch (unless the values were written
        since)."""
        if domain.P is not None and getattr(domain, "_p_matrix_key", None) == ops.pressure_matrix_key(domain, flags):
            return
        ops.SetupPressureMatrix(domain, time_step, flags, False)

    def _pressure_initial_guess(self, cstep: int, reference: torch.Tensor):
        if not self.pressure_warm_start:
            return None
        g = self._pressure_guess.get(cstep)
        if g is None or g.shape != reference.shape or not bool(torch.isfinite(g).all()):
            return None
        return g.to(dtype=reference.dtype, device=reference.device).clone()

    def _store_pressure_guess(self, cstep: int, result: torch.Tensor) -> None:
        if self.pressure_warm_start:
            g = self._pressure_guess.get(cstep)
            r = result.detach()
            if g is not None and g.shape == r.shape and g.dtype == r.dtype and g.device == r.device:
                g.copy_(r)  # the stored buffer is reused
            else:
                self._pressure_guess[cstep] = r.clone()

    def _pressure_rank_deficient(self, P) -> bool:
        """Pure-Neumann pressure operator: every row sum vanishes (:2165-2220)."""
        if P is None:
            return False
        cached = self._pressure_rank_deficient_cache
        if cached is not None and cached[0] is P.index and cached[2] == P.getRows():
            return cached[1]
        value = P.value.reshape(-1)[: P.getNnz()]
        n = P.getRows()
        counts = (P.row[1:] - P.row[:-1]).to(torch.int64)
        seg = torch.repeat_interleave(torch.arange(n, device=value.device), counts)
        row_sum = torch.zeros(n, dtype=value.dtype, device=value.device).index_add_(0, seg, value)
        # the reference's threshold of 1e-10 of the largest entry is below FP32 rounding, which
        # would call every FP32 pressure operator regular; scale it with the dtype instead
        rel = max(1e-10, 64.0 * torch.finfo(value.dtype).eps)
        deficient = bool((row_sum.abs() > rel * value.abs().max()).sum().item() == 0)
        self._pressure_rank_deficient_cache = (P.index, deficient, n)
        return deficient

    def _get_pressure_amg_hierarchy(self, P):
        """Pressure AMG hierarchy, refreshed once per change of the advection diagonal (the
        only input of P that changes between steps): the finest level every change, the
        coarse Galerkin operators every ``coarse_refresh_interval`` changes or when the
        iteration counts rose (:2222-2392)."""
        if self.pressure_use_BiCG:
            return None
        if P.getBatchSize() > 1:
            return self._get_batched_pressure_amg_hierarchy(P, self.pressure_use_amg)
        use_amg = self.pressure_use_amg
        if use_amg is None:
            # the automatic choice never asks for AMG over independent instances (their
            # segmented solves are Krylov-only)
            if P.getRows() < self.PRESSURE_AMG_AUTO_MIN_ROWS or self.domain.hasIndependentBlocks():
                return None
        elif not use_amg:
            return None
        version = self.domain._a_version
        if self._pressure_amg_hierarchy is not None and self._pressure_amg_version == version:
            return self._pressure_amg_hierarchy
        options = dict(self._pressure_amg_options)
        interval = int(options.pop("coarse_refresh_interval", 4))
        refreshes = self._pressure_amg_refreshes
        previous = self._pressure_amg_hierarchy
        if previous is not None and interval > 1 and refreshes % interval != 0 and not self._pressure_amg_stale:
            self._pressure_amg_hierarchy = amg_mod.galerkin_refresh(previous, P, coarse=False)
        else:
            self._pressure_amg_hierarchy = None
            self._pressure_amg_hierarchy = amg_mod.hierarchy_for(P, project_constant=self._pressure_rank_deficient(P),
                                                                 **options)
            self._pressure_amg_refreshes = 0
            self._pressure_amg_stale = False
            self._pressure_amg_baseline_iters = None
            refreshes = 0
        self._pressure_amg_refreshes = refreshes + 1
        self._pressure_amg_version = version
        return self._pressure_amg_hierarchy

    def _get_batched_pressure_amg_hierarchy(self, P, use_amg):
        """Batched counterpart (:2294-2392): the environments share the interpolation and their
        coarse operators are refreshed together; in auto mode AMG is used from
        ``PRESSURE_AMG_AUTO_MIN_ROWS_BATCHED`` unknowns over all environments."""
        batch = P.getBatchSize()
        if use_amg is None:
            use_amg = (P.getRows() * batch >= self.PRESSURE_AMG_AUTO_MIN_ROWS_BATCHED
                       and not self.domain.hasIndependentBlocks())
        if not use_amg:
            return None
        version = self.domain._a_version
        previous = self._pressure_amg_hierarchy
        batched = previous is not None and previous.batch_size == batch
        if batched and self._pressure_amg_version == version:
            return previous
        options = dict(self._pressure_amg_options)
        interval = int(options.pop("coarse_refresh_interval", 4))
        coarse = (not batched or interval <= 1 or self._pressure_amg_refreshes % interval == 0
                  or self._pressure_amg_stale)
        self._pressure_amg_hierarchy = amg_mod.batched_hierarchy_for(
            P, batch, project_constant=self._pressure_rank_deficient(P), previous=previous if batched else None,
            coarse=coarse, **options)
        if coarse:
            self._pressure_amg_refreshes = 0
            self._pressure_amg_stale = False
            self._pressure_amg_baseline_iters = None
        self._pressure_amg_refreshes += 1
        self._pressure_amg_version = version
        return self._pressure_amg_hierarchy

    def _note_pressure_amg_iterations(self, infos) -> None:
        iters = max(int(i.usedIterations) for i in infos)
        base = self._pressure_amg_baseline_iters
        if base is None or self._pressure_amg_refreshes <= 1:
            self._pressure_amg_baseline_iters = iters if base is None else max(base, iters)
        elif iters > 2 * base + 5:
            self._pressure_amg_stale = True

    def _record(self, tag: str, infos) -> None:
        if self.record_solver_stats:
            self.solver_log.append((self.total_step, tag, [(i.usedIterations, i.finalResidual, i.converged) for i in infos]))

    # ---- linear solves (piso_simulation.py:2834-3199)
    def linear_solve(self, csrMat, rhs: torch.Tensor, x: Optional[torch.Tensor] = None, use_BiCG: bool = False,
                     tol=None, max_iter: Optional[int] = None, matrix_rank_deficient: bool = False,
                     residual_reset_step: int = 0, return_best_result: bool = False, amg_hierarchy=None,
                     tag: str = "unknown"):
        if max_iter is None:
            max_iter = self.linear_solve_max_iterations
        if self.domain is not None and self.domain.hasIndependentBlocks():
            return self._segmented_solve(csrMat, rhs, x, use_BiCG, tol, max_iter, matrix_rank_deficient,
                                         amg_hierarchy, tag)
        batch = self.domain.getBatchSize() if self.domain is not None else 1
        abs_tol = resolve_tolerance(tol, rhs, batch_size=batch)
        if isinstance(abs_tol, list):
            rows = rhs.numel() // csrMat.getRows()
            abs_tol = [t for t in abs_tol for _ in range(rows // batch)]
        if abs_tol is None:
            abs_tol = default_tolerance(rhs.dtype)
        if amg_hierarchy is not None:
            if x is None:
                x = torch.zeros_like(rhs)
            infos = amg_mod.amg_pcg_solve(rhs, x, amg_hierarchy, tol=abs_tol, max_iter=max_iter,
                                          return_best_result=return_best_result)
            if amg_hierarchy is self._pressure_amg_hierarchy:
                self._note_pressure_amg_iterations(infos)
            self._record(tag, infos)
            tv = max(abs_tol) if isinstance(abs_tol, list) else float(abs_tol)
            check_solver_return_infos(infos, tv, max_iter, return_best_result, tag)
            return x, True
        if x is None:
            x = torch.zeros_like(rhs)
        # the zero-RHS shortcut (piso_diff.py:867): systems above the on-chip size take it on the
        # device (no host round trip); small ones are checked here
        if csrMat.getRows() > ops.SMALL_SYSTEM_ROWS or not bool(rhs.eq(0).all()):
            infos = ops.SolveLinear(csrMat, rhs, x, maxIterations=max_iter, tolerance=abs_tol,
                                    convergenceCriterion=ops.ConvergenceCriterion.NORM2_NORMALIZED, useBiCG=use_BiCG,
                                    matrixRankDeficient=matrix_rank_deficient, residualResetSteps=residual_reset_step,
                                    transposeA=False, returnBestResult=return_best_result)
            self._record(tag, infos)
            tv = max(abs_tol) if isinstance(abs_tol, list) else float(abs_tol)
            check_solver_return_infos(infos, tv, max_iter, return_best_result, tag)
        else:
            x.zero_()
        return x, True

    def _segmented_solve(self, csrMat, rhs, x, use_BiCG, tol, max_iter, matrix_rank_deficient, amg_hierarchy, tag):
        """A solve over a domain of independent blocks: every instance (and velocity component)
        with its own recurrence, tolerance and stopping rules, one launch per solver phase."""
        if amg_hierarchy is not None:
            raise NotImplementedError("AMG-preconditioned solves of independent blocks are not implemented: use CG")
        if x is None:
            x = torch.zeros_like(rhs)
        tols = segmented_tolerances(tol, rhs, self.domain)
        tile = self.__dict__.get("segment_tile_rows")
        if tile is None:
            tile = self.segment_tile_rows = tune_segment_tiles(csrMat, self.domain.getSegments(), rhs, use_BiCG)
        infos = ops.SolveSegmented(csrMat, self.domain.getSegments(), rhs, x, maxIterations=max_iter, tolerance=tols,
                                   useBiCG=use_BiCG, matrixRankDeficient=matrix_rank_deficient, tileRows=tile)
        self._record(tag, infos)
        check_solver_return_infos(infos, max(tols), max_iter, False, tag)
        return x, True

    # ---- the step
    def single_step(self, static: bool = False) -> bool:
        """One step: flux-balance check, then ``substeps`` PISO substeps, adaptive CFL substeps
        (-1) or substeps fixed from the current velocity (-2) (simulation.py:262-333)."""
        balance = self.domain.GetBoundaryFluxBalance()
        if float(balance.detach().abs().max()) > self.flux_balance_tol:
            raise RuntimeError(f"Domain boundary fluxes not balanced, cannot proceed with simulation step. "
                               f"Flux balance: {balance}, flux_balance_tol: {self.flux_balance_tol}")
        substeps = self.substeps
        time_step = None
        adaptive = False
        if substeps > 0:
            pass
        elif substeps == -1:
            adaptive = True
        elif substeps == -2:
            # substeps from the current velocity, kept for the whole step
            ts, substeps = get_max_time_step(self.domain, self.time_step, 0.8, with_transformations=True)
            time_step = torch.tensor([float(ts)], dtype=self.domain.getDtype())
        else:
            raise ValueError("Invalid substeps")
        try:
            if adaptive:
                return self._PISO_adaptive_step()
            if self.differentiable:
                return self._PISO_split_step_diff(substeps, time_step=time_step)
            return self._PISO_split_step(substeps, time_step=time_step)
        except LinsolveError:
            _LOG.exception("Simulation failed in step (total step %d):", self.total_step)
            return False

    def _PISO_split_step(self, iterations: int, time_step: Optional[torch.Tensor] = None) -> bool:
        domain = self.domain
        if domain is None or not domain.IsInitialized():
            raise RuntimeError("the simulation needs a prepared domain (domain.PrepareSolve())")
        if time_step is None:
            time_step = self._time_step_tensor()
        solve_ok = True
        flags = self._non_ortho_flags
        vcv = self._velocity_corrector_version
        if vcv != 1:
            raise NotImplementedError("only the FD velocity corrector is implemented")
        for step in range(int(iterations)):
            self._run_hook(Hook.PRE, local_step=step, time_step=time_step)
            domain.UpdateDomainData()
            matrix_for_velocity = False
            if self.advect_passive_scalar and domain.hasPassiveScalar():
                matrix_for_velocity, solve_ok = self._advect_scalar(step, time_step, flags)
                if not solve_ok:
                    return solve_ok
            # momentum prediction
            self._run_hook(Hook.PRE_VELOCITY_SETUP, local_step=step, time_step=time_step)
            if not matrix_for_velocity:
                ops.SetupAdvectionMatrix(domain, time_step, flags)
            domain._a_version += 1
            if not self.non_orthogonal:
                ops.SetupAdvectionVelocity(domain, time_step, 0, False)
                self._run_hook(Hook.POST_VELOCITY_SETUP, no_step=0, local_step=step, time_step=time_step)
                vel, solve_ok = self.linear_solve(domain.C, domain.velocityRHS, x=domain.velocityResult,
                                                  use_BiCG=self.advection_use_BiCG, tol=self.advection_tol,
                                                  tag="velocity")
                domain.setVelocityResult(vel)
                domain.UpdateDomainData()
            else:
                ops.CopyVelocityResultFromBlocks(domain)
                for no_step in range(self.advect_non_ortho_steps):
                    ops.SetupAdvectionVelocity(domain, time_step, flags, False)
                    self._run_hook(Hook.POST_VELOCITY_SETUP, no_step=no_step, local_step=step, time_step=time_step)
                    x = None if no_step == 0 else domain.velocityResult
                    vel, solve_ok = self.linear_solve(domain.C, domain.velocityRHS, x=x,
                                                      use_BiCG=self.advection_use_BiCG, tol=self.advection_tol,
                                                      tag="velocity")
                    domain.setVelocityResult(vel)
                    domain.UpdateDomainData()
            if not solve_ok:
                return solve_ok
            self._run_hook(Hook.POST_PREDICTION, local_step=step, time_step=time_step)
            # pressure correctors
            for cstep in range(self.corrector_steps):
                if not self.non_orthogonal:
                    self._setup_pressure_matrix(domain, time_step, 0)
                    ops.SetupPressureRHS(domain, time_step, 0, False, timeStepNorm=self.pressure_time_step_normalized)
                    self._run_hook(Hook.POST_PRESSURE_SETUP, local_step=step, time_step=time_step)
                    # solved in place in the domain's pressure vector (zero start, or the warm start)
                    x = domain.pressureResult
                    g = self._pressure_guess.get(cstep) if self.pressure_warm_start else None
                    if g is None or g.shape != x.shape or not bool(torch.isfinite(g).all()):
                        x.zero_()
                    else:
                        x.copy_(g)
                    p, solve_ok = self.linear_solve(domain.P, domain.pressureRHSdiv, x=x,
                                                    use_BiCG=self.pressure_use_BiCG,
                                                    tol=self._pressure_tol_for(cstep, 0, self.corrector_steps),
                                                    return_best_result=self.pressure_return_best_result,
                                                    amg_hierarchy=self._get_pressure_amg_hierarchy(domain.P),
                                                    tag="pressure")
                    if not solve_ok:
                        return solve_ok
                    if self.normalize_pressure_result:
                        p = center_solution(domain, p, inplace=True)
                    self._store_pressure_guess(cstep, p)
                    domain.setPressureResult(p)
                    domain.UpdateDomainData()
                    self._run_hook(Hook.POST_PRESSURE_RESULT, local_step=step, time_step=time_step)
                else:
                    self._setup_pressure_matrix(domain, time_step, flags)
                    for pstep in range(self.pressure_non_ortho_steps):
                        if pstep == 0:
                            ops.SetupPressureRHS(domain, time_step, flags, False,
                                                 timeStepNorm=self.pressure_time_step_normalized)
                        else:
                            ops.SetupPressureRHSdiv(domain, time_step, flags, False,
                                                    timeStepNorm=self.pressure_time_step_normalized)
                        self._run_hook(Hook.POST_PRESSURE_SETUP, local_step=step, time_step=time_step)
                        x = (domain.pressureResult.clone() if pstep > 0
                             else self._pressure_initial_guess(cstep, domain.pressureRHSdiv))
                        p, solve_ok = self.linear_solve(domain.P, domain.pressureRHSdiv, x=x,
                                                        matrix_rank_deficient=False, residual_reset_step=100,
                                                        use_BiCG=self.pressure_use_BiCG,
                                                        tol=self._pressure_tol_for(cstep, pstep, self.corrector_steps),
                                                        return_best_result=self.pressure_return_best_result,
                                                        amg_hierarchy=self._get_pressure_amg_hierarchy(domain.P),
                                                        tag="pressure")
                        if self.normalize_pressure_result:
                            p = center_solution(domain, p)
                        if pstep == self.pressure_non_ortho_steps - 1:
                            self._store_pressure_guess(cstep, p)
                        domain.setPressureResult(p)
                        domain.UpdateDomainData()
                        self._run_hook(Hook.POST_PRESSURE_RESULT, local_step=step, time_step=time_step)
                        if not solve_ok:
                            return solve_ok
                    self._run_hook(Hook.POST_PRESSURE_NON_ORTHO, local_step=step, time_step=time_step)
                ops.CopyPressureResultToBlocks(domain)
                ops.CorrectVelocity(domain, time_step, version=vcv, timeStepNorm=self.pressure_time_step_normalized)
                self._run_hook(Hook.POST_VELOCITY_CORRECTION, local_step=step, time_step=time_step)
            ops.CopyVelocityResultToBlocks(domain)
            self._run_hook(Hook.POST, local_step=step, time_step=time_step)
            if not solve_ok:
                break
            self.end_step(time_step)
        return solve_ok

    # ---- passive scalar (piso_simulation.py:3588-3775)
    def _advect_scalar(self, step: int, time_step, flags: int) -> tuple:
        """Implicit advection-diffusion of every passive-scalar channel: one matrix for all
        channels, or one per channel when their diffusivities differ; the scalar matrix is
        also the momentum matrix when neither the scalar nor a block has its own viscosity.
        Returns (the scalar matrix serves the momentum step, solves converged)."""
        domain = self.domain
        B, C, n = domain.getBatchSize(), domain.getPassiveScalarChannels(), domain.getTotalSize()
        split = not domain.isPassiveScalarViscosityStatic()
        for_velocity = not (domain.hasPassiveScalarViscosity() or any(b.viscosity is not None for b in domain.getBlocks())
                            or domain.hasPassiveScalarBlockViscosity() or split)
        matrices = []
        if split:
            for ch in range(C):
                ops.SetupAdvectionMatrix(domain, time_step, flags, forPassiveScalar=True, passiveScalarChannel=ch)
                matrices.append(CSRmatrix(domain.C.value.clone(), domain.C.index, domain.C.row))
        else:
            ops.SetupAdvectionMatrix(domain, time_step, flags, forPassiveScalar=True, passiveScalarChannel=0)
        if self.non_orthogonal:
            ops.CopyScalarResultFromBlocks(domain)
        ok = True
        for no_step in range(self.advect_non_ortho_steps if self.non_orthogonal else 1):
            ops.SetupAdvectionScalar(domain, time_step, flags)
            self._run_hook(Hook.POST_SCALAR_SETUP, no_step=no_step, local_step=step, time_step=time_step)
            if split:
                rhs = domain.scalarRHS.view(B, C, n)
                x0 = domain.scalarResult.view(B, C, n) if (self.non_orthogonal and no_step > 0) else None
                out = torch.empty_like(rhs)
                for ch in range(C):
                    r, ok = self.linear_solve(matrices[ch], rhs[:, ch].contiguous().view(-1),
                                              x=None if x0 is None else x0[:, ch].contiguous().view(-1),
                                              use_BiCG=self.advection_use_BiCG, tol=self.advection_tol, tag="scalar")
                    out[:, ch] = r.view(B, n)
                result = out.view(-1)
            else:
                x0 = domain.scalarResult.clone() if (self.non_orthogonal and no_step > 0) else None
                result, ok = self.linear_solve(domain.C, domain.scalarRHS, x=x0, use_BiCG=self.advection_use_BiCG,
                                               tol=self.advection_tol, tag="scalar")
            domain.setScalarResult(result)
            domain.UpdateDomainData()
            if not ok:
                return for_velocity, ok
        ops.CopyScalarResultToBlocks(domain)
        return for_velocity, ok

    # ---- the differentiable step (piso_diff.py; piso_simulation.py:3525-4169 with
    # differentiable operators)
    def _diff_solver(self, csr, use_bicg: bool, tol, tag: str, amg_hierarchy=None, rank_deficient: bool = False):
        def solve(b):
            x, _ = self.linear_solve(csr, b, x=torch.zeros_like(b), use_BiCG=use_bicg, tol=tol,
                                     matrix_rank_deficient=rank_deficient, amg_hierarchy=amg_hierarchy, tag=tag)
            return x
        return solve

    def _PISO_split_step_diff(self, iterations: int, time_step: Optional[torch.Tensor] = None) -> bool:
        """The orthogonal PISO step on autograd-tracked domain tensors: the blocks'
        velocities, the fixed faces' velocities and the velocity sources are inputs of
        differentiable operators whose backward runs the library's adjoint kernels and
        transposed solves. The new block velocities and pressures carry the graph."""
        from . import adjoint as adj

        domain = self.domain
        if domain is None or not domain.IsInitialized():
            raise RuntimeError("the simulation needs a prepared domain (domain.PrepareSolve())")
        if time_step is None:
            time_step = self._time_step_tensor()
        d, B = domain.getSpatialDims(), domain.getBatchSize()
        if self.advect_passive_scalar and domain.hasPassiveScalar():
            raise NotImplementedError("the differentiable step does not advect passive scalars")
        if int(domain.getAdvectionScheme()) != 0:
            raise NotImplementedError("the differentiable step covers the CENTRAL advection scheme")
        if self.non_orthogonal and not domain.isOrthogonal():
            # with non_orthogonal (the reference's default, piso_simulation.py:1469) the step is
            # the orthogonal one on grids whose metric is orthogonal to 1e-10 relative
            # (rectilinear grids: the face transforms of fixed faces carry off-diagonal terms of
            # rounding size only, whose corner terms are negligible)
            raise NotImplementedError("the differentiable step covers grids whose non-orthogonal corner terms vanish "
                                      "(rectilinear grids); this grid is skewed")
        tsn = bool(self.pressure_time_step_normalized)
        for step in range(int(iterations)):
            self._run_hook(Hook.PRE, local_step=step, time_step=time_step)
            domain.UpdateDomainData()
            self._run_hook(Hook.PRE_VELOCITY_SETUP, local_step=step, time_step=time_step)
            blocks = domain.getBlocks()
            vels = [b.velocity for b in blocks]
            bvels = [bd.velocity for _, _, bd in adj.fixed_faces(domain)]
            srcs = [b.velocitySource for b in blocks]
            viscs = adj.viscosity_inputs(domain)
            nv, nb = len(vels), len(bvels)
            A, Cv = adj.AdvectionMatrixFunction.apply(domain, time_step, *vels, *viscs)
            domain._a_version += 1
            rhs = adj.AdvectionRHSFunction.apply(domain, time_step, nv, nb, *vels, *bvels, *srcs, *viscs)
            self._run_hook(Hook.POST_VELOCITY_SETUP, no_step=0, local_step=step, time_step=time_step)
            solve_v = self._diff_solver(domain.C, self.advection_use_BiCG, self.advection_tol, "velocity")
            U = adj.LinearSolveFunction.apply(solve_v, domain.C, Cv, rhs, d, False, False, self.adjoint_rtol)
            self._run_hook(Hook.POST_PREDICTION, local_step=step, time_step=time_step)
            p = None
            for cstep in range(self.corrector_steps):
                domain.setVelocityResult(U.detach().clone())
                domain.setA(A.detach().clone())
                domain.UpdateDomainData()
                Pv = adj.PressureMatrixFunction.apply(domain, time_step, A)
                H, div = adj.PressureRHSFunction.apply(domain, time_step, tsn, A, Cv, U, nv, nb, *vels, *bvels,
                                                       *srcs, *viscs)
                self._run_hook(Hook.POST_PRESSURE_SETUP, local_step=step, time_step=time_step)
                singular = self._pressure_rank_deficient(domain.P)
                solve_p = self._diff_solver(domain.P, self.pressure_use_BiCG,
                                            self._pressure_tol_for(cstep, 0, self.corrector_steps), "pressure",
                                            amg_hierarchy=self._get_pressure_amg_hierarchy(domain.P))
                p = adj.LinearSolveFunction.apply(solve_p, domain.P, Pv, div, 1, singular, True, self.adjoint_rtol)
                if self.normalize_pressure_result:
                    p = adj._project(p, B)
                self._run_hook(Hook.POST_PRESSURE_RESULT, local_step=step, time_step=time_step)
                U = adj.CorrectVelocityFunction.apply(domain, time_step, tsn, H, A, p)
                self._run_hook(Hook.POST_VELOCITY_CORRECTION, local_step=step, time_step=time_step)
            for b, v, pr in zip(blocks, adj.flat_to_blocks(domain, U, d), adj.flat_to_blocks(domain, p, 1)):
                b.velocity, b.pressure = v, pr
            domain._changed = True
            domain.setVelocityResult(U.detach().clone())
            domain.setPressureResult(p.detach().clone())
            domain.UpdateDomainData()
            self._diff_velocity_result = U
            self._run_hook(Hook.POST, local_step=step, time_step=time_step)
            self.end_step(time_step)
        return True

    # ---- adaptive CFL substeps
    def _PISO_adaptive_step(self, CFL_cond: Optional[float] = None, max_subsetps: int = 1000) -> bool:
        """Advance by ``time_step`` in substeps re-derived before each one from the current
        maximum computational velocity (piso_simulation.py:4171-4240): one reduction and one
        4-byte read-back per substep."""
        domain = self.domain
        if domain is None or not domain.IsInitialized():
            raise RuntimeError("the simulation needs a prepared domain (domain.PrepareSolve())")
        CFL_cond = CFL_cond if CFL_cond is not None else self.adaptive_CFL
        if domain.getBatchSize() > 1 and self.adaptive_CFL_per_env and not self.differentiable:
            return self._PISO_adaptive_step_per_env(CFL_cond, max_subsetps)
        remaining = self.time_step
        substep = 0
        warned = False
        while remaining > 0 and not np.isclose(remaining, 0):
            max_vel = domain.getMaxVelocity(True, True).cpu().numpy()
            ts, substeps = _cfl_substep(remaining, max_vel, CFL_cond)
            remaining -= ts
            ts_t = torch.tensor([float(ts)], dtype=domain.getDtype())
            if substeps > max_subsetps and not warned:
                _LOG.warning("adaptive step (CFL=%.02f) results in more than %d substeps (%d).", CFL_cond,
                             max_subsetps, substeps)
                warned = True
            elif substep == 0 and self.print_adaptive_step_info:
                _LOG.debug("Adaptive step %d substeps: %d. From CFL = %.02f, max vel = %.03e, time step = %.03e.",
                           substep, substeps, CFL_cond, float(max_vel), remaining)
            ok = self._PISO_split_step(1, time_step=ts_t)
            substep += 1
            if not ok or self._check_stop():
                return False
        _LOG.debug("Adaptive time step %.03e (CFL=%.02f) used %d substeps.", self.time_step, CFL_cond, substep)
        return True

    def _PISO_adaptive_step_per_env(self, CFL_cond: float, max_subsetps: int) -> bool:
        """Adaptive substeps per batched environment (piso_simulation.py:4242-4340): every
        environment splits ``time_step`` by its own maximum velocity; the batch runs until the
        one with the most substeps is done, and an environment that is already done runs the
        remaining substeps at its last substep size with its state restored afterwards
        (:class:`phipict_amd.batching.EnvStateSnapshot`)."""
        from .batching import EnvStateSnapshot

        domain = self.domain
        B = domain.getBatchSize()
        start_time = self.total_time
        remaining = [self.time_step] * B
        time_steps = [self.time_step] * B
        taken = [0] * B
        substep = 0
        warned = False
        while True:
            active = [r > 0 and not np.isclose(r, 0) for r in remaining]
            if not any(active):
                break
            max_vel = domain.getMaxVelocityPerEnv(True, True).cpu().numpy()
            most = 0
            for b in range(B):
                if not active[b]:
                    continue  # keeps its last substep size
                time_steps[b], n = _cfl_substep(remaining[b], max_vel[b], CFL_cond)
                remaining[b] -= time_steps[b]
                taken[b] += 1
                most = max(most, n)
            if most > max_subsetps and not warned:
                _LOG.warning("adaptive step (CFL=%.02f) results in more than %d substeps (%d).", CFL_cond,
                             max_subsetps, most)
                warned = True
            done = [b for b in range(B) if not active[b]]
            snapshot = EnvStateSnapshot(domain, torch.tensor(done), extra=self._env_state_extras()) if done else None
            ts_t = torch.tensor([float(t) for t in time_steps], dtype=domain.getDtype())
            ok = self._PISO_split_step(1, time_step=ts_t)
            if snapshot is not None:
                snapshot.restore()
            substep += 1
            self.total_time = start_time + self.time_step - float(min(remaining))
            if not ok or self._check_stop():
                return False
        self.total_time = start_time + self.time_step
        _LOG.debug("Adaptive time step %.03e (CFL=%.02f) used %d batched substeps, per environment %s.",
                   self.time_step, CFL_cond, substep, taken)
        return True

    def _env_state_extras(self) -> list:
        """Getters of flat per-environment state outside the domain, restored with it for
        environments that sit out a substep (piso_simulation.py:4342-4353)."""
        return []

    def make_divergence_free(self, max_iter: int = 1000) -> bool:
        """Project the current velocity onto a divergence-free field: one pressure solve and
        correction with the current advection diagonal (piso_simulation.py:3369-3401)."""
        domain = self.domain
        ts = self._time_step_tensor()
        ops.CopyVelocityResultFromBlocks(domain)
        ops.SetupAdvectionMatrix(domain, ts, 0)
        domain._a_version += 1
        # project the block velocity: HbyA = u (pressureRHS = velocityResult), div, solve, correct
        domain.setPressureRHS(domain.velocityResult.clone())
        ops.SetupPressureMatrix(domain, ts, 0)
        ops.SetupPressureRHSdiv(domain, ts, 0)
        p, ok = self.linear_solve(domain.P, domain.pressureRHSdiv, use_BiCG=self.pressure_use_BiCG, tol=self.pressure_tol,
                                  max_iter=max_iter, tag="pressure")
        p = center_solution(domain, p)
        domain.setPressureResult(p)
        ops.CopyPressureResultToBlocks(domain)
        ops.CorrectVelocity(domain, ts, version=1)
        ops.CopyVelocityResultToBlocks(domain)
        return ok

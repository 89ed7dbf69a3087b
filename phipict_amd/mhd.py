"""Inductionless MHD on top of the PISO step (src/phipict/simulation/mhd.py:450-1002).

After every PISO substep (``Hook.POST``) the electric potential is solved from
``div(grad phi) = div(u x e_b)``; before every momentum assembly (``Hook.PRE_VELOCITY_SETUP``)
the Lorentz force ``N (J x e_b)`` with the face-based, discretely divergence-free current
density J is added to the blocks' velocity sources.
"""
from __future__ import annotations

import logging
from typing import Any, Optional

import torch

from . import ops
from .domain import Domain
from .hooks import Hook, Hooks
from .simulation import Simulation, subtract_env_mean
from .solvers import amg as amg_mod

_LOG = logging.getLogger("phipict_amd.mhd")


def epot_matrix_is_anchored(domain: Domain) -> bool:
    """True if some face cell prescribes phi (Dirichlet): only then is the potential matrix
    full rank (mhd.py:57-90)."""
    for block in domain.getBlocks():
        for _f, bound in block.getFixedBoundaries():
            if bound.hasPotentialDirichlet():
                return True
    return False


def compute_u_cross_eb_flat(vel_flat: torch.Tensor, e_b: torch.Tensor, total_size: int, dims: int) -> torch.Tensor:
    """u x e_b for a flat ``[B][d][n]`` velocity, returned as ``[B][3][n]`` (2D lifted to 3D,
    mhd.py:309-355). ``e_b``: uniform ``(3,)`` or one direction per batch ``(B, 3)``."""
    v = vel_flat.reshape(-1, dims, total_size)
    if dims == 2:
        v = torch.cat([v, torch.zeros_like(v[:, :1])], dim=1)
    eb = e_b.to(device=vel_flat.device, dtype=vel_flat.dtype)
    if eb.dim() == 1:
        eb = eb.reshape(1, 3, 1)
    else:
        eb = eb.reshape(-1, 3, 1)
    # (v x e)_x = v_y e_z - v_z e_y, ...
    cx = v[:, 1] * eb[:, 2] - v[:, 2] * eb[:, 1]
    cy = v[:, 2] * eb[:, 0] - v[:, 0] * eb[:, 2]
    cz = v[:, 0] * eb[:, 1] - v[:, 1] * eb[:, 0]
    return torch.stack([cx, cy, cz], dim=1).reshape(-1)


def _flat_to_block(block, flat: torch.Tensor, dims: int, total: int, batch: int) -> torch.Tensor:
    """Block view [B, d, (z), y, x] of a flat [B][d][n] field."""
    off, n = block.globalOffset, block.getCells()
    f = flat.reshape(batch, dims, total)[:, :, off:off + n]
    return f.reshape(batch, dims, *block.getSizesZYX())


class MHDSimulation(Simulation):
    """MHD simulation in the inductionless approximation (mhd.py:450-596)."""

    def __init__(self, domain: Domain, dt: float, stuart_number: torch.Tensor, e_b: torch.Tensor,
                 potential_tol=None, potential_use_BiCG: bool = False, potential_normalize: bool = True,
                 potential_non_ortho_flags: int = 0, potential_use_face_transform: bool = False,
                 potential_solve_max_iter: int = 5000, potential_solve_dtype=None,
                 exclude_potential_solve_gradients: bool = False, potential_reuse_result: bool = True,
                 potential_return_best_result: bool = True, potential_use_preconditioner: bool = False,
                 potential_amg_options: Optional[dict] = None, **kwargs: Any) -> None:
        hooks, prep_fn = kwargs.pop("hooks", None), kwargs.pop("prep_fn", None)
        if hooks is not None and prep_fn is not None:
            raise ValueError("Pass hooks or its former name prep_fn, not both.")
        hooks = Hooks(hooks if hooks is not None else prep_fn)
        super().__init__(domain=domain, dt=dt, hooks=hooks, **kwargs)
        if not isinstance(stuart_number, torch.Tensor) or stuart_number.numel() not in (1, domain.getBatchSize()):
            raise ValueError("stuart_number must be a scalar tensor (or one value per environment).")
        if potential_use_BiCG:
            raise NotImplementedError("the potential solve uses CG")
        if potential_solve_dtype is not None and potential_solve_dtype != domain.getDtype():
            raise NotImplementedError("a potential solve dtype different from the domain's is not implemented")
        self._stuart_number = stuart_number
        e_b = torch.as_tensor(e_b, dtype=domain.getDtype())
        if e_b.dim() == 1 and e_b.numel() != 3:
            raise ValueError("e_b must have 3 components")
        if e_b.dim() == 2 and (e_b.size(1) != 3 or e_b.size(0) != domain.getBatchSize()):
            raise ValueError("per-environment e_b must be [batch, 3]")
        if e_b.dim() > 2:
            raise NotImplementedError("spatially varying magnetic fields are not implemented")
        self._e_b = e_b
        self._e_b_device = e_b.to(domain.getDevice())
        self._potential_tol = potential_tol
        self._potential_normalize = potential_normalize
        self._potential_non_ortho_flags = potential_non_ortho_flags
        self._potential_solve_max_iter = potential_solve_max_iter
        self._potential_return_best_result = potential_return_best_result
        self._potential_use_preconditioner = potential_use_preconditioner
        self._potential_amg_options = dict(potential_amg_options or {})
        self._potential_reuse_result = potential_reuse_result
        self._epot_amg_hierarchy = None
        self._epot_amg_matrix = None
        self._dims = domain.getSpatialDims()
        domain.SetupEpotOnDomain(potential_non_ortho_flags, potential_use_face_transform)
        self._potential_rank_deficient = not epot_matrix_is_anchored(domain)
        self._potential_stale = False
        self._u_cross_eb_flat = compute_u_cross_eb_flat(domain.velocityResult, self._e_b_device, domain.getTotalSize(),
                                                        self._dims)
        self.hooks.append(Hook.POST, self._magnetic_step_callback)
        self.hooks.append(Hook.PRE_VELOCITY_SETUP, self._add_lorentz_force)

    @property
    def magnetic_field(self) -> torch.Tensor:
        return self._e_b

    def _stuart(self, like: torch.Tensor) -> torch.Tensor:
        return self._stuart_number.to(device=like.device, dtype=like.dtype)

    def _add_lorentz_force(self, domain: Domain, **kwargs: Any) -> None:
        """F_L = N (J x e_b) added to the velocity source of every block (mhd.py:543-594)."""
        if self._potential_stale:
            self._magnetic_step_callback(domain)
            self._potential_stale = False
        total, batch = domain.getTotalSize(), domain.getBatchSize()
        J = ops.ComputeCurrentDensityFaceBased(domain, domain.epotResult, self._u_cross_eb_flat)
        Jv = J.reshape(batch, 3, total)
        eb = self._e_b_device.to(Jv.dtype)
        eb = eb.reshape(1, 3, 1) if eb.dim() == 1 else eb.reshape(-1, 3, 1)
        N = self._stuart(Jv).reshape(-1, 1, 1)
        Fx = Jv[:, 1] * eb[:, 2] - Jv[:, 2] * eb[:, 1]
        Fy = Jv[:, 2] * eb[:, 0] - Jv[:, 0] * eb[:, 2]
        Fz = Jv[:, 0] * eb[:, 1] - Jv[:, 1] * eb[:, 0]
        FL = (N * torch.stack([Fx, Fy, Fz], dim=1))[:, : self._dims].reshape(-1).contiguous()
        for block in domain.getBlocks():
            fb = _flat_to_block(block, FL, self._dims, total, batch)
            S = block.velocitySource
            if S is not None:
                if S.dim() == 2:
                    S = S.reshape(*S.shape, *([1] * (fb.dim() - 2)))
                fb = fb + S
            block.setVelocitySource(fb.contiguous())
        domain.UpdateDomainData()

    def _epot_initial_guess(self, domain: Domain, rhs: torch.Tensor):
        if not self._potential_reuse_result:
            return None
        prev = domain.epotResult
        if prev is None or prev.numel() != rhs.numel():
            return None
        x = prev.detach().reshape(rhs.shape).to(dtype=rhs.dtype).clone()
        if self._potential_rank_deficient:
            x = subtract_env_mean(x, domain.getBatchSize())
        return x

    def _get_epot_amg_hierarchy(self, mat):
        if self._epot_amg_matrix is mat and self._epot_amg_hierarchy is not None:
            return self._epot_amg_hierarchy
        options = dict(self._potential_amg_options)
        options.setdefault("project_constant", self._potential_rank_deficient)
        self._epot_amg_hierarchy = amg_mod.hierarchy_for(mat, **options)
        self._epot_amg_matrix = mat
        return self._epot_amg_hierarchy

    def _magnetic_step_callback(self, domain: Domain, time_step=None, max_iter: Optional[int] = None, **kwargs) -> bool:
        """u x e_b, Poisson solve for phi, copy to blocks (mhd.py:877-1002)."""
        total, batch = domain.getTotalSize(), domain.getBatchSize()
        self._u_cross_eb_flat = compute_u_cross_eb_flat(domain.velocityResult, self._e_b_device, total, self._dims)
        inplane = self._u_cross_eb_flat.reshape(batch, 3, total)[:, : self._dims].reshape(-1).contiguous()
        rhs = ops.ComputeEpotRHS(domain, inplane)
        domain.setEpotRHS(rhs)
        domain.UpdateDomainData()
        rhs = domain.epotRHS
        if self._potential_rank_deficient:
            rhs = subtract_env_mean(rhs, batch)
        x = self._epot_initial_guess(domain, rhs)
        phi, ok = self.linear_solve(domain.Epot, rhs, x=x, use_BiCG=False, tol=self._potential_tol,
                                    matrix_rank_deficient=self._potential_rank_deficient,
                                    return_best_result=self._potential_return_best_result,
                                    max_iter=self._potential_solve_max_iter if max_iter is None else max_iter,
                                    amg_hierarchy=(self._get_epot_amg_hierarchy(domain.Epot)
                                                   if self._potential_use_preconditioner else None),
                                    tag="epot")
        if self._potential_normalize and self._potential_rank_deficient:
            phi = subtract_env_mean(phi, batch)
        domain.setEpotResult(phi)
        domain.UpdateDomainData()
        ops.CopyEpotResultToBlocks(domain)
        return bool(ok)

    def make_current_divergence_free(self, max_iter: int = 1000) -> bool:
        return self._magnetic_step_callback(self.domain, max_iter=max_iter)

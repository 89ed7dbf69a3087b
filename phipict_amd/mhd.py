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

from . import _lib, ops
from .domain import Domain
from .hooks import Hook, Hooks
from .simulation import Simulation, center_solution, subtract_env_mean, subtract_env_mean_
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
    mhd.py:309-355), in one library launch. ``e_b``: uniform ``(3,)`` or one direction per
    batch ``(B, 3)``."""
    if not vel_flat.is_contiguous():
        raise ValueError("the velocity must be contiguous")
    batch = vel_flat.numel() // (dims * total_size)
    eb = e_b.to(device=vel_flat.device, dtype=vel_flat.dtype).contiguous()
    out = torch.empty(batch * 3 * total_size, dtype=vel_flat.dtype, device=vel_flat.device)
    _lib.call("pp_cross_eb", _lib.ptr(vel_flat), dims, total_size, batch, _lib.ptr(eb), 1 if eb.dim() == 2 else 0,
              _lib.ptr(out), _lib.dtype_code(vel_flat.dtype), _lib.stream_ptr(vel_flat.device))
    return out


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
        per_instance = domain.hasIndependentBlocks() and isinstance(stuart_number, torch.Tensor) and \
            stuart_number.numel() == domain.getNumBlocks() and domain.getBatchSize() == 1
        if not isinstance(stuart_number, torch.Tensor) or not (stuart_number.numel() in (1, domain.getBatchSize())
                                                               or per_instance):
            raise ValueError("stuart_number must be a scalar tensor (or one value per environment, or per instance "
                             "of a domain of independent blocks).")
        self._stuart_per_block = per_instance and stuart_number.numel() > 1
        if potential_solve_dtype is not None and potential_solve_dtype not in (torch.float32, torch.float64):
            raise ValueError("potential_solve_dtype must be torch.float32 or torch.float64")
        self._potential_use_BiCG = bool(potential_use_BiCG)
        # the Poisson system may be solved in another precision than the domain's: up-cast
        # matrix and right-hand side, solve, cast back (mhd.py:940-949)
        self._potential_solve_dtype = (potential_solve_dtype if potential_solve_dtype != domain.getDtype() else None)
        self._epot_cast = None  # (source matrix, its values' version, the cast matrix)
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
        """N on the device of ``like`` (cached: a host-to-device copy waits for the stream)."""
        key = (like.device, like.dtype, self._stuart_number.data_ptr(), self._stuart_number._version)
        cached = self.__dict__.get("_stuart_cache")
        if cached is None or cached[0] != key:
            cached = (key, self._stuart_number.to(device=like.device, dtype=like.dtype))
            self._stuart_cache = cached
        return cached[1]

    def _add_lorentz_force(self, domain: Domain, **kwargs: Any) -> None:
        """F_L = N (J x e_b) added to the velocity source of every block (mhd.py:543-594)."""
        if self.differentiable:
            self._add_lorentz_force_diff(domain)
            return
        if self._potential_stale:
            self._magnetic_step_callback(domain)
            self._potential_stale = False
        total, batch = domain.getTotalSize(), domain.getBatchSize()
        J = ops.ComputeCurrentDensityFaceBased(domain, domain.epotResult, self._u_cross_eb_flat)
        dtype, dev = J.dtype, J.device
        eb = self._e_b_device.to(dtype).contiguous()
        N_all = self._stuart(J).reshape(-1).contiguous()
        code, stream = _lib.dtype_code(dtype), _lib.stream_ptr(dev)
        for bi, block in enumerate(domain.getBlocks()):
            N = N_all[bi:bi + 1] if self._stuart_per_block else N_all
            # F_L = N (J x e_b)[:d] over the block's cells plus its current source, one launch
            cells = block.getCells()
            out = torch.empty(batch, self._dims, *block.getSizesZYX(), dtype=dtype, device=dev)
            S = block.velocitySource
            kind, S_arg, S_batched = 0, None, 0
            if S is not None:
                S = S.to(device=dev, dtype=dtype).contiguous()
                kind = 1 if S.dim() == 2 else 2
                if S.numel() != S.size(0) * self._dims * (1 if kind == 1 else cells) or S.size(0) not in (1, batch):
                    raise ValueError("velocity source must be [1|B, d] or [1|B, d, (z), y, x]")
                S_arg, S_batched = S, 1 if S.size(0) > 1 else 0
            _lib.call("pp_lorentz_source", _lib.ptr(J), total, batch, self._dims, _lib.ptr(eb),
                      1 if eb.dim() == 2 else 0, _lib.ptr(N), 1 if N.numel() > 1 else 0, block.globalOffset, cells,
                      _lib.ptr(S_arg) if S_arg is not None else None, kind, S_batched, _lib.ptr(out), code, stream)
            block.setVelocitySource(out)
        domain.UpdateDomainData()

    def _epot_matrix_cast(self, mat):
        """The potential matrix in the solve dtype, cast once per set of matrix values."""
        key = (mat.value.data_ptr(), mat.value._version)
        if self._epot_cast is None or self._epot_cast[0] != key:
            self._epot_cast = (key, mat.toType(self._potential_solve_dtype))
        return self._epot_cast[1]

    def _epot_initial_guess(self, domain: Domain, rhs: torch.Tensor):
        if not self._potential_reuse_result:
            return None
        prev = domain.epotResult
        if prev is None or prev.numel() != rhs.numel():
            return None
        if prev.is_contiguous() and prev.dtype == rhs.dtype and not prev.requires_grad:
            x = prev.view(rhs.shape)  # solved in place: the result replaces it anyway
        else:
            x = prev.detach().reshape(rhs.shape).to(dtype=rhs.dtype).clone()
        if self._potential_rank_deficient:
            x = center_solution(domain, x, inplace=True)
        return x

    def _get_epot_amg_hierarchy(self, mat):
        if self._epot_amg_matrix is mat and self._epot_amg_hierarchy is not None:
            return self._epot_amg_hierarchy
        options = dict(self._potential_amg_options)
        options.setdefault("project_constant", self._potential_rank_deficient)
        self._epot_amg_hierarchy = amg_mod.hierarchy_for(mat, **options)
        self._epot_amg_matrix = mat
        return self._epot_amg_hierarchy

    def _env_state_extras(self) -> list:
        """Adds ``u x e_b`` of the last magnetic step, which the next substep's Lorentz force
        uses (mhd.py:864-875)."""
        return [*super()._env_state_extras(), lambda: self._u_cross_eb_flat]

    # ---- the differentiable magnetic step: u x e_b, the potential's right-hand side and solve,
    # the current density and the Lorentz force as tracked operators (piso_diff.py:2454-2740)
    def _diff_state(self, domain: Domain):
        """(phi, u x e_b) the next Lorentz force uses, as tracked tensors."""
        phi = self.__dict__.get("_diff_phi")
        if phi is None:
            phi = domain.epotResult.detach().clone()
        return phi, self._u_cross_eb_flat

    def _add_lorentz_force_diff(self, domain: Domain) -> None:
        from . import adjoint as adj

        phi, uxb = self._diff_state(domain)
        J = adj.CurrentDensityFunction.apply(domain, phi, uxb, *adj.potential_values(domain))
        e_b = self._e_b.to(device=domain.getDevice())
        N = self._stuart_number.to(device=domain.getDevice(), dtype=domain.getDtype())
        srcs = [b.velocitySource for b in domain.getBlocks()]
        for block, F in zip(domain.getBlocks(), adj.lorentz_sources(domain, J, e_b, N, srcs)):
            block.setVelocitySource(F)
        domain.UpdateDomainData()

    def _magnetic_step_diff(self, domain: Domain, max_iter: Optional[int] = None) -> bool:
        from . import adjoint as adj

        total, batch = domain.getTotalSize(), domain.getBatchSize()
        U = self.__dict__.get("_diff_velocity_result")
        if U is None:
            U = domain.velocityResult
        e_b = self._e_b.to(device=domain.getDevice())
        uxb = adj.CrossEbFunction.apply(U, e_b, total, self._dims)
        inplane = uxb.reshape(batch, 3, total)[:, : self._dims].reshape(-1)
        rhs = adj.EpotRHSFunction.apply(domain, inplane, *adj.potential_values(domain))
        singular = self._potential_rank_deficient
        if singular:
            rhs = adj._project(rhs, batch)
        hier = self._get_epot_amg_hierarchy(domain.Epot) if self._potential_use_preconditioner else None

        def solve(b):
            x, _ = self.linear_solve(domain.Epot, b, x=torch.zeros_like(b), use_BiCG=False, tol=self._potential_tol,
                                     matrix_rank_deficient=singular, return_best_result=False,
                                     max_iter=self._potential_solve_max_iter if max_iter is None else max_iter,
                                     amg_hierarchy=hier, tag="epot")
            return x

        phi = adj.LinearSolveFunction.apply(solve, domain.Epot, domain.Epot.value, rhs, 1, singular, True,
                                            self.adjoint_rtol)
        if self._potential_normalize and singular:
            phi = adj._project(phi, batch)
        self._u_cross_eb_flat = uxb
        self._diff_phi = phi
        domain.setEpotRHS(rhs.detach().clone())
        domain.setEpotResult(phi.detach().clone())
        # the blocks' potential: the copy of the result, as tracked tensors of their own (a loss of
        # block.epot reaches the inputs; later steps replace them instead of writing into them)
        for blk, e in zip(domain.getBlocks(), adj.flat_to_blocks(domain, phi, 1)):
            if blk.hasEpot():
                blk.setEpot(e.reshape(blk.epot.shape))
        domain.UpdateDomainData()
        return True

    def _magnetic_step_callback(self, domain: Domain, time_step=None, max_iter: Optional[int] = None, **kwargs) -> bool:
        """u x e_b, Poisson solve for phi, copy to blocks (mhd.py:877-1002)."""
        if self.differentiable:
            return self._magnetic_step_diff(domain, max_iter)
        self._diff_phi = None
        total, batch = domain.getTotalSize(), domain.getBatchSize()
        self._u_cross_eb_flat = compute_u_cross_eb_flat(domain.velocityResult, self._e_b_device, total, self._dims)
        inplane = self._u_cross_eb_flat.reshape(batch, 3, total)[:, : self._dims].reshape(-1).contiguous()
        # written into the domain's vectors in place: their descriptors stay valid
        prev = domain.epotRHS
        if prev is not None and prev.numel() == total * batch and prev.is_contiguous():
            rhs = ops.ComputeEpotRHS(domain, inplane, out=prev)
        else:
            rhs = ops.ComputeEpotRHS(domain, inplane)
            domain.setEpotRHS(rhs)
            domain.UpdateDomainData()
        rhs = domain.epotRHS
        mat = domain.Epot
        if self._potential_solve_dtype is not None:
            mat = self._epot_matrix_cast(mat)
            rhs = rhs.to(self._potential_solve_dtype)
        if self._potential_rank_deficient:
            rhs = center_solution(domain, rhs)
        x = self._epot_initial_guess(domain, rhs)
        # BiCGStab runs on the projected right-hand side without the CG's projected mode
        phi, ok = self.linear_solve(mat, rhs, x=x, use_BiCG=self._potential_use_BiCG, tol=self._potential_tol,
                                    matrix_rank_deficient=self._potential_rank_deficient and not self._potential_use_BiCG,
                                    return_best_result=self._potential_return_best_result,
                                    max_iter=self._potential_solve_max_iter if max_iter is None else max_iter,
                                    amg_hierarchy=(self._get_epot_amg_hierarchy(mat)
                                                   if self._potential_use_preconditioner else None),
                                    tag="epot")
        if phi.dtype != domain.getDtype():
            phi = phi.to(domain.getDtype())
        if self._potential_normalize and self._potential_rank_deficient:
            phi = center_solution(domain, phi, inplace=True)
        cur = domain.epotResult
        if cur is not None and cur.numel() == phi.numel() and cur.is_contiguous() and cur.data_ptr() != phi.data_ptr():
            cur.copy_(phi)
        else:
            domain.setEpotResult(phi)
        domain.UpdateDomainData()
        ops.CopyEpotResultToBlocks(domain)
        return bool(ok)

    def make_current_divergence_free(self, max_iter: int = 1000) -> bool:
        return self._magnetic_step_callback(self.domain, max_iter=max_iter)

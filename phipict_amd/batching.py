"""Batched environments: B copies of one domain stepped together, sharing the grid, the block
and boundary layout and the sparsity patterns, each with its own state (``[B, ...]`` block
fields) -- stacking single-environment domains into one, copying an environment's state, and
saving the state of some environments across a substep they sit out
(src/phipict/batching.py:19-320)."""
from __future__ import annotations

from typing import Callable, Optional, Sequence

import torch

from .domain import Domain, FixedBoundary

__all__ = ["EnvStateSnapshot", "copy_env_state", "stack_domains"]


def _copy_boundary_data(dst: torch.Tensor, src: torch.Tensor, env_idx: int, batch_size: int,
                        setter: Callable[[torch.Tensor], None]) -> None:
    """Write boundary data ``src`` (batch 1) into environment ``env_idx``; shared data equal to
    ``src`` stays shared, otherwise it becomes per-environment first (batching.py:129-150)."""
    if dst.shape[1:] != src.shape[1:]:
        raise ValueError(f"Boundary data shapes differ: {tuple(dst.shape[1:])} vs {tuple(src.shape[1:])}.")
    if dst.size(0) == 1:
        if torch.equal(dst, src):
            return
        dst = dst.expand(batch_size, *dst.shape[1:]).clone()
        dst[env_idx].copy_(src[0])
        setter(dst)
    else:
        dst[env_idx].copy_(src[0])


def copy_env_state(dst: Domain, env_idx: int, src: Domain) -> None:
    """Copy the state of a single-environment domain into environment ``env_idx`` of a batch
    (batching.py:19-127): block velocity, pressure and potential in place, and the velocity of
    every fixed face that differs from ``dst``'s (making it per-environment). The faces'
    Dirichlet potential values are one array shared by all environments here: values that
    differ between the domains raise."""
    if src.getBatchSize() != 1:
        raise ValueError(f"src must have batch size 1, got {src.getBatchSize()}.")
    if not 0 <= env_idx < dst.getBatchSize():
        raise ValueError(f"env_idx {env_idx} out of range for batch size {dst.getBatchSize()}.")
    if src.getNumBlocks() != dst.getNumBlocks():
        raise ValueError("src and dst have a different number of blocks.")
    with torch.no_grad():
        for db, sb in zip(dst.getBlocks(), src.getBlocks()):
            fields = [(db.velocity, sb.velocity), (db.pressure, sb.pressure)]
            if db.epot is not None and sb.epot is not None:
                fields.append((db.epot, sb.epot))
            for df, sf in fields:
                if df.shape[1:] != sf.shape[1:]:
                    raise ValueError(f"Block field shapes differ: {tuple(df.shape[1:])} vs {tuple(sf.shape[1:])}.")
                df[env_idx].copy_(sf[0])
            for face in range(2 * dst.getSpatialDims()):
                dbound, sbound = db.getBoundary(face), sb.getBoundary(face)
                if not isinstance(dbound, FixedBoundary) or not isinstance(sbound, FixedBoundary):
                    continue
                _copy_boundary_data(dbound.velocity, sbound.velocity, env_idx, dst.getBatchSize(), dbound.setVelocity)
                dv, sv = dbound.potentialValues, sbound.potentialValues
                if (dv is None) != (sv is None) or (dv is not None and not torch.equal(dv, sv)):
                    raise NotImplementedError("per-environment Dirichlet potential values are not supported")
            db._domain._changed = True
    dst.UpdateDomainData()


def stack_domains(domains: Sequence[Domain]) -> Domain:
    """Combine single-environment domains of one setup into a batched domain: the first
    becomes the batch (in place, prepared for solving) and environment ``b`` gets the state of
    ``domains[b]`` (batching.py:152-177)."""
    if len(domains) == 0:
        raise ValueError("stack_domains needs at least one domain.")
    batched = domains[0]
    if batched.getBatchSize() != 1:
        raise ValueError(f"Domains must have batch size 1, got {batched.getBatchSize()}.")
    batched.setBatchSize(len(domains))
    batched.PrepareSolve()
    for env_idx, domain in enumerate(domains[1:], start=1):
        copy_env_state(batched, env_idx, domain)
    return batched


class EnvStateSnapshot:
    """State of some environments of a batched domain, to undo a substep for them.

    Saved: the block fields (velocity, pressure, potential), the domain's solution vectors
    (the next solve's initial guess and the potential the Lorentz force uses), ``extra`` flat
    per-environment tensors, the velocity sources and the fixed faces' velocities. Fields are
    read through getters when restoring, since a step may replace a tensor rather than write
    it; sources and face velocities are restored only if the step replaced or wrote them, and
    restoring some environments of a shared one (batch 1) makes it per-environment.
    Potential values of fixed faces are one array shared by all environments here and are not
    part of the per-environment state.
    """

    def __init__(self, domain: Domain, env_idx: torch.Tensor,
                 extra: Sequence[Callable[[], Optional[torch.Tensor]]] = ()) -> None:
        self._domain = domain
        self._batch = domain.getBatchSize()
        self._idx = env_idx.to(device=domain.getDevice(), dtype=torch.long)
        getters: list = []
        for block in domain.getBlocks():
            getters += [lambda block=block: block.velocity, lambda block=block: block.pressure,
                        lambda block=block: block.epot]
        getters += [lambda: domain.velocityResult, lambda: domain.pressureResult, lambda: domain.epotResult, *extra]
        self._fields: list = []
        self._bounds: list = []
        with torch.no_grad():
            for getter in getters:
                t = getter()
                if t is None or t.numel() == 0:
                    continue
                self._fields.append((getter, self._env_rows(t).index_select(0, self._idx)))
            for block in domain.getBlocks():
                src = block.velocitySource
                self._bounds.append((lambda block=block: block.velocitySource, block.setVelocitySource,
                                     0 if src is None else src.data_ptr(), 0 if src is None else src._version,
                                     None if src is None else src.clone()))
                for bound in block.getBoundaries():
                    if not isinstance(bound, FixedBoundary):
                        continue
                    t = bound.velocity
                    self._bounds.append((lambda bound=bound: bound.velocity, bound.setVelocity, t.data_ptr(),
                                         t._version, t.clone()))

    def _env_rows(self, t: torch.Tensor) -> torch.Tensor:
        """``t`` as ``[B, -1]``: block fields are ``[B, ...]``, flat vectors ``B`` parts."""
        if t.numel() % self._batch != 0:
            raise ValueError(f"Tensor of {t.numel()} elements has no slice per environment "
                             f"(batch size {self._batch}).")
        return t.view(self._batch, -1)

    def restore(self) -> None:
        """Write the saved state back into the saved environments."""
        B = self._batch
        with torch.no_grad():
            for getter, rows in self._fields:
                t = getter()
                if t is None or t.numel() == 0:
                    continue
                self._env_rows(t).index_copy_(0, self._idx, rows)
            changed = False
            for getter, setter, ptr, version, saved in self._bounds:
                t = getter()
                if t is None or (t.data_ptr() == ptr and t._version == version):
                    continue  # untouched by the step
                if saved is None:
                    saved = torch.zeros_like(t[:1])  # there was no source
                if t.shape[1:] != saved.shape[1:]:
                    raise ValueError(f"Boundary data changed shape during the step: "
                                     f"{tuple(saved.shape)} to {tuple(t.shape)}.")
                restored = t.expand(B, *t.shape[1:]).clone()
                restored[self._idx] = saved.expand(B, *saved.shape[1:])[self._idx]
                setter(restored)
                changed = True
        if changed:
            self._domain.UpdateDomainData()

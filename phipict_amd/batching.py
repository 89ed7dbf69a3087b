"""Batched environments: B copies of one domain stepped together, sharing the grid, the block
and boundary layout and the sparsity patterns, each with its own state (``[B, ...]`` block
fields) -- stacking single-environment domains into one, copying an environment's state, and
saving the state of some environments across a substep they sit out
(src/phipict/batching.py:19-325)."""
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
    every fixed face that differs from ``dst``'s (making it per-environment), and likewise the
    faces' Dirichlet potential values (one slice per environment once they differ)."""
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
                if (dv is None) != (sv is None):
                    raise NotImplementedError("Dirichlet potential values on one side only")
                if dv is not None and not torch.equal(dv, sv):
                    # the environment gets its own slice of the values
                    fc, nb = dbound.face_cells(), dst.getBatchSize()
                    if sv.numel() != fc:
                        raise ValueError("the source environment's potential values must be one slice")
                    full = (dv.reshape(nb, fc) if dbound.potentialValuesBatched() else dv.reshape(1, fc).expand(nb, fc)).clone()
                    full[env_idx] = sv
                    dbound.setPotentialValues(full)
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


class _SavedRows:
    """Per-environment rows of one tensor that a step may write in place or replace: the
    tensor is looked up again when restoring (``lookup``), viewed as ``[B, -1]`` (block fields
    are ``[B, ...]``, flat solve vectors hold B equal parts) and the saved rows written back."""

    __slots__ = ("lookup", "rows")

    def __init__(self, lookup: Callable[[], Optional[torch.Tensor]], rows: torch.Tensor) -> None:
        self.lookup = lookup
        self.rows = rows


class _SavedInput:
    """A per-block input a step may replace through its setter (a velocity source or a fixed
    face's velocity): kept as the tensor identity (pointer and version) seen when saving plus a
    copy, so that restoring touches only inputs the step changed."""

    __slots__ = ("lookup", "assign", "ptr", "version", "copy")

    def __init__(self, lookup, assign, t: Optional[torch.Tensor]) -> None:
        self.lookup, self.assign = lookup, assign
        self.ptr = 0 if t is None else t.data_ptr()
        self.version = 0 if t is None else t._version
        self.copy = None if t is None else t.clone()

    def unchanged(self, t: Optional[torch.Tensor]) -> bool:
        return t is None or (t.data_ptr() == self.ptr and t._version == self.version)


class EnvStateSnapshot:
    """The state of a subset of a batch's environments, saved before a substep those
    environments sit out and written back after it (batching.py:178-325).

    Saved per environment: every block's velocity, pressure and potential, the domain's
    solution vectors (the next solve's warm start and the potential the Lorentz force reads),
    any ``extra`` per-environment tensors (given as zero-argument lookups), and -- only if the
    step replaces or writes them -- the blocks' velocity sources and the fixed faces'
    velocities. Tensors are looked up again when restoring, since a step may swap a tensor for
    a new one. A shared input (batch 1) the step changed comes back per-environment: the other
    environments keep the new value. Dirichlet potential values of fixed faces are one array
    for all environments here and are not saved.
    """

    def __init__(self, domain: Domain, env_idx: torch.Tensor,
                 extra: Sequence[Callable[[], Optional[torch.Tensor]]] = ()) -> None:
        self._domain = domain
        self._batch = domain.getBatchSize()
        self._idx = env_idx.to(device=domain.getDevice(), dtype=torch.long)
        with torch.no_grad():
            self._rows = [_SavedRows(lookup, self._per_env(t).index_select(0, self._idx))
                          for lookup in self._state_lookups(domain, extra)
                          for t in (lookup(),) if t is not None and t.numel() > 0]
            self._inputs = list(self._input_records(domain))

    @staticmethod
    def _state_lookups(domain: Domain, extra) -> list:
        lookups = []
        for blk in domain.getBlocks():
            for name in ("velocity", "pressure", "epot"):
                lookups.append(lambda blk=blk, name=name: getattr(blk, name))
        for name in ("velocityResult", "pressureResult", "epotResult"):
            lookups.append(lambda name=name: getattr(domain, name))
        return lookups + list(extra)

    @staticmethod
    def _input_records(domain: Domain):
        for blk in domain.getBlocks():
            yield _SavedInput(lambda blk=blk: blk.velocitySource, blk.setVelocitySource, blk.velocitySource)
            for face in blk.getBoundaries():
                if isinstance(face, FixedBoundary):
                    yield _SavedInput(lambda face=face: face.velocity, face.setVelocity, face.velocity)

    def _per_env(self, t: torch.Tensor) -> torch.Tensor:
        if t.numel() % self._batch:
            raise ValueError(f"A tensor of {t.numel()} elements cannot be split into {self._batch} environments.")
        return t.view(self._batch, -1)

    def restore(self) -> None:
        """Write the saved state back into the saved environments."""
        with torch.no_grad():
            for rec in self._rows:
                t = rec.lookup()
                if t is not None and t.numel() > 0:
                    self._per_env(t).index_copy_(0, self._idx, rec.rows)
            replaced = [rec for rec in self._inputs if not rec.unchanged(rec.lookup())]
            for rec in replaced:
                self._put_back(rec, rec.lookup())
        if replaced:
            self._domain.UpdateDomainData()

    def _put_back(self, rec: _SavedInput, now: torch.Tensor) -> None:
        """The changed input ``now`` with the saved environments' rows restored (an input that
        did not exist when saving restores as zeros)."""
        old = rec.copy if rec.copy is not None else torch.zeros_like(now[:1])
        if old.shape[1:] != now.shape[1:]:
            raise ValueError(f"An input changed its shape during the step: {tuple(old.shape)} -> {tuple(now.shape)}.")
        full = now.expand(self._batch, *now.shape[1:]).clone()
        full.index_copy_(0, self._idx, old.expand(self._batch, *old.shape[1:]).index_select(0, self._idx))
        rec.assign(full)

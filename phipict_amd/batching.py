"""Per-environment state of a batched domain, saved before a substep and written back
afterwards for the environments that sit it out (src/phipict/batching.py:178-320)."""
from __future__ import annotations

from typing import Callable, Optional, Sequence

import torch

from .domain import Domain, FixedBoundary


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

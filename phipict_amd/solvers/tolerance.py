"""Relative/absolute solver tolerances (src/phipict/solvers/tolerance.py:116-354)."""
from __future__ import annotations

import math
import numbers
from dataclasses import dataclass
from typing import Any, Optional

import torch

#: the resolved tolerance is floored at this many machine epsilons of the RHS scale
PRECISION_FLOOR_FACTOR = 8.0


def default_tolerance(dtype) -> Optional[float]:
    """Default absolute tolerance: 1e-5 for float32, 1e-8 for float64 (piso_diff.py:396-421)."""
    if dtype == torch.float64:
        return 1e-8
    if dtype == torch.float32:
        return 1e-5
    return None


def rhs_scale(rhs: torch.Tensor) -> float:
    """``||b||_2 / sqrt(n)`` accumulated in float64 (tolerance.py:116-138)."""
    n = rhs.numel()
    if n == 0:
        return 0.0
    return float(torch.linalg.vector_norm(rhs.detach().to(torch.float64)).item()) / math.sqrt(n)


@dataclass(frozen=True)
class SolverTolerance:
    """Tolerance relative to the right-hand side: ``max(rtol * ||b||/sqrt(n), atol,
    8 eps ||b||/sqrt(n))`` (tolerance.py:142-265)."""

    rtol: Optional[float] = None
    atol: Optional[float] = None

    def __post_init__(self) -> None:
        if self.rtol is None and self.atol is None:
            raise ValueError("SolverTolerance needs at least one of rtol, atol.")
        for name in ("rtol", "atol"):
            v = getattr(self, name)
            if v is None:
                continue
            if not isinstance(v, numbers.Real):
                raise TypeError(f"SolverTolerance.{name} must be a float or None.")
            if not float(v) > 0:
                raise ValueError(f"SolverTolerance.{name} must be positive.")

    def _resolve_scale(self, scale: float, dtype) -> float:
        tol = 0.0
        if self.rtol is not None:
            tol = self.rtol * scale
        if self.atol is not None:
            tol = max(tol, self.atol)
        tol = max(tol, PRECISION_FLOOR_FACTOR * torch.finfo(dtype).eps * scale)
        if not tol > 0.0:
            tol = default_tolerance(dtype)
        return float(tol)

    def resolve(self, rhs: torch.Tensor) -> float:
        return self._resolve_scale(rhs_scale(rhs), rhs.dtype)

    def resolve_batched(self, rhs: torch.Tensor, batch_size: int) -> list:
        slices = rhs.detach().reshape(batch_size, -1)
        n = slices.shape[1]
        norms = torch.linalg.vector_norm(slices.to(torch.float64), dim=1)
        scales = (norms / math.sqrt(n)).tolist() if n > 0 else [0.0] * batch_size
        return [self._resolve_scale(s, rhs.dtype) for s in scales]

    def __str__(self) -> str:
        parts = []
        if self.rtol is not None:
            parts.append(f"rtol={self.rtol:.03e}")
        if self.atol is not None:
            parts.append(f"atol={self.atol:.03e}")
        return f"SolverTolerance({', '.join(parts)})"


def resolve_tolerance(tol: Any, rhs: torch.Tensor, batch_size: int = 1) -> Any:
    """Turn a :class:`SolverTolerance` into the absolute value(s) the solvers take; other
    values pass through unchanged (tolerance.py:326-354)."""
    if isinstance(tol, SolverTolerance):
        if batch_size > 1:
            return tol.resolve_batched(rhs, batch_size)
        return tol.resolve(rhs)
    return tol

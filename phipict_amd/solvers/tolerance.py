"""Solver tolerances given relative to the right-hand side.

The solvers stop on ``||r||_2 / sqrt(n) < tol`` (NORM2_NORMALIZED). A :class:`SolverTolerance`
names that bound relative to the same norm of the right-hand side ``b`` and resolves it, per
environment, to

    tol = max(rtol * s, atol, 8 * eps(dtype) * s),    s = ||b||_2 / sqrt(n)

falling back to the dtype default when the result is not positive (a zero right-hand side);
plain floats are absolute tolerances and pass through unchanged. The formula and the
defaults follow src/phipict/solvers/tolerance.py:116-354 and piso_diff.py:396-421.

Every resolution goes through one vectorised path: the per-environment scales are computed in
float64 on the device in one reduction, moved to the host once, and bounded elementwise.
"""
from __future__ import annotations

import math
import numbers
from dataclasses import dataclass
from typing import Any, List, Mapping, Optional

import numpy as np
import torch

__all__ = ["SolverTolerance", "default_tolerance", "parse_tolerance", "resolve_tolerance", "rhs_scale",
           "PRECISION_FLOOR_FACTOR"]

#: resolved tolerances never go below this many machine epsilons of the right-hand side's scale
PRECISION_FLOOR_FACTOR = 8.0

_DTYPE_DEFAULTS = {torch.float32: 1e-5, torch.float64: 1e-8}


def default_tolerance(dtype) -> Optional[float]:
    """The absolute tolerance a solve uses when none is given: 1e-5 in float32, 1e-8 in
    float64, None for other dtypes (piso_diff.py:396-421)."""
    return _DTYPE_DEFAULTS.get(dtype)


def _scales(rhs: torch.Tensor, parts: int) -> np.ndarray:
    """``||b_k||_2 / sqrt(m)`` of the ``parts`` equal slices ``b_k`` of ``rhs`` (m elements
    each), accumulated in float64, as one host array."""
    flat = rhs.detach().reshape(parts, -1)
    m = flat.shape[1]
    if m == 0:
        return np.zeros(parts)
    norms = torch.linalg.vector_norm(flat.to(torch.float64), dim=1)
    return norms.cpu().numpy() / math.sqrt(m)


def rhs_scale(rhs: torch.Tensor) -> float:
    """The criterion norm of the right-hand side, ``||b||_2 / sqrt(n)`` in float64
    (tolerance.py:116-138)."""
    if rhs.numel() == 0:
        return 0.0
    return float(_scales(rhs, 1)[0])


def _check_bound(kind: str, value) -> None:
    if value is None:
        return
    if isinstance(value, bool) or not isinstance(value, numbers.Real):
        raise TypeError(f"{kind} of a SolverTolerance is a real number or None, got {type(value).__name__}.")
    if not float(value) > 0.0:
        raise ValueError(f"{kind} of a SolverTolerance must be > 0, got {value!r}.")


@dataclass(frozen=True)
class SolverTolerance:
    """A stopping bound relative to the right-hand side (``rtol``), with an absolute floor
    (``atol``); at least one of them is required (tolerance.py:142-265)."""

    rtol: Optional[float] = None
    atol: Optional[float] = None

    def __post_init__(self) -> None:
        if self.rtol is None and self.atol is None:
            raise ValueError("A SolverTolerance needs rtol, atol or both.")
        _check_bound("rtol", self.rtol)
        _check_bound("atol", self.atol)

    def _bound(self, scales: np.ndarray, dtype) -> List[float]:
        """The resolved tolerance for each right-hand-side scale."""
        floor = PRECISION_FLOOR_FACTOR * torch.finfo(dtype).eps * scales
        tol = np.maximum(floor, 0.0 if self.rtol is None else float(self.rtol) * scales)
        if self.atol is not None:
            tol = np.maximum(tol, float(self.atol))
        fallback = default_tolerance(dtype)
        return [float(t) if t > 0.0 else fallback for t in tol]

    def resolve(self, rhs: torch.Tensor) -> float:
        """One absolute tolerance for the whole right-hand side."""
        scale = np.array([rhs_scale(rhs)])
        return self._bound(scale, rhs.dtype)[0]

    def resolve_batched(self, rhs: torch.Tensor, batch_size: int) -> list:
        """One absolute tolerance per environment: ``rhs`` is ``batch_size`` equal slices."""
        return self._bound(_scales(rhs, batch_size), rhs.dtype)

    def __str__(self) -> str:
        given = [(k, v) for k, v in (("rtol", self.rtol), ("atol", self.atol)) if v is not None]
        return "SolverTolerance(" + ", ".join(f"{k}={v:.03e}" for k, v in given) + ")"


def parse_tolerance(spec: Any) -> Any:
    """A tolerance from a configuration value: None and numbers pass through, a mapping with
    the keys ``rtol`` / ``atol`` becomes a :class:`SolverTolerance` (other keys raise
    ``ValueError``), a :class:`SolverTolerance` is returned as it is (tolerance.py:268-323)."""
    if spec is None or isinstance(spec, SolverTolerance):
        return spec
    if isinstance(spec, numbers.Real) and not isinstance(spec, bool):
        return float(spec)
    if isinstance(spec, Mapping) or hasattr(spec, "keys"):  # config nodes behave like mappings
        unknown = set(spec.keys()) - {"rtol", "atol"}
        if unknown:
            raise ValueError(f"Unknown tolerance keys {sorted(unknown)}; allowed are 'rtol' and 'atol'.")
        bounds = {k: (None if spec.get(k) is None else float(spec.get(k))) for k in ("rtol", "atol")}
        return SolverTolerance(**bounds)
    raise TypeError(f"Cannot read a tolerance from {type(spec).__name__}.")


def resolve_tolerance(tol: Any, rhs: torch.Tensor, batch_size: int = 1) -> Any:
    """The absolute value(s) a solve takes: a :class:`SolverTolerance` resolves against
    ``rhs`` (a list with one entry per environment when ``batch_size > 1``); anything else is
    returned unchanged (tolerance.py:326-354)."""
    if not isinstance(tol, SolverTolerance):
        return tol
    return tol.resolve_batched(rhs, batch_size) if batch_size > 1 else tol.resolve(rhs)

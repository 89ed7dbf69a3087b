"""Typed boundary-condition helpers (src/phipict/bc.py:71-506)."""
from __future__ import annotations

from dataclasses import dataclass
from enum import IntEnum
from typing import Optional

import torch

from .domain import Block, FixedBoundary, PotentialBC


class Face(IntEnum):
    X_MINUS = 0
    X_PLUS = 1
    Y_MINUS = 2
    Y_PLUS = 3
    Z_MINUS = 4
    Z_PLUS = 5


class Potential:
    """Electric-potential face conditions."""

    @dataclass(frozen=True)
    class Insulating:
        pass

    @dataclass(frozen=True)
    class Open:
        pass

    @dataclass(frozen=True)
    class Dirichlet:
        value: Optional[torch.Tensor] = None

    @dataclass(frozen=True)
    class ThinWall:
        cw: float = 0.0


def set_bc(block: Block, face, condition) -> FixedBoundary:
    """Apply a potential condition to a prescribed face of ``block`` (bc.py:464-506)."""
    bound = block.getBoundary(int(face))
    if not isinstance(bound, FixedBoundary):
        raise ValueError("potential conditions apply to prescribed (FIXED) faces; close the face first")
    if isinstance(condition, Potential.Insulating):
        bound.setPotentialBC(PotentialBC.INSULATING)
    elif isinstance(condition, Potential.Open):
        bound.setPotentialBC(PotentialBC.OPEN)
    elif isinstance(condition, Potential.Dirichlet):
        bound.setPotentialBC(PotentialBC.DIRICHLET)
        if condition.value is not None:
            v = torch.as_tensor(condition.value, dtype=block.getParentDomain().getDtype())
            if v.numel() == 1:
                v = v.reshape(1).expand(bound.face_cells())
            bound.setPotentialValues(v)
    elif isinstance(condition, Potential.ThinWall):
        bound.setPotentialBC(PotentialBC.THIN_WALL, cw=condition.cw)
    else:
        raise TypeError(f"unknown potential condition {condition!r}")
    return bound


def set_potential_values(block: Block, face, values: torch.Tensor) -> None:
    bound = block.getBoundary(int(face))
    if not isinstance(bound, FixedBoundary):
        raise ValueError("potential values apply to prescribed faces")
    bound.setPotentialValues(values)

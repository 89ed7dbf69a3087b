"""Grid vertex generation for rectilinear (graded) blocks.

Weights follow src/phipict/grid/shapes.py (make_weights_tanh :938-993, make_weights :996-1036);
vertex tensors use the reference layout ``[1, d, (z+1), y+1, x+1]`` that
``Domain.CreateBlock(vertexCoordinates=...)`` takes.
"""
from __future__ import annotations

from typing import Optional, Sequence

import numpy as np
import torch


def make_weights_tanh(res: int, A: float, refinement: str) -> list:
    """``res + 1`` weights in [0, 1] with tanh clustering (shapes.py:938-993)."""
    if abs(A) < 1e-10:
        return np.linspace(0.0, 1.0, res + 1).tolist()
    if refinement == "START":
        xi = np.linspace(-1, 0, res + 1)
        w = np.tanh(A * xi) / np.tanh(A) + 1.0
    elif refinement == "END":
        xi = np.linspace(0, 1, res + 1)
        w = np.tanh(A * xi) / np.tanh(A)
    elif refinement == "BOTH":
        xi = np.linspace(-1, 1, res + 1)
        w = 0.5 * (np.tanh(A * xi) / np.tanh(A)) + 0.5
    else:
        raise ValueError(f"Unknown refinement '{refinement}'. Use 'START', 'END', or 'BOTH'.")
    w[0] = 0.0
    w[-1] = 1.0
    return w.tolist()


def _geometric_sizes(n: int, r: float) -> list:
    if abs(r - 1.0) < 1e-10:
        return [1.0 / n] * n
    q = r ** (1.0 / (n - 1))
    a = 1.0 / sum(q ** i for i in range(n))
    return [a * q ** i for i in range(n)]


def make_weights_simple_grading(res: int, grading: float, refinement: str) -> list:
    """OpenFOAM-style simpleGrading weights (shapes.py:795-867)."""
    if refinement == "START":
        sizes = _geometric_sizes(res, grading)
    elif refinement == "END":
        sizes = _geometric_sizes(res, grading)[::-1]
    elif refinement == "BOTH":
        half = res // 2
        sizes = _geometric_sizes(half, grading) + _geometric_sizes(res - half, grading)[::-1]
    else:
        raise ValueError(f"Unknown refinement '{refinement}'.")
    total = sum(sizes)
    sizes = [s / total for s in sizes]
    return [0.0] + list(np.cumsum(sizes))


def make_weights(grading_type: str, res: int, grading: float, refinement: str) -> list:
    if grading_type == "tanh":
        return make_weights_tanh(res, grading, refinement)
    if grading_type == "simple":
        return make_weights_simple_grading(res, grading, refinement)
    if grading_type == "linear":
        return np.linspace(0.0, 1.0, res + 1).tolist()
    raise ValueError(f"Unknown grading type '{grading_type}'.")


def rectilinear_vertices(axes: Sequence[Sequence[float]], dtype=torch.float64, device=None) -> torch.Tensor:
    """Vertex coordinates ``[1, d, (z+1), y+1, x+1]`` of the tensor-product grid of the given
    per-axis coordinates ``[xs, ys, (zs)]``."""
    dims = len(axes)
    arrs = [torch.as_tensor(np.asarray(a, dtype=np.float64)) for a in axes]
    if dims == 2:
        Y, X = torch.meshgrid(arrs[1], arrs[0], indexing="ij")
        v = torch.stack([X, Y])
    elif dims == 3:
        Z, Y, X = torch.meshgrid(arrs[2], arrs[1], arrs[0], indexing="ij")
        v = torch.stack([X, Y, Z])
    else:
        raise ValueError("2 or 3 axes are supported")
    return v.unsqueeze(0).to(dtype=dtype, device=device).contiguous()


def axis_coords(start: float, end: float, weights: Sequence[float]) -> list:
    return [start * (1 - w) + end * w for w in weights]


def cell_centers(vertices: torch.Tensor) -> torch.Tensor:
    """Cell centres ``[d, (z), y, x]`` of a vertex tensor (mean of the cell's corners)."""
    v = vertices[0]
    dims = v.size(0)
    if dims == 2:
        return 0.25 * (v[:, :-1, :-1] + v[:, 1:, :-1] + v[:, :-1, 1:] + v[:, 1:, 1:])
    return 0.125 * (v[:, :-1, :-1, :-1] + v[:, 1:, :-1, :-1] + v[:, :-1, 1:, :-1] + v[:, :-1, :-1, 1:]
                    + v[:, 1:, 1:, :-1] + v[:, 1:, :-1, 1:] + v[:, :-1, 1:, 1:] + v[:, 1:, 1:, 1:])

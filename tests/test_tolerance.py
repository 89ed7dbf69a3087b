"""Tolerance resolution on the host (no GPU): the closed form max(rtol*s, atol, 8 eps s),
s = ||b||/sqrt(n), checked on right-hand sides whose scale is known exactly, and the
configuration parser; the cases follow tests/test_solver_tolerance.py:39-95 of the
reference."""
import math

import pytest
import torch

from phipict_amd.solvers.tolerance import (SolverTolerance, default_tolerance, parse_tolerance, resolve_tolerance,
                                           rhs_scale)


def test_rhs_scale_is_the_criterion_norm():
    assert rhs_scale(torch.full((1000,), 3.0, dtype=torch.float64)) == pytest.approx(3.0)
    # a vector with entries +-2 and zeros: ||b|| / sqrt(n) = 2 sqrt(k / n)
    b = torch.zeros(400, dtype=torch.float32)
    b[::4] = 2.0
    b[1::4] = -2.0
    assert rhs_scale(b) == pytest.approx(2.0 * math.sqrt(200 / 400))
    assert rhs_scale(torch.zeros(0)) == 0.0


def test_relative_and_absolute_bounds():
    b = torch.full((1000,), 3.0, dtype=torch.float64)
    assert SolverTolerance(rtol=1e-6).resolve(b) == pytest.approx(3e-6)
    assert SolverTolerance(rtol=1e-12, atol=1e-9).resolve(b) == pytest.approx(1e-9)
    assert SolverTolerance(atol=1e-7).resolve(b) == pytest.approx(1e-7)
    small = torch.ones(1000, dtype=torch.float64)
    tol = SolverTolerance(rtol=1e-6)
    assert tol.resolve(100.0 * small) == pytest.approx(100.0 * tol.resolve(small))


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_round_off_floor_and_zero_rhs(dtype):
    b = torch.full((1000,), 3.0, dtype=dtype)
    eps = torch.finfo(dtype).eps
    assert SolverTolerance(rtol=1e-30).resolve(b) == pytest.approx(8 * eps * 3.0)
    assert SolverTolerance(rtol=1e-6).resolve(torch.zeros(10, dtype=dtype)) == default_tolerance(dtype)


def test_batched_resolution_per_environment():
    b = torch.cat([torch.full((50,), 1.0), torch.full((50,), 4.0), torch.zeros(50)]).to(torch.float64)
    tols = SolverTolerance(rtol=1e-3, atol=2e-3).resolve_batched(b, 3)
    assert tols == pytest.approx([2e-3, 4e-3, 2e-3])
    assert resolve_tolerance(SolverTolerance(rtol=1e-3), b, batch_size=3) == pytest.approx([1e-3, 4e-3, 1e-8])


def test_parse_and_pass_through():
    assert parse_tolerance(None) is None
    assert parse_tolerance(1e-5) == 1e-5
    assert parse_tolerance({"rtol": 1e-6, "atol": 1e-12}) == SolverTolerance(rtol=1e-6, atol=1e-12)
    with pytest.raises(ValueError):
        parse_tolerance({"rtol": 1e-6, "nonsense": 1.0})
    with pytest.raises(TypeError):
        parse_tolerance("tight")
    with pytest.raises(ValueError):
        SolverTolerance()
    with pytest.raises(ValueError):
        SolverTolerance(rtol=-1.0)
    b = torch.ones(10, dtype=torch.float64)
    assert resolve_tolerance(1e-5, b) == 1e-5
    assert resolve_tolerance(None, b) is None
    assert str(SolverTolerance(rtol=1e-6)) == "SolverTolerance(rtol=1.000e-06)"

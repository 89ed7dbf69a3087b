"""The allocation-free solve entry point, pp_solve_linear_noalloc, called through the C ABI:
two calls in a row (a large system on the multi-workgroup engine, then a smaller one on the
on-chip engine, each with a return-best CG or a BiCGStab) give the same solutions and results
as pp_solve_linear, bit for bit (both run the same kernels on the same grids, whose reductions
are deterministic); a system beyond the scratch maximum is refused with a clear error."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda", 0)


def _poisson(nx, ny, shift):
    """5-point operator of an nx x ny grid with Dirichlet walls, diagonal shifted by shift."""
    n = nx * ny
    idx = np.arange(n).reshape(ny, nx)
    rows, cols, vals = [], [], []
    for dy, dx in ((0, 0), (0, -1), (0, 1), (-1, 0), (1, 0)):
        dst = idx[max(0, dy):ny + min(0, dy), max(0, dx):nx + min(0, dx)]
        src = idx[max(0, -dy):ny + min(0, -dy), max(0, -dx):nx + min(0, -dx)]
        rows.append(src.reshape(-1))
        cols.append(dst.reshape(-1))
        vals.append(np.full(src.size, 4.0 + shift if (dx, dy) == (0, 0) else -1.0))
    r, c, v = (np.concatenate(a) for a in (rows, cols, vals))
    order = np.lexsort((c, r))
    r, c, v = r[order], c[order], v[order]
    row = np.concatenate([[0], np.cumsum(np.bincount(r, minlength=n))]).astype(np.int32)
    return row, c.astype(np.int32), v


def _desc(row, col, val):
    from phipict_amd import _lib

    d = _lib.CsrDesc()
    d.row, d.col, d.val = row.data_ptr(), col.data_ptr(), val.data_ptr()
    d.n, d.nnz, d.num_matrices = row.numel() - 1, col.numel(), 1
    return d


def _call(name, d, b, x, nrhs, tol, use_bicg, return_best):
    from phipict_amd import _lib

    lib = _lib.LIB
    fn = getattr(lib, name)
    fn.restype = C.c_int
    res = (_lib.SolveResult * nrhs)()
    tols = (C.c_double * 1)(tol)
    stream = C.c_void_p(torch.cuda.current_stream(DEV).cuda_stream)
    rc = fn(C.byref(d), C.c_void_p(b.data_ptr()), C.c_void_p(x.data_ptr()), nrhs, _lib.PP_F64, 4000, tols, 1,
            _lib.PP_CONV_NORM2_NORMALIZED, int(use_bicg), 0, 0, 0, int(return_best), stream, res)
    torch.cuda.synchronize()
    return rc, [(r.final_residual, r.used_iterations, r.converged, r.is_finite_residual, r.pad) for r in res]


@pytest.mark.parametrize("use_bicg", [False, True], ids=["cg", "bicgstab"])
def test_noalloc_matches_solve_linear(use_bicg):
    print(f"[noalloc] use_bicg={use_bicg}", flush=True)
    for nx, ny, nrhs in ((240, 180, 2), (40, 30, 3)):  # 43200 rows (multi-workgroup), then 1200 (on-chip)
        row, col, val = _poisson(nx, ny, 0.02)
        if use_bicg:  # a non-symmetric operator: upwind-like skew on the x neighbours
            r = np.repeat(np.arange(row.size - 1), np.diff(row))
            val = np.where(col == r - 1, val * 1.3, np.where(col == r + 1, val * 0.7, val))
        t = [torch.as_tensor(a, device=DEV) for a in (row, col, val)]
        d = _desc(*t)
        n = row.size - 1
        b = torch.as_tensor(np.random.default_rng(nx).standard_normal(nrhs * n), device=DEV)
        x_ref = torch.zeros(nrhs * n, dtype=torch.float64, device=DEV)
        x_new = torch.zeros_like(x_ref)
        rc1, res1 = _call("pp_solve_linear", d, b, x_ref, nrhs, 1e-9, use_bicg, not use_bicg)
        rc2, res2 = _call("pp_solve_linear_noalloc", d, b, x_new, nrhs, 1e-9, use_bicg, not use_bicg)
        assert rc1 == 0 and rc2 == 0
        print(f"  n={n}: {res2}", flush=True)
        assert all(r[2] for r in res2)
        assert res1 == res2
        assert torch.equal(x_ref, x_new)


def test_noalloc_refuses_beyond_scratch():
    from phipict_amd import _lib

    n = 1_300_000  # seven work vectors of n doubles exceed the 56 MiB of work-vector scratch
    row = torch.arange(n + 1, dtype=torch.int32, device=DEV)
    col = torch.arange(n, dtype=torch.int32, device=DEV)
    val = torch.ones(n, dtype=torch.float64, device=DEV)
    b = torch.ones(n, dtype=torch.float64, device=DEV)
    x = torch.zeros(n, dtype=torch.float64, device=DEV)
    rc, _ = _call("pp_solve_linear_noalloc", _desc(row, col, val), b, x, 1, 1e-8, False, False)
    assert rc != 0
    msg = _lib.LIB.pp_last_error
    msg.restype = C.c_char_p
    assert b"scratch maximum" in msg()
    assert float(x.abs().max()) == 0.0  # nothing ran

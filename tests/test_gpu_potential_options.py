"""Options of the MHD potential solve (src/phipict/simulation/mhd.py:457-486, 940-949):
``potential_use_BiCG`` solves the potential with BiCGStab instead of CG, and
``potential_solve_dtype`` solves it in another dtype (the operator and right-hand side are cast
before the solve, the solution cast back). A 2D channel in an oblique field (a non-trivial
potential), two PISO + MHD steps."""
from __future__ import annotations

import numpy as np
import pytest
import torch

import cases
from cases import build_domain, rel_l2, to_np

pytestmark = pytest.mark.gpu


def _run(dtype, steps=2, **kw):
    import phipict_amd as pa

    specs = cases.channel2d(nx=24, ny=16, gradient=4.0, noise=1e-2, seed=3)
    dom = build_domain(specs, 0.05, 5e-3, dtype)
    src = dom.getBlocks()[0].velocitySource.clone()

    def forcing(domain, **k):
        domain.getBlocks()[0].setVelocitySource(src.clone())
        domain.UpdateDomainData()

    tol = kw.pop("tol", 1e-11)
    sim = pa.MHDSimulation(domain=dom, dt=5e-3, stuart_number=torch.tensor(4.0),
                           e_b=torch.tensor([0.0, 0.6, 0.8], dtype=dtype), non_orthogonal=False,
                           hooks={"PRE_VELOCITY_SETUP": forcing}, advection_tol=kw.pop("advection_tol", tol),
                           pressure_tol=kw.pop("pressure_tol", tol), potential_tol=kw.pop("potential_tol", tol), **kw)
    sim.record_solver_stats = True
    for _ in range(steps):
        assert sim.single_step()
    blk = dom.getBlocks()[0]
    phi = to_np(dom.epotResult).astype(np.float64)
    return to_np(blk.velocity).astype(np.float64), phi - phi.mean(), sim


def test_potential_bicgstab_equals_cg():
    print("[potential] BiCGStab vs CG, f64", flush=True)
    u_cg, phi_cg, _ = _run(torch.float64)
    u_bi, phi_bi, sim = _run(torch.float64, potential_use_BiCG=True)
    assert np.abs(phi_cg).max() > 1e-4
    assert rel_l2(phi_bi, phi_cg) < 1e-8
    assert rel_l2(u_bi, u_cg) < 1e-10
    assert any(tag == "epot" for _, tag, _ in sim.solver_log)


def test_potential_solve_dtype_float64_on_a_float32_domain():
    """An FP32 domain whose potential is solved in FP64 to an absolute tolerance of 1e-10 (below
    FP32's reach): the solve converges to it, and the FP32 potential equals the FP64 solution
    of the FP32 operator and right-hand side (cast up, solved by the oracle's projected CG to
    1e-12, cast back) to FP32 rounding."""
    print("[potential] potential_solve_dtype=float64 on float32", flush=True)
    _u32, phi32, sim = _run(torch.float32, advection_tol=1e-6, pressure_tol=1e-6, potential_tol=1e-10,
                            potential_solve_dtype=torch.float64)
    dom = sim.domain
    assert dom.epotResult.dtype == torch.float32
    last = [infos for _, tag, infos in sim.solver_log if tag == "epot"][-1]
    assert all(conv and res <= 1e-10 for _it, res, conv in last), last
    from oracle import oracle as O

    E = dom.Epot
    row, col = to_np(E.row).astype(np.int32), to_np(E.index).astype(np.int32)
    val = to_np(E.value).astype(np.float64)
    b = to_np(dom.epotRHS).astype(np.float64)
    x = np.zeros_like(b)
    ref = O.cg(row, col, val, b.copy(), x, 1e-12, project=True)
    assert ref.converged
    x -= x.mean()
    assert rel_l2(phi32, x) < 1e-6

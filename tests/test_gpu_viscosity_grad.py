"""Gradients with respect to the viscosity through the differentiable step: the domain's value
and a block's own (static or per cell) are tracked inputs of the advection matrix and of the
momentum and pressure right-hand sides (their wall terms), as in the reference (VISCOSITY and
VISCOSITY_BLOCK, src/phipict/core/piso_diff.py:1458-1466, 1811-1817, 2771-2779, 3158-3166).
All three operators are affine in a per-cell viscosity, so their adjoints pass dot-product
tests to rounding; the gradient of a loss after two MHD steps with respect to the domain's
viscosity and to a per-cell block viscosity matches central differences. FP64."""
from __future__ import annotations

import numpy as np
import pytest
import torch

import cases
from cases import build_domain

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda", 0)
F64 = torch.float64
DT = 2e-2


def _dot(a, b):
    return float((a.reshape(-1).to(F64) * b.reshape(-1).to(F64)).sum())


def test_operators_adjoint_in_a_per_cell_viscosity():
    """The closed box with walls of every kind, a per-cell block viscosity and moving faces."""
    from phipict_amd import adjoint as adj
    from phipict_amd import ops

    print("[viscosity] dot-product tests of the three operators", flush=True)
    dom = build_domain(cases.potential_box3d(nx=6, ny=5, nz=4), 0.02, DT, F64, epot=True)
    blk = dom.getBlocks()[0]
    dt = torch.tensor([DT], dtype=F64)
    gen = torch.Generator("cpu").manual_seed(3)
    rnd = lambda like, s=1.0: (s * torch.randn(like.shape, generator=gen, dtype=F64)).to(like.device)  # noqa: E731
    nu0 = blk.viscosity.detach().clone()
    dnu = rnd(nu0, 0.01)
    vels = [b.velocity for b in dom.getBlocks()]
    bvels = [bd.velocity for _, _, bd in adj.fixed_faces(dom)]
    srcs = [b.velocitySource for b in dom.getBlocks()]
    nv, nb = len(vels), len(bvels)

    def set_nu(nu):
        blk.setViscosity(nu)
        dom.UpdateDomainData()
        return adj.viscosity_inputs(dom)

    ops.SetupAdvectionMatrix(dom, dt, 0)
    A0, C0 = dom.A.detach().clone(), dom.C.value.detach().clone()
    V0 = rnd(dom.velocityResult)
    operators = {
        "advection matrix": lambda viscs: list(adj.AdvectionMatrixFunction.apply(dom, dt, *vels, *viscs)),
        "momentum rhs": lambda viscs: [adj.AdvectionRHSFunction.apply(dom, dt, nv, nb, *vels, *bvels, *srcs, *viscs)],
    }

    def pressure_rhs(viscs):
        dom.setA(A0.clone())
        dom.C.value.copy_(C0)
        dom.setVelocityResult(V0.clone())
        dom.UpdateDomainData()
        return [adj.PressureRHSFunction.apply(dom, dt, False, A0, C0, V0, nv, nb, *vels, *bvels, *srcs, *viscs)[0]]

    operators["pressure rhs"] = pressure_rhs
    for name, op in operators.items():
        with torch.no_grad():
            plus = [t.detach().clone() for t in op(set_nu(nu0 + dnu))]
            base = [t.detach().clone() for t in op(set_nu(nu0))]
        ws = [rnd(t) for t in base]
        lhs = sum(_dot(w, p - b) for w, p, b in zip(ws, plus, base))
        nu = nu0.clone().requires_grad_(True)
        outs = op(set_nu(nu))
        (g,) = torch.autograd.grad(outs, nu, ws)
        rhs = _dot(g, dnu)
        print(f"  {name}: <w, F(nu+v)-F(nu)> {lhs:.12e}  <J^T w, v> {rhs:.12e}", flush=True)
        assert abs(lhs) > 1e-8, name
        assert abs(lhs - rhs) <= 1e-10 * (abs(lhs) + abs(rhs)), (name, lhs, rhs)
    set_nu(nu0)


def _loss(nu_dom, nu_blk=None, differentiable=False):
    import phipict_amd as pa

    dom = build_domain(cases.outflow2d(nx=14, ny=8, seed=7), 0.05, DT, F64)
    dom.setViscosity(nu_dom)
    blk = dom.getBlocks()[0]
    if nu_blk is not None:
        blk.setViscosity(nu_blk)
    pa.simulation.balance_boundary_fluxes(dom, [blk.getBoundary(1)])
    dom.UpdateDomainData()
    src = torch.tensor([[0.5, 0.1]], dtype=F64, device=DEV)

    def reset(domain, **kw):
        domain.getBlocks()[0].setVelocitySource(src)
        domain.UpdateDomainData()

    tol = pa.SolverTolerance(rtol=1e-13)
    sim = pa.MHDSimulation(domain=dom, dt=DT, stuart_number=torch.tensor(3.0, dtype=F64),
                           e_b=torch.tensor([0.0, 0.6, 0.8], dtype=F64), non_orthogonal=False, advection_tol=tol,
                           pressure_tol=tol, potential_tol=tol, differentiable=differentiable,
                           hooks={"PRE_VELOCITY_SETUP": reset}, flux_balance_tol=1e-3)
    for _ in range(2):
        assert sim.single_step()
    rng = np.random.default_rng(21)
    w_u = torch.as_tensor(rng.standard_normal(tuple(blk.velocity.shape)), dtype=F64, device=DEV)
    w_p = torch.as_tensor(rng.standard_normal(tuple(blk.pressure.shape)), dtype=F64, device=DEV)
    return (w_u * blk.velocity).sum() + 10.0 * (w_p * blk.pressure).sum()


@pytest.mark.parametrize("which", ["domain", "block per cell"])
def test_viscosity_gradient_through_mhd_steps(which):
    print(f"[viscosity] d loss / d {which} viscosity through two MHD steps", flush=True)
    nu_dom = torch.tensor([0.05], dtype=F64)
    cells = cases.outflow2d(nx=14, ny=8, seed=7)[0].cells_zyx
    nu_blk = None
    if which == "block per cell":
        nu_blk = (0.04 + 0.02 * torch.rand(int(np.prod(cells)), generator=torch.Generator("cpu").manual_seed(4),
                                           dtype=F64)).to(DEV)
    target = nu_dom if which == "domain" else nu_blk
    direction = torch.ones_like(target) if which == "domain" else \
        torch.randn(target.shape, generator=torch.Generator("cpu").manual_seed(5), dtype=F64).to(DEV)
    x = target.clone().requires_grad_(True)
    loss = _loss(x, None, True) if which == "domain" else _loss(nu_dom, x, True)
    (g,) = torch.autograd.grad(loss, x)
    ad = float((g.to(direction.device) * direction).sum())
    eps = 1e-6
    with torch.no_grad():
        if which == "domain":
            fd = float((_loss(target + eps * direction) - _loss(target - eps * direction)) / (2 * eps))
        else:
            fd = float((_loss(nu_dom, target + eps * direction) - _loss(nu_dom, target - eps * direction)) / (2 * eps))
    print(f"  AD {ad:.10e} FD {fd:.10e}", flush=True)
    assert abs(fd) > 1e-4
    assert abs(ad - fd) <= 1e-5 * abs(fd), (ad, fd)

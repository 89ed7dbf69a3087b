"""Dot-product tests of the adjoint kernels (pp_*_grad, through the differentiable operators of
phipict_amd/adjoint.py): for every operator y = F(x) of the step, random directions v of its
inputs and weights w of its outputs satisfy <w, J v> = <J^T w, v>, with J v from a central
difference of the forward kernels and J^T w from the adjoint kernels (autograd). Operators
affine in their inputs (the advection matrix and right-hand side in the velocities, face
velocities and sources; the potential's right-hand side; the current density; the velocity
correction in HbyA and p) agree to rounding; the others (1/A in the pressure matrix, HbyA and
the correction; C times V in HbyA) to the difference's O(eps^2). FP64, a 2D channel with a
per-face-cell inflow and outflow and a 3D two-block pair with a crosswise connection."""
from __future__ import annotations

import numpy as np
import pytest
import torch

import cases
from cases import build_domain

pytestmark = pytest.mark.gpu

F64 = torch.float64
DT = 2e-2


def _domain(case):
    specs = cases.outflow2d(nx=12, ny=8, seed=3) if case == "outflow2d" else cases.rotated_pair3d()
    dom = build_domain(specs, 0.05, DT, F64, epot=True)
    return dom


def _dot(a, b):
    return float((a.reshape(-1).to(F64) * b.reshape(-1).to(F64)).sum())


def _check(op, xs, vs, eps, tol, what):
    """<w, (F(x + eps v) - F(x - eps v)) / (2 eps)> against <grad of <w, F(x)>, v>."""
    with torch.no_grad():
        plus = [t.detach().clone() for t in op([x + eps * v for x, v in zip(xs, vs)])]
        minus = [t.detach().clone() for t in op([x - eps * v for x, v in zip(xs, vs)])]
    gen = torch.Generator("cpu").manual_seed(17)
    ws = [torch.randn(p.shape, generator=gen, dtype=F64).to(p.device) for p in plus]
    lhs = sum(_dot(w, (p - m) / (2 * eps)) for w, p, m in zip(ws, plus, minus))
    xt = [x.detach().clone().requires_grad_(True) for x in xs]
    outs = op(xt)
    grads = torch.autograd.grad(outs, xt, ws, allow_unused=True)
    rhs = sum(_dot(g, v) for g, v in zip(grads, vs) if g is not None)
    print(f"  {what}: <w, Jv> {lhs:.12e}  <J^T w, v> {rhs:.12e}", flush=True)
    assert abs(lhs) > 1e-8, f"{what}: the output does not depend on the inputs"
    assert abs(lhs - rhs) <= tol * (abs(lhs) + abs(rhs)), (what, lhs, rhs)
    return lhs, rhs


def _rand(like, scale=1.0, seed=0):
    gen = torch.Generator("cpu").manual_seed(seed)
    return (scale * torch.randn(like.shape, generator=gen, dtype=F64)).to(like.device)


@pytest.mark.parametrize("case", ["outflow2d", "rotated3d"])
def test_adjoint_kernels_dot_product(case):
    from phipict_amd import adjoint as adj
    from phipict_amd import ops

    print(f"[adjoint-dot] {case}", flush=True)
    dom = _domain(case)
    blocks = dom.getBlocks()
    d = dom.getSpatialDims()
    dt = torch.tensor([DT], dtype=F64)
    faces = adj.fixed_faces(dom)
    u0 = [b.velocity.detach().clone() + _rand(b.velocity, 0.1, 1 + i) for i, b in enumerate(blocks)]
    bv0 = [bd.velocity.detach().clone() for _, _, bd in faces]
    s0 = [torch.tensor([[0.3, -0.2, 0.1][:d]], dtype=F64, device=dom.getDevice()) for _ in blocks]
    nv, nb = len(blocks), len(faces)

    def set_state(us, bvs=None, srcs=None):
        for b, u in zip(blocks, us):
            b.velocity = u.detach().clone()
        if bvs is not None:
            for (_, _, bd), v in zip(faces, bvs):
                bd.setVelocity(v.detach().clone())
        if srcs is not None:
            for b, s in zip(blocks, srcs):
                b.setVelocitySource(s.detach().clone())
        dom._changed = True
        dom.UpdateDomainData()

    # advection matrix: affine in the block velocities
    def op_cmat(xs):
        set_state(xs)
        return list(adj.AdvectionMatrixFunction.apply(dom, dt, *xs))

    vu = [_rand(u, 1.0, 11 + i) for i, u in enumerate(u0)]
    _check(op_cmat, u0, vu, 1e-3, 1e-10, "advection matrix (A, C) / velocities")

    # momentum right-hand side: affine in velocities, face velocities and sources
    def op_crhs(xs):
        us, bvs, srcs = xs[:nv], xs[nv:nv + nb], xs[nv + nb:]
        set_state(us, bvs, srcs)
        return [adj.AdvectionRHSFunction.apply(dom, dt, nv, nb, *us, *bvs, *srcs)]

    xs = u0 + bv0 + s0
    vs = vu + [_rand(v, 1.0, 31 + i) for i, v in enumerate(bv0)] + [_rand(s, 1.0, 41 + i) for i, s in enumerate(s0)]
    _check(op_crhs, xs, vs, 1e-3, 1e-10, "momentum rhs / velocities, face velocities, sources")

    # the base state of the pressure operators
    set_state(u0, bv0, s0)
    ops.SetupAdvectionMatrix(dom, dt, 0)
    A0, C0 = dom.A.detach().clone(), dom.C.value.detach().clone()
    V0 = _rand(dom.velocityResult, 1.0, 51)

    # pressure matrix: 1 / A
    def op_pmat(xs):
        dom.setA(xs[0].detach().clone())
        dom.UpdateDomainData()
        return [adj.PressureMatrixFunction.apply(dom, dt, xs[0])]

    _check(op_pmat, [A0], [A0 * _rand(A0, 0.3, 61)], 1e-5, 1e-7, "pressure matrix / A")

    # HbyA and its divergence: 1 / A, C values times the predicted velocity, block velocities
    def op_prhs(xs):
        A, Cv, V = xs[:3]
        us = xs[3:3 + nv]
        set_state(us)
        dom.setA(A.detach().clone())
        dom.C.value.copy_(Cv.detach())
        dom.setVelocityResult(V.detach().clone())
        dom.UpdateDomainData()
        return list(adj.PressureRHSFunction.apply(dom, dt, False, A, Cv, V, nv, 0, *us))

    xs = [A0, C0, V0] + u0
    dirs = [A0 * _rand(A0, 0.3, 71), _rand(C0, 0.1, 72), _rand(V0, 1.0, 73)] + [_rand(u, 1.0, 74 + i)
                                                                               for i, u in enumerate(u0)]
    names = ["A", "C", "V"] + [f"velocity {i}" for i in range(nv)]
    failed = []
    for out, label in ((0, "HbyA"), (1, "divergence")):
        for k, name in enumerate(names):
            vs = [dv if j == k else torch.zeros_like(dv) for j, dv in enumerate(dirs)]
            try:
                _check(lambda xs_, o=out: [op_prhs(xs_)[o]], xs, vs, 1e-5 if name == "A" else 1e-3,
                       1e-7 if name == "A" else 1e-10, f"{label} / {name}")
            except AssertionError as exc:
                failed.append(str(exc))
    assert not failed, failed

    # FD velocity correction: affine in HbyA and p, 1 / A
    H0 = _rand(dom.pressureRHS, 1.0, 81)
    p0 = _rand(dom.pressureResult, 1.0, 82)

    def op_corr(xs):
        return [adj.CorrectVelocityFunction.apply(dom, dt, False, *xs)]

    _check(op_corr, [H0, A0, p0], [_rand(H0, 1.0, 83), torch.zeros_like(A0), _rand(p0, 1.0, 84)], 1e-3, 1e-10,
           "velocity correction / HbyA, p")
    _check(op_corr, [H0, A0, p0], [torch.zeros_like(H0), A0 * _rand(A0, 0.3, 85), torch.zeros_like(p0)], 1e-5, 1e-7,
           "velocity correction / A")

    # the potential's right-hand side (linear) and the current density (affine)
    f0 = _rand(dom.velocityResult, 1.0, 91)
    _check(lambda xs: [adj.EpotRHSFunction.apply(dom, xs[0])], [f0], [_rand(f0, 1.0, 92)], 1e-3, 1e-10,
           "potential rhs / u x e_b")
    n = dom.getTotalSize()
    phi0 = torch.zeros(n, dtype=F64, device=dom.getDevice()) + _rand(torch.zeros(n), 1.0, 93).to(dom.getDevice())
    x0 = _rand(torch.zeros(3 * n), 1.0, 94).to(dom.getDevice())
    _check(lambda xs: [adj.CurrentDensityFunction.apply(dom, *xs)], [phi0, x0],
           [_rand(phi0, 1.0, 95), _rand(x0, 1.0, 96)], 1e-3, 1e-10, "current density / phi, u x e_b")

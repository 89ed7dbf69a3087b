"""Dirichlet potential values as per-environment data and as differentiable inputs (ref
tests/test_potential_bc.py:605-716): a batch whose environments carry their own values (built
by stack_domains / copy_env_state) steps like single domains with those values; the potential's
right-hand side and the current density, as functions of the values (shared by the environments
or one slice each), pass dot-product tests of their adjoint kernels; and the gradient of a loss
after two differentiable MHD steps with respect to per-environment values matches central
differences. FP64, the closed 3D box with Dirichlet values on -y and per-cell mixed types
(Dirichlet among them) on -x, its +z face at rest."""
from __future__ import annotations

import numpy as np
import pytest
import torch

import cases
from cases import build_domain, rel_l2, to_np

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda", 0)
F64 = torch.float64
DT = 1e-2
E_B = (0.0, 0.6, 0.8)


def _specs():
    return cases.potential_box3d(nx=6, ny=5, nz=4)


def _domain(values=None, batch=1):
    dom = build_domain(_specs(), 0.02, DT, F64, batch=batch, epot=True)
    # the +z face of the box moves through itself: stop it so the box holds its volume (the
    # step refuses unbalanced boundary fluxes); the +x face keeps its tangential motion
    top = dom.getBlocks()[0].getBoundary(5)
    top.setVelocity(torch.zeros_like(top.velocity))
    if values is not None:
        dom.getBlocks()[0].getBoundary(2).setPotentialValues(values)
        dom.UpdateDomainData()
    return dom


def _sim(dom, differentiable=False, tol=1e-13):
    import phipict_amd as pa

    t = pa.SolverTolerance(rtol=tol, atol=1e-15)
    return pa.MHDSimulation(domain=dom, dt=DT, stuart_number=torch.tensor(4.0, dtype=F64),
                            e_b=torch.tensor(E_B, dtype=F64), non_orthogonal=False, differentiable=differentiable,
                            advection_tol=t, pressure_tol=t, potential_tol=t, potential_use_preconditioner=False,
                            potential_return_best_result=False)


def _face_values(seed):
    fc = _domain().getBlocks()[0].getBoundary(2).face_cells()
    return torch.as_tensor(np.random.default_rng(seed).standard_normal(fc), dtype=F64, device=DEV)


def test_per_environment_values_step_like_single_domains():
    from phipict_amd.batching import stack_domains

    print("[potential values] batch of two with their own -y values", flush=True)
    vals = [_face_values(1), _face_values(2)]
    singles = []
    for v in vals:
        dom = _domain(v)
        sim = _sim(dom)
        assert sim.single_step()
        singles.append(dom)
    batched = stack_domains([_domain(v) for v in vals])
    bound = batched.getBlocks()[0].getBoundary(2)
    assert bound.potentialValuesBatched() and bound.potentialValues.numel() == 2 * vals[0].numel()
    sim = _sim(batched)
    assert sim.single_step()
    n = batched.getTotalSize()
    for e, dom in enumerate(singles):
        bv, sv = batched.getBlocks()[0], dom.getBlocks()[0]
        errs = {"velocity": rel_l2(to_np(bv.velocity)[e], to_np(sv.velocity)[0]),
                "pressure": rel_l2(to_np(bv.pressure)[e], to_np(sv.pressure)[0]),
                "potential": rel_l2(to_np(batched.epotResult)[e * n:(e + 1) * n], to_np(dom.epotResult))}
        print(f"  env {e}: {errs}", flush=True)
        assert max(errs.values()) < 1e-10, errs
    # the environments really differ
    assert rel_l2(to_np(batched.epotResult)[:n], to_np(batched.epotResult)[n:]) > 1e-3


def _dot(a, b):
    return float((a.reshape(-1).to(F64) * b.reshape(-1).to(F64)).sum())


@pytest.mark.parametrize("shared", [True, False])
def test_potential_operators_adjoint_in_the_values(shared):
    """Both operators are affine in the values: <w, F(g + v) - F(g)> = <J^T w, v> to rounding."""
    from phipict_amd import adjoint as adj

    print(f"[potential values] dot-product tests, shared={shared}", flush=True)
    batch = 2
    dom = _domain(batch=batch)
    fc = dom.getBlocks()[0].getBoundary(2).face_cells()
    gen = torch.Generator("cpu").manual_seed(7)
    rnd = lambda *s: torch.randn(*s, generator=gen, dtype=F64).to(DEV)  # noqa: E731
    n = dom.getTotalSize()
    field = rnd(batch * dom.getSpatialDims() * n)
    phi, uxb = rnd(batch * n), rnd(batch * 3 * n)
    faces = adj.potential_value_faces(dom)
    assert len(faces) == 2  # -x (mixed types, Dirichlet among them) and -y
    base = [bd.potentialValues.detach().clone() for _, _, bd in faces]
    g0 = base[1] if shared else base[1].repeat(batch)
    dirv = rnd(g0.numel())

    def with_values(g):
        faces[1][2].setPotentialValues(g)
        dom.UpdateDomainData()
        return [bd.potentialValues for _, _, bd in faces]

    for name, op in (("epot rhs", lambda vals: adj.EpotRHSFunction.apply(dom, field, *vals)),
                     ("current density", lambda vals: adj.CurrentDensityFunction.apply(dom, phi, uxb, *vals))):
        with torch.no_grad():
            f_plus = op(with_values(g0 + dirv)).clone()
            f0 = op(with_values(g0)).clone()
        w = rnd(f0.numel())
        lhs = _dot(w, f_plus - f0)
        g = g0.clone().requires_grad_(True)
        vals = with_values(g)
        out = op(vals)
        (grad,) = torch.autograd.grad(out, g, w.reshape(out.shape))
        rhs = _dot(grad, dirv)
        print(f"  {name}: <w, F(g+v)-F(g)> {lhs:.12e}  <J^T w, v> {rhs:.12e}", flush=True)
        assert abs(lhs) > 1e-6
        assert abs(lhs - rhs) <= 1e-11 * (abs(lhs) + abs(rhs)), (name, lhs, rhs)
        # only the Dirichlet cells of the face carry a gradient (-y: every cell)
        assert grad.shape == g0.shape and bool((grad != 0).any())


def _loss(values, steps=2, differentiable=False):
    dom = _domain(batch=2)
    dom.getBlocks()[0].getBoundary(2).setPotentialValues(values)
    dom.UpdateDomainData()
    sim = _sim(dom, differentiable=differentiable)
    for _ in range(steps):
        assert sim.single_step()
    blk = dom.getBlocks()[0]
    rng = np.random.default_rng(8)
    w_u = torch.as_tensor(rng.standard_normal(tuple(blk.velocity.shape)), dtype=F64, device=DEV)
    w_phi = torch.as_tensor(rng.standard_normal(tuple(blk.epot.shape)), dtype=F64, device=DEV)
    return (w_u * blk.velocity ** 2).sum() + (w_phi * blk.epot).sum()


def test_values_gradient_through_mhd_steps():
    print("[potential values] d loss / d per-environment values through two MHD steps", flush=True)
    fc = _face_values(0).numel()
    gen = torch.Generator("cpu").manual_seed(9)
    values = torch.randn(2, fc, generator=gen, dtype=F64).to(DEV)
    direction = torch.randn(2, fc, generator=gen, dtype=F64).to(DEV)
    g = values.clone().requires_grad_(True)
    loss = _loss(g, differentiable=True)
    (grad,) = torch.autograd.grad(loss, g)
    assert not torch.allclose(grad[0], grad[1])
    eps = 1e-4
    with torch.no_grad():
        fd = float((_loss(values + eps * direction) - _loss(values - eps * direction)) / (2 * eps))
    ad = float((grad * direction).sum())
    print(f"  AD {ad:.10e} FD {fd:.10e}", flush=True)
    assert abs(fd) > 1e-6
    assert abs(ad - fd) <= 1e-6 * abs(fd), (ad, fd)

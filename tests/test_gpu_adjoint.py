"""The differentiable step against central finite differences of the forward step on the GPU
(FP64): directional derivatives of a loss of the velocity and pressure after two PISO + MHD
steps with respect to the initial velocity, the inflow profile, the velocity source, the field
direction e_b and the Stuart number agree to 1e-5 relative (the reference checks its adjoint
the same way, test_potential_bc.py:605-716, runscripts/validation/gradcheck.py:23-26); the
solve adjoint satisfies lam = A^{-T} g."""
from __future__ import annotations

import numpy as np
import pytest
import torch

import cases
from cases import build_domain

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda", 0)
F64 = torch.float64


def _specs():
    return cases.outflow2d(nx=14, ny=8, seed=7)


def _run(pert: dict, eps: float, steps: int = 2, differentiable: bool = False, grads: dict = None,
         non_orthogonal: bool = False):
    """Two MHD steps of the outflow channel with every input moved by eps times its direction;
    returns the loss (and, differentiable, the inputs holding gradients)."""
    import phipict_amd as pa

    specs = _specs()
    dom = build_domain(specs, 0.05, 2e-2, F64)
    blk = dom.getBlocks()[0]
    rng = np.random.default_rng(21)
    u0 = blk.velocity.detach().clone() + eps * pert["u"]
    inlet = blk.getBoundary(0)
    vb = inlet.velocity.detach().clone() + eps * pert["vb"]
    src = torch.tensor([[0.5, 0.1]], dtype=F64, device=DEV) + eps * pert["src"]
    e_b = torch.tensor([0.0, 0.6, 0.8], dtype=F64) + eps * pert["eb"].cpu()
    N = torch.tensor(3.0, dtype=F64) + eps * float(pert["N"])
    if differentiable:
        for t in (u0, vb, src, e_b, N):
            t.requires_grad_(True)
    blk.setVelocity(u0)
    inlet.setVelocity(vb)
    # the plug outflow is rescaled to the parabolic inflow's flux (tracked: it depends on vb)
    pa.simulation.balance_boundary_fluxes(dom, [blk.getBoundary(1)])
    dom.UpdateDomainData()

    def reset(domain, **kw):
        domain.getBlocks()[0].setVelocitySource(src)
        domain.UpdateDomainData()

    tol = pa.SolverTolerance(rtol=1e-13)
    sim = pa.MHDSimulation(domain=dom, dt=2e-2, stuart_number=N, e_b=e_b, non_orthogonal=non_orthogonal,
                           advection_tol=tol, pressure_tol=tol, potential_tol=tol, differentiable=differentiable,
                           hooks={"PRE_VELOCITY_SETUP": reset}, flux_balance_tol=1e-3)
    for _ in range(steps):
        assert sim.single_step()
    w_u = torch.as_tensor(rng.standard_normal(tuple(blk.velocity.shape)), dtype=F64, device=DEV)
    w_p = torch.as_tensor(rng.standard_normal(tuple(blk.pressure.shape)), dtype=F64, device=DEV)
    loss = (w_u * blk.velocity).sum() + 10.0 * (w_p * blk.pressure).sum()
    if differentiable:
        loss.backward()
        grads.update(u=u0.grad, vb=vb.grad, src=src.grad, eb=e_b.grad, N=N.grad)
    return float(loss)


def _directions():
    specs = _specs()
    rng = np.random.default_rng(5)
    s = specs[0]
    u = torch.as_tensor(rng.standard_normal([1, 2] + s.cells_zyx), dtype=F64, device=DEV)
    fshape = cases.face_shape(s.cells_zyx, 0)
    vb = torch.zeros([1, 2] + fshape, dtype=F64, device=DEV)
    prof = rng.standard_normal(fshape[0])
    # the normal component moves without changing the net inflow (face heights dy weight the
    # flux), so the domain stays balanced; the tangential component is free
    dy = np.diff(s.vertices[1, :, 0])
    prof -= (prof * dy).sum() / dy.sum()
    vb[0, 0, :, 0] = torch.as_tensor(prof, dtype=F64, device=DEV)
    vb[0, 1, :, 0] = torch.as_tensor(rng.standard_normal(fshape[0]), dtype=F64, device=DEV)
    src = torch.as_tensor(rng.standard_normal((1, 2)), dtype=F64, device=DEV)
    eb = torch.as_tensor(rng.standard_normal(3), dtype=F64, device=DEV)
    return dict(u=u, vb=vb, src=src, eb=eb, N=0.7)


@pytest.mark.parametrize("which", ["u", "vb", "src", "eb", "N"])
def test_step_gradient_matches_finite_differences(which):
    print(f"[adjoint] gradient w.r.t. {which}", flush=True)
    dirs = _directions()
    zero = {k: (v * 0 if isinstance(v, torch.Tensor) else 0.0) for k, v in dirs.items()}
    grads: dict = {}
    _run(zero, 0.0, differentiable=True, grads=grads)
    g = grads[which]
    assert g is not None, f"no gradient for {which}"
    v = dirs[which]
    analytic = float((g.to(DEV) * v).sum()) if isinstance(v, torch.Tensor) else float(g) * v
    eps = 1e-6
    plus = dict(zero)
    plus[which] = v
    fd = (_run(plus, eps) - _run(plus, -eps)) / (2 * eps)
    print(f"analytic {analytic:.10e} fd {fd:.10e}", flush=True)
    assert abs(fd) > 1e-4, "the loss does not depend on this input"
    assert abs(analytic - fd) <= 1e-5 * abs(fd), (analytic, fd)


def test_default_non_orthogonal_mode_on_a_rectilinear_grid():
    """differentiable=True with the reference's default non_orthogonal=True
    (piso_simulation.py:1469): on a rectilinear grid the loss equals that of
    non_orthogonal=False bit for bit and every gradient to the rounding of the adjoint kernels'
    atomic sums; a skewed grid is refused."""
    import phipict_amd as pa

    print("[adjoint] non_orthogonal=True on a rectilinear grid", flush=True)
    zero = {k: (v * 0 if isinstance(v, torch.Tensor) else 0.0) for k, v in _directions().items()}
    runs = []
    for non_ortho in (False, True):
        grads: dict = {}
        loss = _run(zero, 0.0, differentiable=True, grads=grads, non_orthogonal=non_ortho)
        runs.append((loss, grads))
    assert runs[0][0] == runs[1][0]
    for k in runs[0][1]:
        # the adjoint kernels scatter with atomic adds: equal up to the order of those sums
        a, b = runs[0][1][k].double(), runs[1][1][k].double()
        assert float((a - b).norm()) <= 1e-12 * float(a.norm()), k
    specs = cases.skew(_specs(), 0.2, seed=3)
    dom = build_domain(specs, 0.05, 2e-2, F64)
    pa.simulation.balance_boundary_fluxes(dom, [dom.getBlocks()[0].getBoundary(1)])
    dom.UpdateDomainData()
    sim = pa.MHDSimulation(domain=dom, dt=2e-2, stuart_number=torch.tensor(3.0, dtype=F64),
                           e_b=torch.tensor([0.0, 0.6, 0.8], dtype=F64), differentiable=True, flux_balance_tol=1e-3)
    with pytest.raises(NotImplementedError, match="skewed"):
        sim.single_step()


def _run_blocks(case: str, dirs, eps: float, differentiable: bool = False, grads: list = None, batch: int = 1):
    """Two steps of a multi-block domain (connected, periodic and fixed faces) with the initial
    velocities moved by eps times dirs; MHD with an oblique field except for the rotated pair."""
    import phipict_amd as pa

    specs = {"two_blocks": lambda: cases.split_channel2d(12, 10, 5, gradient=4.0)[1],
             "duct3d": lambda: cases.duct3d(nx=6, ny=5, nz=4, noise=0.1, gradient=5.0),
             "rotated_pair3d": cases.rotated_pair3d}[case]()
    dom = build_domain(specs, 0.05, 1e-2, F64, batch=batch)
    us = []
    for blk, dv in zip(dom.getBlocks(), dirs):
        u = blk.velocity.detach().clone() + eps * dv
        if differentiable:
            u.requires_grad_(True)
        blk.setVelocity(u)
        us.append(u)
    dom.UpdateDomainData()
    tol = pa.SolverTolerance(rtol=1e-13)
    kw = dict(domain=dom, dt=1e-2, non_orthogonal=False, advection_tol=tol, pressure_tol=tol,
              differentiable=differentiable)
    if case == "rotated_pair3d":
        sim = pa.Simulation(**kw)
    else:
        sim = pa.MHDSimulation(stuart_number=torch.tensor(2.0, dtype=F64), e_b=torch.tensor([0.0, 0.6, 0.8], dtype=F64),
                               potential_tol=tol, **kw)
    for _ in range(2):
        assert sim.single_step()
    rng = np.random.default_rng(33)
    loss = 0.0
    for blk in dom.getBlocks():
        w = torch.as_tensor(rng.standard_normal(tuple(blk.velocity.shape)), dtype=F64, device=DEV)
        wp = torch.as_tensor(rng.standard_normal(tuple(blk.pressure.shape)), dtype=F64, device=DEV)
        loss = loss + (w * blk.velocity).sum() + 10.0 * (wp * blk.pressure).sum()
    if differentiable:
        loss.backward()
        grads.extend(u.grad for u in us)
    return float(loss)


@pytest.mark.parametrize("case", ["two_blocks", "duct3d", "rotated_pair3d"])
@pytest.mark.parametrize("batch", [1, 2])
def test_multiblock_gradient_matches_finite_differences(case, batch):
    """Connected (also with crossed tangential axes), periodic and wall faces in the adjoint
    kernels, 2D and 3D, one environment and a batch of two: the gradient of the initial
    velocities against central differences."""
    print(f"[adjoint] multi-block {case} batch={batch}", flush=True)
    import phipict_amd as pa  # noqa: F401

    probe = build_domain({"two_blocks": lambda: cases.split_channel2d(12, 10, 5, gradient=4.0)[1],
                          "duct3d": lambda: cases.duct3d(nx=6, ny=5, nz=4, noise=0.1, gradient=5.0),
                          "rotated_pair3d": cases.rotated_pair3d}[case](), 0.05, 1e-2, F64, batch=batch)
    rng = np.random.default_rng(8)
    dirs = [torch.as_tensor(rng.standard_normal(tuple(b.velocity.shape)), dtype=F64, device=DEV)
            for b in probe.getBlocks()]
    grads: list = []
    _run_blocks(case, [d * 0 for d in dirs], 0.0, differentiable=True, grads=grads, batch=batch)
    analytic = sum(float((g * d).sum()) for g, d in zip(grads, dirs))
    eps = 1e-6
    fd = (_run_blocks(case, dirs, eps, batch=batch) - _run_blocks(case, dirs, -eps, batch=batch)) / (2 * eps)
    print(f"analytic {analytic:.10e} fd {fd:.10e}", flush=True)
    assert abs(fd) > 1e-4
    assert abs(analytic - fd) <= 1e-5 * abs(fd), (analytic, fd)


@pytest.mark.parametrize("singular", [False, True])
def test_linear_solve_adjoint_identity(singular):
    """The solve's backward is lam = A^{-T} g (scipy on the host), db = lam and dA = -lam x^T on
    the pattern; singular systems on the mean-free subspace."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla

    import phipict_amd as pa
    from oracle import oracle as O
    from phipict_amd import adjoint as adj

    od = O.OracleDomain(cases.channel2d(12, 10, u0=2.0, seed=3), 0.02, 0.05)
    col, val, A_diag = od.advection_matrix()
    row = od.row
    if singular:
        col, val = od.pressure_matrix(A_diag)
    n = row.size - 1
    A = pa.CSRmatrix(torch.as_tensor(val, device=DEV), torch.as_tensor(col, device=DEV),
                     torch.as_tensor(row, device=DEV))
    rng = np.random.default_rng(4)
    b = rng.standard_normal(n)
    if singular:
        b -= b.mean()
    value = A.value.clone().requires_grad_(True)
    bt = torch.as_tensor(b, device=DEV).requires_grad_(True)
    sim = pa.Simulation(domain=None, dt=0.1)

    def solve(rhs):
        x, _ = sim.linear_solve(A, rhs, x=torch.zeros_like(rhs), use_BiCG=not singular, tol=1e-14,
                                matrix_rank_deficient=singular)
        return x

    x = adj.LinearSolveFunction.apply(solve, A, value, bt, 1, singular, singular, 1e-13)
    g = rng.standard_normal(n)
    (x * torch.as_tensor(g, device=DEV)).sum().backward()
    M = sp.csr_matrix((val, col, row), shape=(n, n))
    if singular:
        g = g - g.mean()
        lam = np.linalg.lstsq(M.T.toarray(), g, rcond=None)[0]
        lam -= lam.mean()
    else:
        lam = spla.spsolve(M.T.tocsc(), g)
    got = bt.grad.cpu().numpy()
    assert np.linalg.norm(got - lam) <= 1e-9 * np.linalg.norm(lam)
    xs = x.detach().cpu().numpy()
    rows = np.repeat(np.arange(n), np.diff(row))
    np.testing.assert_allclose(value.grad.cpu().numpy(), -lam[rows] * xs[col], rtol=1e-7,
                               atol=1e-9 * np.abs(lam).max() * np.abs(xs).max())


def _tjunction_run(n: int, steps: int, every: int, d_in, d_b0: float, eps: float, mode: str):
    """MHD T-junction rollout with advected, flux-balanced outlets; loss = MSE of the branch
    outlet's cell-layer velocity against a seeded target. mode: "fd" (untracked forward),
    "full" (one tracked rollout) or "ckpt" (checkpointed every ``every`` steps)."""
    import phipict_amd as pa
    from phipict_amd import adjoint as adj
    from phipict_amd.simulation import _face_slice

    specs = cases.tjunction3d(n=n)
    dom = build_domain(specs, 1.0 / 200, 5e-3, F64)
    b0, b1, b2 = dom.getBlocks()
    inlet = b0.getBoundary(0)
    vin = inlet.velocity.detach().clone() + eps * d_in
    B0 = torch.tensor(1.0 + eps * d_b0, dtype=F64)
    if mode != "fd":
        vin.requires_grad_(True)
        B0.requires_grad_(True)
    inlet.setVelocity(vin)
    outs = [b1.getBoundary(2), b2.getBoundary(3)]
    velm = torch.tensor([[0.0, 0.5, 0.0]], dtype=F64, device=DEV)
    pa.simulation.balance_boundary_fluxes(dom, outs)
    dom.UpdateDomainData()

    def advect(domain, **kw):
        pa.simulation.update_advective_boundaries(domain, outs, velm, 5e-3)

    def no_source(domain, **kw):
        # the Lorentz force is added to the current source: start every step from none
        for b in domain.getBlocks():
            b.clearVelocitySource()
        domain.UpdateDomainData()

    tol = pa.SolverTolerance(rtol=1e-13)
    e_b = B0 * torch.tensor([0.0, 1.0, 0.0], dtype=F64)
    sim = pa.MHDSimulation(domain=dom, dt=5e-3, stuart_number=torch.tensor(2.0, dtype=F64), e_b=e_b,
                           non_orthogonal=False, advection_tol=tol, pressure_tol=tol, potential_tol=tol,
                           hooks={"PRE": advect, "PRE_VELOCITY_SETUP": no_source}, differentiable=(mode == "full"),
                           flux_balance_tol=1e-6)
    target = torch.as_tensor(np.random.default_rng(5).standard_normal([1, 3, n, 1, n]), dtype=F64, device=DEV)

    def loss_fn(domain):
        layer = _face_slice(domain.getBlocks()[2], 3)
        assert layer.shape == target.shape
        return ((layer - target) ** 2).mean()

    if mode == "ckpt":
        loss, g0 = adj.checkpointed_gradient(sim, steps, every, loss_fn)
        return loss, vin.grad, B0.grad
    for _ in range(steps):
        assert sim.single_step()
    loss = loss_fn(dom)
    if mode == "full":
        loss.backward()
        return float(loss), vin.grad, B0.grad
    return float(loss)


@pytest.mark.parametrize("n", [4, 24])
def test_checkpointed_tjunction_gradients(n):
    """T-junction (3 connected blocks, advected outlets): gradients of the outlet MSE after 4
    MHD steps w.r.t. the inflow field and B0 from the checkpointed backward (segments of 2)
    equal the single-segment backward and central differences. At n = 24 the domain has
    55296 cells, so every forward and transposed solve runs the multi-workgroup engine."""
    print(f"[adjoint] checkpointed T-junction n={n}", flush=True)
    steps = 4
    rng = np.random.default_rng(12)
    fshape = cases.face_shape([n, n, 2 * n], 0)
    d_in = torch.zeros([1, 3] + fshape, dtype=F64, device=DEV)
    # a net inflow change too: the outlets are rescaled against it before and during the rollout
    d_in[0, 0] = torch.as_tensor(rng.standard_normal(fshape), dtype=F64, device=DEV)
    d_in[0, 1] = torch.as_tensor(rng.standard_normal(fshape), dtype=F64, device=DEV)
    loss_c, g_in_c, g_b0_c = _tjunction_run(n, steps, 2, d_in, 0.0, 0.0, "ckpt")
    loss_f, g_in_f, g_b0_f = _tjunction_run(n, steps, steps, d_in, 0.0, 0.0, "full")
    assert abs(loss_c - loss_f) <= 1e-12 * abs(loss_f)
    assert float((g_in_c - g_in_f).norm()) <= 1e-8 * float(g_in_f.norm())
    assert abs(float(g_b0_c) - float(g_b0_f)) <= 1e-8 * abs(float(g_b0_f))
    eps = 1e-6
    fd_in = (_tjunction_run(n, steps, 0, d_in, 0.0, eps, "fd") - _tjunction_run(n, steps, 0, d_in, 0.0, -eps, "fd")) / (2 * eps)
    fd_b0 = (_tjunction_run(n, steps, 0, d_in * 0, 1.0, eps, "fd")
             - _tjunction_run(n, steps, 0, d_in * 0, 1.0, -eps, "fd")) / (2 * eps)
    an_in = float((g_in_c * d_in).sum())
    an_b0 = float(g_b0_c)
    print(f"inflow: analytic {an_in:.10e} fd {fd_in:.10e}; B0: analytic {an_b0:.10e} fd {fd_b0:.10e}", flush=True)
    assert abs(fd_in) > 1e-8 and abs(fd_b0) > 1e-8
    assert abs(an_in - fd_in) <= 1e-5 * abs(fd_in)
    assert abs(an_b0 - fd_b0) <= 1e-5 * abs(fd_b0)

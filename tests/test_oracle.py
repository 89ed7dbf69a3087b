"""Pins the oracle against what the mathematics fixes: closed forms on uniform periodic grids
(metrics, Fourier-mode actions of the assembled operators), invariants the reference's own tests
check (zero row sums and symmetry of the Poisson operators, test_amg.py:73-86; constant
potential with a Dirichlet face, test_potential_bc.py:456-471; a block split and reconnected,
test_batching.py:22-37) and direct solves for the Krylov and AMG paths (test_krylov.py:13-35,
test_amg.py:88-131). CPU only."""
from __future__ import annotations

import math

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import cases
from oracle import oracle as O


def dense(row, col, val, n=None):
    n = n or row.size - 1
    return sp.csr_matrix((np.asarray(val, np.float64), col, row), shape=(n, n)).toarray()


def periodic_box(n, lengths, vel=None, dims=None):
    """Uniform periodic block with cell counts n=[nx, ny(, nz)]."""
    axes = [np.linspace(0.0, L, k + 1) for k, L in zip(n, lengths)]
    v = cases.vertices(axes)
    d = len(n)
    cz = cases.cell_shape(v)
    velocity = np.zeros([d] + cz)
    if vel is not None:
        velocity[...] = np.asarray(vel).reshape([d] + [1] * d)
    return [O.BlockSpec(v, [O.FaceSpec("periodic") for _ in range(2 * d)], velocity=velocity, pressure=np.zeros(cz))]


def centers_x(spec):
    v = spec.vertices
    d = v.shape[0]
    x = v[0]
    for ax in range(d):
        x = 0.5 * (np.take(x, range(x.shape[ax] - 1), axis=ax) + np.take(x, range(1, x.shape[ax]), axis=ax))
    return x.reshape(-1)


# ----------------------------------------------------------------------------- metrics
@pytest.mark.parametrize("d", [2, 3])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_transforms_affine_closed_form(d, dtype):
    """For vertices x = L (h * xi) the cell and face Jacobians are L diag(h) everywhere."""
    rng = np.random.default_rng(7)
    n = [4, 3, 5][:d]
    h = np.array([0.3, 0.7, 0.45][:d])
    L = np.eye(d) + 0.2 * rng.standard_normal((d, d))
    base = cases.vertices([np.arange(k + 1, dtype=np.float64) for k in n])
    v = np.einsum("ij,j...->i...", L @ np.diag(h), base)
    M = L @ np.diag(h)
    T = O.transforms(v, dtype)
    tol = 1e-12 if dtype == np.float64 else 2e-5
    np.testing.assert_allclose(T[:, :d * d], np.tile(M.reshape(-1), (T.shape[0], 1)), atol=tol)
    np.testing.assert_allclose(T[:, d * d:2 * d * d], np.tile(np.linalg.inv(M).reshape(-1), (T.shape[0], 1)),
                               atol=tol * 10)
    np.testing.assert_allclose(T[:, -1], np.linalg.det(M), rtol=tol * 10)
    FT = O.face_transforms(v, dtype)
    for axis in range(d):
        # valid face positions: every tangential vertex index below the last
        idx = [axis] + [slice(None)] * d
        for dd in range(d):
            if dd != axis:
                idx[1 + (d - 1 - dd)] = slice(0, -1)
        ft = FT[tuple(idx)].reshape(-1, 2 * d * d + 1)
        np.testing.assert_allclose(ft[:, :d * d], np.tile(M.reshape(-1), (ft.shape[0], 1)), atol=tol)
        np.testing.assert_allclose(ft[:, -1], np.linalg.det(M), rtol=tol * 10)


def test_transforms_rectilinear_tanh():
    """Rectilinear graded grid: M = diag(dx, dy), det = dx dy, Minv = diag(1/dx, 1/dy)."""
    xs = np.linspace(0, 2, 9)
    ys = cases.tanh_coords(6, -1, 1, 1.5)
    T = O.transforms(cases.vertices([xs, ys]))
    dx = np.tile(np.diff(xs), 6)
    dy = np.repeat(np.diff(ys), 8)
    np.testing.assert_allclose(T[:, 0], dx, rtol=1e-14)
    np.testing.assert_allclose(T[:, 3], dy, rtol=1e-14)
    np.testing.assert_allclose(T[:, [1, 2, 5, 6]], 0, atol=0)
    np.testing.assert_allclose(T[:, 4], 1 / dx, rtol=1e-14)
    np.testing.assert_allclose(T[:, 7], 1 / dy, rtol=1e-14)
    np.testing.assert_allclose(T[:, 8], dx * dy, rtol=1e-14)


# ----------------------------------------------------------------------------- assembly closed forms
@pytest.mark.parametrize("d", [2, 3])
def test_advection_matrix_fourier(d):
    """Constant velocity (a, b, c) on a uniform periodic grid: C cos(k x) =
    (1/dt + 2 nu (1 - cos kh)/h^2) cos(k x) - a sin(kh)/h sin(k x)  (central fluxes)."""
    n = [16, 6, 5][:d]
    lengths = [2.0, 1.3, 0.9][:d]
    vel = [0.7, -0.4, 0.25][:d]
    nu, dt = 0.03, 0.05
    specs = periodic_box(n, lengths, vel)
    od = O.OracleDomain(specs, nu, dt)
    col, val, A = od.advection_matrix()
    D = dense(od.row, col, val)
    h = lengths[0] / n[0]
    k = 2 * math.pi * 3 / lengths[0]
    x = centers_x(specs[0])
    phi = np.cos(k * x)
    expect = (1 / dt + 2 * nu * (1 - math.cos(k * h)) / h ** 2) * phi - vel[0] * math.sin(k * h) / h * np.sin(k * x)
    np.testing.assert_allclose(D @ phi, expect, atol=1e-11 * np.abs(expect).max())
    hs = [L / m for L, m in zip(lengths, n)]
    np.testing.assert_allclose(A, 1 / dt + sum(2 * nu / hh ** 2 for hh in hs), rtol=1e-13)
    np.testing.assert_allclose(np.diag(D), A, rtol=0)


@pytest.mark.parametrize("d", [2, 3])
def test_pressure_and_epot_matrix_fourier(d):
    """P cos(kx) = -(alpha_x/A)(2 - 2cos kh) cos(kx), Epot cos(kx) = -alpha_x (2 - 2cos kh) cos(kx)
    with alpha_x = det/h_x^2."""
    n = [12, 5, 4][:d]
    lengths = [1.5, 1.0, 0.8][:d]
    specs = periodic_box(n, lengths, [0.2, 0.1, -0.3][:d])
    od = O.OracleDomain(specs, 0.02, 0.01)
    _, _, A = od.advection_matrix()
    pc, pv = od.pressure_matrix(A)
    ec, ev = od.epot_matrix()
    hs = [L / m for L, m in zip(lengths, n)]
    det = float(np.prod(hs))
    alpha = det / hs[0] ** 2
    k = 2 * math.pi * 2 / lengths[0]
    x = centers_x(specs[0])
    phi = np.cos(k * x)
    lam = -alpha * (2 - 2 * math.cos(k * hs[0]))
    np.testing.assert_allclose(dense(od.row, pc, pv) @ phi, lam / A[0] * phi, atol=1e-12 * abs(lam / A[0]))
    np.testing.assert_allclose(dense(od.row, ec, ev) @ phi, lam * phi, atol=1e-12 * abs(lam))


def test_advection_rhs_and_hbya_closed_form():
    """Periodic, constant fields: rhs = u/dt + S; HbyA = (u^n/dt - (1/dt - A) u* + S)/A."""
    n, lengths = [8, 6], [1.0, 2.0]
    un = np.array([0.3, -0.2])
    specs = periodic_box(n, lengths, un)
    S = np.array([0.5, 1.5])
    specs[0].vsource = S
    dt = 0.02
    od = O.OracleDomain(specs, 0.01, dt)
    col, val, A = od.advection_matrix()
    rhs = od.advection_rhs().reshape(2, -1)
    np.testing.assert_allclose(rhs, (un / dt + S)[:, None] * np.ones_like(rhs), rtol=1e-13)
    ustar = np.array([0.25, 0.05])
    vres = np.repeat(ustar, od.ncells)
    prhs = od.pressure_rhs(A, col, val, vres).reshape(2, -1)
    expect = (un[:, None] / dt - (1 / dt - A[None, :]) * ustar[:, None] + S[:, None]) / A[None, :]
    np.testing.assert_allclose(prhs, expect, rtol=1e-12)


def test_linear_upwind_closed_form():
    """LINEAR_UPWIND (advection.cu:129-140, 893-937): first-order upwind in the matrix plus the
    deferred correction -F 0.5 psi (phi_D - phi_U), psi = (phi_U - phi_UU) / (phi_D - phi_U), on
    the faces whose far-upwind cell lies in the block. On a uniform grid with flow along +x,
    R = C u - rhs then differs from CENTRAL's by (-0.5 F_e D2u_i - 0.5 F_w' D2u_(i-1)) / det
    (F_e the east face's outward flux, F_w' the west face's, D2u_i = u_(i+1) - 2 u_i + u_(i-1))
    on the cells whose faces are all corrected: zero for a field linear in x, the second
    difference for a quadratic one. The scheme's advective off-diagonals are never positive."""
    nx, ny, lx, ly = 16, 4, 2.0, 1.0
    dx, dy = lx / nx, ly / ny
    x = (np.arange(nx) + 0.5) * dx
    for coef in ((1.0, 0.3, 0.0), (1.0, 0.3, 0.8)):
        ux = coef[0] + coef[1] * x + coef[2] * x * x
        specs = periodic_box([nx, ny], [lx, ly])
        specs[0].velocity = np.stack([np.broadcast_to(ux, (ny, nx)), np.zeros((ny, nx))]).copy()
        R = {}
        for scheme in (O.ADV_CENTRAL, O.ADV_LINEAR_UPWIND):
            od = O.OracleDomain(specs, 1e-4, 0.01, advection_scheme=scheme)
            col, val, A = od.advection_matrix()
            u = specs[0].velocity.reshape(2, -1)
            Cm = dense(od.row, col, val)
            R[scheme] = np.stack([Cm @ u[k] for k in range(2)]) - od.advection_rhs().reshape(2, -1)
            if scheme == O.ADV_LINEAR_UPWIND:
                rows = np.repeat(np.arange(od.ncells), np.diff(od.row))
                off = col != rows
                ix = rows % nx
                inner = (ix >= 2) & (ix <= nx - 3)
                # viscous off-diagonals are -nu alpha; no advective part makes them positive
                assert (val[off & inner] <= 0).all()
        diff = (R[O.ADV_LINEAR_UPWIND] - R[O.ADV_CENTRAL]).reshape(2, ny, nx)
        Fe = dy * 0.5 * (ux[2:-2] + ux[3:-1])          # outward flux through the east face
        Fw = -dy * 0.5 * (ux[1:-3] + ux[2:-2])         # outward flux through the west face
        d2 = ux[2:] - 2 * ux[1:-1] + ux[:-2]           # D2u at cells 1..nx-2
        expect = (-0.5 * Fe * d2[1:-1] - 0.5 * Fw * d2[:-2]) / (dx * dy)
        np.testing.assert_allclose(diff[0][:, 2:-2], np.broadcast_to(expect, (ny, nx - 4)), rtol=1e-9, atol=1e-11)
        np.testing.assert_allclose(diff[1], 0.0, atol=1e-11)
        if coef[2] == 0.0:
            assert np.abs(diff[0][:, 2:-2]).max() < 1e-10
        else:
            assert np.abs(expect).min() > 1e-3


def test_divergence_gradient_current_closed_form():
    """Face-interpolated divergence, FD gradient correction and face-based current density of
    Fourier modes on a uniform periodic grid."""
    n, lengths = [16, 4], [2.0, 1.0]
    specs = periodic_box(n, lengths)
    od = O.OracleDomain(specs, 0.01, 0.1)
    _, _, A = od.advection_matrix()
    hx, hy = lengths[0] / n[0], lengths[1] / n[1]
    k = 2 * math.pi / lengths[0]
    x = centers_x(specs[0])
    N = od.ncells
    h = np.concatenate([np.sin(k * x), np.zeros(N)])
    div = od.pressure_div(h)
    np.testing.assert_allclose(div, hy * np.cos(k * x) * math.sin(k * hx), atol=1e-13)
    # epot rhs of chi = (sin kx, 0) is the same face divergence
    np.testing.assert_allclose(od.epot_rhs(h), hy * np.cos(k * x) * math.sin(k * hx), atol=1e-13)
    # velocity correction: u = HbyA - dp/dx / A with central differences
    p = np.sin(k * x)
    od.pressure[0][...] = p
    hb = np.concatenate([np.full(N, 0.3), np.full(N, -0.1)])
    u = od.correct_velocity(A, hb).reshape(2, -1)
    np.testing.assert_allclose(u[0], 0.3 - np.cos(k * x) * math.sin(k * hx) / hx / A, atol=1e-12)
    np.testing.assert_allclose(u[1], -0.1, atol=1e-13)
    # current density: J = -grad_h(phi) + chi
    chi = np.concatenate([np.full(N, 0.2), np.full(N, -0.7), np.full(N, 0.4)])
    J = od.current_density(p, chi).reshape(3, -1)
    np.testing.assert_allclose(J[0], 0.2 - np.cos(k * x) * math.sin(k * hx) / hx, atol=1e-12)
    np.testing.assert_allclose(J[1], -0.7, atol=1e-12)
    np.testing.assert_allclose(J[2], 0.4, atol=1e-13)


# ----------------------------------------------------------------------------- invariants
@pytest.mark.parametrize("name", list(cases.CASES))
def test_poisson_operators_symmetric(name):
    specs = cases.CASES[name]()
    od = O.OracleDomain(specs, 0.01, 0.02)
    _, _, A = od.advection_matrix()
    pc, pv = od.pressure_matrix(A)
    P = dense(od.row, pc, pv)
    np.testing.assert_allclose(P, P.T, atol=1e-14 * np.abs(P).max())
    # the pressure operator is pure Neumann: every row sums to zero
    np.testing.assert_allclose(P.sum(axis=1), 0, atol=1e-13 * np.abs(P).max())
    ec, ev = od.epot_matrix()
    E = dense(od.row, ec, ev)
    np.testing.assert_allclose(E, E.T, atol=1e-14 * np.abs(E).max())
    assert np.all(np.diag(E) <= 0) and np.all(np.diag(P) < 0)


def test_epot_row_sums_only_at_anchored_faces():
    """Rows sum to zero except cells on Dirichlet or thin-wall faces, whose row sums are negative."""
    for specs in (cases.potential_box3d(), cases.potential_box2d()):
        od = O.OracleDomain(specs, 0.01, 0.02)
        ec, ev = od.epot_matrix()
        E = dense(od.row, ec, ev)
        rs = E.sum(axis=1)
        assert np.all(rs <= 1e-12)
        assert np.any(rs < -1e-6)
    od = O.OracleDomain(cases.duct3d(), 0.01, 0.02)
    ec, ev = od.epot_matrix()
    np.testing.assert_allclose(dense(od.row, ec, ev).sum(axis=1), 0, atol=1e-12)


@pytest.mark.parametrize("mixed", [False, True], ids=["whole_face", "mixed_face"])
def test_thin_wall_surface_laplacian_closed_form(mixed):
    """Thin-wall condition dphi/dn = C_w lap_tau(phi) at a wall (potential.cu:164-218): on a
    rectilinear box, periodic in x with unequal cell widths dx_i, the potential matrix with a
    thin wall on -y differs from the insulating one only in the wall cells' rows, by the 1D
    surface Laplacian C_w / 2 (1/dx_i + 1/dx_j) (phi_j - phi_i) over the tangential neighbours j
    (alpha_x |Minv_y| = (dx dy / dx^2) / dy = 1 / dx); on a mixed face only between two
    thin-wall face cells."""
    rng = np.random.default_rng(3)
    nx, ny, cw = 10, 5, 0.4
    xs = np.concatenate([[0.0], np.cumsum(rng.uniform(0.5, 1.5, nx))])
    ys = np.array([-1.0, -0.7, -0.2, 0.3, 0.6, 1.0])
    v = cases.vertices([xs, ys])
    cz = cases.cell_shape(v)
    thin = np.zeros(nx, bool)
    thin[:6 if mixed else nx] = True

    def domain(with_thin):
        if with_thin and mixed:
            types = np.where(thin, O.POT_THIN_WALL, O.POT_INSULATING).astype(np.int8)
            bottom = O.FaceSpec("fixed", pot_types=types, cw=cw)
        elif with_thin:
            bottom = O.FaceSpec("fixed", pot=O.POT_THIN_WALL, cw=cw)
        else:
            bottom = O.FaceSpec("fixed")
        faces = [O.FaceSpec("periodic"), O.FaceSpec("periodic"), bottom, O.FaceSpec("fixed")]
        od = O.OracleDomain([O.BlockSpec(v, faces, velocity=np.zeros([2] + cz), pressure=np.zeros(cz))], 0.01, 0.02)
        ec, ev = od.epot_matrix()
        return dense(od.row, ec, ev)

    D = domain(True) - domain(False)
    dx = np.diff(xs)
    expect = np.zeros_like(D)
    for i in range(nx):
        for j in ((i - 1) % nx, (i + 1) % nx):
            if thin[i] and thin[j]:
                c = cw * 0.5 * (1.0 / dx[i] + 1.0 / dx[j])
                expect[i, j] += c
                expect[i, i] -= c
    np.testing.assert_allclose(D, expect, atol=1e-13)
    assert np.abs(expect).max() > 0.1


def test_constant_potential_with_dirichlet_face():
    """phi = g everywhere with Dirichlet value g on a face and chi = 0: Epot phi == rhs and
    J == 0 (test_potential_bc.py:456-471)."""
    specs = cases.channel2d(nx=6, ny=5, gradient=None)
    g = 0.75
    fs = specs[0].faces[2]
    fs.pot = O.POT_DIRICHLET
    fs.pot_values = np.full(cases.face_shape(specs[0].cells_zyx, 2), g)
    od = O.OracleDomain(specs, 0.01, 0.02)
    ec, ev = od.epot_matrix()
    N = od.ncells
    phi = np.full(N, g)
    rhs = od.epot_rhs(np.zeros(2 * N))
    np.testing.assert_allclose(dense(od.row, ec, ev) @ phi, rhs, atol=1e-13)
    J = od.current_density(phi, np.zeros(3 * N))
    np.testing.assert_allclose(J, 0, atol=1e-13)
    # Dirichlet values enter the RHS linearly and never the matrix
    fs.pot_values = np.full_like(fs.pot_values, 2 * g)
    od2 = O.OracleDomain(specs, 0.01, 0.02)
    ec2, ev2 = od2.epot_matrix()
    np.testing.assert_array_equal(ev2, ev)
    np.testing.assert_allclose(od2.epot_rhs(np.zeros(2 * N)), 2 * rhs, rtol=1e-14)


def _perm_split(nx, ny, split):
    perm = np.zeros(nx * ny, np.int64)
    for j in range(ny):
        for i in range(nx):
            if i < split:
                m = j * split + i
            else:
                m = ny * split + j * (nx - split) + (i - split)
            perm[j * nx + i] = m
    return perm


def test_split_block_equals_single_block():
    """A channel split into two connected blocks assembles the same operators and fields."""
    nx, ny, split = 16, 12, 6
    single, pair = cases.split_channel2d(nx, ny, split)
    o1 = O.OracleDomain(single, 0.02, 0.03)
    o2 = O.OracleDomain(pair, 0.02, 0.03)
    perm = _perm_split(nx, ny, split)
    c1, v1, A1 = o1.advection_matrix()
    c2, v2, A2 = o2.advection_matrix()
    D1, D2 = dense(o1.row, c1, v1), dense(o2.row, c2, v2)
    np.testing.assert_allclose(D2[np.ix_(perm, perm)], D1, atol=1e-13 * np.abs(D1).max())
    np.testing.assert_allclose(A2[perm], A1, rtol=1e-14)
    np.testing.assert_allclose(o2.advection_rhs().reshape(2, -1)[:, perm], o1.advection_rhs().reshape(2, -1), rtol=1e-13)
    p1c, p1v = o1.pressure_matrix(A1)
    p2c, p2v = o2.pressure_matrix(A2)
    np.testing.assert_allclose(dense(o2.row, p2c, p2v)[np.ix_(perm, perm)], dense(o1.row, p1c, p1v), atol=1e-15)
    e1c, e1v = o1.epot_matrix()
    e2c, e2v = o2.epot_matrix()
    np.testing.assert_allclose(dense(o2.row, e2c, e2v)[np.ix_(perm, perm)], dense(o1.row, e1c, e1v), atol=1e-13)
    rng = np.random.default_rng(3)
    vr = rng.standard_normal(2 * nx * ny)
    vr2 = np.zeros_like(vr)
    vr2.reshape(2, -1)[:, perm] = vr.reshape(2, -1)
    pr1 = o1.pressure_rhs(A1, c1, v1, vr)
    pr2 = o2.pressure_rhs(A2, c2, v2, vr2)
    np.testing.assert_allclose(pr2.reshape(2, -1)[:, perm], pr1.reshape(2, -1), rtol=1e-12, atol=1e-13)
    np.testing.assert_allclose(o2.pressure_div(pr2)[perm], o1.pressure_div(pr1), atol=1e-12)


# ----------------------------------------------------------------------------- Krylov
def _poisson(nx, ny, shift=0.0):
    od = O.OracleDomain(cases.channel2d(nx, ny), 0.02, 0.01)
    _, _, A = od.advection_matrix()
    pc, pv = od.pressure_matrix(A)
    if shift:
        pv = pv.copy()
        for i in range(od.ncells):
            for k in range(od.row[i], od.row[i + 1]):
                if pc[k] == i:
                    pv[k] -= shift
    return od.row, pc, pv


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_cg_matches_direct_solve(dtype):
    row, col, val = _poisson(12, 10, shift=0.5)
    n = row.size - 1
    rng = np.random.default_rng(0)
    b = rng.standard_normal(n).astype(dtype)
    x = np.zeros(n, dtype)
    tol = 1e-10 if dtype == np.float64 else 1e-5
    r = O.cg(row, col, val.astype(dtype), b, x, tol)
    assert r.converged and r.is_finite
    M = sp.csr_matrix((val, col, row), shape=(n, n))
    xd = spla.spsolve(M.tocsc(), b.astype(np.float64))
    res = np.linalg.norm(b - M @ x) / math.sqrt(n)
    assert res < 10 * tol
    assert abs(res - r.final_residual) < 2 * tol + 1e-3 * res
    assert cases.rel_l2(x, xd) < (1e-8 if dtype == np.float64 else 1e-3)


def test_cg_distinct_eigenvalues_and_exact_start():
    """CG on a diagonal matrix with 3 distinct eigenvalues converges within 3 iterations; an
    exact initial guess stops before the first iteration."""
    n = 30
    d = np.repeat([1.0, 3.0, 7.0], 10)
    row = np.arange(n + 1, dtype=np.int32)
    col = np.arange(n, dtype=np.int32)
    b = np.random.default_rng(1).standard_normal(n)
    x = np.zeros(n)
    r = O.cg(row, col, d, b, x, 1e-12)
    assert r.converged and r.used_iterations <= 3
    np.testing.assert_allclose(x, b / d, rtol=1e-12)
    x2 = b / d
    r2 = O.cg(row, col, d, b, x2, 1e-12)
    assert r2.converged and r2.used_iterations == 0


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_bicgstab_matches_direct_solve(dtype):
    od = O.OracleDomain(cases.channel2d(12, 10, u0=3.0), 0.02, 0.05)
    col, val, A = od.advection_matrix()
    n = od.ncells
    M = sp.csr_matrix((val, col, od.row), shape=(n, n))
    assert abs(M - M.T).max() > 1e-3  # non-symmetric
    b = np.random.default_rng(2).standard_normal(n).astype(dtype)
    x = np.zeros(n, dtype)
    tol = 1e-10 if dtype == np.float64 else 1e-5
    r = O.bicgstab(od.row, col, val.astype(dtype), b, x, tol)
    assert r.converged
    xd = spla.spsolve(M.tocsc(), b.astype(np.float64))
    assert np.linalg.norm(b - M @ x) / math.sqrt(n) < 10 * tol
    assert cases.rel_l2(x, xd) < (1e-8 if dtype == np.float64 else 1e-4)
    # converged start reports -1 iterations (krylov.cu:800-805)
    r2 = O.bicgstab(od.row, col, val.astype(dtype), b, x.copy(), 1e3)
    assert r2.converged and r2.used_iterations == -1


def test_cg_return_best_reports_its_residual():
    row, col, val = _poisson(16, 12)
    n = row.size - 1
    b = np.random.default_rng(4).standard_normal(n)
    b -= b.mean()
    x = np.zeros(n)
    r = O.cg(row, col, val, b, x, 1e-14, maxit=7, return_best=True)
    assert not r.converged
    M = sp.csr_matrix((val, col, row), shape=(n, n))
    true_res = np.linalg.norm(b - M @ x) / math.sqrt(n)
    assert abs(true_res - r.final_residual) < 1e-10 * max(1.0, true_res)


def test_projected_cg_on_singular_system():
    """Pure-Neumann operator with a mean-free RHS: projected CG returns the least-squares
    solution up to a constant."""
    row, col, val = _poisson(10, 8)
    n = row.size - 1
    b = np.random.default_rng(5).standard_normal(n)
    b -= b.mean()
    x = np.zeros(n)
    r = O.cg(row, col, val, b, x, 1e-11, project=True)
    assert r.converged
    M = dense(row, col, val)
    xl = np.linalg.lstsq(M, b, rcond=None)[0]
    np.testing.assert_allclose(x - x.mean(), xl - xl.mean(), atol=1e-7 * np.abs(xl).max())


def test_csr_transpose_matches_scipy():
    od = O.OracleDomain(cases.rotated_pair3d(), 0.02, 0.05)
    col, val, _ = od.advection_matrix()
    tr, tc, tv = O.csr_transpose(od.row, col, val)
    M = sp.csr_matrix((val, col, od.row), shape=(od.ncells,) * 2)
    np.testing.assert_array_equal(sp.csr_matrix((tv, tc, tr), shape=M.shape).toarray(), M.T.toarray())


# ----------------------------------------------------------------------------- AMG
def _as_sp(t, shape):
    r, c, v = t
    return sp.csr_matrix((v, c, r), shape=shape)


@pytest.mark.parametrize("which", ["pressure2d", "epot3d"])
def test_amg_hierarchy_invariants(which):
    if which == "pressure2d":
        row, col, val = _poisson(40, 32)
    else:
        od = O.OracleDomain(cases.duct3d(16, 12, 12), 0.01, 0.02)
        col, val = od.epot_matrix()
        row = od.row
    n = row.size - 1
    H = O.OracleHierarchy(row, col, val)
    L = H.levels
    assert len(L) >= 2 and L[-1]["n"] <= 64
    nnz = [lv["A"][1].size for lv in L]
    assert sum(nnz) / nnz[0] < 3.0  # operator complexity (test_amg.py:119)
    for l in range(len(L) - 1):
        nf, nc = L[l]["n"], L[l + 1]["n"]
        A = _as_sp(L[l]["A"], (nf, nf))
        P = _as_sp(L[l]["P"], (nf, nc))
        R = _as_sp(L[l]["R"], (nc, nf))
        assert abs(R - P.T).max() == 0
        Ac = _as_sp(L[l + 1]["A"], (nc, nc))
        np.testing.assert_allclose((R @ A @ P).toarray(), Ac.toarray(), atol=1e-12 * abs(A).max())
        split = L[l]["split"]
        assert split.sum() == nc and nc < nf
        # C points are injected
        cpts = np.flatnonzero(split == 1)
        Pd = P.toarray()
        np.testing.assert_array_equal(Pd[cpts], np.eye(nc))
        # constants are interpolated exactly where A has zero row sums
        zero_rows = np.abs(np.asarray(A.sum(axis=1)).ravel()) < 1e-12 * abs(A).max()
        if zero_rows.all():
            np.testing.assert_allclose(Pd.sum(axis=1), 1.0, atol=1e-12)
        # every F point with strong connections has a strong C neighbour (RS first pass)
        Ad = A.toarray()
        for i in np.flatnonzero(split == 0):
            off = -Ad[i].copy()
            off[i] = 0
            if np.abs(off).max() == 0:
                continue
            strong = np.abs(off) >= 0.25 * np.abs(off).max()
            assert (strong & (split == 1)).any(), i


def test_amg_pcg_matches_direct_and_converges_fast():
    row, col, val = _poisson(48, 40, shift=1e-3)
    n = row.size - 1
    from oracle.piso_oracle import OracleHierarchyState

    hs = OracleHierarchyState(row, col, val, project=False, hdtype=np.float64)
    b = np.random.default_rng(6).standard_normal(n)
    x = np.zeros(n)
    r = O.amg_pcg((row, col, val), hs.native, hs.pinv, b, x, 1e-10, hdtype=np.float64)
    assert r.converged and r.used_iterations < 40
    M = sp.csr_matrix((val, col, row), shape=(n, n))
    xd = spla.spsolve(M.tocsc(), b)
    assert cases.rel_l2(x, xd) < 1e-6
    # plain CG needs many more iterations
    x2 = np.zeros(n)
    r2 = O.cg(row, col, val, b, x2, 1e-10)
    assert r2.used_iterations > 2 * r.used_iterations


def test_amg_pcg_singular_projected():
    row, col, val = _poisson(32, 24)
    n = row.size - 1
    from oracle.piso_oracle import OracleHierarchyState

    hs = OracleHierarchyState(row, col, val, project=True)
    b = np.random.default_rng(8).standard_normal(n)
    b -= b.mean()
    x = np.zeros(n)
    r = O.amg_pcg((row, col, val), hs.native, hs.pinv, b, x, 1e-9, project=True, hdtype=np.float32)
    assert r.converged
    xl = np.linalg.lstsq(dense(row, col, val), b, rcond=None)[0]
    np.testing.assert_allclose(x - x.mean(), xl - xl.mean(), atol=1e-6 * np.abs(xl).max())


# ----------------------------------------------------------------------------- PISO
def test_oracle_hartmann_profile():
    """Steady Hartmann flow between insulating walls: u(y) = G/N (1 - cosh(Ha y)/cosh Ha)
    (examples/hartmann.py:85). Coarse 2D grid, periodic in x; the error is the discretisation
    error of the wall-refined grid."""
    from oracle.piso_oracle import MHDSpec, OraclePISO

    Ha, Re = 4.0, 1.0
    N = Ha ** 2 / Re
    G = N
    specs = cases.channel2d(nx=4, ny=32, stretch=1.5, noise=0.0, gradient=G)
    od = O.OracleDomain(specs, 1.0 / Re, 0.05)
    base = specs[0].vsource.copy()

    def reset_source(sim):
        sim.od.set_source(0, base)

    sim = OraclePISO(od, 0.05, mhd=MHDSpec(stuart=N, e_b=(0.0, 1.0, 0.0), potential_tol=1e-11),
                     pre_velocity_setup=reset_source, advection_tol=1e-11, pressure_tol=1e-11)
    for _ in range(150):
        sim.step()
    y = 0.5 * (specs[0].vertices[1, 1:, 0] + specs[0].vertices[1, :-1, 0])
    u = od.velocity[0].reshape(2, 32, 4)[0].mean(axis=1)
    exact = G / N * (1 - np.cosh(Ha * y) / math.cosh(Ha))
    err = np.abs(u - exact).max() / np.abs(exact).max()
    assert err < 5e-3, err


def test_advective_outflow_is_implicit_euler():
    """The outflow update is the implicit-Euler step of du/dt + U du/dn = 0 over the half cell
    between the last cell centre and the face: u' = (u + a c) / (1 + a), a = 2 U dt / dx; and
    the flux balance leaves a zero net boundary flux (and an already balanced domain unchanged)."""
    from oracle.piso_oracle import balance_boundary_fluxes, net_fixed_flux, update_advective_boundaries

    nx, ny, U, dt = 8, 6, 0.7, 3e-2
    specs = cases.outflow2d(nx=nx, ny=ny)
    od = O.OracleDomain(specs, 0.02, dt)
    rng = np.random.default_rng(4)
    od.velocity[0][...] = rng.standard_normal(od.velocity[0].shape)
    v0 = rng.standard_normal((2, ny))
    od.set_face_velocity(0, 1, v0)
    layer = od.velocity[0].reshape(2, ny, nx)[:, :, -1]
    update_advective_boundaries(od, [(0, 1)], np.array([U, 0.0]), dt, tol=1e30)  # balance off
    a = 2 * U * dt / (4.0 / nx)
    np.testing.assert_allclose(od.face_velocity(0, 1), (v0 + a * layer) / (1 + a), rtol=1e-13, atol=1e-14)
    balance_boundary_fluxes(od, [(0, 1)])
    assert abs(net_fixed_flux(od, [(0, 0), (0, 1), (0, 2), (0, 3)])) < 1e-13
    before = od.face_velocity(0, 1).copy()
    balance_boundary_fluxes(od, [(0, 1)])
    np.testing.assert_array_equal(od.face_velocity(0, 1), before)


@pytest.mark.parametrize("d", [2, 3])
def test_max_velocity_affine_closed_form(d):
    """On vertices x = L diag(h) xi every cell and face has Minv = (L diag(h))^-1, so the
    computational velocity is a linear solve per cell: the maximum over cells and fixed faces
    (static and per-face-cell velocities) of |Minv u| and |Minv u|_2, and of |u| in physical
    space; a face velocity larger than every cell's decides the maximum only with bounds."""
    rng = np.random.default_rng(11 + d)
    n = [5, 4, 3][:d]
    h = np.array([0.3, 0.7, 0.45][:d])
    L = np.eye(d) + 0.3 * rng.standard_normal((d, d))
    M = L @ np.diag(h)
    base = cases.vertices([np.arange(k + 1, dtype=np.float64) for k in n])
    v = np.einsum("ij,j...->i...", M, base)
    cz = cases.cell_shape(v)
    vel = rng.standard_normal([d] + cz)
    static = 0.5 * rng.standard_normal(d)
    varying = rng.standard_normal([d] + cases.face_shape(cz, 3))
    faces = [O.FaceSpec("periodic"), O.FaceSpec("periodic"), O.FaceSpec("fixed", velocity=static),
             O.FaceSpec("fixed", velocity=varying)] + [O.FaceSpec("periodic")] * (2 * d - 4)
    od = O.OracleDomain([O.BlockSpec(v, faces, velocity=vel)], 0.1, 1e-2)
    Minv = np.linalg.inv(M)
    cell_c = np.einsum("ik,k...->i...", Minv, vel).reshape(d, -1)
    face_c = [Minv @ static, np.einsum("ik,k...->i...", Minv, varying).reshape(d, -1)]
    assert od.max_velocity(False, False) == pytest.approx(np.abs(vel).max(), rel=1e-15)
    assert od.max_velocity(False, True) == pytest.approx(np.abs(cell_c).max(), rel=1e-12)
    assert od.max_velocity(False, True, magnitude=True) == pytest.approx(np.linalg.norm(cell_c, axis=0).max(), rel=1e-12)
    with_faces = max(np.abs(cell_c).max(), np.abs(face_c[0]).max(), np.abs(face_c[1]).max())
    assert od.max_velocity(True, True) == pytest.approx(with_faces, rel=1e-12)
    mags = max(np.linalg.norm(cell_c, axis=0).max(), np.linalg.norm(face_c[0]), np.linalg.norm(face_c[1], axis=0).max())
    assert od.max_velocity(True, True, magnitude=True) == pytest.approx(mags, rel=1e-12)
    # a fast static face decides the maximum with bounds only
    od.set_face_velocity(0, 2, 100.0 * np.eye(d)[0])
    assert od.max_velocity(True, False) == pytest.approx(100.0)
    assert od.max_velocity(True, True) == pytest.approx(np.abs(Minv[:, 0]).max() * 100.0, rel=1e-12)
    assert od.max_velocity(False, True) == pytest.approx(np.abs(cell_c).max(), rel=1e-12)
    od.velocity[0][0, 1] = np.nan
    assert np.isnan(od.max_velocity(True, True))


def test_cfl_substep_splits():
    """The substeps of the remaining time cover it exactly, each within the CFL bound, and are
    the fewest that do; a resting or slow flow takes the whole time in one substep."""
    from oracle.piso_oracle import cfl_substep

    assert cfl_substep(0.3, 0.0, 0.8) == (0.3, 1)
    assert cfl_substep(0.3, 1.0, 0.8) == (0.3, 1)
    for t, v, c in [(1.0, 7.3, 0.8), (0.25, 8.0, 0.8), (2.0, 0.41, 0.5), (0.1, 123.4, 0.3)]:
        ts, n = cfl_substep(t, v, c)
        assert ts * n == pytest.approx(t, rel=1e-15)
        assert ts * v <= c * (1 + 1e-15)
        assert n == 1 or (t / (n - 1)) * v > c


def test_oracle_adaptive_step_substeps():
    """An adaptive step of a channel flow at 1.0 (computational velocity 1/dx = 8 in x) and
    CFL 0.8 splits 0.25 into ceil(0.25 * 8 / 0.8) = 3 equal substeps at first, and the substeps
    of a step always add up to it."""
    from oracle.piso_oracle import OraclePISO

    specs = cases.channel2d(nx=16, ny=12, u0=1.0, noise=0.0, gradient=None)
    od = O.OracleDomain(specs, 0.05, 0.25)
    assert od.max_velocity(True, True) == pytest.approx(8.0, rel=1e-12)
    osim = OraclePISO(od, 0.25, advection_tol=1e-12, pressure_tol=1e-12)
    sizes = osim.adaptive_step(0.25, 0.8)
    assert len(sizes) == 3 and sizes[0] == pytest.approx(0.25 / 3, rel=1e-15)
    assert sum(sizes) == pytest.approx(0.25, rel=1e-14)
    sizes2 = osim.adaptive_step(0.25, 0.8)
    assert sum(sizes2) == pytest.approx(0.25, rel=1e-14)


def _scalar_domain(ny, U, top, bottom, dt, nx=4, visc=0.1):
    """A 2D slab (uniform grid on [0, 1]^2, periodic x unless U != 0) with one passive scalar."""
    v = cases.vertices([np.linspace(0.0, 1.0, nx + 1), np.linspace(0.0, 1.0, ny + 1)])
    cz = cases.cell_shape(v)
    if U:
        side = [O.FaceSpec("fixed", velocity=np.array([U, 0.0]), scalar=np.array([1.5])),
                O.FaceSpec("fixed", velocity=np.array([U, 0.0]), scalar_neumann=[True])]
    else:
        side = [O.FaceSpec("periodic"), O.FaceSpec("periodic")]
    faces = side + [bottom, top]
    vel = np.zeros([2] + cz)
    vel[0] = U
    spec = O.BlockSpec(v, faces, velocity=vel, pressure=np.zeros(cz),
                       scalar=np.random.default_rng(3).standard_normal([1] + cz))
    return O.OracleDomain([spec], visc, dt, scalar_visc=[visc]), v


def test_scalar_steady_linear_profile():
    """Pure diffusion between a Dirichlet wall (value a at y = 0) and a prescribed outward
    gradient g at y = 1 has the steady profile a + g y, which the finite-volume scheme represents
    exactly: one implicit step with dt = 1e12 lands on it (ghost-cell coefficient 2 alpha at the
    Dirichlet face, g times the face area at the Neumann face, their signs)."""
    a, g = 2.0, 0.5
    od, v = _scalar_domain(16, 0.0, O.FaceSpec("fixed", scalar=np.array([g]), scalar_neumann=[True]),
                           O.FaceSpec("fixed", scalar=np.array([a])), 1e12)
    col, val, _ = od.scalar_matrix(0)
    rhs = od.scalar_rhs()[0]
    n = od.ncells
    phi = spla.spsolve(sp.csr_matrix((val, col, od.row), shape=(n, n)).tocsc(), rhs)
    yc = 0.5 * (v[1, 1:, 1:] + v[1, :-1, :-1]).reshape(-1)
    np.testing.assert_allclose(phi, a + g * yc, rtol=0, atol=1e-9)


def test_scalar_steady_advected_inflow_value():
    """Advection-diffusion through a slab with a Dirichlet inflow value s, a zero-gradient
    outflow (the cell value carried out, explicitly) and zero-flux walls keeps phi = s: a step
    from phi = s returns phi = s (a wrong sign of the inflow or outflow term leaves a gradient)."""
    od, _ = _scalar_domain(6, 1.0, O.FaceSpec("fixed", scalar_neumann=[True]),
                           O.FaceSpec("fixed", scalar_neumann=[True]), 0.1, nx=10, visc=0.05)
    od.scalar[0][...] = 1.5
    col, val, _ = od.scalar_matrix(0)
    rhs = od.scalar_rhs()[0]
    n = od.ncells
    phi = spla.spsolve(sp.csr_matrix((val, col, od.row), shape=(n, n)).tocsc(), rhs)
    np.testing.assert_allclose(phi, 1.5, rtol=0, atol=1e-12)
    # and a different inflow value moves the field towards it
    od.specs[0].faces[0].scalar = np.array([2.5])
    phi2 = spla.spsolve(sp.csr_matrix((val, col, od.row), shape=(n, n)).tocsc(), od.scalar_rhs()[0])
    assert phi2.max() > 1.5 + 1e-3 and phi2.min() >= 1.5 - 1e-12


# ----------------------------------------------------------------------------- non-orthogonal terms
def _sheared(d: int, n: int, shear):
    """An affine sheared block (x = xi + s0 eta (+ s1 zeta), y = eta (+ s2 zeta), z = zeta; cell
    size 1/n): periodic along x, fixed faces elsewhere, zero velocity."""
    h = 1.0 / n
    sizes = [n + 2, n, n - 1][:d]
    grids = np.meshgrid(*[np.arange(s + 1) * h for s in reversed(sizes)], indexing="ij")  # z, y, x order
    xi = grids[-1]
    eta = grids[-2]
    if d == 2:
        v = np.stack([xi + shear[0] * eta, eta])
        J = np.array([[1.0, shear[0]], [0.0, 1.0]])
    else:
        zeta = grids[0]
        v = np.stack([xi + shear[0] * eta + shear[1] * zeta, eta + shear[2] * zeta, zeta])
        J = np.array([[1.0, shear[0], shear[1]], [0.0, 1.0, shear[2]], [0.0, 0.0, 1.0]])
    faces = [cases.FaceSpec("periodic"), cases.FaceSpec("periodic")] + [
        cases.FaceSpec("fixed", velocity=np.zeros(d)) for _ in range(2 * d - 2)]
    spec = cases.BlockSpec(v, faces, velocity=np.zeros([d] + cases.cell_shape(v)))
    return spec, J, h


def _cell_centers(spec) -> np.ndarray:
    v = spec.vertices
    d = v.shape[0]
    c = v
    for ax in range(1, d + 1):
        sl0 = [slice(None)] * (d + 1)
        sl1 = [slice(None)] * (d + 1)
        sl0[ax] = slice(0, -1)
        sl1[ax] = slice(1, None)
        c = 0.5 * (c[tuple(sl0)] + c[tuple(sl1)])
    return c.reshape(d, -1)  # [d][cells], x fastest


@pytest.mark.parametrize("d", [2, 3])
@pytest.mark.parametrize("field", ["quadratic", "wall-normal"])
def test_nonortho_laplacian_exact_for_quadratics(d, field):
    """On an affine sheared grid the implicit corner terms of the pressure and diffusion matrices
    plus the explicit diagonal-corner terms of the right-hand sides (the reference's default
    mode) form a Laplacian that is exact for quadratic fields: L(p) = det * (tr(H) p) in every
    cell whose corner stencil stays off the fixed faces. The wall-normal field depends only on
    (y, z), so it is periodic along x and the check includes the cells at the periodic seam."""
    spec, J, h = _sheared(d, 8, [0.45, -0.3, 0.25])
    od = O.OracleDomain([spec], 0.37, 0.05, np.float64)
    n = od.ncells
    X = _cell_centers(spec)
    rng = np.random.default_rng(3)
    Hm = rng.standard_normal((d, d))
    Hm = Hm + Hm.T
    g = rng.standard_normal(d)
    if field == "wall-normal":
        Hm[0, :] = 0
        Hm[:, 0] = 0
        g[0] = 0
    p = 0.5 * np.einsum("ic,ij,jc->c", X, Hm, X) + g @ X
    lap = np.trace(Hm) * h ** d  # det * laplacian; the shear keeps the cell volume h^d
    idx = np.indices(list(reversed(spec.size_xyz[:d]))).reshape(d, -1)[::-1]  # x, y, (z) per cell
    size = spec.size_xyz
    ok = np.ones(n, bool)
    for ax in range(1, d):
        ok &= (idx[ax] >= 2) & (idx[ax] <= size[ax] - 3)
    if field == "quadratic":
        ok &= (idx[0] >= 2) & (idx[0] <= size[0] - 3)
    flags = O.NO_DEFAULT_MODE
    # pressure: P p - S(p), with A = 1
    ones = np.ones(n)
    col, val = od.pressure_matrix(ones, flags)
    Lp = O.spmv(od.row, col, val, p) - od.nonortho_rhs(O.OK_PRES, flags, 0, p, ones)
    np.testing.assert_allclose(Lp[ok], lap, rtol=1e-9, atol=1e-12)
    if field == "quadratic":
        # without the corner terms the operator is not exact on the sheared grid (for the
        # wall-normal field the cross terms cancel between opposite faces)
        col0, val0 = od.pressure_matrix(ones, 0)
        assert np.abs(O.spmv(od.row, col0, val0, p)[ok] - lap).max() > 1e-3 * abs(lap)
    # diffusion of each velocity component: det/dt u - det (C u) - S_vel(u) = nu L(u)
    det = h ** d
    u = np.concatenate([p * (k + 1) for k in range(d)])
    ccol, cval, _ = od.advection_matrix(flags)
    for k in range(d):
        uk = u[k * n:(k + 1) * n]
        Lu = det / od.dom.dt * uk - det * O.spmv(od.row, ccol, cval, uk) - od.nonortho_rhs(O.OK_VEL, flags, k, u)
        np.testing.assert_allclose(Lu[ok], 0.37 * lap * (k + 1), rtol=1e-9, atol=1e-12)


@pytest.mark.parametrize("setup", ["oracle", "library"])
def test_ruge_stueben_1d_laplacian_closed_form(setup):
    """Classical Ruge-Stueben on the 1D Dirichlet Laplacian tridiag(-1, 2, -1) with 2^8 - 1
    points: every off-diagonal is strong, so the first pass alternates C and F points (F at both
    ends); an F point interpolates 1/2 from each C neighbour (-a_ij / a_ii times the ratio of its
    off-diagonal sum to the C part, both -2 inside and -1 at the ends); so every coarse
    operator R A P is again the Laplacian, halved per level: 2^-l tridiag(-1, 2, -1) on 2^(8-l) - 1
    points, until at most 64 remain. Independent of both restatements of the setup; a wrong
    weight, a dropped normalisation or a transposed R fails it."""
    n = 2 ** 8 - 1
    i = np.arange(n)
    rows = np.concatenate([i, i[1:], i[:-1]])
    cols = np.concatenate([i, i[1:] - 1, i[:-1] + 1])
    vals = np.concatenate([np.full(n, 2.0), np.full(n - 1, -1.0), np.full(n - 1, -1.0)])
    order = np.lexsort((cols, rows))
    row = np.concatenate([[0], np.cumsum(np.bincount(rows, minlength=n))]).astype(np.int32)
    col, val = cols[order].astype(np.int32), vals[order]
    if setup == "oracle":
        levels = O.OracleHierarchy(row, col, val).levels
    else:
        from phipict_amd.solvers.amg import rs_setup_host

        levels = rs_setup_host(row, col, val)
    assert [lv["n"] for lv in levels] == [255, 127, 63]
    for l, lv in enumerate(levels):
        m = lv["n"]
        A = _as_sp(lv["A"], (m, m)).toarray()
        lap = 2.0 ** -l * (2 * np.eye(m) - np.eye(m, k=1) - np.eye(m, k=-1))
        np.testing.assert_allclose(A, lap, rtol=0, atol=1e-14)
        if l == len(levels) - 1:
            break
        split = np.asarray(lv["split"])
        np.testing.assert_array_equal(split, np.arange(m) % 2)  # C at the odd points
        mc = levels[l + 1]["n"]
        P = _as_sp(lv["P"], (m, mc)).toarray()
        expect = np.zeros((m, mc))
        for j in range(mc):
            c = 2 * j + 1
            expect[c, j] = 1.0
            expect[c - 1, j] = 0.5
            expect[c + 1, j] = 0.5
        np.testing.assert_allclose(P, expect, rtol=0, atol=1e-15)
        np.testing.assert_array_equal(_as_sp(lv["R"], (mc, m)).toarray(), P.T)

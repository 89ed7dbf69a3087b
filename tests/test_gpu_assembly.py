"""GPU parity of the assembly kernels (through the C ABI) with the oracle on the same seeded
domains: CSR pattern (exact), advection matrix and diagonal, momentum RHS, pressure matrix,
HbyA RHS and divergence, FD velocity correction, Epot matrix and RHS, face-based current
density, result copies and grid metrics; FP64 and FP32; single and batched environments."""
from __future__ import annotations

import numpy as np
import pytest
import torch

import cases
from cases import build_domain, rel_l2, to_np
from oracle import oracle as O

pytestmark = pytest.mark.gpu

NU, DT = 0.013, 0.021
TOL = {np.float64: 1e-12, np.float32: 2e-6}


def _tdt(dtype):
    return torch.float64 if dtype == np.float64 else torch.float32


def close(a, b, dtype, what):
    a = np.asarray(a, np.float64).reshape(-1)
    b = np.asarray(b, np.float64).reshape(-1)
    assert a.shape == b.shape, what
    err = rel_l2(a, b)
    assert err <= TOL[dtype], f"{what}: rel L2 {err:.3e}"
    scale = max(np.abs(b).max(), 1e-30)
    bad = np.abs(a - b) > (TOL[dtype] * 50) * scale + (TOL[dtype] * 10) * np.abs(b)
    assert not bad.any(), f"{what}: {bad.sum()} entries differ, worst {np.abs(a - b).max():.3e} of scale {scale:.3e}"


@pytest.fixture(params=[np.float64, np.float32], ids=["f64", "f32"])
def dtype(request):
    return request.param


@pytest.mark.parametrize("name", list(cases.CASES))
def test_assembly_matches_oracle(name, dtype):
    print(f"[assembly] case={name} dtype={np.dtype(dtype).name}", flush=True)
    import phipict_amd as pa

    specs = cases.CASES[name]()
    od = O.OracleDomain(specs, NU, DT, dtype)
    dom = build_domain(specs, NU, DT, _tdt(dtype), epot=True)
    ts = torch.tensor([DT], dtype=_tdt(dtype))
    n, d = od.ncells, od.dims
    # pattern
    np.testing.assert_array_equal(to_np(dom.C.row), od.row)
    pa.SetupAdvectionMatrix(dom, ts, 0)
    ccol, cval, A = od.advection_matrix()
    np.testing.assert_array_equal(to_np(dom.C.index), ccol)
    close(to_np(dom.C.value), cval, dtype, "advection matrix")
    close(to_np(dom.A), A, dtype, "advection diagonal")
    for apg in (False, True):
        pa.SetupAdvectionVelocity(dom, ts, 0, apg)
        close(to_np(dom.velocityRHS), od.advection_rhs(apg), dtype, f"momentum rhs (pressure gradient {apg})")
    pa.SetupPressureMatrix(dom, ts, 0)
    pcol, pval = od.pressure_matrix(A)
    np.testing.assert_array_equal(to_np(dom.P.index), pcol)
    close(to_np(dom.P.value), pval, dtype, "pressure matrix")
    rng = np.random.default_rng(11)
    vres = rng.standard_normal(d * n).astype(dtype)
    dom.setVelocityResult(torch.as_tensor(vres, device=dom.getDevice()))
    pa.SetupPressureRHS(dom, ts, 0)
    prhs = od.pressure_rhs(A, ccol, cval, vres)
    close(to_np(dom.pressureRHS), prhs, dtype, "HbyA")
    close(to_np(dom.pressureRHSdiv), od.pressure_div(prhs), dtype, "HbyA divergence")
    for tsn in (False, True):
        pa.SetupPressureRHSdiv(dom, ts, 0, False, timeStepNorm=tsn)
        close(to_np(dom.pressureRHSdiv), od.pressure_div(prhs, tsn), dtype, f"divergence (time-step norm {tsn})")
        pa.CorrectVelocity(dom, ts, version=1, timeStepNorm=tsn)
        close(to_np(dom.velocityResult), od.correct_velocity(A, prhs, tsn), dtype, f"velocity correction ({tsn})")
    ecol, evalv = od.epot_matrix()
    np.testing.assert_array_equal(to_np(dom.Epot.index), ecol)
    close(to_np(dom.Epot.value), evalv, dtype, "Epot matrix")
    vf = rng.standard_normal(d * n).astype(dtype)
    close(to_np(pa.ComputeEpotRHS(dom, torch.as_tensor(vf, device=dom.getDevice()))), od.epot_rhs(vf), dtype, "Epot rhs")
    phi = rng.standard_normal(n).astype(dtype)
    uxb = rng.standard_normal(3 * n).astype(dtype)
    J = pa.ComputeCurrentDensityFaceBased(dom, torch.as_tensor(phi, device=dom.getDevice()),
                                          torch.as_tensor(uxb, device=dom.getDevice()))
    close(to_np(J), od.current_density(phi, uxb), dtype, "current density")


def test_result_copies(dtype):
    import phipict_amd as pa

    specs = cases.rotated_pair3d()
    od = O.OracleDomain(specs, NU, DT, dtype)
    dom = build_domain(specs, NU, DT, _tdt(dtype), epot=True)
    n, d = od.ncells, od.dims
    flat_v = od.flat_from_blocks(od.velocity, d)
    pa.CopyVelocityResultFromBlocks(dom)
    np.testing.assert_array_equal(to_np(dom.velocityResult), flat_v)
    flat_p = od.flat_from_blocks(od.pressure, 1)
    pa.CopyPressureResultFromBlocks(dom)
    np.testing.assert_array_equal(to_np(dom.pressureResult), flat_p)
    rng = np.random.default_rng(3)
    new_v = rng.standard_normal(d * n).astype(dtype)
    dom.setVelocityResult(torch.as_tensor(new_v, device=dom.getDevice()))
    pa.CopyVelocityResultToBlocks(dom)
    got = np.concatenate([to_np(b.velocity).reshape(d, -1) for b in dom.getBlocks()], axis=1).reshape(-1)
    np.testing.assert_array_equal(got, new_v)
    new_e = rng.standard_normal(n).astype(dtype)
    dom.setEpotResult(torch.as_tensor(new_e, device=dom.getDevice()))
    pa.CopyEpotResultToBlocks(dom)
    got = np.concatenate([to_np(b.epot).reshape(-1) for b in dom.getBlocks()])
    np.testing.assert_array_equal(got, new_e)


@pytest.mark.parametrize("d", [2, 3])
def test_metrics_match_oracle(d, dtype):
    import phipict_amd as pa

    axes = [np.linspace(0, 1, 7), cases.tanh_coords(5, 0, 2, 1.3), np.linspace(0, 0.5, 4)][:d]
    v = cases.skewed_vertices(axes, 0.2, seed=d)
    dev = torch.device("cuda", 0)
    coords = torch.as_tensor(v, dtype=_tdt(dtype), device=dev).unsqueeze(0)
    close(to_np(pa.CoordsToTransforms(coords)).reshape(-1), O.transforms(v, dtype).reshape(-1), dtype, "transforms")
    close(to_np(pa.CoordsToFaceTransforms(coords)).reshape(-1), O.face_transforms(v, dtype).reshape(-1), dtype,
          "face transforms")


def test_batched_assembly_per_env(dtype):
    """Two environments with different velocities, sources and time steps assemble what two
    separate oracle domains assemble (a shared pattern, per-environment values)."""
    import phipict_amd as pa

    specs = cases.channel2d(12, 9, per_cell_source=True)
    rng = np.random.default_rng(21)
    specs2 = cases.channel2d(12, 9, per_cell_source=True, seed=5)
    dts = [DT, 1.7 * DT]
    dom = build_domain(specs, NU, DT, _tdt(dtype), batch=2)
    blk = dom.getBlocks()[0]
    v = blk.velocity.clone()
    v[1] = torch.as_tensor(specs2[0].velocity, dtype=v.dtype)
    blk.setVelocity(v)
    src = torch.stack([torch.as_tensor(specs[0].vsource), torch.as_tensor(specs2[0].vsource)]).to(v)
    blk.setVelocitySource(src)
    dom.setTimeStep(torch.tensor(dts, dtype=_tdt(dtype)))
    pa.SetupAdvectionMatrix(dom, None, 0)
    pa.SetupAdvectionVelocity(dom, None, 0, False)
    pa.SetupPressureMatrix(dom, None, 0)
    n = dom.getTotalSize()
    vres = rng.standard_normal(2 * 2 * n).astype(dtype)
    dom.setVelocityResult(torch.as_tensor(vres, device=dom.getDevice()))
    pa.SetupPressureRHS(dom, None, 0)
    nnz = dom.C.getNnz()
    for e, (sp_, dt_) in enumerate(zip([specs, specs2], dts)):
        od = O.OracleDomain(sp_, NU, dt_, dtype)
        ccol, cval, A = od.advection_matrix()
        close(to_np(dom.C.value)[e * nnz:(e + 1) * nnz], cval, dtype, f"env {e} advection matrix")
        close(to_np(dom.A)[e * n:(e + 1) * n], A, dtype, f"env {e} diagonal")
        close(to_np(dom.velocityRHS)[e * 2 * n:(e + 1) * 2 * n], od.advection_rhs(), dtype, f"env {e} rhs")
        pcol, pval = od.pressure_matrix(A)
        close(to_np(dom.P.value)[e * nnz:(e + 1) * nnz], pval, dtype, f"env {e} pressure matrix")
        vr = vres[e * 2 * n:(e + 1) * 2 * n]
        prhs = od.pressure_rhs(A, ccol, cval, vr)
        close(to_np(dom.pressureRHS)[e * 2 * n:(e + 1) * 2 * n], prhs, dtype, f"env {e} HbyA")
        close(to_np(dom.pressureRHSdiv)[e * n:(e + 1) * n], od.pressure_div(prhs), dtype, f"env {e} divergence")


NONORTHO_CASES = ["channel2d", "duct3d", "inflow2d_varying", "split2d", "rotated3d", "tjunction3d", "heated2d"]


def test_hbya_reads_values_written_after_the_assembly(dtype):
    """HbyA reads C through the ELL copy the assembly writes; values written into C afterwards
    (a hook adding terms, an adjoint's perturbation) reach it as they reach the reference's CSR
    read (pressure.cu:349-522)."""
    import phipict_amd as pa

    specs = cases.inflow2d()
    od = O.OracleDomain(specs, NU, DT, dtype)
    dom = build_domain(specs, NU, DT, _tdt(dtype))
    ts = torch.tensor([DT], dtype=_tdt(dtype))
    pa.SetupAdvectionMatrix(dom, ts, 0)
    ccol, cval, A = od.advection_matrix()
    cval = (cval * (1.0 + 0.1 * np.random.default_rng(2).standard_normal(cval.size))).astype(dtype)
    dom.C.value.copy_(torch.as_tensor(cval, device=dom.getDevice()))
    vres = np.random.default_rng(3).standard_normal(od.dims * od.ncells).astype(dtype)
    dom.setVelocityResult(torch.as_tensor(vres, device=dom.getDevice()))
    pa.SetupPressureRHS(dom, ts, 0)
    close(to_np(dom.pressureRHS), od.pressure_rhs(A, ccol, cval, vres), dtype, "HbyA with modified C")


@pytest.mark.parametrize("name", list(cases.CASES))
def test_linear_upwind_assembly_matches_oracle(name, dtype):
    """Domain.setAdvectionScheme(LINEAR_UPWIND) (advection.cu:129-140, 893-937, 650-684): the
    upwinded advection matrix and diagonal, the momentum right-hand side with the deferred
    correction, and for passive scalars their matrices and right-hand sides, against the
    oracle."""
    print(f"[assembly] LINEAR_UPWIND case={name} dtype={np.dtype(dtype).name}", flush=True)
    import phipict_amd as pa

    specs = cases.CASES[name]()
    od = O.OracleDomain(specs, NU, DT, dtype, advection_scheme=O.ADV_LINEAR_UPWIND)
    dom = build_domain(specs, NU, DT, _tdt(dtype))
    dom.setAdvectionScheme(pa.AdvectionScheme.LINEAR_UPWIND)
    assert dom.getAdvectionScheme() == pa.AdvectionScheme.LINEAR_UPWIND
    ts = torch.tensor([DT], dtype=_tdt(dtype))
    pa.SetupAdvectionMatrix(dom, ts, 0)
    ccol, cval, A = od.advection_matrix()
    close(to_np(dom.C.value), cval, dtype, "advection matrix")
    close(to_np(dom.A), A, dtype, "advection diagonal")
    pa.SetupAdvectionVelocity(dom, ts, 0, False)
    close(to_np(dom.velocityRHS), od.advection_rhs(False), dtype, "momentum rhs")
    central = O.OracleDomain(specs, NU, DT, dtype)
    assert rel_l2(od.advection_rhs(False), central.advection_rhs(False)) > 0
    if specs[0].scalar is not None:
        C_ = od.channels
        for ch in range(C_):
            pa.SetupAdvectionMatrix(dom, ts, 0, forPassiveScalar=True, passiveScalarChannel=ch)
            _, sval, _ = od.scalar_matrix(ch)
            close(to_np(dom.C.value), sval, dtype, f"scalar matrix {ch}")
        pa.SetupAdvectionScalar(dom, ts, 0)
        close(to_np(dom.scalarRHS), od.scalar_rhs().reshape(-1), dtype, "scalar rhs")


@pytest.mark.parametrize("flags", [O.NO_DEFAULT_MODE, O.NO_DEFAULT_MODE | O.NO_DIRECT_RHS, O.NO_DIRECT_MATRIX])
@pytest.mark.parametrize("name", NONORTHO_CASES)
def test_nonortho_assembly_matches_oracle(name, flags, dtype):
    """Skewed grids (jittered interior vertices): the corner terms of the advection, scalar and
    pressure matrices and the explicit corner terms of the momentum, scalar and pressure
    right-hand sides (walks through periodic seams, block connections with crossed axes and
    fixed faces) against the oracle, on random result vectors."""
    print(f"[nonortho] case={name} flags={flags} dtype={np.dtype(dtype).name}", flush=True)
    import phipict_amd as pa

    specs = cases.skew(cases.CASES[name](), 0.25, seed=7)
    od = O.OracleDomain(specs, NU, DT, dtype)
    dom = build_domain(specs, NU, DT, _tdt(dtype))
    assert not dom.isOrthogonal()
    ts = torch.tensor([DT], dtype=_tdt(dtype))
    n, d = od.ncells, od.dims
    dev = dom.getDevice()
    rng = np.random.default_rng(5)
    pa.SetupAdvectionMatrix(dom, ts, flags)
    ccol, cval, A = od.advection_matrix(flags)
    close(to_np(dom.C.value), cval, dtype, "advection matrix")
    close(to_np(dom.A), A, dtype, "advection diagonal")
    vres = rng.standard_normal(d * n).astype(dtype)
    dom.setVelocityResult(torch.as_tensor(vres, device=dev))
    pa.SetupAdvectionVelocity(dom, ts, flags, False)
    close(to_np(dom.velocityRHS), od.advection_rhs(False, flags, vres), dtype, "momentum rhs")
    pa.SetupPressureMatrix(dom, ts, flags)
    pcol, pval = od.pressure_matrix(A, flags)
    close(to_np(dom.P.value), pval, dtype, "pressure matrix")
    pres = rng.standard_normal(n).astype(dtype)
    dom.setPressureResult(torch.as_tensor(pres, device=dev))
    pa.SetupPressureRHS(dom, ts, flags)
    prhs = od.pressure_rhs(A, ccol, cval, vres)
    div = od.pressure_div(prhs)
    if flags & (O.NO_DIRECT_RHS | O.NO_DIAGONAL_RHS):
        div = div + od.nonortho_rhs(O.OK_PRES, flags, 0, pres, A)
    close(to_np(dom.pressureRHSdiv), div, dtype, "pressure rhs divergence with corner terms")
    pa.SetupPressureRHSdiv(dom, ts, flags)
    close(to_np(dom.pressureRHSdiv), div, dtype, "divergence with corner terms")
    if specs[0].scalar is not None:
        C_ = od.channels
        for ch in range(C_):
            pa.SetupAdvectionMatrix(dom, ts, flags, forPassiveScalar=True, passiveScalarChannel=ch)
            _, sval, sA = od.scalar_matrix(ch, flags)
            close(to_np(dom.C.value), sval, dtype, f"scalar matrix {ch}")
        sres = rng.standard_normal(C_ * n).astype(dtype)
        dom.setScalarResult(torch.as_tensor(sres, device=dev))
        pa.SetupAdvectionScalar(dom, ts, flags)
        close(to_np(dom.scalarRHS), od.scalar_rhs(flags, sres.reshape(C_, n)).reshape(-1), dtype, "scalar rhs")


def test_nonortho_flags_skip_exactly_orthogonal_grids():
    """Rectilinear grids take the orthogonal kernels (their corner terms are exactly zero), a
    skewed grid does not, and the potential matrix's corner terms are refused while the
    advection diagonal they weigh the cells with is not assembled."""
    import phipict_amd as pa

    rect = build_domain(cases.channel2d(8, 6), NU, DT, torch.float64, epot=True)
    assert rect.isOrthogonal() and rect.nonOrthoTermsVanish()
    ts = torch.tensor([DT], dtype=torch.float64)
    pa.SetupAdvectionMatrix(rect, ts, 0)
    c0 = rect.C.value.clone()
    pa.SetupAdvectionMatrix(rect, ts, O.NO_DEFAULT_MODE)
    assert torch.equal(c0, rect.C.value)
    pa.SetupEpotMatrix(rect, O.NO_DEFAULT_MODE)
    skewed = build_domain(cases.skew(cases.channel2d(8, 6), 0.2, seed=1), NU, DT, torch.float64, epot=True)
    assert not skewed.nonOrthoTermsVanish()
    with pytest.raises(ValueError):
        pa.SetupEpotMatrix(skewed, O.NO_DEFAULT_MODE)


@pytest.mark.parametrize("name", ["potbox2d", "potbox3d", "tjunction3d", "rotated3d"])
def test_epot_matrix_corner_terms(dtype, name):
    """The potential matrix with its non-orthogonal corner terms on skewed grids (potential.cu:
    103-150: the face coefficients weigh both cells with 1/A of the advection diagonal, the
    orthogonal part with 1; walks through connections, periodic seams and fixed faces, thin-wall
    and Dirichlet potential faces) against the oracle, for each matrix flag."""
    print(f"[nonortho] epot case={name} dtype={np.dtype(dtype).name}", flush=True)
    import phipict_amd as pa

    specs = cases.skew(cases.CASES[name](), 0.25, seed=3)
    od = O.OracleDomain(specs, NU, DT, dtype)
    dom = build_domain(specs, NU, DT, _tdt(dtype), epot=True)
    assert not dom.nonOrthoTermsVanish()
    ts = torch.tensor([DT], dtype=_tdt(dtype))
    for flags in (O.NO_DIRECT_MATRIX, O.NO_CENTER_MATRIX, O.NO_DEFAULT_MODE):
        pa.SetupAdvectionMatrix(dom, ts, flags)
        _, _, A = od.advection_matrix(flags)
        pa.SetupEpotMatrix(dom, flags)
        ecol, evalv = od.epot_matrix(flags, A)
        np.testing.assert_array_equal(to_np(dom.Epot.index), ecol)
        close(to_np(dom.Epot.value), evalv, dtype, f"Epot matrix with corner terms (flags {flags})")
        plain = od.epot_matrix()[1]
        assert np.abs(evalv - plain).max() > 1e-6 * np.abs(plain).max()


@pytest.mark.parametrize("dims", [2, 3])
@pytest.mark.parametrize("per_env", [False, True])
def test_mhd_cross_products(dtype, dims, per_env):
    """pp_cross_eb (u x e_b lifted to 3D) and pp_lorentz_source (N (J x e_b)[:d] over a block
    plus its static, per-cell or absent source) against numpy's cross product."""
    from phipict_amd import _lib

    td = _tdt(dtype)
    dev = torch.device("cuda", 0)
    rng = np.random.default_rng(11)
    B, total, g0, cells = 2, 1000, 137, 611
    v = rng.standard_normal((B, dims, total))
    eb = rng.standard_normal((B, 3)) if per_env else rng.standard_normal(3)
    ebB = eb if per_env else np.broadcast_to(eb, (B, 3))
    v3 = np.concatenate([v, np.zeros((B, 3 - dims, total))], axis=1)
    want = np.stack([np.cross(v3[b].T, ebB[b]).T for b in range(B)])
    vt = torch.as_tensor(v, dtype=td, device=dev).contiguous()
    ebt = torch.as_tensor(np.ascontiguousarray(eb), dtype=td, device=dev)
    out = torch.empty(B * 3 * total, dtype=td, device=dev)
    s = _lib.stream_ptr(dev)
    _lib.call("pp_cross_eb", _lib.ptr(vt), dims, total, B, _lib.ptr(ebt), int(per_env), _lib.ptr(out),
              _lib.dtype_code(td), s)
    close(to_np(out).reshape(B, 3, total), want, dtype, "u x e_b")

    J = rng.standard_normal((B, 3, total))
    N = rng.uniform(1, 30, B) if per_env else np.array([17.0])
    NB = N if per_env else np.repeat(N, B)
    force = np.stack([NB[b] * np.cross(J[b].T, ebB[b]).T for b in range(B)])[:, :dims, g0:g0 + cells]
    Jt = torch.as_tensor(J, dtype=td, device=dev).contiguous()
    Nt = torch.as_tensor(N, dtype=td, device=dev)
    sources = {0: (None, 0, np.zeros((1, dims, 1))),
               1: (rng.standard_normal((B if per_env else 1, dims)), 1, None),
               2: (rng.standard_normal((B if per_env else 1, dims, cells)), 2, None)}
    for kind, (S, k, _) in sources.items():
        St = torch.as_tensor(S, dtype=td, device=dev).contiguous() if S is not None else None
        o = torch.empty(B * dims * cells, dtype=td, device=dev)
        _lib.call("pp_lorentz_source", _lib.ptr(Jt), total, B, dims, _lib.ptr(ebt), int(per_env), _lib.ptr(Nt),
                  int(per_env), g0, cells, _lib.ptr(St) if St is not None else None, k,
                  int(S is not None and S.shape[0] > 1), _lib.ptr(o), _lib.dtype_code(td), s)
        exp = force.copy()
        if kind == 1:
            exp = exp + S[:, :, None]
        elif kind == 2:
            exp = exp + S
        close(to_np(o).reshape(B, dims, cells), exp, dtype, f"Lorentz source kind {kind}")
    # a block reaching past the domain is refused before any launch
    with pytest.raises(RuntimeError):
        _lib.call("pp_lorentz_source", _lib.ptr(Jt), total, B, dims, _lib.ptr(ebt), 0, _lib.ptr(Nt), 0, g0, total,
                  None, 0, 0, _lib.ptr(out), _lib.dtype_code(td), s)


def test_nonortho_batched_per_env(dtype):
    """Two environments of a skewed channel with different velocity and pressure results: the
    corner terms of each environment's matrices and right-hand sides equal those of a single
    oracle domain with that environment's state."""
    import phipict_amd as pa

    flags = O.NO_DEFAULT_MODE | O.NO_DIRECT_RHS
    specs = cases.skew(cases.channel2d(12, 9), 0.25, seed=2)
    specs2 = cases.skew(cases.channel2d(12, 9, seed=5), 0.25, seed=2)
    dom = build_domain(specs, NU, DT, _tdt(dtype), batch=2)
    blk = dom.getBlocks()[0]
    v = blk.velocity.clone()
    v[1] = torch.as_tensor(specs2[0].velocity, dtype=v.dtype)
    blk.setVelocity(v)
    dom.UpdateDomainData()
    ts = torch.tensor([DT, DT], dtype=_tdt(dtype))
    n = dom.getTotalSize()
    rng = np.random.default_rng(31)
    vres = rng.standard_normal(2 * 2 * n).astype(dtype)
    pres = rng.standard_normal(2 * n).astype(dtype)
    pa.SetupAdvectionMatrix(dom, ts, flags)
    dom.setVelocityResult(torch.as_tensor(vres, device=dom.getDevice()))
    pa.SetupAdvectionVelocity(dom, ts, flags, False)
    pa.SetupPressureMatrix(dom, ts, flags)
    dom.setPressureResult(torch.as_tensor(pres, device=dom.getDevice()))
    pa.SetupPressureRHS(dom, ts, flags)
    nnz = dom.C.getNnz()
    for e, sp_ in enumerate([specs, specs2]):
        od = O.OracleDomain(sp_, NU, DT, dtype)
        ccol, cval, A = od.advection_matrix(flags)
        close(to_np(dom.C.value)[e * nnz:(e + 1) * nnz], cval, dtype, f"env {e} advection matrix")
        vr = vres[e * 2 * n:(e + 1) * 2 * n]
        close(to_np(dom.velocityRHS)[e * 2 * n:(e + 1) * 2 * n], od.advection_rhs(False, flags, vr), dtype,
              f"env {e} momentum rhs")
        pcol, pval = od.pressure_matrix(A, flags)
        close(to_np(dom.P.value)[e * nnz:(e + 1) * nnz], pval, dtype, f"env {e} pressure matrix")
        div = od.pressure_div(od.pressure_rhs(A, ccol, cval, vr)) + od.nonortho_rhs(O.OK_PRES, flags, 0,
                                                                                     pres[e * n:(e + 1) * n], A)
        close(to_np(dom.pressureRHSdiv)[e * n:(e + 1) * n], div, dtype, f"env {e} divergence with corner terms")

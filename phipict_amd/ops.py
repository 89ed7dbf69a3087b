"""The reference's ``phipict._C`` operator functions for the PISO path, backed by the gfx950
library. Names, argument order, defaults and return values follow
csrc/phipict/bindings.cpp:639-821; every call runs on the current torch stream.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Optional, Sequence

import torch

from . import _lib
from .domain import ConvergenceCriterion, CSRmatrix, Domain

# systems up to this many rows run the whole solve in one workgroup per RHS group
SMALL_SYSTEM_ROWS = 32768

# non-orthogonal flag bits (csrc/phipict/piso/non_ortho_flags.h)
NON_ORTHO_DIRECT_MATRIX = 1
NON_ORTHO_DIRECT_RHS = 2
NON_ORTHO_DIAGONAL_MATRIX = 4
NON_ORTHO_DIAGONAL_RHS = 8
NON_ORTHO_CENTER_MATRIX = 16


def _prep(domain: Domain, timeStep=None) -> _lib.DomainDesc:
    if not domain.IsInitialized():
        raise RuntimeError("Domain is not initialized.")
    if domain.getNumBlocks() == 0:
        raise RuntimeError("Domain does not contain blocks.")
    if timeStep is not None:
        domain.setTimeStep(timeStep)
    return domain.desc()


def _flags(domain: Domain, nonOrthoFlags: int) -> int:
    """The non-orthogonal flags the kernels get: 0 when the terms are exactly zero on this
    grid (every metric diagonal), so rectilinear grids take the kernels' orthogonal path."""
    nonOrthoFlags = int(nonOrthoFlags)
    if nonOrthoFlags and domain.nonOrthoTermsVanish():
        return 0
    return nonOrthoFlags


def _stream(domain: Domain) -> int:
    return _lib.stream_ptr(domain.getDevice())


def SetupAdvectionMatrix(domain: Domain, timeStep, nonOrthoFlags: int = 0, forPassiveScalar: bool = False,
                         passiveScalarChannel: int = 0) -> None:
    """Implicit advection-diffusion matrix C and its diagonal A (bindings.cpp:639-641)."""
    d = _prep(domain, timeStep)
    flags = _flags(domain, nonOrthoFlags)
    if forPassiveScalar:
        if not domain.hasPassiveScalar():
            raise RuntimeError("the domain has no passive scalar")
        _lib.call("pp_setup_advection_matrix_scalar", C.byref(d), flags, int(passiveScalarChannel),
                  _stream(domain))
    else:
        _lib.call("pp_setup_advection_matrix", C.byref(d), flags, _stream(domain))
    domain.C.mark_ell_valid()
    domain._a_version += 1  # A changed: a pressure matrix built from the old one is stale


def SetupAdvectionScalar(domain: Domain, timeStep, nonOrthoFlags: int = 0) -> None:
    """Passive-scalar right-hand side of every channel into ``domain.scalarRHS``
    (bindings.cpp:642-644 -> advection.cu:1301-1373)."""
    if not domain.hasPassiveScalar():
        raise RuntimeError("the domain has no passive scalar")
    d = _prep(domain, timeStep)
    _lib.call("pp_setup_advection_scalar", C.byref(d), _flags(domain, nonOrthoFlags), _stream(domain))


def SetupAdvectionVelocity(domain: Domain, timeStep, nonOrthoFlags: int = 0, applyPressureGradient: bool = False) -> None:
    """Momentum right-hand side into ``domain.velocityRHS`` (bindings.cpp:645-647)."""
    d = _prep(domain, timeStep)
    _lib.call("pp_setup_advection_velocity", C.byref(d), _flags(domain, nonOrthoFlags), int(bool(applyPressureGradient)),
              _stream(domain))


def SetupPressureMatrix(domain: Domain, timeStep, nonOrthoMode: int = 0, useFaceTransform: bool = False) -> None:
    if useFaceTransform:
        raise NotImplementedError("face-transform pressure Laplacians are not implemented")
    d = _prep(domain, timeStep)
    _lib.call("pp_setup_pressure_matrix", C.byref(d), _flags(domain, nonOrthoMode), _stream(domain))
    domain.P.mark_ell_valid()
    domain._p_matrix_key = pressure_matrix_key(domain, nonOrthoMode)


def pressure_matrix_key(domain: Domain, nonOrthoMode: int) -> tuple:
    """What the pressure matrix's values depend on besides the (static) metrics: the advection
    diagonal's version, the effective flags, and the value tensor (replaced or written since)."""
    P, A = domain.P, domain.A
    return (domain._a_version, A.data_ptr(), A._version, _flags(domain, nonOrthoMode), P.value.data_ptr(),
            P.value._version, domain.getBatchSize())


def SetupPressureRHS(domain: Domain, timeStep, nonOrthoMode: int = 0, useFaceTransform: bool = False,
                     timeStepNorm: bool = False) -> None:
    if useFaceTransform:
        raise NotImplementedError("face-transform fluxes are not implemented")
    if domain.C is not None:
        # HbyA reads C through its ELL copy: values written since the assembly are copied over
        domain.C.refresh_ell()
    d = _prep(domain, timeStep)
    _lib.call("pp_setup_pressure_rhs", C.byref(d), _flags(domain, nonOrthoMode), int(bool(timeStepNorm)), _stream(domain))


def SetupPressureRHSdiv(domain: Domain, timeStep, nonOrthoMode: int = 0, useFaceTransform: bool = False,
                        timeStepNorm: bool = False) -> None:
    if useFaceTransform:
        raise NotImplementedError("face-transform fluxes are not implemented")
    d = _prep(domain, timeStep)
    _lib.call("pp_setup_pressure_rhs_div", C.byref(d), _flags(domain, nonOrthoMode), int(bool(timeStepNorm)),
              _stream(domain))


def SetupPressureCorrection(domain: Domain, timeStep, nonOrthoMode: int = 0, useFaceTransform: bool = False,
                            timeStepNorm: bool = False) -> None:
    """Matrix, HbyA right-hand side and its divergence (pressure.cu:882-945)."""
    SetupPressureMatrix(domain, timeStep, nonOrthoMode, useFaceTransform)
    SetupPressureRHS(domain, timeStep, nonOrthoMode, useFaceTransform, timeStepNorm)


def CorrectVelocity(domain: Domain, timeStep, version: int = 0, timeStepNorm: bool = False) -> None:
    """u = HbyA - grad(p)/A with finite-difference gradients (versions 0 and 1,
    velocity_correction.cu:216-271)."""
    if version not in (0, 1):
        raise NotImplementedError("only the finite-difference velocity corrector (version 0/1) is implemented")
    d = _prep(domain, timeStep)
    _lib.call("pp_correct_velocity", C.byref(d), int(bool(timeStepNorm)), _stream(domain))


def SetupEpotMatrix(domain: Domain, nonOrthoFlags: int = 0, useFaceTransform: bool = False) -> None:
    if domain.Epot is None:
        raise RuntimeError("domain.Epot is not allocated. Call domain.SetupEpotOnDomain() first.")
    flags = _flags(domain, nonOrthoFlags)
    if flags and (domain.A is None or not bool((domain.A != 0).all())):
        # on an A not yet assembled the reference's matrix would hold infinities
        raise ValueError("the potential's non-orthogonal corner terms weigh the cells with 1/A (potential.cu:"
                         "103-150): assemble the advection matrix (domain.A) first")
    d = _prep(domain, None)
    _lib.call("pp_setup_epot_matrix", C.byref(d), int(flags), _stream(domain))
    domain.Epot.mark_ell_valid()


def ComputeEpotRHS(domain: Domain, vectorField: torch.Tensor, useFaceTransform: bool = False,
                   out: Optional[torch.Tensor] = None) -> torch.Tensor:
    """div(u x B) per cell, ``[B * n]`` (potential.cu:325-370); ``vectorField`` is ``[B][d][n]``.
    ``out``: a contiguous ``[B * n]`` vector to write instead of a new one."""
    n, B, dims = domain.getTotalSize(), domain.getBatchSize(), domain.getSpatialDims()
    if vectorField.dim() != 1 or vectorField.numel() != n * dims * B:
        raise ValueError("vectorField must be flat [batch][dims][totalSize]")
    if vectorField.dtype != domain.getDtype():
        raise TypeError("vectorField dtype must match the domain dtype")
    vf = vectorField.contiguous()
    if out is None:
        out = torch.empty(n * B, dtype=domain.getDtype(), device=domain.getDevice())
    elif out.numel() != n * B or not out.is_contiguous() or out.dtype != domain.getDtype():
        raise ValueError("out must be a contiguous [batch * totalSize] vector of the domain dtype")
    d = _prep(domain)
    _lib.call("pp_compute_epot_rhs", C.byref(d), _lib.ptr(vf), _lib.ptr(out), _stream(domain))
    return out


def ComputeCurrentDensityFaceBased(domain: Domain, epotField: torch.Tensor, uCrossBField: torch.Tensor) -> torch.Tensor:
    """Face-based current density J, ``[B][3][n]`` (potential.cu:918-960). ``uCrossBField`` is
    ``[B][3][n]`` (3 components also in 2D)."""
    n, B, dims = domain.getTotalSize(), domain.getBatchSize(), domain.getSpatialDims()
    if epotField.dim() != 1 or epotField.numel() != n * B:
        raise ValueError("epotField size must equal domain totalSize (* batch size).")
    if uCrossBField.numel() == n * dims * B and dims == 3:
        uxb = uCrossBField.contiguous()
    elif uCrossBField.numel() == n * 3 * B:
        uxb = uCrossBField.contiguous()
    else:
        raise ValueError("uCrossBField size must be totalSize * 3 (* batch size)")
    if dims == 2 and uxb.numel() != n * 3 * B:
        raise ValueError("2D domains need the 3-component u x B field")
    phi = epotField.contiguous()
    out = torch.empty(n * 3 * B, dtype=domain.getDtype(), device=domain.getDevice())  # every entry is written
    d = _prep(domain)
    _lib.call("pp_compute_current_density", C.byref(d), _lib.ptr(phi), _lib.ptr(uxb), _lib.ptr(out), _stream(domain))
    return out


def _copy(domain: Domain, field: int, to_blocks: int) -> None:
    d = _prep(domain)
    _lib.call("pp_copy_result", C.byref(d), field, to_blocks, _stream(domain))


def CopyVelocityResultToBlocks(domain: Domain) -> None: _copy(domain, 0, 1)
def CopyVelocityResultFromBlocks(domain: Domain) -> None: _copy(domain, 0, 0)
def CopyPressureResultToBlocks(domain: Domain) -> None: _copy(domain, 1, 1)
def CopyPressureResultFromBlocks(domain: Domain) -> None: _copy(domain, 1, 0)
def CopyEpotResultToBlocks(domain: Domain) -> None: _copy(domain, 2, 1)
def CopyScalarResultToBlocks(domain: Domain) -> None: _copy(domain, 3, 1)
def CopyScalarResultFromBlocks(domain: Domain) -> None: _copy(domain, 3, 0)
def CopyEpotResultFromBlocks(domain: Domain) -> None: _copy(domain, 2, 0)


def CoordsToTransforms(coords: torch.Tensor) -> torch.Tensor:
    dims = coords.dim() - 2
    cells = [s - 1 for s in reversed(coords.shape[2:])] + [1] * (3 - dims)
    out = torch.empty(1, *[s - 1 for s in coords.shape[2:]], 2 * dims * dims + 1, dtype=coords.dtype, device=coords.device)
    c = coords.contiguous()
    _lib.call("pp_coords_to_transforms", dims, (C.c_int32 * 3)(*cells), _lib.ptr(c), _lib.ptr(out),
              _lib.dtype_code(c.dtype), _lib.stream_ptr(c.device))
    return out


def CoordsToFaceTransforms(coords: torch.Tensor) -> torch.Tensor:
    dims = coords.dim() - 2
    cells = [s - 1 for s in reversed(coords.shape[2:])] + [1] * (3 - dims)
    out = torch.empty(1, dims, *coords.shape[2:], 2 * dims * dims + 1, dtype=coords.dtype, device=coords.device)
    c = coords.contiguous()
    _lib.call("pp_coords_to_face_transforms", dims, (C.c_int32 * 3)(*cells), _lib.ptr(c), _lib.ptr(out),
              _lib.dtype_code(c.dtype), _lib.stream_ptr(c.device))
    return out


# ---------------------------------------------------------------- linear solvers
@dataclass
class LinearSolverResultInfo:
    """csrc/phipict/solvers/linear_solvers.h:180-185"""

    finalResidual: float
    usedIterations: int
    converged: bool
    isFiniteResidual: bool

    def __str__(self) -> str:
        return (f"LinearSolverResultInfo(finalResidual={self.finalResidual:.3e}, usedIterations={self.usedIterations}, "
                f"converged={self.converged}, isFiniteResidual={self.isFiniteResidual})")


@dataclass
class KrylovSettings:
    """The reference's krylov::GlobalSettings (krylov.h:68-74)."""

    enabled: bool = True
    checkInterval: int = 8
    useGraphs: bool = True
    cgPreconditioner: str = "none"
    bicgPreconditioner: str = "none"


_KRYLOV = KrylovSettings()
_PRECONDITIONERS = {"none": 0, "jacobi": 1}


def _precond_code(name) -> int:
    if isinstance(name, str) and name in _PRECONDITIONERS:
        return _PRECONDITIONERS[name]
    raise ValueError(f"Unknown preconditioner '{name}', expected 'none' or 'jacobi'.")


def SetKrylovSettings(enabled: bool = True, checkInterval: int = 8, useGraphs: bool = True, cgPreconditioner: str = "none",
                      bicgPreconditioner: str = "none") -> None:
    """Configure the fused solvers (bindings.cpp:739-752): ``checkInterval`` caps the iterations
    the host enqueues per chunk, ``useGraphs`` replays captured chunks of iterations from HIP
    graphs (segmented solves), ``cgPreconditioner`` / ``bicgPreconditioner`` are "none" or
    "jacobi". ``enabled=False`` (the reference's legacy cuBLAS solvers) is recorded; the library
    has one solver path."""
    cg, bicg = _precond_code(cgPreconditioner), _precond_code(bicgPreconditioner)
    if int(checkInterval) < 1:
        raise ValueError("checkInterval must be positive")
    _lib.call("pp_set_krylov_settings", int(checkInterval), int(bool(useGraphs)), cg, bicg)
    _KRYLOV.enabled = bool(enabled)


def GetKrylovSettings() -> dict:
    """The current settings as the reference returns them (bindings.cpp:753-763)."""
    ci, ug, cg, bicg = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32()
    _lib.call("pp_get_krylov_settings", C.byref(ci), C.byref(ug), C.byref(cg), C.byref(bicg))
    names = {v: k for k, v in _PRECONDITIONERS.items()}
    return {"enabled": _KRYLOV.enabled, "checkInterval": int(ci.value), "useGraphs": bool(ug.value),
            "cgPreconditioner": names[cg.value], "bicgPreconditioner": names[bicg.value]}


def _results(res, n) -> list:
    return [LinearSolverResultInfo(float(r.final_residual), int(r.used_iterations), bool(r.converged),
                                   bool(r.is_finite_residual)) for r in res[:n]]


def _host_scalar_list(t, dtype) -> list:
    if isinstance(t, torch.Tensor):
        if t.device.type != "cpu":
            raise ValueError("tolerance and maxit must be host tensors")
        return [float(v) for v in t.reshape(-1).tolist()]
    if isinstance(t, (list, tuple)):
        return [float(v) for v in t]
    return [float(t)]


def SolveLinear(A: CSRmatrix, RHS: torch.Tensor, x: torch.Tensor, maxIterations=1000, tolerance=1e-8,
                convergenceCriterion=ConvergenceCriterion.NORM2_NORMALIZED, useBiCG: bool = False,
                matrixRankDeficient: bool = False, residualResetSteps: int = 0, transposeA: bool = False,
                printResidual: bool = False, returnBestResult: bool = False, BiCGwithPreconditioner: bool = True) -> list:
    """Solve A x = RHS in place in ``x`` (linear_solve.cu:131-164) with the fused CG or BiCGStab.

    ``RHS``/``x`` hold ``k * n`` entries; RHS k uses matrix ``k / (k_total / A.getBatchSize())``.
    ``matrixRankDeficient`` marks a singular operator with the constant nullspace: the solve is
    run on the mean-free right-hand side (the reference routes this flag to its legacy cuBLAS
    CG instead). ``BiCGwithPreconditioner`` is accepted for signature compatibility; the fused
    BiCGStab is unpreconditioned, as the reference's default settings select."""
    n = A.getRows()
    if RHS.dim() != 1 or x.dim() != 1:
        raise ValueError("RHS and x must be vectors.")
    if RHS.numel() % n != 0:
        raise ValueError("Size of RHS must be a multiple of the matrix size.")
    if x.numel() != RHS.numel():
        raise ValueError("Size of x must match RHS.")
    if RHS.dtype != A.getDtype() or x.dtype != A.getDtype():
        raise TypeError("Data type of RHS and x must match A.")
    if not x.is_contiguous():
        raise ValueError("x must be contiguous (the solve writes it in place)")
    if convergenceCriterion not in (ConvergenceCriterion.NORM2, ConvergenceCriterion.NORM2_NORMALIZED):
        raise NotImplementedError("only the NORM2 and NORM2_NORMALIZED criteria are implemented")
    nrhs = RHS.numel() // n
    maxit = int(_host_scalar_list(maxIterations, torch.int32)[0])
    tols = _host_scalar_list(tolerance, A.getDtype())
    if len(tols) not in (1, nrhs):
        raise ValueError("tol must be a scalar or one value per right-hand side.")
    rhs = RHS.contiguous()
    res = (_lib.SolveResult * nrhs)()
    tarr = (C.c_double * len(tols))(*tols)
    desc = A.desc()
    _lib.call("pp_solve_linear", C.byref(desc), _lib.ptr(rhs), _lib.ptr(x), nrhs, _lib.dtype_code(A.getDtype()), maxit,
              tarr, len(tols), int(convergenceCriterion), int(bool(useBiCG)), int(bool(matrixRankDeficient)),
              int(residualResetSteps), int(bool(transposeA)), int(bool(returnBestResult)),
              _lib.stream_ptr(x.device), res)
    return _results(res, nrhs)


def SolveSegmented(A: CSRmatrix, segments: Sequence[int], RHS: torch.Tensor, x: torch.Tensor, maxIterations=1000,
                   tolerance=1e-8, useBiCG: bool = False, matrixRankDeficient: bool = False, tileRows: int = 0) -> list:
    """Solve the independent instances of a block-diagonal system in one launch per phase
    (``pp_solve_segmented``): instance s owns rows ``segments[s] .. segments[s+1]``; ``RHS`` and
    ``x`` hold ``k`` right-hand sides of ``A.getRows()`` entries each (component-major), every
    instance solved with the reference's CG or BiCGStab recurrences and stopping rules on its
    own rows (NORM2_NORMALIZED over them). ``tolerance``: one value or one per (component,
    instance), component-major; ``matrixRankDeficient``: each instance's right-hand side made
    mean-free first. Returns one LinearSolverResultInfo per (component, instance)."""
    n = A.getRows()
    if A.getBatchSize() != 1:
        raise ValueError("a segmented solve takes one block-diagonal matrix")
    if RHS.numel() % n or x.numel() != RHS.numel() or not x.is_contiguous():
        raise ValueError("RHS and x must hold whole right-hand sides and x must be contiguous")
    if RHS.dtype != A.getDtype() or x.dtype != A.getDtype():
        raise TypeError("Data type of RHS and x must match A.")
    ncomp = RHS.numel() // n
    segs = [int(v) for v in segments]
    S = len(segs) - 1
    tols = _host_scalar_list(tolerance, A.getDtype())
    if len(tols) not in (1, ncomp * S):
        raise ValueError("tolerance must be a scalar or one value per instance and component")
    rhs = RHS.contiguous()
    res = (_lib.SolveResult * (ncomp * S))()
    seg_arr = (C.c_int32 * len(segs))(*segs)
    tarr = (C.c_double * len(tols))(*tols)
    desc = A.desc()
    _lib.call("pp_solve_segmented", C.byref(desc), seg_arr, S, ncomp, _lib.ptr(rhs), _lib.ptr(x),
              _lib.dtype_code(A.getDtype()), int(_host_scalar_list(maxIterations, torch.int32)[0]), tarr, len(tols),
              int(bool(useBiCG)), int(bool(matrixRankDeficient)), int(tileRows), _lib.stream_ptr(x.device), res)
    return _results(res, ncomp * S)


def pack_amg_levels(levels: Sequence[Sequence[torch.Tensor]], hd, extras: Optional[Sequence[dict]] = None) -> tuple:
    """ctypes level descriptors for ``levels`` (per level: A row, col, val, dinv, R row, col,
    val, P row, col, val, all values in ``hd``) and optional per-level coalesced copies
    (``extras[i]``: "ell" = (columns, values, width) of the finest A, "a_sell" / "r_sell" /
    "p_sell" = (slice offsets, slice steps, columns, values, lanes, entries per operator) of
    A / R / P in the wave-interleaved sliced layout); returns (array, tensors kept alive)."""
    lv = (_lib.AmgLevel * len(levels))()
    keep = []
    for i, t in enumerate(levels):
        if len(t) != 10:
            raise ValueError("each AMG level needs (A row, col, val, dinv, R row, col, val, P row, col, val)")
        ex = extras[i] if extras is not None else {}
        if t[2] is None and not ((i == 0 and "ell" in ex) or "a_sell" in ex):
            raise ValueError("a level may omit its CSR values only with a sliced copy (or, the finest, an ELL copy)")
        if (t[2] is not None and t[2].dtype != hd) or t[3].dtype != hd or t[6].dtype != hd or t[9].dtype != hd:
            raise TypeError("all hierarchy values must have the dtype of the coarse pseudo-inverse")
        L = lv[i]
        L.n = t[0].numel() - 1
        L.nnz_a = t[1].numel()
        L.n_coarse = t[4].numel() - 1
        L.nnz_r = t[5].numel()
        L.nnz_p = t[8].numel()
        L.a_batched = 1 if t[3].numel() != L.n else 0
        for name, tt in zip(("a_row", "a_col", "a_val", "dinv", "r_row", "r_col", "r_val", "p_row", "p_col", "p_val"), t):
            if tt is None:
                continue
            tt = tt.contiguous()
            keep.append(tt)
            setattr(L, name, _lib.ptr(tt))
        if "ell" in ex:
            col, val, width = ex["ell"]
            if val.dtype != hd or col.numel() != width * L.n or val.numel() % (width * L.n):
                raise ValueError("ELL copy does not match the level operator")
            keep.extend((col, val))
            L.a_ell_width, L.a_ell_col, L.a_ell_val = int(width), _lib.ptr(col), _lib.ptr(val)
        for key, field_ in (("a_sell", "a_sell"), ("r_sell", "r_sell"), ("p_sell", "p_sell")):
            if key in ex:
                off, steps, col, val, lanes, total = ex[key]
                if val.dtype != hd or col.numel() != total or val.numel() % total:
                    raise ValueError("sliced ELL copy does not match the level")
                keep.extend((off, steps, col, val))
                setattr(L, field_, _lib.SellDesc(_lib.ptr(off), _lib.ptr(steps), _lib.ptr(col), _lib.ptr(val), int(lanes), 0,
                                                 int(total)))
    return lv, keep


def amg_pcg_packed(aRow, aCol, aVal, packed: tuple, coarsePinv: torch.Tensor, rhs: torch.Tensor, x: torch.Tensor,
                   maxIterations: int, tolerance: float, normalized: bool = True, returnBest: bool = False,
                   projectConstant: bool = False, preSweeps: int = 1, postSweeps: int = 1, omega: float = 2.0 / 3.0,
                   tolerances: Sequence[float] = (), ell: Optional[tuple] = None) -> list:
    """AMGPCGSolve on levels already packed by ``pack_amg_levels``; ``ell`` = (columns, values,
    width) of an ELL copy of the fine matrix, if one is valid."""
    lv, _keep = packed
    n = aRow.numel() - 1
    if rhs.numel() % n != 0 or rhs.numel() != x.numel():
        raise ValueError("rhs and x must be contiguous and of equal size, a multiple of the matrix size")
    if not x.is_contiguous():
        raise ValueError("x must be contiguous")
    nrhs = rhs.numel() // n
    hd = coarsePinv.dtype
    nc = coarsePinv.size(-1)
    batched_coarse = 1 if (coarsePinv.dim() == 3 and nrhs > 1) else 0
    pinv = coarsePinv.contiguous()
    desc = _lib.CsrDesc()
    desc.row, desc.col, desc.val = _lib.ptr(aRow), _lib.ptr(aCol), _lib.ptr(aVal)
    desc.n, desc.nnz = n, aCol.numel()
    desc.num_matrices = max(1, aVal.numel() // max(1, aCol.numel()))
    if ell is not None:
        desc.ell_col, desc.ell_val, desc.ell_width = _lib.ptr(ell[0]), _lib.ptr(ell[1]), int(ell[2])
    res = (_lib.SolveResult * nrhs)()
    tl = list(tolerances)
    tarr = (C.c_double * max(1, len(tl)))(*(tl if tl else [float(tolerance)]))
    rhs_c = rhs.contiguous()
    _lib.call("pp_amg_pcg_solve", C.byref(desc), lv, len(lv), _lib.ptr(pinv), nc, batched_coarse, _lib.dtype_code(hd),
              _lib.ptr(rhs_c), _lib.ptr(x), nrhs, _lib.dtype_code(aVal.dtype), int(maxIterations), float(tolerance),
              int(bool(normalized)), int(bool(returnBest)), int(bool(projectConstant)), int(preSweeps), int(postSweeps),
              float(omega), tarr, len(tl), _lib.stream_ptr(x.device), res)
    return _results(res, nrhs)


def AMGPCGSolve(aRow, aCol, aVal, levels: Sequence[Sequence[torch.Tensor]], coarsePinv: torch.Tensor, rhs: torch.Tensor,
                x: torch.Tensor, maxIterations: int, tolerance: float, normalized: bool = True,
                returnBestResult: bool = False, projectConstant: bool = False, preSweeps: int = 1, postSweeps: int = 1,
                omega: float = 2.0 / 3.0, tolerances: Sequence[float] = ()) -> list:
    """AMG-preconditioned CG (bindings.cpp:764-776 -> krylov.cu:1565-1612). ``levels`` holds
    per level ``(A row, col, val, dinv, R row, col, val, P row, col, val)``; the hierarchy
    dtype is that of ``coarsePinv``; ``x`` is updated in place."""
    packed = pack_amg_levels(levels, coarsePinv.dtype)
    return amg_pcg_packed(aRow, aCol, aVal, packed, coarsePinv, rhs, x, maxIterations, tolerance, normalized,
                          returnBestResult, projectConstant, preSweeps, postSweeps, omega, tolerances)


def ReleaseKrylovWorkspaces() -> None:
    _lib.call("pp_release_workspaces")


def SparseOuterProduct(a: torch.Tensor, b: torch.Tensor, out_pattern: CSRmatrix) -> None:
    """out.value[k] = a[row(k)] * b[col(k)] on the pattern of ``out_pattern`` (linear_solve.cu:166-190)."""
    n = out_pattern.getRows()
    if a.numel() != n or b.numel() != n:
        raise ValueError("Size of a and b must match matrix size.")
    desc = out_pattern.desc()
    _lib.call("pp_sparse_outer_product", _lib.ptr(a.contiguous()), _lib.ptr(b.contiguous()), C.byref(desc),
              _lib.dtype_code(out_pattern.getDtype()), _lib.stream_ptr(a.device))


def spmv(A: CSRmatrix, x: torch.Tensor) -> torch.Tensor:
    """y = A x for every right-hand side held in ``x``."""
    n = A.getRows()
    nrhs = x.numel() // n
    y = torch.empty_like(x)
    desc = A.desc()
    _lib.call("pp_spmv", C.byref(desc), _lib.ptr(x.contiguous()), _lib.ptr(y), nrhs, _lib.dtype_code(A.getDtype()),
              _lib.stream_ptr(x.device))
    return y

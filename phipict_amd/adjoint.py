"""Differentiable PISO (+ inductionless MHD) step: every operator of the orthogonal step is a
``torch.autograd.Function`` whose backward runs the library's hand-written adjoint kernels
(``pp_*_grad``), and every linear solve is differentiated by a transposed solve
(src/phipict/core/piso_diff.py:1154-1419 LinearSolveFunction, :1427-3685 the per-operator
Functions).

Tracked inputs: the blocks' velocities, the fixed faces' velocities, the blocks' velocity
sources, the domain's viscosity and the blocks' own viscosities (one value or one per cell), and
for MHD the field direction ``e_b``, the Stuart number and the Dirichlet potential values of the
fixed faces (shared by the environments or one slice each). Grid metrics are constants. Gradients of the linear solves:
for ``A x = b`` with an upstream gradient ``g``, ``A^T lam = g``, ``db = lam`` and
``dA = -lam (x)^T`` on the sparsity pattern; singular (pure-Neumann) systems work on the
mean-free subspace (``_project_out_constant``, piso_diff.py:1000-1026).
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Sequence

import torch

from . import _lib, ops
from .domain import CSRmatrix, Domain, FixedBoundary
from .solvers.tolerance import default_tolerance


def fixed_faces(domain: Domain) -> list:
    """(block index, face, boundary) of every fixed face, in block and face order."""
    out = []
    for bi, b in enumerate(domain.getBlocks()):
        for f, bound in b.getFixedBoundaries():
            out.append((bi, f, bound))
    return out


def blocks_to_flat(domain: Domain, fields: Sequence[torch.Tensor], channels: int) -> torch.Tensor:
    """Block fields ``[B, C, ...]`` to the flat ``[B][C][n]`` layout (differentiable)."""
    B = domain.getBatchSize()
    return torch.cat([f.reshape(B, channels, -1) for f in fields], dim=2).reshape(-1)


def flat_to_blocks(domain: Domain, flat: torch.Tensor, channels: int) -> list:
    """The flat ``[B][C][n]`` layout to block fields ``[B, C, (z), y, x]`` (differentiable)."""
    B = domain.getBatchSize()
    parts = torch.split(flat.reshape(B, channels, -1), [b.getCells() for b in domain.getBlocks()], dim=2)
    return [p.reshape(B, channels, *b.getSizesZYX()).contiguous() for p, b in zip(parts, domain.getBlocks())]


def _project(v: torch.Tensor, batch: int) -> torch.Tensor:
    """Remove the mean of every environment's slice (_project_out_constant)."""
    r = v.reshape(batch, -1)
    return (r - r.mean(dim=1, keepdim=True)).reshape(-1)


class _Bound:
    """Bind the forward's descriptor-read inputs (time step, face velocities, block
    pressures) for an adjoint launch, and restore the domain's own afterwards."""

    def __init__(self, domain: Domain, dt, bvels=None, pressures=None):
        self.domain, self.dt, self.bvels, self.pressures = domain, dt, bvels, pressures

    def __enter__(self):
        dom = self.domain
        self.saved_dt = dom.getTimeStep()
        faces = fixed_faces(dom)
        self.saved_bv = [bd.velocity for _, _, bd in faces]
        self.saved_p = [b.pressure for b in dom.getBlocks()]
        dom.setTimeStep(self.dt)
        if self.bvels is not None:
            for (_, _, bd), v in zip(faces, self.bvels):
                if v.data_ptr() != bd.velocity.data_ptr():
                    bd.velocity = v
                    dom._changed = True
        if self.pressures is not None:
            for b, p in zip(dom.getBlocks(), self.pressures):
                if p.data_ptr() != b.pressure.data_ptr():
                    b.pressure = p
                    dom._changed = True
        return dom.desc()

    def __exit__(self, *exc):
        dom = self.domain
        for (_, _, bd), v in zip(fixed_faces(dom), self.saved_bv):
            if v.data_ptr() != bd.velocity.data_ptr():
                bd.velocity = v
                dom._changed = True
        for b, p in zip(dom.getBlocks(), self.saved_p):
            if p.data_ptr() != b.pressure.data_ptr():
                b.pressure = p
                dom._changed = True
        if self.saved_dt is not None:
            dom.setTimeStep(self.saved_dt)
        return False


def _bound_grad_table(domain: Domain, want: bool):
    """Zeroed [B][d][face cells] gradient buffers of the fixed faces and the device table of
    their pointers (entry block*6 + face) the adjoint kernels take."""
    B, d = domain.getBatchSize(), domain.getSpatialDims()
    dev, dt = domain.getDevice(), domain.getDtype()
    faces = fixed_faces(domain)
    bufs = [torch.zeros(B, d, bd.face_cells(), dtype=dt, device=dev) if want else None for _, _, bd in faces]
    table = torch.zeros(domain.getNumBlocks() * 6, dtype=torch.int64)
    for (bi, f, _), buf in zip(faces, bufs):
        if buf is not None:
            table[bi * 6 + f] = buf.data_ptr()
    return bufs, table.to(dev)


def _reduce_bound_grad(v: torch.Tensor, g: torch.Tensor) -> torch.Tensor:
    """Gradient of a face velocity input (static ``[vb, d]`` or ``[vb, d, face...]``) from the
    per-face-cell, per-environment gradient ``[B, d, face cells]``."""
    if v.dim() == 2:
        g = g.sum(dim=2)
    else:
        g = g.reshape(g.size(0), g.size(1), *v.shape[2:])
    if v.size(0) == 1 and g.size(0) > 1:
        g = g.sum(dim=0, keepdim=True)
    return g.to(v.dtype)


def _reduce_source_grads(domain: Domain, sources: Sequence[Optional[torch.Tensor]], gflat: torch.Tensor) -> list:
    """Gradients of the blocks' velocity sources (static ``[sb, d]`` or per cell) from the
    per-cell gradient ``[B][d][n]``."""
    d = domain.getSpatialDims()
    parts = flat_to_blocks(domain, gflat, d)
    out = []
    for s, g in zip(sources, parts):
        if s is None:
            out.append(None)
            continue
        if s.dim() == 2:
            g = g.reshape(g.size(0), d, -1).sum(dim=2)
        if s.size(0) == 1 and g.size(0) > 1:
            g = g.sum(dim=0, keepdim=True)
        out.append(g)
    return out


def viscosity_inputs(domain: Domain) -> list:
    """The viscosities the momentum assembly reads, as tracked inputs: the domain's, then each
    block's own (None for a block that uses the domain's)."""
    return [domain.viscosity] + [b.viscosity for b in domain.getBlocks()]


def _visc_grad_buffer(domain: Domain, viscs: Sequence[Optional[torch.Tensor]], needs: Sequence[bool]):
    """A zeroed [B][n] buffer for the gradient with respect to the viscosity as a per-cell field,
    or None when no viscosity input needs a gradient."""
    if not viscs or not any(needs):
        return None
    return _zeros_flat(domain, 1)


def _reduce_visc_grads(domain: Domain, viscs: Sequence[Optional[torch.Tensor]], g: Optional[torch.Tensor]) -> list:
    """The per-cell viscosity gradient [B][n] to the inputs of ``viscosity_inputs``: the domain's
    value collects the cells of the blocks without their own viscosity, a block's static value
    its cells, a per-cell viscosity its cells one by one; the viscosities are shared by the
    environments, so their gradients sum over them."""
    if not viscs:
        return []
    if g is None:
        return [None] * len(viscs)
    blocks = domain.getBlocks()
    per_cell = g.reshape(domain.getBatchSize(), -1).sum(dim=0)
    parts = torch.split(per_cell, [b.getCells() for b in blocks])
    out = []
    dom_v = viscs[0]
    if dom_v is not None:
        total = sum((p.sum() for b, p in zip(blocks, parts) if b.viscosity is None), per_cell.new_zeros(()))
        out.append(total.reshape(dom_v.shape).to(device=dom_v.device, dtype=dom_v.dtype))
    else:
        out.append(None)
    for v, p in zip(viscs[1:], parts):
        if v is None:
            out.append(None)
        else:
            out.append((p.sum() if v.numel() == 1 else p).reshape(v.shape).to(v.dtype))
    return out


def _zeros_flat(domain: Domain, channels: int) -> torch.Tensor:
    return torch.zeros(domain.getBatchSize() * channels * domain.getTotalSize(), dtype=domain.getDtype(),
                       device=domain.getDevice())


def _stream(domain: Domain) -> int:
    return _lib.stream_ptr(domain.getDevice())


# ----------------------------------------------------------------------------- assembly
class AdvectionMatrixFunction(torch.autograd.Function):
    """(A, C values) of the implicit advection-diffusion matrix from the block velocities and,
    when given after them, the ``viscosity_inputs`` (piso_diff.py:1427-1596)."""

    @staticmethod
    def forward(ctx, domain: Domain, dt: torch.Tensor, *inputs):
        ops.SetupAdvectionMatrix(domain, dt, 0)
        nv = domain.getNumBlocks()
        ctx.domain, ctx.dt = domain, dt.clone()
        ctx.n_vels, ctx.viscs = nv, [None if v is None else v.detach() for v in inputs[nv:]]
        return domain.A.clone(), domain.C.value.clone()

    @staticmethod
    def backward(ctx, gA, gC):
        dom = ctx.domain
        nv = ctx.n_vels
        if not any(ctx.needs_input_grad[2:]):
            return (None, None) + (None,) * (nv + len(ctx.viscs))
        gU = _zeros_flat(dom, dom.getSpatialDims())
        gNu = _visc_grad_buffer(dom, ctx.viscs, ctx.needs_input_grad[2 + nv:])
        if gC is None:
            gC = torch.zeros_like(dom.C.value)
        with _Bound(dom, ctx.dt) as d:
            _lib.call("pp_setup_advection_matrix_grad", C.byref(d), _lib.ptr(gC.contiguous()),
                      _lib.ptr(gA.contiguous()) if gA is not None else None, _lib.ptr(gU),
                      _lib.ptr(gNu) if gNu is not None else None, _stream(dom))
        return (None, None, *flat_to_blocks(dom, gU, dom.getSpatialDims()), *_reduce_visc_grads(dom, ctx.viscs, gNu))


class AdvectionRHSFunction(torch.autograd.Function):
    """Momentum right-hand side from the block velocities, face velocities and sources
    (piso_diff.py:1772-1943, without the pressure-gradient term)."""

    @staticmethod
    def forward(ctx, domain: Domain, dt: torch.Tensor, nv: int, nb: int, *inputs):
        ops.SetupAdvectionVelocity(domain, dt, 0, False)
        ns = domain.getNumBlocks()
        ctx.domain, ctx.dt, ctx.nv, ctx.nb = domain, dt.clone(), nv, nb
        bvels = inputs[nv:nv + nb]
        srcs = inputs[nv + nb:nv + nb + ns]
        ctx.bvels = [v.detach() for v in bvels]
        ctx.src_meta = [None if s is None else s.detach() for s in srcs]
        ctx.viscs = [None if v is None else v.detach() for v in inputs[nv + nb + ns:]]
        return domain.velocityRHS.clone()

    @staticmethod
    def backward(ctx, gR):
        dom = ctx.domain
        d = dom.getSpatialDims()
        gU = _zeros_flat(dom, d)
        want_s = any(s is not None for s in ctx.src_meta)
        gS = _zeros_flat(dom, d) if want_s else None
        gNu = _visc_grad_buffer(dom, ctx.viscs, ctx.needs_input_grad[4 + ctx.nv + ctx.nb + len(ctx.src_meta):])
        bufs, table = _bound_grad_table(dom, ctx.nb > 0)
        with _Bound(dom, ctx.dt, bvels=ctx.bvels) as desc:
            _lib.call("pp_setup_advection_velocity_grad", C.byref(desc), _lib.ptr(gR.contiguous()), _lib.ptr(gU),
                      _lib.ptr(gS) if gS is not None else None, _lib.ptr(table) if ctx.nb else None,
                      _lib.ptr(gNu) if gNu is not None else None, _stream(dom))
        gv = flat_to_blocks(dom, gU, d)
        gb = [_reduce_bound_grad(v, g) for v, g in zip(ctx.bvels, bufs)] if ctx.nb else []
        gs = _reduce_source_grads(dom, ctx.src_meta, gS) if want_s else [None] * len(ctx.src_meta)
        return (None, None, None, None, *gv, *gb, *gs, *_reduce_visc_grads(dom, ctx.viscs, gNu))


class PressureMatrixFunction(torch.autograd.Function):
    """Pressure Laplacian values from the advection diagonal (piso_diff.py:2973-3113)."""

    @staticmethod
    def forward(ctx, domain: Domain, dt: torch.Tensor, A: torch.Tensor):
        ops.SetupPressureMatrix(domain, dt, 0)
        ctx.domain, ctx.dt = domain, dt.clone()
        ctx.save_for_backward(A)
        return domain.P.value.clone()

    @staticmethod
    def backward(ctx, gP):
        dom = ctx.domain
        (A,) = ctx.saved_tensors
        gA = torch.zeros_like(A)
        with _Bound(dom, ctx.dt) as d:
            _lib.call("pp_setup_pressure_matrix_grad", C.byref(d), _lib.ptr(A.contiguous()), _lib.ptr(gP.contiguous()),
                      _lib.ptr(gA), _stream(dom))
        return None, None, gA


class PressureRHSFunction(torch.autograd.Function):
    """HbyA and its face divergence from A, C, the predicted velocity, the block velocities,
    face velocities and sources (piso_diff.py:3115-3524)."""

    @staticmethod
    def forward(ctx, domain: Domain, dt: torch.Tensor, tsn: bool, A, Cv, V, nv: int, nb: int, *inputs):
        ops.SetupPressureRHS(domain, dt, 0, False, timeStepNorm=tsn)
        H, div = domain.pressureRHS.clone(), domain.pressureRHSdiv.clone()
        ns = domain.getNumBlocks()
        ctx.domain, ctx.dt, ctx.tsn, ctx.nv, ctx.nb = domain, dt.clone(), tsn, nv, nb
        ctx.bvels = [v.detach() for v in inputs[nv:nv + nb]]
        ctx.src_meta = [None if s is None else s.detach() for s in inputs[nv + nb:nv + nb + ns]]
        ctx.viscs = [None if v is None else v.detach() for v in inputs[nv + nb + ns:]]
        ctx.save_for_backward(A, Cv, V, H)
        return H, div

    @staticmethod
    def backward(ctx, gH, gDiv):
        dom = ctx.domain
        d = dom.getSpatialDims()
        A, Cv, V, H = ctx.saved_tensors
        gHt = gH.clone().contiguous() if gH is not None else _zeros_flat(dom, d)
        bufs, table = _bound_grad_table(dom, ctx.nb > 0)
        gA, gC, gV, gU = torch.zeros_like(A), torch.zeros_like(Cv), torch.zeros_like(V), _zeros_flat(dom, d)
        want_s = any(s is not None for s in ctx.src_meta)
        gS = _zeros_flat(dom, d) if want_s else None
        # inputs after (domain, dt, tsn, A, Cv, V, nv, nb): velocities, face velocities, sources, viscosities
        gNu = _visc_grad_buffer(dom, ctx.viscs, ctx.needs_input_grad[8 + ctx.nv + ctx.nb + len(ctx.src_meta):])
        with _Bound(dom, ctx.dt, bvels=ctx.bvels) as desc:
            tp = _lib.ptr(table) if ctx.nb else None
            if gDiv is not None:
                _lib.call("pp_setup_pressure_rhs_div_grad", C.byref(desc), _lib.ptr(gDiv.contiguous()), int(ctx.tsn),
                          _lib.ptr(gHt), tp, _stream(dom))
            _lib.call("pp_setup_pressure_rhs_grad", C.byref(desc), _lib.ptr(gHt), _lib.ptr(A), _lib.ptr(Cv),
                      _lib.ptr(V), _lib.ptr(H), _lib.ptr(gA), _lib.ptr(gC), _lib.ptr(gV), _lib.ptr(gU),
                      _lib.ptr(gS) if gS is not None else None, tp, _lib.ptr(gNu) if gNu is not None else None,
                      _stream(dom))
        gv = flat_to_blocks(dom, gU, d)
        gb = [_reduce_bound_grad(v, g) for v, g in zip(ctx.bvels, bufs)] if ctx.nb else []
        gs = _reduce_source_grads(dom, ctx.src_meta, gS) if want_s else [None] * len(ctx.src_meta)
        return (None, None, None, gA, gC, gV, None, None, *gv, *gb, *gs, *_reduce_visc_grads(dom, ctx.viscs, gNu))


class CorrectVelocityFunction(torch.autograd.Function):
    """u = HbyA - grad(p) / A with the finite-difference gradient (piso_diff.py:3526-3686)."""

    @staticmethod
    def forward(ctx, domain: Domain, dt: torch.Tensor, tsn: bool, H, A, p):
        # private copies: later forward operators write the domain's vectors in place
        domain.setPressureRHS(H.detach().clone())
        domain.setA(A.detach().clone())
        domain.setPressureResult(p.detach().clone())
        for b in domain.getBlocks():
            b.pressure = torch.empty_like(b.pressure)
        domain._changed = True
        domain.UpdateDomainData()
        ops.CopyPressureResultToBlocks(domain)
        ops.CorrectVelocity(domain, dt, version=1, timeStepNorm=tsn)
        ctx.domain, ctx.dt, ctx.tsn = domain, dt.clone(), tsn
        ctx.pressures = [b.pressure.clone() for b in domain.getBlocks()]
        ctx.save_for_backward(A)
        return domain.velocityResult.clone()

    @staticmethod
    def backward(ctx, gOut):
        dom = ctx.domain
        (A,) = ctx.saved_tensors
        d = dom.getSpatialDims()
        gH, gA = _zeros_flat(dom, d), torch.zeros_like(A)
        gP = _zeros_flat(dom, 1)
        with _Bound(dom, ctx.dt, pressures=ctx.pressures) as desc:
            _lib.call("pp_correct_velocity_grad", C.byref(desc), _lib.ptr(gOut.contiguous()), _lib.ptr(A),
                      int(ctx.tsn), _lib.ptr(gH), _lib.ptr(gA), _lib.ptr(gP), _stream(dom))
        return None, None, None, gH, gA, gP


# ----------------------------------------------------------------------------- linear solves
class LinearSolveFunction(torch.autograd.Function):
    """x = A^{-1} b for the matrix values ``value`` on ``pattern`` (B matrices, ``rpm``
    right-hand sides each), differentiable in b and the values (piso_diff.py:1154-1419). With
    ``singular`` the right-hand side and the result are taken mean-free per right-hand side
    and the adjoint system is solved on the mean-free subspace."""

    @staticmethod
    def forward(ctx, solve, pattern: CSRmatrix, value: torch.Tensor, b: torch.Tensor, rpm: int, singular: bool,
                symmetric: bool, bwd_rtol: float):
        x = solve(b.detach().contiguous())
        ctx.solve, ctx.pattern, ctx.rpm = solve, pattern, rpm
        ctx.singular, ctx.symmetric, ctx.bwd_rtol = singular, symmetric, bwd_rtol
        ctx.save_for_backward(value, x)
        return x

    @staticmethod
    def backward(ctx, gx):
        value, x = ctx.saved_tensors
        pat = ctx.pattern
        n, nnz = pat.getRows(), pat.getNnz()
        nrhs = x.numel() // n
        g = gx.contiguous()
        if ctx.singular:
            g = _project(g, nrhs)
        A = CSRmatrix(value.detach().contiguous(), pat.index, pat.row)
        lam = torch.zeros_like(g)
        if bool(g.ne(0).any()):
            scale = torch.linalg.vector_norm(g.reshape(nrhs, n), dim=1) / max(1.0, n ** 0.5)
            floor = default_tolerance(g.dtype) * 1e-6
            tols = [max(float(s) * ctx.bwd_rtol, floor) for s in scale.cpu()]
            infos = ops.SolveLinear(A, g, lam, maxIterations=20000, tolerance=tols,
                                    convergenceCriterion=ops.ConvergenceCriterion.NORM2_NORMALIZED,
                                    useBiCG=not ctx.symmetric, matrixRankDeficient=ctx.singular,
                                    transposeA=not ctx.symmetric, returnBestResult=True)
            if not all(i.isFiniteResidual for i in infos):
                raise RuntimeError("adjoint solve: non-finite residual")
        if ctx.singular:
            lam = _project(lam, nrhs)
        gvalue = None
        if ctx.needs_input_grad[2]:
            B = value.numel() // nnz
            gvalue = torch.zeros(B * nnz, dtype=value.dtype, device=value.device)
            out = CSRmatrix(torch.zeros(nnz, dtype=value.dtype, device=value.device), pat.index, pat.row)
            lr, xr = lam.reshape(nrhs, n), x.reshape(nrhs, n)
            for r in range(nrhs):
                ops.SparseOuterProduct(lr[r], xr[r], out)
                m = r // ctx.rpm if B > 1 else 0
                gvalue[m * nnz:(m + 1) * nnz] -= out.value
        return None, None, gvalue, lam, None, None, None, None


# ----------------------------------------------------------------------------- MHD
def _cross(a: torch.Tensor, b: torch.Tensor, dim: int) -> torch.Tensor:
    return torch.linalg.cross(a, b, dim=dim)


class CrossEbFunction(torch.autograd.Function):
    """u x e_b of the flat velocity, ``[B][3][n]`` (mhd.py:309-355)."""

    @staticmethod
    def forward(ctx, v: torch.Tensor, e_b: torch.Tensor, n: int, dims: int):
        from .mhd import compute_u_cross_eb_flat

        ctx.n, ctx.dims = n, dims
        ctx.save_for_backward(v, e_b)
        return compute_u_cross_eb_flat(v.detach().contiguous(), e_b.detach(), n, dims)

    @staticmethod
    def backward(ctx, g):
        v, e_b = ctx.saved_tensors
        n, dims = ctx.n, ctx.dims
        B = v.numel() // (dims * n)
        g3 = g.reshape(B, 3, n)
        v3 = torch.zeros(B, 3, n, dtype=v.dtype, device=v.device)
        v3[:, :dims] = v.reshape(B, dims, n)
        e = e_b.to(device=v.device, dtype=v.dtype)
        eb = e.reshape(-1, 3, 1) if e.dim() == 2 else e.reshape(1, 3, 1)
        # (v x e) . g = v . (e x g) = e . (g x v)
        gv = _cross(eb.expand(B, 3, n), g3, 1)[:, :dims].reshape(-1)
        ge = None
        if ctx.needs_input_grad[1]:
            ge = _cross(g3, v3, 1).sum(dim=2)
            ge = ge if e_b.dim() == 2 else ge.sum(dim=0)
            ge = ge.to(device=e_b.device, dtype=e_b.dtype)
        return gv, ge, None, None


def potential_value_faces(domain: Domain) -> list:
    """(block index, face, boundary) of the fixed faces whose Dirichlet potential values reach the
    potential kernels (_potential_value_bounds, piso_diff.py:2455-2472): the values enter the
    potential's right-hand side and the current density linearly, never the matrix."""
    return [(bi, f, bd) for bi, f, bd in fixed_faces(domain) if bd.hasPotentialValues() and bd.hasPotentialDirichlet()]


def potential_values(domain: Domain) -> list:
    """The value tensors of ``potential_value_faces``, in the same order: the inputs the
    potential operators track."""
    return [bd.potentialValues for _, _, bd in potential_value_faces(domain)]


def _value_grad_table(domain: Domain, faces: list, needs: Sequence[bool]):
    """Zeroed [B][face cells] gradient buffers of the faces whose values need a gradient and the
    device table of their pointers (entry block * 6 + face; None when no face needs one)."""
    if not any(needs):
        return [None] * len(faces), None
    B = domain.getBatchSize()
    bufs = [torch.zeros(B, bd.face_cells(), dtype=domain.getDtype(), device=domain.getDevice()) if need else None
            for (_, _, bd), need in zip(faces, needs)]
    table = torch.zeros(domain.getNumBlocks() * 6, dtype=torch.int64)
    for (bi, f, _), buf in zip(faces, bufs):
        if buf is not None:
            table[bi * 6 + f] = buf.data_ptr()
    return bufs, table.to(domain.getDevice())


def _reduce_value_grads(values: Sequence[torch.Tensor], bufs: Sequence[Optional[torch.Tensor]]) -> list:
    """Per-environment value gradients [B][face cells] to the shape of each value input: values
    shared by the environments get the sum over them (_potential_values_grad, piso_diff.py:2474-2530)."""
    out = []
    for v, g in zip(values, bufs):
        if g is None:
            out.append(None)
            continue
        if v.numel() == g.size(1):
            g = g.sum(dim=0)
        out.append(g.reshape(v.shape).to(v.dtype))
    return out


class EpotRHSFunction(torch.autograd.Function):
    """Face divergence of the in-plane u x e_b for the potential equation, with the Dirichlet
    values of the faces ``potential_value_faces`` (piso_diff.py:2541-2635)."""

    @staticmethod
    def forward(ctx, domain: Domain, field: torch.Tensor, *values):
        ctx.domain = domain
        ctx.faces = potential_value_faces(domain)
        if len(ctx.faces) != len(values):
            raise ValueError("EpotRHSFunction takes the values of every potential_value_faces face")
        ctx.values = [v.detach() for v in values]
        return ops.ComputeEpotRHS(domain, field.detach().contiguous())

    @staticmethod
    def backward(ctx, g):
        dom = ctx.domain
        gF = _zeros_flat(dom, dom.getSpatialDims())
        bufs, table = _value_grad_table(dom, ctx.faces, ctx.needs_input_grad[2:])
        d = dom.desc()
        _lib.call("pp_compute_epot_rhs_grad", C.byref(d), _lib.ptr(g.contiguous()), _lib.ptr(gF),
                  _lib.ptr(table) if table is not None else None, _stream(dom))
        return (None, gF, *_reduce_value_grads(ctx.values, bufs))


class CurrentDensityFunction(torch.autograd.Function):
    """Face-based current density from phi and u x e_b, with the Dirichlet values of the faces
    ``potential_value_faces`` (piso_diff.py:2637-2740)."""

    @staticmethod
    def forward(ctx, domain: Domain, phi: torch.Tensor, uxb: torch.Tensor, *values):
        ctx.domain = domain
        ctx.faces = potential_value_faces(domain)
        if len(ctx.faces) != len(values):
            raise ValueError("CurrentDensityFunction takes the values of every potential_value_faces face")
        ctx.values = [v.detach() for v in values]
        return ops.ComputeCurrentDensityFaceBased(domain, phi.detach().contiguous(), uxb.detach().contiguous())

    @staticmethod
    def backward(ctx, gJ):
        dom = ctx.domain
        gphi, gx = _zeros_flat(dom, 1), _zeros_flat(dom, 3)
        bufs, table = _value_grad_table(dom, ctx.faces, ctx.needs_input_grad[3:])
        d = dom.desc()
        _lib.call("pp_compute_current_density_grad", C.byref(d), _lib.ptr(gJ.contiguous()), _lib.ptr(gphi),
                  _lib.ptr(gx), _lib.ptr(table) if table is not None else None, _stream(dom))
        return (None, gphi, gx, *_reduce_value_grads(ctx.values, bufs))


def lorentz_sources(domain: Domain, J: torch.Tensor, e_b: torch.Tensor, N: torch.Tensor,
                    sources: Sequence[Optional[torch.Tensor]]) -> list:
    """Per block ``N (J x e_b)[:d] + S`` as differentiable tensor operations (mhd.py:543-594);
    the forward of the non-differentiable step runs the same formula in one library launch."""
    B, n, d = domain.getBatchSize(), domain.getTotalSize(), domain.getSpatialDims()
    J3 = J.reshape(B, 3, n)
    e = e_b.to(device=J.device, dtype=J.dtype)
    eb = e.reshape(-1, 3, 1) if e.dim() == 2 else e.reshape(1, 3, 1)
    Nn = N.to(device=J.device, dtype=J.dtype).reshape(-1, 1, 1)
    F = (Nn * _cross(J3, eb.expand(B, 3, n), 1))[:, :d].reshape(-1)
    out = []
    for blk, Fb, S in zip(domain.getBlocks(), flat_to_blocks(domain, F, d), sources):
        if S is not None:
            Fb = Fb + (S.reshape(S.size(0), d, *([1] * (Fb.dim() - 2))) if S.dim() == 2 else S)
        out.append(Fb)
    return out


# ----------------------------------------------------------------------------- checkpointing
def state_tensors(sim) -> dict:
    """The state one step hands to the next: block velocities, fixed-face velocities and, for
    MHD, the potential and u x e_b the next Lorentz force uses."""
    dom = sim.domain
    st = {"velocity": [b.velocity for b in dom.getBlocks()],
          "faces": [bd.velocity for _, _, bd in fixed_faces(dom)],
          "sources": [b.velocitySource for b in dom.getBlocks()]}
    if hasattr(sim, "_u_cross_eb_flat"):
        phi = sim.__dict__.get("_diff_phi") if sim.differentiable else None
        st["phi"] = [phi if phi is not None else dom.epotResult]
        st["uxb"] = [sim._u_cross_eb_flat]
    return st


def _set_state(sim, st: dict) -> None:
    dom = sim.domain
    for b, t in zip(dom.getBlocks(), st["velocity"]):
        b.velocity = t
    for (_, _, bd), t in zip(fixed_faces(dom), st["faces"]):
        bd.velocity = t
        dom._bc_generation += 1
    for b, t in zip(dom.getBlocks(), st["sources"]):
        b.velocitySource = t
    if "phi" in st:
        sim._diff_phi = st["phi"][0]
        sim._u_cross_eb_flat = st["uxb"][0]
        dom.setEpotResult(st["phi"][0].detach().clone())
    dom._changed = True
    dom.UpdateDomainData()


def _clone_state(st: dict, requires_grad: bool = False) -> dict:
    out = {}
    for k, ts in st.items():
        out[k] = [None if t is None else t.detach().clone().requires_grad_(requires_grad) for t in ts]
    return out


def checkpointed_gradient(sim, steps: int, every: int, loss_fn) -> tuple:
    """Backpropagate ``loss_fn(domain)`` after ``steps`` steps through the whole rollout while
    keeping only every ``every``-th state: the forward runs untracked and stores the state at
    the start of each segment; the backward re-runs one segment at a time with the
    differentiable step, from the last to the first, and hands each segment the gradient of its
    end state. Parameters used by every step (e_b, N, tensors the hooks set) collect their
    gradients across the segments; so do tracked tensors the initial state was computed from
    (the initial state's gradients are pushed into their graph at the end). Returns (loss,
    gradients of the initial state)."""
    if every < 1 or steps < 1:
        raise ValueError("steps and every must be positive")
    was_diff = sim.differentiable
    initial = state_tensors(sim)
    ckpts = []
    sim.differentiable = False
    try:
        with torch.no_grad():
            for k in range(0, steps, every):
                ckpts.append((_clone_state(state_tensors(sim)), sim.total_step, sim.total_time, min(every, steps - k)))
                for _ in range(ckpts[-1][3]):
                    if not sim.single_step():
                        raise RuntimeError("checkpointed rollout: a forward step failed")
        final = _clone_state(state_tensors(sim))
        final_counters = (sim.total_step, sim.total_time)
        sim.differentiable = True
        g_next, loss_value = None, None
        for ci in reversed(range(len(ckpts))):
            st, ts, tt, nsteps = ckpts[ci]
            leaves = _clone_state(st, requires_grad=True)
            _set_state(sim, leaves)
            sim.total_step, sim.total_time = ts, tt
            for _ in range(nsteps):
                if not sim.single_step():
                    raise RuntimeError("checkpointed rollout: a recomputed step failed")
            if ci == len(ckpts) - 1:
                obj = loss_fn(sim.domain)
                loss_value = float(obj.detach())
            else:
                end = state_tensors(sim)
                obj = 0.0
                for k, ts_end in end.items():
                    for t, g in zip(ts_end, g_next.get(k, [])):
                        if g is not None and t is not None and t.requires_grad:
                            obj = obj + (t * g).sum()
            if isinstance(obj, torch.Tensor) and obj.requires_grad:
                # parameters computed once outside the rollout (e.g. e_b = B0 * direction) are
                # part of every segment's graph
                obj.backward(retain_graph=True)
            g_next = {k: [None if t is None else t.grad for t in ts] for k, ts in leaves.items()}
        # continue into whatever the initial state was computed from (a tracked inflow, the
        # outflow balanced against it before the rollout)
        outs, grads = [], []
        for k, ts in initial.items():
            for t, g in zip(ts, g_next.get(k, [])):
                if t is not None and g is not None and t.requires_grad:
                    outs.append(t)
                    grads.append(g)
        if outs:
            torch.autograd.backward(outs, grads, retain_graph=True)
        _set_state(sim, final)
        sim.total_step, sim.total_time = final_counters
        return loss_value, g_next
    finally:
        sim.differentiable = was_diff

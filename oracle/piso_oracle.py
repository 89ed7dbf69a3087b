"""ORACLE — test infrastructure only.

The PISO (+ inductionless MHD) step of the reference, driven on the CPU with the oracle's C
kernels and solvers: the step order of ``_PISO_split_step`` (src/phipict/core/piso_simulation.py:
3525-4169, orthogonal and non-orthogonal branches), the magnetic step and Lorentz force of
``MHDSimulation`` (src/phipict/simulation/mhd.py:309-355, 543-594, 877-1002). One environment,
absolute tolerances, fixed or adaptive CFL substeps.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Callable, Optional

import numpy as np

from . import oracle as O


@dataclass
class MHDSpec:
    stuart: float
    e_b: tuple = (0.0, 1.0, 0.0)
    potential_tol: float = 1e-8
    potential_max_iter: int = 5000
    potential_return_best: bool = True
    use_amg: bool = False


def resolve_tol(tol, b: np.ndarray) -> float:
    """Absolute tolerance of a solve: a float is absolute; ``("rtol", r)`` is relative to the RHS
    scale ||b||/sqrt(n), floored at 8 machine epsilons of that scale (tolerance.py:180-215)."""
    if not isinstance(tol, tuple):
        return float(tol)
    scale = float(np.linalg.norm(b.astype(np.float64))) / np.sqrt(max(b.size, 1))
    t = max(tol[1] * scale, 8.0 * float(np.finfo(b.dtype).eps) * scale)
    if not t > 0:
        t = 1e-8 if b.dtype == np.float64 else 1e-5
    return t


def cross(v: np.ndarray, e: np.ndarray) -> np.ndarray:
    """Row-wise v x e for v [3, n], e [3]."""
    return np.stack([v[1] * e[2] - v[2] * e[1], v[2] * e[0] - v[0] * e[2], v[0] * e[1] - v[1] * e[0]])


# ---- boundary driving (piso_simulation.py:308-838), restated on the oracle's face arrays
def face_flux(od: O.OracleDomain, bi: int, f: int) -> np.ndarray:
    """Contravariant outward-axis flux det * Minv[axis] . u per face cell of a fixed face
    (FixedBoundary::GetFluxes, boundaries.cpp; zero for a non-Dirichlet face)."""
    d = od.dims
    axis = f >> 1
    fc = od.face_cells(bi, f)
    v = od.face_velocity(bi, f)
    vel = np.broadcast_to(v.reshape(d, 1), (d, fc)) if v.ndim == 1 else v
    if od.specs[bi].faces[f].neumann:
        return np.zeros(fc, od.dtype)
    t = od.face_transform(bi, f)
    minv = t[:, d * d:2 * d * d].reshape(fc, d, d)[:, axis, :]
    det = t[:, 2 * d * d]
    return det * (minv * vel.T).sum(axis=1)


def net_fixed_flux(od: O.OracleDomain, faces) -> float:
    """Net outward flux of (block, face) pairs (get_fixed_boundary_fluxes :344-382)."""
    total = 0.0
    for bi, f in faces:
        fl = float(np.sum(face_flux(od, bi, f), dtype=np.float64))
        total = total - fl if f % 2 == 0 else total + fl
    return total


def balance_boundary_fluxes(od: O.OracleDomain, free, tol: Optional[float] = None) -> None:
    """Rescale the free faces' velocities so the net flux vanishes (:533-610)."""
    fixed_faces = [(bi, f) for bi, s in enumerate(od.specs) for f, fs in enumerate(s.faces)
                   if fs.kind == "fixed" and (bi, f) not in free]
    fix = net_fixed_flux(od, fixed_faces)
    var = net_fixed_flux(od, free)
    btol = tol if tol is not None else (1e-8 if od.dtype == np.float64 else 1e-5)
    if abs(fix + var) <= btol * 0.01:
        return
    scale = -fix / var
    for bi, f in free:
        od.set_face_velocity(bi, f, od.face_velocity(bi, f) * od.dtype.type(scale))


def update_advective_boundaries(od: O.OracleDomain, bounds, velm: np.ndarray, dt: float,
                                tol: Optional[float] = None) -> None:
    """u_b <- u_b - t (u_b - u_cell), t = 1 - 1/(1 + 2 dt Minv[axis].velm) per face cell, then
    the flux balance (:611-838, get_advective_velocity :455-530)."""
    d = od.dims
    for bi, f in bounds:
        axis, upper = f >> 1, f & 1
        v = od.face_velocity(bi, f)
        if v.ndim == 1:
            raise ValueError("outflow faces must be varying")
        t = od.face_transform(bi, f)
        t_inv_row = t[:, d * d + axis * d: d * d + (axis + 1) * d]  # [fc, d]
        adv = (t_inv_row * np.asarray(velm, od.dtype)[None, :]).sum(axis=1)  # [fc]
        alpha = od.dtype.type(dt) * 2 * adv
        w = 1 - 1 / (1 + alpha)
        shape = list(reversed(od.specs[bi].size_xyz[:d]))  # (z), y, x
        cells = od.velocity[bi].reshape([d] + shape)
        idx = [slice(None)] * (d + 1)
        idx[1 + d - 1 - axis] = slice(-1, None) if upper else slice(0, 1)
        layer = cells[tuple(idx)].reshape(d, -1)
        od.set_face_velocity(bi, f, v - w[None, :] * (v - layer))
    balance_boundary_fluxes(od, list(bounds), tol)


def cfl_substep(remaining: float, max_vel: float, cfl: float) -> tuple:
    """(next substep size, substeps of the remaining time) at the maximum velocity
    (_cfl_substep, src/phipict/core/piso_simulation.py:186-215)."""
    if np.isclose(max_vel, 0):
        max_ts = remaining
    else:
        max_ts = cfl / max_vel
    if max_ts >= remaining:
        return remaining, 1
    n = int(np.ceil(remaining / max_ts))
    return remaining / n, n


class OracleHierarchyState:
    """AMG hierarchy for the oracle's PCG: RS setup once per pattern, full Galerkin refresh
    for a new fine operator, hierarchy values in ``hdtype`` (amg.py:450-598, 749-847)."""

    def __init__(self, row, col, val, project, hdtype=np.float32, base: "Optional[OracleHierarchyState]" = None):
        # one Ruge-Stueben setup per sparsity pattern; later operators of the pattern reuse its
        # interpolation (hierarchy_for, amg.py:952-1045)
        self.h = base.h if base is not None else O.OracleHierarchy(row, col, val.astype(np.float64))
        self.project = project
        self.hdtype = np.dtype(hdtype)
        self.set_fine(row, col, val)

    def set_fine(self, row, col, val, coarse=True):
        levels = self.h.levels
        A = (row, col, val.astype(np.float64))
        self.native = []
        for l, lv in enumerate(levels[:-1]):
            if l > 0 and coarse:
                A = O.galerkin(prev_A, levels[l - 1]["P"], levels[l - 1]["R"], lv["n"])
            elif l > 0:
                A = self.native_fine[l]
            dinv = O.inverse_diagonal(A[0], A[1], A[2].astype(self.hdtype))
            self.native.append(dict(n=lv["n"], A=(A[0], A[1], A[2].astype(self.hdtype)), dinv=dinv,
                                    R=(lv["R"][0], lv["R"][1], lv["R"][2].astype(self.hdtype)),
                                    P=(lv["P"][0], lv["P"][1], lv["P"][2].astype(self.hdtype))))
            prev_A = A
        if coarse:
            Ac = O.galerkin(prev_A, levels[-2]["P"], levels[-2]["R"], levels[-1]["n"])
            nc = levels[-1]["n"]
            D = np.zeros((nc, nc))
            for i in range(nc):
                for k in range(Ac[0][i], Ac[0][i + 1]):
                    D[i, Ac[1][k]] += Ac[2][k]
            self.pinv = np.linalg.pinv(D)
        self.native_fine = {l: lv["A"] for l, lv in enumerate(self.native)}


class OraclePISO:
    """PISO step on an OracleDomain."""

    def __init__(self, od: O.OracleDomain, dt: float, corrector_steps: int = 2, non_orthogonal: bool = False,
                 advection_tol: Optional[float] = None, pressure_tol: Optional[float] = None,
                 pressure_use_amg: bool = False, mhd: Optional[MHDSpec] = None,
                 pre_velocity_setup: Optional[Callable] = None, normalize_pressure: bool = True,
                 max_iter: int = 5000, advect_scalar: bool = True):
        self.od = od
        self.advect_scalar = advect_scalar
        od.set_dt(dt)
        self.dt = dt
        self.corrector_steps = corrector_steps
        self.non_orthogonal = non_orthogonal
        # the reference's default mode (piso_simulation.py:1112-1126): corner terms of the
        # matrices' center and direct neighbours, diagonal neighbours on the right-hand side
        self.flags = O.NO_DEFAULT_MODE if non_orthogonal else 0
        default = 1e-8 if od.dtype == np.float64 else 1e-5
        self.adv_tol = advection_tol if advection_tol is not None else default
        self.p_tol = pressure_tol if pressure_tol is not None else default
        self.pressure_use_amg = pressure_use_amg
        self.mhd = mhd
        self.pre_velocity_setup = pre_velocity_setup
        self.normalize_pressure = normalize_pressure
        self.max_iter = max_iter
        n, d = od.ncells, od.dims
        self.vel_result = np.zeros(d * n, od.dtype)
        self.pressure_result = np.zeros(n, od.dtype)
        self.iters = []
        self.p_hier = None
        self.interp = None  # first hierarchy of the (shared) sparsity pattern
        self.p_refreshes = 0
        self.p_baseline = None
        self.p_stale = False
        if mhd is not None:
            self.e_col, self.e_val = od.epot_matrix()
            self.epot_result = np.zeros(n, od.dtype)
            # pure Neumann unless some face cell prescribes phi (mhd.py:57-90)
            anchored = any(fs.kind == "fixed" and (fs.pot == O.POT_DIRICHLET or
                                                   (fs.pot_types is not None and (fs.pot_types == O.POT_DIRICHLET).any()))
                           for s in od.specs for fs in s.faces)
            self.epot_singular = not anchored
            self.uxb = self._uxb(self.vel_result)
            self.e_hier = None

    # ---- helpers
    def _uxb(self, vel):
        n, d = self.od.ncells, self.od.dims
        v = vel.reshape(d, n)
        if d == 2:
            v = np.concatenate([v, np.zeros((1, n), v.dtype)])
        return cross(v, np.asarray(self.mhd.e_b, v.dtype)).astype(self.od.dtype).reshape(-1)

    def _blocks_velocity_flat(self):
        return self.od.flat_from_blocks(self.od.velocity, self.od.dims)

    def _solve_cg(self, row, col, val, b, x, tol, reset=0, return_best=False, project=False, hier=None):
        if not np.any(b != 0):
            x[...] = 0
            return None
        tol = resolve_tol(tol, b)
        if hier is not None:
            r = O.amg_pcg((row, col, val), hier.native, hier.pinv, b, x, tol, self.max_iter, return_best,
                          hier.project, hdtype=hier.hdtype)
        else:
            r = O.cg(row, col, val, b, x, tol, self.max_iter, True, reset, return_best, project)
        self.iters.append(("cg" if hier is None else "amg", r.used_iterations, r.final_residual, r.converged))
        if not r.is_finite:
            raise RuntimeError("oracle solve: non-finite residual")
        if not r.converged and not return_best:
            raise RuntimeError(f"oracle solve did not converge ({r.used_iterations}, {r.final_residual})")
        return r

    # ---- MHD hooks
    def _lorentz(self):
        od, n, d = self.od, self.od.ncells, self.od.dims
        J = od.current_density(self.epot_result, self.uxb).reshape(3, n)
        F = self.mhd.stuart * cross(J, np.asarray(self.mhd.e_b, J.dtype))
        F = F[:d].astype(od.dtype)
        for bi, s in enumerate(od.specs):
            o, nc = od.offsets[bi], s.ncells
            fb = F[:, o:o + nc]
            base = od.vsource[bi]
            if base is not None:
                if base.size == d:
                    fb = fb + base.reshape(d, 1)
                else:
                    fb = fb + base.reshape(d, nc)
            od.set_source(bi, np.ascontiguousarray(fb.reshape(d, -1)))

    def _magnetic_step(self):
        od = self.od
        n, d = od.ncells, od.dims
        self.uxb = self._uxb(self.vel_result)
        inplane = self.uxb.reshape(3, n)[:d].reshape(-1).copy()
        rhs = od.epot_rhs(inplane)
        if self.epot_singular:
            rhs = rhs - np.mean(rhs)
        x = self.epot_result.copy()
        if self.epot_singular:
            x = x - np.mean(x)
        x = np.ascontiguousarray(x, od.dtype)
        rhs = np.ascontiguousarray(rhs, od.dtype)
        hier = None
        if self.mhd.use_amg:
            if self.e_hier is None:
                self.e_hier = OracleHierarchyState(od.row, self.e_col, self.e_val, self.epot_singular, base=self.interp)
                if self.interp is None:
                    self.interp = self.e_hier
            hier = self.e_hier
        self._solve_cg(od.row, self.e_col, self.e_val, rhs, x, self.mhd.potential_tol,
                       return_best=self.mhd.potential_return_best, project=self.epot_singular, hier=hier)
        if self.epot_singular:
            x = x - np.mean(x)
        self.epot_result = np.ascontiguousarray(x, od.dtype)

    # ---- step
    def _advect_scalar(self):
        """Passive-scalar advection-diffusion before the momentum step (piso_simulation.py:
        3588-3775): one matrix for all channels or one per channel when their diffusivities
        differ; the scalar matrix doubles as the momentum matrix when neither the scalar nor a
        block has its own viscosity. Returns that matrix or None."""
        od = self.od
        n, C_ = od.ncells, od.channels
        split = od.scalar_visc is not None and len(od.scalar_visc) > 1
        for_velocity = not (od.scalar_visc is not None or any(sp.viscosity is not None for sp in od.specs) or split)
        sres = None
        if self.flags:
            sres = np.zeros((C_, n), od.dtype)
            for bi, spec in enumerate(od.specs):
                sres[:, od.offsets[bi]:od.offsets[bi] + spec.ncells] = od.scalar[bi]
        rhs = od.scalar_rhs(self.flags, sres)
        tol = resolve_tol(self.adv_tol, rhs.reshape(-1))
        mats = ([od.scalar_matrix(ch, self.flags) for ch in range(C_)] if split
                else [od.scalar_matrix(0, self.flags)] * C_)
        result = np.zeros_like(rhs)
        for ch in range(C_):
            col, val, _ = mats[ch]
            x = np.zeros(n, od.dtype)
            if np.any(rhs[ch] != 0):
                r = O.bicgstab(od.row, col, val, np.ascontiguousarray(rhs[ch]), x, tol, self.max_iter)
                self.iters.append(("bicg-scalar", r.used_iterations, r.final_residual, r.converged))
                if not r.converged:
                    raise RuntimeError("oracle scalar solve did not converge")
            result[ch] = x
        for bi, spec in enumerate(od.specs):
            o = od.offsets[bi]
            od.scalar[bi][...] = result[:, o:o + spec.ncells]
        return mats[0] if for_velocity else None

    def step(self):
        od = self.od
        n, d = od.ncells, od.dims
        scalar_matrix = self._advect_scalar() if (self.advect_scalar and od.channels) else None
        if self.pre_velocity_setup is not None:
            self.pre_velocity_setup(self)
        if self.mhd is not None:
            self._lorentz()
        ccol, cval, A = scalar_matrix if scalar_matrix is not None else od.advection_matrix(self.flags)
        if not self.non_orthogonal:
            rhs = od.advection_rhs()
            x = self.vel_result
        else:
            self.vel_result = self._blocks_velocity_flat()
            rhs = od.advection_rhs(flags=self.flags, vel_result=self.vel_result)
            x = np.zeros_like(self.vel_result)
        if np.any(rhs != 0):
            adv_tol = resolve_tol(self.adv_tol, rhs)
            for k in range(d):
                xk = np.ascontiguousarray(x[k * n:(k + 1) * n])
                r = O.bicgstab(od.row, ccol, cval, np.ascontiguousarray(rhs[k * n:(k + 1) * n]), xk, adv_tol,
                               self.max_iter)
                self.iters.append(("bicg", r.used_iterations, r.final_residual, r.converged))
                if not r.converged:
                    raise RuntimeError("oracle velocity solve did not converge")
                x[k * n:(k + 1) * n] = xk
        else:
            x[...] = 0
        self.vel_result = x
        for cstep in range(self.corrector_steps):
            pcol, pval = od.pressure_matrix(A, self.flags)
            prhs = od.pressure_rhs(A, ccol, cval, self.vel_result)
            div = od.pressure_div(prhs)
            if self.flags and od.dims > 1:
                # k_pressureRHSaddNonOrthoComponents, pressure.cu:838-860
                div = div + od.nonortho_rhs(O.OK_PRES, self.flags, 0, self.pressure_result, A)
            p = np.zeros(n, od.dtype)
            hier = None
            if self.pressure_use_amg:
                if cstep == 0:
                    # one refresh per new advection diagonal: the finest level every time, the
                    # coarse operators every 4th time or when iterations rose
                    # (piso_simulation.py:2222-2392)
                    if self.p_hier is not None and self.p_refreshes % 4 != 0 and not self.p_stale:
                        self.p_hier.set_fine(od.row, pcol, pval, coarse=False)
                    else:
                        self.p_hier = OracleHierarchyState(od.row, pcol, pval, self._p_singular(pval), base=self.interp)
                        if self.interp is None:
                            self.interp = self.p_hier
                        self.p_refreshes = 0
                        self.p_stale = False
                        self.p_baseline = None
                    self.p_refreshes += 1
                hier = self.p_hier
            r = self._solve_cg(od.row, pcol, pval, div, p, self.p_tol, reset=100 if self.non_orthogonal else 0, hier=hier)
            if hier is not None and r is not None:
                it = r.used_iterations
                if self.p_baseline is None or self.p_refreshes <= 1:
                    self.p_baseline = it if self.p_baseline is None else max(self.p_baseline, it)
                elif it > 2 * self.p_baseline + 5:
                    self.p_stale = True
            if self.normalize_pressure:
                p = p - np.mean(p)
            # the last corrector's pressure system, for residual checks of other solutions
            self.last_pressure_system = (pcol, pval, div)
            self.pressure_result = np.ascontiguousarray(p, od.dtype)
            od.blocks_from_flat(self.pressure_result, od.pressure, 1)
            self.vel_result = od.correct_velocity(A, prhs)
        od.blocks_from_flat(self.vel_result, od.velocity, d)
        if self.mhd is not None:
            # the Lorentz force stays in the velocity source, as in the reference: a caller that
            # drives the flow with a fixed source resets it in ``pre_velocity_setup``
            self._magnetic_step()

    def adaptive_step(self, time_step: float, cfl: float = 0.8) -> list:
        """One step of ``time_step`` in adaptive CFL substeps, each sized from the maximum
        computational velocity (cells and fixed faces) at its start
        (_PISO_adaptive_step, piso_simulation.py:4171-4240). Returns the substep sizes."""
        remaining = time_step
        sizes = []
        while remaining > 0 and not np.isclose(remaining, 0):
            vmax = np.asarray(self.od.max_velocity(True, True), self.od.dtype)
            ts, _ = cfl_substep(remaining, vmax, cfl)
            remaining -= ts
            self.od.set_dt(float(ts))
            self.dt = float(ts)
            self.step()
            sizes.append(float(ts))
        return sizes

    def _p_singular(self, pval):
        """Pure-Neumann pressure operator: every row sum vanishes to within the rounding of the
        value dtype (piso_simulation.py:2165-2220 uses 1e-10 of the largest entry)."""
        row = self.od.row
        sums = np.add.reduceat(pval.astype(np.float64), row[:-1]) if pval.size else np.zeros(0)
        rel = max(1e-10, 64.0 * float(np.finfo(pval.dtype).eps))
        return bool((np.abs(sums) > rel * np.abs(pval).max()).sum() == 0)

"""ORACLE — test infrastructure only.

ctypes wrapper of the C restatement (``oracle/oracle.c``, ``oracle/oracle_amg.c``) of the
reference's PISO kernels, Krylov recurrences and AMG, plus an independent host-side domain
description (vertex coordinates, face specifications) from which oracle domains are built.
Only tests/, ``__graft_entry__.smoke()`` and bench.py's CPU baseline use this module, and only
as the checker.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from dataclasses import dataclass, field
from pathlib import Path
from typing import Optional, Sequence

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "_build" / "liboracle.so"
#: the OpenMP build of the same sources: timing of the CPU baseline only (its sums add in a
#: thread-dependent order, so the single-threaded library is the checker)
LIB_PATH_OMP = HERE / "_build" / "liboracle_omp.so"

OB_CONNECTED, OB_PERIODIC, OB_FIXED = 20, 21, 30
ADV_CENTRAL, ADV_LINEAR_UPWIND = 0, 1  # AdvectionScheme (domain_structs_gpu.h:65-68)
POT_INSULATING, POT_OPEN, POT_DIRICHLET, POT_THIN_WALL = 0, 1, 2, 3


def build() -> Path:
    """Compile the oracle with its Makefile (gcc)."""
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


class OBound(C.Structure):
    _fields_ = [("type", C.c_int32), ("vel_neumann", C.c_int32), ("vel_static", C.c_int32),
                ("has_transform", C.c_int32), ("pot_type", C.c_int32), ("conn_block", C.c_int32),
                ("conn_axes", C.c_int32 * 3), ("pad", C.c_int32), ("vel", C.c_void_p), ("transform", C.c_void_p),
                ("pot_types", C.c_void_p), ("pot_values", C.c_void_p), ("pot_cw", C.c_double)]


class OBlock(C.Structure):
    _fields_ = [("size", C.c_int32 * 3), ("global_offset", C.c_int32), ("has_transform", C.c_int32),
                ("vsource_mode", C.c_int32), ("visc_mode", C.c_int32), ("pad", C.c_int32),
                ("transform", C.c_void_p), ("velocity", C.c_void_p), ("pressure", C.c_void_p),
                ("vsource", C.c_void_p), ("viscosity", C.c_void_p), ("b", OBound * 6)]


class ODomain(C.Structure):
    _fields_ = [("dims", C.c_int32), ("nblocks", C.c_int32), ("ncells", C.c_int32), ("scheme", C.c_int32),
                ("viscosity", C.c_double), ("dt", C.c_double), ("blocks", C.POINTER(OBlock))]


class OFaceScalar(C.Structure):
    _fields_ = [("values", C.c_void_p), ("is_static", C.c_int32), ("neumann", C.c_int32)]


# non-orthogonal flag bits (non_ortho_flags.h) and the reference's default mode
NO_DIRECT_MATRIX, NO_DIRECT_RHS, NO_DIAGONAL_RHS, NO_CENTER_MATRIX = 1, 2, 8, 16
NO_DEFAULT_MODE = NO_CENTER_MATRIX | NO_DIRECT_MATRIX | NO_DIAGONAL_RHS
OK_VEL, OK_PRES, OK_SCAL = 1, 2, 3


class OResult(C.Structure):
    _fields_ = [("final_residual", C.c_double), ("used_iterations", C.c_int32), ("converged", C.c_int32),
                ("is_finite", C.c_int32), ("pad", C.c_int32)]


_LIBS: dict = {}
_THREADED = False


def use_threads(on: bool = True) -> None:
    """Route later oracle calls to the OpenMP build (True) or back to the single-threaded
    checker (False). The thread count is OpenMP's (OMP_NUM_THREADS)."""
    global _THREADED
    _THREADED = bool(on)


def lib() -> C.CDLL:
    path = LIB_PATH_OMP if _THREADED else LIB_PATH
    L = _LIBS.get(path)
    if L is None:
        if not path.exists():
            build()
        L = C.CDLL(str(path))
        L.o_amg_setup.restype = C.c_void_p
        L.o_amg_setup.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_int, C.c_int]
        L.o_amg_num_levels.argtypes = [C.c_void_p]
        L.o_amg_level_sizes.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
        L.o_amg_level_copy.argtypes = [C.c_void_p, C.c_int] + [C.c_void_p] * 10
        L.o_amg_free.argtypes = [C.c_void_p]
        L.o_galerkin.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int] + [C.c_void_p] * 9 + [C.c_int]
        _LIBS[path] = L
    return L


def _p(a: Optional[np.ndarray]) -> Optional[C.c_void_p]:
    return None if a is None else C.c_void_p(a.ctypes.data)


def sfx(dtype) -> str:
    return "_f64" if np.dtype(dtype) == np.float64 else "_f32"


# ----------------------------------------------------------------------------- case spec
@dataclass
class FaceSpec:
    """One face of a block: 'periodic', 'fixed' (Dirichlet/Neumann velocity, potential BC) or
    'connect' (to face ``other_face`` of block ``other`` with tangential target faces)."""

    kind: str = "periodic"
    velocity: Optional[np.ndarray] = None      # [d] static or [d, (z), y, x] face-shaped
    neumann: bool = False
    pot: int = POT_INSULATING
    pot_types: Optional[np.ndarray] = None     # int8 per face cell
    pot_values: Optional[np.ndarray] = None    # per face cell
    cw: float = 0.0
    other: int = -1
    other_face: int = -1
    axis1: int = -1
    axis2: int = -1
    scalar: Optional[np.ndarray] = None        # passive scalar face data: [C] or [C, (z), y, x]
    scalar_neumann: Optional[list] = None      # per channel; default: Neumann on 'neumann' faces


@dataclass
class BlockSpec:
    vertices: np.ndarray                       # [d, (z+1), y+1, x+1]
    faces: list = field(default_factory=list)  # 2d FaceSpec
    velocity: Optional[np.ndarray] = None      # [d, (z), y, x]
    pressure: Optional[np.ndarray] = None      # [(z), y, x]
    vsource: Optional[np.ndarray] = None       # [d] or [d, (z), y, x]
    viscosity: Optional[np.ndarray] = None     # [1] or [(z), y, x]
    scalar: Optional[np.ndarray] = None        # passive scalar [C, (z), y, x]

    @property
    def dims(self) -> int:
        return self.vertices.shape[0]

    @property
    def cells_zyx(self) -> list:
        return [s - 1 for s in self.vertices.shape[1:]]

    @property
    def size_xyz(self) -> list:
        s = list(reversed(self.cells_zyx))
        return s + [1] * (3 - len(s))

    @property
    def ncells(self) -> int:
        return int(np.prod(self.cells_zyx))


def connection_axes(dims: int, f1: int, f2: int, c1: int, c2: int):
    """Face-format axes of both sides of a block connection (boundaries.cpp:857-891)."""
    if c1 < 0:
        c1 = (((f2 >> 1) + 1) % dims) << 1
    if c2 < 0 and dims > 2:
        c2 = (((f2 >> 1) + 2) % dims) << 1
    a1, a2 = [f2], [f1]
    if dims > 1:
        a1.append(c1)
        d1, d2 = f1 >> 1, f2 >> 1
        swapped = False
        if dims == 2 or (c1 >> 1) == (d2 + 1) % dims:
            a2.append((((d1 + 1) % dims) << 1) | (c1 & 1))
        else:
            a2.append((((d1 + 2) % dims) << 1) | (c2 & 1))
            swapped = True
        if dims > 2:
            a1.append(c2)
            a2.append((((d1 + 1) % dims) << 1) | (c1 & 1) if swapped else (((d1 + 2) % dims) << 1) | (c2 & 1))
    return a1 + [0] * (3 - len(a1)), a2 + [0] * (3 - len(a2))


def transforms(coords: np.ndarray, dtype=np.float64) -> np.ndarray:
    """Cell transforms [cells][2d^2+1] of vertex coordinates [d, (z+1), y+1, x+1]."""
    d = coords.shape[0]
    cells = [s - 1 for s in reversed(coords.shape[1:])] + [1] * (3 - d)
    c = np.ascontiguousarray(coords, dtype=dtype)
    out = np.zeros((int(np.prod(cells)), 2 * d * d + 1), dtype=dtype)
    getattr(lib(), "o_transforms" + sfx(dtype))(d, (C.c_int32 * 3)(*cells), _p(c), _p(out))
    return out


def face_transforms(coords: np.ndarray, dtype=np.float64) -> np.ndarray:
    """Face transforms [d, (z+1), y+1, x+1, 2d^2+1]."""
    d = coords.shape[0]
    cells = [s - 1 for s in reversed(coords.shape[1:])] + [1] * (3 - d)
    c = np.ascontiguousarray(coords, dtype=dtype)
    out = np.zeros((d, *coords.shape[1:], 2 * d * d + 1), dtype=dtype)
    getattr(lib(), "o_face_transforms" + sfx(dtype))(d, (C.c_int32 * 3)(*cells), _p(c), _p(out))
    return out


def boundary_slice(ft: np.ndarray, face: int) -> np.ndarray:
    """Face transforms of one block face (Block::GetFaceTransformBoundarySlice, block.cpp:790-816)."""
    d = ft.shape[0]
    axis, upper = face >> 1, face & 1
    a = ft[axis]
    idx = []
    for dd in range(d - 1, -1, -1):
        if dd == axis:
            idx.append(slice(-1, None) if upper else slice(0, 1))
        else:
            idx.append(slice(0, -1))
    idx.append(slice(None))
    return np.ascontiguousarray(a[tuple(idx)])


class OracleDomain:
    """Host domain for the oracle kernels, built from BlockSpecs in one dtype."""

    def __init__(self, specs: Sequence[BlockSpec], viscosity: float, dt: float, dtype=np.float64,
                 scalar_visc: Optional[Sequence[float]] = None, advection_scheme: int = ADV_CENTRAL):
        self.scalar_visc = None if scalar_visc is None else [float(v) for v in np.atleast_1d(scalar_visc)]
        self.advection_scheme = int(advection_scheme)
        self.dtype = np.dtype(dtype)
        self.specs = list(specs)
        self.dims = specs[0].dims
        self.keep = []
        self.blocks = (OBlock * len(specs))()
        off = 0
        self.offsets = []
        self.velocity, self.pressure, self.vsource = [], [], []
        self.face_vel, self.face_tr = {}, {}  # (block, face) -> face velocity / face transforms
        self.cell_tr = []  # per block: cell transforms [cells][2d^2+1]
        for bi, s in enumerate(specs):
            b = self.blocks[bi]
            for i in range(3):
                b.size[i] = s.size_xyz[i]
            b.global_offset = off
            self.offsets.append(off)
            off += s.ncells
            T = transforms(s.vertices, self.dtype)
            FT = face_transforms(s.vertices, self.dtype)
            self.keep += [T, FT]
            self.cell_tr.append(T)
            b.has_transform = 1
            b.transform = _p(T)
            # the state arrays are the oracle's own copies: a step never writes into the case
            vel = np.zeros((self.dims, s.ncells), self.dtype) if s.velocity is None else \
                np.array(s.velocity.reshape(self.dims, -1), self.dtype)
            prs = np.zeros(s.ncells, self.dtype) if s.pressure is None else \
                np.array(s.pressure.reshape(-1), self.dtype)
            self.velocity.append(vel)
            self.pressure.append(prs)
            b.velocity, b.pressure = _p(vel), _p(prs)
            if s.vsource is None:
                b.vsource_mode = 0
                self.vsource.append(None)
            else:
                vs = np.array(s.vsource, self.dtype)
                b.vsource_mode = 1 if vs.ndim == 1 else 2
                vs = vs.reshape(-1)
                self.vsource.append(vs)
                b.vsource = _p(vs)
            if s.viscosity is None:
                b.visc_mode = 0
            else:
                vv = np.ascontiguousarray(np.asarray(s.viscosity, self.dtype).reshape(-1))
                b.visc_mode = 1 if vv.size == 1 else 2
                self.keep.append(vv)
                b.viscosity = _p(vv)
            for f in range(2 * self.dims):
                fs = s.faces[f]
                ob = b.b[f]
                if fs.kind == "periodic":
                    ob.type = OB_PERIODIC
                elif fs.kind == "connect":
                    ob.type = OB_CONNECTED
                    ob.conn_block = fs.other
                    if (bi, f) < (fs.other, fs.other_face):
                        axes, _ = connection_axes(self.dims, f, fs.other_face, fs.axis1, fs.axis2)
                    else:
                        # the connection is made from the partner's side with the partner's axes
                        ps = specs[fs.other].faces[fs.other_face]
                        _, axes = connection_axes(self.dims, fs.other_face, f, ps.axis1, ps.axis2)
                    for i in range(3):
                        ob.conn_axes[i] = axes[i]
                else:
                    ob.type = OB_FIXED
                    ob.vel_neumann = 1 if fs.neumann else 0
                    vel_b = np.zeros(self.dims, self.dtype) if fs.velocity is None else \
                        np.array(fs.velocity, self.dtype)
                    ob.vel_static = 1 if vel_b.ndim == 1 else 0
                    vel_b = vel_b.reshape(-1)
                    bt = boundary_slice(FT, f).reshape(-1)
                    self.keep += [vel_b, bt]
                    ob.vel = _p(vel_b)
                    ob.has_transform = 1
                    ob.transform = _p(bt)
                    self.face_vel[(bi, f)] = vel_b
                    self.face_tr[(bi, f)] = bt
                    ob.pot_type = fs.pot
                    if fs.pot_types is not None:
                        pt = np.ascontiguousarray(fs.pot_types.reshape(-1), np.int8)
                        self.keep.append(pt)
                        ob.pot_types = _p(pt)
                    if fs.pot_values is not None:
                        pv = np.ascontiguousarray(fs.pot_values.reshape(-1), self.dtype)
                        self.keep.append(pv)
                        ob.pot_values = _p(pv)
                    ob.pot_cw = fs.cw
        self.ncells = off
        self.viscosity = float(viscosity)
        self.scalar = [None if s.scalar is None else np.array(s.scalar.reshape(s.scalar.shape[0], -1), self.dtype)
                       for s in specs]
        self.channels = 0 if self.scalar[0] is None else self.scalar[0].shape[0]
        self.dom = ODomain(self.dims, len(specs), off, self.advection_scheme, float(viscosity), float(dt), self.blocks)
        self.row = np.zeros(off + 1, np.int32)
        rc = lib().o_csr_rows_f64(C.byref(self.dom), _p(self.row))
        assert rc == 0
        self.nnz = int(self.row[-1])

    def face_cells(self, bi: int, f: int) -> int:
        size = list(self.specs[bi].size_xyz[:self.dims])
        size[f >> 1] = 1
        return int(np.prod(size))

    def face_velocity(self, bi: int, f: int) -> np.ndarray:
        """Velocity of a fixed face: [d] (static) or [d, face cells] (varying)."""
        v = self.face_vel[(bi, f)]
        return v if self.blocks[bi].b[f].vel_static else v.reshape(self.dims, -1)

    def set_face_velocity(self, bi: int, f: int, vel: np.ndarray) -> None:
        """Replace a fixed face's velocity: [d] (static) or [d, face cells] (varying)."""
        v = np.ascontiguousarray(vel, self.dtype)
        static = v.ndim == 1
        if not static and v.reshape(self.dims, -1).shape[1] != self.face_cells(bi, f):
            raise ValueError("face velocity must have one value per face cell")
        v = v.reshape(-1)
        ob = self.blocks[bi].b[f]
        ob.vel_static = 1 if static else 0
        ob.vel = _p(v)
        self.face_vel[(bi, f)] = v

    def face_transform(self, bi: int, f: int) -> np.ndarray:
        """Face transforms of a fixed face, [face cells, 2d^2+1]."""
        return self.face_tr[(bi, f)].reshape(-1, 2 * self.dims * self.dims + 1)

    def max_velocity(self, with_bounds: bool = True, computational: bool = False, magnitude: bool = False,
                     block: Optional[int] = None) -> float:
        """Largest absolute velocity component (or magnitude) over the blocks' cells and, with
        bounds, their fixed faces' face cells; computational: Minv u with the cell transforms
        and the face transforms (Block::getMaxVelocity[Magnitude], block.cpp:267-273, 505-600;
        FixedBoundary::getVelocity, boundaries.cpp:302-318; Domain::getMaxVelocity,
        domain.cpp:212-243). A NaN velocity gives NaN."""
        d = self.dims
        ts = 2 * d * d + 1

        def measure(v, tr):
            if computational:
                minv = tr.reshape(-1, ts)[:, d * d:2 * d * d].reshape(-1, d, d)
                v = np.einsum("cik,kc->ic", minv, v)
            m = np.sqrt((v * v).sum(axis=0)) if magnitude else np.abs(v)
            return float(np.max(m)) if m.size else 0.0

        best = 0.0
        for bi, s in enumerate(self.specs):
            if block is not None and bi != block:
                continue
            vals = [measure(self.velocity[bi], self.cell_tr[bi])]
            if with_bounds:
                for f in range(2 * d):
                    if (bi, f) not in self.face_vel:
                        continue
                    v = self.face_velocity(bi, f)
                    fc = self.face_cells(bi, f)
                    v = np.repeat(v.reshape(d, 1), fc, axis=1) if v.ndim == 1 else v
                    vals.append(measure(v, self.face_tr[(bi, f)]))
            for x in vals:
                best = x if (x > best or np.isnan(x)) and not np.isnan(best) else best
        return best

    # ---- passive scalar (restated in numpy; the matrix through the C advection matrix)
    def scalar_is_neumann(self, bi: int, f: int, ch: int) -> bool:
        fs = self.specs[bi].faces[f]
        if fs.scalar_neumann is not None:
            return bool(fs.scalar_neumann[ch])
        return bool(fs.neumann) and fs.scalar is None

    def scalar_diffusivity(self, ch: int) -> float:
        """getViscosity with forPassiveScalar (discretization.cuh:1141-1151)."""
        if self.scalar_visc is None:
            return self.viscosity
        return self.scalar_visc[0] if len(self.scalar_visc) == 1 else self.scalar_visc[ch]

    def _face_scalar_table(self, ch: int):
        """Per block face (block * 6 + face) the channel's fixed-face values and condition."""
        table = (OFaceScalar * (6 * len(self.specs)))()
        keep = []
        for bi, spec in enumerate(self.specs):
            for f in range(2 * self.dims):
                fsp = spec.faces[f]
                if fsp.kind != "fixed":
                    continue
                e = table[bi * 6 + f]
                e.neumann = 1 if self.scalar_is_neumann(bi, f, ch) else 0
                if fsp.scalar is not None:
                    a = np.asarray(fsp.scalar, self.dtype)
                    e.is_static = 1 if a.ndim == 1 else 0
                    v = np.ascontiguousarray(a.reshape(self.channels, -1)[ch])
                    keep.append(v)
                    e.values = v.ctypes.data
        return table, keep

    def nonortho_rhs(self, kind: int, flags: int, w: int, data: np.ndarray, A: Optional[np.ndarray] = None) -> np.ndarray:
        """Explicit non-orthogonal corner terms S of every cell (o_nonortho_rhs): kind OK_VEL
        (component w of a [d][n] velocity result), OK_PRES (pressure result, needs A) or OK_SCAL
        (channel w of a [C][n] scalar result)."""
        out = np.zeros(self.ncells, self.dtype)
        table, keep = self._face_scalar_table(w) if kind == OK_SCAL else (None, [])
        data = np.ascontiguousarray(data, self.dtype)
        A = None if A is None else np.ascontiguousarray(A, self.dtype)
        assert self._fn("o_nonortho_rhs")(C.byref(self.dom), int(kind), int(flags), int(w), _p(data), _p(A),
                                          None if table is None else C.cast(table, C.c_void_p), _p(out)) == 0
        del keep
        return out

    def scalar_matrix(self, ch: int, flags: int = 0):
        """The advection-diffusion matrix of passive-scalar channel ch (PISO_build_matrix with
        forPassiveScalar, advection.cu:26-306): the same fluxes with the channel's diffusivity
        for every cell and its boundary condition at the prescribed faces -- the C matrix of
        the oracle on a copy of the domain whose viscosity and wall conditions are replaced."""
        blocks = (OBlock * len(self.specs))()
        C.memmove(blocks, self.blocks, C.sizeof(blocks))
        for bi in range(len(self.specs)):
            blocks[bi].visc_mode = 0
            for f in range(2 * self.dims):
                if blocks[bi].b[f].type == OB_FIXED:
                    blocks[bi].b[f].vel_neumann = 1 if self.scalar_is_neumann(bi, f, ch) else 0
        dom = ODomain(self.dims, len(self.specs), self.ncells, self.advection_scheme, self.scalar_diffusivity(ch),
                      self.dom.dt, blocks)
        col = np.zeros(self.nnz, np.int32)
        val = np.zeros(self.nnz, self.dtype)
        A = np.zeros(self.ncells, self.dtype)
        assert self._fn("o_advection_matrix")(C.byref(dom), int(flags), 1, _p(self.row), _p(col), _p(val), _p(A)) == 0
        return col, val, A

    def face_cell_index(self, bi: int, f: int) -> np.ndarray:
        """Block cell (flat) next to every face cell of face f, in face-cell order."""
        sz = list(self.specs[bi].size_xyz)
        axis, upper = f >> 1, f & 1
        fs = list(sz)
        fs[axis] = 1
        q = np.indices((fs[2], fs[1], fs[0])).reshape(3, -1)  # z, y, x in face-flat order
        pos = {0: q[2], 1: q[1], 2: q[0]}
        pos[axis] = np.full(q.shape[1], sz[axis] - 1 if upper else 0)
        return pos[0] + sz[0] * (pos[1] + sz[1] * pos[2])

    def scalar_rhs(self, flags: int = 0, scalar_result: Optional[np.ndarray] = None) -> np.ndarray:
        """kPISO_build_scalar_advection_RHS, advection.cu:544-700: [C][n]; with LINEAR_UPWIND the
        deferred correction of each channel (``o_deferred_field``) is added after the prescribed
        faces' terms; with the flags' right-hand-side bits the corner terms of ``scalar_result``
        [C][n] times the diffusivity are subtracted before the division by det."""
        use_no = bool(flags & (NO_DIRECT_RHS | NO_DIAGONAL_RHS)) and self.dims > 1
        S = [self.nonortho_rhs(OK_SCAL, flags, ch, scalar_result) for ch in range(self.channels)] if use_no else None
        d, ts = self.dims, 2 * self.dims * self.dims + 1
        C_ = self.channels
        dt = self.dtype.type(self.dom.dt)
        out = np.zeros((C_, self.ncells), self.dtype)
        corr = None
        if self.advection_scheme == ADV_LINEAR_UPWIND:
            glob = np.concatenate([self.scalar[bi] for bi in range(len(self.specs))], axis=1)
            corr = np.zeros((C_, self.ncells), self.dtype)
            for ch in range(C_):
                g = np.ascontiguousarray(glob[ch])
                assert self._fn("o_deferred_field")(C.byref(self.dom), _p(g), _p(corr[ch])) == 0
        for bi, spec in enumerate(self.specs):
            T = self.cell_tr[bi].reshape(-1, ts)
            det = T[:, -1]
            phi = self.scalar[bi]
            o = self.offsets[bi]
            for ch in range(C_):
                t = det * phi[ch] / dt
                for f in range(2 * d):
                    fsp = spec.faces[f]
                    if fsp.kind != "fixed":
                        continue
                    cells = self.face_cell_index(bi, f)
                    fc = cells.size
                    ft = self.face_transform(bi, f)
                    minv = ft[:, d * d:2 * d * d].reshape(fc, d, d)[:, f >> 1, :]
                    detf = ft[:, -1]
                    fn = self.dtype.type((f & 1) * 2 - 1)
                    from .piso_oracle import face_flux
                    flux = face_flux(self, bi, f) * fn
                    visc = self.dtype.type(self.scalar_diffusivity(ch))
                    sv = np.zeros(fc, self.dtype) if fsp.scalar is None else \
                        np.broadcast_to(np.asarray(fsp.scalar, self.dtype).reshape(C_, -1)[ch], (fc,))
                    if not self.scalar_is_neumann(bi, f, ch):
                        alpha = detf * (minv * minv).sum(axis=1)
                        np.add.at(t, cells, -sv * flux + sv * visc * 2 * alpha)
                    else:
                        area = detf * np.sqrt((minv * minv).sum(axis=1))
                        np.add.at(t, cells, -phi[ch][cells] * flux + sv * visc * area)
                if corr is not None:
                    t = t + corr[ch][o:o + spec.ncells]
                if S is not None:
                    t = t - S[ch][o:o + spec.ncells] * self.dtype.type(self.scalar_diffusivity(ch))
                out[ch, o:o + spec.ncells] = t / det
        return out

    def set_dt(self, dt: float) -> None:
        self.dom.dt = float(dt)

    def set_source(self, bi: int, src: Optional[np.ndarray]) -> None:
        b = self.blocks[bi]
        if src is None:
            b.vsource_mode = 0
            self.vsource[bi] = None
            return
        vs = np.ascontiguousarray(src, self.dtype)
        b.vsource_mode = 1 if vs.ndim == 1 else 2
        vs = vs.reshape(-1)
        self.vsource[bi] = vs
        b.vsource = _p(vs)

    def set_velocity_block(self, bi: int, vel: np.ndarray) -> None:
        self.velocity[bi][...] = vel.reshape(self.dims, -1)

    def flat_from_blocks(self, fields, channels) -> np.ndarray:
        out = np.zeros((channels, self.ncells), self.dtype)
        for bi, f in enumerate(fields):
            o = self.offsets[bi]
            out[:, o:o + f.size // channels] = f.reshape(channels, -1)
        return out.reshape(-1)

    def blocks_from_flat(self, flat: np.ndarray, fields, channels) -> None:
        v = flat.reshape(channels, self.ncells)
        for bi, f in enumerate(fields):
            o = self.offsets[bi]
            n = f.size // channels
            f.reshape(channels, -1)[...] = v[:, o:o + n]

    # --- kernels
    def _fn(self, name):
        return getattr(lib(), name + sfx(self.dtype))

    def advection_matrix(self, flags: int = 0):
        col = np.zeros(self.nnz, np.int32)
        val = np.zeros(self.nnz, self.dtype)
        A = np.zeros(self.ncells, self.dtype)
        assert self._fn("o_advection_matrix")(C.byref(self.dom), int(flags), 0, _p(self.row), _p(col), _p(val),
                                              _p(A)) == 0
        return col, val, A

    def advection_rhs(self, apply_pressure_gradient: bool = False, flags: int = 0,
                      vel_result: Optional[np.ndarray] = None) -> np.ndarray:
        """Momentum right-hand side [d][n]; with the flags' right-hand-side bits the corner terms
        of ``vel_result`` [d][n] are subtracted before the division by det."""
        rhs = np.zeros(self.dims * self.ncells, self.dtype)
        S = None
        if flags & (NO_DIRECT_RHS | NO_DIAGONAL_RHS) and self.dims > 1:
            S = np.concatenate([self.nonortho_rhs(OK_VEL, flags, k, vel_result) for k in range(self.dims)])
        assert self._fn("o_advection_rhs")(C.byref(self.dom), int(apply_pressure_gradient), _p(S), _p(rhs)) == 0
        return rhs

    def pressure_matrix(self, A: np.ndarray, flags: int = 0):
        col = np.zeros(self.nnz, np.int32)
        val = np.zeros(self.nnz, self.dtype)
        assert self._fn("o_pressure_matrix")(C.byref(self.dom), _p(A), int(flags), _p(self.row), _p(col), _p(val)) == 0
        return col, val

    def pressure_rhs(self, A, ccol, cval, vel_result) -> np.ndarray:
        prhs = np.zeros(self.dims * self.ncells, self.dtype)
        assert self._fn("o_pressure_rhs")(C.byref(self.dom), _p(A), _p(self.row), _p(ccol), _p(cval),
                                          _p(vel_result), _p(prhs)) == 0
        return prhs

    def pressure_div(self, prhs, time_step_norm: bool = False) -> np.ndarray:
        div = np.zeros(self.ncells, self.dtype)
        assert self._fn("o_pressure_div")(C.byref(self.dom), _p(prhs), int(time_step_norm), _p(div)) == 0
        return div

    def correct_velocity(self, A, prhs, time_step_norm: bool = False) -> np.ndarray:
        vel = np.zeros(self.dims * self.ncells, self.dtype)
        assert self._fn("o_correct_velocity")(C.byref(self.dom), _p(A), _p(prhs), int(time_step_norm), _p(vel)) == 0
        return vel

    def epot_matrix(self, flags: int = 0, A: np.ndarray = None):
        """The potential's Laplacian; with non-orthogonal ``flags`` also its corner terms,
        whose face coefficients weigh the cells with 1/A (A: the advection diagonal)."""
        col = np.zeros(self.nnz, np.int32)
        val = np.zeros(self.nnz, self.dtype)
        a = None if A is None else np.ascontiguousarray(A, self.dtype)
        assert self._fn("o_epot_matrix")(C.byref(self.dom), int(flags), None if a is None else _p(a), _p(self.row),
                                         _p(col), _p(val)) == 0
        return col, val

    def epot_rhs(self, vf: np.ndarray) -> np.ndarray:
        div = np.zeros(self.ncells, self.dtype)
        v = np.ascontiguousarray(vf, self.dtype)
        assert self._fn("o_epot_rhs")(C.byref(self.dom), _p(v), _p(div)) == 0
        return div

    def current_density(self, phi, uxb) -> np.ndarray:
        J = np.zeros(3 * self.ncells, self.dtype)
        assert self._fn("o_current_density")(C.byref(self.dom), _p(np.ascontiguousarray(phi, self.dtype)),
                                             _p(np.ascontiguousarray(uxb, self.dtype)), _p(J)) == 0
        return J


# ----------------------------------------------------------------------------- solvers
def spmv(row, col, val, x):
    y = np.zeros(row.size - 1, val.dtype)
    getattr(lib(), "o_spmv" + sfx(val.dtype))(row.size - 1, _p(row), _p(col), _p(val), _p(x), _p(y))
    return y


def cg(row, col, val, b, x, tol, maxit=5000, normalized=True, reset_every=0, return_best=False, project=False,
       jacobi=False):
    """Fused CG of the reference, optionally Jacobi-preconditioned (returns the OResult; x
    updated in place)."""
    r = OResult()
    getattr(lib(), "o_cg" + sfx(val.dtype))(row.size - 1, _p(row), _p(col), _p(val), _p(b), _p(x), int(maxit),
                                            C.c_double(tol), int(normalized), int(reset_every), int(return_best),
                                            int(project), int(jacobi), C.byref(r))
    return r


def bicgstab(row, col, val, b, x, tol, maxit=5000, normalized=True, jacobi=False):
    """Fused BiCGStab of the reference, optionally right-preconditioned with Jacobi."""
    r = OResult()
    getattr(lib(), "o_bicgstab" + sfx(val.dtype))(row.size - 1, _p(row), _p(col), _p(val), _p(b), _p(x), int(maxit),
                                                  C.c_double(tol), int(normalized), int(jacobi), C.byref(r))
    return r


def csr_transpose(row, col, val):
    n = row.size - 1
    tr = np.zeros(n + 1, np.int32)
    tc = np.zeros(col.size, np.int32)
    tv = np.zeros(val.size, val.dtype)
    getattr(lib(), "o_csr_transpose" + sfx(val.dtype))(n, col.size, _p(row), _p(col), _p(val), _p(tr), _p(tc), _p(tv))
    return tr, tc, tv


class OracleHierarchy:
    """Ruge-Stueben hierarchy of the oracle (double precision)."""

    def __init__(self, row, col, val, theta=0.25, max_levels=20, max_coarse=64):
        L = lib()
        r = np.ascontiguousarray(row, np.int32)
        c = np.ascontiguousarray(col, np.int32)
        v = np.ascontiguousarray(val, np.float64)
        h = L.o_amg_setup(r.size - 1, _p(r), _p(c), _p(v), float(theta), int(max_levels), int(max_coarse))
        self.levels = []
        nl = L.o_amg_num_levels(h)
        for l in range(nl):
            sz = np.zeros(4, np.int32)
            L.o_amg_level_sizes(h, l, _p(sz))
            n, na, npp, nr = [int(s) for s in sz]
            ar, ac, av = np.zeros(n + 1, np.int32), np.zeros(max(na, 1), np.int32), np.zeros(max(na, 1))
            lv = {"n": n}
            if l < nl - 1:
                nn = np.zeros(4, np.int32)
                L.o_amg_level_sizes(h, l + 1, _p(nn))
                pr, pc, pv = np.zeros(n + 1, np.int32), np.zeros(max(npp, 1), np.int32), np.zeros(max(npp, 1))
                rr, rc, rv = np.zeros(nn[0] + 1, np.int32), np.zeros(max(nr, 1), np.int32), np.zeros(max(nr, 1))
                sp = np.zeros(n, np.int32)
                L.o_amg_level_copy(h, l, _p(ar), _p(ac), _p(av), _p(pr), _p(pc), _p(pv), _p(rr), _p(rc), _p(rv), _p(sp))
                lv.update(P=(pr, pc[:npp], pv[:npp]), R=(rr, rc[:nr], rv[:nr]), split=sp)
            else:
                L.o_amg_level_copy(h, l, _p(ar), _p(ac), _p(av), None, None, None, None, None, None, None)
            lv["A"] = (ar, ac[:na], av[:na])
            self.levels.append(lv)
        L.o_amg_free(h)
        An = self.levels[-1]["A"]
        nc = self.levels[-1]["n"]
        D = np.zeros((nc, nc))
        for i in range(nc):
            for k in range(An[0][i], An[0][i + 1]):
                D[i, An[1][k]] += An[2][k]
        self.coarse_pinv = np.linalg.pinv(D)


def galerkin(A, P, R, nc):
    """R A P in double with the oracle's product order."""
    ar, ac, av = A
    pr, pc, pv = P
    rr, rc, rv = R
    cap = ac.size * 8 + 1024
    cr, cc, cv = np.zeros(nc + 1, np.int32), np.zeros(cap, np.int32), np.zeros(cap)
    nnz = lib().o_galerkin(ar.size - 1, _p(ar), _p(ac), _p(np.ascontiguousarray(av, np.float64)), nc, _p(pr), _p(pc),
                           _p(pv), _p(rr), _p(rc), _p(rv), _p(cr), _p(cc), _p(cv), cap)
    assert nnz >= 0
    return cr, cc[:nnz], cv[:nnz]


def amg_pcg(A_fine, levels_h, pinv_h, b, x, tol, maxit=5000, return_best=False, project=False, pre=1, post=1,
            omega=2.0 / 3.0, hdtype=np.float32):
    """AMG-PCG of the reference with a given hierarchy (levels_h: list of dicts holding A, dinv,
    R, P arrays in the hierarchy dtype; the coarsest level enters only through pinv_h)."""
    fr, fc, fv = A_fine
    T = fv.dtype
    hd = np.dtype(hdtype)
    name = {("float64", "float64"): "o_amg_pcg_f64_f64", ("float64", "float32"): "o_amg_pcg_f64_f32",
            ("float32", "float32"): "o_amg_pcg_f32_f32"}[(str(T), str(hd))]
    fn = getattr(lib(), name)
    L = len(levels_h)
    keep = []

    def arr_ptrs(key, idx, dt=None):
        a = (C.c_void_p * L)()
        for l, lv in enumerate(levels_h):
            v = lv[key][idx] if idx is not None else lv[key]
            v = np.ascontiguousarray(v, dt) if dt is not None else np.ascontiguousarray(v)
            keep.append(v)
            a[l] = v.ctypes.data
        return a

    ln = (C.c_int32 * L)(*[lv["n"] for lv in levels_h])
    pinv = np.ascontiguousarray(pinv_h, hd)
    res = OResult()
    fn(fr.size - 1, _p(fr), _p(fc), _p(fv), L, ln, arr_ptrs("A", 0, np.int32), arr_ptrs("A", 1, np.int32),
       arr_ptrs("A", 2, hd), arr_ptrs("dinv", None, hd), arr_ptrs("R", 0, np.int32), arr_ptrs("R", 1, np.int32),
       arr_ptrs("R", 2, hd), arr_ptrs("P", 0, np.int32), arr_ptrs("P", 1, np.int32), arr_ptrs("P", 2, hd),
       _p(pinv), pinv.shape[0], _p(b), _p(x), int(maxit), C.c_double(tol), 1, int(return_best), int(project), int(pre),
       int(post), C.c_double(omega), C.byref(res))
    return res


def inverse_diagonal(row, col, val):
    """1 / (sum of the row's diagonal entries), 1 where that sum vanishes (the V-cycle's Jacobi
    weights, krylov.cu:272-283); the entries are added in row order."""
    n = row.size - 1
    rows = np.repeat(np.arange(n), np.diff(row))
    on = col == rows
    d = np.zeros(n, val.dtype)
    np.add.at(d, rows[on], val[on])
    return np.where(np.abs(d) > 1e-30, 1.0 / np.where(d == 0, 1, d), 1.0).astype(val.dtype)

"""Test domains shared by the oracle (CPU) and the gfx950 path: each case is a list of
``oracle.BlockSpec`` built with numpy only, and ``build_domain`` turns the same specs into a
``phipict_amd.Domain`` on a GPU. Geometry, fields and boundary data are seeded."""
from __future__ import annotations

import math
from typing import Optional, Sequence

import numpy as np

from oracle.oracle import BlockSpec, FaceSpec, POT_DIRICHLET, POT_INSULATING, POT_OPEN, POT_THIN_WALL


# ----------------------------------------------------------------------------- grids
def tanh_coords(n: int, lo: float, hi: float, stretch: float) -> np.ndarray:
    """n+1 vertex coordinates clustered towards both ends."""
    xi = np.linspace(-1.0, 1.0, n + 1)
    w = 0.5 * np.tanh(stretch * xi) / math.tanh(stretch) + 0.5 if stretch > 0 else 0.5 * xi + 0.5
    w[0], w[-1] = 0.0, 1.0
    return lo + (hi - lo) * w


def vertices(axes: Sequence[np.ndarray]) -> np.ndarray:
    """Tensor-product vertices [d, (z+1), y+1, x+1] of per-axis coordinates [xs, ys, (zs)]."""
    if len(axes) == 2:
        Y, X = np.meshgrid(axes[1], axes[0], indexing="ij")
        return np.stack([X, Y])
    Z, Y, X = np.meshgrid(axes[2], axes[1], axes[0], indexing="ij")
    return np.stack([X, Y, Z])


def skewed_vertices(axes: Sequence[np.ndarray], amount: float, seed: int) -> np.ndarray:
    """Rectilinear vertices with seeded interior jitter (a non-orthogonal grid)."""
    v = vertices(axes).copy()
    rng = np.random.default_rng(seed)
    d = v.shape[0]
    inner = tuple([slice(None)] + [slice(1, -1)] * d)
    spacing = min(float(np.min(np.diff(a))) for a in axes)
    v[inner] += amount * spacing * rng.uniform(-1, 1, v[inner].shape)
    return v


def skew(specs: Sequence[BlockSpec], amount: float = 0.25, seed: int = 0) -> list:
    """The blocks of ``specs`` with seeded jitter on every vertex off the block faces (the block
    interfaces stay matched): a non-orthogonal grid of the same topology."""
    rng = np.random.default_rng(seed)
    for spec in specs:
        v = spec.vertices.copy()
        d = v.shape[0]
        inner = tuple([slice(None)] + [slice(1, -1)] * d)
        spacing = min(float(np.min(np.abs(np.diff(v[a], axis=d - 1 - a)))) for a in range(d))
        v[inner] += amount * spacing * rng.uniform(-1, 1, v[inner].shape)
        spec.vertices = v
    return list(specs)


def cell_shape(v: np.ndarray) -> list:
    return [s - 1 for s in v.shape[1:]]


def face_shape(cells_zyx: Sequence[int], face: int) -> list:
    d = len(cells_zyx)
    s = list(cells_zyx)
    s[d - 1 - (face >> 1)] = 1
    return s


def _noise(rng, shape, scale):
    return scale * rng.standard_normal(shape)


# ----------------------------------------------------------------------------- cases
def channel2d(nx: int = 16, ny: int = 12, stretch: float = 1.5, length: float = 2.0, seed: int = 0,
              noise: float = 1e-3, u0: float = 0.0, gradient: Optional[float] = 1.0, wall_pot: int = POT_INSULATING,
              per_cell_source: bool = False) -> list:
    """Hartmann-type 2D channel: periodic in x, no-slip walls at y = -1, 1, driven by a uniform
    pressure-gradient source."""
    rng = np.random.default_rng(seed)
    v = vertices([np.linspace(0.0, length, nx + 1), tanh_coords(ny, -1.0, 1.0, stretch)])
    cz = cell_shape(v)
    faces = [FaceSpec("periodic"), FaceSpec("periodic"), FaceSpec("fixed", pot=wall_pot),
             FaceSpec("fixed", pot=wall_pot)]
    vel = _noise(rng, [2] + cz, noise)
    vel[0] += u0
    src = None
    if gradient is not None:
        src = np.array([gradient, 0.0])
        if per_cell_source:
            src = src.reshape(2, 1, 1) * (1.0 + 0.1 * rng.standard_normal([2] + cz))
    return [BlockSpec(v, faces, velocity=vel, pressure=_noise(rng, cz, 1e-3), vsource=src)]


def duct3d(nx: int = 12, ny: int = 8, nz: int = 10, stretch: float = 1.5, seed: int = 1, noise: float = 1e-3,
           gradient: float = 1.0) -> list:
    """Rectangular 3D duct: periodic in x, closed insulating walls in y and z (wall-refined)."""
    rng = np.random.default_rng(seed)
    v = vertices([np.linspace(0.0, 4.0, nx + 1), tanh_coords(ny, -1.0, 1.0, stretch),
                  tanh_coords(nz, -1.0, 1.0, stretch)])
    cz = cell_shape(v)
    faces = [FaceSpec("periodic"), FaceSpec("periodic")] + [FaceSpec("fixed") for _ in range(4)]
    return [BlockSpec(v, faces, velocity=_noise(rng, [3] + cz, noise), pressure=np.zeros(cz),
                      vsource=np.array([gradient, 0.0, 0.0]))]


def inflow2d(nx: int = 20, ny: int = 10, stretch: float = 1.5, seed: int = 2, U: float = 1.0,
             varying: bool = False) -> list:
    """Plug flow through a 2D channel: Dirichlet inflow U at -x, the same outflow at +x (a balanced
    plug profile) or, with ``varying``, a per-face-cell inflow profile and a zero-gradient outflow;
    no-slip walls in y."""
    rng = np.random.default_rng(seed)
    v = vertices([np.linspace(0.0, 4.0, nx + 1), tanh_coords(ny, -1.0, 1.0, stretch)])
    cz = cell_shape(v)
    if varying:
        prof = np.zeros([2] + face_shape(cz, 0))
        yc = 0.5 * (v[1, 1:, 0] + v[1, :-1, 0])
        prof[0, :, 0] = U * (1 - yc ** 2) * 1.5
        inflow = FaceSpec("fixed", velocity=prof, pot=POT_OPEN)
        outflow = FaceSpec("fixed", neumann=True, pot=POT_OPEN)
    else:
        inflow = FaceSpec("fixed", velocity=np.array([U, 0.0]), pot=POT_OPEN)
        outflow = FaceSpec("fixed", velocity=np.array([U, 0.0]), pot=POT_OPEN)
    faces = [inflow, outflow, FaceSpec("fixed"), FaceSpec("fixed")]
    vel = _noise(rng, [2] + cz, 1e-3)
    vel[0] += U
    return [BlockSpec(v, faces, velocity=vel, pressure=np.zeros(cz))]


def outflow2d(nx: int = 24, ny: int = 12, stretch: float = 1.5, seed: int = 5, U: float = 1.0) -> list:
    """A 2D channel with a parabolic Dirichlet inflow at -x, no-slip walls in y and an outflow
    at +x whose per-face-cell velocity (initially the plug U) is advected out of the domain
    and rescaled to a zero net flux every step (update_advective_boundaries)."""
    rng = np.random.default_rng(seed)
    v = vertices([np.linspace(0.0, 4.0, nx + 1), tanh_coords(ny, -1.0, 1.0, stretch)])
    cz = cell_shape(v)
    yc = 0.5 * (v[1, 1:, 0] + v[1, :-1, 0])
    prof = np.zeros([2] + face_shape(cz, 0))
    prof[0, :, 0] = 1.5 * U * (1 - yc ** 2)
    out = np.zeros([2] + face_shape(cz, 1))
    out[0] = U
    faces = [FaceSpec("fixed", velocity=prof, pot=POT_OPEN), FaceSpec("fixed", velocity=out, pot=POT_OPEN),
             FaceSpec("fixed"), FaceSpec("fixed")]
    vel = _noise(rng, [2] + cz, 1e-3)
    vel[0] += U
    return [BlockSpec(v, faces, velocity=vel, pressure=np.zeros(cz))]


def hartmann_c1(nx: int = 64, ny: int = 64, stretch: float = 1.5, length: float = 4.0, seed: int = 0,
                U: float = 1.0) -> list:
    """The 2D Hartmann MHD channel of the CPU-reference workload: x uniform on [0, length], y in
    [-1, 1] tanh-stretched toward both walls; uniform inflow U at -x and an advected
    (zero-gradient) outflow at +x, both with an open potential; insulating no-slip walls in y;
    initial velocity 0 plus seeded N(0, 1e-3) noise."""
    rng = np.random.default_rng(seed)
    v = vertices([np.linspace(0.0, length, nx + 1), tanh_coords(ny, -1.0, 1.0, stretch)])
    cz = cell_shape(v)
    out = np.zeros([2] + face_shape(cz, 1))
    out[0] = U
    faces = [FaceSpec("fixed", velocity=np.array([U, 0.0]), pot=POT_OPEN), FaceSpec("fixed", velocity=out, pot=POT_OPEN),
             FaceSpec("fixed"), FaceSpec("fixed")]
    return [BlockSpec(v, faces, velocity=_noise(rng, [2] + cz, 1e-3), pressure=np.zeros(cz))]


def inflow_duct3d(nx: int = 48, ny: int = 32, nz: int = 24, stretch: float = 1.5, seed: int = 9, U: float = 1.0) -> list:
    """A 3D duct with a uniform inflow U at -x and an advected outflow at +x (open potential at
    both), insulating no-slip walls in y and z (wall-refined); initial plug U plus N(0, 1e-3)."""
    rng = np.random.default_rng(seed)
    v = vertices([np.linspace(0.0, 4.0, nx + 1), tanh_coords(ny, -1.0, 1.0, stretch),
                  tanh_coords(nz, -1.0, 1.0, stretch)])
    cz = cell_shape(v)
    out = np.zeros([3] + face_shape(cz, 1))
    out[0] = U
    faces = [FaceSpec("fixed", velocity=np.array([U, 0.0, 0.0]), pot=POT_OPEN),
             FaceSpec("fixed", velocity=out, pot=POT_OPEN)] + [FaceSpec("fixed") for _ in range(4)]
    vel = _noise(rng, [3] + cz, 1e-3)
    vel[0] += U
    return [BlockSpec(v, faces, velocity=vel, pressure=np.zeros(cz))]


def split_channel2d(nx: int = 16, ny: int = 12, split: int = 6, seed: int = 0, stretch: float = 1.5,
                    gradient: float = 1.0) -> tuple:
    """The channel of ``channel2d`` as one block and as two blocks joined along x (the second
    block's +x face connected back to the first block's -x face). Returns (single, pair)."""
    single = channel2d(nx, ny, stretch=stretch, seed=seed, gradient=gradient)
    s = single[0]
    v = s.vertices
    left = BlockSpec(np.ascontiguousarray(v[:, :, :split + 1]),
                     [FaceSpec("connect", other=1, other_face=1), FaceSpec("connect", other=1, other_face=0),
                      FaceSpec("fixed"), FaceSpec("fixed")],
                     velocity=np.ascontiguousarray(s.velocity[:, :, :split]),
                     pressure=np.ascontiguousarray(s.pressure[:, :split]), vsource=s.vsource)
    right = BlockSpec(np.ascontiguousarray(v[:, :, split:]),
                      [FaceSpec("connect", other=0, other_face=1), FaceSpec("connect", other=0, other_face=0),
                       FaceSpec("fixed"), FaceSpec("fixed")],
                      velocity=np.ascontiguousarray(s.velocity[:, :, split:]),
                      pressure=np.ascontiguousarray(s.pressure[:, split:]), vsource=s.vsource)
    return single, [left, right]


def rotated_pair3d(seed: int = 3) -> list:
    """Two 3D blocks whose connection maps the tangential axes crosswise: block 0's +x face
    (tangential y, z) joins block 1's -x face with block 0's y running along block 1's z and
    block 0's z along block 1's y. All other faces are closed walls."""
    rng = np.random.default_rng(seed)
    n0 = (5, 4, 6)  # x, y, z cells
    n1 = (4, 6, 4)  # block 1: y1 = z0, z1 = y0
    v0 = vertices([np.linspace(0, 1, n0[0] + 1), tanh_coords(n0[1], 0, 1, 1.0), np.linspace(0, 1.5, n0[2] + 1)])
    v1 = vertices([np.linspace(1, 2, n1[0] + 1), np.linspace(0, 1.5, n1[1] + 1), tanh_coords(n1[2], 0, 1, 1.0)])
    f0 = [FaceSpec("fixed"), FaceSpec("connect", other=1, other_face=0, axis1=4, axis2=2)] + \
         [FaceSpec("fixed") for _ in range(4)]
    f1 = [FaceSpec("connect", other=0, other_face=1), FaceSpec("fixed", neumann=True)] + \
         [FaceSpec("fixed") for _ in range(4)]
    c0, c1 = cell_shape(v0), cell_shape(v1)
    b0 = BlockSpec(v0, f0, velocity=_noise(rng, [3] + c0, 0.1), pressure=_noise(rng, c0, 0.1),
                   vsource=np.array([0.5, 0.0, 0.0]))
    b1 = BlockSpec(v1, f1, velocity=_noise(rng, [3] + c1, 0.1), pressure=_noise(rng, c1, 0.1),
                   viscosity=np.array([0.02]))
    return [b0, b1]


def potential_box3d(nx: int = 6, ny: int = 5, nz: int = 4, seed: int = 4, cw: float = 0.1) -> list:
    """A closed 3D box exercising every potential condition: Dirichlet values on -y, thin wall on
    +y, open faces on +-z, and per-cell mixed types (insulating/Dirichlet/thin wall) on -x;
    per-cell viscosity and source."""
    rng = np.random.default_rng(seed)
    v = vertices([tanh_coords(nx, 0, 1, 1.0), np.linspace(0, 1, ny + 1), tanh_coords(nz, 0, 0.8, 0.7)])
    cz = cell_shape(v)
    fs_x = face_shape(cz, 0)
    types = rng.integers(0, 4, size=fs_x).astype(np.int8)
    types[types == POT_OPEN] = POT_INSULATING
    faces = [FaceSpec("fixed", pot_types=types, pot_values=rng.standard_normal(fs_x), cw=cw),
             FaceSpec("fixed", velocity=np.array([0.0, 0.2, 0.1]), pot=POT_INSULATING),
             FaceSpec("fixed", pot=POT_DIRICHLET, pot_values=rng.standard_normal(face_shape(cz, 2))),
             FaceSpec("fixed", pot=POT_THIN_WALL, cw=cw),
             FaceSpec("fixed", pot=POT_OPEN, neumann=True),
             FaceSpec("fixed", pot=POT_OPEN, velocity=rng.standard_normal([3] + face_shape(cz, 5)) * 0.1)]
    return [BlockSpec(v, faces, velocity=_noise(rng, [3] + cz, 0.2), pressure=_noise(rng, cz, 0.2),
                      vsource=_noise(rng, [3] + cz, 0.1), viscosity=0.01 + 0.01 * rng.uniform(size=cz))]


def potential_box2d(nx: int = 7, ny: int = 6, seed: int = 5, cw: float = 0.3) -> list:
    """2D counterpart of ``potential_box3d`` (thin wall, Dirichlet, open, mixed types)."""
    rng = np.random.default_rng(seed)
    v = vertices([tanh_coords(nx, 0, 1, 1.2), np.linspace(0, 1, ny + 1)])
    cz = cell_shape(v)
    fs_x = face_shape(cz, 0)
    types = rng.integers(0, 4, size=fs_x).astype(np.int8)
    faces = [FaceSpec("fixed", pot_types=types, pot_values=rng.standard_normal(fs_x), cw=cw),
             FaceSpec("fixed", pot=POT_THIN_WALL, cw=cw, velocity=np.array([0.3, 0.0])),
             FaceSpec("fixed", pot=POT_DIRICHLET, pot_values=rng.standard_normal(face_shape(cz, 2))),
             FaceSpec("fixed", pot=POT_OPEN, neumann=True)]
    return [BlockSpec(v, faces, velocity=_noise(rng, [2] + cz, 0.2), pressure=_noise(rng, cz, 0.2),
                      vsource=np.array([0.1, -0.2]))]


def heated_channel2d(nx: int = 20, ny: int = 10, channels: int = 2, seed: int = 6, U: float = 1.0,
                     stretch: float = 1.5) -> list:
    """A 2D channel carrying ``channels`` passive scalars: plug inflow U at -x with a per-face-cell
    scalar profile, the same plug leaving at +x (the scalar zero-gradient there), a hot bottom
    wall (Dirichlet scalar) and a top wall with a prescribed scalar flux (Neumann)."""
    rng = np.random.default_rng(seed)
    v = vertices([np.linspace(0.0, 4.0, nx + 1), tanh_coords(ny, -1.0, 1.0, stretch)])
    cz = cell_shape(v)
    C = channels
    inflow_s = 0.2 + 0.1 * rng.standard_normal([C] + face_shape(cz, 0))
    faces = [FaceSpec("fixed", velocity=np.array([U, 0.0]), pot=POT_OPEN, scalar=inflow_s),
             FaceSpec("fixed", velocity=np.array([U, 0.0]), pot=POT_OPEN, scalar_neumann=[True] * C),
             FaceSpec("fixed", scalar=np.linspace(1.0, 0.5, C)),
             FaceSpec("fixed", scalar=np.linspace(0.3, -0.2, C), scalar_neumann=[True] * C)]
    vel = _noise(rng, [2] + cz, 1e-2)
    vel[0] += U
    scal = 0.5 + 0.1 * rng.standard_normal([C] + cz)
    return [BlockSpec(v, faces, velocity=vel, pressure=np.zeros(cz), scalar=scal)]


def tjunction3d(n: int = 4, seed: int = 4, U: float = 1.0) -> list:
    """A 3-block T-junction duct: the main arm (2n x n x n cells, x in [0, 2]) takes an inflow
    u ~ N(U, 0.1 U) (seeded) at -x and joins the junction cube (n^3, x in [2, 3]) whose -y face is
    an outlet and whose +y face continues into the branch cube (n^3, y in [1, 2]) with an outlet
    at +y; all other faces are insulating no-slip walls. The outlets start with 45% of the inflow
    each (the first flux balance rescales them, so the balanced outflow depends on the inflow
    smoothly) and are advected and flux-balanced every step. At n = 32: 65 536 + 2 x 32 768 cells."""
    rng = np.random.default_rng(seed)
    h = 1.0 / n
    v0 = vertices([np.linspace(0.0, 2.0, 2 * n + 1), np.linspace(0.0, 1.0, n + 1), np.linspace(0.0, 1.0, n + 1)])
    v1 = vertices([np.linspace(2.0, 3.0, n + 1), np.linspace(0.0, 1.0, n + 1), np.linspace(0.0, 1.0, n + 1)])
    v2 = vertices([np.linspace(2.0, 3.0, n + 1), np.linspace(1.0, 2.0, n + 1), np.linspace(0.0, 1.0, n + 1)])
    c0 = cell_shape(v0)
    inflow = np.zeros([3] + face_shape(c0, 0))
    inflow[0] = U * (1.0 + 0.1 * rng.standard_normal(face_shape(c0, 0)))
    half = 0.45 * float(inflow[0].sum()) * h * h  # 45% of the inflow through each outlet (area 1)
    out1 = np.zeros([3] + face_shape(cell_shape(v1), 2))
    out1[1] = -half
    out2 = np.zeros([3] + face_shape(cell_shape(v2), 3))
    out2[1] = half
    wall = FaceSpec("fixed")
    f0 = [FaceSpec("fixed", velocity=inflow, pot=POT_OPEN), FaceSpec("connect", other=1, other_face=0),
          wall, wall, wall, wall]
    f1 = [FaceSpec("connect", other=0, other_face=1), wall, FaceSpec("fixed", velocity=out1, pot=POT_OPEN),
          FaceSpec("connect", other=2, other_face=2), wall, wall]
    f2 = [wall, wall, FaceSpec("connect", other=1, other_face=3), FaceSpec("fixed", velocity=out2, pot=POT_OPEN),
          wall, wall]
    blocks = []
    for v, f in ((v0, f0), (v1, f1), (v2, f2)):
        cz = cell_shape(v)
        vel = _noise(rng, [3] + cz, 1e-2)
        blocks.append(BlockSpec(v, f, velocity=vel, pressure=np.zeros(cz)))
    return blocks


def heated_periodic2d(channels: int = 2, seed: int = 8) -> list:
    """``channel2d`` carrying passive scalars: a hot bottom wall (Dirichlet) and a top wall with
    a Dirichlet first channel and a prescribed flux (Neumann) on the others."""
    specs = channel2d(16, 12, seed=seed, u0=0.5)
    rng = np.random.default_rng(seed)
    cz = cell_shape(specs[0].vertices)
    specs[0].faces[2].scalar = np.linspace(1.0, 0.5, channels)
    specs[0].faces[3].scalar = np.linspace(0.0, 0.2, channels)
    specs[0].faces[3].scalar_neumann = [False] + [True] * (channels - 1)
    specs[0].scalar = 0.5 + 0.1 * rng.standard_normal([channels] + cz)
    return specs


CASES = {
    "channel2d": lambda: channel2d(),
    "duct3d": lambda: duct3d(),
    "inflow2d": lambda: inflow2d(),
    "inflow2d_varying": lambda: inflow2d(varying=True),
    "split2d": lambda: split_channel2d()[1],
    "rotated3d": lambda: rotated_pair3d(),
    "potbox3d": lambda: potential_box3d(),
    "potbox2d": lambda: potential_box2d(),
    "tjunction3d": lambda: tjunction3d(),
    "heated2d": lambda: heated_channel2d(),
    "heated_periodic2d": lambda: heated_periodic2d(),
}


# ----------------------------------------------------------------------------- product domain
def build_domain(specs: Sequence[BlockSpec], viscosity: float, dt: float, dtype=None, device=None, batch: int = 1,
                 epot: bool = False, scalar_visc=None):
    """``phipict_amd.Domain`` with the blocks, boundaries and fields of ``specs``."""
    import torch

    import phipict_amd as pa

    dtype = dtype or torch.float64
    device = device or torch.device("cuda", 0)
    dims = specs[0].dims
    channels = specs[0].scalar.shape[0] if specs[0].scalar is not None else 1
    dom = pa.Domain(dims, torch.tensor([viscosity], dtype=dtype), dtype=dtype, device=device,
                    passiveScalarChannels=channels,
                    scalarViscosity=None if scalar_visc is None else torch.tensor(np.atleast_1d(scalar_visc), dtype=dtype))

    def T(a):
        return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device=device)

    blocks = []
    for s in specs:
        vel = T(s.velocity if s.velocity is not None else np.zeros([dims] + s.cells_zyx)).unsqueeze(0)
        prs = T(s.pressure if s.pressure is not None else np.zeros(s.cells_zyx)).reshape(1, 1, *s.cells_zyx)
        blk = dom.CreateBlock(velocity=vel, pressure=prs, vertexCoordinates=T(s.vertices).unsqueeze(0),
                              passiveScalar=None if s.scalar is None else T(s.scalar).unsqueeze(0))
        if s.vsource is not None:
            vs = np.asarray(s.vsource)
            blk.setVelocitySource(T(vs).reshape(1, dims) if vs.ndim == 1 else T(vs).unsqueeze(0))
        if s.viscosity is not None:
            vv = np.asarray(s.viscosity)
            blk.setViscosity(T(vv).reshape(-1) if vv.size == 1 else T(vv).reshape(1, 1, *s.cells_zyx))
        blocks.append(blk)
    for bi, s in enumerate(specs):
        blk = blocks[bi]
        for f, fs in enumerate(s.faces):
            if fs.kind == "periodic":
                continue
            if fs.kind == "connect":
                if (bi, f) < (fs.other, fs.other_face):
                    a1 = "" if fs.axis1 < 0 else fs.axis1
                    a2 = "" if fs.axis2 < 0 else fs.axis2
                    blk.ConnectBlock(f, blocks[fs.other], fs.other_face, a1, a2)
                continue
            sc = None
            if fs.scalar is not None:
                sv = np.asarray(fs.scalar)
                sc = T(sv).reshape(1, -1) if sv.ndim == 1 else T(sv).unsqueeze(0)
            if fs.neumann:
                bound = blk.OpenBoundary(f, passiveScalar=sc)
            else:
                vel = None
                if fs.velocity is not None:
                    fv = np.asarray(fs.velocity)
                    vel = T(fv).reshape(1, dims) if fv.ndim == 1 else T(fv).unsqueeze(0)
                bound = blk.CloseBoundary(f, velocity=vel, passiveScalar=sc)
            if fs.scalar_neumann is not None:
                bound.setPassiveScalarType([pa.BoundaryConditionType.NEUMANN if nm else pa.BoundaryConditionType.DIRICHLET
                                            for nm in fs.scalar_neumann])
            if fs.pot_types is not None:
                bound.setPotentialTypes(torch.as_tensor(fs.pot_types.reshape(-1), dtype=torch.int8),
                                        cw=fs.cw if (fs.pot_types == POT_THIN_WALL).any() else None)
            else:
                bound.setPotentialBC(pa.PotentialBC(fs.pot), cw=fs.cw if fs.pot == POT_THIN_WALL else None)
            if fs.pot_values is not None:
                bound.setPotentialValues(T(fs.pot_values).reshape(-1))
    if batch > 1:
        dom.setBatchSize(batch)
    dom.PrepareSolve()
    dom.setTimeStep(dt)
    if epot:
        dom.SetupEpotOnDomain()
    return dom


def rel_l2(a, b) -> float:
    a = np.asarray(a, np.float64).reshape(-1)
    b = np.asarray(b, np.float64).reshape(-1)
    den = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (den if den > 0 else 1.0))


def to_np(t) -> np.ndarray:
    return t.detach().cpu().numpy()

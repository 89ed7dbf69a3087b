"""Multi-block domain data model: ``Domain``, ``Block``, boundaries and ``CSRmatrix``.

Mirrors the classes the reference binds in ``phipict._C`` (csrc/phipict/bindings.cpp:92-625,
csrc/phipict/domain/domain_structs.h) for the PISO path: the same names, tensor layouts and
methods, with torch tensors owning every field. ``Domain.desc()`` packs the device-side
descriptor table (include/phipict_hip.h) that the gfx950 kernels read; it is rebuilt only
when a tensor of the domain changed (``UpdateDomainData``), like the reference's atlas
(csrc/phipict/domain/domain.cpp:1053-1342).
"""
from __future__ import annotations

import ctypes as C
from enum import IntEnum
from typing import Optional, Sequence

import torch

from . import _lib

_p = _lib.ptr


class BoundaryType(IntEnum):
    """Face types (csrc/phipict/domain/domain_structs_gpu.h:146-161)."""

    CONNECTED_GRID = 20
    PERIODIC = 21
    FIXED = 30


class BoundaryConditionType(IntEnum):
    DIRICHLET = 0
    NEUMANN = 1


class PotentialBC(IntEnum):
    """Electric-potential face types (domain_structs_gpu.h:217-223)."""

    INSULATING = 0
    OPEN = 1
    DIRICHLET = 2
    THIN_WALL = 3


class ConvergenceCriterion(IntEnum):
    NORM2 = 0
    NORM2_NORMALIZED = 1


class AdvectionScheme(IntEnum):
    CENTRAL = 0
    LINEAR_UPWIND = 1


_FACES = {"-x": 0, "+x": 1, "-y": 2, "+y": 3, "-z": 4, "+z": 5}
_AXES = {"x": 0, "y": 1, "z": 2}


def face_index(face) -> int:
    """Face string ('-x', '+y', ...) or index -> index [-x,+x,-y,+y,-z,+z]."""
    if isinstance(face, str):
        if face not in _FACES:
            raise ValueError(f"unknown face {face!r}")
        return _FACES[face]
    return int(face)


def axis_index(axis) -> int:
    if isinstance(axis, str):
        return _AXES[axis]
    return int(axis)


def transform_size(dims: int) -> int:
    return 2 * dims * dims + 1


def _metric_orthogonality(t: torch.Tensor, dims: int) -> tuple:
    """(off-diagonal ``Minv Minv^T`` within 1e-10 relative, exactly zero) of transforms
    ``[..., 2d^2+1]``. The exact test uses the products ``Minv[i][k] Minv[j][k]`` the assembly
    sums, so an exact zero here means the non-orthogonal terms are exactly zero."""
    minv = t.reshape(-1, transform_size(dims))[:, dims * dims:2 * dims * dims].reshape(-1, dims, dims)
    g = (minv.to(torch.float64) @ minv.to(torch.float64).transpose(1, 2))
    diag = torch.diagonal(g, dim1=1, dim2=2)
    off = g - torch.diag_embed(diag)
    scale = diag.abs().amax(dim=1, keepdim=True).clamp_min(1e-300)
    near = bool(((off.abs().amax(dim=2) / scale) <= 1e-10).all())
    prod = minv.unsqueeze(2) * minv.unsqueeze(1)  # [cells, i, j, k]
    eye = torch.eye(dims, dtype=torch.bool, device=t.device)
    exact = bool((prod[:, ~eye] == 0).all())
    return near, exact


def _metric_diagonal(t: torch.Tensor, dims: int) -> bool:
    """Every off-diagonal entry of M and Minv is exactly 0 in the transforms ``[..., 2d^2+1]``
    (rectilinear grids): the kernels may then read the diagonal alone, with the same results."""
    m = t.reshape(-1, transform_size(dims))[:, :2 * dims * dims].reshape(-1, 2, dims, dims)
    off = ~torch.eye(dims, dtype=torch.bool, device=t.device)
    return bool((m[:, :, off] == 0).all())


def transform_vectors(vectors: torch.Tensor, transforms: torch.Tensor, inverse: bool) -> torch.Tensor:
    """Vectors ``[N, d, (z), y, x]`` mapped by the cell transforms ``[1|N, (z), y, x, 2d^2+1]``:
    ``Minv v`` (inverse, physical to computational) or ``M v`` (TransformVectors,
    csrc/phipict/grid/transform_vectors.cu:84-136)."""
    if not 2 < vectors.dim() < 6 or vectors.size(0) < 1:
        raise ValueError("vectors must have batch and channel dimension and be 1-3D.")
    dims = vectors.dim() - 2
    if vectors.size(1) != dims:
        raise ValueError("vectors channel dimension must match spatial dimensionality.")
    if transforms.dim() != vectors.dim():
        raise ValueError("dimensionality of transforms must match vectors.")
    if transforms.size(0) not in (1, vectors.size(0)):
        raise ValueError("transforms batch dimension must be 1 or match the vectors.")
    if transforms.size(-1) != transform_size(dims):
        raise ValueError("transforms channels must match spatial dimensions.")
    if list(transforms.shape[1:-1]) != list(vectors.shape[2:]):
        raise ValueError("spatial dimensions of transforms must match vectors.")
    if transforms.dtype != vectors.dtype or transforms.device != vectors.device:
        raise ValueError("vectors and transforms must share dtype and device.")
    v = vectors.contiguous()
    t = transforms.contiguous()
    out = torch.empty_like(v)
    cells = 1
    for n in vectors.shape[2:]:
        cells *= n
    _lib.call("pp_transform_vectors", dims, v.size(0), cells, _p(v), _p(t), int(t.size(0) > 1), int(bool(inverse)),
              _p(out), _lib.dtype_code(v.dtype), _lib.stream_ptr(v.device))
    return out


# ------------------------------------------------------------------------------------------
class CSRmatrix:
    """CSR matrix with int32 ``row``/``index`` and one or more value arrays of one pattern
    (csrc/phipict/domain/domain_structs.h:76-110; ``getBatchSize() = |value| / nnz``)."""

    def __init__(self, value, index=None, row=None, dtype=None, device=None):
        if isinstance(value, int):
            # CSRmatrix(numValues, numRows, dtype, device)
            nnz, n = value, index
            self.value = torch.zeros(nnz, dtype=dtype, device=device)
            self.index = torch.zeros(nnz, dtype=torch.int32, device=device)
            self.row = torch.zeros(n + 1, dtype=torch.int32, device=device)
        else:
            if value.dim() != 1 or index.dim() != 1 or row.dim() != 1:
                raise ValueError("value, index and row must be 1D")
            if index.dtype != torch.int32 or row.dtype != torch.int32:
                raise TypeError("index and row must be int32")
            if index.numel() > 0 and value.numel() % index.numel() != 0:
                raise ValueError("value size must be a multiple of the index size")
            self.value, self.index, self.row = value, index, row
        # optional ELL copy written by the library's assembly kernels (coalesced SpMV layout):
        # valid while the value tensor is the one it was written with and was not modified
        self.ell_col: Optional[torch.Tensor] = None
        self.ell_val: Optional[torch.Tensor] = None
        self.ell_width = 0
        self._ell_key = None

    def _value_key(self) -> tuple:
        return (self.value.data_ptr(), self.value._version, self.value.numel())

    def mark_ell_valid(self) -> None:
        """The ELL copy now holds the current values (called after the library wrote both)."""
        if self.ell_val is not None:
            self._ell_key = self._value_key()

    def ell_valid(self) -> bool:
        return self.ell_val is not None and self._ell_key == self._value_key()

    def refresh_ell(self) -> None:
        """Rewrite the ELL copy's values from the CSR values when they were changed outside the
        library's assembly (kernels that read the matrix through its ELL copy then see them)."""
        if self.ell_val is None or self.ell_valid():
            return
        nnz = self.index.numel()
        nmat = self.value.numel() // max(nnz, 1)
        _lib.call("pp_ell_from_csr", _lib.ptr(self.row), _lib.ptr(self.index), _lib.ptr(self.value), self.getRows(),
                  nnz, nmat, self.ell_width, _lib.dtype_code(self.value.dtype), None, _lib.ptr(self.ell_val),
                  _lib.dtype_code(self.ell_val.dtype), _lib.stream_ptr(self.value.device))
        self.mark_ell_valid()

    def share_ell_from(self, other: "CSRmatrix") -> None:
        """Use ``other``'s ELL copy for this matrix (same pattern and value storage)."""
        if other.ell_valid() and other.value.data_ptr() == self.value.data_ptr():
            self.ell_col, self.ell_val, self.ell_width = other.ell_col, other.ell_val, other.ell_width
            self._ell_key = self._value_key()

    def getRows(self) -> int:
        return self.row.numel() - 1

    def getNnz(self) -> int:
        return self.index.numel()

    def getSize(self) -> int:
        return self.index.numel()

    def getBatchSize(self) -> int:
        return max(1, self.value.numel() // max(1, self.index.numel()))

    def getDevice(self):
        return self.value.device

    def getDtype(self):
        return self.value.dtype

    def setValue(self, t: torch.Tensor) -> None:
        if t.numel() != self.value.numel():
            raise ValueError("value must have the correct size")
        self.value = t

    def CreateValue(self) -> None:
        self.value = torch.zeros_like(self.value)

    def WithZeroValue(self) -> "CSRmatrix":
        return CSRmatrix(torch.zeros_like(self.value), self.index, self.row)

    def toType(self, dtype) -> "CSRmatrix":
        return CSRmatrix(self.value.to(dtype), self.index, self.row)

    def clone(self) -> "CSRmatrix":
        return CSRmatrix(self.value.clone(), self.index.clone(), self.row.clone())

    def copy(self) -> "CSRmatrix":
        return CSRmatrix(self.value, self.index, self.row)

    def detach(self) -> None:
        self.value = self.value.detach()

    def desc(self, ell_output: bool = False) -> _lib.CsrDesc:
        """Library descriptor. The ELL copy is included when valid, or always with
        ``ell_output`` (assembly kernels that write it)."""
        d = _lib.CsrDesc()
        d.row = _lib.ptr(self.row)
        d.col = _lib.ptr(self.index)
        d.val = _lib.ptr(self.value)
        d.n = self.getRows()
        d.nnz = self.getNnz()
        d.num_matrices = self.getBatchSize()
        if self.ell_val is not None and (ell_output or self.ell_valid()):
            d.ell_width = self.ell_width
            d.ell_col = _lib.ptr(self.ell_col)
            d.ell_val = _lib.ptr(self.ell_val)
        return d

    def __str__(self) -> str:
        return f"CSRmatrix(rows={self.getRows()}, nnz={self.getNnz()}, batch={self.getBatchSize()}, dtype={self.value.dtype})"


# ------------------------------------------------------------------------------------------
class Boundary:
    type: BoundaryType

    def __init__(self, domain: "Domain"):
        self._domain = domain

    def getParentDomain(self) -> "Domain":
        return self._domain


class PeriodicBoundary(Boundary):
    type = BoundaryType.PERIODIC


class ConnectedBoundary(Boundary):
    """Connection to a face of another block; ``axes`` in the reference's face format
    (csrc/phipict/domain/boundaries.cpp:857-891)."""

    type = BoundaryType.CONNECTED_GRID

    def __init__(self, domain, block: "Block", axes: Sequence[int]):
        super().__init__(domain)
        self._block = block
        self.axes = [int(a) for a in axes]

    def getConnectedBlock(self) -> "Block":
        return self._block


class FixedBoundary(Boundary):
    """Prescribed face: velocity (Dirichlet or Neumann, static ``[1|B, d]`` or varying
    ``[1|B, d, (z), y, x]`` with size 1 along the normal), the face transform and the
    electric-potential condition (csrc/phipict/domain/boundaries.cpp)."""

    type = BoundaryType.FIXED

    def __init__(self, domain, block: "Block", face: int, velocity: Optional[torch.Tensor] = None,
                 velocity_type: BoundaryConditionType = BoundaryConditionType.DIRICHLET):
        super().__init__(domain)
        self._block = block
        self.face = face
        dims = domain.getSpatialDims()
        if velocity is None:
            velocity = torch.zeros(1, dims, dtype=domain.getDtype(), device=domain.getDevice())
        self.setVelocity(velocity)
        self.velocityType = BoundaryConditionType(velocity_type)
        self.transform: Optional[torch.Tensor] = None
        self._potentialType = PotentialBC.INSULATING
        self._potentialCw = 0.0
        self.potentialTypes: Optional[torch.Tensor] = None
        self.potentialValues: Optional[torch.Tensor] = None
        # passive scalar: face data (Dirichlet value or Neumann normal gradient), static
        # [1|B, C] or varying [1|B, C, (z), y, x], and the condition of every channel
        self.passiveScalar: Optional[torch.Tensor] = None
        self._scalarTypes: list = []
        self._scalarDefault = BoundaryConditionType.DIRICHLET
        self._scalarTypesDevice: Optional[torch.Tensor] = None

    # velocity
    def setVelocity(self, velocity: torch.Tensor) -> None:
        dims = self._domain.getSpatialDims()
        if velocity.dim() not in (2, dims + 2) or velocity.size(1) != dims:
            raise ValueError("boundary velocity must be [1|B, d] or [1|B, d, (z), y, x]")
        if velocity.dim() == dims + 2:
            expect = self.face_shape()
            if list(velocity.shape[2:]) != expect:
                raise ValueError(f"varying boundary velocity must have spatial shape {expect}")
        self.velocity = velocity.to(device=self._domain.getDevice(), dtype=self._domain.getDtype()).contiguous()
        self._domain._changed = True
        self._domain._bc_generation += 1

    def isVelocityStatic(self) -> bool:
        return self.velocity.dim() == 2

    def makeVelocityVarying(self) -> None:
        if self.isVelocityStatic():
            shape = [self.velocity.size(0), self.velocity.size(1)] + self.face_shape()
            v = self.velocity.reshape(*self.velocity.shape, *([1] * (len(shape) - 2))).expand(*shape)
            self.setVelocity(v.contiguous())

    def getVelocityVarying(self) -> torch.Tensor:
        """The velocity as ``[vb, d, (z), y, x]``; a static velocity expanded over the face."""
        if not self.isVelocityStatic():
            return self.velocity
        shape = [self.velocity.size(0), self.velocity.size(1)] + self.face_shape()
        return self.velocity.reshape(*self.velocity.shape, *([1] * (len(shape) - 2))).expand(*shape).contiguous()

    def getVelocity(self, computational: bool = False) -> torch.Tensor:
        """The velocity; with ``computational`` and a face transform, ``Minv u`` per face cell
        (FixedBoundary::getVelocity, boundaries.cpp:302-318)."""
        if computational and self.hasTransform():
            return transform_vectors(self.getVelocityVarying(), self.transform, True)
        return self.velocity

    # passive scalar (FixedBoundary::setPassiveScalar / setPassiveScalarType, boundaries.cpp)
    def setPassiveScalar(self, t: torch.Tensor) -> None:
        """Static ``[1|B, C]`` or varying ``[1|B, C, (z), y, x]`` face data of the passive scalar."""
        dims = self._domain.getSpatialDims()
        C = self._domain.getPassiveScalarChannels()
        if t.dim() not in (2, dims + 2) or t.size(1) != C:
            raise ValueError(f"passive scalar must be [1|B, {C}] or [1|B, {C}, (z), y, x]")
        if t.dim() == dims + 2 and list(t.shape[2:]) != self.face_shape():
            raise ValueError(f"varying passive scalar must have spatial shape {self.face_shape()}")
        if t.size(0) not in (1, self._domain.getBatchSize()):
            raise ValueError("passive scalar batch must be 1 or the batch size")
        self.passiveScalar = t.to(device=self._domain.getDevice(), dtype=self._domain.getDtype()).contiguous()
        self._domain._changed = True

    def CreatePassiveScalar(self) -> None:
        C = self._domain.getPassiveScalarChannels()
        self.setPassiveScalar(torch.zeros(1, C, dtype=self._domain.getDtype(), device=self._domain.getDevice()))

    def hasPassiveScalar(self) -> bool:
        return self.passiveScalar is not None

    def isPassiveScalarStatic(self) -> bool:
        return self.passiveScalar is None or self.passiveScalar.dim() == 2

    def getPassiveScalarChannels(self) -> int:
        return 0 if self.passiveScalar is None else self.passiveScalar.size(1)

    def setPassiveScalarType(self, types) -> None:
        """One condition for every channel, or one per channel."""
        if isinstance(types, (list, tuple)):
            self._scalarTypes = [BoundaryConditionType(t) for t in types]
        else:
            self._scalarDefault = BoundaryConditionType(types)
            self._scalarTypes = []
        self._scalarTypesDevice = None
        self._domain._changed = True

    def getPassiveScalarTypes(self) -> list:
        C = self._domain.getPassiveScalarChannels()
        if self._scalarTypes:
            if len(self._scalarTypes) != C:
                raise ValueError(f"{len(self._scalarTypes)} passive scalar boundary types for {C} channels")
            return list(self._scalarTypes)
        return [self._scalarDefault] * C

    @property
    def passiveScalarTypes(self) -> list:
        return self.getPassiveScalarTypes()

    def isPassiveScalarBoundaryTypeStatic(self) -> bool:
        return True

    def _scalar_types_device(self) -> Optional[torch.Tensor]:
        if self._domain.getPassiveScalarChannels() == 0:
            return None
        types = self.getPassiveScalarTypes()
        t = self._scalarTypesDevice
        if t is None or t.numel() != len(types):
            t = torch.tensor([int(x) for x in types], dtype=torch.int8).to(self._domain.getDevice())
            self._scalarTypesDevice = t
        return t

    def face_shape(self) -> list:
        """Spatial shape [(z), y, x] of the face (size 1 along the normal)."""
        sizes = list(self._block.getSizesZYX())
        dims = self._domain.getSpatialDims()
        axis = self.face >> 1
        sizes[dims - 1 - axis] = 1
        return sizes

    def face_cells(self) -> int:
        n = 1
        for s in self.face_shape():
            n *= s
        return n

    def setTransform(self, t: Optional[torch.Tensor]) -> None:
        self.transform = t
        self._domain._changed = True
        self._domain._bc_generation += 1
        self._domain._transform_generation += 1

    def hasTransform(self) -> bool:
        return self.transform is not None

    # electric potential (csrc/phipict/domain/boundaries.cpp:896-993)
    def setPotentialBC(self, type: PotentialBC, cw: Optional[float] = None) -> None:
        type = PotentialBC(type)
        if type == PotentialBC.THIN_WALL:
            if cw is None or not cw > 0:
                raise ValueError("A THIN_WALL potential BC needs a wall conductance ratio cw > 0.")
            self._potentialCw = float(cw)
        else:
            if cw is not None and cw != 0.0:
                raise ValueError("cw is only valid for a THIN_WALL potential BC.")
            self._potentialCw = 0.0
        self._potentialType = type
        self.potentialTypes = None
        self._domain._changed = True

    def getPotentialBC(self) -> PotentialBC:
        return self._potentialType

    def getPotentialCw(self) -> float:
        return self._potentialCw

    def setPotentialTypes(self, types: torch.Tensor, cw: Optional[float] = None) -> None:
        if types.is_floating_point():
            raise TypeError("Potential BC types must be an integer tensor of PotentialBC values.")
        if types.numel() != self.face_cells():
            raise ValueError("Potential BC types must have one entry per face cell")
        if int(types.min()) < 0 or int(types.max()) > 3:
            raise ValueError("Potential BC types contain values that are not PotentialBC members.")
        if bool((types == int(PotentialBC.THIN_WALL)).any()):
            if cw is None or not cw > 0:
                raise ValueError("Potential BC types contain THIN_WALL cells but no wall conductance ratio cw was given.")
            self._potentialCw = float(cw)
        else:
            self._potentialCw = 0.0
        self.potentialTypes = types.to(device=self._domain.getDevice(), dtype=torch.int8).contiguous().reshape(-1)
        self._domain._changed = True

    def clearPotentialTypes(self) -> None:
        self.potentialTypes = None
        self._domain._changed = True

    def setPotentialValues(self, values: torch.Tensor) -> None:
        """Dirichlet potential values of the face: one per face cell, shared by the
        environments, or one slice per environment (``[batch, face...]``). A tensor that
        requires grad stays in its graph: the differentiable magnetic step returns gradients
        for it (bc.set_potential_values, src/phipict/bc.py; potentialValuesGrad)."""
        fc, batch = self.face_cells(), self._domain.getBatchSize()
        if values.numel() != fc and not (batch > 1 and values.numel() == batch * fc):
            raise ValueError("Potential values must have one entry per face cell (or one slice per environment)")
        self.potentialValues = values.to(device=self._domain.getDevice(), dtype=self._domain.getDtype()).contiguous().reshape(-1)
        self._domain._changed = True

    def potentialValuesBatched(self) -> bool:
        """The values hold one slice per environment."""
        return self.potentialValues is not None and self.potentialValues.numel() != self.face_cells()

    def clearPotentialValues(self) -> None:
        self.potentialValues = None
        self._domain._changed = True

    def hasPotentialDirichlet(self) -> bool:
        if self.potentialTypes is not None:
            return bool((self.potentialTypes == int(PotentialBC.DIRICHLET)).any())
        return self._potentialType == PotentialBC.DIRICHLET

    def hasPotentialTypes(self) -> bool:
        return self.potentialTypes is not None

    def hasPotentialValues(self) -> bool:
        return self.potentialValues is not None

    def hasPotentialThinWall(self) -> bool:
        if self.potentialTypes is not None:
            return bool((self.potentialTypes == int(PotentialBC.THIN_WALL)).any())
        return self._potentialType == PotentialBC.THIN_WALL

    # legacy flag interface, thin wrappers over the PotentialBC above
    # (domain_structs.h:242-249, boundaries.cpp:968-991)
    def setEpotCw(self, cw: float) -> None:
        """cw > 0 makes the face THIN_WALL; 0 turns a THIN_WALL face back to INSULATING."""
        if not cw >= 0.0:
            raise ValueError("Epot wall conductance Cw must be non-negative.")
        if cw > 0.0:
            self.setPotentialBC(PotentialBC.THIN_WALL, cw)
        elif self._potentialType == PotentialBC.THIN_WALL:
            self.setPotentialBC(PotentialBC.INSULATING)

    def hasEpotCw(self) -> bool:
        return self.hasPotentialThinWall() and self._potentialCw > 0.0

    def getEpotCw(self) -> float:
        return self._potentialCw if self.hasPotentialThinWall() else 0.0

    def setEpotDirichlet(self, dirichlet: bool) -> None:
        """True makes the face DIRICHLET; False turns a DIRICHLET face back to INSULATING."""
        if dirichlet:
            self.setPotentialBC(PotentialBC.DIRICHLET)
        elif self._potentialType == PotentialBC.DIRICHLET:
            self.setPotentialBC(PotentialBC.INSULATING)

    def hasEpotDirichlet(self) -> bool:
        return self.hasPotentialDirichlet()

    def setEpotInsulating(self, insulating: bool) -> None:
        """False makes the face OPEN; True turns an OPEN face back to INSULATING."""
        if not insulating:
            self.setPotentialBC(PotentialBC.OPEN)
        elif self._potentialType == PotentialBC.OPEN:
            self.setPotentialBC(PotentialBC.INSULATING)

    def isEpotInsulating(self) -> bool:
        return self._potentialType != PotentialBC.OPEN

    def GetFluxes(self) -> torch.Tensor:
        """Contravariant boundary fluxes det*Minv[axis].u per face cell, [vb, face cells]."""
        dims = self._domain.getSpatialDims()
        axis = self.face >> 1
        vb = self.velocity.size(0)
        fc = self.face_cells()
        if self.isVelocityStatic():
            vel = self.velocity.reshape(vb, dims, 1).expand(vb, dims, fc)
        else:
            vel = self.velocity.reshape(vb, dims, fc)
        if self.velocityType != BoundaryConditionType.DIRICHLET and self.transform is not None:
            vel = torch.zeros_like(vel)
        if self.transform is None:
            return vel[:, axis, :]
        t = self.transform.reshape(fc, transform_size(dims))
        minv = t[:, dims * dims:2 * dims * dims].reshape(fc, dims, dims)[:, axis, :]  # [fc, d]
        det = t[:, 2 * dims * dims]
        return det[None, :] * (minv.t()[None, :, :] * vel).sum(dim=1)


# ------------------------------------------------------------------------------------------
class Block:
    """One structured block: fields ``velocity [B, d, (z), y, x]``, ``pressure [B, 1, ...]``,
    optional ``epot``, ``velocitySource`` and ``viscosity``, metrics from vertex coordinates,
    and one boundary per face (csrc/phipict/domain/block.cpp)."""

    def __init__(self, domain: "Domain", velocity: torch.Tensor, pressure: Optional[torch.Tensor], name: str):
        self._domain = domain
        self.name = name
        dims = domain.getSpatialDims()
        if velocity.dim() != dims + 2 or velocity.size(1) != dims:
            raise ValueError(f"velocity must have shape [B, {dims}, (z), y, x]")
        self.velocity = velocity.to(device=domain.getDevice(), dtype=domain.getDtype()).contiguous()
        spatial = list(self.velocity.shape[2:])
        if pressure is None:
            pressure = torch.zeros(self.velocity.size(0), 1, *spatial, dtype=domain.getDtype(), device=domain.getDevice())
        self.pressure = pressure.to(device=domain.getDevice(), dtype=domain.getDtype()).contiguous()
        self.epot: Optional[torch.Tensor] = None
        self.passiveScalar: Optional[torch.Tensor] = None
        self.velocitySource: Optional[torch.Tensor] = None
        self.viscosity: Optional[torch.Tensor] = None
        self.vertexCoordinates: Optional[torch.Tensor] = None
        self.transform: Optional[torch.Tensor] = None
        self.faceTransform: Optional[torch.Tensor] = None
        self.globalOffset = 0
        self.csrOffset = 0
        self.boundaries: list[Boundary] = [PeriodicBoundary(domain) for _ in range(2 * dims)]

    # sizes
    def getSpatialDims(self) -> int:
        return self._domain.getSpatialDims()

    def getSizesZYX(self) -> list:
        return list(self.velocity.shape[2:])

    def getSizes(self) -> list:
        """[x, y, z] cell counts (1 for unused dimensions)."""
        s = list(reversed(self.getSizesZYX()))
        return s + [1] * (3 - len(s))

    def getStrides(self) -> int:
        n = 1
        for s in self.getSizesZYX():
            n *= s
        return n

    def getCells(self) -> int:
        return self.getStrides()

    def getBatchSize(self) -> int:
        return self.velocity.size(0)

    def getDevice(self):
        return self._domain.getDevice()

    def getParentDomain(self) -> "Domain":
        return self._domain

    # fields
    def setVelocity(self, v: torch.Tensor) -> None:
        if v.shape != self.velocity.shape:
            raise ValueError("velocity shape mismatch")
        self.velocity = v.contiguous()
        self._domain._changed = True

    def setPressure(self, p: torch.Tensor) -> None:
        if p.shape != self.pressure.shape:
            raise ValueError("pressure shape mismatch")
        self.pressure = p.contiguous()
        self._domain._changed = True

    def CreateEpot(self) -> None:
        self.epot = torch.zeros_like(self.pressure)
        self._domain._changed = True

    def setEpot(self, e: torch.Tensor) -> None:
        self.epot = e.contiguous()
        self._domain._changed = True

    def hasEpot(self) -> bool:
        return self.epot is not None

    # passive scalar [B, C, (z), y, x] (Block::setPassiveScalar, block.cpp)
    def setPassiveScalar(self, t: torch.Tensor) -> None:
        C = self._domain.getPassiveScalarChannels()
        if list(t.shape) != [self.getBatchSize(), C] + self.getSizesZYX():
            raise ValueError(f"passive scalar must be [{self.getBatchSize()}, {C}, (z), y, x]")
        self.passiveScalar = t.to(device=self._domain.getDevice(), dtype=self._domain.getDtype()).contiguous()
        self._domain._changed = True

    def CreatePassiveScalar(self) -> None:
        C = self._domain.getPassiveScalarChannels()
        self.setPassiveScalar(torch.zeros(self.getBatchSize(), C, *self.getSizesZYX(), dtype=self._domain.getDtype(),
                                          device=self._domain.getDevice()))

    def hasPassiveScalar(self) -> bool:
        return self.passiveScalar is not None

    def getPassiveScalarChannels(self) -> int:
        return 0 if self.passiveScalar is None else self.passiveScalar.size(1)

    def hasPassiveScalarViscosity(self) -> bool:
        return False

    def isAllFixedBoundariesPassiveScalarTypeStatic(self) -> bool:
        return True

    def setVelocitySource(self, s: torch.Tensor) -> None:
        """Static ``[1|B, d]`` or per-cell ``[1|B, d, (z), y, x]`` source (block.cpp:275-300)."""
        dims = self.getSpatialDims()
        s = s.to(device=self._domain.getDevice(), dtype=self._domain.getDtype())
        if s.dim() == 2:
            if s.size(1) != dims:
                raise ValueError("static velocity source must be [1|B, d]")
        elif s.dim() == dims + 2:
            if list(s.shape[1:]) != [dims] + self.getSizesZYX():
                raise ValueError("per-cell velocity source must be [1|B, d, (z), y, x]")
        else:
            raise ValueError("velocity source must be [1|B, d] or [1|B, d, (z), y, x]")
        if s.size(0) not in (1, self.getBatchSize()):
            raise ValueError("velocity source batch must be 1 or the batch size")
        self.velocitySource = s.contiguous()
        self._domain._changed = True

    def clearVelocitySource(self) -> None:
        self.velocitySource = None
        self._domain._changed = True

    def hasVelocitySource(self) -> bool:
        return self.velocitySource is not None

    @property
    def isVelocitySourceStatic(self) -> bool:
        return self.velocitySource is not None and self.velocitySource.dim() == 2

    def setViscosity(self, v: torch.Tensor) -> None:
        v = v.to(device=self._domain.getDevice(), dtype=self._domain.getDtype()).contiguous()
        if v.numel() not in (1, self.getCells()):
            raise ValueError("block viscosity must be static [1] or per cell")
        self.viscosity = v
        self._domain._changed = True

    def clearViscosity(self) -> None:
        self.viscosity = None
        self._domain._changed = True

    def hasViscosity(self) -> bool:
        return self.viscosity is not None

    def hasVertexCoordinates(self) -> bool:
        return self.vertexCoordinates is not None

    def isViscosityStatic(self) -> bool:
        return self.viscosity is not None and self.viscosity.numel() == 1

    # metrics
    def setVertexCoordinates(self, coords: torch.Tensor) -> None:
        """Vertex coordinates ``[1, d, (z+1), y+1, x+1]`` -> cell transforms ``[1, (z), y, x, 2d^2+1]``
        and face transforms ``[1, d, (z+1), y+1, x+1, 2d^2+1]`` (Block::setVertexCoordinates,
        block.cpp:995-1011, kernels grid_gen.cu:298-527)."""
        dims = self.getSpatialDims()
        if coords.dim() != dims + 2 or coords.size(0) != 1 or coords.size(1) != dims:
            raise ValueError("vertex coordinates must be [1, d, (z+1), y+1, x+1]")
        if [s - 1 for s in coords.shape[2:]] != self.getSizesZYX():
            raise ValueError("vertex coordinates spatial dimensions must match velocity's +1")
        dev = self._domain.getDevice()
        if dev.type != "cuda":
            raise RuntimeError("metrics are computed by the gfx950 library; the domain must live on a GPU")
        coords = coords.to(device=dev, dtype=self._domain.getDtype()).contiguous()
        self.vertexCoordinates = coords
        ts = transform_size(dims)
        cells = (C.c_int32 * 3)(*self.getSizes())
        self.transform = torch.empty(1, *self.getSizesZYX(), ts, dtype=coords.dtype, device=dev)
        vshape = list(coords.shape[2:])
        self.faceTransform = torch.empty(1, dims, *vshape, ts, dtype=coords.dtype, device=dev)
        code = _lib.dtype_code(coords.dtype)
        s = _lib.stream_ptr(dev)
        _lib.call("pp_coords_to_transforms", dims, cells, _lib.ptr(coords), _lib.ptr(self.transform), code, s)
        _lib.call("pp_coords_to_face_transforms", dims, cells, _lib.ptr(coords), _lib.ptr(self.faceTransform), code, s)
        for b in self.boundaries:
            if isinstance(b, FixedBoundary):
                b.setTransform(self.faceTransformBoundarySlice(b.face))
        self._domain._changed = True
        self._domain._transform_generation += 1

    def setTransform(self, transform: torch.Tensor, faceTransform: Optional[torch.Tensor] = None) -> None:
        dims = self.getSpatialDims()
        if list(transform.shape) != [1] + self.getSizesZYX() + [transform_size(dims)]:
            raise ValueError("transform must be [1, (z), y, x, 2d^2+1]")
        self.vertexCoordinates = None
        self.transform = transform.to(device=self._domain.getDevice(), dtype=self._domain.getDtype()).contiguous()
        self.faceTransform = None if faceTransform is None else faceTransform.contiguous()
        for b in self.boundaries:
            if isinstance(b, FixedBoundary):
                b.setTransform(self.faceTransformBoundarySlice(b.face) if self.faceTransform is not None else None)
        self._domain._changed = True
        self._domain._transform_generation += 1

    def hasTransform(self) -> bool:
        return self.transform is not None

    def getVelocity(self, computational: bool = False) -> torch.Tensor:
        """The velocity; with ``computational``, ``Minv u`` per cell (Block::getVelocity,
        block.cpp:267-273)."""
        if computational:
            if not self.hasTransform():
                raise RuntimeError("Coordinates or Transformations are required to compute computational velocities.")
            return transform_vectors(self.velocity, self.transform, True)
        return self.velocity

    def getMaxVelocity(self, withBounds: bool = True, computational: bool = False) -> torch.Tensor:
        """Largest absolute velocity component over the block (and its fixed faces), a 0-d
        tensor (Block::getMaxVelocity, block.cpp:505-535)."""
        return self._domain._max_velocity(self, withBounds, computational, 0)[-1]

    def getMaxVelocityPerEnv(self, withBounds: bool = True, computational: bool = False) -> torch.Tensor:
        """Largest absolute velocity component of every environment, ``[B]``
        (Block::getMaxVelocityPerEnv, block.cpp:537-568)."""
        return self._domain._max_velocity(self, withBounds, computational, 0)[:-1]

    def getMaxVelocityMagnitude(self, withBounds: bool = True, computational: bool = False) -> torch.Tensor:
        """Largest velocity magnitude over the block (and its fixed faces), a 0-d tensor
        (Block::getMaxVelocityMagnitude, block.cpp:575-600)."""
        return self._domain._max_velocity(self, withBounds, computational, 1)[-1]

    def _transform_component_major(self) -> Optional[torch.Tensor]:
        """The cell transforms as [2d^2+1][cells] for the library's descriptors; rebuilt when
        the transform tensor is replaced or written."""
        t = self.transform
        if t is None:
            return None
        key = (t.data_ptr(), t._version, t.dtype)
        cached = getattr(self, "_tcm", None)
        if cached is not None and cached[0] == key:
            return cached[1]
        ts = transform_size(self.getSpatialDims())
        out = torch.empty(ts, self.getCells(), dtype=t.dtype, device=t.device)
        src = t.contiguous()
        _lib.call("pp_transpose_2d", _lib.ptr(src), self.getCells(), ts, _lib.dtype_code(t.dtype), _lib.ptr(out),
                  _lib.stream_ptr(t.device))
        self._tcm = (key, out, src)
        return out

    def hasFaceTransform(self) -> bool:
        return self.faceTransform is not None

    def isOrthogonal(self) -> bool:
        """True when the metric ``Minv Minv^T`` of every cell is diagonal to 1e-10 relative, i.e.
        the non-orthogonal correction terms of the reference's assembly are negligible (one device
        sync per new transform)."""
        return self._orthogonality()[0]

    def nonOrthoTermsVanish(self) -> bool:
        """True when the non-orthogonal correction terms are exactly zero: the off-diagonal
        products ``Minv[i] . Minv[j]`` of every cell and of every fixed face's transform are 0."""
        if not self._orthogonality()[1]:
            return False
        faces = [b.transform for b in self.getBoundaries() if isinstance(b, FixedBoundary) and b.transform is not None]
        key = (self._domain._transform_generation,) + tuple((t.data_ptr(), t._version) for t in faces)
        cached = getattr(self, "_face_ortho", None)
        if cached is None or cached[0] != key:
            # one device sync per new set of face transforms
            cached = (key, all(_metric_orthogonality(t, self.getSpatialDims())[1] for t in faces))
            self._face_ortho = cached
        return cached[1]

    def isTransformDiagonal(self) -> bool:
        """True when every cell's M and Minv are diagonal matrices (off-diagonal entries exactly
        0): the library's kernels then read the diagonal of the metric alone."""
        return self._orthogonality()[2]

    def _orthogonality(self) -> tuple:
        """(near-orthogonal, corner terms exactly zero, diagonal metric), cached per transform
        (one device sync per new transform)."""
        t = self.transform
        if t is None:
            return True, True, True
        key = (self._domain._transform_generation, t.data_ptr(), t._version)
        if getattr(self, "_ortho_key", None) != key:
            self._ortho = _metric_orthogonality(t, self.getSpatialDims()) + (
                _metric_diagonal(t, self.getSpatialDims()),)
            self._ortho_key = key
        return self._ortho

    def faceTransformBoundarySlice(self, face: int) -> torch.Tensor:
        """Face transforms of one block face, ``[1, (z), y, x, T]`` with size 1 along the
        normal (Block::GetFaceTransformBoundarySlice, block.cpp:790-816)."""
        dims = self.getSpatialDims()
        axis, upper = face >> 1, face & 1
        ft = self.faceTransform[:, axis]
        idx = [slice(None)]
        for d in range(dims - 1, -1, -1):
            if d == axis:
                idx.append(slice(-1, None) if upper else slice(0, 1))
            else:
                idx.append(slice(0, -1))
        idx.append(slice(None))
        return ft[tuple(idx)].contiguous()

    # boundaries
    def getBoundary(self, face) -> Boundary:
        return self.boundaries[face_index(face)]

    def getBoundaries(self) -> list:
        return list(self.boundaries)

    def getFixedBoundaries(self) -> list:
        return [(i, b) for i, b in enumerate(self.boundaries) if isinstance(b, FixedBoundary)]

    def _setBoundary(self, face, boundary: Boundary) -> None:
        self.boundaries[face_index(face)] = boundary
        self._domain._changed = True
        self._domain._bc_generation += 1

    def _make_fixed(self, face: int, velocity, vtype) -> FixedBoundary:
        fb = FixedBoundary(self._domain, self, face, velocity, vtype)
        if self.faceTransform is not None:
            fb.setTransform(self.faceTransformBoundarySlice(face))
        self._setBoundary(face, fb)
        return fb

    def _disconnect(self, face: int) -> None:
        b = self.boundaries[face]
        if isinstance(b, ConnectedBoundary):
            other = b.getConnectedBlock()
            of = b.axes[0]
            if isinstance(other.boundaries[of], ConnectedBoundary) and other.boundaries[of].getConnectedBlock() is self:
                other.boundaries[of] = PeriodicBoundary(self._domain)
                other._make_fixed(of, None, BoundaryConditionType.DIRICHLET)

    def CloseBoundary(self, face, velocity: Optional[torch.Tensor] = None, passiveScalar=None) -> FixedBoundary:
        """No-slip (or prescribed-velocity) Dirichlet face; the passive scalar is Dirichlet, zero
        or ``passiveScalar`` (Block::CloseBoundary, block.cpp:960-964)."""
        f = face_index(face)
        self._disconnect(f)
        bound = self._make_fixed(f, velocity, BoundaryConditionType.DIRICHLET)
        bound.setPassiveScalarType(BoundaryConditionType.DIRICHLET)
        if passiveScalar is not None:
            bound.setPassiveScalar(passiveScalar)
        elif self._domain.getPassiveScalarChannels() > 0:
            bound.CreatePassiveScalar()
        return bound

    def OpenBoundary(self, face, passiveScalar=None) -> FixedBoundary:
        """Zero-gradient (Neumann) velocity face; the passive scalar is zero-gradient unless
        ``passiveScalar`` gives Dirichlet values (Block::OpenBoundary, block.cpp:969-977)."""
        f = face_index(face)
        self._disconnect(f)
        bound = self._make_fixed(f, None, BoundaryConditionType.NEUMANN)
        if passiveScalar is not None:
            bound.setPassiveScalarType(BoundaryConditionType.DIRICHLET)
            bound.setPassiveScalar(passiveScalar)
        else:
            bound.setPassiveScalarType(BoundaryConditionType.NEUMANN)
            if self._domain.getPassiveScalarChannels() > 0:
                bound.CreatePassiveScalar()
        return bound

    def CloseAllBoundaries(self) -> None:
        for f in range(2 * self.getSpatialDims()):
            self.CloseBoundary(f)

    def MakePeriodic(self, axis) -> None:
        a = axis_index(axis)
        for f in (2 * a, 2 * a + 1):
            self._disconnect(f)
            if not isinstance(self.boundaries[f], PeriodicBoundary):
                self._setBoundary(f, PeriodicBoundary(self._domain))

    def MakeAllPeriodic(self) -> None:
        for a in range(self.getSpatialDims()):
            self.MakePeriodic(a)

    def ConnectBlock(self, face, other: "Block", otherFace, axis1="", axis2="") -> None:
        """Connect face ``face`` to ``otherFace`` of ``other``; ``axis1``/``axis2`` name the
        faces of ``other`` the tangential axes map to (ConnectBlocks, boundaries.cpp:857-891)."""
        f1, f2 = face_index(face), face_index(otherFace)
        dims = self.getSpatialDims()

        def ax(a, k):
            if a == "" or a is None or (isinstance(a, int) and a < 0):
                # default: the k-th tangential axis of other's face, lower side
                return (((f2 >> 1) + k) % dims) << 1
            return face_index(a)

        c1 = ax(axis1, 1)
        c2 = ax(axis2, 2) if dims > 2 else 0
        self._disconnect(f1)
        other._disconnect(f2)
        axes1 = [f2]
        axes2 = [f1]
        if dims > 1:
            axes1.append(c1)
            d1, d2 = f1 >> 1, f2 >> 1
            swapped = False
            if dims == 2 or (c1 >> 1) == (d2 + 1) % dims:
                axes2.append((((d1 + 1) % dims) << 1) | (c1 & 1))
            else:
                if (c2 >> 1) != (d2 + 1) % dims:
                    raise ValueError("Invalid connection.")
                axes2.append((((d1 + 2) % dims) << 1) | (c2 & 1))
                swapped = True
            if dims > 2:
                axes1.append(c2)
                if not swapped:
                    axes2.append((((d1 + 2) % dims) << 1) | (c2 & 1))
                else:
                    axes2.append((((d1 + 1) % dims) << 1) | (c1 & 1))
        self._setBoundary(f1, ConnectedBoundary(self._domain, other, axes1))
        other._setBoundary(f2, ConnectedBoundary(self._domain, self, axes2))

    def ComputeCSRSize(self) -> int:
        """Non-zeros of the block's rows (Block::ComputeCSRSize, block.cpp:1145-1155)."""
        dims = self.getSpatialDims()
        size = self.getSizes()
        n = self.getCells() * (2 * dims + 1)
        for d in range(dims):
            area = size[(d + 1) % 3] * size[(d + 2) % 3]
            if isinstance(self.boundaries[2 * d], FixedBoundary):
                n -= area
            if isinstance(self.boundaries[2 * d + 1], FixedBoundary):
                n -= area
        return n

    def __str__(self) -> str:
        return f"Block[{self.name!r} {self.getSpatialDims()}D, size={self.getSizes()}, batch={self.getBatchSize()}]"


# ------------------------------------------------------------------------------------------
class Domain:
    """Multi-block domain (csrc/phipict/domain/domain.cpp): blocks, the flat solve vectors
    ``[B][channel][totalSize]`` and the matrices C (advection), P (pressure) and Epot."""

    def __init__(self, spatialDims: int, viscosity, name: str = "Domain", dtype=torch.float32, device=None,
                 passiveScalarChannels: int = 1, scalarViscosity: Optional[torch.Tensor] = None):
        if spatialDims not in (2, 3):
            raise ValueError("spatialDims must be 2 or 3")
        self._dims = int(spatialDims)
        self._dtype = dtype
        self._device = torch.device(device) if device is not None else torch.device("cuda")
        if self._device.type == "cuda" and self._device.index is None:
            self._device = torch.device("cuda", torch.cuda.current_device())
        if not isinstance(viscosity, torch.Tensor):
            viscosity = torch.tensor([float(viscosity)], dtype=dtype)
        if viscosity.dtype != dtype:
            raise TypeError("Data type of viscosity does not match.")
        # a viscosity that requires grad stays in its graph (the differentiable step returns its
        # gradient); the kernels read its value through the descriptor
        self.viscosity = viscosity.cpu().reshape(-1)
        self.name = name
        self._blocks: list[Block] = []
        self._batch = 1
        self._initialized = False
        self._desc_generation = 0
        self._changed = True
        self._atlas_key = None
        self._dom_key = None
        self.advectionScheme = AdvectionScheme.CENTRAL
        self.totalSize = 0
        self.C: Optional[CSRmatrix] = None
        self.P: Optional[CSRmatrix] = None
        self.Epot: Optional[CSRmatrix] = None
        self.A: Optional[torch.Tensor] = None
        self.velocityRHS = self.velocityResult = None
        self.pressureRHS = self.pressureRHSdiv = self.pressureResult = None
        self.epotRHS: Optional[torch.Tensor] = None
        self.epotResult: Optional[torch.Tensor] = None
        self._epot_enabled = False
        self._epot_flags = (0, False)
        self._time_step: Optional[torch.Tensor] = None
        self._time_step_host: Optional[torch.Tensor] = None
        self._atlas: Optional[torch.Tensor] = None
        self._atlas_bytes: Optional[bytes] = None
        # device copies of the block descriptors: a ring of slots filled by asynchronous copies
        # from pinned staging buffers, so that a changed descriptor never waits for the GPU
        self._atlas_ring: Optional[list] = None
        self._atlas_slot = -1
        self._flux_cache: Optional[tuple] = None
        self._bc_generation = 0  # bumped by every change of a boundary object, velocity or transform
        # bumped by every new cell or face transform: keys the orthogonality caches, which a
        # tensor's address and version alone cannot (a new tensor may reuse a freed address)
        self._transform_generation = 0
        self._desc = _lib.DomainDesc()
        # bumped by every assembly of the advection matrix: the pressure operator depends on
        # the advection diagonal only, so solvers key their pressure caches on it
        self._a_version = 0
        if passiveScalarChannels < 0:
            raise ValueError("passiveScalarChannels must not be negative")
        self._scalar_channels = int(passiveScalarChannels)
        self.passiveScalarViscosity: Optional[torch.Tensor] = None
        self._scalar_visc_device: Optional[torch.Tensor] = None
        self.scalarRHS: Optional[torch.Tensor] = None
        self.scalarResult: Optional[torch.Tensor] = None
        if scalarViscosity is not None:
            self.setScalarViscosity(scalarViscosity)

    # --- basic accessors
    def getSpatialDims(self) -> int:
        return self._dims

    def getDtype(self):
        return self._dtype

    def getDevice(self):
        return self._device

    def getBatchSize(self) -> int:
        return self._batch

    def getTotalSize(self) -> int:
        return self.totalSize

    def getBlocks(self) -> list:
        return list(self._blocks)

    def getBlock(self, i: int) -> Block:
        return self._blocks[i]

    def getNumBlocks(self) -> int:
        return len(self._blocks)

    def IsInitialized(self) -> bool:
        return self._initialized

    def hasEpot(self) -> bool:
        return self.Epot is not None and self.epotRHS is not None and self.epotResult is not None

    # --- passive scalar (domain.cpp; bindings.cpp:498-538)
    def getPassiveScalarChannels(self) -> int:
        return self._scalar_channels

    def hasPassiveScalar(self) -> bool:
        return bool(self._blocks) and self._scalar_channels > 0 and all(b.hasPassiveScalar() for b in self._blocks)

    def CreatePassiveScalarOnBlocks(self) -> None:
        for b in self._blocks:
            if not b.hasPassiveScalar():
                b.CreatePassiveScalar()
        self._initialized = False

    def setViscosity(self, viscosity: torch.Tensor) -> None:
        """The domain's kinematic viscosity (one value; blocks may carry their own). A tensor that
        requires grad stays in its graph: the differentiable step returns its gradient."""
        v = viscosity if isinstance(viscosity, torch.Tensor) else torch.tensor([float(viscosity)], dtype=self._dtype)
        if v.dtype != self._dtype:
            raise TypeError("Data type of viscosity does not match.")
        if v.numel() != 1:
            raise ValueError("the domain viscosity is one value")
        self.viscosity = v.cpu().reshape(-1)
        self._changed = True

    def setScalarViscosity(self, viscosity: torch.Tensor) -> None:
        """Diffusivity of the passive scalar: one value for every channel or one per channel."""
        v = torch.as_tensor(viscosity, dtype=self._dtype).detach().cpu().reshape(-1)
        if v.numel() not in (1, self._scalar_channels):
            raise ValueError("scalar viscosity must have 1 or passiveScalarChannels entries")
        self.passiveScalarViscosity = v
        self._scalar_visc_device = v.to(self._device)
        self._changed = True

    def clearScalarViscosity(self) -> None:
        self.passiveScalarViscosity = None
        self._scalar_visc_device = None
        self._changed = True

    def hasPassiveScalarViscosity(self) -> bool:
        return self.passiveScalarViscosity is not None

    def isPassiveScalarViscosityStatic(self) -> bool:
        return self.passiveScalarViscosity is None or self.passiveScalarViscosity.numel() == 1

    def hasPassiveScalarBlockViscosity(self) -> bool:
        return False

    def isAllFixedBoundariesPassiveScalarTypeStatic(self) -> bool:
        return True

    def setScalarRHS(self, t): self.scalarRHS = self._check_flat(t, self._scalar_channels, "scalarRHS")
    def setScalarResult(self, t): self.scalarResult = self._check_flat(t, self._scalar_channels, "scalarResult")

    def CreateScalarRHS(self) -> None:
        self.scalarRHS = torch.zeros(self._batch * self._scalar_channels * self.totalSize, dtype=self._dtype,
                                     device=self._device)
        self._changed = True

    def CreateScalarResult(self) -> None:
        self.scalarResult = torch.zeros(self._batch * self._scalar_channels * self.totalSize, dtype=self._dtype,
                                        device=self._device)
        self._changed = True

    def setAdvectionScheme(self, scheme) -> None:
        """Convective scheme of the momentum and passive-scalar equations (Domain::setAdvectionScheme,
        bindings.cpp:518-521): CENTRAL, or LINEAR_UPWIND (implicit first-order upwind plus an
        explicit second-order upwind correction on the right-hand side)."""
        self.advectionScheme = AdvectionScheme(scheme)
        self._changed = True

    def getAdvectionScheme(self) -> AdvectionScheme:
        return self.advectionScheme

    # --- blocks
    def CreateBlock(self, velocity: Optional[torch.Tensor] = None, pressure: Optional[torch.Tensor] = None,
                    passiveScalar=None, vertexCoordinates: Optional[torch.Tensor] = None, name: str = "Block") -> Block:
        if velocity is None:
            if vertexCoordinates is None:
                raise ValueError("velocity or vertexCoordinates is required")
            spatial = [s - 1 for s in vertexCoordinates.shape[2:]]
            velocity = torch.zeros(self._batch, self._dims, *spatial, dtype=self._dtype, device=self._device)
        block = Block(self, velocity, pressure, name)
        if vertexCoordinates is not None:
            block.setVertexCoordinates(vertexCoordinates)
        if passiveScalar is not None:
            block.setPassiveScalar(passiveScalar)
        self._blocks.append(block)
        self._initialized = False
        self._changed = True
        return block

    def CreateBlockWithSize(self, size, name: str = "Block") -> Block:
        sx, sy, sz = (list(size) + [1, 1, 1])[:3]
        spatial = [sz, sy, sx][3 - self._dims:]
        return self.CreateBlock(velocity=torch.zeros(self._batch, self._dims, *spatial, dtype=self._dtype,
                                                     device=self._device), name=name)

    # --- batching (Domain::setBatchSize, domain.cpp:1348-1368)
    def setBatchSize(self, batchSize: int) -> None:
        if batchSize < 1:
            raise ValueError("batch size must be positive")
        if batchSize == self._batch:
            return
        for b in self._blocks:
            if b.velocity.size(0) != 1 and b.velocity.size(0) != batchSize:
                raise ValueError("cannot change a batch size > 1 to a different one")
            b.velocity = b.velocity[:1].repeat(batchSize, *([1] * (b.velocity.dim() - 1))).contiguous()
            b.pressure = b.pressure[:1].repeat(batchSize, *([1] * (b.pressure.dim() - 1))).contiguous()
            if b.epot is not None:
                b.epot = b.epot[:1].repeat(batchSize, *([1] * (b.epot.dim() - 1))).contiguous()
            if b.passiveScalar is not None:
                b.passiveScalar = b.passiveScalar[:1].repeat(batchSize, *([1] * (b.passiveScalar.dim() - 1))).contiguous()
        self._batch = batchSize
        self._initialized = False
        self._changed = True

    # --- independent blocks: uneven-size batching (no reference counterpart)
    def setIndependentBlocks(self, independent: bool = True) -> None:
        """Treat every block as an independent instance: the domain then holds instances of
        different grid sizes side by side (no block may be connected to another), whose linear
        systems are the diagonal blocks of the domain's matrices and are solved segmented --
        each instance with its own recurrence, stopping rules, tolerance and mean projection,
        all of them in one launch per solver phase."""
        if independent:
            for b in self._blocks:
                if any(isinstance(bd, ConnectedBoundary) for bd in b.boundaries):
                    raise ValueError("independent blocks cannot be connected to each other")
        self._independent = bool(independent)
        self._segment_cache = None

    def hasIndependentBlocks(self) -> bool:
        return bool(getattr(self, "_independent", False))

    def getSegments(self):
        """Instance row offsets (the blocks' global offsets and the total size) of a domain of
        independent blocks, else None."""
        if not self.hasIndependentBlocks():
            return None
        if not self._initialized:
            self.PrepareSolve()
        return tuple(b.globalOffset for b in self._blocks) + (self.totalSize,)

    def segmentIds(self) -> torch.Tensor:
        """Instance index of every cell (device, int64), cached."""
        segs = self.getSegments()
        cache = getattr(self, "_segment_cache", None)
        if cache is None or cache[0] != segs:
            lengths = torch.tensor([segs[i + 1] - segs[i] for i in range(len(segs) - 1)])
            ids = torch.repeat_interleave(torch.arange(len(segs) - 1), lengths).to(self._device)
            cache = (segs, ids, lengths.to(device=self._device, dtype=self._dtype))
            self._segment_cache = cache
        return cache[1]

    def segmentLengths(self) -> torch.Tensor:
        self.segmentIds()
        return self._segment_cache[2]

    # --- setup (Domain::PrepareSolve, domain.cpp:426-561)
    def PrepareSolve(self) -> None:
        if not self._blocks:
            raise RuntimeError("Domain does not contain any blocks.")
        dims, B = self._dims, self._batch
        csr, total = 0, 0
        for b in self._blocks:
            b.csrOffset = csr
            b.globalOffset = total
            csr += b.ComputeCSRSize()
            total += b.getCells()
            if b.getBatchSize() != B:
                raise RuntimeError("All blocks must have the same batch size.")
        for bi, b in enumerate(self._blocks):
            for f, bound in enumerate(b.boundaries):
                if isinstance(bound, ConnectedBoundary):
                    other = bound.getConnectedBlock()
                    if other not in self._blocks:
                        raise RuntimeError("Connected block is not part of this domain.")
                    ob = other.boundaries[bound.axes[0]]
                    if not (isinstance(ob, ConnectedBoundary) and ob.getConnectedBlock() is b and ob.axes[0] == f):
                        raise RuntimeError("Mismatch in block connection.")
                if isinstance(bound, PeriodicBoundary) and not isinstance(b.boundaries[f ^ 1], PeriodicBoundary):
                    raise RuntimeError("Opposite boundary must also be periodic.")
        self.totalSize = total
        dev, dt = self._device, self._dtype
        row = torch.zeros(total + 1, dtype=torch.int32, device=dev)
        index = torch.zeros(csr, dtype=torch.int32, device=dev)
        self.C = CSRmatrix(torch.zeros(csr * B, dtype=dt, device=dev), index, row)
        self.P = CSRmatrix(torch.zeros(csr * B, dtype=dt, device=dev), index, row)
        width = 2 * dims + 1
        self._ell_col = torch.empty(width * total, dtype=torch.int32, device=dev)
        for M in (self.C, self.P):
            M.ell_col, M.ell_width = self._ell_col, width
            M.ell_val = torch.zeros(B * width * total, dtype=dt, device=dev)
        self.A = torch.zeros(B * total, dtype=dt, device=dev)
        self.velocityRHS = torch.zeros(B * total * dims, dtype=dt, device=dev)
        self.velocityResult = torch.zeros(B * total * dims, dtype=dt, device=dev)
        self.pressureRHS = torch.zeros(B * total * dims, dtype=dt, device=dev)
        self.pressureRHSdiv = torch.zeros(B * total, dtype=dt, device=dev)
        self.pressureResult = torch.zeros(B * total, dtype=dt, device=dev)
        if self.hasPassiveScalar():
            self.CreateScalarRHS()
            self.CreateScalarResult()
        if self._time_step is None or self._time_step.numel() != B:
            self.setTimeStep(torch.ones(B, dtype=dt))
        self._initialized = True
        self._changed = True
        self.UpdateDomainData()
        _lib.call("pp_csr_pattern", C.byref(self._desc), _lib.ptr(row), _lib.ptr(index), _lib.stream_ptr(dev))
        _lib.call("pp_ell_from_csr", _lib.ptr(row), _lib.ptr(index), None, total, csr, 1, width, _lib.dtype_code(dt),
                  _lib.ptr(self._ell_col), None, _lib.dtype_code(dt), _lib.stream_ptr(dev))
        if self._epot_enabled:
            self.SetupEpotOnDomain(*self._epot_flags)

    # --- flat vectors
    def _check_flat(self, t: torch.Tensor, channels: int, name: str) -> torch.Tensor:
        if t.dim() != 1 or t.numel() != self._batch * self.totalSize * channels:
            raise ValueError(f"{name} must be flat with {self._batch * self.totalSize * channels} entries")
        if t.dtype != self._dtype:
            raise TypeError(f"{name} dtype mismatch")
        t = t.contiguous()
        cur = getattr(self, name, None)
        # the descriptors hold pointers only: the same storage again changes nothing
        if cur is None or cur.data_ptr() != t.data_ptr() or cur.numel() != t.numel():
            self._changed = True
        return t

    def setVelocityRHS(self, t): self.velocityRHS = self._check_flat(t, self._dims, "velocityRHS")
    def setVelocityResult(self, t): self.velocityResult = self._check_flat(t, self._dims, "velocityResult")
    def setPressureRHS(self, t): self.pressureRHS = self._check_flat(t, self._dims, "pressureRHS")
    def setPressureRHSdiv(self, t): self.pressureRHSdiv = self._check_flat(t, 1, "pressureRHSdiv")
    def setPressureResult(self, t): self.pressureResult = self._check_flat(t, 1, "pressureResult")
    def setEpotRHS(self, t): self.epotRHS = self._check_flat(t, 1, "epotRHS")
    def setEpotResult(self, t): self.epotResult = self._check_flat(t, 1, "epotResult")

    def setA(self, t):
        self.A = self._check_flat(t, 1, "A")

    def CreateVelocityResult(self):
        self.velocityResult = torch.zeros(self._batch * self.totalSize * self._dims, dtype=self._dtype, device=self._device)
        self._changed = True

    def CreatePressureResult(self):
        self.pressureResult = torch.zeros(self._batch * self.totalSize, dtype=self._dtype, device=self._device)
        self._changed = True

    # --- time step (Domain::setTimeStep, domain.cpp:370-400): one value or one per environment
    def setTimeStep(self, timeStep) -> None:
        if not isinstance(timeStep, torch.Tensor):
            timeStep = torch.tensor([float(timeStep)], dtype=self._dtype)
        ts = timeStep.detach().to("cpu", self._dtype).reshape(-1)
        if ts.numel() == 1:
            ts = ts.repeat(self._batch)
        if ts.numel() != self._batch:
            raise ValueError("time step must have 1 or batch-size entries")
        if self._time_step_host is not None and torch.equal(ts, self._time_step_host):
            return
        self._time_step_host = ts.clone()
        dev_ts = ts.to(self._device)
        if self._time_step is not None and self._time_step.numel() == dev_ts.numel():
            self._time_step.copy_(dev_ts)
        else:
            self._time_step = dev_ts
            self._changed = True

    def getTimeStep(self) -> Optional[torch.Tensor]:
        return self._time_step_host

    # --- electric potential (Domain::SetupEpotOnDomain, domain.cpp:756-781)
    def SetupEpotOnDomain(self, nonOrthoFlags: int = 0, useFaceTransform: bool = False) -> None:
        if not self._initialized:
            raise RuntimeError("Domain is not initialized. Run domain.PrepareSolve() first.")
        if useFaceTransform:
            raise NotImplementedError("face-transform Laplacians are not implemented")
        if nonOrthoFlags and not self.nonOrthoTermsVanish() and not bool((self.A != 0).all()):
            # the corner terms weigh the cells with 1/A (potential.cu:103-150); on an A not yet
            # assembled the reference's matrix would hold infinities
            raise ValueError("the potential's non-orthogonal corner terms weigh the cells with 1/A: assemble the "
                             "advection matrix (domain.A) before SetupEpotOnDomain with non-orthogonal flags")
        self._epot_enabled = True
        self._epot_flags = (nonOrthoFlags, useFaceTransform)
        if self.epotRHS is None or self.epotRHS.numel() != self._batch * self.totalSize:
            self.epotRHS = torch.zeros(self._batch * self.totalSize, dtype=self._dtype, device=self._device)
        if self.epotResult is None or self.epotResult.numel() != self._batch * self.totalSize:
            self.epotResult = torch.zeros(self._batch * self.totalSize, dtype=self._dtype, device=self._device)
        self.Epot = CSRmatrix(torch.zeros(self.P.getNnz(), dtype=self._dtype, device=self._device), self.P.index, self.P.row)
        self.Epot.ell_col, self.Epot.ell_width = self._ell_col, self.P.ell_width
        self.Epot.ell_val = torch.zeros(self.P.ell_width * self.totalSize, dtype=self._dtype, device=self._device)
        for b in self._blocks:
            if b.epot is None:
                b.CreateEpot()
        self._changed = True
        self.UpdateDomainData()
        flags = 0 if self.nonOrthoTermsVanish() else int(nonOrthoFlags)
        _lib.call("pp_setup_epot_matrix", C.byref(self._desc), flags, _lib.stream_ptr(self._device))
        self.Epot.mark_ell_valid()

    # --- descriptors
    def _block_desc(self, b: Block) -> _lib.BlockDesc:
        d = _lib.BlockDesc()
        size = b.getSizes()
        for i in range(3):
            d.size[i] = size[i]
        d.cells = b.getCells()
        d.global_offset = b.globalOffset
        d.csr_offset = b.csrOffset
        d.has_transform = 1 if b.transform is not None else 0
        d.transform_diagonal = 1 if (b.transform is not None and b.isTransformDiagonal()) else 0
        d.transform = _lib.ptr(b._transform_component_major())
        d.velocity = _lib.ptr(b.velocity)
        d.pressure = _lib.ptr(b.pressure)
        d.epot = _lib.ptr(b.epot)
        d.scalar = _lib.ptr(b.passiveScalar) if self.hasPassiveScalar() else 0
        if b.velocitySource is None:
            d.vsource_mode = 0
        else:
            d.vsource_mode = 1 if b.velocitySource.dim() == 2 else 2
            d.vsource_batched = 1 if b.velocitySource.size(0) > 1 else 0
            d.vsource = _lib.ptr(b.velocitySource)
        if b.viscosity is None:
            d.visc_mode = 0
        else:
            d.visc_mode = 1 if b.viscosity.numel() == 1 else 2
            d.viscosity = _lib.ptr(b.viscosity)
        for f in range(2 * self._dims):
            bound = b.boundaries[f]
            bd = d.bnd[f]
            bd.type = int(bound.type)
            if isinstance(bound, FixedBoundary):
                bd.vel_type = int(bound.velocityType)
                bd.vel_static = 1 if bound.isVelocityStatic() else 0
                bd.vel_batched = 1 if bound.velocity.size(0) > 1 else 0
                bd.vel = _lib.ptr(bound.velocity)
                bd.has_transform = 1 if bound.transform is not None else 0
                bd.transform = _lib.ptr(bound.transform)
                bd.pot_type = int(bound.getPotentialBC())
                bd.pot_types = _lib.ptr(bound.potentialTypes)
                bd.pot_values = _lib.ptr(bound.potentialValues)
                bd.pot_values_batched = 1 if bound.potentialValuesBatched() else 0
                if bd.pot_values_batched and bound.potentialValues.numel() != self.getBatchSize() * bound.face_cells():
                    raise ValueError("per-environment potential values do not match the batch size")
                bd.pot_cw = bound.getPotentialCw()
                if self._scalar_channels > 0:
                    ps = bound.passiveScalar
                    bd.scalar = _lib.ptr(ps) if ps is not None else 0
                    bd.scalar_static = 1 if (ps is None or ps.dim() == 2) else 0
                    bd.scalar_batched = 1 if (ps is not None and ps.size(0) > 1) else 0
                    bd.scalar_types = _lib.ptr(bound._scalar_types_device())
            elif isinstance(bound, ConnectedBoundary):
                bd.conn_block = self._blocks.index(bound.getConnectedBlock())
                for i in range(3):
                    bd.conn_axes[i] = bound.axes[i] if i < len(bound.axes) else 0
        return d

    def _block_key(self, b: Block) -> tuple:
        """What a block descriptor depends on beyond the descriptor generation: the tensors it
        points to (replaced tensors change the pointer, in-place writes to the transform its
        version)."""
        t, s, v = b.transform, b.velocitySource, b.viscosity
        return (_p(b.velocity), _p(b.pressure), _p(b.epot), _p(s), -1 if s is None else s.dim(),
                -1 if s is None else s.size(0), _p(v), -1 if v is None else v.numel(), _p(t),
                -1 if t is None else t._version, _p(b.passiveScalar))

    def UpdateDomainData(self) -> None:
        """Repack the device descriptors if any tensor of the domain changed (a no-op when the
        descriptor generation and every referenced tensor are unchanged)."""
        if not self._initialized:
            raise RuntimeError("Domain is not initialized.")
        gen = self._desc_generation
        akey = (gen, tuple(self._block_key(b) for b in self._blocks))
        if akey != self._atlas_key:
            arr = (_lib.BlockDesc * len(self._blocks))(*[self._block_desc(b) for b in self._blocks])
            raw = bytes(arr)
            if raw != self._atlas_bytes:
                self._upload_atlas(raw)
                self._atlas_bytes = raw
            self._atlas_key = akey
        dkey = (gen, _p(self._atlas), _p(self._time_step), _p(self.A), _p(self.velocityRHS), _p(self.velocityResult),
                _p(self.pressureRHS), _p(self.pressureRHSdiv), _p(self.pressureResult), _p(self.epotRHS),
                _p(self.epotResult), float(self.viscosity[0]), _p(self.scalarRHS), _p(self.scalarResult),
                _p(self._scalar_visc_device), self.hasPassiveScalar(),
                *(None if M is None else (_p(M.value), _p(M.ell_val), M.ell_valid()) for M in (self.C, self.P, self.Epot)))
        if dkey == self._dom_key:
            self._changed = False
            return
        self._dom_key = dkey
        d = self._desc
        d.dims = self._dims
        d.num_blocks = len(self._blocks)
        d.num_cells = self.totalSize
        d.batch = self._batch
        d.dtype = _lib.dtype_code(self._dtype)
        d.max_block_cells = max(b.getCells() for b in self._blocks)
        d.viscosity = float(self.viscosity[0])
        d.blocks = _lib.ptr(self._atlas)
        d.time_step = _lib.ptr(self._time_step)
        d.A = _lib.ptr(self.A)
        d.velocity_rhs = _lib.ptr(self.velocityRHS)
        d.velocity_result = _lib.ptr(self.velocityResult)
        d.pressure_rhs = _lib.ptr(self.pressureRHS)
        d.pressure_rhs_div = _lib.ptr(self.pressureRHSdiv)
        d.pressure_result = _lib.ptr(self.pressureResult)
        d.epot_rhs = _lib.ptr(self.epotRHS)
        d.epot_result = _lib.ptr(self.epotResult)
        d.C = self.C.desc(ell_output=True) if self.C is not None else _lib.CsrDesc()
        d.P = self.P.desc(ell_output=True) if self.P is not None else _lib.CsrDesc()
        d.E = self.Epot.desc(ell_output=True) if self.Epot is not None else _lib.CsrDesc()
        scalar = self.hasPassiveScalar()
        d.scalar_channels = self._scalar_channels if scalar else 0
        vs = self._scalar_visc_device
        d.scalar_visc_count = 0 if (vs is None or not scalar) else vs.numel()
        d.scalar_viscosity = _lib.ptr(vs) if scalar else None
        d.scalar_rhs = _lib.ptr(self.scalarRHS) if scalar else None
        d.scalar_result = _lib.ptr(self.scalarResult) if scalar else None
        d.advection_scheme = int(self.advectionScheme)
        self._changed = False

    ATLAS_RING = 16

    def _upload_atlas(self, raw: bytes) -> None:
        """Copy the packed block descriptors to the next device slot of the ring. The copy is
        asynchronous on the current stream; a slot is reused only after the work that was
        enqueued while it was current has finished (an event recorded when it was retired)."""
        import numpy as np

        size = len(raw)
        if self._device.type != "cuda":
            self._atlas = torch.frombuffer(bytearray(raw), dtype=torch.uint8).clone()
            return
        stream = torch.cuda.current_stream(self._device)
        if self._atlas_ring is None or self._atlas_ring[0][0].numel() != size:
            if self._atlas_ring is not None:
                torch.cuda.synchronize(self._device)
            self._atlas_ring = [[torch.empty(size, dtype=torch.uint8, pin_memory=True),
                                 torch.empty(size, dtype=torch.uint8, device=self._device), None]
                                for _ in range(self.ATLAS_RING)]
            self._atlas_slot = -1
        if self._atlas_slot >= 0:
            retired = self._atlas_ring[self._atlas_slot]
            if retired[2] is None:
                retired[2] = torch.cuda.Event()
            retired[2].record(stream)
        slot = (self._atlas_slot + 1) % self.ATLAS_RING
        host, dev, ev = self._atlas_ring[slot]
        if ev is not None:
            ev.synchronize()
        host.numpy()[:] = np.frombuffer(raw, dtype=np.uint8)
        dev.copy_(host, non_blocking=True)
        self._atlas = dev
        self._atlas_slot = slot

    def IsTensorChanged(self) -> bool:
        return self._changed

    # Every mutator of a block, boundary or domain tensor sets _changed; setting it also bumps
    # the descriptor generation that the cached descriptors are keyed on.
    @property
    def _changed(self) -> bool:
        return self._changed_flag

    @_changed.setter
    def _changed(self, value: bool) -> None:
        if value:
            self._desc_generation = getattr(self, "_desc_generation", 0) + 1
        self._changed_flag = bool(value)

    def desc(self) -> _lib.DomainDesc:
        d = self._desc
        if (self._changed or (self.C is not None and d.C.val != self.C.value.data_ptr())
                or (self.P is not None and d.P.val != self.P.value.data_ptr())
                or (self.Epot is not None and d.E.val != self.Epot.value.data_ptr())):
            self.UpdateDomainData()
        return self._desc

    # --- maximum velocity (CFL substeps)
    def _max_velocity(self, block: Optional[Block], withBounds: bool, computational: bool, norm: int) -> torch.Tensor:
        """``[B + 1]`` maxima from one reduction launch: every environment, then all of them."""
        if not self._blocks:
            raise RuntimeError("Domain does not contain any blocks.")
        if not self._initialized:
            raise RuntimeError("Domain is not initialized. Run domain.PrepareSolve() first.")
        blocks = self._blocks if block is None else [block]
        if computational and not all(b.hasTransform() for b in blocks):
            raise RuntimeError("Coordinates or Transformations are required to compute computational velocities.")
        idx = -1 if block is None else self._blocks.index(block)
        out = torch.empty(self._batch + 1, dtype=self._dtype, device=self._device)
        _lib.call("pp_max_velocity", C.byref(self.desc()), idx, int(bool(withBounds)), int(bool(computational)), int(norm),
                  _p(out), _lib.stream_ptr(self._device))
        return out

    def getMaxVelocity(self, withBounds: bool = True, computational: bool = False) -> torch.Tensor:
        """Largest absolute velocity component over all blocks (and fixed faces) and
        environments, a 0-d tensor (Domain::getMaxVelocity, domain.cpp:212-221)."""
        return self._max_velocity(None, withBounds, computational, 0)[-1]

    def getMaxVelocityPerEnv(self, withBounds: bool = True, computational: bool = False) -> torch.Tensor:
        """Largest absolute velocity component of every environment, ``[B]``
        (Domain::getMaxVelocityPerEnv, domain.cpp:223-232)."""
        return self._max_velocity(None, withBounds, computational, 0)[:-1]

    def getMaxVelocityMagnitude(self, withBounds: bool = True, computational: bool = False) -> torch.Tensor:
        """Largest velocity magnitude over all blocks (and fixed faces), a 0-d tensor
        (Domain::getMaxVelocityMagnitude, domain.cpp:234-243)."""
        return self._max_velocity(None, withBounds, computational, 1)[-1]

    # --- diagnostics
    def GetBoundaryFluxBalance(self) -> torch.Tensor:
        """Net outward flux through all prescribed faces, one value per environment
        (Domain::GetGlobalFluxBalance, domain.cpp:296-331)."""
        # cached while no boundary was replaced or written (in-place writes bump _version)
        key = [self._batch, self._bc_generation, len(self._blocks)]
        for b in self._blocks:
            for f, bound in b.getFixedBoundaries():
                t = bound.transform
                key.append((f, bound.velocity._version, int(bound.velocityType), None if t is None else t._version))
        key = tuple(key)
        if self._flux_cache is not None and self._flux_cache[0] == key:
            return self._flux_cache[1].clone()
        total = torch.zeros(self._batch, dtype=torch.float64)
        for b in self._blocks:
            for f, bound in b.getFixedBoundaries():
                sign = 1.0 if (f & 1) else -1.0
                fl = bound.GetFluxes().to(torch.float64).sum(dim=-1).cpu()
                total += sign * (fl if fl.numel() == self._batch else fl.expand(self._batch))
        self._flux_cache = (key, total.clone())
        return total

    def isOrthogonal(self) -> bool:
        return all(b.isOrthogonal() for b in self._blocks)

    def nonOrthoTermsVanish(self) -> bool:
        """True when every block's non-orthogonal correction terms are exactly zero (the
        assembly then skips them)."""
        return all(b.nonOrthoTermsVanish() for b in self._blocks)

    def getVertexCoordinates(self) -> list:
        return [b.vertexCoordinates for b in self._blocks]

    def __str__(self) -> str:
        return f"Domain[{self.name!r} {self._dims}D, blocks={len(self._blocks)}, cells={self.totalSize}, batch={self._batch}]"

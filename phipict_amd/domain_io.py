"""Saving and loading multi-block domains as a ``<name>.json`` (structure, settings) plus
``<name>.npz`` (tensors, referenced by index) pair in the reference's format version 2
(src/phipict/io/domain_io.py:51-760): tensors shared between objects are stored once, each file
is replaced atomically, and both carry a shared save id so that a pair from an interrupted
save is refused on load.
"""
from __future__ import annotations

import json
import os
import tempfile
import uuid
from pathlib import Path
from typing import Any, Optional

import numpy as np
import torch

from .domain import BoundaryConditionType, ConnectedBoundary, Domain, FixedBoundary, PeriodicBoundary, PotentialBC

__all__ = ["load_domain", "save_domain", "FORMAT_VERSION"]

FORMAT_VERSION = 2
_SUFFIXES = (".json", ".npz")
_SAVE_ID = "save_id"
_FACES = ("-x", "+x", "-y", "+y", "-z", "+z")
_DTYPES = {torch.float32: "float32", torch.float64: "float64", torch.int8: "int8"}


def _pair_paths(path: Path) -> tuple:
    if not isinstance(path, Path):
        raise TypeError(f"path must be a pathlib.Path, got {type(path).__name__}: {path!r}.")
    if path.suffix in _SUFFIXES:
        path = path.with_suffix("")
    return path.with_name(path.name + ".json"), path.with_name(path.name + ".npz")


def _atomic_write(target: Path, write) -> None:
    fd, tmp = tempfile.mkstemp(dir=target.parent, prefix=f".{target.name}.", suffix=".tmp")
    try:
        with os.fdopen(fd, "wb") as f:
            write(f)
            f.flush()
            os.fsync(f.fileno())
        os.replace(tmp, target)
    except BaseException:
        Path(tmp).unlink(missing_ok=True)
        raise


def _bc_name(t) -> str:
    return "DIRICHLET" if BoundaryConditionType(t) == BoundaryConditionType.DIRICHLET else "NEUMANN"


def _bc_type(name: str) -> BoundaryConditionType:
    if name == "DIRICHLET":
        return BoundaryConditionType.DIRICHLET
    if name == "NEUMANN":
        return BoundaryConditionType.NEUMANN
    raise TypeError("Unsupported boundary condition type.")


def save_domain(domain: Domain, path: Path) -> None:
    """Write ``domain`` to ``path.json`` + ``path.npz`` (save_domain, domain_io.py:289-453)."""
    json_path, npz_path = _pair_paths(path)
    data: list = []
    index: dict = {}

    def add(t: torch.Tensor, d: dict, name: str) -> None:
        if t.dtype not in _DTYPES:
            raise TypeError(f"Unsupported dtype {t.dtype}.")
        key = (t.device, t.dtype, t.data_ptr(), tuple(t.shape), tuple(t.stride()))
        if key not in index:
            index[key] = len(data)
            data.append(t)
        d[name] = str(index[key])

    dd: dict = {"format_version": FORMAT_VERSION, "name": domain.name, "spatialDims": domain.getSpatialDims()}
    add(domain.viscosity, dd, "viscosity")
    dd["passiveScalarChannels"] = domain.getPassiveScalarChannels()
    if domain.hasPassiveScalarViscosity():
        add(domain.passiveScalarViscosity, dd, "passiveScalarViscosity")
    blocks = domain.getBlocks()
    dd["blocks"] = []
    for block in blocks:
        bd: dict = {"name": block.name}
        if block.hasViscosity():
            add(block.viscosity, bd, "viscosity")
        add(block.velocity, bd, "velocity")
        add(block.pressure, bd, "pressure")
        if block.hasPassiveScalar():
            add(block.passiveScalar, bd, "scalar")
        if block.hasEpot():
            add(block.epot, bd, "epot")
        if block.hasVelocitySource():
            add(block.velocitySource, bd, "velocitySource")
        if block.hasVertexCoordinates():
            add(block.vertexCoordinates, bd, "vertexCoordinates")
        elif block.hasTransform():
            add(block.transform, bd, "transform")
            if block.hasFaceTransform():
                add(block.faceTransform, bd, "faceTransform")
        bd["boundaries"] = []
        for f in range(2 * domain.getSpatialDims()):
            bound = block.getBoundary(f)
            b: dict = {}
            if isinstance(bound, FixedBoundary):
                b["type"] = "FIXED"
                b["velocityType"] = _bc_name(bound.velocityType)
                add(bound.velocity, b, "velocity")
                if bound.hasPassiveScalar():
                    b["passiveScalarType"] = [_bc_name(t) for t in bound.passiveScalarTypes]
                    add(bound.passiveScalar, b, "scalar")
                if bound.hasPotentialTypes():
                    add(bound.potentialTypes, b, "potentialTypes")
                if bound.hasPotentialTypes() or bound.getPotentialBC() != PotentialBC.INSULATING:
                    b["potentialBC"] = bound.getPotentialBC().name
                if bound.hasPotentialThinWall():
                    b["potentialCw"] = bound.getPotentialCw()
                if bound.hasPotentialValues():
                    add(bound.potentialValues, b, "potentialValues")
                if bound.hasTransform():
                    add(bound.transform, b, "transform")
            elif isinstance(bound, ConnectedBoundary):
                b["type"] = "CONNECTED"
                b["connectedBlock"] = blocks.index(bound.getConnectedBlock())
                b["axes"] = list(bound.axes)
            elif isinstance(bound, PeriodicBoundary):
                b["type"] = "PERIODIC"
            else:
                raise TypeError(f"Block '{block.name}' face {_FACES[f]}: unknown boundary type {bound.type}.")
            bd["boundaries"].append(b)
        dd["blocks"].append(bd)
    save_id = uuid.uuid4().hex
    dd[_SAVE_ID] = save_id
    dd["data_info"] = {str(i): {"shape": list(t.shape), "dtype": _DTYPES[t.dtype], "device": str(t.device.type)}
                       for i, t in enumerate(data)}
    arrays = {str(i): t.detach().cpu().numpy() for i, t in enumerate(data)}
    arrays[_SAVE_ID] = np.array(save_id)
    json_path.parent.mkdir(parents=True, exist_ok=True)
    # the tensors first: the JSON names them, so it is written last
    _atomic_write(npz_path, lambda f: np.savez_compressed(f, **arrays))
    _atomic_write(json_path, lambda f: f.write(json.dumps(dd, indent=1).encode()))


def _read_pair(path: Path) -> tuple:
    json_path, npz_path = _pair_paths(path)
    missing = [str(p) for p in (json_path, npz_path) if not p.is_file()]
    if missing:
        raise FileNotFoundError(f"Saved domain incomplete, missing: {', '.join(missing)}.")
    try:
        dd = json.loads(json_path.read_text())
    except json.JSONDecodeError as e:
        raise ValueError(f"{json_path} is not valid JSON: {e}.") from e
    version = int(dd.get("format_version", 1))
    if version > FORMAT_VERSION:
        raise ValueError(f"{json_path} has format version {version}; this build reads up to {FORMAT_VERSION}.")
    for key in ("spatialDims", "name", "blocks", "data_info", "viscosity"):
        if key not in dd:
            raise ValueError(f"{json_path} is missing the '{key}' entry.")
    info = dd["data_info"]
    try:
        with np.load(npz_path) as npz:
            if _SAVE_ID in dd:
                npz_id = str(npz[_SAVE_ID]) if _SAVE_ID in npz else None
                if npz_id != dd[_SAVE_ID]:
                    raise ValueError(f"{json_path} and {npz_path} belong to different saves (interrupted save?).")
            lacking = [k for k in info if k not in npz]
            if lacking:
                raise ValueError(f"{npz_path} lacks tensors {lacking} listed in {json_path}.")
            arrays = {k: npz[k] for k in info}
    except (OSError, EOFError) as e:
        raise ValueError(f"{npz_path} is not a readable .npz file: {e}.") from e
    for k, i in info.items():
        if "shape" in i and list(arrays[k].shape) != list(i["shape"]):
            raise ValueError(f"Tensor {k} in {npz_path} has shape {list(arrays[k].shape)}, {json_path} expects "
                             f"{i['shape']}.")
    return dd, arrays


def load_domain(path: Path, dtype: Optional[torch.dtype] = None, device=None, with_scalar: bool = True) -> Domain:
    """Rebuild a domain from a saved pair (load_domain, domain_io.py:497-760). ``dtype`` and
    ``device`` default to those of the saved viscosity and block fields; host tensors stay on
    the host, integer tensors keep their dtype. The domain is returned unprepared."""
    dd, arrays = _read_pair(path)
    info = dd["data_info"]
    if dtype is None:
        dtype = getattr(torch, info[dd["viscosity"]].get("dtype", "float32"))
    if device is None:
        bds = dd["blocks"]
        vkey = bds[0].get("velocity") if bds else None
        device = info[vkey].get("device", "cuda") if vkey else "cuda"
    target = torch.device(device)
    tensors = {}
    for k, a in arrays.items():
        t = torch.from_numpy(np.ascontiguousarray(a))
        on_host = info[k].get("device", "cuda") == "cpu"
        tensors[k] = t.to(device="cpu" if on_host else target, dtype=dtype if t.is_floating_point() else None)

    def get(d: dict, name: str) -> Any:
        return tensors[d[name]] if name in d else None

    domain = Domain(dd["spatialDims"], get(dd, "viscosity"), dd["name"], dtype=dtype, device=target,
                    passiveScalarChannels=dd.get("passiveScalarChannels", 1) if with_scalar else 0,
                    scalarViscosity=get(dd, "passiveScalarViscosity") if with_scalar else None)
    nb = 2 * dd["spatialDims"]
    for bd in dd["blocks"]:
        if len(bd.get("boundaries", [])) != nb:
            raise ValueError(f"Block '{bd.get('name')}' lists {len(bd.get('boundaries', []))} boundaries, "
                             f"expected {nb}.")
        coords = get(bd, "vertexCoordinates")
        block = domain.CreateBlock(get(bd, "velocity"), get(bd, "pressure"),
                                   get(bd, "scalar") if with_scalar else None, vertexCoordinates=coords,
                                   name=bd["name"])
        if "viscosity" in bd:
            block.setViscosity(get(bd, "viscosity"))
        if "velocitySource" in bd:
            block.setVelocitySource(get(bd, "velocitySource"))
        if "epot" in bd:
            block.setEpot(get(bd, "epot"))
        if coords is None and "transform" in bd:
            block.setTransform(get(bd, "transform"), get(bd, "faceTransform"))
    blocks = domain.getBlocks()
    for bd, block in zip(dd["blocks"], blocks):
        for f, b in enumerate(bd["boundaries"]):
            kind = b["type"]
            if kind == "FIXED":
                _load_fixed(block, f, b, get, with_scalar)
            elif kind == "CONNECTED":
                block.ConnectBlock(_FACES[f], blocks[b["connectedBlock"]], *[_FACES[x] for x in b["axes"]])
            elif kind == "PERIODIC":
                block.MakePeriodic(f // 2)
            elif kind in ("NEUMANN", "DIRICHLET", "DIRICHLET_VARYING"):
                raise NotImplementedError(f"Block '{block.name}' face {_FACES[f]}: {kind} boundaries cannot be loaded.")
            else:
                raise TypeError(f"Block '{block.name}' face {_FACES[f]}: unknown boundary type {kind!r}.")
    return domain


def _load_fixed(block, f: int, b: dict, get, with_scalar: bool) -> None:
    has_scalar = with_scalar and "scalar" in b
    if b.get("velocityType", "DIRICHLET") == "NEUMANN":
        stype = b.get("passiveScalarType", "NEUMANN")
        dirichlet = any(t == "DIRICHLET" for t in stype) if isinstance(stype, list) else stype == "DIRICHLET"
        bound = block.OpenBoundary(f, get(b, "scalar") if has_scalar and dirichlet else None)
    else:
        bound = block.CloseBoundary(f, get(b, "velocity"), get(b, "scalar") if has_scalar else None)
        if has_scalar:
            st = b["passiveScalarType"]
            bound.setPassiveScalarType([_bc_type(t) for t in st] if isinstance(st, list) else _bc_type(st))
    values, types = get(b, "potentialValues"), get(b, "potentialTypes")
    if values is not None:
        bound.setPotentialValues(values)
    cw = b.get("potentialCw")
    if types is not None:
        if types.is_floating_point():
            types = types.round()
        bound.setPotentialTypes(types.to(torch.int8), cw=cw)
    elif "potentialBC" in b:
        bound.setPotentialBC(PotentialBC[b["potentialBC"]], cw=cw)
    else:
        # files written before the PotentialBC type held independent flags (domain_io.py:202-209)
        if "epotCw" in b:
            bound.setEpotCw(b["epotCw"])
        if "epotInsulating" in b:
            bound.setEpotInsulating(b["epotInsulating"])
        if b.get("epotDirichlet", False):
            bound.setEpotDirichlet(True)

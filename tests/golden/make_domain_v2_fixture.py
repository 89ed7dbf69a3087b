"""Write ``domain_v2_case.json`` + ``domain_v2_case.npz``: a saved domain laid out field by field
as the reference's ``save_domain`` writes it (src/phipict/io/domain_io.py:298-452, format version
2; dtype names domain_io.py:57-82, boundary-condition names domain_io.py:85-106).

The pair is built here from numpy and json alone, without this package's own ``save_domain``, so
that loading it checks the loader against the reference's schema and not against a writer that
shares the loader's assumptions:

* the JSON keys in the order the reference writes them: ``format_version``, ``name``,
  ``spatialDims``, ``viscosity``, ``passiveScalarChannels``, ``passiveScalarViscosity``,
  ``blocks`` (``name``, ``viscosity``, ``velocity``, ``pressure``, ``scalar``, ``epot``,
  ``velocitySource``, ``vertexCoordinates`` | ``transform``, ``boundaries``), ``save_id`` and
  ``data_info`` (``shape``, ``dtype``, ``device`` per tensor);
* every tensor referenced by the string of its index, a tensor shared by several objects stored
  once (the outflow's velocity and scalar share one zero tensor, as the reference's
  ``_tensor_key`` dedups shared storage);
* FIXED faces with ``velocityType``, ``passiveScalarType`` as a list of names, ``scalar``,
  ``potentialTypes`` (int8), ``potentialBC``, ``potentialCw`` and ``potentialValues``; CONNECTED
  faces with ``connectedBlock`` and ``axes``; PERIODIC faces;
* the NPZ written with ``np.savez_compressed`` and keyed ``"0"``, ``"1"``, ... plus ``save_id``.

Run ``python tests/golden/make_domain_v2_fixture.py`` to rewrite the pair; the values are fixed
by a seeded generator and the save id is a constant, so a rerun reproduces the pair.
"""
import json
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
NAME = "domain_v2_case"
SAVE_ID = "5f0c2b8e9a7d4e61b3c2d1e0f9a8b7c6"


def build():
    rng = np.random.default_rng(20261203)
    data = []

    def add(arr, d, key):
        for i, a in enumerate(data):
            if a is arr:
                d[key] = str(i)
                return
        data.append(arr)
        d[key] = str(len(data) - 1)

    f32 = np.float32
    dom = {"format_version": 2, "name": "two_block_channel", "spatialDims": 2}
    add(np.array([0.015], f32), dom, "viscosity")
    dom["passiveScalarChannels"] = 2
    add(np.array([0.1, 0.05], f32), dom, "passiveScalarViscosity")
    dom["blocks"] = []

    # left block: 4 x 3 cells on a stretched vertex grid, block viscosity, source and potential
    xs = np.array([0.0, 0.2, 0.45, 0.75, 1.1], f32)
    ys = np.array([0.0, 0.3, 0.7, 1.0], f32)
    vx, vy = np.meshgrid(xs, ys)
    left = {"name": "inlet_block"}
    add(np.full((1, 1, 3, 4), 0.02, f32), left, "viscosity")
    add(rng.standard_normal((1, 2, 3, 4)).astype(f32), left, "velocity")
    add(rng.standard_normal((1, 1, 3, 4)).astype(f32), left, "pressure")
    add(rng.standard_normal((1, 2, 3, 4)).astype(f32), left, "scalar")
    add(rng.standard_normal((1, 1, 3, 4)).astype(f32), left, "epot")
    add(np.stack([np.full((3, 4), 0.5, f32), np.zeros((3, 4), f32)])[None], left, "velocitySource")
    add(np.stack([vx, vy])[None].astype(f32), left, "vertexCoordinates")

    # right block: 3 x 3 uniform cells, no optional fields
    right = {"name": "outlet_block"}
    add(rng.standard_normal((1, 2, 3, 3)).astype(f32), right, "velocity")
    add(rng.standard_normal((1, 1, 3, 3)).astype(f32), right, "pressure")
    add(rng.standard_normal((1, 2, 3, 3)).astype(f32), right, "scalar")

    inflow = {"type": "FIXED", "velocityType": "DIRICHLET"}
    add(np.array([[1.0, 0.0]], f32), inflow, "velocity")
    inflow["passiveScalarType"] = ["DIRICHLET", "DIRICHLET"]
    add(np.array([[1.0, 0.25]], f32), inflow, "scalar")
    inflow["potentialBC"] = "OPEN"
    to_right = {"type": "CONNECTED", "connectedBlock": 1, "axes": [0, 2]}
    left["boundaries"] = [inflow, to_right, {"type": "PERIODIC"}, {"type": "PERIODIC"}]

    to_left = {"type": "CONNECTED", "connectedBlock": 0, "axes": [1, 2]}
    outflow = {"type": "FIXED", "velocityType": "NEUMANN"}
    zeros = np.zeros((1, 2), f32)
    add(zeros, outflow, "velocity")
    outflow["passiveScalarType"] = ["NEUMANN", "NEUMANN"]
    add(zeros, outflow, "scalar")
    add(np.array([3, 2, 1], np.int8), outflow, "potentialTypes")
    outflow["potentialBC"] = "THIN_WALL"
    outflow["potentialCw"] = 0.25
    add(np.array([0.0, 0.5, 0.0], f32), outflow, "potentialValues")
    right["boundaries"] = [to_left, outflow, {"type": "PERIODIC"}, {"type": "PERIODIC"}]

    dom["blocks"] = [left, right]
    dom["save_id"] = SAVE_ID
    dom["data_info"] = {str(i): {"shape": list(a.shape), "dtype": str(a.dtype), "device": "cpu"}
                        for i, a in enumerate(data)}
    arrays = {str(i): a for i, a in enumerate(data)}
    arrays["save_id"] = np.array(SAVE_ID)
    return dom, arrays


def main():
    dom, arrays = build()
    np.savez_compressed(HERE / f"{NAME}.npz", **arrays)
    (HERE / f"{NAME}.json").write_text(json.dumps(dom, indent=1))


if __name__ == "__main__":
    main()

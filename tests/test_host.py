"""Host-side checks of the product (CPU only, no kernel launches): the C-ABI library loads and
exports every entry point the header declares, the ctypes descriptors match the C layout, the
host Ruge-Stueben setup equals the oracle's, and the host logic mirrors the reference
(connection axes, grid weights, tolerances, hooks, option validation)."""
from __future__ import annotations

import ctypes as C
import math
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest
import torch

import cases
from oracle import oracle as O

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "phipict_hip.h"


def declared_entry_points() -> list:
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pp_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from phipict_amd import _lib

    names = declared_entry_points()
    assert len(names) >= 30
    for name in names:
        assert hasattr(_lib.LIB, name), name
    # every declared entry point has a ctypes signature (pp_last_error is typed separately)
    assert set(names) - {"pp_last_error"} == set(_lib.SIGNATURES)


def test_version_and_error_reporting_without_gpu():
    from phipict_amd import _lib

    assert _lib.LIB.pp_version() > 0
    # argument checks fail before any device call
    rc = _lib.LIB.pp_cast(None, 8, None, 4, 10, None)
    assert rc != 0
    assert b"null" in _lib.LIB.pp_last_error()


def test_descriptor_layout_matches_header(tmp_path):
    from phipict_amd import _lib

    structs = {"pp_boundary_desc": _lib.BoundaryDesc, "pp_block_desc": _lib.BlockDesc, "pp_csr": _lib.CsrDesc,
               "pp_domain": _lib.DomainDesc, "pp_solve_result": _lib.SolveResult, "pp_amg_level": _lib.AmgLevel,
               "pp_sell": _lib.SellDesc}
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void) {"]
    for cname, ct in structs.items():
        lines.append(f'printf("{cname} __sizeof__ %zu\\n", sizeof({cname}));')
        for fname, _ in ct._fields_:
            lines.append(f'printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-o", str(exe), str(src)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {}
    for line in out:
        if line:
            s, f, v = line.split()
            got[(s, f)] = int(v)
    for cname, ct in structs.items():
        assert got[(cname, "__sizeof__")] == C.sizeof(ct), cname
        for fname, _ in ct._fields_:
            assert got[(cname, fname)] == getattr(ct, fname).offset, (cname, fname)


@pytest.mark.parametrize("which", ["pressure2d", "epot3d", "advection2d"])
def test_host_rs_setup_equals_oracle(which):
    """The library's C++ Ruge-Stueben setup makes the same splittings and the same P, R and
    coarse operators as the oracle's restatement."""
    from phipict_amd.solvers.amg import rs_setup_host

    if which == "pressure2d":
        od = O.OracleDomain(cases.channel2d(40, 30), 0.02, 0.01)
        _, _, A = od.advection_matrix()
        col, val = od.pressure_matrix(A)
    elif which == "epot3d":
        od = O.OracleDomain(cases.duct3d(14, 10, 12), 0.01, 0.02)
        col, val = od.epot_matrix()
    else:
        od = O.OracleDomain(cases.inflow2d(30, 20), 0.02, 0.05)
        col, val, _ = od.advection_matrix()
    row = od.row
    mine = rs_setup_host(row, col, val)
    ref = O.OracleHierarchy(row, col, val).levels
    assert len(mine) == len(ref)
    for a, b in zip(mine, ref):
        assert a["n"] == b["n"]
        for key in ("A", "P", "R"):
            if key not in b:
                continue
            for x, y in zip(a[key][:2], b[key][:2]):
                np.testing.assert_array_equal(x, y)
            np.testing.assert_allclose(a[key][2], b[key][2], rtol=1e-13, atol=1e-300)
        if "split" in b:
            np.testing.assert_array_equal(a["split"], b["split"])


def _cpu_domain(dims):
    import phipict_amd as pa

    return pa.Domain(dims, torch.tensor([0.01], dtype=torch.float64), dtype=torch.float64, device="cpu")


@pytest.mark.parametrize("dims", [2, 3])
def test_connection_axes_match_restatement(dims):
    """Block.ConnectBlock stores the same face-format axes as the oracle's restatement of
    ConnectBlocks (boundaries.cpp:857-891) for every face pair and valid axis choice."""
    tested = 0
    for f1 in range(2 * dims):
        for f2 in range(2 * dims):
            d2 = f2 >> 1
            tang = [((d2 + k) % dims) for k in range(1, dims)]
            choices = [(-1, -1)]
            if dims == 2:
                choices += [((tang[0] << 1) | s, -1) for s in (0, 1)]
            else:
                for s1 in (0, 1):
                    for s2 in (0, 1):
                        choices.append(((tang[0] << 1) | s1, (tang[1] << 1) | s2))
                        choices.append(((tang[1] << 1) | s1, (tang[0] << 1) | s2))
            for c1, c2 in choices:
                dom = _cpu_domain(dims)
                n = [4] * dims
                b0 = dom.CreateBlockWithSize(n)
                b1 = dom.CreateBlockWithSize(n)
                b0.ConnectBlock(f1, b1, f2, "" if c1 < 0 else c1, "" if c2 < 0 else c2)
                a1, a2 = O.connection_axes(dims, f1, f2, c1, c2)
                assert b0.getBoundary(f1).axes == a1[:dims], (f1, f2, c1, c2)
                assert b1.getBoundary(f2).axes == a2[:dims], (f1, f2, c1, c2)
                tested += 1
    assert tested > 10


def test_csr_size_formula():
    """nnz = n (2d+1) - sum over prescribed faces of the face area (block.cpp:1145-1155)."""
    for name, make in cases.CASES.items():
        specs = make()
        od = O.OracleDomain(specs, 0.01, 0.01)
        dom = _cpu_domain(specs[0].dims)
        total = 0
        for s in specs:
            b = dom.CreateBlockWithSize(s.size_xyz[: s.dims])
            for f, fs in enumerate(s.faces):
                if fs.kind == "fixed":
                    b.CloseBoundary(f)
            total += b.ComputeCSRSize()
        assert total == od.nnz, name


def test_grid_weights():
    from phipict_amd import grid

    w = np.array(grid.make_weights_tanh(20, 1.5, "BOTH"))
    assert w[0] == 0 and w[-1] == 1 and np.all(np.diff(w) > 0)
    np.testing.assert_allclose(w + w[::-1], 1.0, atol=1e-15)
    d = np.diff(w)
    assert d[0] < d[10] and d[-1] < d[10]
    np.testing.assert_allclose(grid.make_weights_tanh(8, 0.0, "BOTH"), np.linspace(0, 1, 9))
    s = np.array(grid.make_weights_simple_grading(10, 4.0, "START"))
    ds = np.diff(s)
    np.testing.assert_allclose(ds[-1] / ds[0], 4.0, rtol=1e-12)
    np.testing.assert_allclose(s[-1], 1.0)
    v = grid.rectilinear_vertices([np.linspace(0, 1, 4), np.linspace(0, 2, 3)])
    assert list(v.shape) == [1, 2, 3, 4]
    np.testing.assert_allclose(v[0, 1, :, 0].numpy(), [0, 1, 2])


def test_solver_tolerance_resolution():
    from phipict_amd.solvers.tolerance import SolverTolerance, default_tolerance, resolve_tolerance

    b = torch.full((100,), 2.0, dtype=torch.float64)
    t = SolverTolerance(rtol=1e-3, atol=1e-9)
    assert math.isclose(resolve_tolerance(t, b), 2e-3)
    assert math.isclose(SolverTolerance(atol=1e-4).resolve(b), 1e-4)
    # the precision floor: 8 eps of the RHS scale
    assert math.isclose(SolverTolerance(rtol=1e-30).resolve(b), 8 * torch.finfo(torch.float64).eps * 2.0)
    bb = torch.cat([torch.full((50,), 1.0), torch.full((50,), 4.0)]).double()
    assert resolve_tolerance(t, bb, batch_size=2) == pytest.approx([1e-3, 4e-3])
    assert resolve_tolerance(1e-7, b) == 1e-7
    assert default_tolerance(torch.float32) == 1e-5 and default_tolerance(torch.float64) == 1e-8
    with pytest.raises(ValueError):
        SolverTolerance()
    with pytest.raises(ValueError):
        SolverTolerance(rtol=-1.0)


def test_hooks_order_and_validation():
    from phipict_amd.hooks import Hook, Hooks

    seen = []
    h = Hooks({"PRE": lambda **k: seen.append("a")})
    h.append(Hook.PRE, lambda **k: seen.append("b")).prepend("PRE", lambda **k: seen.append("c"))
    h.run(Hook.PRE)
    assert seen == ["c", "a", "b"]
    assert Hook.PRE in h and Hook.POST not in h
    with pytest.raises(ValueError):
        h.append("NOT_A_HOOK", print)


def test_simulation_rejects_unimplemented_options():
    import phipict_amd as pa

    dom = _cpu_domain(2)
    with pytest.raises(NotImplementedError):
        pa.Simulation(domain=dom, dt=0.1, differentiable=True, velocity_corrector="FVM")
    # the reference's default non_orthogonal=True is accepted with differentiable=True: the step
    # refuses skewed grids when it runs
    pa.Simulation(domain=dom, dt=0.1, differentiable=True)
    with pytest.raises(TypeError):
        pa.Simulation(domain=dom, dt=0.1, no_such_option=1)
    with pytest.raises(ValueError):
        pa.Simulation(domain=dom, dt=0.1, velocity_corrector="SPECTRAL")
    pa.Simulation(domain=dom, dt=0.1, non_orthogonal=False)


def test_product_does_not_import_oracle():
    """The package never reaches the oracle: no import of it anywhere under phipict_amd/."""
    for p in (ROOT / "phipict_amd").rglob("*.py"):
        text = p.read_text()
        assert "oracle" not in re.sub(r"#.*", "", text).replace('"""', ""), p


@pytest.mark.parametrize("which", ["pressure2d", "epot3d"])
def test_ap_pattern_is_the_symbolic_product(which):
    """The host pattern of A P (pp_ap_pattern), on every level of the setup, is the symbolic
    product of the fine pattern and the interpolation (scipy's product of 0/1 matrices, so no
    entry cancels), sorted per row; the oracle's Galerkin product R A P fits in the coarse pattern
    (its pattern is the setup's), and a coarse pattern with one entry of R (A P) removed is
    refused."""
    import scipy.sparse as sp

    from phipict_amd.solvers.amg import ap_pattern_host, rs_setup_host

    if which == "pressure2d":
        od = O.OracleDomain(cases.channel2d(40, 30), 0.02, 0.01)
        _, _, A = od.advection_matrix()
        col, val = od.pressure_matrix(A)
    else:
        od = O.OracleDomain(cases.duct3d(14, 10, 12), 0.01, 0.02)
        col, val = od.epot_matrix()
    row = od.row
    levels = rs_setup_host(row, col, val)
    fine = (row, col, val)
    for l in range(len(levels) - 1):
        lv, nxt = levels[l], levels[l + 1]
        frow, fcol = fine[0], fine[1]
        n, nc = frow.size - 1, nxt["n"]
        ap_row, ap_col = ap_pattern_host(frow, fcol, lv["P"], lv["R"], nxt["A"][0], nxt["A"][1])
        ones = lambda c: np.ones(c.size)  # noqa: E731
        Ab = sp.csr_matrix((ones(fcol), fcol, frow), shape=(n, n))
        Pb = sp.csr_matrix((ones(lv["P"][1]), lv["P"][1], lv["P"][0]), shape=(n, nc))
        ref = (Ab @ Pb).tocsr()
        ref.sort_indices()
        np.testing.assert_array_equal(ap_row, ref.indptr)
        np.testing.assert_array_equal(ap_col, ref.indices)
        rap = O.galerkin(fine, lv["P"], lv["R"], nc)
        np.testing.assert_array_equal(rap[0], nxt["A"][0])
        np.testing.assert_array_equal(rap[1], nxt["A"][1])
        fine = rap
    # drop the first off-diagonal entry of coarse row 0 of the first level
    lv, nxt = levels[0], levels[1]
    cr, cc = nxt["A"][0].copy(), nxt["A"][1].copy()
    drop = next(e for e in range(cr[0], cr[1]) if cc[e] != 0)
    cc2 = np.delete(cc, drop)
    cr2 = cr.copy()
    cr2[1:] -= 1
    from phipict_amd._lib import PhipictHipError

    with pytest.raises(PhipictHipError, match="misses an entry"):
        ap_pattern_host(row, col, lv["P"], lv["R"], cr2, cc2)


def test_counter_traffic_table(tmp_path):
    """Counter summaries: kernel names split into base and template arguments, counters scaled
    by the copy kernel's reading, early-exit dispatches kept apart."""
    from phipict_amd import profiling

    assert profiling.split_kernel_name("void pp::k_jacobi_sweep<1, double, float, double>(pp::Csr<float>, int)") == \
        ("k_jacobi_sweep", ("1", "double", "float", "double"))
    assert profiling.split_kernel_name("pp::k_copy_bytes(uint4 const*, uint4*, unsigned long)") == ("k_copy_bytes", ())
    assert profiling.split_kernel_name("void pp::k_x<pp::Csr<float>, 2>(int)") == ("k_x", ("pp::Csr<float>", "2"))

    def write(path, counter, rows):
        with open(path, "w") as f:
            f.write("Dispatch_Id,Kernel_Name,Grid_Size,Counter_Name,Counter_Value\n")
            for did, name, grid, val in rows:
                # two dimensions of one counter: summed
                f.write(f'{did},"{name}",{grid},{counter},{val / 2}\n{did},"{name}",{grid},{counter},{val / 2}\n')

    copy = "pp::k_copy_bytes(uint4 const*, uint4*, unsigned long)"
    kern = "void pp::k_cg_direction<double, 1, 1>(pp::Csr<double>)"
    # the copy of 1 MiB reads 2048 KiB and writes 1536 KiB on the counters: scales 2 and 1.5
    write(tmp_path / "f.csv", "FETCH_SIZE", [(1, copy, 100, 2048), (2, kern, 512, 400), (3, kern, 512, 600),
                                             (4, kern, 512, 1)])
    write(tmp_path / "w.csv", "WRITE_SIZE", [(1, copy, 100, 1536), (2, kern, 512, 150), (3, kern, 512, 150),
                                             (4, kern, 512, 0)])
    tab = profiling.traffic_table(tmp_path / "f.csv", tmp_path / "w.csv", 1 << 20)
    assert tab["fetch_scale"] == pytest.approx(2.0) and tab["write_scale"] == pytest.approx(1.5)
    r = tab["kernels"][("k_cg_direction", ("double", "1", "1"), 512)]
    assert r["dispatches"] == 3 and r["early_exits"] == 1
    assert r["fetch_bytes"] == pytest.approx(500 * 1024 / 2)
    assert r["write_bytes"] == pytest.approx(150 * 1024 / 1.5)


def test_substeps_setter_and_cfl_split():
    """substeps takes a positive integer or "ADAPTIVE" (-1), anything else raises like the
    reference's setter (piso_simulation.py:1596-1619); the CFL split of the product and the
    oracle's agree on the same inputs, including a float32 maximum velocity."""
    import phipict_amd as pa
    from phipict_amd.simulation import _cfl_substep
    from oracle.piso_oracle import cfl_substep

    sim = pa.Simulation(domain=None, dt=0.1)
    assert sim.substeps == 1
    sim.substeps = "adaptive"
    assert sim.substeps == -1
    sim.substeps = 3
    assert sim.substeps == 3
    with pytest.raises(ValueError):
        sim.substeps = 0
    with pytest.raises(TypeError):
        sim.substeps = 2.0
    with pytest.raises(TypeError):
        pa.Simulation(domain=None, substeps="fast")
    for t, v in [(0.25, np.asarray(8.0, np.float32)), (0.25, np.asarray(7.77, np.float64)), (1.0, np.float32(0.0)),
                 (0.1, np.asarray(3.3, np.float32))]:
        assert _cfl_substep(t, v, 0.8) == cfl_substep(t, v, 0.8)


def test_domain_save_load_round_trip(tmp_path):
    """The JSON + NPZ pair (format version 2) restores blocks, fields, connections, periodic
    faces, face velocities, passive scalars with per-channel conditions and potential settings;
    shared tensors are stored once; a pair from different saves, a missing file and a newer
    format are refused."""
    import json

    import phipict_amd as pa
    from phipict_amd.domain_io import load_domain, save_domain

    dom = pa.Domain(2, torch.tensor([0.02], dtype=torch.float64), name="case", dtype=torch.float64, device="cpu",
                    passiveScalarChannels=2, scalarViscosity=torch.tensor([0.1, 0.2], dtype=torch.float64))
    b0 = dom.CreateBlockWithSize([6, 4], name="left")
    b1 = dom.CreateBlockWithSize([5, 4], name="right")
    g = torch.Generator().manual_seed(3)
    for b in (b0, b1):
        b.setVelocity(torch.randn(b.velocity.shape, generator=g, dtype=torch.float64))
        b.setPassiveScalar(torch.randn(1, 2, *b.getSizesZYX(), generator=g, dtype=torch.float64))
    b0.ConnectBlock("+x", b1, "-x")
    inlet = b0.CloseBoundary("-x", velocity=torch.tensor([[1.0, 0.0]], dtype=torch.float64),
                             passiveScalar=torch.tensor([[0.5, 0.25]], dtype=torch.float64))
    inlet.setPotentialBC(pa.PotentialBC.OPEN)
    out = b1.OpenBoundary("+x")
    wall = b1.CloseBoundary("-y")
    wall.setPassiveScalarType([pa.BoundaryConditionType.DIRICHLET, pa.BoundaryConditionType.NEUMANN])
    wall.setPotentialTypes(torch.tensor([3, 3, 1, 0, 2], dtype=torch.int8), cw=0.4)
    wall.setPotentialValues(torch.linspace(0, 1, 5, dtype=torch.float64))
    for b in (b0, b1):
        b.CloseBoundary("+y")
    b0.CloseBoundary("-y")
    save_domain(dom, tmp_path / "case")
    meta = json.loads((tmp_path / "case.json").read_text())
    assert meta["format_version"] == 2 and meta["passiveScalarChannels"] == 2
    back = load_domain(tmp_path / "case.json")
    assert back.name == "case" and back.getDtype() == torch.float64 and back.getPassiveScalarChannels() == 2
    torch.testing.assert_close(back.passiveScalarViscosity, dom.passiveScalarViscosity)
    for a, b in zip(dom.getBlocks(), back.getBlocks()):
        assert a.name == b.name
        for name in ("velocity", "pressure", "passiveScalar"):
            torch.testing.assert_close(getattr(b, name), getattr(a, name), rtol=0, atol=0)
        for f in range(4):
            ba, bb = a.getBoundary(f), b.getBoundary(f)
            assert type(ba) is type(bb), (a.name, f)
            if isinstance(ba, pa.FixedBoundary):
                assert ba.velocityType == bb.velocityType
                torch.testing.assert_close(bb.velocity, ba.velocity, rtol=0, atol=0)
                if ba.hasPassiveScalar():
                    torch.testing.assert_close(bb.passiveScalar, ba.passiveScalar, rtol=0, atol=0)
                assert bb.passiveScalarTypes == ba.passiveScalarTypes
                assert bb.getPotentialBC() == ba.getPotentialBC()
                assert bb.getPotentialCw() == ba.getPotentialCw()
                assert (bb.potentialTypes is None) == (ba.potentialTypes is None)
                if ba.potentialTypes is not None:
                    assert torch.equal(bb.potentialTypes, ba.potentialTypes)
                    torch.testing.assert_close(bb.potentialValues, ba.potentialValues, rtol=0, atol=0)
            elif isinstance(ba, pa.ConnectedBoundary):
                assert list(bb.axes) == list(ba.axes)
                assert dom.getBlocks().index(ba.getConnectedBlock()) == back.getBlocks().index(bb.getConnectedBlock())
    # a pair from two saves is refused
    other = tmp_path / "other"
    save_domain(dom, other)
    (tmp_path / "case.npz").write_bytes((tmp_path / "other.npz").read_bytes())
    with pytest.raises(ValueError, match="different saves"):
        load_domain(tmp_path / "case")
    (tmp_path / "other.npz").unlink()
    with pytest.raises(FileNotFoundError):
        load_domain(other)
    meta["format_version"] = 3
    (tmp_path / "v3.json").write_text(json.dumps(meta))
    (tmp_path / "v3.npz").write_bytes((tmp_path / "case.npz").read_bytes())
    with pytest.raises(ValueError, match="format version"):
        load_domain(tmp_path / "v3")


def test_load_domain_reference_schema_fixture(tmp_path):
    """A saved pair laid out field by field as the reference's save_domain writes it
    (domain_io.py:298-452; written by tests/golden/make_domain_v2_fixture.py from numpy and json
    alone) loads with every field in place: domain and per-channel scalar viscosity, a block with
    viscosity, source and potential, the connection axes, periodic faces,
    an inflow with Dirichlet scalars and an OPEN potential, an outflow whose velocity and scalar
    share one stored tensor, per-cell potential types with cw and Dirichlet values; the legacy
    potential flags of older files (domain_io.py:202-209) map to the same PotentialBC values."""
    import json
    import shutil

    import phipict_amd as pa
    from phipict_amd.domain_io import load_domain

    golden = Path(__file__).resolve().parent / "golden"
    meta = json.loads((golden / "domain_v2_case.json").read_text())
    with np.load(golden / "domain_v2_case.npz") as npz:
        arr = {k: npz[k] for k in npz.files}
    # the metrics of vertex coordinates are computed by the library on the GPU (the GPU test
    # loads the pair as it is); here the first block falls back to the unit grid
    assert meta["blocks"][0]["vertexCoordinates"] in meta["data_info"]
    host_meta = json.loads(json.dumps(meta))
    host_meta["blocks"][0].pop("vertexCoordinates")
    (tmp_path / "host.json").write_text(json.dumps(host_meta))
    shutil.copy(golden / "domain_v2_case.npz", tmp_path / "host.npz")
    dom = load_domain(tmp_path / "host.json")
    assert dom.name == "two_block_channel" and dom.getSpatialDims() == 2
    assert dom.getDtype() == torch.float32 and dom.getPassiveScalarChannels() == 2
    assert dom.getDevice().type == "cpu"
    torch.testing.assert_close(dom.viscosity.cpu().reshape(-1), torch.tensor([0.015]), rtol=0, atol=0)
    torch.testing.assert_close(dom.passiveScalarViscosity.cpu(), torch.tensor([0.1, 0.05]), rtol=0, atol=0)
    left, right = dom.getBlocks()
    assert (left.name, right.name) == ("inlet_block", "outlet_block")
    assert left.getSizesZYX() == [3, 4] and right.getSizesZYX() == [3, 3]

    def stored(d, key):
        return torch.from_numpy(arr[d[key]])

    lb, rb = meta["blocks"]
    for block, bd in ((left, lb), (right, rb)):
        for field, key in (("velocity", "velocity"), ("pressure", "pressure"), ("passiveScalar", "scalar")):
            torch.testing.assert_close(getattr(block, field).cpu(), stored(bd, key), rtol=0, atol=0)
    torch.testing.assert_close(left.viscosity.cpu(), stored(lb, "viscosity"), rtol=0, atol=0)
    torch.testing.assert_close(left.velocitySource.cpu(), stored(lb, "velocitySource"), rtol=0, atol=0)
    torch.testing.assert_close(left.epot.cpu(), stored(lb, "epot"), rtol=0, atol=0)
    assert not right.hasViscosity() and not right.hasEpot() and not right.hasVelocitySource()

    inflow, to_right, ly0, ly1 = (left.getBoundary(f) for f in range(4))
    to_left, outflow, ry0, ry1 = (right.getBoundary(f) for f in range(4))
    assert all(isinstance(b, pa.PeriodicBoundary) for b in (ly0, ly1, ry0, ry1))
    assert isinstance(to_right, pa.ConnectedBoundary) and to_right.getConnectedBlock() is right
    assert list(to_right.axes) == [0, 2] and list(to_left.axes) == [1, 2]
    assert isinstance(inflow, pa.FixedBoundary) and inflow.velocityType == pa.BoundaryConditionType.DIRICHLET
    torch.testing.assert_close(inflow.velocity.cpu().reshape(-1), torch.tensor([1.0, 0.0]), rtol=0, atol=0)
    torch.testing.assert_close(inflow.passiveScalar.cpu().reshape(-1), torch.tensor([1.0, 0.25]), rtol=0, atol=0)
    assert list(inflow.passiveScalarTypes) == [pa.BoundaryConditionType.DIRICHLET] * 2
    assert inflow.getPotentialBC() == pa.PotentialBC.OPEN and not inflow.hasPotentialTypes()
    assert isinstance(outflow, pa.FixedBoundary) and outflow.velocityType == pa.BoundaryConditionType.NEUMANN
    assert list(outflow.passiveScalarTypes) == [pa.BoundaryConditionType.NEUMANN] * 2
    assert meta["blocks"][1]["boundaries"][1]["velocity"] == meta["blocks"][1]["boundaries"][1]["scalar"]
    assert outflow.potentialTypes.tolist() == [3, 2, 1] and outflow.getPotentialCw() == 0.25
    assert outflow.hasPotentialThinWall() and outflow.hasPotentialDirichlet()
    torch.testing.assert_close(outflow.potentialValues.cpu(), torch.tensor([0.0, 0.5, 0.0]), rtol=0, atol=0)

    # the same pair with the pre-PotentialBC flags of older files
    legacy = json.loads(json.dumps(host_meta))
    out_d = legacy["blocks"][1]["boundaries"][1]
    for key in ("potentialTypes", "potentialBC", "potentialCw", "potentialValues"):
        out_d.pop(key)
    out_d["epotCw"] = 0.5
    in_d = legacy["blocks"][0]["boundaries"][0]
    in_d.pop("potentialBC")
    in_d["epotInsulating"] = False
    (tmp_path / "legacy.json").write_text(json.dumps(legacy))
    shutil.copy(golden / "domain_v2_case.npz", tmp_path / "legacy.npz")
    old = load_domain(tmp_path / "legacy", dtype=torch.float64)
    assert old.getDtype() == torch.float64
    assert old.getBlocks()[1].getBoundary(1).getPotentialBC() == pa.PotentialBC.THIN_WALL
    assert old.getBlocks()[1].getBoundary(1).getEpotCw() == 0.5
    assert old.getBlocks()[0].getBoundary(0).getPotentialBC() == pa.PotentialBC.OPEN
    out_d.pop("epotCw")
    out_d["epotDirichlet"] = True
    (tmp_path / "legacy.json").write_text(json.dumps(legacy))
    assert load_domain(tmp_path / "legacy").getBlocks()[1].getBoundary(1).getPotentialBC() == pa.PotentialBC.DIRICHLET


def test_legacy_potential_flags_map_to_potential_bc():
    """setEpotCw / setEpotInsulating / setEpotDirichlet set and undo THIN_WALL, OPEN and
    DIRICHLET (boundaries.cpp:968-991; the reference's test_potential_bc.py:360-370)."""
    import phipict_amd as pa

    dom = pa.Domain(2, torch.tensor([0.02]), name="flags", device="cpu")
    bound = dom.CreateBlockWithSize([4, 3]).CloseBoundary("-y")
    bound.setEpotCw(0.3)
    assert bound.getPotentialBC() == pa.PotentialBC.THIN_WALL and bound.getEpotCw() == 0.3 and bound.hasEpotCw()
    bound.setEpotCw(0.0)
    assert bound.getPotentialBC() == pa.PotentialBC.INSULATING and not bound.hasEpotCw()
    bound.setEpotInsulating(False)
    assert bound.getPotentialBC() == pa.PotentialBC.OPEN and not bound.isEpotInsulating()
    bound.setEpotDirichlet(True)
    assert bound.getPotentialBC() == pa.PotentialBC.DIRICHLET and bound.hasEpotDirichlet()
    bound.setEpotDirichlet(False)
    assert bound.getPotentialBC() == pa.PotentialBC.INSULATING
    with pytest.raises(ValueError):
        bound.setEpotCw(-1.0)


def test_krylov_settings_keywords():
    """SetKrylovSettings / GetKrylovSettings with the reference's keywords and preconditioner
    names (bindings.cpp:739-763); host-only calls of the library."""
    import phipict_amd as pa

    pa.SetKrylovSettings(enabled=True, checkInterval=4, useGraphs=False, cgPreconditioner="jacobi",
                         bicgPreconditioner="none")
    got = pa.GetKrylovSettings()
    assert got == {"enabled": True, "checkInterval": 4, "useGraphs": False, "cgPreconditioner": "jacobi",
                   "bicgPreconditioner": "none"}
    with pytest.raises(ValueError):
        pa.SetKrylovSettings(cgPreconditioner="ilu")
    pa.SetKrylovSettings()
    assert pa.GetKrylovSettings()["cgPreconditioner"] == "none"


def test_amg_pcg_solve_reference_keywords():
    """AMGPCGSolve takes the reference's keyword names (bindings.cpp:764-776), including
    returnBestResult."""
    import inspect

    import phipict_amd as pa

    names = list(inspect.signature(pa.AMGPCGSolve).parameters)
    assert names == ["aRow", "aCol", "aVal", "levels", "coarsePinv", "rhs", "x", "maxIterations", "tolerance", "normalized",
                     "returnBestResult", "projectConstant", "preSweeps", "postSweeps", "omega", "tolerances"]

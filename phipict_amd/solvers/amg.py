"""Algebraic multigrid for the pressure and electric-potential Poisson systems.

Mirrors src/phipict/solvers/amg.py: a classical Ruge-Stueben setup on the host
(``build_amg_hierarchy``, amg.py:450-598; here the library's own C++ setup instead of
pyamg), an interpolation cache per sparsity pattern (``hierarchy_for``, :952-1045), Galerkin
refreshes of the coarse operators on the device (``galerkin_refresh``, :749-847) and the
AMG-preconditioned CG (``amg_pcg_solve``, :1544-1627) with the V-cycle on the device in the
hierarchy's native dtype (float32 by default, :956).
"""
from __future__ import annotations

import ctypes as C
import os
import time
from dataclasses import dataclass, field
from typing import Any, Optional, Sequence

import numpy as np
import torch

from .. import _lib
from ..domain import CSRmatrix
from .. import ops

JACOBI_OMEGA = 2.0 / 3.0
MIN_ABS_DIAG = 1e-30


def _host_csr(csr: CSRmatrix):
    row = csr.row.detach().cpu().numpy().astype(np.int32)
    col = csr.index.detach().cpu().numpy().astype(np.int32)
    val = csr.value.detach().reshape(-1)[: csr.getNnz()].to(torch.float64).cpu().numpy()
    return row, col, val


@dataclass
class Level:
    """One level of a hierarchy. ``A`` holds the operator in the hierarchy dtype (for the
    finest level: the system matrix itself); R and P are kept in double for refreshes and in
    the native dtype for the V-cycle. The coarsest level has no R/P."""

    n: int
    A: CSRmatrix
    R_row: Optional[torch.Tensor] = None
    R_col: Optional[torch.Tensor] = None
    R_val: Optional[torch.Tensor] = None  # float64
    P_row: Optional[torch.Tensor] = None
    P_col: Optional[torch.Tensor] = None
    P_val: Optional[torch.Tensor] = None  # float64
    # pattern of this level's A P (row offsets, columns) for the Galerkin refresh of the next
    # level, with the longest rows of A, P, A P, R and the next level's operator (int32 [5], host)
    ap: Optional[tuple] = None
    native: dict = field(default_factory=dict)
    # per sparsity pattern, shared by the refreshes of a hierarchy: sliced ELL layouts of the
    # operator ("A") and of R ("R") (SellLayout), and R / P values per V-cycle dtype ("RP", dtype)
    shared: dict = field(default_factory=dict)
    native_extra: dict = field(default_factory=dict)


#: lanes per row the autotuner tries for a sliced copy (a wavefront reads 64 / lanes rows)
SELL_LANES = (1, 2, 4, 8, 16)
#: levels of at most this many rows run in the library's one-workgroup coarse tail, on CSR
TAIL_ROWS = 128
#: time the candidate lanes per row on the device (PHIPICT_AMD_AUTOTUNE=0: row-length rule)
AUTOTUNE = os.environ.get("PHIPICT_AMD_AUTOTUNE", "1") != "0"
#: fixed lanes per row for every sliced copy (tests), or None
FORCE_LANES: Optional[int] = None
#: autotuner decisions: (rows, entries, {lanes: microseconds}, chosen lanes)
TUNING_LOG: list = []
#: probe timings within this fraction of the fastest count as ties
TUNE_MARGIN = 0.1
# The refreshes invert the coarsest operator on the device (pp_coarse_inverse) when it has at
# most this many rows and the device inverse of the first operator of the pattern matches the
# host pseudo-inverse to DEVICE_PINV_RTOL of its largest entry; PHIPICT_AMD_DEVICE_PINV=0 keeps
# every refresh on the host.
DEVICE_PINV_MAX = 128
DEVICE_PINV_RTOL = 1e-9
DEVICE_PINV = os.environ.get("PHIPICT_AMD_DEVICE_PINV", "1") != "0"


class SellLayout:
    """Wave-interleaved sliced ELL layout of a CSR pattern (the pp_sell layout): slices of
    64 / lanes rows, each padded to its longest row rounded up to a multiple of ``lanes``;
    entry k * lanes + j of row r of slice s at slice_off[s] + 64 k + r * lanes + j. Holds the
    device slice offsets, steps and columns, plus value copies per dtype for matrices whose
    values never change (R, P)."""

    def __init__(self, row_host: np.ndarray, row_dev: torch.Tensor, col_dev: torch.Tensor, lanes: int = 1):
        if lanes not in (1, 2, 4, 8, 16, 32):
            raise ValueError(f"lanes must be a power of two <= 32, got {lanes}")
        n = row_host.size - 1
        per = 64 // lanes
        lengths = np.diff(row_host.astype(np.int64))
        slices = (n + per - 1) // per
        padded = np.zeros(slices * per, np.int64)
        padded[:n] = lengths
        steps = -(-padded.reshape(slices, per).max(axis=1) // lanes)
        off = np.zeros(slices, np.int64)
        off[1:] = np.cumsum(64 * steps[:-1])
        self.total = max(int(64 * steps.sum()), 64)
        if self.total >= 2 ** 31:
            raise ValueError("sliced ELL copy too large for 32-bit offsets")
        dev = row_dev.device
        self.n, self.nnz, self.lanes = n, int(lengths.sum()), lanes
        self.slice_off = _dev(off.astype(np.int32), dev)
        self.slice_steps = _dev(steps.astype(np.int32), dev)
        self.col = torch.zeros(self.total, dtype=torch.int32, device=dev)
        self.row, self.src_col = row_dev, col_dev
        _lib.call("pp_sell_fill", _lib.ptr(row_dev), _lib.ptr(col_dev), None, n, self.nnz, 1, lanes,
                  _lib.ptr(self.slice_off), _lib.ptr(self.slice_steps), self.total, _lib.PP_F32, _lib.ptr(self.col), None,
                  _lib.PP_F32, _lib.stream_ptr(dev))
        self.values: dict = {}

    def fill(self, val: torch.Tensor, dtype, nmat: int = 1) -> torch.Tensor:
        """Values of ``nmat`` operators of this pattern (``val``: nmat x nnz) in the layout."""
        out = torch.empty(nmat * self.total, dtype=dtype, device=val.device)
        _lib.call("pp_sell_fill", _lib.ptr(self.row), _lib.ptr(self.src_col), _lib.ptr(val), self.n, self.nnz, nmat,
                  self.lanes, _lib.ptr(self.slice_off), _lib.ptr(self.slice_steps), self.total,
                  _lib.dtype_code(val.dtype), None, _lib.ptr(out), _lib.dtype_code(dtype), _lib.stream_ptr(val.device))
        return out

    def cached_values(self, val: torch.Tensor, dtype) -> torch.Tensor:
        if dtype not in self.values:
            self.values[dtype] = self.fill(val, dtype)
        return self.values[dtype]

    def desc(self, val: torch.Tensor) -> tuple:
        return (self.slice_off, self.slice_steps, self.col, val, self.lanes, self.total)


def _rule_lanes(avg: float) -> int:
    """Lanes per row without timing: about eight entries per lane."""
    lanes = 1
    while lanes < 16 and avg > 8.0 * lanes:
        lanes *= 2
    return lanes


def tuned_layout(row_host: np.ndarray, row_dev: torch.Tensor, col_dev: torch.Tensor, val_dev: torch.Tensor,
                 n_cols: int, dtype) -> SellLayout:
    """The sliced layout of a level matrix with the lanes per row that runs its product fastest.

    The autotuner builds the layout for every candidate in SELL_LANES up to twice the mean row
    length, times y = M x through it on the device (``pp_sell_probe``, 20 launches after a
    warm-up) and keeps the fastest; decisions are appended to TUNING_LOG. Without a GPU, with
    AUTOTUNE off or FORCE_LANES set, one layout is built by rule."""
    n = row_host.size - 1
    avg = float(row_host[-1] - row_host[0]) / max(1, n)
    if FORCE_LANES is not None:
        return SellLayout(row_host, row_dev, col_dev, FORCE_LANES)
    if not AUTOTUNE or not val_dev.is_cuda:
        return SellLayout(row_host, row_dev, col_dev, _rule_lanes(avg))
    timings, layouts = {}, {}
    for lanes in SELL_LANES:
        if lanes > 1 and lanes > 2 * avg:
            break
        lay = SellLayout(row_host, row_dev, col_dev, lanes)
        vals = lay.fill(val_dev, dtype)
        desc = _lib.SellDesc(_lib.ptr(lay.slice_off), _lib.ptr(lay.slice_steps), _lib.ptr(lay.col), _lib.ptr(vals), lanes,
                             0, lay.total)
        us = C.c_double()
        _lib.call("pp_sell_probe", C.byref(desc), n, int(n_cols), _lib.dtype_code(dtype), 20,
                  _lib.stream_ptr(val_dev.device), C.byref(us))
        timings[lanes], layouts[lanes] = us.value, lay
    # Timings within TUNE_MARGIN of the fastest do not tell the candidates apart (small matrices,
    # whose probe is launch-bound, or a profiler's per-dispatch cost): among those, the lanes
    # closest to the rule's, whose lanes walk about eight entries each
    fastest = min(timings.values())
    close = [l for l, t in timings.items() if t <= fastest * (1.0 + TUNE_MARGIN)]
    rule = _rule_lanes(avg)
    chosen = min(close, key=lambda l: (abs(np.log2(l) - np.log2(rule)), -l))
    TUNING_LOG.append((n, int(row_host[-1]), timings, chosen))
    return layouts[chosen]


class AMGHierarchy:
    """A V-cycle hierarchy: levels, the dense pseudo-inverse of the coarsest operator, the
    smoothing sweeps and whether the constant mode is projected out (amg.py:260-448)."""

    def __init__(self, levels: list, coarse_A: torch.Tensor, project_constant: bool = False, num_pre_smooth: int = 1,
                 num_post_smooth: int = 1, native_dtype=torch.float32, setup_seconds: float = 0.0):
        self.levels = levels
        self.project_constant = bool(project_constant)
        self.num_pre_smooth = num_pre_smooth
        self.num_post_smooth = num_post_smooth
        self.native_dtype = native_dtype
        self.setup_seconds = setup_seconds
        self.set_coarse(coarse_A)
        self.device_pinv = self._device_pinv_matches(coarse_A)

    def set_coarse(self, coarse_A: torch.Tensor, device: bool = False) -> None:
        """Dense pseudo-inverse of the coarsest operator, ``[nc, nc]`` or ``[B, nc, nc]`` for
        batched operators, in float64: numpy's on the host (amg.py:583-587, 1354-1362), or with
        ``device`` the device inverse, which the host does not wait for."""
        if device:
            self.coarse_pinv64 = device_coarse_inverse(coarse_A, self.project_constant)
        else:
            dense = coarse_A.detach().to(torch.float64).cpu().numpy()
            self.coarse_pinv64 = torch.from_numpy(np.linalg.pinv(dense)).to(coarse_A.device)
        self._pinv_native = {}

    def _device_pinv_matches(self, coarse_A: torch.Tensor) -> bool:
        """Whether refreshes of this pattern may use the device inverse: the coarsest operator
        is small enough and, for this first operator, the device inverse (plain, or with the
        constant null space removed when the constant mode is projected out) equals numpy's
        pseudo-inverse to DEVICE_PINV_RTOL. A refresh has the same pattern and the same assembly,
        so it keeps the symmetry and null space this check relies on."""
        n = coarse_A.shape[-1]
        if not DEVICE_PINV or coarse_A.device.type != "cuda" or n > DEVICE_PINV_MAX:
            return False
        dev = device_coarse_inverse(coarse_A, self.project_constant)
        scale = float(self.coarse_pinv64.abs().max())
        err = float((dev - self.coarse_pinv64).abs().max())
        return scale > 0.0 and np.isfinite(err) and err <= DEVICE_PINV_RTOL * scale

    @property
    def dtype(self):
        return self.levels[0].A.getDtype()

    @property
    def batch_size(self) -> int:
        """Operators per level: 1, or one per batched environment (shared interpolation)."""
        return self.levels[0].A.getBatchSize()

    def coarse_pinv(self, dtype) -> torch.Tensor:
        if dtype not in self._pinv_native:
            self._pinv_native[dtype] = self.coarse_pinv64.to(dtype).contiguous()
        return self._pinv_native[dtype]

    def native_levels(self, dtype) -> list:
        """Per level above the coarsest: (A row, col, val, dinv, R row, col, val, P row, col, val)
        in ``dtype`` (the layout _C.AMGPCGSolve takes, bindings.cpp:764-776)."""
        out = []
        for lvl in self.levels[:-1]:
            nat = lvl.native.get(dtype)
            if nat is None:
                A = lvl.A
                if lvl is self.levels[0] and A.ell_col is not None and A.ell_width > 0:
                    aval = None  # the V-cycle reads the finest operator through its ELL copy only
                elif lvl is not self.levels[0] and "A" in lvl.shared:
                    aval = None  # ... and the other levels outside the tail through their sliced copies
                else:
                    aval = A.value if A.value.dtype == dtype else A.value.to(dtype)
                dinv = torch.empty(A.getRows() * A.getBatchSize(), dtype=dtype, device=A.value.device)
                d = A.desc()
                _lib.call("pp_inverse_diagonal", C.byref(d), _lib.dtype_code(A.getDtype()), _lib.ptr(dinv),
                          _lib.dtype_code(dtype), _lib.stream_ptr(A.value.device))
                # R and P never change: their copies in the V-cycle dtype are shared by refreshes
                rp = lvl.shared.get(("RP", dtype))
                if rp is None:
                    rp = lvl.shared[("RP", dtype)] = (lvl.R_val.to(dtype), lvl.P_val.to(dtype))
                nat = (A.row, A.index, aval, dinv, lvl.R_row, lvl.R_col, rp[0], lvl.P_row, lvl.P_col, rp[1])
                lvl.native[dtype] = nat
            out.append(nat)
        return out

    def native_extras(self, dtype) -> list:
        """Per level above the coarsest: coalesced copies for the V-cycle kernels in ``dtype``
        -- "ell": (columns, values, width) of the finest operator, "a_sell" / "r_sell" /
        "p_sell": wave-interleaved sliced copies (slice offsets, steps, columns, values, lanes,
        entries per operator) of the operators, restrictions and interpolations outside the
        coarse tail."""
        out = []
        for i, lvl in enumerate(self.levels[:-1]):
            ex = lvl.native_extra.get(dtype)
            if ex is None:
                ex = {}
                A = lvl.A
                dev = A.value.device
                if i == 0 and A.ell_col is not None and A.ell_width > 0:
                    ev = torch.empty(A.ell_width * A.getRows() * A.getBatchSize(), dtype=dtype, device=dev)
                    _lib.call("pp_ell_from_csr", _lib.ptr(A.row), _lib.ptr(A.index), _lib.ptr(A.value), A.getRows(),
                              A.getNnz(), A.getBatchSize(), A.ell_width, _lib.dtype_code(A.getDtype()), None,
                              _lib.ptr(ev), _lib.dtype_code(dtype), _lib.stream_ptr(dev))
                    ex["ell"] = (A.ell_col, ev, A.ell_width)
                layout = lvl.shared.get("A")
                if i > 0 and layout is not None:
                    nmat = A.getBatchSize()
                    ex["a_sell"] = layout.desc(layout.fill(A.value.reshape(-1)[: nmat * A.getNnz()], dtype, nmat))
                layout = lvl.shared.get("R")
                if layout is not None:
                    ex["r_sell"] = layout.desc(layout.cached_values(lvl.R_val, dtype))
                layout = lvl.shared.get("P")
                if layout is not None:
                    ex["p_sell"] = layout.desc(layout.cached_values(lvl.P_val, dtype))
                lvl.native_extra[dtype] = ex
            out.append(ex)
        return out

    def packed_levels(self, dtype) -> tuple:
        """Level descriptors in ``dtype`` packed for the library (cached per dtype)."""
        cache = self.__dict__.setdefault("_packed", {})
        if dtype not in cache:
            parent = self.__dict__.get("_packed_parent")
            pcache = getattr(parent, "_packed", {}) if parent is not None else {}
            if dtype in pcache and len(parent.levels) == len(self.levels):
                # a finest-level refresh: copy the parent's descriptors, point level 0 at the
                # new operator values and inverse diagonal
                arr, keep = pcache[dtype]
                arr = type(arr).from_buffer_copy(arr)
                nat = self.native_levels(dtype)[0]
                arr[0].a_val = _lib.ptr(nat[2])  # null when the finest level runs on its ELL copy
                arr[0].dinv = nat[3].data_ptr()
                ex = self.native_extras(dtype)[0]
                if "ell" in ex:
                    arr[0].a_ell_col, arr[0].a_ell_val = ex["ell"][0].data_ptr(), ex["ell"][1].data_ptr()
                    arr[0].a_ell_width = ex["ell"][2]
                else:
                    arr[0].a_ell_width, arr[0].a_ell_col, arr[0].a_ell_val = 0, None, None
                # keep the tensors of levels >= 1 and of level 0's R copy (shared with the
                # parent) and the new level 0
                base = keep[0] if isinstance(keep, tuple) else keep
                new = [t for t in nat if isinstance(t, torch.Tensor)]
                new += [t for v in ex.values() for t in v if isinstance(t, torch.Tensor)]
                cache[dtype] = (arr, (base, new))
            else:
                cache[dtype] = ops.pack_amg_levels(self.native_levels(dtype), dtype, self.native_extras(dtype))
            self.__dict__.pop("_packed_parent", None)
        return cache[dtype]

    def operator_complexity(self) -> float:
        fine = self.levels[0].A.getNnz()
        return sum(l.A.getNnz() for l in self.levels) / max(1, fine)

    def describe(self) -> str:
        rows = [f"level {i}: n={l.n}, nnz={l.A.getNnz()}" for i, l in enumerate(self.levels)]
        return "\n".join(rows + [f"operator complexity {self.operator_complexity():.3f}"])


class _SetupHandle:
    def __init__(self, n, row, col, val, theta, max_levels, max_coarse):
        h = C.c_void_p()
        _lib.call("pp_amg_setup", int(n), row.ctypes.data, col.ctypes.data, val.ctypes.data, float(theta),
                  int(max_levels), int(max_coarse), C.byref(h))
        self.h = h

    def levels(self):
        L = _lib.LIB.pp_amg_num_levels(self.h)
        out = []
        for l in range(L):
            n, na, npp, nr = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32()
            _lib.call("pp_amg_level_sizes", self.h, l, C.byref(n), C.byref(na), C.byref(npp), C.byref(nr))
            ar = np.empty(n.value + 1, np.int32); ac = np.empty(max(na.value, 1), np.int32); av = np.empty(max(na.value, 1))
            has_t = l < L - 1
            if has_t:
                pr = np.empty(n.value + 1, np.int32); pc = np.empty(max(npp.value, 1), np.int32); pv = np.empty(max(npp.value, 1))
            args = [ar.ctypes.data, ac.ctypes.data, av.ctypes.data]
            if has_t:
                # R has n_coarse rows: size known from the next level
                nn = C.c_int32()
                _lib.call("pp_amg_level_sizes", self.h, l + 1, C.byref(nn), None, None, None)
                rr = np.empty(nn.value + 1, np.int32); rc = np.empty(max(nr.value, 1), np.int32); rv = np.empty(max(nr.value, 1))
                split = np.empty(n.value, np.int32)
                args += [pr.ctypes.data, pc.ctypes.data, pv.ctypes.data, rr.ctypes.data, rc.ctypes.data, rv.ctypes.data,
                         split.ctypes.data]
            else:
                args += [None] * 7
            _lib.call("pp_amg_level_copy", self.h, l, *args)
            lv = {"n": n.value, "A": (ar, ac[: na.value], av[: na.value])}
            if has_t:
                lv["P"] = (pr, pc[: npp.value], pv[: npp.value])
                lv["R"] = (rr, rc[: nr.value], rv[: nr.value])
                lv["split"] = split
            out.append(lv)
        return out

    def __del__(self):
        if getattr(self, "h", None):
            _lib.LIB.pp_amg_free(self.h)
            self.h = None


def rs_setup_host(row: np.ndarray, col: np.ndarray, val: np.ndarray, theta: float = 0.25, max_levels: int = 20,
                  max_coarse: int = 64) -> list:
    """Run the library's Ruge-Stueben setup on a host CSR matrix (float64); returns per level
    dicts with numpy arrays (A, P, R, C/F splitting)."""
    n = row.size - 1
    return _SetupHandle(n, np.ascontiguousarray(row, np.int32), np.ascontiguousarray(col, np.int32),
                        np.ascontiguousarray(val, np.float64), theta, max_levels, max_coarse).levels()


def ap_pattern_host(fine_row, fine_col, P, R, coarse_row, coarse_col) -> tuple:
    """Host pattern of A P for one level, (row offsets, sorted columns); checks that the coarse
    pattern holds every entry of R (A P)."""
    pr, pc, rr, rc = (np.ascontiguousarray(a, np.int32) for a in (P[0], P[1], R[0], R[1]))
    fr, fc = np.ascontiguousarray(fine_row, np.int32), np.ascontiguousarray(fine_col, np.int32)
    cr, cc = np.ascontiguousarray(coarse_row, np.int32), np.ascontiguousarray(coarse_col, np.int32)
    n, nc = fr.size - 1, cr.size - 1
    ap_row = np.empty(n + 1, np.int32)
    _lib.call("pp_ap_pattern", n, fr.ctypes.data, fc.ctypes.data, nc, pr.ctypes.data, pc.ctypes.data, None, None,
              None, None, ap_row.ctypes.data, None)
    ap_col = np.empty(max(int(ap_row[-1]), 1), np.int32)
    _lib.call("pp_ap_pattern", n, fr.ctypes.data, fc.ctypes.data, nc, pr.ctypes.data, pc.ctypes.data, rr.ctypes.data,
              rc.ctypes.data, cr.ctypes.data, cc.ctypes.data, ap_row.ctypes.data, ap_col.ctypes.data)
    return ap_row, ap_col[: int(ap_row[-1])]


#: kernels of the two row products of a refresh (PP_RAP_* of the library) and their fit:
#: (name, largest output row, largest Y row) -- None means no limit
RAP_KERNELS = {1: ("reg16", 16, None), 2: ("reg32", 32, None), 3: ("rows8", None, 32), 4: ("rows16", None, 64),
               5: ("rows32", None, 128), 6: ("rows64", None, 256), 7: ("long", None, None)}
#: refresh kernel decisions: (level rows, product, {kernel: microseconds}, chosen)
RAP_TUNING_LOG: list = []


def _rap_fits(kind: int, max_y: int, max_out: int) -> bool:
    _, out_cap, y_cap = RAP_KERNELS[kind]
    return (out_cap is None or max_out <= out_cap) and (y_cap is None or max_y <= y_cap)


def galerkin_rows(prev_A: CSRmatrix, level: "Level", coarse_A: CSRmatrix, c_val: torch.Tensor, ap_val: torch.Tensor,
                  batch: int, plan=None, stages: int = 3) -> None:
    """C = R (A P) of ``level`` into ``c_val`` (the pattern of ``coarse_A``) for ``batch`` operators
    of ``prev_A``'s pattern: the library's two row products, with the level's kernel plan."""
    ap_row, ap_col, longest = level.ap[:3]
    plan_arr = np.asarray(plan if plan is not None else level.ap[3], np.int32)
    d = prev_A.desc()
    d.num_matrices, d.ell_width, d.ell_col, d.ell_val = batch, 0, None, None
    _lib.call("pp_galerkin_rows", C.byref(d), _lib.dtype_code(prev_A.getDtype()), _lib.ptr(level.P_row),
              _lib.ptr(level.P_col), _lib.ptr(level.P_val), _lib.ptr(ap_row), _lib.ptr(ap_col), ap_col.numel(),
              _lib.ptr(ap_val), _lib.ptr(level.R_row), _lib.ptr(level.R_col), _lib.ptr(level.R_val),
              _lib.ptr(coarse_A.row), _lib.ptr(coarse_A.index), _lib.ptr(c_val), coarse_A.getRows(), coarse_A.getNnz(),
              _lib.dtype_code(c_val.dtype), longest.ctypes.data, plan_arr.ctypes.data, int(stages),
              _lib.stream_ptr(c_val.device))


def tune_rap_plan(prev_A: CSRmatrix, level: "Level", coarse_A: CSRmatrix) -> np.ndarray:
    """Kernel of each row product of this level's refresh: every kernel that fits the rows is
    timed on the device (one warm-up, then two launches) on the level's own operator, the fastest
    kept; decisions go to RAP_TUNING_LOG. Without a GPU or with AUTOTUNE off: the library's rule."""
    plan = np.zeros(2, np.int32)
    if not AUTOTUNE or not prev_A.value.is_cuda:
        return plan
    _, ap_col, longest = level.ap[:3]
    ap_val = torch.empty(max(ap_col.numel(), 1), dtype=torch.float64, device=prev_A.value.device)
    c_val = torch.empty(max(coarse_A.getNnz(), 1), dtype=prev_A.getDtype(), device=prev_A.value.device)
    for stage, (max_y, max_out) in enumerate(((int(longest[1]), int(longest[2])), (int(longest[2]), int(longest[4])))):
        cands = [k for k in RAP_KERNELS if _rap_fits(k, max_y, max_out)]
        if len(cands) > 1 and 7 in cands:
            cands.remove(7)  # the one-lane fallback only where nothing else fits
        timings = {}
        for k in cands:
            trial = plan.copy()
            trial[stage] = k
            galerkin_rows(prev_A, level, coarse_A, c_val, ap_val, 1, trial, 1 << stage)
            start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            start.record()
            for _ in range(2):
                galerkin_rows(prev_A, level, coarse_A, c_val, ap_val, 1, trial, 1 << stage)
            stop.record()
            stop.synchronize()
            timings[k] = 1e3 * start.elapsed_time(stop) / 2
        best = min(timings, key=timings.get)
        plan[stage] = best
        if stage == 0:
            # the second product reads A P: leave the fastest kernel's values in the workspace
            galerkin_rows(prev_A, level, coarse_A, c_val, ap_val, 1, plan, 1)
        RAP_TUNING_LOG.append((level.n, ("A P", "R (A P)")[stage],
                               {RAP_KERNELS[k][0]: round(t, 1) for k, t in timings.items()}, RAP_KERNELS[best][0]))
    return plan


def _dev(a: np.ndarray, device, dtype=None) -> torch.Tensor:
    t = torch.from_numpy(np.ascontiguousarray(a))
    return t.to(device=device, dtype=dtype) if dtype is not None else t.to(device)


def build_amg_hierarchy(csr_mat: CSRmatrix, *, method: str = "ruge_stuben", max_levels: int = 20, max_coarse: int = 64,
                        strength: float = 0.25, num_pre_smooth: int = 1, num_post_smooth: int = 1,
                        project_constant: bool = False, native_dtype=torch.float32) -> AMGHierarchy:
    """Classical Ruge-Stueben hierarchy for ``csr_mat`` (first matrix of a batch) built on the
    host and uploaded (amg.py:450-598)."""
    if method != "ruge_stuben":
        raise NotImplementedError("only the Ruge-Stueben coarsening is implemented")
    t0 = time.perf_counter()
    device, dtype = csr_mat.value.device, csr_mat.value.dtype
    row, col, val = _host_csr(csr_mat)
    host = rs_setup_host(row, col, val, strength, max_levels, max_coarse)
    levels = []
    for i, lv in enumerate(host):
        if i == 0:
            A = CSRmatrix(csr_mat.value.reshape(-1)[: csr_mat.getNnz()], csr_mat.index, csr_mat.row)
            A.share_ell_from(csr_mat)
        else:
            ar, ac, av = lv["A"]
            A = CSRmatrix(_dev(av, device, dtype), _dev(ac, device), _dev(ar, device))
        level = Level(n=lv["n"], A=A)
        if "P" in lv:
            pr, pc, pv = lv["P"]
            rr, rc, rv = lv["R"]
            level.P_row, level.P_col, level.P_val = _dev(pr, device), _dev(pc, device), _dev(pv, device, torch.float64)
            level.R_row, level.R_col, level.R_val = _dev(rr, device), _dev(rc, device), _dev(rv, device, torch.float64)
            fine_row, fine_col = (row, col) if i == 0 else lv["A"][:2]
            ap_row, ap_col = ap_pattern_host(fine_row, fine_col, lv["P"], lv["R"], host[i + 1]["A"][0],
                                             host[i + 1]["A"][1])
            longest = np.array([np.diff(r).max(initial=0) for r in (fine_row, pr, ap_row, rr, host[i + 1]["A"][0])],
                               np.int32)
            level.ap = (_dev(ap_row, device), _dev(ap_col, device), longest, np.zeros(2, np.int32))
            if lv["n"] > TAIL_ROWS:
                # transfers out of and into a level outside the coarse tail run through
                # sliced copies with autotuned lanes per row
                nc = host[i + 1]["n"]
                level.shared["R"] = tuned_layout(rr, level.R_row, level.R_col, level.R_val, lv["n"], native_dtype)
                level.shared["P"] = tuned_layout(pr, level.P_row, level.P_col, level.P_val, nc, native_dtype)
        if i > 0 and lv["n"] > TAIL_ROWS and "P" in lv:
            level.shared["A"] = tuned_layout(lv["A"][0], A.row, A.index, A.value, lv["n"], native_dtype)
        levels.append(level)
    for lvl, nxt in zip(levels[:-1], levels[1:]):
        # the refresh kernels of every level, timed on the level's own operator
        lvl.ap = lvl.ap[:3] + (tune_rap_plan(lvl.A, lvl, nxt.A),)
    return AMGHierarchy(levels, _dense(levels[-1].A), project_constant, num_pre_smooth, num_post_smooth, native_dtype,
                        time.perf_counter() - t0)


_DENSE_ROWS: dict = {}


def device_coarse_inverse(coarse_A: torch.Tensor, rank_one: bool) -> torch.Tensor:
    """pp_coarse_inverse of a dense ``[n, n]`` or ``[B, n, n]`` operator on the device: its
    inverse, or with ``rank_one`` its pseudo-inverse for a symmetric operator whose null space
    is the constant vector; float64, queued on the current stream."""
    C_ = coarse_A.detach().to(torch.float64).contiguous()
    n = C_.shape[-1]
    out = torch.empty_like(C_)
    _lib.call("pp_coarse_inverse", _lib.ptr(C_), _lib.ptr(out), n, C_.numel() // (n * n), int(bool(rank_one)),
              _lib.stream_ptr(C_.device))
    return out


def _dense(A: CSRmatrix) -> torch.Tensor:
    """The operator as a dense float64 ``[n, n]``, or ``[B, n, n]`` for B batched operators."""
    n = A.getRows()
    key = (A.row.data_ptr(), n, A.getNnz())
    rows = _DENSE_ROWS.get(key)
    if rows is None or rows[0] is not A.row:
        row = A.row.to(torch.int64)
        counts = row[1:] - row[:-1]
        # row index per entry; computed once per pattern (repeat_interleave waits for the GPU)
        rows = (A.row, torch.repeat_interleave(torch.arange(n, device=row.device), counts))
        _DENSE_ROWS[key] = rows
    rows = rows[1]
    B, nnz = A.getBatchSize(), A.getNnz()
    if B > 1:
        D = torch.zeros(B, n * n, dtype=torch.float64, device=A.value.device)
        D.index_add_(1, rows * n + A.index.to(torch.int64), A.value.reshape(-1)[: B * nnz].reshape(B, nnz).to(torch.float64))
        return D.reshape(B, n, n)
    D = torch.zeros(n, n, dtype=torch.float64, device=A.value.device)
    D.index_put_((rows, A.index.to(torch.int64)), A.value.reshape(-1)[:nnz].to(torch.float64), accumulate=True)
    return D


def galerkin_refresh(hierarchy: AMGHierarchy, csr_mat: CSRmatrix, *, project_constant: Optional[bool] = None,
                     coarse: bool = True) -> AMGHierarchy:
    """Keep P and R, replace the finest operator by ``csr_mat`` and (with ``coarse``) recompute
    the coarse operators as device Galerkin products (amg.py:749-847). A batched ``csr_mat``
    (B operators of one pattern) gives B operators on every level from the same interpolation,
    refreshed together (BatchedAMGHierarchy.refresh, amg.py:1325-1368); a finest-only refresh
    needs a hierarchy of the same batch size."""
    t0 = time.perf_counter()
    fine = hierarchy.levels[0]
    if csr_mat.getRows() != fine.n:
        raise ValueError("operator does not match the finest AMG level")
    B = csr_mat.getBatchSize()
    if not coarse and B != hierarchy.batch_size:
        raise ValueError("a finest-level refresh needs a hierarchy of the same batch size")
    A = CSRmatrix(csr_mat.value.reshape(-1)[: B * csr_mat.getNnz()], csr_mat.index, csr_mat.row)
    A.share_ell_from(csr_mat)
    levels = [Level(n=fine.n, A=A, R_row=fine.R_row, R_col=fine.R_col, R_val=fine.R_val, P_row=fine.P_row,
                    P_col=fine.P_col, P_val=fine.P_val, shared=fine.shared)]
    levels[0].ap = fine.ap
    if coarse:
        for i in range(1, len(hierarchy.levels)):
            prev = levels[-1]
            old = hierarchy.levels[i]
            cval = torch.empty(B * old.A.getNnz(), dtype=A.getDtype(), device=A.value.device)
            # A P of all B operators (double), then R (A P) into the coarse pattern
            ap_val = torch.empty(B * max(hierarchy.levels[i - 1].ap[1].numel(), 1), dtype=torch.float64,
                                 device=A.value.device)
            galerkin_rows(prev.A, hierarchy.levels[i - 1], old.A, cval, ap_val, B)
            del ap_val
            levels.append(Level(n=old.n, A=CSRmatrix(cval, old.A.index, old.A.row), R_row=old.R_row, R_col=old.R_col,
                                R_val=old.R_val, P_row=old.P_row, P_col=old.P_col, P_val=old.P_val, ap=old.ap,
                                shared=old.shared))
        coarse_A = _dense(levels[-1].A)
    else:
        for old in hierarchy.levels[1:]:
            levels.append(old)
        coarse_A = None
    h = AMGHierarchy.__new__(AMGHierarchy)
    h.levels = levels
    h.project_constant = hierarchy.project_constant if project_constant is None else bool(project_constant)
    h.num_pre_smooth, h.num_post_smooth = hierarchy.num_pre_smooth, hierarchy.num_post_smooth
    h.native_dtype = hierarchy.native_dtype
    h.setup_seconds = time.perf_counter() - t0
    h.device_pinv = hierarchy.device_pinv and h.project_constant == hierarchy.project_constant
    if coarse_A is not None:
        h.set_coarse(coarse_A, device=h.device_pinv)
    else:
        h.coarse_pinv64 = hierarchy.coarse_pinv64
        h._pinv_native = hierarchy._pinv_native
        h._packed_parent = hierarchy
    return h


# interpolation cache per sparsity pattern (amg.py:857-868, 952-1045)
_INTERPOLATION_CACHE: dict = {}


def _sparsity_key(csr: CSRmatrix) -> tuple:
    return (str(csr.value.device), csr.getRows(), csr.getNnz(), csr.row.data_ptr(), csr.index.data_ptr())


def clear_interpolation_cache() -> None:
    _INTERPOLATION_CACHE.clear()


def release_pattern(csr_mat: CSRmatrix) -> bool:
    """Drop the cached interpolation (and the hierarchy holding it) of ``csr_mat``'s sparsity
    pattern, freeing its device memory once no hierarchy refreshed from it is alive; True if
    an entry was cached."""
    key = _sparsity_key(csr_mat)
    found = False
    for dtype in (torch.float32, torch.float64):
        found |= _INTERPOLATION_CACHE.pop((key, dtype), None) is not None
    return found


def hierarchy_for(csr_mat: CSRmatrix, *, project_constant: bool = False, reuse_interpolation: bool = True,
                  native_dtype=torch.float32, **build_kwargs) -> AMGHierarchy:
    """Hierarchy for ``csr_mat``: a new host setup for the first operator of a pattern, a
    device Galerkin refresh with the cached interpolation for every further one."""
    if not reuse_interpolation:
        return build_amg_hierarchy(csr_mat, project_constant=project_constant, native_dtype=native_dtype, **build_kwargs)
    # one interpolation per pattern and operator dtype (a float32 pressure and a potential solved
    # in float64 on the same grid keep one each)
    key = (_sparsity_key(csr_mat), csr_mat.getDtype())
    base = _INTERPOLATION_CACHE.get(key)
    if base is None:
        h = build_amg_hierarchy(csr_mat, project_constant=project_constant, native_dtype=native_dtype, **build_kwargs)
        _INTERPOLATION_CACHE[key] = h
        return h
    h = galerkin_refresh(base, csr_mat, project_constant=project_constant)
    h.native_dtype = native_dtype
    return h


def batched_hierarchy_for(csr_mat: CSRmatrix, batch_size: int, *, project_constant: bool = False,
                          previous: Optional[AMGHierarchy] = None, coarse: bool = True, native_dtype=torch.float32,
                          **build_kwargs) -> AMGHierarchy:
    """Hierarchy of the B operators of a batched ``csr_mat`` (amg.py:1414-1482): the cached
    interpolation of the sparsity pattern (set up from the first operator if the pattern is
    new) shared by all environments, every level holding B operators refreshed together.
    ``previous`` (same pattern and batch size) allows a finest-only refresh (``coarse=False``)."""
    nnz = csr_mat.getNnz()
    if csr_mat.getBatchSize() != batch_size:
        raise ValueError(f"matrix holds {csr_mat.getBatchSize()} operators, not {batch_size}")
    if (previous is None or previous.batch_size != batch_size or previous.levels[0].A.getNnz() != nnz
            or previous.levels[0].A.row is not csr_mat.row):
        key = (_sparsity_key(csr_mat), csr_mat.getDtype())
        base = _INTERPOLATION_CACHE.get(key)
        if base is None:
            first = CSRmatrix(csr_mat.value.reshape(-1)[:nnz], csr_mat.index, csr_mat.row)
            base = build_amg_hierarchy(first, project_constant=project_constant, native_dtype=native_dtype,
                                       **build_kwargs)
            _INTERPOLATION_CACHE[key] = base
        previous = base
        coarse = True
    h = galerkin_refresh(previous, csr_mat, project_constant=project_constant, coarse=coarse)
    h.native_dtype = native_dtype
    return h


@dataclass
class AMGSolveInfo:
    finalResidual: float
    usedIterations: int
    converged: bool
    isFiniteResidual: bool


def amg_pcg_solve(rhs: torch.Tensor, x: torch.Tensor, hierarchy: AMGHierarchy, *, tol, max_iter: int,
                  return_best_result: bool = False) -> list:
    """Solve ``A x = rhs`` in place with AMG-preconditioned CG; ``A`` is the finest level
    (amg.py:1544-1627). Non-finite results are zeroed, as the reference's contract."""
    A = hierarchy.levels[0].A
    n = A.getRows()
    if rhs.numel() % n:
        raise ValueError(f"RHS size {rhs.numel()} is not a multiple of the matrix size {n}.")
    pdtype = hierarchy.native_dtype or rhs.dtype
    packed = hierarchy.packed_levels(pdtype)
    pinv = hierarchy.coarse_pinv(pdtype)
    tols = [float(t) for t in tol] if isinstance(tol, (list, tuple)) else []
    tol_scalar = max(tols) if tols else float(tol)
    same_dtype = A.value.dtype == rhs.dtype
    ell = (A.ell_col, A.ell_val, A.ell_width) if (same_dtype and A.ell_valid()) else None
    infos = ops.amg_pcg_packed(A.row, A.index, A.value if same_dtype else A.value.to(rhs.dtype),
                               packed, pinv, rhs.reshape(-1), x.view(-1), int(max_iter), tol_scalar, True,
                               bool(return_best_result), hierarchy.project_constant, hierarchy.num_pre_smooth,
                               hierarchy.num_post_smooth, JACOBI_OMEGA, tols, ell=ell)
    for i, info in enumerate(infos):
        if not info.isFiniteResidual:
            x.view(len(infos), -1)[i].zero_()
    return infos

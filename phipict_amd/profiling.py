"""Sampled per-kernel timing through the library's event timer (``pp_kernel_timer_*``): a
pseudo-random 1/k of the launches of each instrumented kernel is bracketed by HIP events on the
stream it runs on; and summaries of rocprofv3 memory-counter runs."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib

KERNEL_NAMES = {
    1: "advection_matrix", 2: "advection_rhs", 3: "pressure_matrix", 4: "pressure_rhs", 5: "pressure_div",
    6: "correct_velocity", 7: "epot_rhs", 8: "current_density", 9: "copy_result",
    20: "residual", 21: "cg_direction", 22: "cg_update", 23: "bicg_p", 24: "bicg_v", 25: "bicg_t", 26: "bicg_update",
    27: "cg_small", 28: "bicg_small",
    30: "pcg_dots", 31: "pcg_direction", 32: "pcg_update", 33: "amg_presmooth", 34: "amg_jacobi",
    35: "amg_level_residual", 36: "amg_restrict", 37: "amg_prolong", 38: "amg_coarse", 39: "mean", 40: "subtract_mean",
    41: "amg_post_jacobi", 42: "amg_tail", 50: "spmv", 51: "galerkin", 52: "galerkin_ap",
    60: "seg_cg_direction", 61: "seg_cg_update", 62: "seg_bicg",
}


@dataclass
class KernelTiming:
    kernel_id: int
    name: str
    rows: int
    launches: int
    sampled: int
    sampled_ms: float

    @property
    def avg_ms(self) -> float:
        return self.sampled_ms / self.sampled if self.sampled else float("nan")

    @property
    def est_total_ms(self) -> float:
        """Total time of all launches, extrapolated from the sampled ones."""
        return self.avg_ms * self.launches if self.sampled else 0.0


def start(sample_every: int = 1) -> None:
    _lib.call("pp_kernel_timer_start", int(sample_every))


def stop() -> None:
    _lib.call("pp_kernel_timer_stop")


def read() -> list:
    count = C.c_int32(0)
    _lib.call("pp_kernel_timer_read", 0, None, None, None, None, None, C.byref(count))
    n = count.value
    ids = np.zeros(max(n, 1), np.int32)
    rows = np.zeros(max(n, 1), np.int32)
    ms = np.zeros(max(n, 1), np.float64)
    sampled = np.zeros(max(n, 1), np.int64)
    launches = np.zeros(max(n, 1), np.int64)
    _lib.call("pp_kernel_timer_read", n, ids.ctypes.data, rows.ctypes.data, ms.ctypes.data, sampled.ctypes.data,
              launches.ctypes.data, C.byref(count))
    return [KernelTiming(int(ids[i]), KERNEL_NAMES.get(int(ids[i]), f"kernel{ids[i]}"), int(rows[i]), int(launches[i]),
                         int(sampled[i]), float(ms[i])) for i in range(n)]


# ----------------------------------------------------------------------------- memory counters
# rocprofv3 names of the library kernels behind the timer's kernel names (the two refresh
# products run whichever of k_rap_reg, k_rap_rows and k_rap_rows_long the level's plan chose)
TIMER_TO_KERNEL = {
    "advection_matrix": "k_advection_matrix", "advection_rhs": "k_advection_rhs", "pressure_matrix": "k_pressure_matrix",
    "pressure_rhs": "k_pressure_rhs", "pressure_div": "k_pressure_div", "correct_velocity": "k_correct_velocity",
    "epot_rhs": "k_epot_rhs", "current_density": "k_current_density", "copy_result": "k_copy_result",
    "cg_direction": "k_cg_turn", "cg_update": "k_cg_step", "bicg_p": "k_bicg_dir", "bicg_v": "k_bicg_apply",
    "bicg_t": "k_bicg_stab", "bicg_update": "k_bicg_step", "pcg_direction": "k_pcg_dir", "pcg_update": "k_cg_step",
    "amg_presmooth": "k_level", "amg_jacobi": "k_level", "amg_post_jacobi": "k_level", "amg_restrict": "k_level",
    "amg_prolong": "k_level", "amg_tail": "k_vcycle_tail", "galerkin": "k_rap_", "galerkin_ap": "k_rap_",
    "seg_cg_direction": "k_cg_turn", "seg_cg_update": "k_cg_step",
}
#: the level pass of each AMG timer: k_level's first template argument (OP_FIRST, OP_SWEEP,
#: OP_APPLY); the finest level's first sweep runs inside k_pcg_decide
AMG_TIMER_OP = {"amg_presmooth": "0", "amg_jacobi": "1", "amg_post_jacobi": "1", "amg_restrict": "3",
                "amg_prolong": "3"}


def split_kernel_name(name: str) -> tuple:
    """'void pp::k_x<1, double, float>(...)' -> ('k_x', ('1', 'double', 'float'))."""
    head = name.split("(")[0].strip()
    if head.startswith("void "):
        head = head[5:]
    args = ()
    if "<" in head:
        base, rest = head.split("<", 1)
        depth, cur, out = 0, "", []
        for ch in rest:
            if ch == "<":
                depth += 1
            elif ch == ">":
                if depth == 0:
                    break
                depth -= 1
            if ch == "," and depth == 0:
                out.append(cur.strip())
                cur = ""
            else:
                cur += ch
        out.append(cur.strip())
        args = tuple(out)
        head = base
    return head.split("::")[-1], args


def read_counter_csv(path) -> dict:
    """Per dispatch id: (kernel name, grid size, {counter: value}) from a rocprofv3
    ``counter_collection.csv`` (values summed over the counter's dimensions)."""
    import csv

    out = {}
    with open(path, newline="") as f:
        for r in csv.DictReader(f):
            did = int(r["Dispatch_Id"])
            name = r["Kernel_Name"]
            grid = int(r.get("Grid_Size", 0) or 0)
            ent = out.setdefault(did, (name, grid, {}))
            ent[2][r["Counter_Name"]] = ent[2].get(r["Counter_Name"], 0.0) + float(r["Counter_Value"])
    return out


def traffic_table(fetch_csv, write_csv, copy_bytes: int, idle_fraction: float = 0.02) -> dict:
    """Measured HBM bytes per dispatch from a FETCH_SIZE run and a WRITE_SIZE run (KiB per
    dispatch each), scaled by what the same counters read for ``k_copy_bytes`` moving
    ``copy_bytes`` bytes. Dispatches whose fetch is below ``idle_fraction`` of the group's
    largest are counted as early exits (a solve that had already converged) and kept apart.
    Keys: (kernel, template arguments, grid size in work-items)."""
    fetch = read_counter_csv(fetch_csv)
    write = read_counter_csv(write_csv)

    def counter(d, key):
        return next((v for k, v in d.items() if k.startswith(key)), None)

    def calib(table, key):
        copies = [(grid, counter(c, key)) for name, grid, c in table.values()
                  if split_kernel_name(name)[0] == "k_copy_bytes"]
        # the calibration copies are the widest k_copy_bytes launches (markers are one workgroup)
        widest = max((g for g, _ in copies), default=0)
        vals = [v for g, v in copies if g == widest and v]
        if not vals:
            raise ValueError("no k_copy_bytes dispatch in the counter run")
        return float(np.median(vals)) * 1024.0 / copy_bytes

    ff, fw = calib(fetch, "FETCH_SIZE"), calib(write, "WRITE_SIZE")
    groups = {}
    for tab, key, scale, slot in ((fetch, "FETCH_SIZE", ff, 0), (write, "WRITE_SIZE", fw, 1)):
        for name, grid, c in tab.values():
            v = counter(c, key)
            if v is None:
                continue
            g = groups.setdefault(split_kernel_name(name) + (grid,), ([], []))
            g[slot].append(v * 1024.0 / scale)
    def busy_mean(vals):
        v = np.asarray(vals, dtype=np.float64)
        top = v.max()
        busy = v >= idle_fraction * top if top > 0 else np.ones(v.size, dtype=bool)
        return float(v[busy].mean()), int((~busy).sum())

    rows = {}
    for key_, (fb, wb) in groups.items():
        if not fb or not wb:
            continue
        fmean, fidle = busy_mean(fb)
        wmean, _ = busy_mean(wb)
        rows[key_] = {"dispatches": len(fb), "early_exits": fidle, "fetch_bytes": fmean, "write_bytes": wmean}
    return {"fetch_scale": ff, "write_scale": fw, "kernels": rows}

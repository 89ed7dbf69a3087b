"""Sampled per-kernel timing through the library's event timer (``pp_kernel_timer_*``): every
k-th launch of each instrumented kernel is bracketed by HIP events on the stream it runs on."""
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
    41: "amg_post_jacobi", 42: "amg_tail", 50: "spmv", 51: "galerkin",
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

"""ctypes binding of the gfx950 library (C ABI in ``include/phipict_hip.h``).

The library is the only compute path of this package: importing this module fails loudly
when ``libphipict_hip.so`` is missing, and there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

# PHIPICT_HIP_LIBRARY may name another build of the same library (e.g. a tuning variant)
_LIB_PATH = Path(os.environ.get("PHIPICT_HIP_LIBRARY") or Path(__file__).resolve().parent / "_lib" / "libphipict_hip.so")

PP_F32, PP_F64 = 4, 8
PP_BOUND_CONNECTED, PP_BOUND_PERIODIC, PP_BOUND_FIXED = 20, 21, 30
PP_BC_DIRICHLET, PP_BC_NEUMANN = 0, 1
PP_CONV_NORM2, PP_CONV_NORM2_NORMALIZED = 0, 1


class BoundaryDesc(C.Structure):
    _fields_ = [
        ("type", C.c_int32),
        ("vel_type", C.c_int32),
        ("vel_static", C.c_int32),
        ("vel_batched", C.c_int32),
        ("has_transform", C.c_int32),
        ("pot_type", C.c_int32),
        ("conn_block", C.c_int32),
        ("conn_axes", C.c_int32 * 3),
        ("scalar_static", C.c_int32),
        ("scalar_batched", C.c_int32),
        ("pot_values_batched", C.c_int32),
        ("reserved", C.c_int32),
        ("vel", C.c_uint64),
        ("transform", C.c_uint64),
        ("pot_types", C.c_uint64),
        ("pot_values", C.c_uint64),
        ("pot_cw", C.c_double),
        ("scalar", C.c_uint64),
        ("scalar_types", C.c_uint64),
    ]


class BlockDesc(C.Structure):
    _fields_ = [
        ("size", C.c_int32 * 3),
        ("cells", C.c_int32),
        ("global_offset", C.c_int32),
        ("csr_offset", C.c_int32),
        ("has_transform", C.c_int32),
        ("vsource_mode", C.c_int32),
        ("vsource_batched", C.c_int32),
        ("visc_mode", C.c_int32),
        ("transform_diagonal", C.c_int32),
        ("pad1", C.c_int32),
        ("transform", C.c_uint64),
        ("velocity", C.c_uint64),
        ("pressure", C.c_uint64),
        ("epot", C.c_uint64),
        ("vsource", C.c_uint64),
        ("viscosity", C.c_uint64),
        ("scalar", C.c_uint64),
        ("bnd", BoundaryDesc * 6),
    ]


class CsrDesc(C.Structure):
    _fields_ = [
        ("row", C.c_void_p),
        ("col", C.c_void_p),
        ("val", C.c_void_p),
        ("n", C.c_int32),
        ("nnz", C.c_int32),
        ("num_matrices", C.c_int32),
        ("ell_width", C.c_int32),
        ("ell_col", C.c_void_p),
        ("ell_val", C.c_void_p),
    ]


class DomainDesc(C.Structure):
    _fields_ = [
        ("dims", C.c_int32),
        ("num_blocks", C.c_int32),
        ("num_cells", C.c_int32),
        ("batch", C.c_int32),
        ("dtype", C.c_int32),
        ("max_block_cells", C.c_int32),
        ("viscosity", C.c_double),
        ("blocks", C.c_void_p),
        ("time_step", C.c_void_p),
        ("A", C.c_void_p),
        ("velocity_rhs", C.c_void_p),
        ("velocity_result", C.c_void_p),
        ("pressure_rhs", C.c_void_p),
        ("pressure_rhs_div", C.c_void_p),
        ("pressure_result", C.c_void_p),
        ("epot_rhs", C.c_void_p),
        ("epot_result", C.c_void_p),
        ("C", CsrDesc),
        ("P", CsrDesc),
        ("E", CsrDesc),
        ("scalar_channels", C.c_int32),
        ("scalar_visc_count", C.c_int32),
        ("scalar_viscosity", C.c_void_p),
        ("scalar_rhs", C.c_void_p),
        ("scalar_result", C.c_void_p),
        ("advection_scheme", C.c_int32),
        ("reserved", C.c_int32),
    ]


class SolveResult(C.Structure):
    _fields_ = [
        ("final_residual", C.c_double),
        ("used_iterations", C.c_int32),
        ("converged", C.c_int32),
        ("is_finite_residual", C.c_int32),
        ("pad", C.c_int32),
    ]


class SellDesc(C.Structure):
    _fields_ = [
        ("slice_off", C.c_void_p),
        ("slice_steps", C.c_void_p),
        ("col", C.c_void_p),
        ("val", C.c_void_p),
        ("lanes", C.c_int32),
        ("pad", C.c_int32),
        ("total", C.c_int64),
    ]


class AmgLevel(C.Structure):
    _fields_ = [
        ("n", C.c_int32),
        ("nnz_a", C.c_int32),
        ("n_coarse", C.c_int32),
        ("nnz_r", C.c_int32),
        ("nnz_p", C.c_int32),
        ("a_batched", C.c_int32),
        ("a_row", C.c_void_p),
        ("a_col", C.c_void_p),
        ("a_val", C.c_void_p),
        ("dinv", C.c_void_p),
        ("r_row", C.c_void_p),
        ("r_col", C.c_void_p),
        ("r_val", C.c_void_p),
        ("p_row", C.c_void_p),
        ("p_col", C.c_void_p),
        ("p_val", C.c_void_p),
        ("a_ell_width", C.c_int32),
        ("pad", C.c_int32),
        ("a_ell_col", C.c_void_p),
        ("a_ell_val", C.c_void_p),
        ("a_sell", SellDesc),
        ("r_sell", SellDesc),
        ("p_sell", SellDesc),
    ]


assert C.sizeof(BoundaryDesc) == 112, C.sizeof(BoundaryDesc)
assert C.sizeof(BlockDesc) == 776, C.sizeof(BlockDesc)

_i32, _i64, _vp, _dbl = C.c_int32, C.c_int64, C.c_void_p, C.c_double
_P = C.POINTER

# name -> (argtypes); every entry point returns int
SIGNATURES = {
    "pp_version": [],
    "pp_device_sync": [],
    "pp_coords_to_transforms": [_i32, _P(_i32), _vp, _vp, _i32, _vp],
    "pp_coords_to_face_transforms": [_i32, _P(_i32), _vp, _vp, _i32, _vp],
    "pp_transpose_2d": [_vp, _i32, _i32, _i32, _vp, _vp],
    "pp_csr_pattern": [_P(DomainDesc), _vp, _vp, _vp],
    "pp_setup_advection_matrix": [_P(DomainDesc), _i32, _vp],
    "pp_setup_advection_velocity": [_P(DomainDesc), _i32, _i32, _vp],
    "pp_setup_pressure_matrix": [_P(DomainDesc), _i32, _vp],
    "pp_setup_pressure_rhs": [_P(DomainDesc), _i32, _i32, _vp],
    "pp_setup_pressure_rhs_div": [_P(DomainDesc), _i32, _i32, _vp],
    "pp_correct_velocity": [_P(DomainDesc), _i32, _vp],
    "pp_setup_epot_matrix": [_P(DomainDesc), _i32, _vp],
    "pp_compute_epot_rhs": [_P(DomainDesc), _vp, _vp, _vp],
    "pp_compute_current_density": [_P(DomainDesc), _vp, _vp, _vp, _vp],
    "pp_copy_result": [_P(DomainDesc), _i32, _i32, _vp],
    "pp_max_velocity": [_P(DomainDesc), _i32, _i32, _i32, _i32, _vp, _vp],
    "pp_setup_advection_matrix_scalar": [_P(DomainDesc), _i32, _i32, _vp],
    "pp_setup_advection_scalar": [_P(DomainDesc), _i32, _vp],
    "pp_setup_advection_matrix_grad": [_P(DomainDesc), _vp, _vp, _vp, _vp, _vp],
    "pp_setup_advection_velocity_grad": [_P(DomainDesc), _vp, _vp, _vp, _vp, _vp, _vp],
    "pp_setup_pressure_matrix_grad": [_P(DomainDesc), _vp, _vp, _vp, _vp],
    "pp_setup_pressure_rhs_grad": [_P(DomainDesc), _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "pp_setup_pressure_rhs_div_grad": [_P(DomainDesc), _vp, _i32, _vp, _vp, _vp],
    "pp_correct_velocity_grad": [_P(DomainDesc), _vp, _vp, _i32, _vp, _vp, _vp, _vp],
    "pp_compute_epot_rhs_grad": [_P(DomainDesc), _vp, _vp, _vp, _vp],
    "pp_compute_current_density_grad": [_P(DomainDesc), _vp, _vp, _vp, _vp, _vp],
    "pp_transform_vectors": [_i32, _i32, _i32, _vp, _vp, _i32, _i32, _vp, _i32, _vp],
    "pp_solve_linear": [_P(CsrDesc), _vp, _vp, _i32, _i32, _i32, _P(_dbl), _i32, _i32, _i32, _i32, _i32, _i32, _i32,
                        _vp, _P(SolveResult)],
    "pp_solve_linear_noalloc": [_P(CsrDesc), _vp, _vp, _i32, _i32, _i32, _P(_dbl), _i32, _i32, _i32, _i32, _i32, _i32, _i32,
                        _vp, _P(SolveResult)],
    "pp_amg_pcg_solve": [_P(CsrDesc), _P(AmgLevel), _i32, _vp, _i32, _i32, _i32, _vp, _vp, _i32, _i32, _i32, _dbl,
                         _i32, _i32, _i32, _i32, _i32, _dbl, _P(_dbl), _i32, _vp, _P(SolveResult)],
    "pp_release_workspaces": [],
    "pp_sparse_outer_product": [_vp, _vp, _P(CsrDesc), _i32, _vp],
    "pp_csr_transpose": [_P(CsrDesc), _i32, _vp, _vp, _vp, _vp],
    "pp_spmv": [_P(CsrDesc), _vp, _vp, _i32, _i32, _vp],
    "pp_amg_setup": [_i32, _vp, _vp, _vp, _dbl, _i32, _i32, _P(_vp)],
    "pp_amg_num_levels": [_vp],
    "pp_amg_level_sizes": [_vp, _i32, _P(_i32), _P(_i32), _P(_i32), _P(_i32)],
    "pp_amg_level_copy": [_vp, _i32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "pp_amg_free": [_vp],
    "pp_galerkin_rows": [_P(CsrDesc), _i32, _vp, _vp, _vp, _vp, _vp, _i32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i32,
                         _i32, _i32, _vp, _vp, _i32, _vp],
    "pp_inverse_diagonal": [_P(CsrDesc), _i32, _vp, _i32, _vp],
    "pp_cast": [_vp, _i32, _vp, _i32, _i64, _vp],
    "pp_ell_from_csr": [_vp, _vp, _vp, _i32, _i32, _i32, _i32, _i32, _vp, _vp, _i32, _vp],
    "pp_ap_pattern": [_i32, _vp, _vp, _i32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "pp_coarse_inverse": [_vp, _vp, _i32, _i32, _i32, _vp],
    "pp_kernel_timer_start": [_i32],
    "pp_kernel_timer_stop": [],
    "pp_kernel_timer_read": [_i32, _vp, _vp, _vp, _vp, _vp, _P(_i32)],
    "pp_copy_bytes": [_vp, _vp, _i64, _vp],
    "pp_sell_fill": [_vp, _vp, _vp, _i32, _i32, _i32, _i32, _vp, _vp, _i64, _i32, _vp, _vp, _i32, _vp],
    "pp_sell_probe": [_P(SellDesc), _i32, _i32, _i32, _i32, _vp, _P(_dbl)],
    "pp_set_krylov_settings": [_i32, _i32, _i32],
    "pp_solve_segmented": [_P(CsrDesc), _vp, _i32, _i32, _vp, _vp, _i32, _i32, _P(_dbl), _i32, _i32, _i32, _i32, _vp,
                           _vp],
    "pp_get_krylov_settings": [_P(_i32), _P(_i32), _P(_i32)],
    "pp_cross_eb": [_vp, _i32, _i32, _i32, _vp, _i32, _vp, _i32, _vp],
    "pp_lorentz_source": [_vp, _i32, _i32, _i32, _vp, _i32, _vp, _i32, _i32, _i32, _vp, _i32, _i32, _vp, _i32, _vp],
}


class PhipictHipError(RuntimeError):
    """A failure reported by the gfx950 library."""


def _load() -> C.CDLL:
    if not _LIB_PATH.exists():
        raise ImportError(
            f"phipict_amd: the gfx950 library {_LIB_PATH} is not built. Run "
            "`python -m phipict_amd.build` (or __graft_entry__.build()) first; there is no CPU fallback."
        )
    # torch must be imported first so that the library shares torch's HIP runtime
    import torch  # noqa: F401

    lib = C.CDLL(str(_LIB_PATH), mode=os.RTLD_NOW | os.RTLD_GLOBAL)
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = C.c_int
    lib.pp_last_error.argtypes = []
    lib.pp_last_error.restype = C.c_char_p
    return lib


LIB = _load()
LIB_PATH = _LIB_PATH


def check(rc: int, what: str = "") -> None:
    """Raise PhipictHipError for a non-zero return code of an entry point."""
    if rc != 0:
        msg = LIB.pp_last_error()
        raise PhipictHipError(f"{what or 'phipict_hip'} failed ({rc}): {msg.decode() if msg else ''}")


def call(name: str, *args) -> int:
    rc = getattr(LIB, name)(*args)
    check(rc, name)
    return rc


def dtype_code(dtype) -> int:
    import torch

    if dtype == torch.float64:
        return PP_F64
    if dtype == torch.float32:
        return PP_F32
    raise TypeError(f"unsupported dtype {dtype}; float32 and float64 are supported")


def stream_ptr(device=None) -> int:
    """The current torch stream of ``device`` as a raw hipStream_t value."""
    import torch

    return torch.cuda.current_stream(device).cuda_stream


def ptr(t) -> int:
    return 0 if t is None else int(t.data_ptr())

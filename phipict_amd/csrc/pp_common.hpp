// Shared host/device helpers of the gfx950 PISO library: error handling, dtype dispatch,
// index arithmetic and per-cell boundary resolution on the domain descriptors of
// include/phipict_hip.h.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>

#include "../../include/phipict_hip.h"

namespace pp {

// ---- error handling ----------------------------------------------------------------
void set_error(const std::string &msg);
const char *get_error();

#define PP_CHECK_ARG(cond, msg)                         \
    do {                                                \
        if (!(cond)) {                                  \
            ::pp::set_error(std::string(__func__) + ": " + (msg)); \
            return PP_ERR_ARG;                          \
        }                                               \
    } while (0)

#define PP_HIP(call)                                                                    \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess) {                                                         \
            ::pp::set_error(std::string(__func__) + ": " #call " -> " + hipGetErrorString(e_)); \
            return PP_ERR_HIP;                                                          \
        }                                                                               \
    } while (0)

#define PP_LAUNCH_CHECK()                                                                \
    do {                                                                                 \
        hipError_t e_ = hipGetLastError();                                               \
        if (e_ != hipSuccess) {                                                          \
            ::pp::set_error(std::string(__func__) + ": launch -> " + hipGetErrorString(e_)); \
            return PP_ERR_HIP;                                                           \
        }                                                                                \
    } while (0)

#define PP_DISPATCH(DTYPE, ...)                                   \
    do {                                                          \
        if ((DTYPE) == PP_F64) { using T = double; __VA_ARGS__; } \
        else if ((DTYPE) == PP_F32) { using T = float; __VA_ARGS__; } \
        else { ::pp::set_error(std::string(__func__) + ": unsupported dtype"); return PP_ERR_ARG; } \
    } while (0)

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

// ---- sampled kernel timer ---------------------------------------------------------------
// When enabled (pp_kernel_timer_start), every k-th launch of an instrumented kernel is
// bracketed by two HIP events on its own stream; durations are summed per (kernel, rows).
struct TimerSlot;
TimerSlot *timer_begin(int kernel_id, int rows, hipStream_t s);
bool timer_enabled(); // launches are being timed (their events must not be captured in graphs)
void timer_end(TimerSlot *slot, hipStream_t s);

#define PP_TIMED(ID, ROWS, STREAM, ...)                                       \
    do {                                                                      \
        ::pp::TimerSlot *ts_ = ::pp::timer_begin((ID), (ROWS), (STREAM));     \
        __VA_ARGS__;                                                          \
        if (ts_) ::pp::timer_end(ts_, (STREAM));                              \
    } while (0)

// ---- device helpers ------------------------------------------------------------------
template <typename T> __device__ __forceinline__ T dsqrt(T x);
template <> __device__ __forceinline__ float dsqrt<float>(float x) { return sqrtf(x); }
template <> __device__ __forceinline__ double dsqrt<double>(double x) { return sqrt(x); }

struct Pos { int a[3]; };

__device__ __forceinline__ int flat3(const int *p, const int32_t *s) { return p[0] + s[0] * (p[1] + s[1] * p[2]); }

// Block containing global cell g (blocks sorted by global_offset).
__device__ __forceinline__ int find_block(const pp_block_desc *__restrict__ blocks, int nb, int g) {
    int lo = 0, hi = nb - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (blocks[mid].global_offset <= g) lo = mid; else hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ void unflat(int f, const int32_t *s, int *p) {
    p[0] = f % s[0];
    int r = f / s[0];
    p[1] = r % s[1];
    p[2] = r / s[1];
}

__device__ __forceinline__ bool at_bound(const int *p, int bound, const pp_block_desc &b) {
    const int axis = bound >> 1;
    return (bound & 1) ? p[axis] == b.size[axis] - 1 : p[axis] == 0;
}
// prescribed face (the reference's isEmptyBound, topology.cuh:24-31, FIXED only here)
__device__ __forceinline__ bool is_fixed(const pp_block_desc &b, int bound) { return b.bnd[bound].type == PP_BOUND_FIXED; }
__device__ __forceinline__ int face_sign(int bound) { return (bound & 1) * 2 - 1; }

__device__ __forceinline__ int face_cells(const pp_block_desc &b, int bound) {
    const int a = bound >> 1;
    return b.cells / b.size[a];
}
__device__ __forceinline__ int face_flat(const int *p, int bound, const pp_block_desc &b) {
    int s[3] = {b.size[0], b.size[1], b.size[2]};
    int q[3] = {p[0], p[1], p[2]};
    s[bound >> 1] = 1;
    q[bound >> 1] = 0;
    return q[0] + s[0] * (q[1] + s[1] * q[2]);
}
__device__ __forceinline__ int pos_mod(int a, int m) { int r = a % m; return r < 0 ? r + m : r; }

// Position in the connected block across face `bdim` (computeConnectedPos, topology.cuh:246-266).
// `border` cells in from the connected face; a block one cell thick there clamps it to 0 (the
// reference would read outside the block).
__device__ __forceinline__ void conn_pos(const pp_block_desc *__restrict__ blocks, int d, const int *pos, int bdim,
                                         const pp_boundary_desc &cb, int border, int *out) {
    const pp_block_desc &ob = blocks[cb.conn_block];
    out[0] = pos[0]; out[1] = pos[1]; out[2] = pos[2];
    int ca = cb.conn_axes[0] >> 1;
    if (border > ob.size[ca] - 1) border = ob.size[ca] - 1;
    out[ca] = (cb.conn_axes[0] & 1) ? ob.size[ca] - 1 - border : border;
    if (d > 1) {
        int axis = (bdim + 1) % d;
        ca = cb.conn_axes[1] >> 1;
        out[ca] = (cb.conn_axes[1] & 1) ? ob.size[ca] - 1 - pos[axis] : pos[axis];
        if (d > 2) {
            axis = (bdim + 2) % d;
            ca = cb.conn_axes[2] >> 1;
            out[ca] = (cb.conn_axes[2] & 1) ? ob.size[ca] - 1 - pos[axis] : pos[axis];
        }
    }
}
__device__ __forceinline__ int conn_channel(int d, int w, int bdim, const pp_boundary_desc &cb) {
    return cb.conn_axes[pos_mod(w - bdim, d)] >> 1;
}

// Neighbour across face `bound`: interior, periodic wrap or connected block. Returns the
// neighbour block index, its position and the channel of the face normal there.
__device__ __forceinline__ int neighbor(const pp_block_desc *__restrict__ blocks, int bi, int d, const int *p,
                                        int bound, int *np, int *nw) {
    const pp_block_desc &b = blocks[bi];
    const int dim = bound >> 1;
    np[0] = p[0]; np[1] = p[1]; np[2] = p[2];
    *nw = dim;
    const bool atb = at_bound(p, bound, b);
    if (atb && b.bnd[bound].type == PP_BOUND_CONNECTED) {
        const pp_boundary_desc &cb = b.bnd[bound];
        conn_pos(blocks, d, p, dim, cb, 0, np);
        *nw = conn_channel(d, dim, dim, cb);
        return cb.conn_block;
    }
    if (atb && b.bnd[bound].type == PP_BOUND_PERIODIC) np[dim] = (bound & 1) ? 0 : b.size[dim] - 1;
    else np[dim] = p[dim] + face_sign(bound);
    return bi;
}

// Strided view of the 2*D*D+1 scalars {M, Minv, det} of one cell or face cell.
template <typename T>
struct TView {
    const T *p;
    int stride;
    __device__ __forceinline__ T operator[](int k) const { return p[(size_t)k * stride]; }
};

// Per-cell metric access. Cell transforms are component-major, [2*D*D+1][cells] (neighbouring
// work-items read neighbouring addresses); boundary face transforms are [face cells][2*D*D+1].
template <typename T, int D>
struct Metric {
    static constexpr int TS = 2 * D * D + 1;
    __device__ static __forceinline__ TView<T> at(const pp_block_desc &b, int f) {
        return TView<T>{reinterpret_cast<const T *>(b.transform) + f, b.cells};
    }
    __device__ static __forceinline__ TView<T> face(const pp_boundary_desc &fb, int ff) {
        return TView<T>{reinterpret_cast<const T *>(fb.transform) + (size_t)ff * TS, 1};
    }
    __device__ static __forceinline__ T det(const pp_block_desc &b, int f) {
        if (!b.has_transform) return T(1);
        return at(b, f)[2 * D * D];
    }
    // det * |Minv[w]|^2 (getLaplaceCoefficientOrthogonal, discretization.cuh:359-374)
    __device__ static __forceinline__ T alpha_t(const TView<T> &t, int w) {
        T r = 0;
#pragma unroll
        for (int i = 0; i < D; ++i) {
            const T m = t[D * D + w * D + i];
            r += m * m;
        }
        return t[2 * D * D] * r;
    }
    // Blocks whose metrics are diagonal (transform_diagonal: every off-diagonal entry of M and
    // Minv is 0, e.g. rectilinear grids) read only the diagonal and det: the same values as the
    // full rows, whose other products are exact zeros, from a third of the loads.
    __device__ static __forceinline__ T alpha(const pp_block_desc &b, int f, int w) {
        if (!b.has_transform) return T(1);
        const TView<T> t = at(b, f);
        if (b.transform_diagonal) {
            const T m = t[D * D + w * D + w];
            return t[2 * D * D] * (m * m);
        }
        return alpha_t(t, w);
    }
    // det * Minv[w] . v (VelocityToContravariantComponent, velocity.cuh:24-39)
    __device__ static __forceinline__ T contra_t(const TView<T> &t, int w, const T *v) {
        T r = 0;
#pragma unroll
        for (int k = 0; k < D; ++k) r += t[D * D + w * D + k] * v[k];
        return t[2 * D * D] * r;
    }
    __device__ static __forceinline__ T minv_norm(const pp_block_desc &b, int f, int axis) {
        if (!b.has_transform) return T(1);
        const TView<T> t = at(b, f);
        if (b.transform_diagonal) {
            const T m = t[D * D + axis * D + axis];
            return dsqrt<T>(m * m);
        }
        T r = 0;
#pragma unroll
        for (int i = 0; i < D; ++i) {
            const T m = t[D * D + axis * D + i];
            r += m * m;
        }
        return dsqrt<T>(r);
    }
};

template <typename T>
__device__ __forceinline__ T visc_at(const pp_domain &dom, const pp_block_desc &b, int f) {
    if (b.visc_mode == 0) return T(dom.viscosity);
    const T *v = reinterpret_cast<const T *>(b.viscosity);
    return b.visc_mode == 1 ? v[0] : v[f];
}

// Passive-scalar diffusivity of channel ch (getViscosity with forPassiveScalar,
// discretization.cuh:1141-1151): the domain's scalar viscosity (one value or one per channel),
// else the domain viscosity
template <typename T>
__device__ __forceinline__ T scalar_visc(const pp_domain &dom, int ch) {
    if (dom.scalar_visc_count == 0) return T(dom.viscosity);
    return reinterpret_cast<const T *>(dom.scalar_viscosity)[dom.scalar_visc_count == 1 ? 0 : ch];
}
// passive-scalar condition of a fixed face for channel ch (Dirichlet when no types are given)
__device__ __forceinline__ bool scalar_neumann(const pp_boundary_desc &fb, int ch) {
    return fb.scalar_types && reinterpret_cast<const int8_t *>(fb.scalar_types)[ch] == PP_BC_NEUMANN;
}
// passive-scalar face data of channel ch at face cell ff (0 when the face has none)
template <typename T>
__device__ __forceinline__ T fixed_scalar(const pp_boundary_desc &fb, int C, int fc, int ff, int ch, int env) {
    if (!fb.scalar) return T(0);
    const T *data = reinterpret_cast<const T *>(fb.scalar);
    const int e = fb.scalar_batched ? env : 0;
    if (fb.scalar_static) return data[e * C + ch];
    return data[((size_t)e * C + ch) * fc + ff];
}

// Velocity component k of cell f of block b (env e): from the block field (glob == nullptr)
// or from a flat [batch][d][n] vector whose env slice starts at glob.
template <typename T>
__device__ __forceinline__ T vel_at(const pp_domain &dom, const pp_block_desc &b, int f, int k, const T *glob,
                                    int env) {
    if (glob) return glob[(size_t)k * dom.num_cells + b.global_offset + f];
    return reinterpret_cast<const T *>(b.velocity)[((size_t)env * dom.dims + k) * b.cells + f];
}

template <typename T, int D>
__device__ __forceinline__ T contra(const pp_domain &dom, const pp_block_desc &b, int f, int w, const T *glob, int env) {
    if (!b.has_transform) return vel_at<T>(dom, b, f, w, glob, env);
    if (b.transform_diagonal) {
        const TView<T> t = Metric<T, D>::at(b, f);
        return t[2 * D * D] * (t[D * D + w * D + w] * vel_at<T>(dom, b, f, w, glob, env));
    }
    T v[D];
#pragma unroll
    for (int k = 0; k < D; ++k) v[k] = vel_at<T>(dom, b, f, k, glob, env);
    return Metric<T, D>::contra_t(Metric<T, D>::at(b, f), w, v);
}

// All D contravariant components of cell f (one read of the velocity and the metric).
template <typename T, int D>
__device__ __forceinline__ void contra_all(const pp_domain &dom, const pp_block_desc &b, int f, const T *glob, int env,
                                           T *out) {
    T v[D];
#pragma unroll
    for (int k = 0; k < D; ++k) v[k] = vel_at<T>(dom, b, f, k, glob, env);
    if (!b.has_transform) {
#pragma unroll
        for (int k = 0; k < D; ++k) out[k] = v[k];
        return;
    }
    const TView<T> t = Metric<T, D>::at(b, f);
    if (b.transform_diagonal) {
        const T det = t[2 * D * D];
#pragma unroll
        for (int w = 0; w < D; ++w) out[w] = det * (t[D * D + w * D + w] * v[w]);
        return;
    }
#pragma unroll
    for (int w = 0; w < D; ++w) out[w] = Metric<T, D>::contra_t(t, w, v);
}

// FIXED boundary velocity component k at face cell ff (getFixedBoundaryData, block_data.cuh:123-196)
template <typename T>
__device__ __forceinline__ T fixed_vel(const pp_boundary_desc &fb, int d, int fc, int ff, int k, int env) {
    const T *data = reinterpret_cast<const T *>(fb.vel);
    const int e = fb.vel_batched ? env : 0;
    if (fb.vel_static) return data[e * d + k];
    return data[((size_t)e * d + k) * fc + ff];
}
// contravariant component of the boundary velocity (velocity.cuh:245-265)
template <typename T, int D>
__device__ __forceinline__ T contra_fixed(const pp_boundary_desc &fb, int fc, int ff, int w, int env) {
    if (!fb.has_transform) return fixed_vel<T>(fb, D, fc, ff, w, env);
    T v[D];
#pragma unroll
    for (int k = 0; k < D; ++k) v[k] = fb.vel_type == PP_BC_DIRICHLET ? fixed_vel<T>(fb, D, fc, ff, k, env) : T(0);
    return Metric<T, D>::contra_t(Metric<T, D>::face(fb, ff), w, v);
}
template <typename T, int D>
__device__ __forceinline__ T alpha_fixed(const pp_boundary_desc &fb, int ff, int w) {
    if (!fb.has_transform) return T(1);
    return Metric<T, D>::alpha_t(Metric<T, D>::face(fb, ff), w);
}
template <typename T>
__device__ __forceinline__ T vsource(const pp_domain &dom, const pp_block_desc &b, int f, int k, int env) {
    if (b.vsource_mode == 0) return T(0);
    const T *s = reinterpret_cast<const T *>(b.vsource);
    const int e = b.vsource_batched ? env : 0;
    if (b.vsource_mode == 1) return s[e * dom.dims + k];
    return s[((size_t)e * dom.dims + k) * b.cells + f];
}
__device__ __forceinline__ int pot_bc(const pp_block_desc &b, int bound, const int *p) {
    const pp_boundary_desc &fb = b.bnd[bound];
    if (!fb.pot_types) return fb.pot_type;
    return reinterpret_cast<const int8_t *>(fb.pot_types)[face_flat(p, bound, b)];
}
template <typename T>
__device__ __forceinline__ T pot_value(const pp_block_desc &b, int bound, const int *p, int env) {
    const pp_boundary_desc &fb = b.bnd[bound];
    if (!fb.pot_values) return T(0);
    const size_t e = fb.pot_values_batched ? (size_t)env * face_cells(b, bound) : 0;
    return reinterpret_cast<const T *>(fb.pot_values)[e + face_flat(p, bound, b)];
}
__device__ __forceinline__ bool is_epot_dirichlet(const pp_block_desc &b, int bound, const int *p) {
    return b.bnd[bound].type == PP_BOUND_FIXED && pot_bc(b, bound, p) == PP_POT_DIRICHLET;
}
__device__ __forceinline__ bool is_insulating_wall(const pp_block_desc &b, int bound, const int *p) {
    if (b.bnd[bound].type != PP_BOUND_FIXED) return false;
    const int t = pot_bc(b, bound, p);
    return t == PP_POT_INSULATING || t == PP_POT_THIN_WALL;
}
__device__ __forceinline__ bool is_thin_wall(const pp_block_desc &b, int bound, const int *p) {
    return b.bnd[bound].type == PP_BOUND_FIXED && pot_bc(b, bound, p) == PP_POT_THIN_WALL;
}

} // namespace pp

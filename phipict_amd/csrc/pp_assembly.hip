// PISO assembly kernels for gfx950: one work-item per (cell, environment), the grid's y
// dimension enumerates batched environments. Each kernel restates one reference kernel
// (cited at its definition) for the CENTRAL advection scheme; the non-orthogonal corner terms
// of skewed grids are added where the flags ask for them.
//
// Matrix kernels write the values of rows whose pattern (row offsets, ascending columns)
// is fixed per domain; k_csr_pattern writes that pattern once per domain.
#include "pp_common.hpp"
#include "pp_nonortho.hpp"

#include <vector>

namespace pp {

constexpr int ASM_BLOCK = 256;

// Face fluxes of cell f (not multiplied by the face sign): computeFluxesNDLoop,
// discretization.cuh:758-852; with EPOT the prescribed faces follow
// computeEpotFluxesNDLoop :856-912 (insulating wall 0, open face velC).
template <typename T, int D, bool EPOT>
__device__ __forceinline__ void cell_fluxes(const pp_domain &dom, const pp_block_desc *__restrict__ blocks, int bi, const int *p,
                            int f, const T *glob, int env, T *fl) {
    const pp_block_desc &b = blocks[bi];
    T velCd[D];
    contra_all<T, D>(dom, b, f, glob, env, velCd);
#pragma unroll
    for (int bound = 0; bound < 2 * D; ++bound) {
        const int dim = bound >> 1;
        const int up = bound & 1;
        const T velC = velCd[dim];
        if (at_bound(p, bound, b)) {
            const pp_boundary_desc &bd = b.bnd[bound];
            if (bd.type == PP_BOUND_FIXED) {
                if (EPOT) {
                    fl[bound] = is_insulating_wall(b, bound, p) ? T(0) : velC;
                } else {
                    fl[bound] = contra_fixed<T, D>(bd, face_cells(b, bound), face_flat(p, bound, b), dim, env);
                }
            } else if (bd.type == PP_BOUND_CONNECTED) {
                int np[3];
                conn_pos(blocks, D, p, dim, bd, 0, np);
                const pp_block_desc &nb = blocks[bd.conn_block];
                T velN = contra<T, D>(dom, nb, flat3(np, nb.size), bd.conn_axes[0] >> 1, glob, env);
                if ((bd.conn_axes[0] & 1) == up) velN = -velN;
                fl[bound] = (velN + velC) * T(0.5f);
            } else if (bd.type == PP_BOUND_PERIODIC) {
                int np[3] = {p[0], p[1], p[2]};
                np[dim] = up ? 0 : b.size[dim] - 1;
                const T velN = contra<T, D>(dom, b, flat3(np, b.size), dim, glob, env);
                fl[bound] = (velN + velC) * T(0.5f);
            } else {
                fl[bound] = T(0);
            }
        } else {
            int np[3] = {p[0], p[1], p[2]};
            np[dim] += face_sign(bound);
            const T velN = contra<T, D>(dom, b, flat3(np, b.size), dim, glob, env);
            fl[bound] = (velN + velC) * T(0.5f);
        }
    }
}

// Emit the 2D+1 candidate entries of a row in ascending column order, first occurrence
// first among equal columns (advection.cu:286-300). Returns the number of entries. With
// ell_out the values are also written to the ELL copy (entry k at ell_out[k * n]), the row
// padded with zeros to 2D+1 entries.
template <typename T, int D>
__device__ __forceinline__ int emit_row(int *idx, T *vals, bool divide, T det, int32_t *col_out, T *val_out,
                                        T *ell_out = nullptr, int n = 0) {
    int written = 0;
#pragma unroll
    for (int it = 0; it < 2 * D + 1; ++it) {
        int best = -1, bv = 0x7fffffff;
#pragma unroll
        for (int i = 0; i < 2 * D + 1; ++i)
            if (idx[i] >= 0 && idx[i] < bv) { bv = idx[i]; best = i; }
        if (best < 0) break;
        T v = vals[best];
        if (divide) { v = v / det; vals[best] = v; }
        if (col_out) col_out[written] = idx[best];
        val_out[written] = v;
        if (ell_out) ell_out[(size_t)written * n] = v;
        idx[best] = -1;
        ++written;
    }
    if (ell_out)
        for (int k = written; k < 2 * D + 1; ++k) ell_out[(size_t)k * n] = T(0);
    return written;
}

// hasFarUpwindCell, shared_stencils.cuh:35-43: the face's neighbour and the far-upwind cell
// (one further against the flow along the face normal) both lie inside the block.
__device__ __forceinline__ bool far_upwind(const int *p, int bound, bool outflow, const pp_block_desc &b) {
    const int axis = bound >> 1, fs = face_sign(bound), last = b.size[axis] - 1;
    const int iD = p[axis] + fs;
    if (iD < 0 || iD > last) return false;
    const int iUU = outflow ? p[axis] - fs : p[axis] + 2 * fs;
    return iUU >= 0 && iUU <= last;
}

// LINEAR_UPWIND's deferred correction of one field component at cell p (advection.cu:893-937
// for the velocity, :650-684 for the passive scalar): the matrix holds first-order upwind, so
// -F 0.5 psi (phi_D - phi_U) with psi = r = (phi_U - phi_UU) / (phi_D - phi_U) is subtracted
// for every face whose far-upwind cell exists, skipping flat faces. fl: the cell's face fluxes
// (before the face sign), phi: the component over the block.
template <typename T, int D>
__device__ __forceinline__ void deferred_correction(const pp_block_desc &b, const int *p, const T *fl,
                                                    const T *__restrict__ phi, T &t) {
#pragma unroll
    for (int bound = 0; bound < 2 * D; ++bound) {
        const int axis = bound >> 1, fs = face_sign(bound);
        const T ffo = T(fs) * fl[bound];
        const bool outflow = ffo >= T(0);
        if (!far_upwind(p, bound, outflow, b)) continue;
        int q[3] = {p[0], p[1], p[2]};
        const T phiP = phi[flat3(q, b.size)];
        q[axis] = p[axis] + fs;
        const T phiN = phi[flat3(q, b.size)];
        q[axis] = outflow ? p[axis] - fs : p[axis] + 2 * fs;
        const T phiUU = phi[flat3(q, b.size)];
        const T phiU = outflow ? phiP : phiN;
        const T phiD = outflow ? phiN : phiP;
        const T dD = phiD - phiU;
        if ((dD < T(0) ? -dD : dD) < T(1e-20)) continue;
        const T psi = (phiU - phiUU) / dD;
        t -= ffo * T(0.5) * psi * dD;
    }
}

#ifndef PP_ASM_XCD_SPAN
#define PP_ASM_XCD_SPAN 0
#endif
constexpr int ASM_XCDS = 8;

// Workgroup index in cell order. The dispatcher deals workgroups to the XCDs in turn
// (workgroup w to XCD w mod 8); with PP_ASM_XCD_SPAN each XCD instead takes one contiguous
// span of tiles, so the neighbour rows and planes a stencil reads were mostly fetched into
// the same XCD's L2 by its own earlier workgroups.
__device__ __forceinline__ int asm_tile() {
    if (!PP_ASM_XCD_SPAN) return blockIdx.x;
    const int n = gridDim.x, x = blockIdx.x % ASM_XCDS, slot = blockIdx.x / ASM_XCDS;
    const int per = n / ASM_XCDS, extra = n % ASM_XCDS;
    return x * per + (x < extra ? x : extra) + slot;
}

// Workgroups never straddle blocks: tile = block * wpb + tile in block, so the block index and
// its descriptor are uniform over the workgroup (scalar loads).
#define CELL_PROLOGUE                                                          \
    const int wpb_ = (dom.max_block_cells + ASM_BLOCK - 1) / ASM_BLOCK;        \
    const int tile_ = asm_tile();                                              \
    const int bi = __builtin_amdgcn_readfirstlane(tile_ / wpb_);               \
    const int f = (tile_ - bi * wpb_) * ASM_BLOCK + threadIdx.x;               \
    const pp_block_desc *__restrict__ blocks = dom.blocks;                     \
    const pp_block_desc &b = blocks[bi];                                       \
    if (f >= b.cells) return;                                                  \
    const int g = b.global_offset + f;                                         \
    const int env = blockIdx.y;                                                \
    int p[3];                                                                  \
    unflat(f, b.size, p);

// CSR pattern: row offsets and ascending neighbour columns of every row (shared by the
// advection, pressure and potential matrices; row layout of topology.cuh:168-233).
template <int D>
__global__ void __launch_bounds__(ASM_BLOCK) k_csr_pattern(pp_domain dom, int32_t *row, int32_t *col) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= dom.num_cells) return;
    const pp_block_desc *__restrict__ blocks = dom.blocks;
    const int bi = find_block(blocks, dom.num_blocks, g);
    const pp_block_desc &b = blocks[bi];
    const int f = g - b.global_offset;
    int p[3];
    unflat(f, b.size, p);
    // row start: block csr offset + sum of the row sizes of the block's earlier cells,
    // in closed form (each prescribed face removes one entry per face cell before f)
    int start = f * (2 * D + 1);
    for (int bound = 0; bound < 2 * D; ++bound) {
        if (!is_fixed(b, bound)) continue;
        const int a = bound >> 1;
        const int up = bound & 1;
        // cells before f whose coordinate along a equals the face coordinate
        const int stride_a = a == 0 ? 1 : (a == 1 ? b.size[0] : b.size[0] * b.size[1]);
        const int outer = f / (stride_a * b.size[a]);                 // full slabs before
        const int inner = f % stride_a;                               // position inside the slab
        const int ca = p[a];
        const int target = up ? b.size[a] - 1 : 0;
        int cnt = outer * stride_a;
        if (ca > target) cnt += stride_a;
        else if (ca == target) cnt += inner;
        start -= cnt;
    }
    start += b.csr_offset;
    int idx[2 * D + 1];
    float dummy[2 * D + 1];
    idx[0] = g;
    for (int bound = 0; bound < 2 * D; ++bound) {
        if (at_bound(p, bound, b) && is_fixed(b, bound)) { idx[bound + 1] = -1; continue; }
        int np[3], nw;
        const int nbi = neighbor(blocks, bi, D, p, bound, np, &nw);
        idx[bound + 1] = flat3(np, blocks[nbi].size) + blocks[nbi].global_offset;
    }
    for (int i = 0; i < 2 * D + 1; ++i) dummy[i] = 0.f;
    const int w = emit_row<float, D>(idx, dummy, false, 1.f, col + start, dummy);
    if (g == 0) row[0] = 0;
    row[g + 1] = start + w;
}

// PISO_build_matrix, advection.cu:26-306: values of C and the diagonal A. With sc >= 0 the
// matrix of passive-scalar channel sc (forPassiveScalar): the channel's diffusivity and its
// boundary condition at the prescribed faces. LINEAR_UPWIND puts first-order upwind on the
// faces whose far-upwind cell lies in the block (advection.cu:129-140), CENTRAL elsewhere.
template <typename T, int D, bool NO>
__global__ void __launch_bounds__(ASM_BLOCK) k_advection_matrix(pp_domain dom, int sc, int flags) {
    CELL_PROLOGUE
    const T dt = reinterpret_cast<const T *>(dom.time_step)[env];
    T fl[2 * D];
    cell_fluxes<T, D, false>(dom, blocks, bi, p, f, nullptr, env, fl);
    int idx[2 * D + 1];
    T rv[2 * D + 1];
#pragma unroll
    for (int i = 0; i < 2 * D + 1; ++i) rv[i] = T(0);
    const T det = Metric<T, D>::det(b, f);
    T diag = det / dt;
    idx[0] = g;
    const T vp = sc < 0 ? visc_at<T>(dom, b, f) : scalar_visc<T>(dom, sc);
    T ap[D];
#pragma unroll
    for (int k = 0; k < D; ++k) ap[k] = Metric<T, D>::alpha(b, f, k);
#pragma unroll
    for (int bound = 0; bound < 2 * D; ++bound) {
        const int dim = bound >> 1;
        if (!at_bound(p, bound, b) || !is_fixed(b, bound)) {
            const T ffo = T(face_sign(bound)) * fl[bound];
            if (dom.advection_scheme == PP_ADV_LINEAR_UPWIND && far_upwind(p, bound, ffo >= T(0), b)) {
                if (ffo >= T(0)) diag += ffo;
                else rv[bound + 1] += ffo;
            } else {
                const T ff = T(0.5) * ffo;
                diag += ff;
                rv[bound + 1] += ff;
            }
            int np[3], nw;
            const int nbi = neighbor(blocks, bi, D, p, bound, np, &nw);
            const pp_block_desc &nb = blocks[nbi];
            const int nf = flat3(np, nb.size);
            const T an = Metric<T, D>::alpha(nb, nf, nw);
            const T vn = sc < 0 ? visc_at<T>(dom, nb, nf) : vp;
            idx[bound + 1] = nf + nb.global_offset;
            const T vc = (ap[dim] * vp + an * vn) * T(0.5);
            diag += vc;
            rv[bound + 1] -= vc;
            if (NO && flags) {
                if (sc < 0) {
                    const CornerData<T, D, NK_VELOCITY_TYPES> cd{dom, blocks, 0, env};
                    matrix_corner_terms<T, D, NK_VELOCITY_TYPES>(cd, bi, p, f, bound, flags, vp, true, nullptr, T(1),
                                                                  false, T(1), diag, rv);
                } else {
                    // scalar: coefficients without viscosity, times the channel's diffusivity
                    const CornerData<T, D, NK_SCALAR_TYPES> cd{dom, blocks, sc, env};
                    matrix_corner_terms<T, D, NK_SCALAR_TYPES>(cd, bi, p, f, bound, flags, T(1), false, nullptr, vp,
                                                                false, T(1), diag, rv);
                }
            }
        } else {
            const int slip = sc < 0 ? (b.bnd[bound].vel_type == PP_BC_DIRICHLET ? 0 : 1)
                                    : (scalar_neumann(b.bnd[bound], sc) ? 1 : 0);
            diag += T(1 - slip) * 2 * vp * ap[dim];
            idx[bound + 1] = -1;
        }
    }
    rv[0] = diag;
    const int32_t start = dom.C.row[g];
    T *val = reinterpret_cast<T *>(dom.C.val) + (size_t)env * dom.C.nnz + start;
    T *ell = dom.C.ell_val ? reinterpret_cast<T *>(dom.C.ell_val) + ((size_t)env * (2 * D + 1)) * dom.num_cells + g
                           : nullptr;
    emit_row<T, D>(idx, rv, true, det, nullptr, val, ell, dom.num_cells);
    reinterpret_cast<T *>(dom.A)[(size_t)env * dom.num_cells + g] = rv[0];
}

// Pressure at a position that may lie one cell outside the block (getPressureAtWithBounds,
// nonortho_laplace.cuh:302-378).
template <typename T>
__device__ T pressure_at(const pp_domain &dom, const pp_block_desc *__restrict__ blocks, int bi, const int *pin,
                         int env) {
    const pp_block_desc &b = blocks[bi];
    int t[3] = {pin[0], pin[1], pin[2]};
    const int D = dom.dims;
    for (int dim = 0; dim < D; ++dim) {
        const bool lower = pin[dim] < 0;
        const bool upper = pin[dim] >= b.size[dim];
        if (!lower && !upper) continue;
        const pp_boundary_desc &bd = b.bnd[dim * 2 + (upper ? 1 : 0)];
        if (bd.type == PP_BOUND_FIXED) {
            t[dim] = lower ? 0 : b.size[dim] - 1;
            return reinterpret_cast<const T *>(b.pressure)[(size_t)env * b.cells + flat3(t, b.size)];
        }
        if (bd.type == PP_BOUND_CONNECTED) {
            int op[3];
            conn_pos(blocks, D, t, dim, bd, lower ? -(pin[dim] + 1) : pin[dim] - b.size[dim], op);
            const pp_block_desc &nb = blocks[bd.conn_block];
            return reinterpret_cast<const T *>(nb.pressure)[(size_t)env * nb.cells + flat3(op, nb.size)];
        }
        if (bd.type == PP_BOUND_PERIODIC) {
            t[dim] += lower ? b.size[dim] : -b.size[dim];
            return reinterpret_cast<const T *>(b.pressure)[(size_t)env * b.cells + flat3(t, b.size)];
        }
        return T(0);
    }
    return reinterpret_cast<const T *>(b.pressure)[(size_t)env * b.cells + flat3(pin, b.size)];
}

// getPressureGradient, velocity.cuh:347-382 (FD, transformed with Minv^T)
template <typename T, int D>
__device__ void pressure_grad(const pp_domain &dom, const pp_block_desc *__restrict__ blocks, int bi, const int *p,
                              int f, int env, T *g) {
    const pp_block_desc &b = blocks[bi];
    T gr[D];
#pragma unroll
    for (int dim = 0; dim < D; ++dim) {
        int t[3] = {p[0], p[1], p[2]};
        T fac = T(0.5);
        if (!(p[dim] == 0 && is_fixed(b, dim * 2))) t[dim] = p[dim] - 1;
        else fac = T(1);
        const T vn = pressure_at<T>(dom, blocks, bi, t, env);
        if (!(p[dim] == b.size[dim] - 1 && is_fixed(b, dim * 2 + 1))) t[dim] = p[dim] + 1;
        else { t[dim] = p[dim]; fac = T(1); }
        const T vp = pressure_at<T>(dom, blocks, bi, t, env);
        gr[dim] = (vp - vn) * fac;
    }
    if (b.has_transform && b.transform_diagonal) {
        const TView<T> t = Metric<T, D>::at(b, f);
#pragma unroll
        for (int j = 0; j < D; ++j) gr[j] = t[D * D + j * D + j] * gr[j];
    } else if (b.has_transform) {
        const TView<T> t = Metric<T, D>::at(b, f);
        T r[D];
#pragma unroll
        for (int j = 0; j < D; ++j) r[j] = T(0);
#pragma unroll
        for (int i = 0; i < D; ++i)
#pragma unroll
            for (int j = 0; j < D; ++j) r[j] += t[D * D + i * D + j] * gr[i];
#pragma unroll
        for (int j = 0; j < D; ++j) gr[j] = r[j];
    }
#pragma unroll
    for (int j = 0; j < D; ++j) g[j] = gr[j];
}

// Boundary source of the momentum equation at cell f for component dim: advective inflow
// and wall viscous terms of prescribed faces (advection.cu:831-853, pressure.cu:410-432).
template <typename T, int D>
__device__ __forceinline__ T boundary_source(const pp_domain &dom, const pp_block_desc &b, const int *p, int f, int dim,
                                             int env) {
    T s = T(0);
#pragma unroll
    for (int bound = 0; bound < 2 * D; ++bound) {
        if (!at_bound(p, bound, b)) continue;
        const pp_boundary_desc &fb = b.bnd[bound];
        if (fb.type != PP_BOUND_FIXED) continue;
        const T fn = T((bound & 1) * 2 - 1);
        const int fc = face_cells(b, bound), ff = face_flat(p, bound, b);
        const T vel = fixed_vel<T>(fb, D, fc, ff, dim, env);
        const T alpha = alpha_fixed<T, D>(fb, ff, bound >> 1);
        const T flux = contra_fixed<T, D>(fb, fc, ff, bound >> 1, env) * fn;
        s -= vel * flux;
        const T bv = visc_at<T>(dom, b, f);
        const T slip = fb.vel_type == PP_BC_DIRICHLET ? T(0) : T(1);
        s += vel * (1 - slip) * bv * 2 * alpha;
    }
    return s;
}

// kPISO_build_advection_RHS, advection.cu:799-955 (with LINEAR_UPWIND the deferred correction
// of the block velocity, lagged to the start of the step)
template <typename T, int D, bool NO>
__global__ void __launch_bounds__(ASM_BLOCK) k_advection_rhs(pp_domain dom, int apply_pg, int flags) {
    CELL_PROLOGUE
    const T dt = reinterpret_cast<const T *>(dom.time_step)[env];
    const T det = Metric<T, D>::det(b, f);
    const T *vel = reinterpret_cast<const T *>(b.velocity) + (size_t)env * D * b.cells;
    const bool deferred = dom.advection_scheme == PP_ADV_LINEAR_UPWIND;
    T fl[2 * D];
    if (deferred) cell_fluxes<T, D, false>(dom, blocks, bi, p, f, nullptr, env, fl);
    T *rhs = reinterpret_cast<T *>(dom.velocity_rhs) + (size_t)env * D * dom.num_cells;
    T pg[D];
    if (apply_pg) pressure_grad<T, D>(dom, blocks, bi, p, f, env, pg);
    // explicit corner terms of every component (one walk per corner)
    T Sv[D];
    if (NO && flags) {
        const CornerData<T, D, NK_VELOCITY> cd{dom, blocks, 0, env};
        nonortho_rhs<T, D, NK_VELOCITY>(cd, bi, p, f, flags, true, nullptr, Sv);
    }
#pragma unroll
    for (int dim = 0; dim < D; ++dim) {
        T t = det * vel[(size_t)dim * b.cells + f] / dt;
        // the reference accumulates the boundary terms one face at a time on t
#pragma unroll
        for (int bound = 0; bound < 2 * D; ++bound) {
            if (!at_bound(p, bound, b)) continue;
            const pp_boundary_desc &fb = b.bnd[bound];
            if (fb.type != PP_BOUND_FIXED) continue;
            const T fn = T((bound & 1) * 2 - 1);
            const int fc = face_cells(b, bound), ff = face_flat(p, bound, b);
            const T v = fixed_vel<T>(fb, D, fc, ff, dim, env);
            const T alpha = alpha_fixed<T, D>(fb, ff, bound >> 1);
            const T flux = contra_fixed<T, D>(fb, fc, ff, bound >> 1, env) * fn;
            t -= v * flux;
            const T bv = visc_at<T>(dom, b, f);
            const T slip = fb.vel_type == PP_BC_DIRICHLET ? T(0) : T(1);
            t += v * (1 - slip) * bv * 2 * alpha;
        }
        if (deferred) deferred_correction<T, D>(b, p, fl, vel + (size_t)dim * b.cells, t);
        if (NO && flags) t -= Sv[dim];
        t /= det;
        t += vsource<T>(dom, b, f, dim, env);
        if (apply_pg) t -= pg[dim] * dt;
        rhs[(size_t)dim * dom.num_cells + g] = t;
    }
}

// kPISO_build_scalar_advection_RHS, advection.cu:544-700: every passive-scalar channel;
// prescribed faces give the advected face value and the diffusive flux (Dirichlet: the face
// value with the ghost-cell coefficient 2 alpha; Neumann: the cell value advected and the
// prescribed normal gradient times the face area); with LINEAR_UPWIND the deferred correction
template <typename T, int D, bool NO>
__global__ void __launch_bounds__(ASM_BLOCK) k_scalar_rhs(pp_domain dom, int flags) {
    CELL_PROLOGUE
    const int C = dom.scalar_channels;
    const T dt = reinterpret_cast<const T *>(dom.time_step)[env];
    const T det = Metric<T, D>::det(b, f);
    const T *phi = reinterpret_cast<const T *>(b.scalar) + (size_t)env * C * b.cells;
    T *rhs = reinterpret_cast<T *>(dom.scalar_rhs) + (size_t)env * C * dom.num_cells;
    const bool deferred = dom.advection_scheme == PP_ADV_LINEAR_UPWIND;
    T fl[2 * D];
    if (deferred) cell_fluxes<T, D, false>(dom, blocks, bi, p, f, nullptr, env, fl);
    for (int ch = 0; ch < C; ++ch) {
        const T phiP = phi[(size_t)ch * b.cells + f];
        T t = det * phiP / dt;
#pragma unroll
        for (int bound = 0; bound < 2 * D; ++bound) {
            if (!at_bound(p, bound, b) || !is_fixed(b, bound)) continue;
            const pp_boundary_desc &fb = b.bnd[bound];
            const T fn = T((bound & 1) * 2 - 1);
            const int fc = face_cells(b, bound), ff = face_flat(p, bound, b), axis = bound >> 1;
            const T s = fixed_scalar<T>(fb, C, fc, ff, ch, env);
            const T flux = contra_fixed<T, D>(fb, fc, ff, axis, env) * fn;
            const T visc = scalar_visc<T>(dom, ch);
            if (!scalar_neumann(fb, ch)) {
                t -= s * flux;
                t += s * visc * 2 * alpha_fixed<T, D>(fb, ff, axis);
            } else {
                t -= phiP * flux;
                T area = T(1);
                if (fb.has_transform) {
                    const TView<T> tf = Metric<T, D>::face(fb, ff);
                    T r = T(0);
#pragma unroll
                    for (int k = 0; k < D; ++k) r += tf[D * D + axis * D + k] * tf[D * D + axis * D + k];
                    area = tf[2 * D * D] * dsqrt<T>(r);
                }
                t += s * visc * area;
            }
        }
        if (deferred) deferred_correction<T, D>(b, p, fl, phi + (size_t)ch * b.cells, t);
        if (NO && flags) {
            const CornerData<T, D, NK_SCALAR> cd{dom, blocks, ch, env};
            T S[1];
            nonortho_rhs<T, D, NK_SCALAR>(cd, bi, p, f, flags, false, nullptr, S);
            t -= S[0] * scalar_visc<T>(dom, ch);
        }
        rhs[(size_t)ch * dom.num_cells + g] = t / det;
    }
}

// PISO_build_pressure_matrix, pressure.cu:25-193 (useFaceTransform = false)
template <typename T, int D, bool NO>
__global__ void __launch_bounds__(ASM_BLOCK) k_pressure_matrix(pp_domain dom, int flags) {
    CELL_PROLOGUE
    const T *A = reinterpret_cast<const T *>(dom.A) + (size_t)env * dom.num_cells;
    int idx[2 * D + 1];
    T rv[2 * D + 1];
#pragma unroll
    for (int i = 0; i < 2 * D + 1; ++i) rv[i] = T(0);
    T ap[D];
#pragma unroll
    for (int k = 0; k < D; ++k) ap[k] = Metric<T, D>::alpha(b, f, k);
    const T raP = T(1.0) / A[g];
    idx[0] = g;
#pragma unroll
    for (int bound = 0; bound < 2 * D; ++bound) {
        const int dim = bound >> 1;
        if (!(at_bound(p, bound, b) && is_fixed(b, bound))) {
            int np[3], nw;
            const int nbi = neighbor(blocks, bi, D, p, bound, np, &nw);
            const pp_block_desc &nb = blocks[nbi];
            const int nf = flat3(np, nb.size);
            const T an = Metric<T, D>::alpha(nb, nf, nw);
            const int ng = nf + nb.global_offset;
            const T raN = T(1.0) / A[ng];
            const T coef = T(0.5) * (ap[dim] * raP + an * raN);
            rv[0] -= coef;
            rv[bound + 1] += coef;
            idx[bound + 1] = ng;
            if (NO && flags) {
                const CornerData<T, D, NK_FIXED_ONLY> cd{dom, blocks, 0, env};
                matrix_corner_terms<T, D, NK_FIXED_ONLY>(cd, bi, p, f, bound, flags, raP, false, A, T(1), true, T(-1),
                                                         rv[0], rv);
            }
        } else {
            idx[bound + 1] = -1;
        }
    }
    const int32_t start = dom.P.row[g];
    T *val = reinterpret_cast<T *>(dom.P.val) + (size_t)env * dom.P.nnz + start;
    T *ell = dom.P.ell_val ? reinterpret_cast<T *>(dom.P.ell_val) + ((size_t)env * (2 * D + 1)) * dom.num_cells + g
                           : nullptr;
    emit_row<T, D>(idx, rv, false, T(1), nullptr, val, ell, dom.num_cells);
}

// PISO_build_pressure_rhs, pressure.cu:349-522 (CENTRAL, boundary sources on)
template <typename T, int D>
__global__ void __launch_bounds__(ASM_BLOCK) k_pressure_rhs(pp_domain dom) {
    CELL_PROLOGUE
    const T dt = reinterpret_cast<const T *>(dom.time_step)[env];
    const T rDiag = 1 / reinterpret_cast<const T *>(dom.A)[(size_t)env * dom.num_cells + g];
    const T det = Metric<T, D>::det(b, f);
    const T *vres = reinterpret_cast<const T *>(dom.velocity_result) + (size_t)env * D * dom.num_cells;
    const T *vel = reinterpret_cast<const T *>(b.velocity) + (size_t)env * D * b.cells;
    T *prhs = reinterpret_cast<T *>(dom.pressure_rhs) + (size_t)env * D * dom.num_cells;
    // off-diagonal part of C times the predicted velocity, all components in one pass over the row
    T Hs[D];
#pragma unroll
    for (int dim = 0; dim < D; ++dim) Hs[dim] = T(0);
    if (dom.C.ell_val) {
        const int n = dom.num_cells;
        const T *ev = reinterpret_cast<const T *>(dom.C.ell_val) + (size_t)env * (2 * D + 1) * n + g;
        const int32_t *ec = dom.C.ell_col + g;
#pragma unroll
        for (int k = 0; k < 2 * D + 1; ++k) {
            const int ix = ec[(size_t)k * n];
            const T cv = ev[(size_t)k * n];
            if (ix != g)
#pragma unroll
                for (int dim = 0; dim < D; ++dim) Hs[dim] += cv * vres[ix + (size_t)n * dim];
        }
    } else {
        const int32_t cs = dom.C.row[g];
        const int rs = dom.C.row[g + 1] - cs;
        const T *cval = reinterpret_cast<const T *>(dom.C.val) + (size_t)env * dom.C.nnz + cs;
        const int32_t *ccol = dom.C.col + cs;
        for (int i = 0; i < rs && i < 2 * D + 1; ++i) {
            const int ix = ccol[i];
            if (ix != g)
#pragma unroll
                for (int dim = 0; dim < D; ++dim) Hs[dim] += cval[i] * vres[ix + (size_t)dom.num_cells * dim];
        }
    }
#pragma unroll
    for (int dim = 0; dim < D; ++dim) {
        const T velOld = vel[(size_t)dim * b.cells + f] / dt;
        const T H = Hs[dim];
        T S = boundary_source<T, D>(dom, b, p, f, dim, env);
        S /= det;
        S += vsource<T>(dom, b, f, dim, env);
        prhs[g + (size_t)dom.num_cells * dim] = rDiag * (velOld - H + S);
    }
}

// k_computePressureRHSdivergenceFromFlux, pressure.cu:755-803 (useFaceTransform = false)
template <typename T, int D, bool NO>
// With the flags' right-hand-side bits it also adds the pressure's explicit corner terms
// (k_pressureRHSaddNonOrthoComponents, pressure.cu:838-860).
__global__ void __launch_bounds__(ASM_BLOCK) k_pressure_div(pp_domain dom, int time_step_norm, int flags) {
    CELL_PROLOGUE
    const T *prhs = reinterpret_cast<const T *>(dom.pressure_rhs) + (size_t)env * D * dom.num_cells;
    T fl[2 * D];
    cell_fluxes<T, D, false>(dom, blocks, bi, p, f, prhs, env, fl);
    T dv = T(0);
#pragma unroll
    for (int dim = 0; dim < D; ++dim) dv += fl[dim * 2 + 1] - fl[dim * 2];
    if (time_step_norm) dv /= reinterpret_cast<const T *>(dom.time_step)[env];
    if (NO && (flags & (NO_DIRECT_RHS | NO_DIAGONAL_RHS))) {
        const CornerData<T, D, NK_PRESSURE> cd{dom, blocks, 0, env};
        T S[1];
        nonortho_rhs<T, D, NK_PRESSURE>(cd, bi, p, f, flags, false,
                                        reinterpret_cast<const T *>(dom.A) + (size_t)env * dom.num_cells, S);
        dv += S[0];
    }
    reinterpret_cast<T *>(dom.pressure_rhs_div)[(size_t)env * dom.num_cells + g] = dv;
}

// PISO_update_velocity, velocity_correction.cu:26-60
template <typename T, int D>
__global__ void __launch_bounds__(ASM_BLOCK) k_correct_velocity(pp_domain dom, int time_step_norm) {
    CELL_PROLOGUE
    const T rDiag = 1 / reinterpret_cast<const T *>(dom.A)[(size_t)env * dom.num_cells + g];
    T gr[D];
    pressure_grad<T, D>(dom, blocks, bi, p, f, env, gr);
    const T dt = reinterpret_cast<const T *>(dom.time_step)[env];
    const T *prhs = reinterpret_cast<const T *>(dom.pressure_rhs) + (size_t)env * D * dom.num_cells;
    T *vres = reinterpret_cast<T *>(dom.velocity_result) + (size_t)env * D * dom.num_cells;
#pragma unroll
    for (int dim = 0; dim < D; ++dim) {
        T u = -rDiag * gr[dim];
        if (time_step_norm) u *= dt;
        const size_t o = g + (size_t)dom.num_cells * dim;
        u += prhs[o];
        vres[o] = u;
    }
}

// PISO_build_epot_matrix, potential.cu:29-233 (useFaceTransform = false)
template <typename T, int D>
__global__ void __launch_bounds__(ASM_BLOCK) k_epot_matrix(pp_domain dom, int flags) {
    CELL_PROLOGUE
    if (env != 0) return; // one matrix shared by all environments
    // the corner terms weigh both cells with 1/A, the advection diagonal of the first
    // environment, although the orthogonal part uses 1 (interpolateNonOrthoLaplaceComponents
    // with withA set, potential.cu:103-150)
    const T *A = reinterpret_cast<const T *>(dom.A);
    int idx[2 * D + 1];
    T rv[2 * D + 1];
#pragma unroll
    for (int i = 0; i < 2 * D + 1; ++i) rv[i] = T(0);
    T ap[D];
#pragma unroll
    for (int k = 0; k < D; ++k) ap[k] = Metric<T, D>::alpha(b, f, k);
    const T raP = T(1.0);
    idx[0] = g;
#pragma unroll
    for (int bound = 0; bound < 2 * D; ++bound) {
        const int dim = bound >> 1;
        if (!(at_bound(p, bound, b) && is_fixed(b, bound))) {
            int np[3], nw;
            const int nbi = neighbor(blocks, bi, D, p, bound, np, &nw);
            const pp_block_desc &nb = blocks[nbi];
            const int nf = flat3(np, nb.size);
            const T an = Metric<T, D>::alpha(nb, nf, nw);
            const T raN = T(1.0);
            const T coef = T(0.5) * (ap[dim] * raP + an * raN);
            rv[0] -= coef;
            rv[bound + 1] += coef;
            idx[bound + 1] = nf + nb.global_offset;
            if (flags) {
                const CornerData<T, D, NK_FIXED_ONLY> cd{dom, blocks, 0, 0};
                matrix_corner_terms<T, D, NK_FIXED_ONLY>(cd, bi, p, f, bound, flags, T(1.0) / A[g], false, A, T(1),
                                                         true, T(-1), rv[0], rv);
            }
        } else {
            idx[bound + 1] = -1;
            if (is_epot_dirichlet(b, bound, p)) rv[0] -= ap[dim] * raP;
        }
    }
    // thin-wall Robin correction, potential.cu:164-218
    for (int tw = 0; tw < 2 * D; ++tw) {
        const int twDim = tw >> 1;
        if (!at_bound(p, tw, b)) continue;
        if (!is_thin_wall(b, tw, p)) continue;
        const T Cw = T(b.bnd[tw].pot_cw);
        const T mnP = Metric<T, D>::minv_norm(b, f, twDim);
        for (int tDim = 0; tDim < D; ++tDim) {
            if (tDim == twDim) continue;
            for (int ts = 0; ts < 2; ++ts) {
                const int tb = tDim * 2 + ts;
                if (idx[tb + 1] < 0) continue;
                int tn[3] = {p[0], p[1], p[2]};
                if (at_bound(p, tb, b)) {
                    if (b.bnd[tb].type == PP_BOUND_PERIODIC) tn[tDim] = ts == 1 ? 0 : b.size[tDim] - 1;
                    else continue;
                } else {
                    tn[tDim] = p[tDim] + face_sign(tb);
                }
                if (!is_thin_wall(b, tw, tn)) continue;
                const int tf = flat3(tn, b.size);
                const T ant = Metric<T, D>::alpha(b, tf, tDim);
                const T mnN = Metric<T, D>::minv_norm(b, tf, twDim);
                const T sc = Cw * T(0.5) * (ap[tDim] * mnP + ant * mnN);
                rv[0] -= sc;
                rv[tb + 1] += sc;
            }
        }
    }
    const int32_t start = dom.E.row[g];
    T *val = reinterpret_cast<T *>(dom.E.val) + start;
    T *ell = dom.E.ell_val ? reinterpret_cast<T *>(dom.E.ell_val) + g : nullptr;
    emit_row<T, D>(idx, rv, false, T(1), nullptr, val, ell, dom.num_cells);
}

// k_computeEpotRHS, potential.cu:272-322. vf: [batch][d][n] -> div [batch][n]
template <typename T, int D>
__global__ void __launch_bounds__(ASM_BLOCK) k_epot_rhs(pp_domain dom, const T *__restrict__ vf, T *__restrict__ div) {
    CELL_PROLOGUE
    const T *v = vf + (size_t)env * D * dom.num_cells;
    T fl[2 * D];
    cell_fluxes<T, D, true>(dom, blocks, bi, p, f, v, env, fl);
    T dv = T(0);
#pragma unroll
    for (int dim = 0; dim < D; ++dim) dv += fl[dim * 2 + 1] - fl[dim * 2];
#pragma unroll
    for (int bound = 0; bound < 2 * D; ++bound) {
        if (!at_bound(p, bound, b) || !is_fixed(b, bound)) continue;
        if (!is_epot_dirichlet(b, bound, p)) continue;
        const T gv = pot_value<T>(b, bound, p, env);
        if (gv == T(0)) continue;
        dv -= Metric<T, D>::alpha(b, f, bound >> 1) * gv;
    }
    div[(size_t)env * dom.num_cells + g] = dv;
}

// computeCurrentDensityFaceBased, potential.cu:708-885. phi [batch][n], uxb [batch][3][n]
// (the in-plane components are the first D), J [batch][3][n].
template <typename T, int D>
__global__ void __launch_bounds__(ASM_BLOCK) k_current_density(pp_domain dom, const T *__restrict__ phi_all,
                                                             const T *__restrict__ uxb_all, T *__restrict__ J) {
    CELL_PROLOGUE
    const int nc = dom.num_cells;
    const T *phi = phi_all + (size_t)env * nc;
    const T *uxb = uxb_all + (size_t)env * 3 * nc;
    const T phiP = phi[g];
    T ap[D], cvp[D];
#pragma unroll
    for (int k = 0; k < D; ++k) {
        ap[k] = Metric<T, D>::alpha(b, f, k);
        cvp[k] = contra<T, D>(dom, b, f, k, uxb, env);
    }
    T jf[2 * D];
#pragma unroll
    for (int bound = 0; bound < 2 * D; ++bound) {
        const int dim = bound >> 1;
        const int up = bound & 1;
        const T fs = T(face_sign(bound));
        const bool atb = at_bound(p, bound, b);
        if (atb && is_fixed(b, bound)) {
            if (is_insulating_wall(b, bound, p)) jf[bound] = T(0);
            else if (is_epot_dirichlet(b, bound, p)) {
                const T gv = pot_value<T>(b, bound, p, env);
                jf[bound] = ap[dim] * fs * (phiP - gv) + cvp[dim];
            } else jf[bound] = cvp[dim];
            continue;
        }
        T phiN, aN, cvN;
        if (atb) {
            const pp_boundary_desc &bd = b.bnd[bound];
            if (bd.type == PP_BOUND_CONNECTED) {
                int np[3];
                conn_pos(blocks, D, p, dim, bd, 0, np);
                const int ca = conn_channel(D, dim, dim, bd);
                const pp_block_desc &nb = blocks[bd.conn_block];
                const int nf = flat3(np, nb.size);
                phiN = phi[nf + nb.global_offset];
                aN = Metric<T, D>::alpha(nb, nf, ca);
                cvN = contra<T, D>(dom, nb, nf, ca, uxb, env);
                if ((bd.conn_axes[0] & 1) == up) cvN = -cvN;
            } else if (bd.type == PP_BOUND_PERIODIC) {
                int np[3] = {p[0], p[1], p[2]};
                np[dim] = up ? 0 : b.size[dim] - 1;
                const int nf = flat3(np, b.size);
                phiN = phi[nf + b.global_offset];
                aN = Metric<T, D>::alpha(b, nf, dim);
                cvN = contra<T, D>(dom, b, nf, dim, uxb, env);
            } else {
                jf[bound] = T(0);
                continue;
            }
        } else {
            int np[3] = {p[0], p[1], p[2]};
            np[dim] += face_sign(bound);
            const int nf = flat3(np, b.size);
            phiN = phi[nf + b.global_offset];
            aN = Metric<T, D>::alpha(b, nf, dim);
            cvN = contra<T, D>(dom, b, nf, dim, uxb, env);
        }
        const T coef = T(0.5) * (ap[dim] + aN);
        const T chi = T(0.5) * (cvp[dim] + cvN);
        jf[bound] = -coef * fs * (phiN - phiP) + chi;
    }
    T ja[D];
#pragma unroll
    for (int k = 0; k < D; ++k) ja[k] = T(0.5) * (jf[k * 2] + jf[k * 2 + 1]);
    T Jv[3] = {T(0), T(0), T(0)};
    if (b.has_transform && b.transform_diagonal) {
        const TView<T> t = Metric<T, D>::at(b, f);
        const T inv = T(1) / t[2 * D * D];
#pragma unroll
        for (int k = 0; k < D; ++k) Jv[k] = (t[k * D + k] * ja[k]) * inv;
        if constexpr (D == 2) Jv[2] = uxb[g + (size_t)nc * 2];
    } else if (b.has_transform) {
        const TView<T> t = Metric<T, D>::at(b, f);
        const T inv = T(1) / t[2 * D * D];
        if constexpr (D == 2) {
            Jv[0] = (t[0] * ja[0] + t[1] * ja[1]) * inv;
            Jv[1] = (t[2] * ja[0] + t[3] * ja[1]) * inv;
            Jv[2] = uxb[g + (size_t)nc * 2];
        } else {
            Jv[0] = (t[0] * ja[0] + t[1] * ja[1] + t[2] * ja[2]) * inv;
            Jv[1] = (t[3] * ja[0] + t[4] * ja[1] + t[5] * ja[2]) * inv;
            Jv[2] = (t[6] * ja[0] + t[7] * ja[1] + t[8] * ja[2]) * inv;
        }
    } else {
#pragma unroll
        for (int k = 0; k < D; ++k) Jv[k] = ja[k];
        if constexpr (D == 2) Jv[2] = uxb[g + (size_t)nc * 2];
    }
    T *Jo = J + (size_t)env * 3 * nc;
    Jo[g] = Jv[0];
    Jo[g + nc] = Jv[1];
    Jo[g + (size_t)nc * 2] = Jv[2];
}

// flat [batch][ch][n] <-> block fields [batch][ch][cells]
template <typename T>
__global__ void __launch_bounds__(ASM_BLOCK) k_copy_result(pp_domain dom, int field, int ch, int to_blocks,
                                                         const T *__restrict__ flat_in, T *__restrict__ flat_out) {
    const int wpb_ = (dom.max_block_cells + ASM_BLOCK - 1) / ASM_BLOCK;
    const int bi = __builtin_amdgcn_readfirstlane(blockIdx.x / wpb_);
    const int f = (blockIdx.x - bi * wpb_) * ASM_BLOCK + threadIdx.x;
    const pp_block_desc &b = dom.blocks[bi];
    if (f >= b.cells) return;
    const int g = b.global_offset + f;
    const int env = blockIdx.y;
    T *bf = reinterpret_cast<T *>(field == 0 ? b.velocity : (field == 1 ? b.pressure : (field == 2 ? b.epot : b.scalar)));
    if (!bf) return;
    for (int c = 0; c < ch; ++c) {
        const size_t fo = ((size_t)env * ch + c) * dom.num_cells + g;
        const size_t bo = ((size_t)env * ch + c) * b.cells + f;
        if (to_blocks) bf[bo] = flat_in[fo];
        else flat_out[fo] = bf[bo];
    }
}

// det and adjugate*(1/det) of MatrixSquare (common/transformations.h:437-495, 558-562)
template <typename T, int D>
__device__ __forceinline__ void det_inverse(const T *m, T &det, T *inv) {
    T a[D * D];
    if constexpr (D == 2) {
        det = m[0] * m[3] - m[2] * m[1];
        a[0] = m[3]; a[1] = -m[1]; a[2] = -m[2]; a[3] = m[0];
    } else {
        det = m[0] * (m[4] * m[8] - m[7] * m[5]) - m[1] * (m[3] * m[8] - m[6] * m[5]) + m[2] * (m[3] * m[7] - m[6] * m[4]);
        a[0] = m[4] * m[8] - m[7] * m[5];
        a[1] = m[2] * m[7] - m[1] * m[8];
        a[2] = m[1] * m[5] - m[2] * m[4];
        a[3] = m[5] * m[6] - m[3] * m[8];
        a[4] = m[0] * m[8] - m[2] * m[6];
        a[5] = m[3] * m[2] - m[0] * m[5];
        a[6] = m[3] * m[7] - m[6] * m[4];
        a[7] = m[6] * m[1] - m[0] * m[7];
        a[8] = m[0] * m[4] - m[3] * m[1];
    }
    const T s = T(1) / det;
#pragma unroll
    for (int i = 0; i < D * D; ++i) inv[i] = a[i] * s;
}

// k_CoordsToTransforms, grid_gen.cu:298-354
template <typename T, int D>
__global__ void __launch_bounds__(ASM_BLOCK) k_coords_to_transforms(const T *__restrict__ coords, int cx, int cy, int cz,
                                                                  T *__restrict__ out) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    const int ncell = cx * cy * cz;
    if (f >= ncell) return;
    const int32_t cs[3] = {cx, cy, cz};
    const int32_t vs[3] = {cx + 1, D > 1 ? cy + 1 : 1, D > 2 ? cz + 1 : 1};
    const size_t vstride = (size_t)vs[0] * vs[1] * vs[2];
    int p[3];
    unflat(f, cs, p);
    T fc[2 * D][D];
#pragma unroll
    for (int i = 0; i < 2 * D; ++i)
#pragma unroll
        for (int k = 0; k < D; ++k) fc[i][k] = T(0);
#pragma unroll
    for (int v = 0; v < (1 << D); ++v) {
        int vp[3] = {p[0] + (v & 1), p[1] + ((v & 2) != 0), p[2] + ((v & 4) != 0)};
        T co[D];
#pragma unroll
        for (int k = 0; k < D; ++k) co[k] = coords[k * vstride + flat3(vp, vs)];
#pragma unroll
        for (int fi = 0; fi < D; ++fi) {
            const int up = (v & (1 << fi)) != 0;
#pragma unroll
            for (int k = 0; k < D; ++k) fc[fi * 2 + up][k] += co[k];
        }
    }
    const T norm = T(1.0) / T(1 << (D - 1));
#pragma unroll
    for (int i = 0; i < 2 * D; ++i)
#pragma unroll
        for (int k = 0; k < D; ++k) fc[i][k] *= norm;
    T *t = out + (size_t)f * (2 * D * D + 1);
    T m[D * D];
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
        for (int k = 0; k < D; ++k) m[i * D + k] = fc[k * 2 + 1][i] - fc[k * 2][i];
    T det;
    det_inverse<T, D>(m, det, t + D * D);
#pragma unroll
    for (int i = 0; i < D * D; ++i) t[i] = m[i];
    t[2 * D * D] = det;
}

// k_CoordsToFaceTransforms, grid_gen.cu:398-527; out [d][vertex grid][2d^2+1]
template <typename T, int D>
__global__ void __launch_bounds__(ASM_BLOCK) k_coords_to_face_transforms(const T *__restrict__ coords, int cx, int cy,
                                                                       int cz, T *__restrict__ out) {
    const int32_t vs[3] = {cx + 1, D > 1 ? cy + 1 : 1, D > 2 ? cz + 1 : 1};
    const int vstride = vs[0] * vs[1] * vs[2];
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= vstride) return;
    int p[3];
    unflat(f, vs, p);
    constexpr int NFV = 1 << (D - 1), NEV = 1 << (D - 2);
    constexpr int TS = 2 * D * D + 1;
#pragma unroll
    for (int dim = 0; dim < D; ++dim) {
        T *o = out + ((size_t)dim * vstride + f) * TS;
        bool skip = false;
#pragma unroll
        for (int i = 1; i < D; ++i) {
            const int td = (dim + i) % D;
            if (p[td] == vs[td] - 1) skip = true;
        }
        if (skip) {
            for (int i = 0; i < TS; ++i) o[i] = T(0);
            continue;
        }
        T t[D * D];
#pragma unroll
        for (int i = 0; i < D * D; ++i) t[i] = T(0);
        T nw = T(0.5);
#pragma unroll
        for (int up = 0; up < 2; ++up) {
            const int fs = up * 2 - 1;
            const bool atb = up ? p[dim] == vs[dim] - 1 : p[dim] == 0;
            int fp[3] = {p[0], p[1], p[2]};
            if (!atb) fp[dim] += fs;
            else nw = T(1.0);
#pragma unroll
            for (int v = 0; v < NFV; ++v) {
                int vp[3] = {fp[0], fp[1], fp[2]};
#pragma unroll
                for (int i = 1; i < D; ++i) vp[(dim + i) % D] += (v & i) != 0;
#pragma unroll
                for (int k = 0; k < D; ++k) t[k * D + dim] += coords[(size_t)k * vstride + flat3(vp, vs)] * T(fs);
            }
        }
        nw /= T(NFV);
#pragma unroll
        for (int k = 0; k < D; ++k) t[k * D + dim] *= nw;
        const T tnw = T(1.0) / T(NEV);
#pragma unroll
        for (int kk = 1; kk < D; ++kk) {
            const int td = (dim + kk) % D;
#pragma unroll
            for (int v = 0; v < NFV; ++v) {
                int vp[3] = {p[0], p[1], p[2]};
#pragma unroll
                for (int i = 1; i < D; ++i) vp[(dim + i) % D] += (v & i) != 0;
                const T tfs = (v & kk) ? T(1) : T(-1);
#pragma unroll
                for (int k = 0; k < D; ++k) t[k * D + td] += coords[(size_t)k * vstride + flat3(vp, vs)] * tfs;
            }
#pragma unroll
            for (int k = 0; k < D; ++k) t[k * D + td] *= tnw;
        }
        T det;
        det_inverse<T, D>(t, det, o + D * D);
#pragma unroll
        for (int i = 0; i < D * D; ++i) o[i] = t[i];
        o[2 * D * D] = det;
    }
}

// out[c][r] = in[r][c] through 32x32 LDS tiles (row-major [rows][cols] -> [cols][rows]).
template <typename T>
__global__ void __launch_bounds__(256) k_transpose_2d(const T *__restrict__ in, int rows, int cols, T *__restrict__ out) {
    __shared__ T tile[32][33];
    const int r0 = blockIdx.x * 32, c0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5; // 32 x 8
    for (int k = ty; k < 32; k += 8) {
        const int r = r0 + k, c = c0 + tx;
        if (r < rows && c < cols) tile[k][tx] = in[(size_t)r * cols + c];
    }
    __syncthreads();
    for (int k = ty; k < 32; k += 8) {
        const int c = c0 + k, r = r0 + tx;
        if (r < rows && c < cols) out[(size_t)c * rows + r] = tile[tx][k];
    }
}

// ---- adjoints of the assembly kernels (the reference's _GRAD kernels, advection.cu:309-542,
// :957-1200, pressure.cu:195-348, :524-755, :805-880, velocity_correction.cu:300-375,
// potential.cu:980-1329), orthogonal CENTRAL path. Each restates its forward kernel above in
// reverse: per cell, the gradient of the cell's outputs is pulled back to the inputs the
// outputs were computed from and added to their gradients. Inputs shared by several cells
// (neighbour velocities, pressures, the diagonal) take atomic adds; a cell's own entries are
// written by its work-item alone. Gradient buffers are flat [batch][channels][n] (global cell
// index), fixed faces' velocity gradients [batch][d][face cells] per (block, face).

// Position of each candidate entry in the row written by emit_row (ascending column, the
// lower candidate first among equal columns); -1 for candidates that were not written.
template <int D>
__device__ __forceinline__ void row_positions(const int *idx, int *pos) {
#pragma unroll
    for (int s = 0; s < 2 * D + 1; ++s) {
        int r = 0;
#pragma unroll
        for (int t = 0; t < 2 * D + 1; ++t)
            if (idx[t] >= 0 && (idx[t] < idx[s] || (idx[t] == idx[s] && t < s))) ++r;
        pos[s] = idx[s] >= 0 ? r : -1;
    }
}

// gradient of the contravariant component w of cell f (global index gi) of a vector field:
// d(det Minv[w] . v)/dv_k = det Minv[w][k]; gv points at the field's [channels][n] gradient
template <typename T, int D>
__device__ __forceinline__ void add_contra_grad(const pp_block_desc &b, int f, int gi, int w, int n, T gval,
                                                T *__restrict__ gv) {
    if (!b.has_transform) {
        atomicAdd(gv + (size_t)w * n + gi, gval);
        return;
    }
    const TView<T> t = Metric<T, D>::at(b, f);
    const T s = gval * t[2 * D * D];
    if (b.transform_diagonal) {
        atomicAdd(gv + (size_t)w * n + gi, s * t[D * D + w * D + w]);
        return;
    }
#pragma unroll
    for (int k = 0; k < D; ++k) atomicAdd(gv + (size_t)k * n + gi, s * t[D * D + w * D + k]);
}

// gradient of contra_fixed (component w at face cell ff) with respect to the face velocity
template <typename T, int D>
__device__ __forceinline__ void contra_fixed_grad(const pp_boundary_desc &fb, int ff, int w, T gval, T *gv) {
    if (!fb.has_transform) {
        gv[w] += gval;
        return;
    }
    if (fb.vel_type != PP_BC_DIRICHLET) return;
    const TView<T> t = Metric<T, D>::face(fb, ff);
    const T s = gval * t[2 * D * D];
#pragma unroll
    for (int k = 0; k < D; ++k) gv[k] += s * t[D * D + w * D + k];
}

// Pull the gradients gfl[bound] of the face fluxes of cell f (cell_fluxes, non-EPOT and EPOT
// forms) back to the vector field (gfield: [D|3][n] of this environment) and, for the
// non-EPOT form, to the fixed faces' velocities (gbv: per face, D values for this face cell).
template <typename T, int D, bool EPOT>
__device__ void cell_fluxes_grad(const pp_domain &dom, const pp_block_desc *__restrict__ blocks, int bi, const int *p,
                                 int f, int gi, const T *gfl, T *__restrict__ gfield, T (*gbv)[D]) {
    const pp_block_desc &b = blocks[bi];
    const int n = dom.num_cells;
#pragma unroll
    for (int bound = 0; bound < 2 * D; ++bound) {
        const int dim = bound >> 1;
        const int up = bound & 1;
        const T gf = gfl[bound];
        if (at_bound(p, bound, b)) {
            const pp_boundary_desc &bd = b.bnd[bound];
            if (bd.type == PP_BOUND_FIXED) {
                if (EPOT) {
                    if (!is_insulating_wall(b, bound, p)) add_contra_grad<T, D>(b, f, gi, dim, n, gf, gfield);
                } else if (gbv) {
                    contra_fixed_grad<T, D>(bd, face_flat(p, bound, b), dim, gf, gbv[bound]);
                }
            } else if (bd.type == PP_BOUND_CONNECTED) {
                int np[3];
                conn_pos(blocks, D, p, dim, bd, 0, np);
                const pp_block_desc &nb = blocks[bd.conn_block];
                const int nf = flat3(np, nb.size);
                const T sgn = (bd.conn_axes[0] & 1) == up ? T(-1) : T(1);
                add_contra_grad<T, D>(nb, nf, nf + nb.global_offset, bd.conn_axes[0] >> 1, n, T(0.5) * gf * sgn, gfield);
                add_contra_grad<T, D>(b, f, gi, dim, n, T(0.5) * gf, gfield);
            } else if (bd.type == PP_BOUND_PERIODIC) {
                int np[3] = {p[0], p[1], p[2]};
                np[dim] = up ? 0 : b.size[dim] - 1;
                const int nf = flat3(np, b.size);
                add_contra_grad<T, D>(b, nf, nf + b.global_offset, dim, n, T(0.5) * gf, gfield);
                add_contra_grad<T, D>(b, f, gi, dim, n, T(0.5) * gf, gfield);
            }
        } else {
            int np[3] = {p[0], p[1], p[2]};
            np[dim] += face_sign(bound);
            const int nf = flat3(np, b.size);
            add_contra_grad<T, D>(b, nf, nf + b.global_offset, dim, n, T(0.5) * gf, gfield);
            add_contra_grad<T, D>(b, f, gi, dim, n, T(0.5) * gf, gfield);
        }
    }
}

// the candidate columns of row g of C / P (as the forward kernels build them)
template <typename T, int D>
__device__ __forceinline__ void row_columns(const pp_block_desc *__restrict__ blocks, int bi, const int *p, int g,
                                            int *idx) {
    const pp_block_desc &b = blocks[bi];
    idx[0] = g;
#pragma unroll
    for (int bound = 0; bound < 2 * D; ++bound) {
        if (!at_bound(p, bound, b) || !is_fixed(b, bound)) {
            int np[3], nw;
            const int nbi = neighbor(blocks, bi, D, p, bound, np, &nw);
            const pp_block_desc &nb = blocks[nbi];
            idx[bound + 1] = flat3(np, nb.size) + nb.global_offset;
        } else {
            idx[bound + 1] = -1;
        }
    }
}

// store the fixed faces' velocity gradients of one cell (each face cell borders one cell)
template <typename T, int D>
__device__ __forceinline__ void store_face_grads(const pp_block_desc &b, const int *p, int bi, int env,
                                                 T *const *__restrict__ gbnd, T (*gbv)[D]) {
#pragma unroll
    for (int bound = 0; bound < 2 * D; ++bound) {
        if (!at_bound(p, bound, b) || !is_fixed(b, bound)) continue;
        T *gb = gbnd[bi * 6 + bound];
        if (!gb) continue;
        const int fc = face_cells(b, bound), ff = face_flat(p, bound, b);
#pragma unroll
        for (int k = 0; k < D; ++k) gb[((size_t)env * D + k) * fc + ff] += gbv[bound][k];
    }
}

// adjoint of the momentum boundary source (boundary_source / the loop in k_advection_rhs)
// scaled by q: adds to gbv[bound][k] for every fixed face of the cell
template <typename T, int D>
__device__ __forceinline__ void boundary_source_grad(const pp_domain &dom, const pp_block_desc &b, const int *p, int f,
                                                     int dim, int env, T q, T (*gbv)[D], T &gnu) {
#pragma unroll
    for (int bound = 0; bound < 2 * D; ++bound) {
        if (!at_bound(p, bound, b)) continue;
        const pp_boundary_desc &fb = b.bnd[bound];
        if (fb.type != PP_BOUND_FIXED) continue;
        const T fn = T((bound & 1) * 2 - 1);
        const int fc = face_cells(b, bound), ff = face_flat(p, bound, b);
        const T vel = fixed_vel<T>(fb, D, fc, ff, dim, env);
        const T alpha = alpha_fixed<T, D>(fb, ff, bound >> 1);
        const T flux = contra_fixed<T, D>(fb, fc, ff, bound >> 1, env) * fn;
        const T bv = visc_at<T>(dom, b, f);
        const T slip = fb.vel_type == PP_BC_DIRICHLET ? T(0) : T(1);
        // s += -vel * flux + vel * (1 - slip) * bv * 2 * alpha
        gbv[bound][dim] += q * ((1 - slip) * bv * 2 * alpha - flux);
        gnu += q * vel * (1 - slip) * 2 * alpha;
        contra_fixed_grad<T, D>(fb, ff, bound >> 1, -q * vel * fn, gbv[bound]);
    }
}

// SetupAdvectionMatrixEulerImplicit_GRAD (advection.cu:309-542): gC [B][nnz], gA [B][n] (or
// null) -> gU [B][D][n] (the block velocity; C depends on it through the face fluxes only)
// gNu [B][n] (or null): the gradient with respect to the viscosity as a per-cell field (the
// viscous coefficient of a face is (alpha_P nu_P + alpha_N nu_N) / 2, of a no-slip wall
// 2 alpha_P nu_P); the caller reduces it to the domain's, a block's or a cell's viscosity.
template <typename T, int D>
__global__ void __launch_bounds__(ASM_BLOCK) k_advection_matrix_grad(pp_domain dom, const T *__restrict__ gC,
                                                                     const T *__restrict__ gA, T *__restrict__ gU,
                                                                     T *__restrict__ gNu) {
    CELL_PROLOGUE
    const T det = Metric<T, D>::det(b, f);
    int idx[2 * D + 1], pos[2 * D + 1];
    row_columns<T, D>(blocks, bi, p, g, idx);
    row_positions<D>(idx, pos);
    const T *gc = gC + (size_t)env * dom.C.nnz + dom.C.row[g];
    T grow[2 * D + 1];
#pragma unroll
    for (int s = 0; s < 2 * D + 1; ++s) grow[s] = pos[s] >= 0 ? gc[pos[s]] / det : T(0);
    if (gA) grow[0] += gA[(size_t)env * dom.num_cells + g] / det;
    T gfl[2 * D];
#pragma unroll
    for (int bound = 0; bound < 2 * D; ++bound)
        gfl[bound] = idx[bound + 1] >= 0 ? T(0.5) * T(face_sign(bound)) * (grow[0] + grow[bound + 1]) : T(0);
    cell_fluxes_grad<T, D, false>(dom, blocks, bi, p, f, g, gfl, gU + (size_t)env * D * dom.num_cells, nullptr);
    if (gNu) {
        // diag += vc, off-diagonal -= vc with vc = (ap nu_P + an nu_N) / 2; walls: diag += (1 - slip) 2 ap nu_P
        T *gn = gNu + (size_t)env * dom.num_cells;
        T own = T(0);
#pragma unroll
        for (int bound = 0; bound < 2 * D; ++bound) {
            const int dim = bound >> 1;
            const T ap = Metric<T, D>::alpha(b, f, dim);
            if (idx[bound + 1] >= 0) {
                int np[3], nw;
                const int nbi = neighbor(blocks, bi, D, p, bound, np, &nw);
                const pp_block_desc &nb = blocks[nbi];
                const int nf = flat3(np, nb.size);
                const T dvc = grow[0] - grow[bound + 1];
                own += dvc * ap * T(0.5);
                atomicAdd(gn + nf + nb.global_offset, dvc * Metric<T, D>::alpha(nb, nf, nw) * T(0.5));
            } else {
                const T slip = b.bnd[bound].vel_type == PP_BC_DIRICHLET ? T(0) : T(1);
                own += grow[0] * (1 - slip) * 2 * ap;
            }
        }
        atomicAdd(gn + g, own);
    }
}

// SetupAdvectionVelocityEulerImplicitRHS_GRAD (advection.cu:957-1200), without the pressure
// gradient term: gR [B][D][n] -> gU (block velocity), gS (per-cell source, or null), fixed
// faces' velocities (gbnd, per block and face, or null entries)
template <typename T, int D>
__global__ void __launch_bounds__(ASM_BLOCK) k_advection_rhs_grad(pp_domain dom, const T *__restrict__ gR,
                                                                  T *__restrict__ gU, T *__restrict__ gS,
                                                                  T *const *__restrict__ gbnd, T *__restrict__ gNu) {
    CELL_PROLOGUE
    const int n = dom.num_cells;
    const T dt = reinterpret_cast<const T *>(dom.time_step)[env];
    const T det = Metric<T, D>::det(b, f);
    T gbv[2 * D][D];
    T gnu = T(0);
#pragma unroll
    for (int i = 0; i < 2 * D; ++i)
#pragma unroll
        for (int k = 0; k < D; ++k) gbv[i][k] = T(0);
#pragma unroll
    for (int dim = 0; dim < D; ++dim) {
        const size_t o = ((size_t)env * D + dim) * n + g;
        const T gv = gR[o];
        gU[o] += gv * det / dt / det;
        if (gS) gS[o] += gv;
        boundary_source_grad<T, D>(dom, b, p, f, dim, env, gv / det, gbv, gnu);
    }
    if (gbnd) store_face_grads<T, D>(b, p, bi, env, gbnd, gbv);
    if (gNu && gnu != T(0)) atomicAdd(gNu + (size_t)env * n + g, gnu);
}

// SetupPressureMatrix_GRAD (pressure.cu:195-348): gP [B][nnz] -> gA [B][n] (A: the diagonal
// the matrix was built from)
template <typename T, int D>
__global__ void __launch_bounds__(ASM_BLOCK) k_pressure_matrix_grad(pp_domain dom, const T *__restrict__ Aval,
                                                                    const T *__restrict__ gP, T *__restrict__ gA) {
    CELL_PROLOGUE
    const T *A = Aval + (size_t)env * dom.num_cells;
    T *ga = gA + (size_t)env * dom.num_cells;
    int idx[2 * D + 1], pos[2 * D + 1];
    row_columns<T, D>(blocks, bi, p, g, idx);
    row_positions<D>(idx, pos);
    const T *gp = gP + (size_t)env * dom.P.nnz + dom.P.row[g];
    const T raP = T(1.0) / A[g];
    T gaP = T(0);
#pragma unroll
    for (int bound = 0; bound < 2 * D; ++bound) {
        if (idx[bound + 1] < 0) continue;
        const int dim = bound >> 1;
        int np[3], nw;
        const int nbi = neighbor(blocks, bi, D, p, bound, np, &nw);
        const pp_block_desc &nb = blocks[nbi];
        const int nf = flat3(np, nb.size);
        const T an = Metric<T, D>::alpha(nb, nf, nw);
        const int ng = nf + nb.global_offset;
        const T raN = T(1.0) / A[ng];
        // coef = 0.5 (ap raP + an raN); diagonal -= coef, entry += coef
        const T gcoef = gp[pos[bound + 1]] - gp[pos[0]];
        gaP -= gcoef * T(0.5) * Metric<T, D>::alpha(b, f, dim) * raP * raP;
        atomicAdd(ga + ng, -gcoef * T(0.5) * an * raN * raN);
    }
    atomicAdd(ga + g, gaP);
}

// SetupPressureRHS_GRAD (pressure.cu:524-755): gH [B][D][n] (gradient of HbyA) -> gA, gC
// (off-diagonal entries of C), gV (the predicted velocity), gU (block velocity), gS
// (per-cell source or null), fixed faces' velocities. Aval, Cval, V and H are the forward
// inputs and output.
template <typename T, int D>
__global__ void __launch_bounds__(ASM_BLOCK) k_pressure_rhs_grad(pp_domain dom, const T *__restrict__ gH,
                                                                 const T *__restrict__ Aval, const T *__restrict__ Cval,
                                                                 const T *__restrict__ V, const T *__restrict__ H,
                                                                 T *__restrict__ gA, T *__restrict__ gC,
                                                                 T *__restrict__ gV, T *__restrict__ gU,
                                                                 T *__restrict__ gS, T *const *__restrict__ gbnd,
                                                                 T *__restrict__ gNu) {
    CELL_PROLOGUE
    const int n = dom.num_cells;
    const T dt = reinterpret_cast<const T *>(dom.time_step)[env];
    const T A = Aval[(size_t)env * n + g];
    const T det = Metric<T, D>::det(b, f);
    T q[D];
    T ga = T(0), gnu = T(0);
    T gbv[2 * D][D];
#pragma unroll
    for (int i = 0; i < 2 * D; ++i)
#pragma unroll
        for (int k = 0; k < D; ++k) gbv[i][k] = T(0);
#pragma unroll
    for (int dim = 0; dim < D; ++dim) {
        const size_t o = ((size_t)env * D + dim) * n + g;
        // HbyA = (u/dt - H + S) / A
        q[dim] = gH[o] / A;
        ga -= gH[o] * H[o] / A;
        gU[o] += q[dim] / dt;
        if (gS) gS[o] += q[dim];
        boundary_source_grad<T, D>(dom, b, p, f, dim, env, q[dim] / det, gbv, gnu);
    }
    atomicAdd(gA + (size_t)env * n + g, ga);
    if (gNu && gnu != T(0)) atomicAdd(gNu + (size_t)env * n + g, gnu);
    const int32_t cs = dom.C.row[g];
    const int rs = dom.C.row[g + 1] - cs;
    const T *cv = Cval + (size_t)env * dom.C.nnz + cs;
    T *gc = gC + (size_t)env * dom.C.nnz + cs;
    const T *v = V + (size_t)env * D * n;
    T *gv = gV + (size_t)env * D * n;
    for (int i = 0; i < rs && i < 2 * D + 1; ++i) {
        const int ix = dom.C.col[cs + i];
        if (ix == g) continue;
        T s = T(0);
#pragma unroll
        for (int dim = 0; dim < D; ++dim) {
            s -= q[dim] * v[ix + (size_t)n * dim];
            atomicAdd(gv + ix + (size_t)n * dim, -cv[i] * q[dim]);
        }
        gc[i] += s;
    }
    if (gbnd) store_face_grads<T, D>(b, p, bi, env, gbnd, gbv);
}

// SetupPressureRHSdiv_GRAD (pressure.cu:805-880): gDiv [B][n] -> gH [B][D][n] (HbyA) and the
// fixed faces' velocities
template <typename T, int D>
__global__ void __launch_bounds__(ASM_BLOCK) k_pressure_div_grad(pp_domain dom, const T *__restrict__ gDiv,
                                                                 int time_step_norm, T *__restrict__ gH,
                                                                 T *const *__restrict__ gbnd) {
    CELL_PROLOGUE
    T gd = gDiv[(size_t)env * dom.num_cells + g];
    if (time_step_norm) gd /= reinterpret_cast<const T *>(dom.time_step)[env];
    T gfl[2 * D];
#pragma unroll
    for (int bound = 0; bound < 2 * D; ++bound) gfl[bound] = T(face_sign(bound)) * gd;
    T gbv[2 * D][D];
#pragma unroll
    for (int i = 0; i < 2 * D; ++i)
#pragma unroll
        for (int k = 0; k < D; ++k) gbv[i][k] = T(0);
    cell_fluxes_grad<T, D, false>(dom, blocks, bi, p, f, g, gfl, gH + (size_t)env * D * dom.num_cells, gbv);
    if (gbnd) store_face_grads<T, D>(b, p, bi, env, gbnd, gbv);
}

// global index of the cell pressure_at reads for a position that may lie one cell outside
// the block; -1 where it reads zero
__device__ int pressure_at_index(const pp_domain &dom, const pp_block_desc *__restrict__ blocks, int bi,
                                 const int *pin) {
    const pp_block_desc &b = blocks[bi];
    int t[3] = {pin[0], pin[1], pin[2]};
    const int D = dom.dims;
    for (int dim = 0; dim < D; ++dim) {
        const bool lower = pin[dim] < 0;
        const bool upper = pin[dim] >= b.size[dim];
        if (!lower && !upper) continue;
        const pp_boundary_desc &bd = b.bnd[dim * 2 + (upper ? 1 : 0)];
        if (bd.type == PP_BOUND_FIXED) {
            t[dim] = lower ? 0 : b.size[dim] - 1;
            return b.global_offset + flat3(t, b.size);
        }
        if (bd.type == PP_BOUND_CONNECTED) {
            int op[3];
            conn_pos(blocks, D, t, dim, bd, lower ? -(pin[dim] + 1) : pin[dim] - b.size[dim], op);
            const pp_block_desc &nb = blocks[bd.conn_block];
            return nb.global_offset + flat3(op, nb.size);
        }
        if (bd.type == PP_BOUND_PERIODIC) {
            t[dim] += lower ? b.size[dim] : -b.size[dim];
            return b.global_offset + flat3(t, b.size);
        }
        return -1;
    }
    return b.global_offset + flat3(pin, b.size);
}

// CorrectVelocity_GRAD (velocity_correction.cu:300-375), version 1 (FD): gOut [B][D][n] ->
// gH (HbyA), gA, gP [B][n] (the pressure the gradient was taken of, global cell index)
template <typename T, int D>
__global__ void __launch_bounds__(ASM_BLOCK) k_correct_velocity_grad(pp_domain dom, const T *__restrict__ gOut,
                                                                     const T *__restrict__ Aval, int time_step_norm,
                                                                     T *__restrict__ gH, T *__restrict__ gA,
                                                                     T *__restrict__ gP) {
    CELL_PROLOGUE
    const int n = dom.num_cells;
    const T A = Aval[(size_t)env * n + g];
    const T rDiag = 1 / A;
    T gr[D];
    pressure_grad<T, D>(dom, blocks, bi, p, f, env, gr);
    const T s = time_step_norm ? reinterpret_cast<const T *>(dom.time_step)[env] : T(1);
    T ggr[D];
    T ga = T(0);
#pragma unroll
    for (int dim = 0; dim < D; ++dim) {
        const size_t o = ((size_t)env * D + dim) * n + g;
        const T go = gOut[o];
        gH[o] += go;
        // u = -rDiag * gr * s + HbyA
        ga += go * gr[dim] * s * rDiag * rDiag;
        ggr[dim] = -rDiag * go * s;
    }
    atomicAdd(gA + (size_t)env * n + g, ga);
    // gr_j = sum_i Minv[i][j] raw_i  ->  graw_i = sum_j Minv[i][j] ggr_j
    T graw[D];
    if (b.has_transform) {
        const TView<T> t = Metric<T, D>::at(b, f);
#pragma unroll
        for (int i = 0; i < D; ++i) {
            T acc = T(0);
#pragma unroll
            for (int j = 0; j < D; ++j) acc += t[D * D + i * D + j] * ggr[j];
            graw[i] = acc;
        }
    } else {
#pragma unroll
        for (int i = 0; i < D; ++i) graw[i] = ggr[i];
    }
    T *gp = gP + (size_t)env * n;
#pragma unroll
    for (int dim = 0; dim < D; ++dim) {
        int t[3] = {p[0], p[1], p[2]};
        T fac = T(0.5);
        if (!(p[dim] == 0 && is_fixed(b, dim * 2))) t[dim] = p[dim] - 1;
        else fac = T(1);
        const int in_ = pressure_at_index(dom, blocks, bi, t);
        if (!(p[dim] == b.size[dim] - 1 && is_fixed(b, dim * 2 + 1))) t[dim] = p[dim] + 1;
        else { t[dim] = p[dim]; fac = T(1); }
        const int ip = pressure_at_index(dom, blocks, bi, t);
        if (ip >= 0) atomicAdd(gp + ip, graw[dim] * fac);
        if (in_ >= 0) atomicAdd(gp + in_, -graw[dim] * fac);
    }
}

// ComputeEpotRHSGrad (potential.cu:980-1040): gDiv [B][n] -> gF [B][D][n], the in-plane
// vector field whose face fluxes the divergence was taken of
// Gradient of a Dirichlet potential value at the face cell next to cell p: gv holds one
// [batch][face cells] buffer per (block, face) (entry block * 6 + face, null: not wanted); the
// face cell has this one cell as its only neighbour, so no atomics are needed.
template <typename T>
__device__ __forceinline__ void put_value_grad(T *const *gv, int bi, int bound, const pp_block_desc &b, const int *p,
                                               int env, T v) {
    if (!gv) return;
    T *buf = gv[bi * 6 + bound];
    if (buf) buf[(size_t)env * face_cells(b, bound) + face_flat(p, bound, b)] += v;
}

// ComputeEpotRHSGrad (potential.cu:980-1040): gDiv [B][n] -> gF [B][d][n] (in-plane u x e_b) and
// the Dirichlet values' gradients (div -= alpha g at a Dirichlet face, potential.cu:272-322).
template <typename T, int D>
__global__ void __launch_bounds__(ASM_BLOCK) k_epot_rhs_grad(pp_domain dom, const T *__restrict__ gDiv,
                                                             T *__restrict__ gF, T *const *gv) {
    CELL_PROLOGUE
    const T gd = gDiv[(size_t)env * dom.num_cells + g];
    T gfl[2 * D];
#pragma unroll
    for (int bound = 0; bound < 2 * D; ++bound) gfl[bound] = T(face_sign(bound)) * gd;
    cell_fluxes_grad<T, D, true>(dom, blocks, bi, p, f, g, gfl, gF + (size_t)env * D * dom.num_cells, nullptr);
    if (gv) {
#pragma unroll
        for (int bound = 0; bound < 2 * D; ++bound) {
            if (!at_bound(p, bound, b) || !is_fixed(b, bound) || !is_epot_dirichlet(b, bound, p)) continue;
            put_value_grad<T>(gv, bi, bound, b, p, env, -Metric<T, D>::alpha(b, f, bound >> 1) * gd);
        }
    }
}

// ComputeCurrentDensityFaceBasedGrad (potential.cu:1248-1329): gJ [B][3][n] -> gPhi [B][n],
// gX [B][3][n] (u x e_b)
// and the Dirichlet values' gradients (gv as in k_epot_rhs_grad)
template <typename T, int D>
__global__ void __launch_bounds__(ASM_BLOCK) k_current_density_grad(pp_domain dom, const T *__restrict__ gJall,
                                                                    T *__restrict__ gPhiAll, T *__restrict__ gXall,
                                                                    T *const *gv) {
    CELL_PROLOGUE
    const int nc = dom.num_cells;
    const T *gJ = gJall + (size_t)env * 3 * nc;
    T *gphi = gPhiAll + (size_t)env * nc;
    T *gx = gXall + (size_t)env * 3 * nc;
    const T gJv[3] = {gJ[g], gJ[g + nc], gJ[g + (size_t)nc * 2]};
    T gja[D];
    if (b.has_transform) {
        const TView<T> t = Metric<T, D>::at(b, f);
        const T inv = T(1) / t[2 * D * D];
#pragma unroll
        for (int k = 0; k < D; ++k) {
            T acc = T(0);
#pragma unroll
            for (int i = 0; i < D; ++i) acc += t[i * D + k] * gJv[i];
            gja[k] = acc * inv;
        }
    } else {
#pragma unroll
        for (int k = 0; k < D; ++k) gja[k] = gJv[k];
    }
    if constexpr (D == 2) atomicAdd(gx + g + (size_t)nc * 2, gJv[2]);
    T ap[D];
#pragma unroll
    for (int k = 0; k < D; ++k) ap[k] = Metric<T, D>::alpha(b, f, k);
    T gphiP = T(0);
#pragma unroll
    for (int bound = 0; bound < 2 * D; ++bound) {
        const int dim = bound >> 1;
        const int up = bound & 1;
        const T fs = T(face_sign(bound));
        const T gjf = T(0.5) * gja[dim];
        const bool atb = at_bound(p, bound, b);
        if (atb && is_fixed(b, bound)) {
            if (is_insulating_wall(b, bound, p)) continue;
            if (is_epot_dirichlet(b, bound, p)) {
                // jf = ap fs (phiP - value) + cvp
                gphiP += ap[dim] * fs * gjf;
                put_value_grad<T>(gv, bi, bound, b, p, env, -ap[dim] * fs * gjf);
            }
            add_contra_grad<T, D>(b, f, g, dim, nc, gjf, gx);
            continue;
        }
        int ng, nf, ca;
        const pp_block_desc *nbp;
        T sgn = T(1);
        if (atb) {
            const pp_boundary_desc &bd = b.bnd[bound];
            if (bd.type == PP_BOUND_CONNECTED) {
                int np[3];
                conn_pos(blocks, D, p, dim, bd, 0, np);
                ca = conn_channel(D, dim, dim, bd);
                nbp = &blocks[bd.conn_block];
                nf = flat3(np, nbp->size);
                if ((bd.conn_axes[0] & 1) == up) sgn = T(-1);
            } else if (bd.type == PP_BOUND_PERIODIC) {
                int np[3] = {p[0], p[1], p[2]};
                np[dim] = up ? 0 : b.size[dim] - 1;
                nbp = &b;
                nf = flat3(np, b.size);
                ca = dim;
            } else {
                continue;
            }
        } else {
            int np[3] = {p[0], p[1], p[2]};
            np[dim] += face_sign(bound);
            nbp = &b;
            nf = flat3(np, b.size);
            ca = dim;
        }
        ng = nf + nbp->global_offset;
        const T aN = Metric<T, D>::alpha(*nbp, nf, ca);
        const T coef = T(0.5) * (ap[dim] + aN);
        // jf = -coef fs (phiN - phiP) + 0.5 (cvp + sgn cvN)
        atomicAdd(gphi + ng, -coef * fs * gjf);
        gphiP += coef * fs * gjf;
        add_contra_grad<T, D>(b, f, g, dim, nc, T(0.5) * gjf, gx);
        add_contra_grad<T, D>(*nbp, nf, ng, ca, nc, T(0.5) * gjf * sgn, gx);
    }
    atomicAdd(gphi + g, gphiP);
}

// ---- maximum velocity (CFL substeps) and vector transforms

// Unsigned integers that order non-negative values (and one canonical NaN above +inf) like
// the values themselves: the running maxima are combined with integer atomic max.
__device__ __forceinline__ unsigned long long order_bits(double v) {
    return v != v ? 0x7ff8000000000000ull : static_cast<unsigned long long>(__double_as_longlong(v));
}
__device__ __forceinline__ unsigned int order_bits(float v) { return v != v ? 0x7fc00000u : __float_as_uint(v); }

__device__ __forceinline__ void fold_max(double *p, double v) {
    atomicMax(reinterpret_cast<unsigned long long *>(p), order_bits(v));
}
__device__ __forceinline__ void fold_max(float *p, float v) { atomicMax(reinterpret_cast<unsigned int *>(p), order_bits(v)); }

// larger of a and b; a NaN on either side wins, as torch::max propagates NaN
template <typename T>
__device__ __forceinline__ T nan_max(T a, T b) { return (b > a || b != b) ? b : a; }

template <typename T>
__global__ void __launch_bounds__(256) k_fill(T *__restrict__ p, int n, T v) {
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) p[i] = v;
}

// v <- Minv v with the strided transform t (TransformVectors with inverse = true,
// transform_vectors.cu:43-72: rows of Minv at D*D + row*D)
template <typename T, int D>
__device__ __forceinline__ void to_computational(const TView<T> &t, T *v) {
    T r[D];
#pragma unroll
    for (int i = 0; i < D; ++i) {
        T s = 0;
#pragma unroll
        for (int k = 0; k < D; ++k) s += t[D * D + i * D + k] * v[k];
        r[i] = s;
    }
#pragma unroll
    for (int i = 0; i < D; ++i) v[i] = r[i];
}

template <typename T, int D>
__device__ __forceinline__ T velocity_measure(const T *v, int norm) {
    if (norm) {
        T s = 0;
#pragma unroll
        for (int i = 0; i < D; ++i) s += v[i] * v[i];
        return dsqrt<T>(s);
    }
    T m = v[0] < T(0) ? -v[0] : v[0];
#pragma unroll
    for (int i = 1; i < D; ++i) m = nan_max<T>(m, v[i] < T(0) ? -v[i] : v[i]);
    return m;
}

// Block::getMaxVelocity / getMaxVelocityPerEnv / getMaxVelocityMagnitude, block.cpp:505-600:
// max over the block's cells and, with bounds, the face cells of its fixed faces of the
// largest absolute component (norm 0) or magnitude (norm 1) of the velocity, optionally in
// computational space (Minv u: block.cpp:267-273, boundaries.cpp:302-318). A block's
// workgroups sweep its cells followed by its fixed faces' face cells; every workgroup folds
// its maximum into out[env] and out[batch] (the maximum over environments) with one atomic
// max each. out must hold zeros (k_fill) beforehand.
template <typename T, int D>
__global__ void __launch_bounds__(ASM_BLOCK) k_max_velocity(pp_domain dom, int firstBlock, int withBounds,
                                                            int computational, int norm, T *__restrict__ out) {
    const int wpb = (dom.max_block_cells + ASM_BLOCK - 1) / ASM_BLOCK;
    const int bi = firstBlock + __builtin_amdgcn_readfirstlane(blockIdx.x / wpb);
    const int tile = blockIdx.x - (bi - firstBlock) * wpb;
    const pp_block_desc &b = dom.blocks[bi];
    const int env = blockIdx.y;
    const int cells = b.cells;
    int faceStart[2 * D + 1];
    int units = cells;
#pragma unroll
    for (int f = 0; f < 2 * D; ++f) {
        faceStart[f] = units;
        if (withBounds && b.bnd[f].type == PP_BOUND_FIXED) units += cells / b.size[f >> 1];
    }
    faceStart[2 * D] = units;
    const T *vel = reinterpret_cast<const T *>(b.velocity) + (size_t)env * D * cells;
    const bool cellTransform = computational && b.has_transform;
    T m = 0;
    for (int u = tile * ASM_BLOCK + threadIdx.x; u < units; u += wpb * ASM_BLOCK) {
        T v[D];
        if (u < cells) {
#pragma unroll
            for (int c = 0; c < D; ++c) v[c] = vel[(size_t)c * cells + u];
            if (cellTransform) to_computational<T, D>(Metric<T, D>::at(b, u), v);
        } else {
            int f = 0;
#pragma unroll
            for (int k = 1; k < 2 * D; ++k)
                if (u >= faceStart[k]) f = k;
            const pp_boundary_desc &fb = b.bnd[f];
            const int ff = u - faceStart[f];
            const int fc = faceStart[f + 1] - faceStart[f];
            const int vb = fb.vel_batched ? env : 0;
            const T *fv = reinterpret_cast<const T *>(fb.vel);
#pragma unroll
            for (int c = 0; c < D; ++c)
                v[c] = fb.vel_static ? fv[(size_t)vb * D + c] : fv[((size_t)vb * D + c) * fc + ff];
            if (computational && fb.has_transform) to_computational<T, D>(Metric<T, D>::face(fb, ff), v);
        }
        m = nan_max<T>(m, velocity_measure<T, D>(v, norm));
    }
    // wavefront, then workgroup maximum
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = nan_max<T>(m, __shfl_xor(m, o, 64));
    __shared__ T wmax[ASM_BLOCK / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) wmax[wave] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        T r = wmax[0];
#pragma unroll
        for (int w = 1; w < ASM_BLOCK / 64; ++w) r = nan_max<T>(r, wmax[w]);
        fold_max(out + env, r);
        fold_max(out + dom.batch, r);
    }
}

// TransformVectors, transform_vectors.cu:43-136: out[n][i][c] = sum_k T[i][k] v[n][k][c] with
// T = Minv (inverse) or M of cell c; vectors [N][D][cells], transforms [1|N][cells][2D^2+1]
template <typename T, int D>
__global__ void __launch_bounds__(256) k_transform_vectors(const T *__restrict__ v, const T *__restrict__ tr,
                                                           int cells, int batchedTransforms, int inverse,
                                                           T *__restrict__ out) {
    const int n = blockIdx.y;
    const T *vn = v + (size_t)n * D * cells;
    T *on = out + (size_t)n * D * cells;
    const T *tn = tr + (batchedTransforms ? (size_t)n * cells * (2 * D * D + 1) : 0);
    const int off = inverse ? D * D : 0;
    for (int c = blockIdx.x * 256 + threadIdx.x; c < cells; c += gridDim.x * 256) {
        const T *t = tn + (size_t)c * (2 * D * D + 1) + off;
        T a[D];
#pragma unroll
        for (int k = 0; k < D; ++k) a[k] = vn[(size_t)k * cells + c];
#pragma unroll
        for (int i = 0; i < D; ++i) {
            T s = 0;
#pragma unroll
            for (int k = 0; k < D; ++k) s += t[i * D + k] * a[k];
            on[(size_t)i * cells + c] = s;
        }
    }
}

static int check_domain(const pp_domain *dom, const char *fn) {
    if (!dom || !dom->blocks || dom->num_blocks <= 0 || dom->num_cells <= 0 || dom->batch <= 0 ||
        dom->max_block_cells <= 0 ||
        (dom->dims != 2 && dom->dims != 3) || (dom->dtype != PP_F32 && dom->dtype != PP_F64)) {
        set_error(std::string(fn) + ": invalid domain descriptor");
        return PP_ERR_ARG;
    }
    if (dom->batch > 65535) {
        set_error(std::string(fn) + ": batch exceeds the grid limit");
        return PP_ERR_ARG;
    }
    return PP_OK;
}

static inline dim3 cell_grid(const pp_domain *dom, int envs) {
    return dim3((dom->num_cells + ASM_BLOCK - 1) / ASM_BLOCK, envs);
}
// grid of CELL_PROLOGUE kernels: ceil(max_block_cells / ASM_BLOCK) workgroups per block
static inline dim3 block_grid(const pp_domain *dom, int envs) {
    const int wpb = (dom->max_block_cells + ASM_BLOCK - 1) / ASM_BLOCK;
    return dim3(wpb * dom->num_blocks, envs);
}

#define DIMS_DISPATCH(DIMV, ...)                                   \
    if ((DIMV) == 2) { constexpr int D = 2; __VA_ARGS__; }         \
    else { constexpr int D = 3; __VA_ARGS__; }

} // namespace pp

using namespace pp;

extern "C" {

int pp_transpose_2d(const void *in, int32_t rows, int32_t cols, int32_t dtype, void *out, void *stream) {
    PP_CHECK_ARG(in && out && rows >= 0 && cols >= 0, "bad arguments");
    if (rows == 0 || cols == 0) return PP_OK;
    PP_CHECK_ARG((cols + 31) / 32 <= 65535, "too many columns for the grid");
    const dim3 g((rows + 31) / 32, (cols + 31) / 32);
    PP_DISPATCH(dtype, hipLaunchKernelGGL((k_transpose_2d<T>), g, dim3(256), 0, as_stream(stream),
                                          reinterpret_cast<const T *>(in), rows, cols, reinterpret_cast<T *>(out)));
    PP_LAUNCH_CHECK();
    return PP_OK;
}

int pp_csr_pattern(const pp_domain *dom, int32_t *row, int32_t *col, void *stream) {
    if (int e = check_domain(dom, __func__)) return e;
    PP_CHECK_ARG(row && col, "row/col must not be null");
    DIMS_DISPATCH(dom->dims,
        hipLaunchKernelGGL((k_csr_pattern<D>), cell_grid(dom, 1), dim3(ASM_BLOCK), 0, as_stream(stream), *dom, row, col));
    PP_LAUNCH_CHECK();
    return PP_OK;
}

// Flags without a corner term on either side (0, or only NON_ORTHO_DIAGONAL_MATRIX, which the
// reference does not implement) take the orthogonal path.
static inline int matrix_flags(int32_t f) { return f & (NO_DIRECT_MATRIX | NO_CENTER_MATRIX); }
static inline int rhs_flags(int32_t f) { return f & (NO_DIRECT_RHS | NO_DIAGONAL_RHS); }

int pp_setup_advection_matrix(const pp_domain *dom, int32_t non_ortho_flags, void *stream) {
    if (int e = check_domain(dom, __func__)) return e;
    const int flags = matrix_flags(non_ortho_flags);
    PP_CHECK_ARG(dom->C.val && dom->C.row && dom->A && dom->time_step, "C, A and time_step must be set");
    PP_DISPATCH(dom->dtype, DIMS_DISPATCH(dom->dims,
        PP_TIMED(PP_K_ADVECTION_MATRIX, dom->num_cells, as_stream(stream), hipLaunchKernelGGL((flags ? k_advection_matrix<T, D, true> : k_advection_matrix<T, D, false>), block_grid(dom, dom->batch), dim3(ASM_BLOCK), 0, as_stream(stream), *dom, -1, flags))));
    PP_LAUNCH_CHECK();
    return PP_OK;
}

int pp_setup_advection_matrix_scalar(const pp_domain *dom, int32_t non_ortho_flags, int32_t channel, void *stream) {
    if (int e = check_domain(dom, __func__)) return e;
    const int flags = matrix_flags(non_ortho_flags);
    PP_CHECK_ARG(dom->C.val && dom->C.row && dom->A && dom->time_step, "C, A and time_step must be set");
    PP_CHECK_ARG(channel >= 0 && channel < dom->scalar_channels, "passive scalar channel out of range");
    PP_CHECK_ARG(dom->scalar_visc_count == 0 || dom->scalar_viscosity, "scalar viscosity not set");
    PP_DISPATCH(dom->dtype, DIMS_DISPATCH(dom->dims,
        hipLaunchKernelGGL((flags ? k_advection_matrix<T, D, true> : k_advection_matrix<T, D, false>), block_grid(dom, dom->batch), dim3(ASM_BLOCK), 0, as_stream(stream), *dom, (int)channel, flags)));
    PP_LAUNCH_CHECK();
    return PP_OK;
}

int pp_setup_advection_scalar(const pp_domain *dom, int32_t non_ortho_flags, void *stream) {
    if (int e = check_domain(dom, __func__)) return e;
    const int flags = rhs_flags(non_ortho_flags);
    PP_CHECK_ARG(dom->scalar_channels > 0 && dom->scalar_rhs && dom->time_step, "passive scalar fields must be set");
    PP_CHECK_ARG(!flags || dom->scalar_result, "the non-orthogonal terms read scalar_result");
    PP_CHECK_ARG(dom->scalar_visc_count == 0 || dom->scalar_viscosity, "scalar viscosity not set");
    PP_DISPATCH(dom->dtype, DIMS_DISPATCH(dom->dims,
        hipLaunchKernelGGL((flags ? k_scalar_rhs<T, D, true> : k_scalar_rhs<T, D, false>), block_grid(dom, dom->batch), dim3(ASM_BLOCK), 0, as_stream(stream), *dom, flags)));
    PP_LAUNCH_CHECK();
    return PP_OK;
}

int pp_setup_advection_velocity(const pp_domain *dom, int32_t non_ortho_flags, int32_t apply_pressure_gradient, void *stream) {
    if (int e = check_domain(dom, __func__)) return e;
    const int flags = rhs_flags(non_ortho_flags);
    PP_CHECK_ARG(dom->velocity_rhs && dom->time_step, "velocity_rhs and time_step must be set");
    PP_CHECK_ARG(!flags || dom->velocity_result, "the non-orthogonal terms read velocity_result");
    PP_DISPATCH(dom->dtype, DIMS_DISPATCH(dom->dims,
        PP_TIMED(PP_K_ADVECTION_RHS, dom->num_cells, as_stream(stream), hipLaunchKernelGGL((flags ? k_advection_rhs<T, D, true> : k_advection_rhs<T, D, false>), block_grid(dom, dom->batch), dim3(ASM_BLOCK), 0, as_stream(stream), *dom,
                           (int)apply_pressure_gradient, flags))));
    PP_LAUNCH_CHECK();
    return PP_OK;
}

int pp_setup_pressure_matrix(const pp_domain *dom, int32_t non_ortho_flags, void *stream) {
    if (int e = check_domain(dom, __func__)) return e;
    const int flags = matrix_flags(non_ortho_flags);
    PP_CHECK_ARG(dom->P.val && dom->P.row && dom->A, "P and A must be set");
    PP_DISPATCH(dom->dtype, DIMS_DISPATCH(dom->dims,
        PP_TIMED(PP_K_PRESSURE_MATRIX, dom->num_cells, as_stream(stream), hipLaunchKernelGGL((flags ? k_pressure_matrix<T, D, true> : k_pressure_matrix<T, D, false>), block_grid(dom, dom->batch), dim3(ASM_BLOCK), 0, as_stream(stream), *dom, flags))));
    PP_LAUNCH_CHECK();
    return PP_OK;
}

int pp_setup_pressure_rhs(const pp_domain *dom, int32_t non_ortho_flags, int32_t time_step_norm, void *stream) {
    if (int e = check_domain(dom, __func__)) return e;
    const int flags = rhs_flags(non_ortho_flags);
    PP_CHECK_ARG(dom->pressure_rhs && dom->pressure_rhs_div && dom->C.val && dom->velocity_result, "pressure fields must be set");
    PP_CHECK_ARG(!flags || (dom->pressure_result && dom->A), "the non-orthogonal terms read pressure_result and A");
    PP_DISPATCH(dom->dtype, DIMS_DISPATCH(dom->dims,
        PP_TIMED(PP_K_PRESSURE_RHS, dom->num_cells, as_stream(stream), hipLaunchKernelGGL((k_pressure_rhs<T, D>), block_grid(dom, dom->batch), dim3(ASM_BLOCK), 0, as_stream(stream), *dom));
        PP_TIMED(PP_K_PRESSURE_DIV, dom->num_cells, as_stream(stream), hipLaunchKernelGGL((flags ? k_pressure_div<T, D, true> : k_pressure_div<T, D, false>), block_grid(dom, dom->batch), dim3(ASM_BLOCK), 0, as_stream(stream), *dom,
                           (int)time_step_norm, flags))));
    PP_LAUNCH_CHECK();
    return PP_OK;
}

int pp_setup_pressure_rhs_div(const pp_domain *dom, int32_t non_ortho_flags, int32_t time_step_norm, void *stream) {
    if (int e = check_domain(dom, __func__)) return e;
    const int flags = rhs_flags(non_ortho_flags);
    PP_CHECK_ARG(dom->pressure_rhs && dom->pressure_rhs_div, "pressure fields must be set");
    PP_CHECK_ARG(!flags || (dom->pressure_result && dom->A), "the non-orthogonal terms read pressure_result and A");
    PP_DISPATCH(dom->dtype, DIMS_DISPATCH(dom->dims,
        PP_TIMED(PP_K_PRESSURE_DIV, dom->num_cells, as_stream(stream), hipLaunchKernelGGL((flags ? k_pressure_div<T, D, true> : k_pressure_div<T, D, false>), block_grid(dom, dom->batch), dim3(ASM_BLOCK), 0, as_stream(stream), *dom,
                           (int)time_step_norm, flags))));
    PP_LAUNCH_CHECK();
    return PP_OK;
}

int pp_correct_velocity(const pp_domain *dom, int32_t time_step_norm, void *stream) {
    if (int e = check_domain(dom, __func__)) return e;
    PP_CHECK_ARG(dom->velocity_result && dom->pressure_rhs && dom->A, "velocity_result, pressure_rhs and A must be set");
    PP_DISPATCH(dom->dtype, DIMS_DISPATCH(dom->dims,
        PP_TIMED(PP_K_CORRECT_VELOCITY, dom->num_cells, as_stream(stream), hipLaunchKernelGGL((k_correct_velocity<T, D>), block_grid(dom, dom->batch), dim3(ASM_BLOCK), 0, as_stream(stream), *dom,
                           (int)time_step_norm))));
    PP_LAUNCH_CHECK();
    return PP_OK;
}

int pp_setup_epot_matrix(const pp_domain *dom, int32_t non_ortho_flags, void *stream) {
    if (int e = check_domain(dom, __func__)) return e;
    PP_CHECK_ARG(dom->E.val && dom->E.row, "E must be set");
    const int flags = non_ortho_flags & (NO_DIRECT_MATRIX | NO_CENTER_MATRIX);
    PP_CHECK_ARG(!flags || dom->A, "the potential's corner terms need the advection diagonal A");
    PP_DISPATCH(dom->dtype, DIMS_DISPATCH(dom->dims,
        hipLaunchKernelGGL((k_epot_matrix<T, D>), block_grid(dom, 1), dim3(ASM_BLOCK), 0, as_stream(stream), *dom,
                           flags)));
    PP_LAUNCH_CHECK();
    return PP_OK;
}

int pp_compute_epot_rhs(const pp_domain *dom, const void *vector_field, void *divergence, void *stream) {
    if (int e = check_domain(dom, __func__)) return e;
    PP_CHECK_ARG(vector_field && divergence, "vector_field and divergence must not be null");
    PP_DISPATCH(dom->dtype, DIMS_DISPATCH(dom->dims,
        PP_TIMED(PP_K_EPOT_RHS, dom->num_cells, as_stream(stream), hipLaunchKernelGGL((k_epot_rhs<T, D>), block_grid(dom, dom->batch), dim3(ASM_BLOCK), 0, as_stream(stream), *dom,
                           reinterpret_cast<const T *>(vector_field), reinterpret_cast<T *>(divergence)))));
    PP_LAUNCH_CHECK();
    return PP_OK;
}

int pp_compute_current_density(const pp_domain *dom, const void *phi, const void *uxb, void *j_out, void *stream) {
    if (int e = check_domain(dom, __func__)) return e;
    PP_CHECK_ARG(phi && uxb && j_out, "phi, uxb and j_out must not be null");
    PP_DISPATCH(dom->dtype, DIMS_DISPATCH(dom->dims,
        PP_TIMED(PP_K_CURRENT_DENSITY, dom->num_cells, as_stream(stream), hipLaunchKernelGGL((k_current_density<T, D>), block_grid(dom, dom->batch), dim3(ASM_BLOCK), 0, as_stream(stream), *dom,
                           reinterpret_cast<const T *>(phi), reinterpret_cast<const T *>(uxb), reinterpret_cast<T *>(j_out)))));
    PP_LAUNCH_CHECK();
    return PP_OK;
}

int pp_copy_result(const pp_domain *dom, int32_t field, int32_t to_blocks, void *stream) {
    if (int e = check_domain(dom, __func__)) return e;
    PP_CHECK_ARG(field >= 0 && field <= 3, "field must be 0, 1, 2 or 3");
    void *flat = field == 0 ? dom->velocity_result
                            : (field == 1 ? dom->pressure_result : (field == 2 ? dom->epot_result : dom->scalar_result));
    PP_CHECK_ARG(flat != nullptr, "result vector not set");
    PP_CHECK_ARG(field != 3 || dom->scalar_channels > 0, "the domain has no passive scalar");
    const int ch = field == 0 ? dom->dims : (field == 3 ? dom->scalar_channels : 1);
    PP_DISPATCH(dom->dtype,
        PP_TIMED(PP_K_COPY_RESULT, dom->num_cells, as_stream(stream), hipLaunchKernelGGL((k_copy_result<T>), block_grid(dom, dom->batch), dim3(ASM_BLOCK), 0, as_stream(stream), *dom, (int)field, ch,
                           (int)to_blocks, reinterpret_cast<const T *>(flat), reinterpret_cast<T *>(flat))));
    PP_LAUNCH_CHECK();
    return PP_OK;
}

#define PP_GRAD_LAUNCH(KERNEL, ...)                                                                              \
    PP_DISPATCH(dom->dtype, DIMS_DISPATCH(dom->dims,                                                            \
        hipLaunchKernelGGL((KERNEL<T, D>), block_grid(dom, dom->batch), dim3(ASM_BLOCK), 0, as_stream(stream), \
                           *dom, __VA_ARGS__)))
#define PP_T(X) reinterpret_cast<T *>(X)
#define PP_CT(X) reinterpret_cast<const T *>(X)

int pp_setup_advection_matrix_grad(const pp_domain *dom, const void *grad_c, const void *grad_a, void *grad_u,
                                   void *grad_visc, void *stream) {
    if (int e = check_domain(dom, __func__)) return e;
    PP_CHECK_ARG(grad_c && grad_u && dom->C.row, "grad_c, grad_u and the C pattern must be set");
    PP_GRAD_LAUNCH(k_advection_matrix_grad, PP_CT(grad_c), PP_CT(grad_a), PP_T(grad_u), PP_T(grad_visc));
    PP_LAUNCH_CHECK();
    return PP_OK;
}

int pp_setup_advection_velocity_grad(const pp_domain *dom, const void *grad_rhs, void *grad_u, void *grad_source,
                                     void *const *grad_bound, void *grad_visc, void *stream) {
    if (int e = check_domain(dom, __func__)) return e;
    PP_CHECK_ARG(grad_rhs && grad_u, "grad_rhs and grad_u must be set");
    PP_GRAD_LAUNCH(k_advection_rhs_grad, PP_CT(grad_rhs), PP_T(grad_u), PP_T(grad_source),
                   reinterpret_cast<T *const *>(grad_bound), PP_T(grad_visc));
    PP_LAUNCH_CHECK();
    return PP_OK;
}

int pp_setup_pressure_matrix_grad(const pp_domain *dom, const void *a, const void *grad_p, void *grad_a,
                                  void *stream) {
    if (int e = check_domain(dom, __func__)) return e;
    PP_CHECK_ARG(a && grad_p && grad_a && dom->P.row, "a, grad_p, grad_a and the P pattern must be set");
    PP_GRAD_LAUNCH(k_pressure_matrix_grad, PP_CT(a), PP_CT(grad_p), PP_T(grad_a));
    PP_LAUNCH_CHECK();
    return PP_OK;
}

int pp_setup_pressure_rhs_grad(const pp_domain *dom, const void *grad_hbya, const void *a, const void *c_val,
                               const void *velocity_result, const void *hbya, void *grad_a, void *grad_c,
                               void *grad_velocity_result, void *grad_u, void *grad_source, void *const *grad_bound,
                               void *grad_visc, void *stream) {
    if (int e = check_domain(dom, __func__)) return e;
    PP_CHECK_ARG(grad_hbya && a && c_val && velocity_result && hbya && grad_a && grad_c && grad_velocity_result &&
                     grad_u && dom->C.row,
                 "null argument");
    PP_GRAD_LAUNCH(k_pressure_rhs_grad, PP_CT(grad_hbya), PP_CT(a), PP_CT(c_val), PP_CT(velocity_result), PP_CT(hbya),
                   PP_T(grad_a), PP_T(grad_c), PP_T(grad_velocity_result), PP_T(grad_u), PP_T(grad_source),
                   reinterpret_cast<T *const *>(grad_bound), PP_T(grad_visc));
    PP_LAUNCH_CHECK();
    return PP_OK;
}

int pp_setup_pressure_rhs_div_grad(const pp_domain *dom, const void *grad_div, int32_t time_step_norm, void *grad_hbya,
                                   void *const *grad_bound, void *stream) {
    if (int e = check_domain(dom, __func__)) return e;
    PP_CHECK_ARG(grad_div && grad_hbya, "grad_div and grad_hbya must be set");
    PP_GRAD_LAUNCH(k_pressure_div_grad, PP_CT(grad_div), (int)time_step_norm, PP_T(grad_hbya),
                   reinterpret_cast<T *const *>(grad_bound));
    PP_LAUNCH_CHECK();
    return PP_OK;
}

int pp_correct_velocity_grad(const pp_domain *dom, const void *grad_out, const void *a, int32_t time_step_norm,
                             void *grad_hbya, void *grad_a, void *grad_pressure, void *stream) {
    if (int e = check_domain(dom, __func__)) return e;
    PP_CHECK_ARG(grad_out && a && grad_hbya && grad_a && grad_pressure, "null argument");
    PP_GRAD_LAUNCH(k_correct_velocity_grad, PP_CT(grad_out), PP_CT(a), (int)time_step_norm, PP_T(grad_hbya),
                   PP_T(grad_a), PP_T(grad_pressure));
    PP_LAUNCH_CHECK();
    return PP_OK;
}

int pp_compute_epot_rhs_grad(const pp_domain *dom, const void *grad_div, void *grad_field, void *const *grad_values,
                             void *stream) {
    if (int e = check_domain(dom, __func__)) return e;
    PP_CHECK_ARG(grad_div && grad_field, "null argument");
    PP_GRAD_LAUNCH(k_epot_rhs_grad, PP_CT(grad_div), PP_T(grad_field), reinterpret_cast<T *const *>(grad_values));
    PP_LAUNCH_CHECK();
    return PP_OK;
}

int pp_compute_current_density_grad(const pp_domain *dom, const void *grad_j, void *grad_phi, void *grad_uxb,
                                    void *const *grad_values, void *stream) {
    if (int e = check_domain(dom, __func__)) return e;
    PP_CHECK_ARG(grad_j && grad_phi && grad_uxb, "null argument");
    PP_GRAD_LAUNCH(k_current_density_grad, PP_CT(grad_j), PP_T(grad_phi), PP_T(grad_uxb),
                   reinterpret_cast<T *const *>(grad_values));
    PP_LAUNCH_CHECK();
    return PP_OK;
}
#undef PP_GRAD_LAUNCH
#undef PP_T
#undef PP_CT

int pp_max_velocity(const pp_domain *dom, int32_t block, int32_t with_bounds, int32_t computational, int32_t norm,
                    void *out, void *stream) {
    if (int e = check_domain(dom, __func__)) return e;
    PP_CHECK_ARG(out, "out must not be null");
    PP_CHECK_ARG(block >= -1 && block < dom->num_blocks, "block index out of range");
    PP_CHECK_ARG(norm == 0 || norm == 1, "norm must be 0 or 1");
    const int first = block < 0 ? 0 : block;
    const int nblocks = block < 0 ? dom->num_blocks : 1;
    const int wpb = (dom->max_block_cells + ASM_BLOCK - 1) / ASM_BLOCK;
    PP_DISPATCH(dom->dtype, DIMS_DISPATCH(dom->dims,
        hipLaunchKernelGGL((k_fill<T>), dim3(1), dim3(256), 0, as_stream(stream), static_cast<T *>(out),
                           dom->batch + 1, T(0));
        hipLaunchKernelGGL((k_max_velocity<T, D>), dim3(wpb * nblocks, dom->batch), dim3(ASM_BLOCK), 0,
                           as_stream(stream), *dom, first, (int)(with_bounds != 0), (int)(computational != 0),
                           (int)norm, static_cast<T *>(out))));
    PP_LAUNCH_CHECK();
    return PP_OK;
}

int pp_transform_vectors(int32_t dims, int32_t batch, int32_t cells, const void *vectors, const void *transforms,
                         int32_t batched_transforms, int32_t inverse, void *out, int32_t dtype, void *stream) {
    PP_CHECK_ARG(dims == 2 || dims == 3, "dims must be 2 or 3");
    PP_CHECK_ARG(batch >= 1 && batch <= 65535 && cells >= 0, "bad sizes");
    PP_CHECK_ARG(vectors && transforms && out, "null argument");
    PP_CHECK_ARG(out != vectors, "out must not alias the vectors");
    if (cells == 0) return PP_OK;
    const dim3 g(std::min((cells + 255) / 256, 4096), batch);
    PP_DISPATCH(dtype, DIMS_DISPATCH(dims,
        hipLaunchKernelGGL((k_transform_vectors<T, D>), g, dim3(256), 0, as_stream(stream),
                           static_cast<const T *>(vectors), static_cast<const T *>(transforms), (int)cells,
                           (int)(batched_transforms != 0), (int)(inverse != 0), static_cast<T *>(out))));
    PP_LAUNCH_CHECK();
    return PP_OK;
}

int pp_coords_to_transforms(int32_t dims, const int32_t *cells, const void *coords, void *out, int32_t dtype, void *stream) {
    PP_CHECK_ARG(dims == 2 || dims == 3, "dims must be 2 or 3");
    PP_CHECK_ARG(cells && coords && out, "null argument");
    const int cx = cells[0], cy = dims > 1 ? cells[1] : 1, cz = dims > 2 ? cells[2] : 1;
    PP_CHECK_ARG(cx > 0 && cy > 0 && cz > 0, "cell counts must be positive");
    const int n = cx * cy * cz;
    PP_DISPATCH(dtype, DIMS_DISPATCH(dims,
        hipLaunchKernelGGL((k_coords_to_transforms<T, D>), dim3((n + ASM_BLOCK - 1) / ASM_BLOCK), dim3(ASM_BLOCK), 0, as_stream(stream),
                           reinterpret_cast<const T *>(coords), cx, cy, cz, reinterpret_cast<T *>(out))));
    PP_LAUNCH_CHECK();
    return PP_OK;
}

int pp_coords_to_face_transforms(int32_t dims, const int32_t *cells, const void *coords, void *out, int32_t dtype, void *stream) {
    PP_CHECK_ARG(dims == 2 || dims == 3, "dims must be 2 or 3");
    PP_CHECK_ARG(cells && coords && out, "null argument");
    const int cx = cells[0], cy = dims > 1 ? cells[1] : 1, cz = dims > 2 ? cells[2] : 1;
    PP_CHECK_ARG(cx > 0 && cy > 0 && cz > 0, "cell counts must be positive");
    const int nv = (cx + 1) * (dims > 1 ? cy + 1 : 1) * (dims > 2 ? cz + 1 : 1);
    PP_DISPATCH(dtype, DIMS_DISPATCH(dims,
        hipLaunchKernelGGL((k_coords_to_face_transforms<T, D>), dim3((nv + ASM_BLOCK - 1) / ASM_BLOCK), dim3(ASM_BLOCK), 0,
                           as_stream(stream), reinterpret_cast<const T *>(coords), cx, cy, cz, reinterpret_cast<T *>(out))));
    PP_LAUNCH_CHECK();
    return PP_OK;
}

} // extern "C"

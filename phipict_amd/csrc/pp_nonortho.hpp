// Non-orthogonal Laplace terms of skewed grids: the tangential metric coefficients of a face,
// the walk around a cell's corner, and the corner terms the matrices take implicitly and the
// right-hand sides explicitly (nonortho_laplace.cuh:53-131, block_data.cuh:280-430, :477-616,
// nonortho_rhs.cuh:28-200). Only the flag bits of the reference's default mode and
// NON_ORTHO_DIRECT_RHS are implemented; NON_ORTHO_DIAGONAL_MATRIX is not implemented there
// either.
#pragma once
#include "pp_common.hpp"

namespace pp {

enum { NO_DIRECT_MATRIX = 1, NO_DIRECT_RHS = 2, NO_DIAGONAL_MATRIX = 4, NO_DIAGONAL_RHS = 8, NO_CENTER_MATRIX = 16 };

// The data a corner term reads (GridDataType of the reference). The matrix kinds read only
// the boundary condition types of fixed faces (velocity, or scalar channel w), no values.
enum { NK_FIXED_ONLY = 0, NK_VELOCITY = 1, NK_PRESSURE = 2, NK_SCALAR = 3, NK_VELOCITY_TYPES = 4, NK_SCALAR_TYPES = 5 };

// alpha_{c1 c2} = det Minv[c1].Minv[c2] (getLaplaceCoefficientsFull / SingleAxis,
// discretization.cuh:419-470).
template <typename T, int D>
__device__ __forceinline__ T lap_coef_t(const TView<T> &t, int c1, int c2) {
    T r = T(0);
#pragma unroll
    for (int i = 0; i < D; ++i) r += t[D * D + c1 * D + i] * t[D * D + c2 * D + i];
    return t[2 * D * D] * r;
}
template <typename T, int D>
__device__ __forceinline__ T lap_coef(const pp_block_desc &b, int f, int c1, int c2) {
    if (!b.has_transform) return c1 == c2 ? T(1) : T(0);
    return lap_coef_t<T, D>(Metric<T, D>::at(b, f), c1, c2);
}

// computeConnectedDir, topology.cuh:236-244: the face-format direction `dir` seen from the
// block across connection cb of face axis bdim.
__device__ __forceinline__ int conn_dir(int d, int dir, int bdim, const pp_boundary_desc &cb) {
    return cb.conn_axes[pos_mod((dir >> 1) - bdim, d)] ^ (dir & 1);
}

// Where the walk is: block and cell.
struct WalkPos {
    int bi;
    int p[3];
};

// Reads of one data kind: cell values (result vectors), values on fixed faces, and the kind's
// boundary condition type on a fixed face.
template <typename T, int D, int KIND>
struct CornerData {
    const pp_domain &dom;
    const pp_block_desc *__restrict__ blocks;
    int w;   // velocity component or scalar channel
    int env;
    __device__ T cell(int bi, const int *p) const {
        const pp_block_desc &b = blocks[bi];
        const size_t g = (size_t)b.global_offset + flat3(p, b.size);
        const size_t n = dom.num_cells;
        if (KIND == NK_VELOCITY)
            return reinterpret_cast<const T *>(dom.velocity_result)[((size_t)env * D + w) * n + g];
        if (KIND == NK_PRESSURE) return reinterpret_cast<const T *>(dom.pressure_result)[(size_t)env * n + g];
        if (KIND == NK_SCALAR)
            return reinterpret_cast<const T *>(dom.scalar_result)[((size_t)env * dom.scalar_channels + w) * n + g];
        return T(0);
    }
    // getFixedBoundaryData (block_data.cuh:136-199): the reference keeps no pressure values on
    // fixed faces, so the pressure kind reads 0 there.
    __device__ T fixed(int bi, int bound, const int *p) const {
        const pp_block_desc &b = blocks[bi];
        const pp_boundary_desc &fb = b.bnd[bound];
        const int fc = face_cells(b, bound), ff = face_flat(p, bound, b);
        if (KIND == NK_VELOCITY) return fixed_vel<T>(fb, D, fc, ff, w, env);
        if (KIND == NK_SCALAR) return fixed_scalar<T>(fb, dom.scalar_channels, fc, ff, w, env);
        return T(0);
    }
    // getFixedBoundaryType (block_data.cuh:110-134)
    __device__ bool neumann(int bi, int bound) const {
        const pp_boundary_desc &fb = blocks[bi].bnd[bound];
        if (KIND == NK_VELOCITY || KIND == NK_VELOCITY_TYPES) return fb.vel_type != PP_BC_DIRICHLET;
        if (KIND == NK_SCALAR || KIND == NK_SCALAR_TYPES) return scalar_neumann(fb, w);
        return false;
    }
};

template <typename T>
struct CornerValue {
    T data;
    int cells;    // cells the value averages; 0 when it comes from a fixed face
    bool neumann; // boundary condition type of that fixed face
};

// getCornerValue, block_data.cuh:501-616, without the center cell (includeDepth0 false) and to
// depth 2: the two walks start towards dir1 and dir2 and turn around the corner until they
// meet, reach a fixed face (whose values give the corner value), or end. Data of depth-1
// cells is summed only with `depth1`. A walk across a block connection lands on the second
// cell layer of the connected block (computeConnectedPos with border offset 1), as in the
// reference.
template <typename T, int D, int KIND>
__device__ CornerValue<T> corner_value(const CornerData<T, D, KIND> &cd, int bi, const int *p, int dir1, int dir2,
                                       bool depth1) {
    const pp_block_desc *__restrict__ blocks = cd.blocks;
    T data = T(0);
    int num = 1;
    bool neu = false;
    WalkPos cyc[2];
    int cdir[2][2] = {{dir1, dir2}, {dir2, dir1}};
#pragma unroll
    for (int d = 0; d < 2; ++d) {
        cyc[d].bi = bi;
        cyc[d].p[0] = p[0]; cyc[d].p[1] = p[1]; cyc[d].p[2] = p[2];
    }
    for (int depth = 1; depth <= 2; ++depth) {
        for (int d = 0; d < 2; ++d) {
            const int da = cdir[d][0], db = cdir[d][1];
            const int axis = da >> 1;
            const pp_block_desc &cb = blocks[cyc[d].bi];
            if (at_bound(cyc[d].p, da, cb)) {
                const pp_boundary_desc &bd = cb.bnd[da];
                if (bd.type == PP_BOUND_FIXED) {
                    neu = cd.neumann(cyc[d].bi, da);
                    if (at_bound(cyc[d].p, db, cb)) return {cd.fixed(cyc[d].bi, da, cyc[d].p), 0, neu};
                    int q[3] = {cyc[d].p[0], cyc[d].p[1], cyc[d].p[2]};
                    q[db >> 1] += face_sign(db);
                    return {(cd.fixed(cyc[d].bi, da, cyc[d].p) + cd.fixed(cyc[d].bi, da, q)) * T(0.5), 0, neu};
                } else if (bd.type == PP_BOUND_CONNECTED) {
                    int q[3];
                    conn_pos(blocks, D, cyc[d].p, axis, bd, 1, q);
                    cyc[d].bi = bd.conn_block;
                    cyc[d].p[0] = q[0]; cyc[d].p[1] = q[1]; cyc[d].p[2] = q[2];
                    cdir[d][0] = conn_dir(D, db, axis, bd);
                    cdir[d][1] = conn_dir(D, da, axis, bd) ^ 1;
                } else if (bd.type == PP_BOUND_PERIODIC) {
                    cyc[d].p[axis] = (da & 1) ? 0 : cb.size[axis] - 1;
                    cdir[d][0] = db;
                    cdir[d][1] = da ^ 1;
                }
            } else {
                cyc[d].p[axis] += face_sign(da);
                cdir[d][0] = db;
                cdir[d][1] = da ^ 1;
            }
            const int o = d ^ 1;
            if (cyc[d].bi == cyc[o].bi && cyc[d].p[0] == cyc[o].p[0] && cyc[d].p[1] == cyc[o].p[1] &&
                cyc[d].p[2] == cyc[o].p[2])
                return {data / T(num), num, neu};
            if (depth > 1 || depth1) data += cd.cell(cyc[d].bi, cyc[d].p);
            ++num;
        }
    }
    return {data / T(num), num, neu};
}

// getBlockDataNeighbor (zeroBound false), block_data.cuh:280-320
template <typename T, int D, int KIND>
__device__ T neighbor_data(const CornerData<T, D, KIND> &cd, int bi, const int *p, int dir) {
    const pp_block_desc &b = cd.blocks[bi];
    const int axis = dir >> 1;
    int q[3] = {p[0], p[1], p[2]};
    if (at_bound(p, dir, b)) {
        const pp_boundary_desc &bd = b.bnd[dir];
        if (bd.type == PP_BOUND_FIXED) return cd.fixed(bi, dir, p);
        if (bd.type == PP_BOUND_CONNECTED) {
            conn_pos(cd.blocks, D, p, axis, bd, 1, q);
            return cd.cell(bd.conn_block, q);
        }
        if (bd.type == PP_BOUND_PERIODIC) {
            q[axis] = (dir & 1) ? 0 : b.size[axis] - 1;
            return cd.cell(bi, q);
        }
        return T(0);
    }
    q[axis] += face_sign(dir);
    return cd.cell(bi, q);
}

// getBlockDataNeighborDiagonal (zeroBound false), block_data.cuh:372-420: two steps, the
// prescribed direction second.
template <typename T, int D, int KIND>
__device__ T diagonal_data(const CornerData<T, D, KIND> &cd, int bi, const int *p, int dir1, int dir2) {
    const bool first_fixed = is_fixed(cd.blocks[bi], dir1);
    const int dirs[2] = {first_fixed ? dir2 : dir1, first_fixed ? dir1 : dir2};
    int cbi = bi;
    int q[3] = {p[0], p[1], p[2]};
    for (int i = 0; i < 2; ++i) {
        const int dir = dirs[i];
        const int axis = dir >> 1;
        const pp_block_desc &b = cd.blocks[cbi];
        if ((dir & 1) ? q[axis] == b.size[axis] - 1 : q[axis] == 0) {
            const pp_boundary_desc &bd = b.bnd[dir];
            if (bd.type == PP_BOUND_FIXED) return cd.fixed(cbi, dir, q);
            if (bd.type == PP_BOUND_CONNECTED) {
                int r[3];
                conn_pos(cd.blocks, D, q, axis, bd, 1, r);
                cbi = bd.conn_block;
                q[0] = r[0]; q[1] = r[1]; q[2] = r[2];
            } else if (bd.type == PP_BOUND_PERIODIC) {
                q[axis] = (dir & 1) ? 0 : b.size[axis] - 1;
            } else {
                return T(0);
            }
        } else {
            q[axis] += face_sign(dir);
        }
    }
    return cd.cell(cbi, q);
}

// Matrix variant of the tangential coefficients of face `face` towards axis tAxis
// (interpolateNonOrthoLaplaceComponents, nonortho_laplace.cuh:53-131, useFaceTransform false):
// (alpha_P[axis, tAxis] raP + alpha_N[tAxis] raN) / 2 with ra = (1/A) * viscosity as selected.
// As in the reference it is 0 for every face of a cell at either end of the face's axis when
// that face is prescribed, and the neighbour's coefficients use unmapped axes.
template <typename T, int D>
__device__ T face_alpha_matrix(const pp_domain &dom, const pp_block_desc *__restrict__ blocks, int bi, const int *p,
                               int f, int face, int tAxis, T raP, bool withVisc, const T *A) {
    const pp_block_desc &b = blocks[bi];
    const int axis = face >> 1;
    if (!((0 < p[axis] && p[axis] < b.size[axis] - 1) || !is_fixed(b, face))) return T(0);
    int np[3], nw;
    const int nbi = neighbor(blocks, bi, D, p, face, np, &nw);
    const pp_block_desc &nb = blocks[nbi];
    const int nf = flat3(np, nb.size);
    T raN = A ? T(1) / A[nb.global_offset + nf] : T(1);
    if (withVisc) raN = raN * visc_at<T>(dom, nb, nf);
    const T aP = lap_coef<T, D>(b, f, axis, tAxis);
    const T aN = lap_coef<T, D>(nb, nf, tAxis, axis);
    return (aP * raP + aN * raN) * T(0.5);
}

// Corner terms of face `bound` of a matrix row (PISO_build_matrix, advection.cu:177-240;
// PISO_build_pressure_matrix, pressure.cu:104-160). `sign` +1 adds them as the advection matrix
// does (diffusion positive on the diagonal), -1 as the pressure matrix (Laplacian). vscale
// multiplies the coefficient (the scalar diffusivity of a passive-scalar matrix).
template <typename T, int D, int KIND>
__device__ void matrix_corner_terms(const CornerData<T, D, KIND> &cd, int bi, const int *p, int f, int bound,
                                    int flags, T raP, bool withVisc, const T *A, T vscale, bool always_gradient,
                                    T sign, T &diag, T *rv) {
    const bool nbr = flags & NO_DIRECT_MATRIX, ctr = flags & NO_CENTER_MATRIX;
    if (D < 2 || !(nbr || ctr)) return;
    const int dim = bound >> 1;
    const int fs = face_sign(bound);
#pragma unroll
    for (int i = 1; i < D; ++i) {
        const int tAxis = (dim + i) % D;
        const T alpha = face_alpha_matrix<T, D>(cd.dom, cd.blocks, bi, p, f, bound, tAxis, raP, withVisc, A) * vscale;
        if (alpha == T(0)) continue;
#pragma unroll
        for (int tUp = 0; tUp < 2; ++tUp) {
            const int tFace = tAxis * 2 + tUp;
            const int tfs = face_sign(tFace);
            const CornerValue<T> cv = corner_value<T, D, KIND>(cd, bi, p, bound, tFace, false);
            if (cv.cells < 1) {
                if (always_gradient || cv.neumann) {
                    const T c = T(fs * tfs) * alpha * T(0.25);
                    if (ctr) diag -= sign * (3 * c);
                    if (nbr) {
                        rv[bound + 1] -= sign * (3 * c);
                        rv[(tFace ^ 1) + 1] += sign * c;
                    }
                }
            } else {
                const T c = T(fs * tfs) * alpha * (T(1.0) / T(cv.cells));
                if (ctr) diag -= sign * c;
                if (nbr) {
                    rv[bound + 1] -= sign * c;
                    rv[tFace + 1] -= sign * c;
                }
            }
        }
    }
}

// Explicit corner terms of a right-hand side (getNonOrthoLaplaceRHS_v2, nonortho_rhs.cuh:28-200,
// useFaceTransform false), for the Laplace term positive on the left-hand side. withVisc: ra
// includes the viscosity (velocity); A: ra includes 1/A (pressure).
template <typename T, int D, int KIND>
__device__ T nonortho_rhs(const CornerData<T, D, KIND> &cd, int bi, const int *p, int f, int flags, bool withVisc,
                          const T *A) {
    if (D < 2 || !(flags & (NO_DIRECT_RHS | NO_DIAGONAL_RHS))) return T(0);
    const pp_block_desc *__restrict__ blocks = cd.blocks;
    const pp_block_desc &b = blocks[bi];
    T S = T(0);
    const T viscP = withVisc ? visc_at<T>(cd.dom, b, f) : T(1);
    const T raP = (A ? T(1) / A[b.global_offset + f] : T(1)) * viscP;
#pragma unroll
    for (int face = 0; face < 2 * D; ++face) {
        const int axis = face >> 1;
        const int fs = face_sign(face);
        const bool atb = at_bound(p, face, b);
        if (atb && is_fixed(b, face)) {
            if (KIND == NK_PRESSURE || cd.neumann(bi, face)) continue;
            // tangential derivative of the prescribed values along the face
            const pp_boundary_desc &fb = b.bnd[face];
            const int ff = face_flat(p, face, b);
#pragma unroll
            for (int dd = 1; dd < D; ++dd) {
                const int tAxis = (axis + dd) % D;
                const T bA = fb.has_transform ? lap_coef_t<T, D>(Metric<T, D>::face(fb, ff), tAxis, axis)
                                              : (tAxis == axis ? T(1) : T(0));
                const bool tLo = p[tAxis] == 0, tHi = p[tAxis] == b.size[tAxis] - 1;
                int lo[3] = {p[0], p[1], p[2]}, hi[3] = {p[0], p[1], p[2]};
                if (!tLo) lo[tAxis] -= 1;
                if (!tHi) hi[tAxis] += 1;
                const T df = (tLo || tHi) ? T(1.0) : T(0.5);
                const T g = df * (cd.fixed(bi, face, hi) - cd.fixed(bi, face, lo));
                S -= T(fs) * bA * g * viscP;
            }
            continue;
        }
        // neighbour across the face (resolveNeighborCell, topology.cuh:352-398: border 0, axes
        // mapped through a connection)
        int np[3], nw;
        const int nbi = neighbor(blocks, bi, D, p, face, np, &nw);
        const pp_block_desc &nb = blocks[nbi];
        const int nf = flat3(np, nb.size);
        int map[3] = {0, 2, 4};
        if (atb && b.bnd[face].type == PP_BOUND_CONNECTED)
#pragma unroll
            for (int k = 0; k < D; ++k) map[k] = conn_dir(D, map[k], axis, b.bnd[face]);
        T raN = A ? T(1) / A[nb.global_offset + nf] : T(1);
        if (withVisc) raN = raN * visc_at<T>(cd.dom, nb, nf);
#pragma unroll
        for (int dd = 1; dd < D; ++dd) {
            const int tAxis = (axis + dd) % D;
            T aN;
            if (nb.has_transform) aN = lap_coef<T, D>(nb, nf, map[tAxis] >> 1, map[axis] >> 1);
            else aN = tAxis == axis ? T(1) : T(0);
            const T faceAlpha = (lap_coef<T, D>(b, f, axis, tAxis) * raP + aN * raN) * T(0.5);
            T tg = T(0);
#pragma unroll
            for (int tUp = 0; tUp < 2; ++tUp) {
                const int tFace = tAxis * 2 + tUp;
                const int tfs = face_sign(tFace);
                const CornerValue<T> cv = corner_value<T, D, KIND>(cd, bi, p, face, tFace, flags & NO_DIRECT_RHS);
                if (cv.cells == 0 && (cv.neumann || KIND == NK_PRESSURE)) {
                    const int tOther = tFace ^ 1;
                    if (flags & NO_DIRECT_RHS) {
                        tg += T(tfs) * neighbor_data<T, D, KIND>(cd, bi, p, face) * T(0.75);
                        tg += T(-tfs) * neighbor_data<T, D, KIND>(cd, bi, p, tOther) * T(0.25);
                    }
                    if (flags & NO_DIAGONAL_RHS)
                        tg += T(-tfs) * diagonal_data<T, D, KIND>(cd, bi, p, face, tOther) * T(0.25);
                } else {
                    tg += T(tfs) * cv.data;
                }
            }
            S -= T(fs) * faceAlpha * tg;
        }
    }
    return S;
}

} // namespace pp

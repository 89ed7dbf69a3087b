// Non-orthogonal Laplace terms of skewed grids: the tangential metric coefficients of a face,
// the walk around a cell's corner, and the corner terms the matrices take implicitly and the
// right-hand sides explicitly (nonortho_laplace.cuh:53-131, block_data.cuh:280-430, :477-616,
// nonortho_rhs.cuh:28-200). Only the flag bits of the reference's default mode and
// NON_ORTHO_DIRECT_RHS are implemented; NON_ORTHO_DIAGONAL_MATRIX is not implemented there
// either.
//
// Positions are three scalars updated through selects (a cell index along a run-time axis
// would otherwise keep them in scratch memory), and the velocity's corner terms carry all
// components through one walk.
#pragma once
#include "pp_common.hpp"

namespace pp {

enum { NO_DIRECT_MATRIX = 1, NO_DIRECT_RHS = 2, NO_DIAGONAL_MATRIX = 4, NO_DIAGONAL_RHS = 8, NO_CENTER_MATRIX = 16 };

// The data a corner term reads (GridDataType of the reference). The matrix kinds read only
// the boundary condition types of fixed faces (velocity, or scalar channel w), no values.
enum { NK_FIXED_ONLY = 0, NK_VELOCITY = 1, NK_PRESSURE = 2, NK_SCALAR = 3, NK_VELOCITY_TYPES = 4, NK_SCALAR_TYPES = 5 };

struct P3 {
    int x, y, z;
    __device__ __forceinline__ int get(int a) const { return a == 0 ? x : (a == 1 ? y : z); }
    __device__ __forceinline__ void set(int a, int v) {
        x = a == 0 ? v : x;
        y = a == 1 ? v : y;
        z = a == 2 ? v : z;
    }
    __device__ __forceinline__ void add(int a, int v) { set(a, get(a) + v); }
    __device__ __forceinline__ bool operator==(const P3 &o) const { return x == o.x && y == o.y && z == o.z; }
};

__device__ __forceinline__ int bsize(const pp_block_desc &b, int a) { return a == 0 ? b.size[0] : (a == 1 ? b.size[1] : b.size[2]); }
__device__ __forceinline__ bool at_bnd(const P3 &p, int dir, const pp_block_desc &b) {
    const int a = dir >> 1;
    return (dir & 1) ? p.get(a) == bsize(b, a) - 1 : p.get(a) == 0;
}
__device__ __forceinline__ int flat_p(const P3 &p, const pp_block_desc &b) { return p.x + b.size[0] * (p.y + b.size[1] * p.z); }
// face cell of position p on face `bound` (the face axis dropped)
__device__ __forceinline__ int face_flat_p(const P3 &p, int bound, const pp_block_desc &b) {
    const int a = bound >> 1;
    const int sx = a == 0 ? 1 : b.size[0], sy = a == 1 ? 1 : b.size[1];
    const int qx = a == 0 ? 0 : p.x, qy = a == 1 ? 0 : p.y, qz = a == 2 ? 0 : p.z;
    return qx + sx * (qy + sy * qz);
}
// computeConnectedPos (topology.cuh:246-266), `border` cells in from the connected face; a
// block one cell thick there clamps it to 0 (the reference would read outside the block)
template <int D>
__device__ __forceinline__ P3 conn_p(const pp_block_desc *__restrict__ blocks, const P3 &pos, int bdim,
                                     const pp_boundary_desc &cb, int border) {
    const pp_block_desc &ob = blocks[cb.conn_block];
    P3 out = pos;
    int ca = cb.conn_axes[0] >> 1;
    const int sz = bsize(ob, ca);
    if (border > sz - 1) border = sz - 1;
    out.set(ca, (cb.conn_axes[0] & 1) ? sz - 1 - border : border);
    if (D > 1) {
        int axis = (bdim + 1) % D;
        ca = cb.conn_axes[1] >> 1;
        out.set(ca, (cb.conn_axes[1] & 1) ? bsize(ob, ca) - 1 - pos.get(axis) : pos.get(axis));
        if (D > 2) {
            axis = (bdim + 2) % D;
            ca = cb.conn_axes[2] >> 1;
            out.set(ca, (cb.conn_axes[2] & 1) ? bsize(ob, ca) - 1 - pos.get(axis) : pos.get(axis));
        }
    }
    return out;
}

// Neighbour across face `bound` (interior, periodic wrap or connected block, border 0): the
// `neighbor` of the orthogonal kernels on P3 positions. Returns the neighbour's block.
template <int D>
__device__ __forceinline__ int neighbor_p(const pp_block_desc *__restrict__ blocks, int bi, const P3 &p, int bound,
                                          P3 &np) {
    const pp_block_desc &b = blocks[bi];
    const int dim = bound >> 1;
    np = p;
    const bool atb = at_bnd(p, bound, b);
    if (atb && b.bnd[bound].type == PP_BOUND_CONNECTED) {
        np = conn_p<D>(blocks, p, dim, b.bnd[bound], 0);
        return b.bnd[bound].conn_block;
    }
    if (atb && b.bnd[bound].type == PP_BOUND_PERIODIC) np.set(dim, (bound & 1) ? 0 : bsize(b, dim) - 1);
    else np.add(dim, face_sign(bound));
    return bi;
}

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
    const int k = pos_mod((dir >> 1) - bdim, d);
    return (k == 0 ? cb.conn_axes[0] : (k == 1 ? cb.conn_axes[1] : cb.conn_axes[2])) ^ (dir & 1);
}

// Reads of one data kind: NC values per cell (velocity: all components; otherwise component
// or channel w), values on fixed faces, and the kind's boundary condition type on a fixed face.
template <typename T, int D, int KIND>
struct CornerData {
    static constexpr int NC = KIND == NK_VELOCITY ? D : 1;
    const pp_domain &dom;
    const pp_block_desc *__restrict__ blocks;
    int w;   // scalar channel
    int env;
    __device__ __forceinline__ void cell(int bi, const P3 &p, T *out) const {
        const pp_block_desc &b = blocks[bi];
        const size_t g = (size_t)b.global_offset + flat_p(p, b);
        const size_t n = dom.num_cells;
        if (KIND == NK_VELOCITY) {
#pragma unroll
            for (int k = 0; k < NC; ++k)
                out[k] = reinterpret_cast<const T *>(dom.velocity_result)[((size_t)env * D + k) * n + g];
        } else if (KIND == NK_PRESSURE) {
            out[0] = reinterpret_cast<const T *>(dom.pressure_result)[(size_t)env * n + g];
        } else if (KIND == NK_SCALAR) {
            out[0] = reinterpret_cast<const T *>(dom.scalar_result)[((size_t)env * dom.scalar_channels + w) * n + g];
        } else {
            out[0] = T(0);
        }
    }
    // getFixedBoundaryData (block_data.cuh:136-199): the reference keeps no pressure values on
    // fixed faces, so the pressure kind reads 0 there.
    __device__ __forceinline__ void fixed(int bi, int bound, const P3 &p, T *out) const {
        const pp_block_desc &b = blocks[bi];
        const pp_boundary_desc &fb = b.bnd[bound];
        const int fc = face_cells(b, bound), ff = face_flat_p(p, bound, b);
        if (KIND == NK_VELOCITY) {
#pragma unroll
            for (int k = 0; k < NC; ++k) out[k] = fixed_vel<T>(fb, D, fc, ff, k, env);
        } else if (KIND == NK_SCALAR) {
            out[0] = fixed_scalar<T>(fb, dom.scalar_channels, fc, ff, w, env);
        } else {
            out[0] = T(0);
        }
    }
    // getFixedBoundaryType (block_data.cuh:110-134)
    __device__ __forceinline__ bool neumann(int bi, int bound) const {
        const pp_boundary_desc &fb = blocks[bi].bnd[bound];
        if (KIND == NK_VELOCITY || KIND == NK_VELOCITY_TYPES) return fb.vel_type != PP_BC_DIRICHLET;
        if (KIND == NK_SCALAR || KIND == NK_SCALAR_TYPES) return scalar_neumann(fb, w);
        return false;
    }
};

template <typename T, int NC>
struct CornerValue {
    T data[NC];
    int cells;    // cells the value averages; 0 when it comes from a fixed face
    bool neumann; // boundary condition type of that fixed face
};

// getCornerValue, block_data.cuh:501-616, without the center cell (includeDepth0 false) and to
// depth 2: the two walks start towards dir1 and dir2 and turn around the corner until they
// meet, reach a fixed face (whose values give the corner value), or end. Data of depth-1
// cells is summed only with `depth1`. A walk across a block connection lands on the second
// cell layer of the connected block (computeConnectedPos with border offset 1), as in the
// reference.
template <typename T, int D, int KIND, int NC = CornerData<T, D, KIND>::NC>
__device__ __forceinline__ CornerValue<T, NC> corner_value(const CornerData<T, D, KIND> &cd, int bi, const P3 &p, int dir1, int dir2,
                                           bool depth1) {
    const pp_block_desc *__restrict__ blocks = cd.blocks;
    CornerValue<T, NC> r;
#pragma unroll
    for (int k = 0; k < NC; ++k) r.data[k] = T(0);
    r.neumann = false;
    int num = 1;
    int cb[2] = {bi, bi};
    P3 cp[2] = {p, p};
    int da_[2] = {dir1, dir2}, db_[2] = {dir2, dir1};
#pragma unroll
    for (int depth = 1; depth <= 2; ++depth) {
#pragma unroll
        for (int d = 0; d < 2; ++d) {
            const int da = da_[d], db = db_[d];
            const int axis = da >> 1;
            const pp_block_desc &b = blocks[cb[d]];
            if (at_bnd(cp[d], da, b)) {
                const pp_boundary_desc &bd = b.bnd[da];
                if (bd.type == PP_BOUND_FIXED) {
                    r.neumann = cd.neumann(cb[d], da);
                    r.cells = 0;
                    cd.fixed(cb[d], da, cp[d], r.data);
                    if (!at_bnd(cp[d], db, b)) {
                        P3 q = cp[d];
                        q.add(db >> 1, face_sign(db));
                        T o[NC];
                        cd.fixed(cb[d], da, q, o);
#pragma unroll
                        for (int k = 0; k < NC; ++k) r.data[k] = (r.data[k] + o[k]) * T(0.5);
                    }
                    return r;
                } else if (bd.type == PP_BOUND_CONNECTED) {
                    cp[d] = conn_p<D>(blocks, cp[d], axis, bd, 1);
                    cb[d] = bd.conn_block;
                    da_[d] = conn_dir(D, db, axis, bd);
                    db_[d] = conn_dir(D, da, axis, bd) ^ 1;
                } else if (bd.type == PP_BOUND_PERIODIC) {
                    cp[d].set(axis, (da & 1) ? 0 : bsize(b, axis) - 1);
                    da_[d] = db;
                    db_[d] = da ^ 1;
                }
            } else {
                cp[d].add(axis, face_sign(da));
                da_[d] = db;
                db_[d] = da ^ 1;
            }
            if (cb[d] == cb[d ^ 1] && cp[d] == cp[d ^ 1]) {
                r.cells = num;
#pragma unroll
                for (int k = 0; k < NC; ++k) r.data[k] = r.data[k] / T(num);
                return r;
            }
            if (depth > 1 || depth1) {
                T v[NC];
                cd.cell(cb[d], cp[d], v);
#pragma unroll
                for (int k = 0; k < NC; ++k) r.data[k] += v[k];
            }
            ++num;
        }
    }
    r.cells = num;
#pragma unroll
    for (int k = 0; k < NC; ++k) r.data[k] = r.data[k] / T(num);
    return r;
}

// getBlockDataNeighbor (zeroBound false), block_data.cuh:280-320
template <typename T, int D, int KIND, int NC = CornerData<T, D, KIND>::NC>
__device__ __forceinline__ void neighbor_data(const CornerData<T, D, KIND> &cd, int bi, const P3 &p, int dir, T *out) {
    const pp_block_desc &b = cd.blocks[bi];
    const int axis = dir >> 1;
    P3 q = p;
    if (at_bnd(p, dir, b)) {
        const pp_boundary_desc &bd = b.bnd[dir];
        if (bd.type == PP_BOUND_FIXED) { cd.fixed(bi, dir, p, out); return; }
        if (bd.type == PP_BOUND_CONNECTED) { cd.cell(bd.conn_block, conn_p<D>(cd.blocks, p, axis, bd, 1), out); return; }
        if (bd.type == PP_BOUND_PERIODIC) {
            q.set(axis, (dir & 1) ? 0 : bsize(b, axis) - 1);
            cd.cell(bi, q, out);
            return;
        }
#pragma unroll
        for (int k = 0; k < NC; ++k) out[k] = T(0);
        return;
    }
    q.add(axis, face_sign(dir));
    cd.cell(bi, q, out);
}

// getBlockDataNeighborDiagonal (zeroBound false), block_data.cuh:372-420: two steps, the
// prescribed direction second.
template <typename T, int D, int KIND, int NC = CornerData<T, D, KIND>::NC>
__device__ __forceinline__ void diagonal_data(const CornerData<T, D, KIND> &cd, int bi, const P3 &p, int dir1, int dir2, T *out) {
    const bool first_fixed = is_fixed(cd.blocks[bi], dir1);
    const int dirs[2] = {first_fixed ? dir2 : dir1, first_fixed ? dir1 : dir2};
    int cbi = bi;
    P3 q = p;
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const int dir = dirs[i];
        const int axis = dir >> 1;
        const pp_block_desc &b = cd.blocks[cbi];
        if (at_bnd(q, dir, b)) {
            const pp_boundary_desc &bd = b.bnd[dir];
            if (bd.type == PP_BOUND_FIXED) { cd.fixed(cbi, dir, q, out); return; }
            if (bd.type == PP_BOUND_CONNECTED) {
                q = conn_p<D>(cd.blocks, q, axis, bd, 1);
                cbi = bd.conn_block;
            } else if (bd.type == PP_BOUND_PERIODIC) {
                q.set(axis, (dir & 1) ? 0 : bsize(b, axis) - 1);
            } else {
#pragma unroll
                for (int k = 0; k < NC; ++k) out[k] = T(0);
                return;
            }
        } else {
            q.add(axis, face_sign(dir));
        }
    }
    cd.cell(cbi, q, out);
}

// Matrix variant of the tangential coefficients of face `face` towards axis tAxis
// (interpolateNonOrthoLaplaceComponents, nonortho_laplace.cuh:53-131, useFaceTransform false):
// (alpha_P[axis, tAxis] raP + alpha_N[tAxis] raN) / 2 with ra = (1/A) * viscosity as selected.
// As in the reference it is 0 for every face of a cell at either end of the face's axis when
// that face is prescribed, and the neighbour's coefficients use unmapped axes.
template <typename T, int D>
__device__ __forceinline__ T face_alpha_matrix(const pp_domain &dom, const pp_block_desc *__restrict__ blocks, int bi, const int *p,
                               int f, int face, int tAxis, T raP, bool withVisc, const T *A) {
    const pp_block_desc &b = blocks[bi];
    const int axis = face >> 1;
    if (!((0 < p[axis] && p[axis] < b.size[axis] - 1) || !is_fixed(b, face))) return T(0);
    P3 np;
    const int nbi = neighbor_p<D>(blocks, bi, P3{p[0], p[1], p[2]}, face, np);
    const pp_block_desc &nb = blocks[nbi];
    const int nf = flat_p(np, nb);
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
__device__ __forceinline__ void matrix_corner_terms(const CornerData<T, D, KIND> &cd, int bi, const int *p, int f, int bound,
                                    int flags, T raP, bool withVisc, const T *A, T vscale, bool always_gradient,
                                    T sign, T &diag, T *rv) {
    const bool nbr = flags & NO_DIRECT_MATRIX, ctr = flags & NO_CENTER_MATRIX;
    if (D < 2 || !(nbr || ctr)) return;
    const int dim = bound >> 1;
    const int fs = face_sign(bound);
    const P3 pp{p[0], p[1], p[2]};
#pragma unroll
    for (int i = 1; i < D; ++i) {
        const int tAxis = (dim + i) % D;
        const T alpha = face_alpha_matrix<T, D>(cd.dom, cd.blocks, bi, p, f, bound, tAxis, raP, withVisc, A) * vscale;
        if (alpha == T(0)) continue;
#pragma unroll
        for (int tUp = 0; tUp < 2; ++tUp) {
            const int tFace = tAxis * 2 + tUp;
            const int tfs = face_sign(tFace);
            const auto cv = corner_value<T, D, KIND>(cd, bi, pp, bound, tFace, false);
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
// useFaceTransform false), for the Laplace term positive on the left-hand side, into S[NC].
// withVisc: ra includes the viscosity (velocity); A: ra includes 1/A (pressure).
template <typename T, int D, int KIND, int NC = CornerData<T, D, KIND>::NC>
__device__ __forceinline__ void nonortho_rhs(const CornerData<T, D, KIND> &cd, int bi, const int *pin, int f, int flags, bool withVisc,
                             const T *A, T *S) {
#pragma unroll
    for (int k = 0; k < NC; ++k) S[k] = T(0);
    if (D < 2 || !(flags & (NO_DIRECT_RHS | NO_DIAGONAL_RHS))) return;
    const pp_block_desc *__restrict__ blocks = cd.blocks;
    const pp_block_desc &b = blocks[bi];
    const P3 p{pin[0], pin[1], pin[2]};
    const T viscP = withVisc ? visc_at<T>(cd.dom, b, f) : T(1);
    const T raP = (A ? T(1) / A[b.global_offset + f] : T(1)) * viscP;
#pragma unroll
    for (int face = 0; face < 2 * D; ++face) {
        const int axis = face >> 1;
        const int fs = face_sign(face);
        const bool atb = at_bound(pin, face, b);
        if (atb && is_fixed(b, face)) {
            if (KIND == NK_PRESSURE || cd.neumann(bi, face)) continue;
            // tangential derivative of the prescribed values along the face
            const pp_boundary_desc &fb = b.bnd[face];
            const int ff = face_flat(pin, face, b);
#pragma unroll
            for (int dd = 1; dd < D; ++dd) {
                const int tAxis = (axis + dd) % D;
                const T bA = fb.has_transform ? lap_coef_t<T, D>(Metric<T, D>::face(fb, ff), tAxis, axis)
                                              : (tAxis == axis ? T(1) : T(0));
                const bool tLo = pin[tAxis] == 0, tHi = pin[tAxis] == b.size[tAxis] - 1;
                P3 lo = p, hi = p;
                if (!tLo) lo.add(tAxis, -1);
                if (!tHi) hi.add(tAxis, 1);
                const T df = (tLo || tHi) ? T(1.0) : T(0.5);
                T vh[NC], vl[NC];
                cd.fixed(bi, face, hi, vh);
                cd.fixed(bi, face, lo, vl);
#pragma unroll
                for (int k = 0; k < NC; ++k) S[k] -= T(fs) * bA * (df * (vh[k] - vl[k])) * viscP;
            }
            continue;
        }
        // neighbour across the face (resolveNeighborCell, topology.cuh:352-398: border 0, axes
        // mapped through a connection)
        P3 np;
        const int nbi = neighbor_p<D>(blocks, bi, p, face, np);
        const pp_block_desc &nb = blocks[nbi];
        const int nf = flat_p(np, nb);
        const bool conn = atb && b.bnd[face].type == PP_BOUND_CONNECTED;
        T raN = A ? T(1) / A[nb.global_offset + nf] : T(1);
        if (withVisc) raN = raN * visc_at<T>(cd.dom, nb, nf);
#pragma unroll
        for (int dd = 1; dd < D; ++dd) {
            const int tAxis = (axis + dd) % D;
            T aN;
            if (nb.has_transform) {
                const int mt = conn ? conn_dir(D, tAxis * 2, axis, b.bnd[face]) >> 1 : tAxis;
                const int ma = conn ? conn_dir(D, axis * 2, axis, b.bnd[face]) >> 1 : axis;
                aN = lap_coef<T, D>(nb, nf, mt, ma);
            } else {
                aN = tAxis == axis ? T(1) : T(0);
            }
            const T faceAlpha = (lap_coef<T, D>(b, f, axis, tAxis) * raP + aN * raN) * T(0.5);
            T tg[NC];
#pragma unroll
            for (int k = 0; k < NC; ++k) tg[k] = T(0);
#pragma unroll
            for (int tUp = 0; tUp < 2; ++tUp) {
                const int tFace = tAxis * 2 + tUp;
                const int tfs = face_sign(tFace);
                const auto cv = corner_value<T, D, KIND>(cd, bi, p, face, tFace, flags & NO_DIRECT_RHS);
                if (cv.cells == 0 && (cv.neumann || KIND == NK_PRESSURE)) {
                    const int tOther = tFace ^ 1;
                    if (flags & NO_DIRECT_RHS) {
                        T a[NC], c[NC];
                        neighbor_data<T, D, KIND>(cd, bi, p, face, a);
                        neighbor_data<T, D, KIND>(cd, bi, p, tOther, c);
#pragma unroll
                        for (int k = 0; k < NC; ++k) {
                            tg[k] += T(tfs) * a[k] * T(0.75);
                            tg[k] += T(-tfs) * c[k] * T(0.25);
                        }
                    }
                    if (flags & NO_DIAGONAL_RHS) {
                        T e[NC];
                        diagonal_data<T, D, KIND>(cd, bi, p, face, tOther, e);
#pragma unroll
                        for (int k = 0; k < NC; ++k) tg[k] += T(-tfs) * e[k] * T(0.25);
                    }
                } else {
#pragma unroll
                    for (int k = 0; k < NC; ++k) tg[k] += T(tfs) * cv.data[k];
                }
            }
#pragma unroll
            for (int k = 0; k < NC; ++k) S[k] -= T(fs) * faceAlpha * tg[k];
        }
    }
}

} // namespace pp

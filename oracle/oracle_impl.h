/*
 * ORACLE — test infrastructure only (see oracle.h). Typed body, included by oracle.c once
 * with REAL=double (suffix _f64) and once with REAL=float (suffix _f32).
 *
 * Restates, for one environment and CENTRAL advection, the reference kernels:
 *   grid metrics      csrc/phipict/grid/grid_gen.cu:298-354 (cells), :398-527 (faces)
 *   advection matrix  csrc/phipict/piso/advection.cu:26-306
 *   advection RHS     csrc/phipict/piso/advection.cu:799-955
 *   pressure matrix   csrc/phipict/piso/pressure.cu:25-193
 *   pressure RHS      csrc/phipict/piso/pressure.cu:349-522
 *   divergence        csrc/phipict/piso/pressure.cu:755-803
 *   correction        csrc/phipict/piso/velocity_correction.cu:26-60
 *   Epot matrix       csrc/phipict/mhd/potential.cu:29-233
 *   Epot RHS          csrc/phipict/mhd/potential.cu:272-322
 *   current density   csrc/phipict/mhd/potential.cu:708-885
 *   fused CG          csrc/phipict/solvers/krylov/krylov.cu:291-568, 1058-1125
 *   fused BiCGStab    csrc/phipict/solvers/krylov/krylov.cu:598-824, 1129-1186
 * Non-orthogonal corner terms (piso/device/nonortho_*.cuh) vanish identically on
 * rectilinear grids and are not restated.
 */

#define FN2(a, b) a##b
#define FN1(a, b) FN2(a, b)
#define FN(name) FN1(name, SFX)

typedef struct {
    const ob_domain *dom;
    int d;
    int ts; /* transform size 2d^2+1 */
} FN(octx);

static int FN(ncells)(const ob_block *b) { return b->size[0] * b->size[1] * b->size[2]; }
static int FN(flat)(const int *p, const int *s) { return p[0] + s[0] * (p[1] + s[1] * p[2]); }
static void FN(unflat)(int idx, const int *s, int *p) {
    p[0] = idx % s[0];
    p[1] = (idx / s[0]) % s[1];
    p[2] = idx / (s[0] * s[1]);
}
static int FN(at_bound)(const int *p, int bound, const ob_block *b) {
    int axis = bound >> 1;
    return (bound & 1) ? p[axis] == b->size[axis] - 1 : p[axis] == 0;
}
/* isEmptyBound, piso/device/topology.cuh:24-31: only FIXED among the prescribed types */
static int FN(is_empty)(const ob_block *b, int bound) { return b->b[bound].type == OB_FIXED; }
static int FN(face_sign)(int bound) { return (bound & 1) * 2 - 1; }
/* face-cell index of cell p on face `bound` (normal coordinate 0, face grid size 1) */
static int FN(face_flat)(const int *p, int bound, const ob_block *b) {
    int s[3] = {b->size[0], b->size[1], b->size[2]};
    int q[3] = {p[0], p[1], p[2]};
    s[bound >> 1] = 1;
    q[bound >> 1] = 0;
    return FN(flat)(q, s);
}
static int FN(face_cells)(int bound, const ob_block *b) {
    int s[3] = {b->size[0], b->size[1], b->size[2]};
    s[bound >> 1] = 1;
    return s[0] * s[1] * s[2];
}
static int FN(pos_mod)(int a, int m) { int r = a % m; return r < 0 ? r + m : r; }

/* computeConnectedPos, topology.cuh:246-266 */
static void FN(conn_pos)(const FN(octx) *c, const int *pos, int bdim, const ob_bound *cb, int border, int *out) {
    const ob_block *ob = &c->dom->blocks[cb->conn_block];
    out[0] = pos[0]; out[1] = pos[1]; out[2] = pos[2];
    int ca = cb->conn_axes[0] >> 1;
    if (border > ob->size[ca] - 1) border = ob->size[ca] - 1; /* one cell thick: stay inside */
    out[ca] = (cb->conn_axes[0] & 1) ? ob->size[ca] - 1 - border : border;
    if (c->d > 1) {
        int axis = (bdim + 1) % c->d;
        ca = cb->conn_axes[1] >> 1;
        out[ca] = (cb->conn_axes[1] & 1) ? ob->size[ca] - 1 - pos[axis] : pos[axis];
        if (c->d > 2) {
            axis = (bdim + 2) % c->d;
            ca = cb->conn_axes[2] >> 1;
            out[ca] = (cb->conn_axes[2] & 1) ? ob->size[ca] - 1 - pos[axis] : pos[axis];
        }
    }
}
/* channel of computeConnectedPosWithChannel, topology.cuh:270-281 */
static int FN(conn_channel)(const FN(octx) *c, int w, int bdim, const ob_bound *cb) {
    return cb->conn_axes[FN(pos_mod)(w - bdim, c->d)] >> 1;
}

static const REAL *FN(T)(const ob_block *b, const int *p, int ts) {
    return (const REAL *)b->transform + (size_t)FN(flat)(p, b->size) * ts;
}
/* getDeterminant, discretization.cuh:279-289 */
static REAL FN(det)(const FN(octx) *c, const ob_block *b, const int *p) {
    if (!b->has_transform) return (REAL)1;
    return FN(T)(b, p, c->ts)[2 * c->d * c->d];
}
/* getLaplaceCoefficientOrthogonal, discretization.cuh:359-374: det * |Minv[w]|^2 */
static REAL FN(alpha)(const FN(octx) *c, const ob_block *b, const int *p, int w) {
    if (!b->has_transform) return (REAL)1;
    const REAL *t = FN(T)(b, p, c->ts);
    const REAL *mi = t + c->d * c->d + w * c->d;
    REAL r = 0;
    for (int i = 0; i < c->d; ++i) r += mi[i] * mi[i];
    return t[2 * c->d * c->d] * r;
}
/* getMinvNorm, discretization.cuh:677-686 */
static REAL FN(minv_norm)(const FN(octx) *c, const ob_block *b, const int *p, int axis) {
    if (!b->has_transform) return (REAL)1;
    const REAL *t = FN(T)(b, p, c->ts);
    const REAL *mi = t + c->d * c->d + axis * c->d;
    REAL r = 0;
    for (int i = 0; i < c->d; ++i) r += mi[i] * mi[i];
    return sizeof(REAL) == 4 ? (REAL)sqrtf((float)r) : (REAL)sqrt((double)r);
}
/* getViscosityBlock, discretization.cuh:1152-1172 */
static REAL FN(visc)(const FN(octx) *c, const ob_block *b, const int *p) {
    if (b->visc_mode == 0) return (REAL)c->dom->viscosity;
    if (b->visc_mode == 1) return ((const REAL *)b->viscosity)[0];
    return ((const REAL *)b->viscosity)[FN(flat)(p, b->size)];
}
/* velocity component k of a cell, from the block field or a flat global vector */
static REAL FN(vel_at)(const FN(octx) *c, const ob_block *b, const int *p, int k, const REAL *glob) {
    int f = FN(flat)(p, b->size);
    if (glob) return glob[f + b->global_offset + (size_t)c->dom->ncells * k];
    return ((const REAL *)b->velocity)[(size_t)k * FN(ncells)(b) + f];
}
/* getContravariantComponent, velocity.cuh:222-234 with VelocityToContravariantComponent :24-39 */
static REAL FN(contra)(const FN(octx) *c, const ob_block *b, const int *p, int w, const REAL *glob) {
    if (!b->has_transform) return FN(vel_at)(c, b, p, w, glob);
    REAL v[3];
    for (int k = 0; k < c->d; ++k) v[k] = FN(vel_at)(c, b, p, k, glob);
    const REAL *t = FN(T)(b, p, c->ts);
    const REAL *mi = t + c->d * c->d + w * c->d;
    REAL r = 0;
    for (int k = 0; k < c->d; ++k) r += mi[k] * v[k];
    return t[2 * c->d * c->d] * r;
}
/* getFixedBoundaryData (VELOCITY), block_data.cuh:123-196 */
static REAL FN(fixed_vel)(const FN(octx) *c, const ob_block *b, int bound, const int *p, int k) {
    const ob_bound *fb = &b->b[bound];
    const REAL *data = (const REAL *)fb->vel;
    if (fb->vel_static) return data[k];
    return data[(size_t)k * FN(face_cells)(bound, b) + FN(face_flat)(p, bound, b)];
}
/* getContravariantComponentBoundaryFixed, velocity.cuh:245-265 with getVelocityFromBoundaryFixed :134-153 */
static REAL FN(contra_fixed)(const FN(octx) *c, const ob_block *b, int bound, const int *p, int w) {
    const ob_bound *fb = &b->b[bound];
    if (!fb->has_transform) return FN(fixed_vel)(c, b, bound, p, w);
    REAL v[3] = {0, 0, 0};
    if (!fb->vel_neumann)
        for (int k = 0; k < c->d; ++k) v[k] = FN(fixed_vel)(c, b, bound, p, k);
    const REAL *t = (const REAL *)fb->transform + (size_t)FN(face_flat)(p, bound, b) * c->ts;
    const REAL *mi = t + c->d * c->d + w * c->d;
    REAL r = 0;
    for (int k = 0; k < c->d; ++k) r += mi[k] * v[k];
    return t[2 * c->d * c->d] * r;
}
/* getLaplaceCoefficientOrthogonalBoundary, discretization.cuh:603-618 */
static REAL FN(alpha_fixed)(const FN(octx) *c, const ob_block *b, int bound, const int *p, int w) {
    const ob_bound *fb = &b->b[bound];
    if (!fb->has_transform) return (REAL)1;
    const REAL *t = (const REAL *)fb->transform + (size_t)FN(face_flat)(p, bound, b) * c->ts;
    const REAL *mi = t + c->d * c->d + w * c->d;
    REAL r = 0;
    for (int i = 0; i < c->d; ++i) r += mi[i] * mi[i];
    return t[2 * c->d * c->d] * r;
}
/* getBlockVelocitySource, nonortho_laplace.cuh:382-393 */
static REAL FN(source)(const ob_block *b, const int *p, int k) {
    if (b->vsource_mode == 0) return 0;
    if (b->vsource_mode == 1) return ((const REAL *)b->vsource)[k];
    return ((const REAL *)b->vsource)[(size_t)k * FN(ncells)(b) + FN(flat)(p, b->size)];
}
static int FN(pot_bc)(const ob_block *b, int bound, const int *p) {
    const ob_bound *fb = &b->b[bound];
    if (!fb->pot_types) return fb->pot_type;
    return fb->pot_types[FN(face_flat)(p, bound, b)];
}
static REAL FN(pot_value)(const ob_block *b, int bound, const int *p) {
    const ob_bound *fb = &b->b[bound];
    if (!fb->pot_values) return 0;
    return ((const REAL *)fb->pot_values)[FN(face_flat)(p, bound, b)];
}
static int FN(is_epot_dirichlet)(const ob_block *b, int bound, const int *p) {
    return b->b[bound].type == OB_FIXED && FN(pot_bc)(b, bound, p) == OB_POT_DIRICHLET;
}
static int FN(is_insulating_wall)(const ob_block *b, int bound, const int *p) {
    if (b->b[bound].type != OB_FIXED) return 0;
    int t = FN(pot_bc)(b, bound, p);
    return t == OB_POT_INSULATING || t == OB_POT_THIN_WALL;
}
static int FN(is_thin_wall)(const ob_block *b, int bound, const int *p) {
    return b->b[bound].type == OB_FIXED && FN(pot_bc)(b, bound, p) == OB_POT_THIN_WALL;
}

/* Neighbour of cell p across face `bound` (interior, periodic or connected): block, position
 * and the channel of the face-normal direction in the neighbour block. Used by the matrix
 * kernels (advection.cu:148-169, pressure.cu:53-67). */
static const ob_block *FN(neighbor)(const FN(octx) *c, const ob_block *b, const int *p, int bound, int *np, int *nw) {
    int dim = bound >> 1, up = bound & 1;
    np[0] = p[0]; np[1] = p[1]; np[2] = p[2];
    *nw = dim;
    if (FN(at_bound)(p, bound, b) && b->b[bound].type == OB_CONNECTED) {
        const ob_bound *cb = &b->b[bound];
        FN(conn_pos)(c, p, dim, cb, 0, np);
        *nw = FN(conn_channel)(c, dim, dim, cb);
        return &c->dom->blocks[cb->conn_block];
    }
    if (FN(at_bound)(p, bound, b) && b->b[bound].type == OB_PERIODIC) np[dim] = up ? 0 : b->size[dim] - 1;
    else np[dim] = p[dim] + FN(face_sign)(bound);
    return b;
}

/* computeFluxesNDLoop, discretization.cuh:758-852 (fluxes not multiplied by the face sign).
 * epot: computeEpotFluxesNDLoop :856-912 (prescribed faces: insulating 0, open velC). */
static void FN(fluxes)(const FN(octx) *c, const ob_block *b, const int *p, const REAL *glob, int epot, REAL *fl) {
    for (int bound = 0; bound < 2 * c->d; ++bound) {
        int dim = bound >> 1, up = bound & 1;
        REAL velC = FN(contra)(c, b, p, dim, glob);
        if (FN(at_bound)(p, bound, b)) {
            const ob_bound *bd = &b->b[bound];
            if (bd->type == OB_FIXED) {
                if (epot) fl[bound] = FN(is_insulating_wall)(b, bound, p) ? 0 : velC;
                else fl[bound] = FN(contra_fixed)(c, b, bound, p, dim);
            } else if (bd->type == OB_CONNECTED) {
                int np[3];
                const ob_block *nb = &c->dom->blocks[bd->conn_block];
                FN(conn_pos)(c, p, dim, bd, 0, np);
                int nw = bd->conn_axes[0] >> 1;
                REAL velN = FN(contra)(c, nb, np, nw, glob);
                if ((bd->conn_axes[0] & 1) == up) velN = -velN;
                fl[bound] = (velN + velC) * (REAL)0.5f;
            } else if (bd->type == OB_PERIODIC) {
                int np[3] = {p[0], p[1], p[2]};
                np[dim] = up ? 0 : b->size[dim] - 1;
                REAL velN = FN(contra)(c, b, np, dim, glob);
                fl[bound] = (velN + velC) * (REAL)0.5f;
            } else {
                fl[bound] = 0;
            }
        } else {
            int np[3] = {p[0], p[1], p[2]};
            np[dim] += FN(face_sign)(bound);
            REAL velN = FN(contra)(c, b, np, dim, glob);
            fl[bound] = (velN + velC) * (REAL)0.5f;
        }
    }
}

/* CSR row emission, advection.cu:286-300 / csr_rows.cuh:23-33: entries in ascending
 * column order, first occurrence first among equal columns. Returns entries written. */
static int FN(emit_row)(int *idx, REAL *vals, int n, int32_t *col, REAL *val, int divide, REAL det) {
    int written = 0;
    for (;;) {
        int best = -1, bv = 0x7fffffff;
        for (int i = 0; i < n; ++i)
            if (idx[i] >= 0 && idx[i] < bv) { bv = idx[i]; best = i; }
        if (best < 0) break;
        if (divide) vals[best] /= det;
        col[written] = idx[best];
        val[written] = vals[best];
        idx[best] = -1;
        ++written;
    }
    return written;
}

/* hasFarUpwindCell, shared_stencils.cuh:35-43: the face's neighbour and the far-upwind cell
 * (one further against the flow along the face normal) both lie inside the block. */
static int FN(far_upwind)(const int *p, int bound, int outflow, const ob_block *b) {
    const int axis = bound >> 1, fs = FN(face_sign)(bound), last = b->size[axis] - 1;
    const int iD = p[axis] + fs;
    if (iD < 0 || iD > last) return 0;
    const int iUU = outflow ? p[axis] - fs : p[axis] + 2 * fs;
    return iUU >= 0 && iUU <= last;
}

/* The LINEAR_UPWIND deferred correction of one field component at cell p (advection.cu:893-937
 * for the velocity, :650-684 for the passive scalar): for every face whose far-upwind cell
 * exists, t -= F 0.5 psi (phi_D - phi_U) with psi = r = (phi_U - phi_UU) / (phi_D - phi_U),
 * skipped on flat faces (|phi_D - phi_U| < 1e-20). fl: the cell's face fluxes (not multiplied
 * by the face sign); phi: the component's values over the block. */
static void FN(deferred_correction)(const ob_block *b, int d, const int *p, const REAL *fl, const REAL *phi, REAL *t) {
    for (int bound = 0; bound < 2 * d; ++bound) {
        const int axis = bound >> 1, fs = FN(face_sign)(bound);
        const REAL ffo = (REAL)fs * fl[bound];
        const int outflow = ffo >= 0;
        if (!FN(far_upwind)(p, bound, outflow, b)) continue;
        int q[3] = {p[0], p[1], p[2]};
        const REAL phiP = phi[FN(flat)(q, b->size)];
        q[axis] = p[axis] + fs;
        const REAL phiN = phi[FN(flat)(q, b->size)];
        q[axis] = outflow ? p[axis] - fs : p[axis] + 2 * fs;
        const REAL phiUU = phi[FN(flat)(q, b->size)];
        const REAL phiU = outflow ? phiP : phiN;
        const REAL phiD = outflow ? phiN : phiP;
        const REAL dD = phiD - phiU;
        const REAL absDD = dD < 0 ? -dD : dD;
        if (absDD < (REAL)1e-20) continue;
        const REAL psi = (phiU - phiUU) / dD;
        *t -= ffo * (REAL)0.5 * psi * dD;
    }
}

static int FN(domain_ok)(const ob_domain *dom) { return dom && dom->dims >= 2 && dom->dims <= 3 && dom->nblocks > 0; }

/* ---- non-orthogonal corner terms of skewed grids ----------------------------------------
 * nonortho_laplace.cuh:53-131 (face coefficients of the matrices), block_data.cuh:280-420
 * (neighbour and diagonal data), :477-616 (corner walk), nonortho_rhs.cuh:28-200 (explicit
 * corner terms), with useFaceTransform false. Flag bits of non_ortho_flags.h. */

typedef struct {
    const FN(octx) *c;
    int kind;
    int w;
    const REAL *data;            /* VEL, SCAL: [..][ncells] (component/channel w); PRES: [ncells] */
    const ob_face_scalar *fscal; /* SCAL: [nblocks*6] face data of channel w */
    int scalar_types;            /* TYPES: the face conditions are the scalar channel's (fscal) */
} FN(ndata);

/* getLaplaceCoefficientsFull / SingleAxis, discretization.cuh:419-470: det Minv[c1].Minv[c2] */
static REAL FN(lapc_t)(const FN(octx) *c, const REAL *t, int c1, int c2) {
    REAL r = 0;
    for (int i = 0; i < c->d; ++i) r += t[c->d * c->d + c1 * c->d + i] * t[c->d * c->d + c2 * c->d + i];
    return t[2 * c->d * c->d] * r;
}
static REAL FN(lapc)(const FN(octx) *c, const ob_block *b, const int *p, int c1, int c2) {
    if (!b->has_transform) return c1 == c2 ? (REAL)1 : (REAL)0;
    return FN(lapc_t)(c, FN(T)(b, p, c->ts), c1, c2);
}
/* computeConnectedDir, topology.cuh:236-244 */
static int FN(conn_dir)(const FN(octx) *c, int dir, int bdim, const ob_bound *cb) {
    return cb->conn_axes[FN(pos_mod)((dir >> 1) - bdim, c->d)] ^ (dir & 1);
}
static int FN(block_index)(const FN(octx) *c, const ob_block *b) { return (int)(b - c->dom->blocks); }

static REAL FN(ncell)(const FN(ndata) *nd, const ob_block *b, const int *p) {
    const size_t g = (size_t)b->global_offset + FN(flat)(p, b->size);
    const size_t n = (size_t)nd->c->dom->ncells;
    if (nd->kind == OK_VEL || nd->kind == OK_SCAL) return nd->data[(size_t)nd->w * n + g];
    if (nd->kind == OK_PRES) return nd->data[g];
    return 0;
}
/* getFixedBoundaryData (block_data.cuh:136-199); no pressure values live on fixed faces */
static REAL FN(nfixed)(const FN(ndata) *nd, const ob_block *b, int bound, const int *p) {
    if (nd->kind == OK_VEL) return FN(fixed_vel)(nd->c, b, bound, p, nd->w);
    if (nd->kind == OK_SCAL) {
        const ob_face_scalar *fs = &nd->fscal[FN(block_index)(nd->c, b) * 6 + bound];
        if (!fs->values) return 0;
        return ((const REAL *)fs->values)[fs->is_static ? 0 : FN(face_flat)(p, bound, b)];
    }
    return 0;
}
/* getFixedBoundaryType (block_data.cuh:110-134) */
static int FN(nneumann)(const FN(ndata) *nd, const ob_block *b, int bound) {
    if (nd->kind == OK_SCAL || (nd->kind == OK_TYPES && nd->scalar_types))
        return nd->fscal[FN(block_index)(nd->c, b) * 6 + bound].neumann;
    if (nd->kind == OK_VEL || nd->kind == OK_TYPES) return b->b[bound].vel_neumann;
    return 0;
}

/* getCornerValue, block_data.cuh:501-616 (includeDepth0 false, maxDepth 2). Returns the value;
 * *cells = cells averaged (0: from a fixed face), *neu = that face's condition is Neumann. */
static REAL FN(corner)(const FN(ndata) *nd, const ob_block *b, const int *p, int dir1, int dir2, int depth1,
                       int *cells, int *neu) {
    const FN(octx) *c = nd->c;
    REAL data = 0;
    int num = 1;
    *neu = 0;
    const ob_block *cb[2] = {b, b};
    int pos[2][3] = {{p[0], p[1], p[2]}, {p[0], p[1], p[2]}};
    int dir[2][2] = {{dir1, dir2}, {dir2, dir1}};
    for (int depth = 1; depth <= 2; ++depth) {
        for (int d = 0; d < 2; ++d) {
            int da = dir[d][0], db = dir[d][1], axis = da >> 1;
            if (FN(at_bound)(pos[d], da, cb[d])) {
                const ob_bound *bd = &cb[d]->b[da];
                if (bd->type == OB_FIXED) {
                    *neu = FN(nneumann)(nd, cb[d], da);
                    *cells = 0;
                    if (FN(at_bound)(pos[d], db, cb[d])) return FN(nfixed)(nd, cb[d], da, pos[d]);
                    int q[3] = {pos[d][0], pos[d][1], pos[d][2]};
                    q[db >> 1] += FN(face_sign)(db);
                    return (FN(nfixed)(nd, cb[d], da, pos[d]) + FN(nfixed)(nd, cb[d], da, q)) * (REAL)0.5;
                } else if (bd->type == OB_CONNECTED) {
                    int q[3];
                    FN(conn_pos)(c, pos[d], axis, bd, 1, q);
                    cb[d] = &c->dom->blocks[bd->conn_block];
                    pos[d][0] = q[0]; pos[d][1] = q[1]; pos[d][2] = q[2];
                    dir[d][0] = FN(conn_dir)(c, db, axis, bd);
                    dir[d][1] = FN(conn_dir)(c, da, axis, bd) ^ 1;
                } else if (bd->type == OB_PERIODIC) {
                    pos[d][axis] = (da & 1) ? 0 : cb[d]->size[axis] - 1;
                    dir[d][0] = db;
                    dir[d][1] = da ^ 1;
                }
            } else {
                pos[d][axis] += FN(face_sign)(da);
                dir[d][0] = db;
                dir[d][1] = da ^ 1;
            }
            int o = d ^ 1;
            if (cb[d] == cb[o] && pos[d][0] == pos[o][0] && pos[d][1] == pos[o][1] && pos[d][2] == pos[o][2]) {
                *cells = num;
                return data / (REAL)num;
            }
            if (depth > 1 || depth1) data += FN(ncell)(nd, cb[d], pos[d]);
            ++num;
        }
    }
    *cells = num;
    return data / (REAL)num;
}

/* getBlockDataNeighbor (zeroBound false), block_data.cuh:276-320 */
static REAL FN(nb_data)(const FN(ndata) *nd, const ob_block *b, const int *p, int dir) {
    int axis = dir >> 1;
    int q[3] = {p[0], p[1], p[2]};
    if (FN(at_bound)(p, dir, b)) {
        const ob_bound *bd = &b->b[dir];
        if (bd->type == OB_FIXED) return FN(nfixed)(nd, b, dir, p);
        if (bd->type == OB_CONNECTED) {
            FN(conn_pos)(nd->c, p, axis, bd, 1, q);
            return FN(ncell)(nd, &nd->c->dom->blocks[bd->conn_block], q);
        }
        if (bd->type == OB_PERIODIC) {
            q[axis] = (dir & 1) ? 0 : b->size[axis] - 1;
            return FN(ncell)(nd, b, q);
        }
        return 0;
    }
    q[axis] += FN(face_sign)(dir);
    return FN(ncell)(nd, b, q);
}

/* getBlockDataNeighborDiagonal (zeroBound false), block_data.cuh:372-420 */
static REAL FN(diag_data)(const FN(ndata) *nd, const ob_block *b, const int *p, int dir1, int dir2) {
    int first_fixed = FN(is_empty)(b, dir1);
    int dirs[2] = {first_fixed ? dir2 : dir1, first_fixed ? dir1 : dir2};
    const ob_block *cb = b;
    int q[3] = {p[0], p[1], p[2]};
    for (int i = 0; i < 2; ++i) {
        int dir = dirs[i], axis = dir >> 1;
        if ((dir & 1) ? q[axis] == cb->size[axis] - 1 : q[axis] == 0) {
            const ob_bound *bd = &cb->b[dir];
            if (bd->type == OB_FIXED) return FN(nfixed)(nd, cb, dir, q);
            if (bd->type == OB_CONNECTED) {
                int r[3];
                FN(conn_pos)(nd->c, q, axis, bd, 1, r);
                cb = &nd->c->dom->blocks[bd->conn_block];
                q[0] = r[0]; q[1] = r[1]; q[2] = r[2];
            } else if (bd->type == OB_PERIODIC) {
                q[axis] = (dir & 1) ? 0 : cb->size[axis] - 1;
            } else {
                return 0;
            }
        } else {
            q[axis] += FN(face_sign)(dir);
        }
    }
    return FN(ncell)(nd, cb, q);
}

/* interpolateNonOrthoLaplaceComponents, nonortho_laplace.cuh:53-131: the coefficient of face
 * `face` towards tAxis, (alpha_P[axis,tAxis] raP + alpha_N[tAxis] raN) / 2; 0 for every face
 * of a cell at either end of the face's axis when that face is prescribed (as the reference's
 * test reads) */
static REAL FN(face_alpha_m)(const FN(octx) *c, const ob_block *b, const int *p, int face, int tAxis, REAL raP,
                             int with_visc, const REAL *A) {
    int axis = face >> 1;
    if (!((0 < p[axis] && p[axis] < b->size[axis] - 1) || !FN(is_empty)(b, face))) return 0;
    int np[3], nw;
    const ob_block *nb = FN(neighbor)(c, b, p, face, np, &nw);
    REAL raN = A ? (REAL)1 / A[nb->global_offset + FN(flat)(np, nb->size)] : (REAL)1;
    if (with_visc) raN = raN * FN(visc)(c, nb, np);
    REAL aP = FN(lapc)(c, b, p, axis, tAxis);
    REAL aN = FN(lapc)(c, nb, np, tAxis, axis);
    return (aP * raP + aN * raN) * (REAL)0.5;
}

/* corner terms of face `bound` of a matrix row: PISO_build_matrix, advection.cu:177-240 (sign
 * +1, diag = the diagonal accumulator); PISO_build_pressure_matrix, pressure.cu:104-160 (sign
 * -1, diag = rv[0]) */
static void FN(matrix_corner)(const FN(ndata) *nd, const ob_block *b, const int *p, int bound, int flags, REAL raP,
                              int with_visc, const REAL *A, REAL vscale, int always_gradient, REAL sign, REAL *diag,
                              REAL *rv) {
    const FN(octx) *c = nd->c;
    int nbr = flags & ONO_DIRECT_MATRIX, ctr = flags & ONO_CENTER_MATRIX;
    if (c->d < 2 || !(nbr || ctr)) return;
    int dim = bound >> 1, fs = FN(face_sign)(bound);
    for (int i = 1; i < c->d; ++i) {
        int tAxis = (dim + i) % c->d;
        REAL alpha = FN(face_alpha_m)(c, b, p, bound, tAxis, raP, with_visc, A) * vscale;
        if (alpha == 0) continue;
        for (int tUp = 0; tUp < 2; ++tUp) {
            int tFace = tAxis * 2 + tUp, tfs = FN(face_sign)(tFace);
            int cells, neu;
            FN(corner)(nd, b, p, bound, tFace, 0, &cells, &neu);
            if (cells < 1) {
                if (always_gradient || neu) {
                    REAL k = (REAL)(fs * tfs) * alpha * (REAL)0.25;
                    if (ctr) *diag -= sign * (3 * k);
                    if (nbr) {
                        rv[bound + 1] -= sign * (3 * k);
                        rv[(tFace ^ 1) + 1] += sign * k;
                    }
                }
            } else {
                REAL k = (REAL)(fs * tfs) * alpha * ((REAL)1.0 / (REAL)cells);
                if (ctr) *diag -= sign * k;
                if (nbr) {
                    rv[bound + 1] -= sign * k;
                    rv[tFace + 1] -= sign * k;
                }
            }
        }
    }
}

/* getNonOrthoLaplaceRHS_v2, nonortho_rhs.cuh:28-200 (useFaceTransform false) */
static REAL FN(nonortho_rhs)(const FN(ndata) *nd, const ob_block *b, const int *p, int flags, int with_visc,
                             const REAL *A) {
    const FN(octx) *c = nd->c;
    if (c->d < 2 || !(flags & (ONO_DIRECT_RHS | ONO_DIAGONAL_RHS))) return 0;
    REAL S = 0;
    int f = FN(flat)(p, b->size);
    REAL viscP = with_visc ? FN(visc)(c, b, p) : (REAL)1;
    REAL raP = (A ? (REAL)1 / A[b->global_offset + f] : (REAL)1) * viscP;
    for (int face = 0; face < 2 * c->d; ++face) {
        int axis = face >> 1, fs = FN(face_sign)(face);
        int atb = FN(at_bound)(p, face, b);
        if (atb && FN(is_empty)(b, face)) {
            if (nd->kind == OK_PRES || FN(nneumann)(nd, b, face)) continue;
            const ob_bound *fb = &b->b[face];
            for (int dd = 1; dd < c->d; ++dd) {
                int tAxis = (axis + dd) % c->d;
                REAL bA = fb->has_transform
                              ? FN(lapc_t)(c, (const REAL *)fb->transform + (size_t)FN(face_flat)(p, face, b) * c->ts,
                                           tAxis, axis)
                              : (tAxis == axis ? (REAL)1 : (REAL)0);
                int tLo = p[tAxis] == 0, tHi = p[tAxis] == b->size[tAxis] - 1;
                int lo[3] = {p[0], p[1], p[2]}, hi[3] = {p[0], p[1], p[2]};
                if (!tLo) lo[tAxis] -= 1;
                if (!tHi) hi[tAxis] += 1;
                REAL df = (tLo || tHi) ? (REAL)1.0 : (REAL)0.5;
                REAL g = df * (FN(nfixed)(nd, b, face, hi) - FN(nfixed)(nd, b, face, lo));
                S -= (REAL)fs * bA * g * viscP;
            }
            continue;
        }
        /* resolveNeighborCell, topology.cuh:352-398 */
        int np[3], nw;
        const ob_block *nb = FN(neighbor)(c, b, p, face, np, &nw);
        int map[3] = {0, 2, 4};
        if (atb && b->b[face].type == OB_CONNECTED)
            for (int k = 0; k < c->d; ++k) map[k] = FN(conn_dir)(c, map[k], axis, &b->b[face]);
        REAL raN = A ? (REAL)1 / A[nb->global_offset + FN(flat)(np, nb->size)] : (REAL)1;
        if (with_visc) raN = raN * FN(visc)(c, nb, np);
        for (int dd = 1; dd < c->d; ++dd) {
            int tAxis = (axis + dd) % c->d;
            REAL aN = nb->has_transform ? FN(lapc)(c, nb, np, map[tAxis] >> 1, map[axis] >> 1)
                                        : (tAxis == axis ? (REAL)1 : (REAL)0);
            REAL faceAlpha = (FN(lapc)(c, b, p, axis, tAxis) * raP + aN * raN) * (REAL)0.5;
            REAL tg = 0;
            for (int tUp = 0; tUp < 2; ++tUp) {
                int tFace = tAxis * 2 + tUp, tfs = FN(face_sign)(tFace);
                int cells, neu;
                REAL v = FN(corner)(nd, b, p, face, tFace, (flags & ONO_DIRECT_RHS) != 0, &cells, &neu);
                if (cells == 0 && (neu || nd->kind == OK_PRES)) {
                    int tOther = tFace ^ 1;
                    if (flags & ONO_DIRECT_RHS) {
                        tg += (REAL)tfs * FN(nb_data)(nd, b, p, face) * (REAL)0.75;
                        tg += (REAL)(-tfs) * FN(nb_data)(nd, b, p, tOther) * (REAL)0.25;
                    }
                    if (flags & ONO_DIAGONAL_RHS) tg += (REAL)(-tfs) * FN(diag_data)(nd, b, p, face, tOther) * (REAL)0.25;
                } else {
                    tg += (REAL)tfs * v;
                }
            }
            S -= (REAL)fs * faceAlpha * tg;
        }
    }
    return S;
}

/* S of every cell (the explicit corner terms; k_pressureRHSaddNonOrthoComponents adds them to
 * the pressure divergence, kPISO_build_advection_RHS and the scalar right-hand side subtract
 * them before dividing by det). kind OK_VEL (component w of vel_result [d][n], viscosity in
 * the coefficients), OK_PRES (pressure result, 1/A in the coefficients), OK_SCAL (channel w of
 * [C][n]; the caller multiplies by the diffusivity). */
int FN(o_nonortho_rhs)(const ob_domain *dom, int kind, int flags, int w, const REAL *data, const REAL *adiag,
                       const ob_face_scalar *fscal, REAL *out) {
    if (!FN(domain_ok)(dom)) return -1;
    if (kind == OK_PRES && !adiag) return -1;
    if (kind == OK_SCAL && !fscal) return -1;
    FN(octx) c = {dom, dom->dims, 2 * dom->dims * dom->dims + 1};
    FN(ndata) nd = {&c, kind, w, data, fscal, 0};
    for (int bi = 0; bi < dom->nblocks; ++bi) {
        const ob_block *b = &dom->blocks[bi];
        int n = FN(ncells)(b);
        O_PAR
        for (int f = 0; f < n; ++f) {
            int p[3];
            FN(unflat)(f, b->size, p);
            out[b->global_offset + f] = FN(nonortho_rhs)(&nd, b, p, flags, kind == OK_VEL, kind == OK_PRES ? adiag : NULL);
        }
    }
    return 0;
}

/* CSR pattern row offsets: rows in global cell order, row sizes 2d+1 minus the prescribed
 * faces of the cell (getCSRMatrixRowEndOffsetFromBlockBoundaries3D, topology.cuh:168-233). */
int FN(o_csr_rows)(const ob_domain *dom, int32_t *row) {
    if (!FN(domain_ok)(dom)) return -1;
    int d = dom->dims;
    row[0] = 0;
    for (int bi = 0; bi < dom->nblocks; ++bi) {
        const ob_block *b = &dom->blocks[bi];
        int n = FN(ncells)(b);
        for (int f = 0; f < n; ++f) {
            int p[3];
            FN(unflat)(f, b->size, p);
            int rs = 2 * d + 1;
            for (int bound = 0; bound < 2 * d; ++bound)
                if (FN(at_bound)(p, bound, b) && FN(is_empty)(b, bound)) --rs;
            int g = b->global_offset + f;
            row[g + 1] = row[g] + rs;
        }
    }
    return 0;
}

/* PISO_build_matrix, advection.cu:26-306 (CENTRAL scheme, orthogonal part) */
/* flags: non-orthogonal flag bits; scalar_mode: the passive-scalar variant, whose corner
 * coefficients carry no viscosity and are multiplied by the (domain) diffusivity */
int FN(o_advection_matrix)(const ob_domain *dom, int flags, int scalar_mode, const int32_t *row, int32_t *col,
                           REAL *val, REAL *adiag) {
    if (!FN(domain_ok)(dom)) return -1;
    FN(octx) c = {dom, dom->dims, 2 * dom->dims * dom->dims + 1};
    FN(ndata) nd = {&c, OK_TYPES, 0, NULL, NULL, 0};
    const REAL dt = (REAL)dom->dt;
    int bad = 0;
    for (int bi = 0; bi < dom->nblocks; ++bi) {
        const ob_block *b = &dom->blocks[bi];
        int n = FN(ncells)(b);
        O_PAR_OR(bad)
        for (int f = 0; f < n; ++f) {
            int p[3];
            FN(unflat)(f, b->size, p);
            REAL fl[6];
            FN(fluxes)(&c, b, p, NULL, 0, fl);
            int idx[7];
            REAL rv[7] = {0, 0, 0, 0, 0, 0, 0};
            const REAL det = b->has_transform ? FN(det)(&c, b, p) : (REAL)1;
            REAL diag = det / dt;
            idx[0] = f + b->global_offset;
            const REAL vp = FN(visc)(&c, b, p);
            REAL ap[3] = {1, 1, 1};
            for (int dd = 0; dd < c.d; ++dd) ap[dd] = FN(alpha)(&c, b, p, dd);
            for (int bound = 0; bound < 2 * c.d; ++bound) {
                int dim = bound >> 1;
                REAL fs = (REAL)FN(face_sign)(bound);
                int atb = FN(at_bound)(p, bound, b);
                if (!atb || !FN(is_empty)(b, bound)) {
                    REAL ffo = fs * fl[bound];
                    if (dom->scheme == OB_ADV_LINEAR_UPWIND && FN(far_upwind)(p, bound, ffo >= 0, b)) {
                        /* first-order upwind implicitly (advection.cu:129-140); the rest of the
                         * face value is the deferred correction of the right-hand side */
                        if (ffo >= 0) diag += ffo;
                        else rv[bound + 1] += ffo;
                    } else {
                        REAL ff = (REAL)0.5 * ffo;
                        diag += ff;
                        rv[bound + 1] += ff;
                    }
                    int np[3], nw;
                    const ob_block *nb = FN(neighbor)(&c, b, p, bound, np, &nw);
                    REAL an = FN(alpha)(&c, nb, np, nw);
                    REAL vn = FN(visc)(&c, nb, np);
                    idx[bound + 1] = FN(flat)(np, nb->size) + nb->global_offset;
                    REAL vc = (ap[dim] * vp + an * vn) * (REAL)0.5;
                    diag += vc;
                    rv[bound + 1] -= vc;
                    if (flags)
                        FN(matrix_corner)(&nd, b, p, bound, flags, scalar_mode ? (REAL)1 : vp, !scalar_mode, NULL,
                                          scalar_mode ? vp : (REAL)1, 0, (REAL)1, &diag, rv);
                } else {
                    int slip = b->b[bound].vel_neumann ? 1 : 0;
                    diag += (REAL)(1 - slip) * 2 * vp * ap[dim];
                    idx[bound + 1] = -1;
                }
            }
            for (int dd = c.d; dd < 3; ++dd) { idx[dd * 2 + 1] = -1; idx[dd * 2 + 2] = -1; }
            rv[0] = diag;
            int g = b->global_offset + f;
            int w = FN(emit_row)(idx, rv, 7, col + row[g], val + row[g], 1, det);
            if (w != row[g + 1] - row[g]) bad = 1;
            adiag[g] = rv[0];
        }
    }
    return bad ? -2 : 0;
}

/* kPISO_build_advection_RHS, advection.cu:799-955 (CENTRAL, orthogonal part) */
int FN(o_advection_rhs)(const ob_domain *dom, int apply_pressure_gradient, const REAL *S, REAL *rhs);
/* getPressureGradient + getPressureAtWithBounds, velocity.cuh:347-382, nonortho_laplace.cuh:302-378 */
static REAL FN(pressure_at)(const FN(octx) *c, const ob_block *b, const int *pin) {
    int t[3] = {pin[0], pin[1], pin[2]};
    for (int dim = 0; dim < c->d; ++dim) {
        int bound = dim * 2;
        if (pin[dim] < 0) {
            const ob_bound *bd = &b->b[bound];
            if (bd->type == OB_FIXED) { t[dim] = 0; return ((const REAL *)b->pressure)[FN(flat)(t, b->size)]; }
            if (bd->type == OB_CONNECTED) {
                int op[3];
                const ob_block *nb = &c->dom->blocks[bd->conn_block];
                FN(conn_pos)(c, t, dim, bd, -(pin[dim] + 1), op);
                return ((const REAL *)nb->pressure)[FN(flat)(op, nb->size)];
            }
            if (bd->type == OB_PERIODIC) { t[dim] += b->size[dim]; return ((const REAL *)b->pressure)[FN(flat)(t, b->size)]; }
            return 0;
        }
        bound = dim * 2 + 1;
        if (pin[dim] >= b->size[dim]) {
            const ob_bound *bd = &b->b[bound];
            if (bd->type == OB_FIXED) { t[dim] = b->size[dim] - 1; return ((const REAL *)b->pressure)[FN(flat)(t, b->size)]; }
            if (bd->type == OB_CONNECTED) {
                int op[3];
                const ob_block *nb = &c->dom->blocks[bd->conn_block];
                FN(conn_pos)(c, t, dim, bd, pin[dim] - b->size[dim], op);
                return ((const REAL *)nb->pressure)[FN(flat)(op, nb->size)];
            }
            if (bd->type == OB_PERIODIC) { t[dim] -= b->size[dim]; return ((const REAL *)b->pressure)[FN(flat)(t, b->size)]; }
            return 0;
        }
    }
    return ((const REAL *)b->pressure)[FN(flat)(pin, b->size)];
}
static void FN(pressure_grad)(const FN(octx) *c, const ob_block *b, const int *p, REAL *g) {
    REAL gr[3];
    for (int dim = 0; dim < c->d; ++dim) {
        int t[3] = {p[0], p[1], p[2]};
        REAL fac = (REAL)0.5;
        if (!(p[dim] == 0 && FN(is_empty)(b, dim * 2))) t[dim] = p[dim] - 1;
        else fac = 1;
        REAL vn = FN(pressure_at)(c, b, t);
        if (!(p[dim] == b->size[dim] - 1 && FN(is_empty)(b, dim * 2 + 1))) t[dim] = p[dim] + 1;
        else { t[dim] = p[dim]; fac = 1; }
        REAL vpp = FN(pressure_at)(c, b, t);
        gr[dim] = (vpp - vn) * fac;
    }
    if (b->has_transform) {
        const REAL *mi = FN(T)(b, p, c->ts) + c->d * c->d;
        REAL r[3] = {0, 0, 0};
        for (int i = 0; i < c->d; ++i)
            for (int j = 0; j < c->d; ++j) r[j] += mi[i * c->d + j] * gr[i];
        for (int j = 0; j < c->d; ++j) gr[j] = r[j];
    }
    for (int j = 0; j < c->d; ++j) g[j] = gr[j];
}

/* S: the explicit non-orthogonal corner terms [d][n] (o_nonortho_rhs with OK_VEL) or NULL */
int FN(o_advection_rhs)(const ob_domain *dom, int apply_pressure_gradient, const REAL *S, REAL *rhs) {
    if (!FN(domain_ok)(dom)) return -1;
    FN(octx) c = {dom, dom->dims, 2 * dom->dims * dom->dims + 1};
    const REAL dt = (REAL)dom->dt;
    for (int bi = 0; bi < dom->nblocks; ++bi) {
        const ob_block *b = &dom->blocks[bi];
        int n = FN(ncells)(b);
        O_PAR
        for (int f = 0; f < n; ++f) {
            int p[3];
            FN(unflat)(f, b->size, p);
            const REAL det = b->has_transform ? FN(det)(&c, b, p) : (REAL)1;
            REAL fl[6];
            if (dom->scheme == OB_ADV_LINEAR_UPWIND) FN(fluxes)(&c, b, p, NULL, 0, fl);
            for (int dim = 0; dim < c.d; ++dim) {
                REAL t = det * ((const REAL *)b->velocity)[(size_t)dim * n + f] / dt;
                for (int bound = 0; bound < 2 * c.d; ++bound) {
                    if (!FN(at_bound)(p, bound, b)) continue;
                    if (b->b[bound].type != OB_FIXED) continue;
                    REAL fn = (REAL)((bound & 1) * 2 - 1);
                    REAL vel = FN(fixed_vel)(&c, b, bound, p, dim);
                    REAL alpha = FN(alpha_fixed)(&c, b, bound, p, bound >> 1);
                    REAL flux = FN(contra_fixed)(&c, b, bound, p, bound >> 1) * fn;
                    t -= vel * flux;
                    REAL bv = FN(visc)(&c, b, p);
                    REAL slip = b->b[bound].vel_neumann ? 1 : 0;
                    t += vel * (1 - slip) * bv * 2 * alpha;
                }
                if (dom->scheme == OB_ADV_LINEAR_UPWIND)
                    FN(deferred_correction)(b, c.d, p, fl, (const REAL *)b->velocity + (size_t)dim * n, &t);
                if (S) t -= S[(size_t)dim * dom->ncells + b->global_offset + f];
                t /= det;
                t += FN(source)(b, p, dim);
                if (apply_pressure_gradient) {
                    REAL g[3];
                    FN(pressure_grad)(&c, b, p, g);
                    t -= g[dim] * dt;
                }
                rhs[(size_t)dim * dom->ncells + b->global_offset + f] = t;
            }
        }
    }
    return 0;
}

/* The LINEAR_UPWIND deferred correction of a passive-scalar channel (advection.cu:650-684)
 * alone, for the oracle's numpy restatement of the scalar right-hand side: out[g] = the sum
 * the kernel subtracts face by face (before the division by det), phi in global cell order. */
int FN(o_deferred_field)(const ob_domain *dom, const REAL *phi, REAL *out) {
    if (!FN(domain_ok)(dom)) return -1;
    FN(octx) c = {dom, dom->dims, 2 * dom->dims * dom->dims + 1};
    for (int bi = 0; bi < dom->nblocks; ++bi) {
        const ob_block *b = &dom->blocks[bi];
        int n = FN(ncells)(b);
        O_PAR
        for (int f = 0; f < n; ++f) {
            int p[3];
            FN(unflat)(f, b->size, p);
            REAL fl[6];
            FN(fluxes)(&c, b, p, NULL, 0, fl);
            REAL t = 0;
            FN(deferred_correction)(b, c.d, p, fl, phi + b->global_offset, &t);
            out[b->global_offset + f] = t;
        }
    }
    return 0;
}

/* PISO_build_pressure_matrix, pressure.cu:25-193 (orthogonal part, useFaceTransform=false) */
int FN(o_pressure_matrix)(const ob_domain *dom, const REAL *adiag, int flags, const int32_t *row, int32_t *col,
                          REAL *val) {
    if (!FN(domain_ok)(dom)) return -1;
    FN(octx) c = {dom, dom->dims, 2 * dom->dims * dom->dims + 1};
    FN(ndata) nd = {&c, OK_TYPES, 0, NULL, NULL, 0};
    int bad = 0;
    for (int bi = 0; bi < dom->nblocks; ++bi) {
        const ob_block *b = &dom->blocks[bi];
        int n = FN(ncells)(b);
        O_PAR_OR(bad)
        for (int f = 0; f < n; ++f) {
            int p[3];
            FN(unflat)(f, b->size, p);
            int g = b->global_offset + f;
            int idx[7];
            REAL rv[7] = {0, 0, 0, 0, 0, 0, 0};
            REAL ap[3];
            for (int dd = 0; dd < c.d; ++dd) ap[dd] = FN(alpha)(&c, b, p, dd);
            const REAL raP = (REAL)1.0 / adiag[g];
            idx[0] = g;
            for (int bound = 0; bound < 2 * c.d; ++bound) {
                int dim = bound >> 1;
                if (!(FN(at_bound)(p, bound, b) && FN(is_empty)(b, bound))) {
                    int np[3], nw;
                    const ob_block *nb = FN(neighbor)(&c, b, p, bound, np, &nw);
                    REAL an = FN(alpha)(&c, nb, np, nw);
                    int ng = FN(flat)(np, nb->size) + nb->global_offset;
                    REAL raN = (REAL)1.0 / adiag[ng];
                    REAL coef = (REAL)0.5 * (ap[dim] * raP + an * raN);
                    rv[0] -= coef;
                    rv[bound + 1] += coef;
                    idx[bound + 1] = ng;
                    if (flags) FN(matrix_corner)(&nd, b, p, bound, flags, raP, 0, adiag, (REAL)1, 1, (REAL)-1, &rv[0], rv);
                } else {
                    idx[bound + 1] = -1;
                }
            }
            for (int dd = c.d; dd < 3; ++dd) { idx[dd * 2 + 1] = -1; idx[dd * 2 + 2] = -1; }
            int w = FN(emit_row)(idx, rv, 7, col + row[g], val + row[g], 0, 1);
            if (w != row[g + 1] - row[g]) bad = 1;
        }
    }
    return bad ? -2 : 0;
}

/* PISO_build_pressure_rhs, pressure.cu:349-522 (CENTRAL, boundary sources on) */
int FN(o_pressure_rhs)(const ob_domain *dom, const REAL *adiag, const int32_t *crow, const int32_t *ccol,
                       const REAL *cval, const REAL *vel_result, REAL *prhs) {
    if (!FN(domain_ok)(dom)) return -1;
    FN(octx) c = {dom, dom->dims, 2 * dom->dims * dom->dims + 1};
    const REAL dt = (REAL)dom->dt;
    for (int bi = 0; bi < dom->nblocks; ++bi) {
        const ob_block *b = &dom->blocks[bi];
        int n = FN(ncells)(b);
        O_PAR
        for (int f = 0; f < n; ++f) {
            int p[3];
            FN(unflat)(f, b->size, p);
            int g = b->global_offset + f;
            const REAL rDiag = 1 / adiag[g];
            const REAL det = b->has_transform ? FN(det)(&c, b, p) : (REAL)1;
            int cs = crow[g], rs = crow[g + 1] - cs;
            for (int dim = 0; dim < c.d; ++dim) {
                const REAL velOld = ((const REAL *)b->velocity)[(size_t)dim * n + f] / dt;
                REAL H = 0;
                for (int i = 0; i < rs && i < 7; ++i) {
                    int ix = ccol[cs + i];
                    if (ix != g) H += cval[cs + i] * vel_result[ix + (size_t)dom->ncells * dim];
                }
                REAL S = 0;
                for (int bound = 0; bound < 2 * c.d; ++bound) {
                    if (!FN(at_bound)(p, bound, b)) continue;
                    if (b->b[bound].type != OB_FIXED) continue;
                    REAL fn = (REAL)((bound & 1) * 2 - 1);
                    REAL vel = FN(fixed_vel)(&c, b, bound, p, dim);
                    REAL alpha = FN(alpha_fixed)(&c, b, bound, p, bound >> 1);
                    REAL flux = FN(contra_fixed)(&c, b, bound, p, bound >> 1) * fn;
                    S -= vel * flux;
                    REAL bv = FN(visc)(&c, b, p);
                    REAL slip = b->b[bound].vel_neumann ? 1 : 0;
                    S += vel * (1 - slip) * bv * 2 * alpha;
                }
                S /= det;
                S += FN(source)(b, p, dim);
                prhs[g + (size_t)dom->ncells * dim] = rDiag * (velOld - H + S);
            }
        }
    }
    return 0;
}

/* k_computePressureRHSdivergenceFromFlux, pressure.cu:755-803 (useFaceTransform=false) */
int FN(o_pressure_div)(const ob_domain *dom, const REAL *prhs, int time_step_norm, REAL *div) {
    if (!FN(domain_ok)(dom)) return -1;
    FN(octx) c = {dom, dom->dims, 2 * dom->dims * dom->dims + 1};
    for (int bi = 0; bi < dom->nblocks; ++bi) {
        const ob_block *b = &dom->blocks[bi];
        int n = FN(ncells)(b);
        O_PAR
        for (int f = 0; f < n; ++f) {
            int p[3];
            FN(unflat)(f, b->size, p);
            REAL fl[6];
            FN(fluxes)(&c, b, p, prhs, 0, fl);
            REAL dv = 0;
            for (int dim = 0; dim < c.d; ++dim) dv += fl[dim * 2 + 1] - fl[dim * 2];
            if (time_step_norm) dv /= (REAL)dom->dt;
            div[f + b->global_offset] = dv;
        }
    }
    return 0;
}

/* PISO_update_velocity, velocity_correction.cu:26-60 */
int FN(o_correct_velocity)(const ob_domain *dom, const REAL *adiag, const REAL *prhs, int time_step_norm, REAL *vel_result) {
    if (!FN(domain_ok)(dom)) return -1;
    FN(octx) c = {dom, dom->dims, 2 * dom->dims * dom->dims + 1};
    for (int bi = 0; bi < dom->nblocks; ++bi) {
        const ob_block *b = &dom->blocks[bi];
        int n = FN(ncells)(b);
        O_PAR
        for (int f = 0; f < n; ++f) {
            int p[3];
            FN(unflat)(f, b->size, p);
            int g = b->global_offset + f;
            const REAL rDiag = 1 / adiag[g];
            REAL gr[3];
            FN(pressure_grad)(&c, b, p, gr);
            for (int dim = 0; dim < c.d; ++dim) {
                REAL u = -rDiag * gr[dim];
                if (time_step_norm) u *= (REAL)dom->dt;
                size_t o = g + (size_t)dom->ncells * dim;
                u += prhs[o];
                vel_result[o] = u;
            }
        }
    }
    return 0;
}

/* PISO_build_epot_matrix, potential.cu:29-233 (useFaceTransform=false). flags: the
 * non-orthogonal corner terms (potential.cu:103-150), whose face coefficients weigh both cells
 * with 1/adiag (interpolateNonOrthoLaplaceComponents with withA set, potential.cu:108) although
 * the orthogonal part uses 1 */
int FN(o_epot_matrix)(const ob_domain *dom, int flags, const REAL *adiag, const int32_t *row, int32_t *col,
                      REAL *val) {
    if (!FN(domain_ok)(dom)) return -1;
    if (flags && !adiag) return -1;
    FN(octx) c = {dom, dom->dims, 2 * dom->dims * dom->dims + 1};
    FN(ndata) nd = {&c, OK_TYPES, 0, NULL, NULL, 0};
    int bad = 0;
    for (int bi = 0; bi < dom->nblocks; ++bi) {
        const ob_block *b = &dom->blocks[bi];
        int n = FN(ncells)(b);
        O_PAR_OR(bad)
        for (int f = 0; f < n; ++f) {
            int p[3];
            FN(unflat)(f, b->size, p);
            int g = b->global_offset + f;
            int idx[7];
            REAL rv[7] = {0, 0, 0, 0, 0, 0, 0};
            REAL ap[3];
            for (int dd = 0; dd < c.d; ++dd) ap[dd] = FN(alpha)(&c, b, p, dd);
            const REAL raP = (REAL)1.0;
            idx[0] = g;
            for (int bound = 0; bound < 2 * c.d; ++bound) {
                int dim = bound >> 1;
                if (!(FN(at_bound)(p, bound, b) && FN(is_empty)(b, bound))) {
                    int np[3], nw;
                    const ob_block *nb = FN(neighbor)(&c, b, p, bound, np, &nw);
                    REAL an = FN(alpha)(&c, nb, np, nw);
                    int ng = FN(flat)(np, nb->size) + nb->global_offset;
                    const REAL raN = (REAL)1.0;
                    REAL coef = (REAL)0.5 * (ap[dim] * raP + an * raN);
                    rv[0] -= coef;
                    rv[bound + 1] += coef;
                    idx[bound + 1] = ng;
                    if (flags)
                        FN(matrix_corner)(&nd, b, p, bound, flags, (REAL)1.0 / adiag[g], 0, adiag, (REAL)1, 1, (REAL)-1,
                                          &rv[0], rv);
                } else {
                    idx[bound + 1] = -1;
                    if (FN(is_epot_dirichlet)(b, bound, p)) rv[0] -= ap[dim] * raP;
                }
            }
            for (int dd = c.d; dd < 3; ++dd) { idx[dd * 2 + 1] = -1; idx[dd * 2 + 2] = -1; }
            /* thin-wall Robin correction, potential.cu:164-218 */
            for (int tw = 0; tw < 2 * c.d; ++tw) {
                int twDim = tw >> 1;
                if (!FN(at_bound)(p, tw, b)) continue;
                if (!FN(is_thin_wall)(b, tw, p)) continue;
                REAL Cw = (REAL)b->b[tw].pot_cw;
                REAL mnP = FN(minv_norm)(&c, b, p, twDim);
                for (int tDim = 0; tDim < c.d; ++tDim) {
                    if (tDim == twDim) continue;
                    for (int ts = 0; ts < 2; ++ts) {
                        int tb = tDim * 2 + ts;
                        if (idx[tb + 1] < 0) continue;
                        int tfs = FN(face_sign)(tb);
                        int tn[3] = {p[0], p[1], p[2]};
                        if (FN(at_bound)(p, tb, b)) {
                            if (b->b[tb].type == OB_PERIODIC) tn[tDim] = ts == 1 ? 0 : b->size[tDim] - 1;
                            else continue;
                        } else {
                            tn[tDim] = p[tDim] + tfs;
                        }
                        if (!FN(is_thin_wall)(b, tw, tn)) continue;
                        REAL ant = FN(alpha)(&c, b, tn, tDim);
                        REAL mnN = FN(minv_norm)(&c, b, tn, twDim);
                        REAL sc = Cw * (REAL)0.5 * (ap[tDim] * mnP + ant * mnN);
                        rv[0] -= sc;
                        rv[tb + 1] += sc;
                    }
                }
            }
            int w = FN(emit_row)(idx, rv, 7, col + row[g], val + row[g], 0, 1);
            if (w != row[g + 1] - row[g]) bad = 1;
        }
    }
    return bad ? -2 : 0;
}

/* k_computeEpotRHS, potential.cu:272-322 (useFaceTransform=false). vf: [d][n] */
int FN(o_epot_rhs)(const ob_domain *dom, const REAL *vf, REAL *div) {
    if (!FN(domain_ok)(dom)) return -1;
    FN(octx) c = {dom, dom->dims, 2 * dom->dims * dom->dims + 1};
    for (int bi = 0; bi < dom->nblocks; ++bi) {
        const ob_block *b = &dom->blocks[bi];
        int n = FN(ncells)(b);
        O_PAR
        for (int f = 0; f < n; ++f) {
            int p[3];
            FN(unflat)(f, b->size, p);
            REAL fl[6];
            FN(fluxes)(&c, b, p, vf, 1, fl);
            REAL dv = 0;
            for (int dim = 0; dim < c.d; ++dim) dv += fl[dim * 2 + 1] - fl[dim * 2];
            for (int bound = 0; bound < 2 * c.d; ++bound) {
                if (!FN(at_bound)(p, bound, b) || !FN(is_empty)(b, bound)) continue;
                if (!FN(is_epot_dirichlet)(b, bound, p)) continue;
                REAL gv = FN(pot_value)(b, bound, p);
                if (gv == 0) continue;
                dv -= FN(alpha)(&c, b, p, bound >> 1) * gv;
            }
            div[f + b->global_offset] = dv;
        }
    }
    return 0;
}

/* computeCurrentDensityFaceBased, potential.cu:708-885. phi [n], uxb [3][n], J [3][n] */
int FN(o_current_density)(const ob_domain *dom, const REAL *phi, const REAL *uxb, REAL *J) {
    if (!FN(domain_ok)(dom)) return -1;
    FN(octx) c = {dom, dom->dims, 2 * dom->dims * dom->dims + 1};
    const int nc = dom->ncells;
    for (int bi = 0; bi < dom->nblocks; ++bi) {
        const ob_block *b = &dom->blocks[bi];
        int n = FN(ncells)(b);
        O_PAR
        for (int f = 0; f < n; ++f) {
            int p[3];
            FN(unflat)(f, b->size, p);
            int g = f + b->global_offset;
            REAL phiP = phi[g];
            REAL ap[3] = {0, 0, 0}, cvp[3] = {0, 0, 0};
            for (int dd = 0; dd < c.d; ++dd) {
                ap[dd] = FN(alpha)(&c, b, p, dd);
                cvp[dd] = FN(contra)(&c, b, p, dd, uxb);
            }
            REAL jf[6] = {0, 0, 0, 0, 0, 0};
            for (int bound = 0; bound < 2 * c.d; ++bound) {
                int dim = bound >> 1, up = bound & 1;
                REAL fs = (REAL)FN(face_sign)(bound);
                int atb = FN(at_bound)(p, bound, b);
                if (atb && FN(is_empty)(b, bound)) {
                    if (FN(is_insulating_wall)(b, bound, p)) jf[bound] = 0;
                    else if (FN(is_epot_dirichlet)(b, bound, p)) {
                        REAL gv = FN(pot_value)(b, bound, p);
                        jf[bound] = ap[dim] * fs * (phiP - gv) + cvp[dim];
                    } else jf[bound] = cvp[dim];
                    continue;
                }
                REAL phiN, aN, cvN;
                if (atb) {
                    const ob_bound *bd = &b->b[bound];
                    if (bd->type == OB_CONNECTED) {
                        int np[3];
                        const ob_block *nb = &dom->blocks[bd->conn_block];
                        FN(conn_pos)(&c, p, dim, bd, 0, np);
                        int ca = FN(conn_channel)(&c, dim, dim, bd);
                        phiN = phi[FN(flat)(np, nb->size) + nb->global_offset];
                        aN = FN(alpha)(&c, nb, np, ca);
                        cvN = FN(contra)(&c, nb, np, ca, uxb);
                        if ((bd->conn_axes[0] & 1) == up) cvN = -cvN;
                    } else if (bd->type == OB_PERIODIC) {
                        int np[3] = {p[0], p[1], p[2]};
                        np[dim] = up ? 0 : b->size[dim] - 1;
                        phiN = phi[FN(flat)(np, b->size) + b->global_offset];
                        aN = FN(alpha)(&c, b, np, dim);
                        cvN = FN(contra)(&c, b, np, dim, uxb);
                    } else {
                        jf[bound] = 0;
                        continue;
                    }
                } else {
                    int np[3] = {p[0], p[1], p[2]};
                    np[dim] += FN(face_sign)(bound);
                    phiN = phi[FN(flat)(np, b->size) + b->global_offset];
                    aN = FN(alpha)(&c, b, np, dim);
                    cvN = FN(contra)(&c, b, np, dim, uxb);
                }
                REAL coef = (REAL)0.5 * (ap[dim] + aN);
                REAL chi = (REAL)0.5 * (cvp[dim] + cvN);
                jf[bound] = -coef * fs * (phiN - phiP) + chi;
            }
            REAL ja[3] = {0, 0, 0};
            for (int dd = 0; dd < c.d; ++dd) ja[dd] = (REAL)0.5 * (jf[dd * 2] + jf[dd * 2 + 1]);
            REAL Jv[3] = {0, 0, 0};
            if (b->has_transform) {
                const REAL *t = FN(T)(b, p, c.ts);
                const REAL *M = t;
                REAL inv = (REAL)1 / t[2 * c.d * c.d];
                if (c.d == 2) {
                    Jv[0] = (M[0] * ja[0] + M[1] * ja[1]) * inv;
                    Jv[1] = (M[2] * ja[0] + M[3] * ja[1]) * inv;
                    Jv[2] = uxb[g + (size_t)nc * 2];
                } else {
                    Jv[0] = (M[0] * ja[0] + M[1] * ja[1] + M[2] * ja[2]) * inv;
                    Jv[1] = (M[3] * ja[0] + M[4] * ja[1] + M[5] * ja[2]) * inv;
                    Jv[2] = (M[6] * ja[0] + M[7] * ja[1] + M[8] * ja[2]) * inv;
                }
            } else {
                for (int dd = 0; dd < c.d; ++dd) Jv[dd] = ja[dd];
                if (c.d == 2) Jv[2] = uxb[g + (size_t)nc * 2];
            }
            J[g] = Jv[0];
            J[g + nc] = Jv[1];
            J[g + (size_t)nc * 2] = Jv[2];
        }
    }
    return 0;
}

/* det / inverse of MatrixSquare, common/transformations.h:437-495, 558-562 */
static REAL FN(mdet)(int d, const REAL *m) {
    if (d == 2) return m[0] * m[3] - m[2] * m[1];
    return m[0] * (m[4] * m[8] - m[7] * m[5]) - m[1] * (m[3] * m[8] - m[6] * m[5]) + m[2] * (m[3] * m[7] - m[6] * m[4]);
}
static void FN(minv)(int d, const REAL *m, REAL det, REAL *a) {
    if (d == 2) {
        a[0] = m[3]; a[1] = -m[1]; a[2] = -m[2]; a[3] = m[0];
    } else {
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
    REAL s = (REAL)1 / det;
    for (int i = 0; i < d * d; ++i) a[i] *= s;
}

/* k_CoordsToTransforms, grid_gen.cu:298-354. cells[3]: cell counts (1 for unused);
 * coords [d][vz][vy][vx]; out [cells][2d^2+1]. */
int FN(o_transforms)(int d, const int32_t *cells, const REAL *coords, REAL *out) {
    int vs[3] = {cells[0] + 1, d > 1 ? cells[1] + 1 : 1, d > 2 ? cells[2] + 1 : 1};
    size_t vstride = (size_t)vs[0] * vs[1] * vs[2];
    int cs[3] = {cells[0], d > 1 ? cells[1] : 1, d > 2 ? cells[2] : 1};
    int ncell = cs[0] * cs[1] * cs[2];
    int ts = 2 * d * d + 1;
    for (int f = 0; f < ncell; ++f) {
        int p[3];
        FN(unflat)(f, cs, p);
        REAL fc[6][3];
        memset(fc, 0, sizeof(fc));
        for (int v = 0; v < (1 << d); ++v) {
            int vp[3] = {p[0] + (v & 1), p[1] + ((v & 2) != 0), p[2] + ((v & 4) != 0)};
            REAL co[3];
            for (int k = 0; k < d; ++k) co[k] = coords[k * vstride + FN(flat)(vp, vs)];
            for (int fi = 0; fi < d; ++fi) {
                int up = (v & (1 << fi)) != 0;
                for (int k = 0; k < d; ++k) fc[fi * 2 + up][k] += co[k];
            }
        }
        REAL norm = (REAL)1.0 / (REAL)(1 << (d - 1));
        for (int fi = 0; fi < 2 * d; ++fi)
            for (int k = 0; k < d; ++k) fc[fi][k] *= norm;
        REAL *t = out + (size_t)f * ts;
        for (int i = 0; i < d; ++i)
            for (int k = 0; k < d; ++k) t[i * d + k] = fc[k * 2 + 1][i] - fc[k * 2][i];
        REAL det = FN(mdet)(d, t);
        FN(minv)(d, t, det, t + d * d);
        t[2 * d * d] = det;
    }
    return 0;
}

/* k_CoordsToFaceTransforms, grid_gen.cu:398-527. out [d][vz][vy][vx][2d^2+1] (zeros where
 * no face exists). */
int FN(o_face_transforms)(int d, const int32_t *cells, const REAL *coords, REAL *out) {
    int vs[3] = {cells[0] + 1, d > 1 ? cells[1] + 1 : 1, d > 2 ? cells[2] + 1 : 1};
    size_t vstride = (size_t)vs[0] * vs[1] * vs[2];
    int ts = 2 * d * d + 1;
    int nfv = 1 << (d - 1), nev = 1 << (d - 2);
    memset(out, 0, sizeof(REAL) * vstride * d * ts);
    for (size_t f = 0; f < vstride; ++f) {
        int p[3];
        FN(unflat)((int)f, vs, p);
        for (int dim = 0; dim < d; ++dim) {
            int skip = 0;
            for (int i = 1; i < d; ++i) {
                int td = (dim + i) % d;
                if (p[td] == vs[td] - 1) skip = 1;
            }
            if (skip) continue;
            REAL t[9];
            memset(t, 0, sizeof(t));
            REAL nw = (REAL)0.5;
            for (int up = 0; up < 2; ++up) {
                int fs = up * 2 - 1;
                int atb = up ? p[dim] == vs[dim] - 1 : p[dim] == 0;
                int fp[3] = {p[0], p[1], p[2]};
                if (!atb) fp[dim] += fs;
                else nw = (REAL)1.0;
                for (int v = 0; v < nfv; ++v) {
                    int vp[3] = {fp[0], fp[1], fp[2]};
                    for (int i = 1; i < d; ++i) vp[(dim + i) % d] += (v & i) != 0;
                    for (int k = 0; k < d; ++k) t[k * d + dim] += coords[k * vstride + FN(flat)(vp, vs)] * (REAL)fs;
                }
            }
            nw /= (REAL)nfv;
            for (int k = 0; k < d; ++k) t[k * d + dim] *= nw;
            REAL tnw = (REAL)1.0 / (REAL)nev;
            for (int kk = 1; kk < d; ++kk) {
                int td = (dim + kk) % d;
                for (int v = 0; v < nfv; ++v) {
                    int vp[3] = {p[0], p[1], p[2]};
                    for (int i = 1; i < d; ++i) vp[(dim + i) % d] += (v & i) != 0;
                    REAL tfs = (v & kk) ? (REAL)1 : (REAL)-1;
                    for (int k = 0; k < d; ++k) t[k * d + td] += coords[k * vstride + FN(flat)(vp, vs)] * tfs;
                }
                for (int k = 0; k < d; ++k) t[k * d + td] *= tnw;
            }
            REAL *o = out + ((size_t)dim * vstride + f) * ts;
            for (int i = 0; i < d * d; ++i) o[i] = t[i];
            REAL det = FN(mdet)(d, t);
            FN(minv)(d, t, det, o + d * d);
            o[2 * d * d] = det;
        }
    }
    return 0;
}

/* ---------------- sparse linear algebra ---------------- */

int FN(o_spmv)(int n, const int32_t *row, const int32_t *col, const REAL *val, const REAL *x, REAL *y) {
    O_PAR
    for (int i = 0; i < n; ++i) {
        REAL s = 0;
        for (int j = row[i]; j < row[i + 1]; ++j) s += val[j] * x[col[j]];
        y[i] = s;
    }
    return 0;
}

#ifndef ORACLE_ST_ENUM
#define ORACLE_ST_ENUM
enum { ST_RUNNING = 0, ST_CONVERGED = 1, ST_NONFINITE = 2, ST_MAXIT = 3, ST_RISING = 4 };
#endif

/* Inverse diagonal of the Jacobi preconditioner (kInverseDiagonal, krylov.cu:272-283): the sum
 * of the row's diagonal entries, 1 where it is zero. */
static void FN(o_jacobi_dinv)(int n, const int32_t *row, const int32_t *col, const REAL *val, REAL *dinv) {
    for (int i = 0; i < n; ++i) {
        REAL d = 0;
        for (int j = row[i]; j < row[i + 1]; ++j)
            if (col[j] == i) d += val[j];
        dinv[i] = d != (REAL)0 ? (REAL)1 / d : (REAL)1;
    }
}

/* Fused CG of krylov::cgSolve, krylov.cu:1058-1125 (kResidual :291-366, kCgDirection
 * :427-498, kCgUpdate :501-568, kRestoreBest :369-375, collectResults :958-980), one RHS.
 * jacobi: z = diag(A)^-1 r (PM_JACOBI, krylov.cu:414-425), rho = (r, z); otherwise z = r.
 * project: subtract the mean of b first (singular consistent systems). */
int FN(o_cg)(int n, const int32_t *row, const int32_t *col, const REAL *val, const REAL *b_in, REAL *x,
             int maxit, double tol, int normalized, int reset_every, int return_best, int project, int jacobi,
             ob_result *res) {
    const double ns = normalized ? 1.0 / sqrt((double)n) : 1.0;
    REAL *r = (REAL *)malloc(sizeof(REAL) * n), *pa = (REAL *)malloc(sizeof(REAL) * n);
    REAL *pb = (REAL *)malloc(sizeof(REAL) * n), *q = (REAL *)malloc(sizeof(REAL) * n);
    REAL *bw = (REAL *)malloc(sizeof(REAL) * n), *xbest = (REAL *)malloc(sizeof(REAL) * n);
    REAL *dv = (REAL *)malloc(sizeof(REAL) * n);
    if (jacobi) FN(o_jacobi_dinv)(n, row, col, val, dv);
    else for (int i = 0; i < n; ++i) dv[i] = (REAL)1;
    memcpy(bw, b_in, sizeof(REAL) * n);
    if (project) {
        double m = 0;
        for (int i = 0; i < n; ++i) m += (double)bw[i];
        REAL mm = (REAL)(m / (double)n);
        for (int i = 0; i < n; ++i) bw[i] -= mm;
    }
    double rho = 0, alpha = 0, beta = 0, crit = 0, best = INFINITY, last = 0;
    int it = 0, bestIt = -1, rising = 0, done = ST_RUNNING, restart = 1, saveBest = 0, usedIt = 0;
    /* initial residual */
    {
        double acc = 0, accz = 0;
        O_PAR_SUM(acc, accz)
        for (int i = 0; i < n; ++i) {
            REAL s = 0;
            for (int j = row[i]; j < row[i + 1]; ++j) s += val[j] * x[col[j]];
            REAL ri = bw[i] - s;
            r[i] = ri;
            acc += (double)ri * (double)ri;
            accz += (double)ri * (double)(ri * dv[i]);
        }
        rho = jacobi ? accz : acc;
        restart = 1;
        crit = sqrt(acc) * ns;
        if (!isfinite(crit)) { done = ST_NONFINITE; usedIt = 0; }
        else if (crit == 0.0) { done = ST_CONVERGED; usedIt = 0; }
    }
    REAL *pin = pa, *pout = pb;
    long guard = 0;
    while (done == ST_RUNNING && guard++ < (long)maxit + 4) {
        if (reset_every > 0 && it > 0 && (it % reset_every) == 0) {
            double acc = 0, accz = 0;
            O_PAR_SUM(acc, accz)
            for (int i = 0; i < n; ++i) {
                REAL s = 0;
                for (int j = row[i]; j < row[i + 1]; ++j) s += val[j] * x[col[j]];
                REAL ri = bw[i] - s;
                r[i] = ri;
                acc += (double)ri * (double)ri;
                accz += (double)ri * (double)(ri * dv[i]);
            }
            rho = jacobi ? accz : acc;
            restart = 1;
        }
        /* direction */
        double pq = 0;
        const REAL bt = (REAL)beta;
        O_PAR
        for (int i = 0; i < n; ++i) {
            const REAL zi = jacobi ? r[i] * dv[i] : r[i];
            pout[i] = restart ? zi : zi + bt * pin[i];
        }
        O_PAR_SUM(pq)
        for (int i = 0; i < n; ++i) {
            REAL s = 0;
            for (int j = row[i]; j < row[i + 1]; ++j) s += val[j] * pout[col[j]];
            q[i] = s;
            pq += (double)pout[i] * (double)s;
        }
        {
            double a = rho / pq;
            restart = 0;
            if (!isfinite(a)) { done = ST_NONFINITE; usedIt = it; break; }
            alpha = a;
        }
        /* update */
        double rr = 0, rz = 0;
        const REAL al = (REAL)alpha;
        O_PAR_SUM(rr, rz)
        for (int i = 0; i < n; ++i) {
            REAL xi = x[i];
            if (saveBest && return_best) xbest[i] = xi;
            x[i] = xi + al * pout[i];
            REAL ri = r[i] - al * q[i];
            r[i] = ri;
            rr += (double)ri * (double)ri;
            rz += (double)ri * (double)(ri * dv[i]);
        }
        crit = sqrt(rr) * ns;
        ++it;
        saveBest = 0;
        usedIt = it - 1;
        if (!isfinite(crit)) done = ST_NONFINITE;
        else {
            if (return_best) {
                if (crit < best) { best = crit; bestIt = it - 1; saveBest = 1; }
                if (it > 1 && crit >= last) ++rising; else rising = 0;
                last = crit;
            }
            if (crit < tol) done = ST_CONVERGED;
            else if (it >= maxit) done = ST_MAXIT;
            else if (return_best && rising >= 100) done = ST_RISING;
            else {
                double gamma = jacobi ? rz : rr;
                beta = gamma / rho;
                rho = gamma;
                if (!isfinite(beta)) done = ST_NONFINITE;
            }
        }
        REAL *tmp = pin; pin = pout; pout = tmp;
    }
    if (return_best && (done == ST_MAXIT || done == ST_RISING) && bestIt >= 0 && best < crit)
        memcpy(x, xbest, sizeof(REAL) * n);
    res->final_residual = crit;
    res->used_iterations = usedIt;
    res->converged = done == ST_CONVERGED;
    res->is_finite = isfinite(crit) ? 1 : 0;
    if (return_best && (done == ST_MAXIT || done == ST_RISING) && bestIt >= 0 && best < crit) {
        res->used_iterations = bestIt;
        res->final_residual = best;
    }
    free(r); free(pa); free(pb); free(q); free(bw); free(xbest); free(dv);
    return 0;
}

/* Fused BiCGStab of krylov::bicgstabSolve, krylov.cu:1129-1186 (kResidual with r0,
 * kBicgP :598-610, kBicgV :613-670, kBicgT :674-750, kBicgUpdate :754-824), one RHS.
 * jacobi: right preconditioning with diag(A)^-1 -- v = A D^-1 p, t = A D^-1 s and the
 * solution updated with D^-1 (alpha p + omega s) (loadPre, krylov.cu:243-247). */
int FN(o_bicgstab)(int n, const int32_t *row, const int32_t *col, const REAL *val, const REAL *b, REAL *x,
                   int maxit, double tol, int normalized, int jacobi, ob_result *res) {
    const double ns = normalized ? 1.0 / sqrt((double)n) : 1.0;
    REAL *r = (REAL *)malloc(sizeof(REAL) * n), *r0 = (REAL *)malloc(sizeof(REAL) * n);
    REAL *p = (REAL *)calloc(n, sizeof(REAL)), *v = (REAL *)calloc(n, sizeof(REAL));
    REAL *t = (REAL *)malloc(sizeof(REAL) * n), *s = (REAL *)malloc(sizeof(REAL) * n);
    REAL *dv = (REAL *)malloc(sizeof(REAL) * n);
    if (jacobi) FN(o_jacobi_dinv)(n, row, col, val, dv);
    else for (int i = 0; i < n; ++i) dv[i] = (REAL)1;
    double rho = 0, alpha = 0, beta = 0, omega = 0, crit = 0, halfCrit = 0;
    int it = 0, done = ST_RUNNING, restart = 1, halfDone = 0, usedIt = 0;
    {
        double acc = 0;
        O_PAR_SUM(acc)
        for (int i = 0; i < n; ++i) {
            REAL sm = 0;
            for (int j = row[i]; j < row[i + 1]; ++j) sm += val[j] * x[col[j]];
            REAL ri = b[i] - sm;
            r[i] = ri;
            r0[i] = ri;
            acc += (double)ri * (double)ri;
        }
        rho = acc;
        restart = 1;
        crit = sqrt(acc) * ns;
        if (!isfinite(crit)) { done = ST_NONFINITE; usedIt = 0; }
        else if (crit < tol) { done = ST_CONVERGED; usedIt = -1; }
    }
    long guard = 0;
    while (done == ST_RUNNING && guard++ < (long)maxit + 4) {
        {
            const REAL bt = (REAL)beta, om = (REAL)omega;
            O_PAR
            for (int i = 0; i < n; ++i) p[i] = restart ? r[i] : r[i] + bt * (p[i] - om * v[i]);
        }
        {
            double acc = 0;
            O_PAR_SUM(acc)
            for (int i = 0; i < n; ++i) {
                REAL sm = 0;
                for (int j = row[i]; j < row[i + 1]; ++j) sm += val[j] * (jacobi ? p[col[j]] * dv[col[j]] : p[col[j]]);
                v[i] = sm;
                acc += (double)r0[i] * (double)sm;
            }
            double a = rho / acc;
            restart = 0;
            if (!isfinite(a)) { done = ST_NONFINITE; usedIt = it; break; }
            alpha = a;
        }
        {
            const REAL al = (REAL)alpha;
            O_PAR
            for (int i = 0; i < n; ++i) s[i] = r[i] - al * v[i];
            double ts_ = 0, tt = 0, ss = 0;
            O_PAR_SUM(ts_, tt, ss)
            for (int i = 0; i < n; ++i) {
                REAL sm = 0;
                for (int j = row[i]; j < row[i + 1]; ++j) sm += val[j] * (jacobi ? s[col[j]] * dv[col[j]] : s[col[j]]);
                t[i] = sm;
                ts_ += (double)sm * (double)s[i];
                tt += (double)sm * (double)sm;
                ss += (double)s[i] * (double)s[i];
            }
            double critS = sqrt(ss) * ns;
            halfCrit = critS;
            if (!isfinite(critS)) { done = ST_NONFINITE; crit = critS; usedIt = it; break; }
            else if (critS < tol) { halfDone = 1; omega = 0.0; }
            else {
                double om = ts_ / tt;
                if (!isfinite(om)) { done = ST_NONFINITE; crit = critS; usedIt = it; break; }
                omega = om;
            }
        }
        {
            const REAL al = (REAL)alpha, om = (REAL)omega;
            double rr0 = 0, rr = 0;
            O_PAR_SUM(rr0, rr)
            for (int i = 0; i < n; ++i) {
                REAL si = r[i] - al * v[i];
                if (halfDone) {
                    x[i] += al * p[i] * dv[i];
                    r[i] = si;
                } else {
                    x[i] += (al * p[i] + om * si) * dv[i];
                    REAL ri = si - om * t[i];
                    r[i] = ri;
                    rr0 += (double)r0[i] * (double)ri;
                    rr += (double)ri * (double)ri;
                }
            }
            if (halfDone) { done = ST_CONVERGED; crit = halfCrit; usedIt = it; }
            else {
                crit = sqrt(rr) * ns;
                if (!isfinite(crit)) { done = ST_NONFINITE; usedIt = it; }
                else if (crit < tol) { done = ST_CONVERGED; usedIt = it + 1; }
                else {
                    ++it;
                    usedIt = it - 1;
                    double bnew = (rr0 / rho) * (alpha / omega);
                    rho = rr0;
                    beta = bnew;
                    if (!isfinite(beta)) done = ST_NONFINITE;
                    else if (it >= maxit) done = ST_MAXIT;
                }
            }
        }
    }
    res->final_residual = crit;
    res->used_iterations = usedIt;
    res->converged = done == ST_CONVERGED;
    res->is_finite = isfinite(crit) ? 1 : 0;
    free(r); free(r0); free(p); free(v); free(t); free(s); free(dv);
    return 0;
}

/* CSR transpose (closed form of the transpose krylov::csrTranspose computes with
 * cusparseCsr2cscEx2, krylov.cu:1620-1652): entries of a column in ascending row order. */
int FN(o_csr_transpose)(int n, int nnz, const int32_t *row, const int32_t *col, const REAL *val,
                        int32_t *trow, int32_t *tcol, REAL *tval) {
    for (int i = 0; i <= n; ++i) trow[i] = 0;
    for (int k = 0; k < nnz; ++k) trow[col[k] + 1]++;
    for (int i = 0; i < n; ++i) trow[i + 1] += trow[i];
    int *fill = (int *)malloc(sizeof(int) * (n + 1));
    for (int i = 0; i < n; ++i) fill[i] = trow[i];
    for (int i = 0; i < n; ++i)
        for (int k = row[i]; k < row[i + 1]; ++k) {
            int c = col[k];
            int o = fill[c]++;
            tcol[o] = i;
            tval[o] = val[k];
        }
    free(fill);
    return 0;
}

#undef FN
#undef FN1
#undef FN2

/*
 * phipict_hip.h — C ABI of the MI355X (gfx950) PISO-step and linear-solver library.
 *
 * Every entry point takes plain pointers and sizes: device pointers for field data,
 * host pointers for descriptors and result records, and a hipStream_t passed as
 * `void *stream` (NULL = the null stream). Scalars that the reference takes as
 * 1-element host tensors (timeStep, maxit, tol) are plain values or host arrays here.
 * All entry points return 0 on success and a negative PP_ERR_* code otherwise;
 * pp_last_error() returns the message of the last failure on the calling thread.
 *
 * Each declaration names the reference interface it replaces (paths relative to the
 * phiPICT source tree). The reference binds those as pybind11 members of `phipict._C`
 * (csrc/phipict/bindings.cpp:44-874); a maintainer binds these with ctypes, see the
 * stub in INTEGRATION.md.
 *
 * Layout conventions (identical to the reference, csrc/phipict/piso/device/launch.cuh:134-140,
 * csrc/phipict/domain/domain_structs_gpu.h:186-194):
 *  - block fields are [batch][channel][z][y][x] (x fastest), a block's cells are numbered
 *    x + sx*(y + sy*z);
 *  - flat solve vectors are [batch][channel][num_cells], block b's cells start at
 *    its global_offset;
 *  - CSR matrices: int32 row[n+1], int32 col[nnz] (ascending column per row),
 *    values [num_matrices][nnz];
 *  - a cell transform is {M[d][d], Minv[d][d], det} row-major, 2d^2+1 scalars.
 */
#ifndef PHIPICT_HIP_H
#define PHIPICT_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- enums (values are the reference's, csrc/phipict/domain/domain_structs_gpu.h:146-223) ---- */
enum { PP_F32 = 4, PP_F64 = 8 };
enum { PP_BOUND_CONNECTED = 20, PP_BOUND_PERIODIC = 21, PP_BOUND_FIXED = 30 };
enum { PP_BC_DIRICHLET = 0, PP_BC_NEUMANN = 1 };
enum { PP_POT_INSULATING = 0, PP_POT_OPEN = 1, PP_POT_DIRICHLET = 2, PP_POT_THIN_WALL = 3 };
enum { PP_CONV_NORM2 = 0, PP_CONV_NORM2_NORMALIZED = 1 };
enum {
    PP_OK = 0,
    PP_ERR_ARG = -1,      /* invalid argument / inconsistent sizes */
    PP_ERR_HIP = -2,      /* HIP runtime error */
    PP_ERR_UNSUPPORTED = -3,
    PP_ERR_NOMEM = -4,
};

/* ---- domain descriptors ---- */

/* One face of a block, 112 bytes. Face order [-x,+x,-y,+y,-z,+z]. */
typedef struct pp_boundary_desc {
    int32_t type;          /* PP_BOUND_* */
    int32_t vel_type;      /* FIXED: PP_BC_DIRICHLET or PP_BC_NEUMANN (velocity) */
    int32_t vel_static;    /* FIXED: 1 = velocity [vb][d]; 0 = [vb][d][face cells] */
    int32_t vel_batched;   /* FIXED: 1 = velocity has one slice per environment (vb = batch) */
    int32_t has_transform; /* FIXED: face transforms [face cells][2d^2+1] given */
    int32_t pot_type;      /* FIXED: PP_POT_* (used when pot_types == 0) */
    int32_t conn_block;    /* CONNECTED: index of the connected block */
    int32_t conn_axes[3];  /* CONNECTED: face-format axes of the connection */
    int32_t scalar_static; /* FIXED: passive scalar [sb][C] (1) or [sb][C][face cells] (0) */
    int32_t scalar_batched;/* FIXED: passive scalar has one slice per environment */
    int32_t pot_values_batched; /* FIXED: pot_values has one slice per environment */
    int32_t reserved;      /* zero */
    uint64_t vel;          /* device pointer (dtype) */
    uint64_t transform;    /* device pointer (dtype) or 0 */
    uint64_t pot_types;    /* device pointer int8 [face cells] or 0 */
    uint64_t pot_values;   /* device pointer (dtype) [face cells] ([batch][face cells]) or 0 */
    double   pot_cw;       /* thin-wall conductance ratio */
    uint64_t scalar;       /* FIXED: passive scalar face data (Dirichlet value or Neumann gradient) or 0 */
    uint64_t scalar_types; /* FIXED: int8 [C] PP_BC_* per channel, or 0 (all Dirichlet) */
} pp_boundary_desc;

/* One block, 104 + 6*112 = 776 bytes. */
typedef struct pp_block_desc {
    int32_t size[3];        /* cells in x, y, z (1 for unused dimensions) */
    int32_t cells;          /* size[0]*size[1]*size[2] */
    int32_t global_offset;  /* first global cell index */
    int32_t csr_offset;     /* first CSR entry of the block's rows */
    int32_t has_transform;
    int32_t vsource_mode;   /* 0 none, 1 static [vb][d], 2 per cell [vb][d][cells] */
    int32_t vsource_batched;/* source has one slice per environment */
    int32_t visc_mode;      /* 0 domain viscosity, 1 static [1], 2 per cell [cells] */
    int32_t transform_diagonal; /* 1: the off-diagonal entries of every cell's M and Minv are 0 */
    int32_t pad1;
    uint64_t transform;     /* component-major [2d^2+1][cells]: Block.transform transposed */
    uint64_t velocity;      /* [batch][d][cells] */
    uint64_t pressure;      /* [batch][cells] */
    uint64_t epot;          /* [batch][cells] or 0 */
    uint64_t vsource;       /* see vsource_mode */
    uint64_t viscosity;     /* see visc_mode */
    uint64_t scalar;        /* passive scalar [batch][C][cells] or 0 */
    pp_boundary_desc bnd[6];
} pp_block_desc;

typedef struct pp_csr {
    const int32_t *row;     /* [n+1] */
    const int32_t *col;     /* [nnz] */
    void *val;              /* [num_matrices][nnz] */
    int32_t n;
    int32_t nnz;
    int32_t num_matrices;   /* 1 = shared by all right-hand sides */
    int32_t ell_width;      /* > 0: the ELL copy below holds the same matrix */
    /* Optional ELL copy (no reference counterpart; the library's own coalesced layout of the
     * same entries): entry k of row i at [k][i], rows padded with zero values on the row's
     * own column. Column order within a row equals the CSR order. */
    const int32_t *ell_col; /* [ell_width][n] */
    void *ell_val;          /* [num_matrices][ell_width][n] */
} pp_csr;

/* Host-side view of a domain (the reference's DomainGPU, domain_structs_gpu.h:334-445). */
/* Convective schemes of the momentum and passive-scalar equations (AdvectionScheme,
 * csrc/phipict/domain/domain_structs_gpu.h:65-68): CENTRAL assembles the face value into the
 * matrix; LINEAR_UPWIND puts first-order upwind into the matrix and the second-order upwind
 * remainder on the right-hand side (deferred correction). */
enum { PP_ADV_CENTRAL = 0, PP_ADV_LINEAR_UPWIND = 1 };

typedef struct pp_domain {
    int32_t dims;           /* 2 or 3 */
    int32_t num_blocks;
    int32_t num_cells;      /* per environment */
    int32_t batch;          /* environments */
    int32_t dtype;          /* PP_F32 / PP_F64 */
    int32_t max_block_cells;
    double  viscosity;      /* domain viscosity */
    const pp_block_desc *blocks; /* DEVICE pointer to num_blocks descriptors */
    const void *time_step;  /* DEVICE pointer, [batch] (dtype) */
    void *A;                /* [batch][n] advection diagonal */
    void *velocity_rhs;     /* [batch][d][n] */
    void *velocity_result;  /* [batch][d][n] */
    void *pressure_rhs;     /* [batch][d][n] */
    void *pressure_rhs_div; /* [batch][n] */
    void *pressure_result;  /* [batch][n] */
    void *epot_rhs;         /* [batch][n] */
    void *epot_result;      /* [batch][n] */
    pp_csr C;               /* advection matrix, values [batch][nnz] */
    pp_csr P;               /* pressure matrix, values [batch][nnz] */
    pp_csr E;               /* electric potential matrix, values [nnz] */
    int32_t scalar_channels;  /* passive scalar channels C (0: none) */
    int32_t scalar_visc_count;/* 0: the domain viscosity; 1: one value for all channels; C */
    const void *scalar_viscosity; /* DEVICE pointer [scalar_visc_count] (dtype) */
    void *scalar_rhs;       /* [batch][C][n] */
    void *scalar_result;    /* [batch][C][n] */
    int32_t advection_scheme; /* PP_ADV_CENTRAL / PP_ADV_LINEAR_UPWIND (Domain::setAdvectionScheme) */
    int32_t reserved;
} pp_domain;

/* The reference's LinearSolverResultInfo (csrc/phipict/solvers/linear_solvers.h:180-185). */
typedef struct pp_solve_result {
    double  final_residual;
    int32_t used_iterations;
    int32_t converged;
    int32_t is_finite_residual;
    int32_t pad;
} pp_solve_result;

/* Wave-interleaved sliced ELL copy of a CSR matrix (no reference counterpart): rows in slices
 * of 64 / lanes; slice s holds slice_steps[s] columns of 64 entries, entry k * lanes + j of row
 * r of the slice at slice_off[s] + 64 k + r * lanes + j (col, val), padding with column 0 and
 * value 0. A wavefront reads one slice with `lanes` lanes per row. total: entries per operator
 * (the value stride of batched operators). The four pointers set, or all null. */
typedef struct pp_sell {
    const int32_t *slice_off, *slice_steps, *col;
    const void *val;
    int32_t lanes, pad;
    int64_t total;
} pp_sell;

/* One AMG level above the coarsest (reference layout of _C.AMGPCGSolve levels,
 * csrc/phipict/bindings.cpp:764-776: A row,col,val,dinv, R row,col,val, P row,col,val). */
typedef struct pp_amg_level {
    int32_t n, nnz_a, n_coarse, nnz_r, nnz_p, a_batched; /* a_batched: A values [B][nnz_a], dinv [B][n] */
    const int32_t *a_row, *a_col; const void *a_val; const void *dinv;
    const int32_t *r_row, *r_col; const void *r_val;
    const int32_t *p_row, *p_col; const void *p_val;
    /* optional ELL copy of A (layout of pp_csr.ell_col/ell_val; no reference counterpart) */
    int32_t a_ell_width, pad;
    const int32_t *a_ell_col; const void *a_ell_val;
    /* optional sliced ELL copies of A, R and P (values in the hierarchy dtype; A only when
     * shared by all right-hand sides) */
    pp_sell a_sell, r_sell, p_sell;
} pp_amg_level;

/* ---- library ---- */
const char *pp_last_error(void);
int pp_version(void);
int pp_device_sync(void);

/* ---- inductionless MHD cross products ----
 * pp_cross_eb replaces _compute_u_cross_eb_flat (src/phipict/simulation/mhd.py:309-355):
 * v [batch][vdims][n] (2D lifted to 3D with a zero z component), e_b [3] or [batch][3]
 * (eb_per_batch) -> out [batch][3][n] = v x e_b.
 * pp_lorentz_source replaces the force part of add_lorentz_force (mhd.py:543-594) for one
 * block: out [batch][dims][block_cells] = N (J x e_b)[:dims] over the block's cells
 * (block_offset .. + block_cells of J [batch][3][total]) plus the block's current velocity
 * source: src_kind 0 none, 1 static [1|batch][dims], 2 per cell [1|batch][dims][block_cells]
 * (src_batched: one per environment). stuart: N, one value or one per environment. */
int pp_cross_eb(const void *v, int32_t vdims, int32_t n, int32_t batch, const void *e_b, int32_t eb_per_batch, void *out,
                int32_t dtype, void *stream);
int pp_lorentz_source(const void *J, int32_t total, int32_t batch, int32_t dims, const void *e_b, int32_t eb_per_batch,
                      const void *stuart, int32_t stuart_per_batch, int32_t block_offset, int32_t block_cells,
                      const void *src, int32_t src_kind, int32_t src_batched, void *out, int32_t dtype, void *stream);

/* ---- measurement (no counterpart in the reference) ----
 * A sampled per-kernel timer: while started, a pseudo-random 1/sample_every of the launches of
 * each instrumented kernel (all of them for 1) is bracketed by HIP events on the stream it is
 * launched on.
 * pp_kernel_timer_read returns, per (kernel id, rows per launch), the summed duration of the
 * sampled launches, their number and the number of all launches; *count is the number of
 * entries (may exceed capacity). */
enum pp_kernel_id {
    PP_K_ADVECTION_MATRIX = 1, PP_K_ADVECTION_RHS = 2, PP_K_PRESSURE_MATRIX = 3, PP_K_PRESSURE_RHS = 4,
    PP_K_PRESSURE_DIV = 5, PP_K_CORRECT_VELOCITY = 6, PP_K_EPOT_RHS = 7, PP_K_CURRENT_DENSITY = 8,
    PP_K_COPY_RESULT = 9,
    PP_K_RESIDUAL = 20, PP_K_CG_DIRECTION = 21, PP_K_CG_UPDATE = 22, PP_K_BICG_P = 23, PP_K_BICG_V = 24,
    PP_K_BICG_T = 25, PP_K_BICG_UPDATE = 26, PP_K_CG_SMALL = 27, PP_K_BICG_SMALL = 28,
    PP_K_PCG_DOTS = 30, PP_K_PCG_DIRECTION = 31, PP_K_PCG_UPDATE = 32, PP_K_AMG_PRESMOOTH = 33,
    PP_K_AMG_JACOBI = 34, PP_K_AMG_LEVEL_RESIDUAL = 35, PP_K_AMG_RESTRICT = 36, PP_K_AMG_PROLONG = 37,
    PP_K_AMG_COARSE = 38, PP_K_MEAN = 39, PP_K_SUBTRACT_MEAN = 40, PP_K_AMG_POST_JACOBI = 41,
    PP_K_AMG_TAIL = 42,
    PP_K_SPMV = 50, PP_K_GALERKIN = 51, PP_K_GALERKIN_AP = 52,
    PP_K_SEG_CG_DIRECTION = 60, PP_K_SEG_CG_UPDATE = 61, PP_K_SEG_BICG = 62
};
int pp_kernel_timer_start(int32_t sample_every);
int pp_kernel_timer_stop(void);
int pp_kernel_timer_read(int32_t capacity, int32_t *kernel_ids, int32_t *rows, double *total_ms, int64_t *sampled,
                         int64_t *launches, int32_t *count);
/* A plain device copy of `bytes` (a multiple of 16) bytes by the kernel k_copy_bytes: the
 * reference point that scales hardware memory counters (FETCH_SIZE, WRITE_SIZE) to bytes. */
int pp_copy_bytes(const void *src, void *dst, int64_t bytes, void *stream);

/* ---- grid metrics ----
 * replaces CoordsToTransforms (csrc/phipict/grid/grid_gen.cu:356-391) and
 * CoordsToFaceTransforms (grid_gen.cu:529-): coords [d][(sz+1)][sy+1][sx+1] vertices ->
 * out [cells][2d^2+1] resp. [d][(sz+1)][sy+1][sx+1][2d^2+1]. */
int pp_coords_to_transforms(int32_t dims, const int32_t *cells, const void *coords, void *out,
                            int32_t dtype, void *stream);
int pp_coords_to_face_transforms(int32_t dims, const int32_t *cells, const void *coords, void *out,
                                 int32_t dtype, void *stream);
/* replaces TransformVectors (csrc/phipict/grid/transform_vectors.cu:84-136): vectors
 * [batch][dims][cells], transforms [1|batch][cells][2d^2+1] -> out[n][i][c] =
 * sum_k T[i][k] vectors[n][k][c] with T = Minv of cell c (inverse) or M. */
int pp_transform_vectors(int32_t dims, int32_t batch, int32_t cells, const void *vectors, const void *transforms,
                         int32_t batched_transforms, int32_t inverse, void *out, int32_t dtype, void *stream);

/* ---- PISO assembly (non-orthogonal correction flags accepted; the corner terms are
 * zero on the rectilinear grids these entry points are validated on) ---- */

/* out[c][r] = in[r][c] for a row-major [rows][cols] array (Block.transform [cells][2d^2+1]
 * to the component-major layout the block descriptors take); no reference counterpart. */
int pp_transpose_2d(const void *in, int32_t rows, int32_t cols, int32_t dtype, void *out, void *stream);

/* CSR pattern shared by C, P and E: row[n+1] and ascending columns (the row layout the
 * reference's matrix kernels write, csrc/phipict/piso/device/topology.cuh:168-233, and
 * the block offsets of Domain::PrepareSolve, csrc/phipict/domain/domain.cpp:426-561).
 * Needs dom->blocks with csr_offset set; written once per domain. */
int pp_csr_pattern(const pp_domain *dom, int32_t *row, int32_t *col, void *stream);

/* non_ortho_flags of the setup entry points below: the bits of csrc/phipict/piso/
 * non_ortho_flags.h. NON_ORTHO_DIRECT_MATRIX (1) and NON_ORTHO_CENTER_MATRIX (16) add the
 * corner terms to the matrices (nonortho_laplace.cuh:53-131, advection.cu:177-240,
 * pressure.cu:104-160); NON_ORTHO_DIRECT_RHS (2) and NON_ORTHO_DIAGONAL_RHS (8) add the explicit
 * corner terms of velocity_result / scalar_result / pressure_result to the right-hand sides
 * (nonortho_rhs.cuh:28-200), which then must be set; NON_ORTHO_DIAGONAL_MATRIX (4) is ignored,
 * as in the reference. pp_setup_epot_matrix takes the matrix bits (potential.cu:103-150); its
 * corner-term face coefficients weigh both cells with 1/A of the domain's advection diagonal
 * (first environment), as the reference's interpolation with withA set, so A must be set. */

/* replaces SetupAdvectionMatrixEulerImplicit (csrc/phipict/piso/advection.cu:1230-1250),
 * kernel PISO_build_matrix advection.cu:26-306: writes C (row, col, values) and A. */
int pp_setup_advection_matrix(const pp_domain *dom, int32_t non_ortho_flags, void *stream);
/* SetupAdvectionMatrix with forPassiveScalar (advection.cu:26-306): the matrix of passive
 * scalar channel `channel` (the channel's diffusivity and its boundary conditions) into C, A */
int pp_setup_advection_matrix_scalar(const pp_domain *dom, int32_t non_ortho_flags, int32_t channel, void *stream);
/* replaces SetupAdvectionScalar (advection.cu:544-700, 1301-1373): scalar_rhs [B][C][n] */
int pp_setup_advection_scalar(const pp_domain *dom, int32_t non_ortho_flags, void *stream);
/* replaces SetupAdvectionVelocityEulerImplicitRHS (advection.cu:1375-1406),
 * kernel kPISO_build_advection_RHS advection.cu:799-955: writes velocity_rhs. */
int pp_setup_advection_velocity(const pp_domain *dom, int32_t non_ortho_flags,
                                int32_t apply_pressure_gradient, void *stream);
/* replaces SetupPressureMatrix (csrc/phipict/piso/pressure.cu:947-975),
 * kernel PISO_build_pressure_matrix pressure.cu:25-193: writes P. */
int pp_setup_pressure_matrix(const pp_domain *dom, int32_t non_ortho_flags, void *stream);
/* replaces SetupPressureRHS (pressure.cu:978-1027): kernels PISO_build_pressure_rhs
 * pressure.cu:349-522 and k_computePressureRHSdivergenceFromFlux :755-803. */
int pp_setup_pressure_rhs(const pp_domain *dom, int32_t non_ortho_flags, int32_t time_step_norm,
                          void *stream);
/* replaces SetupPressureRHSdiv (pressure.cu:1030-1070): divergence only; with the flags'
 * right-hand-side bits plus k_pressureRHSaddNonOrthoComponents (pressure.cu:838-860), which
 * pp_setup_pressure_rhs applies too. */
int pp_setup_pressure_rhs_div(const pp_domain *dom, int32_t non_ortho_flags, int32_t time_step_norm,
                              void *stream);
/* replaces CorrectVelocity version 1 (csrc/phipict/piso/velocity_correction.cu:216-271),
 * kernel PISO_update_velocity velocity_correction.cu:26-60. */
int pp_correct_velocity(const pp_domain *dom, int32_t time_step_norm, void *stream);
/* replaces SetupEpotMatrix (csrc/phipict/mhd/potential.cu:250-262),
 * kernel PISO_build_epot_matrix potential.cu:29-233: writes E. */
int pp_setup_epot_matrix(const pp_domain *dom, int32_t non_ortho_flags, void *stream);
/* replaces ComputeEpotRHS (potential.cu:325-370), kernel k_computeEpotRHS potential.cu:272-322:
 * vector_field [batch][d][n] -> divergence [batch][n]. */
int pp_compute_epot_rhs(const pp_domain *dom, const void *vector_field, void *divergence, void *stream);
/* replaces ComputeCurrentDensityFaceBased (potential.cu:918-960), kernel
 * k_computeCurrentDensityFaceBased potential.cu:888-916: phi [batch][n],
 * u x B [batch][3][n] (2D: 3 components) -> J [batch][3][n]. */
int pp_compute_current_density(const pp_domain *dom, const void *phi, const void *uxb, void *j_out,
                               void *stream);
/* replaces Copy{Velocity,Pressure,Epot}Result{To,From}Blocks (csrc/phipict/piso/result_copy.cu:19-475).
 * field: 0 velocity, 1 pressure, 2 epot, 3 passive scalar; to_blocks: 1 flat->blocks, 0 blocks->flat. */
int pp_copy_result(const pp_domain *dom, int32_t field, int32_t to_blocks, void *stream);
/* ---- adjoints of the assembly (orthogonal CENTRAL path; the reference's *Grad bindings,
 * csrc/phipict/bindings.cpp:781-814). Gradients are ADDED to the output buffers (zero them
 * first): flat [batch][channels][n] by global cell index; grad_bound (device array of
 * num_blocks*6 pointers, entry block*6+face, null where unused) points at [batch][d][face
 * cells] gradients of the fixed faces' velocities; grad_source (or null) is per cell
 * [batch][d][n]. The descriptor must show the forward's inputs (block velocity, pressure,
 * boundary data, time step); values the forward read from the domain's flat vectors are
 * passed explicitly. grad_visc (or null) is [batch][n]: the gradient with respect to the
 * viscosity as a per-cell field (the reference's viscosityGrad and block viscosityGrad), which
 * the caller sums over the cells that use the domain's or a block's viscosity. */
/* replaces SetupAdvectionMatrixGrad (advection.cu:309-542): d(C values [B][nnz], A [B][n] or
 * null) -> d(block velocity), d(viscosity) */
int pp_setup_advection_matrix_grad(const pp_domain *dom, const void *grad_c, const void *grad_a, void *grad_u,
                                   void *grad_visc, void *stream);
/* replaces SetupAdvectionVelocityGrad (advection.cu:957-1200, applyPressureGradient off):
 * d(velocity RHS) -> d(block velocity), d(source), d(fixed-face velocities), d(viscosity) */
int pp_setup_advection_velocity_grad(const pp_domain *dom, const void *grad_rhs, void *grad_u, void *grad_source,
                                     void *const *grad_bound, void *grad_visc, void *stream);
/* replaces SetupPressureMatrixGrad (pressure.cu:195-348): d(P values) -> d(A); a = forward A */
int pp_setup_pressure_matrix_grad(const pp_domain *dom, const void *a, const void *grad_p, void *grad_a,
                                  void *stream);
/* replaces SetupPressureRHSGrad (pressure.cu:524-755): d(HbyA) -> d(A), d(C values),
 * d(velocity result), d(block velocity), d(source), d(fixed-face velocities), d(viscosity); a,
 * c_val, velocity_result and hbya are the forward's inputs and output */
int pp_setup_pressure_rhs_grad(const pp_domain *dom, const void *grad_hbya, const void *a, const void *c_val,
                               const void *velocity_result, const void *hbya, void *grad_a, void *grad_c,
                               void *grad_velocity_result, void *grad_u, void *grad_source, void *const *grad_bound,
                               void *grad_visc, void *stream);
/* replaces SetupPressureRHSdivGrad (pressure.cu:805-880): d(div) -> d(HbyA), d(fixed-face
 * velocities) */
int pp_setup_pressure_rhs_div_grad(const pp_domain *dom, const void *grad_div, int32_t time_step_norm, void *grad_hbya,
                                   void *const *grad_bound, void *stream);
/* replaces CorrectVelocityGrad (velocity_correction.cu:300-375, version 1): d(corrected
 * velocity) -> d(HbyA), d(A), d(pressure [B][n]); the blocks must hold the forward's pressure */
int pp_correct_velocity_grad(const pp_domain *dom, const void *grad_out, const void *a, int32_t time_step_norm,
                             void *grad_hbya, void *grad_a, void *grad_pressure, void *stream);
/* replaces ComputeEpotRHSGrad (potential.cu:980-1040): d(div [B][n]) -> d(in-plane field [B][d][n])
 * and, with grad_values (a device table, entry block * 6 + face, of [B][face cells] buffers or
 * null entries; the table itself may be null), d(Dirichlet potential values) of the faces
 * (the reference's potentialValuesGrad, potential.cu:1000-1040) */
int pp_compute_epot_rhs_grad(const pp_domain *dom, const void *grad_div, void *grad_field, void *const *grad_values,
                             void *stream);
/* replaces ComputeCurrentDensityFaceBasedGrad (potential.cu:1248-1329): d(J [B][3][n]) ->
 * d(phi [B][n]), d(u x e_b [B][3][n]), d(Dirichlet potential values) as above */
int pp_compute_current_density_grad(const pp_domain *dom, const void *grad_j, void *grad_phi, void *grad_uxb,
                                    void *const *grad_values, void *stream);
/* replaces Domain::getMaxVelocity / getMaxVelocityPerEnv / getMaxVelocityMagnitude
 * (csrc/phipict/domain/domain.cpp:212-245) and the Block methods of the same names
 * (csrc/phipict/domain/block.cpp:505-600): block -1 = every block, else that block only;
 * with_bounds adds the face cells of the fixed faces; computational measures Minv u (cells
 * of blocks with transforms, faces with face transforms); norm 0 = largest absolute
 * component, 1 = Euclidean magnitude. out[batch + 1] (dtype): out[b] for environment b,
 * out[batch] = the maximum over all environments; a NaN velocity gives NaN. */
int pp_max_velocity(const pp_domain *dom, int32_t block, int32_t with_bounds, int32_t computational, int32_t norm,
                    void *out, void *stream);

/* ---- linear solvers ---- */

/* replaces SolveLinear (csrc/phipict/solvers/linear_solve.cu:131-164) with the fused
 * solvers krylov::cgSolve / krylov::bicgstabSolve (csrc/phipict/solvers/krylov/krylov.cu:1058-1186).
 * rhs, x: [nrhs][n] device; RHS k uses matrix k / (nrhs / A->num_matrices).
 * tols: host array of 1 or nrhs absolute tolerances. matrix_rank_deficient: the operator
 * has the constant nullspace; the solve keeps the residual mean-free (projection).
 * results: host array [nrhs]. */
int pp_solve_linear(const pp_csr *A, const void *rhs, void *x, int32_t nrhs, int32_t dtype,
                    int32_t maxit, const double *tols, int32_t ntols, int32_t criterion,
                    int32_t use_bicg, int32_t matrix_rank_deficient, int32_t residual_reset_steps,
                    int32_t transpose_a, int32_t return_best, void *stream, pp_solve_result *results);

/* replaces AMGPCGSolve (csrc/phipict/bindings.cpp:764-776 -> krylov::amgPcgSolveTorch,
 * krylov.cu:1565-1612): AMG-preconditioned CG with one V-cycle per iteration.
 * coarse_pinv: [nc][nc] or [nrhs][nc][nc] (coarse_batched) in hdtype. */
int pp_amg_pcg_solve(const pp_csr *A, const pp_amg_level *levels, int32_t num_levels,
                     const void *coarse_pinv, int32_t nc, int32_t coarse_batched, int32_t hdtype,
                     const void *rhs, void *x, int32_t nrhs, int32_t dtype, int32_t maxit,
                     double tol, int32_t normalized, int32_t return_best, int32_t project_constant,
                     int32_t pre_sweeps, int32_t post_sweeps, double omega,
                     const double *tols, int32_t ntols, void *stream, pp_solve_result *results);

/* Segmented solve (no reference counterpart; the uneven-size batching form): nseg independent
 * systems of different sizes, the diagonal blocks of one CSR matrix A (instance s owns rows
 * seg_rows[s] .. seg_rows[s+1]; its rows reference only its own columns), with ncomp
 * right-hand sides each (rhs, x: [ncomp][n]; the velocity components of every instance share
 * its block). CG or BiCGStab with the reference's recurrences and stopping rules per instance
 * and component (NORM2_NORMALIZED over the instance's rows), one launch per phase for all;
 * project: each instance's right-hand side made mean-free first (rank-deficient CG).
 * tols: 1 or ncomp * nseg values (component-major); tile_rows: rows per workgroup tile, a
 * multiple of 256 (0: 1024). results: ncomp * nseg entries, component-major. */
int pp_solve_segmented(const pp_csr *A, const int32_t *seg_rows, int32_t nseg, int32_t ncomp, const void *rhs, void *x,
                       int32_t dtype, int32_t maxit, const double *tols, int32_t ntols, int32_t use_bicg,
                       int32_t project, int32_t tile_rows, void *stream, pp_solve_result *results);

/* No counterpart in the reference: pp_solve_linear for callers that cannot allocate (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). Same
 * arguments, results and kernels as pp_solve_linear, but every work buffer (vectors, partial
 * sums, solve records, tolerances) comes from one static device array of 63.1 MiB per GPU (56
 * MiB of work vectors: e.g. seven vectors of 1.0 M doubles), the host snapshots from a static
 * host array (at most 2048 right-hand sides) and the snapshot events are created at the first
 * call on a device; it calls no hipMalloc, hipFree or other allocating routine. Larger solves
 * and transposed BiCGStab solves are refused with an error. Calls on one GPU share the array:
 * calls that overlap on one GPU (from other streams or threads) race. */
int pp_solve_linear_noalloc(const pp_csr *A, const void *rhs, void *x, int32_t nrhs, int32_t dtype,
                            int32_t maxit, const double *tols, int32_t ntols, int32_t criterion,
                            int32_t use_bicg, int32_t matrix_rank_deficient, int32_t residual_reset_steps,
                            int32_t transpose_a, int32_t return_best, void *stream, pp_solve_result *results);

/* replaces ReleaseKrylovWorkspaces (bindings.cpp:777, krylov.cu:921-924). */
int pp_release_workspaces(void);

/* replace SetKrylovSettings / GetKrylovSettings (bindings.cpp:739-763, krylov.h:68-74):
 * check_interval = iterations per launch chunk at most (the host's pacing of the multi-
 * workgroup engines); use_graphs: replay captured chunks of iterations from HIP graphs (the
 * segmented engine; the reference replays its iteration chunks from CUDA graphs,
 * krylov.cu:984-1042); cg_jacobi / bicg_jacobi select the Jacobi preconditioner of the fused
 * CG (z = diag(A)^-1 r, krylov.cu:414-425) and of the fused BiCGStab (right preconditioning,
 * krylov.cu:243-247) that pp_solve_linear runs. */
int pp_set_krylov_settings(int32_t check_interval, int32_t use_graphs, int32_t cg_jacobi, int32_t bicg_jacobi);
int pp_get_krylov_settings(int32_t *check_interval, int32_t *use_graphs, int32_t *cg_jacobi, int32_t *bicg_jacobi);

/* replaces SparseOuterProduct (linear_solve.cu:166-190, SDDMM legacy_cg.cu:888-927):
 * out.val[k] = a[row(k)] * b[col(k)] on the pattern of out. */
int pp_sparse_outer_product(const void *a, const void *b, const pp_csr *out, int32_t dtype, void *stream);

/* replaces krylov::csrTranspose (krylov.cu:1620-1652): pattern and values of A^T into
 * t_row[n+1], t_col[nnz], t_val[num_matrices][nnz] (caller-allocated, device). */
int pp_csr_transpose(const pp_csr *A, int32_t dtype, int32_t *t_row, int32_t *t_col, void *t_val,
                     void *stream);

/* Sparse matrix-vector product y = A x for nrhs right-hand sides (x, y: [nrhs][n]). */
int pp_spmv(const pp_csr *A, const void *x, void *y, int32_t nrhs, int32_t dtype, void *stream);

/* ---- AMG setup (host) ----
 * replaces build_amg_hierarchy (src/phipict/solvers/amg.py:450-598), whose classical
 * Ruge-Stueben coarsening comes from pyamg (ruge_stuben_solver, strength
 * ('classical', theta)). Host CSR in double precision. Returns a handle. */
int pp_amg_setup(int32_t n, const int32_t *row, const int32_t *col, const double *val,
                 double theta, int32_t max_levels, int32_t max_coarse, void **handle);
/* number of levels including the coarsest; sizes of level l (l < levels-1 has R, P). */
int pp_amg_num_levels(void *handle);
int pp_amg_level_sizes(void *handle, int32_t level, int32_t *n, int32_t *nnz_a, int32_t *nnz_p,
                       int32_t *nnz_r);
/* copy level l into caller buffers (host); any pointer may be NULL. */
int pp_amg_level_copy(void *handle, int32_t level, int32_t *a_row, int32_t *a_col, double *a_val,
                      int32_t *p_row, int32_t *p_col, double *p_val, int32_t *r_row, int32_t *r_col,
                      double *r_val, int32_t *splitting);
int pp_amg_free(void *handle);

/* Galerkin refresh on the device (replaces the products of galerkin_refresh,
 * src/phipict/solvers/amg.py:749-847, and the product map of _galerkin_map, :1114-1180):
 * C = R (A P) as two row products, first A P into the pattern (ap_row, ap_col) of
 * pp_ap_pattern (ap_val: a double workspace of A->num_matrices x ap_nnz values, overwritten),
 * then R (A P) into the coarse pattern (c_row, c_col) with c_n rows; every entry accumulated in
 * double in a fixed order. R and P values are double; A holds A->num_matrices value arrays
 * (a_dtype); C gets one value array per A matrix (c_dtype). max_rows (host): the longest rows
 * of A, P, A P, R and C. plan (host, [2], or NULL for the built-in rule): the kernel of each
 * product, PP_RAP_* (a kernel that does not fit the rows is refused); stages: bit 0 runs A P,
 * bit 1 R (A P) (the autotuner times them one at a time). No memory beyond the arguments. */
enum { PP_RAP_AUTO = 0, PP_RAP_REG16, PP_RAP_REG32, PP_RAP_ROWS8, PP_RAP_ROWS16, PP_RAP_ROWS32, PP_RAP_ROWS64,
       PP_RAP_LONG };
int pp_galerkin_rows(const pp_csr *A, int32_t a_dtype, const int32_t *p_row, const int32_t *p_col, const double *p_val,
                     const int32_t *ap_row, const int32_t *ap_col, int32_t ap_nnz, double *ap_val,
                     const int32_t *r_row, const int32_t *r_col, const double *r_val, const int32_t *c_row,
                     const int32_t *c_col, void *c_val, int32_t c_n, int32_t c_nnz, int32_t c_dtype,
                     const int32_t *max_rows, const int32_t *plan, int32_t stages, void *stream);
/* ELL copy of a CSR matrix (the layout of pp_csr.ell_col / ell_val): entry k of row i at
 * [k][i], rows padded to `width` entries with the row's own column and zero values. Either
 * output may be NULL; no reference counterpart. */
int pp_ell_from_csr(const int32_t *row, const int32_t *col, const void *val, int32_t n, int32_t nnz, int32_t nmat,
                    int32_t width, int32_t dtype, int32_t *ell_col, void *ell_val, int32_t ell_dtype, void *stream);
/* Fill a sliced ELL copy (pp_sell layout with `lanes` lanes per row) of the CSR matrix (row,
 * col, val) with n rows and nnz entries, whose slice offsets and steps the caller computed from
 * the row lengths: columns (sell_col, may be null) and the values of nmat operators (stride nnz
 * in, total out) converted from dtype to sell_dtype (sell_val, may be null). */
int pp_sell_fill(const int32_t *row, const int32_t *col, const void *val, int32_t n, int32_t nnz, int32_t nmat,
                 int32_t lanes, const int32_t *slice_off, const int32_t *slice_steps, int64_t total, int32_t dtype,
                 int32_t *sell_col, void *sell_val, int32_t sell_dtype, void *stream);

/* Autotuner probe (no reference counterpart): mean duration in microseconds of y = M x through
 * the sliced copy M (rows x cols, values in hdtype), over reps launches after a warm-up. */
int pp_sell_probe(const pp_sell *M, int32_t rows, int32_t cols, int32_t hdtype, int32_t reps, void *stream, double *us);

/* Symbolic pattern of A P for the refresh (host; the pattern side of _galerkin_map,
 * src/phipict/solvers/amg.py:1114-1180): fine pattern (a_row, a_col) with n rows, interpolation
 * pattern (p_row, p_col) with nc columns. Row i holds the union of the interpolation rows of row
 * i's columns, sorted. First call with ap_col NULL: writes ap_row[n+1]; second call with ap_col
 * of ap_row[n] entries: writes the columns and, when R's (r_row, r_col) and the coarse pattern
 * (c_row, c_col) are given, checks that the coarse pattern holds every entry of R (A P). */
int pp_ap_pattern(int32_t n, const int32_t *a_row, const int32_t *a_col, int32_t nc, const int32_t *p_row,
                  const int32_t *p_col, const int32_t *r_row, const int32_t *r_col, const int32_t *c_row,
                  const int32_t *c_col, int32_t *ap_row, int32_t *ap_col);

/* Inverse of the dense coarsest AMG operator on the device, for a refresh that should not wait
 * for the host (replaces np.linalg.pinv of the coarsest operator, src/phipict/solvers/amg.py:
 * 583-587 and 1354-1362, where the operator allows it): C and out are batch row-major n x n
 * double matrices, n <= 128, one workgroup each, Gauss-Jordan with partial pivoting. rank_one
 * = 0: out = inv(C). rank_one = 1: out = inv(C + (s/n) 1 1^T) - 1/(s n) 1 1^T with
 * s = trace(C)/n, which is pinv(C) when C is symmetric with the constant vector spanning its
 * null space. No memory beyond the arguments. */
int pp_coarse_inverse(const double *C, double *out, int32_t n, int32_t batch, int32_t rank_one, void *stream);
/* inverse diagonal of a CSR matrix (rows without a usable diagonal get 1):
 * dinv [num_matrices][n] in dtype_out from values in dtype. */
int pp_inverse_diagonal(const pp_csr *A, int32_t dtype, void *dinv, int32_t dtype_out, void *stream);
/* values cast: out[i] = (dtype_out) in[i] */
int pp_cast(const void *in, int32_t dtype_in, void *out, int32_t dtype_out, int64_t count, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* PHIPICT_HIP_H */

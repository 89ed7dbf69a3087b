/*
 * ORACLE — test infrastructure only. A plain-C, single-threaded restatement of the
 * phiPICT reference algorithms on the PISO hot path (per-cell assembly kernels, fused
 * CG/BiCGStab recurrences, Ruge-Stueben AMG and its V-cycle PCG). Only tests/,
 * __graft_entry__.smoke() and bench.py's CPU baseline may load it, and only as the
 * checker; the product path never calls it.
 *
 * Every function cites the reference file:line it restates (paths relative to the
 * phiPICT source tree). Descriptors here are the oracle's own (host pointers, one
 * environment), independent of the product's include/phipict_hip.h.
 *
 * The same sources also build a multi-threaded copy for timing the CPU baseline
 * (_build/liboracle_omp.so, compiled with -fopenmp): the cell and row loops below are
 * marked O_PAR / O_PAR_SUM, which share them among threads there and expand to nothing in
 * the checker build. Threaded sums add in a thread-dependent order, so only the
 * single-threaded build checks results.
 */
#ifndef PHIPICT_ORACLE_H
#define PHIPICT_ORACLE_H
#include <stdint.h>

#ifdef _OPENMP
#define O_PRAGMA(x) _Pragma(#x)
#define O_PAR O_PRAGMA(omp parallel for schedule(static))
#define O_PAR_SUM(...) O_PRAGMA(omp parallel for schedule(static) reduction(+ : __VA_ARGS__))
#define O_PAR_OR(v) O_PRAGMA(omp parallel for schedule(static) reduction(| : v))
#else
#define O_PAR
#define O_PAR_SUM(...)
#define O_PAR_OR(v)
#endif

enum { OB_CONNECTED = 20, OB_PERIODIC = 21, OB_FIXED = 30 };
enum { OB_POT_INSULATING = 0, OB_POT_OPEN = 1, OB_POT_DIRICHLET = 2, OB_POT_THIN_WALL = 3 };

typedef struct ob_bound {
    int32_t type;          /* OB_* */
    int32_t vel_neumann;   /* FIXED: velocity BC type is NEUMANN */
    int32_t vel_static;    /* FIXED: velocity [d] (1) or [d][face cells] (0) */
    int32_t has_transform; /* FIXED: face transform [face cells][2d^2+1] */
    int32_t pot_type;      /* FIXED: potential BC (when pot_types is NULL) */
    int32_t conn_block;    /* CONNECTED */
    int32_t conn_axes[3];  /* CONNECTED: face-format axes */
    int32_t pad;
    const void *vel;
    const void *transform;
    const int8_t *pot_types;
    const void *pot_values;
    double pot_cw;
} ob_bound;

typedef struct ob_block {
    int32_t size[3];
    int32_t global_offset;
    int32_t has_transform;
    int32_t vsource_mode;  /* 0 none, 1 [d], 2 [d][cells] */
    int32_t visc_mode;     /* 0 domain, 1 [1], 2 [cells] */
    int32_t pad;
    const void *transform; /* [cells][2d^2+1] */
    const void *velocity;  /* [d][cells] */
    const void *pressure;  /* [cells] */
    const void *vsource;
    const void *viscosity;
    ob_bound b[6];
} ob_block;

/* convective schemes (AdvectionScheme, domain_structs_gpu.h:65-68) */
enum { OB_ADV_CENTRAL = 0, OB_ADV_LINEAR_UPWIND = 1 };

typedef struct ob_domain {
    int32_t dims;
    int32_t nblocks;
    int32_t ncells;
    int32_t scheme;        /* OB_ADV_* */
    double viscosity;
    double dt;
    const ob_block *blocks;
} ob_domain;

/* non-orthogonal flag bits (non_ortho_flags.h) and what a corner term reads: no values
 * (matrices), velocity result component w, the pressure result, or passive-scalar result
 * channel w */
enum { ONO_DIRECT_MATRIX = 1, ONO_DIRECT_RHS = 2, ONO_DIAGONAL_RHS = 8, ONO_CENTER_MATRIX = 16 };
enum { OK_TYPES = 0, OK_VEL = 1, OK_PRES = 2, OK_SCAL = 3 };

/* Passive-scalar data of one fixed face for one channel (non-orthogonal right-hand sides). */
typedef struct ob_face_scalar {
    const void *values;    /* [1] (is_static) or [face cells]; NULL: zero */
    int32_t is_static;
    int32_t neumann;       /* the channel's condition on this face is Neumann */
} ob_face_scalar;

typedef struct ob_result {
    double final_residual;
    int32_t used_iterations;
    int32_t converged;
    int32_t is_finite;
    int32_t pad;
} ob_result;

#endif

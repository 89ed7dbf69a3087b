// Shared declarations of the Krylov solvers: matrix views, the per-right-hand-side solve
// record, wavefront / workgroup reductions and the small-system engine's launcher.
#pragma once

#include "pp_common.hpp"

namespace pp {

// How a right-hand side's solve ended (0 while it runs). The numbers are reported to the
// host in pp_solve_result.pad.
enum Fate : int32_t { F_ACTIVE = 0, F_MET = 1, F_NONFINITE = 2, F_EXHAUSTED = 3, F_STALLED = 4 };

// Record flags.
enum RecFlag : int32_t {
    R_FRESH = 1, // the next direction is the (preconditioned) residual itself
    R_HALF = 2,  // BiCGStab met the tolerance at the half step
    R_KEEP = 4,  // the current iterate is the best so far: save it before it changes
    R_RESET = 8, // CG: the residual was recomputed as b - A x; its (r, r) is in the reset partials
};

// Return-best solves stop after this many consecutive iterations without a falling residual.
constexpr int RISE_LIMIT = 100;

// Per right-hand side state of a running solve, kept on the device; the host reads copies of
// it (snapshots) to schedule launches and to report results.
struct SolveRec {
    double tol;
    double rho;   // inner product of the recurrence: (r, r) CG, (r, z) PCG, (r0, r) BiCGStab
    double alpha; // BiCGStab: the last step lengths, needed by the next direction
    double omega;
    double crit;  // residual criterion of the current iterate
    double best;  // return-best: smallest criterion so far
    double prev;  // criterion one iteration earlier (rise detection)
    double half;  // BiCGStab: criterion of the half-step residual s
    int32_t iter, maxit;
    int32_t report; // usedIterations to report
    int32_t bestAt; // iteration of the best iterate, -1 if none
    int32_t climb;  // consecutive iterations whose residual did not fall
    int32_t fate;
    int32_t flags;
    int32_t zero; // the right-hand side is exactly zero
};

template <typename T>
struct Csr {
    const int32_t *__restrict__ row;
    const int32_t *__restrict__ col;
    const T *__restrict__ val;
    int n;
    int valStride; // 0: matrix shared by all right-hand sides
    int rpm;       // right-hand sides per matrix
    int ell_w;     // > 0: ELL copy of the same entries (entry k of row i at [k][i])
    const int32_t *__restrict__ ell_col;
    const T *__restrict__ ell_val;
};

// Matrix view of a descriptor; the ELL copy is used when present.
template <typename T>
inline Csr<T> csr_view(const pp_csr *A, int rpm) {
    Csr<T> C{A->row, A->col, reinterpret_cast<const T *>(A->val), A->n, A->num_matrices > 1 ? A->nnz : 0, rpm, 0,
             nullptr, nullptr};
    if (A->ell_width > 0 && A->ell_col && A->ell_val) {
        C.ell_w = A->ell_width;
        C.ell_col = A->ell_col;
        C.ell_val = reinterpret_cast<const T *>(A->ell_val);
    }
    return C;
}

// ---------------- reductions (wave64 shuffles + LDS tree) ----------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}

// Sum vals[NV] over the workgroup; result valid in thread 0. sm: >= NV * 16 doubles.
template <int NV>
__device__ __forceinline__ void block_sum(double (&vals)[NV], double *sm) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
    for (int q = 0; q < NV; ++q) {
        double v = wave_sum(vals[q]);
        if (lane == 0) sm[q * 16 + wid] = v;
    }
    __syncthreads();
    if (wid == 0) {
#pragma unroll
        for (int q = 0; q < NV; ++q) {
            double v = lane < nw ? sm[q * 16 + lane] : 0.0;
            vals[q] = wave_sum(v);
        }
    }
    __syncthreads();
}

// ---------------- stopping rules ----------------
// The decisions of the reference's solvers (krylov.cu:252-268, 354-366, 485-498, 531-568,
// 583-593, 650-670, 730-750, 785-824): which right-hand sides stop, why, and the iteration
// count reported. Both engines apply them to a register copy of a record.
__host__ __device__ __forceinline__ bool rec_active(const SolveRec &s) { return s.fate == F_ACTIVE; }

// Initial residual r = b - A x with (r, r) = rr: CG stops only on an exactly zero residual,
// BiCGStab on the tolerance (reporting -1 iterations).
__device__ __forceinline__ void rule_open(SolveRec &s, double rr, double scale, bool bicg) {
    s.rho = rr;
    s.flags |= R_FRESH;
    s.crit = sqrt(rr) * scale;
    if (!isfinite(s.crit)) { s.fate = F_NONFINITE; s.report = 0; }
    else if (bicg ? s.crit < s.tol : s.crit == 0.0) { s.fate = F_MET; s.report = bicg ? -1 : 0; }
}

// CG / PCG: the residual (r, r) = rr of the iterate just formed. Counts the iteration, keeps
// the return-best bookkeeping and decides met / exhausted / stalled.
__device__ __forceinline__ void rule_step(SolveRec &s, double rr, double scale, bool keepBest) {
    s.iter += 1;
    s.crit = sqrt(rr) * scale;
    s.flags &= ~R_KEEP;
    s.report = s.iter - 1;
    if (!isfinite(s.crit)) { s.fate = F_NONFINITE; return; }
    if (keepBest) {
        if (s.crit < s.best) { s.best = s.crit; s.bestAt = s.iter - 1; s.flags |= R_KEEP; }
        s.climb = (s.iter > 1 && s.crit >= s.prev) ? s.climb + 1 : 0;
        s.prev = s.crit;
    }
    if (s.crit < s.tol) s.fate = F_MET;
    else if (s.iter >= s.maxit) s.fate = F_EXHAUSTED;
    else if (keepBest && s.climb >= RISE_LIMIT) s.fate = F_STALLED;
}

// Unpreconditioned CG after rule_step: beta = rr / rho, rho = rr (returns beta).
__device__ __forceinline__ double rule_cg_beta(SolveRec &s, double rr) {
    const double beta = rr / s.rho;
    s.rho = rr;
    if (!isfinite(beta)) s.fate = F_NONFINITE;
    return beta;
}

// Step length rho / denominator; a non-finite one ends the solve (reporting the iterations
// done so far, the criterion unchanged).
__device__ __forceinline__ double rule_alpha(SolveRec &s, double denom) {
    const double alpha = s.rho / denom;
    if (!isfinite(alpha)) { s.fate = F_NONFINITE; s.report = s.iter; }
    return alpha;
}

// PCG: rho = (r, z) of the new preconditioned residual; returns beta (0 for a fresh
// direction).
__device__ __forceinline__ double rule_pcg_rho(SolveRec &s, double rz) {
    const double beta = (s.flags & R_FRESH) ? 0.0 : rz / s.rho;
    s.rho = rz;
    if (!isfinite(rz) || !isfinite(beta)) { s.fate = F_NONFINITE; s.report = s.iter; }
    return beta;
}

// BiCGStab half step: ts = (t, s), tt = (t, t), ss = (s, s). Returns omega; a half-step
// residual below the tolerance sets R_HALF (omega 0).
__device__ __forceinline__ double rule_bicg_half(SolveRec &s, double ts, double tt, double ss, double scale) {
    const double cs = sqrt(ss) * scale;
    s.half = cs;
    if (!isfinite(cs)) { s.fate = F_NONFINITE; s.crit = cs; s.report = s.iter; return 0.0; }
    if (cs < s.tol) { s.flags |= R_HALF; return 0.0; }
    const double omega = ts / tt;
    if (!isfinite(omega)) { s.fate = F_NONFINITE; s.crit = cs; s.report = s.iter; }
    return omega;
}

// BiCGStab after the half step met the tolerance.
__device__ __forceinline__ void rule_bicg_half_met(SolveRec &s) {
    s.fate = F_MET;
    s.crit = s.half;
    s.report = s.iter;
}

// BiCGStab full step: r0r = (r0, r), rr = (r, r) of the new residual; s.alpha and s.omega
// hold the step's lengths. Returns beta for the next direction.
__device__ __forceinline__ double rule_bicg_full(SolveRec &s, double r0r, double rr, double scale) {
    s.crit = sqrt(rr) * scale;
    if (!isfinite(s.crit)) { s.fate = F_NONFINITE; s.report = s.iter; return 0.0; }
    if (s.crit < s.tol) { s.fate = F_MET; s.report = s.iter + 1; return 0.0; }
    s.iter += 1;
    s.report = s.iter - 1;
    const double beta = (r0r / s.rho) * (s.alpha / s.omega);
    s.rho = r0r;
    if (!isfinite(beta)) s.fate = F_NONFINITE;
    else if (s.iter >= s.maxit) s.fate = F_EXHAUSTED;
    return beta;
}

// A fresh record (kInitState, krylov.cu:252-268).
__host__ __device__ __forceinline__ SolveRec rec_init(double tol, int maxit) {
    SolveRec s{};
    s.tol = tol;
    s.maxit = maxit;
    s.best = INFINITY;
    s.bestAt = -1;
    s.flags = R_FRESH;
    return s;
}

// ---------------- small systems: one workgroup per group of right-hand sides ----------------
constexpr int SMALL_BLOCK = 1024;
constexpr int SMALL_N = 1 << 15;
constexpr int SMALL_SLOTS = 9; // work vectors per right-hand side

struct SmallArgs {
    int n, B, nb_per_group, rpm, maxit, normalized, returnBest, project, resetEvery, bicg;
    double normScale;
    int dinvStride; // Jacobi preconditioner: n per matrix of a batch, 0 for a shared matrix
};

// Launch the on-chip engine: `groups` workgroups, records rec[B] initialised by the caller,
// work vectors (9 per right-hand side, n each) at work; dinv (may be null): inverse diagonals
// of the Jacobi preconditioner.
template <typename T>
void launch_small(const Csr<T> &A, const T *b, T *x, void *work, SolveRec *rec, const T *dinv, const SmallArgs &a,
                  int groups, hipStream_t s);

// Values of A^T on A's (structurally symmetric) pattern into tval; *bad set if a mirror entry
// is missing.
template <typename T>
int launch_transpose_symmetric(const pp_csr *A, T *tval, int *bad, hipStream_t s);

} // namespace pp

// Krylov solvers of the gfx950 library, part 2: the multi-workgroup engine for large systems
// (CG, BiCGStab, AMG-preconditioned CG) and the AMG V-cycle.
//
// Design: replicated reductions. Every iteration phase is one kernel over a grid of G
// workgroups. Each workgroup writes the partial sums of its rows into a partial set
// [rhs][quantity][G]; the NEXT phase kernel starts with every workgroup folding the partials
// it needs itself, in one fixed order, so all workgroups derive bit-identical scalars
// (step lengths, convergence decisions) without a finish kernel, an atomic or a device-scope
// fence. The kernel boundary is the only grid-wide synchronisation: CG runs two launches per
// iteration, BiCGStab four, and the PCG outer loop two besides the V-cycle, whose first
// kernel also takes the convergence decision.
//
// Per-right-hand-side records (SolveRec) come in two copies: a deciding kernel reads copy
// `a` and writes copy `b` from workgroup 0 (no other workgroup of that launch reads `b`), a
// later kernel of the iteration commits back into `a`. The host snapshots copy `b` after each
// deciding kernel.
//
// Recurrences, stopping rules and reported iteration counts restate the reference's fused
// solvers (csrc/phipict/solvers/krylov/krylov.cu: CG :291-568, BiCGStab :598-824, AMG-PCG
// :1389-1546 with the V-cycle :1247-1385); the rules themselves are in pp_krylov.hpp.
#include "pp_krylov.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

namespace pp {

constexpr int WG = 256; // threads per workgroup of every phase and level kernel
// Workgroups per CU of the phase kernels (the width every workgroup folds at the start of the
// next phase) and of the level passes.
#ifndef PP_PHASE_WG_PER_CU
#define PP_PHASE_WG_PER_CU 8
#endif
#ifndef PP_LEVEL_WG_PER_CU
#define PP_LEVEL_WG_PER_CU 8
#endif

// ======================================================================================
// Partial sets and the replicated fold
// ======================================================================================
struct Parts {
    double *p; // [rhs][nq][G]
    int nq, G;
    __device__ __forceinline__ double *at(int rhs, int q) const { return p + ((size_t)rhs * nq + q) * G; }
};

// Write this workgroup's sums acc[k][q] of right-hand sides first..first+count-1.
template <int NB, int NQ>
__device__ __forceinline__ void put_parts(const Parts &P, double (&acc)[NB][NQ], int first, int count, double *lds) {
    double flat[NB * NQ];
#pragma unroll
    for (int k = 0; k < NB; ++k)
#pragma unroll
        for (int q = 0; q < NQ; ++q) flat[k * NQ + q] = acc[k][q];
    block_sum<NB * NQ>(flat, lds);
    if (threadIdx.x == 0)
        for (int k = 0; k < count; ++k)
            for (int q = 0; q < NQ; ++q) P.at(first + k, q)[blockIdx.x] = flat[k * NQ + q];
}

// Totals of quantities q0..q0+NQ-1 of right-hand sides first..first+count-1, folded over
// the G partials in a fixed order and broadcast to every thread: every workgroup of a launch
// obtains the same bits.
template <int NB, int NQ>
__device__ __forceinline__ void fold(const Parts &P, int q0, int first, int count, double (&out)[NB][NQ], double *lds) {
    double acc[NB * NQ];
#pragma unroll
    for (int k = 0; k < NB; ++k)
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            double s = 0.0;
            if (k < count) {
                const double *src = P.at(first + k, q0 + q);
                for (int g = threadIdx.x; g < P.G; g += WG) s += src[g];
            }
            acc[k * NQ + q] = s;
        }
    block_sum<NB * NQ>(acc, lds);
    __shared__ double share[16];
    if (threadIdx.x == 0)
#pragma unroll
        for (int v = 0; v < NB * NQ; ++v) share[v] = acc[v];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NB; ++k)
#pragma unroll
        for (int q = 0; q < NQ; ++q) out[k][q] = share[k * NQ + q];
    __syncthreads();
}

// ======================================================================================
// Row access of the system matrix (shared by NB right-hand sides of one matrix)
// ======================================================================================
// The right-hand sides of grid row y: matrix m, right-hand sides first .. first+count-1.
struct Span { int m, first, count; };
template <int NB>
__device__ __forceinline__ Span span_of(int y, int rpm, int B) {
    const int chunks = (rpm + NB - 1) / NB;
    Span s;
    s.m = y / chunks;
    const int lead = (y % chunks) * NB;
    s.first = s.m * rpm + lead;
    s.count = max(0, min(min(NB, rpm - lead), B - s.first));
    return s;
}

// The entries of row i: with EW > 0 the ELL copy of compile-time width (all loads issued
// before any use), with EW == 0 the ELL copy of runtime width or the CSR arrays.
template <int EW, typename T>
struct RowLoad {
    static constexpr int CAP = EW > 0 ? EW : 1;
    int c[CAP];
    T a[CAP];
};

template <int EW, typename T, typename F>
__device__ __forceinline__ void row_each(const Csr<T> &A, const T *__restrict__ val, const T *__restrict__ ellv, int i, F f) {
    const int n = A.n;
    if constexpr (EW > 0) {
        RowLoad<EW, T> L;
#pragma unroll
        for (int j = 0; j < EW; ++j) {
            L.c[j] = A.ell_col[(size_t)j * n + i];
            L.a[j] = ellv[(size_t)j * n + i];
        }
#pragma unroll
        for (int j = 0; j < EW; ++j) f(L.c[j], L.a[j]);
    } else if (A.ell_w > 0) {
        for (int j = 0; j < A.ell_w; ++j) f(A.ell_col[(size_t)j * n + i], ellv[(size_t)j * n + i]);
    } else {
        const int end = A.row[i + 1];
        for (int e = A.row[i]; e < end; ++e) f(A.col[e], val[e]);
    }
}

// Grid-wide row sweep of a phase kernel: workgroup x takes rows x*WG + t, + G*WG, ...
#define PP_ROWS(i, n) for (int i = blockIdx.x * WG + (int)threadIdx.x; i < (n); i += gridDim.x * WG)

// Records of the span, read into registers.
template <int NB>
__device__ __forceinline__ void load_recs(const SolveRec *recs, const Span &sp, SolveRec (&R)[NB], bool (&act)[NB],
                                          bool &any) {
    any = false;
#pragma unroll
    for (int k = 0; k < NB; ++k) {
        if (k < sp.count) R[k] = recs[sp.first + k];
        act[k] = k < sp.count && rec_active(R[k]);
        any |= act[k];
    }
}

template <int NB>
__device__ __forceinline__ void store_recs(SolveRec *recs, const Span &sp, const SolveRec (&R)[NB]) {
    if (recs && blockIdx.x == 0 && threadIdx.x == 0)
        for (int k = 0; k < sp.count; ++k) recs[sp.first + k] = R[k];
}

// Control of a phase kernel, passed by value.
struct Phase {
    const SolveRec *in; // records read at the start
    SolveRec *out;      // records written by workgroup 0 (null: none)
    int B;              // right-hand sides
    double scale;       // 1 / sqrt(n) (NORM2_NORMALIZED) or 1
    int keepBest;       // return-best bookkeeping
};

// Jacobi preconditioner of the fused CG / BiCGStab (krylov.cu:414-425, 243-247): the inverse
// diagonal of each matrix, d[m * stride + i]; with JAC false it is never read.
template <typename T>
struct Jac {
    const T *d;
    int stride; // n per matrix of a batch, 0 when the matrix is shared
    __device__ __forceinline__ const T *of(int m) const { return d + (size_t)m * stride; }
};

// ======================================================================================
// Opening: partial sums of b, and the initial residual r = b - A x
// ======================================================================================
// Partials of sum(v) per right-hand side (the projection of a rank-deficient b).
template <typename T>
__global__ void __launch_bounds__(WG) k_sum(const T *__restrict__ v, int n, Parts out) {
    __shared__ double lds[16];
    const int b = blockIdx.y;
    double acc[1][1] = {{0.0}};
    PP_ROWS(i, n) acc[0][0] += double(v[(size_t)b * n + i]);
    put_parts<1, 1>(out, acc, b, 1, lds);
}

// r = bw - A x with bw = b - mean(b) when bSum is given (bw written to bwOut for later
// residual resets), r0 = r for BiCGStab. Partials: (r, r), (bw, bw) and sum(r), or with the
// Jacobi preconditioner (r, D^-1 r) in the last place.
template <typename T, int NB, int EW, bool JAC>
__global__ void __launch_bounds__(WG) k_open(Csr<T> A, const T *__restrict__ b, Parts bSum, T *__restrict__ bwOut,
                                             const T *__restrict__ x, T *__restrict__ r, T *__restrict__ r0, int B,
                                             Parts out, Jac<T> jac) {
    __shared__ double lds[16 * NB * 3];
    const int n = A.n;
    const Span sp = span_of<NB>(blockIdx.y, A.rpm, B);
    double mean[NB][1];
    if (bSum.p) {
        fold<NB, 1>(bSum, 0, sp.first, sp.count, mean, lds);
#pragma unroll
        for (int k = 0; k < NB; ++k) mean[k][0] /= double(n);
    }
    const T *val = A.val + (size_t)sp.m * A.valStride;
    const T *ellv = A.ell_val + (A.valStride ? (size_t)sp.m * A.ell_w * n : 0);
    double acc[NB][3];
#pragma unroll
    for (int k = 0; k < NB; ++k) acc[k][0] = acc[k][1] = acc[k][2] = 0.0;
    PP_ROWS(i, n) {
        T ax[NB];
#pragma unroll
        for (int k = 0; k < NB; ++k) ax[k] = T(0);
        row_each<EW>(A, val, ellv, i, [&](int c, T a) {
#pragma unroll
            for (int k = 0; k < NB; ++k)
                if (k < sp.count) ax[k] += a * x[(size_t)(sp.first + k) * n + c];
        });
#pragma unroll
        for (int k = 0; k < NB; ++k)
            if (k < sp.count) {
                const size_t o = (size_t)(sp.first + k) * n + i;
                const T bi = bSum.p ? b[o] - T(mean[k][0]) : b[o];
                if (bwOut) bwOut[o] = bi;
                const T ri = bi - ax[k];
                r[o] = ri;
                if (r0) r0[o] = ri;
                acc[k][0] += double(ri) * double(ri);
                acc[k][1] += double(bi) * double(bi);
                if constexpr (JAC) acc[k][2] += double(ri) * double(ri * jac.of(sp.m)[i]);
                else acc[k][2] += double(ri);
            }
    }
    put_parts<NB, 3>(out, acc, sp.first, sp.count, lds);
}

__global__ void k_init_recs(SolveRec *a, int B, double tol, const double *tols, int maxit) {
    for (int b = threadIdx.x; b < B; b += blockDim.x) a[b] = rec_init(tols ? tols[b] : tol, maxit);
}

// ======================================================================================
// CG: two launches per iteration
// ======================================================================================
enum TurnMode { TURN_OPEN = 0, TURN_STEP = 1, TURN_STEP_RESET = 2, TURN_AFTER_RESET = 3 };

// Decide, then turn: fold (r, r) of the latest residual and apply the rules (TURN_OPEN:
// initial residual, TURN_STEP: after an update), then the new direction p' = r + beta pIn
// (r when fresh), q = A p', partials (p', q). p' is formed at every stencil neighbour as it is
// read, so p' and q come out of one pass.
// TURN_STEP_RESET decides like TURN_STEP and then, instead of a direction, replaces the
// residual by bw - A x (partials of its (r, r) into resetOut, records marked fresh and
// reset); TURN_AFTER_RESET only forms the fresh direction from the records already decided.
// JAC: z = D^-1 r in place of r, rho = (r, z) (the last quantity of rrIn).
template <typename T, int NB, int EW, bool JAC>
__global__ void __launch_bounds__(WG) k_cg_turn(Csr<T> A, T *__restrict__ r, const T *__restrict__ pIn,
                                                T *__restrict__ pOut, T *__restrict__ q, const T *__restrict__ x,
                                                const T *__restrict__ bw, Parts rrIn, Parts pqOut, Parts resetOut,
                                                Phase ph, int mode, Jac<T> jac) {
    __shared__ double lds[16 * NB * 3];
    const int n = A.n;
    const Span sp = span_of<NB>(blockIdx.y, A.rpm, ph.B);
    SolveRec R[NB];
    bool act[NB], any;
    load_recs<NB>(ph.in, sp, R, act, any);
    T beta[NB];
#pragma unroll
    for (int k = 0; k < NB; ++k) beta[k] = T(0);
    if (mode != TURN_AFTER_RESET) {
        double tot[NB][3];
        fold<NB, 3>(rrIn, 0, sp.first, sp.count, tot, lds);
        any = false;
#pragma unroll
        for (int k = 0; k < NB; ++k) {
            if (!act[k]) continue;
            if (mode == TURN_OPEN) {
                rule_open(R[k], tot[k][0], ph.scale, false);
                R[k].zero = tot[k][1] == 0.0;
                if (JAC) R[k].rho = tot[k][2];
            } else {
                rule_step(R[k], tot[k][0], ph.scale, ph.keepBest != 0);
                if (rec_active(R[k])) beta[k] = T(rule_cg_beta(R[k], JAC ? tot[k][2] : tot[k][0]));
            }
            act[k] = rec_active(R[k]);
            any |= act[k];
        }
        if (mode == TURN_STEP_RESET)
#pragma unroll
            for (int k = 0; k < NB; ++k)
                if (act[k]) R[k].flags |= R_FRESH | R_RESET;
        store_recs<NB>(ph.out, sp, R);
    }
    const T *val = A.val + (size_t)sp.m * A.valStride;
    const T *ellv = A.ell_val + (A.valStride ? (size_t)sp.m * A.ell_w * n : 0);
    const T *dv = JAC ? jac.of(sp.m) : nullptr;
    if (mode == TURN_STEP_RESET) {
        double acc[NB][2];
#pragma unroll
        for (int k = 0; k < NB; ++k) acc[k][0] = acc[k][1] = 0.0;
        if (any)
            PP_ROWS(i, n) {
                T ax[NB];
#pragma unroll
                for (int k = 0; k < NB; ++k) ax[k] = T(0);
                row_each<EW>(A, val, ellv, i, [&](int c, T a) {
#pragma unroll
                    for (int k = 0; k < NB; ++k)
                        if (act[k]) ax[k] += a * x[(size_t)(sp.first + k) * n + c];
                });
#pragma unroll
                for (int k = 0; k < NB; ++k)
                    if (act[k]) {
                        const size_t o = (size_t)(sp.first + k) * n + i;
                        const T ri = bw[o] - ax[k];
                        r[o] = ri;
                        acc[k][0] += double(ri) * double(ri);
                        if constexpr (JAC) acc[k][1] += double(ri) * double(ri * dv[i]);
                    }
            }
        put_parts<NB, 2>(resetOut, acc, sp.first, sp.count, lds);
        return;
    }
    bool fresh[NB];
#pragma unroll
    for (int k = 0; k < NB; ++k) fresh[k] = (R[k].flags & R_FRESH) != 0;
    double acc[NB][1];
#pragma unroll
    for (int k = 0; k < NB; ++k) acc[k][0] = 0.0;
    if (any)
        PP_ROWS(i, n) {
            T ap[NB];
#pragma unroll
            for (int k = 0; k < NB; ++k) ap[k] = T(0);
            row_each<EW>(A, val, ellv, i, [&](int c, T a) {
#pragma unroll
                for (int k = 0; k < NB; ++k)
                    if (act[k]) {
                        const size_t o = (size_t)(sp.first + k) * n + c;
                        const T zc = JAC ? r[o] * dv[c] : r[o];
                        ap[k] += a * (fresh[k] ? zc : zc + beta[k] * pIn[o]);
                    }
            });
#pragma unroll
            for (int k = 0; k < NB; ++k)
                if (act[k]) {
                    const size_t o = (size_t)(sp.first + k) * n + i;
                    const T zi = JAC ? r[o] * dv[i] : r[o];
                    const T pi = fresh[k] ? zi : zi + beta[k] * pIn[o];
                    pOut[o] = pi;
                    q[o] = ap[k];
                    acc[k][0] += double(pi) * double(ap[k]);
                }
        }
    put_parts<NB, 1>(pqOut, acc, sp.first, sp.count, lds);
}

// Advance: alpha = rho / (p', q) from the folded partials; x += alpha p', r -= alpha q
// (saving x first when it is the best iterate); partials (r, r), 0 and sum(r), or (r, D^-1 r)
// with JAC. Workgroup 0 commits the records (fresh and reset marks cleared; a reset's rho from
// resetIn).
template <typename T, int NB, bool JAC>
__global__ void __launch_bounds__(WG) k_cg_step(int n, int rpm, T *__restrict__ r, const T *__restrict__ p,
                                                const T *__restrict__ q, T *__restrict__ x, T *__restrict__ xbest,
                                                Parts pqIn, Parts resetIn, Parts rrOut, Phase ph, Jac<T> jac) {
    __shared__ double lds[16 * NB * 3];
    const Span sp = span_of<NB>(blockIdx.y, rpm, ph.B);
    SolveRec R[NB];
    bool act[NB], any;
    load_recs<NB>(ph.in, sp, R, act, any);
    bool reset = false;
#pragma unroll
    for (int k = 0; k < NB; ++k) reset |= act[k] && (R[k].flags & R_RESET);
    if (reset) {
        double rr[NB][2];
        fold<NB, 2>(resetIn, 0, sp.first, sp.count, rr, lds);
#pragma unroll
        for (int k = 0; k < NB; ++k)
            if (act[k] && (R[k].flags & R_RESET)) R[k].rho = JAC ? rr[k][1] : rr[k][0];
    }
    double pq[NB][1];
    fold<NB, 1>(pqIn, 0, sp.first, sp.count, pq, lds);
    T alpha[NB];
    bool keep[NB];
#pragma unroll
    for (int k = 0; k < NB; ++k) {
        alpha[k] = T(0);
        keep[k] = false;
        if (!act[k]) continue;
        alpha[k] = T(rule_alpha(R[k], pq[k][0]));
        R[k].flags &= ~(R_FRESH | R_RESET);
        act[k] = rec_active(R[k]);
        keep[k] = act[k] && xbest && (R[k].flags & R_KEEP);
    }
    store_recs<NB>(ph.out, sp, R);
    double acc[NB][3];
#pragma unroll
    for (int k = 0; k < NB; ++k) acc[k][0] = acc[k][1] = acc[k][2] = 0.0;
#pragma unroll
    for (int k = 0; k < NB; ++k) {
        if (!act[k]) continue;
        const size_t base = (size_t)(sp.first + k) * n;
        PP_ROWS(i, n) {
            const size_t o = base + i;
            const T xi = x[o];
            if (keep[k]) xbest[o] = xi;
            x[o] = xi + alpha[k] * p[o];
            const T ri = r[o] - alpha[k] * q[o];
            r[o] = ri;
            acc[k][0] += double(ri) * double(ri);
            if constexpr (JAC) acc[k][2] += double(ri) * double(ri * jac.of(sp.m)[i]);
            else acc[k][2] += double(ri);
        }
    }
    put_parts<NB, 3>(rrOut, acc, sp.first, sp.count, lds);
}

// ======================================================================================
// BiCGStab: four launches per iteration
// ======================================================================================
// Decide on the latest residual (partials (r0, r), (r, r) of the update, or the opening's
// (r, r), (b, b)), then p = r + beta (p - omega v) (r when fresh) row by row.
template <typename T>
__global__ void __launch_bounds__(WG) k_bicg_dir(int n, const T *__restrict__ r, T *__restrict__ p, const T *__restrict__ v,
                                                 Parts rrIn, Phase ph, int open) {
    __shared__ double lds[16 * 2];
    const int b = blockIdx.y;
    SolveRec R = ph.in[b];
    bool act = rec_active(R);
    double tot[1][2];
    fold<1, 2>(rrIn, 0, b, 1, tot, lds);
    T beta = T(0);
    if (act) {
        if (open) {
            rule_open(R, tot[0][0], ph.scale, true);
            R.zero = tot[0][1] == 0.0;
        } else {
            beta = T(rule_bicg_full(R, tot[0][0], tot[0][1], ph.scale));
        }
        act = rec_active(R);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) ph.out[b] = R;
    if (!act) return;
    const bool fresh = R.flags & R_FRESH;
    const T omega = T(R.omega);
    const size_t base = (size_t)b * n;
    if (fresh) PP_ROWS(i, n) p[base + i] = r[base + i];
    else PP_ROWS(i, n) p[base + i] = r[base + i] + beta * (p[base + i] - omega * v[base + i]);
}

// v = A p (A D^-1 p with JAC); partials (r0, v).
template <typename T, int NB, int EW, bool JAC>
__global__ void __launch_bounds__(WG) k_bicg_apply(Csr<T> A, const T *__restrict__ p, T *__restrict__ v,
                                                   const T *__restrict__ r0, Parts r0vOut, Phase ph, Jac<T> jac) {
    __shared__ double lds[16 * NB];
    const int n = A.n;
    const Span sp = span_of<NB>(blockIdx.y, A.rpm, ph.B);
    SolveRec R[NB];
    bool act[NB], any;
    load_recs<NB>(ph.in, sp, R, act, any);
    const T *val = A.val + (size_t)sp.m * A.valStride;
    const T *ellv = A.ell_val + (A.valStride ? (size_t)sp.m * A.ell_w * n : 0);
    double acc[NB][1];
#pragma unroll
    for (int k = 0; k < NB; ++k) acc[k][0] = 0.0;
    if (any)
        PP_ROWS(i, n) {
            T ap[NB];
#pragma unroll
            for (int k = 0; k < NB; ++k) ap[k] = T(0);
            row_each<EW>(A, val, ellv, i, [&](int c, T a) {
#pragma unroll
                for (int k = 0; k < NB; ++k)
                    if (act[k]) {
                        const T pc = p[(size_t)(sp.first + k) * n + c];
                        ap[k] += a * (JAC ? pc * jac.of(sp.m)[c] : pc);
                    }
            });
#pragma unroll
            for (int k = 0; k < NB; ++k)
                if (act[k]) {
                    const size_t o = (size_t)(sp.first + k) * n + i;
                    v[o] = ap[k];
                    acc[k][0] += double(r0[o]) * double(ap[k]);
                }
        }
    put_parts<NB, 1>(r0vOut, acc, sp.first, sp.count, lds);
}

// alpha = rho / (r0, v) folded; s = r - alpha v (formed at every neighbour as it is read),
// t = A s (A D^-1 s with JAC); partials (t, s), (t, t), (s, s).
template <typename T, int NB, int EW, bool JAC>
__global__ void __launch_bounds__(WG) k_bicg_stab(Csr<T> A, const T *__restrict__ r, const T *__restrict__ v,
                                                  T *__restrict__ t, Parts r0vIn, Parts tsOut, Phase ph, Jac<T> jac) {
    __shared__ double lds[16 * NB * 3];
    const int n = A.n;
    const Span sp = span_of<NB>(blockIdx.y, A.rpm, ph.B);
    SolveRec R[NB];
    bool act[NB], any;
    load_recs<NB>(ph.in, sp, R, act, any);
    double r0v[NB][1];
    fold<NB, 1>(r0vIn, 0, sp.first, sp.count, r0v, lds);
    T alpha[NB];
    any = false;
#pragma unroll
    for (int k = 0; k < NB; ++k) {
        alpha[k] = T(0);
        if (!act[k]) continue;
        alpha[k] = T(rule_alpha(R[k], r0v[k][0]));
        act[k] = rec_active(R[k]);
        any |= act[k];
    }
    const T *val = A.val + (size_t)sp.m * A.valStride;
    const T *ellv = A.ell_val + (A.valStride ? (size_t)sp.m * A.ell_w * n : 0);
    double acc[NB][3];
#pragma unroll
    for (int k = 0; k < NB; ++k) acc[k][0] = acc[k][1] = acc[k][2] = 0.0;
    if (any)
        PP_ROWS(i, n) {
            T as[NB];
#pragma unroll
            for (int k = 0; k < NB; ++k) as[k] = T(0);
            row_each<EW>(A, val, ellv, i, [&](int c, T a) {
#pragma unroll
                for (int k = 0; k < NB; ++k)
                    if (act[k]) {
                        const size_t o = (size_t)(sp.first + k) * n + c;
                        const T sc = r[o] - alpha[k] * v[o];
                        as[k] += a * (JAC ? sc * jac.of(sp.m)[c] : sc);
                    }
            });
#pragma unroll
            for (int k = 0; k < NB; ++k)
                if (act[k]) {
                    const size_t o = (size_t)(sp.first + k) * n + i;
                    const T si = r[o] - alpha[k] * v[o];
                    t[o] = as[k];
                    acc[k][0] += double(as[k]) * double(si);
                    acc[k][1] += double(as[k]) * double(as[k]);
                    acc[k][2] += double(si) * double(si);
                }
        }
    put_parts<NB, 3>(tsOut, acc, sp.first, sp.count, lds);
}

// alpha and omega from the folded partials (half-step test on (s, s)); x += alpha p + omega s,
// r = s - omega t (half step: x += alpha p, r = s), the x increment scaled by D^-1 with JAC;
// partials (r0, r), (r, r). Workgroup 0 commits the records.
template <typename T, bool JAC>
__global__ void __launch_bounds__(WG) k_bicg_step(int n, int rpm, T *__restrict__ r, const T *__restrict__ r0,
                                                  const T *__restrict__ p, const T *__restrict__ v,
                                                  const T *__restrict__ t, T *__restrict__ x, Parts r0vIn, Parts tsIn,
                                                  Parts rrOut, Phase ph, Jac<T> jac) {
    __shared__ double lds[16 * 3];
    const int b = blockIdx.y;
    SolveRec R = ph.in[b];
    bool act = rec_active(R);
    double r0v[1][1], ts[1][3];
    fold<1, 1>(r0vIn, 0, b, 1, r0v, lds);
    fold<1, 3>(tsIn, 0, b, 1, ts, lds);
    bool half = false;
    if (act) {
        R.alpha = rule_alpha(R, r0v[0][0]);
        if (rec_active(R)) R.omega = rule_bicg_half(R, ts[0][0], ts[0][1], ts[0][2], ph.scale);
        act = rec_active(R);
        half = act && (R.flags & R_HALF);
        if (half) rule_bicg_half_met(R);
        R.flags &= ~R_FRESH;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) ph.out[b] = R;
    double acc[1][2] = {{0.0, 0.0}};
    if (act) {
        const T alpha = T(R.alpha), omega = T(R.omega);
        const size_t base = (size_t)b * n;
        const T *dv = JAC ? jac.of(b / rpm) : nullptr;
        PP_ROWS(i, n) {
            const size_t o = base + i;
            const T si = r[o] - alpha * v[o];
            if (half) {
                x[o] += JAC ? alpha * p[o] * dv[i] : alpha * p[o];
                r[o] = si;
            } else {
                x[o] += JAC ? (alpha * p[o] + omega * si) * dv[i] : alpha * p[o] + omega * si;
                const T ri = si - omega * t[o];
                r[o] = ri;
                acc[0][0] += double(r0[o]) * double(ri);
                acc[0][1] += double(ri) * double(ri);
            }
        }
    }
    put_parts<1, 2>(rrOut, acc, b, 1, lds);
}

// Inverse diagonal of each matrix for the Jacobi preconditioner (kInverseDiagonal,
// krylov.cu:272-283): the row's diagonal entries summed, 1 where the sum is zero.
template <typename T>
__global__ void k_jacobi_dinv(Csr<T> A, int nmat, T *__restrict__ dinv) {
    const int m = blockIdx.y;
    const T *val = A.val + (size_t)m * A.valStride;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < A.n; i += gridDim.x * blockDim.x) {
        T d = T(0);
        for (int e = A.row[i]; e < A.row[i + 1]; ++e)
            if (A.col[e] == i) d += val[e];
        dinv[(size_t)m * A.n + i] = d != T(0) ? T(1) / d : T(1);
    }
}

// x = 0 for every right-hand side (all right-hand sides were exactly zero).
template <typename T>
__global__ void k_fill_zero(T *x, size_t count) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x)
        x[i] = T(0);
}

// Best iterate back into x for right-hand sides that stopped on the iteration limit or a
// stalled residual and had a better iterate earlier.
template <typename T>
__global__ void k_take_best(T *x, const T *xbest, const SolveRec *rec, int n, int B) {
    const int b = blockIdx.y;
    const SolveRec s = rec[b];
    if (!((s.fate == F_EXHAUSTED || s.fate == F_STALLED) && s.bestAt >= 0 && s.best < s.crit)) return;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        x[(size_t)b * n + i] = xbest[(size_t)b * n + i];
}

// ======================================================================================
// AMG V-cycle: level matrices and one generic level pass
// ======================================================================================
// A level matrix (operator A_l, restriction R_l or interpolation P_l) in one of three layouts:
//  * LAY_ELL (finest level only): entry k of row i at [k * n + i], width ellW;
//  * LAY_SELL: wave-interleaved sliced ELL -- slices of 64 / lanes rows, entry k * lanes + j of
//    row r of slice s at off[s] + 64 k + r lanes + j, so a wavefront-wide load covers `lanes`
//    entries of each of the slice's rows and a row's entries are split over `lanes` lanes;
//  * LAY_CSR: row pointers and columns, `lanes` lanes per row over consecutive entries.
// `lanes` (a power of two <= 32) is chosen per matrix by the autotuner for the sliced layouts
// and by the row length for plain CSR.
enum Layout { LAY_CSR = 0, LAY_SELL = 1, LAY_ELL = 2 };

template <typename H>
struct LvMat {
    const int32_t *row;   // CSR row pointers
    const int32_t *col;   // columns (CSR, ELL or SELL order)
    const H *val;         // values (same order)
    const int32_t *off;   // SELL slice offsets
    const int32_t *steps; // SELL slice steps (entries per lane)
    int n, lanes, ellW;
    int64_t valStride;    // values per batched operator (0: shared)
};

// Sum over row i of M(i, c) * g(c), split over `lanes` lanes of the wavefront (sub-lane j);
// after the exchange every lane of the row's group holds the total (fixed order).
template <int LAY, int EW, typename H, typename Gth>
__device__ __forceinline__ H lv_row(const LvMat<H> &M, const H *__restrict__ val, int i, int j, bool valid, Gth g) {
    H sum = H(0);
    if constexpr (LAY == LAY_ELL) {
        if (valid) {
            if constexpr (EW > 0) {
                int c[EW];
                H a[EW];
#pragma unroll
                for (int k = 0; k < EW; ++k) {
                    c[k] = M.col[(size_t)k * M.n + i];
                    a[k] = val[(size_t)k * M.n + i];
                }
#pragma unroll
                for (int k = 0; k < EW; ++k) sum += a[k] * g(c[k]);
            } else {
                for (int k = 0; k < M.ellW; ++k) sum += val[(size_t)k * M.n + i] * g(M.col[(size_t)k * M.n + i]);
            }
        }
        return sum;
    } else {
        if (valid) {
            if constexpr (LAY == LAY_SELL) {
                const int per = 64 / M.lanes;
                const int s = i / per;
                const int64_t base = M.off[s] + (int64_t)(i - s * per) * M.lanes + j;
                const int w = M.steps[s];
#pragma unroll 4
                for (int k = 0; k < w; ++k) sum += val[base + 64 * k] * g(M.col[base + 64 * k]);
            } else {
                const int end = M.row[i + 1];
                for (int e = M.row[i] + j; e < end; e += M.lanes) sum += val[e] * g(M.col[e]);
            }
        }
        for (int o = M.lanes >> 1; o > 0; o >>= 1) sum += __shfl_xor(sum, o, 64);
        return sum;
    }
}

enum LvOp { OP_FIRST = 0, OP_SWEEP = 1, OP_RESID = 2, OP_APPLY = 3 };

// One pass over a level. With b the level's right-hand side (minus `shift` on the finest
// level of a projected solve), d the inverse diagonal and w the damping:
//  OP_FIRST: first Jacobi sweep from a zero guess, x = w d b, fused with its residual
//            res = b - A x = b - w A (d b) (the neighbours' d b are gathered);
//  OP_SWEEP: out = x + w d (b - A x);
//  OP_RESID: res = b - A x;
//  OP_APPLY: out = M x (restriction / interpolation), or out += M x with accumulate.
// OP_SWEEP with sums.p set also leaves partials (b, out) and sum(out) of the unshifted b.
template <typename TB, typename H, typename TO>
struct LvPass {
    LvMat<H> M;
    const H *dinv;
    const TB *b;
    const H *x;
    TO *out;
    H *res;
    const double *shift; // per batch (null: none)
    int64_t dinvStride, bStride, xStride, outStride;
    H omega;
    int accumulate;
    const SolveRec *rec;
    Parts sums;
};

template <int OP, int LAY, int EW, typename TB, typename H, typename TO>
__device__ __forceinline__ void level_pass(const LvPass<TB, H, TO> &a, int batch, double shift, double *lds) {
    const LvMat<H> &M = a.M;
    const int n = M.n;
    const int lanes = LAY == LAY_ELL ? 1 : M.lanes;
    const int per = 64 / lanes;
    const int lane = threadIdx.x & 63, j = lane % lanes;
    const int wave = (blockIdx.x * WG + (int)threadIdx.x) >> 6, waves = gridDim.x * (WG / 64);
    const H *val = M.val + (size_t)batch * M.valStride;
    const H *d = a.dinv ? a.dinv + (size_t)batch * a.dinvStride : nullptr;
    const TB *b = a.b ? a.b + (size_t)batch * a.bStride : nullptr;
    const H *x = a.x + (size_t)batch * a.xStride;
    TO *out = a.out ? a.out + (size_t)batch * a.outStride : nullptr;
    H *res = a.res ? a.res + (size_t)batch * n : nullptr;
    auto rhs = [&](int c) { return a.shift ? H(double(b[c]) - shift) : H(b[c]); };
    double acc[1][2] = {{0.0, 0.0}};
    for (int tile = wave; tile * per < n; tile += waves) {
        const int i = tile * per + lane / lanes;
        const bool valid = i < n;
        if constexpr (OP == OP_FIRST) {
            const H s = lv_row<LAY, EW>(M, val, i, j, valid, [&](int c) { return d[c] * rhs(c); });
            if (valid && j == 0) {
                const H bi = rhs(i);
                out[i] = TO(a.omega * d[i] * bi);
                res[i] = bi - a.omega * s;
            }
        } else if constexpr (OP == OP_SWEEP) {
            const H s = lv_row<LAY, EW>(M, val, i, j, valid, [&](int c) { return x[c]; });
            if (valid && j == 0) {
                const TO o = TO(x[i] + a.omega * d[i] * (rhs(i) - s));
                out[i] = o;
                if (a.sums.p) {
                    acc[0][0] += double(b[i]) * double(o);
                    acc[0][1] += double(o);
                }
            }
        } else if constexpr (OP == OP_RESID) {
            const H s = lv_row<LAY, EW>(M, val, i, j, valid, [&](int c) { return x[c]; });
            if (valid && j == 0) res[i] = rhs(i) - s;
        } else {
            const H s = lv_row<LAY, EW>(M, val, i, j, valid, [&](int c) { return x[c]; });
            if (valid && j == 0) out[i] = a.accumulate ? TO(out[i] + s) : TO(s);
        }
    }
    if constexpr (OP == OP_SWEEP)
        if (a.sums.p) put_parts<1, 2>(a.sums, acc, batch, 1, lds);
}

template <int OP, int LAY, int EW, typename TB, typename H, typename TO>
__global__ void __launch_bounds__(WG) k_level(LvPass<TB, H, TO> a) {
    __shared__ double lds[32];
    const int batch = blockIdx.y;
    if (!rec_active(a.rec[batch])) return;
    level_pass<OP, LAY, EW>(a, batch, a.shift ? a.shift[batch] : 0.0, lds);
}

// ======================================================================================
// AMG-PCG outer loop
// ======================================================================================
// Decide and start the V-cycle: fold the latest residual's partials and apply the rules
// (open: the initial residual; otherwise first the step length of the update that formed it,
// whose non-finite value ends the solve, then the iteration's criterion), publish the
// residual's mean for the projection (meanOut), then the finest level's first Jacobi sweep
// with residual.
template <int LAY, int EW, typename T, typename H>
__global__ void __launch_bounds__(WG) k_pcg_decide(LvPass<T, H, H> a, Parts rrIn, Parts pqIn, Phase ph, double *meanOut,
                                                   int project, int open) {
    __shared__ double lds[32];
    const int b = blockIdx.y;
    SolveRec R = ph.in[b];
    bool act = rec_active(R);
    double rr[1][3];
    fold<1, 3>(rrIn, 0, b, 1, rr, lds);
    if (act) {
        if (open) rule_open(R, rr[0][0], ph.scale, false);
        else {
            double pq[1][1];
            fold<1, 1>(pqIn, 0, b, 1, pq, lds);
            SolveRec probe = R;
            (void)rule_alpha(probe, pq[0][0]);
            if (!rec_active(probe)) R = probe;
            else rule_step(R, rr[0][0], ph.scale, ph.keepBest != 0);
        }
        act = rec_active(R);
    }
    const double mean = project ? rr[0][2] / double(a.M.n) : 0.0;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        ph.out[b] = R;
        meanOut[b] = mean;
    }
    if (!act) return;
    level_pass<OP_FIRST, LAY, EW>(a, b, mean, lds);
}

// Direction of the preconditioned solve: rho = (r, z) with z = u - mean(u) (projection) or
// z = u, from the folded partials of the last finest-level sweep ((r, u), sum(u)) and of the
// update (sum(r)); then p' = z + beta pIn (z when fresh) formed at every neighbour as it is
// read, q = A p', partials (p', q). Workgroup 0 writes the records with rho (fresh cleared).
template <typename T, int EW>
__global__ void __launch_bounds__(WG) k_pcg_dir(Csr<T> A, const T *__restrict__ u, const T *__restrict__ pIn,
                                                T *__restrict__ pOut, T *__restrict__ q, Parts postIn, Parts rrIn,
                                                Parts pqOut, Phase ph, int project) {
    __shared__ double lds[32];
    const int n = A.n;
    const int b = blockIdx.y;
    SolveRec R = ph.in[b];
    bool act = rec_active(R);
    double ru[1][2];
    fold<1, 2>(postIn, 0, b, 1, ru, lds);
    double zmean = 0.0, rz = ru[0][0];
    if (project) {
        double rs[1][1];
        fold<1, 1>(rrIn, 2, b, 1, rs, lds);
        zmean = ru[0][1] / double(n);
        rz = ru[0][0] - zmean * rs[0][0];
    }
    T beta = T(0);
    bool fresh = false;
    if (act) {
        fresh = R.flags & R_FRESH;
        beta = T(rule_pcg_rho(R, rz));
        R.flags &= ~R_FRESH;
        act = rec_active(R);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) ph.out[b] = R;
    double acc[1][1] = {{0.0}};
    if (act) {
        const T zs = T(zmean);
        const T *val = A.val + (size_t)b * A.valStride;
        const T *ellv = A.ell_val + (A.valStride ? (size_t)b * A.ell_w * n : 0);
        const size_t base = (size_t)b * n;
        auto dir = [&](int c) {
            const T z = project ? u[base + c] - zs : u[base + c];
            return fresh ? z : z + beta * pIn[base + c];
        };
        PP_ROWS(i, n) {
            T ap = T(0);
            row_each<EW>(A, val, ellv, i, [&](int c, T av) { ap += av * dir(c); });
            const T pi = dir(i);
            pOut[base + i] = pi;
            q[base + i] = ap;
            acc[0][0] += double(pi) * double(ap);
        }
    }
    put_parts<1, 1>(pqOut, acc, b, 1, lds);
}

// Dense coarse solve x = pinv b: one wavefront per row of the pseudo-inverse.
template <typename H>
__global__ void __launch_bounds__(256) k_coarse_solve(const H *__restrict__ pinv, int nc, const H *__restrict__ b,
                                                      H *__restrict__ x, const SolveRec *__restrict__ rec, int batchedPinv) {
    const int batch = blockIdx.y;
    if (!rec_active(rec[batch])) return;
    if (batchedPinv) pinv += (size_t)batch * nc * nc;
    const int i = blockIdx.x * 4 + threadIdx.x / 64, lane = threadIdx.x % 64;
    double sum = 0;
    if (i < nc)
        for (int c = lane; c < nc; c += 64) sum += double(pinv[(size_t)i * nc + c]) * double(b[(size_t)batch * nc + c]);
    sum = wave_sum(sum);
    if (i < nc && lane == 0) x[(size_t)batch * nc + i] = H(sum);
}

// ---- coarse tail of the V-cycle: the smallest levels in one workgroup ----
// The levels lt..L-1, the dense coarse solve and the way back up run as phases of a single
// workgroup per batch (barriers between phases), instead of ~4 latency-bound launches per
// level. A workgroup has little memory-level parallelism, so only levels that one pass covers
// with at least 8 lanes per row qualify (coarse rows are long: 40-50 entries).
constexpr int TAIL_MAX = 12;
constexpr int TAIL_BLOCK = 1024;
constexpr int TAIL_ROWS = TAIL_BLOCK / 8; // larger tails measured slower (207 us at 750 rows)
constexpr int TAIL_CW = 16; // lanes per row of the dense coarse solve
// lanes per row in the tail: the widest power of two (<= 32) that covers the level in one pass
static int tail_width(int rows) {
    int w = 32;
    while (w > 1 && rows * w > TAIL_BLOCK) w >>= 1;
    return w;
}

template <typename H>
struct TailArgs {
    int nl, nc, pre, post, coarseBatched; // nl tail levels; index k = level - first tail level
    H omega;
    const H *pinv;
    H *cb, *cx;
    int n[TAIL_MAX], wa[TAIL_MAX], wr[TAIL_MAX], wp[TAIL_MAX];
    Csr<H> A[TAIL_MAX], R[TAIL_MAX], P[TAIL_MAX];
    const H *dinv[TAIL_MAX];
    H *x[TAIL_MAX], *y[TAIL_MAX], *res[TAIL_MAX], *b[TAIL_MAX];
};

template <typename H, typename G>
__device__ __forceinline__ H tail_row_sum(const Csr<H> &A, const H *val, int i, bool valid, int lane, int W, G gather) {
    H sum = H(0);
    if (valid) {
        const int e = A.row[i + 1];
        for (int j = A.row[i] + lane; j < e; j += W) sum += val[j] * gather(A.col[j]);
    }
    for (int o = W / 2; o > 0; o >>= 1) sum += __shfl_down(sum, o, W);
    return sum;
}

// The level table lives in device memory (written once per solve): indexing a by-value kernel
// argument with the runtime level index would copy the whole table into per-lane scratch.
// Level k's iterate after j Jacobi updates is x[k] for even j and y[k] for odd j.
template <typename H>
__global__ void __launch_bounds__(TAIL_BLOCK) k_vcycle_tail(const TailArgs<H> *__restrict__ ap,
                                                            const SolveRec *__restrict__ rec) {
    const int batch = blockIdx.y;
    if (!rec_active(rec[batch])) return;
    const TailArgs<H> &a = *ap;
    const int nl = a.nl, pre = a.pre, post = a.post;
    const H omega = a.omega;
    auto iterate = [&](int k, int updates) { return (updates & 1) ? a.y[k] : a.x[k]; };
    for (int k = 0; k < nl; ++k) {
        const int n = a.n[k], W = a.wa[k];
        const size_t off = (size_t)batch * n;
        const Csr<H> &A = a.A[k];
        const H *val = A.val + (size_t)batch * A.valStride;
        const H *dv = a.dinv[k] + (A.valStride ? off : 0);
        const H *bl = a.b[k] + off;
        H *res = a.res[k] + off;
        const int lane = threadIdx.x % W, grp = threadIdx.x / W, gpb = TAIL_BLOCK / W;
        {
            H *xo = a.x[k] + off;
            for (int base = 0; base < n; base += gpb) {
                const int i = base + grp;
                const bool valid = i < n;
                const H sum = tail_row_sum<H>(A, val, i, valid, lane, W, [&](int c) { return dv[c] * bl[c]; });
                if (valid && lane == 0) {
                    xo[i] = omega * dv[i] * bl[i];
                    res[i] = bl[i] - omega * sum;
                }
            }
        }
        __syncthreads();
        for (int sw = 1; sw < pre; ++sw) {
            const H *xc = iterate(k, sw - 1) + off;
            H *xo = iterate(k, sw) + off;
            for (int base = 0; base < n; base += gpb) {
                const int i = base + grp;
                const bool valid = i < n;
                const H sum = tail_row_sum<H>(A, val, i, valid, lane, W, [&](int c) { return xc[c]; });
                if (valid && lane == 0) xo[i] = xc[i] + omega * dv[i] * (bl[i] - sum);
            }
            __syncthreads();
        }
        if (pre > 1) {
            const H *xc = iterate(k, pre - 1) + off;
            for (int base = 0; base < n; base += gpb) {
                const int i = base + grp;
                const bool valid = i < n;
                const H sum = tail_row_sum<H>(A, val, i, valid, lane, W, [&](int c) { return xc[c]; });
                if (valid && lane == 0) res[i] = bl[i] - sum;
            }
            __syncthreads();
        }
        // restriction into the next level's right-hand side (or the coarsest one)
        const Csr<H> &R = a.R[k];
        const int ncR = R.n, WR = a.wr[k];
        H *bn = (k + 1 < nl ? a.b[k + 1] : a.cb) + (size_t)batch * ncR;
        const int laneR = threadIdx.x % WR, grpR = threadIdx.x / WR, gpbR = TAIL_BLOCK / WR;
        for (int base = 0; base < ncR; base += gpbR) {
            const int i = base + grpR;
            const bool valid = i < ncR;
            const H sum = tail_row_sum<H>(R, R.val, i, valid, laneR, WR, [&](int c) { return res[c]; });
            if (valid && laneR == 0) bn[i] = sum;
        }
        __syncthreads();
    }
    {
        // dense coarse solve: TAIL_CW lanes per row read consecutive entries of the row
        const int nc = a.nc;
        const H *pv = a.pinv + (a.coarseBatched ? (size_t)batch * nc * nc : 0);
        const H *cb = a.cb + (size_t)batch * nc;
        H *cx = a.cx + (size_t)batch * nc;
        const int lane = threadIdx.x % TAIL_CW, grp = threadIdx.x / TAIL_CW;
        for (int base = 0; base < nc; base += TAIL_BLOCK / TAIL_CW) {
            const int i = base + grp;
            double sum = 0;
            if (i < nc)
                for (int j = lane; j < nc; j += TAIL_CW) sum += double(pv[(size_t)i * nc + j]) * double(cb[j]);
#pragma unroll
            for (int o = TAIL_CW / 2; o > 0; o >>= 1) sum += __shfl_down(sum, o, TAIL_CW);
            if (i < nc && lane == 0) cx[i] = H(sum);
        }
    }
    __syncthreads();
    for (int k = nl - 1; k >= 0; --k) {
        const int n = a.n[k];
        const size_t off = (size_t)batch * n;
        const int ncoarse = k + 1 < nl ? a.n[k + 1] : a.nc;
        const H *xc = (k + 1 < nl ? iterate(k + 1, pre - 1 + post) : a.cx) + (size_t)batch * ncoarse;
        const Csr<H> &P = a.P[k];
        const int WP = a.wp[k];
        {
            H *xk = iterate(k, pre - 1) + off;
            const int lane = threadIdx.x % WP, grp = threadIdx.x / WP, gpb = TAIL_BLOCK / WP;
            for (int base = 0; base < n; base += gpb) {
                const int i = base + grp;
                const bool valid = i < n;
                const H sum = tail_row_sum<H>(P, P.val, i, valid, lane, WP, [&](int c) { return xc[c]; });
                if (valid && lane == 0) xk[i] += sum;
            }
        }
        __syncthreads();
        const Csr<H> &A = a.A[k];
        const H *val = A.val + (size_t)batch * A.valStride;
        const H *dv = a.dinv[k] + (A.valStride ? off : 0);
        const H *bl = a.b[k] + off;
        const int W = a.wa[k];
        const int lane = threadIdx.x % W, grp = threadIdx.x / W, gpb = TAIL_BLOCK / W;
        for (int sw = 0; sw < post; ++sw) {
            const H *xcur = iterate(k, pre - 1 + sw) + off;
            H *xo = iterate(k, pre + sw) + off;
            for (int base = 0; base < n; base += gpb) {
                const int i = base + grp;
                const bool valid = i < n;
                const H sum = tail_row_sum<H>(A, val, i, valid, lane, W, [&](int c) { return xcur[c]; });
                if (valid && lane == 0) xo[i] = xcur[i] + omega * dv[i] * (bl[i] - sum);
            }
            __syncthreads();
        }
    }
}


// ======================================================================================
// Host side: workspaces, the launch pacer, drivers
// ======================================================================================
struct Space {
    void *vec = nullptr;   size_t vecCap = 0;   // work vectors
    void *parts = nullptr; size_t partsCap = 0; // partial sets
    void *recs = nullptr;  size_t recsCap = 0;  // two record copies + per-batch means
    void *aux = nullptr;   size_t auxCap = 0;   // V-cycle level vectors
    void *tols = nullptr;  size_t tolsCap = 0;
    void *tailDev = nullptr; size_t tailCap = 0;
    void *tailHost = nullptr;
    char *snap = nullptr;  size_t snapCap = 0;  // pinned snapshots: two slots of B records
    bool fixed = false;    // carved from the static scratch array: never grows, never freed
    hipEvent_t ev[2] = {nullptr, nullptr}; // snapshot events created once (fixed spaces)
};
// Workspaces belong to the calling host thread: solves issued from different threads (each
// on its own stream) run concurrently without sharing buffers. g_lock guards what is shared
// (settings, the graph cache, the capture streams, the registry of workspaces, the static
// scratch of the allocation-free entry point).
static std::mutex g_lock;

// The workspaces of one host thread: CG, BiCGStab, AMG-PCG, the segmented engine (with its
// uploaded tile layout) and the transposed-matrix buffer.
struct SegSpace {
    Space sp;
    int32_t *layoutDev = nullptr; size_t layoutCap = 0;
    int32_t *layoutHost = nullptr; size_t layoutHostCap = 0;
};
struct ThreadSpaces {
    Space cg, bicg, amg;
    SegSpace seg;
    void *tbuf = nullptr; size_t tcap = 0;
    int *tbad = nullptr;
};
static std::vector<ThreadSpaces *> g_all_spaces; // every thread's, for pp_release_workspaces

static ThreadSpaces &my_spaces() {
    static thread_local ThreadSpaces *mine = nullptr;
    if (!mine) {
        mine = new ThreadSpaces();
        std::lock_guard<std::mutex> lock(g_lock);
        g_all_spaces.push_back(mine);
    }
    return *mine;
}

static int grow(void **ptr, size_t *cap, size_t bytes) {
    bytes = std::max<size_t>(bytes, 256);
    if (*ptr && *cap >= bytes) return PP_OK;
    if (*ptr) (void)hipFree(*ptr);
    *ptr = nullptr;
    *cap = 0;
    PP_HIP(hipMalloc(ptr, bytes));
    *cap = bytes;
    return PP_OK;
}

static int cu_count() {
    static int cus = 0;
    if (!cus) {
        int dev = 0;
        (void)hipGetDevice(&dev);
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) cus = 256;
    }
    return cus;
}

// Workgroups per grid row of a phase kernel: enough for every row, at most 8 per CU shared by
// the grid rows (the partial sets are folded by every workgroup of the next launch, so the
// width also bounds the fold).
static int phase_width(long rows, int gridRows) {
    const long want = std::max<long>(1, (rows + WG - 1) / WG);
    const long cap = std::max<long>(1, (long)cu_count() * PP_PHASE_WG_PER_CU / std::max(1, gridRows));
    return (int)std::min(want, cap);
}

// Carves partial sets out of one allocation.
struct PartsCarver {
    double *base;
    size_t used = 0;
    Parts take(int B, int nq, int G) {
        Parts p{base + used, nq, G};
        used += (size_t)B * nq * G;
        return p;
    }
};

static int prepare(Space &sp, size_t vecBytes, size_t partDoubles, int B, const double *tols, int ntols, hipStream_t s) {
    if (sp.fixed) {
        const size_t recBytes = 2 * sizeof(SolveRec) * (size_t)B + sizeof(double) * B;
        if (vecBytes > sp.vecCap || partDoubles * sizeof(double) > sp.partsCap || recBytes > sp.recsCap ||
            2 * sizeof(SolveRec) * (size_t)B > sp.snapCap || sizeof(double) * (size_t)ntols > sp.tolsCap) {
            set_error("solve exceeds the static scratch maximum (" + std::to_string(sp.vecCap >> 20) +
                      " MiB of work vectors, " + std::to_string(sp.snapCap / (2 * sizeof(SolveRec))) +
                      " right-hand sides): use pp_solve_linear");
            return PP_ERR_ARG;
        }
        if (ntols > 1) PP_HIP(hipMemcpyAsync(sp.tols, tols, sizeof(double) * ntols, hipMemcpyHostToDevice, s));
        return PP_OK;
    }
    if (int e = grow(&sp.vec, &sp.vecCap, vecBytes)) return e;
    if (int e = grow(&sp.parts, &sp.partsCap, partDoubles * sizeof(double))) return e;
    if (int e = grow(&sp.recs, &sp.recsCap, 2 * sizeof(SolveRec) * B + sizeof(double) * B)) return e;
    const size_t snapBytes = 2 * sizeof(SolveRec) * (size_t)B;
    if (sp.snapCap < snapBytes) {
        if (sp.snap) (void)hipHostFree(sp.snap);
        sp.snap = nullptr;
        sp.snapCap = 0;
        PP_HIP(hipHostMalloc((void **)&sp.snap, snapBytes, hipHostMallocDefault));
        sp.snapCap = snapBytes;
    }
    if (ntols > 1) {
        if (int e = grow(&sp.tols, &sp.tolsCap, sizeof(double) * ntols)) return e;
        PP_HIP(hipMemcpyAsync(sp.tols, tols, sizeof(double) * ntols, hipMemcpyHostToDevice, s));
    }
    return PP_OK;
}

static SolveRec *recs_a(Space &sp) { return reinterpret_cast<SolveRec *>(sp.recs); }
static SolveRec *recs_b(Space &sp, int B) { return recs_a(sp) + B; }
static double *recs_means(Space &sp, int B) { return reinterpret_cast<double *>(recs_a(sp) + 2 * B); }

static void report(const SolveRec *recs, int B, int returnBest, pp_solve_result *out) {
    for (int b = 0; b < B; ++b) {
        const SolveRec &s = recs[b];
        pp_solve_result r;
        r.final_residual = s.crit;
        r.used_iterations = s.report;
        r.converged = s.fate == F_MET;
        r.is_finite_residual = std::isfinite(s.crit) ? 1 : 0;
        r.pad = s.fate;
        if (returnBest && (s.fate == F_EXHAUSTED || s.fate == F_STALLED) && s.bestAt >= 0 && s.best < s.crit) {
            r.used_iterations = s.bestAt;
            r.final_residual = s.best;
        }
        out[b] = r;
    }
}

static void report_zero(int B, pp_solve_result *out) {
    for (int b = 0; b < B; ++b) out[b] = pp_solve_result{0.0, 0, 1, 1, F_MET};
}

// The launch pacer. After the opening launches the host observes the decided records once,
// synchronously (a solve that starts converged enqueues nothing more). Then it enqueues
// iterations in chunks; each chunk ends with an asynchronous copy of the decided records into
// one of two pinned slots and an event, and at most two chunks are unobserved at a time. With
// two observations of a right-hand side the host fits a log-linear decay of its criterion and
// estimates how many iterations remain; it keeps that many, less `margin`, enqueued (at most
// `cap` per chunk) and otherwise waits for the next observation. Without an estimate, `eager`
// (cheap iterations) doubles the chunk from 1 up to `cap`; otherwise one iteration is enqueued
// per observation. Launches only follow the device records, so a wrong estimate costs time,
// never results.
struct Pace {
    int cap, margin;
    bool eager;
};

// Captured chunks of iterations (the useGraphs setting): a chunk of k iterations starting at
// an even or odd iteration is captured once into a HIP graph, keyed by everything its kernels
// are given (pointers, sizes, grids), and replayed with one graph launch instead of 2k-4k
// kernel launches. The cache keeps the most recent entries.
static int g_use_graphs = 1;
static std::mutex g_graph_lock; // the graph cache and the capture streams
struct ChunkGraph {
    std::vector<uintptr_t> key;
    int k, parity;
    hipGraphExec_t exec;
};
static std::vector<ChunkGraph> g_graphs;
constexpr size_t GRAPH_CACHE = 96;

static void drop_graphs() {
    for (auto &g : g_graphs) (void)hipGraphExecDestroy(g.exec);
    g_graphs.clear();
}

template <typename Body>
static int launch_chunk(hipStream_t s, int first, int k, const std::vector<uintptr_t> *key, Body &body) {
    if (!key || !g_use_graphs || timer_enabled()) {
        for (int j = 0; j < k; ++j) body(s, first + j);
        return PP_OK;
    }
    const int parity = first & 1;
    std::lock_guard<std::mutex> glock(g_graph_lock);
    for (size_t i = 0; i < g_graphs.size(); ++i)
        if (g_graphs[i].k == k && g_graphs[i].parity == parity && g_graphs[i].key == *key) {
            PP_HIP(hipGraphLaunch(g_graphs[i].exec, s));
            return PP_OK;
        }
    // capture on a private stream of this device (the caller's may be the legacy default
    // stream, which cannot capture); the graph is then launched on the caller's stream
    static hipStream_t capture[16] = {};
    int dev = 0;
    PP_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= 16) {
        for (int j = 0; j < k; ++j) body(s, first + j);
        return PP_OK;
    }
    if (!capture[dev]) PP_HIP(hipStreamCreateWithFlags(&capture[dev], hipStreamNonBlocking));
    hipStream_t cs = capture[dev];
    hipGraph_t graph = nullptr;
    PP_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
    for (int j = 0; j < k; ++j) body(cs, first + j);
    const hipError_t ce = hipStreamEndCapture(cs, &graph);
    if (ce != hipSuccess || !graph) {
        set_error(std::string("launch_chunk: capture failed: ") + hipGetErrorString(ce));
        return PP_ERR_HIP;
    }
    hipGraphExec_t exec = nullptr;
    const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ie != hipSuccess) {
        set_error(std::string("launch_chunk: instantiate failed: ") + hipGetErrorString(ie));
        return PP_ERR_HIP;
    }
    if (g_graphs.size() >= GRAPH_CACHE) {
        (void)hipGraphExecDestroy(g_graphs.front().exec);
        g_graphs.erase(g_graphs.begin());
    }
    g_graphs.push_back(ChunkGraph{*key, k, parity, exec});
    PP_HIP(hipGraphLaunch(exec, s));
    return PP_OK;
}

template <typename Open, typename Body>
static int pace_solve(Space &sp, hipStream_t s, const SolveRec *decided, int B, int loopsMax, Pace pace, int *lastSlot,
                      Open stopAfterOpening, Body body, const std::vector<uintptr_t> *graphKey = nullptr) {
    const size_t bytes = sizeof(SolveRec) * (size_t)B;
    const SolveRec *slots[2] = {reinterpret_cast<SolveRec *>(sp.snap), reinterpret_cast<SolveRec *>(sp.snap + bytes)};
    hipEvent_t ev[2] = {sp.ev[0], sp.ev[1]};
    const bool ownEvents = !(ev[0] && ev[1]); // fixed spaces bring events created at set-up
    int rc = PP_OK;
    auto cleanup = [&]() {
        if (ownEvents)
            for (auto &e : ev)
                if (e) (void)hipEventDestroy(e);
    };
    if (ownEvents)
        for (auto &e : ev)
            if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
                cleanup();
                set_error("pace_solve: event creation failed");
                return PP_ERR_HIP;
            }
    auto snap = [&](int slot) -> int {
        if (hipMemcpyAsync((void *)slots[slot], decided, bytes, hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipEventRecord(ev[slot], s) != hipSuccess) {
            set_error("pace_solve: snapshot failed");
            return PP_ERR_HIP;
        }
        return PP_OK;
    };
    auto wait = [&](int slot) -> int {
        if (hipEventSynchronize(ev[slot]) != hipSuccess) {
            set_error("pace_solve: synchronisation failed");
            return PP_ERR_HIP;
        }
        return PP_OK;
    };
    auto allStopped = [&](int slot) {
        for (int b = 0; b < B; ++b)
            if (rec_active(slots[slot][b])) return false;
        return true;
    };
    std::vector<int> seenIt(B, -1);
    std::vector<double> seenCrit(B, 0.0);
    // iterations still needed after an observation (-1: no estimate for some right-hand side)
    auto remaining = [&](int slot) -> int {
        int most = 0;
        bool known = true;
        for (int b = 0; b < B; ++b) {
            const SolveRec &x = slots[slot][b];
            if (!rec_active(x)) continue;
            int need = -1;
            if (seenIt[b] >= 0 && x.iter > seenIt[b] && seenCrit[b] > 0.0 && x.crit > 0.0 && x.tol > 0.0 &&
                std::isfinite(x.crit) && x.crit < seenCrit[b]) {
                const double perIt = std::log(x.crit / seenCrit[b]) / double(x.iter - seenIt[b]);
                const double est = std::log(x.tol / x.crit) / perIt;
                if (std::isfinite(est)) need = std::max(1, (int)std::ceil(est - 1e-9));
            }
            if (x.iter != seenIt[b]) {
                seenIt[b] = x.iter;
                seenCrit[b] = x.crit;
            }
            if (need < 0) known = false;
            else most = std::max(most, std::min(need, std::max(1, x.maxit - x.iter)));
        }
        return known ? most : -1;
    };
    do {
        if ((rc = snap(0)) || (rc = wait(0))) break;
        *lastSlot = 0;
        if (allStopped(0) || stopAfterOpening(slots[0])) break;
        (void)remaining(0);
        const int cap = std::max(1, pace.cap);
        int done = 0, target = -1, chunk = 1, slot = 0;
        int pendSlot[2], pendLoops[2], pending = 0;
        while (true) {
            int next;
            if (target >= 0) next = std::min(target - pace.margin - done, cap);
            else if (pace.eager) next = pending < 2 ? chunk : 0;
            else next = pending == 0 ? 1 : 0;
            if (next <= 0) next = pending == 0 ? 1 : 0;
            next = std::min(next, loopsMax - done);
            if (next > 0) {
                if ((rc = launch_chunk(s, done, next, graphKey, body))) break;
                done += next;
                if (hipGetLastError() != hipSuccess) {
                    set_error("pace_solve: kernel launch failed");
                    rc = PP_ERR_HIP;
                    break;
                }
                slot ^= 1;
                if ((rc = snap(slot))) break;
                pendSlot[pending] = slot;
                pendLoops[pending] = done;
                ++pending;
                if (target < 0 && pace.eager) chunk = std::min(2 * chunk, cap);
            }
            if (pending == 0) break; // every allowed iteration enqueued and observed
            if (pending == 2 || next == 0) {
                const int obs = pendSlot[0], obsLoops = pendLoops[0];
                if ((rc = wait(obs))) break;
                *lastSlot = obs;
                pendSlot[0] = pendSlot[1];
                pendLoops[0] = pendLoops[1];
                --pending;
                if (allStopped(obs)) break;
                const int rem = remaining(obs);
                target = rem >= 0 ? obsLoops + rem : -1;
            }
        }
    } while (false);
    cleanup();
    return rc;
}

static const SolveRec *slot_recs(Space &sp, int slot, int B) {
    return reinterpret_cast<const SolveRec *>(sp.snap + (size_t)slot * sizeof(SolveRec) * B);
}

static bool all_zero_rhs(const SolveRec *r, int B) {
    for (int b = 0; b < B; ++b)
        if (!r[b].zero) return false;
    return true;
}

#define PP_NB_CASE(NBV, ...)                                  \
    switch (NBV) {                                            \
        case 1: { constexpr int NB = 1; __VA_ARGS__; break; } \
        case 2: { constexpr int NB = 2; __VA_ARGS__; break; } \
        case 3: { constexpr int NB = 3; __VA_ARGS__; break; } \
        default: { constexpr int NB = 4; __VA_ARGS__; break; } \
    }
#define PP_EW_CASE(EWV, ...)                                  \
    switch (EWV) {                                            \
        case 7: { constexpr int EW = 7; __VA_ARGS__; break; } \
        case 5: { constexpr int EW = 5; __VA_ARGS__; break; } \
        default: { constexpr int EW = 0; __VA_ARGS__; break; } \
    }

// Settings of the fused solvers (SetKrylovSettings, bindings.cpp:739-752): iterations per
// launch chunk at most, and the Jacobi preconditioner of CG and of BiCGStab.
static int g_chunk_cap = 8;
static int g_cg_jacobi = 0, g_bicg_jacobi = 0;

#define PP_JAC_CASE(ON, ...)                                   \
    if (ON) { constexpr bool JAC = true; __VA_ARGS__; }        \
    else { constexpr bool JAC = false; __VA_ARGS__; }

// Inverse diagonals of the Jacobi preconditioner into dv (numMat x n).
template <typename T>
static void launch_jacobi(const Csr<T> &A, int numMat, T *dv, hipStream_t s) {
    hipLaunchKernelGGL((k_jacobi_dinv<T>), dim3(std::min(1024, (A.n + 255) / 256), numMat), dim3(256), 0, s, A, numMat, dv);
}

// ---------------- CG (large systems) ----------------
template <typename T>
static int cg_large(const Csr<T> &A, const T *b, T *x, int B, int maxit, const double *tols, int ntols, double scale,
                    int resetEvery, int returnBest, int project, int jacobi, Space &sp, hipStream_t s,
                    pp_solve_result *res) {
    const int n = A.n;
    const int NBv = std::min(A.rpm, 4);
    const int numMat = (B + A.rpm - 1) / A.rpm;
    const int rows = numMat * ((A.rpm + NBv - 1) / NBv);
    const int G = phase_width(n, rows);
    enum { V_R = 0, V_Q, V_P0, V_P1, V_XB, V_BW, V_COUNT };
    const size_t vecBytes = sizeof(T) * ((size_t)V_COUNT * n * B + (size_t)numMat * n);
    const size_t partDoubles = (size_t)B * G * (1 + 3 + 1 + 2);
    if (int e = prepare(sp, vecBytes, partDoubles, B, tols, ntols, s)) return e;
    T *vec = reinterpret_cast<T *>(sp.vec);
    auto V = [&](int k) { return vec + (size_t)k * n * B; };
    T *r = V(V_R), *q = V(V_Q), *p[2] = {V(V_P0), V(V_P1)};
    T *xbest = returnBest ? V(V_XB) : nullptr;
    T *bw = project ? V(V_BW) : const_cast<T *>(b);
    const Jac<T> jac{V(V_COUNT), A.valStride ? n : 0};
    PartsCarver pc{reinterpret_cast<double *>(sp.parts)};
    const Parts sSum = pc.take(B, 1, G), sRR = pc.take(B, 3, G), sPQ = pc.take(B, 1, G), sRst = pc.take(B, 2, G);
    SolveRec *ra = recs_a(sp), *rb = recs_b(sp, B);
    const dim3 grid(G, rows), block(WG);
    hipLaunchKernelGGL(k_init_recs, dim3(1), dim3(64), 0, s, ra, B, tols[0], ntols > 1 ? (const double *)sp.tols : nullptr, maxit);
    if (jacobi) launch_jacobi<T>(A, numMat, V(V_COUNT), s);
    if (project) hipLaunchKernelGGL((k_sum<T>), dim3(G, B), block, 0, s, b, n, sSum);
    const Phase openPh{ra, rb, B, scale, returnBest};
    PP_JAC_CASE(jacobi, PP_NB_CASE(NBv, PP_EW_CASE(A.ell_w,
        hipLaunchKernelGGL((k_open<T, NB, EW, JAC>), grid, block, 0, s, A, b, project ? sSum : Parts{nullptr, 1, G},
                           project ? bw : (T *)nullptr, x, r, (T *)nullptr, B, sRR, jac);
        hipLaunchKernelGGL((k_cg_turn<T, NB, EW, JAC>), grid, block, 0, s, A, r, p[0], p[1], q, x, bw, sRR, sPQ, sRst,
                           openPh, (int)TURN_OPEN, jac))));
    PP_LAUNCH_CHECK();
    const Phase stepPh{rb, ra, B, scale, returnBest}, turnPh{ra, rb, B, scale, returnBest};
    const Phase afterPh{rb, nullptr, B, scale, returnBest};
    auto body = [&](hipStream_t st, int j) {
        const int cur = 1 ^ (j & 1); // the opening turn wrote p[1]
        const bool reset = resetEvery > 0 && (j + 1) % resetEvery == 0;
        PP_JAC_CASE(jacobi, PP_NB_CASE(NBv,
            PP_TIMED(PP_K_CG_UPDATE, n, st,
                hipLaunchKernelGGL((k_cg_step<T, NB, JAC>), grid, block, 0, st, n, A.rpm, r, p[cur], q, x, xbest, sPQ,
                                   sRst, sRR, stepPh, jac));
            PP_EW_CASE(A.ell_w,
                if (reset) {
                    hipLaunchKernelGGL((k_cg_turn<T, NB, EW, JAC>), grid, block, 0, st, A, r, p[cur], p[cur ^ 1], q, x, bw,
                                       sRR, sPQ, sRst, turnPh, (int)TURN_STEP_RESET, jac);
                    hipLaunchKernelGGL((k_cg_turn<T, NB, EW, JAC>), grid, block, 0, st, A, r, p[cur], p[cur ^ 1], q, x, bw,
                                       sRR, sPQ, sRst, afterPh, (int)TURN_AFTER_RESET, jac);
                } else {
                    PP_TIMED(PP_K_CG_DIRECTION, n, st,
                        hipLaunchKernelGGL((k_cg_turn<T, NB, EW, JAC>), grid, block, 0, st, A, r, p[cur], p[cur ^ 1], q, x,
                                           bw, sRR, sPQ, sRst, turnPh, (int)TURN_STEP, jac));
                })));
    };
    int last = 0;
    bool zero = false;
    auto opening = [&](const SolveRec *recs) { return zero = all_zero_rhs(recs, B); };
    if (int e = pace_solve(sp, s, rb, B, maxit + 2, Pace{g_chunk_cap, 0, true}, &last, opening, body)) return e;
    if (zero) {
        hipLaunchKernelGGL((k_fill_zero<T>), dim3(std::min(4096, (n * B + 255) / 256)), dim3(256), 0, s, x, (size_t)n * B);
        PP_LAUNCH_CHECK();
        report_zero(B, res);
        return PP_OK;
    }
    if (returnBest) hipLaunchKernelGGL((k_take_best<T>), dim3(std::min(1024, (n + 255) / 256), B), dim3(256), 0, s, x, xbest, rb, n, B);
    PP_LAUNCH_CHECK();
    report(slot_recs(sp, last, B), B, returnBest, res);
    return PP_OK;
}

// ---------------- BiCGStab (large systems) ----------------
template <typename T>
static int bicg_large(const Csr<T> &A, const T *b, T *x, int B, int maxit, const double *tols, int ntols, double scale,
                      int jacobi, Space &sp, hipStream_t s, pp_solve_result *res) {
    const int n = A.n;
    const int NBv = std::min(A.rpm, 4);
    const int numMat = (B + A.rpm - 1) / A.rpm;
    const int rows = numMat * ((A.rpm + NBv - 1) / NBv);
    const int G = phase_width(n, std::max(rows, B));
    enum { V_R = 0, V_R0, V_P, V_V, V_T, V_COUNT };
    const size_t vecBytes = sizeof(T) * ((size_t)V_COUNT * n * B + (size_t)numMat * n);
    const size_t partDoubles = (size_t)B * G * (3 + 1 + 3);
    if (int e = prepare(sp, vecBytes, partDoubles, B, tols, ntols, s)) return e;
    T *vec = reinterpret_cast<T *>(sp.vec);
    auto V = [&](int k) { return vec + (size_t)k * n * B; };
    T *r = V(V_R), *r0 = V(V_R0), *p = V(V_P), *v = V(V_V), *t = V(V_T);
    const Jac<T> jac{V(V_COUNT), A.valStride ? n : 0};
    PartsCarver pc{reinterpret_cast<double *>(sp.parts)};
    const Parts sRR = pc.take(B, 3, G), sV = pc.take(B, 1, G), sT = pc.take(B, 3, G);
    SolveRec *ra = recs_a(sp), *rb = recs_b(sp, B);
    const dim3 grid(G, rows), gridB(G, B), block(WG);
    hipLaunchKernelGGL(k_init_recs, dim3(1), dim3(64), 0, s, ra, B, tols[0], ntols > 1 ? (const double *)sp.tols : nullptr, maxit);
    if (jacobi) launch_jacobi<T>(A, numMat, V(V_COUNT), s);
    PP_NB_CASE(NBv, PP_EW_CASE(A.ell_w,
        hipLaunchKernelGGL((k_open<T, NB, EW, false>), grid, block, 0, s, A, b, Parts{nullptr, 1, G}, (T *)nullptr, x, r, r0, B,
                           sRR, jac)));
    hipLaunchKernelGGL((k_bicg_dir<T>), gridB, block, 0, s, n, r, p, v, sRR, Phase{ra, rb, B, scale, 0}, 1);
    PP_LAUNCH_CHECK();
    const Phase useB{rb, nullptr, B, scale, 0}, commit{rb, ra, B, scale, 0}, decide{ra, rb, B, scale, 0};
    auto body = [&](hipStream_t st, int) {
        PP_JAC_CASE(jacobi,
            PP_NB_CASE(NBv, PP_EW_CASE(A.ell_w,
                PP_TIMED(PP_K_BICG_V, n, st,
                    hipLaunchKernelGGL((k_bicg_apply<T, NB, EW, JAC>), grid, block, 0, st, A, p, v, r0, sV, useB, jac));
                PP_TIMED(PP_K_BICG_T, n, st,
                    hipLaunchKernelGGL((k_bicg_stab<T, NB, EW, JAC>), grid, block, 0, st, A, r, v, t, sV, sT, useB, jac))));
            PP_TIMED(PP_K_BICG_UPDATE, n, st,
                hipLaunchKernelGGL((k_bicg_step<T, JAC>), gridB, block, 0, st, n, A.rpm, r, r0, p, v, t, x, sV, sT, sRR,
                                   commit, jac)));
        PP_TIMED(PP_K_BICG_P, n, st, hipLaunchKernelGGL((k_bicg_dir<T>), gridB, block, 0, st, n, r, p, v, sRR, decide, 0));
    };
    int last = 0;
    bool zero = false;
    auto opening = [&](const SolveRec *recs) { return zero = all_zero_rhs(recs, B); };
    if (int e = pace_solve(sp, s, rb, B, maxit + 2, Pace{g_chunk_cap, 0, true}, &last, opening, body)) return e;
    if (zero) {
        hipLaunchKernelGGL((k_fill_zero<T>), dim3(std::min(4096, (n * B + 255) / 256)), dim3(256), 0, s, x, (size_t)n * B);
        PP_LAUNCH_CHECK();
        report_zero(B, res);
        return PP_OK;
    }
    report(slot_recs(sp, last, B), B, 0, res);
    return PP_OK;
}

// ---------------- small systems ----------------
template <typename T>
static int small_solve(const Csr<T> &A, const T *b, T *x, int B, int maxit, const double *tols, int ntols, double scale,
                       int resetEvery, int returnBest, int project, int bicg, int jacobi, Space &sp, hipStream_t s,
                       pp_solve_result *res) {
    const int n = A.n;
    const int numMat = (B + A.rpm - 1) / A.rpm;
    if (int e = prepare(sp, sizeof(T) * ((size_t)SMALL_SLOTS * n * B + (size_t)numMat * n), 8, B, tols, ntols, s)) return e;
    SolveRec *ra = recs_a(sp);
    T *dinv = jacobi ? reinterpret_cast<T *>(sp.vec) + (size_t)SMALL_SLOTS * n * B : nullptr;
    hipLaunchKernelGGL(k_init_recs, dim3(1), dim3(64), 0, s, ra, B, tols[0], ntols > 1 ? (const double *)sp.tols : nullptr, maxit);
    if (jacobi) launch_jacobi<T>(A, numMat, dinv, s);
    SmallArgs a;
    a.n = n; a.B = B; a.rpm = A.rpm; a.maxit = maxit; a.normalized = 1; a.returnBest = returnBest; a.project = project;
    a.resetEvery = resetEvery; a.bicg = bicg; a.normScale = scale;
    a.dinvStride = A.valStride ? n : 0;
    a.nb_per_group = std::min(A.rpm, 4);
    const int groups = numMat * ((A.rpm + a.nb_per_group - 1) / a.nb_per_group);
    PP_TIMED(bicg ? PP_K_BICG_SMALL : PP_K_CG_SMALL, n, s, launch_small<T>(A, b, x, sp.vec, ra, dinv, a, groups, s));
    PP_LAUNCH_CHECK();
    SolveRec *host = reinterpret_cast<SolveRec *>(sp.snap);
    PP_HIP(hipMemcpyAsync(host, ra, sizeof(SolveRec) * B, hipMemcpyDeviceToHost, s));
    PP_HIP(hipStreamSynchronize(s));
    report(host, B, returnBest, res);
    return PP_OK;
}

// ---------------- AMG-PCG ----------------
// Lanes per row for a plain CSR level matrix (the autotuned sliced copies carry their own):
// about four entries per lane.
static int csr_lanes(long nnz, long rows) {
    const double avg = double(nnz) / double(std::max<long>(1, rows));
    int lanes = 1;
    while (lanes < 16 && avg > 4.0 * lanes) lanes <<= 1;
    return lanes;
}

template <typename H>
static LvMat<H> lv_csr(const int32_t *row, const int32_t *col, const void *val, int n, long nnz, int64_t stride) {
    LvMat<H> M{};
    M.row = row; M.col = col; M.val = reinterpret_cast<const H *>(val);
    M.n = n; M.lanes = csr_lanes(nnz, n); M.valStride = stride;
    return M;
}

template <typename H>
static LvMat<H> lv_sell(const pp_sell &sd, int n, bool batched) {
    LvMat<H> M{};
    M.col = sd.col; M.val = reinterpret_cast<const H *>(sd.val); M.off = sd.slice_off; M.steps = sd.slice_steps;
    M.n = n; M.lanes = sd.lanes; M.valStride = batched ? sd.total : 0;
    return M;
}

// Grid of a level pass over n rows of matrix M: one wavefront per slice of 64 / lanes rows.
template <typename H>
static dim3 level_grid(const LvMat<H> &M, int lay, int B) {
    const int per = lay == LAY_ELL ? 64 : 64 / M.lanes;
    const long waves = (M.n + per - 1) / per;
    const long wgs = std::max<long>(1, (waves + WG / 64 - 1) / (WG / 64));
    return dim3((unsigned)std::min<long>(wgs, std::max(1, cu_count() * PP_LEVEL_WG_PER_CU / B)), B);
}

template <int OP, typename TB, typename H, typename TO>
static void launch_level(const LvPass<TB, H, TO> &a, int lay, int B, hipStream_t s) {
    const dim3 g = level_grid(a.M, lay, B);
    if (lay == LAY_SELL) hipLaunchKernelGGL((k_level<OP, LAY_SELL, 0, TB, H, TO>), g, dim3(WG), 0, s, a);
    else if (lay == LAY_CSR) hipLaunchKernelGGL((k_level<OP, LAY_CSR, 0, TB, H, TO>), g, dim3(WG), 0, s, a);
    else if (a.M.ellW == 7) hipLaunchKernelGGL((k_level<OP, LAY_ELL, 7, TB, H, TO>), g, dim3(WG), 0, s, a);
    else if (a.M.ellW == 5) hipLaunchKernelGGL((k_level<OP, LAY_ELL, 5, TB, H, TO>), g, dim3(WG), 0, s, a);
    else hipLaunchKernelGGL((k_level<OP, LAY_ELL, 0, TB, H, TO>), g, dim3(WG), 0, s, a);
}

template <typename T, typename H>
static void launch_decide(const LvPass<T, H, H> &a, int lay, const Parts &rr, const Parts &pq, const Phase &ph,
                          double *means, int project, int open, int B, hipStream_t s) {
    const dim3 g = level_grid(a.M, lay, B);
    if (lay == LAY_CSR)
        hipLaunchKernelGGL((k_pcg_decide<LAY_CSR, 0, T, H>), g, dim3(WG), 0, s, a, rr, pq, ph, means, project, open);
    else if (a.M.ellW == 7)
        hipLaunchKernelGGL((k_pcg_decide<LAY_ELL, 7, T, H>), g, dim3(WG), 0, s, a, rr, pq, ph, means, project, open);
    else if (a.M.ellW == 5)
        hipLaunchKernelGGL((k_pcg_decide<LAY_ELL, 5, T, H>), g, dim3(WG), 0, s, a, rr, pq, ph, means, project, open);
    else hipLaunchKernelGGL((k_pcg_decide<LAY_ELL, 0, T, H>), g, dim3(WG), 0, s, a, rr, pq, ph, means, project, open);
}

template <typename T, typename H>
static int amg_pcg(const Csr<T> &Af, const pp_amg_level *lv, int L, const H *pinv, int nc, int coarseBatched, const T *b,
                   T *x, int B, int maxit, const double *tols, int ntols, double scale, int returnBest, int project,
                   int pre, int post, double omegaD, hipStream_t s, pp_solve_result *res) {
    Space &sp = my_spaces().amg;
    const int n = Af.n;
    pre = std::max(1, pre);
    post = std::max(1, post);
    const H omega = H(omegaD);
    // level matrices
    std::vector<LvMat<H>> MA(L), MR(L), MP(L);
    std::vector<int> layA(L), layR(L), layP(L);
    for (int l = 0; l < L; ++l) {
        const pp_amg_level &v = lv[l];
        if (l == 0 && v.a_ell_width > 0) {
            LvMat<H> M{};
            M.col = v.a_ell_col; M.val = reinterpret_cast<const H *>(v.a_ell_val); M.n = v.n; M.lanes = 1;
            M.ellW = v.a_ell_width; M.valStride = v.a_batched ? (int64_t)v.a_ell_width * v.n : 0;
            MA[l] = M;
            layA[l] = LAY_ELL;
        } else if (v.a_sell.slice_off) {
            MA[l] = lv_sell<H>(v.a_sell, v.n, v.a_batched);
            layA[l] = LAY_SELL;
        } else {
            MA[l] = lv_csr<H>(v.a_row, v.a_col, v.a_val, v.n, v.nnz_a, v.a_batched ? v.nnz_a : 0);
            layA[l] = LAY_CSR;
        }
        if (v.r_sell.slice_off) { MR[l] = lv_sell<H>(v.r_sell, v.n_coarse, false); layR[l] = LAY_SELL; }
        else { MR[l] = lv_csr<H>(v.r_row, v.r_col, v.r_val, v.n_coarse, v.nnz_r, 0); layR[l] = LAY_CSR; }
        if (v.p_sell.slice_off) { MP[l] = lv_sell<H>(v.p_sell, v.n, false); layP[l] = LAY_SELL; }
        else { MP[l] = lv_csr<H>(v.p_row, v.p_col, v.p_val, v.n, v.nnz_p, 0); layP[l] = LAY_CSR; }
    }
    const int G = phase_width(n, B);
    const int G0 = (int)level_grid(MA[0], layA[0], B).x;
    enum { V_R = 0, V_Q, V_P0, V_P1, V_U, V_XB, V_COUNT };
    const size_t partDoubles = (size_t)B * G * (1 + 3 + 1) + (size_t)B * 2 * G0;
    if (int e = prepare(sp, sizeof(T) * (size_t)V_COUNT * n * B, partDoubles, B, tols, ntols, s)) return e;
    T *vec = reinterpret_cast<T *>(sp.vec);
    auto V = [&](int k) { return vec + (size_t)k * n * B; };
    T *r = V(V_R), *q = V(V_Q), *p[2] = {V(V_P0), V(V_P1)}, *u = V(V_U);
    T *xbest = returnBest ? V(V_XB) : nullptr;
    PartsCarver pc{reinterpret_cast<double *>(sp.parts)};
    const Parts sSum = pc.take(B, 1, G), sRR = pc.take(B, 3, G), sPQ = pc.take(B, 1, G), sPost = pc.take(B, 2, G0);
    SolveRec *ra = recs_a(sp), *rb = recs_b(sp, B);
    double *means = recs_means(sp, B);
    // level vectors: x, res, y, b per level, then the coarsest b and x
    std::vector<size_t> offs(L + 1);
    size_t total = 0;
    for (int l = 0; l < L; ++l) { offs[l] = total; total += (size_t)4 * lv[l].n * B; }
    offs[L] = total;
    total += (size_t)2 * nc * B;
    if (int e = grow(&sp.aux, &sp.auxCap, total * sizeof(H))) return e;
    H *aux = reinterpret_cast<H *>(sp.aux);
    auto lx = [&](int l) { return aux + offs[l]; };
    auto lres = [&](int l) { return aux + offs[l] + (size_t)lv[l].n * B; };
    auto ly = [&](int l) { return aux + offs[l] + (size_t)2 * lv[l].n * B; };
    auto lb = [&](int l) { return l < L ? aux + offs[l] + (size_t)3 * lv[l].n * B : aux + offs[L]; };
    H *cx = aux + offs[L] + (size_t)nc * B;
    // the smallest levels (never the finest) run in the one-workgroup tail
    int lt = L;
    for (int l = 1; l < L; ++l)
        if (lv[l].n <= TAIL_ROWS && L - l <= TAIL_MAX) { lt = l; break; }
    if (lt < L) {
        TailArgs<H> tail{};
        tail.nl = L - lt; tail.nc = nc; tail.pre = pre; tail.post = post; tail.coarseBatched = coarseBatched;
        tail.omega = omega; tail.pinv = pinv; tail.cb = lb(L); tail.cx = cx;
        for (int l = lt; l < L; ++l) {
            const int k = l - lt;
            tail.n[k] = lv[l].n;
            tail.wa[k] = tail_width(lv[l].n);
            tail.wr[k] = tail_width(lv[l].n_coarse);
            tail.wp[k] = tail_width(lv[l].n);
            tail.A[k] = Csr<H>{lv[l].a_row, lv[l].a_col, reinterpret_cast<const H *>(lv[l].a_val), lv[l].n,
                               lv[l].a_batched ? lv[l].nnz_a : 0, 1, 0, nullptr, nullptr};
            tail.R[k] = Csr<H>{lv[l].r_row, lv[l].r_col, reinterpret_cast<const H *>(lv[l].r_val), lv[l].n_coarse, 0, 1, 0,
                               nullptr, nullptr};
            tail.P[k] = Csr<H>{lv[l].p_row, lv[l].p_col, reinterpret_cast<const H *>(lv[l].p_val), lv[l].n, 0, 1, 0,
                               nullptr, nullptr};
            tail.dinv[k] = reinterpret_cast<const H *>(lv[l].dinv);
            tail.x[k] = lx(l); tail.y[k] = ly(l); tail.res[k] = lres(l); tail.b[k] = lb(l);
        }
        // the pinned staging copy is rewritten only by a later solve, after this one's final
        // synchronisation
        if (!sp.tailHost) PP_HIP(hipHostMalloc(&sp.tailHost, sizeof(TailArgs<double>), hipHostMallocDefault));
        if (int e = grow(&sp.tailDev, &sp.tailCap, sizeof(TailArgs<double>))) return e;
        std::memcpy(sp.tailHost, &tail, sizeof(tail));
        PP_HIP(hipMemcpyAsync(sp.tailDev, sp.tailHost, sizeof(tail), hipMemcpyHostToDevice, s));
    }
    const TailArgs<H> *tailDev = reinterpret_cast<const TailArgs<H> *>(sp.tailDev);
    const double *shift = project ? means : nullptr;
    // pass descriptors
    auto pass_a = [&](int l) {
        LvPass<H, H, H> a{};
        a.M = MA[l];
        a.dinv = reinterpret_cast<const H *>(lv[l].dinv);
        a.dinvStride = lv[l].a_batched ? lv[l].n : 0;
        a.b = lb(l); a.bStride = lv[l].n;
        a.xStride = lv[l].n; a.outStride = lv[l].n;
        a.omega = omega;
        a.rec = rb;
        return a;
    };
    auto pass_fine = [&]() {
        LvPass<T, H, H> a{};
        a.M = MA[0];
        a.dinv = reinterpret_cast<const H *>(lv[0].dinv);
        a.dinvStride = lv[0].a_batched ? n : 0;
        a.b = r; a.bStride = n; a.shift = shift;
        a.xStride = n; a.outStride = n;
        a.omega = omega;
        a.rec = rb;
        return a;
    };
    auto pass_m = [&](const LvMat<H> &M, const H *in, int64_t inStride, H *out, int accumulate) {
        LvPass<H, H, H> a{};
        a.M = M;
        a.x = in; a.xStride = inStride;
        a.out = out; a.outStride = M.n;
        a.accumulate = accumulate;
        a.rec = rb;
        return a;
    };
    std::vector<H *> cur(L), oth(L);
    // the finest level's first sweep, launched by the deciding kernel
    auto decide = [&](hipStream_t st, const Phase &ph, int open) {
        LvPass<T, H, H> a = pass_fine();
        a.out = lx(0);
        a.res = lres(0);
        PP_TIMED(PP_K_AMG_PRESMOOTH, n, st, launch_decide<T, H>(a, layA[0], sRR, sPQ, ph, means, project, open, B, st));
    };
    // the rest of the V-cycle after the finest level's first sweep
    auto vcycle_rest = [&](hipStream_t st) {
        for (int l = 0; l < lt; ++l) {
            cur[l] = lx(l);
            oth[l] = ly(l);
            if (l > 0) {
                LvPass<H, H, H> a = pass_a(l);
                a.out = lx(l);
                a.res = lres(l);
                PP_TIMED(PP_K_AMG_PRESMOOTH, lv[l].n, st, launch_level<OP_FIRST>(a, layA[l], B, st));
            }
            for (int k = 1; k < pre; ++k) {
                if (l == 0) {
                    LvPass<T, H, H> a = pass_fine();
                    a.x = cur[0]; a.out = oth[0];
                    launch_level<OP_SWEEP>(a, layA[0], B, st);
                } else {
                    LvPass<H, H, H> a = pass_a(l);
                    a.x = cur[l]; a.out = oth[l];
                    launch_level<OP_SWEEP>(a, layA[l], B, st);
                }
                std::swap(cur[l], oth[l]);
            }
            if (pre > 1) {
                if (l == 0) {
                    LvPass<T, H, H> a = pass_fine();
                    a.x = cur[0]; a.res = lres(0);
                    launch_level<OP_RESID>(a, layA[0], B, st);
                } else {
                    LvPass<H, H, H> a = pass_a(l);
                    a.x = cur[l]; a.res = lres(l);
                    launch_level<OP_RESID>(a, layA[l], B, st);
                }
            }
            PP_TIMED(PP_K_AMG_RESTRICT, lv[l].n_coarse, st,
                     launch_level<OP_APPLY>(pass_m(MR[l], lres(l), lv[l].n, lb(l + 1), 0), layR[l], B, st));
        }
        if (lt < L) {
            PP_TIMED(PP_K_AMG_TAIL, lv[lt].n, st,
                     hipLaunchKernelGGL((k_vcycle_tail<H>), dim3(1, B), dim3(TAIL_BLOCK), 0, st, tailDev, rb));
            // the tail leaves each of its levels' iterate in x after an even number of sweeps
            const bool inX = ((pre - 1) + post) % 2 == 0;
            for (int l = lt; l < L; ++l) {
                cur[l] = inX ? lx(l) : ly(l);
                oth[l] = inX ? ly(l) : lx(l);
            }
        } else {
            PP_TIMED(PP_K_AMG_COARSE, nc, st,
                     hipLaunchKernelGGL((k_coarse_solve<H>), dim3((nc + 3) / 4, B), dim3(256), 0, st, pinv, nc, lb(L), cx, rb,
                                        coarseBatched));
        }
        for (int l = std::min(L, lt) - 1; l >= 0; --l) {
            const H *xc = l + 1 < L ? cur[l + 1] : cx;
            const int ncoarse = l + 1 < L ? lv[l + 1].n : nc;
            PP_TIMED(PP_K_AMG_PROLONG, lv[l].n, st,
                     launch_level<OP_APPLY>(pass_m(MP[l], xc, ncoarse, cur[l], 1), layP[l], B, st));
            for (int k = 0; k < post; ++k) {
                if (l == 0 && k == post - 1) {
                    // the preconditioned residual u, with partials (r, u) and sum(u)
                    LvPass<T, H, T> a{};
                    a.M = MA[0];
                    a.dinv = reinterpret_cast<const H *>(lv[0].dinv);
                    a.dinvStride = lv[0].a_batched ? n : 0;
                    a.b = r; a.bStride = n; a.shift = shift;
                    a.x = cur[0]; a.xStride = n;
                    a.out = u; a.outStride = n;
                    a.omega = omega;
                    a.rec = rb;
                    a.sums = sPost;
                    PP_TIMED(PP_K_AMG_POST_JACOBI, n, st, launch_level<OP_SWEEP>(a, layA[0], B, st));
                } else if (l == 0) {
                    LvPass<T, H, H> a = pass_fine();
                    a.x = cur[0]; a.out = oth[0];
                    PP_TIMED(PP_K_AMG_JACOBI, n, st, launch_level<OP_SWEEP>(a, layA[0], B, st));
                } else {
                    LvPass<H, H, H> a = pass_a(l);
                    a.x = cur[l]; a.out = oth[l];
                    PP_TIMED(PP_K_AMG_POST_JACOBI, lv[l].n, st, launch_level<OP_SWEEP>(a, layA[l], B, st));
                }
                std::swap(cur[l], oth[l]);
            }
        }
    };
    hipLaunchKernelGGL(k_init_recs, dim3(1), dim3(64), 0, s, ra, B, tols[0], ntols > 1 ? (const double *)sp.tols : nullptr, maxit);
    if (project) hipLaunchKernelGGL((k_sum<T>), dim3(G, B), dim3(WG), 0, s, b, n, sSum);
    PP_EW_CASE(Af.ell_w,
        hipLaunchKernelGGL((k_open<T, 1, EW, false>), dim3(G, B), dim3(WG), 0, s, Af, b, project ? sSum : Parts{nullptr, 1, G},
                           (T *)nullptr, x, r, (T *)nullptr, B, sRR, Jac<T>{nullptr, 0}));
    decide(s, Phase{ra, rb, B, scale, returnBest}, 1);
    PP_LAUNCH_CHECK();
    const Phase dirPh{rb, ra, B, scale, returnBest}, updPh{ra, nullptr, B, scale, returnBest},
        decPh{ra, rb, B, scale, returnBest};
    auto body = [&](hipStream_t st, int j) {
        const int cur2 = j & 1;
        vcycle_rest(st);
        PP_TIMED(PP_K_PCG_DIRECTION, n, st,
            PP_EW_CASE(Af.ell_w,
                hipLaunchKernelGGL((k_pcg_dir<T, EW>), dim3(G, B), dim3(WG), 0, st, Af, u, p[cur2], p[cur2 ^ 1], q, sPost, sRR,
                                   sPQ, dirPh, project)));
        PP_TIMED(PP_K_PCG_UPDATE, n, st,
            hipLaunchKernelGGL((k_cg_step<T, 1, false>), dim3(G, B), dim3(WG), 0, st, n, 1, r, p[cur2 ^ 1], q, x, xbest,
                               sPQ, Parts{nullptr, 2, G}, sRR, updPh, Jac<T>{nullptr, 0}));
        decide(st, decPh, 0);
    };
    int last = 0;
    auto opening = [&](const SolveRec *) { return false; };
    if (int e = pace_solve(sp, s, rb, B, maxit + 2, Pace{2, 1, false}, &last, opening, body)) return e;
    if (returnBest) hipLaunchKernelGGL((k_take_best<T>), dim3(std::min(1024, (n + 255) / 256), B), dim3(256), 0, s, x, xbest, rb, n, B);
    PP_LAUNCH_CHECK();
    report(slot_recs(sp, last, B), B, returnBest, res);
    return PP_OK;
}

// ---------------- autotuner probe ----------------
// Mean duration in microseconds of y = M x through a sliced copy with its lanes, over `reps`
// launches after one warm-up (x = 1, y scratch), for the autotuner of the lanes per row.
template <typename H>
static int probe_sell(const pp_sell &sd, int rows, int cols, int reps, hipStream_t s, double *us) {
    H *buf = nullptr;
    PP_HIP(hipMalloc((void **)&buf, sizeof(H) * ((size_t)rows + cols)));
    H *xin = buf + rows;
    hipLaunchKernelGGL((k_fill_zero<H>), dim3(std::min(1024, (rows + cols + 255) / 256)), dim3(256), 0, s, buf,
                       (size_t)rows + cols);
    hipEvent_t e0, e1;
    PP_HIP(hipEventCreate(&e0));
    PP_HIP(hipEventCreate(&e1));
    SolveRec *rec = nullptr;
    PP_HIP(hipMalloc((void **)&rec, sizeof(SolveRec)));
    hipLaunchKernelGGL(k_init_recs, dim3(1), dim3(64), 0, s, rec, 1, 1.0, (const double *)nullptr, 1);
    LvPass<H, H, H> a{};
    a.M = lv_sell<H>(sd, rows, false);
    a.x = xin; a.out = buf; a.outStride = rows;
    a.rec = rec;
    launch_level<OP_APPLY>(a, LAY_SELL, 1, s);
    PP_HIP(hipEventRecord(e0, s));
    for (int k = 0; k < reps; ++k) launch_level<OP_APPLY>(a, LAY_SELL, 1, s);
    PP_HIP(hipEventRecord(e1, s));
    PP_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    PP_HIP(hipEventElapsedTime(&ms, e0, e1));
    *us = 1000.0 * ms / std::max(1, reps);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(rec);
    (void)hipFree(buf);
    return PP_OK;
}

// ======================================================================================
// Segmented solves: independent systems of different sizes in one launch per phase
// ======================================================================================
// The instances' operators are the diagonal blocks of one CSR matrix (instance s owns matrix
// rows seg[s] .. seg[s+1]); a solve of ncomp right-hand sides per instance has K = ncomp * S
// pieces, piece k = c * S + s with its rows at vector offset base[k] = c * N + seg[s]. Tiles of
// tileRows rows map workgroups to pieces (a tile never straddles two pieces); partial sums are
// written per tile, [quantity][tile], and every tile folds the tiles of its own piece, so each
// piece runs its own recurrence, stopping rules and record, while one launch per phase covers
// all of them. The first tile of a piece writes its record.
namespace seg {

struct Pieces {
    const int32_t *base, *len, *mrow; // [K]
    const int32_t *tilePiece, *tileFirst; // [T]
    const int32_t *tileOff;           // [K + 1]
    int K, T, tileRows, N;
};

struct SegParts {
    double *p; // [nq][T]
    int nq, T;
    __device__ __forceinline__ double *at(int q) const { return p + (size_t)q * T; }
};

template <int NQ>
__device__ __forceinline__ void put(const SegParts &P, double (&acc)[NQ], double *lds) {
    block_sum<NQ>(acc, lds);
    if (threadIdx.x == 0)
#pragma unroll
        for (int q = 0; q < NQ; ++q) P.at(q)[blockIdx.x] = acc[q];
}

// Totals over the tiles t0 .. t1-1 of quantities q0.., the same bits in every tile.
template <int NQ>
__device__ __forceinline__ void fold(const SegParts &P, int q0, int t0, int t1, double (&out)[NQ], double *lds) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        double s = 0.0;
        const double *src = P.at(q0 + q);
        for (int t = t0 + threadIdx.x; t < t1; t += WG) s += src[t];
        out[q] = s;
    }
    block_sum<NQ>(out, lds);
    __shared__ double share[8];
    if (threadIdx.x == 0)
#pragma unroll
        for (int q = 0; q < NQ; ++q) share[q] = out[q];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NQ; ++q) out[q] = share[q];
    __syncthreads();
}

// This tile: its piece, the piece's tile range, the rows it covers and the vector offset.
struct Tile {
    int k, t0, t1, first, last, base, mrow, len;
    bool lead;
};
__device__ __forceinline__ Tile tile_of(const Pieces &P) {
    Tile t;
    t.k = P.tilePiece[blockIdx.x];
    t.t0 = P.tileOff[t.k];
    t.t1 = P.tileOff[t.k + 1];
    t.first = P.tileFirst[blockIdx.x];
    t.len = P.len[t.k];
    t.last = min(t.first + P.tileRows, t.len);
    t.base = P.base[t.k];
    t.mrow = P.mrow[t.k];
    t.lead = t.first == 0;
    return t;
}

#define PP_TILE_ROWS(i, t) for (int i = (t).first + (int)threadIdx.x; i < (t).last; i += WG)

// Entries of matrix row gr (ELL copy when present) with their vector columns.
template <typename T, typename F>
__device__ __forceinline__ void seg_row(const Csr<T> &A, int N, int gr, F f) {
    if (A.ell_w > 0) {
        for (int j = 0; j < A.ell_w; ++j) f(A.ell_col[(size_t)j * N + gr], A.ell_val[(size_t)j * N + gr]);
    } else {
        const int end = A.row[gr + 1];
        for (int e = A.row[gr]; e < end; ++e) f(A.col[e], A.val[e]);
    }
}

template <typename T>
__global__ void __launch_bounds__(WG) k_sum(const T *__restrict__ v, Pieces P, SegParts out) {
    __shared__ double lds[16];
    const Tile t = tile_of(P);
    double acc[1] = {0.0};
    PP_TILE_ROWS(i, t) acc[0] += double(v[t.base + i]);
    put<1>(out, acc, lds);
}

// r = bw - A x (bw = b minus the piece's mean with bSum), r0 = r for BiCGStab; partials
// (r, r), (bw, bw), sum(r).
template <typename T>
__global__ void __launch_bounds__(WG) k_open(Csr<T> A, const T *__restrict__ b, SegParts bSum, T *__restrict__ bwOut,
                                             const T *__restrict__ x, T *__restrict__ r, T *__restrict__ r0, Pieces P,
                                             SegParts out) {
    __shared__ double lds[16 * 3];
    const Tile t = tile_of(P);
    double mean[1] = {0.0};
    if (bSum.p) {
        fold<1>(bSum, 0, t.t0, t.t1, mean, lds);
        mean[0] /= double(t.len);
    }
    double acc[3] = {0.0, 0.0, 0.0};
    PP_TILE_ROWS(i, t) {
        T ax = T(0);
        seg_row(A, P.N, t.mrow + i, [&](int c, T a) { ax += a * x[t.base + (c - t.mrow)]; });
        const int o = t.base + i;
        const T bi = bSum.p ? b[o] - T(mean[0]) : b[o];
        if (bwOut) bwOut[o] = bi;
        const T ri = bi - ax;
        r[o] = ri;
        if (r0) r0[o] = ri;
        acc[0] += double(ri) * double(ri);
        acc[1] += double(bi) * double(bi);
        acc[2] += double(ri);
    }
    put<3>(out, acc, lds);
}

// CG: decide on the folded (r, r) (open: the initial residual), then p' = r + beta pIn (r when
// fresh) formed at the neighbours, q = A p', partials (p', q).
template <typename T>
__global__ void __launch_bounds__(WG) k_cg_turn(Csr<T> A, const T *__restrict__ r, const T *__restrict__ pIn,
                                                T *__restrict__ pOut, T *__restrict__ q, Pieces P, SegParts rrIn,
                                                SegParts pqOut, const SolveRec *in, SolveRec *out, int open) {
    __shared__ double lds[16 * 3];
    const Tile t = tile_of(P);
    SolveRec R = in[t.k];
    bool act = rec_active(R);
    double tot[2];
    fold<2>(rrIn, 0, t.t0, t.t1, tot, lds);
    T beta = T(0);
    const double scale = 1.0 / sqrt(double(t.len));
    if (act) {
        if (open) {
            rule_open(R, tot[0], scale, false);
            R.zero = tot[1] == 0.0;
        } else {
            rule_step(R, tot[0], scale, false);
            if (rec_active(R)) beta = T(rule_cg_beta(R, tot[0]));
        }
        act = rec_active(R);
    }
    if (t.lead && threadIdx.x == 0) out[t.k] = R;
    double acc[1] = {0.0};
    if (act) {
        const bool fresh = R.flags & R_FRESH;
        PP_TILE_ROWS(i, t) {
            T ap = T(0);
            seg_row(A, P.N, t.mrow + i, [&](int c, T a) {
                const int o = t.base + (c - t.mrow);
                ap += a * (fresh ? r[o] : r[o] + beta * pIn[o]);
            });
            const int o = t.base + i;
            const T pi = fresh ? r[o] : r[o] + beta * pIn[o];
            pOut[o] = pi;
            q[o] = ap;
            acc[0] += double(pi) * double(ap);
        }
    }
    put<1>(pqOut, acc, lds);
}

// CG: alpha from the folded (p', q); x += alpha p', r -= alpha q; partials (r, r), 0, sum(r);
// the lead tile commits the record.
template <typename T>
__global__ void __launch_bounds__(WG) k_cg_step(T *__restrict__ r, const T *__restrict__ p, const T *__restrict__ q,
                                                T *__restrict__ x, Pieces P, SegParts pqIn, SegParts rrOut,
                                                const SolveRec *in, SolveRec *out) {
    __shared__ double lds[16 * 3];
    const Tile t = tile_of(P);
    SolveRec R = in[t.k];
    bool act = rec_active(R);
    double pq[1];
    fold<1>(pqIn, 0, t.t0, t.t1, pq, lds);
    T alpha = T(0);
    if (act) {
        alpha = T(rule_alpha(R, pq[0]));
        R.flags &= ~R_FRESH;
        act = rec_active(R);
    }
    if (t.lead && threadIdx.x == 0) out[t.k] = R;
    double acc[3] = {0.0, 0.0, 0.0};
    if (act)
        PP_TILE_ROWS(i, t) {
            const int o = t.base + i;
            x[o] += alpha * p[o];
            const T ri = r[o] - alpha * q[o];
            r[o] = ri;
            acc[0] += double(ri) * double(ri);
            acc[2] += double(ri);
        }
    put<3>(rrOut, acc, lds);
}

// BiCGStab: decide (open: initial residual; otherwise the full step's (r0, r), (r, r)), then
// p = r + beta (p - omega v) row by row.
template <typename T>
__global__ void __launch_bounds__(WG) k_bi_dir(const T *__restrict__ r, T *__restrict__ p, const T *__restrict__ v,
                                               Pieces P, SegParts rrIn, const SolveRec *in, SolveRec *out, int open) {
    __shared__ double lds[16 * 2];
    const Tile t = tile_of(P);
    SolveRec R = in[t.k];
    bool act = rec_active(R);
    double tot[2];
    fold<2>(rrIn, 0, t.t0, t.t1, tot, lds);
    const double scale = 1.0 / sqrt(double(t.len));
    T beta = T(0);
    if (act) {
        if (open) {
            rule_open(R, tot[0], scale, true);
            R.zero = tot[1] == 0.0;
        } else {
            beta = T(rule_bicg_full(R, tot[0], tot[1], scale));
        }
        act = rec_active(R);
    }
    if (t.lead && threadIdx.x == 0) out[t.k] = R;
    if (!act) return;
    const bool fresh = R.flags & R_FRESH;
    const T omega = T(R.omega);
    PP_TILE_ROWS(i, t) {
        const int o = t.base + i;
        p[o] = fresh ? r[o] : r[o] + beta * (p[o] - omega * v[o]);
    }
}

// v = A p; partials (r0, v).
template <typename T>
__global__ void __launch_bounds__(WG) k_bi_apply(Csr<T> A, const T *__restrict__ p, T *__restrict__ v,
                                                 const T *__restrict__ r0, Pieces P, SegParts out, const SolveRec *in) {
    __shared__ double lds[16];
    const Tile t = tile_of(P);
    double acc[1] = {0.0};
    if (rec_active(in[t.k]))
        PP_TILE_ROWS(i, t) {
            T ap = T(0);
            seg_row(A, P.N, t.mrow + i, [&](int c, T a) { ap += a * p[t.base + (c - t.mrow)]; });
            const int o = t.base + i;
            v[o] = ap;
            acc[0] += double(r0[o]) * double(ap);
        }
    put<1>(out, acc, lds);
}

// alpha from the folded (r0, v); s = r - alpha v at the neighbours, t = A s; partials
// (t, s), (t, t), (s, s).
template <typename T>
__global__ void __launch_bounds__(WG) k_bi_stab(Csr<T> A, const T *__restrict__ r, const T *__restrict__ v,
                                                T *__restrict__ tv, Pieces P, SegParts r0vIn, SegParts out,
                                                const SolveRec *in) {
    __shared__ double lds[16 * 3];
    const Tile t = tile_of(P);
    SolveRec R = in[t.k];
    double r0v[1];
    fold<1>(r0vIn, 0, t.t0, t.t1, r0v, lds);
    T alpha = T(0);
    bool act = rec_active(R);
    if (act) {
        alpha = T(rule_alpha(R, r0v[0]));
        act = rec_active(R);
    }
    double acc[3] = {0.0, 0.0, 0.0};
    if (act)
        PP_TILE_ROWS(i, t) {
            T as = T(0);
            seg_row(A, P.N, t.mrow + i, [&](int c, T a) {
                const int o = t.base + (c - t.mrow);
                as += a * (r[o] - alpha * v[o]);
            });
            const int o = t.base + i;
            const T si = r[o] - alpha * v[o];
            tv[o] = as;
            acc[0] += double(as) * double(si);
            acc[1] += double(as) * double(as);
            acc[2] += double(si) * double(si);
        }
    put<3>(out, acc, lds);
}

// alpha, omega from the folded partials; x and r updated; partials (r0, r), (r, r); the lead
// tile commits the record.
template <typename T>
__global__ void __launch_bounds__(WG) k_bi_step(T *__restrict__ r, const T *__restrict__ r0, const T *__restrict__ p,
                                                const T *__restrict__ v, const T *__restrict__ tv, T *__restrict__ x,
                                                Pieces P, SegParts r0vIn, SegParts tsIn, SegParts rrOut,
                                                const SolveRec *in, SolveRec *out) {
    __shared__ double lds[16 * 3];
    const Tile t = tile_of(P);
    SolveRec R = in[t.k];
    bool act = rec_active(R);
    double r0v[1], ts[3];
    fold<1>(r0vIn, 0, t.t0, t.t1, r0v, lds);
    fold<3>(tsIn, 0, t.t0, t.t1, ts, lds);
    bool half = false;
    if (act) {
        R.alpha = rule_alpha(R, r0v[0]);
        if (rec_active(R)) R.omega = rule_bicg_half(R, ts[0], ts[1], ts[2], 1.0 / sqrt(double(t.len)));
        act = rec_active(R);
        half = act && (R.flags & R_HALF);
        if (half) rule_bicg_half_met(R);
        R.flags &= ~R_FRESH;
    }
    if (t.lead && threadIdx.x == 0) out[t.k] = R;
    double acc[2] = {0.0, 0.0};
    if (act) {
        const T alpha = T(R.alpha), omega = T(R.omega);
        PP_TILE_ROWS(i, t) {
            const int o = t.base + i;
            const T si = r[o] - alpha * v[o];
            if (half) {
                x[o] += alpha * p[o];
                r[o] = si;
            } else {
                x[o] += alpha * p[o] + omega * si;
                const T ri = si - omega * tv[o];
                r[o] = ri;
                acc[0] += double(r0[o]) * double(ri);
                acc[1] += double(ri) * double(ri);
            }
        }
    }
    put<2>(rrOut, acc, lds);
}

template <typename T>
__global__ void k_fill_pieces_zero(T *x, Pieces P, const SolveRec *rec) {
    const Tile t = tile_of(P);
    if (!rec[t.k].zero) return;
    PP_TILE_ROWS(i, t) x[t.base + i] = T(0);
}

} // namespace seg

// Host side of the segmented solve: tiles built from the pieces, uploaded with one copy.
struct SegLayout {
    std::vector<int32_t> host; // base[K] len[K] mrow[K] tileOff[K+1] tilePiece[T] tileFirst[T]
    int K = 0, T = 0;
};

static SegLayout seg_layout(const int32_t *segRows, int S, int ncomp, int N, int tileRows) {
    SegLayout L;
    L.K = S * ncomp;
    std::vector<int32_t> base(L.K), len(L.K), mrow(L.K), off(L.K + 1, 0), piece, first;
    for (int c = 0; c < ncomp; ++c)
        for (int s = 0; s < S; ++s) {
            const int k = c * S + s;
            base[k] = c * N + segRows[s];
            len[k] = segRows[s + 1] - segRows[s];
            mrow[k] = segRows[s];
            const int tiles = std::max(1, (len[k] + tileRows - 1) / tileRows);
            off[k + 1] = off[k] + tiles;
            for (int j = 0; j < tiles; ++j) {
                piece.push_back(k);
                first.push_back(j * tileRows);
            }
        }
    L.T = (int)piece.size();
    for (auto *v : {&base, &len, &mrow, &off, &piece, &first}) L.host.insert(L.host.end(), v->begin(), v->end());
    return L;
}


template <typename T>
static int seg_solve(const Csr<T> &A, const int32_t *segRows, int S, int ncomp, const T *b, T *x, int maxit,
                     const double *tols, int ntols, int useBicg, int project, int tileRows, hipStream_t s,
                     pp_solve_result *res) {
    const int N = A.n;
    SegLayout L = seg_layout(segRows, S, ncomp, N, tileRows);
    const int K = L.K, Tt = L.T;
    SegSpace &g_seg = my_spaces().seg;
    Space &sp = g_seg.sp;
    const size_t len = (size_t)ncomp * N;
    const int nvec = useBicg ? 5 : 5;
    const size_t partDoubles = (size_t)Tt * (useBicg ? (3 + 1 + 3) : (1 + 3 + 1));
    if (int e = prepare(sp, sizeof(T) * nvec * len, partDoubles, K, tols, ntols, s)) return e;
    // layout upload through a pinned staging copy (rewritten only after this solve's final wait)
    const size_t lbytes = sizeof(int32_t) * L.host.size();
    if (g_seg.layoutHostCap < lbytes) {
        if (g_seg.layoutHost) (void)hipHostFree(g_seg.layoutHost);
        g_seg.layoutHost = nullptr;
        PP_HIP(hipHostMalloc((void **)&g_seg.layoutHost, lbytes, hipHostMallocDefault));
        g_seg.layoutHostCap = lbytes;
    }
    if (int e = grow((void **)&g_seg.layoutDev, &g_seg.layoutCap, lbytes)) return e;
    std::memcpy(g_seg.layoutHost, L.host.data(), lbytes);
    PP_HIP(hipMemcpyAsync(g_seg.layoutDev, g_seg.layoutHost, lbytes, hipMemcpyHostToDevice, s));
    const int32_t *ld = g_seg.layoutDev;
    seg::Pieces P{ld, ld + K, ld + 2 * K, ld + 4 * K + 1, ld + 4 * K + 1 + Tt, ld + 3 * K, K, Tt, tileRows, N};
    // everything the iteration kernels are given, for the captured chunks
    uint64_t layoutHash = 1469598103934665603ull;
    for (int32_t v : L.host) layoutHash = (layoutHash ^ (uint32_t)v) * 1099511628211ull;
    const std::vector<uintptr_t> key = {(uintptr_t)A.row, (uintptr_t)A.col, (uintptr_t)A.val, (uintptr_t)A.ell_col,
                                        (uintptr_t)A.ell_val, (uintptr_t)A.ell_w, (uintptr_t)b, (uintptr_t)x,
                                        (uintptr_t)sp.vec, (uintptr_t)sp.parts, (uintptr_t)sp.recs, (uintptr_t)ld,
                                        (uintptr_t)layoutHash, (uintptr_t)K, (uintptr_t)Tt, (uintptr_t)N,
                                        (uintptr_t)tileRows, (uintptr_t)useBicg, (uintptr_t)project, sizeof(T)};
    T *vec = reinterpret_cast<T *>(sp.vec);
    auto V = [&](int k) { return vec + (size_t)k * len; };
    double *pd = reinterpret_cast<double *>(sp.parts);
    size_t used = 0;
    auto take = [&](int nq) { seg::SegParts q{pd + used, nq, Tt}; used += (size_t)nq * Tt; return q; };
    SolveRec *ra = recs_a(sp), *rb = recs_b(sp, K);
    hipLaunchKernelGGL(k_init_recs, dim3(1), dim3(256), 0, s, ra, K, tols[0], ntols > 1 ? (const double *)sp.tols : nullptr, maxit);
    const dim3 grid(Tt), block(WG);
    int last = 0;
    bool zero = false;
    auto opening = [&](const SolveRec *recs) {
        zero = false;
        for (int k = 0; k < K; ++k) zero |= recs[k].zero != 0;
        return false;
    };
    if (!useBicg) {
        T *r = V(0), *q = V(1), *p[2] = {V(2), V(3)}, *bw = project ? V(4) : const_cast<T *>(b);
        const seg::SegParts sSum = take(1), sRR = take(3), sPQ = take(1);
        if (project) hipLaunchKernelGGL((seg::k_sum<T>), grid, block, 0, s, b, P, sSum);
        hipLaunchKernelGGL((seg::k_open<T>), grid, block, 0, s, A, b, project ? sSum : seg::SegParts{nullptr, 1, Tt},
                           project ? bw : (T *)nullptr, x, r, (T *)nullptr, P, sRR);
        hipLaunchKernelGGL((seg::k_cg_turn<T>), grid, block, 0, s, A, r, p[0], p[1], q, P, sRR, sPQ, ra, rb, 1);
        PP_LAUNCH_CHECK();
        auto body = [&](hipStream_t st, int j) {
            const int cur = 1 ^ (j & 1); // the opening turn wrote p[1]
            PP_TIMED(PP_K_SEG_CG_UPDATE, N, st,
                     hipLaunchKernelGGL((seg::k_cg_step<T>), grid, block, 0, st, r, p[cur], q, x, P, sPQ, sRR, rb, ra));
            PP_TIMED(PP_K_SEG_CG_DIRECTION, N, st,
                     hipLaunchKernelGGL((seg::k_cg_turn<T>), grid, block, 0, st, A, r, p[cur], p[cur ^ 1], q, P, sRR, sPQ, ra,
                                        rb, 0));
        };
        if (int e = pace_solve(sp, s, rb, K, maxit + 2, Pace{g_chunk_cap, 0, true}, &last, opening, body, &key)) return e;
    } else {
        T *r = V(0), *r0 = V(1), *p = V(2), *v = V(3), *tv = V(4);
        const seg::SegParts sRR = take(3), sV = take(1), sT = take(3);
        hipLaunchKernelGGL((seg::k_open<T>), grid, block, 0, s, A, b, seg::SegParts{nullptr, 1, Tt}, (T *)nullptr, x, r, r0,
                           P, sRR);
        hipLaunchKernelGGL((seg::k_bi_dir<T>), grid, block, 0, s, r, p, v, P, sRR, ra, rb, 1);
        PP_LAUNCH_CHECK();
        auto body = [&](hipStream_t st, int) {
            PP_TIMED(PP_K_SEG_BICG, N, st,
                hipLaunchKernelGGL((seg::k_bi_apply<T>), grid, block, 0, st, A, p, v, r0, P, sV, rb);
                hipLaunchKernelGGL((seg::k_bi_stab<T>), grid, block, 0, st, A, r, v, tv, P, sV, sT, rb);
                hipLaunchKernelGGL((seg::k_bi_step<T>), grid, block, 0, st, r, r0, p, v, tv, x, P, sV, sT, sRR, rb, ra);
                hipLaunchKernelGGL((seg::k_bi_dir<T>), grid, block, 0, st, r, p, v, P, sRR, ra, rb, 0));
        };
        if (int e = pace_solve(sp, s, rb, K, maxit + 2, Pace{g_chunk_cap, 0, true}, &last, opening, body, &key)) return e;
    }
    // pieces whose right-hand side is exactly zero get x = 0 (converged, no iterations)
    const SolveRec *recs = slot_recs(sp, last, K);
    if (zero) hipLaunchKernelGGL((seg::k_fill_pieces_zero<T>), grid, block, 0, s, x, P, rb);
    PP_LAUNCH_CHECK();
    report(recs, K, 0, res);
    for (int k = 0; k < K; ++k)
        if (recs[k].zero) res[k] = pp_solve_result{0.0, 0, 1, 1, F_MET};
    return PP_OK;
}

// The body of pp_solve_linear and of its allocation-free variant (spaces: the CG and the
// BiCGStab workspaces; allowTranspose: the transposed-operator buffer may be allocated).
static int solve_linear(const pp_csr *A, const void *rhs, void *x, int32_t nrhs, int32_t dtype, int32_t maxit,
                        const double *tols, int32_t ntols, int32_t criterion, int32_t use_bicg,
                        int32_t matrix_rank_deficient, int32_t residual_reset_steps, int32_t transpose_a,
                        int32_t return_best, void *stream, pp_solve_result *results, Space &cgSpace, Space &bicgSpace,
                        bool allowTranspose) {
    PP_CHECK_ARG(A && A->row && A->col && A->val && rhs && x && tols && results, "null argument");
    PP_CHECK_ARG(A->n > 0 && nrhs > 0, "empty system");
    PP_CHECK_ARG(A->num_matrices >= 1 && nrhs % A->num_matrices == 0,
                 "number of right-hand sides must be a multiple of the number of matrices");
    PP_CHECK_ARG(ntols == 1 || ntols == nrhs, "tolerance must be a scalar or one value per right-hand side");
    PP_CHECK_ARG(criterion == PP_CONV_NORM2 || criterion == PP_CONV_NORM2_NORMALIZED, "unsupported convergence criterion");
    PP_CHECK_ARG(maxit >= 0, "maxit must be non-negative");
    PP_CHECK_ARG(!(use_bicg && matrix_rank_deficient), "rank-deficient mode is a CG mode");
    PP_CHECK_ARG(allowTranspose || !(transpose_a && use_bicg),
                 "transposed BiCGStab solves need a transposed copy of the matrix: use pp_solve_linear");
    hipStream_t s = as_stream(stream);
    const int rpm = nrhs / A->num_matrices;
    const double scale = criterion == PP_CONV_NORM2_NORMALIZED ? 1.0 / std::sqrt(double(A->n)) : 1.0;
    if (use_bicg) return_best = 0;
    const int jacobi = use_bicg ? g_bicg_jacobi : g_cg_jacobi;
    ThreadSpaces &ts = my_spaces();
    void *&tbuf = ts.tbuf;
    size_t &tcap = ts.tcap;
    int *&tbad = ts.tbad;
    PP_DISPATCH(dtype, {
        Csr<T> C = csr_view<T>(A, rpm);
        if (transpose_a && use_bicg) {
            C.ell_w = 0;
            const size_t bytes = sizeof(T) * (size_t)A->nnz * A->num_matrices;
            if (int e = grow(&tbuf, &tcap, bytes)) return e;
            if (!tbad) PP_HIP(hipMalloc((void **)&tbad, sizeof(int)));
            PP_HIP(hipMemsetAsync(tbad, 0, sizeof(int), s));
            if (int e = launch_transpose_symmetric<T>(A, reinterpret_cast<T *>(tbuf), tbad, s)) return e;
            int bad = 0;
            PP_HIP(hipMemcpyAsync(&bad, tbad, sizeof(int), hipMemcpyDeviceToHost, s));
            PP_HIP(hipStreamSynchronize(s));
            PP_CHECK_ARG(!bad, "transposed solves need a structurally symmetric pattern");
            C.val = reinterpret_cast<const T *>(tbuf);
        }
        const T *b = reinterpret_cast<const T *>(rhs);
        T *xx = reinterpret_cast<T *>(x);
        Space &sp = use_bicg ? bicgSpace : cgSpace;
        if (A->n <= SMALL_N)
            return small_solve<T>(C, b, xx, nrhs, maxit, tols, ntols, scale, use_bicg ? 0 : residual_reset_steps,
                                  return_best, matrix_rank_deficient, use_bicg, jacobi, sp, s, results);
        if (use_bicg) return bicg_large<T>(C, b, xx, nrhs, maxit, tols, ntols, scale, jacobi, sp, s, results);
        return cg_large<T>(C, b, xx, nrhs, maxit, tols, ntols, scale, residual_reset_steps, return_best,
                           matrix_rank_deficient, jacobi, sp, s, results);
    });
    return PP_OK;
}

// ---------------- the allocation-free variant ----------------
// Work vectors, partial sets, records and tolerances of pp_solve_linear_noalloc live in one
// static device array per GPU, carved into fixed regions at the first call on a device (its
// address from hipGetSymbolAddress, kept per device); host snapshots live in a static host
// array and the snapshot events are created at that set-up too.
constexpr size_t SCR_VEC = size_t(56) << 20, SCR_PARTS = size_t(6) << 20, SCR_RECS = size_t(1) << 20,
                 SCR_TOLS = size_t(64) << 10;
constexpr size_t SCRATCH_BYTES = SCR_VEC + SCR_PARTS + SCR_RECS + SCR_TOLS;
constexpr int SCRATCH_RHS = 2048;     // right-hand sides per call at most
constexpr int SCRATCH_DEVICES = 16;   // devices of one process with a scratch table entry
static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];
static Space g_fixed[SCRATCH_DEVICES];
static SolveRec g_fixed_snap[2 * SCRATCH_RHS];

static int fixed_space(Space **out) {
    int dev = 0;
    PP_HIP(hipGetDevice(&dev));
    PP_CHECK_ARG(dev >= 0 && dev < SCRATCH_DEVICES, "device index beyond the scratch table");
    Space &sp = g_fixed[dev];
    if (!sp.fixed) {
        unsigned char *base = nullptr;
        PP_HIP(hipGetSymbolAddress((void **)&base, HIP_SYMBOL(scratch)));
        PP_HIP(hipEventCreateWithFlags(&sp.ev[0], hipEventDisableTiming));
        PP_HIP(hipEventCreateWithFlags(&sp.ev[1], hipEventDisableTiming));
        sp.vec = base;                                    sp.vecCap = SCR_VEC;
        sp.parts = base + SCR_VEC;                        sp.partsCap = SCR_PARTS;
        sp.recs = base + SCR_VEC + SCR_PARTS;             sp.recsCap = SCR_RECS;
        sp.tols = base + SCR_VEC + SCR_PARTS + SCR_RECS;  sp.tolsCap = SCR_TOLS;
        sp.snap = reinterpret_cast<char *>(g_fixed_snap); sp.snapCap = sizeof(g_fixed_snap);
        sp.fixed = true;
    }
    *out = &sp;
    return PP_OK;
}

} // namespace pp

using namespace pp;

extern "C" {

int pp_solve_linear(const pp_csr *A, const void *rhs, void *x, int32_t nrhs, int32_t dtype, int32_t maxit,
                    const double *tols, int32_t ntols, int32_t criterion, int32_t use_bicg, int32_t matrix_rank_deficient,
                    int32_t residual_reset_steps, int32_t transpose_a, int32_t return_best, void *stream,
                    pp_solve_result *results) {
    ThreadSpaces &ts = my_spaces();
    return solve_linear(A, rhs, x, nrhs, dtype, maxit, tols, ntols, criterion, use_bicg, matrix_rank_deficient,
                        residual_reset_steps, transpose_a, return_best, stream, results, ts.cg, ts.bicg, true);
}

int pp_solve_linear_noalloc(const pp_csr *A, const void *rhs, void *x, int32_t nrhs, int32_t dtype, int32_t maxit,
                            const double *tols, int32_t ntols, int32_t criterion, int32_t use_bicg,
                            int32_t matrix_rank_deficient, int32_t residual_reset_steps, int32_t transpose_a,
                            int32_t return_best, void *stream, pp_solve_result *results) {
    std::lock_guard<std::mutex> lock(g_lock);
    Space *sp = nullptr;
    if (int e = fixed_space(&sp)) return e;
    return solve_linear(A, rhs, x, nrhs, dtype, maxit, tols, ntols, criterion, use_bicg, matrix_rank_deficient,
                        residual_reset_steps, transpose_a, return_best, stream, results, *sp, *sp, false);
}

int pp_set_krylov_settings(int32_t check_interval, int32_t use_graphs, int32_t cg_jacobi, int32_t bicg_jacobi) {
    PP_CHECK_ARG(check_interval >= 1, "checkInterval must be positive");
    std::lock_guard<std::mutex> lock(g_lock);
    g_chunk_cap = check_interval;
    if (!use_graphs && g_use_graphs) {
        std::lock_guard<std::mutex> glock(g_graph_lock);
        drop_graphs();
    }
    g_use_graphs = use_graphs ? 1 : 0;
    g_cg_jacobi = cg_jacobi ? 1 : 0;
    g_bicg_jacobi = bicg_jacobi ? 1 : 0;
    return PP_OK;
}

int pp_get_krylov_settings(int32_t *check_interval, int32_t *use_graphs, int32_t *cg_jacobi, int32_t *bicg_jacobi) {
    std::lock_guard<std::mutex> lock(g_lock);
    if (check_interval) *check_interval = g_chunk_cap;
    if (use_graphs) *use_graphs = g_use_graphs;
    if (cg_jacobi) *cg_jacobi = g_cg_jacobi;
    if (bicg_jacobi) *bicg_jacobi = g_bicg_jacobi;
    return PP_OK;
}

int pp_solve_segmented(const pp_csr *A, const int32_t *seg_rows, int32_t nseg, int32_t ncomp, const void *rhs, void *x,
                       int32_t dtype, int32_t maxit, const double *tols, int32_t ntols, int32_t use_bicg,
                       int32_t project, int32_t tile_rows, void *stream, pp_solve_result *results) {
    PP_CHECK_ARG(A && A->row && A->col && A->val && seg_rows && rhs && x && tols && results, "null argument");
    PP_CHECK_ARG(A->num_matrices == 1, "a segmented solve takes one block-diagonal matrix");
    PP_CHECK_ARG(nseg >= 1 && ncomp >= 1, "need at least one instance and one component");
    PP_CHECK_ARG(seg_rows[0] == 0 && seg_rows[nseg] == A->n, "instance rows must cover the matrix rows");
    for (int s = 0; s < nseg; ++s) PP_CHECK_ARG(seg_rows[s + 1] > seg_rows[s], "instance row ranges must be non-empty and ascending");
    PP_CHECK_ARG(ntols == 1 || ntols == nseg * ncomp, "tolerance must be a scalar or one value per instance and component");
    PP_CHECK_ARG(maxit >= 0, "maxit must be non-negative");
    PP_CHECK_ARG(!(use_bicg && project), "the projected (rank-deficient) mode is a CG mode");
    if (tile_rows <= 0) tile_rows = 1024;
    PP_CHECK_ARG(tile_rows % WG == 0 && tile_rows <= (1 << 16), "tile_rows must be a multiple of 256, at most 65536");
    
    PP_DISPATCH(dtype, {
        const Csr<T> C = csr_view<T>(A, 1);
        return seg_solve<T>(C, seg_rows, nseg, ncomp, reinterpret_cast<const T *>(rhs), reinterpret_cast<T *>(x), maxit,
                            tols, ntols, use_bicg, project, tile_rows, as_stream(stream), results);
    });
    return PP_OK;
}

int pp_amg_pcg_solve(const pp_csr *A, const pp_amg_level *levels, int32_t num_levels, const void *coarse_pinv, int32_t nc,
                     int32_t coarse_batched, int32_t hdtype, const void *rhs, void *x, int32_t nrhs, int32_t dtype,
                     int32_t maxit, double tol, int32_t normalized, int32_t return_best, int32_t project_constant,
                     int32_t pre_sweeps, int32_t post_sweeps, double omega, const double *tols, int32_t ntols, void *stream,
                     pp_solve_result *results) {
    PP_CHECK_ARG(A && levels && coarse_pinv && rhs && x && results, "null argument");
    PP_CHECK_ARG(num_levels >= 1, "AMG hierarchy needs at least one level above the coarsest");
    PP_CHECK_ARG(levels[0].n == A->n, "finest AMG level does not match the system matrix");
    PP_CHECK_ARG(nc > 0, "empty coarse level");
    PP_CHECK_ARG(A->num_matrices == 1 || A->num_matrices == nrhs, "system matrix must be shared or one per right-hand side");
    for (int l = 0; l + 1 < num_levels; ++l) PP_CHECK_ARG(levels[l].n_coarse == levels[l + 1].n, "inconsistent level sizes");
    auto sell_ok = [](const pp_sell &sd) {
        const bool any = sd.slice_off || sd.slice_steps || sd.col || sd.val;
        const bool lanesOk = sd.lanes == 1 || sd.lanes == 2 || sd.lanes == 4 || sd.lanes == 8 || sd.lanes == 16 || sd.lanes == 32;
        return !any || (sd.slice_off && sd.slice_steps && sd.col && sd.val && lanesOk && sd.total > 0);
    };
    for (int l = 0; l < num_levels; ++l) {
        const pp_amg_level &v = levels[l];
        PP_CHECK_ARG(v.a_ell_width == 0 || (l == 0 && v.a_ell_width > 0 && v.a_ell_col && v.a_ell_val),
                     "an ELL copy is accepted for the finest level only and needs a positive width, columns and values");
        PP_CHECK_ARG(sell_ok(v.a_sell) && sell_ok(v.r_sell) && sell_ok(v.p_sell),
                     "sliced ELL copy needs offsets, steps, columns, values, lanes in {1,2,...,32} and its size");
        PP_CHECK_ARG(v.a_val || v.a_sell.slice_off || (l == 0 && v.a_ell_width > 0),
                     "level values missing (a level needs CSR values or a sliced / finest-level ELL copy)");
        PP_CHECK_ARG(v.a_row && v.a_col && v.r_row && v.r_col && v.r_val && v.p_row && v.p_col && v.p_val && v.dinv,
                     "level CSR arrays and inverse diagonal are required");
        PP_CHECK_ARG(l == 0 || v.n > TAIL_ROWS || v.a_val, "levels of the coarse tail need CSR values");
    }
    PP_CHECK_ARG(!levels[0].a_sell.slice_off, "the finest level runs on its ELL copy or its CSR arrays, not a sliced copy");
    PP_CHECK_ARG(levels[num_levels - 1].n_coarse == nc, "coarsest level size mismatch");
    PP_CHECK_ARG(ntols == 0 || ntols == 1 || ntols == nrhs, "tolerance count");
    
    hipStream_t s = as_stream(stream);
    const double scale = normalized ? 1.0 / std::sqrt(double(A->n)) : 1.0;
    double t1 = tol;
    const double *tp = ntols > 1 ? tols : &t1;
    const int nt = ntols > 1 ? ntols : 1;
#define PP_AMG_CALL(TT, HH)                                                                                            \
    return amg_pcg<TT, HH>(csr_view<TT>(A, 1), levels, num_levels, reinterpret_cast<const HH *>(coarse_pinv), nc,      \
                           coarse_batched, reinterpret_cast<const TT *>(rhs), reinterpret_cast<TT *>(x), nrhs, maxit, tp, \
                           nt, scale, return_best, project_constant, pre_sweeps, post_sweeps, omega, s, results)
    if (dtype == PP_F64 && hdtype == PP_F32) PP_AMG_CALL(double, float);
    if (dtype == PP_F64 && hdtype == PP_F64) PP_AMG_CALL(double, double);
    if (dtype == PP_F32 && hdtype == PP_F32) PP_AMG_CALL(float, float);
#undef PP_AMG_CALL
    set_error("pp_amg_pcg_solve: unsupported (dtype, hierarchy dtype) combination");
    return PP_ERR_UNSUPPORTED;
}

int pp_release_workspaces(void) {
    // frees every thread's workspaces: no solve may be running
    {
        std::lock_guard<std::mutex> glock(g_graph_lock);
        drop_graphs();
    }
    std::lock_guard<std::mutex> lock(g_lock);
    for (ThreadSpaces *ts : g_all_spaces) {
        if (ts->seg.layoutDev) (void)hipFree(ts->seg.layoutDev);
        if (ts->seg.layoutHost) (void)hipHostFree(ts->seg.layoutHost);
        ts->seg.layoutDev = nullptr; ts->seg.layoutHost = nullptr; ts->seg.layoutCap = ts->seg.layoutHostCap = 0;
        if (ts->tbuf) (void)hipFree(ts->tbuf);
        if (ts->tbad) (void)hipFree(ts->tbad);
        ts->tbuf = nullptr; ts->tbad = nullptr; ts->tcap = 0;
        for (Space *spp : {&ts->cg, &ts->bicg, &ts->amg, &ts->seg.sp}) {
            Space &sp = *spp;
            for (void *p : {sp.vec, sp.parts, sp.recs, sp.aux, sp.tols, sp.tailDev})
                if (p) (void)hipFree(p);
            if (sp.tailHost) (void)hipHostFree(sp.tailHost);
            if (sp.snap) (void)hipHostFree(sp.snap);
            sp = Space();
        }
    }
    return PP_OK;
}

int pp_sell_probe(const pp_sell *M, int32_t rows, int32_t cols, int32_t hdtype, int32_t reps, void *stream, double *us) {
    PP_CHECK_ARG(M && M->slice_off && M->slice_steps && M->col && M->val && us && rows > 0 && cols > 0, "bad arguments");
    PP_CHECK_ARG(M->lanes == 1 || M->lanes == 2 || M->lanes == 4 || M->lanes == 8 || M->lanes == 16 || M->lanes == 32,
                 "lanes must be a power of two <= 32");
    std::lock_guard<std::mutex> lock(g_lock);
    if (hdtype == PP_F32) return probe_sell<float>(*M, rows, cols, std::max(1, reps), as_stream(stream), us);
    if (hdtype == PP_F64) return probe_sell<double>(*M, rows, cols, std::max(1, reps), as_stream(stream), us);
    set_error("pp_sell_probe: dtype");
    return PP_ERR_ARG;
}

} // extern "C"

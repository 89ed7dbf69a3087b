// Fused Krylov solvers for gfx950 (CG, BiCGStab, AMG-preconditioned CG).
//
// Recurrences, stopping rules and reported iteration counts restate the reference's fused
// solvers (csrc/phipict/solvers/krylov/krylov.cu): kResidual :291-366, kCgDirection
// :427-498, kCgUpdate :501-568, kCgDots :571-593, kBicgP/V/T/Update :598-824, the V-cycle
// kernels :1247-1385 and the drivers :1058-1186, :1389-1546. Reductions accumulate in
// double, as there.
//
// Two execution engines:
//  * small systems (n <= SMALL_N): one 1024-thread workgroup per group of right-hand sides
//    runs the whole solve on-chip -- every reduction is an LDS tree inside the workgroup,
//    so a solve is one launch with no host round trip. Many groups run concurrently.
//  * large systems: one kernel per iteration phase over a multi-workgroup grid; partial
//    sums per workgroup are combined in a fixed order by the last workgroup to finish
//    (deterministic), which also updates the per-RHS scalars on the device. The host
//    launches chunks of iterations and polls a device "all done" flag one chunk behind.
#include "pp_common.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

namespace pp {

enum DoneStatus : int { RUNNING = 0, CONVERGED = 1, NONFINITE = 2, MAXIT = 3, RISING = 4 };
constexpr int RISING_CUTOFF = 100;

struct BatchState {
    double tol, rho, alpha, beta, omega, crit, best, last, halfCrit;
    int it, maxit, bestIt, rising, done, restart, saveBest, halfDone, usedIt, pad;
};

struct Control {
    unsigned int counter;
    int allDone;
    int zeroRhs; // every right-hand side is exactly zero: x = 0 without iterating
    int pad;
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
    // sliced ELL copy (pp_sell layout; AMG levels only): used when sell_off is set
    const int32_t *__restrict__ sell_off;
    const int32_t *__restrict__ sell_w;
    const int32_t *__restrict__ sell_col;
    const T *__restrict__ sell_val;
};

// Matrix view of a descriptor; the ELL copy is used when present.
template <typename T>
static Csr<T> csr_view(const pp_csr *A, int rpm) {
    Csr<T> C{A->row, A->col, reinterpret_cast<const T *>(A->val), A->n, A->num_matrices > 1 ? A->nnz : 0, rpm, 0,
             nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
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

// Sum vals[NV] over the workgroup; result valid in thread 0. sm: >= NV * (blockDim/64) doubles.
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

__device__ __forceinline__ bool finite_d(double v) { return isfinite(v); }

// ======================================================================================
// Small-system engine: one workgroup per group of <= NB right-hand sides sharing a matrix.
// ======================================================================================
constexpr int SMALL_BLOCK = 1024;
constexpr int SMALL_N = 1 << 15;

struct SmallArgs {
    int n, B, nb_per_group, rpm, maxit, normalized, returnBest, project, resetEvery, bicg;
    double normScale;
};

// Work vectors per RHS: [k][n] for k in {r, p0, p1, q, xbest, bw, r0, v, t}
template <typename T>
struct SmallWork {
    T *base;
    int n;
    __device__ T *vec(int rhs, int k) const { return base + ((size_t)rhs * 9 + k) * n; }
};

template <typename T, int NB>
__device__ __forceinline__ void small_spmv_gather(const Csr<T> &A, const T *val, int i, const T *const *src, T *acc,
                                                  int nb) {
    for (int k = 0; k < NB; ++k) acc[k] = T(0);
    const int e = A.row[i + 1];
    for (int j = A.row[i]; j < e; ++j) {
        const int c = A.col[j];
        const T a = val[j];
#pragma unroll
        for (int k = 0; k < NB; ++k)
            if (k < nb) acc[k] += a * src[k][c];
    }
}

// CG, one workgroup per group (krylov.cu:291-568 in workgroup form).
template <typename T, int NB>
__global__ void __launch_bounds__(SMALL_BLOCK) k_cg_small(Csr<T> A, const T *__restrict__ b, T *__restrict__ x,
                                                         SmallWork<T> W, BatchState *st, SmallArgs a) {
    __shared__ double sm[16 * 2 * NB];
    __shared__ int s_done[NB];
    const int gpm = (a.rpm + NB - 1) / NB;
    const int m = blockIdx.x / gpm;
    const int j = blockIdx.x - m * gpm;
    const int b0 = m * a.rpm + j * NB;
    const int nb = max(0, min(min(NB, a.rpm - j * NB), a.B - b0));
    if (nb <= 0) return;
    const int n = a.n;
    const T *val = A.val + (size_t)m * A.valStride;
    const int tid = threadIdx.x;

    T *r[NB], *pa[NB], *pb[NB], *q[NB], *xb[NB], *bw[NB];
#pragma unroll
    for (int k = 0; k < NB; ++k) {
        const int rr = b0 + (k < nb ? k : 0);
        r[k] = W.vec(rr, 0); pa[k] = W.vec(rr, 1); pb[k] = W.vec(rr, 2); q[k] = W.vec(rr, 3);
        xb[k] = W.vec(rr, 4); bw[k] = W.vec(rr, 5);
    }
    // b (optionally mean-projected) into bw
    double mean[NB];
#pragma unroll
    for (int k = 0; k < NB; ++k) mean[k] = 0.0;
    if (a.project) {
        for (int i = tid; i < n; i += blockDim.x)
#pragma unroll
            for (int k = 0; k < NB; ++k)
                if (k < nb) mean[k] += double(b[(size_t)(b0 + k) * n + i]);
        block_sum<NB>(mean, sm);
        if (tid == 0)
#pragma unroll
            for (int k = 0; k < NB; ++k) sm[2 * 16 * NB - 1 - k] = mean[k] / double(n);
        __syncthreads();
#pragma unroll
        for (int k = 0; k < NB; ++k) mean[k] = sm[2 * 16 * NB - 1 - k];
        __syncthreads();
    }
    for (int i = tid; i < n; i += blockDim.x)
#pragma unroll
        for (int k = 0; k < NB; ++k)
            if (k < nb) {
                const T bi = b[(size_t)(b0 + k) * n + i];
                bw[k][i] = a.project ? bi - T(mean[k]) : bi;
            }
    __syncthreads();
    // initial residual (kResidual, initial call)
    {
        double acc[NB];
#pragma unroll
        for (int k = 0; k < NB; ++k) acc[k] = 0.0;
        const T *xs[NB];
#pragma unroll
        for (int k = 0; k < NB; ++k) xs[k] = x + (size_t)(b0 + (k < nb ? k : 0)) * n;
        for (int i = tid; i < n; i += blockDim.x) {
            T s[NB];
            small_spmv_gather<T, NB>(A, val, i, xs, s, nb);
#pragma unroll
            for (int k = 0; k < NB; ++k)
                if (k < nb) {
                    const T ri = bw[k][i] - s[k];
                    r[k][i] = ri;
                    acc[k] += double(ri) * double(ri);
                }
        }
        block_sum<NB>(acc, sm);
        if (tid == 0) {
            for (int k = 0; k < nb; ++k) {
                BatchState &s = st[b0 + k];
                s.rho = acc[k];
                s.restart = 1;
                const double crit = sqrt(acc[k]) * a.normScale;
                s.crit = crit;
                if (!finite_d(crit)) { s.done = NONFINITE; s.usedIt = 0; }
                else if (crit == 0.0) { s.done = CONVERGED; s.usedIt = 0; }
            }
        }
        __syncthreads();
    }
    int parity = 0;
    for (int iter = 0; iter < a.maxit + 2; ++iter) {
        if (tid < NB) s_done[tid] = tid < nb ? st[b0 + tid].done : 1;
        __syncthreads();
        bool active[NB], anyActive = false;
#pragma unroll
        for (int k = 0; k < NB; ++k) { active[k] = k < nb && s_done[k] == RUNNING; anyActive |= active[k]; }
        __syncthreads();
        if (!anyActive) break;
        T *xk[NB];
#pragma unroll
        for (int k = 0; k < NB; ++k) xk[k] = x + (size_t)(b0 + (k < nb ? k : 0)) * n;
        // residual replacement every resetEvery iterations
        if (a.resetEvery > 0) {
            bool doReset[NB], any = false;
#pragma unroll
            for (int k = 0; k < NB; ++k) {
                const BatchState &s = st[b0 + (k < nb ? k : 0)];
                doReset[k] = active[k] && s.it > 0 && (s.it % a.resetEvery) == 0;
                any |= doReset[k];
            }
            if (any) {
                double acc[NB];
#pragma unroll
                for (int k = 0; k < NB; ++k) acc[k] = 0.0;
                for (int i = tid; i < n; i += blockDim.x) {
                    T s[NB];
                    small_spmv_gather<T, NB>(A, val, i, (const T *const *)xk, s, nb);
#pragma unroll
                    for (int k = 0; k < NB; ++k)
                        if (doReset[k]) {
                            const T ri = bw[k][i] - s[k];
                            r[k][i] = ri;
                            acc[k] += double(ri) * double(ri);
                        }
                }
                block_sum<NB>(acc, sm);
                if (tid == 0)
                    for (int k = 0; k < nb; ++k)
                        if (doReset[k]) { st[b0 + k].rho = acc[k]; st[b0 + k].restart = 1; }
                __syncthreads();
            }
        }
        T *pin[NB], *pout[NB];
#pragma unroll
        for (int k = 0; k < NB; ++k) { pin[k] = parity ? pb[k] : pa[k]; pout[k] = parity ? pa[k] : pb[k]; }
        // direction: p' = r + beta p (restart: r); q = A p'; (p', q)
        {
            bool restart[NB];
            T beta[NB];
#pragma unroll
            for (int k = 0; k < NB; ++k) {
                const BatchState &s = st[b0 + (k < nb ? k : 0)];
                restart[k] = active[k] && s.restart;
                beta[k] = active[k] ? T(s.beta) : T(0);
            }
            double acc[NB];
#pragma unroll
            for (int k = 0; k < NB; ++k) acc[k] = 0.0;
            for (int i = tid; i < n; i += blockDim.x) {
                T sum[NB];
#pragma unroll
                for (int k = 0; k < NB; ++k) sum[k] = T(0);
                const int e = A.row[i + 1];
                for (int jj = A.row[i]; jj < e; ++jj) {
                    const int c = A.col[jj];
                    const T av = val[jj];
#pragma unroll
                    for (int k = 0; k < NB; ++k)
                        if (active[k]) {
                            const T zc = r[k][c];
                            sum[k] += av * (restart[k] ? zc : zc + beta[k] * pin[k][c]);
                        }
                }
#pragma unroll
                for (int k = 0; k < NB; ++k)
                    if (active[k]) {
                        const T zi = r[k][i];
                        const T pi = restart[k] ? zi : zi + beta[k] * pin[k][i];
                        pout[k][i] = pi;
                        q[k][i] = sum[k];
                        acc[k] += double(pi) * double(sum[k]);
                    }
            }
            block_sum<NB>(acc, sm);
            if (tid == 0)
                for (int k = 0; k < nb; ++k) {
                    if (!active[k]) continue;
                    BatchState &s = st[b0 + k];
                    const double alpha = s.rho / acc[k];
                    s.restart = 0;
                    if (!finite_d(alpha)) { s.done = NONFINITE; s.usedIt = s.it; }
                    else s.alpha = alpha;
                }
            __syncthreads();
        }
        // update: x += alpha p'; r -= alpha q; (r, r)
        {
            T alpha[NB];
            bool act2[NB], saveBest[NB];
#pragma unroll
            for (int k = 0; k < NB; ++k) {
                const BatchState &s = st[b0 + (k < nb ? k : 0)];
                act2[k] = active[k] && s.done == RUNNING;
                alpha[k] = T(s.alpha);
                saveBest[k] = act2[k] && s.saveBest && a.returnBest;
            }
            double acc[NB];
#pragma unroll
            for (int k = 0; k < NB; ++k) acc[k] = 0.0;
            for (int i = tid; i < n; i += blockDim.x) {
#pragma unroll
                for (int k = 0; k < NB; ++k)
                    if (act2[k]) {
                        const T xi = xk[k][i];
                        if (saveBest[k]) xb[k][i] = xi;
                        xk[k][i] = xi + alpha[k] * pout[k][i];
                        const T ri = r[k][i] - alpha[k] * q[k][i];
                        r[k][i] = ri;
                        acc[k] += double(ri) * double(ri);
                    }
            }
            block_sum<NB>(acc, sm);
            if (tid == 0)
                for (int k = 0; k < nb; ++k) {
                    if (!act2[k]) continue;
                    BatchState &s = st[b0 + k];
                    const double crit = sqrt(acc[k]) * a.normScale;
                    ++s.it;
                    s.crit = crit;
                    s.saveBest = 0;
                    s.usedIt = s.it - 1;
                    if (!finite_d(crit)) { s.done = NONFINITE; continue; }
                    if (a.returnBest) {
                        if (crit < s.best) { s.best = crit; s.bestIt = s.it - 1; s.saveBest = 1; }
                        if (s.it > 1 && crit >= s.last) ++s.rising; else s.rising = 0;
                        s.last = crit;
                    }
                    if (crit < s.tol) s.done = CONVERGED;
                    else if (s.it >= s.maxit) s.done = MAXIT;
                    else if (a.returnBest && s.rising >= RISING_CUTOFF) s.done = RISING;
                    else {
                        const double gamma = acc[k];
                        s.beta = gamma / s.rho;
                        s.rho = gamma;
                        if (!finite_d(s.beta)) s.done = NONFINITE;
                    }
                }
            __syncthreads();
        }
        parity ^= 1;
    }
    // best iterate restore (kRestoreBest)
    if (a.returnBest) {
#pragma unroll
        for (int k = 0; k < NB; ++k) {
            if (k >= nb) break;
            const BatchState &s = st[b0 + k];
            if ((s.done == MAXIT || s.done == RISING) && s.bestIt >= 0 && s.best < s.crit) {
                T *xk = x + (size_t)(b0 + k) * n;
                for (int i = tid; i < n; i += blockDim.x) xk[i] = xb[k][i];
            }
        }
    }
}

// BiCGStab, one workgroup per group (krylov.cu:598-824 in workgroup form).
template <typename T, int NB>
__global__ void __launch_bounds__(SMALL_BLOCK) k_bicg_small(Csr<T> A, const T *__restrict__ b, T *__restrict__ x,
                                                           SmallWork<T> W, BatchState *st, SmallArgs a) {
    __shared__ double sm[16 * 3 * NB];
    __shared__ int s_done[NB];
    const int gpm = (a.rpm + NB - 1) / NB;
    const int m = blockIdx.x / gpm;
    const int j = blockIdx.x - m * gpm;
    const int b0 = m * a.rpm + j * NB;
    const int nb = max(0, min(min(NB, a.rpm - j * NB), a.B - b0));
    if (nb <= 0) return;
    const int n = a.n;
    const T *val = A.val + (size_t)m * A.valStride;
    const int tid = threadIdx.x;
    T *r[NB], *r0[NB], *p[NB], *v[NB], *t[NB], *xk[NB];
#pragma unroll
    for (int k = 0; k < NB; ++k) {
        const int rr = b0 + (k < nb ? k : 0);
        r[k] = W.vec(rr, 0); p[k] = W.vec(rr, 1); v[k] = W.vec(rr, 3); r0[k] = W.vec(rr, 6); t[k] = W.vec(rr, 7);
        xk[k] = x + (size_t)rr * n;
    }
    // initial residual with shadow residual (kResidual with r0)
    {
        double acc[NB];
#pragma unroll
        for (int k = 0; k < NB; ++k) acc[k] = 0.0;
        for (int i = tid; i < n; i += blockDim.x) {
            T s[NB];
            small_spmv_gather<T, NB>(A, val, i, (const T *const *)xk, s, nb);
#pragma unroll
            for (int k = 0; k < NB; ++k)
                if (k < nb) {
                    const T ri = b[(size_t)(b0 + k) * n + i] - s[k];
                    r[k][i] = ri;
                    r0[k][i] = ri;
                    p[k][i] = T(0);
                    v[k][i] = T(0);
                    acc[k] += double(ri) * double(ri);
                }
        }
        block_sum<NB>(acc, sm);
        if (tid == 0)
            for (int k = 0; k < nb; ++k) {
                BatchState &s = st[b0 + k];
                s.rho = acc[k];
                s.restart = 1;
                const double crit = sqrt(acc[k]) * a.normScale;
                s.crit = crit;
                if (!finite_d(crit)) { s.done = NONFINITE; s.usedIt = 0; }
                else if (crit < s.tol) { s.done = CONVERGED; s.usedIt = -1; }
            }
        __syncthreads();
    }
    for (int iter = 0; iter < a.maxit + 2; ++iter) {
        if (tid < NB) s_done[tid] = tid < nb ? st[b0 + tid].done : 1;
        __syncthreads();
        bool active[NB], anyActive = false;
#pragma unroll
        for (int k = 0; k < NB; ++k) { active[k] = k < nb && s_done[k] == RUNNING; anyActive |= active[k]; }
        __syncthreads();
        if (!anyActive) break;
        // p = r + beta (p - omega v)
        {
            bool restart[NB];
            T beta[NB], omega[NB];
#pragma unroll
            for (int k = 0; k < NB; ++k) {
                const BatchState &s = st[b0 + (k < nb ? k : 0)];
                restart[k] = s.restart; beta[k] = T(s.beta); omega[k] = T(s.omega);
            }
            for (int i = tid; i < n; i += blockDim.x)
#pragma unroll
                for (int k = 0; k < NB; ++k)
                    if (active[k]) p[k][i] = restart[k] ? r[k][i] : r[k][i] + beta[k] * (p[k][i] - omega[k] * v[k][i]);
            __syncthreads();
        }
        // v = A p; (r0, v)
        {
            double acc[NB];
#pragma unroll
            for (int k = 0; k < NB; ++k) acc[k] = 0.0;
            for (int i = tid; i < n; i += blockDim.x) {
                T sum[NB];
#pragma unroll
                for (int k = 0; k < NB; ++k) sum[k] = T(0);
                const int e = A.row[i + 1];
                for (int jj = A.row[i]; jj < e; ++jj) {
                    const int c = A.col[jj];
                    const T av = val[jj];
#pragma unroll
                    for (int k = 0; k < NB; ++k)
                        if (active[k]) sum[k] += av * p[k][c];
                }
#pragma unroll
                for (int k = 0; k < NB; ++k)
                    if (active[k]) { v[k][i] = sum[k]; acc[k] += double(r0[k][i]) * double(sum[k]); }
            }
            block_sum<NB>(acc, sm);
            if (tid == 0)
                for (int k = 0; k < nb; ++k) {
                    if (!active[k]) continue;
                    BatchState &s = st[b0 + k];
                    const double alpha = s.rho / acc[k];
                    s.restart = 0;
                    if (!finite_d(alpha)) { s.done = NONFINITE; s.usedIt = s.it; }
                    else s.alpha = alpha;
                }
            __syncthreads();
        }
        // t = A s, s = r - alpha v; (t,s), (t,t), (s,s)
        bool act2[NB];
        {
            T alpha[NB];
#pragma unroll
            for (int k = 0; k < NB; ++k) {
                const BatchState &s = st[b0 + (k < nb ? k : 0)];
                act2[k] = active[k] && s.done == RUNNING;
                alpha[k] = T(s.alpha);
            }
            double acc[NB * 3];
#pragma unroll
            for (int k = 0; k < NB * 3; ++k) acc[k] = 0.0;
            for (int i = tid; i < n; i += blockDim.x) {
                T sum[NB];
#pragma unroll
                for (int k = 0; k < NB; ++k) sum[k] = T(0);
                const int e = A.row[i + 1];
                for (int jj = A.row[i]; jj < e; ++jj) {
                    const int c = A.col[jj];
                    const T av = val[jj];
#pragma unroll
                    for (int k = 0; k < NB; ++k)
                        if (act2[k]) sum[k] += av * (r[k][c] - alpha[k] * v[k][c]);
                }
#pragma unroll
                for (int k = 0; k < NB; ++k)
                    if (act2[k]) {
                        const T si = r[k][i] - alpha[k] * v[k][i];
                        t[k][i] = sum[k];
                        acc[k * 3 + 0] += double(sum[k]) * double(si);
                        acc[k * 3 + 1] += double(sum[k]) * double(sum[k]);
                        acc[k * 3 + 2] += double(si) * double(si);
                    }
            }
            block_sum<NB * 3>(acc, sm);
            if (tid == 0)
                for (int k = 0; k < nb; ++k) {
                    if (!act2[k]) continue;
                    BatchState &s = st[b0 + k];
                    const double critS = sqrt(acc[k * 3 + 2]) * a.normScale;
                    s.halfCrit = critS;
                    if (!finite_d(critS)) { s.done = NONFINITE; s.crit = critS; s.usedIt = s.it; }
                    else if (critS < s.tol) { s.halfDone = 1; s.omega = 0.0; }
                    else {
                        const double omega = acc[k * 3 + 0] / acc[k * 3 + 1];
                        if (!finite_d(omega)) { s.done = NONFINITE; s.crit = critS; s.usedIt = s.it; }
                        else s.omega = omega;
                    }
                }
            __syncthreads();
        }
        // x += alpha p + omega s; r = s - omega t; (r0, r), (r, r)
        {
            T alpha[NB], omega[NB];
            bool act3[NB], half[NB];
#pragma unroll
            for (int k = 0; k < NB; ++k) {
                const BatchState &s = st[b0 + (k < nb ? k : 0)];
                act3[k] = act2[k] && s.done == RUNNING;
                alpha[k] = T(s.alpha); omega[k] = T(s.omega); half[k] = s.halfDone;
            }
            double acc[NB * 2];
#pragma unroll
            for (int k = 0; k < NB * 2; ++k) acc[k] = 0.0;
            for (int i = tid; i < n; i += blockDim.x)
#pragma unroll
                for (int k = 0; k < NB; ++k)
                    if (act3[k]) {
                        const T si = r[k][i] - alpha[k] * v[k][i];
                        if (half[k]) {
                            xk[k][i] += alpha[k] * p[k][i] * T(1);
                            r[k][i] = si;
                        } else {
                            xk[k][i] += (alpha[k] * p[k][i] + omega[k] * si) * T(1);
                            const T ri = si - omega[k] * t[k][i];
                            r[k][i] = ri;
                            acc[k * 2 + 0] += double(r0[k][i]) * double(ri);
                            acc[k * 2 + 1] += double(ri) * double(ri);
                        }
                    }
            block_sum<NB * 2>(acc, sm);
            if (tid == 0)
                for (int k = 0; k < nb; ++k) {
                    if (!act3[k]) continue;
                    BatchState &s = st[b0 + k];
                    if (s.halfDone) { s.done = CONVERGED; s.crit = s.halfCrit; s.usedIt = s.it; continue; }
                    const double crit = sqrt(acc[k * 2 + 1]) * a.normScale;
                    s.crit = crit;
                    if (!finite_d(crit)) { s.done = NONFINITE; s.usedIt = s.it; }
                    else if (crit < s.tol) { s.done = CONVERGED; s.usedIt = s.it + 1; }
                    else {
                        ++s.it;
                        s.usedIt = s.it - 1;
                        const double rhoNew = acc[k * 2 + 0];
                        const double beta = (rhoNew / s.rho) * (s.alpha / s.omega);
                        s.rho = rhoNew;
                        s.beta = beta;
                        if (!finite_d(beta)) s.done = NONFINITE;
                        else if (s.it >= s.maxit) s.done = MAXIT;
                    }
                }
            __syncthreads();
        }
    }
}

// ======================================================================================
// Large-system engine: kernel per phase, last-workgroup reductions.
// ======================================================================================
constexpr int BLOCK = 256;

// Rows of a workgroup: a grid-wide sweep (workgroup b takes rows b*BLOCK + k*G*BLOCK). Giving
// each XCD a contiguous span of rows (so that stencil neighbours hit its own L2) measured
// slower: the sweep's DRAM locality matters more than the cross-XCD reads, which the
// Infinity Cache serves.
#define FOR_BLOCK_ROWS(i, n) for (int i = blockIdx.x * BLOCK + (int)threadIdx.x; i < (n); i += gridDim.x * BLOCK)

// Entries (column, value) of row i of A, from its ELL copy or its CSR arrays; with EW > 0 the
// ELL width is a compile-time constant and all column and value loads are issued up front.
template <int EW, typename T, typename F>
__device__ __forceinline__ void for_row_entries(const Csr<T> &A, const T *__restrict__ ellv, const T *__restrict__ val,
                                                int n, int i, F f) {
    if constexpr (EW > 0) {
        int cs[EW];
        T as[EW];
#pragma unroll
        for (int j = 0; j < EW; ++j) {
            cs[j] = A.ell_col[(size_t)j * n + i];
            as[j] = ellv[(size_t)j * n + i];
        }
#pragma unroll
        for (int j = 0; j < EW; ++j) f(cs[j], as[j]);
    } else {
        const int ew = A.ell_w;
        const int jb = ew ? 0 : A.row[i], je = ew ? ew : A.row[i + 1];
        for (int j = jb; j < je; ++j) {
            const size_t eo = (size_t)j * n + i;
            f(ew ? A.ell_col[eo] : A.col[j], ew ? ellv[eo] : val[j]);
        }
    }
}

// Per-workgroup partial sums, written with plain stores. The one-workgroup finish kernel that
// follows on the same stream combines them in a fixed order (deterministic); the kernel
// boundary makes them visible across XCDs. (A last-workgroup-done scheme would need a
// device-scope release fence per workgroup, which on gfx950 writes back the XCD's L2.)
template <int NB, int NQ>
__device__ __forceinline__ void store_partials(double (&acc)[NB][NQ], int b0, int nb, double *partial, double *sm) {
    double flat[NB * NQ];
#pragma unroll
    for (int k = 0; k < NB; ++k)
#pragma unroll
        for (int q = 0; q < NQ; ++q) flat[k * NQ + q] = acc[k][q];
    block_sum<NB * NQ>(flat, sm);
    if (threadIdx.x == 0) {
        const int G = gridDim.x;
        for (int k = 0; k < nb; ++k)
            for (int q = 0; q < NQ; ++q) partial[((size_t)(b0 + k) * G + blockIdx.x) * NQ + q] = flat[k * NQ + q];
    }
}

template <int NQ>
__device__ __forceinline__ void reduce_batch(const double *partial, int b, int G, double (&out)[NQ], double *sm) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) out[q] = 0.0;
    for (int g = threadIdx.x; g < G; g += blockDim.x) {
        const double *pp = partial + ((size_t)b * G + g) * NQ;
#pragma unroll
        for (int q = 0; q < NQ; ++q) out[q] += pp[q];
    }
    block_sum<NQ>(out, sm);
}

__device__ __forceinline__ void finish_last(BatchState *st, int B, Control *ctrl) {
    if (threadIdx.x == 0) {
        int nd = 0;
        for (int b = 0; b < B; ++b) nd += st[b].done != RUNNING;
        ctrl->allDone = nd == B;
    }
}

struct Group { int b0, nb, m; };
template <int NB>
__device__ __forceinline__ Group batch_group(int B, int rpm) {
    const int gpm = (rpm + NB - 1) / NB;
    Group g;
    g.m = blockIdx.y / gpm;
    const int j = blockIdx.y - g.m * gpm;
    g.b0 = g.m * rpm + j * NB;
    g.nb = max(0, min(min(NB, rpm - j * NB), B - g.b0));
    return g;
}

__global__ void k_init_state(BatchState *st, int B, double tol, int maxit, Control *ctrl, const double *tols) {
    for (int b = threadIdx.x; b < B; b += blockDim.x) {
        BatchState s = {};
        s.tol = tols ? tols[b] : tol;
        s.maxit = maxit;
        s.best = INFINITY;
        s.restart = 1;
        s.bestIt = -1;
        st[b] = s;
    }
    if (threadIdx.x == 0) { ctrl->counter = 0; ctrl->allDone = 0; ctrl->zeroRhs = 0; }
}

// r = b - A x (+ r0 = r for BiCGStab); reset mode only acts at multiples of resetEvery.
template <typename T, int NB, int EW = 0>
__global__ void __launch_bounds__(BLOCK) k_residual(Csr<T> A, const T *__restrict__ b, const T *__restrict__ x,
                                                    T *__restrict__ r, T *__restrict__ r0, BatchState *st, int B,
                                                    double *partial, Control *ctrl, double normScale, int resetEvery) {
    __shared__ double sm[16 * NB * 2];
    if (ctrl->allDone) return;
    const int n = A.n;
    const Group grp = batch_group<NB>(B, A.rpm);
    auto isActive = [&](const BatchState &s) {
        return s.done == RUNNING && (resetEvery <= 0 || (s.it > 0 && (s.it % resetEvery) == 0));
    };
    bool active[NB], any = false;
#pragma unroll
    for (int k = 0; k < NB; ++k) { active[k] = k < grp.nb && isActive(st[grp.b0 + k]); any |= active[k]; }
    double acc[NB][2];
#pragma unroll
    for (int k = 0; k < NB; ++k) acc[k][0] = acc[k][1] = 0.0;
    if (any) {
        const T *val = A.val + (size_t)grp.m * A.valStride;
        const T *ellv = A.ell_val + (A.valStride ? (size_t)grp.m * A.ell_w * n : 0);
        FOR_BLOCK_ROWS(i, n) {
            T sum[NB];
#pragma unroll
            for (int k = 0; k < NB; ++k) sum[k] = T(0);
            for_row_entries<EW>(A, ellv, val, n, i, [&](int c, T a) {
#pragma unroll
                for (int k = 0; k < NB; ++k)
                    if (active[k]) sum[k] += a * x[(size_t)(grp.b0 + k) * n + c];
            });
#pragma unroll
            for (int k = 0; k < NB; ++k)
                if (active[k]) {
                    const size_t o = (size_t)(grp.b0 + k) * n + i;
                    const T bo = b[o];
                    const T ri = bo - sum[k];
                    r[o] = ri;
                    if (r0) r0[o] = ri;
                    acc[k][0] += double(ri) * double(ri);
                    acc[k][1] += double(bo) * double(bo);
                }
        }
    }
    store_partials<NB, 2>(acc, grp.b0, grp.nb, partial, sm);
}

// Finish of k_residual: rho, restart and (initial call) the converged-start test. In the
// initial call with zeroShortcut, right-hand sides that are all exactly zero finish at once
// and ctrl->zeroRhs asks for x = 0 (the reference's zero-RHS shortcut, piso_diff.py:867).
__global__ void __launch_bounds__(BLOCK) k_fin_residual(BatchState *st, int B, const double *partial, int G,
                                                        Control *ctrl, double normScale, int resetEvery, int bicg,
                                                        int zeroShortcut = 0) {
    __shared__ double sm[16 * 2];
    if (ctrl->allDone) return;
    if (zeroShortcut && resetEvery <= 0) {
        __shared__ int allZero; // reductions are valid in thread 0 only: broadcast the verdict
        if (threadIdx.x == 0) allZero = 1;
        for (int bb = 0; bb < B; ++bb) {
            double red[2];
            reduce_batch<2>(partial, bb, G, red, sm);
            if (threadIdx.x == 0 && red[1] != 0.0) allZero = 0;
        }
        __syncthreads();
        if (allZero) {
            if (threadIdx.x == 0) {
                for (int bb = 0; bb < B; ++bb) {
                    st[bb].done = CONVERGED;
                    st[bb].crit = 0.0;
                    st[bb].usedIt = 0;
                }
                ctrl->zeroRhs = 1;
                ctrl->allDone = 1;
            }
            return;
        }
    }
    for (int bb = 0; bb < B; ++bb) {
        const BatchState &sb = st[bb];
        if (!(sb.done == RUNNING && (resetEvery <= 0 || (sb.it > 0 && (sb.it % resetEvery) == 0)))) continue;
        double red[2];
        reduce_batch<2>(partial, bb, G, red, sm);
        if (threadIdx.x == 0) {
            BatchState &s = st[bb];
            s.rho = red[0];
            s.restart = 1;
            if (resetEvery <= 0) {
                const double crit = sqrt(red[0]) * normScale;
                s.crit = crit;
                if (!finite_d(crit)) { s.done = NONFINITE; s.usedIt = 0; }
                else if (bicg ? crit < s.tol : crit == 0.0) {
                    s.done = CONVERGED;
                    s.usedIt = bicg ? -1 : 0;
                }
            }
        }
        __syncthreads();
    }
    finish_last(st, B, ctrl);
}

// PM_VECTOR: z given; PM_VECTOR_SHIFTED: z = u - mean(u) formed from u as it is read
enum PrecondMode { PM_NONE = 0, PM_VECTOR = 2, PM_VECTOR_SHIFTED = 3 };

// CG direction: p' = z + beta p (z = r or the preconditioned vector u); q = A p'; (p', q).
template <typename T, int NB, int PM, int EW = 0>
__global__ void __launch_bounds__(BLOCK) k_cg_direction(Csr<T> A, const T *__restrict__ r, const T *__restrict__ z,
                                                        const double *__restrict__ zShift, const T *__restrict__ pIn,
                                                        T *__restrict__ pOut, T *__restrict__ q, BatchState *st, int B,
                                                        double *partial, Control *ctrl) {
    __shared__ double sm[16 * NB];
    if (ctrl->allDone) return;
    const int n = A.n;
    const Group grp = batch_group<NB>(B, A.rpm);
    bool active[NB], restart[NB];
    T beta[NB], shift[NB];
    bool any = false;
#pragma unroll
    for (int k = 0; k < NB; ++k) {
        active[k] = k < grp.nb && st[grp.b0 + k].done == RUNNING;
        restart[k] = active[k] && st[grp.b0 + k].restart;
        beta[k] = active[k] ? T(st[grp.b0 + k].beta) : T(0);
        shift[k] = (active[k] && PM == PM_VECTOR_SHIFTED) ? T(zShift[grp.b0 + k]) : T(0);
        any |= active[k];
    }
    // z = u - mean(u) (the projected preconditioner output) when zShift is given
    const T *zz = PM == PM_NONE ? r : z;
    double acc[NB][1];
#pragma unroll
    for (int k = 0; k < NB; ++k) acc[k][0] = 0.0;
    if (any) {
        const T *val = A.val + (size_t)grp.m * A.valStride;
        const T *ellv = A.ell_val + (A.valStride ? (size_t)grp.m * A.ell_w * n : 0);
        const int ew = A.ell_w;
        FOR_BLOCK_ROWS(i, n) {
            T sum[NB];
#pragma unroll
            for (int k = 0; k < NB; ++k) sum[k] = T(0);
            auto entry = [&](int c, T a) {
#pragma unroll
                for (int k = 0; k < NB; ++k)
                    if (active[k]) {
                        const size_t o = (size_t)(grp.b0 + k) * n + c;
                        const T zc = PM == PM_VECTOR_SHIFTED ? zz[o] - shift[k] : zz[o];
                        sum[k] += a * (restart[k] ? zc : zc + beta[k] * pIn[o]);
                    }
            };
            if constexpr (EW > 0) {
                // ELL rows of a compile-time width: all column and value loads issued up front
                int cs[EW];
                T as[EW];
#pragma unroll
                for (int j = 0; j < EW; ++j) {
                    cs[j] = A.ell_col[(size_t)j * n + i];
                    as[j] = ellv[(size_t)j * n + i];
                }
#pragma unroll
                for (int j = 0; j < EW; ++j) entry(cs[j], as[j]);
            } else {
                const int jb = ew ? 0 : A.row[i], je = ew ? ew : A.row[i + 1];
                for (int j = jb; j < je; ++j) {
                    const size_t eo = (size_t)j * n + i;
                    entry(ew ? A.ell_col[eo] : A.col[j], ew ? ellv[eo] : val[j]);
                }
            }
#pragma unroll
            for (int k = 0; k < NB; ++k)
                if (active[k]) {
                    const size_t o = (size_t)(grp.b0 + k) * n + i;
                    const T zi = PM == PM_VECTOR_SHIFTED ? zz[o] - shift[k] : zz[o];
                    const T pi = restart[k] ? zi : zi + beta[k] * pIn[o];
                    pOut[o] = pi;
                    q[o] = sum[k];
                    acc[k][0] += double(pi) * double(sum[k]);
                }
        }
    }
    store_partials<NB, 1>(acc, grp.b0, grp.nb, partial, sm);
}

// Finish of a direction phase (CG and BiCGStab's v = A p): alpha = rho / (d, A p).
__global__ void __launch_bounds__(BLOCK) k_fin_alpha(BatchState *st, int B, const double *partial, int G,
                                                     Control *ctrl) {
    __shared__ double sm[16];
    if (ctrl->allDone) return;
    for (int bb = 0; bb < B; ++bb) {
        if (st[bb].done != RUNNING) continue;
        double red[1];
        reduce_batch<1>(partial, bb, G, red, sm);
        if (threadIdx.x == 0) {
            BatchState &s = st[bb];
            const double alpha = s.rho / red[0];
            s.restart = 0;
            if (!finite_d(alpha)) { s.done = NONFINITE; s.usedIt = s.it; }
            else s.alpha = alpha;
        }
        __syncthreads();
    }
    finish_last(st, B, ctrl);
}

// x += alpha p; r -= alpha q; (r, r). Last workgroup: convergence and beta.
template <typename T, int NB, int PM>
__global__ void __launch_bounds__(BLOCK) k_cg_update(int n, int rpm, T *__restrict__ r, const T *__restrict__ p,
                                                     const T *__restrict__ q, T *__restrict__ x, T *__restrict__ xbest,
                                                     BatchState *st, int B, double *partial, Control *ctrl,
                                                     double normScale, int returnBest, double *sumPartial) {
    // sumPartial (optional): per-workgroup sums of the new r (its mean for the next projection)
    __shared__ double sm[16 * NB];
    if (ctrl->allDone) return;
    const Group grp = batch_group<NB>(B, rpm);
    double acc[NB][1], rsum[NB][1];
#pragma unroll
    for (int k = 0; k < NB; ++k) acc[k][0] = rsum[k][0] = 0.0;
#pragma unroll
    for (int k = 0; k < NB; ++k) {
        if (k >= grp.nb) break;
        const BatchState &bs = st[grp.b0 + k];
        if (bs.done != RUNNING) continue;
        const T alpha = T(bs.alpha);
        const bool saveBest = bs.saveBest && xbest != nullptr;
        const size_t off = (size_t)(grp.b0 + k) * n;
        FOR_BLOCK_ROWS(i, n) {
            const size_t o = off + i;
            const T xi = x[o];
            if (saveBest) xbest[o] = xi;
            x[o] = xi + alpha * p[o];
            const T ri = r[o] - alpha * q[o];
            r[o] = ri;
            acc[k][0] += double(ri) * double(ri);
            rsum[k][0] += double(ri);
        }
    }
    store_partials<NB, 1>(acc, grp.b0, grp.nb, partial, sm);
    if (sumPartial) store_partials<NB, 1>(rsum, grp.b0, grp.nb, sumPartial, sm);
}

// Finish of k_cg_update: iteration count, convergence, best iterate, beta (unpreconditioned).
template <int PM>
__global__ void __launch_bounds__(BLOCK) k_fin_cg_update(BatchState *st, int B, const double *partial, int G,
                                                         Control *ctrl, double normScale, int returnBest,
                                                         const double *sumPartial, int n, double *meanOut) {
    // with sumPartial: also meanOut[b] = mean of the new r (for the next projected V-cycle)
    __shared__ double sm[16];
    if (ctrl->allDone) return;
    {
        for (int bb = 0; bb < B; ++bb) {
            if (st[bb].done != RUNNING) continue;
            if (sumPartial) {
                double rs[1];
                reduce_batch<1>(sumPartial, bb, G, rs, sm);
                if (threadIdx.x == 0) meanOut[bb] = rs[0] / double(n);
            }
            double red[1];
            reduce_batch<1>(partial, bb, G, red, sm);
            if (threadIdx.x == 0) {
                BatchState &s = st[bb];
                const double crit = sqrt(red[0]) * normScale;
                ++s.it;
                s.crit = crit;
                s.saveBest = 0;
                s.usedIt = s.it - 1;
                if (!finite_d(crit)) s.done = NONFINITE;
                else {
                    if (returnBest) {
                        if (crit < s.best) { s.best = crit; s.bestIt = s.it - 1; s.saveBest = 1; }
                        if (s.it > 1 && crit >= s.last) ++s.rising; else s.rising = 0;
                        s.last = crit;
                    }
                    if (crit < s.tol) s.done = CONVERGED;
                    else if (s.it >= s.maxit) s.done = MAXIT;
                    else if (returnBest && s.rising >= RISING_CUTOFF) s.done = RISING;
                    else if (PM == PM_NONE) {
                        const double gamma = red[0];
                        s.beta = gamma / s.rho;
                        s.rho = gamma;
                        if (!finite_d(s.beta)) s.done = NONFINITE;
                    }
                }
            }
            __syncthreads();
        }
        finish_last(st, B, ctrl);
    }
}

// (r, z) for an explicit preconditioned vector z; last workgroup: rho and beta (kCgDots).
template <typename T, bool SHIFT>
__global__ void __launch_bounds__(BLOCK) k_cg_dots(int n, const T *__restrict__ r, const T *__restrict__ z,
                                                   const double *__restrict__ zShift, BatchState *st, int B,
                                                   double *partial, Control *ctrl) {
    // (r, z); with SHIFT z = u - mean(u), zShift holding the means (projected preconditioner output)
    __shared__ double sm[16];
    if (ctrl->allDone) return;
    const int b = blockIdx.y;
    double acc[1][1] = {{0.0}};
    if (st[b].done == RUNNING) {
        const T sh = SHIFT ? T(zShift[b]) : T(0);
        FOR_BLOCK_ROWS(i, n) {
            const T zi = SHIFT ? z[(size_t)b * n + i] - sh : z[(size_t)b * n + i];
            acc[0][0] += double(r[(size_t)b * n + i]) * double(zi);
        }
    }
    store_partials<1, 1>(acc, b, 1, partial, sm);
}

// Finish of k_cg_dots: rho = (r, z) and beta.
__global__ void __launch_bounds__(BLOCK) k_fin_cg_dots(BatchState *st, int B, const double *partial, int G,
                                                       Control *ctrl) {
    __shared__ double sm[16];
    if (ctrl->allDone) return;
    {
        for (int bb = 0; bb < B; ++bb) {
            if (st[bb].done != RUNNING) continue;
            double red[1];
            reduce_batch<1>(partial, bb, G, red, sm);
            if (threadIdx.x == 0) {
                BatchState &s = st[bb];
                if (!s.restart) s.beta = red[0] / s.rho;
                s.rho = red[0];
                if (!finite_d(s.rho) || !finite_d(s.beta)) { s.done = NONFINITE; s.usedIt = s.it; }
            }
            __syncthreads();
        }
        finish_last(st, B, ctrl);
    }
}

template <typename T>
__global__ void __launch_bounds__(BLOCK) k_bicg_p(int n, const T *__restrict__ r, T *__restrict__ p,
                                                  const T *__restrict__ v, const BatchState *st) {
    const int b = blockIdx.y;
    const BatchState &bs = st[b];
    if (bs.done != RUNNING) return;
    const bool restart = bs.restart;
    const T beta = T(bs.beta), omega = T(bs.omega);
    const size_t off = (size_t)b * n;
    // the first (restart) direction is r itself: p and v are not read before they are written
    if (restart)
        FOR_BLOCK_ROWS(i, n) p[off + i] = r[off + i];
    else
        FOR_BLOCK_ROWS(i, n) {
            const size_t o = off + i;
            p[o] = r[o] + beta * (p[o] - omega * v[o]);
        }
}

template <typename T, int NB, int EW = 0>
__global__ void __launch_bounds__(BLOCK) k_bicg_v(Csr<T> A, const T *__restrict__ p, T *__restrict__ v,
                                                  const T *__restrict__ r0, BatchState *st, int B, double *partial,
                                                  Control *ctrl) {
    __shared__ double sm[16 * NB];
    if (ctrl->allDone) return;
    const int n = A.n;
    const Group grp = batch_group<NB>(B, A.rpm);
    bool active[NB], any = false;
#pragma unroll
    for (int k = 0; k < NB; ++k) { active[k] = k < grp.nb && st[grp.b0 + k].done == RUNNING; any |= active[k]; }
    double acc[NB][1];
#pragma unroll
    for (int k = 0; k < NB; ++k) acc[k][0] = 0.0;
    if (any) {
        const T *val = A.val + (size_t)grp.m * A.valStride;
        const T *ellv = A.ell_val + (A.valStride ? (size_t)grp.m * A.ell_w * n : 0);
        FOR_BLOCK_ROWS(i, n) {
            T sum[NB];
#pragma unroll
            for (int k = 0; k < NB; ++k) sum[k] = T(0);
            for_row_entries<EW>(A, ellv, val, n, i, [&](int c, T a) {
#pragma unroll
                for (int k = 0; k < NB; ++k)
                    if (active[k]) sum[k] += a * p[(size_t)(grp.b0 + k) * n + c];
            });
#pragma unroll
            for (int k = 0; k < NB; ++k)
                if (active[k]) {
                    const size_t o = (size_t)(grp.b0 + k) * n + i;
                    v[o] = sum[k];
                    acc[k][0] += double(r0[o]) * double(sum[k]);
                }
        }
    }
    store_partials<NB, 1>(acc, grp.b0, grp.nb, partial, sm);
}

template <typename T, int NB, int EW = 0>
__global__ void __launch_bounds__(BLOCK) k_bicg_t(Csr<T> A, const T *__restrict__ r, const T *__restrict__ v,
                                                  T *__restrict__ t, BatchState *st, int B, double *partial,
                                                  Control *ctrl, double normScale) {
    __shared__ double sm[16 * NB * 3];
    if (ctrl->allDone) return;
    const int n = A.n;
    const Group grp = batch_group<NB>(B, A.rpm);
    bool active[NB], any = false;
    T alpha[NB];
#pragma unroll
    for (int k = 0; k < NB; ++k) {
        active[k] = k < grp.nb && st[grp.b0 + k].done == RUNNING;
        alpha[k] = active[k] ? T(st[grp.b0 + k].alpha) : T(0);
        any |= active[k];
    }
    double acc[NB][3];
#pragma unroll
    for (int k = 0; k < NB; ++k) acc[k][0] = acc[k][1] = acc[k][2] = 0.0;
    if (any) {
        const T *val = A.val + (size_t)grp.m * A.valStride;
        const T *ellv = A.ell_val + (A.valStride ? (size_t)grp.m * A.ell_w * n : 0);
        FOR_BLOCK_ROWS(i, n) {
            T sum[NB];
#pragma unroll
            for (int k = 0; k < NB; ++k) sum[k] = T(0);
            for_row_entries<EW>(A, ellv, val, n, i, [&](int c, T a) {
#pragma unroll
                for (int k = 0; k < NB; ++k)
                    if (active[k]) {
                        const size_t o = (size_t)(grp.b0 + k) * n + c;
                        sum[k] += a * (r[o] - alpha[k] * v[o]);
                    }
            });
#pragma unroll
            for (int k = 0; k < NB; ++k)
                if (active[k]) {
                    const size_t o = (size_t)(grp.b0 + k) * n + i;
                    const T si = r[o] - alpha[k] * v[o];
                    t[o] = sum[k];
                    acc[k][0] += double(sum[k]) * double(si);
                    acc[k][1] += double(sum[k]) * double(sum[k]);
                    acc[k][2] += double(si) * double(si);
                }
        }
    }
    store_partials<NB, 3>(acc, grp.b0, grp.nb, partial, sm);
}

// Finish of k_bicg_t: half-step convergence and omega.
__global__ void __launch_bounds__(BLOCK) k_fin_bicg_t(BatchState *st, int B, const double *partial, int G,
                                                      Control *ctrl, double normScale) {
    __shared__ double sm[16 * 3];
    if (ctrl->allDone) return;
    {
        for (int bb = 0; bb < B; ++bb) {
            if (st[bb].done != RUNNING) continue;
            double red[3];
            reduce_batch<3>(partial, bb, G, red, sm);
            if (threadIdx.x == 0) {
                BatchState &s = st[bb];
                const double critS = sqrt(red[2]) * normScale;
                s.halfCrit = critS;
                if (!finite_d(critS)) { s.done = NONFINITE; s.crit = critS; s.usedIt = s.it; }
                else if (critS < s.tol) { s.halfDone = 1; s.omega = 0.0; }
                else {
                    const double omega = red[0] / red[1];
                    if (!finite_d(omega)) { s.done = NONFINITE; s.crit = critS; s.usedIt = s.it; }
                    else s.omega = omega;
                }
            }
            __syncthreads();
        }
        finish_last(st, B, ctrl);
    }
}

template <typename T, int NB>
__global__ void __launch_bounds__(BLOCK) k_bicg_update(int n, int rpm, T *__restrict__ r, const T *__restrict__ r0,
                                                       const T *__restrict__ p, const T *__restrict__ v,
                                                       const T *__restrict__ t, T *__restrict__ x, BatchState *st,
                                                       int B, double *partial, Control *ctrl, double normScale) {
    __shared__ double sm[16 * NB * 2];
    if (ctrl->allDone) return;
    const Group grp = batch_group<NB>(B, rpm);
    double acc[NB][2];
#pragma unroll
    for (int k = 0; k < NB; ++k) acc[k][0] = acc[k][1] = 0.0;
#pragma unroll
    for (int k = 0; k < NB; ++k) {
        if (k >= grp.nb) break;
        const BatchState &bs = st[grp.b0 + k];
        if (bs.done != RUNNING) continue;
        const T alpha = T(bs.alpha), omega = T(bs.omega);
        const bool half = bs.halfDone;
        const size_t off = (size_t)(grp.b0 + k) * n;
        FOR_BLOCK_ROWS(i, n) {
            const size_t o = off + i;
            const T si = r[o] - alpha * v[o];
            if (half) {
                x[o] += alpha * p[o] * T(1);
                r[o] = si;
            } else {
                x[o] += (alpha * p[o] + omega * si) * T(1);
                const T ri = si - omega * t[o];
                r[o] = ri;
                acc[k][0] += double(r0[o]) * double(ri);
                acc[k][1] += double(ri) * double(ri);
            }
        }
    }
    store_partials<NB, 2>(acc, grp.b0, grp.nb, partial, sm);
}

// Finish of k_bicg_update: convergence, iteration count and beta.
__global__ void __launch_bounds__(BLOCK) k_fin_bicg_update(BatchState *st, int B, const double *partial, int G,
                                                           Control *ctrl, double normScale) {
    __shared__ double sm[16 * 2];
    if (ctrl->allDone) return;
    {
        for (int bb = 0; bb < B; ++bb) {
            if (st[bb].done != RUNNING) continue;
            double red[2];
            reduce_batch<2>(partial, bb, G, red, sm);
            if (threadIdx.x == 0) {
                BatchState &s = st[bb];
                if (s.halfDone) {
                    s.done = CONVERGED;
                    s.crit = s.halfCrit;
                    s.usedIt = s.it;
                } else {
                    const double crit = sqrt(red[1]) * normScale;
                    s.crit = crit;
                    if (!finite_d(crit)) { s.done = NONFINITE; s.usedIt = s.it; }
                    else if (crit < s.tol) { s.done = CONVERGED; s.usedIt = s.it + 1; }
                    else {
                        ++s.it;
                        s.usedIt = s.it - 1;
                        const double rhoNew = red[0];
                        const double beta = (rhoNew / s.rho) * (s.alpha / s.omega);
                        s.rho = rhoNew;
                        s.beta = beta;
                        if (!finite_d(beta)) s.done = NONFINITE;
                        else if (s.it >= s.maxit) s.done = MAXIT;
                    }
                }
            }
            __syncthreads();
        }
        finish_last(st, B, ctrl);
    }
}

template <typename T>
__global__ void k_zero_if_zero_rhs(T *x, size_t count, const Control *ctrl) {
    if (!ctrl->zeroRhs) return;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x)
        x[i] = T(0);
}

template <typename T>
__global__ void k_restore_best(T *x, const T *xbest, const BatchState *st, int n, int B) {
    for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < (size_t)n * B; idx += (size_t)gridDim.x * blockDim.x) {
        const BatchState &s = st[idx / n];
        if ((s.done == MAXIT || s.done == RISING) && s.bestIt >= 0 && s.best < s.crit) x[idx] = xbest[idx];
    }
}

// mean per batch into means[b] (kMean); only batches still running are counted
template <typename T>
__global__ void __launch_bounds__(BLOCK) k_mean(const T *__restrict__ v, int n, int B, double *means, double *partial,
                                                Control *ctrl, const BatchState *st, int onlyRunning) {
    __shared__ double sm[16];
    if (onlyRunning && ctrl->allDone) return;
    const int batch = blockIdx.y;
    double acc[1][1] = {{0.0}};
    if (!onlyRunning || st[batch].done == RUNNING)
        FOR_BLOCK_ROWS(i, n) acc[0][0] += double(v[(size_t)batch * n + i]);
    store_partials<1, 1>(acc, batch, 1, partial, sm);
}

// Finish of k_mean: means[b] = sum / n.
__global__ void __launch_bounds__(BLOCK) k_fin_mean(const double *partial, int G, int n, int B, double *means,
                                                    const Control *ctrl, int onlyRunning) {
    __shared__ double sm[16];
    if (onlyRunning && ctrl->allDone) return;
    for (int b = 0; b < B; ++b) {
        double red[1];
        reduce_batch<1>(partial, b, G, red, sm);
        if (threadIdx.x == 0) means[b] = red[0] / double(n);
        __syncthreads();
    }
}

template <typename T>
__global__ void k_subtract_mean(T *__restrict__ v, int n, const double *means, const BatchState *st, int onlyRunning) {
    const int batch = blockIdx.y;
    if (onlyRunning && st[batch].done != RUNNING) return;
    const T m = T(means[batch]);
    FOR_BLOCK_ROWS(i, n) v[(size_t)batch * n + i] -= m;
}

// ---------------- AMG V-cycle level kernels (krylov.cu:1247-1360) ----------------
// Rows are processed by groups of W lanes (CSR-vector) when the level's rows are long.
template <int W, typename H, typename G>
__device__ __forceinline__ H group_row_sum(const int32_t *__restrict__ row, const int32_t *__restrict__ col,
                                           const H *__restrict__ val, int i, bool valid, int lane, G gather) {
    H sum = H(0);
    if (valid) {
        const int e = row[i + 1];
        for (int j = row[i] + lane; j < e; j += W) sum += val[j] * gather(col[j]);
    }
#pragma unroll
    for (int o = W / 2; o > 0; o >>= 1) sum += __shfl_down(sum, o, W);
    return sum;
}

// Row sum through the ELL copy (entry k of row i at [k][i]) when the level has one.
// Row sum of a level matrix through its sliced ELL copy, its ELL copy or its CSR arrays (in
// this order of preference); the copies are read with one lane per row (W = 1).
template <int W, typename H, typename G>
__device__ __forceinline__ H level_row_sum(const Csr<H> &A, const H *__restrict__ val, const H *__restrict__ ellv,
                                           int i, bool valid, int lane, G gather) {
    if (A.sell_off) {
        // W lanes per row split the row's entries (entry k on lane k % W)
        H sum = H(0);
        if (valid) {
            const int sl = i >> 6;
            const int w = A.sell_w[sl];
            const int base = A.sell_off[sl] + (i & 63);
            // eight entries' loads in flight per round trip (the sum keeps its order)
#pragma unroll 8
            for (int k = lane; k < w; k += W) sum += A.sell_val[base + 64 * k] * gather(A.sell_col[base + 64 * k]);
        }
#pragma unroll
        for (int o = W / 2; o > 0; o >>= 1) sum += __shfl_down(sum, o, W);
        return sum;
    }
    if (!A.ell_w) return group_row_sum<W, H>(A.row, A.col, val, i, valid, lane, gather);
    H sum = H(0);
    if (W == 1 && A.ell_w == 7) {
        // the 3D finest level: all column and value loads of the row issued up front
        if (valid) {
            int cs[7];
            H as[7];
#pragma unroll
            for (int k = 0; k < 7; ++k) {
                cs[k] = A.ell_col[(size_t)k * A.n + i];
                as[k] = ellv[(size_t)k * A.n + i];
            }
#pragma unroll
            for (int k = 0; k < 7; ++k) sum += as[k] * gather(cs[k]);
        }
        return sum;
    }
    if (valid)
        for (int k = lane; k < A.ell_w; k += W) {
            const size_t o = (size_t)k * A.n + i;
            sum += ellv[o] * gather(A.ell_col[o]);
        }
#pragma unroll
    for (int o = W / 2; o > 0; o >>= 1) sum += __shfl_down(sum, o, W);
    return sum;
}

// Rows in groups of W lanes, swept over the whole grid; each wavefront takes 64 / W
// consecutive rows per pass.
#define GROUP_ROW_LOOP(W, n)                                                                  \
    const int lane = threadIdx.x % (W);                                                       \
    const int numGroups = gridDim.x * (BLOCK / (W));                                          \
    const int waveFirst = (blockIdx.x * BLOCK + (threadIdx.x & ~63)) / (W);                   \
    const int groupInWave = (threadIdx.x & 63) / (W);                                         \
    for (int base = waveFirst; base < (n); base += numGroups)

template <typename TB, typename H>
__device__ __forceinline__ H load_rhs(const TB *__restrict__ b, size_t o, const double *shift, int batch) {
    return shift ? H(double(b[o]) - shift[batch]) : H(b[o]);
}

template <int W, typename TB, typename H>
__global__ void __launch_bounds__(BLOCK) k_presmooth_residual(Csr<H> A, const H *__restrict__ dinv, const TB *__restrict__ b,
                                                              const double *shift, H *__restrict__ x, H *__restrict__ res,
                                                              H omega, const BatchState *__restrict__ st) {
    if (st[blockIdx.y].done != RUNNING) return;
    const int n = A.n, batch = blockIdx.y;
    const H *val = A.val + (size_t)batch * A.valStride;
    const H *dv = dinv + (A.valStride ? (size_t)batch * n : 0);
    const H *ev = A.ell_val + (A.valStride ? (size_t)batch * A.ell_w * n : 0);
    const size_t off = (size_t)batch * n;
    GROUP_ROW_LOOP(W, n) {
        const int i = base + groupInWave;
        const bool valid = i < n;
        const H sum = level_row_sum<W, H>(A, val, ev, i, valid, lane,
                                          [&](int c) { return dv[c] * load_rhs<TB, H>(b, off + c, shift, batch); });
        if (valid && lane == 0) {
            const H bi = load_rhs<TB, H>(b, off + i, shift, batch);
            x[off + i] = omega * dv[i] * bi;
            res[off + i] = bi - omega * sum;
        }
    }
}

template <int W, typename TB, typename H, typename TO>
__global__ void __launch_bounds__(BLOCK) k_jacobi_sweep(Csr<H> A, const H *__restrict__ dinv, const TB *__restrict__ b,
                                                        const double *shift, const H *__restrict__ x, TO *__restrict__ out,
                                                        H omega, const BatchState *__restrict__ st,
                                                        double *outSumPartial) {
    // outSumPartial (optional): per-workgroup sums of the written values (their mean follows)
    __shared__ double sm[16];
    if (st[blockIdx.y].done != RUNNING) return;
    double osum[1][1] = {{0.0}};
    const int n = A.n, batch = blockIdx.y;
    const H *val = A.val + (size_t)batch * A.valStride;
    const H *dv = dinv + (A.valStride ? (size_t)batch * n : 0);
    const H *ev = A.ell_val + (A.valStride ? (size_t)batch * A.ell_w * n : 0);
    const size_t off = (size_t)batch * n;
    GROUP_ROW_LOOP(W, n) {
        const int i = base + groupInWave;
        const bool valid = i < n;
        const H sum = level_row_sum<W, H>(A, val, ev, i, valid, lane, [&](int c) { return x[off + c]; });
        if (valid && lane == 0) {
            const H bi = load_rhs<TB, H>(b, off + i, shift, batch);
            const TO o = TO(x[off + i] + omega * dv[i] * (bi - sum));
            out[off + i] = o;
            osum[0][0] += double(o);
        }
    }
    if (outSumPartial) store_partials<1, 1>(osum, batch, 1, outSumPartial, sm);
}

template <int W, typename TB, typename H>
__global__ void __launch_bounds__(BLOCK) k_level_residual(Csr<H> A, const TB *__restrict__ b, const double *shift,
                                                          const H *__restrict__ x, H *__restrict__ res,
                                                          const BatchState *__restrict__ st) {
    if (st[blockIdx.y].done != RUNNING) return;
    const int n = A.n, batch = blockIdx.y;
    const H *val = A.val + (size_t)batch * A.valStride;
    const H *ev = A.ell_val + (A.valStride ? (size_t)batch * A.ell_w * n : 0);
    const size_t off = (size_t)batch * n;
    GROUP_ROW_LOOP(W, n) {
        const int i = base + groupInWave;
        const bool valid = i < n;
        const H sum = level_row_sum<W, H>(A, val, ev, i, valid, lane, [&](int c) { return x[off + c]; });
        if (valid && lane == 0) res[off + i] = load_rhs<TB, H>(b, off + i, shift, batch) - sum;
    }
}

template <int W, typename H>
__global__ void __launch_bounds__(BLOCK) k_spmv_rect(Csr<H> M, const H *__restrict__ in, int nIn, H *__restrict__ out,
                                                     int accumulate, const BatchState *__restrict__ st) {
    if (st[blockIdx.y].done != RUNNING) return;
    const int n = M.n, batch = blockIdx.y;
    const H *val = M.val + (size_t)batch * M.valStride;
    const H *vin = in + (size_t)batch * nIn;
    const size_t off = (size_t)batch * n;
    GROUP_ROW_LOOP(W, n) {
        const int i = base + groupInWave;
        const bool valid = i < n;
        const H sum = level_row_sum<W, H>(M, val, (const H *)nullptr, i, valid, lane, [&](int c) { return vin[c]; });
        if (valid && lane == 0) out[off + i] = accumulate ? out[off + i] + sum : sum;
    }
}

template <typename H>
__global__ void __launch_bounds__(256) k_coarse_solve(const H *__restrict__ pinv, int nc, const H *__restrict__ b, H *__restrict__ x,
                                                      const BatchState *__restrict__ st, int batchedPinv) {
    // one wavefront per row of the dense pseudo-inverse, lanes over consecutive columns
    if (st[blockIdx.y].done != RUNNING) return;
    const int batch = blockIdx.y;
    if (batchedPinv) pinv += (size_t)batch * nc * nc;
    const int i = blockIdx.x * 4 + threadIdx.x / 64, lane = threadIdx.x % 64;
    double sum = 0;
    if (i < nc)
        for (int j = lane; j < nc; j += 64) sum += double(pinv[(size_t)i * nc + j]) * double(b[(size_t)batch * nc + j]);
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
                                                            const BatchState *__restrict__ st) {
    const int batch = blockIdx.y;
    if (st[batch].done != RUNNING) return;
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

static int row_group_width(double nnzPerRow) {
    if (nnzPerRow <= 3.0) return 1;
    if (nnzPerRow <= 6.0) return 2;
    if (nnzPerRow <= 12.0) return 4;
    if (nnzPerRow <= 24.0) return 8;
    if (nnzPerRow <= 48.0) return 16;
    return 32;
}

// ELL rows of width 7 (3D) and 5 (2D) get kernels with the width fixed at compile time
#define EW_DISPATCH(EWV, ...)                                               \
    switch (EWV) {                                                          \
        case 7: { constexpr int EW = 7; __VA_ARGS__; break; }               \
        case 5: { constexpr int EW = 5; __VA_ARGS__; break; }               \
        default: { constexpr int EW = 0; __VA_ARGS__; break; }              \
    }

#define W_DISPATCH(WV, ...)                                   \
    switch (WV) {                                             \
        case 1: { constexpr int W = 1; __VA_ARGS__; break; }  \
        case 2: { constexpr int W = 2; __VA_ARGS__; break; }  \
        case 4: { constexpr int W = 4; __VA_ARGS__; break; }  \
        case 8: { constexpr int W = 8; __VA_ARGS__; break; }  \
        case 16: { constexpr int W = 16; __VA_ARGS__; break; } \
        default: { constexpr int W = 32; __VA_ARGS__; break; } \
    }

#define NB_DISPATCH(NBV, ...)                                 \
    switch (NBV) {                                            \
        case 1: { constexpr int NB = 1; __VA_ARGS__; break; } \
        case 2: { constexpr int NB = 2; __VA_ARGS__; break; } \
        case 3: { constexpr int NB = 3; __VA_ARGS__; break; } \
        default: { constexpr int NB = 4; __VA_ARGS__; break; } \
    }

// ---------------- workspace ----------------
struct Workspace {
    void *vecs = nullptr;
    size_t vecBytes = 0;
    double *partial = nullptr;
    size_t partialCount = 0;
    void *state = nullptr; // Control + BatchState[]
    size_t stateBytes = 0;
    void *aux = nullptr;
    size_t auxBytes = 0;
    double *tols = nullptr;
    size_t tolCount = 0;
    BatchState *hostState = nullptr;
    int hostCap = 0;
    char *hostSnap = nullptr; // two pinned snapshots of Control + BatchState[] for the host loop
    size_t snapBytes = 0;
    void *tailDev = nullptr;  // V-cycle tail level table (device copy and pinned staging)
    size_t tailBytes = 0;
    void *tailHost = nullptr;
    int device = -1;
};

static std::mutex g_mutex;
static Workspace g_ws[3]; // CG, BiCGStab, AMG-PCG

static int reserve(void **ptr, size_t *cap, size_t bytes) {
    if (*cap >= bytes && *ptr) return PP_OK;
    if (*ptr) { (void)hipFree(*ptr); *ptr = nullptr; *cap = 0; }
    PP_HIP(hipMalloc(ptr, bytes));
    *cap = bytes;
    return PP_OK;
}

static int num_cus() {
    static int cus = 0;
    if (!cus) {
        int dev = 0;
        (void)hipGetDevice(&dev);
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) cus = 256;
    }
    return cus;
}

static int grid_width(int n, int groups) {
    const int want = std::max(1, (n + BLOCK - 1) / BLOCK);
    const int maxBlocks = std::max(1, num_cus() * 8 / std::max(1, groups));
    return std::min(want, maxBlocks);
}

static int prepare_workspace(Workspace &ws, int nvec, size_t n, int B, size_t elem, size_t partials, const double *tols, int ntols,
                             hipStream_t s) {
    const size_t stateBytes = 128 + sizeof(BatchState) * (size_t)B;
    if (int e = reserve(&ws.vecs, &ws.vecBytes, std::max<size_t>(1, (size_t)nvec * n * B * elem))) return e;
    if (int e = reserve((void **)&ws.partial, &ws.partialCount, std::max<size_t>(8, partials * sizeof(double)))) return e;
    if (int e = reserve(&ws.state, &ws.stateBytes, stateBytes)) return e;
    if (B > ws.hostCap) {
        if (ws.hostState) (void)hipHostFree(ws.hostState);
        PP_HIP(hipHostMalloc((void **)&ws.hostState, sizeof(BatchState) * B, hipHostMallocDefault));
        ws.hostCap = B;
    }
    if (ws.snapBytes < stateBytes) {
        if (ws.hostSnap) (void)hipHostFree(ws.hostSnap);
        ws.hostSnap = nullptr;
        ws.snapBytes = 0;
        PP_HIP(hipHostMalloc((void **)&ws.hostSnap, 2 * stateBytes, hipHostMallocDefault));
        ws.snapBytes = stateBytes;
    }
    if (ntols > 1) {
        if (int e = reserve((void **)&ws.tols, &ws.tolCount, sizeof(double) * ntols)) return e;
        PP_HIP(hipMemcpyAsync(ws.tols, tols, sizeof(double) * ntols, hipMemcpyHostToDevice, s));
    }
    return PP_OK;
}

static Control *ctrl_of(Workspace &ws) { return reinterpret_cast<Control *>(ws.state); }
static BatchState *state_of(Workspace &ws) { return reinterpret_cast<BatchState *>(reinterpret_cast<char *>(ws.state) + 128); }

static void fill_results(const BatchState *states, int B, int returnBest, pp_solve_result *out) {
    for (int b = 0; b < B; ++b) {
        const BatchState &st = states[b];
        pp_solve_result r;
        r.final_residual = st.crit;
        r.used_iterations = st.usedIt;
        r.converged = st.done == CONVERGED;
        r.is_finite_residual = std::isfinite(st.crit) ? 1 : 0;
        r.pad = st.done;
        if (st.done == RUNNING) r.converged = 0;
        if (returnBest && (st.done == MAXIT || st.done == RISING) && st.bestIt >= 0 && st.best < st.crit) {
            r.used_iterations = st.bestIt;
            r.final_residual = st.best;
        }
        out[b] = r;
    }
}

// Results from the device states (one copy and a stream synchronisation).
static int collect(Workspace &ws, int B, int returnBest, hipStream_t s, pp_solve_result *out) {
    PP_HIP(hipMemcpyAsync(ws.hostState, state_of(ws), sizeof(BatchState) * B, hipMemcpyDeviceToHost, s));
    PP_HIP(hipStreamSynchronize(s));
    fill_results(ws.hostState, B, returnBest, out);
    return PP_OK;
}

// Results from the host loop's last observed snapshot: once every right-hand side has stopped
// (or the iteration budget is spent and every chunk observed) no later launch changes a state,
// so no further copy or synchronisation is needed; the launches still queued (restoring the
// best iterate, zeroing x for a zero right-hand side) only write x, in stream order.
static int collect_snapshot(Workspace &ws, int B, int returnBest, int slot, pp_solve_result *out) {
    fill_results(reinterpret_cast<const BatchState *>(ws.hostSnap + slot * ws.snapBytes + 128), B, returnBest, out);
    return PP_OK;
}

// Host loop over chunks of iterations. After each chunk the stream copies Control and the
// per-RHS states into a pinned snapshot; the host observes the snapshots in order. The initial
// residual is observed synchronously, so a converged start costs no iteration. From two
// observations of a right-hand side's residual the host estimates the total iteration count
// (log-linear decay) and keeps that many iterations, less `margin`, enqueued ahead of the GPU,
// at most `cap` per chunk and at most two unobserved chunks; once the estimate is reached it
// waits for the next snapshot instead of enqueuing launches that would only return at once.
// Without an estimate, `speculate` (cheap iterations) doubles the chunks 1, 2, 4, ... up to
// `cap` one chunk ahead, otherwise (expensive iterations) one iteration is enqueued and
// observed before the next. The estimate only schedules launches: every launch checks the
// device state itself, so a wrong estimate costs time, never results.
struct LoopPolicy {
    int cap;        // iterations per chunk at most
    int margin;     // iterations short of the estimate to stop enqueuing at
    bool speculate; // enqueue ahead without an estimate
};

template <typename Body>
static int run_loop(Workspace &ws, hipStream_t s, int B, int maxLoops, LoopPolicy pol, int *finalSlot, Body body) {
    hipEvent_t ev[2];
    PP_HIP(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
    PP_HIP(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
    const size_t bytes = 128 + sizeof(BatchState) * (size_t)B;
    const int cap = std::max(1, pol.cap);
    auto snapCtrl = [&](int slot) { return reinterpret_cast<const Control *>(ws.hostSnap + slot * ws.snapBytes); };
    auto snapState = [&](int slot) {
        return reinterpret_cast<const BatchState *>(ws.hostSnap + slot * ws.snapBytes + 128);
    };
    int rc = PP_OK;
    auto snap = [&](int slot) -> int {
        if (hipMemcpyAsync(ws.hostSnap + slot * ws.snapBytes, ws.state, bytes, hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipEventRecord(ev[slot], s) != hipSuccess) { set_error("run_loop: poll failed"); return PP_ERR_HIP; }
        return PP_OK;
    };
    auto wait = [&](int slot) -> int {
        if (hipEventSynchronize(ev[slot]) != hipSuccess) { set_error("run_loop: sync failed"); return PP_ERR_HIP; }
        return PP_OK;
    };
    std::vector<int> prevIt(B, -1);
    std::vector<double> prevCrit(B, 0.0);
    // iterations still needed after the observed snapshot (-1: no estimate yet)
    auto remaining = [&](int slot) -> int {
        const BatchState *st = snapState(slot);
        int rem = 0;
        bool known = true;
        for (int b = 0; b < B; ++b) {
            const BatchState &x = st[b];
            if (x.done != RUNNING) continue;
            int r = -1;
            if (prevIt[b] >= 0 && x.it > prevIt[b] && prevCrit[b] > 0.0 && x.crit > 0.0 && x.tol > 0.0 &&
                std::isfinite(x.crit) && x.crit < prevCrit[b]) {
                const double rate = std::log(x.crit / prevCrit[b]) / double(x.it - prevIt[b]);
                const double need = std::log(x.tol / x.crit) / rate;
                if (std::isfinite(need)) r = std::max(1, (int)std::ceil(need - 1e-9));
            }
            if (x.it != prevIt[b]) { prevIt[b] = x.it; prevCrit[b] = x.crit; }
            if (r < 0) known = false;
            else rem = std::max(rem, std::min(r, std::max(1, x.maxit - x.it)));
        }
        return known ? rem : -1;
    };
    do {
        if ((rc = snap(0)) || (rc = wait(0))) break;
        *finalSlot = 0;
        if (snapCtrl(0)->allDone) break;
        (void)remaining(0);
        int loops = 0, estTotal = -1, grow = 1, slot = 0;
        int qSlot[2], qLoops[2], q = 0; // unobserved snapshots (oldest first) and the loops they follow
        while (true) {
            int next;
            if (estTotal >= 0) next = std::min(estTotal - pol.margin - loops, cap);
            else if (pol.speculate) next = q < 2 ? grow : 0;
            else next = q == 0 ? 1 : 0;
            if (next <= 0) next = q == 0 ? 1 : 0; // nothing in flight: make progress
            next = std::min(next, maxLoops - loops);
            if (next > 0) {
                for (int k = 0; k < next; ++k) body(s);
                loops += next;
                if (hipGetLastError() != hipSuccess) { set_error("run_loop: kernel launch failed"); rc = PP_ERR_HIP; break; }
                slot ^= 1;
                if ((rc = snap(slot))) break;
                qSlot[q] = slot;
                qLoops[q] = loops;
                ++q;
                if (estTotal < 0 && pol.speculate) grow = std::min(2 * grow, cap);
            }
            if (q == 0) break; // iteration budget spent, everything observed
            if (q == 2 || next == 0) {
                if ((rc = wait(qSlot[0]))) break;
                const int obs = qSlot[0], obsLoops = qLoops[0];
                *finalSlot = obs;
                qSlot[0] = qSlot[1];
                qLoops[0] = qLoops[1];
                --q;
                if (snapCtrl(obs)->allDone) break;
                const int r = remaining(obs);
                estTotal = r >= 0 ? obsLoops + r : -1;
            }
        }
    } while (false);
    (void)hipEventDestroy(ev[0]);
    (void)hipEventDestroy(ev[1]);
    return rc;
}

static int g_check_interval = 8;

template <typename T>
static int cg_grid(const Csr<T> &A, const T *b, T *x, int B, int maxit, const double *tols, int ntols, double normScale,
                   int resetEvery, int returnBest, int project, hipStream_t s, pp_solve_result *res) {
    Workspace &ws = g_ws[0];
    const int n = A.n;
    const int NBv = std::min(A.rpm, 4);
    const int gpm = (A.rpm + NBv - 1) / NBv;
    const int numMat = (B + A.rpm - 1) / A.rpm;
    const int groups = numMat * gpm;
    const int G = grid_width(n, groups);
    const int Gm = grid_width(n, B);
    enum { R = 0, Q, P0, P1, XB, BW, NUMV };
    if (int e = prepare_workspace(ws, NUMV, n, B, sizeof(T), (size_t)B * std::max(G, Gm) * 3 + 2 * B, tols, ntols, s)) return e;
    T *vec = reinterpret_cast<T *>(ws.vecs);
    auto V = [&](int k) { return vec + (size_t)k * n * B; };
    T *r = V(R), *q = V(Q), *p0 = V(P0), *p1 = V(P1), *xbest = returnBest ? V(XB) : nullptr, *bw = V(BW);
    BatchState *st = state_of(ws);
    Control *ctrl = ctrl_of(ws);
    double *partial = ws.partial;
    double *means = partial + (size_t)B * std::max(G, Gm) * 3;
    const dim3 grid(G, groups);
    hipLaunchKernelGGL(k_init_state, dim3(1), dim3(64), 0, s, st, B, ntols == 1 ? tols[0] : 0.0, maxit, ctrl,
                       ntols > 1 ? ws.tols : nullptr);
    PP_HIP(hipMemcpyAsync(bw, b, sizeof(T) * (size_t)n * B, hipMemcpyDeviceToDevice, s));
    if (project) {
        hipLaunchKernelGGL((k_mean<T>), dim3(Gm, B), dim3(BLOCK), 0, s, bw, n, B, means, partial, ctrl, st, 0);
        hipLaunchKernelGGL(k_fin_mean, dim3(1), dim3(BLOCK), 0, s, partial, Gm, n, B, means, ctrl, 0);
        hipLaunchKernelGGL((k_subtract_mean<T>), dim3(Gm, B), dim3(BLOCK), 0, s, bw, n, means, st, 0);
    }
    NB_DISPATCH(NBv,
        EW_DISPATCH(A.ell_w, hipLaunchKernelGGL((k_residual<T, NB, EW>), grid, dim3(BLOCK), 0, s, A, bw, x, r, (T *)nullptr, st, B, partial, ctrl,
                           normScale, 0)));
    hipLaunchKernelGGL(k_fin_residual, dim3(1), dim3(BLOCK), 0, s, st, B, partial, G, ctrl, normScale, 0, 0, 1);
    PP_LAUNCH_CHECK();
    int parity = 0;
    auto body = [&](hipStream_t strm) {
        T *pIn = parity ? p1 : p0, *pOut = parity ? p0 : p1;
        NB_DISPATCH(NBv,
            if (resetEvery > 0) {
                EW_DISPATCH(A.ell_w, hipLaunchKernelGGL((k_residual<T, NB, EW>), grid, dim3(BLOCK), 0, strm, A, bw, x, r, (T *)nullptr, st, B, partial,
                                   ctrl, normScale, resetEvery));
                hipLaunchKernelGGL(k_fin_residual, dim3(1), dim3(BLOCK), 0, strm, st, B, partial, G, ctrl, normScale,
                                   resetEvery, 0);
            }
            PP_TIMED(PP_K_CG_DIRECTION, n, strm,
                EW_DISPATCH(A.ell_w,
                    hipLaunchKernelGGL((k_cg_direction<T, NB, PM_NONE, EW>), grid, dim3(BLOCK), 0, strm, A, r,
                                       (const T *)nullptr, (const double *)nullptr, pIn, pOut, q, st, B, partial, ctrl)));
            hipLaunchKernelGGL(k_fin_alpha, dim3(1), dim3(BLOCK), 0, strm, st, B, partial, G, ctrl);
            PP_TIMED(PP_K_CG_UPDATE, n, strm,
                hipLaunchKernelGGL((k_cg_update<T, NB, PM_NONE>), grid, dim3(BLOCK), 0, strm, n, A.rpm, r, pOut, q, x, xbest,
                                   st, B, partial, ctrl, normScale, returnBest, (double *)nullptr));
            hipLaunchKernelGGL((k_fin_cg_update<PM_NONE>), dim3(1), dim3(BLOCK), 0, strm, st, B, partial, G, ctrl,
                               normScale, returnBest, (const double *)nullptr, n, (double *)nullptr));
        parity ^= 1;
    };
    int fin = 0;
    if (int e = run_loop(ws, s, B, maxit + 2, LoopPolicy{g_check_interval, 0, true}, &fin, body)) return e;
    if (returnBest)
        hipLaunchKernelGGL((k_restore_best<T>), dim3(std::min(4096, (n * B + 255) / 256)), dim3(256), 0, s, x, xbest, st, n, B);
    hipLaunchKernelGGL((k_zero_if_zero_rhs<T>), dim3(std::min(4096, (n * B + 255) / 256)), dim3(256), 0, s, x,
                       (size_t)n * B, ctrl);
    PP_LAUNCH_CHECK();
    return collect_snapshot(ws, B, returnBest, fin, res);
}

template <typename T>
static int bicg_grid(const Csr<T> &A, const T *b, T *x, int B, int maxit, const double *tols, int ntols, double normScale,
                     hipStream_t s, pp_solve_result *res) {
    Workspace &ws = g_ws[1];
    const int n = A.n;
    const int NBv = std::min(A.rpm, 4);
    const int gpm = (A.rpm + NBv - 1) / NBv;
    const int numMat = (B + A.rpm - 1) / A.rpm;
    const int groups = numMat * gpm;
    const int G = grid_width(n, groups);
    const int Gp = grid_width(n, B);
    enum { R = 0, R0, P, V_, TT, NUMV };
    if (int e = prepare_workspace(ws, NUMV, n, B, sizeof(T), (size_t)B * std::max(G, Gp) * 3, tols, ntols, s)) return e;
    T *vec = reinterpret_cast<T *>(ws.vecs);
    auto V = [&](int k) { return vec + (size_t)k * n * B; };
    T *r = V(R), *r0 = V(R0), *p = V(P), *v = V(V_), *t = V(TT);
    BatchState *st = state_of(ws);
    Control *ctrl = ctrl_of(ws);
    double *partial = ws.partial;
    const dim3 grid(G, groups), gridP(Gp, B);
    hipLaunchKernelGGL(k_init_state, dim3(1), dim3(64), 0, s, st, B, ntols == 1 ? tols[0] : 0.0, maxit, ctrl,
                       ntols > 1 ? ws.tols : nullptr);
    NB_DISPATCH(NBv,
        EW_DISPATCH(A.ell_w, hipLaunchKernelGGL((k_residual<T, NB, EW>), grid, dim3(BLOCK), 0, s, A, b, x, r, r0, st, B, partial, ctrl, normScale, 0)));
    hipLaunchKernelGGL(k_fin_residual, dim3(1), dim3(BLOCK), 0, s, st, B, partial, G, ctrl, normScale, 0, 1, 1);
    PP_LAUNCH_CHECK();
    auto body = [&](hipStream_t strm) {
        PP_TIMED(PP_K_BICG_P, n, strm, hipLaunchKernelGGL((k_bicg_p<T>), gridP, dim3(BLOCK), 0, strm, n, r, p, v, st));
        NB_DISPATCH(NBv,
            EW_DISPATCH(A.ell_w, PP_TIMED(PP_K_BICG_V, n, strm,
                hipLaunchKernelGGL((k_bicg_v<T, NB, EW>), grid, dim3(BLOCK), 0, strm, A, p, v, r0, st, B, partial, ctrl)));
            hipLaunchKernelGGL(k_fin_alpha, dim3(1), dim3(BLOCK), 0, strm, st, B, partial, G, ctrl);
            EW_DISPATCH(A.ell_w, PP_TIMED(PP_K_BICG_T, n, strm,
                hipLaunchKernelGGL((k_bicg_t<T, NB, EW>), grid, dim3(BLOCK), 0, strm, A, r, v, t, st, B, partial, ctrl,
                                   normScale)));
            hipLaunchKernelGGL(k_fin_bicg_t, dim3(1), dim3(BLOCK), 0, strm, st, B, partial, G, ctrl, normScale);
            PP_TIMED(PP_K_BICG_UPDATE, n, strm,
                hipLaunchKernelGGL((k_bicg_update<T, NB>), grid, dim3(BLOCK), 0, strm, n, A.rpm, r, r0, p, v, t, x, st, B,
                                   partial, ctrl, normScale));
            hipLaunchKernelGGL(k_fin_bicg_update, dim3(1), dim3(BLOCK), 0, strm, st, B, partial, G, ctrl, normScale));
    };
    int fin = 0;
    if (int e = run_loop(ws, s, B, maxit + 2, LoopPolicy{g_check_interval, 0, true}, &fin, body)) return e;
    hipLaunchKernelGGL((k_zero_if_zero_rhs<T>), dim3(std::min(4096, (n * B + 255) / 256)), dim3(256), 0, s, x,
                       (size_t)n * B, ctrl);
    PP_LAUNCH_CHECK();
    return collect_snapshot(ws, B, 0, fin, res);
}

template <typename T>
static int cg_small(const Csr<T> &A, const T *b, T *x, int B, int maxit, const double *tols, int ntols, double normScale,
                    int resetEvery, int returnBest, int project, int bicg, hipStream_t s, pp_solve_result *res) {
    Workspace &ws = g_ws[bicg ? 1 : 0];
    const int n = A.n;
    if (int e = prepare_workspace(ws, 9, n, B, sizeof(T), 8, tols, ntols, s)) return e;
    BatchState *st = state_of(ws);
    Control *ctrl = ctrl_of(ws);
    hipLaunchKernelGGL(k_init_state, dim3(1), dim3(64), 0, s, st, B, ntols == 1 ? tols[0] : 0.0, maxit, ctrl,
                       ntols > 1 ? ws.tols : nullptr);
    SmallWork<T> W{reinterpret_cast<T *>(ws.vecs), n};
    SmallArgs a;
    a.n = n; a.B = B; a.rpm = A.rpm; a.maxit = maxit; a.normalized = 1; a.returnBest = returnBest; a.project = project;
    a.resetEvery = resetEvery; a.bicg = bicg; a.normScale = normScale;
    const int NBv = std::min(A.rpm, 4);
    a.nb_per_group = NBv;
    const int gpm = (A.rpm + NBv - 1) / NBv;
    const int numMat = (B + A.rpm - 1) / A.rpm;
    const int groups = numMat * gpm;
    NB_DISPATCH(NBv,
        if (bicg) PP_TIMED(PP_K_BICG_SMALL, n, s,
                           hipLaunchKernelGGL((k_bicg_small<T, NB>), dim3(groups), dim3(SMALL_BLOCK), 0, s, A, b, x, W, st, a));
        else PP_TIMED(PP_K_CG_SMALL, n, s,
                      hipLaunchKernelGGL((k_cg_small<T, NB>), dim3(groups), dim3(SMALL_BLOCK), 0, s, A, b, x, W, st, a)));
    PP_LAUNCH_CHECK();
    return collect(ws, B, returnBest, s, res);
}

// ---------------- AMG-PCG driver ----------------
template <typename T, typename H>
static int amg_pcg(const Csr<T> &Af, const pp_amg_level *lv, int L, const H *pinv, int nc, int coarseBatched, const T *b,
                   T *x, int B, int maxit, const double *tols, int ntols, double normScale, int returnBest, int project,
                   int pre, int post, double omegaD, hipStream_t s, pp_solve_result *res) {
    Workspace &ws = g_ws[2];
    const int n = Af.n;
    pre = std::max(1, pre);
    post = std::max(1, post);
    const H omega = H(omegaD);
    const int G = grid_width(n, B);
    enum { R = 0, Q, P0, P1, X_UNUSED, BW, XB, U, NUMV };
    // partial sums: [phase partials | sums of r | sums of u] regions of B * GP each, then the means
    const int GP = std::max(G, std::max(1, num_cus() * 8 / B));
    if (int e = prepare_workspace(ws, NUMV, n, B, sizeof(T), (size_t)B * GP * 3 + 2 * B, tols, ntols, s)) return e;
    // hierarchy vectors: x, res, y, b per level + coarsest b and x
    std::vector<size_t> offs(L + 1);
    size_t total = 0;
    for (int l = 0; l < L; ++l) { offs[l] = total; total += (size_t)4 * lv[l].n * B; }
    offs[L] = total;
    total += (size_t)2 * nc * B;
    if (int e = reserve(&ws.aux, &ws.auxBytes, std::max<size_t>(1, total * sizeof(H)))) return e;
    H *aux = reinterpret_cast<H *>(ws.aux);
    auto lx = [&](int l) { return aux + offs[l]; };
    auto lres = [&](int l) { return aux + offs[l] + (size_t)lv[l].n * B; };
    auto ly = [&](int l) { return aux + offs[l] + (size_t)2 * lv[l].n * B; };
    auto lb = [&](int l) { return l < L ? aux + offs[l] + (size_t)3 * lv[l].n * B : aux + offs[L]; };
    H *cx = aux + offs[L] + (size_t)nc * B;
    T *vec = reinterpret_cast<T *>(ws.vecs);
    auto V = [&](int k) { return vec + (size_t)k * n * B; };
    T *r = V(R), *q = V(Q), *p0 = V(P0), *p1 = V(P1), *bw = V(BW), *u = V(U), *xbest = returnBest ? V(XB) : nullptr;
    BatchState *st = state_of(ws);
    Control *ctrl = ctrl_of(ws);
    double *partial = ws.partial;
    double *sumR = partial + (size_t)B * GP, *sumU = partial + (size_t)2 * B * GP;
    double *meanR = partial + (size_t)3 * B * GP, *meanU = meanR + B;
    const dim3 grid(G, B);
    hipLaunchKernelGGL(k_init_state, dim3(1), dim3(64), 0, s, st, B, ntols == 1 ? tols[0] : 0.0, maxit, ctrl,
                       ntols > 1 ? ws.tols : nullptr);
    PP_HIP(hipMemcpyAsync(bw, b, sizeof(T) * (size_t)n * B, hipMemcpyDeviceToDevice, s));
    if (project) {
        hipLaunchKernelGGL((k_mean<T>), grid, dim3(BLOCK), 0, s, bw, n, B, meanR, partial, ctrl, st, 1);
        hipLaunchKernelGGL(k_fin_mean, dim3(1), dim3(BLOCK), 0, s, partial, G, n, B, meanR, ctrl, 1);
        hipLaunchKernelGGL((k_subtract_mean<T>), grid, dim3(BLOCK), 0, s, bw, n, meanR, st, 1);
    }
    EW_DISPATCH(Af.ell_w, hipLaunchKernelGGL((k_residual<T, 1, EW>), grid, dim3(BLOCK), 0, s, Af, bw, x, r, (T *)nullptr, st, B, partial, ctrl, normScale, 0));
    hipLaunchKernelGGL(k_fin_residual, dim3(1), dim3(BLOCK), 0, s, st, B, partial, G, ctrl, normScale, 0, 0);
    if (project) {
        // mean of the initial residual for the first projected V-cycle; later ones come with
        // the residual updates (k_cg_update / k_fin_cg_update)
        hipLaunchKernelGGL((k_mean<T>), grid, dim3(BLOCK), 0, s, r, n, B, meanR, partial, ctrl, st, 1);
        hipLaunchKernelGGL(k_fin_mean, dim3(1), dim3(BLOCK), 0, s, partial, G, n, B, meanR, ctrl, 1);
    }
    PP_LAUNCH_CHECK();
    const int cus = num_cus();
    auto lgrid = [&](int rows, int W) {
        const int want = std::max<long>(1, ((long)rows * W + BLOCK - 1) / BLOCK);
        return dim3(std::min(want, std::max(1, cus * 8 / B)), B);
    };
    std::vector<Csr<H>> LA(L), LR(L), LP(L);
    std::vector<int> wa(L), wr(L), wp(L);
    for (int l = 0; l < L; ++l) {
        LA[l] = Csr<H>{lv[l].a_row, lv[l].a_col, reinterpret_cast<const H *>(lv[l].a_val), lv[l].n, lv[l].a_batched ? lv[l].nnz_a : 0, 1};
        if (lv[l].a_ell_width > 0) {
            LA[l].ell_w = lv[l].a_ell_width;
            LA[l].ell_col = lv[l].a_ell_col;
            LA[l].ell_val = reinterpret_cast<const H *>(lv[l].a_ell_val);
        }
        LR[l] = Csr<H>{lv[l].r_row, lv[l].r_col, reinterpret_cast<const H *>(lv[l].r_val), lv[l].n_coarse, 0, 1};
        LP[l] = Csr<H>{lv[l].p_row, lv[l].p_col, reinterpret_cast<const H *>(lv[l].p_val), lv[l].n, 0, 1};
        auto attach = [](Csr<H> &M, const pp_sell &sd) {
            if (!sd.slice_off) return;
            M.sell_off = sd.slice_off;
            M.sell_w = sd.slice_width;
            M.sell_col = sd.col;
            M.sell_val = reinterpret_cast<const H *>(sd.val);
        };
        attach(LA[l], lv[l].a_sell);
        attach(LR[l], lv[l].r_sell);
        attach(LP[l], lv[l].p_sell);
        // ELL and sliced ELL copies are read with one lane per row (coalesced across the
        // wavefront; 2, 4 or 8 lanes per sliced row measured slower on the duct hierarchy)
        wa[l] = (LA[l].ell_w || LA[l].sell_off) ? 1 : row_group_width(double(lv[l].nnz_a) / std::max(1, lv[l].n));
        wr[l] = LR[l].sell_off ? 1 : row_group_width(double(lv[l].nnz_r) / std::max(1, lv[l].n_coarse));
        wp[l] = LP[l].sell_off ? 1 : row_group_width(double(lv[l].nnz_p) / std::max(1, lv[l].n));
    }
    // first level handled by the one-workgroup tail (never the finest level)
    int lt = L;
    for (int l = 1; l < L; ++l)
        if (lv[l].n <= TAIL_ROWS && L - l <= TAIL_MAX) { lt = l; break; }
    TailArgs<H> tail{};
    if (lt < L) {
        tail.nl = L - lt; tail.nc = nc; tail.pre = pre; tail.post = post; tail.coarseBatched = coarseBatched;
        tail.omega = omega; tail.pinv = pinv; tail.cb = lb(L); tail.cx = cx;
        for (int l = lt; l < L; ++l) {
            const int k = l - lt;
            tail.n[k] = lv[l].n;
            tail.wa[k] = tail_width(lv[l].n);
            tail.wr[k] = tail_width(lv[l].n_coarse);
            tail.wp[k] = tail_width(lv[l].n);
            tail.A[k] = LA[l]; tail.R[k] = LR[l]; tail.P[k] = LP[l];
            tail.dinv[k] = reinterpret_cast<const H *>(lv[l].dinv);
            tail.x[k] = lx(l); tail.y[k] = ly(l); tail.res[k] = lres(l); tail.b[k] = lb(l);
        }
        // the staging buffer is rewritten only by the next solve, which starts after this
        // solve's final stream synchronisation
        if (!ws.tailHost) PP_HIP(hipHostMalloc(&ws.tailHost, sizeof(TailArgs<double>), hipHostMallocDefault));
        if (int e = reserve(&ws.tailDev, &ws.tailBytes, sizeof(TailArgs<double>))) return e;
        std::memcpy(ws.tailHost, &tail, sizeof(tail));
        PP_HIP(hipMemcpyAsync(ws.tailDev, ws.tailHost, sizeof(tail), hipMemcpyHostToDevice, s));
    }
    const TailArgs<H> *tailDev = reinterpret_cast<const TailArgs<H> *>(ws.tailDev);
    // With the constant-mode projection the V-cycle runs on r - mean(r) (mean from the last
    // residual update) and its output u is used as z = u - mean(u): the last finest-level sweep
    // leaves per-workgroup sums of u, one finish kernel makes the mean, and the dot and
    // direction kernels subtract it as they read u (the same T-arithmetic subtraction as
    // storing u - mean(u) first).
    const double *shift = project ? meanR : nullptr;
    auto vcycle = [&](hipStream_t strm) {
        std::vector<H *> cur(L), oth(L);
        for (int l = 0; l < L; ++l) {
            if (l == lt) break; // levels lt.. run in the one-workgroup tail
            const dim3 g = lgrid(lv[l].n, wa[l]);
            cur[l] = lx(l);
            oth[l] = ly(l);
            H *resl = lres(l);
            const H *dinv = reinterpret_cast<const H *>(lv[l].dinv);
            W_DISPATCH(wa[l],
                if (l == 0) PP_TIMED(PP_K_AMG_PRESMOOTH, lv[l].n, strm, hipLaunchKernelGGL((k_presmooth_residual<W, T, H>), g, dim3(BLOCK), 0, strm, LA[l], dinv, r, shift, cur[l], resl, omega, st));
                else PP_TIMED(PP_K_AMG_PRESMOOTH, lv[l].n, strm, hipLaunchKernelGGL((k_presmooth_residual<W, H, H>), g, dim3(BLOCK), 0, strm, LA[l], dinv, lb(l), (const double *)nullptr, cur[l], resl, omega, st));
                for (int k = 1; k < pre; ++k) {
                    if (l == 0) hipLaunchKernelGGL((k_jacobi_sweep<W, T, H, H>), g, dim3(BLOCK), 0, strm, LA[l], dinv, r, shift, cur[l], oth[l], omega, st, (double *)nullptr);
                    else hipLaunchKernelGGL((k_jacobi_sweep<W, H, H, H>), g, dim3(BLOCK), 0, strm, LA[l], dinv, lb(l), (const double *)nullptr, cur[l], oth[l], omega, st, (double *)nullptr);
                    std::swap(cur[l], oth[l]);
                }
                if (pre > 1) {
                    if (l == 0) hipLaunchKernelGGL((k_level_residual<W, T, H>), g, dim3(BLOCK), 0, strm, LA[l], r, shift, cur[l], resl, st);
                    else hipLaunchKernelGGL((k_level_residual<W, H, H>), g, dim3(BLOCK), 0, strm, LA[l], lb(l), (const double *)nullptr, cur[l], resl, st);
                });
            W_DISPATCH(wr[l],
                PP_TIMED(PP_K_AMG_RESTRICT, lv[l].n_coarse, strm, hipLaunchKernelGGL((k_spmv_rect<W, H>), lgrid(lv[l].n_coarse, wr[l]), dim3(BLOCK), 0, strm, LR[l], resl, lv[l].n, lb(l + 1), 0, st)));
        }
        if (lt < L) {
            PP_TIMED(PP_K_AMG_TAIL, lv[lt].n, strm,
                hipLaunchKernelGGL((k_vcycle_tail<H>), dim3(1, B), dim3(TAIL_BLOCK), 0, strm, tailDev, st));
            for (int l = lt; l < L; ++l) {
                // the tail leaves level l's correction in x after an even number of sweeps
                const bool inX = ((pre - 1) + post) % 2 == 0;
                cur[l] = inX ? lx(l) : ly(l);
                oth[l] = inX ? ly(l) : lx(l);
            }
        } else {
            PP_TIMED(PP_K_AMG_COARSE, nc, strm, hipLaunchKernelGGL((k_coarse_solve<H>), dim3((nc + 3) / 4, B), dim3(256), 0, strm, pinv, nc, lb(L), cx, st, coarseBatched));
        }
        for (int l = std::min(L, lt) - 1; l >= 0; --l) {
            const dim3 g = lgrid(lv[l].n, wa[l]);
            const H *xc = l + 1 < L ? cur[l + 1] : cx;
            const int ncoarse = l + 1 < L ? lv[l + 1].n : nc;
            const H *dinv = reinterpret_cast<const H *>(lv[l].dinv);
            W_DISPATCH(wp[l],
                PP_TIMED(PP_K_AMG_PROLONG, lv[l].n, strm, hipLaunchKernelGGL((k_spmv_rect<W, H>), lgrid(lv[l].n, wp[l]), dim3(BLOCK), 0, strm, LP[l], xc, ncoarse, cur[l], 1, st)));
            W_DISPATCH(wa[l],
                for (int k = 0; k < post; ++k) {
                    if (l == 0) {
                        if (k == post - 1) PP_TIMED(PP_K_AMG_POST_JACOBI, lv[l].n, strm, hipLaunchKernelGGL((k_jacobi_sweep<W, T, H, T>), g, dim3(BLOCK), 0, strm, LA[l], dinv, r, shift, cur[l], u, omega, st, project ? sumU : (double *)nullptr));
                        else PP_TIMED(PP_K_AMG_JACOBI, lv[l].n, strm, hipLaunchKernelGGL((k_jacobi_sweep<W, T, H, H>), g, dim3(BLOCK), 0, strm, LA[l], dinv, r, shift, cur[l], oth[l], omega, st, (double *)nullptr));
                    } else {
                        PP_TIMED(PP_K_AMG_POST_JACOBI, lv[l].n, strm, hipLaunchKernelGGL((k_jacobi_sweep<W, H, H, H>), g, dim3(BLOCK), 0, strm, LA[l], dinv, lb(l), (const double *)nullptr, cur[l], oth[l], omega, st, (double *)nullptr));
                    }
                    std::swap(cur[l], oth[l]);
                });
        }
        if (project)
            hipLaunchKernelGGL(k_fin_mean, dim3(1), dim3(BLOCK), 0, strm, sumU, (int)lgrid(n, wa[0]).x, n, B, meanU, ctrl, 1);
    };
    int parity = 0;
    auto body = [&](hipStream_t strm) {
        T *pIn = parity ? p1 : p0, *pOut = parity ? p0 : p1;
        vcycle(strm);
        PP_TIMED(PP_K_PCG_DOTS, n, strm,
            if (project) hipLaunchKernelGGL((k_cg_dots<T, true>), grid, dim3(BLOCK), 0, strm, n, r, u, meanU, st, B, partial, ctrl);
            else hipLaunchKernelGGL((k_cg_dots<T, false>), grid, dim3(BLOCK), 0, strm, n, r, u, (const double *)nullptr, st, B,
                                    partial, ctrl));
        hipLaunchKernelGGL(k_fin_cg_dots, dim3(1), dim3(BLOCK), 0, strm, st, B, partial, G, ctrl);
        PP_TIMED(PP_K_PCG_DIRECTION, n, strm,
            EW_DISPATCH(Af.ell_w,
                if (project) hipLaunchKernelGGL((k_cg_direction<T, 1, PM_VECTOR_SHIFTED, EW>), grid, dim3(BLOCK), 0, strm, Af,
                                                r, u, meanU, pIn, pOut, q, st, B, partial, ctrl);
                else hipLaunchKernelGGL((k_cg_direction<T, 1, PM_VECTOR, EW>), grid, dim3(BLOCK), 0, strm, Af, r, u,
                                        (const double *)nullptr, pIn, pOut, q, st, B, partial, ctrl)));
        hipLaunchKernelGGL(k_fin_alpha, dim3(1), dim3(BLOCK), 0, strm, st, B, partial, G, ctrl);
        PP_TIMED(PP_K_PCG_UPDATE, n, strm,
            hipLaunchKernelGGL((k_cg_update<T, 1, PM_VECTOR>), grid, dim3(BLOCK), 0, strm, n, 1, r, pOut, q, x, xbest, st, B,
                               partial, ctrl, normScale, returnBest, project ? sumR : (double *)nullptr));
        hipLaunchKernelGGL((k_fin_cg_update<PM_VECTOR>), dim3(1), dim3(BLOCK), 0, strm, st, B, partial, G, ctrl, normScale,
                           returnBest, project ? sumR : (const double *)nullptr, n, meanR);
        parity ^= 1;
    };
    int fin = 0;
    if (int e = run_loop(ws, s, B, maxit + 2, LoopPolicy{2, 1, false}, &fin, body)) return e;
    if (returnBest)
        hipLaunchKernelGGL((k_restore_best<T>), dim3(std::min(4096, (n * B + 255) / 256)), dim3(256), 0, s, x, xbest, st, n, B);
    PP_LAUNCH_CHECK();
    return collect_snapshot(ws, B, returnBest, fin, res);
}

// ---------------- misc kernels ----------------
template <typename T>
__global__ void k_spmv(Csr<T> A, const T *__restrict__ x, T *__restrict__ y, int B) {
    const int b = blockIdx.y;
    const T *val = A.val + (size_t)(b / A.rpm) * A.valStride;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < A.n; i += gridDim.x * blockDim.x) {
        T s = T(0);
        for (int j = A.row[i]; j < A.row[i + 1]; ++j) s += val[j] * x[(size_t)b * A.n + A.col[j]];
        y[(size_t)b * A.n + i] = s;
    }
}

// Sliced ELL copy (pp_sell layout) of a CSR matrix: one work-item per row slot of the padded
// slices; padding entries get column 0 and value 0.
template <typename T, typename TO>
__global__ void k_sell_fill(const int32_t *__restrict__ row, const int32_t *__restrict__ col, const T *__restrict__ val,
                            int n, const int32_t *__restrict__ sliceOff, const int32_t *__restrict__ sliceW,
                            int32_t *__restrict__ sellCol, TO *__restrict__ sellVal) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (n + 63) / 64 * 64) return;
    const int sl = i >> 6, w = sliceW[sl], base = sliceOff[sl] + (i & 63);
    const int b = i < n ? row[i] : 0, len = i < n ? row[i + 1] - b : 0;
    for (int k = 0; k < w; ++k) {
        const bool has = k < len;
        if (sellCol) sellCol[base + 64 * k] = has ? col[b + k] : 0;
        if (sellVal) sellVal[base + 64 * k] = has ? TO(val[b + k]) : TO(0);
    }
}

// ELL copy of a CSR matrix: entry k of row i at [k][i], rows padded to `width` with the
// row's own column and zero values; rows longer than width are an error left to the caller
// (the library's domain matrices have at most 2d+1 entries per row).
template <typename T, typename TO>
__global__ void k_ell_from_csr(const int32_t *__restrict__ row, const int32_t *__restrict__ col,
                               const T *__restrict__ val, int n, int nnz, int width, int32_t *__restrict__ ellCol,
                               TO *__restrict__ ellVal) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int m = blockIdx.y;
    const int b = row[i], e = min(row[i + 1], b + width);
    for (int k = 0; k < width; ++k) {
        const size_t o = (size_t)k * n + i;
        const bool has = b + k < e;
        if (ellCol && m == 0) ellCol[o] = has ? col[b + k] : i;
        if (ellVal) ellVal[(size_t)m * width * n + o] = has ? TO(val[(size_t)m * nnz + b + k]) : TO(0);
    }
}

// Transpose of a structurally symmetric pattern: A^T has A's pattern, the value at (i, j)
// is A(j, i), found by searching row j (entries of a column in ascending row order, the
// order of the reference's cusparseCsr2cscEx2 transpose).
template <typename T>
__global__ void k_transpose_symmetric(const int32_t *__restrict__ row, const int32_t *__restrict__ col,
                                      const T *__restrict__ val, int n, int nnz, int nmat, T *__restrict__ tval,
                                      int *bad) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int k = row[i]; k < row[i + 1]; ++k) {
        const int j = col[k];
        int found = -1;
        for (int kk = row[j]; kk < row[j + 1]; ++kk)
            if (col[kk] == i) { found = kk; break; }
        if (found < 0) { *bad = 1; continue; }
        for (int m = 0; m < nmat; ++m) tval[(size_t)m * nnz + k] = val[(size_t)m * nnz + found];
    }
}

template <typename T>
__global__ void k_outer_product(const T *__restrict__ a, const T *__restrict__ b, const int32_t *__restrict__ row,
                                const int32_t *__restrict__ col, T *__restrict__ out, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const T ai = a[i];
    for (int k = row[i]; k < row[i + 1]; ++k) out[k] = ai * b[col[k]];
}

template <typename T>
__global__ void k_segmented_gather(const int32_t *__restrict__ seg, const int32_t *__restrict__ src,
                                   const T *__restrict__ w, int nOut, const T *__restrict__ in, T *__restrict__ out,
                                   int inStride, int outStride) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nOut) return;
    const int m = blockIdx.y;
    const T *vin = in + (size_t)m * inStride;
    T s = T(0);
    for (int t = seg[e]; t < seg[e + 1]; ++t) s += w[t] * vin[src[t]];
    out[(size_t)m * outStride + e] = s;
}


// out[e] = sum_{t in seg(e)} w[t] * in[src[t]] accumulated in double (Galerkin map). W lanes
// share an entry: lane j takes terms j, j+W, ... (neighbouring lanes read neighbouring terms),
// then a fixed shuffle tree combines the W partial sums.
template <int W, typename TA, typename TO>
__global__ void __launch_bounds__(256) k_galerkin_apply(const int32_t *__restrict__ seg, const int32_t *__restrict__ src,
                                                        const double *__restrict__ w, int nOut,
                                                        const TA *__restrict__ in, TO *__restrict__ out, int inStride,
                                                        int outStride) {
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    const int e = gid / W;
    const int lane = gid % W;
    const TA *vin = in + (size_t)blockIdx.y * inStride;
    double acc = 0.0;
    if (e < nOut) {
        const int end = seg[e + 1];
        for (int t = seg[e] + lane; t < end; t += W) acc += w[t] * double(vin[src[t]]);
    }
#pragma unroll
    for (int o = W / 2; o > 0; o >>= 1) acc += __shfl_down(acc, o, W);
    if (e < nOut && lane == 0) out[(size_t)blockIdx.y * outStride + e] = TO(acc);
}

// Galerkin product C = R A P into the fixed pattern of C (one thread per coarse row,
// accumulation in double; A may hold nmat value arrays of one pattern).
template <typename TA, typename TO>
__global__ void k_galerkin_rap(const int32_t *__restrict__ rrow, const int32_t *__restrict__ rcol, const double *__restrict__ rval,
                               const int32_t *__restrict__ arow, const int32_t *__restrict__ acol, const TA *__restrict__ aval,
                               int annz, const int32_t *__restrict__ prow, const int32_t *__restrict__ pcol,
                               const double *__restrict__ pval, const int32_t *__restrict__ crow, const int32_t *__restrict__ ccol,
                               TO *__restrict__ cval, int cnnz, int nc, int *bad) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nc) return;
    const int m = blockIdx.y;
    const TA *av = aval + (size_t)m * annz;
    const int cb = crow[i], ce = crow[i + 1];
    double acc[64];
    const bool inreg = ce - cb <= 64;
    TO *out = cval + (size_t)m * cnnz;
    if (inreg)
        for (int t = 0; t < ce - cb; ++t) acc[t] = 0.0;
    else
        for (int t = cb; t < ce; ++t) out[t] = TO(0);
    for (int kk = rrow[i]; kk < rrow[i + 1]; ++kk) {
        const int k = rcol[kk];
        const double rv = rval[kk];
        for (int ll = arow[k]; ll < arow[k + 1]; ++ll) {
            const int l = acol[ll];
            const double ra = rv * double(av[ll]);
            for (int jj = prow[l]; jj < prow[l + 1]; ++jj) {
                const int j = pcol[jj];
                int lo = cb, hi = ce - 1;
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if (ccol[mid] < j) lo = mid + 1; else hi = mid;
                }
                if (lo >= ce || ccol[lo] != j) { *bad = 1; continue; }
                if (inreg) acc[lo - cb] += ra * pval[jj];
                else out[lo] = TO(double(out[lo]) + ra * pval[jj]);
            }
        }
    }
    if (inreg)
        for (int t = 0; t < ce - cb; ++t) out[cb + t] = TO(acc[t]);
}

template <typename T, typename TO>
__global__ void k_inverse_diagonal(const int32_t *__restrict__ row, const int32_t *__restrict__ col,
                                   const T *__restrict__ val, int n, int nnz, int nmat, TO *__restrict__ dinv) {
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (size_t)n * nmat) return;
    const int m = idx / n, i = idx % n;
    const T *v = val + (size_t)m * nnz;
    T d = T(0);
    for (int k = row[i]; k < row[i + 1]; ++k)
        if (col[k] == i) d += v[k];
    TO dd = TO(d);
    const double ad = fabs(double(dd));
    dinv[idx] = ad > 1e-30 ? TO(1) / dd : TO(1);
}

// The same through the ELL copy (entry k of row i at [k][i]; padding adds zeros on the diagonal)
template <typename T, typename TO>
__global__ void k_inverse_diagonal_ell(const int32_t *__restrict__ ellCol, const T *__restrict__ ellVal, int n,
                                       int width, int nmat, TO *__restrict__ dinv) {
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (size_t)n * nmat) return;
    const int m = idx / n, i = idx % n;
    const T *v = ellVal + (size_t)m * width * n;
    T d = T(0);
    for (int k = 0; k < width; ++k)
        if (ellCol[(size_t)k * n + i] == i) d += v[(size_t)k * n + i];
    TO dd = TO(d);
    const double ad = fabs(double(dd));
    dinv[idx] = ad > 1e-30 ? TO(1) / dd : TO(1);
}

template <typename TI, typename TO>
__global__ void k_cast(const TI *__restrict__ in, TO *__restrict__ out, int64_t count) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = TO(in[i]);
}

} // namespace pp

using namespace pp;

extern "C" {

int pp_solve_linear(const pp_csr *A, const void *rhs, void *x, int32_t nrhs, int32_t dtype, int32_t maxit,
                    const double *tols, int32_t ntols, int32_t criterion, int32_t use_bicg, int32_t matrix_rank_deficient,
                    int32_t residual_reset_steps, int32_t transpose_a, int32_t return_best, void *stream,
                    pp_solve_result *results) {
    PP_CHECK_ARG(A && A->row && A->col && A->val && rhs && x && tols && results, "null argument");
    PP_CHECK_ARG(A->n > 0 && nrhs > 0, "empty system");
    PP_CHECK_ARG(A->num_matrices >= 1 && nrhs % A->num_matrices == 0,
                 "number of right-hand sides must be a multiple of the number of matrices");
    PP_CHECK_ARG(ntols == 1 || ntols == nrhs, "tolerance must be a scalar or one value per right-hand side");
    PP_CHECK_ARG(criterion == PP_CONV_NORM2 || criterion == PP_CONV_NORM2_NORMALIZED, "unsupported convergence criterion");
    PP_CHECK_ARG(maxit >= 0, "maxit must be non-negative");
    PP_CHECK_ARG(!(use_bicg && matrix_rank_deficient), "rank-deficient mode is a CG mode");
    std::lock_guard<std::mutex> lock(g_mutex);
    hipStream_t s = as_stream(stream);
    const int rpm = nrhs / A->num_matrices;
    const double normScale = criterion == PP_CONV_NORM2_NORMALIZED ? 1.0 / std::sqrt(double(A->n)) : 1.0;
    if (use_bicg) return_best = 0;
    static void *tbuf = nullptr;
    static size_t tcap = 0;
    static int *tbad = nullptr;
    PP_DISPATCH(dtype, {
        Csr<T> C = csr_view<T>(A, rpm);
        if (transpose_a && use_bicg) {
            C.ell_w = 0;
            const size_t bytes = sizeof(T) * (size_t)A->nnz * A->num_matrices;
            if (int e = reserve(&tbuf, &tcap, std::max<size_t>(bytes, 8))) return e;
            if (!tbad) PP_HIP(hipMalloc((void **)&tbad, sizeof(int)));
            PP_HIP(hipMemsetAsync(tbad, 0, sizeof(int), s));
            hipLaunchKernelGGL((k_transpose_symmetric<T>), dim3((A->n + 255) / 256), dim3(256), 0, s, A->row, A->col,
                               reinterpret_cast<const T *>(A->val), A->n, A->nnz, A->num_matrices, reinterpret_cast<T *>(tbuf), tbad);
            int bad = 0;
            PP_HIP(hipMemcpyAsync(&bad, tbad, sizeof(int), hipMemcpyDeviceToHost, s));
            PP_HIP(hipStreamSynchronize(s));
            PP_CHECK_ARG(!bad, "transposed solves need a structurally symmetric pattern");
            C.val = reinterpret_cast<const T *>(tbuf);
        }
        const T *b = reinterpret_cast<const T *>(rhs);
        T *xx = reinterpret_cast<T *>(x);
        if (A->n <= SMALL_N)
            return cg_small<T>(C, b, xx, nrhs, maxit, tols, ntols, normScale, use_bicg ? 0 : residual_reset_steps,
                               return_best, matrix_rank_deficient, use_bicg, s, results);
        if (use_bicg) return bicg_grid<T>(C, b, xx, nrhs, maxit, tols, ntols, normScale, s, results);
        return cg_grid<T>(C, b, xx, nrhs, maxit, tols, ntols, normScale, residual_reset_steps, return_best,
                          matrix_rank_deficient, s, results);
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
        const bool any = sd.slice_off || sd.slice_width || sd.col || sd.val;
        return !any || (sd.slice_off && sd.slice_width && sd.col && sd.val);
    };
    for (int l = 0; l < num_levels; ++l) {
        PP_CHECK_ARG(levels[l].a_ell_width == 0 || (levels[l].a_ell_width > 0 && levels[l].a_ell_col && levels[l].a_ell_val),
                     "level ELL copy needs a positive width, columns and values");
        PP_CHECK_ARG(sell_ok(levels[l].a_sell) && sell_ok(levels[l].r_sell), "sliced ELL copy needs all four arrays");
        PP_CHECK_ARG(!levels[l].a_sell.slice_off || !levels[l].a_batched, "sliced ELL copy of a batched level operator");
        PP_CHECK_ARG(levels[l].a_val || (l == 0 && levels[l].a_ell_width > 0),
                     "level values missing (only the finest level with an ELL copy may omit its CSR values)");
    }
    PP_CHECK_ARG(levels[num_levels - 1].n_coarse == nc, "coarsest level size mismatch");
    PP_CHECK_ARG(ntols == 0 || ntols == 1 || ntols == nrhs, "tolerance count");
    std::lock_guard<std::mutex> lock(g_mutex);
    hipStream_t s = as_stream(stream);
    const double normScale = normalized ? 1.0 / std::sqrt(double(A->n)) : 1.0;
    double t1 = tol;
    const double *tp = ntols > 1 ? tols : &t1;
    const int nt = ntols > 1 ? ntols : 1;
    if (dtype == PP_F64 && hdtype == PP_F32) {
        const Csr<double> C = csr_view<double>(A, 1);
        return amg_pcg<double, float>(C, levels, num_levels, reinterpret_cast<const float *>(coarse_pinv), nc, coarse_batched,
                                      reinterpret_cast<const double *>(rhs), reinterpret_cast<double *>(x), nrhs, maxit, tp, nt,
                                      normScale, return_best, project_constant, pre_sweeps, post_sweeps, omega, s, results);
    }
    if (dtype == PP_F64 && hdtype == PP_F64) {
        const Csr<double> C = csr_view<double>(A, 1);
        return amg_pcg<double, double>(C, levels, num_levels, reinterpret_cast<const double *>(coarse_pinv), nc, coarse_batched,
                                       reinterpret_cast<const double *>(rhs), reinterpret_cast<double *>(x), nrhs, maxit, tp, nt,
                                       normScale, return_best, project_constant, pre_sweeps, post_sweeps, omega, s, results);
    }
    if (dtype == PP_F32 && hdtype == PP_F32) {
        const Csr<float> C = csr_view<float>(A, 1);
        return amg_pcg<float, float>(C, levels, num_levels, reinterpret_cast<const float *>(coarse_pinv), nc, coarse_batched,
                                     reinterpret_cast<const float *>(rhs), reinterpret_cast<float *>(x), nrhs, maxit, tp, nt,
                                     normScale, return_best, project_constant, pre_sweeps, post_sweeps, omega, s, results);
    }
    set_error("pp_amg_pcg_solve: unsupported (dtype, hierarchy dtype) combination");
    return PP_ERR_UNSUPPORTED;
}

int pp_release_workspaces(void) {
    std::lock_guard<std::mutex> lock(g_mutex);
    for (auto &ws : g_ws) {
        if (ws.vecs) (void)hipFree(ws.vecs);
        if (ws.partial) (void)hipFree(ws.partial);
        if (ws.state) (void)hipFree(ws.state);
        if (ws.aux) (void)hipFree(ws.aux);
        if (ws.tols) (void)hipFree(ws.tols);
        if (ws.hostState) (void)hipHostFree(ws.hostState);
        if (ws.hostSnap) (void)hipHostFree(ws.hostSnap);
        if (ws.tailDev) (void)hipFree(ws.tailDev);
        if (ws.tailHost) (void)hipHostFree(ws.tailHost);
        ws = Workspace();
    }
    return PP_OK;
}

int pp_spmv(const pp_csr *A, const void *x, void *y, int32_t nrhs, int32_t dtype, void *stream) {
    PP_CHECK_ARG(A && x && y && nrhs > 0, "null argument");
    PP_CHECK_ARG(A->num_matrices >= 1 && nrhs % A->num_matrices == 0, "rhs count");
    const int rpm = nrhs / A->num_matrices;
    PP_DISPATCH(dtype,
        Csr<T> C{A->row, A->col, reinterpret_cast<const T *>(A->val), A->n, A->num_matrices > 1 ? A->nnz : 0, rpm};
        hipLaunchKernelGGL((k_spmv<T>), dim3(std::min(4096, (A->n + 255) / 256), nrhs), dim3(256), 0, as_stream(stream), C,
                           reinterpret_cast<const T *>(x), reinterpret_cast<T *>(y), nrhs));
    PP_LAUNCH_CHECK();
    return PP_OK;
}

int pp_sparse_outer_product(const void *a, const void *b, const pp_csr *out, int32_t dtype, void *stream) {
    PP_CHECK_ARG(a && b && out && out->val, "null argument");
    PP_DISPATCH(dtype,
        hipLaunchKernelGGL((k_outer_product<T>), dim3((out->n + 255) / 256), dim3(256), 0, as_stream(stream),
                           reinterpret_cast<const T *>(a), reinterpret_cast<const T *>(b), out->row, out->col,
                           reinterpret_cast<T *>(out->val), out->n));
    PP_LAUNCH_CHECK();
    return PP_OK;
}

int pp_csr_transpose(const pp_csr *A, int32_t dtype, int32_t *t_row, int32_t *t_col, void *t_val, void *stream) {
    PP_CHECK_ARG(A && t_row && t_col && t_val, "null argument");
    hipStream_t s = as_stream(stream);
    int *bad = nullptr;
    PP_HIP(hipMalloc((void **)&bad, sizeof(int)));
    PP_HIP(hipMemsetAsync(bad, 0, sizeof(int), s));
    PP_HIP(hipMemcpyAsync(t_row, A->row, sizeof(int32_t) * (A->n + 1), hipMemcpyDeviceToDevice, s));
    PP_HIP(hipMemcpyAsync(t_col, A->col, sizeof(int32_t) * A->nnz, hipMemcpyDeviceToDevice, s));
    PP_DISPATCH(dtype,
        hipLaunchKernelGGL((k_transpose_symmetric<T>), dim3((A->n + 255) / 256), dim3(256), 0, s, A->row, A->col,
                           reinterpret_cast<const T *>(A->val), A->n, A->nnz, A->num_matrices, reinterpret_cast<T *>(t_val), bad));
    int hbad = 0;
    PP_HIP(hipMemcpyAsync(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost, s));
    PP_HIP(hipStreamSynchronize(s));
    (void)hipFree(bad);
    PP_CHECK_ARG(!hbad, "pattern is not structurally symmetric");
    return PP_OK;
}

int pp_segmented_gather(const int32_t *seg, const int32_t *src, const void *w, int32_t n_out, const void *in, void *out,
                        int32_t nmat, int32_t in_stride, int32_t out_stride, int32_t dtype, void *stream) {
    PP_CHECK_ARG(seg && src && w && in && out && n_out >= 0 && nmat >= 1, "null argument");
    if (n_out == 0) return PP_OK;
    PP_DISPATCH(dtype,
        hipLaunchKernelGGL((k_segmented_gather<T>), dim3((n_out + 255) / 256, nmat), dim3(256), 0, as_stream(stream), seg, src,
                           reinterpret_cast<const T *>(w), n_out, reinterpret_cast<const T *>(in), reinterpret_cast<T *>(out),
                           in_stride, out_stride));
    PP_LAUNCH_CHECK();
    return PP_OK;
}


int pp_ell_from_csr(const int32_t *row, const int32_t *col, const void *val, int32_t n, int32_t nnz, int32_t nmat,
                    int32_t width, int32_t dtype, int32_t *ell_col, void *ell_val, int32_t ell_dtype, void *stream) {
    PP_CHECK_ARG(row && n >= 0 && width > 0 && nmat >= 1, "bad arguments");
    PP_CHECK_ARG(!ell_col || col, "ell_col needs col");
    PP_CHECK_ARG(!ell_val || val, "ell_val needs val");
    if (n == 0) return PP_OK;
    const dim3 g((n + 255) / 256, nmat);
    hipStream_t s = as_stream(stream);
    if (dtype == PP_F64 && ell_dtype == PP_F64)
        hipLaunchKernelGGL((k_ell_from_csr<double, double>), g, dim3(256), 0, s, row, col, (const double *)val, n, nnz, width, ell_col, (double *)ell_val);
    else if (dtype == PP_F64 && ell_dtype == PP_F32)
        hipLaunchKernelGGL((k_ell_from_csr<double, float>), g, dim3(256), 0, s, row, col, (const double *)val, n, nnz, width, ell_col, (float *)ell_val);
    else if (dtype == PP_F32 && ell_dtype == PP_F32)
        hipLaunchKernelGGL((k_ell_from_csr<float, float>), g, dim3(256), 0, s, row, col, (const float *)val, n, nnz, width, ell_col, (float *)ell_val);
    else if (dtype == PP_F32 && ell_dtype == PP_F64)
        hipLaunchKernelGGL((k_ell_from_csr<float, double>), g, dim3(256), 0, s, row, col, (const float *)val, n, nnz, width, ell_col, (double *)ell_val);
    else { set_error("pp_ell_from_csr: dtype"); return PP_ERR_ARG; }
    PP_LAUNCH_CHECK();
    return PP_OK;
}

int pp_sell_fill(const int32_t *row, const int32_t *col, const void *val, int32_t n, const int32_t *slice_off,
                 const int32_t *slice_width, int32_t dtype, int32_t *sell_col, void *sell_val, int32_t sell_dtype,
                 void *stream) {
    PP_CHECK_ARG(row && slice_off && slice_width && n >= 0, "bad arguments");
    PP_CHECK_ARG(!sell_col || col, "sell_col needs col");
    PP_CHECK_ARG(!sell_val || val, "sell_val needs val");
    if (n == 0) return PP_OK;
    const dim3 g(((n + 63) / 64 * 64 + 255) / 256);
    hipStream_t s = as_stream(stream);
    if (!sell_val)
        hipLaunchKernelGGL((k_sell_fill<float, float>), g, dim3(256), 0, s, row, col, (const float *)nullptr, n, slice_off,
                           slice_width, sell_col, (float *)nullptr);
    else if (dtype == PP_F64 && sell_dtype == PP_F64)
        hipLaunchKernelGGL((k_sell_fill<double, double>), g, dim3(256), 0, s, row, col, (const double *)val, n, slice_off,
                           slice_width, sell_col, (double *)sell_val);
    else if (dtype == PP_F64 && sell_dtype == PP_F32)
        hipLaunchKernelGGL((k_sell_fill<double, float>), g, dim3(256), 0, s, row, col, (const double *)val, n, slice_off,
                           slice_width, sell_col, (float *)sell_val);
    else if (dtype == PP_F32 && sell_dtype == PP_F32)
        hipLaunchKernelGGL((k_sell_fill<float, float>), g, dim3(256), 0, s, row, col, (const float *)val, n, slice_off,
                           slice_width, sell_col, (float *)sell_val);
    else if (dtype == PP_F32 && sell_dtype == PP_F64)
        hipLaunchKernelGGL((k_sell_fill<float, double>), g, dim3(256), 0, s, row, col, (const float *)val, n, slice_off,
                           slice_width, sell_col, (double *)sell_val);
    else { set_error("pp_sell_fill: dtype"); return PP_ERR_ARG; }
    PP_LAUNCH_CHECK();
    return PP_OK;
}

int pp_galerkin_apply(const int32_t *seg, const int32_t *src, const double *w, int32_t n_out, int64_t nterms,
                      const void *a_val,
                      int32_t a_dtype, void *c_val, int32_t c_dtype, int32_t nmat, int32_t a_stride, int32_t c_stride,
                      void *stream) {
    PP_CHECK_ARG(seg && src && w && a_val && c_val && n_out >= 0 && nmat >= 1, "null argument");
    if (n_out == 0) return PP_OK;
    hipStream_t s = as_stream(stream);
    // lanes per entry from the mean number of terms
    const double mean = double(nterms) / double(n_out);
    const int W = mean > 24 ? 16 : (mean > 12 ? 8 : (mean > 6 ? 4 : (mean > 3 ? 2 : 1)));
    const dim3 g((unsigned)(((size_t)n_out * W + 255) / 256), nmat);
#define GAL_LAUNCH(WV, TA, TO)                                                                                  \
    PP_TIMED(PP_K_GALERKIN, n_out, s,                                                                          \
             hipLaunchKernelGGL((k_galerkin_apply<WV, TA, TO>), g, dim3(256), 0, s, seg, src, w, n_out,          \
                                (const TA *)a_val, (TO *)c_val, a_stride, c_stride))
#define GAL_W(TA, TO)                                                                                           \
    switch (W) {                                                                                                \
    case 1: GAL_LAUNCH(1, TA, TO); break;                                                                      \
    case 2: GAL_LAUNCH(2, TA, TO); break;                                                                      \
    case 4: GAL_LAUNCH(4, TA, TO); break;                                                                      \
    case 8: GAL_LAUNCH(8, TA, TO); break;                                                                      \
    default: GAL_LAUNCH(16, TA, TO); break;                                                                    \
    }
    if (a_dtype == PP_F64 && c_dtype == PP_F64) { GAL_W(double, double) }
    else if (a_dtype == PP_F64 && c_dtype == PP_F32) { GAL_W(double, float) }
    else if (a_dtype == PP_F32 && c_dtype == PP_F32) { GAL_W(float, float) }
    else if (a_dtype == PP_F32 && c_dtype == PP_F64) { GAL_W(float, double) }
    else { set_error("pp_galerkin_apply: dtype"); return PP_ERR_ARG; }
#undef GAL_W
#undef GAL_LAUNCH
    PP_LAUNCH_CHECK();
    return PP_OK;
}

int pp_galerkin_rap(const int32_t *r_row, const int32_t *r_col, const double *r_val, const pp_csr *A, int32_t a_dtype,
                    const int32_t *p_row, const int32_t *p_col, const double *p_val, const int32_t *c_row,
                    const int32_t *c_col, void *c_val, int32_t c_n, int32_t c_nnz, int32_t c_dtype, void *stream) {
    PP_CHECK_ARG(r_row && r_col && r_val && A && p_row && p_col && p_val && c_row && c_col && c_val, "null argument");
    if (c_n == 0) return PP_OK;
    hipStream_t s = as_stream(stream);
    static int *bad = nullptr;
    if (!bad) PP_HIP(hipMalloc((void **)&bad, sizeof(int)));
    PP_HIP(hipMemsetAsync(bad, 0, sizeof(int), s));
    const dim3 g((c_n + 127) / 128, A->num_matrices);
    if (a_dtype == PP_F64 && c_dtype == PP_F64)
        hipLaunchKernelGGL((k_galerkin_rap<double, double>), g, dim3(128), 0, s, r_row, r_col, r_val, A->row, A->col, (const double *)A->val, A->nnz, p_row, p_col, p_val, c_row, c_col, (double *)c_val, c_nnz, c_n, bad);
    else if (a_dtype == PP_F64 && c_dtype == PP_F32)
        hipLaunchKernelGGL((k_galerkin_rap<double, float>), g, dim3(128), 0, s, r_row, r_col, r_val, A->row, A->col, (const double *)A->val, A->nnz, p_row, p_col, p_val, c_row, c_col, (float *)c_val, c_nnz, c_n, bad);
    else if (a_dtype == PP_F32 && c_dtype == PP_F32)
        hipLaunchKernelGGL((k_galerkin_rap<float, float>), g, dim3(128), 0, s, r_row, r_col, r_val, A->row, A->col, (const float *)A->val, A->nnz, p_row, p_col, p_val, c_row, c_col, (float *)c_val, c_nnz, c_n, bad);
    else if (a_dtype == PP_F32 && c_dtype == PP_F64)
        hipLaunchKernelGGL((k_galerkin_rap<float, double>), g, dim3(128), 0, s, r_row, r_col, r_val, A->row, A->col, (const float *)A->val, A->nnz, p_row, p_col, p_val, c_row, c_col, (double *)c_val, c_nnz, c_n, bad);
    else { set_error("pp_galerkin_rap: dtype"); return PP_ERR_ARG; }
    PP_LAUNCH_CHECK();
    return PP_OK;
}

int pp_inverse_diagonal(const pp_csr *A, int32_t dtype, void *dinv, int32_t dtype_out, void *stream) {
    PP_CHECK_ARG(A && dinv, "null argument");
    const size_t total = (size_t)A->n * A->num_matrices;
    const dim3 g((total + 255) / 256);
    hipStream_t s = as_stream(stream);
    const bool ell = A->ell_width > 0 && A->ell_col && A->ell_val;
    if (ell && dtype == PP_F64 && dtype_out == PP_F64)
        hipLaunchKernelGGL((k_inverse_diagonal_ell<double, double>), g, dim3(256), 0, s, A->ell_col, (const double *)A->ell_val, A->n, A->ell_width, A->num_matrices, (double *)dinv);
    else if (ell && dtype == PP_F64 && dtype_out == PP_F32)
        hipLaunchKernelGGL((k_inverse_diagonal_ell<double, float>), g, dim3(256), 0, s, A->ell_col, (const double *)A->ell_val, A->n, A->ell_width, A->num_matrices, (float *)dinv);
    else if (ell && dtype == PP_F32 && dtype_out == PP_F32)
        hipLaunchKernelGGL((k_inverse_diagonal_ell<float, float>), g, dim3(256), 0, s, A->ell_col, (const float *)A->ell_val, A->n, A->ell_width, A->num_matrices, (float *)dinv);
    else if (ell && dtype == PP_F32 && dtype_out == PP_F64)
        hipLaunchKernelGGL((k_inverse_diagonal_ell<float, double>), g, dim3(256), 0, s, A->ell_col, (const float *)A->ell_val, A->n, A->ell_width, A->num_matrices, (double *)dinv);
    else if (dtype == PP_F64 && dtype_out == PP_F64)
        hipLaunchKernelGGL((k_inverse_diagonal<double, double>), g, dim3(256), 0, s, A->row, A->col, (const double *)A->val, A->n, A->nnz, A->num_matrices, (double *)dinv);
    else if (dtype == PP_F64 && dtype_out == PP_F32)
        hipLaunchKernelGGL((k_inverse_diagonal<double, float>), g, dim3(256), 0, s, A->row, A->col, (const double *)A->val, A->n, A->nnz, A->num_matrices, (float *)dinv);
    else if (dtype == PP_F32 && dtype_out == PP_F32)
        hipLaunchKernelGGL((k_inverse_diagonal<float, float>), g, dim3(256), 0, s, A->row, A->col, (const float *)A->val, A->n, A->nnz, A->num_matrices, (float *)dinv);
    else if (dtype == PP_F32 && dtype_out == PP_F64)
        hipLaunchKernelGGL((k_inverse_diagonal<float, double>), g, dim3(256), 0, s, A->row, A->col, (const float *)A->val, A->n, A->nnz, A->num_matrices, (double *)dinv);
    else { set_error("pp_inverse_diagonal: dtype"); return PP_ERR_ARG; }
    PP_LAUNCH_CHECK();
    return PP_OK;
}

int pp_cast(const void *in, int32_t dtype_in, void *out, int32_t dtype_out, int64_t count, void *stream) {
    PP_CHECK_ARG(in && out && count >= 0, "null argument");
    if (count == 0) return PP_OK;
    const dim3 g((unsigned)std::min<int64_t>(4096, (count + 255) / 256));
    hipStream_t s = as_stream(stream);
    if (dtype_in == PP_F64 && dtype_out == PP_F32) hipLaunchKernelGGL((k_cast<double, float>), g, dim3(256), 0, s, (const double *)in, (float *)out, count);
    else if (dtype_in == PP_F32 && dtype_out == PP_F64) hipLaunchKernelGGL((k_cast<float, double>), g, dim3(256), 0, s, (const float *)in, (double *)out, count);
    else if (dtype_in == PP_F64 && dtype_out == PP_F64) hipLaunchKernelGGL((k_cast<double, double>), g, dim3(256), 0, s, (const double *)in, (double *)out, count);
    else if (dtype_in == PP_F32 && dtype_out == PP_F32) hipLaunchKernelGGL((k_cast<float, float>), g, dim3(256), 0, s, (const float *)in, (float *)out, count);
    else { set_error("pp_cast: dtype"); return PP_ERR_ARG; }
    PP_LAUNCH_CHECK();
    return PP_OK;
}

} // extern "C"

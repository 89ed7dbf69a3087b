// Krylov solvers of the gfx950 library, part 1: the stopping rules live in pp_krylov.hpp; this
// file holds the small-system engine (a whole solve inside one workgroup) and the sparse
// utility kernels (SpMV, ELL / sliced ELL copies, transposes, outer products, Galerkin maps,
// inverse diagonals). The multi-workgroup engine and the AMG V-cycle are in pp_solve.hip.
//
// The recurrences, stopping rules and reported iteration counts restate the reference's fused
// solvers (csrc/phipict/solvers/krylov/krylov.cu: CG :291-568, BiCGStab :598-824); reductions
// accumulate in double, as there.
#include "pp_krylov.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace pp {

// ======================================================================================
// Small systems (n <= SMALL_N): one 1024-thread workgroup solves a group of up to NB
// right-hand sides sharing a matrix, start to finish. Every reduction is an LDS tree inside
// the workgroup, the records live in LDS while the solve runs (thread 0 applies the rules),
// so a solve is one launch and many groups run side by side.
// ======================================================================================

// The right-hand sides of workgroup `wg`: matrix m, right-hand sides first .. first+count-1.
struct SmallSpan { int m, first, count; };
__device__ __forceinline__ SmallSpan small_span(int wg, int nbPerGroup, int rpm, int B) {
    const int chunks = (rpm + nbPerGroup - 1) / nbPerGroup; // groups per matrix
    SmallSpan s;
    s.m = wg / chunks;
    const int lead = (wg % chunks) * nbPerGroup;
    s.first = s.m * rpm + lead;
    s.count = max(0, min(min(nbPerGroup, rpm - lead), B - s.first));
    return s;
}

// Work vectors of one right-hand side: [slot][n] for the slots below.
enum SmallSlot { S_R = 0, S_P0 = 1, S_P1 = 2, S_Q = 3, S_XB = 4, S_BW = 5, S_R0 = 6, S_V = 7, S_T = 8, S_SLOTS = SMALL_SLOTS };

template <typename T>
__device__ __forceinline__ T *small_vec(T *work, int n, int rhs, int slot) {
    return work + ((size_t)rhs * S_SLOTS + slot) * n;
}

// Row i of the group's matrix applied to NB vectors: out[k] = sum_j a_ij * f(k, j).
template <typename T, int NB, typename F>
__device__ __forceinline__ void small_row(const Csr<T> &A, const T *__restrict__ val, int i, T (&out)[NB], F f) {
#pragma unroll
    for (int k = 0; k < NB; ++k) out[k] = T(0);
    const int end = A.row[i + 1];
    for (int e = A.row[i]; e < end; ++e) {
        const int c = A.col[e];
        const T a = val[e];
#pragma unroll
        for (int k = 0; k < NB; ++k) out[k] += a * f(k, c);
    }
}

// Workgroup sums of acc[NV] broadcast to every thread through LDS.
template <int NV>
__device__ __forceinline__ void small_allsum(double (&acc)[NV], double *red) {
    block_sum<NV>(acc, red);
    __shared__ double bcast[16];
    if (threadIdx.x == 0)
#pragma unroll
        for (int q = 0; q < NV; ++q) bcast[q] = acc[q];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NV; ++q) acc[q] = bcast[q];
    __syncthreads();
}

// CG; with dinv the Jacobi-preconditioned variant (z = D^-1 r, rho = (r, z)).
template <typename T, int NB>
__global__ void __launch_bounds__(SMALL_BLOCK) k_small_cg(Csr<T> A, const T *__restrict__ b, T *__restrict__ x,
                                                          T *__restrict__ work, SolveRec *rec, const T *__restrict__ dinv,
                                                          SmallArgs a) {
    __shared__ SolveRec R[NB];
    __shared__ double red[16 * 2 * NB];
    __shared__ double beta[NB], alpha[NB];
    static_assert(2 * NB <= 16, "reduction width");
    const SmallSpan sp = small_span(blockIdx.x, NB, a.rpm, a.B);
    if (sp.count <= 0) return;
    const int n = a.n, tid = threadIdx.x;
    const T *val = A.val + (size_t)sp.m * A.valStride;
    const T *dv = dinv ? dinv + (size_t)sp.m * a.dinvStride : nullptr;
    auto zv = [&](const T *rk, int i) { return dv ? rk[i] * dv[i] : rk[i]; }; // preconditioned residual
    T *r[NB], *pv[2][NB], *q[NB], *xb[NB], *bw[NB], *xs[NB];
    bool on[NB];
#pragma unroll
    for (int k = 0; k < NB; ++k) {
        const int h = sp.first + (k < sp.count ? k : 0);
        on[k] = k < sp.count;
        r[k] = small_vec(work, n, h, S_R);
        pv[0][k] = small_vec(work, n, h, S_P0);
        pv[1][k] = small_vec(work, n, h, S_P1);
        q[k] = small_vec(work, n, h, S_Q);
        xb[k] = small_vec(work, n, h, S_XB);
        bw[k] = small_vec(work, n, h, S_BW);
        xs[k] = x + (size_t)h * n;
    }
    if (tid < NB && on[tid]) R[tid] = rec[sp.first + tid];
    // working right-hand side: b, or b minus its mean (rank-deficient systems)
    double mean[NB];
#pragma unroll
    for (int k = 0; k < NB; ++k) mean[k] = 0.0;
    if (a.project) {
        for (int i = tid; i < n; i += SMALL_BLOCK)
#pragma unroll
            for (int k = 0; k < NB; ++k)
                if (on[k]) mean[k] += double(b[(size_t)(sp.first + k) * n + i]);
        small_allsum<NB>(mean, red);
    }
    for (int i = tid; i < n; i += SMALL_BLOCK)
#pragma unroll
        for (int k = 0; k < NB; ++k)
            if (on[k]) {
                const T bi = b[(size_t)(sp.first + k) * n + i];
                bw[k][i] = a.project ? bi - T(mean[k] / double(n)) : bi;
            }
    __syncthreads();
    // r = bw - A x; (r, r) and (r, z) of the selected right-hand sides
    auto residual = [&](const bool (&sel)[NB], double (&rr)[2 * NB]) {
#pragma unroll
        for (int k = 0; k < 2 * NB; ++k) rr[k] = 0.0;
        for (int i = tid; i < n; i += SMALL_BLOCK) {
            T ax[NB];
            small_row<T, NB>(A, val, i, ax, [&](int k, int c) { return sel[k] ? xs[k][c] : T(0); });
#pragma unroll
            for (int k = 0; k < NB; ++k)
                if (sel[k]) {
                    const T ri = bw[k][i] - ax[k];
                    r[k][i] = ri;
                    rr[2 * k] += double(ri) * double(ri);
                    rr[2 * k + 1] += double(ri) * double(zv(r[k], i));
                }
        }
        small_allsum<2 * NB>(rr, red);
    };
    {
        double rr[2 * NB];
        residual(on, rr);
        if (tid == 0)
            for (int k = 0; k < sp.count; ++k) {
                rule_open(R[k], rr[2 * k], a.normScale, false);
                if (dv) R[k].rho = rr[2 * k + 1];
            }
        __syncthreads();
    }
    for (int loop = 0, cur = 0; loop < a.maxit + 2; ++loop, cur ^= 1) {
        bool act[NB], any = false;
#pragma unroll
        for (int k = 0; k < NB; ++k) { act[k] = on[k] && rec_active(R[k]); any |= act[k]; }
        if (!any) break;
        if (a.resetEvery > 0) {
            // residual replacement: r = bw - A x, a fresh direction
            bool due[NB], anyDue = false;
#pragma unroll
            for (int k = 0; k < NB; ++k) {
                due[k] = act[k] && R[k].iter > 0 && R[k].iter % a.resetEvery == 0;
                anyDue |= due[k];
            }
            if (anyDue) {
                double rr[2 * NB];
                residual(due, rr);
                if (tid == 0)
                    for (int k = 0; k < sp.count; ++k)
                        if (due[k]) { R[k].rho = dv ? rr[2 * k + 1] : rr[2 * k]; R[k].flags |= R_FRESH; }
                __syncthreads();
            }
        }
        // direction p' = r + beta p (r when fresh), q = A p', alpha = rho / (p', q)
        {
            bool fresh[NB];
            T bt[NB];
#pragma unroll
            for (int k = 0; k < NB; ++k) { fresh[k] = (R[k].flags & R_FRESH) != 0; bt[k] = T(beta[k]); }
            T *const *pin = pv[cur];
            T *const *pout = pv[cur ^ 1];
            double pq[NB];
#pragma unroll
            for (int k = 0; k < NB; ++k) pq[k] = 0.0;
            auto dir = [&](int k, int c) { return fresh[k] ? zv(r[k], c) : zv(r[k], c) + bt[k] * pin[k][c]; };
            for (int i = tid; i < n; i += SMALL_BLOCK) {
                T ap[NB];
                small_row<T, NB>(A, val, i, ap, [&](int k, int c) { return act[k] ? dir(k, c) : T(0); });
#pragma unroll
                for (int k = 0; k < NB; ++k)
                    if (act[k]) {
                        const T pi = dir(k, i);
                        pout[k][i] = pi;
                        q[k][i] = ap[k];
                        pq[k] += double(pi) * double(ap[k]);
                    }
            }
            small_allsum<NB>(pq, red);
            if (tid == 0)
                for (int k = 0; k < sp.count; ++k)
                    if (act[k]) {
                        alpha[k] = rule_alpha(R[k], pq[k]);
                        R[k].flags &= ~R_FRESH;
                    }
            __syncthreads();
        }
        // x += alpha p'; r -= alpha q; the rules on (r, r)
        {
            bool go[NB], keep[NB];
            T al[NB];
#pragma unroll
            for (int k = 0; k < NB; ++k) {
                go[k] = act[k] && rec_active(R[k]);
                keep[k] = go[k] && a.returnBest && (R[k].flags & R_KEEP);
                al[k] = T(alpha[k]);
            }
            T *const *p = pv[cur ^ 1];
            double rr[2 * NB];
#pragma unroll
            for (int k = 0; k < 2 * NB; ++k) rr[k] = 0.0;
            for (int i = tid; i < n; i += SMALL_BLOCK)
#pragma unroll
                for (int k = 0; k < NB; ++k)
                    if (go[k]) {
                        const T xi = xs[k][i];
                        if (keep[k]) xb[k][i] = xi;
                        xs[k][i] = xi + al[k] * p[k][i];
                        const T ri = r[k][i] - al[k] * q[k][i];
                        r[k][i] = ri;
                        rr[2 * k] += double(ri) * double(ri);
                        rr[2 * k + 1] += double(ri) * double(zv(r[k], i));
                    }
            small_allsum<2 * NB>(rr, red);
            if (tid == 0)
                for (int k = 0; k < sp.count; ++k)
                    if (go[k]) {
                        rule_step(R[k], rr[2 * k], a.normScale, a.returnBest != 0);
                        if (rec_active(R[k])) beta[k] = rule_cg_beta(R[k], dv ? rr[2 * k + 1] : rr[2 * k]);
                    }
            __syncthreads();
        }
    }
    // a stopped return-best solve whose best iterate beat the last one gets it back
    if (a.returnBest)
#pragma unroll
        for (int k = 0; k < NB; ++k) {
            const SolveRec s = R[k];
            if (on[k] && (s.fate == F_EXHAUSTED || s.fate == F_STALLED) && s.bestAt >= 0 && s.best < s.crit)
                for (int i = tid; i < n; i += SMALL_BLOCK) xs[k][i] = xb[k][i];
        }
    if (tid < NB && on[tid]) rec[sp.first + tid] = R[tid];
}

// BiCGStab; with dinv right-preconditioned by Jacobi (v = A D^-1 p, t = A D^-1 s, the
// solution updated with D^-1 (alpha p + omega s)).
template <typename T, int NB>
__global__ void __launch_bounds__(SMALL_BLOCK) k_small_bicg(Csr<T> A, const T *__restrict__ b, T *__restrict__ x,
                                                            T *__restrict__ work, SolveRec *rec, const T *__restrict__ dinv,
                                                            SmallArgs a) {
    __shared__ SolveRec R[NB];
    __shared__ double red[16 * 3 * NB];
    __shared__ double beta[NB];
    const SmallSpan sp = small_span(blockIdx.x, NB, a.rpm, a.B);
    if (sp.count <= 0) return;
    const int n = a.n, tid = threadIdx.x;
    const T *val = A.val + (size_t)sp.m * A.valStride;
    const T *dv = dinv ? dinv + (size_t)sp.m * a.dinvStride : nullptr;
    auto pre = [&](T v_, int i) { return dv ? v_ * dv[i] : v_; };
    T *r[NB], *r0[NB], *p[NB], *v[NB], *t[NB], *xs[NB];
    bool on[NB];
#pragma unroll
    for (int k = 0; k < NB; ++k) {
        const int h = sp.first + (k < sp.count ? k : 0);
        on[k] = k < sp.count;
        r[k] = small_vec(work, n, h, S_R);
        p[k] = small_vec(work, n, h, S_P0);
        v[k] = small_vec(work, n, h, S_V);
        r0[k] = small_vec(work, n, h, S_R0);
        t[k] = small_vec(work, n, h, S_T);
        xs[k] = x + (size_t)h * n;
    }
    if (tid < NB && on[tid]) R[tid] = rec[sp.first + tid];
    {
        // r = r0 = b - A x; p = v = 0
        double rr[NB];
#pragma unroll
        for (int k = 0; k < NB; ++k) rr[k] = 0.0;
        for (int i = tid; i < n; i += SMALL_BLOCK) {
            T ax[NB];
            small_row<T, NB>(A, val, i, ax, [&](int k, int c) { return on[k] ? xs[k][c] : T(0); });
#pragma unroll
            for (int k = 0; k < NB; ++k)
                if (on[k]) {
                    const T ri = b[(size_t)(sp.first + k) * n + i] - ax[k];
                    r[k][i] = ri;
                    r0[k][i] = ri;
                    p[k][i] = T(0);
                    v[k][i] = T(0);
                    rr[k] += double(ri) * double(ri);
                }
        }
        small_allsum<NB>(rr, red);
        if (tid == 0)
            for (int k = 0; k < sp.count; ++k) rule_open(R[k], rr[k], a.normScale, true);
        __syncthreads();
    }
    for (int loop = 0; loop < a.maxit + 2; ++loop) {
        bool act[NB], any = false;
#pragma unroll
        for (int k = 0; k < NB; ++k) { act[k] = on[k] && rec_active(R[k]); any |= act[k]; }
        if (!any) break;
        // p = r + beta (p - omega v) (r when fresh); v = A p; alpha = rho / (r0, v)
        {
            bool fresh[NB];
            T bt[NB], om[NB];
#pragma unroll
            for (int k = 0; k < NB; ++k) {
                fresh[k] = (R[k].flags & R_FRESH) != 0;
                bt[k] = T(beta[k]);
                om[k] = T(R[k].omega);
            }
            for (int i = tid; i < n; i += SMALL_BLOCK)
#pragma unroll
                for (int k = 0; k < NB; ++k)
                    if (act[k]) p[k][i] = fresh[k] ? r[k][i] : r[k][i] + bt[k] * (p[k][i] - om[k] * v[k][i]);
            __syncthreads();
            double r0v[NB];
#pragma unroll
            for (int k = 0; k < NB; ++k) r0v[k] = 0.0;
            for (int i = tid; i < n; i += SMALL_BLOCK) {
                T ap[NB];
                small_row<T, NB>(A, val, i, ap, [&](int k, int c) { return act[k] ? pre(p[k][c], c) : T(0); });
#pragma unroll
                for (int k = 0; k < NB; ++k)
                    if (act[k]) {
                        v[k][i] = ap[k];
                        r0v[k] += double(r0[k][i]) * double(ap[k]);
                    }
            }
            small_allsum<NB>(r0v, red);
            if (tid == 0)
                for (int k = 0; k < sp.count; ++k)
                    if (act[k]) {
                        R[k].alpha = rule_alpha(R[k], r0v[k]);
                        R[k].flags &= ~R_FRESH;
                    }
            __syncthreads();
        }
        // s = r - alpha v, t = A s; omega from (t, s) / (t, t), half-step test on (s, s)
        bool go[NB];
        {
            T al[NB];
#pragma unroll
            for (int k = 0; k < NB; ++k) { go[k] = act[k] && rec_active(R[k]); al[k] = T(R[k].alpha); }
            double acc[3 * NB];
#pragma unroll
            for (int k = 0; k < 3 * NB; ++k) acc[k] = 0.0;
            for (int i = tid; i < n; i += SMALL_BLOCK) {
                T as[NB];
                small_row<T, NB>(A, val, i, as, [&](int k, int c) { return go[k] ? pre(r[k][c] - al[k] * v[k][c], c) : T(0); });
#pragma unroll
                for (int k = 0; k < NB; ++k)
                    if (go[k]) {
                        const T si = r[k][i] - al[k] * v[k][i];
                        t[k][i] = as[k];
                        acc[3 * k] += double(as[k]) * double(si);
                        acc[3 * k + 1] += double(as[k]) * double(as[k]);
                        acc[3 * k + 2] += double(si) * double(si);
                    }
            }
            small_allsum<3 * NB>(acc, red);
            if (tid == 0)
                for (int k = 0; k < sp.count; ++k)
                    if (go[k]) R[k].omega = rule_bicg_half(R[k], acc[3 * k], acc[3 * k + 1], acc[3 * k + 2], a.normScale);
            __syncthreads();
        }
        // x += alpha p + omega s, r = s - omega t (the half step: x += alpha p, r = s)
        {
            bool fin[NB], half[NB];
            T al[NB], om[NB];
#pragma unroll
            for (int k = 0; k < NB; ++k) {
                fin[k] = go[k] && rec_active(R[k]);
                half[k] = (R[k].flags & R_HALF) != 0;
                al[k] = T(R[k].alpha);
                om[k] = T(R[k].omega);
            }
            double acc[2 * NB];
#pragma unroll
            for (int k = 0; k < 2 * NB; ++k) acc[k] = 0.0;
            for (int i = tid; i < n; i += SMALL_BLOCK)
#pragma unroll
                for (int k = 0; k < NB; ++k)
                    if (fin[k]) {
                        const T si = r[k][i] - al[k] * v[k][i];
                        if (half[k]) {
                            xs[k][i] += pre(al[k] * p[k][i], i);
                            r[k][i] = si;
                        } else {
                            xs[k][i] += pre(al[k] * p[k][i] + om[k] * si, i);
                            const T ri = si - om[k] * t[k][i];
                            r[k][i] = ri;
                            acc[2 * k] += double(r0[k][i]) * double(ri);
                            acc[2 * k + 1] += double(ri) * double(ri);
                        }
                    }
            small_allsum<2 * NB>(acc, red);
            if (tid == 0)
                for (int k = 0; k < sp.count; ++k)
                    if (fin[k]) {
                        if (half[k]) rule_bicg_half_met(R[k]);
                        else beta[k] = rule_bicg_full(R[k], acc[2 * k], acc[2 * k + 1], a.normScale);
                    }
            __syncthreads();
        }
    }
    if (tid < NB && on[tid]) rec[sp.first + tid] = R[tid];
}

template <typename T>
void launch_small(const Csr<T> &A, const T *b, T *x, void *work, SolveRec *rec, const T *dinv, const SmallArgs &a,
                  int groups, hipStream_t s) {
    T *w = reinterpret_cast<T *>(work);
    switch (a.nb_per_group) {
#define PP_SMALL_CASE(NBV)                                                                                                \
    case NBV:                                                                                                             \
        if (a.bicg) hipLaunchKernelGGL((k_small_bicg<T, NBV>), dim3(groups), dim3(SMALL_BLOCK), 0, s, A, b, x, w, rec, dinv, a); \
        else hipLaunchKernelGGL((k_small_cg<T, NBV>), dim3(groups), dim3(SMALL_BLOCK), 0, s, A, b, x, w, rec, dinv, a);    \
        break;
        PP_SMALL_CASE(1)
        PP_SMALL_CASE(2)
        PP_SMALL_CASE(3)
        default: PP_SMALL_CASE(4)
#undef PP_SMALL_CASE
    }
}
template void launch_small<float>(const Csr<float> &, const float *, float *, void *, SolveRec *, const float *,
                                  const SmallArgs &, int, hipStream_t);
template void launch_small<double>(const Csr<double> &, const double *, double *, void *, SolveRec *, const double *,
                                   const SmallArgs &, int, hipStream_t);

// ======================================================================================
// Sparse utility kernels
// ======================================================================================

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

// Wave-interleaved sliced ELL copy (pp_sell layout) of a CSR matrix: slices of 64 / lanes
// rows; entry k * lanes + j of row r of slice s sits at slice_off[s] + 64 k + r lanes + j
// (k < slice_steps[s]), so one wavefront-wide load covers `lanes` consecutive entries of each
// of the slice's rows. One work-item per slot of a slice's 64-wide columns; padding entries
// get column 0 and value 0. Values of nmat matrices (stride nnz) go to stride total.
template <typename T, typename TO>
__global__ void k_sell_fill(const int32_t *__restrict__ row, const int32_t *__restrict__ col, const T *__restrict__ val,
                            int n, int nnz, int lanes, const int32_t *__restrict__ sliceOff,
                            const int32_t *__restrict__ sliceSteps, int64_t total, int32_t *__restrict__ sellCol,
                            TO *__restrict__ sellVal) {
    const int rowsPerSlice = 64 / lanes;
    const int slices = (n + rowsPerSlice - 1) / rowsPerSlice;
    const int slot = blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= slices * 64) return;
    const int s = slot >> 6, l = slot & 63;
    const int i = s * rowsPerSlice + l / lanes, j = l % lanes;
    const int m = blockIdx.y;
    const int first = i < n ? row[i] : 0, len = i < n ? row[i + 1] - first : 0;
    const int64_t base = sliceOff[s] + l;
    for (int k = 0; k < sliceSteps[s]; ++k) {
        const int e = k * lanes + j;
        const bool has = e < len;
        if (sellCol && m == 0) sellCol[base + 64 * k] = has ? col[first + e] : 0;
        if (sellVal) sellVal[(size_t)m * total + base + 64 * k] = has ? TO(val[(size_t)m * nnz + first + e]) : TO(0);
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
int launch_transpose_symmetric(const pp_csr *A, T *tval, int *bad, hipStream_t s) {
    hipLaunchKernelGGL((k_transpose_symmetric<T>), dim3((A->n + 255) / 256), dim3(256), 0, s, A->row, A->col,
                       reinterpret_cast<const T *>(A->val), A->n, A->nnz, A->num_matrices, tval, bad);
    PP_LAUNCH_CHECK();
    return PP_OK;
}
template int launch_transpose_symmetric<float>(const pp_csr *, float *, int *, hipStream_t);
template int launch_transpose_symmetric<double>(const pp_csr *, double *, int *, hipStream_t);

template <typename T>
__global__ void k_outer_product(const T *__restrict__ a, const T *__restrict__ b, const int32_t *__restrict__ row,
                                const int32_t *__restrict__ col, T *__restrict__ out, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const T ai = a[i];
    for (int k = row[i]; k < row[i + 1]; ++k) out[k] = ai * b[col[k]];
}

// Galerkin refresh C = R (A P) as two row products, both by k_rap_rows: O = X Y into the fixed
// pattern of O, O(i, J) = sum_t X(i, t) Y(t, J) over row i of X in column order, every entry
// accumulated in double by one lane (no atomics, the same bits on every run). A group of G lanes
// (one wavefront or a part of one) takes an output row; per chunk of the row of X it stages the
// X entries and the rows of Y they select in LDS with coalesced loads, and each lane then sums
// its own output entries (e = lane + G r) over the staged rows, finding J by binary search in
// each (sorted) row of Y. First product X = A, Y = P into the pattern of A P (double
// workspace); second X = R, Y = A P into the coarse pattern.
constexpr int RAP_WG = 256;
constexpr int RAP_OWN = 4;  // output entries per lane and chunk, X entries per lane and chunk
constexpr int RAP_TPL = 4;  // staged Y entries per lane

// Orders the LDS accesses of a wavefront's lanes around this point (groups never span
// wavefronts, so no workgroup barrier is needed).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int G, typename TX, typename TO>
__global__ void __launch_bounds__(RAP_WG) k_rap_rows(int nOut, const int32_t *__restrict__ xrow,
                                                     const int32_t *__restrict__ xcol, const TX *__restrict__ xval,
                                                     int64_t xStride, const int32_t *__restrict__ yrow,
                                                     const int32_t *__restrict__ ycol, const double *__restrict__ yval,
                                                     int64_t yStride, const int32_t *__restrict__ orow,
                                                     const int32_t *__restrict__ ocol, TO *__restrict__ oval,
                                                     int64_t oStride) {
    constexpr int GPW = RAP_WG / G, CAP = RAP_OWN * G, TCAP = RAP_TPL * G;
    __shared__ int sYb[GPW][CAP];
    __shared__ int sPre[GPW][CAP + 1];
    __shared__ double sXv[GPW][CAP];
    __shared__ int sYc[GPW][TCAP];
    __shared__ double sYv[GPW][TCAP];
    const int grp = threadIdx.x / G, lane = threadIdx.x % G;
    const int i = blockIdx.x * GPW + grp;
    if (i >= nOut) return;
    int *yb = sYb[grp], *pre = sPre[grp], *yc = sYc[grp];
    double *xv = sXv[grp], *yv = sYv[grp];
    const TX *X = xval + (int64_t)blockIdx.y * xStride;
    const double *Y = yval + (int64_t)blockIdx.y * yStride;
    TO *O = oval + (int64_t)blockIdx.y * oStride;
    const int ob = orow[i], m = orow[i + 1] - ob;
    const int xb = xrow[i], nx = xrow[i + 1] - xb;
    for (int oc = 0; oc < m; oc += CAP) {
        int J[RAP_OWN];
        double acc[RAP_OWN];
#pragma unroll
        for (int r = 0; r < RAP_OWN; ++r) {
            const int e = oc + lane + G * r;
            J[r] = e < m ? ocol[ob + e] : -1;
            acc[r] = 0.0;
        }
        for (int xc = 0; xc < nx; xc += CAP) {
            const int nxc = nx - xc < CAP ? nx - xc : CAP;
            for (int t = lane; t < nxc; t += G) {
                const int k = xcol[xb + xc + t];
                const int b0 = yrow[k];
                yb[t] = b0;
                pre[t + 1] = yrow[k + 1] - b0;
                xv[t] = double(X[xb + xc + t]);
            }
            wave_sync();
            // offsets of the staged rows: a blocked scan over the group's lanes
            int len[RAP_OWN], tot = 0;
#pragma unroll
            for (int r = 0; r < RAP_OWN; ++r) {
                const int t = lane * RAP_OWN + r;
                len[r] = t < nxc ? pre[t + 1] : 0;
                tot += len[r];
            }
            int inc = tot;
#pragma unroll
            for (int o = 1; o < G; o <<= 1) {
                const int v = __shfl_up(inc, o, G);
                if (lane >= o) inc += v;
            }
            wave_sync();
            int run = inc - tot;
#pragma unroll
            for (int r = 0; r < RAP_OWN; ++r) {
                const int t = lane * RAP_OWN + r;
                if (t < nxc) {
                    run += len[r];
                    pre[t + 1] = run;
                }
            }
            if (lane == 0) pre[0] = 0;
            wave_sync();
            for (int t0 = 0; t0 < nxc;) {
                // the staged rows t0..t1-1 fit in TCAP entries (a row of Y is at most TCAP long)
                const int base = pre[t0];
                int lo = t0 + 1, hi = nxc;
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (pre[mid] - base <= TCAP) lo = mid;
                    else hi = mid - 1;
                }
                const int t1 = lo, cnt = pre[t1] - base;
                for (int j = lane; j < cnt; j += G) {
                    int a = t0, z = t1 - 1; // the row holding entry j: the last t with pre[t] - base <= j
                    while (a < z) {
                        const int mid = (a + z + 1) >> 1;
                        if (pre[mid] - base <= j) a = mid;
                        else z = mid - 1;
                    }
                    const int q = yb[a] + (j + base - pre[a]);
                    yc[j] = ycol[q];
                    yv[j] = Y[q];
                }
                wave_sync();
#pragma unroll
                for (int r = 0; r < RAP_OWN; ++r) {
                    if (J[r] < 0) continue;
                    double s = acc[r];
                    for (int t = t0; t < t1; ++t) {
                        int a = pre[t] - base, z = pre[t + 1] - base;
                        while (a < z) {
                            const int mid = (a + z) >> 1;
                            if (yc[mid] < J[r]) a = mid + 1;
                            else z = mid;
                        }
                        if (a < pre[t + 1] - base && yc[a] == J[r]) s += xv[t] * yv[a];
                    }
                    acc[r] = s;
                }
                wave_sync();
                t0 = t1;
            }
        }
#pragma unroll
        for (int r = 0; r < RAP_OWN; ++r)
            if (J[r] >= 0) O[ob + oc + lane + G * r] = TO(acc[r]);
    }
}

// The same product with one lane per output row, for output rows of at most MAXR entries: the
// row's columns and sums stay in registers, and each term finds its entry by an unrolled compare
// against all MAXR columns (no indexed register access). The row of X is taken RAP_RU entries at
// a time: their columns, values and Y row bounds are loaded together, then the Y rows are walked
// side by side (entry s of each), so every round of dependent loads serves RAP_RU terms. A sum's
// terms arrive in this fixed (chunk, s, u) order on every run.
constexpr int RAP_RU = 4;
template <int MAXR, typename TX, typename TO>
__global__ void __launch_bounds__(RAP_WG) k_rap_reg(int nOut, const int32_t *__restrict__ xrow,
                                                    const int32_t *__restrict__ xcol, const TX *__restrict__ xval,
                                                    int64_t xStride, const int32_t *__restrict__ yrow,
                                                    const int32_t *__restrict__ ycol, const double *__restrict__ yval,
                                                    int64_t yStride, const int32_t *__restrict__ orow,
                                                    const int32_t *__restrict__ ocol, TO *__restrict__ oval,
                                                    int64_t oStride) {
    const int i = blockIdx.x * RAP_WG + threadIdx.x;
    if (i >= nOut) return;
    const TX *X = xval + (int64_t)blockIdx.y * xStride;
    const double *Y = yval + (int64_t)blockIdx.y * yStride;
    TO *O = oval + (int64_t)blockIdx.y * oStride;
    const int ob = orow[i], m = orow[i + 1] - ob;
    int cols[MAXR];
    double acc[MAXR];
#pragma unroll
    for (int r = 0; r < MAXR; ++r) {
        cols[r] = r < m ? ocol[ob + r] : -1;
        acc[r] = 0.0;
    }
    const int xe = xrow[i + 1];
    for (int t0 = xrow[i]; t0 < xe; t0 += RAP_RU) {
        int yb[RAP_RU], len[RAP_RU], k[RAP_RU];
        double xv[RAP_RU];
#pragma unroll
        for (int u = 0; u < RAP_RU; ++u) {
            const bool ok = t0 + u < xe;
            k[u] = ok ? xcol[t0 + u] : -1;
            xv[u] = ok ? double(X[t0 + u]) : 0.0;
        }
        int mlen = 0;
#pragma unroll
        for (int u = 0; u < RAP_RU; ++u) {
            yb[u] = k[u] >= 0 ? yrow[k[u]] : 0;
            len[u] = k[u] >= 0 ? yrow[k[u] + 1] - yb[u] : 0;
            mlen = len[u] > mlen ? len[u] : mlen;
        }
        for (int s = 0; s < mlen; ++s) {
            int c[RAP_RU];
            double v[RAP_RU];
#pragma unroll
            for (int u = 0; u < RAP_RU; ++u) {
                const bool ok = s < len[u];
                c[u] = ok ? ycol[yb[u] + s] : -1;
                v[u] = ok ? xv[u] * Y[yb[u] + s] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < RAP_RU; ++u)
#pragma unroll
                for (int r = 0; r < MAXR; ++r)
                    if (cols[r] == c[u]) acc[r] += v[u];
        }
    }
#pragma unroll
    for (int r = 0; r < MAXR; ++r)
        if (r < m) O[ob + r] = TO(acc[r]);
}

// The same product for rows of Y longer than the staging area: one lane per output row, J found
// by binary search in the rows of Y in global memory (slow; no such rows arise in Ruge-Stueben
// hierarchies of the PISO systems).
template <typename TX, typename TO>
__global__ void __launch_bounds__(RAP_WG) k_rap_rows_long(int nOut, const int32_t *__restrict__ xrow,
                                                          const int32_t *__restrict__ xcol, const TX *__restrict__ xval,
                                                          int64_t xStride, const int32_t *__restrict__ yrow,
                                                          const int32_t *__restrict__ ycol,
                                                          const double *__restrict__ yval, int64_t yStride,
                                                          const int32_t *__restrict__ orow,
                                                          const int32_t *__restrict__ ocol, TO *__restrict__ oval,
                                                          int64_t oStride) {
    const int i = blockIdx.x * RAP_WG + threadIdx.x;
    if (i >= nOut) return;
    const TX *X = xval + (int64_t)blockIdx.y * xStride;
    const double *Y = yval + (int64_t)blockIdx.y * yStride;
    TO *O = oval + (int64_t)blockIdx.y * oStride;
    for (int e = orow[i]; e < orow[i + 1]; ++e) {
        const int J = ocol[e];
        double s = 0.0;
        for (int t = xrow[i]; t < xrow[i + 1]; ++t) {
            const int k = xcol[t];
            int a = yrow[k], z = yrow[k + 1];
            while (a < z) {
                const int mid = (a + z) >> 1;
                if (ycol[mid] < J) a = mid + 1;
                else z = mid;
            }
            if (a < yrow[k + 1] && ycol[a] == J) s += double(X[t]) * Y[a];
        }
        O[e] = TO(s);
    }
}

// Kernels of one row product: RAP_AUTO chooses by the longest rows; the others name a kernel
// (the host autotuner times the ones that fit and keeps the fastest per level and product).
enum RapKernel { RAP_AUTO = 0, RAP_REG16, RAP_REG32, RAP_ROWS8, RAP_ROWS16, RAP_ROWS32, RAP_ROWS64, RAP_LONG, RAP_KINDS };

// Whether kernel `kind` fits rows of X up to maxX, of Y up to maxY and outputs up to maxOut
// entries (k_rap_rows stages whole rows of Y: at most RAP_TPL G entries).
static bool rap_fits(int kind, int maxY, int maxOut) {
    switch (kind) {
    case RAP_REG16: return maxOut <= 16;
    case RAP_REG32: return maxOut <= 32;
    case RAP_ROWS8: return maxY <= RAP_TPL * 8;
    case RAP_ROWS16: return maxY <= RAP_TPL * 16;
    case RAP_ROWS32: return maxY <= RAP_TPL * 32;
    case RAP_ROWS64: return maxY <= RAP_TPL * 64;
    case RAP_LONG: return true;
    default: return false;
    }
}

// The rule without tuning: one lane per row for short rows of X and outputs of up to 32
// entries; otherwise the fewest lanes per row whose chunks hold a row of Y and cover the longer
// rows in a few chunks; the one-lane fallback for rows of Y beyond the staging area.
static int rap_rule(int maxX, int maxY, int maxOut) {
    if (maxOut <= 32 && maxX <= 8) return maxOut <= 16 ? RAP_REG16 : RAP_REG32;
    if (maxY > RAP_TPL * 64) return RAP_LONG;
    const int row = std::max(maxX, maxOut);
    int G = 8, kind = RAP_ROWS8;
    while (G < 64 && (RAP_TPL * G < maxY || RAP_OWN * G < row)) {
        G <<= 1;
        ++kind;
    }
    return kind;
}

template <typename TX, typename TO>
static int launch_rap(int kind, int nOut, int maxX, int maxY, int maxOut, int nmat, const int32_t *xrow,
                      const int32_t *xcol, const TX *xval, int64_t xStride, const int32_t *yrow, const int32_t *ycol,
                      const double *yval, int64_t yStride, const int32_t *orow, const int32_t *ocol, TO *oval,
                      int64_t oStride, hipStream_t s) {
    if (kind == RAP_AUTO) kind = rap_rule(maxX, maxY, maxOut);
    PP_CHECK_ARG(kind > RAP_AUTO && kind < RAP_KINDS, "unknown row-product kernel");
    PP_CHECK_ARG(rap_fits(kind, maxY, maxOut), "the row-product kernel does not fit these rows");
#define RAP_REG(MR)                                                                                               \
    hipLaunchKernelGGL((k_rap_reg<MR, TX, TO>), dim3((nOut + RAP_WG - 1) / RAP_WG, nmat), dim3(RAP_WG), 0, s, nOut, \
                       xrow, xcol, xval, xStride, yrow, ycol, yval, yStride, orow, ocol, oval, oStride)
#define RAP_GO(GV)                                                                                                \
    hipLaunchKernelGGL((k_rap_rows<GV, TX, TO>), dim3((nOut + RAP_WG / GV - 1) / (RAP_WG / GV), nmat), dim3(RAP_WG), \
                       0, s, nOut, xrow, xcol, xval, xStride, yrow, ycol, yval, yStride, orow, ocol, oval, oStride)
    switch (kind) {
    case RAP_REG16: RAP_REG(16); break;
    case RAP_REG32: RAP_REG(32); break;
    case RAP_ROWS8: RAP_GO(8); break;
    case RAP_ROWS16: RAP_GO(16); break;
    case RAP_ROWS32: RAP_GO(32); break;
    case RAP_ROWS64: RAP_GO(64); break;
    default:
        hipLaunchKernelGGL((k_rap_rows_long<TX, TO>), dim3((nOut + RAP_WG - 1) / RAP_WG, nmat), dim3(RAP_WG), 0, s,
                           nOut, xrow, xcol, xval, xStride, yrow, ycol, yval, yStride, orow, ocol, oval, oStride);
        break;
    }
#undef RAP_GO
#undef RAP_REG
    return PP_OK;
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

// Inverse of a dense coarsest operator, one workgroup per operator: Gauss-Jordan elimination
// with partial pivoting on [M | I] held in registers. Thread t owns column t % 256 of the 2n
// augmented columns and the rows t / 256 + 4 j (j < 32), so n <= 128. Pivot rows are not
// swapped: a pivot row is marked used, and row r of the right half ends as row iperm[r] of
// the inverse. With rank_one, M = C + (s / n) 1 1^T with s = trace(C) / n and the result is
// inv(M) - 1 / (s n) 1 1^T: the pseudo-inverse of a symmetric C whose null space is the
// constant vector (for C + s e e^T with unit e spanning both null spaces of C,
// inv(C + s e e^T) = pinv(C) + e e^T / s).
constexpr int CINV_MAX = 128;
constexpr int CINV_THREADS = 1024;
constexpr int CINV_ROWS = CINV_MAX / (CINV_THREADS / (2 * CINV_MAX)); // rows per thread: 32

__global__ void __launch_bounds__(CINV_THREADS) k_coarse_inverse(const double *__restrict__ Cm,
                                                                 double *__restrict__ out, int n, int rank_one) {
    __shared__ double colk[CINV_MAX];
    __shared__ double prow[2 * CINV_MAX];
    __shared__ int iperm[CINV_MAX];
    __shared__ int used[CINV_MAX];
    __shared__ int piv_s;
    __shared__ double shift_s;
    const int t = threadIdx.x, c = t & (2 * CINV_MAX - 1), rg = t / (2 * CINV_MAX);
    const int groups = CINV_THREADS / (2 * CINV_MAX);
    const size_t nn = (size_t)n * n;
    const double *C = Cm + (size_t)blockIdx.x * nn;
    double *o = out + (size_t)blockIdx.x * nn;
    if (t < n) {
        colk[t] = C[(size_t)t * n + t];
        used[t] = 0;
    }
    __syncthreads();
    if (t == 0) {
        double tr = 0.0;
        for (int i = 0; i < n; ++i) tr += colk[i];
        shift_s = rank_one ? tr / n : 0.0;
    }
    __syncthreads();
    const double add = shift_s / n;
    double a[CINV_ROWS];
#pragma unroll
    for (int j = 0; j < CINV_ROWS; ++j) {
        const int r = rg + groups * j;
        double v = 0.0;
        if (r < n && c < 2 * n) v = c < n ? C[(size_t)r * n + c] + add : (c - n == r ? 1.0 : 0.0);
        a[j] = v;
    }
    for (int k = 0; k < n; ++k) {
        __syncthreads(); // the previous step's reads of colk and prow are done
        if (c == k) {
#pragma unroll
            for (int j = 0; j < CINV_ROWS; ++j) {
                const int r = rg + groups * j;
                if (r < n) colk[r] = a[j];
            }
        }
        __syncthreads();
        if (t < 64) {
            // pivot: the unused row of largest |entry| in column k, the lowest index on ties
            // (a NaN entry counts as 0, so every step finds an unused row: bi < n)
            double best = -1.0;
            int bi = CINV_MAX;
            for (int r = t; r < n; r += 64) {
                const double v = fabs(colk[r]);
                const double key = v == v ? v : 0.0;
                if (!used[r] && key > best) {
                    best = key;
                    bi = r;
                }
            }
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1) {
                const double ob = __shfl_xor(best, off, 64);
                const int oi = __shfl_xor(bi, off, 64);
                if (ob > best || (ob == best && oi < bi)) {
                    best = ob;
                    bi = oi;
                }
            }
            if (t == 0) {
                piv_s = bi;
                used[bi] = 1;
                iperm[bi] = k;
            }
        }
        __syncthreads();
        const int piv = piv_s;
        const double rp = 1.0 / colk[piv];
#pragma unroll
        for (int j = 0; j < CINV_ROWS; ++j)
            if (rg + groups * j == piv) {
                a[j] *= rp;
                if (c < 2 * n) prow[c] = a[j];
            }
        __syncthreads();
        const double pc = c < 2 * n ? prow[c] : 0.0;
#pragma unroll
        for (int j = 0; j < CINV_ROWS; ++j) {
            const int r = rg + groups * j;
            if (r < n && r != piv) a[j] -= colk[r] * pc;
        }
    }
    __syncthreads();
    if (c >= n && c < 2 * n) {
        const double corr = rank_one ? 1.0 / (shift_s * n) : 0.0;
#pragma unroll
        for (int j = 0; j < CINV_ROWS; ++j) {
            const int r = rg + groups * j;
            if (r < n) o[(size_t)iperm[r] * n + (c - n)] = a[j] - corr;
        }
    }
}

} // namespace pp

using namespace pp;

extern "C" {

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

int pp_sell_fill(const int32_t *row, const int32_t *col, const void *val, int32_t n, int32_t nnz, int32_t nmat,
                 int32_t lanes, const int32_t *slice_off, const int32_t *slice_steps, int64_t total, int32_t dtype,
                 int32_t *sell_col, void *sell_val, int32_t sell_dtype, void *stream) {
    PP_CHECK_ARG(row && slice_off && slice_steps && n >= 0 && nmat >= 1, "bad arguments");
    PP_CHECK_ARG(lanes == 1 || lanes == 2 || lanes == 4 || lanes == 8 || lanes == 16 || lanes == 32, "lanes must be a power of two <= 32");
    PP_CHECK_ARG(!sell_col || col, "sell_col needs col");
    PP_CHECK_ARG(!sell_val || val, "sell_val needs val");
    if (n == 0) return PP_OK;
    const int slices = (n + 64 / lanes - 1) / (64 / lanes);
    const dim3 g((slices * 64 + 255) / 256, sell_val ? nmat : 1);
    hipStream_t s = as_stream(stream);
#define PP_FILL(TI, TO_) hipLaunchKernelGGL((k_sell_fill<TI, TO_>), g, dim3(256), 0, s, row, col, (const TI *)val, n, nnz, lanes, \
                                            slice_off, slice_steps, total, sell_col, (TO_ *)sell_val)
    if (!sell_val) PP_FILL(float, float);
    else if (dtype == PP_F64 && sell_dtype == PP_F64) PP_FILL(double, double);
    else if (dtype == PP_F64 && sell_dtype == PP_F32) PP_FILL(double, float);
    else if (dtype == PP_F32 && sell_dtype == PP_F32) PP_FILL(float, float);
    else if (dtype == PP_F32 && sell_dtype == PP_F64) PP_FILL(float, double);
    else { set_error("pp_sell_fill: dtype"); return PP_ERR_ARG; }
#undef PP_FILL
    PP_LAUNCH_CHECK();
    return PP_OK;
}

int pp_galerkin_rows(const pp_csr *A, int32_t a_dtype, const int32_t *p_row, const int32_t *p_col, const double *p_val,
                     const int32_t *ap_row, const int32_t *ap_col, int32_t ap_nnz, double *ap_val,
                     const int32_t *r_row, const int32_t *r_col, const double *r_val, const int32_t *c_row,
                     const int32_t *c_col, void *c_val, int32_t c_n, int32_t c_nnz, int32_t c_dtype,
                     const int32_t *max_rows, const int32_t *plan, int32_t stages, void *stream) {
    PP_CHECK_ARG(A && A->row && A->col && A->val && p_row && p_col && p_val && ap_row && ap_col && ap_val && r_row &&
                 r_col && r_val && c_row && c_col && c_val && max_rows, "null argument");
    PP_CHECK_ARG(A->n >= 0 && c_n >= 0 && ap_nnz >= 0 && c_nnz >= 0 && A->num_matrices >= 1, "sizes");
    PP_CHECK_ARG(a_dtype == PP_F64 || a_dtype == PP_F32, "a_dtype");
    PP_CHECK_ARG(c_dtype == PP_F64 || c_dtype == PP_F32, "c_dtype");
    PP_CHECK_ARG(stages >= 1 && stages <= 3, "stages");
    if (c_n == 0 || A->n == 0) return PP_OK;
    hipStream_t s = as_stream(stream);
    const int nmat = A->num_matrices;
    // max_rows: the longest rows of A, P, A P, R and C
    const int k1 = plan ? plan[0] : RAP_AUTO, k2 = plan ? plan[1] : RAP_AUTO;
    int rc = PP_OK;
    if (stages & 1)
        PP_TIMED(PP_K_GALERKIN_AP, A->n, s, {
            if (a_dtype == PP_F64)
                rc = launch_rap<double, double>(k1, A->n, max_rows[0], max_rows[1], max_rows[2], nmat, A->row, A->col,
                                                (const double *)A->val, (int64_t)A->nnz, p_row, p_col, p_val, 0,
                                                ap_row, ap_col, ap_val, (int64_t)ap_nnz, s);
            else
                rc = launch_rap<float, double>(k1, A->n, max_rows[0], max_rows[1], max_rows[2], nmat, A->row, A->col,
                                               (const float *)A->val, (int64_t)A->nnz, p_row, p_col, p_val, 0, ap_row,
                                               ap_col, ap_val, (int64_t)ap_nnz, s);
        });
    if (rc != PP_OK) return rc;
    if (stages & 2)
        PP_TIMED(PP_K_GALERKIN, c_n, s, {
            if (c_dtype == PP_F64)
                rc = launch_rap<double, double>(k2, c_n, max_rows[3], max_rows[2], max_rows[4], nmat, r_row, r_col,
                                                r_val, 0, ap_row, ap_col, ap_val, (int64_t)ap_nnz, c_row, c_col,
                                                (double *)c_val, (int64_t)c_nnz, s);
            else
                rc = launch_rap<double, float>(k2, c_n, max_rows[3], max_rows[2], max_rows[4], nmat, r_row, r_col,
                                               r_val, 0, ap_row, ap_col, ap_val, (int64_t)ap_nnz, c_row, c_col,
                                               (float *)c_val, (int64_t)c_nnz, s);
        });
    if (rc != PP_OK) return rc;
    PP_LAUNCH_CHECK();
    return PP_OK;
}

int pp_coarse_inverse(const double *C, double *out, int32_t n, int32_t batch, int32_t rank_one, void *stream) {
    PP_CHECK_ARG(C && out, "null argument");
    PP_CHECK_ARG(n >= 1 && n <= CINV_MAX && batch >= 1, "n must lie in [1, 128] and batch be positive");
    hipLaunchKernelGGL(k_coarse_inverse, dim3(batch), dim3(CINV_THREADS), 0, as_stream(stream), C, out, (int)n,
                       (int)(rank_one != 0));
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


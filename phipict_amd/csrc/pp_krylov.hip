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
    static thread_local int *bad = nullptr; // one flag per host thread
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


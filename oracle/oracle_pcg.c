/*
 * ORACLE — test infrastructure only (see oracle.h).
 *
 * AMG-preconditioned CG of the reference: krylov::amgPcgSolve,
 * csrc/phipict/solvers/krylov/krylov.cu:1389-1546, with the V-cycle kernels :1247-1321 (damped
 * Jacobi, first pre-sweep fused with the residual), dense coarse solve :1349-1360 and
 * constant-mode projection :1363-1385. Compiled with multiply-add contraction like the Krylov
 * restatement (the hierarchy setup in oracle_amg.c is host code and is compiled without it).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

/* ---------------- AMG-PCG (krylov.cu:1389-1546) ----------------
 * T: Krylov type, H: hierarchy type. Level arrays are passed per level:
 * n[l], A (row, col, val in H), dinv (H), R, P (H) for l < L; the coarsest level is the
 * dense pinv [nc][nc] (H). */
#define PCG_BODY(T, H, NAME)                                                                         \
int NAME(int n, const int32_t *frow, const int32_t *fcol, const T *fval, int L,                       \
         const int32_t *ln, const int32_t **larow, const int32_t **lacol, const H **laval,            \
         const H **ldinv, const int32_t **lrrow, const int32_t **lrcol, const H **lrval,             \
         const int32_t **lprow, const int32_t **lpcol, const H **lpval, const H *pinv, int nc,        \
         const T *b_in, T *x, int maxit, double tol, int normalized, int return_best, int project,    \
         int pre, int post, double omega_d, ob_result *res) {                                         \
    const double ns = normalized ? 1.0 / sqrt((double)n) : 1.0;                                       \
    const H omega = (H)omega_d;                                                                       \
    if (pre < 1) pre = 1;                                                                             \
    if (post < 1) post = 1;                                                                           \
    T *r = (T *)malloc(sizeof(T) * n), *q = (T *)malloc(sizeof(T) * n), *u = (T *)malloc(sizeof(T) * n); \
    T *pa = (T *)calloc(n, sizeof(T)), *pb = (T *)calloc(n, sizeof(T)), *bw = (T *)malloc(sizeof(T) * n); \
    T *xbest = (T *)malloc(sizeof(T) * n);                                                            \
    H *lx[32], *ly[32], *lres[32], *lb[33];                                                           \
    for (int l = 0; l < L; ++l) {                                                                     \
        lx[l] = (H *)calloc(ln[l], sizeof(H)); ly[l] = (H *)calloc(ln[l], sizeof(H));                 \
        lres[l] = (H *)calloc(ln[l], sizeof(H)); lb[l] = (H *)calloc(ln[l], sizeof(H));               \
    }                                                                                                 \
    lb[L] = (H *)calloc(nc, sizeof(H));                                                               \
    H *cx = (H *)calloc(nc, sizeof(H));                                                               \
    memcpy(bw, b_in, sizeof(T) * n);                                                                  \
    if (project) {                                                                                    \
        double m = 0;                                                                                 \
        for (int i = 0; i < n; ++i) m += (double)bw[i];                                               \
        m /= (double)n;                                                                               \
        T mm = (T)m;                                                                                  \
        for (int i = 0; i < n; ++i) bw[i] -= mm;                                                      \
    }                                                                                                 \
    double rho = 0, alpha = 0, beta = 0, crit = 0, best = INFINITY, last = 0;                         \
    int it = 0, bestIt = -1, rising = 0, done = 0, restart = 1, saveBest = 0, usedIt = 0;             \
    {                                                                                                 \
        double acc = 0;                                                                           \
        O_PAR_SUM(acc)                                                                            \
        for (int i = 0; i < n; ++i) {                                                             \
            T s = 0;                                                                                  \
            for (int j = frow[i]; j < frow[i + 1]; ++j) s += fval[j] * x[fcol[j]];                    \
            T ri = bw[i] - s; r[i] = ri; acc += (double)ri * (double)ri;                              \
        }                                                                                             \
        rho = acc; restart = 1; crit = sqrt(acc) * ns;                                                \
        if (!isfinite(crit)) { done = 2; usedIt = 0; }                                                \
        else if (crit == 0.0) { done = 1; usedIt = 0; }                                               \
    }                                                                                                 \
    T *pin = pa, *pout = pb;                                                                          \
    long guard = 0;                                                                                   \
    while (done == 0 && guard++ < (long)maxit + 4) {                                                  \
        /* V-cycle: u = M r */                                                                        \
        double shift = 0;                                                                             \
        if (project) {                                                                            \
            double m = 0;                                                                         \
            O_PAR_SUM(m)                                                                          \
            for (int i = 0; i < n; ++i) m += (double)r[i];                                        \
            shift = m / (double)n;                                                                \
        }                                                                                         \
        H *cur[32], *oth[32];                                                                         \
        for (int l = 0; l < L; ++l) {                                                                 \
            int nl = ln[l];                                                                           \
            cur[l] = lx[l]; oth[l] = ly[l];                                                           \
            const H *bl = lb[l];                                                                      \
            /* fused first pre-sweep + residual (kPresmoothResidual) */                               \
            O_PAR                                                                                 \
            for (int i = 0; i < nl; ++i) {                                                        \
                H s = 0;                                                                              \
                for (int j = larow[l][i]; j < larow[l][i + 1]; ++j) {                                 \
                    int c = lacol[l][j];                                                              \
                    H bc = l == 0 ? (project ? (H)((double)r[c] - shift) : (H)r[c]) : bl[c];          \
                    s += laval[l][j] * (ldinv[l][c] * bc);                                            \
                }                                                                                     \
                H bi = l == 0 ? (project ? (H)((double)r[i] - shift) : (H)r[i]) : bl[i];              \
                cur[l][i] = omega * ldinv[l][i] * bi;                                                 \
                lres[l][i] = bi - omega * s;                                                          \
            }                                                                                         \
            for (int k = 1; k < pre; ++k) {                                                           \
                O_PAR                                                                             \
                for (int i = 0; i < nl; ++i) {                                                    \
                    H s = 0;                                                                          \
                    for (int j = larow[l][i]; j < larow[l][i + 1]; ++j) s += laval[l][j] * cur[l][lacol[l][j]]; \
                    H bi = l == 0 ? (project ? (H)((double)r[i] - shift) : (H)r[i]) : bl[i];          \
                    oth[l][i] = cur[l][i] + omega * ldinv[l][i] * (bi - s);                           \
                }                                                                                     \
                H *t_ = cur[l]; cur[l] = oth[l]; oth[l] = t_;                                         \
            }                                                                                         \
            if (pre > 1) {                                                                            \
                O_PAR                                                                             \
                for (int i = 0; i < nl; ++i) {                                                    \
                    H s = 0;                                                                          \
                    for (int j = larow[l][i]; j < larow[l][i + 1]; ++j) s += laval[l][j] * cur[l][lacol[l][j]]; \
                    H bi = l == 0 ? (project ? (H)((double)r[i] - shift) : (H)r[i]) : bl[i];          \
                    lres[l][i] = bi - s;                                                              \
                }                                                                                     \
            }                                                                                         \
            int ncl = l + 1 < L ? ln[l + 1] : nc;                                                     \
            O_PAR                                                                                 \
            for (int i = 0; i < ncl; ++i) {                                                       \
                H s = 0;                                                                              \
                for (int j = lrrow[l][i]; j < lrrow[l][i + 1]; ++j) s += lrval[l][j] * lres[l][lrcol[l][j]]; \
                lb[l + 1][i] = s;                                                                     \
            }                                                                                         \
        }                                                                                             \
        O_PAR                                                                                     \
        for (int i = 0; i < nc; ++i) {                                                            \
            double s = 0;                                                                         \
            for (int j = 0; j < nc; ++j) s += (double)pinv[(size_t)i * nc + j] * (double)lb[L][j];    \
            cx[i] = (H)s;                                                                             \
        }                                                                                             \
        for (int l = L - 1; l >= 0; --l) {                                                            \
            int nl = ln[l];                                                                           \
            const H *xc = l + 1 < L ? cur[l + 1] : cx;                                                \
            O_PAR                                                                                 \
            for (int i = 0; i < nl; ++i) {                                                        \
                H s = 0;                                                                              \
                for (int j = lprow[l][i]; j < lprow[l][i + 1]; ++j) s += lpval[l][j] * xc[lpcol[l][j]]; \
                cur[l][i] = cur[l][i] + s;                                                            \
            }                                                                                         \
            const H *bl = lb[l];                                                                      \
            for (int k = 0; k < post; ++k) {                                                          \
                O_PAR                                                                             \
                for (int i = 0; i < nl; ++i) {                                                    \
                    H s = 0;                                                                          \
                    for (int j = larow[l][i]; j < larow[l][i + 1]; ++j) s += laval[l][j] * cur[l][lacol[l][j]]; \
                    H bi = l == 0 ? (project ? (H)((double)r[i] - shift) : (H)r[i]) : bl[i];          \
                    H o = cur[l][i] + omega * ldinv[l][i] * (bi - s);                                 \
                    if (l == 0 && k == post - 1) u[i] = (T)o; else oth[l][i] = o;                     \
                }                                                                                     \
                if (!(l == 0 && k == post - 1)) { H *t_ = cur[l]; cur[l] = oth[l]; oth[l] = t_; }      \
            }                                                                                         \
        }                                                                                             \
        if (project) {                                                                                \
            double m = 0;                                                                             \
            O_PAR_SUM(m)                                                                          \
            for (int i = 0; i < n; ++i) m += (double)u[i];                                        \
            T mm = (T)(m / (double)n);                                                                \
            for (int i = 0; i < n; ++i) u[i] -= mm;                                                   \
        }                                                                                             \
        /* kCgDots */                                                                                 \
        {                                                                                             \
            double rz = 0;                                                                            \
            O_PAR_SUM(rz)                                                                         \
            for (int i = 0; i < n; ++i) rz += (double)r[i] * (double)u[i];                        \
            if (!restart) beta = rz / rho;                                                            \
            rho = rz;                                                                                 \
            if (!isfinite(rho) || !isfinite(beta)) { done = 2; usedIt = it; break; }                  \
        }                                                                                             \
        /* kCgDirection<PM_VECTOR> */                                                                 \
        {                                                                                             \
            const T bt = (T)beta;                                                                     \
            O_PAR                                                                                 \
            for (int i = 0; i < n; ++i) pout[i] = restart ? u[i] : u[i] + bt * pin[i];            \
            double pq = 0;                                                                        \
            O_PAR_SUM(pq)                                                                         \
            for (int i = 0; i < n; ++i) {                                                         \
                T s = 0;                                                                              \
                for (int j = frow[i]; j < frow[i + 1]; ++j) s += fval[j] * pout[fcol[j]];             \
                q[i] = s; pq += (double)pout[i] * (double)s;                                          \
            }                                                                                         \
            double a = rho / pq;                                                                      \
            restart = 0;                                                                              \
            if (!isfinite(a)) { done = 2; usedIt = it; break; }                                       \
            alpha = a;                                                                                \
        }                                                                                             \
        /* kCgUpdate<PM_VECTOR> */                                                                    \
        {                                                                                             \
            double rr = 0;                                                                            \
            const T al = (T)alpha;                                                                \
            O_PAR_SUM(rr)                                                                         \
            for (int i = 0; i < n; ++i) {                                                         \
                T xi = x[i];                                                                          \
                if (saveBest && return_best) xbest[i] = xi;                                           \
                x[i] = xi + al * pout[i];                                                             \
                T ri = r[i] - al * q[i]; r[i] = ri; rr += (double)ri * (double)ri;                    \
            }                                                                                         \
            crit = sqrt(rr) * ns; ++it; saveBest = 0; usedIt = it - 1;                                \
            if (!isfinite(crit)) done = 2;                                                            \
            else {                                                                                    \
                if (return_best) {                                                                    \
                    if (crit < best) { best = crit; bestIt = it - 1; saveBest = 1; }                  \
                    if (it > 1 && crit >= last) ++rising; else rising = 0;                            \
                    last = crit;                                                                      \
                }                                                                                     \
                if (crit < tol) done = 1;                                                             \
                else if (it >= maxit) done = 3;                                                       \
                else if (return_best && rising >= 100) done = 4;                                      \
            }                                                                                         \
        }                                                                                             \
        { T *t_ = pin; pin = pout; pout = t_; }                                                        \
    }                                                                                                 \
    int restore = return_best && (done == 3 || done == 4) && bestIt >= 0 && best < crit;              \
    if (restore) memcpy(x, xbest, sizeof(T) * n);                                                     \
    res->final_residual = restore ? best : crit;                                                      \
    res->used_iterations = restore ? bestIt : usedIt;                                                 \
    res->converged = done == 1;                                                                       \
    res->is_finite = isfinite(crit) ? 1 : 0;                                                          \
    for (int l = 0; l < L; ++l) { free(lx[l]); free(ly[l]); free(lres[l]); free(lb[l]); }             \
    free(lb[L]); free(cx);                                                                            \
    free(r); free(q); free(u); free(pa); free(pb); free(bw); free(xbest);                             \
    return 0;                                                                                         \
}

PCG_BODY(double, double, o_amg_pcg_f64_f64)
PCG_BODY(double, float, o_amg_pcg_f64_f32)
PCG_BODY(float, float, o_amg_pcg_f32_f32)

/*
 * ORACLE — test infrastructure only (see oracle.h).
 *
 * Classical Ruge-Stueben AMG setup and the AMG-preconditioned CG of the reference.
 *
 * Setup: the reference calls pyamg (>=4.0, an optional dependency that is not vendored,
 * pyproject.toml:511-512) through build_amg_hierarchy (src/phipict/solvers/amg.py:450-598)
 * with ruge_stuben_solver(A, max_levels=20, max_coarse=64, strength=('classical',
 * {'theta': 0.25})). pyamg is absent here, so this restates the published algorithm:
 *   1. classical strength, abs norm: j != i is strong for i if |a_ij| >= theta * max_k!=i |a_ik|
 *   2. RS first-pass C/F splitting with lambda buckets (pyamg amg_core rs_cf_splitting):
 *      lambda_i = |S^T_i|; nodes with lambda 0 become F; repeatedly take the unassigned
 *      node of largest lambda as C, make its unassigned strong dependents F and raise the
 *      lambda of their unassigned strong influences; lower lambda of i's own influences.
 *   3. modified classical interpolation (pyamg classical_interpolation, modified=True):
 *      strong F-F links without a common strong C neighbour are treated as weak;
 *      w_ij = -(a_ij + sum_{k in F_i^s} a_ik abar_kj / sum_{m in C_i} abar_km)
 *             / (a_ii + sum_{weak} a_in), abar_km = a_km where sign(a_km) != sign(a_kk).
 *   4. R = P^T, A_c = R A P; recurse while n > max_coarse and levels < max_levels.
 * Exact parity with pyamg's interpolation is unpinned (pyamg is not available); the
 * product's setup is checked against this restatement, and both are checked by the
 * invariants the reference's tests hold (Galerkin exactness, contraction, complexity).
 *
 * PCG: krylov::amgPcgSolve, csrc/phipict/solvers/krylov/krylov.cu:1389-1546, with the
 * V-cycle kernels :1247-1321 (damped Jacobi, first pre-sweep fused with the residual),
 * dense coarse solve :1349-1360 and constant-mode projection :1363-1385.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

typedef struct {
    int n, nnz;
    int *row, *col;
    double *val;
} ocsr;

typedef struct {
    int nlev;        /* levels including the coarsest */
    ocsr A[32];
    ocsr P[32], R[32];
    int *split[32];  /* 1 = C */
} ohier;

static void csr_free(ocsr *m) { free(m->row); free(m->col); free(m->val); memset(m, 0, sizeof(*m)); }

static ocsr csr_transpose(const ocsr *a, int ncols) {
    ocsr t;
    t.n = ncols; t.nnz = a->nnz;
    t.row = (int *)calloc(ncols + 1, sizeof(int));
    t.col = (int *)malloc(sizeof(int) * (a->nnz > 0 ? a->nnz : 1));
    t.val = (double *)malloc(sizeof(double) * (a->nnz > 0 ? a->nnz : 1));
    for (int k = 0; k < a->nnz; ++k) t.row[a->col[k] + 1]++;
    for (int i = 0; i < ncols; ++i) t.row[i + 1] += t.row[i];
    int *fill = (int *)malloc(sizeof(int) * (ncols + 1));
    memcpy(fill, t.row, sizeof(int) * (ncols + 1));
    for (int i = 0; i < a->n; ++i)
        for (int k = a->row[i]; k < a->row[i + 1]; ++k) {
            int o = fill[a->col[k]]++;
            t.col[o] = i;
            t.val[o] = a->val[k];
        }
    free(fill);
    return t;
}

/* C = A * B (A: n x m, B: m x p), rows with ascending columns, duplicates merged.
 * Accumulation order: for each entry a_ik in row order, for each b_kj in row order. */
static ocsr csr_matmul(const ocsr *a, const ocsr *b, int p) {
    ocsr c;
    c.n = a->n;
    c.row = (int *)calloc(a->n + 1, sizeof(int));
    int *mark = (int *)malloc(sizeof(int) * (p > 0 ? p : 1));
    double *acc = (double *)calloc(p > 0 ? p : 1, sizeof(double));
    for (int j = 0; j < p; ++j) mark[j] = -1;
    int cap = a->nnz + 16, nnz = 0;
    c.col = (int *)malloc(sizeof(int) * cap);
    c.val = (double *)malloc(sizeof(double) * cap);
    int *cols = (int *)malloc(sizeof(int) * (p > 0 ? p : 1));
    for (int i = 0; i < a->n; ++i) {
        int nc = 0;
        for (int kk = a->row[i]; kk < a->row[i + 1]; ++kk) {
            int k = a->col[kk];
            double av = a->val[kk];
            for (int jj = b->row[k]; jj < b->row[k + 1]; ++jj) {
                int j = b->col[jj];
                if (mark[j] != i) { mark[j] = i; acc[j] = 0; cols[nc++] = j; }
                acc[j] += av * b->val[jj];
            }
        }
        /* insertion sort of the row's columns */
        for (int x = 1; x < nc; ++x) {
            int v = cols[x], y = x - 1;
            while (y >= 0 && cols[y] > v) { cols[y + 1] = cols[y]; --y; }
            cols[y + 1] = v;
        }
        if (nnz + nc > cap) {
            while (nnz + nc > cap) cap *= 2;
            c.col = (int *)realloc(c.col, sizeof(int) * cap);
            c.val = (double *)realloc(c.val, sizeof(double) * cap);
        }
        for (int x = 0; x < nc; ++x) { c.col[nnz] = cols[x]; c.val[nnz] = acc[cols[x]]; ++nnz; }
        c.row[i + 1] = nnz;
    }
    c.nnz = nnz;
    free(mark); free(acc); free(cols);
    return c;
}

/* strong[k] for every entry k of A (diagonal never strong) */
static void strength(const ocsr *a, double theta, char *strong) {
    for (int i = 0; i < a->n; ++i) {
        double mx = 0;
        for (int k = a->row[i]; k < a->row[i + 1]; ++k)
            if (a->col[k] != i && fabs(a->val[k]) > mx) mx = fabs(a->val[k]);
        double thr = theta * mx;
        for (int k = a->row[i]; k < a->row[i + 1]; ++k)
            strong[k] = a->col[k] != i && fabs(a->val[k]) > 0 && fabs(a->val[k]) >= thr;
    }
}

enum { U_NODE = -1, F_NODE = 0, C_NODE = 1 };

/* RS first pass with lambda intervals (pyamg amg_core rs_cf_splitting). S: strong
 * dependencies per row (Sp, Sj), T = S^T (Tp, Tj), both with ascending columns. */
static void rs_split(int n, const int *Sp, const int *Sj, const int *Tp, const int *Tj, int *split) {
    int *lambda = (int *)malloc(sizeof(int) * n);
    int *iptr = (int *)calloc(n + 1, sizeof(int));
    int *icnt = (int *)calloc(n + 1, sizeof(int));
    int *i2n = (int *)malloc(sizeof(int) * n);
    int *n2i = (int *)malloc(sizeof(int) * n);
    for (int i = 0; i < n; ++i) lambda[i] = Tp[i + 1] - Tp[i];
    for (int i = 0; i < n; ++i) icnt[lambda[i]]++;
    for (int i = 0, cs = 0; i < n; ++i) { iptr[i] = cs; cs += icnt[i]; icnt[i] = 0; }
    for (int i = 0; i < n; ++i) {
        int li = lambda[i];
        int idx = iptr[li] + icnt[li];
        i2n[idx] = i; n2i[i] = idx;
        icnt[li]++;
    }
    for (int i = 0; i < n; ++i) split[i] = U_NODE;
    for (int i = 0; i < n; ++i)
        if (lambda[i] == 0 || (lambda[i] == 1 && Tj[Tp[i]] == i)) split[i] = F_NODE;
    for (int top = n - 1; top > -1; --top) {
        int i = i2n[top];
        int li = lambda[i];
        icnt[li]--;
        if (split[i] == F_NODE) continue;
        split[i] = C_NODE;
        for (int jj = Tp[i]; jj < Tp[i + 1]; ++jj) {
            int j = Tj[jj];
            if (split[j] != U_NODE) continue;
            split[j] = F_NODE;
            for (int kk = Sp[j]; kk < Sp[j + 1]; ++kk) {
                int k = Sj[kk];
                if (split[k] != U_NODE) continue;
                if (lambda[k] >= n - 1) continue;
                int lk = lambda[k];
                int op = n2i[k];
                int np = iptr[lk] + icnt[lk] - 1;
                n2i[i2n[op]] = np;
                n2i[i2n[np]] = op;
                int tmp = i2n[op]; i2n[op] = i2n[np]; i2n[np] = tmp;
                icnt[lk] -= 1;
                icnt[lk + 1] += 1;
                iptr[lk + 1] = np;
                lambda[k]++;
            }
        }
        for (int jj = Sp[i]; jj < Sp[i + 1]; ++jj) {
            int j = Sj[jj];
            if (split[j] != U_NODE) continue;
            if (lambda[j] == 0) continue;
            int lj = lambda[j];
            int op = n2i[j];
            int np = iptr[lj];
            n2i[i2n[op]] = np;
            n2i[i2n[np]] = op;
            int tmp = i2n[op]; i2n[op] = i2n[np]; i2n[np] = tmp;
            icnt[lj] -= 1;
            icnt[lj - 1] += 1;
            iptr[lj] += 1;
            iptr[lj - 1] = iptr[lj] - icnt[lj - 1];
            lambda[j]--;
        }
    }
    for (int i = 0; i < n; ++i) if (split[i] == U_NODE) split[i] = F_NODE;
    free(lambda); free(iptr); free(icnt); free(i2n); free(n2i);
}

static double entry(const ocsr *a, int i, int j) {
    double s = 0;
    for (int k = a->row[i]; k < a->row[i + 1]; ++k) if (a->col[k] == j) s += a->val[k];
    return s;
}
static double diag_of(const ocsr *a, int i) { return entry(a, i, i); }

/* modified classical interpolation (see the file header) */
static ocsr interpolation(const ocsr *a, const char *strong, const int *split, int *nc_out) {
    int n = a->n;
    int *cidx = (int *)malloc(sizeof(int) * n);
    int nc = 0;
    for (int i = 0; i < n; ++i) cidx[i] = split[i] == C_NODE ? nc++ : -1;
    *nc_out = nc;
    ocsr p;
    p.n = n;
    p.row = (int *)calloc(n + 1, sizeof(int));
    int cap = n * 4 + 16, nnz = 0;
    p.col = (int *)malloc(sizeof(int) * cap);
    p.val = (double *)malloc(sizeof(double) * cap);
    char *fstrong = (char *)malloc(a->nnz > 0 ? a->nnz : 1);
    for (int i = 0; i < n; ++i) {
        if (nnz + 64 > cap) {
            cap *= 2;
            p.col = (int *)realloc(p.col, sizeof(int) * cap);
            p.val = (double *)realloc(p.val, sizeof(double) * cap);
        }
        if (split[i] == C_NODE) {
            p.col[nnz] = cidx[i]; p.val[nnz] = 1.0; ++nnz;
            p.row[i + 1] = nnz;
            continue;
        }
        /* strong F neighbours that share a strong C neighbour with i stay strong */
        for (int kk = a->row[i]; kk < a->row[i + 1]; ++kk) {
            fstrong[kk] = 0;
            int k = a->col[kk];
            if (!strong[kk] || split[k] == C_NODE) continue;
            int common = 0;
            for (int jj = a->row[i]; jj < a->row[i + 1] && !common; ++jj) {
                if (!strong[jj] || split[a->col[jj]] != C_NODE) continue;
                int j = a->col[jj];
                for (int mm = a->row[k]; mm < a->row[k + 1]; ++mm)
                    if (a->col[mm] == j && strong[mm]) { common = 1; break; }
            }
            fstrong[kk] = (char)common;
        }
        double den = 0;
        for (int kk = a->row[i]; kk < a->row[i + 1]; ++kk) {
            int k = a->col[kk];
            if (k == i) den += a->val[kk];
            else if (!(strong[kk] && (split[k] == C_NODE || fstrong[kk]))) den += a->val[kk];
        }
        /* F-F links whose inner denominator vanishes are lumped into the diagonal */
        for (int kk = a->row[i]; kk < a->row[i + 1]; ++kk) {
            if (!fstrong[kk]) continue;
            int k = a->col[kk];
            double akk = diag_of(a, k), inner = 0;
            for (int jj = a->row[i]; jj < a->row[i + 1]; ++jj) {
                if (!strong[jj] || split[a->col[jj]] != C_NODE) continue;
                double akm = entry(a, k, a->col[jj]);
                if ((akm < 0) != (akk < 0) && akm != 0) inner += akm;
            }
            if (inner == 0) { den += a->val[kk]; fstrong[kk] = 0; }
        }
        int start = nnz;
        for (int jj = a->row[i]; jj < a->row[i + 1]; ++jj) {
            int j = a->col[jj];
            if (!strong[jj] || split[j] != C_NODE) continue;
            double num = a->val[jj];
            for (int kk = a->row[i]; kk < a->row[i + 1]; ++kk) {
                if (!fstrong[kk]) continue;
                int k = a->col[kk];
                double akk = diag_of(a, k), inner = 0;
                for (int mm = a->row[i]; mm < a->row[i + 1]; ++mm) {
                    if (!strong[mm] || split[a->col[mm]] != C_NODE) continue;
                    double akm = entry(a, k, a->col[mm]);
                    if ((akm < 0) != (akk < 0) && akm != 0) inner += akm;
                }
                double akj = entry(a, k, j);
                if (!((akj < 0) != (akk < 0) && akj != 0)) akj = 0;
                num += a->val[kk] * akj / inner;
            }
            p.col[nnz] = cidx[j];
            p.val[nnz] = -num / den;
            ++nnz;
        }
        /* ascending coarse columns (C indices follow fine order, so already ascending
         * when the fine row is; sort defensively) */
        for (int x = start + 1; x < nnz; ++x) {
            int cv = p.col[x]; double vv = p.val[x]; int y = x - 1;
            while (y >= start && p.col[y] > cv) { p.col[y + 1] = p.col[y]; p.val[y + 1] = p.val[y]; --y; }
            p.col[y + 1] = cv; p.val[y + 1] = vv;
        }
        p.row[i + 1] = nnz;
    }
    p.nnz = nnz;
    free(cidx); free(fstrong);
    return p;
}

static int setup_level(ohier *h, int l, double theta) {
    const ocsr *a = &h->A[l];
    int n = a->n;
    char *strong = (char *)malloc(a->nnz > 0 ? a->nnz : 1);
    strength(a, theta, strong);
    int *Sp = (int *)calloc(n + 1, sizeof(int));
    int ns = 0;
    for (int k = 0; k < a->nnz; ++k) ns += strong[k];
    int *Sj = (int *)malloc(sizeof(int) * (ns > 0 ? ns : 1));
    ns = 0;
    for (int i = 0; i < n; ++i) {
        for (int k = a->row[i]; k < a->row[i + 1]; ++k) if (strong[k]) Sj[ns++] = a->col[k];
        Sp[i + 1] = ns;
    }
    ocsr s = {n, ns, Sp, Sj, (double *)calloc(ns > 0 ? ns : 1, sizeof(double))};
    ocsr t = csr_transpose(&s, n);
    int *split = (int *)malloc(sizeof(int) * n);
    rs_split(n, Sp, Sj, t.row, t.col, split);
    int nc = 0;
    ocsr p = interpolation(a, strong, split, &nc);
    free(strong);
    csr_free(&s); csr_free(&t);
    if (nc == 0 || nc == n) { free(split); csr_free(&p); return 0; }
    ocsr r = csr_transpose(&p, nc);
    ocsr ap = csr_matmul(a, &p, nc);
    ocsr rap = csr_matmul(&r, &ap, nc);
    csr_free(&ap);
    h->P[l] = p; h->R[l] = r; h->split[l] = split;
    h->A[l + 1] = rap;
    return 1;
}

void *o_amg_setup(int n, const int32_t *row, const int32_t *col, const double *val, double theta, int max_levels, int max_coarse) {
    ohier *h = (ohier *)calloc(1, sizeof(ohier));
    int nnz = row[n];
    h->A[0].n = n; h->A[0].nnz = nnz;
    h->A[0].row = (int *)malloc(sizeof(int) * (n + 1));
    h->A[0].col = (int *)malloc(sizeof(int) * (nnz > 0 ? nnz : 1));
    h->A[0].val = (double *)malloc(sizeof(double) * (nnz > 0 ? nnz : 1));
    memcpy(h->A[0].row, row, sizeof(int) * (n + 1));
    int w = 0;
    for (int i = 0; i < n; ++i) {
        for (int k = row[i]; k < row[i + 1]; ++k) {
            if (col[k] < 0) continue; /* sentinel entries are dropped (amg.py:116-135) */
            h->A[0].col[w] = col[k]; h->A[0].val[w] = val[k]; ++w;
        }
        h->A[0].row[i + 1] = w;
    }
    h->A[0].nnz = w;
    h->nlev = 1;
    if (max_levels > 32) max_levels = 32;
    while (h->nlev < max_levels && h->A[h->nlev - 1].n > max_coarse) {
        if (!setup_level(h, h->nlev - 1, theta)) break;
        h->nlev++;
    }
    return h;
}

int o_amg_num_levels(void *hp) { return ((ohier *)hp)->nlev; }

void o_amg_level_sizes(void *hp, int l, int32_t *out) {
    ohier *h = (ohier *)hp;
    out[0] = h->A[l].n; out[1] = h->A[l].nnz;
    out[2] = l < h->nlev - 1 ? h->P[l].nnz : 0;
    out[3] = l < h->nlev - 1 ? h->R[l].nnz : 0;
}

void o_amg_level_copy(void *hp, int l, int32_t *arow, int32_t *acol, double *aval, int32_t *prow, int32_t *pcol,
                      double *pval, int32_t *rrow, int32_t *rcol, double *rval, int32_t *split) {
    ohier *h = (ohier *)hp;
    const ocsr *a = &h->A[l];
    if (arow) memcpy(arow, a->row, sizeof(int) * (a->n + 1));
    if (acol) memcpy(acol, a->col, sizeof(int) * a->nnz);
    if (aval) memcpy(aval, a->val, sizeof(double) * a->nnz);
    if (l < h->nlev - 1) {
        const ocsr *p = &h->P[l], *r = &h->R[l];
        if (prow) memcpy(prow, p->row, sizeof(int) * (p->n + 1));
        if (pcol) memcpy(pcol, p->col, sizeof(int) * p->nnz);
        if (pval) memcpy(pval, p->val, sizeof(double) * p->nnz);
        if (rrow) memcpy(rrow, r->row, sizeof(int) * (r->n + 1));
        if (rcol) memcpy(rcol, r->col, sizeof(int) * r->nnz);
        if (rval) memcpy(rval, r->val, sizeof(double) * r->nnz);
        if (split) for (int i = 0; i < a->n; ++i) split[i] = h->split[l][i] == C_NODE;
    }
}

void o_amg_free(void *hp) {
    ohier *h = (ohier *)hp;
    for (int l = 0; l < h->nlev; ++l) {
        csr_free(&h->A[l]);
        if (l < h->nlev - 1) { csr_free(&h->P[l]); csr_free(&h->R[l]); free(h->split[l]); }
    }
    free(h);
}

/* Galerkin product R A P in double, with the accumulation order of csr_matmul. */
int o_galerkin(int n, const int32_t *arow, const int32_t *acol, const double *aval, int nc,
               const int32_t *prow, const int32_t *pcol, const double *pval,
               const int32_t *rrow, const int32_t *rcol, const double *rval,
               int32_t *crow_out, int32_t *ccol_out, double *cval_out, int cap) {
    ocsr a = {n, arow[n], (int *)arow, (int *)acol, (double *)aval};
    ocsr p = {n, prow[n], (int *)prow, (int *)pcol, (double *)pval};
    ocsr r = {nc, rrow[nc], (int *)rrow, (int *)rcol, (double *)rval};
    ocsr ap = csr_matmul(&a, &p, nc);
    ocsr c = csr_matmul(&r, &ap, nc);
    csr_free(&ap);
    if (c.nnz > cap) { csr_free(&c); return -1; }
    memcpy(crow_out, c.row, sizeof(int) * (nc + 1));
    memcpy(ccol_out, c.col, sizeof(int) * c.nnz);
    memcpy(cval_out, c.val, sizeof(double) * c.nnz);
    int nnz = c.nnz;
    csr_free(&c);
    return nnz;
}

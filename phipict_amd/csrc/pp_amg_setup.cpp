// Host-side classical Ruge-Stueben AMG setup (replaces the pyamg call of
// build_amg_hierarchy, src/phipict/solvers/amg.py:450-598: ruge_stuben_solver with
// strength ('classical', theta), max_levels, max_coarse).
//
// Steps per level: classical strength of connection (abs norm), RS first-pass C/F
// splitting with lambda intervals, modified classical interpolation (strong F-F links
// without a common strong C neighbour count as weak, F-F distribution only through
// coefficients of opposite sign to the diagonal), R = P^T, A_c = R (A P) in double.
// Every step uses dense scatter arrays, so the cost is linear in the non-zeros.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/phipict_hip.h"

namespace pp {
void set_error(const std::string &msg);
}

namespace {

struct Csr {
    int n = 0;
    std::vector<int32_t> row, col;
    std::vector<double> val;
    int nnz() const { return (int)col.size(); }
};

struct Hierarchy {
    std::vector<Csr> A, P, R;
    std::vector<std::vector<int32_t>> split;
};

Csr transpose(const Csr &a, int ncols) {
    Csr t;
    t.n = ncols;
    t.row.assign(ncols + 1, 0);
    t.col.resize(a.nnz());
    t.val.resize(a.nnz());
    for (int k = 0; k < a.nnz(); ++k) t.row[a.col[k] + 1]++;
    for (int i = 0; i < ncols; ++i) t.row[i + 1] += t.row[i];
    std::vector<int32_t> fill(t.row.begin(), t.row.end() - 1);
    for (int i = 0; i < a.n; ++i)
        for (int k = a.row[i]; k < a.row[i + 1]; ++k) {
            const int o = fill[a.col[k]]++;
            t.col[o] = i;
            t.val[o] = a.val[k];
        }
    return t;
}

// C = A B with the accumulation order: entries a_ik in row order, then b_kj in row order.
Csr matmul(const Csr &a, const Csr &b, int p) {
    Csr c;
    c.n = a.n;
    c.row.assign(a.n + 1, 0);
    std::vector<int32_t> mark(std::max(p, 1), -1), cols;
    std::vector<double> acc(std::max(p, 1), 0.0);
    cols.reserve(64);
    c.col.reserve((size_t)a.nnz() * 2);
    c.val.reserve((size_t)a.nnz() * 2);
    for (int i = 0; i < a.n; ++i) {
        cols.clear();
        for (int kk = a.row[i]; kk < a.row[i + 1]; ++kk) {
            const int k = a.col[kk];
            const double av = a.val[kk];
            for (int jj = b.row[k]; jj < b.row[k + 1]; ++jj) {
                const int j = b.col[jj];
                if (mark[j] != i) { mark[j] = i; acc[j] = 0.0; cols.push_back(j); }
                acc[j] += av * b.val[jj];
            }
        }
        std::sort(cols.begin(), cols.end());
        for (int j : cols) { c.col.push_back(j); c.val.push_back(acc[j]); }
        c.row[i + 1] = (int32_t)c.col.size();
    }
    return c;
}

enum { U_NODE = -1, F_NODE = 0, C_NODE = 1 };

void rs_split(int n, const std::vector<int32_t> &Sp, const std::vector<int32_t> &Sj, const std::vector<int32_t> &Tp,
              const std::vector<int32_t> &Tj, std::vector<int32_t> &split) {
    std::vector<int> lambda(n), iptr(n + 1, 0), icnt(n + 1, 0), i2n(n), n2i(n);
    for (int i = 0; i < n; ++i) lambda[i] = Tp[i + 1] - Tp[i];
    for (int i = 0; i < n; ++i) icnt[lambda[i]]++;
    for (int i = 0, cs = 0; i < n; ++i) { iptr[i] = cs; cs += icnt[i]; icnt[i] = 0; }
    for (int i = 0; i < n; ++i) {
        const int li = lambda[i];
        const int idx = iptr[li] + icnt[li];
        i2n[idx] = i;
        n2i[i] = idx;
        icnt[li]++;
    }
    split.assign(n, U_NODE);
    for (int i = 0; i < n; ++i)
        if (lambda[i] == 0 || (lambda[i] == 1 && Tj[Tp[i]] == i)) split[i] = F_NODE;
    for (int top = n - 1; top > -1; --top) {
        const int i = i2n[top];
        icnt[lambda[i]]--;
        if (split[i] == F_NODE) continue;
        split[i] = C_NODE;
        for (int jj = Tp[i]; jj < Tp[i + 1]; ++jj) {
            const int j = Tj[jj];
            if (split[j] != U_NODE) continue;
            split[j] = F_NODE;
            for (int kk = Sp[j]; kk < Sp[j + 1]; ++kk) {
                const int k = Sj[kk];
                if (split[k] != U_NODE) continue;
                if (lambda[k] >= n - 1) continue;
                const int lk = lambda[k];
                const int op = n2i[k];
                const int np = iptr[lk] + icnt[lk] - 1;
                n2i[i2n[op]] = np;
                n2i[i2n[np]] = op;
                std::swap(i2n[op], i2n[np]);
                icnt[lk] -= 1;
                icnt[lk + 1] += 1;
                iptr[lk + 1] = np;
                lambda[k]++;
            }
        }
        for (int jj = Sp[i]; jj < Sp[i + 1]; ++jj) {
            const int j = Sj[jj];
            if (split[j] != U_NODE) continue;
            if (lambda[j] == 0) continue;
            const int lj = lambda[j];
            const int op = n2i[j];
            const int np = iptr[lj];
            n2i[i2n[op]] = np;
            n2i[i2n[np]] = op;
            std::swap(i2n[op], i2n[np]);
            icnt[lj] -= 1;
            icnt[lj - 1] += 1;
            iptr[lj] += 1;
            iptr[lj - 1] = iptr[lj] - icnt[lj - 1];
            lambda[j]--;
        }
    }
    for (int i = 0; i < n; ++i)
        if (split[i] == U_NODE) split[i] = F_NODE;
}

bool setup_level(Hierarchy &h, double theta) {
    const Csr &a = h.A.back();
    const int n = a.n;
    const int nnz = a.nnz();
    std::vector<char> strong(std::max(nnz, 1), 0);
    for (int i = 0; i < n; ++i) {
        double mx = 0.0;
        for (int k = a.row[i]; k < a.row[i + 1]; ++k)
            if (a.col[k] != i && std::fabs(a.val[k]) > mx) mx = std::fabs(a.val[k]);
        const double thr = theta * mx;
        for (int k = a.row[i]; k < a.row[i + 1]; ++k)
            strong[k] = a.col[k] != i && std::fabs(a.val[k]) > 0 && std::fabs(a.val[k]) >= thr;
    }
    Csr s;
    s.n = n;
    s.row.assign(n + 1, 0);
    for (int i = 0; i < n; ++i) {
        for (int k = a.row[i]; k < a.row[i + 1]; ++k)
            if (strong[k]) s.col.push_back(a.col[k]);
        s.row[i + 1] = (int32_t)s.col.size();
    }
    s.val.assign(s.col.size(), 0.0);
    Csr t = transpose(s, n);
    std::vector<int32_t> split;
    rs_split(n, s.row, s.col, t.row, t.col, split);

    std::vector<int32_t> cidx(n, -1);
    int nc = 0;
    for (int i = 0; i < n; ++i)
        if (split[i] == C_NODE) cidx[i] = nc++;
    if (nc == 0 || nc == n) return false;

    // modified classical interpolation
    Csr p;
    p.n = n;
    p.row.assign(n + 1, 0);
    p.col.reserve((size_t)n * 4);
    p.val.reserve((size_t)n * 4);
    std::vector<double> dense(n, 0.0);   // scatter of a row k: value per column
    std::vector<char> dset(n, 0);
    std::vector<int32_t> cmark(n, -1);   // strong C neighbours of the current row i
    std::vector<char> smark(n, 0);       // strong entries of row k, by column
    std::vector<char> fstrong;
    std::vector<double> inner;
    for (int i = 0; i < n; ++i) {
        if (split[i] == C_NODE) {
            p.col.push_back(cidx[i]);
            p.val.push_back(1.0);
            p.row[i + 1] = (int32_t)p.col.size();
            continue;
        }
        const int rb = a.row[i], re = a.row[i + 1];
        fstrong.assign(re - rb, 0);
        inner.assign(re - rb, 0.0);
        for (int jj = rb; jj < re; ++jj)
            if (strong[jj] && split[a.col[jj]] == C_NODE) cmark[a.col[jj]] = i;
        // strong F neighbours sharing a strong C neighbour with i stay strong
        for (int kk = rb; kk < re; ++kk) {
            const int k = a.col[kk];
            if (!strong[kk] || split[k] == C_NODE) continue;
            bool common = false;
            for (int mm = a.row[k]; mm < a.row[k + 1] && !common; ++mm)
                if (strong[mm] && cmark[a.col[mm]] == i) common = true;
            fstrong[kk - rb] = common;
        }
        double den = 0.0;
        for (int kk = rb; kk < re; ++kk) {
            const int k = a.col[kk];
            if (k == i) den += a.val[kk];
            else if (!(strong[kk] && (split[k] == C_NODE || fstrong[kk - rb]))) den += a.val[kk];
        }
        // inner denominators, in the order of row i's strong C entries
        for (int kk = rb; kk < re; ++kk) {
            if (!fstrong[kk - rb]) continue;
            const int k = a.col[kk];
            for (int mm = a.row[k]; mm < a.row[k + 1]; ++mm) {
                const int c = a.col[mm];
                if (!dset[c]) { dset[c] = 1; dense[c] = 0.0; }
                dense[c] += a.val[mm];
            }
            const double akk = dset[k] ? dense[k] : 0.0;
            double in = 0.0;
            for (int jj = rb; jj < re; ++jj) {
                if (!strong[jj] || split[a.col[jj]] != C_NODE) continue;
                const int m = a.col[jj];
                const double akm = dset[m] ? dense[m] : 0.0;
                if ((akm < 0) != (akk < 0) && akm != 0) in += akm;
            }
            for (int mm = a.row[k]; mm < a.row[k + 1]; ++mm) dset[a.col[mm]] = 0;
            inner[kk - rb] = in;
            if (in == 0) { den += a.val[kk]; fstrong[kk - rb] = 0; }
        }
        const size_t start = p.col.size();
        // numerators: a_ij + sum_k a_ik abar_kj / inner_k, k in row order
        std::vector<double> num;
        std::vector<int32_t> jcol;
        for (int jj = rb; jj < re; ++jj) {
            const int j = a.col[jj];
            if (!strong[jj] || split[j] != C_NODE) continue;
            num.push_back(a.val[jj]);
            jcol.push_back(j);
        }
        for (int kk = rb; kk < re; ++kk) {
            if (!fstrong[kk - rb]) continue;
            const int k = a.col[kk];
            for (int mm = a.row[k]; mm < a.row[k + 1]; ++mm) {
                const int c = a.col[mm];
                if (!dset[c]) { dset[c] = 1; dense[c] = 0.0; }
                dense[c] += a.val[mm];
            }
            const double akk = dset[k] ? dense[k] : 0.0;
            for (size_t q = 0; q < jcol.size(); ++q) {
                double akj = dset[jcol[q]] ? dense[jcol[q]] : 0.0;
                if (!((akj < 0) != (akk < 0) && akj != 0)) akj = 0;
                num[q] += a.val[kk] * akj / inner[kk - rb];
            }
            for (int mm = a.row[k]; mm < a.row[k + 1]; ++mm) dset[a.col[mm]] = 0;
        }
        for (size_t q = 0; q < jcol.size(); ++q) {
            p.col.push_back(cidx[jcol[q]]);
            p.val.push_back(-num[q] / den);
        }
        // ascending coarse columns (stable insertion sort, as the oracle)
        for (size_t x = start + 1; x < p.col.size(); ++x) {
            const int32_t cv = p.col[x];
            const double vv = p.val[x];
            size_t y = x;
            while (y > start && p.col[y - 1] > cv) { p.col[y] = p.col[y - 1]; p.val[y] = p.val[y - 1]; --y; }
            p.col[y] = cv;
            p.val[y] = vv;
        }
        p.row[i + 1] = (int32_t)p.col.size();
    }
    (void)smark;
    Csr r = transpose(p, nc);
    Csr ap = matmul(a, p, nc);
    Csr rap = matmul(r, ap, nc);
    h.P.push_back(std::move(p));
    h.R.push_back(std::move(r));
    h.split.push_back(std::move(split));
    h.A.push_back(std::move(rap));
    return true;
}

} // namespace

extern "C" {

int pp_amg_setup(int32_t n, const int32_t *row, const int32_t *col, const double *val, double theta, int32_t max_levels,
                 int32_t max_coarse, void **handle) {
    if (!row || !col || !val || !handle || n <= 0) {
        pp::set_error("pp_amg_setup: invalid argument");
        return PP_ERR_ARG;
    }
    auto h = std::make_unique<Hierarchy>();
    Csr a;
    a.n = n;
    a.row.assign(n + 1, 0);
    for (int i = 0; i < n; ++i) {
        for (int k = row[i]; k < row[i + 1]; ++k) {
            if (col[k] < 0) continue; // sentinel entries are dropped (amg.py:116-135)
            a.col.push_back(col[k]);
            a.val.push_back(val[k]);
        }
        a.row[i + 1] = (int32_t)a.col.size();
    }
    h->A.push_back(std::move(a));
    max_levels = std::min<int32_t>(max_levels, 32);
    while ((int)h->A.size() < max_levels && h->A.back().n > max_coarse) {
        if (!setup_level(*h, theta)) break;
    }
    *handle = h.release();
    return PP_OK;
}

int pp_amg_num_levels(void *handle) {
    if (!handle) return PP_ERR_ARG;
    return (int)static_cast<Hierarchy *>(handle)->A.size();
}

int pp_amg_level_sizes(void *handle, int32_t level, int32_t *n, int32_t *nnz_a, int32_t *nnz_p, int32_t *nnz_r) {
    Hierarchy *h = static_cast<Hierarchy *>(handle);
    if (!h || level < 0 || level >= (int)h->A.size()) { pp::set_error("pp_amg_level_sizes: level"); return PP_ERR_ARG; }
    if (n) *n = h->A[level].n;
    if (nnz_a) *nnz_a = h->A[level].nnz();
    const bool hasT = level < (int)h->P.size();
    if (nnz_p) *nnz_p = hasT ? h->P[level].nnz() : 0;
    if (nnz_r) *nnz_r = hasT ? h->R[level].nnz() : 0;
    return PP_OK;
}

int pp_amg_level_copy(void *handle, int32_t level, int32_t *a_row, int32_t *a_col, double *a_val, int32_t *p_row,
                      int32_t *p_col, double *p_val, int32_t *r_row, int32_t *r_col, double *r_val, int32_t *splitting) {
    Hierarchy *h = static_cast<Hierarchy *>(handle);
    if (!h || level < 0 || level >= (int)h->A.size()) { pp::set_error("pp_amg_level_copy: level"); return PP_ERR_ARG; }
    const Csr &a = h->A[level];
    if (a_row) std::memcpy(a_row, a.row.data(), sizeof(int32_t) * (a.n + 1));
    if (a_col) std::memcpy(a_col, a.col.data(), sizeof(int32_t) * a.nnz());
    if (a_val) std::memcpy(a_val, a.val.data(), sizeof(double) * a.nnz());
    if (level < (int)h->P.size()) {
        const Csr &p = h->P[level], &r = h->R[level];
        if (p_row) std::memcpy(p_row, p.row.data(), sizeof(int32_t) * (p.n + 1));
        if (p_col) std::memcpy(p_col, p.col.data(), sizeof(int32_t) * p.nnz());
        if (p_val) std::memcpy(p_val, p.val.data(), sizeof(double) * p.nnz());
        if (r_row) std::memcpy(r_row, r.row.data(), sizeof(int32_t) * (r.n + 1));
        if (r_col) std::memcpy(r_col, r.col.data(), sizeof(int32_t) * r.nnz());
        if (r_val) std::memcpy(r_val, r.val.data(), sizeof(double) * r.nnz());
        if (splitting) for (int i = 0; i < a.n; ++i) splitting[i] = h->split[level][i] == C_NODE;
    }
    return PP_OK;
}

// ---- pattern of A P ------------------------------------------------------------------
// The device refresh computes C = R (A P) as two row products (pp_galerkin_rows), the first into
// the symbolic pattern of A P: row i holds the union of the interpolation rows of row i's
// columns, sorted. Built once per interpolation, with a dense marker over the coarse columns.
// The second call (ap_col given) also checks that every entry of R (A P) lies in the coarse
// pattern, so the refresh needs no check of its own.
int pp_ap_pattern(int32_t n, const int32_t *a_row, const int32_t *a_col, int32_t nc, const int32_t *p_row,
                  const int32_t *p_col, const int32_t *r_row, const int32_t *r_col, const int32_t *c_row,
                  const int32_t *c_col, int32_t *ap_row, int32_t *ap_col) {
    if (!a_row || !a_col || !p_row || !p_col || !ap_row || n < 0 || nc < 0) {
        pp::set_error("pp_ap_pattern: null argument");
        return PP_ERR_ARG;
    }
    std::vector<int32_t> mark((size_t)nc, -1), cols;
    int64_t total = 0;
    if (!ap_col) ap_row[0] = 0;
    for (int i = 0; i < n; ++i) {
        cols.clear();
        for (int a = a_row[i]; a < a_row[i + 1]; ++a) {
            const int k = a_col[a];
            if (k < 0 || k >= n) { pp::set_error("pp_ap_pattern: column out of range"); return PP_ERR_ARG; }
            for (int q = p_row[k]; q < p_row[k + 1]; ++q) {
                const int J = p_col[q];
                if (J < 0 || J >= nc) { pp::set_error("pp_ap_pattern: interpolation column out of range"); return PP_ERR_ARG; }
                if (mark[J] != i) { mark[J] = i; cols.push_back(J); }
            }
        }
        total += (int64_t)cols.size();
        if (total > INT32_MAX) { pp::set_error("pp_ap_pattern: more than 2^31 entries"); return PP_ERR_ARG; }
        if (!ap_col) {
            ap_row[i + 1] = (int32_t)total;
        } else {
            if (ap_row[i + 1] - ap_row[i] != (int32_t)cols.size()) {
                pp::set_error("pp_ap_pattern: row offsets of another pattern");
                return PP_ERR_ARG;
            }
            std::sort(cols.begin(), cols.end());
            std::copy(cols.begin(), cols.end(), ap_col + ap_row[i]);
        }
    }
    if (!ap_col || !r_row || !r_col || !c_row || !c_col) return PP_OK;
    std::fill(mark.begin(), mark.end(), -1);
    for (int I = 0; I < nc; ++I) {
        for (int e = c_row[I]; e < c_row[I + 1]; ++e) mark[c_col[e]] = I;
        for (int t = r_row[I]; t < r_row[I + 1]; ++t) {
            const int i = r_col[t];
            if (i < 0 || i >= n) { pp::set_error("pp_ap_pattern: R column out of range"); return PP_ERR_ARG; }
            for (int e = ap_row[i]; e < ap_row[i + 1]; ++e)
                if (mark[ap_col[e]] != I) {
                    pp::set_error("pp_ap_pattern: coarse pattern misses an entry of R A P");
                    return PP_ERR_ARG;
                }
        }
    }
    return PP_OK;
}

int pp_amg_free(void *handle) {
    delete static_cast<Hierarchy *>(handle);
    return PP_OK;
}

} // extern "C"

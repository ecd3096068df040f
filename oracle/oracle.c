/*
 * oracle.c -- plain-C restatement of RAPtor's AMG solve path and Galerkin
 * product, used ONLY as a checker by tests/, the smoke check and the
 * cpu_baseline leg of bench.py.  The product (libraptor_amd.so) never links
 * or calls this file.
 *
 * Every function names the reference lines it restates (paths relative to
 * the reference's raptor/ directory).  The arithmetic order is the
 * reference's: sequential left-to-right sums from 0.0 (or from b_i for the
 * residual), one rounding per + and per *, no fused multiply-add (this file
 * is built with -ffp-contract=off).
 *
 * Pinning (what checks each function beyond its own formula) is recorded in
 * the tests that use it: reference golden vectors (the test_data directory), the
 * reference compiled from source (oracle/_ref), closed forms and brute force.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ZERO_TOL 1e-16 /* core/types.hpp:26 */

/* CSR_spmv, util/linalg/spmv.cpp:52-67 */
void orc_spmv(int n_rows, const int* rowptr, const int* col, const double* val,
              const double* x, double* y)
{
    for (int i = 0; i < n_rows; i++) {
        double s = 0;
        for (int j = rowptr[i]; j < rowptr[i + 1]; j++) s += val[j] * x[col[j]];
        y[i] = s;
    }
}

/* CSR_append, util/linalg/spmv.cpp:88-103 (row sum formed first, then added) */
void orc_spmv_append(int n_rows, const int* rowptr, const int* col, const double* val,
                     const double* x, double* y)
{
    for (int i = 0; i < n_rows; i++) {
        double s = 0;
        for (int j = rowptr[i]; j < rowptr[i + 1]; j++) s += val[j] * x[col[j]];
        y[i] += s;
    }
}

/* CSR_residual, util/linalg/spmv.cpp:69-85, via ParMatrix::residual
 * util/linalg/par_spmv.cpp:215-252 */
void orc_residual(int n_rows, const int* rowptr, const int* col, const double* val,
                  const double* x, const double* b, double* r)
{
    for (int i = 0; i < n_rows; i++) {
        double s = b[i];
        for (int j = rowptr[i]; j < rowptr[i + 1]; j++) s -= val[j] * x[col[j]];
        r[i] = s;
    }
}

/* Matrix::mult_T core/matrix.hpp:342-348 (zero b, then CSR_append_T
 * util/linalg/spmv.cpp:150-163: rows ascending, b[col] += v * x[row]) */
void orc_spmv_t(int n_rows, int n_cols, const int* rowptr, const int* col, const double* val,
                const double* x, double* y)
{
    for (int c = 0; c < n_cols; c++) y[c] = 0.0;
    for (int i = 0; i < n_rows; i++)
        for (int j = rowptr[i]; j < rowptr[i + 1]; j++) y[col[j]] += val[j] * x[i];
}

/* Weighted Jacobi: the serial body util/linalg/relax.cpp:10-43, which is the
 * formula the parallel jacobi_helper (util/linalg/par_relax.cpp:106-152)
 * intends; the parallel one never assigns its diagonal and so never updates
 * x.  With the diagonal stored first both visit the off-diagonals in the
 * same order. */
void orc_jacobi(int n_rows, const int* rowptr, const int* col, const double* val,
                const double* b, double* x, double* tmp, int sweeps, double omega)
{
    for (int it = 0; it < sweeps; it++) {
        for (int i = 0; i < n_rows; i++) tmp[i] = x[i];
        for (int i = 0; i < n_rows; i++) {
            int s = rowptr[i], e = rowptr[i + 1];
            if (s == e) continue;
            double row_sum = 0, diag = 0;
            for (int j = s; j < e; j++) {
                if (col[j] == i) diag = val[j];
                else row_sum += val[j] * tmp[col[j]];
            }
            if (fabs(diag) > ZERO_TOL)
                x[i] = ((1.0 - omega) * tmp[i]) + (omega * ((b[i] - row_sum) / diag));
        }
    }
}

/* Hybrid SOR: sor_helper util/linalg/par_relax.cpp:154-166 calling
 * SOR_forward :29-68 (one rank: the off_proc block is empty).  Restated
 * only so the V-cycle and PCG drivers can be compared with the reference's
 * default smoother; it is not part of the GPU path. */
void orc_sor(int n_rows, const int* rowptr, const int* col, const double* val,
             const double* b, double* x, int sweeps, double omega)
{
    for (int it = 0; it < sweeps; it++) {
        int start_on = 0;
        for (int i = 0; i < n_rows; i++) {
            double row_sum = 0, diag;
            int end_on = rowptr[i + 1];
            if (col[start_on] == i) {
                diag = val[start_on];
                start_on++;
            } else {
                continue; /* reference keeps start_on unchanged here */
            }
            for (int j = start_on; j < end_on; j++) row_sum += val[j] * x[col[j]];
            start_on = end_on;
            x[i] = (x[i] + omega * (b[i] - x[i] - row_sum)) / diag;
        }
    }
}

/* Vector::inner_product core/vector.cpp:155-165 */
double orc_dot(int n, const double* x, const double* y)
{
    double r = 0.0;
    for (int i = 0; i < n; i++) r += x[i] * y[i];
    return r;
}

/* ParVector::norm(2) core/par_vector.cpp:88-98 over Vector::norm
 * core/vector.cpp:107-118 (entries with |v| <= 1e-16 skipped; the local
 * root is undone with pow(.,2) and re-taken) */
double orc_norm2(int n, const double* x)
{
    double r = 0.0;
    for (int i = 0; i < n; i++)
        if (fabs(x[i]) > ZERO_TOL) r += pow(x[i], 2);
    r = pow(r, 1.0 / 2);
    r = pow(r, 2);
    return pow(r, 1. / 2);
}

/* Dense LU with partial pivoting, LAPACK dgetrf_ semantics (column-major,
 * first maximal |pivot|), applied to the row-major coarse matrix exactly as
 * ParMultilevel::duplicate_coarse does (multilevel/par_multilevel.hpp:
 * 295-331): the column-major reading of a row-major array is the
 * transpose, so this factors A_c^T.  Right-looking unblocked form. */
int orc_lu_factor(int n, double* a, int* piv)
{
    int info = 0;
    for (int j = 0; j < n; j++) {
        int p = j;
        double amax = fabs(a[j + j * n]);
        for (int i = j + 1; i < n; i++)
            if (fabs(a[i + j * n]) > amax) { amax = fabs(a[i + j * n]); p = i; }
        piv[j] = p;
        if (a[p + j * n] != 0.0) {
            if (p != j)
                for (int k = 0; k < n; k++) {
                    double t = a[j + k * n]; a[j + k * n] = a[p + k * n]; a[p + k * n] = t;
                }
            double r = 1.0 / a[j + j * n];
            for (int i = j + 1; i < n; i++) a[i + j * n] *= r;
        } else if (info == 0) {
            info = j + 1;
        }
        for (int k = j + 1; k < n; k++) {
            double u = a[j + k * n];
            for (int i = j + 1; i < n; i++) a[i + k * n] -= a[i + j * n] * u;
        }
    }
    return info;
}

/* dgetrs_('N') of the factor above (multilevel/par_multilevel.hpp:360-374):
 * row swaps, unit-lower forward substitution, upper back substitution. */
void orc_lu_solve(int n, const double* lu, const int* piv, double* b)
{
    for (int j = 0; j < n; j++)
        if (piv[j] != j) { double t = b[j]; b[j] = b[piv[j]]; b[piv[j]] = t; }
    for (int j = 0; j < n; j++) {
        double bj = b[j];
        for (int i = j + 1; i < n; i++) b[i] -= bj * lu[i + j * n];
    }
    for (int j = n - 1; j >= 0; j--) {
        b[j] /= lu[j + j * n];
        double bj = b[j];
        for (int i = 0; i < j; i++) b[i] -= bj * lu[i + j * n];
    }
}

/* In-row sort by column: sort_helper(CSRMatrix*) core/matrix.cpp:587-618
 * with vec_sort core/utilities.hpp:8-43.  Insertion sort keyed on column;
 * it is only ever applied to rows whose columns are distinct, where every
 * correct sort gives the same result. */
void orc_sort_rows(int n_rows, const int* rowptr, int* col, double* val)
{
    for (int i = 0; i < n_rows; i++) {
        for (int j = rowptr[i] + 1; j < rowptr[i + 1]; j++) {
            int c = col[j];
            double v = val ? val[j] : 0.0;
            int k = j - 1;
            while (k >= rowptr[i] && col[k] > c) {
                col[k + 1] = col[k];
                if (val) val[k + 1] = val[k];
                k--;
            }
            col[k + 1] = c;
            if (val) val[k + 1] = v;
        }
    }
}

/* move_diag_helper(CSRMatrix*) core/matrix.cpp:729-787: the diagonal entry
 * moves to the front of its row, the others keep their order. */
void orc_move_diag(int n_rows, const int* rowptr, int* col, double* val)
{
    for (int i = 0; i < n_rows; i++) {
        int s = rowptr[i];
        for (int j = s; j < rowptr[i + 1]; j++) {
            if (col[j] == i) {
                double t = val ? val[j] : 0.0;
                for (int k = j; k > s; k--) {
                    col[k] = col[k - 1];
                    if (val) val[k] = val[k - 1];
                }
                col[s] = i;
                if (val) val[s] = t;
                break;
            }
        }
    }
}

/* remove_duplicates_helper(CSRMatrix*) core/matrix.cpp:922-986 for rows
 * that are sorted: equal columns are summed in order, and an entry whose
 * magnitude is below 1e-16 is dropped.  Compacts in place; returns nnz. */
static int64_t remove_duplicates(int n_rows, int* rowptr, int* col, double* val)
{
    int orig_start = rowptr[0];
    for (int row = 0; row < n_rows; row++) {
        int new_start = rowptr[row];
        int orig_end = rowptr[row + 1];
        if (orig_end - orig_start == 0) {
            orig_start = orig_end;
            rowptr[row + 1] = rowptr[row];
            continue;
        }
        int c = col[orig_start];
        col[new_start] = c;
        val[new_start] = val[orig_start];
        int prev = c, ctr = 1;
        for (int j = orig_start + 1; j < orig_end; j++) {
            c = col[j];
            if (c == prev) {
                val[ctr - 1 + new_start] += val[j];
            } else {
                if (fabs(val[ctr - 1 + new_start]) < ZERO_TOL) ctr--;
                col[ctr + new_start] = c;
                val[ctr + new_start] = val[j];
                ctr++;
                prev = c;
            }
        }
        if (fabs(val[ctr - 1 + new_start]) < ZERO_TOL) ctr--;
        orig_start = orig_end;
        rowptr[row + 1] = rowptr[row] + ctr;
    }
    return rowptr[n_rows];
}

/* Grows a malloc'd CSR body. */
static int reserve(int** col, double** val, int64_t* cap, int64_t need)
{
    if (need <= *cap) return 0;
    int64_t nc = *cap ? *cap : 1024;
    while (nc < need) nc *= 2;
    int* c2 = (int*)realloc(*col, (size_t)nc * sizeof(int));
    if (!c2) return -1;
    *col = c2;
    double* v2 = (double*)realloc(*val, (size_t)nc * sizeof(double));
    if (!v2) return -1;
    *val = v2;
    *cap = nc;
    return 0;
}

/* C = A*B: spgemm_helper util/linalg/matmult.cpp:71-137 (dense sums, linked
 * list of touched columns, keep |sum| > 1e-16), then the one-rank
 * ParCSRMatrix::mult post-pass util/linalg/par_matmult.cpp:73-107,244-366 ->
 * add_append util/linalg/add.cpp:39-78 (sort rows, remove_duplicates).
 * Output rows are sorted ascending by column. */
int64_t orc_spgemm(int a_rows, const int* a_ptr, const int* a_col, const double* a_val,
                   int b_rows, int b_cols, const int* b_ptr, const int* b_col, const double* b_val,
                   int* c_ptr, int** c_col, double** c_val)
{
    (void)b_rows;
    int* next = (int*)malloc((size_t)(b_cols > 0 ? b_cols : 1) * sizeof(int));
    double* sums = (double*)calloc((size_t)(b_cols > 0 ? b_cols : 1), sizeof(double));
    int* col = NULL;
    double* val = NULL;
    int64_t cap = 0, nnz = 0;
    for (int k = 0; k < b_cols; k++) next[k] = -1;
    c_ptr[0] = 0;
    for (int i = 0; i < a_rows; i++) {
        int head = -2, length = 0;
        for (int j = a_ptr[i]; j < a_ptr[i + 1]; j++) {
            int ca = a_col[j];
            double va = a_val[j];
            for (int k = b_ptr[ca]; k < b_ptr[ca + 1]; k++) {
                int cb = b_col[k];
                sums[cb] += va * b_val[k];
                if (next[cb] == -1) { next[cb] = head; head = cb; length++; }
            }
        }
        if (reserve(&col, &val, &cap, nnz + length)) { nnz = -1; break; }
        for (int j = 0; j < length; j++) {
            if (fabs(sums[head]) > ZERO_TOL) { col[nnz] = head; val[nnz] = sums[head]; nnz++; }
            int t = head;
            head = next[head];
            next[t] = -1;
            sums[t] = 0;
        }
        c_ptr[i + 1] = (int)nnz;
    }
    free(next);
    free(sums);
    if (nnz >= 0) {
        orc_sort_rows(a_rows, c_ptr, col, val);
        nnz = remove_duplicates(a_rows, c_ptr, col, val);
    }
    *c_col = col;
    *c_val = val;
    return nnz;
}

/* C = P^T * B with P given as CSR: CSR_to_CSC core/matrix.cpp:492-538 (rows
 * ascending within each column), spgemm_T_helper util/linalg/matmult.cpp:
 * 139-204, then the one-rank mult_T_combine post-pass
 * util/linalg/par_matmult.cpp:387-570 -> add_append util/linalg/add.cpp:
 * 39-78.  Output rows are sorted ascending; callers that need the
 * reference's coarse layout apply orc_move_diag
 * (ruge_stuben/par_ruge_stuben_solver.hpp:142-143). */
int64_t orc_spgemm_tn(int p_rows, int p_cols, const int* p_ptr, const int* p_col, const double* p_val,
                      int b_cols, const int* b_ptr, const int* b_col, const double* b_val,
                      int* c_ptr, int** c_col, double** c_val)
{
    /* CSC view of P */
    int64_t pnnz = p_ptr[p_rows];
    int* t_ptr = (int*)calloc((size_t)p_cols + 1, sizeof(int));
    int* t_row = (int*)malloc((size_t)(pnnz > 0 ? pnnz : 1) * sizeof(int));
    double* t_val = (double*)malloc((size_t)(pnnz > 0 ? pnnz : 1) * sizeof(double));
    int* ctr = (int*)calloc((size_t)(p_cols > 0 ? p_cols : 1), sizeof(int));
    for (int64_t j = 0; j < pnnz; j++) t_ptr[p_col[j] + 1]++;
    for (int c = 1; c <= p_cols; c++) t_ptr[c] += t_ptr[c - 1];
    for (int i = 0; i < p_rows; i++)
        for (int j = p_ptr[i]; j < p_ptr[i + 1]; j++) {
            int c = p_col[j];
            int idx = t_ptr[c] + ctr[c]++;
            t_row[idx] = i;
            t_val[idx] = p_val[j];
        }
    free(ctr);

    int* next = (int*)malloc((size_t)(b_cols > 0 ? b_cols : 1) * sizeof(int));
    double* sums = (double*)calloc((size_t)(b_cols > 0 ? b_cols : 1), sizeof(double));
    int* col = NULL;
    double* val = NULL;
    int64_t cap = 0, nnz = 0;
    for (int k = 0; k < b_cols; k++) next[k] = -1;
    c_ptr[0] = 0;
    for (int i = 0; i < p_cols; i++) {
        int head = -2, length = 0;
        for (int j = t_ptr[i]; j < t_ptr[i + 1]; j++) {
            int k = t_row[j];
            double pv = t_val[j];
            for (int l = b_ptr[k]; l < b_ptr[k + 1]; l++) {
                int cb = b_col[l];
                sums[cb] += pv * b_val[l];
                if (next[cb] == -1) { next[cb] = head; head = cb; length++; }
            }
        }
        if (reserve(&col, &val, &cap, nnz + length)) { nnz = -1; break; }
        for (int j = 0; j < length; j++) {
            if (fabs(sums[head]) > ZERO_TOL) { col[nnz] = head; val[nnz] = sums[head]; nnz++; }
            int t = head;
            head = next[head];
            next[t] = -1;
            sums[t] = 0;
        }
        c_ptr[i + 1] = (int)nnz;
    }
    free(next);
    free(sums);
    free(t_ptr);
    free(t_row);
    free(t_val);
    if (nnz >= 0) {
        orc_sort_rows(p_cols, c_ptr, col, val);
        nnz = remove_duplicates(p_cols, c_ptr, col, val);
    }
    *c_col = col;
    *c_val = val;
    return nnz;
}

void orc_free(void* p) { free(p); }

/* ---------------------------------------------------------------------- */
/* Hierarchy drivers                                                        */
/* ---------------------------------------------------------------------- */

/* ---- multicolour smoothers (this build's labelled non-reference smoothers) ----
 * Greedy colouring of the graph of A + A^T without the diagonal: rows in
 * ascending order take the smallest colour no already-coloured neighbour
 * holds. */
int orc_greedy_colors(int n, const int* rowptr, const int* col, int* color)
{
    int* tp = (int*)calloc((size_t)n + 1, sizeof(int));
    for (int i = 0; i < n; i++)
        for (int j = rowptr[i]; j < rowptr[i + 1]; j++)
            if (col[j] != i) tp[col[j] + 1]++;
    for (int i = 0; i < n; i++) tp[i + 1] += tp[i];
    int* trow = (int*)malloc((size_t)(tp[n] > 0 ? tp[n] : 1) * sizeof(int));
    int* pos = (int*)malloc((size_t)(n > 0 ? n : 1) * sizeof(int));
    for (int i = 0; i < n; i++) pos[i] = tp[i];
    for (int i = 0; i < n; i++)
        for (int j = rowptr[i]; j < rowptr[i + 1]; j++)
            if (col[j] != i) trow[pos[col[j]]++] = i;
    int* taken = (int*)calloc((size_t)n + 1, sizeof(int)); /* taken[c] == i + 1: used next to row i */
    int ncolors = 0;
    for (int i = 0; i < n; i++) color[i] = -1;
    for (int i = 0; i < n; i++) {
        for (int j = rowptr[i]; j < rowptr[i + 1]; j++)
            if (col[j] != i && color[col[j]] >= 0) taken[color[col[j]]] = i + 1;
        for (int j = tp[i]; j < tp[i + 1]; j++)
            if (color[trow[j]] >= 0) taken[color[trow[j]]] = i + 1;
        int c = 0;
        while (c < ncolors && taken[c] == i + 1) c++;
        color[i] = c;
        if (c == ncolors) ncolors++;
    }
    free(tp); free(trow); free(pos); free(taken);
    return ncolors;
}

/* Multicolour sweeps: colours in increasing order (then decreasing for the
 * symmetric form), rows of a colour in ascending order; each row takes
 * the reference's per-row update with the current x -- SOR_forward's
 * x_i = (x_i + w (y_i - x_i - sum)) / a_ii on the way up and SOR_backward's
 * x_i = (1 - w) x_i + w ((y_i - sum) / a_ii) on the way down
 * (util/linalg/par_relax.cpp:29-104), sum over the off-diagonals in stored
 * order; a row whose first entry is not its diagonal keeps x.  Rows of one
 * colour are never neighbours, so any order within a colour gives these
 * bits. */
void orc_mc_sor(int n, const int* rowptr, const int* col, const double* val, const double* b, double* x,
                int sweeps, double omega, int symmetric)
{
    int* color = (int*)malloc((size_t)(n > 0 ? n : 1) * sizeof(int));
    const int nc = orc_greedy_colors(n, rowptr, col, color);
    for (int it = 0; it < sweeps; it++)
        for (int pass = 0; pass < (symmetric ? 2 : 1); pass++)
            for (int k = 0; k < nc; k++) {
                const int c = pass == 0 ? k : nc - 1 - k;
                for (int i = 0; i < n; i++) {
                    if (color[i] != c) continue;
                    int s0 = rowptr[i], e0 = rowptr[i + 1];
                    if (s0 == e0 || col[s0] != i) continue;
                    const double diag = val[s0];
                    double row_sum = 0;
                    for (int j = s0 + 1; j < e0; j++) row_sum += val[j] * x[col[j]];
                    if (pass == 0) x[i] = (x[i] + omega * (b[i] - x[i] - row_sum)) / diag;
                    else x[i] = ((1.0 - omega) * x[i]) + (omega * ((b[i] - row_sum) / diag));
                }
            }
    free(color);
}

static void relax(const orc_hier* h, int l, double* x, const double* b, double* tmp)
{
    const orc_csr* A = &h->A[l];
    if (h->relax_type == 0)
        orc_jacobi(A->n_rows, A->rowptr, A->col, A->val, b, x, tmp, h->num_sweeps, h->relax_weight);
    else if (h->relax_type == 3 || h->relax_type == 4)
        orc_mc_sor(A->n_rows, A->rowptr, A->col, A->val, b, x, h->num_sweeps, h->relax_weight, h->relax_type == 4);
    else
        orc_sor(A->n_rows, A->rowptr, A->col, A->val, b, x, h->num_sweeps, h->relax_weight);
}

/* ParMultilevel::cycle multilevel/par_multilevel.hpp:341-457: V(1,1) cycle.
 * Level vectors x_l, b_l, tmp_l are allocated per call. */
static void cycle_level(const orc_hier* h, int l, double* x, const double* b)
{
    const orc_csr* A = &h->A[l];
    if (l == h->num_levels - 1) {
        /* :353-386 -- dgetrs_ on the gathered b, copied to x */
        int n = h->coarse_n;
        double* bd = (double*)malloc((size_t)(n > 0 ? n : 1) * sizeof(double));
        memcpy(bd, b, (size_t)n * sizeof(double));
        orc_lu_solve(n, h->coarse_lu, h->coarse_piv, bd);
        memcpy(x, bd, (size_t)n * sizeof(double));
        free(bd);
        return;
    }
    const orc_csr* P = &h->P[l];
    int n = A->n_rows, nc = P->n_cols;
    double* tmp = (double*)malloc((size_t)(n > 0 ? n : 1) * sizeof(double));
    double* xc = (double*)calloc((size_t)(nc > 0 ? nc : 1), sizeof(double)); /* :389 */
    double* bc = (double*)malloc((size_t)(nc > 0 ? nc : 1) * sizeof(double));
    relax(h, l, x, b, tmp);                                                      /* :392-406 */
    orc_residual(n, A->rowptr, A->col, A->val, x, b, tmp);                       /* :409 */
    orc_spmv_t(P->n_rows, nc, P->rowptr, P->col, P->val, tmp, bc);                /* :411 */
    cycle_level(h, l + 1, xc, bc);                                               /* :423 */
    orc_spmv_append(P->n_rows, P->rowptr, P->col, P->val, xc, x);                /* :430 */
    relax(h, l, x, b, tmp);                                                      /* :432-446 */
    free(tmp);
    free(xc);
    free(bc);
}

void orc_cycle(const orc_hier* h, double* x, const double* b) { cycle_level(h, 0, x, b); }

/* ---- single-precision V-cycle ----
 * The reference has no single-precision cycle; this is cycle_level's
 * operation sequence (multilevel/par_multilevel.hpp:341-457) with every
 * matrix value rounded once to float and every vector, product and sum in
 * float, one rounding per + and per *.  The coarsest system is solved in
 * double on the widened right-hand side (dgetrs_ as above) and the solution
 * rounded once to float.  Jacobi's diagonal test compares with 1e-16
 * rounded to float. */
static void cycle_level_f(const orc_hier* h, int l, float* x, const float* b)
{
    const orc_csr* A = &h->A[l];
    if (l == h->num_levels - 1) {
        int n = h->coarse_n;
        double* bd = (double*)malloc((size_t)(n > 0 ? n : 1) * sizeof(double));
        for (int i = 0; i < n; i++) bd[i] = (double)b[i];
        orc_lu_solve(n, h->coarse_lu, h->coarse_piv, bd);
        for (int i = 0; i < n; i++) x[i] = (float)bd[i];
        free(bd);
        return;
    }
    const orc_csr* P = &h->P[l];
    const int n = A->n_rows, nc = P->n_cols;
    const float w = (float)h->relax_weight, tol = (float)ZERO_TOL;
    float* tmp = (float*)malloc((size_t)(n > 0 ? n : 1) * sizeof(float));
    float* xc = (float*)calloc((size_t)(nc > 0 ? nc : 1), sizeof(float));
    float* bc = (float*)malloc((size_t)(nc > 0 ? nc : 1) * sizeof(float));
    for (int pass = 0; pass < 2; pass++) {
        if (pass == 1) {
            /* Pt r into bc, recursion, then x += P xc (row sum first) */
            for (int i = 0; i < n; i++) {
                float s = b[i];
                for (int j = A->rowptr[i]; j < A->rowptr[i + 1]; j++) s -= (float)A->val[j] * x[A->col[j]];
                tmp[i] = s;
            }
            for (int c = 0; c < nc; c++) bc[c] = 0.0f;
            for (int i = 0; i < P->n_rows; i++)
                for (int j = P->rowptr[i]; j < P->rowptr[i + 1]; j++) bc[P->col[j]] += (float)P->val[j] * tmp[i];
            cycle_level_f(h, l + 1, xc, bc);
            for (int i = 0; i < P->n_rows; i++) {
                float s = 0.0f;
                for (int j = P->rowptr[i]; j < P->rowptr[i + 1]; j++) s += (float)P->val[j] * xc[P->col[j]];
                x[i] += s;
            }
        }
        for (int it = 0; it < h->num_sweeps; it++) {
            for (int i = 0; i < n; i++) tmp[i] = x[i];
            for (int i = 0; i < n; i++) {
                int s0 = A->rowptr[i], e0 = A->rowptr[i + 1];
                if (s0 == e0) continue;
                float row_sum = 0.0f, diag = 0.0f;
                for (int j = s0; j < e0; j++) {
                    if (A->col[j] == i) diag = (float)A->val[j];
                    else row_sum += (float)A->val[j] * tmp[A->col[j]];
                }
                if (fabsf(diag) > tol) x[i] = ((1.0f - w) * tmp[i]) + (w * ((b[i] - row_sum) / diag));
            }
        }
    }
    free(tmp);
    free(xc);
    free(bc);
}

void orc_cycle_fp32(const orc_hier* h, float* x, const float* b) { cycle_level_f(h, 0, x, b); }

/* ParMultilevel::solve multilevel/par_multilevel.hpp:459-538 */
int orc_amg_solve(const orc_hier* h, double* x, const double* b, double solve_tol,
                  int max_iterations, double* residuals)
{
    const orc_csr* A = &h->A[0];
    int n = A->n_rows, iter = 0;
    double b_norm = orc_norm2(n, b), r_norm;
    double* r = (double*)malloc((size_t)(n > 0 ? n : 1) * sizeof(double));
    orc_residual(n, A->rowptr, A->col, A->val, x, b, r);
    r_norm = fabs(b_norm) > ZERO_TOL ? orc_norm2(n, r) / b_norm : orc_norm2(n, r);
    if (residuals) residuals[iter] = r_norm;
    while (r_norm > solve_tol && iter < max_iterations) {
        orc_cycle(h, x, b);
        iter++;
        orc_residual(n, A->rowptr, A->col, A->val, x, b, r);
        r_norm = fabs(b_norm) > ZERO_TOL ? orc_norm2(n, r) / b_norm : orc_norm2(n, r);
        if (residuals) residuals[iter] = r_norm;
    }
    free(r);
    return iter;
}

static void push(double* res, int max_res, int* n_res, double v)
{
    if (*n_res < max_res) res[*n_res] = v;
    (*n_res)++;
}

/* PCG krylov/par_cg.cpp:121-263, including its quirks: an extra V-cycle on
 * b fixes the tolerance scale (tol *= sqrt(b.z0)); res[0] = sqrt(r0.z0) and
 * later entries are (r.z)/(b.z0); the stop test is r.z < tol; every 8th
 * iteration recomputes the true residual and restarts p = z.  Returns the
 * iteration count, or -1 where the reference exits on p.Ap < 0. */
int orc_pcg(const orc_csr* A, const orc_hier* h, double* x, const double* b,
            double tol, int max_iter, double* res, int max_res, int* n_res)
{
    int n = A->n_rows;
    size_t bytes = (size_t)(n > 0 ? n : 1) * sizeof(double);
    double *r = (double*)malloc(bytes), *z = (double*)malloc(bytes);
    double *p = (double*)malloc(bytes), *Ap = (double*)malloc(bytes);
    int iter = 0, recompute_r = 8, rc;
    *n_res = 0;
    if (max_iter <= 0) max_iter = ((int)(1.3 * n)) + 2;

    memset(z, 0, bytes);
    orc_cycle(h, z, b);
    double b_inner = orc_dot(n, b, z);
    double norm_b = sqrt(b_inner);
    if (norm_b > ZERO_TOL) tol = tol * norm_b;

    orc_residual(n, A->rowptr, A->col, A->val, x, b, r);
    memset(z, 0, bytes);
    orc_cycle(h, z, r);
    memcpy(p, z, bytes);
    double rz = orc_dot(n, r, z);
    push(res, max_res, n_res, sqrt(rz));

    while (iter < max_iter) {
        iter++;
        orc_spmv(n, A->rowptr, A->col, A->val, p, Ap);
        double App = orc_dot(n, Ap, p);
        if (App < 0.0) { iter = -1; break; }
        double alpha = rz / App;
        for (int i = 0; i < n; i++) x[i] += p[i] * alpha;
        int full_r = recompute_r && iter % recompute_r == 0;
        if (full_r) orc_residual(n, A->rowptr, A->col, A->val, x, b, r);
        else for (int i = 0; i < n; i++) r[i] += Ap[i] * (-1.0 * alpha);
        memset(z, 0, bytes);
        orc_cycle(h, z, r);
        double next = orc_dot(n, r, z);
        double beta = next / rz;
        push(res, max_res, n_res, next / b_inner);
        if (next < tol) break;
        if (full_r) memcpy(p, z, bytes);
        else {
            for (int i = 0; i < n; i++) p[i] *= beta;
            for (int i = 0; i < n; i++) p[i] += z[i] * 1.0;
        }
        rz = next;
    }
    rc = iter;
    free(r); free(z); free(p); free(Ap);
    return rc;
}

/* CG krylov/par_cg.cpp:8-118 (unpreconditioned; res entries are
 * ||r||/||b||; recompute the true residual at iteration 0 and every 8th). */
int orc_cg(const orc_csr* A, double* x, const double* b, double tol, int max_iter,
           double* res, int max_res, int* n_res)
{
    int n = A->n_rows;
    size_t bytes = (size_t)(n > 0 ? n : 1) * sizeof(double);
    double *r = (double*)malloc(bytes), *p = (double*)malloc(bytes), *Ap = (double*)malloc(bytes);
    int iter = 0, recompute_r = 8;
    *n_res = 0;
    double b_norm = orc_norm2(n, b);
    if (b_norm < ZERO_TOL) b_norm = 1.0;
    if (max_iter <= 0) max_iter = ((int)(1.3 * n)) + 2;

    orc_residual(n, A->rowptr, A->col, A->val, x, b, r);
    memcpy(p, r, bytes);
    double rr = orc_dot(n, r, r);
    double norm_r = sqrt(rr);
    push(res, max_res, n_res, norm_r / b_norm);
    if (norm_r != 0.0) tol = tol * norm_r;

    while (norm_r > tol && iter < max_iter) {
        orc_spmv(n, A->rowptr, A->col, A->val, p, Ap);
        double App = orc_dot(n, Ap, p);
        if (App < 0.0) { iter = -1; break; }
        double alpha = rr / App;
        for (int i = 0; i < n; i++) x[i] += p[i] * alpha;
        if ((iter % recompute_r) && iter > 0)
            for (int i = 0; i < n; i++) r[i] += Ap[i] * (-1.0 * alpha);
        else
            orc_residual(n, A->rowptr, A->col, A->val, x, b, r);
        double next = orc_dot(n, r, r);
        double beta = next / rr;
        for (int i = 0; i < n; i++) p[i] *= beta;
        for (int i = 0; i < n; i++) p[i] += r[i] * 1.0;
        rr = next;
        norm_r = sqrt(rr);
        push(res, max_res, n_res, norm_r / b_norm);
        iter++;
    }
    free(r); free(p); free(Ap);
    return iter;
}

/* ---- BiCGStab ----
 * krylov/par_bicgstab.cpp:11-110 (BiCGStab) and :240-360 (Pre_BiCGStab,
 * the V-cycle from zero as the preconditioner on p and s) at one rank.
 * Every vector update is the reference's sequence of copy / axpy / scale
 * passes (axpy: y_i += a x_i; scale: p_i *= a), the scalars are formed in
 * its order, res[] holds ||r_i||_2 (ParVector::norm, entries <= 1e-16
 * skipped) starting with ||r_0||, and the loop runs while
 * ||r|| > tol ||r_0|| and iter < max_iter (max_iter <= 0: 1.3 n + 2).
 * Pre_BiCGStab in the reference leaves its iteration counter uninitialised
 * and its p_hat / s_hat unsized; here the counter starts at 0 and both
 * vectors have n entries (the algorithm the routine spells out). */
static int bicgstab_run(const orc_csr* A, const orc_hier* h, double* x, const double* b, double tol,
                        int max_iter, double* res, int max_res, int* n_res)
{
    const int n = A->n_rows;
    const size_t bytes = (size_t)(n > 0 ? n : 1) * sizeof(double);
    double *r = (double*)malloc(bytes), *rs = (double*)malloc(bytes), *s = (double*)malloc(bytes);
    double *p = (double*)malloc(bytes), *Ap = (double*)malloc(bytes), *As = (double*)malloc(bytes);
    double *ph = (double*)malloc(bytes), *sh = (double*)malloc(bytes);
    *n_res = 0;
    if (max_iter <= 0) max_iter = ((int)(1.3 * n)) + 2;
    orc_residual(n, A->rowptr, A->col, A->val, x, b, r);
    memcpy(rs, r, bytes);
    memcpy(p, r, bytes);
    double rr_inner = orc_dot(n, r, rs);
    double norm_r = orc_norm2(n, r);
    push(res, max_res, n_res, norm_r);
    if (norm_r != 0.0) tol = tol * norm_r;
    int iter = 0;
    while (norm_r > tol && iter < max_iter) {
        const double* pd = p;
        if (h) {
            memset(ph, 0, bytes);
            orc_cycle(h, ph, p);
            pd = ph;
        }
        orc_spmv(n, A->rowptr, A->col, A->val, pd, Ap);
        const double Apr_inner = orc_dot(n, Ap, rs);
        const double alpha = rr_inner / Apr_inner;
        for (int i = 0; i < n; i++) s[i] = r[i];
        for (int i = 0; i < n; i++) s[i] += (-1.0 * alpha) * Ap[i];
        const double* sd = s;
        if (h) {
            memset(sh, 0, bytes);
            orc_cycle(h, sh, s);
            sd = sh;
        }
        orc_spmv(n, A->rowptr, A->col, A->val, sd, As);
        const double As_inner = orc_dot(n, As, s);
        const double AsAs_inner = orc_dot(n, As, As);
        const double omega = As_inner / AsAs_inner;
        for (int i = 0; i < n; i++) x[i] += alpha * pd[i];
        for (int i = 0; i < n; i++) x[i] += omega * sd[i];
        for (int i = 0; i < n; i++) r[i] = s[i];
        for (int i = 0; i < n; i++) r[i] += (-1.0 * omega) * As[i];
        const double next_inner = orc_dot(n, r, rs);
        const double beta = (next_inner / rr_inner) * (alpha / omega);
        for (int i = 0; i < n; i++) p[i] *= beta;
        for (int i = 0; i < n; i++) p[i] += 1.0 * r[i];
        for (int i = 0; i < n; i++) p[i] += (-1.0 * beta * omega) * Ap[i];
        rr_inner = next_inner;
        norm_r = orc_norm2(n, r);
        push(res, max_res, n_res, norm_r);
        iter++;
    }
    free(r); free(rs); free(s); free(p); free(Ap); free(As); free(ph); free(sh);
    return iter;
}

int orc_bicgstab(const orc_csr* A, double* x, const double* b, double tol, int max_iter, double* res,
                 int max_res, int* n_res)
{
    return bicgstab_run(A, NULL, x, b, tol, max_iter, res, max_res, n_res);
}

int orc_pre_bicgstab(const orc_csr* A, const orc_hier* h, double* x, const double* b, double tol, int max_iter,
                     double* res, int max_res, int* n_res)
{
    return bicgstab_run(A, h, x, b, tol, max_iter, res, max_res, n_res);
}

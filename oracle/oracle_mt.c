/*
 * oracle_mt.c -- the oracle's V-cycle and a flexible PCG on every host core
 * (OpenMP), for the CPU baseline that bench.py times beside the GPU.  TEST
 * INFRASTRUCTURE ONLY, like oracle.c: the product never links or calls it.
 *
 * Row-parallel forms of oracle.c's routines with the same per-row operation
 * sequence, so every vector the cycle produces is bit-identical to
 * orc_cycle's whatever the thread count:
 *   - SpMV, residual, Jacobi and x += P e split rows over threads (each row
 *     is summed left to right as in util/linalg/spmv.cpp:52-103 and
 *     util/linalg/relax.cpp:10-43);
 *   - the restriction b_c = P^T r (util/linalg/spmv.cpp:150-163, a scatter
 *     over rows ascending) is a gather over R = P^T built by a counting sort
 *     with rows ascending (core/matrix.cpp:492-538), which adds each b_c
 *     entry's terms in the scatter's order.
 * Inner products are summed over fixed 4096-row chunks, chunk sums added in
 * chunk order: deterministic for any thread count (not the serial
 * left-to-right order of core/vector.cpp:155-165).
 */
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define ZERO_TOL 1e-16 /* core/types.hpp:26 */
#define CHUNK 4096
#define PAR_MIN 20000 /* rows below which a loop runs on one thread */

typedef struct {
    int n_rows, n_cols;
    int* rowptr;
    int* col;
    double* val;
} mt_csr;

struct orc_mt_hier {
    const orc_hier* h;
    mt_csr* R; /* R_l = P_l^T, num_levels - 1 entries */
    double** x;
    double** b;
    double** t;
};

void orc_mt_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }

int orc_mt_threads(void) { return omp_get_max_threads(); }

/* CSR_to_CSC, core/matrix.cpp:492-538: entries of column c in ascending row order */
static mt_csr transpose(const orc_csr* P)
{
    mt_csr R;
    const int nnz = P->rowptr[P->n_rows];
    R.n_rows = P->n_cols;
    R.n_cols = P->n_rows;
    R.rowptr = (int*)calloc((size_t)P->n_cols + 1, sizeof(int));
    R.col = (int*)malloc((size_t)(nnz > 0 ? nnz : 1) * sizeof(int));
    R.val = (double*)malloc((size_t)(nnz > 0 ? nnz : 1) * sizeof(double));
    for (int j = 0; j < nnz; j++) R.rowptr[P->col[j] + 1]++;
    for (int c = 0; c < P->n_cols; c++) R.rowptr[c + 1] += R.rowptr[c];
    int* pos = (int*)malloc((size_t)(P->n_cols > 0 ? P->n_cols : 1) * sizeof(int));
    memcpy(pos, R.rowptr, (size_t)P->n_cols * sizeof(int));
    for (int i = 0; i < P->n_rows; i++)
        for (int j = P->rowptr[i]; j < P->rowptr[i + 1]; j++) {
            const int q = pos[P->col[j]]++;
            R.col[q] = i;
            R.val[q] = P->val[j];
        }
    free(pos);
    return R;
}

orc_mt_hier* orc_mt_prepare(const orc_hier* h)
{
    orc_mt_hier* m = (orc_mt_hier*)calloc(1, sizeof(orc_mt_hier));
    const int L = h->num_levels;
    m->h = h;
    m->R = (mt_csr*)calloc((size_t)(L > 1 ? L - 1 : 1), sizeof(mt_csr));
    m->x = (double**)calloc((size_t)L, sizeof(double*));
    m->b = (double**)calloc((size_t)L, sizeof(double*));
    m->t = (double**)calloc((size_t)L, sizeof(double*));
    for (int l = 0; l < L; l++) {
        const size_t n = (size_t)(h->A[l].n_rows > 0 ? h->A[l].n_rows : 1);
        m->x[l] = (double*)malloc(n * sizeof(double));
        m->b[l] = (double*)malloc(n * sizeof(double));
        m->t[l] = (double*)malloc(n * sizeof(double));
        if (l < L - 1) m->R[l] = transpose(&h->P[l]);
    }
    return m;
}

void orc_mt_free(orc_mt_hier* m)
{
    if (!m) return;
    for (int l = 0; l < m->h->num_levels; l++) {
        free(m->x[l]);
        free(m->b[l]);
        free(m->t[l]);
        if (l < m->h->num_levels - 1) {
            free(m->R[l].rowptr);
            free(m->R[l].col);
            free(m->R[l].val);
        }
    }
    free(m->R);
    free(m->x);
    free(m->b);
    free(m->t);
    free(m);
}

/* y = A x (op 0), y = b - A x (op 1), y += A x (op 2), row by row */
static void mt_rows(int op, int n, const int* rp, const int* col, const double* val, const double* x,
                    const double* b, double* y)
{
#pragma omp parallel for schedule(static) if (n >= PAR_MIN)
    for (int i = 0; i < n; i++) {
        double s = op == 1 ? b[i] : 0.0;
        if (op == 1)
            for (int j = rp[i]; j < rp[i + 1]; j++) s -= val[j] * x[col[j]];
        else
            for (int j = rp[i]; j < rp[i + 1]; j++) s += val[j] * x[col[j]];
        if (op == 2) y[i] += s;
        else y[i] = s;
    }
}

/* orc_jacobi's sweeps (util/linalg/relax.cpp:10-43) */
static void mt_jacobi(const orc_csr* A, const double* b, double* x, double* tmp, int sweeps, double omega)
{
    const int n = A->n_rows;
    for (int it = 0; it < sweeps; it++) {
#pragma omp parallel for schedule(static) if (n >= PAR_MIN)
        for (int i = 0; i < n; i++) tmp[i] = x[i];
#pragma omp parallel for schedule(static) if (n >= PAR_MIN)
        for (int i = 0; i < n; i++) {
            const int s = A->rowptr[i], e = A->rowptr[i + 1];
            if (s == e) continue;
            double row_sum = 0, diag = 0;
            for (int j = s; j < e; j++) {
                if (A->col[j] == i) diag = A->val[j];
                else row_sum += A->val[j] * tmp[A->col[j]];
            }
            if (fabs(diag) > ZERO_TOL) x[i] = ((1.0 - omega) * tmp[i]) + (omega * ((b[i] - row_sum) / diag));
        }
    }
}

/* cycle_level of oracle.c (multilevel/par_multilevel.hpp:341-457), Jacobi smoothing */
static void mt_cycle_level(orc_mt_hier* m, int l, double* x, const double* b)
{
    const orc_hier* h = m->h;
    const orc_csr* A = &h->A[l];
    if (l == h->num_levels - 1) {
        const int n = h->coarse_n;
        double* bd = m->t[l];
        memcpy(bd, b, (size_t)n * sizeof(double));
        orc_lu_solve(n, h->coarse_lu, h->coarse_piv, bd);
        memcpy(x, bd, (size_t)n * sizeof(double));
        return;
    }
    const orc_csr* P = &h->P[l];
    const mt_csr* R = &m->R[l];
    const int n = A->n_rows, nc = P->n_cols;
    double* tmp = m->t[l];
    double* xc = m->x[l + 1];
    double* bc = m->b[l + 1];
    memset(xc, 0, (size_t)nc * sizeof(double));
    mt_jacobi(A, b, x, tmp, h->num_sweeps, h->relax_weight);
    mt_rows(1, n, A->rowptr, A->col, A->val, x, b, tmp);
    mt_rows(0, nc, R->rowptr, R->col, R->val, tmp, NULL, bc);
    mt_cycle_level(m, l + 1, xc, bc);
    mt_rows(2, P->n_rows, P->rowptr, P->col, P->val, xc, NULL, x);
    mt_jacobi(A, b, x, tmp, h->num_sweeps, h->relax_weight);
}

/* orc_cycle on all threads (Jacobi hierarchies only: relax_type 0) */
int orc_mt_cycle(orc_mt_hier* m, double* x, const double* b)
{
    if (m->h->relax_type != 0) return -1;
    mt_cycle_level(m, 0, x, b);
    return 0;
}

/* sum of a_i b_i over fixed CHUNK-row chunks, added in chunk order */
static double mt_dot(int n, const double* a, const double* b)
{
    const int nch = (n + CHUNK - 1) / CHUNK;
    double* part = (double*)malloc((size_t)(nch > 0 ? nch : 1) * sizeof(double));
#pragma omp parallel for schedule(static) if (n >= PAR_MIN)
    for (int c = 0; c < nch; c++) {
        const int e = (c + 1) * CHUNK < n ? (c + 1) * CHUNK : n;
        double s = 0.0;
        for (int i = c * CHUNK; i < e; i++) s += a[i] * b[i];
        part[c] = s;
    }
    double s = 0.0;
    for (int c = 0; c < nch; c++) s += part[c];
    free(part);
    return s;
}

/* Flexible PCG in FP64 preconditioned by one V-cycle from zero, stopping on
 * the true relative residual ||b - A x|| / ||b|| <= rtol: the algorithm of
 * the product's FP64-cycle ramd_fpcg (r -= alpha A p with a true-residual
 * confirmation at the stop; beta = -alpha (z.Ap) / (r.z)_old).  The outer
 * operator A is the caller's matrix.  Returns the iteration count (-1 on
 * p.Ap <= 0) and the final true relative residual in *relres. */
int orc_mt_fpcg(const orc_csr* A, orc_mt_hier* m, double* x, const double* b, double rtol, int max_iter,
                double* relres)
{
    const int n = A->n_rows;
    const size_t bytes = (size_t)(n > 0 ? n : 1) * sizeof(double);
    double *r = (double*)malloc(bytes), *z = (double*)malloc(bytes);
    double *p = (double*)malloc(bytes), *Ap = (double*)malloc(bytes);
    const double bn = sqrt(mt_dot(n, b, b));
    const double scale = bn > 0 ? bn : 1.0;
    mt_rows(1, n, A->rowptr, A->col, A->val, x, b, r);
    double rel = sqrt(mt_dot(n, r, r)) / scale;
    int it = 0;
    if (rel > rtol) {
        memset(z, 0, bytes);
        orc_mt_cycle(m, z, r);
        memcpy(p, z, bytes);
        double rz = mt_dot(n, r, z);
        while (it < max_iter) {
            mt_rows(0, n, A->rowptr, A->col, A->val, p, NULL, Ap);
            const double pap = mt_dot(n, p, Ap);
            if (!(pap > 0.0)) {
                it = -1;
                break;
            }
            const double alpha = rz / pap;
#pragma omp parallel for schedule(static) if (n >= PAR_MIN)
            for (int i = 0; i < n; i++) {
                x[i] += alpha * p[i];
                r[i] -= alpha * Ap[i];
            }
            it++;
            rel = sqrt(mt_dot(n, r, r)) / scale;
            if (rel <= rtol) {
                mt_rows(1, n, A->rowptr, A->col, A->val, x, b, r);
                rel = sqrt(mt_dot(n, r, r)) / scale;
                if (rel <= rtol) break;
            }
            memset(z, 0, bytes);
            orc_mt_cycle(m, z, r);
            const double zap = mt_dot(n, z, Ap);
            const double rzn = mt_dot(n, r, z);
            const double beta = -alpha * zap / rz;
#pragma omp parallel for schedule(static) if (n >= PAR_MIN)
            for (int i = 0; i < n; i++) p[i] = z[i] + beta * p[i];
            rz = rzn;
        }
    }
    *relres = rel;
    free(r);
    free(z);
    free(p);
    free(Ap);
    return it;
}

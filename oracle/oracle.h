/*
 * oracle.h -- CPU restatement of RAPtor's AMG solve path and Galerkin
 * product.  TEST INFRASTRUCTURE ONLY: tests/, the smoke check and the
 * cpu_baseline leg of bench.py use it as a checker; the product never
 * links or calls it.
 *
 * All matrices are CSR with int32 row pointers / columns and FP64 values,
 * exactly like raptor::CSRMatrix (core/matrix.hpp:392-394).  Every routine
 * is compiled with -ffp-contract=off so each a*b+c rounds twice, as the
 * reference's x86-64 build without -march does.
 */
#ifndef RAPTOR_ORACLE_H
#define RAPTOR_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- sparse matrix-vector family (util/linalg/spmv.cpp) ---- */
void orc_spmv(int n_rows, const int* rowptr, const int* col, const double* val,
              const double* x, double* y);
void orc_spmv_append(int n_rows, const int* rowptr, const int* col, const double* val,
                     const double* x, double* y);
void orc_residual(int n_rows, const int* rowptr, const int* col, const double* val,
                  const double* x, const double* b, double* r);
void orc_spmv_t(int n_rows, int n_cols, const int* rowptr, const int* col, const double* val,
                const double* x, double* y);

/* ---- relaxation ---- */
void orc_jacobi(int n_rows, const int* rowptr, const int* col, const double* val,
                const double* b, double* x, double* tmp, int sweeps, double omega);
void orc_sor(int n_rows, const int* rowptr, const int* col, const double* val,
             const double* b, double* x, int sweeps, double omega);

/* ---- vectors (core/vector.cpp, core/par_vector.cpp) ---- */
double orc_dot(int n, const double* x, const double* y);
double orc_norm2(int n, const double* x);

/* ---- dense LU of the coarsest level (LAPACK dgetrf_/dgetrs_ semantics on
 *      a row-major array, i.e. the factor/solve is of the transpose) ---- */
int  orc_lu_factor(int n, double* a, int* piv);
void orc_lu_solve(int n, const double* lu, const int* piv, double* b);

/* ---- layout helpers (core/matrix.cpp) ---- */
void orc_sort_rows(int n_rows, const int* rowptr, int* col, double* val);
void orc_move_diag(int n_rows, const int* rowptr, int* col, double* val);

/* ---- Galerkin product (util/linalg/matmult.cpp, add.cpp) ----
 * Output arrays are malloc'd; release them with orc_free. */
int64_t orc_spgemm(int a_rows, const int* a_ptr, const int* a_col, const double* a_val,
                   int b_rows, int b_cols, const int* b_ptr, const int* b_col, const double* b_val,
                   int* c_ptr, int** c_col, double** c_val);
int64_t orc_spgemm_tn(int p_rows, int p_cols, const int* p_ptr, const int* p_col, const double* p_val,
                      int b_cols, const int* b_ptr, const int* b_col, const double* b_val,
                      int* c_ptr, int** c_col, double** c_val);
void orc_free(void* p);

/* multicolour smoothers: greedy colouring of A + A^T, then sweeps colour by
 * colour with the reference's SOR_forward / SOR_backward row updates */
int  orc_greedy_colors(int n, const int* rowptr, const int* col, int* color);
void orc_mc_sor(int n, const int* rowptr, const int* col, const double* val, const double* b, double* x,
                int sweeps, double omega, int symmetric);

/* ---- hierarchy, V-cycle, stationary solve and PCG ---- */
typedef struct {
    int n_rows, n_cols;
    const int* rowptr;
    const int* col;
    const double* val;
} orc_csr;

typedef struct {
    int num_levels;
    const orc_csr* A;        /* num_levels entries, diagonal stored first */
    const orc_csr* P;        /* num_levels-1 entries */
    const double* coarse_lu; /* coarse_n*coarse_n, output of orc_lu_factor */
    const int* coarse_piv;
    int coarse_n;
    int relax_type;          /* 0 = weighted Jacobi, 1 = hybrid SOR, 3 / 4 = multicolour SOR / SSOR */
    int num_sweeps;
    double relax_weight;
} orc_hier;

void orc_cycle(const orc_hier* h, double* x, const double* b);
/* the same cycle in single precision (matrix values rounded once to float) */
void orc_cycle_fp32(const orc_hier* h, float* x, const float* b);
int  orc_amg_solve(const orc_hier* h, double* x, const double* b, double solve_tol,
                   int max_iterations, double* residuals);
int  orc_pcg(const orc_csr* A, const orc_hier* h, double* x, const double* b,
             double tol, int max_iter, double* res, int max_res, int* n_res);
/* BiCGStab and Pre_BiCGStab (krylov/par_bicgstab.cpp:11-110, 240-360) */
int  orc_bicgstab(const orc_csr* A, double* x, const double* b, double tol, int max_iter,
                  double* res, int max_res, int* n_res);
int  orc_pre_bicgstab(const orc_csr* A, const orc_hier* h, double* x, const double* b, double tol,
                      int max_iter, double* res, int max_res, int* n_res);
int  orc_cg(const orc_csr* A, double* x, const double* b, double tol, int max_iter,
            double* res, int max_res, int* n_res);

/* ---- oracle_mt.c: the cycle and a flexible PCG on every host core
 *      (OpenMP, row-parallel; cycle results bit-identical to orc_cycle) ---- */
typedef struct orc_mt_hier orc_mt_hier;
void orc_mt_set_threads(int n);
int  orc_mt_threads(void);
orc_mt_hier* orc_mt_prepare(const orc_hier* h);
void orc_mt_free(orc_mt_hier* m);
int  orc_mt_cycle(orc_mt_hier* m, double* x, const double* b);
int  orc_mt_fpcg(const orc_csr* A, orc_mt_hier* m, double* x, const double* b, double rtol, int max_iter,
                 double* relres);

#ifdef __cplusplus
}
#endif
#endif

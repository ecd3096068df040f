// ref_harness.cpp -- C entry points over the reference compiled from its own
// sources (see Makefile).  TEST INFRASTRUCTURE ONLY: the parity tests call it
// to obtain the reference's own results at one MPI rank; the product never
// links or calls it.  Outputs of variable size are stashed and fetched with
// ref_fetch().
#include <map>

#include "raptor.hpp"
#include "krylov/par_bicgstab.hpp"

using namespace raptor;

static std::vector<int> g_rp, g_col;
static std::vector<double> g_val;
static int g_nrows = 0, g_ncols = 0;
static ParMultilevel* g_ml = nullptr;
static bool g_sa = false;
static std::vector<double> g_weights;

static ParCSRMatrix* make_par(int n_rows, int n_cols, const int* rp, const int* col, const double* val)
{
    ParCSRMatrix* A = new ParCSRMatrix(n_rows, n_cols);
    A->on_proc->idx1[0] = 0;
    A->off_proc->idx1[0] = 0;
    for (int i = 0; i < n_rows; i++) {
        for (int j = rp[i]; j < rp[i + 1]; j++) {
            A->on_proc->idx2.emplace_back(col[j]);
            A->on_proc->vals.emplace_back(val ? val[j] : 1.0);
        }
        A->on_proc->idx1[i + 1] = A->on_proc->idx2.size();
        A->off_proc->idx1[i + 1] = 0;
    }
    A->on_proc->nnz = A->on_proc->idx2.size();
    A->off_proc->nnz = 0;
    A->finalize();
    return A;
}

static long stash(Matrix* M)
{
    g_nrows = M->n_rows;
    g_ncols = M->n_cols;
    g_rp.assign(M->idx1.begin(), M->idx1.begin() + M->n_rows + 1);
    g_col.assign(M->idx2.begin(), M->idx2.begin() + M->nnz);
    g_val.assign(M->vals.begin(), M->vals.begin() + M->nnz);
    return M->nnz;
}

extern "C" {

int ref_init()
{
    int f = 0;
    MPI_Initialized(&f);
    if (!f) {
        int argc = 0;
        char** argv = NULL;
        MPI_Init(&argc, &argv);
    }
    int n;
    MPI_Comm_size(MPI_COMM_WORLD, &n);
    return n;
}

int ref_last_n() { return g_nrows; }
int ref_last_ncols() { return g_ncols; }

void ref_fetch(int* rp, int* col, double* val)
{
    std::copy(g_rp.begin(), g_rp.end(), rp);
    std::copy(g_col.begin(), g_col.end(), col);
    if (val) std::copy(g_val.begin(), g_val.end(), val);
}

// par_stencil_grid with laplace_stencil_27pt (kind 0) or diffusion_stencil_2d (kind 1)
long ref_gallery(int kind, int m, double eps, double theta)
{
    double* st;
    ParCSRMatrix* A;
    if (kind == 0) {
        int grid[3] = {m, m, m};
        st = laplace_stencil_27pt();
        A = par_stencil_grid(st, grid, 3);
    } else {
        int grid[2] = {m, m};
        st = diffusion_stencil_2d(eps, theta);
        A = par_stencil_grid(st, grid, 2);
    }
    long nnz = stash(A->on_proc);
    delete[] st;
    delete A;
    return nnz;
}

long ref_strength(int n, const int* rp, const int* col, const double* val, int type, double theta)
{
    ParCSRMatrix* A = make_par(n, n, rp, col, val);
    ParCSRMatrix* S = A->strength(type == 0 ? Classical : Symmetric, theta);
    long nnz = stash(S->on_proc);
    delete S;
    delete A;
    return nnz;
}

// type: 0 RS, 1 CLJP, 2 Falgout, 3 PMIS, 4 HMIS
int ref_split(int n, const int* rp, const int* col, int type, const double* weights, int* states)
{
    ParCSRMatrix* S = make_par(n, n, rp, col, nullptr);
    aligned_vector<int> st, off;
    std::vector<double> w(weights, weights + n);
    switch (type) {
    case 0: split_rs(S, st, off); break;
    case 1: split_cljp(S, st, off, false, w.data()); break;
    case 2: split_falgout(S, st, off, false, w.data()); break;
    case 3: split_pmis(S, st, off, false, w.data()); break;
    case 4: split_hmis(S, st, off, false, w.data()); break;
    default: delete S; return -1;
    }
    std::copy(st.begin(), st.end(), states);
    delete S;
    return 0;
}

long ref_interp(int n, const int* rp, const int* col, const double* val, const int* srp, const int* scol,
                const double* sval, const int* states, int type, double filter)
{
    ParCSRMatrix* A = make_par(n, n, rp, col, val);
    ParCSRMatrix* S = make_par(n, n, srp, scol, sval);
    aligned_vector<int> st(states, states + n), off;
    ParCSRMatrix* P = nullptr;
    if (type == 0) P = direct_interpolation(A, S, st, off);
    else if (type == 1) P = mod_classical_interpolation(A, S, st, off);
    else P = extended_interpolation(A, S, st, off, filter);
    long nnz = stash(P->on_proc);
    delete P;
    delete S;
    delete A;
    return nnz;
}

// AP = A->mult(P) (ParCSRMatrix::mult, util/linalg/par_matmult.cpp:73-107)
long ref_spgemm(int an, int acols, const int* arp, const int* acol, const double* aval, int bn, int bcols,
                const int* brp, const int* bcol, const double* bval)
{
    ParCSRMatrix* A = make_par(an, acols, arp, acol, aval);
    ParCSRMatrix* B = make_par(bn, bcols, brp, bcol, bval);
    ParCSRMatrix* C = A->mult(B);
    long nnz = stash(C->on_proc);
    delete C;
    delete B;
    delete A;
    return nnz;
}

// Ac = (A P)^T-side product as in extend_hierarchy: AP->mult_T(P), sort, move_diag
long ref_rap(int n, const int* arp, const int* acol, const double* aval, int nc, const int* prp, const int* pcol,
             const double* pval, int move_diag_first)
{
    ParCSRMatrix* A = make_par(n, n, arp, acol, aval);
    A->on_proc->move_diag();
    ParCSRMatrix* P = make_par(n, nc, prp, pcol, pval);
    ParCSRMatrix* AP = A->mult(P);
    ParCSRMatrix* Ac = AP->mult_T(P);
    Ac->sort();
    if (move_diag_first) Ac->on_proc->move_diag();
    long nnz = stash(Ac->on_proc);
    delete Ac;
    delete AP;
    delete P;
    delete A;
    return nnz;
}

// ParRugeStubenSolver setup; relax 0 Jacobi (the reference's), 1 SOR, 2 SSOR.
// weights NULL -> the solver's own form_rand_weights stream.
int ref_hier_setup(int n, const int* rp, const int* col, const double* val, double theta, int coarsen, int interp,
                   int strength, int relax, double relax_weight, int max_coarse, int max_levels,
                   const double* weights)
{
    delete g_ml;
    coarsen_t ct[] = {RS, CLJP, Falgout, PMIS, HMIS};
    interp_t it[] = {Direct, ModClassical, Extended};
    relax_t rt[] = {Jacobi, SOR, SSOR};
    ParCSRMatrix* A = make_par(n, n, rp, col, val);
    g_sa = false;
    g_ml = new ParRugeStubenSolver(theta, ct[coarsen], it[interp], strength == 0 ? Classical : Symmetric, rt[relax]);
    g_ml->relax_weight = relax_weight;
    g_ml->max_coarse = max_coarse;
    g_ml->max_levels = max_levels;
    g_weights.clear();
    if (weights) {
        g_weights.assign(weights, weights + n);
        g_ml->weights = new double[n];
        std::copy(weights, weights + n, g_ml->weights);
    } else {
        g_weights.resize(n);
        srand(2448422);
        for (int i = 0; i < n; i++) g_weights[i] = double(rand()) / RAND_MAX;
    }
    g_ml->setup(A);
    delete A;
    return g_ml->num_levels;
}

// which: 0 = A_l, 1 = P_l
long ref_hier_level(int l, int which)
{
    ParLevel* L = g_ml->levels[l];
    if (which == 0) return stash(L->A->on_proc);
    return L->P ? stash(L->P->on_proc) : -1;
}

// C/F states of level l, recomputed from the stored A_l exactly as
// extend_hierarchy computes them (same strength, splitting and weights)
int ref_hier_states(int l, int* states)
{
    ParCSRMatrix* A = g_ml->levels[l]->A;
    ParCSRMatrix* S = A->strength(g_ml->strength_type, g_ml->strong_threshold);
    aligned_vector<int> st, off;
    if (g_sa) {
        // aggregates, recomputed as extend_hierarchy does
        aligned_vector<int> agg;
        mis2(S, st, off, false, g_weights.data());
        aggregate(A, S, st, off, agg, false);
        // aggregates are global column ids of A; below the fine level the
        // coarse operators keep the fine root rows as their global column
        // ids (fit_candidates copies them into T's column map), so map them
        // back to local indices
        std::map<int, int> local;
        for (int i = 0; i < (int)A->on_proc_column_map.size(); i++) local[A->on_proc_column_map[i]] = i;
        for (size_t i = 0; i < agg.size(); i++) states[i] = agg[i] < 0 ? agg[i] : local.at(agg[i]);
        delete S;
        return (int)agg.size();
    }
    ParRugeStubenSolver* rs = (ParRugeStubenSolver*)g_ml;
    switch (rs->coarsen_type) {
    case RS: if (l < 3) split_rs(S, st, off); else split_falgout(S, st, off, false, g_weights.data()); break;
    case CLJP: split_cljp(S, st, off, false, g_weights.data()); break;
    case Falgout: split_falgout(S, st, off, false, g_weights.data()); break;
    case PMIS: split_pmis(S, st, off, false, g_weights.data()); break;
    case HMIS: split_hmis(S, st, off, false, g_weights.data()); break;
    }
    std::copy(st.begin(), st.end(), states);
    delete S;
    return (int)st.size();
}

// coarse LU from dgetrf_ (column-major) and 1-based pivots
int ref_hier_coarse(double* lu, int* piv)
{
    int n = g_ml->coarse_n;
    if (lu) std::copy(g_ml->A_coarse.begin(), g_ml->A_coarse.end(), lu);
    if (piv) std::copy(g_ml->LU_permute.begin(), g_ml->LU_permute.end(), piv);
    return n;
}

void ref_hier_cycle(double* x, const double* b)
{
    int n = g_ml->levels[0]->A->global_num_rows;
    ParVector xv(n, n), bv(n, n);
    std::copy(x, x + n, xv.local.values.begin());
    std::copy(b, b + n, bv.local.values.begin());
    g_ml->cycle(xv, bv);
    std::copy(xv.local.values.begin(), xv.local.values.end(), x);
}

int ref_hier_solve(double* x, const double* b, double tol, int max_iter, double* res)
{
    int n = g_ml->levels[0]->A->global_num_rows;
    ParVector xv(n, n), bv(n, n);
    std::copy(x, x + n, xv.local.values.begin());
    std::copy(b, b + n, bv.local.values.begin());
    g_ml->solve_tol = tol;
    g_ml->max_iterations = max_iter;
    int it = g_ml->solve(xv, bv);
    std::copy(xv.local.values.begin(), xv.local.values.end(), x);
    aligned_vector<double>& r = g_ml->get_residuals();
    for (int i = 0; i <= it; i++) res[i] = r[i];
    return it;
}

// PCG(A, ml, x, b, res, tol, max_iter) with A the caller's (finalized) matrix
int ref_pcg(int n, const int* rp, const int* col, const double* val, double* x, const double* b, double tol,
            int max_iter, double* res, int max_res)
{
    ParCSRMatrix* A = make_par(n, n, rp, col, val);
    ParVector xv(n, n), bv(n, n);
    std::copy(x, x + n, xv.local.values.begin());
    std::copy(b, b + n, bv.local.values.begin());
    aligned_vector<double> r;
    PCG(A, g_ml, xv, bv, r, tol, max_iter);
    std::copy(xv.local.values.begin(), xv.local.values.end(), x);
    int k = std::min((int)r.size(), max_res);
    for (int i = 0; i < k; i++) res[i] = r[i];
    delete A;
    return (int)r.size();
}

// serial jacobi(CSRMatrix*, b, x, tmp, sweeps, omega), util/linalg/relax.cpp:10-43
void ref_serial_jacobi(int n, const int* rp, const int* col, const double* val, double* x, const double* b,
                       int sweeps, double omega)
{
    CSRMatrix A(n, n, (int)rp[n]);
    A.idx1.assign(rp, rp + n + 1);
    A.idx2.assign(col, col + rp[n]);
    A.vals.assign(val, val + rp[n]);
    A.nnz = rp[n];
    Vector xv(n), bv(n), tmp(n);
    std::copy(x, x + n, xv.values.begin());
    std::copy(b, b + n, bv.values.begin());
    jacobi(&A, bv, xv, tmp, sweeps, omega);
    std::copy(xv.values.begin(), xv.values.end(), x);
}

int ref_cg(int n, const int* rp, const int* col, const double* val, double* x, const double* b, double tol,
           int max_iter, double* res, int max_res)
{
    ParCSRMatrix* A = make_par(n, n, rp, col, val);
    ParVector xv(n, n), bv(n, n);
    std::copy(x, x + n, xv.local.values.begin());
    std::copy(b, b + n, bv.local.values.begin());
    aligned_vector<double> r;
    CG(A, xv, bv, r, tol, max_iter);
    std::copy(xv.local.values.begin(), xv.local.values.end(), x);
    int k = std::min((int)r.size(), max_res);
    for (int i = 0; i < k; i++) res[i] = r[i];
    delete A;
    return (int)r.size();
}

// BiCGStab (krylov/par_bicgstab.cpp:11-110) on one rank; returns the number
// of residuals the reference recorded
int ref_bicgstab(int n, const int* rp, const int* col, const double* val, double* x, const double* b, double tol,
                 int max_iter, double* res, int max_res)
{
    ParCSRMatrix* A = make_par(n, n, rp, col, val);
    ParVector xv(n, n), bv(n, n);
    std::copy(x, x + n, xv.local.values.begin());
    std::copy(b, b + n, bv.local.values.begin());
    aligned_vector<double> r;
    BiCGStab(A, xv, bv, r, tol, max_iter);
    std::copy(xv.local.values.begin(), xv.local.values.end(), x);
    int k = std::min((int)r.size(), max_res);
    for (int i = 0; i < k; i++) res[i] = r[i];
    delete A;
    return (int)r.size();
}

// ParSmoothedAggregationSolver setup (relax as in ref_hier_setup)
int ref_hier_setup_sa(int n, const int* rp, const int* col, const double* val, double theta, int strength,
                      int relax, double relax_weight, int max_coarse, int max_levels, const double* weights,
                      double prolong_weight, int prolong_steps)
{
    delete g_ml;
    relax_t rt[] = {Jacobi, SOR, SSOR};
    ParCSRMatrix* A = make_par(n, n, rp, col, val);
    g_sa = true;
    g_ml = new ParSmoothedAggregationSolver(theta, MIS, JacobiProlongation, strength == 0 ? Classical : Symmetric,
                                            rt[relax], prolong_steps, prolong_weight);
    g_ml->relax_weight = relax_weight;
    g_ml->max_coarse = max_coarse;
    g_ml->max_levels = max_levels;
    g_weights.clear();
    if (weights) {
        g_weights.assign(weights, weights + n);
        g_ml->weights = new double[n];
        std::copy(weights, weights + n, g_ml->weights);
    } else {
        g_weights.resize(n);
        srand(2448422);
        for (int i = 0; i < n; i++) g_weights[i] = double(rand()) / RAND_MAX;
    }
    g_ml->setup(A);
    delete A;
    return g_ml->num_levels;
}

int ref_mis2(int n, const int* srp, const int* scol, const double* weights, int* states)
{
    ParCSRMatrix* S = make_par(n, n, srp, scol, nullptr);
    aligned_vector<int> st, off;
    std::vector<double> w(weights, weights + n);
    mis2(S, st, off, false, w.data());
    std::copy(st.begin(), st.end(), states);
    delete S;
    return 0;
}

int ref_aggregate(int n, const int* rp, const int* col, const double* val, const int* srp, const int* scol,
                  const int* states, const double* rand_vals, int* aggs)
{
    ParCSRMatrix* A = make_par(n, n, rp, col, val);
    ParCSRMatrix* S = make_par(n, n, srp, scol, nullptr);
    aligned_vector<int> st(states, states + n), off, agg;
    std::vector<double> r;
    if (rand_vals) r.assign(rand_vals, rand_vals + n);
    int na = aggregate(A, S, st, off, agg, false, rand_vals ? r.data() : nullptr);
    std::copy(agg.begin(), agg.end(), aggs);
    delete S;
    delete A;
    return na;
}

// T of fit_candidates (stashed); R (n_aggs) written out
long ref_fit_candidates(int n, const int* rp, const int* col, const double* val, const int* aggs, int n_aggs,
                        const double* B, double* R)
{
    ParCSRMatrix* A = make_par(n, n, rp, col, val);
    aligned_vector<int> agg(aggs, aggs + n);
    aligned_vector<double> b(B, B + n), r;
    ParCSRMatrix* T = fit_candidates(A, n_aggs, agg, b, r, 1, false, 1e-10);
    std::copy(r.begin(), r.end(), R);
    long nnz = stash(T->on_proc);
    delete T;
    delete A;
    return nnz;
}

long ref_jacobi_prolongation(int n, const int* rp, const int* col, const double* val, int n_aggs, const int* trp,
                             const int* tcol, const double* tval, double omega, int steps)
{
    // finalize() leaves the rows sorted; the solver hands jacobi_prolongation
    // its level matrix with the diagonal first (strength() and aggregate()
    // sort it and move the diagonal, par_strength.cpp:369-370)
    ParCSRMatrix* A = make_par(n, n, rp, col, val);
    A->on_proc->move_diag();
    ParCSRMatrix* T = make_par(n, n_aggs, trp, tcol, tval);
    ParCSRMatrix* P = jacobi_prolongation(A, T, false, omega, steps);
    long nnz = stash(P->on_proc);
    delete P;
    delete T;
    delete A;
    return nnz;
}

}  // extern "C"

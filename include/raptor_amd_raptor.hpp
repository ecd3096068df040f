/*
 * raptor_amd_raptor.hpp -- drop-in adapter that puts libraptor_amd (the
 * MI355X solve path, include/raptor_amd.h) behind RAPtor's own C++
 * interface.  It is compiled against the reference's headers (the
 * directory holding raptor.hpp on the include path) and linked with
 * libraptor_amd.so beside the reference's library.
 *
 * What it provides, one rank per GPU:
 *
 *   raptor_amd::ParAMDRugeStubenSolver        a raptor::ParRugeStubenSolver
 *   raptor_amd::ParAMDSmoothedAggregationSolver a raptor::ParSmoothedAggregationSolver
 *
 *     - setup(Af) (the virtual ParMultilevel::setup,
 *       multilevel/par_multilevel.hpp:124) builds the hierarchy on the
 *       device with every public option of ParMultilevel
 *       (multilevel/par_multilevel.hpp:79-100,626-658) and of the solver
 *       (ruge_stuben/par_ruge_stuben_solver.hpp:15-25,
 *       aggregation/par_smoothed_aggregation_solver.hpp:17-32) forwarded:
 *       strong_threshold, strength_type, relax_type, num_smooth_sweeps,
 *       relax_weight, max_coarse, max_levels, weights, and coarsen_type /
 *       interp_type / interp_filter or prolong_weight / prolong_smooth_steps;
 *     - the device levels are mirrored into the inherited `levels` (A_l, P_l
 *       as ParCSRMatrix, the dense coarse LU by duplicate_coarse), so every
 *       inherited member keeps working on the host (print_hierarchy, and
 *       cycle / solve / PCG through a ParMultilevel*), unless
 *       mirror_host_levels is set to false before setup;
 *     - cycle(x, b) and solve(x, b) (non-virtual in the reference, so reached
 *       through the adapter's own static type) run on the device;
 *     - with track_times = true, setup_times and solve_times are filled with
 *       per-level seconds in the reference's layout (5 entries per level,
 *       the communication buckets 0 at one rank).
 *
 *   PCG(A, ml, x, b, res, tol, max_iter, precond_t, comm_t)
 *     the reference's PCG (krylov/par_cg.hpp:14-16) for an adapter `ml`:
 *     overload resolution picks it over the reference's PCG whenever ml has
 *     the adapter's static type.  Same rules (krylov/par_cg.cpp:121-263),
 *     same res entries appended, the same two lines printed, exit(-1) on an
 *     indefinite matrix; precond_t / comm_t get the device seconds of the
 *     V-cycles and of the inner products added.
 *
 * Smoothers: the reference's default relax_type is SOR, a sequential
 * sweep with no parallel form of the same bits; the device runs Jacobi
 * (relax_type == Jacobi, the formula of util/linalg/relax.cpp:10-43) or,
 * when gpu_relax_type is set explicitly to RAMD_MC_SOR / RAMD_MC_SSOR, a
 * labelled multicolour SOR/SSOR.  relax_type SOR/SSOR with gpu_relax_type
 * left at -1 is refused by setup (raptor_amd::Error), never silently
 * replaced.
 *
 * Errors: a failing library call throws raptor_amd::Error (a
 * std::runtime_error) carrying ramd_last_error()'s text.
 *
 * Requires one MPI rank (the multi-rank halo exchange is not part of this
 * library), a C++11 compiler and the reference's headers.
 */
#ifndef RAPTOR_AMD_RAPTOR_HPP
#define RAPTOR_AMD_RAPTOR_HPP

#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "aggregation/par_smoothed_aggregation_solver.hpp"
#include "krylov/par_cg.hpp"
#include "multilevel/par_multilevel.hpp"
#include "ruge_stuben/par_ruge_stuben_solver.hpp"

#include "raptor_amd.h"

namespace raptor_amd {

class Error : public std::runtime_error {
  public:
    Error(int code, const std::string& what) : std::runtime_error(what), code(code) {}
    int code;
};

inline void check(int rc, const char* call)
{
    if (rc == RAMD_OK) return;
    char buf[1024];
    ramd_last_error(buf, sizeof buf);
    throw Error(rc, std::string("raptor_amd: ") + call + ": " + buf);
}

// relax_t of the reference -> the device smoother (gpu_relax_type -1 =
// follow relax_type, which must then be Jacobi)
inline int device_relax(raptor::relax_t relax_type, int gpu_relax_type)
{
    if (gpu_relax_type >= 0) {
        if (gpu_relax_type != RAMD_JACOBI && gpu_relax_type != RAMD_MC_SOR && gpu_relax_type != RAMD_MC_SSOR)
            throw Error(RAMD_E_ARG, "raptor_amd: gpu_relax_type must be RAMD_JACOBI, RAMD_MC_SOR or RAMD_MC_SSOR");
        return gpu_relax_type;
    }
    if (relax_type == raptor::Jacobi) return RAMD_JACOBI;
    throw Error(RAMD_E_ARG,
                "raptor_amd: relax_type SOR/SSOR is a sequential sweep with no device form; set relax_type = "
                "Jacobi, or set gpu_relax_type = RAMD_MC_SOR / RAMD_MC_SSOR to choose the multicolour smoother");
}

// solver-specific options (overloads on the reference's solver classes)
inline void solver_opts(const raptor::ParRugeStubenSolver& s, ramd_amg_opts& o)
{
    check(ramd_amg_opts_default(RAMD_SOLVER_RS, &o), "ramd_amg_opts_default");
    o.coarsen_type = (int)s.coarsen_type;
    o.interp_type = (int)s.interp_type;
    o.interp_filter = s.interp_filter;
}

inline void solver_opts(const raptor::ParSmoothedAggregationSolver& s, ramd_amg_opts& o)
{
    check(ramd_amg_opts_default(RAMD_SOLVER_SA, &o), "ramd_amg_opts_default");
    if (s.agg_type != raptor::MIS || s.prolong_type != raptor::JacobiProlongation)
        throw Error(RAMD_E_ARG, "raptor_amd: only MIS aggregation with Jacobi prolongation exists");
    o.prolong_weight = s.prolong_weight;
    o.prolong_smooth_steps = s.prolong_smooth_steps;
}

// A host CSR block as a one-rank ParCSRMatrix (partition, finalize())
inline raptor::ParCSRMatrix* host_par_csr(int n_rows, int n_cols, const std::vector<int>& rp,
                                         const std::vector<int>& col, const std::vector<double>& val)
{
    raptor::ParCSRMatrix* M = new raptor::ParCSRMatrix(n_rows, n_cols);
    raptor::Matrix* on = M->on_proc;
    on->idx1.assign(rp.begin(), rp.end());
    on->idx2.assign(col.begin(), col.end());
    on->vals.assign(val.begin(), val.end());
    on->nnz = (int)col.size();
    M->finalize();
    return M;
}

// Downloads a device CSR handle into host arrays
inline void download(ramd_csr A, int& n_rows, int& n_cols, std::vector<int>& rp, std::vector<int>& col,
                     std::vector<double>& val)
{
    int32_t nr = 0, nc = 0;
    int64_t nnz = 0;
    check(ramd_csr_info(A, &nr, &nc, &nnz), "ramd_csr_info");
    n_rows = nr;
    n_cols = nc;
    rp.resize((size_t)nr + 1);
    col.resize((size_t)nnz);
    val.resize((size_t)nnz);
    check(ramd_csr_download(A, rp.data(), nnz ? col.data() : NULL, nnz ? val.data() : NULL), "ramd_csr_download");
}

template <class Base>
class AMDMultilevel : public Base {
  public:
    // Base's constructor arguments, e.g. (strong_threshold, coarsen_type,
    // interp_type, strength_type, relax_type) for ParRugeStubenSolver
    template <class... Args>
    explicit AMDMultilevel(Args&&... args) : Base(std::forward<Args>(args)...)
    {
        device = 0;
        gpu_relax_type = -1;
        mirror_host_levels = true;
        device_setup = 1;
        ctx = NULL;
        hier = NULL;
        n = 0;
        xd = bd = NULL;
    }

    ~AMDMultilevel()
    {
        release();
        if (ctx) ramd_finalize(ctx);
    }

    // ParMultilevel::setup (multilevel/par_multilevel.hpp:124-212)
    void setup(raptor::ParCSRMatrix* Af)
    {
        int nprocs = 1;
        RAPtor_MPI_Comm_size(RAPtor_MPI_COMM_WORLD, &nprocs);
        if (nprocs != 1) throw Error(RAMD_E_ARG, "raptor_amd: one MPI rank (one GPU) only");
        if (Af->off_proc->nnz != 0) throw Error(RAMD_E_ARG, "raptor_amd: off_proc entries at one rank");
        ramd_amg_opts o;
        solver_opts(*this, o);
        o.strong_threshold = this->strong_threshold;
        o.strength_type = (int)this->strength_type;
        o.num_smooth_sweeps = this->num_smooth_sweeps;
        o.relax_weight = this->relax_weight;
        o.max_coarse = this->max_coarse;
        o.max_levels = this->max_levels;
        o.weights = this->weights;  // NULL: the reference's srand(2448422) stream
        o.relax_type = device_relax(this->relax_type, gpu_relax_type);
        o.device_setup = device_setup;
        if (!ctx) check(ramd_init(device, &ctx), "ramd_init");
        release();
        raptor::Matrix* on = Af->on_proc;
        n = Af->local_num_rows;
        check(ramd_hier_setup(ctx, n, (int64_t)on->nnz, on->idx1.data(), on->idx2.data(), on->vals.data(), &o,
                              &hier),
              "ramd_hier_setup");
        check(ramd_malloc(ctx, sizeof(double) * (size_t)(n > 0 ? n : 1), (void**)&xd), "ramd_malloc");
        check(ramd_malloc(ctx, sizeof(double) * (size_t)(n > 0 ? n : 1), (void**)&bd), "ramd_malloc");
        int nl = 0;
        check(ramd_hier_info(hier, &nl, NULL, NULL, NULL), "ramd_hier_info");
        device_levels = nl;
        // as setup_helper: the weights array is dropped after the setup
        delete[] this->weights;
        this->weights = NULL;
        if (mirror_host_levels) mirror();
        if (this->track_times) {
            delete[] this->setup_times;
            this->setup_times = new double[5 * (size_t)(nl > this->max_levels ? nl : this->max_levels)]();
            check(ramd_hier_level_times(hier, 0, this->setup_times), "ramd_hier_level_times");
            delete[] this->solve_times;
            this->solve_times = new double[5 * (size_t)nl]();
            check(ramd_hier_track_times(hier, 1), "ramd_hier_track_times");
        }
    }

    // ParMultilevel::cycle(x, b, 0) (multilevel/par_multilevel.hpp:341-457),
    // one V-cycle of the whole device hierarchy from the fine level
    void cycle(raptor::ParVector& x, raptor::ParVector& b, int level = 0)
    {
        if (level != 0) throw Error(RAMD_E_ARG, "raptor_amd: the device cycle starts at level 0");
        to_device(x, b);
        check(ramd_vcycle(hier, xd, bd, 0), "ramd_vcycle");
        from_device(x);
        refresh_solve_times();
    }

    // ParMultilevel::solve (multilevel/par_multilevel.hpp:459-538)
    int solve(raptor::ParVector& x, raptor::ParVector& b)
    {
        to_device(x, b);
        std::vector<double> res((size_t)this->max_iterations + 1);
        int32_t iters = 0;
        check(ramd_amg_solve(hier, xd, bd, this->solve_tol, this->max_iterations, res.data(), &iters),
              "ramd_amg_solve");
        from_device(x);
        if (this->store_residuals) {
            this->residuals.resize((size_t)this->max_iterations + 1);
            for (int i = 0; i <= iters; i++) this->residuals[i] = res[(size_t)i];
        }
        refresh_solve_times();
        return iters;
    }

    // the reference's PCG on the device (see ::PCG below)
    void pcg(raptor::ParCSRMatrix* A, raptor::ParVector& x, raptor::ParVector& b, aligned_vector<double>& res,
             double tol, int max_iter, double* precond_t, double* comm_t)
    {
        if (A->local_num_rows != n || A->off_proc->nnz != 0)
            throw Error(RAMD_E_ARG, "raptor_amd: PCG matrix does not match the hierarchy (one rank)");
        raptor::Matrix* on = A->on_proc;
        ramd_csr Ad = NULL;
        check(ramd_csr_create(ctx, n, A->on_proc_num_cols, (int64_t)on->nnz, on->idx1.data(), on->idx2.data(),
                              on->vals.data(), &Ad),
              "ramd_csr_create");
        const int max_res_i = max_iter <= 0 ? (int)(1.3 * b.global_n) + 3 : max_iter + 1;
        std::vector<double> hist((size_t)(max_res_i < (1 << 22) ? max_res_i : (1 << 22)));
        int32_t n_res = 0, iters = 0;
        to_device(x, b);
        const int rc = ramd_pcg_timed(hier, Ad, xd, bd, tol, max_iter, hist.data(), (int32_t)hist.size(), &n_res,
                                      &iters, precond_t, comm_t);
        ramd_csr_destroy(Ad);
        if (rc == RAMD_E_INDEFINITE) {
            // krylov/par_cg.cpp:193-200
            printf("Indefinite matrix detected in CG! Aborting...\n");
            exit(-1);
        }
        check(rc, "ramd_pcg_timed");
        from_device(x);
        for (int i = 0; i < n_res && i < (int)hist.size(); i++) res.emplace_back(hist[(size_t)i]);
        const int max_it = max_iter <= 0 ? ((int)(1.3 * b.global_n)) + 2 : max_iter;
        // krylov/par_cg.cpp:247-260
        if (iters == max_it) printf("Max Iterations Reached.\n");
        else printf("%d Iteration required to converge\n", iters);
        printf("Relative Residual: %lg\n\n", res[(size_t)(iters > 0 ? iters - 1 : 0)]);
        refresh_solve_times();
    }

    ramd_ctx context() const { return ctx; }
    ramd_hier hierarchy() const { return hier; }
    int num_device_levels() const { return device_levels; }

    int device;              // GPU index (set before setup)
    int gpu_relax_type;      // -1: from relax_type (Jacobi only); else RAMD_JACOBI / RAMD_MC_SOR / RAMD_MC_SSOR
    bool mirror_host_levels; // copy the device levels into `levels` at setup
    int device_setup;        // 1: strength / splitting / interpolation kernels on the device

  private:
    void release()
    {
        if (xd) ramd_free(ctx, xd);
        if (bd) ramd_free(ctx, bd);
        xd = bd = NULL;
        if (hier) ramd_hier_destroy(hier);
        hier = NULL;
        device_levels = 0;
    }

    void to_device(raptor::ParVector& x, raptor::ParVector& b)
    {
        if (!hier) throw Error(RAMD_E_STATE, "raptor_amd: setup has not run");
        if (x.local_n != n || b.local_n != n) throw Error(RAMD_E_ARG, "raptor_amd: vector size mismatch");
        check(ramd_memcpy_h2d(ctx, xd, x.local.data(), sizeof(double) * (size_t)n), "ramd_memcpy_h2d");
        check(ramd_memcpy_h2d(ctx, bd, b.local.data(), sizeof(double) * (size_t)n), "ramd_memcpy_h2d");
    }

    void from_device(raptor::ParVector& x)
    {
        check(ramd_memcpy_d2h(ctx, x.local.data(), xd, sizeof(double) * (size_t)n), "ramd_memcpy_d2h");
    }

    void refresh_solve_times()
    {
        if (this->track_times && this->solve_times)
            check(ramd_hier_level_times(hier, 1, this->solve_times), "ramd_hier_level_times");
    }

    // levels[l]->A / P as host ParCSRMatrix copies of the device levels, the
    // level vectors sized, and the dense coarse LU formed by the inherited
    // duplicate_coarse
    void mirror()
    {
        for (size_t i = 0; i < this->levels.size(); i++) delete this->levels[i];
        this->levels.clear();
        std::vector<int> rp, col;
        std::vector<double> val;
        int nr = 0, nc = 0;
        for (int l = 0; l < device_levels; l++) {
            ramd_csr A = NULL, P = NULL;
            check(ramd_hier_level(hier, l, &A, &P), "ramd_hier_level");
            raptor::ParLevel* L = new raptor::ParLevel();
            download(A, nr, nc, rp, col, val);
            L->A = host_par_csr(nr, nc, rp, col, val);
            L->A->on_proc->move_diag();  // finalize() sorted the rows; the levels keep the diagonal first
            L->x.resize(nr, nr);
            L->b.resize(nr, nr);
            L->tmp.resize(nr, nr);
            if (P) {
                download(P, nr, nc, rp, col, val);
                L->P = host_par_csr(nr, nc, rp, col, val);
            }
            this->levels.push_back(L);
        }
        this->num_levels = device_levels;
        this->duplicate_coarse();
    }

    ramd_ctx ctx;
    ramd_hier hier;
    int device_levels = 0;
    int n;
    double* xd;
    double* bd;
};

typedef AMDMultilevel<raptor::ParRugeStubenSolver> ParAMDRugeStubenSolver;
typedef AMDMultilevel<raptor::ParSmoothedAggregationSolver> ParAMDSmoothedAggregationSolver;

}  // namespace raptor_amd

// PCG(A, ml, x, b, res, tol, max_iter, precond_t, comm_t) of
// krylov/par_cg.hpp:14-16 for a device hierarchy: an exact match for an
// adapter pointer, so it is chosen over the reference's ParMultilevel*
// overload.
template <class Base>
void PCG(raptor::ParCSRMatrix* A, raptor_amd::AMDMultilevel<Base>* ml, raptor::ParVector& x, raptor::ParVector& b,
         aligned_vector<double>& res, double tol = 1e-05, int max_iter = -1, double* precond_t = NULL,
         double* comm_t = NULL)
{
    ml->pcg(A, x, b, res, tol, max_iter, precond_t, comm_t);
}

#endif /* RAPTOR_AMD_RAPTOR_HPP */

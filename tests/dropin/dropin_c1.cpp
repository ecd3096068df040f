// Drives RAPtor's own interface with the device hierarchy behind it
// (include/raptor_amd_raptor.hpp): the 100x100 rotated anisotropic
// diffusion problem (gallery/diffusion.hpp diffusion_stencil_2d(0.001, pi/8)
// on par_stencil_grid), b = A * ones, x0 = 0, ParRugeStubenSolver(0.25, CLJP,
// ModClassical, Classical, Jacobi) with relax_weight = 2/3 and track_times,
// then PCG(A, ml, x, b, res, 1e-8, -1, &precond_t, &comm_t) exactly as a
// reference driver writes it (examples/benchmark_solve.cpp).
//
// Usage: dropin_c1 <m> <x_out.bin> [sor]
//   prints RAMD_* lines (levels, iterations, residual history, timing
//   buckets, a check that the mirrored host levels equal the device ones);
//   writes x as raw doubles.  With "sor" the solver keeps the reference's
//   default relax_type (SOR): setup must refuse it (RAMD_REFUSED).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "core/types.hpp"
#include "gallery/diffusion.hpp"
#include "gallery/par_stencil.hpp"
#include "raptor_amd_raptor.hpp"

using namespace raptor;

static bool mirror_equal(raptor_amd::ParAMDRugeStubenSolver* ml)
{
    for (int l = 0; l < ml->num_levels; l++) {
        ramd_csr A = NULL, P = NULL;
        raptor_amd::check(ramd_hier_level(ml->hierarchy(), l, &A, &P), "ramd_hier_level");
        ramd_csr M[2] = {A, P};
        ParCSRMatrix* H[2] = {ml->levels[l]->A, ml->levels[l]->P};
        for (int k = 0; k < 2; k++) {
            if (!M[k]) {
                if (H[k]) return false;
                continue;
            }
            std::vector<int> rp, col;
            std::vector<double> val;
            int nr, nc;
            raptor_amd::download(M[k], nr, nc, rp, col, val);
            Matrix* on = H[k]->on_proc;
            if (H[k]->local_num_rows != nr || H[k]->global_num_cols != nc) return false;
            if (!std::equal(rp.begin(), rp.end(), on->idx1.begin())) return false;
            // P is sorted by finalize(); compare row by row as sets of (col, val)
            for (int i = 0; i < nr; i++)
                for (int j = rp[i]; j < rp[i + 1]; j++) {
                    bool found = false;
                    for (int q = on->idx1[i]; q < on->idx1[i + 1]; q++)
                        if (on->idx2[q] == col[j] && on->vals[q] == val[j]) found = true;
                    if (!found) return false;
                }
        }
    }
    return true;
}

int main(int argc, char** argv)
{
    MPI_Init(&argc, &argv);
    const int m = argc > 1 ? atoi(argv[1]) : 100;
    const char* xout = argc > 2 ? argv[2] : NULL;
    const bool sor = argc > 3 && strcmp(argv[3], "sor") == 0;

    data_t* stencil = diffusion_stencil_2d(0.001, M_PI / 8.0);
    int grid[2] = {m, m};
    ParCSRMatrix* A = par_stencil_grid(stencil, grid, 2);
    ParVector x(A->global_num_rows, A->local_num_rows);
    ParVector b(A->global_num_rows, A->local_num_rows);
    x.set_const_value(1.0);
    A->mult(x, b);
    x.set_const_value(0.0);

    raptor_amd::ParAMDRugeStubenSolver* ml =
        sor ? new raptor_amd::ParAMDRugeStubenSolver(0.25, CLJP, ModClassical, Classical)
            : new raptor_amd::ParAMDRugeStubenSolver(0.25, CLJP, ModClassical, Classical, Jacobi);
    ml->relax_weight = 2.0 / 3.0;
    ml->track_times = true;
    int rc = 0;
    try {
        ml->setup(A);
    } catch (const raptor_amd::Error& e) {
        printf("RAMD_REFUSED %d %s\n", e.code, e.what());
        rc = sor ? 0 : 1;
        delete ml;
        delete A;
        delete[] stencil;
        MPI_Finalize();
        return rc;
    }
    if (sor) {
        printf("RAMD_NOT_REFUSED\n");
        rc = 1;
    }
    ml->print_hierarchy();
    printf("RAMD_LEVELS %d", ml->num_levels);
    for (int l = 0; l < ml->num_levels; l++) printf(" %d", ml->levels[l]->A->global_num_rows);
    printf("\n");
    printf("RAMD_MIRROR %s\n", mirror_equal(ml) ? "equal" : "different");

    aligned_vector<double> res;
    double precond_t = 0.0, comm_t = 0.0;
    // the reference's call, unchanged: overload resolution picks the device PCG
    PCG(A, ml, x, b, res, 1e-8, -1, &precond_t, &comm_t);
    printf("RAMD_ITERS %d\n", (int)res.size() - 1);
    printf("RAMD_RES");
    for (size_t i = 0; i < res.size(); i++) printf(" %.17g", res[i]);
    printf("\n");
    printf("RAMD_TIMES %.9g %.9g\n", precond_t, comm_t);
    printf("RAMD_SETUP_TIMES");
    for (int l = 0; l < ml->num_levels; l++) printf(" %.9g", ml->setup_times[5 * l]);
    printf("\nRAMD_SOLVE_TIMES");
    for (int l = 0; l < ml->num_levels; l++) printf(" %.9g", ml->solve_times[5 * l]);
    printf("\n");
    // ParMultilevel::solve through the adapter (stationary cycles, solve_tol 1e-7)
    ParVector x2(A->global_num_rows, A->local_num_rows);
    x2.set_const_value(0.0);
    const int it2 = ml->solve(x2, b);
    printf("RAMD_SOLVE %d %.17g\n", it2, ml->residuals[it2]);
    if (xout) {
        FILE* f = fopen(xout, "wb");
        if (!f || fwrite(x.local.data(), sizeof(double), x.local_n, f) != (size_t)x.local_n) rc = 1;
        if (f) fclose(f);
    }
    delete ml;
    delete A;
    delete[] stencil;
    MPI_Finalize();
    return rc;
}

// Host side of the AMG setup: the hierarchy pieces that stay on the CPU
// (strength of connection, C/F splitting, interpolation, aggregation) and
// the stencil generators.  Each routine reproduces the reference's one-rank
// result exactly: same visiting order, same tie rules, same floating-point
// operation sequence.
#pragma once

#include <cstdint>
#include <vector>

namespace ramd {
namespace host {

// CSR with int32 indices and FP64 values (raptor::CSRMatrix layout).
struct Csr {
    int n_rows = 0, n_cols = 0;
    std::vector<int> rowptr{0};
    std::vector<int> col;
    std::vector<double> val;
    int64_t nnz() const { return (int64_t)col.size(); }
};

// C/F states, core/types.hpp:31-37
constexpr int kNewSelection = 3, kNewUnselection = 2, kSelected = 1, kUnselected = 0, kUnassigned = -1,
              kNoNeighbors = -2;
constexpr double kZeroTol = 1e-16;  // core/types.hpp:26

// Row layout helpers (core/matrix.cpp:587-618, 729-787).
void sort_rows(Csr& A);
void move_diag(Csr& A);

// glibc's rand() (the TYPE_3 additive-feedback generator behind
// srand/rand), reimplemented so the library reproduces the reference's
// weight stream without touching the process-wide libc generator.
class GlibcRand {
public:
    explicit GlibcRand(unsigned seed);
    int next();
    static constexpr int kRandMax = 2147483647;

private:
    int32_t r_[34];
    int f_, b_;
};

// Weights of ParMultilevel::form_rand_weights, multilevel/par_multilevel.hpp:215-225.
std::vector<double> reference_weights(int n, int first_row = 0);

// Strength (par_strength.cpp).  A is sorted with the diagonal moved first,
// in place, exactly as the reference does before reading it.
Csr classical_strength(Csr& A, double theta);
Csr symmetric_strength(Csr& A, double theta);

// C/F splittings (ruge_stuben/par_cf_splitting.cpp, ruge_stuben/cf_splitting.cpp).
// S gets its diagonal moved first in place.  weights has >= n entries.
void split_cljp(Csr& S, const double* weights, std::vector<int>& states);
void split_pmis(Csr& S, const double* weights, std::vector<int>& states);
void split_hmis(Csr& S, const double* weights, std::vector<int>& states);
void split_falgout(Csr& S, const double* weights, std::vector<int>& states);
void split_rs(Csr& S, std::vector<int>& states);
// split_rs(CSRMatrix*, states, prev_states, second_pass) of cf_splitting.cpp
void split_rs_serial(Csr& S, std::vector<int>& states, bool prev_states, bool second_pass);
void hmis_first_pass(Csr& S, std::vector<int>& states);
void falgout_first_pass(Csr& S, std::vector<int>& states);

// Interpolation (ruge_stuben/par_interpolation.cpp).  P has one column per
// coarse point, numbered in increasing fine index.
Csr direct_interpolation(Csr& A, Csr& S, const std::vector<int>& states);
Csr mod_classical_interpolation(Csr& A, Csr& S, const std::vector<int>& states);
Csr extended_interpolation(Csr& A, Csr& S, const std::vector<int>& states, double filter);
void filter_interp(Csr& P, double filter_threshold);

// Smoothed aggregation (aggregation/*.cpp).  The Jacobi prolongation
// P = T - (omega D^-1 A) T is sa_scaled_operator, a sparse product (done
// on the device) and sa_subtract.
void mis2(const Csr& S, const double* weights, std::vector<int>& states);
int aggregate(Csr& A, Csr& S, const std::vector<int>& states, const double* rand_vals,
              std::vector<int>& aggregates);
Csr fit_candidates(int n_aggs, const std::vector<int>& aggregates, const std::vector<double>& B,
                   std::vector<double>& R);
Csr sa_scaled_operator(const Csr& A, double omega);
Csr sa_subtract(const Csr& P, const Csr& Q);

// Stencil generators (gallery/par_stencil.hpp:16-236).  stencil has 3^dim
// entries, grid has dim entries.
// Greedy colouring of the graph of A + A^T (diagonal ignored): rows in
// ascending order take the smallest colour no already-coloured neighbour
// has.  Returns the number of colours.  No reference counterpart (the
// multicolour smoothers are this build's).
int greedy_colors(int n, const int* rowptr, const int* col, std::vector<int>& color);

Csr stencil_grid(const double* stencil, const int* grid, int dim);
std::vector<double> laplace_stencil_27pt();
std::vector<double> diffusion_stencil_2d(double eps, double theta);
// 7-point finite-volume variable-coefficient diffusion on m^3 (no reference
// generator exists): cell coefficients k = exp(g), g ~ N(0,1) from `seed`,
// harmonic-mean faces, Dirichlet walls at half-cell distance.
Csr fv7_variable(int m, uint64_t seed);

}  // namespace host
}  // namespace ramd

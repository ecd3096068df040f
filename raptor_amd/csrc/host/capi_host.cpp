// C entry points of the host setup pieces (strength, splitting,
// interpolation, stencil generators).  They touch no device state, so they
// also work where no GPU is present; the parity tests call them directly.
#include <cstdlib>
#include <cstring>

#include "../status.hpp"
#include "setup.hpp"

using namespace ramd;
using host::Csr;

static Csr make_csr(int32_t n_rows, int32_t n_cols, const int32_t* rowptr, const int32_t* col, const double* val)
{
    RAMD_CHECK(n_rows >= 0 && n_cols >= 0 && rowptr && rowptr[0] == 0, RAMD_E_ARG, "bad CSR header");
    const int64_t nnz = rowptr[n_rows];
    RAMD_CHECK(nnz == 0 || col, RAMD_E_ARG, "NULL column array");
    for (int i = 0; i < n_rows; i++) RAMD_CHECK(rowptr[i + 1] >= rowptr[i], RAMD_E_ARG, "rowptr decreases at %d", i);
    for (int64_t j = 0; j < nnz; j++)
        RAMD_CHECK(col[j] >= 0 && col[j] < n_cols, RAMD_E_ARG, "column %d out of range [0,%d)", col[j], n_cols);
    Csr A;
    A.n_rows = n_rows;
    A.n_cols = n_cols;
    A.rowptr.assign(rowptr, rowptr + n_rows + 1);
    A.col.assign(col, col + nnz);
    if (val) A.val.assign(val, val + nnz);
    else A.val.assign((size_t)nnz, 0.0);
    return A;
}

template <class T>
static T* export_array(const std::vector<T>& v)
{
    T* p = (T*)std::malloc(sizeof(T) * (v.size() ? v.size() : 1));
    if (!p) throw std::bad_alloc();
    if (!v.empty()) std::memcpy(p, v.data(), sizeof(T) * v.size());
    return p;
}

extern "C" {

void ramd_host_free(void* p) { std::free(p); }

int ramd_host_strength(int32_t n, const int32_t* rowptr, const int32_t* col, const double* val, int strength_type,
                       double theta, int32_t** s_rowptr, int32_t** s_col, int64_t* s_nnz)
{
    return guarded([&] {
        RAMD_CHECK(s_rowptr && s_col && s_nnz && val, RAMD_E_ARG, "ramd_host_strength: NULL argument");
        Csr A = make_csr(n, n, rowptr, col, val);
        Csr S = strength_type == RAMD_SYMMETRIC ? host::symmetric_strength(A, theta)
                                                : host::classical_strength(A, theta);
        *s_rowptr = export_array(S.rowptr);
        *s_col = export_array(S.col);
        *s_nnz = S.nnz();
    });
}

int ramd_host_split(int32_t n, const int32_t* s_rowptr, const int32_t* s_col, int coarsen_type,
                    const double* weights, int32_t* states)
{
    return guarded([&] {
        RAMD_CHECK(states, RAMD_E_ARG, "ramd_host_split: NULL states");
        Csr S = make_csr(n, n, s_rowptr, s_col, nullptr);
        std::vector<double> w;
        if (weights) w.assign(weights, weights + n);
        else w = host::reference_weights(n, 0);
        std::vector<int> st;
        switch (coarsen_type) {
        case RAMD_RS: host::split_rs(S, st); break;
        case RAMD_CLJP: host::split_cljp(S, w.data(), st); break;
        case RAMD_FALGOUT: host::split_falgout(S, w.data(), st); break;
        case RAMD_PMIS: host::split_pmis(S, w.data(), st); break;
        case RAMD_HMIS: host::split_hmis(S, w.data(), st); break;
        default: fail(RAMD_E_ARG, "ramd_host_split: unknown coarsen_type %d", coarsen_type);
        }
        std::copy(st.begin(), st.end(), states);
    });
}

int ramd_host_split_rs_serial(int32_t n, const int32_t* s_rowptr, const int32_t* s_col, int second_pass,
                              int32_t* states)
{
    return guarded([&] {
        RAMD_CHECK(states, RAMD_E_ARG, "ramd_host_split_rs_serial: NULL states");
        Csr S = make_csr(n, n, s_rowptr, s_col, nullptr);
        std::vector<int> st;
        host::split_rs_serial(S, st, false, second_pass != 0);
        std::copy(st.begin(), st.end(), states);
    });
}

int ramd_host_interp(int32_t n, const int32_t* rowptr, const int32_t* col, const double* val,
                     const int32_t* s_rowptr, const int32_t* s_col, const int32_t* states, int interp_type,
                     double filter, int32_t* n_coarse, int32_t** p_rowptr, int32_t** p_col, double** p_val,
                     int64_t* p_nnz)
{
    return guarded([&] {
        RAMD_CHECK(val && states && n_coarse && p_rowptr && p_col && p_val && p_nnz, RAMD_E_ARG,
                   "ramd_host_interp: NULL argument");
        Csr A = make_csr(n, n, rowptr, col, val);
        // S's values are A's values on S's pattern (classical and symmetric
        // strength both copy them); rebuild them so any pattern can be passed.
        Csr S = make_csr(n, n, s_rowptr, s_col, nullptr);
        host::sort_rows(A);
        for (int i = 0; i < n; i++)
            for (int j = S.rowptr[i]; j < S.rowptr[i + 1]; j++) {
                const int c = S.col[j];
                double v = 0.0;
                for (int k = A.rowptr[i]; k < A.rowptr[i + 1]; k++)
                    if (A.col[k] == c) {
                        v = A.val[k];
                        break;
                    }
                S.val[j] = v;
            }
        std::vector<int> st(states, states + n);
        Csr P;
        switch (interp_type) {
        case RAMD_DIRECT: P = host::direct_interpolation(A, S, st); break;
        case RAMD_MOD_CLASSICAL: P = host::mod_classical_interpolation(A, S, st); break;
        case RAMD_EXTENDED: P = host::extended_interpolation(A, S, st, filter); break;
        default: fail(RAMD_E_ARG, "ramd_host_interp: interp_type %d unknown", interp_type);
        }
        *n_coarse = P.n_cols;
        *p_rowptr = export_array(P.rowptr);
        *p_col = export_array(P.col);
        *p_val = export_array(P.val);
        *p_nnz = P.nnz();
    });
}

int ramd_host_gallery(int kind, int32_t m, double eps, double theta, uint64_t seed, int32_t* n, int32_t** rowptr,
                      int32_t** col, double** val, int64_t* nnz)
{
    return guarded([&] {
        RAMD_CHECK(n && rowptr && col && val && nnz && m > 0, RAMD_E_ARG, "ramd_host_gallery: bad argument");
        Csr A;
        if (kind == 0) {
            const int grid[3] = {m, m, m};
            std::vector<double> st = host::laplace_stencil_27pt();
            A = host::stencil_grid(st.data(), grid, 3);
        } else if (kind == 1) {
            const int grid[2] = {m, m};
            std::vector<double> st = host::diffusion_stencil_2d(eps, theta);
            A = host::stencil_grid(st.data(), grid, 2);
        } else if (kind == 2) {
            A = host::fv7_variable(m, seed);
        } else {
            fail(RAMD_E_ARG, "ramd_host_gallery: unknown kind %d", kind);
        }
        *n = A.n_rows;
        *rowptr = export_array(A.rowptr);
        *col = export_array(A.col);
        *val = export_array(A.val);
        *nnz = A.nnz();
    });
}

int ramd_host_greedy_colors(int32_t n, const int32_t* rowptr, const int32_t* col, int32_t* colors,
                            int32_t* n_colors)
{
    return guarded([&] {
        RAMD_CHECK(n >= 0 && rowptr && colors && n_colors && (col || rowptr[n] == 0), RAMD_E_ARG,
                   "ramd_host_greedy_colors: bad argument");
        for (int64_t j = 0; j < rowptr[n]; j++)
            RAMD_CHECK(col[j] >= 0 && col[j] < n, RAMD_E_ARG, "ramd_host_greedy_colors: column %d out of range",
                       col[j]);
        std::vector<int> c;
        *n_colors = host::greedy_colors(n, rowptr, col, c);
        std::copy(c.begin(), c.end(), colors);
    });
}

int ramd_host_mis2(int32_t n, const int32_t* s_rowptr, const int32_t* s_col, const double* weights,
                   int32_t* states)
{
    return guarded([&] {
        RAMD_CHECK(states, RAMD_E_ARG, "ramd_host_mis2: NULL states");
        Csr S = make_csr(n, n, s_rowptr, s_col, nullptr);
        std::vector<double> w;
        if (weights) w.assign(weights, weights + n);
        else w = host::reference_weights(n, 0);
        std::vector<int> st;
        host::mis2(S, w.data(), st);
        std::copy(st.begin(), st.end(), states);
    });
}

int ramd_host_aggregate(int32_t n, const int32_t* rowptr, const int32_t* col, const double* val,
                        const int32_t* s_rowptr, const int32_t* s_col, const int32_t* states,
                        const double* rand_vals, int32_t* aggregates, int32_t* n_aggs)
{
    return guarded([&] {
        RAMD_CHECK(val && states && aggregates && n_aggs, RAMD_E_ARG, "ramd_host_aggregate: NULL argument");
        Csr A = make_csr(n, n, rowptr, col, val);
        Csr S = make_csr(n, n, s_rowptr, s_col, nullptr);
        std::vector<int> st(states, states + n), agg;
        *n_aggs = host::aggregate(A, S, st, rand_vals, agg);
        std::copy(agg.begin(), agg.end(), aggregates);
    });
}

int ramd_host_fit_candidates(int32_t n, const int32_t* aggregates, int32_t n_aggs, const double* B, double* R,
                             int32_t** t_rowptr, int32_t** t_col, double** t_val, int64_t* t_nnz)
{
    return guarded([&] {
        RAMD_CHECK(n >= 0 && n_aggs >= 0 && aggregates && B && R && t_rowptr && t_col && t_val && t_nnz,
                   RAMD_E_ARG, "ramd_host_fit_candidates: bad argument");
        for (int i = 0; i < n; i++)
            RAMD_CHECK(aggregates[i] >= -1 && aggregates[i] < n, RAMD_E_ARG,
                       "ramd_host_fit_candidates: aggregate %d of row %d out of range", aggregates[i], i);
        std::vector<int> agg(aggregates, aggregates + n);
        std::vector<double> b(B, B + n), r;
        Csr T = host::fit_candidates(n_aggs, agg, b, r);
        std::copy(r.begin(), r.end(), R);
        *t_rowptr = export_array(T.rowptr);
        *t_col = export_array(T.col);
        *t_val = export_array(T.val);
        *t_nnz = T.nnz();
    });
}

}  // extern "C"

// Device-resident AMG hierarchy shared by the single-vector path
// (hier.hip) and the multi-vector path (block.hip).
#pragma once

#include <functional>
#include <vector>

#include "common.hpp"

struct ramd_hier_s;

namespace ramd {

ramd_csr_s* spgemm(ramd_csr_s* A, ramd_csr_s* B);
ramd_csr_s* galerkin(ramd_csr_s* A, ramd_csr_s* P);

struct Level {
    ramd_csr_s* A = nullptr;  // diagonal first
    ramd_csr_s* P = nullptr;  // n_l x n_{l+1}
    ramd_csr_s* R = nullptr;  // P^T (owned by P)
    double* diag = nullptr;   // a_ii (0 where absent)
    double *x = nullptr, *b = nullptr, *t = nullptr, *r = nullptr;
    float* diag32 = nullptr;
    float *x32 = nullptr, *b32 = nullptr, *t32 = nullptr, *r32 = nullptr;
    std::vector<int> states;
    int n = 0;
    // multicolour smoothers: rows of colour c are color_rows[color_ptr[c] ..
    // color_ptr[c+1]) (device), ascending
    std::vector<int> color_ptr;
    int* color_rows = nullptr;
};

// Largest coarse system solved by one wavefront (LU staged in LDS).
constexpr int kCoarseWave = 64;

// The substitutions of coarse_solve_wave: lane `lane` of the wavefront
// holds b_i (i = lane) and its pivot pv, the LU factors are in ls; returns
// x_i.  Every lane of the wavefront calls it.
__device__ __forceinline__ double coarse_lu_lane(int n, const double* ls, int pv, double v, int lane)
{
    for (int j = 0; j < n; j++) {
        const int p = __shfl(pv, j);
        if (p != j) {
            const double vj = __shfl(v, j), vp = __shfl(v, p);
            if (lane == j) v = vp;
            else if (lane == p) v = vj;
        }
    }
    for (int j = 0; j < n; j++) {
        const double bj = __shfl(v, j);
        if (lane > j && lane < n) v -= bj * ls[lane + j * n];
    }
    for (int j = n - 1; j >= 0; j--) {
        if (lane == j) v /= ls[j + j * n];
        const double bj = __shfl(v, j);
        if (lane < j) v -= bj * ls[lane + j * n];
    }
    return v;
}

// dgetrs_('N') on one right-hand side by one wavefront, n <= kCoarseWave:
// lane i holds b_i; the row swaps, the unit-lower and the upper substitution
// run in the reference BLAS column order with the same operations per entry
// as the workgroup kernel (b_i -= b_j * l_ij; b_j /= u_jj), so the result is
// bit-identical.  b and x are read and written with element stride `stride`;
// `ls` is n*n doubles of LDS.  All 64 lanes must call it.
template <typename TB, typename TX>
__device__ __forceinline__ void coarse_solve_wave(int n, const double* __restrict__ lu, const int* __restrict__ piv,
                                                  const TB* __restrict__ b, TX* __restrict__ x, size_t stride,
                                                  double* ls)
{
    const int lane = threadIdx.x;
    for (int t = lane; t < n * n; t += 64) ls[t] = lu[t];
    const int pv = lane < n ? piv[lane] : 0;
    double v = lane < n ? (double)b[(size_t)lane * stride] : 0.0;
    __syncthreads();
    v = coarse_lu_lane(n, ls, pv, v, lane);
    if (lane < n) x[(size_t)lane * stride] = (TX)v;
}

// Device setup steps (setup_gpu.hip), bit-identical to the host ones.
ramd_csr_s* gpu_classical_strength(ramd_csr_s* A, double theta);
ramd_csr_s* gpu_symmetric_strength(ramd_csr_s* A, double theta);
// PMIS on the device; prior_states (host, n; nullptr for plain PMIS) are
// states an earlier pass decided (HMIS), applied before the first round
void gpu_split_pmis(const ramd_csr_s* S, const double* weights, int* states, const int* prior_states = nullptr);
// CLJP on the device (prior_states: Falgout's Ruge-Stuben states, or nullptr)
void gpu_split_cljp(const ramd_csr_s* S, const double* weights, int* states, const int* prior_states = nullptr);
ramd_csr_s* gpu_mod_classical_interpolation(const ramd_csr_s* A, const ramd_csr_s* S, const int* states_d,
                                            const std::vector<int>& states_h);
ramd_csr_s* gpu_direct_interpolation(const ramd_csr_s* A, const ramd_csr_s* S, const int* states_d,
                                     const std::vector<int>& states_h);
ramd_csr_s* gpu_extended_interpolation(const ramd_csr_s* A, const ramd_csr_s* S, const int* states_d,
                                       const std::vector<int>& states_h, double filter);
// Smoothed aggregation's setup steps on the device, with the host
// restatement's results: MIS-2 of S (weights: host, n) into states (device);
// aggregation of A and S (both stored sorted, diagonal first) into agg
// (device; the root row or -1; rand_vals: host, n, or nullptr for zeros),
// returning the number of aggregates; the Jacobi prolongation's scaled
// operator and its subtraction P - Q.
void gpu_mis2(const ramd_csr_s* S, const double* weights, int* states);
int gpu_aggregate(const ramd_csr_s* A, const ramd_csr_s* S, const int* states, int* agg,
                  const double* rand_vals = nullptr);
// fit_candidates with one candidate: T (n x n_aggs) from the aggregates
// (device agg_d and its host copy) and the candidate B_d (device, n); the
// aggregate norms go to R_d (device, n_aggs)
ramd_csr_s* gpu_fit_candidates(ramd_ctx ctx, int n, int n_aggs, const int* agg_d, const std::vector<int>& agg_h,
                               const double* B_d, double* R_d);
ramd_csr_s* gpu_sa_scaled_operator(const ramd_csr_s* A, double omega);
ramd_csr_s* gpu_sa_subtract(const ramd_csr_s* P, const ramd_csr_s* Q);

// One FP64 V-cycle from x = 0 (ParMultilevel::cycle on the fine level).
void vcycle_from_zero(ramd_hier_s* h, double* x, const double* b);

struct BlockState;  // multi-vector work space (block.hip)
void block_free(BlockState* b);

// Device-time buckets: pairs of hipEvents recorded on a stream around a
// piece of work, summed per slot when resolved.  Used for the per-level
// solve times (ParMultilevel::track_times) and PCG's precond_t / comm_t.
struct EventSpans {
    struct Span {
        hipEvent_t b, e;
        int slot;
    };
    std::vector<Span> pending;
    std::vector<hipEvent_t> pool;
    // records the start event of a span for `slot`; returns its index
    int open(hipStream_t s, int slot);
    void close(hipStream_t s, int idx);
    // waits for the last recorded event and adds every span's seconds to acc[slot]
    void resolve(std::vector<double>& acc);
    void release();
};

// Stream capture of fn's launches into an instantiated graph (thread-local
// capture mode; no memset node is ever recorded: buffers are cleared by
// kernels).
hipGraphExec_t capture_graph(hipStream_t s, const std::function<void()>& fn);

// Captured flexible-PCG iteration of the single-vector solve: the operator
// step and the preconditioner application (V-cycle and inner products; the
// direction update that writes x runs eagerly), one pair per parity of the
// alternating r.z slots, valid for one (A, hierarchy forms, precision,
// partials).
struct FpcgGraphs {
    hipGraphExec_t gA[2] = {nullptr, nullptr}, gB[2] = {nullptr, nullptr};
    uint64_t a_uid = 0, key = 0;
    int prec = -1;
    const double* part = nullptr;
    uint64_t used = 0;  // last use (least recently used entry is replaced)
};
// captured iterations kept per hierarchy (per outer operator, hierarchy
// forms and precision; no captured kernel touches the solution vector, so
// one capture serves every solution vector)
constexpr int kFpcgGraphSets = 4;
void fpcg_graphs_free(FpcgGraphs* g);

}  // namespace ramd

struct ramd_hier_s {
    ramd_ctx ctx = nullptr;
    ramd_amg_opts opts{};
    std::vector<ramd::Level> levels;
    int coarse_n = 0;
    int coarse_info = 0;   // dgetrf's info for the coarse matrix (> 0: zero pivot)
    double* lu = nullptr;  // column-major LU of the coarse matrix read column-major (A_c^T)
    int* piv = nullptr;
    double* cb = nullptr;  // FP64 staging for the FP32 cycle's coarse solve
    double* cx = nullptr;
    double times[5] = {0, 0, 0, 0, 0};
    // host seconds of each level's setup step (the level's extend_hierarchy;
    // the coarsest level gets the dense factorisation), as the reference's
    // setup_times[5 * l] (multilevel/par_multilevel.hpp:171-208)
    std::vector<double> setup_level_s;
    // per-level device seconds of the cycle's work on that level while
    // tracking is on (solve_times[5 * l], multilevel/par_multilevel.hpp:343-456)
    bool track = false;
    std::vector<double> solve_level_s;
    ramd::EventSpans spans;
    bool fp32 = false;
    // coarse tail (the small coarsest levels' cycle as one workgroup
    // kernel; same bits either way): mode -1 off, 0 automatic (timed per
    // precision at the first cycle: tail_pick, -2 not yet), >= 1 forced to
    // start at that level; tail_force overrides both while timing
    int tail_mode = 0;
    int tail_pick[2] = {-2, -2};
    int tail_force = -2;
    // PCG work vectors (fine level)
    double *pr = nullptr, *pz = nullptr, *pp = nullptr, *pAp = nullptr;
    float *pr32 = nullptr, *pz32 = nullptr;
    ramd::BlockState* blk = nullptr;
    // captured FPCG iteration and its own partial-sum buffer (never
    // reallocated while graphs hold it)
    ramd::FpcgGraphs* fg = nullptr;  // kFpcgGraphSets entries (allocated at first use)
    uint64_t fg_tick = 0;
    double* fpart = nullptr;
    size_t fpart_cap = 0;
};

namespace ramd {
// Identity and buffer counts of A and of every level operator, and the
// coarse-tail switch: captured graphs hold raw pointers to all of their
// buffers and the launches these choose, and are keyed on this.
inline uint64_t forms_key(const ramd_hier_s* h, const ramd_csr_s* A)
{
    uint64_t k = A->uid * 1000003ull + A->forms;
    for (int t : {h->tail_mode, h->tail_pick[0], h->tail_pick[1]}) k = k * 131ull + (uint64_t)(t + 3);
    for (const Level& L : h->levels)
        for (const ramd_csr_s* M : {(const ramd_csr_s*)L.A, (const ramd_csr_s*)L.P, (const ramd_csr_s*)L.R})
            if (M) k = k * 1000003ull + M->uid * 31ull + M->forms;
    return k;
}
}  // namespace ramd

namespace ramd {
// Builds the FP32 copies of every level (values, diagonal, level vectors).
void enable_fp32(ramd_hier_s* h);
// Before cycling in precision prec (0 FP64, 1 FP32): the FP32 copies (prec
// 1) and, in the coarse tail's automatic mode, its timed choice.
void cycle_prepare(ramd_hier_s* h, int prec);
inline dim3 grid1(int64_t n, int bs = 256)
{
    int64_t g = (n + bs - 1) / bs;
    return dim3((unsigned)(g > 0 ? g : 1));
}
}  // namespace ramd

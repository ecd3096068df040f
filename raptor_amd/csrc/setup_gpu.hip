// Hierarchy setup on the device for the RS path the 27-point benchmarks use:
// classical strength (par_strength.cpp:8-340), PMIS splitting
// (ruge_stuben/par_cf_splitting.cpp:75-89, 1229-1381) and extended
// interpolation with its filter (ruge_stuben/par_interpolation.cpp:175-989),
// each with the same decisions and the same per-row arithmetic order as the
// host restatement (host/setup.cpp), so S, the C/F states and P are
// bit-identical.  Row pointers are formed on the host from device counts
// (the CSR handle keeps a host copy of them anyway).
#include <cmath>
#include <climits>
#include <cstring>
#include <stdexcept>
#include <vector>

#include "hier.hpp"
#include "host/setup.hpp"

namespace ramd {

constexpr int kSelectedState = 1, kUnselectedState = 0, kUnassignedState = -1, kNoNeighborsState = -2;
constexpr int kNewSelectionState = 3, kNewUnselectionState = 2;
constexpr double kSetupZeroTol = 1e-16;

// ---------------------------------------------------------------------------
// classical strength: one thread per row.  The row keeps its diagonal slot
// first (the stored diagonal, or (i, 0) when the first entry is not the
// diagonal), then every off-diagonal beyond theta times the row's extreme
// off-diagonal of the sign opposite to the diagonal, in stored order.
// ---------------------------------------------------------------------------
template <bool FILL>
__global__ void strength_kernel(int n, const int* __restrict__ rp, const int* __restrict__ col,
                                const double* __restrict__ val, double theta, int* __restrict__ cnt,
                                const int* __restrict__ srp, int* __restrict__ scol, double* __restrict__ sval)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int s = rp[i];
    const int e = rp[i + 1];
    if (e == s) {
        if (!FILL) cnt[i] = 0;
        return;
    }
    double diag = 0.0;
    if (col[s] == i) {
        diag = val[s];
        s++;
    }
    double row_scale;
    if (diag < 0.0) {
        row_scale = -(double)INT_MAX;
        for (int j = s; j < e; j++)
            if (val[j] > row_scale) row_scale = val[j];
    } else {
        row_scale = (double)INT_MAX;
        for (int j = s; j < e; j++)
            if (val[j] < row_scale) row_scale = val[j];
    }
    const double threshold = row_scale * theta;
    int o = FILL ? srp[i] : 0;
    if (FILL) {
        scol[o] = i;
        sval[o] = diag;
    }
    o++;
    for (int j = s; j < e; j++) {
        const double v = val[j];
        if (diag < 0 ? v > threshold : v < threshold) {
            if (FILL) {
                scol[o] = col[j];
                sval[o] = v;
            }
            o++;
        }
    }
    if (!FILL) cnt[i] = o;
}

// counts -> host exclusive prefix sum -> device row pointers
static std::vector<int> host_rowptr(ramd_ctx ctx, const int* d_cnt, int n)
{
    std::vector<int> cnt((size_t)n), rp((size_t)n + 1, 0);
    if (n) {
        RAMD_HIP(hipMemcpyAsync(cnt.data(), d_cnt, sizeof(int) * n, hipMemcpyDeviceToHost, ctx->stream));
        RAMD_HIP(hipStreamSynchronize(ctx->stream));
    }
    int64_t acc = 0;
    for (int i = 0; i < n; i++) {
        acc += cnt[i];
        RAMD_CHECK(acc <= INT32_MAX, RAMD_E_ARG, "setup product exceeds 2^31 stored entries");
        rp[i + 1] = (int)acc;
    }
    return rp;
}

static dim3 rows_grid(int n) { return dim3((unsigned)std::max(1, (n + 255) / 256)); }

ramd_csr_s* gpu_classical_strength(ramd_csr_s* A, double theta)
{
    ramd_ctx ctx = A->ctx;
    hipStream_t s = ctx->stream;
    const int n = A->n_rows;
    int* cnt = dalloc<int>((size_t)std::max(n, 1));
    std::vector<int> rp;
    try {
        hipLaunchKernelGGL((strength_kernel<false>), rows_grid(n), dim3(256), 0, s, n, A->rowptr, A->col, A->val,
                           theta, cnt, nullptr, nullptr, nullptr);
        RAMD_HIP(hipGetLastError());
        rp = host_rowptr(ctx, cnt, n);
    } catch (...) {
        dfree(cnt);
        throw;
    }
    dfree(cnt);
    const int64_t nnz = rp[n];
    int* d_rp = dalloc<int>((size_t)n + 1);
    int* d_col = dalloc<int>((size_t)std::max<int64_t>(nnz, 1));
    double* d_val = dalloc<double>((size_t)std::max<int64_t>(nnz, 1));
    RAMD_HIP(hipMemcpyAsync(d_rp, rp.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL((strength_kernel<true>), rows_grid(n), dim3(256), 0, s, n, A->rowptr, A->col, A->val, theta,
                       nullptr, d_rp, d_col, d_val);
    RAMD_HIP(hipGetLastError());
    return csr_wrap(ctx, n, A->n_cols, nnz, d_rp, d_col, d_val, rp.data());
}

// ---------------------------------------------------------------------------
// symmetric strength (par_strength.cpp:342-534): each row's sign flag and
// threshold (theta times its extreme off-diagonal of the sign opposite to
// the diagonal), then an entry is strong if it passes the test of its row
// or of its column's row.  The row's first stored entry is kept first as
// the diagonal slot, whatever its column (as there).
// ---------------------------------------------------------------------------
__global__ void sym_scale_kernel(int n, const int* __restrict__ rp, const int* __restrict__ col,
                                 const double* __restrict__ val, double theta, int* __restrict__ neg,
                                 double* __restrict__ scale)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int s = rp[i];
    const int e = rp[i + 1];
    if (e == s) {
        neg[i] = 0;
        scale[i] = 0.0;
        return;
    }
    double diag = 0.0;
    if (col[s] == i) {
        diag = val[s];
        s++;
    }
    double row_scale;
    if (diag < 0.0) {
        row_scale = -(double)INT_MAX;
        for (int j = s; j < e; j++)
            if (val[j] > row_scale) row_scale = val[j];
    } else {
        row_scale = (double)INT_MAX;
        for (int j = s; j < e; j++)
            if (val[j] < row_scale) row_scale = val[j];
    }
    neg[i] = diag < 0.0 ? 1 : 0;
    scale[i] = row_scale * theta;
}

template <bool FILL>
__global__ void sym_strength_kernel(int n, const int* __restrict__ rp, const int* __restrict__ col,
                                    const double* __restrict__ val, const int* __restrict__ neg,
                                    const double* __restrict__ scale, int* __restrict__ cnt,
                                    const int* __restrict__ srp, int* __restrict__ scol, double* __restrict__ sval)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int s = rp[i];
    const int e = rp[i + 1];
    if (e == s) {
        if (!FILL) cnt[i] = 0;
        return;
    }
    const bool nd = neg[i] != 0;
    const double th = scale[i];
    int o = FILL ? srp[i] : 0;
    if (FILL) {
        scol[o] = i;
        sval[o] = val[s];
    }
    o++;
    s++;
    for (int j = s; j < e; j++) {
        const double v = val[j];
        const int c = col[j];
        if ((nd && v > th) || (!nd && v < th) || (neg[c] && v > scale[c]) || (!neg[c] && v < scale[c])) {
            if (FILL) {
                scol[o] = c;
                sval[o] = v;
            }
            o++;
        }
    }
    if (!FILL) cnt[i] = o;
}

ramd_csr_s* gpu_symmetric_strength(ramd_csr_s* A, double theta)
{
    ramd_ctx ctx = A->ctx;
    hipStream_t s = ctx->stream;
    const int n = A->n_rows;
    const size_t m = (size_t)std::max(n, 1);
    struct Bufs {
        int *cnt = nullptr, *neg = nullptr;
        double* scale = nullptr;
        ~Bufs()
        {
            dfree(cnt);
            dfree(neg);
            dfree(scale);
        }
    } b;
    b.cnt = dalloc<int>(m);
    b.neg = dalloc<int>(m);
    b.scale = dalloc<double>(m);
    hipLaunchKernelGGL(sym_scale_kernel, rows_grid(n), dim3(256), 0, s, n, A->rowptr, A->col, A->val, theta, b.neg,
                       b.scale);
    hipLaunchKernelGGL((sym_strength_kernel<false>), rows_grid(n), dim3(256), 0, s, n, A->rowptr, A->col, A->val,
                       b.neg, b.scale, b.cnt, nullptr, nullptr, nullptr);
    RAMD_HIP(hipGetLastError());
    std::vector<int> rp = host_rowptr(ctx, b.cnt, n);
    const int64_t nnz = rp[n];
    int* d_rp = dalloc<int>((size_t)n + 1);
    int* d_col = dalloc<int>((size_t)std::max<int64_t>(nnz, 1));
    double* d_val = dalloc<double>((size_t)std::max<int64_t>(nnz, 1));
    RAMD_HIP(hipMemcpyAsync(d_rp, rp.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL((sym_strength_kernel<true>), rows_grid(n), dim3(256), 0, s, n, A->rowptr, A->col, A->val,
                       b.neg, b.scale, nullptr, d_rp, d_col, d_val);
    RAMD_HIP(hipGetLastError());
    RAMD_HIP(hipStreamSynchronize(s));  // the scratch buffers are freed on return
    return csr_wrap(ctx, n, A->n_cols, nnz, d_rp, d_col, d_val, rp.data());
}

// ---------------------------------------------------------------------------
// PMIS.  Every step of a round reads only what the previous step wrote:
// selection reads weights, marking reads the new selections, the update
// rewrites each vertex from its own state, so one thread per vertex gives
// the sequential loop's result.  The transpose of S (rows depending on a
// column) is built with atomics; only its sets are used.
// ---------------------------------------------------------------------------
__global__ void pmis_init_kernel(int n, const int* __restrict__ srp, const int* __restrict__ scol,
                                 int* __restrict__ states, int* __restrict__ tcnt)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int s = srp[i];
    const int e = srp[i + 1];
    states[i] = (e - s > 1) ? kUnassignedState : kNoNeighborsState;
    if (s < e && scol[s] == i) s++;
    for (int j = s; j < e; j++) atomicAdd(&tcnt[scol[j]], 1);
}

__global__ void pmis_transpose_kernel(int n, const int* __restrict__ srp, const int* __restrict__ scol,
                                      const int* __restrict__ tptr, int* __restrict__ tfill, int* __restrict__ tidx)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int s = srp[i];
    const int e = srp[i + 1];
    if (s < e && scol[s] == i) s++;
    for (int j = s; j < e; j++) {
        const int c = scol[j];
        tidx[tptr[c] + atomicAdd(&tfill[c], 1)] = i;
    }
}

// states already decided (HMIS: the Ruge-Stuben first pass) make the rows
// that strongly depend on a C point F (the loop before the first round of
// the reference's PMIS)
__global__ void pmis_prior_kernel(int n, const int* __restrict__ tptr, const int* __restrict__ tidx,
                                  int* __restrict__ states)
{
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= n || states[u] != kSelectedState) return;
    for (int j = tptr[u]; j < tptr[u + 1]; j++) {
        const int r = tidx[j];
        if (states[r] == kUnassignedState) states[r] = kUnselectedState;
    }
}

// CLJP weight updates after a round's selections (update_row_weights and
// update_local_dist2_weights, ruge_stuben/par_cf_splitting.cpp:497-640).
// Every strong edge loses its mark and lowers its target's weight by 1.0 at
// most once (an atomic exchange on the mark decides which thread does it),
// and a weight that stays >= 1 moves by exact integer steps, so the order of
// the decrements changes no result: a weight that falls below 1 makes its
// vertex F at the end of the round whatever its exact value.
__global__ void cljp_row_weights_kernel(int n, const int* __restrict__ srp, const int* __restrict__ scol,
                                        const int* __restrict__ states, int* __restrict__ edgemark,
                                        double* __restrict__ w, int new_state)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n || states[c] != new_state) return;
    int s = srp[c];
    const int e = srp[c + 1];
    if (s < e && scol[s] == c) s++;
    for (int j = s; j < e; j++) {
        const int idx = scol[j];
        if (states[idx] == kUnassignedState && atomicExch(&edgemark[j], 0) == 1) atomicAdd(&w[idx], -1.0);
    }
}

// row `r` of S (diagonal slot skipped) holds column c
__device__ __forceinline__ bool s_row_has(const int* __restrict__ srp, const int* __restrict__ scol, int r, int c)
{
    int s = srp[r];
    const int e = srp[r + 1];
    if (s < e && scol[s] == r) s++;
    for (int j = s; j < e; j++)
        if (scol[j] == c) return true;
    return false;
}

// for a new C point c: every row idx depending on c (not an earlier C point)
// loses the marks of its edges to unassigned rows that also depend on c
__global__ void cljp_dist2_kernel(int n, const int* __restrict__ srp, const int* __restrict__ scol,
                                  const int* __restrict__ tptr, const int* __restrict__ tidx,
                                  const int* __restrict__ states, int* __restrict__ edgemark, double* __restrict__ w,
                                  int new_state)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n || states[c] != new_state) return;
    for (int t = tptr[c]; t < tptr[c + 1]; t++) {
        const int idx = tidx[t];
        if (states[idx] == kSelectedState) continue;
        int is = srp[idx];
        const int ie = srp[idx + 1];
        if (is < ie && scol[is] == idx) is++;
        for (int k = is; k < ie; k++) {
            const int ik = scol[k];
            if (states[ik] == kUnassignedState && s_row_has(srp, scol, ik, c) && atomicExch(&edgemark[k], 0) == 1)
                atomicAdd(&w[ik], -1.0);
        }
    }
}

// initial_weights: the random value plus 1.0 once per row that strongly
// depends on i (added one at a time, as there); then the first pass over
// unassigned vertices with weight < 1
__global__ void pmis_weights_kernel(int n, const double* __restrict__ rnd, const int* __restrict__ tptr,
                                    int* __restrict__ states, double* __restrict__ w, int first_pass)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double v = rnd[i];
    for (int k = tptr[i]; k < tptr[i + 1]; k++) v += 1;
    if (first_pass && states[i] == kUnassignedState && v < 1) {
        states[i] = kUnselectedState;
        v = 0.0;
    } else if (states[i] != kUnassignedState) {
        v = 0.0;
    }
    w[i] = v;
}

__global__ void pmis_select_kernel(int n, const int* __restrict__ srp, const int* __restrict__ scol,
                                   const int* __restrict__ tptr, const int* __restrict__ tidx,
                                   const double* __restrict__ w, int* __restrict__ states)
{
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= n || states[u] != kUnassignedState) return;
    const double weight = w[u];
    if (0.0 > weight) return;
    int s = srp[u];
    const int e = srp[u + 1];
    if (s < e && scol[s] == u) s++;
    for (int j = s; j < e; j++)
        if (w[scol[j]] > weight) return;
    for (int j = tptr[u]; j < tptr[u + 1]; j++)
        if (w[tidx[j]] > weight) return;
    states[u] = kNewSelectionState;
}

__global__ void pmis_mark_kernel(int n, const int* __restrict__ tptr, const int* __restrict__ tidx,
                                 int* __restrict__ states)
{
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= n || states[u] != kNewSelectionState) return;
    for (int j = tptr[u]; j < tptr[u + 1]; j++) {
        const int r = tidx[j];
        if (states[r] == kUnassignedState) states[r] = kNewUnselectionState;
    }
}

__global__ void pmis_update_kernel(int n, int* __restrict__ states, double* __restrict__ w,
                                   int* __restrict__ remaining)
{
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= n) return;
    const int st = states[u];
    if (st == kNewSelectionState) {
        w[u] = 0.0;
        states[u] = kSelectedState;
    } else if (st == kUnassignedState || st == kNewUnselectionState) {
        if (w[u] < 1.0 || st == kNewUnselectionState) {
            w[u] = 0.0;
            states[u] = kUnselectedState;
        } else {
            atomicAdd(remaining, 1);
        }
    }
}

// states (device, n ints, caller-owned) of PMIS on S; weights = the random
// values (host, n)
__global__ void ones_kernel(int* __restrict__ p, int64_t n)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = 1;
}

// PMIS (cljp = false) or CLJP (cljp = true, cljp_main_loop :1383-1607)
static void gpu_split_mis(const ramd_csr_s* S, const double* weights, int* states, const int* prior_states,
                          bool cljp);

void gpu_split_pmis(const ramd_csr_s* S, const double* weights, int* states, const int* prior_states)
{
    gpu_split_mis(S, weights, states, prior_states, false);
}

void gpu_split_cljp(const ramd_csr_s* S, const double* weights, int* states, const int* prior_states)
{
    gpu_split_mis(S, weights, states, prior_states, true);
}

static void gpu_split_mis(const ramd_csr_s* S, const double* weights, int* states, const int* prior_states,
                          bool cljp)
{
    ramd_ctx ctx = S->ctx;
    hipStream_t s = ctx->stream;
    const int n = S->n_rows;
    struct Bufs {
        int *tcnt = nullptr, *tptr = nullptr, *tidx = nullptr, *rem = nullptr, *mark = nullptr;
        double *rnd = nullptr, *w = nullptr;
        ~Bufs()
        {
            dfree(tcnt);
            dfree(tptr);
            dfree(tidx);
            dfree(rem);
            dfree(mark);
            dfree(rnd);
            dfree(w);
        }
    } b;
    const size_t m = (size_t)std::max(n, 1);
    b.tcnt = dalloc<int>(m);
    b.rnd = dalloc<double>(m);
    b.w = dalloc<double>(m);
    b.rem = dalloc<int>(1);
    RAMD_HIP(hipMemsetAsync(b.tcnt, 0, sizeof(int) * m, s));
    RAMD_HIP(hipMemcpyAsync(b.rnd, weights, sizeof(double) * n, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(pmis_init_kernel, rows_grid(n), dim3(256), 0, s, n, S->rowptr, S->col, states, b.tcnt);
    RAMD_HIP(hipGetLastError());
    std::vector<int> tp = host_rowptr(ctx, b.tcnt, n);
    b.tptr = dalloc<int>((size_t)n + 1);
    b.tidx = dalloc<int>((size_t)std::max(tp[n], 1));
    RAMD_HIP(hipMemcpyAsync(b.tptr, tp.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice, s));
    RAMD_HIP(hipMemsetAsync(b.tcnt, 0, sizeof(int) * m, s));  // reused as fill counters
    hipLaunchKernelGGL(pmis_transpose_kernel, rows_grid(n), dim3(256), 0, s, n, S->rowptr, S->col, b.tptr, b.tcnt,
                       b.tidx);
    bool prior_c = false;
    if (prior_states) {
        RAMD_HIP(hipMemcpyAsync(states, prior_states, sizeof(int) * n, hipMemcpyHostToDevice, s));
        if (!cljp) hipLaunchKernelGGL(pmis_prior_kernel, rows_grid(n), dim3(256), 0, s, n, b.tptr, b.tidx, states);
        for (int i = 0; i < n && !prior_c; i++) prior_c = prior_states[i] == kSelectedState;
    }
    hipLaunchKernelGGL(pmis_weights_kernel, rows_grid(n), dim3(256), 0, s, n, b.rnd, b.tptr, states, b.w,
                       cljp ? 0 : 1);
    RAMD_HIP(hipGetLastError());
    if (cljp) {
        const int64_t ne = std::max<int64_t>(S->nnz, 1);
        b.mark = dalloc<int>((size_t)ne);
        hipLaunchKernelGGL(ones_kernel, dim3((unsigned)std::min<int64_t>((ne + 255) / 256, 8192)), dim3(256), 0, s,
                           b.mark, ne);
        if (prior_c) {  // the C points an earlier pass chose update the weights first
            hipLaunchKernelGGL(cljp_row_weights_kernel, rows_grid(n), dim3(256), 0, s, n, S->rowptr, S->col, states,
                               b.mark, b.w, kSelectedState);
            hipLaunchKernelGGL(cljp_dist2_kernel, rows_grid(n), dim3(256), 0, s, n, S->rowptr, S->col, b.tptr,
                               b.tidx, states, b.mark, b.w, kSelectedState);
        }
        RAMD_HIP(hipGetLastError());
    }
    for (int round = 0;; round++) {
        RAMD_CHECK(round <= n + 1, RAMD_E_NUMERIC, "C/F splitting did not finish");
        RAMD_HIP(hipMemsetAsync(b.rem, 0, sizeof(int), s));
        hipLaunchKernelGGL(pmis_select_kernel, rows_grid(n), dim3(256), 0, s, n, S->rowptr, S->col, b.tptr, b.tidx,
                           b.w, states);
        if (cljp) {
            hipLaunchKernelGGL(cljp_row_weights_kernel, rows_grid(n), dim3(256), 0, s, n, S->rowptr, S->col, states,
                               b.mark, b.w, kNewSelectionState);
            hipLaunchKernelGGL(cljp_dist2_kernel, rows_grid(n), dim3(256), 0, s, n, S->rowptr, S->col, b.tptr,
                               b.tidx, states, b.mark, b.w, kNewSelectionState);
        } else {
            hipLaunchKernelGGL(pmis_mark_kernel, rows_grid(n), dim3(256), 0, s, n, b.tptr, b.tidx, states);
        }
        hipLaunchKernelGGL(pmis_update_kernel, rows_grid(n), dim3(256), 0, s, n, states, b.w, b.rem);
        RAMD_HIP(hipGetLastError());
        int rem = 0;
        RAMD_HIP(hipMemcpyAsync(&rem, b.rem, sizeof(int), hipMemcpyDeviceToHost, s));
        RAMD_HIP(hipStreamSynchronize(s));
        if (rem == 0) break;
    }
}

// ---------------------------------------------------------------------------
// Extended interpolation, one wavefront per row.  The row's columns are
// collected in discovery order (strong C neighbours, and the strong C
// neighbours of strong F neighbours, first occurrence kept) in LDS, with an
// LDS hash table from column to position.  The loops over the row's own
// neighbours run in order with the wavefront in step; the inner loops over a
// neighbour's row give its entries to the lanes (one entry each), and what
// must happen in entry order -- appending new columns, the running column
// sum -- is done in lane order (prefix counts, a serial pass over a ballot),
// while the per-entry additions to P's entries go to distinct entries, so
// every entry and the weak sum receive their additions in the host's
// sequence.  A row with more than kExtCols columns sets *overflow (the
// caller then builds this level on the host).
// ---------------------------------------------------------------------------
constexpr int kExtCols = 256;
constexpr int kExtHash = 512;  // load factor <= 1/2
constexpr int kExtWaves = 4;

// LDS written by one lane becomes visible to the wavefront's other lanes
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ int ext_hash(int c) { return (int)(((unsigned)c * 2654435761u) >> 23); }

// position of column c in the row's list, or -1
__device__ __forceinline__ int ext_lookup(const int* hk, const int* hv, int c)
{
    for (int h = ext_hash(c);; h = (h + 1) & (kExtHash - 1)) {
        const int k = hk[h];
        if (k == c) return hv[h];
        if (k < 0) return -1;
    }
}

// c is not in the table; lanes may insert distinct columns at once
__device__ __forceinline__ void ext_insert(int* hk, int* hv, int c, int p)
{
    for (int h = ext_hash(c);; h = (h + 1) & (kExtHash - 1))
        if (atomicCAS(&hk[h], -1, c) == -1) {
            hv[h] = p;
            return;
        }
}

// lanes below this one whose bit is set in m
__device__ __forceinline__ int lanes_below(unsigned long long m)
{
    return (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
}

template <bool FILL>
__global__ __launch_bounds__(64 * kExtWaves) void ext_interp_kernel(
    int n, const int* __restrict__ arp, const int* __restrict__ acol, const double* __restrict__ aval,
    const int* __restrict__ srp, const int* __restrict__ scol, const int* __restrict__ states,
    const int* __restrict__ cmap, double filter, int* __restrict__ cnt, const int* __restrict__ prp,
    int* __restrict__ pcol, double* __restrict__ pval, int* __restrict__ overflow)
{
    __shared__ int lcol[kExtWaves][kExtCols];
    __shared__ double lval[kExtWaves][kExtCols];
    __shared__ int hkey[kExtWaves][kExtHash];
    __shared__ int hpos[kExtWaves][kExtHash];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int i = blockIdx.x * kExtWaves + wv;
    if (i >= n) return;  // the whole wavefront leaves together
    int* lc = lcol[wv];
    double* lv = lval[wv];
    int* hk = hkey[wv];
    int* hv = hpos[wv];
    const int st = states[i];
    if (st != kUnselectedState) {
        if (lane == 0) {
            if (st == kSelectedState) {
                if (FILL) {
                    pcol[prp[i]] = cmap[i];
                    pval[prp[i]] = 1;
                } else {
                    cnt[i] = 1;
                }
            } else if (!FILL) {
                cnt[i] = 0;
            }
        }
        return;
    }
    for (int h = lane; h < kExtHash; h += 64) hk[h] = -1;
    wave_sync();
    int len = 0;
    bool over = false;
    // discovery (host: the first loop over S row i)
    const int s_end = srp[i + 1];
    for (int j = srp[i] + 1; j < s_end && !over; j++) {
        const int c = scol[j];
        const int sc = states[c];
        if (sc == kSelectedState) {
            if (ext_lookup(hk, hv, c) < 0) {
                if (len == kExtCols) {
                    over = true;
                } else {
                    if (lane == 0) {
                        lc[len] = c;
                        ext_insert(hk, hv, c, len);
                    }
                    wave_sync();
                    len++;
                }
            }
        } else if (sc == kUnselectedState) {
            const int e = srp[c + 1];
            for (int b0 = srp[c] + 1; b0 < e; b0 += 64) {
                const int k = b0 + lane;
                const int ck = k < e ? scol[k] : -1;
                const bool cand = ck >= 0 && states[ck] == kSelectedState && ext_lookup(hk, hv, ck) < 0;
                const unsigned long long m = __ballot(cand);
                const int add = __popcll(m);
                if (len + add > kExtCols) {
                    over = true;
                    break;
                }
                if (cand) {
                    const int p = len + lanes_below(m);
                    lc[p] = ck;
                    ext_insert(hk, hv, ck, p);
                }
                wave_sync();
                len += add;
            }
        }
    }
    if (over) {
        if (lane == 0) {
            atomicOr(overflow, 1);
            if (!FILL) cnt[i] = 0;
        }
        return;
    }
    for (int k = lane; k < len; k += 64) lv[k] = 0.0;
    wave_sync();
    // accumulation (host: the loop over A row i).  Row i's entries are
    // taken 64 at a time, one per lane, with everything they need that does
    // not change in the loop (list position, strength -- the column among
    // S row i's -- state, the neighbour's row bounds and diagonal) loaded at
    // once; the entries are then visited in order.
    int start = arp[i];
    const int end = arp[i + 1];
    double weak_sum = aval[start++];
    for (int c0 = start; c0 < end; c0 += 64) {
        const int jl = c0 + lane;
        const bool inj = jl < end;
        const int cj = inj ? acol[jl] : -1;
        const double vj = inj ? aval[jl] : 0.0;
        const int idxj = inj ? ext_lookup(hk, hv, cj) : -1;
        bool strong = false;
        for (int t0 = srp[i] + 1; t0 < s_end; t0 += 64) {
            const int sc = t0 + lane < s_end ? scol[t0 + lane] : -2;
            const int tn = min(64, s_end - t0);
            for (int t = 0; t < tn; t++) {
                const int st_t = __shfl(sc, t);  // every lane takes part in the shuffle
                strong = strong || st_t == cj;
            }
        }
        const int stj = inj ? states[cj] : kNoNeighborsState;
        const bool fnb = inj && idxj < 0 && strong && stj == kUnselectedState;
        int skj = 0, ekj = 0;
        double dgj = 0.0;
        if (fnb) {
            skj = arp[cj];
            ekj = arp[cj + 1];
            dgj = aval[skj];
        }
        const int nj = min(64, end - c0);
        for (int t = 0; t < nj; t++) {
            const double v = __shfl(vj, t);
            const int idx = __shfl(idxj, t);
            if (idx >= 0) {
                if (lane == 0) lv[idx] += v;
            } else if (__shfl((int)strong, t)) {
                if (!__shfl((int)fnb, t)) continue;  // a strong neighbour that is not F
                const int sk = __shfl(skj, t), ek = __shfl(ekj, t);
                const int sign = __shfl(dgj, t) > 0 ? 1 : -1;
                // entries of the neighbour's row that count: opposite sign to
                // its diagonal, column i or a column of the list; their sum in
                // entry order
                if (ek - sk - 1 <= 64) {
                    const int k = sk + 1 + lane;
                    const bool in = k < ek;
                    const double vk = in ? aval[k] : 0.0;
                    const int ck = in ? acol[k] : -1;
                    const bool neg = in && vk * sign < 0;
                    const int pk = (neg && ck != i) ? ext_lookup(hk, hv, ck) : -1;
                    double col_sum = 0;
                    for (unsigned long long m = __ballot(neg && (ck == i || pk >= 0)); m; m &= m - 1)
                        col_sum += __shfl(vk, __ffsll((long long)m) - 1);
                    if (fabs(col_sum) < kSetupZeroTol) {
                        weak_sum += v;
                    } else {
                        col_sum = v / col_sum;
                        // distinct columns: each entry of the list gets at most one addition here
                        if (pk >= 0) lv[pk] += (col_sum * vk);
                        const unsigned long long mi = __ballot(neg && ck == i);
                        if (mi) weak_sum += (col_sum * __shfl(vk, __ffsll((long long)mi) - 1));
                    }
                } else {
                    double col_sum = 0;
                    for (int b0 = sk + 1; b0 < ek; b0 += 64) {
                        const int k = b0 + lane;
                        const bool in = k < ek;
                        const double vk = in ? aval[k] : 0.0;
                        const int ck = in ? acol[k] : -1;
                        const bool use = in && vk * sign < 0 && (ck == i || ext_lookup(hk, hv, ck) >= 0);
                        for (unsigned long long m = __ballot(use); m; m &= m - 1)
                            col_sum += __shfl(vk, __ffsll((long long)m) - 1);
                    }
                    if (fabs(col_sum) < kSetupZeroTol) {
                        weak_sum += v;
                    } else {
                        col_sum = v / col_sum;
                        for (int b0 = sk + 1; b0 < ek; b0 += 64) {
                            const int k = b0 + lane;
                            const bool in = k < ek;
                            const double vk = in ? aval[k] : 0.0;
                            const int ck = in ? acol[k] : -1;
                            const bool neg = in && vk * sign < 0;
                            const int pk = (neg && ck != i) ? ext_lookup(hk, hv, ck) : -1;
                            if (pk >= 0) lv[pk] += (col_sum * vk);
                            const unsigned long long mi = __ballot(neg && ck == i);
                            if (mi) weak_sum += (col_sum * __shfl(vk, __ffsll((long long)mi) - 1));
                        }
                    }
                }
            } else if (__shfl(stj, t) != kNoNeighborsState) {
                weak_sum += v;
            }
            wave_sync();
        }
    }
    if (fabs(weak_sum) > kSetupZeroTol)
        for (int k = lane; k < len; k += 64) lv[k] /= -weak_sum;
    wave_sync();
    // filter (host: filter_rows), one lane, entries in order
    if (lane == 0) {
        int kept = len;
        if (filter > kSetupZeroTol && filter <= 1) {
            double row_max = 0;
            for (int k = 0; k < len; k++) {
                const double a = fabs(lv[k]);
                if (a > row_max) row_max = a;
            }
            row_max *= filter;
            double row_sum = 0, remain_sum = 0;
            kept = 0;
            for (int k = 0; k < len; k++) {
                const double v = lv[k];
                row_sum += v;
                if (fabs(v) >= row_max) {
                    lc[kept] = lc[k];
                    lv[kept] = v;
                    kept++;
                    remain_sum += v;
                }
            }
            if (fabs(remain_sum) > kSetupZeroTol && fabs(row_sum - remain_sum) > kSetupZeroTol) {
                const double row_scale = row_sum / remain_sum;
                for (int k = 0; k < kept; k++) lv[k] *= row_scale;
            }
        }
        if (FILL) {
            const int o = prp[i];
            for (int k = 0; k < kept; k++) {
                pcol[o + k] = cmap[lc[k]];
                pval[o + k] = lv[k];
            }
        } else {
            cnt[i] = kept;
        }
    }
}

// ---------------------------------------------------------------------------
// Modified classical interpolation (ruge_stuben/par_interpolation.cpp:
// 991-1451), one wavefront per row.  The row's columns are its strong C
// neighbours in S order (initial values the strong entries of A); the loop
// over A's row runs with the wavefront in step, a strong F neighbour's row is
// spread over the lanes, its coarse sum is formed in entry order by a serial
// pass over a ballot, and the additions go to distinct entries -- the host's
// sequence for every entry.  A row with more than kExtCols strong C
// neighbours sets *overflow.
// ---------------------------------------------------------------------------
template <bool FILL>
__global__ __launch_bounds__(64 * kExtWaves) void modc_interp_kernel(
    int n, const int* __restrict__ arp, const int* __restrict__ acol, const double* __restrict__ aval,
    const int* __restrict__ srp, const int* __restrict__ scol, const double* __restrict__ sval,
    const int* __restrict__ states, const int* __restrict__ cmap, int* __restrict__ cnt, const int* __restrict__ prp,
    int* __restrict__ pcol, double* __restrict__ pval, int* __restrict__ overflow)
{
    __shared__ int lcol[kExtWaves][kExtCols];
    __shared__ double lval[kExtWaves][kExtCols];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int i = blockIdx.x * kExtWaves + wv;
    if (i >= n) return;  // the whole wavefront leaves together
    int* lc = lcol[wv];
    double* lv = lval[wv];
    if (states[i] == kSelectedState) {
        if (lane == 0) {
            if (FILL) {
                pcol[prp[i]] = cmap[i];
                pval[prp[i]] = 1;
            } else {
                cnt[i] = 1;
            }
        }
        return;
    }
    // strong C neighbours in S order
    const int s_end = srp[i + 1];
    int len = 0;
    for (int b0 = srp[i] + 1; b0 < s_end; b0 += 64) {
        const int j = b0 + lane;
        const int c = j < s_end ? scol[j] : -1;
        const bool isc = c >= 0 && states[c] == kSelectedState;
        const unsigned long long m = __ballot(isc);
        if (len + __popcll(m) > kExtCols) {
            if (lane == 0) {
                atomicOr(overflow, 1);
                if (!FILL) cnt[i] = 0;
            }
            return;
        }
        if (isc) {
            const int p = len + lanes_below(m);
            lc[p] = c;
            lv[p] = sval[j];
        }
        len += __popcll(m);
    }
    wave_sync();
    // position of column c in the list, or -1 (a scan: the lists are short)
    auto pos_of = [&](int c) -> int {
        for (int t = 0; t < len; t++)
            if (lc[t] == c) return t;
        return -1;
    };
    int start = arp[i];
    const int end = arp[i + 1];
    int ctr = srp[i] + 1;
    double weak_sum = aval[start++];
    for (int j = start; j < end; j++) {
        const int c = acol[j];
        const double v = aval[j];
        if (ctr < s_end && scol[ctr] == c) {
            ctr++;
            if (states[c] == kSelectedState) continue;
            const int sk = arp[c], ek = arp[c + 1];
            const int sign = aval[sk] > 0 ? 1 : -1;
            double coarse_sum = 0;
            for (int b0 = sk + 1; b0 < ek; b0 += 64) {
                const int k = b0 + lane;
                const int ck = k < ek ? acol[k] : -1;
                const double vk = k < ek ? aval[k] : 0.0;
                const bool use = ck >= 0 && states[ck] == kSelectedState && vk * sign < 0 && pos_of(ck) >= 0;
                for (unsigned long long m = __ballot(use); m; m &= m - 1)
                    coarse_sum += __shfl(vk, __ffsll((long long)m) - 1);
            }
            if (fabs(coarse_sum) < kSetupZeroTol) weak_sum += v;
            else coarse_sum = v / coarse_sum;
            if (coarse_sum != 0.0) {
                for (int b0 = sk + 1; b0 < ek; b0 += 64) {
                    const int k = b0 + lane;
                    const int ck = k < ek ? acol[k] : -1;
                    const double vk = k < ek ? aval[k] : 0.0;
                    const int pk = (ck >= 0 && states[ck] == kSelectedState && vk * sign < 0) ? pos_of(ck) : -1;
                    if (pk >= 0) lv[pk] += (coarse_sum * vk);  // distinct columns: distinct entries
                }
            }
        } else if (states[c] != kNoNeighborsState) {
            weak_sum += v;
        }
        wave_sync();
    }
    for (int k = lane; k < len; k += 64) {
        const double pv = lv[k] / -weak_sum;
        if (FILL) {
            pcol[prp[i] + k] = cmap[lc[k]];
            pval[prp[i] + k] = pv;
        }
    }
    if (!FILL && lane == 0) cnt[i] = len;
}

// ---------------------------------------------------------------------------
// Direct interpolation (ruge_stuben/par_interpolation.cpp:1453-1754), one
// thread per row: the strong C neighbours' entries of A (found by walking A's
// row alongside S's), scaled by the ratios of all to strong negative and
// positive sums, in S order.  Every row with a non-zero state (C points and
// rows without neighbours) owns a coarse column, as there.
// ---------------------------------------------------------------------------
template <bool FILL>
__global__ void direct_interp_kernel(int n, const int* __restrict__ arp, const int* __restrict__ acol,
                                     const double* __restrict__ aval, const int* __restrict__ srp,
                                     const int* __restrict__ scol, const int* __restrict__ states,
                                     const int* __restrict__ cmap, int* __restrict__ cnt, const int* __restrict__ prp,
                                     int* __restrict__ pcol, double* __restrict__ pval)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (states[i] == kSelectedState) {
        if (FILL) {
            pcol[prp[i]] = cmap[i];
            pval[prp[i]] = 1;
        } else {
            cnt[i] = 1;
        }
        return;
    }
    int s = srp[i];
    const int e = srp[i + 1];
    if (s < e && scol[s] == i) s++;
    const int as0 = arp[i], ae = arp[i + 1];
    // A's value at column c of row i (S's columns appear in A's row in order;
    // a column missing from A's row reads as 0 instead of running past it)
    int actr = as0;
    auto a_at = [&](int c) {
        while (actr < ae && acol[actr] != c) actr++;
        return actr < ae ? aval[actr] : 0.0;
    };
    double sum_strong_pos = 0, sum_strong_neg = 0, sum_all_pos = 0, sum_all_neg = 0;
    int count = 0;
    for (int j = s; j < e; j++) {
        const int c = scol[j];
        const double v = a_at(c);
        if (states[c] == kSelectedState) {
            if (v < 0) sum_strong_neg += v;
            else sum_strong_pos += v;
            count++;
        }
    }
    if (!FILL) {
        cnt[i] = count;
        return;
    }
    double diag = aval[as0];
    for (int j = as0 + 1; j < ae; j++) {
        const double v = aval[j];
        if (v < 0) sum_all_neg += v;
        else sum_all_pos += v;
    }
    const double alpha = sum_all_neg / sum_strong_neg;
    double beta;
    if (sum_strong_pos == 0) {
        diag += sum_all_pos;
        beta = 0;
    } else {
        beta = sum_all_pos / sum_strong_pos;
    }
    const double neg_coeff = -alpha / diag;
    const double pos_coeff = -beta / diag;
    int o = prp[i];
    actr = as0;
    for (int j = s; j < e; j++) {
        const int c = scol[j];
        const double v = a_at(c);
        if (states[c] == kSelectedState) {
            pcol[o] = cmap[c];
            pval[o] = v < 0 ? neg_coeff * v : pos_coeff * v;
            o++;
        }
    }
}

ramd_csr_s* gpu_direct_interpolation(const ramd_csr_s* A, const ramd_csr_s* S, const int* states_d,
                                     const std::vector<int>& states_h)
{
    ramd_ctx ctx = A->ctx;
    hipStream_t s = ctx->stream;
    const int n = A->n_rows;
    std::vector<int> cmap((size_t)n, -1);
    int nc = 0;
    for (int i = 0; i < n; i++)
        if (states_h[i] != 0) cmap[i] = nc++;
    struct Bufs {
        int *cmap = nullptr, *cnt = nullptr;
        ~Bufs()
        {
            dfree(cmap);
            dfree(cnt);
        }
    } b;
    b.cmap = dalloc<int>((size_t)std::max(n, 1));
    b.cnt = dalloc<int>((size_t)std::max(n, 1));
    RAMD_HIP(hipMemcpyAsync(b.cmap, cmap.data(), sizeof(int) * n, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL((direct_interp_kernel<false>), rows_grid(n), dim3(256), 0, s, n, A->rowptr, A->col, A->val,
                       S->rowptr, S->col, states_d, b.cmap, b.cnt, nullptr, nullptr, nullptr);
    RAMD_HIP(hipGetLastError());
    std::vector<int> rp = host_rowptr(ctx, b.cnt, n);
    const int64_t nnz = rp[n];
    int* d_rp = dalloc<int>((size_t)n + 1);
    int* d_col = dalloc<int>((size_t)std::max<int64_t>(nnz, 1));
    double* d_val = dalloc<double>((size_t)std::max<int64_t>(nnz, 1));
    RAMD_HIP(hipMemcpyAsync(d_rp, rp.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL((direct_interp_kernel<true>), rows_grid(n), dim3(256), 0, s, n, A->rowptr, A->col, A->val,
                       S->rowptr, S->col, states_d, b.cmap, nullptr, d_rp, d_col, d_val);
    RAMD_HIP(hipGetLastError());
    RAMD_HIP(hipStreamSynchronize(s));
    return csr_wrap(ctx, n, nc, nnz, d_rp, d_col, d_val, rp.data());
}

// P on the device (extended with filter, or modified classical), or nullptr
// when a row needs more than kExtCols columns
static ramd_csr_s* gpu_interpolation(const ramd_csr_s* A, const ramd_csr_s* S, const int* states_d,
                                     const std::vector<int>& states_h, double filter, bool modc)
{
    ramd_ctx ctx = A->ctx;
    hipStream_t s = ctx->stream;
    const int n = A->n_rows;
    std::vector<int> cmap((size_t)n, -1);
    int nc = 0;
    for (int i = 0; i < n; i++)
        if (states_h[i] == kSelectedState) cmap[i] = nc++;
    struct Bufs {
        int *cmap = nullptr, *cnt = nullptr, *over = nullptr;
        ~Bufs()
        {
            dfree(cmap);
            dfree(cnt);
            dfree(over);
        }
    } b;
    b.cmap = dalloc<int>((size_t)std::max(n, 1));
    b.cnt = dalloc<int>((size_t)std::max(n, 1));
    b.over = dalloc<int>(1);
    RAMD_HIP(hipMemcpyAsync(b.cmap, cmap.data(), sizeof(int) * n, hipMemcpyHostToDevice, s));
    RAMD_HIP(hipMemsetAsync(b.over, 0, sizeof(int), s));
    const dim3 grid((unsigned)std::max(1, (n + kExtWaves - 1) / kExtWaves)), block(64 * kExtWaves);
    // count pass (cnt) or fill pass (prp, pcol, pval)
    auto launch = [&](int* cnt, const int* prp, int* pcol, double* pval) {
        if (modc) {
            if (cnt)
                hipLaunchKernelGGL((modc_interp_kernel<false>), grid, block, 0, s, n, A->rowptr, A->col, A->val,
                                   S->rowptr, S->col, S->val, states_d, b.cmap, cnt, prp, pcol, pval, b.over);
            else
                hipLaunchKernelGGL((modc_interp_kernel<true>), grid, block, 0, s, n, A->rowptr, A->col, A->val,
                                   S->rowptr, S->col, S->val, states_d, b.cmap, cnt, prp, pcol, pval, b.over);
        } else {
            if (cnt)
                hipLaunchKernelGGL((ext_interp_kernel<false>), grid, block, 0, s, n, A->rowptr, A->col, A->val,
                                   S->rowptr, S->col, states_d, b.cmap, filter, cnt, prp, pcol, pval, b.over);
            else
                hipLaunchKernelGGL((ext_interp_kernel<true>), grid, block, 0, s, n, A->rowptr, A->col, A->val,
                                   S->rowptr, S->col, states_d, b.cmap, filter, cnt, prp, pcol, pval, b.over);
        }
        RAMD_HIP(hipGetLastError());
    };
    launch(b.cnt, nullptr, nullptr, nullptr);
    int over = 0;
    RAMD_HIP(hipMemcpyAsync(&over, b.over, sizeof(int), hipMemcpyDeviceToHost, s));
    std::vector<int> rp = host_rowptr(ctx, b.cnt, n);  // synchronizes
    if (over) return nullptr;
    const int64_t nnz = rp[n];
    int* d_rp = dalloc<int>((size_t)n + 1);
    int* d_col = dalloc<int>((size_t)std::max<int64_t>(nnz, 1));
    double* d_val = dalloc<double>((size_t)std::max<int64_t>(nnz, 1));
    RAMD_HIP(hipMemcpyAsync(d_rp, rp.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice, s));
    launch(nullptr, d_rp, d_col, d_val);
    RAMD_HIP(hipStreamSynchronize(s));
    return csr_wrap(ctx, n, nc, nnz, d_rp, d_col, d_val, rp.data());
}

ramd_csr_s* gpu_extended_interpolation(const ramd_csr_s* A, const ramd_csr_s* S, const int* states_d,
                                       const std::vector<int>& states_h, double filter)
{
    return gpu_interpolation(A, S, states_d, states_h, filter, false);
}

ramd_csr_s* gpu_mod_classical_interpolation(const ramd_csr_s* A, const ramd_csr_s* S, const int* states_d,
                                            const std::vector<int>& states_h)
{
    return gpu_interpolation(A, S, states_d, states_h, 0.0, true);
}


// ---------------------------------------------------------------------------
// Smoothed aggregation on the device: MIS-2 (aggregation/par_mis.cpp:
// 204-645), aggregation (aggregation/par_aggregate.cpp:5-185) and the two
// row passes of the Jacobi prolongation (the scaled operator and the
// subtraction P - Q, aggregation/par_prolongation.cpp:7-103 with
// util/linalg/par_add.cpp:184-310), one thread per row, with the host
// restatement's results.
//
// MIS-2 states: -1 unassigned and 4 tentatively selected (both still in the
// loop), 3 / 2 this iteration's selection / unselection, 1 / 0 settled.  The
// reference's loops over the remaining vertices only ask whether a vertex is
// still in the loop, whether it was just selected, and for maxima over sets,
// so the order in which they visit vertices does not matter.  Its marking of
// the neighbours of new selections keeps two quirks that make the marks
// persistent (:505-593): a vertex marked in an iteration is unmarked at the
// end of it the first time only (the list reset writes next[w] = 0, which
// is not "unassigned"), so from its second marking on it stays marked.
// mark[] holds those bits: 1 marked (C[w] == Selected), 2 marked before,
// 4 marked in this iteration.
// ---------------------------------------------------------------------------
constexpr int kTmpSelectionState = 4;

__global__ void set_int_kernel(int* __restrict__ p, int n, int v)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__device__ __forceinline__ bool mis2_in_loop(int s) { return s == kUnassignedState || s == kTmpSelectionState; }

// tentative selection (:378-415): a vertex in the loop whose neighbours in S
// with smaller random values have all left the loop.  A vertex in the loop
// reads as in the loop whether or not this launch has marked it yet (-1 and
// 4 both count), so the concurrent marks change no decision.
__global__ void mis2_tentative_kernel(int n, const int* __restrict__ srp, const int* __restrict__ scol,
                                      const double* __restrict__ r, int* states, int* remaining)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) *remaining = 0;
    if (i >= n || !mis2_in_loop(states[i])) return;
    const double ri = r[i];
    for (int j = srp[i]; j < srp[i + 1]; j++) {
        const int c = scol[j];
        if (ri > r[c]) {
            const int t = states[c];
            if (t == kUnassignedState || t > kSelectedState) return;
        }
    }
    states[i] = kTmpSelectionState;
}

// largest random value of a tentative vertex in each row of S (:420-446)
__global__ void mis2_row_max_kernel(int n, const int* __restrict__ srp, const int* __restrict__ scol,
                                    const double* __restrict__ r, const int* __restrict__ states,
                                    double* __restrict__ row_max)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double m = -1.0;
    for (int j = srp[i]; j < srp[i + 1]; j++) {
        const int u = scol[j];
        if (states[u] > kSelectedState && r[u] > m) m = r[u];
    }
    row_max[i] = m;
}

// a tentative vertex with no larger tentative value within distance two is
// selected (:449-473)
__global__ void mis2_select_kernel(int n, const int* __restrict__ srp, const int* __restrict__ scol,
                                   const double* __restrict__ r, const double* __restrict__ row_max,
                                   int* __restrict__ states)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || states[i] != kTmpSelectionState) return;
    const double ri = r[i];
    for (int j = srp[i]; j < srp[i + 1]; j++)
        if (row_max[scol[j]] > ri) return;
    states[i] = kNewSelectionState;
}

// rows of S holding a new selection are marked (:505-526; the reference
// walks the columns of the new selections)
__global__ void mis2_mark_kernel(int n, const int* __restrict__ srp, const int* __restrict__ scol,
                                 const int* __restrict__ states, int* __restrict__ mark)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int j = srp[i]; j < srp[i + 1]; j++)
        if (states[scol[j]] == kNewSelectionState) {
            mark[i] |= 5;
            return;
        }
}

// a vertex in the loop next to a new selection or to a marked vertex is
// unselected (:555-586).  No vertex becomes a new selection here, so the
// concurrent writes change no decision.
__global__ void mis2_unselect_kernel(int n, const int* __restrict__ srp, const int* __restrict__ scol,
                                     const int* __restrict__ mark, int* states)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !mis2_in_loop(states[i])) return;
    for (int j = srp[i]; j < srp[i + 1]; j++) {
        const int c = scol[j];
        if (states[c] == kNewSelectionState || (mark[c] & 1)) {
            states[i] = kNewUnselectionState;
            return;
        }
    }
}

// the list reset (:587-593) and the settling of this iteration's decisions
// (:595-612); counts the vertices left in the loop
__global__ void mis2_settle_kernel(int n, int* __restrict__ states, int* __restrict__ mark,
                                   int* __restrict__ remaining)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    bool left = false;
    if (i < n) {
        int m = mark[i];
        if (m & 4) {
            m &= ~4;
            if (!(m & 2)) m = 2;  // first marking: unmarked, remembered
            mark[i] = m;
        }
        const int s = states[i];
        if (s == kNewSelectionState) states[i] = kSelectedState;
        else if (s == kNewUnselectionState) states[i] = kUnselectedState;
        else left = mis2_in_loop(s);
    }
    const unsigned long long b = __ballot(left);
    if ((threadIdx.x & 63) == 0 && b) atomicAdd(remaining, (int)__popcll(b));
}

void gpu_mis2(const ramd_csr_s* S, const double* weights, int* states)
{
    ramd_ctx ctx = S->ctx;
    hipStream_t s = ctx->stream;
    const int n = S->n_rows;
    if (!n) return;
    struct Bufs {
        double *r = nullptr, *row_max = nullptr;
        int *mark = nullptr, *remaining = nullptr;
        ~Bufs()
        {
            dfree(r);
            dfree(row_max);
            dfree(mark);
            dfree(remaining);
        }
    } b;
    b.r = dalloc<double>((size_t)n);
    b.row_max = dalloc<double>((size_t)n);
    b.mark = dalloc<int>((size_t)n);
    b.remaining = dalloc<int>(1);
    RAMD_HIP(hipMemcpyAsync(b.r, weights, sizeof(double) * n, hipMemcpyHostToDevice, s));
    const dim3 g = rows_grid(n), blk(256);
    hipLaunchKernelGGL(set_int_kernel, g, blk, 0, s, states, n, kUnassignedState);
    hipLaunchKernelGGL(set_int_kernel, g, blk, 0, s, b.mark, n, 0);
    int remaining = 0;
    int64_t iter = 0;
    do {
        if (++iter > (int64_t)n + 2) throw std::runtime_error("mis2: independent-set loop does not terminate");
        hipLaunchKernelGGL(mis2_tentative_kernel, g, blk, 0, s, n, S->rowptr, S->col, b.r, states, b.remaining);
        hipLaunchKernelGGL(mis2_row_max_kernel, g, blk, 0, s, n, S->rowptr, S->col, b.r, states, b.row_max);
        hipLaunchKernelGGL(mis2_select_kernel, g, blk, 0, s, n, S->rowptr, S->col, b.r, b.row_max, states);
        hipLaunchKernelGGL(mis2_mark_kernel, g, blk, 0, s, n, S->rowptr, S->col, states, b.mark);
        hipLaunchKernelGGL(mis2_unselect_kernel, g, blk, 0, s, n, S->rowptr, S->col, b.mark, states);
        hipLaunchKernelGGL(mis2_settle_kernel, g, blk, 0, s, n, states, b.mark, b.remaining);
        RAMD_HIP(hipGetLastError());
        RAMD_HIP(hipMemcpyAsync(&remaining, b.remaining, sizeof(int), hipMemcpyDeviceToHost, s));
        RAMD_HIP(hipStreamSynchronize(s));
    } while (remaining > 0);
}

// aggregate: every row with more than its diagonal in S and selected roots
// its own aggregate; a row with at most its diagonal is -n (:45-60)
__global__ void agg_init_kernel(int n, const int* __restrict__ srp, const int* __restrict__ states,
                                int* __restrict__ agg, int* __restrict__ n_aggs)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    bool root = false;
    if (i < n) {
        if (srp[i + 1] - srp[i] <= 1) {
            agg[i] = -n;
        } else if (states[i] == kSelectedState) {
            agg[i] = i;
            root = true;
        } else {
            agg[i] = -1;
        }
    }
    const unsigned long long b = __ballot(root);
    if ((threadIdx.x & 63) == 0 && b) atomicAdd(n_aggs, (int)__popcll(b));
}

// pass 1 (:96-126): an unselected row joins the aggregate of its first
// selected neighbour in S that roots one.  Only selected rows are read and
// only unselected rows written.
__global__ void agg_pass1_kernel(int n, const int* __restrict__ srp, const int* __restrict__ scol,
                                 const int* __restrict__ states, int* agg)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || states[i] == kSelectedState) return;
    for (int j = srp[i]; j < srp[i + 1]; j++) {
        const int c = scol[j];
        if (states[c] == kSelectedState && agg[c] >= 0) {
            agg[i] = agg[c];
            return;
        }
    }
}

// pass 2 for row i (:134-172): the aggregate of its strongest aggregated
// neighbour, stored as -(aggregate + 1); with none, -(-n + 1) = n - 1.  The
// neighbour's weight |a_ij| + r_j (r: the random values, zero when absent,
// as the solver calls it) is found by walking A's row alongside S's (both
// sorted with the diagonal first).  Returns false when an S entry is
// missing from A.
__device__ __forceinline__ bool agg_pass2_row(int i, int n, const int* __restrict__ srp,
                                              const int* __restrict__ scol, const int* __restrict__ arp,
                                              const int* __restrict__ acol, const double* __restrict__ aval,
                                              const double* __restrict__ r, const int* agg, int& out, int& best)
{
    int ctr = arp[i];
    const int a_end = arp[i + 1];
    double max_val = 0.0;
    best = -n;
    for (int j = srp[i]; j < srp[i + 1]; j++) {
        const int c = scol[j];
        while (ctr < a_end && acol[ctr] != c) ctr++;
        if (ctr == a_end) return false;
        const double val = fabs(aval[ctr]) + (r ? r[c] : 0.0);
        if (val > max_val && agg[c] >= 0) {
            max_val = val;
            best = agg[c];
        }
    }
    out = -(best + 1);
    return true;
}

// Pass 2 over every row at once, reading pass 1's values.  The reference
// runs it in row order, so a row also sees the pass-2 values of the rows
// before it; those are negative (never "aggregated") unless a row found no
// aggregated neighbour (n - 1, flag 2), in which case the caller reruns the
// pass in row order.  Flag 1: an S entry missing from A.
__global__ void agg_pass2_kernel(int n, const int* __restrict__ srp, const int* __restrict__ scol,
                                 const int* __restrict__ arp, const int* __restrict__ acol,
                                 const double* __restrict__ aval, const double* __restrict__ r,
                                 const int* __restrict__ agg, int* __restrict__ out, int* __restrict__ flags)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (agg[i] != -1) {
        out[i] = agg[i];
        return;
    }
    int v, best;
    if (!agg_pass2_row(i, n, srp, scol, arp, acol, aval, r, agg, v, best)) {
        atomicOr(flags, 1);
        out[i] = -1;
        return;
    }
    out[i] = v;
    if (best == -n) atomicOr(flags, 2);
}

// pass 2 in the reference's row order by one thread.  Slow for large n, and
// only reached when some row finds no aggregated neighbour: after MIS-2 on a
// symmetric S (the solver's default strength) every row left to pass 2 has
// a neighbour that pass 1 joined to a root, so only a nonsymmetric S, an
// explicit zero of A kept in S (weight 0 is never chosen), or states from
// elsewhere get here.
__global__ void agg_pass2_serial_kernel(int n, const int* __restrict__ srp, const int* __restrict__ scol,
                                        const int* __restrict__ arp, const int* __restrict__ acol,
                                        const double* __restrict__ aval, const double* __restrict__ r, int* agg,
                                        int* __restrict__ flags)
{
    if (blockIdx.x || threadIdx.x) return;
    for (int i = 0; i < n; i++) {
        if (agg[i] != -1) continue;
        int v, best;
        if (!agg_pass2_row(i, n, srp, scol, arp, acol, aval, r, agg, v, best)) {
            atomicOr(flags, 1);
            return;
        }
        agg[i] = v;
    }
}

// decoding (:174-183): <= -n none, other negatives -(v + 1)
__global__ void agg_decode_kernel(int n, const int* src, int* agg)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int v = src[i];
    agg[i] = v <= -n ? -1 : (v < 0 ? -(v + 1) : v);
}

int gpu_aggregate(const ramd_csr_s* A, const ramd_csr_s* S, const int* states, int* agg, const double* rand_vals)
{
    ramd_ctx ctx = A->ctx;
    hipStream_t s = ctx->stream;
    const int n = A->n_rows;
    if (!n) return 0;
    struct Bufs {
        int *out = nullptr, *cnt = nullptr;
        double* r = nullptr;
        ~Bufs()
        {
            dfree(out);
            dfree(cnt);
            dfree(r);
        }
    } b;
    b.out = dalloc<int>((size_t)n);
    b.cnt = dalloc<int>(2);  // aggregates, flags
    if (rand_vals) {
        b.r = dalloc<double>((size_t)n);
        RAMD_HIP(hipMemcpyAsync(b.r, rand_vals, sizeof(double) * n, hipMemcpyHostToDevice, s));
    }
    const dim3 g = rows_grid(n), blk(256);
    hipLaunchKernelGGL(set_int_kernel, dim3(1), dim3(64), 0, s, b.cnt, 2, 0);
    hipLaunchKernelGGL(agg_init_kernel, g, blk, 0, s, n, S->rowptr, states, agg, b.cnt);
    hipLaunchKernelGGL(agg_pass1_kernel, g, blk, 0, s, n, S->rowptr, S->col, states, agg);
    hipLaunchKernelGGL(agg_pass2_kernel, g, blk, 0, s, n, S->rowptr, S->col, A->rowptr, A->col, A->val, b.r, agg,
                       b.out, b.cnt + 1);
    RAMD_HIP(hipGetLastError());
    int h[2] = {0, 0};
    RAMD_HIP(hipMemcpyAsync(h, b.cnt, sizeof(h), hipMemcpyDeviceToHost, s));
    RAMD_HIP(hipStreamSynchronize(s));
    if (h[1] & 1) throw std::runtime_error("aggregate: strength entry missing from A");
    if (h[1] & 2) {
        hipLaunchKernelGGL(agg_pass2_serial_kernel, dim3(1), dim3(64), 0, s, n, S->rowptr, S->col, A->rowptr, A->col,
                           A->val, b.r, agg, b.cnt + 1);
        hipLaunchKernelGGL(agg_decode_kernel, g, blk, 0, s, n, agg, agg);
    } else {
        hipLaunchKernelGGL(agg_decode_kernel, g, blk, 0, s, n, b.out, agg);
    }
    RAMD_HIP(hipGetLastError());
    RAMD_HIP(hipStreamSynchronize(s));
    return h[0];
}

// the scaled operator of the Jacobi prolongation: row i of A times
// omega / sum_j |a_ij| (absolute row sum in stored order; zero rows stay zero)
__global__ void sa_scale_kernel(int n, const int* __restrict__ rp, const double* __restrict__ val, double omega,
                                double* __restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int s = rp[i], e = rp[i + 1];
    double row_sum = 0.0;
    for (int j = s; j < e; j++) row_sum += fabs(val[j]);
    double inv = 0.0;
    if (row_sum != 0.0) inv = (1.0 / fabs(row_sum)) * omega;
    for (int j = s; j < e; j++) out[j] = val[j] * inv;
}

ramd_csr_s* gpu_sa_scaled_operator(const ramd_csr_s* A, double omega)
{
    ramd_ctx ctx = A->ctx;
    hipStream_t s = ctx->stream;
    const int n = A->n_rows;
    const int64_t nnz = A->nnz;
    int* d_rp = dalloc<int>((size_t)n + 1);
    int* d_col = dalloc<int>((size_t)std::max<int64_t>(nnz, 1));
    double* d_val = dalloc<double>((size_t)std::max<int64_t>(nnz, 1));
    RAMD_HIP(hipMemcpyAsync(d_rp, A->rowptr, sizeof(int) * ((size_t)n + 1), hipMemcpyDeviceToDevice, s));
    if (nnz) RAMD_HIP(hipMemcpyAsync(d_col, A->col, sizeof(int) * nnz, hipMemcpyDeviceToDevice, s));
    hipLaunchKernelGGL(sa_scale_kernel, rows_grid(n), dim3(256), 0, s, n, A->rowptr, A->val, omega, d_val);
    RAMD_HIP(hipGetLastError());
    return csr_wrap(ctx, n, A->n_cols, nnz, d_rp, d_col, d_val, A->h_rowptr.data());
}

// Row i of P - Q (the reference's subtract: P's row then Q's row negated,
// stably sorted by column, equal columns summed in that order, |v| < 1e-16
// dropped, the diagonal (column i) moved first).  P's rows hold distinct
// columns, the diagonal first when present and the rest ascending (T, or an
// earlier difference); Q's rows hold distinct ascending columns (the sparse
// product), so one merge gives the sorted order.
template <bool FILL>
__global__ void sa_subtract_kernel(int n, const int* __restrict__ prp, const int* __restrict__ pcol,
                                   const double* __restrict__ pval, const int* __restrict__ qrp,
                                   const int* __restrict__ qcol, const double* __restrict__ qval,
                                   int* __restrict__ cnt, const int* __restrict__ crp, int* __restrict__ ccol,
                                   double* __restrict__ cval)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int ps = prp[i];
    const int pe = prp[i + 1];
    int pd = -1;  // P's diagonal entry, merged at its column
    if (ps < pe && pcol[ps] == i) pd = ps++;
    int q = qrp[i];
    const int qe = qrp[i + 1];
    const int o0 = FILL ? crp[i] : 0;
    int o = o0, diag_at = -1;
    while (true) {
        // next P entry in column order
        int pj = -1;
        if (pd >= 0 && (ps >= pe || i < pcol[ps])) pj = pd;
        else if (ps < pe) pj = ps;
        const int pc = pj >= 0 ? pcol[pj] : INT_MAX;
        const int qc = q < qe ? qcol[q] : INT_MAX;
        if (pc == INT_MAX && qc == INT_MAX) break;
        int c;
        double v;
        if (pc == qc) {
            c = pc;
            v = pval[pj] + (-qval[q]);
        } else if (pc < qc) {
            c = pc;
            v = pval[pj];
        } else {
            c = qc;
            v = -qval[q];
        }
        if (pc <= qc) {
            if (pj == pd) pd = -1;
            else ps++;
        }
        if (qc <= pc) q++;
        if (!(fabs(v) < kSetupZeroTol)) {
            if (FILL) {
                if (c == i) diag_at = o;
                ccol[o] = c;
                cval[o] = v;
            }
            o++;
        }
    }
    if (!FILL) {
        cnt[i] = o;
    } else if (diag_at > o0) {
        const double dv = cval[diag_at];
        for (int m = diag_at; m > o0; m--) {
            ccol[m] = ccol[m - 1];
            cval[m] = cval[m - 1];
        }
        ccol[o0] = i;
        cval[o0] = dv;
    }
}

ramd_csr_s* gpu_sa_subtract(const ramd_csr_s* P, const ramd_csr_s* Q)
{
    ramd_ctx ctx = P->ctx;
    hipStream_t s = ctx->stream;
    const int n = P->n_rows;
    int* cnt = dalloc<int>((size_t)std::max(n, 1));
    std::vector<int> rp;
    try {
        hipLaunchKernelGGL((sa_subtract_kernel<false>), rows_grid(n), dim3(256), 0, s, n, P->rowptr, P->col, P->val,
                           Q->rowptr, Q->col, Q->val, cnt, nullptr, nullptr, nullptr);
        RAMD_HIP(hipGetLastError());
        rp = host_rowptr(ctx, cnt, n);
    } catch (...) {
        dfree(cnt);
        throw;
    }
    dfree(cnt);
    const int64_t nnz = rp[n];
    int* d_rp = dalloc<int>((size_t)n + 1);
    int* d_col = dalloc<int>((size_t)std::max<int64_t>(nnz, 1));
    double* d_val = dalloc<double>((size_t)std::max<int64_t>(nnz, 1));
    RAMD_HIP(hipMemcpyAsync(d_rp, rp.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL((sa_subtract_kernel<true>), rows_grid(n), dim3(256), 0, s, n, P->rowptr, P->col, P->val,
                       Q->rowptr, Q->col, Q->val, nullptr, d_rp, d_col, d_val);
    RAMD_HIP(hipGetLastError());
    RAMD_HIP(hipStreamSynchronize(s));
    return csr_wrap(ctx, n, P->n_cols, nnz, d_rp, d_col, d_val, rp.data());
}

// fit_candidates (aggregation/par_candidates.cpp:6-216, one candidate):
// aggregated row i holds (colmap[agg[i]], B[i]) before scaling
__global__ void fit_fill_kernel(int n, const int* __restrict__ agg, const int* __restrict__ colmap,
                                const int* __restrict__ trp, const double* __restrict__ B, int* __restrict__ tcol,
                                double* __restrict__ tval)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int a = agg[i];
    if (a < 0) return;
    const int j = trp[i];
    tcol[j] = colmap[a];
    tval[j] = B[i];
}

// aggregate k's norm: the sum of B_i^2 over its rows in ascending order (the
// transpose's row k), its square root (R, the next level's candidate) and
// the reciprocal that scales the aggregate's entries
__global__ void fit_norm_kernel(int m, const int* __restrict__ ttrp, const double* __restrict__ ttval,
                                double* __restrict__ R, double* __restrict__ scale)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m) return;
    double sum = 0.0;
    for (int j = ttrp[k]; j < ttrp[k + 1]; j++) sum += ttval[j] * ttval[j];
    const double r = sqrt(sum);
    R[k] = r;
    scale[k] = 1.0 / r;
}

__global__ void fit_scale_kernel(int n, const int* __restrict__ trp, const int* __restrict__ tcol,
                                 const double* __restrict__ B, const double* __restrict__ scale,
                                 double* __restrict__ tval)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int j = trp[i];
    if (j < trp[i + 1]) tval[j] = B[i] * scale[tcol[j]];
}

ramd_csr_s* gpu_fit_candidates(ramd_ctx ctx, int n, int n_aggs, const int* agg_d, const std::vector<int>& agg_h,
                               const double* B_d, double* R_d)
{
    hipStream_t s = ctx->stream;
    // the column of each aggregate (roots in increasing order) and T's row
    // pointers, from the host copy of the aggregates
    std::vector<int> colmap((size_t)std::max(n, 1), 0), rp((size_t)n + 1, 0);
    for (int a : agg_h)
        if (a >= 0) colmap[a] = 1;
    int nc = 0;
    for (int i = 0; i < n; i++)
        if (colmap[i]) colmap[i] = nc++;
    if (nc != n_aggs)
        throw std::runtime_error("fit_candidates: an aggregate has no selected root (the reference's aggregation "
                                 "leaves such a point undefined)");
    for (int i = 0; i < n; i++) rp[i + 1] = rp[i] + (agg_h[i] >= 0 ? 1 : 0);
    const int64_t nnz = rp[n];
    struct Bufs {
        int* colmap = nullptr;
        double* scale = nullptr;
        ramd_csr_s* Tb = nullptr;
        ~Bufs()
        {
            dfree(colmap);
            dfree(scale);
            csr_free(Tb);
        }
    } b;
    b.colmap = dalloc<int>((size_t)std::max(n, 1));
    b.scale = dalloc<double>((size_t)std::max(n_aggs, 1));
    RAMD_HIP(hipMemcpyAsync(b.colmap, colmap.data(), sizeof(int) * n, hipMemcpyHostToDevice, s));
    auto make = [&]() {
        int* d_rp = dalloc<int>((size_t)n + 1);
        int* d_col = dalloc<int>((size_t)std::max<int64_t>(nnz, 1));
        double* d_val = dalloc<double>((size_t)std::max<int64_t>(nnz, 1));
        RAMD_HIP(hipMemcpyAsync(d_rp, rp.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice, s));
        ramd_csr_s* M = csr_wrap(ctx, n, n_aggs, nnz, d_rp, d_col, d_val, rp.data());
        hipLaunchKernelGGL(fit_fill_kernel, rows_grid(n), dim3(256), 0, s, n, agg_d, b.colmap, M->rowptr, B_d, M->col,
                           M->val);
        RAMD_HIP(hipGetLastError());
        return M;
    };
    // unscaled entries, transposed (members of each aggregate in ascending
    // row order) for the norms; then the scaled T
    b.Tb = make();
    ramd_csr_s* Tt = csr_transpose(b.Tb);
    hipLaunchKernelGGL(fit_norm_kernel, rows_grid(n_aggs), dim3(256), 0, s, n_aggs, Tt->rowptr, Tt->val, R_d,
                       b.scale);
    RAMD_HIP(hipGetLastError());
    ramd_csr_s* T = make();
    hipLaunchKernelGGL(fit_scale_kernel, rows_grid(n), dim3(256), 0, s, n, T->rowptr, T->col, B_d, b.scale, T->val);
    RAMD_HIP(hipGetLastError());
    RAMD_HIP(hipStreamSynchronize(s));
    return T;
}

// flags a row that is not stored sorted with its diagonal first (the
// diagonal, when present, then ascending distinct columns)
__global__ void diag_first_sorted_kernel(int n, const int* __restrict__ rp, const int* __restrict__ col,
                                         int* __restrict__ bad)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int s = rp[i];
    const int e = rp[i + 1];
    if (s < e && col[s] == i) s++;
    for (int j = s; j < e; j++)
        if (col[j] == i || (j > s && col[j] <= col[j - 1])) {
            atomicOr(bad, 1);
            return;
        }
}

static bool diag_first_sorted(const ramd_csr_s* M)
{
    hipStream_t s = M->ctx->stream;
    int* bad = dalloc<int>(1);
    int h = 0;
    try {
        hipLaunchKernelGGL(set_int_kernel, dim3(1), dim3(64), 0, s, bad, 1, 0);
        if (M->n_rows)
            hipLaunchKernelGGL(diag_first_sorted_kernel, rows_grid(M->n_rows), dim3(256), 0, s, M->n_rows,
                               M->rowptr, M->col, bad);
        RAMD_HIP(hipGetLastError());
        RAMD_HIP(hipMemcpyAsync(&h, bad, sizeof(int), hipMemcpyDeviceToHost, s));
        RAMD_HIP(hipStreamSynchronize(s));
    } catch (...) {
        dfree(bad);
        throw;
    }
    dfree(bad);
    return h == 0;
}

}  // namespace ramd

using namespace ramd;

extern "C" {

int ramd_strength_gpu(ramd_csr A, double theta, ramd_csr* S)
{
    return guarded([&] {
        RAMD_CHECK(A && S, RAMD_E_ARG, "ramd_strength_gpu: NULL argument");
        RAMD_CHECK(A->n_rows == A->n_cols, RAMD_E_ARG, "ramd_strength_gpu: A is not square");
        A->ctx->set_current();
        *S = gpu_classical_strength(A, theta);
    });
}

int ramd_symmetric_strength_gpu(ramd_csr A, double theta, ramd_csr* S)
{
    return guarded([&] {
        RAMD_CHECK(A && S, RAMD_E_ARG, "ramd_symmetric_strength_gpu: NULL argument");
        RAMD_CHECK(A->n_rows == A->n_cols, RAMD_E_ARG, "ramd_symmetric_strength_gpu: A is not square");
        A->ctx->set_current();
        *S = gpu_symmetric_strength(A, theta);
    });
}

int ramd_split_pmis_gpu(ramd_csr S, const double* weights, int32_t* states)
{
    return guarded([&] {
        RAMD_CHECK(S && states, RAMD_E_ARG, "ramd_split_pmis_gpu: NULL argument");
        RAMD_CHECK(S->n_rows == S->n_cols, RAMD_E_ARG, "ramd_split_pmis_gpu: S is not square");
        ramd_ctx ctx = S->ctx;
        ctx->set_current();
        const int n = S->n_rows;
        std::vector<double> w;
        if (weights) w.assign(weights, weights + n);
        else w = host::reference_weights(n, 0);
        int* d = dalloc<int>((size_t)std::max(n, 1));
        try {
            gpu_split_pmis(S, w.data(), d);
            if (n) RAMD_HIP(hipMemcpyAsync(states, d, sizeof(int) * n, hipMemcpyDeviceToHost, ctx->stream));
            RAMD_HIP(hipStreamSynchronize(ctx->stream));
        } catch (...) {
            dfree(d);
            throw;
        }
        dfree(d);
    });
}

int ramd_split_hmis_gpu(ramd_csr S, const double* weights, int32_t* states)
{
    return guarded([&] {
        RAMD_CHECK(S && states, RAMD_E_ARG, "ramd_split_hmis_gpu: NULL argument");
        RAMD_CHECK(S->n_rows == S->n_cols, RAMD_E_ARG, "ramd_split_hmis_gpu: S is not square");
        ramd_ctx ctx = S->ctx;
        ctx->set_current();
        const int n = S->n_rows;
        std::vector<double> w;
        if (weights) w.assign(weights, weights + n);
        else w = host::reference_weights(n, 0);
        // the Ruge-Stuben first pass on the host copy of S
        host::Csr Sh;
        Sh.n_rows = Sh.n_cols = n;
        Sh.rowptr = S->h_rowptr;
        Sh.col.resize((size_t)S->nnz);
        Sh.val.resize((size_t)S->nnz);
        if (S->nnz) {
            RAMD_HIP(hipMemcpyAsync(Sh.col.data(), S->col, sizeof(int) * S->nnz, hipMemcpyDeviceToHost, ctx->stream));
            RAMD_HIP(hipMemcpyAsync(Sh.val.data(), S->val, sizeof(double) * S->nnz, hipMemcpyDeviceToHost,
                                    ctx->stream));
        }
        RAMD_HIP(hipStreamSynchronize(ctx->stream));
        std::vector<int> prior;
        host::hmis_first_pass(Sh, prior);
        int* d = dalloc<int>((size_t)std::max(n, 1));
        try {
            gpu_split_pmis(S, w.data(), d, prior.data());
            if (n) RAMD_HIP(hipMemcpyAsync(states, d, sizeof(int) * n, hipMemcpyDeviceToHost, ctx->stream));
            RAMD_HIP(hipStreamSynchronize(ctx->stream));
        } catch (...) {
            dfree(d);
            throw;
        }
        dfree(d);
    });
}

int ramd_interp_extended_gpu(ramd_csr A, ramd_csr S, const int32_t* states, double filter, ramd_csr* P)
{
    return guarded([&] {
        RAMD_CHECK(A && S && states && P, RAMD_E_ARG, "ramd_interp_extended_gpu: NULL argument");
        RAMD_CHECK(A->n_rows == S->n_rows && A->n_rows == A->n_cols, RAMD_E_ARG,
                   "ramd_interp_extended_gpu: size mismatch");
        ramd_ctx ctx = A->ctx;
        ctx->set_current();
        const int n = A->n_rows;
        std::vector<int> st(states, states + n);
        int* d = dalloc<int>((size_t)std::max(n, 1));
        ramd_csr_s* out = nullptr;
        try {
            RAMD_HIP(hipMemcpyAsync(d, st.data(), sizeof(int) * n, hipMemcpyHostToDevice, ctx->stream));
            out = gpu_extended_interpolation(A, S, d, st, filter);
        } catch (...) {
            dfree(d);
            throw;
        }
        dfree(d);
        RAMD_CHECK(out, RAMD_E_ARG, "ramd_interp_extended_gpu: a row needs more than %d columns", kExtCols);
        *P = out;
    });
}

int ramd_interp_direct_gpu(ramd_csr A, ramd_csr S, const int32_t* states, ramd_csr* P)
{
    return guarded([&] {
        RAMD_CHECK(A && S && states && P, RAMD_E_ARG, "ramd_interp_direct_gpu: NULL argument");
        RAMD_CHECK(A->n_rows == S->n_rows && A->n_rows == A->n_cols, RAMD_E_ARG,
                   "ramd_interp_direct_gpu: size mismatch");
        ramd_ctx ctx = A->ctx;
        ctx->set_current();
        const int n = A->n_rows;
        std::vector<int> st(states, states + n);
        int* d = dalloc<int>((size_t)std::max(n, 1));
        try {
            RAMD_HIP(hipMemcpyAsync(d, st.data(), sizeof(int) * n, hipMemcpyHostToDevice, ctx->stream));
            *P = gpu_direct_interpolation(A, S, d, st);
        } catch (...) {
            dfree(d);
            throw;
        }
        dfree(d);
    });
}

int ramd_split_cljp_gpu(ramd_csr S, const double* weights, int32_t* states)
{
    return guarded([&] {
        RAMD_CHECK(S && states, RAMD_E_ARG, "ramd_split_cljp_gpu: NULL argument");
        RAMD_CHECK(S->n_rows == S->n_cols, RAMD_E_ARG, "ramd_split_cljp_gpu: S is not square");
        ramd_ctx ctx = S->ctx;
        ctx->set_current();
        const int n = S->n_rows;
        std::vector<double> w;
        if (weights) w.assign(weights, weights + n);
        else w = host::reference_weights(n, 0);
        int* d = dalloc<int>((size_t)std::max(n, 1));
        try {
            gpu_split_cljp(S, w.data(), d);
            if (n) RAMD_HIP(hipMemcpyAsync(states, d, sizeof(int) * n, hipMemcpyDeviceToHost, ctx->stream));
            RAMD_HIP(hipStreamSynchronize(ctx->stream));
        } catch (...) {
            dfree(d);
            throw;
        }
        dfree(d);
    });
}

int ramd_interp_mod_classical_gpu(ramd_csr A, ramd_csr S, const int32_t* states, ramd_csr* P)
{
    return guarded([&] {
        RAMD_CHECK(A && S && states && P, RAMD_E_ARG, "ramd_interp_mod_classical_gpu: NULL argument");
        RAMD_CHECK(A->n_rows == S->n_rows && A->n_rows == A->n_cols, RAMD_E_ARG,
                   "ramd_interp_mod_classical_gpu: size mismatch");
        ramd_ctx ctx = A->ctx;
        ctx->set_current();
        const int n = A->n_rows;
        std::vector<int> st(states, states + n);
        int* d = dalloc<int>((size_t)std::max(n, 1));
        ramd_csr_s* out = nullptr;
        try {
            RAMD_HIP(hipMemcpyAsync(d, st.data(), sizeof(int) * n, hipMemcpyHostToDevice, ctx->stream));
            out = gpu_mod_classical_interpolation(A, S, d, st);
        } catch (...) {
            dfree(d);
            throw;
        }
        dfree(d);
        RAMD_CHECK(out, RAMD_E_ARG, "ramd_interp_mod_classical_gpu: a row has more than %d strong C neighbours",
                   kExtCols);
        *P = out;
    });
}

int ramd_mis2_gpu(ramd_csr S, const double* weights, int32_t* states)
{
    return guarded([&] {
        RAMD_CHECK(S && states, RAMD_E_ARG, "ramd_mis2_gpu: NULL argument");
        RAMD_CHECK(S->n_rows == S->n_cols, RAMD_E_ARG, "ramd_mis2_gpu: S is not square");
        ramd_ctx ctx = S->ctx;
        ctx->set_current();
        const int n = S->n_rows;
        std::vector<double> w;
        if (weights) w.assign(weights, weights + n);
        else w = host::reference_weights(n, 0);
        int* d = dalloc<int>((size_t)std::max(n, 1));
        try {
            gpu_mis2(S, w.data(), d);
            if (n) RAMD_HIP(hipMemcpyAsync(states, d, sizeof(int) * n, hipMemcpyDeviceToHost, ctx->stream));
            RAMD_HIP(hipStreamSynchronize(ctx->stream));
        } catch (...) {
            dfree(d);
            throw;
        }
        dfree(d);
    });
}

int ramd_fit_candidates_gpu(ramd_ctx ctx, int32_t n, const int32_t* aggregates, int32_t n_aggs, const double* B,
                            double* R, ramd_csr* T)
{
    return guarded([&] {
        RAMD_CHECK(ctx && aggregates && B && R && T && n >= 0 && n_aggs >= 0 && n_aggs <= n, RAMD_E_ARG,
                   "ramd_fit_candidates_gpu: bad argument");
        for (int i = 0; i < n; i++)
            RAMD_CHECK(aggregates[i] >= -1 && aggregates[i] < n, RAMD_E_ARG,
                       "ramd_fit_candidates_gpu: aggregate %d of row %d out of range", aggregates[i], i);
        ctx->set_current();
        std::vector<int> agg(aggregates, aggregates + n);
        const size_t m = (size_t)std::max(n, 1);
        int* agg_d = dalloc<int>(m);
        double* B_d = nullptr;
        double* R_d = nullptr;
        try {
            B_d = dalloc<double>(m);
            R_d = dalloc<double>((size_t)std::max(n_aggs, 1));
            if (n) {
                RAMD_HIP(hipMemcpyAsync(agg_d, agg.data(), sizeof(int) * n, hipMemcpyHostToDevice, ctx->stream));
                RAMD_HIP(hipMemcpyAsync(B_d, B, sizeof(double) * n, hipMemcpyHostToDevice, ctx->stream));
            }
            ramd_csr_s* out = gpu_fit_candidates(ctx, n, n_aggs, agg_d, agg, B_d, R_d);
            if (n_aggs)
                RAMD_HIP(hipMemcpyAsync(R, R_d, sizeof(double) * n_aggs, hipMemcpyDeviceToHost, ctx->stream));
            RAMD_HIP(hipStreamSynchronize(ctx->stream));
            *T = out;
        } catch (...) {
            dfree(agg_d);
            dfree(B_d);
            dfree(R_d);
            throw;
        }
        dfree(agg_d);
        dfree(B_d);
        dfree(R_d);
    });
}

int ramd_aggregate_gpu(ramd_csr A, ramd_csr S, const int32_t* states, const double* rand_vals, int32_t* aggregates,
                       int32_t* n_aggs)
{
    return guarded([&] {
        RAMD_CHECK(A && S && states && aggregates && n_aggs, RAMD_E_ARG, "ramd_aggregate_gpu: NULL argument");
        RAMD_CHECK(A->n_rows == A->n_cols && S->n_rows == A->n_rows && S->n_cols == A->n_cols, RAMD_E_ARG,
                   "ramd_aggregate_gpu: size mismatch");
        ramd_ctx ctx = A->ctx;
        ctx->set_current();
        RAMD_CHECK(diag_first_sorted(A) && diag_first_sorted(S), RAMD_E_ARG,
                   "ramd_aggregate_gpu: A and S must store each row sorted with its diagonal first");
        const int n = A->n_rows;
        int* d = dalloc<int>((size_t)std::max(n, 1) * 2);
        try {
            if (n) RAMD_HIP(hipMemcpyAsync(d, states, sizeof(int) * n, hipMemcpyHostToDevice, ctx->stream));
            *n_aggs = gpu_aggregate(A, S, d, d + std::max(n, 1), rand_vals);
            if (n)
                RAMD_HIP(hipMemcpyAsync(aggregates, d + std::max(n, 1), sizeof(int) * n, hipMemcpyDeviceToHost,
                                        ctx->stream));
            RAMD_HIP(hipStreamSynchronize(ctx->stream));
        } catch (...) {
            dfree(d);
            throw;
        }
        dfree(d);
    });
}

}  // extern "C"

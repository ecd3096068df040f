// Two-phase sparse matrix-matrix product C = A B on the device, with the
// reference's accumulation order and drop rule (util/linalg/matmult.cpp:
// 71-204, util/linalg/add.cpp:39-78, core/matrix.cpp:922-986).
//
// Phase 1 (symbolic): one wavefront per row inserts every product's column
// into a hash set (LDS when it fits, global memory otherwise) and counts the
// distinct columns -> an upper bound of the row length.
// Phase 2 (numeric): one wavefront per row walks the row of A in stored
// order; for each a_ij the lanes spread over row j of B.  Columns inside one
// row of B are distinct, so no two lanes ever update the same accumulator in
// the same step, and every c_ik receives its products in exactly the
// reference's order (a_ij in A's order, then b_jk in B's order), starting
// from 0 -- bit-identical sums without atomics on values.  Entries with
// |c_ik| <= 1e-16 are dropped and the survivors are written in ascending
// column order (their rank among the row's kept columns).
// Phase 3: rows that lost entries are compacted.
#include <algorithm>
#include <numeric>

#include "common.hpp"

namespace ramd {

constexpr int kWave = 64;



__device__ __forceinline__ unsigned hash_col(int c) { return (unsigned)c * 2654435761u; }

__device__ __forceinline__ int lane_id() { return threadIdx.x & (kWave - 1); }

__device__ __forceinline__ int wave_count(bool pred)
{
    return __popcll(__ballot(pred));
}

__device__ __forceinline__ int wave_prefix(bool pred)
{
    const unsigned long long m = __ballot(pred);
    const int lane = lane_id();
    return __popcll(m & ((lane == 0) ? 0ull : (~0ull >> (64 - lane))));
}

// products per row: sum over j in A_i of nnz(B_j)
__global__ void count_products(int m, const int* __restrict__ arp, const int* __restrict__ acol,
                               const int* __restrict__ brp, int64_t* __restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    int64_t s = 0;
    for (int j = arp[i]; j < arp[i + 1]; j++) {
        const int c = acol[j];
        s += brp[c + 1] - brp[c];
    }
    out[i] = s;
}

// insert key into open-addressing table (size T, power of two); returns slot
__device__ __forceinline__ int hash_insert(int* keys, int T, int key)
{
    unsigned h = hash_col(key) & (unsigned)(T - 1);
    while (true) {
        int k = keys[h];
        if (k == key) return (int)h;
        if (k == -1) {
            int old = atomicCAS(&keys[h], -1, key);
            if (old == -1 || old == key) return (int)h;
        }
        h = (h + 1) & (unsigned)(T - 1);
    }
}

// ---------------------------------------------------------------------------
// symbolic: rows[] lists the rows of this bin; table size T per wave
// ---------------------------------------------------------------------------
__device__ void symbolic_row(int row, int* keys, int T, const int* __restrict__ arp, const int* __restrict__ acol,
                             const int* __restrict__ brp, const int* __restrict__ bcol, int* __restrict__ counts)
{
    const int lane = lane_id();
    for (int t = lane; t < T; t += kWave) keys[t] = -1;
    __builtin_amdgcn_wave_barrier();
    for (int j = arp[row]; j < arp[row + 1]; j++) {
        const int c = acol[j];
        for (int t = brp[c] + lane; t < brp[c + 1]; t += kWave) hash_insert(keys, T, bcol[t]);
    }
    __builtin_amdgcn_wave_barrier();
    __threadfence_block();
    int cnt = 0;
    for (int t0 = 0; t0 < T; t0 += kWave) {
        const int t = t0 + lane;
        cnt += wave_count(t < T && keys[t] != -1);
    }
    if (lane == 0) counts[row] = cnt;
}

template <int T, int WAVES>
__global__ __launch_bounds__(WAVES * kWave) void symbolic_lds(int nrows, const int* __restrict__ rows,
                                                            const int* __restrict__ arp, const int* __restrict__ acol,
                                                            const int* __restrict__ brp, const int* __restrict__ bcol,
                                                            int* __restrict__ counts)
{
    __shared__ int keys[WAVES][T];
    const int w = threadIdx.x / kWave;
    const int r = blockIdx.x * WAVES + w;
    if (r >= nrows) return;
    symbolic_row(rows[r], keys[w], T, arp, acol, brp, bcol, counts);
}

__global__ __launch_bounds__(kWave) void symbolic_global(int nrows, const int* __restrict__ rows,
                                                         const int64_t* __restrict__ tab_off,
                                                         const int* __restrict__ tab_size, int* __restrict__ pool,
                                                         const int* __restrict__ arp, const int* __restrict__ acol,
                                                         const int* __restrict__ brp, const int* __restrict__ bcol,
                                                         int* __restrict__ counts)
{
    const int r = blockIdx.x;
    if (r >= nrows) return;
    symbolic_row(rows[r], pool + tab_off[r], tab_size[r], arp, acol, brp, bcol, counts);
}

// ---------------------------------------------------------------------------
// numeric
// ---------------------------------------------------------------------------
__device__ void numeric_row(int row, int* keys, double* sums, int* lcol, double* lval, int T,
                            const int* __restrict__ arp, const int* __restrict__ acol, const double* __restrict__ aval,
                            const int* __restrict__ brp, const int* __restrict__ bcol, const double* __restrict__ bval,
                            const int* __restrict__ crp, int* __restrict__ ccol, double* __restrict__ cval,
                            int* __restrict__ clen)
{
    const int lane = lane_id();
    for (int t = lane; t < T; t += kWave) {
        keys[t] = -1;
        sums[t] = 0.0;
    }
    __builtin_amdgcn_wave_barrier();
    for (int j = arp[row]; j < arp[row + 1]; j++) {
        const int c = acol[j];
        const double a = aval[j];
        for (int t = brp[c] + lane; t < brp[c + 1]; t += kWave) {
            const int slot = hash_insert(keys, T, bcol[t]);
            sums[slot] += a * bval[t];
        }
        // every lane's update for this a_ij lands before the next a_ij's
        __builtin_amdgcn_wave_barrier();
        __threadfence_block();
    }
    // gather kept entries (|sum| > 1e-16) into a compact list
    int kept = 0;
    for (int t0 = 0; t0 < T; t0 += kWave) {
        const int t = t0 + lane;
        bool keep = false;
        int k = -1;
        double v = 0.0;
        if (t < T) {
            k = keys[t];
            v = sums[t];
            keep = (k != -1) && (fabs(v) > 1e-16);
        }
        const int p = wave_prefix(keep);
        if (keep) {
            lcol[kept + p] = k;
            lval[kept + p] = v;
        }
        kept += wave_count(keep);
    }
    __builtin_amdgcn_wave_barrier();
    __threadfence_block();
    // rank sort by column, write out in ascending order
    const int base = crp[row];
    for (int e = lane; e < kept; e += kWave) {
        const int ce = lcol[e];
        int rank = 0;
        for (int f = 0; f < kept; f++) rank += (lcol[f] < ce);
        ccol[base + rank] = ce;
        cval[base + rank] = lval[e];
    }
    if (lane == 0) clen[row] = kept;
}

template <int T, int WAVES>
__global__ __launch_bounds__(WAVES * kWave) void numeric_lds(
    int nrows, const int* __restrict__ rows, const int* __restrict__ arp, const int* __restrict__ acol,
    const double* __restrict__ aval, const int* __restrict__ brp, const int* __restrict__ bcol,
    const double* __restrict__ bval, const int* __restrict__ crp, int* __restrict__ ccol,
    double* __restrict__ cval, int* __restrict__ clen)
{
    __shared__ int keys[WAVES][T];
    __shared__ double sums[WAVES][T];
    __shared__ int lcol[WAVES][T / 2];
    __shared__ double lval[WAVES][T / 2];
    const int w = threadIdx.x / kWave;
    const int r = blockIdx.x * WAVES + w;
    if (r >= nrows) return;
    numeric_row(rows[r], keys[w], sums[w], lcol[w], lval[w], T, arp, acol, aval, brp, bcol, bval, crp, ccol, cval,
                clen);
}

__global__ __launch_bounds__(kWave) void numeric_global(
    int nrows, const int* __restrict__ rows, const int64_t* __restrict__ tab_off, const int* __restrict__ tab_size,
    int* __restrict__ kpool, double* __restrict__ spool, int* __restrict__ lcpool, double* __restrict__ lvpool,
    const int* __restrict__ arp, const int* __restrict__ acol, const double* __restrict__ aval,
    const int* __restrict__ brp, const int* __restrict__ bcol, const double* __restrict__ bval,
    const int* __restrict__ crp, int* __restrict__ ccol, double* __restrict__ cval, int* __restrict__ clen)
{
    const int r = blockIdx.x;
    if (r >= nrows) return;
    const int64_t o = tab_off[r];
    numeric_row(rows[r], kpool + o, spool + o, lcpool + o, lvpool + o, tab_size[r], arp, acol, aval, brp, bcol, bval,
                crp, ccol, cval, clen);
}

__global__ void compact_rows(int m, const int* __restrict__ rp_bound, const int* __restrict__ rp_final,
                             const int* __restrict__ col_in, const double* __restrict__ val_in,
                             int* __restrict__ col_out, double* __restrict__ val_out)
{
    const int i = blockIdx.x;
    if (i >= m) return;
    const int s = rp_bound[i], d = rp_final[i], len = rp_final[i + 1] - rp_final[i];
    for (int t = threadIdx.x; t < len; t += blockDim.x) {
        col_out[d + t] = col_in[s + t];
        val_out[d + t] = val_in[s + t];
    }
}

// diagonal moved to the front of each row, others keep their order
__global__ void move_diag_kernel(int m, const int* __restrict__ rp, int* __restrict__ col, double* __restrict__ val)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const int s = rp[i], e = rp[i + 1];
    for (int j = s; j < e; j++) {
        if (col[j] == i) {
            const double t = val[j];
            for (int k = j; k > s; k--) {
                col[k] = col[k - 1];
                val[k] = val[k - 1];
            }
            col[s] = i;
            val[s] = t;
            break;
        }
    }
}

static int next_pow2(int64_t v)
{
    int p = 1;
    while (p < v) p <<= 1;
    return p;
}

// Runs symbolic+numeric for C = A B, returns a new handle.
static ramd_csr_s* spgemm_impl(ramd_csr_s* A, ramd_csr_s* B)
{
    ramd_ctx ctx = A->ctx;
    hipStream_t st = ctx->stream;
    const int m = A->n_rows, n = B->n_cols;
    RAMD_CHECK(A->n_cols == B->n_rows, RAMD_E_ARG, "spgemm: inner dimensions %d and %d differ", A->n_cols, B->n_rows);

    // --- product counts
    int64_t* d_prod = dalloc<int64_t>((size_t)std::max(m, 1));
    if (m) {
        hipLaunchKernelGGL(count_products, dim3((m + 255) / 256), dim3(256), 0, st, m, A->rowptr, A->col, B->rowptr,
                           d_prod);
        RAMD_HIP(hipGetLastError());
    }
    std::vector<int64_t> prod((size_t)m);
    if (m) RAMD_HIP(hipMemcpyAsync(prod.data(), d_prod, sizeof(int64_t) * m, hipMemcpyDeviceToHost, st));
    RAMD_HIP(hipStreamSynchronize(st));
    dfree(d_prod);

    // --- symbolic bins by product count
    static const int kSymT[] = {64, 128, 256, 512, 1024, 2048, 4096, 8192};
    const int nsym = 8;
    std::vector<std::vector<int>> sym_bins(nsym + 1);
    for (int i = 0; i < m; i++) {
        const int64_t p = prod[i];
        int b = 0;
        while (b < nsym && kSymT[b] < p) b++;
        sym_bins[b].push_back(i);
    }
    int* d_counts = dalloc<int>((size_t)std::max(m, 1));
    std::vector<int*> temps;
    auto upload_rows = [&](const std::vector<int>& rows) {
        int* d = dalloc<int>(rows.size());
        temps.push_back(d);
        RAMD_HIP(hipMemcpyAsync(d, rows.data(), sizeof(int) * rows.size(), hipMemcpyHostToDevice, st));
        return d;
    };
    for (int b = 0; b < nsym; b++) {
        const std::vector<int>& rows = sym_bins[b];
        if (rows.empty()) continue;
        const int nr = (int)rows.size();
        int* d_rows = upload_rows(rows);
        switch (kSymT[b]) {
#define SYM_CASE(TT, WW)                                                                                          \
    case TT:                                                                                                      \
        hipLaunchKernelGGL((symbolic_lds<TT, WW>), dim3((nr + WW - 1) / WW), dim3(WW * kWave), 0, st, nr, d_rows, \
                           A->rowptr, A->col, B->rowptr, B->col, d_counts);                                       \
        break;
            SYM_CASE(64, 4)
            SYM_CASE(128, 4)
            SYM_CASE(256, 4)
            SYM_CASE(512, 4)
            SYM_CASE(1024, 4)
            SYM_CASE(2048, 4)
            SYM_CASE(4096, 2)
            SYM_CASE(8192, 1)
#undef SYM_CASE
        }
        RAMD_HIP(hipGetLastError());
    }
    void* sym_pool = nullptr;
    if (!sym_bins[nsym].empty()) {
        const std::vector<int>& rows = sym_bins[nsym];
        const int nr = (int)rows.size();
        std::vector<int64_t> off(nr);
        std::vector<int> sz(nr);
        int64_t tot = 0;
        for (int r = 0; r < nr; r++) {
            sz[r] = next_pow2(prod[rows[r]]);
            off[r] = tot;
            tot += sz[r];
        }
        int* d_rows = upload_rows(rows);
        int64_t* d_off = dalloc<int64_t>(nr);
        int* d_sz = dalloc<int>(nr);
        sym_pool = dmalloc(sizeof(int) * tot);
        RAMD_HIP(hipMemcpyAsync(d_off, off.data(), sizeof(int64_t) * nr, hipMemcpyHostToDevice, st));
        RAMD_HIP(hipMemcpyAsync(d_sz, sz.data(), sizeof(int) * nr, hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(symbolic_global, dim3(nr), dim3(kWave), 0, st, nr, d_rows, d_off, d_sz, (int*)sym_pool,
                           A->rowptr, A->col, B->rowptr, B->col, d_counts);
        RAMD_HIP(hipGetLastError());
        RAMD_HIP(hipStreamSynchronize(st));
        dfree(d_off);
        dfree(d_sz);
    }
    std::vector<int> counts((size_t)m);
    if (m) RAMD_HIP(hipMemcpyAsync(counts.data(), d_counts, sizeof(int) * m, hipMemcpyDeviceToHost, st));
    RAMD_HIP(hipStreamSynchronize(st));
    dfree(sym_pool);
    for (int* p : temps) dfree(p);
    temps.clear();

    // --- bound row pointers
    std::vector<int> rp_bound((size_t)m + 1, 0);
    for (int i = 0; i < m; i++) {
        const int64_t next = (int64_t)rp_bound[i] + counts[i];
        RAMD_CHECK(next <= INT32_MAX, RAMD_E_ARG, "spgemm: product has more than 2^31 entries");
        rp_bound[i + 1] = (int)next;
    }
    const int64_t nnz_bound = rp_bound[m];
    int* d_rpb = dalloc<int>((size_t)m + 1);
    int* d_col = dalloc<int>((size_t)std::max<int64_t>(nnz_bound, 1));
    double* d_val = dalloc<double>((size_t)std::max<int64_t>(nnz_bound, 1));
    RAMD_HIP(hipMemcpyAsync(d_rpb, rp_bound.data(), sizeof(int) * (m + 1), hipMemcpyHostToDevice, st));

    // --- numeric bins by distinct count (table = 2x, power of two)
    static const int kNumT[] = {32, 64, 128, 256, 512, 1024, 2048};
    const int nnum = 7;
    std::vector<std::vector<int>> num_bins(nnum + 1);
    for (int i = 0; i < m; i++) {
        if (counts[i] == 0) continue;
        const int64_t need = 2 * (int64_t)counts[i];
        int b = 0;
        while (b < nnum && kNumT[b] < need) b++;
        num_bins[b].push_back(i);
    }
    int* d_len = dalloc<int>((size_t)std::max(m, 1));
    RAMD_HIP(hipMemsetAsync(d_len, 0, sizeof(int) * std::max(m, 1), st));
    for (int b = 0; b < nnum; b++) {
        const std::vector<int>& rows = num_bins[b];
        if (rows.empty()) continue;
        const int nr = (int)rows.size();
        int* d_rows = upload_rows(rows);
        switch (kNumT[b]) {
#define NUM_CASE(TT, WW)                                                                                          \
    case TT:                                                                                                      \
        hipLaunchKernelGGL((numeric_lds<TT, WW>), dim3((nr + WW - 1) / WW), dim3(WW * kWave), 0, st, nr, d_rows,  \
                           A->rowptr, A->col, A->val, B->rowptr, B->col, B->val, d_rpb, d_col, d_val, d_len);     \
        break;
            NUM_CASE(32, 4)
            NUM_CASE(64, 4)
            NUM_CASE(128, 4)
            NUM_CASE(256, 4)
            NUM_CASE(512, 4)
            NUM_CASE(1024, 2)
            NUM_CASE(2048, 1)
#undef NUM_CASE
        }
        RAMD_HIP(hipGetLastError());
    }
    void *kpool = nullptr, *spool = nullptr, *lcpool = nullptr, *lvpool = nullptr;
    int64_t* d_off = nullptr;
    int* d_sz = nullptr;
    if (!num_bins[nnum].empty()) {
        const std::vector<int>& rows = num_bins[nnum];
        const int nr = (int)rows.size();
        std::vector<int64_t> off(nr);
        std::vector<int> sz(nr);
        int64_t tot = 0;
        for (int r = 0; r < nr; r++) {
            sz[r] = next_pow2(2 * (int64_t)counts[rows[r]]);
            off[r] = tot;
            tot += sz[r];
        }
        int* d_rows = upload_rows(rows);
        d_off = dalloc<int64_t>(nr);
        d_sz = dalloc<int>(nr);
        kpool = dmalloc(sizeof(int) * tot);
        spool = dmalloc(sizeof(double) * tot);
        lcpool = dmalloc(sizeof(int) * tot);
        lvpool = dmalloc(sizeof(double) * tot);
        RAMD_HIP(hipMemcpyAsync(d_off, off.data(), sizeof(int64_t) * nr, hipMemcpyHostToDevice, st));
        RAMD_HIP(hipMemcpyAsync(d_sz, sz.data(), sizeof(int) * nr, hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(numeric_global, dim3(nr), dim3(kWave), 0, st, nr, d_rows, d_off, d_sz, (int*)kpool,
                           (double*)spool, (int*)lcpool, (double*)lvpool, A->rowptr, A->col, A->val, B->rowptr,
                           B->col, B->val, d_rpb, d_col, d_val, d_len);
        RAMD_HIP(hipGetLastError());
    }
    std::vector<int> len((size_t)m);
    if (m) RAMD_HIP(hipMemcpyAsync(len.data(), d_len, sizeof(int) * m, hipMemcpyDeviceToHost, st));
    RAMD_HIP(hipStreamSynchronize(st));
    dfree(kpool);
    dfree(spool);
    dfree(lcpool);
    dfree(lvpool);
    dfree(d_off);
    dfree(d_sz);
    for (int* p : temps) dfree(p);
    temps.clear();

    // --- final row pointers, compaction where entries were dropped
    std::vector<int> rp((size_t)m + 1, 0);
    for (int i = 0; i < m; i++) rp[i + 1] = rp[i] + len[i];
    const int64_t nnz = rp[m];
    int* d_rp = dalloc<int>((size_t)m + 1);
    RAMD_HIP(hipMemcpyAsync(d_rp, rp.data(), sizeof(int) * (m + 1), hipMemcpyHostToDevice, st));
    if (nnz != nnz_bound) {
        int* c2 = dalloc<int>((size_t)std::max<int64_t>(nnz, 1));
        double* v2 = dalloc<double>((size_t)std::max<int64_t>(nnz, 1));
        if (m) {
            hipLaunchKernelGGL(compact_rows, dim3(m), dim3(64), 0, st, m, d_rpb, d_rp, d_col, d_val, c2, v2);
            RAMD_HIP(hipGetLastError());
        }
        RAMD_HIP(hipStreamSynchronize(st));
        dfree(d_col);
        dfree(d_val);
        d_col = c2;
        d_val = v2;
    }
    RAMD_HIP(hipStreamSynchronize(st));
    dfree(d_rpb);
    dfree(d_len);
    dfree(d_counts);
    return csr_wrap(ctx, m, n, nnz, d_rp, d_col, d_val, rp.data());
}

ramd_csr_s* spgemm(ramd_csr_s* A, ramd_csr_s* B) { return spgemm_impl(A, B); }

ramd_csr_s* galerkin(ramd_csr_s* A, ramd_csr_s* P)
{
    ramd_csr_s* AP = spgemm_impl(A, P);
    ramd_csr_s* R = csr_transpose(P);
    ramd_csr_s* Ac = nullptr;
    try {
        Ac = spgemm_impl(R, AP);
    } catch (...) {
        csr_free(AP);
        throw;
    }
    csr_free(AP);
    if (Ac->n_rows) {
        hipLaunchKernelGGL(move_diag_kernel, dim3((Ac->n_rows + 255) / 256), dim3(256), 0, Ac->ctx->stream,
                           Ac->n_rows, Ac->rowptr, Ac->col, Ac->val);
        RAMD_HIP(hipGetLastError());
    }
    RAMD_HIP(hipStreamSynchronize(Ac->ctx->stream));
    Ac->diag_first = true;
    return Ac;
}

}  // namespace ramd

using namespace ramd;

extern "C" {

int ramd_spgemm(ramd_csr A, ramd_csr B, ramd_csr* C)
{
    return guarded([&] {
        RAMD_CHECK(A && B && C, RAMD_E_ARG, "ramd_spgemm: NULL argument");
        RAMD_CHECK(A->ctx == B->ctx, RAMD_E_ARG, "ramd_spgemm: matrices belong to different contexts");
        A->ctx->set_current();
        *C = spgemm(A, B);
    });
}

int ramd_spgemm_tn(ramd_csr P, ramd_csr B, ramd_csr* C)
{
    return guarded([&] {
        RAMD_CHECK(P && B && C, RAMD_E_ARG, "ramd_spgemm_tn: NULL argument");
        RAMD_CHECK(P->ctx == B->ctx, RAMD_E_ARG, "ramd_spgemm_tn: matrices belong to different contexts");
        RAMD_CHECK(P->n_rows == B->n_rows, RAMD_E_ARG, "ramd_spgemm_tn: P has %d rows, B has %d", P->n_rows,
                   B->n_rows);
        P->ctx->set_current();
        *C = spgemm(csr_transpose(P), B);
    });
}

int ramd_galerkin(ramd_csr A, ramd_csr P, ramd_csr* Ac)
{
    return guarded([&] {
        RAMD_CHECK(A && P && Ac, RAMD_E_ARG, "ramd_galerkin: NULL argument");
        RAMD_CHECK(A->ctx == P->ctx, RAMD_E_ARG, "ramd_galerkin: matrices belong to different contexts");
        RAMD_CHECK(A->n_cols == P->n_rows && A->n_rows == A->n_cols, RAMD_E_ARG,
                   "ramd_galerkin: A is %dx%d, P is %dx%d", A->n_rows, A->n_cols, P->n_rows, P->n_cols);
        A->ctx->set_current();
        *Ac = galerkin(A, P);
    });
}

}  // extern "C"

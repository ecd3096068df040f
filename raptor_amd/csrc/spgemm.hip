// Two-phase sparse matrix-matrix product C = A B on the device, with the
// reference's accumulation order and drop rule (util/linalg/matmult.cpp:
// 71-204, util/linalg/add.cpp:39-78, core/matrix.cpp:922-986).
//
// Phase 1 (symbolic): one wavefront per row (a group of 8 lanes for rows of
// at most 128 products) inserts every product's column into a hash set (LDS
// when it fits, global memory otherwise) and counts the distinct columns ->
// an upper bound of the row length.
// Phase 2 (numeric): one wavefront per row (8 lanes for rows of at most 32
// distinct columns, 16 for at most 64) walks the row of A in stored order;
// for each a_ij the row's lanes spread over row j of B.  Columns inside one
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
// lanes per row for the bins of short rows (tables of at most 128 slots)
constexpr int kSmallGroup = 8;
// rows per workgroup of WAVES wavefronts with G lanes per row
template <int WAVES, int G>
constexpr int kGroups = WAVES * kWave / G;



__device__ __forceinline__ unsigned hash_col(int c) { return (unsigned)c * 2654435761u; }

// Lane groups of G lanes (G divides 64) inside a wavefront: one row per
// group.  The group's bits of a ballot, its lane index, and the count and
// prefix of a predicate over the group.
template <int G>
__device__ __forceinline__ int grp_lane()
{
    return threadIdx.x & (G - 1);
}

template <int G>
__device__ __forceinline__ unsigned long long grp_bits(bool pred)
{
    const unsigned long long m = __ballot(pred);
    if (G == kWave) return m;
    return (m >> ((threadIdx.x & (kWave - 1)) & ~(G - 1))) & ((1ull << (G == kWave ? 0 : G)) - 1);
}

template <int G>
__device__ __forceinline__ int grp_count(bool pred)
{
    return __popcll(grp_bits<G>(pred));
}

template <int G>
__device__ __forceinline__ int grp_prefix(bool pred)
{
    const unsigned long long m = grp_bits<G>(pred);
    const int lane = grp_lane<G>();
    return __popcll(m & ((1ull << lane) - 1));
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
// symbolic: rows[] lists the rows of this bin; table size T per row, one
// row per group of G lanes (a whole wavefront for the larger tables)
// ---------------------------------------------------------------------------
template <int G>
__device__ void symbolic_row(int row, int* keys, int T, const int* __restrict__ arp, const int* __restrict__ acol,
                             const int* __restrict__ brp, const int* __restrict__ bcol, int* __restrict__ counts)
{
    const int lane = grp_lane<G>();
    for (int t = lane; t < T; t += G) keys[t] = -1;
    __builtin_amdgcn_wave_barrier();
    for (int j = arp[row]; j < arp[row + 1]; j++) {
        const int c = acol[j];
        for (int t = brp[c] + lane; t < brp[c + 1]; t += G) hash_insert(keys, T, bcol[t]);
    }
    __builtin_amdgcn_wave_barrier();
    __threadfence_block();
    int cnt = 0;
    for (int t0 = 0; t0 < T; t0 += G) {
        const int t = t0 + lane;
        cnt += grp_count<G>(t < T && keys[t] != -1);
    }
    if (lane == 0) counts[row] = cnt;
}

template <int T, int WAVES, int G = kWave>
__global__ __launch_bounds__(WAVES * kWave) void symbolic_lds(int nrows, const int* __restrict__ rows,
                                                            const int* __restrict__ arp, const int* __restrict__ acol,
                                                            const int* __restrict__ brp, const int* __restrict__ bcol,
                                                            int* __restrict__ counts)
{
    constexpr int GROUPS = kGroups<WAVES, G>;
    __shared__ int keys[GROUPS][T];
    const int w = threadIdx.x / G;
    const int r = blockIdx.x * GROUPS + w;
    if (r >= nrows) return;
    symbolic_row<G>(rows[r], keys[w], T, arp, acol, brp, bcol, counts);
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
    symbolic_row<kWave>(rows[r], pool + tab_off[r], tab_size[r], arp, acol, brp, bcol, counts);
}

// ---------------------------------------------------------------------------
// numeric
// ---------------------------------------------------------------------------
template <int G>
__device__ void numeric_row(int row, int* keys, double* sums, int* lcol, double* lval, int T,
                            const int* __restrict__ arp, const int* __restrict__ acol, const double* __restrict__ aval,
                            const int* __restrict__ brp, const int* __restrict__ bcol, const double* __restrict__ bval,
                            const int* __restrict__ crp, int* __restrict__ ccol, double* __restrict__ cval,
                            int* __restrict__ clen)
{
    const int lane = grp_lane<G>();
    for (int t = lane; t < T; t += G) {
        keys[t] = -1;
        sums[t] = 0.0;
    }
    __builtin_amdgcn_wave_barrier();
    for (int j = arp[row]; j < arp[row + 1]; j++) {
        const int c = acol[j];
        const double a = aval[j];
        for (int t = brp[c] + lane; t < brp[c + 1]; t += G) {
            const int slot = hash_insert(keys, T, bcol[t]);
            sums[slot] += a * bval[t];
        }
        // every lane's update for this a_ij lands before the next a_ij's
        __builtin_amdgcn_wave_barrier();
        __threadfence_block();
    }
    // gather kept entries (|sum| > 1e-16) into a compact list
    int kept = 0;
    for (int t0 = 0; t0 < T; t0 += G) {
        const int t = t0 + lane;
        bool keep = false;
        int k = -1;
        double v = 0.0;
        if (t < T) {
            k = keys[t];
            v = sums[t];
            keep = (k != -1) && (fabs(v) > 1e-16);
        }
        const int p = grp_prefix<G>(keep);
        if (keep) {
            lcol[kept + p] = k;
            lval[kept + p] = v;
        }
        kept += grp_count<G>(keep);
    }
    __builtin_amdgcn_wave_barrier();
    __threadfence_block();
    // rank sort by column, write out in ascending order
    const int base = crp[row];
    for (int e = lane; e < kept; e += G) {
        const int ce = lcol[e];
        int rank = 0;
        for (int f = 0; f < kept; f++) rank += (lcol[f] < ce);
        ccol[base + rank] = ce;
        cval[base + rank] = lval[e];
    }
    if (lane == 0) clen[row] = kept;
}

template <int T, int WAVES, int G = kWave>
__global__ __launch_bounds__(WAVES * kWave) void numeric_lds(
    int nrows, const int* __restrict__ rows, const int* __restrict__ arp, const int* __restrict__ acol,
    const double* __restrict__ aval, const int* __restrict__ brp, const int* __restrict__ bcol,
    const double* __restrict__ bval, const int* __restrict__ crp, int* __restrict__ ccol,
    double* __restrict__ cval, int* __restrict__ clen)
{
    constexpr int GROUPS = kGroups<WAVES, G>;
    __shared__ int keys[GROUPS][T];
    __shared__ double sums[GROUPS][T];
    __shared__ int lcol[GROUPS][T / 2];
    __shared__ double lval[GROUPS][T / 2];
    const int w = threadIdx.x / G;
    const int r = blockIdx.x * GROUPS + w;
    if (r >= nrows) return;
    numeric_row<G>(rows[r], keys[w], sums[w], lcol[w], lval[w], T, arp, acol, aval, brp, bcol, bval, crp, ccol, cval,
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
    numeric_row<kWave>(rows[r], kpool + o, spool + o, lcpool + o, lvpool + o, tab_size[r], arp, acol, aval, brp, bcol, bval,
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

// ---------------------------------------------------------------------------
// Rows binned by a per-row key on the device: bin b < nb holds the rows with
// th[b-1] < mult * key <= th[b], bin nb those above th[nb-1]; with skip_zero
// a row whose key is 0 goes into no bin.  Pass 0 counts per bin, pass 1
// writes the rows of bin b from base[b] on (the order inside a bin is not
// fixed: rows are independent, so it changes no result).  Each workgroup
// aggregates in LDS and touches each global counter once.
// ---------------------------------------------------------------------------
constexpr int kMaxBins = 16;
struct BinSpec {
    int64_t th[kMaxBins];
    int nb;
    int mult;
    int skip_zero;
    int base[kMaxBins + 1];
};

template <int PASS, typename K>
__global__ __launch_bounds__(256) void bin_rows_kernel(int m, const K* __restrict__ key, BinSpec sp,
                                                       int* __restrict__ counter, int* __restrict__ out)
{
    __shared__ int hist[kMaxBins + 1];
    __shared__ int gbase[kMaxBins + 1];
    const int t = threadIdx.x;
    if (t <= sp.nb) hist[t] = 0;
    __syncthreads();
    const int i = blockIdx.x * 256 + t;
    int b = -1, pos = 0;
    if (i < m) {
        const int64_t k = (int64_t)key[i];
        if (!(sp.skip_zero && k == 0)) {
            const int64_t v = k * sp.mult;
            b = 0;
            while (b < sp.nb && sp.th[b] < v) b++;
            pos = atomicAdd(&hist[b], 1);
        }
    }
    __syncthreads();
    if (t <= sp.nb && hist[t]) gbase[t] = atomicAdd(&counter[t], hist[t]);
    __syncthreads();
    if (PASS == 1 && b >= 0) out[sp.base[b] + gbase[b] + pos] = i;
}

template <typename K>
__global__ void gather_keys(int nr, const int* __restrict__ rows, const K* __restrict__ key, int64_t* __restrict__ out)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < nr) out[r] = (int64_t)key[rows[r]];
}

// Device buffers released at scope exit (also on an exception).
struct DevTemps {
    std::vector<void*> p;
    template <class T>
    T* get(size_t n)
    {
        T* q = dalloc<T>(std::max<size_t>(n, 1));
        p.push_back(q);
        return q;
    }
    ~DevTemps()
    {
        for (void* q : p) dfree(q);
    }
};

// Bins rows 0..m-1 by key (see bin_rows_kernel) into rows_out (m ints);
// returns the first row-list index and the count of every bin (nb + 1).
template <typename K>
static void bin_rows(ramd_ctx ctx, int m, const K* key, int mult, bool skip_zero, const int* th, int nb,
                     int* rows_out, std::vector<int>& first, std::vector<int>& count, DevTemps& tmp)
{
    hipStream_t s = ctx->stream;
    BinSpec sp{};
    for (int b = 0; b < nb; b++) sp.th[b] = th[b];
    sp.nb = nb;
    sp.mult = mult;
    sp.skip_zero = skip_zero ? 1 : 0;
    int* counter = tmp.get<int>(2 * (kMaxBins + 1));
    RAMD_HIP(hipMemsetAsync(counter, 0, sizeof(int) * 2 * (kMaxBins + 1), s));
    const dim3 g((unsigned)std::max(1, (m + 255) / 256));
    count.assign(nb + 1, 0);
    first.assign(nb + 1, 0);
    if (m) {
        hipLaunchKernelGGL((bin_rows_kernel<0, K>), g, dim3(256), 0, s, m, key, sp, counter, nullptr);
        RAMD_HIP(hipGetLastError());
        RAMD_HIP(hipMemcpyAsync(count.data(), counter, sizeof(int) * (nb + 1), hipMemcpyDeviceToHost, s));
        RAMD_HIP(hipStreamSynchronize(s));
        for (int b = 0, acc = 0; b <= nb; b++) {
            first[b] = acc;
            sp.base[b] = acc;
            acc += count[b];
        }
        hipLaunchKernelGGL((bin_rows_kernel<1, K>), g, dim3(256), 0, s, m, key, sp, counter + kMaxBins + 1,
                           rows_out);
        RAMD_HIP(hipGetLastError());
    }
}

// The keys of rows[0..nr) on the host.
template <typename K>
static std::vector<int64_t> keys_of(ramd_ctx ctx, const int* rows, int nr, const K* key, DevTemps& tmp)
{
    std::vector<int64_t> out((size_t)nr);
    if (!nr) return out;
    int64_t* d = tmp.get<int64_t>((size_t)nr);
    hipLaunchKernelGGL((gather_keys<K>), dim3((nr + 255) / 256), dim3(256), 0, ctx->stream, nr, rows, key, d);
    RAMD_HIP(hipGetLastError());
    RAMD_HIP(hipMemcpyAsync(out.data(), d, sizeof(int64_t) * nr, hipMemcpyDeviceToHost, ctx->stream));
    RAMD_HIP(hipStreamSynchronize(ctx->stream));
    return out;
}

static int next_pow2(int64_t v)
{
    int p = 1;
    while (p < v) p <<= 1;
    return p;
}

// Runs symbolic+numeric for C = A B, returns a new handle.  Everything that
// scales with the rows stays on the device: the host reads per-bin counts,
// the rows of the global-hash bins with their sizes, and the totals.
static ramd_csr_s* spgemm_impl(ramd_csr_s* A, ramd_csr_s* B)
{
    ramd_ctx ctx = A->ctx;
    hipStream_t st = ctx->stream;
    const int m = A->n_rows, n = B->n_cols;
    RAMD_CHECK(A->n_cols == B->n_rows, RAMD_E_ARG, "spgemm: inner dimensions %d and %d differ", A->n_cols, B->n_rows);
    DevTemps tmp;

    // --- product counts, and the rows binned by them
    int64_t* d_prod = tmp.get<int64_t>((size_t)m);
    if (m) {
        hipLaunchKernelGGL(count_products, dim3((m + 255) / 256), dim3(256), 0, st, m, A->rowptr, A->col, B->rowptr,
                           d_prod);
        RAMD_HIP(hipGetLastError());
    }
    static const int kSymT[] = {64, 128, 256, 512, 1024, 2048, 4096, 8192};
    const int nsym = 8;
    int* d_rows = tmp.get<int>((size_t)m);
    std::vector<int> first, count;
    bin_rows<int64_t>(ctx, m, d_prod, 1, false, kSymT, nsym, d_rows, first, count, tmp);

    // --- symbolic: distinct columns per row
    int* d_counts = tmp.get<int>((size_t)m);
    for (int b = 0; b < nsym; b++) {
        const int nr = count[b];
        if (!nr) continue;
        const int* rows = d_rows + first[b];
        switch (kSymT[b]) {
#define SYM_CASE(TT, WW)                                                                                          \
    case TT:                                                                                                      \
        hipLaunchKernelGGL((symbolic_lds<TT, WW>), dim3((nr + WW - 1) / WW), dim3(WW * kWave), 0, st, nr, rows,   \
                           A->rowptr, A->col, B->rowptr, B->col, d_counts);                                       \
        break;
#define SYM_CASE_G(TT, WW, GG)                                                                                    \
    case TT: {                                                                                                    \
        constexpr int groups = kGroups<WW, GG>;                                                                   \
        hipLaunchKernelGGL((symbolic_lds<TT, WW, GG>), dim3((nr + groups - 1) / groups), dim3(WW * kWave), 0, st,  \
                           nr, rows, A->rowptr, A->col, B->rowptr, B->col, d_counts);                             \
        break;                                                                                                    \
    }
            SYM_CASE_G(64, 4, kSmallGroup)
            SYM_CASE_G(128, 4, kSmallGroup)
            SYM_CASE(256, 4)
            SYM_CASE(512, 4)
            SYM_CASE(1024, 4)
            SYM_CASE(2048, 4)
            SYM_CASE(4096, 2)
            SYM_CASE(8192, 1)
#undef SYM_CASE
#undef SYM_CASE_G
        }
        RAMD_HIP(hipGetLastError());
    }
    if (count[nsym]) {
        const int nr = count[nsym];
        const int* rows = d_rows + first[nsym];
        const std::vector<int64_t> prod = keys_of<int64_t>(ctx, rows, nr, d_prod, tmp);
        std::vector<int64_t> off(nr);
        std::vector<int> sz(nr);
        int64_t tot = 0;
        for (int r = 0; r < nr; r++) {
            sz[r] = next_pow2(prod[r]);
            off[r] = tot;
            tot += sz[r];
        }
        int64_t* d_off = tmp.get<int64_t>(nr);
        int* d_sz = tmp.get<int>(nr);
        int* pool = tmp.get<int>((size_t)tot);
        RAMD_HIP(hipMemcpyAsync(d_off, off.data(), sizeof(int64_t) * nr, hipMemcpyHostToDevice, st));
        RAMD_HIP(hipMemcpyAsync(d_sz, sz.data(), sizeof(int) * nr, hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(symbolic_global, dim3(nr), dim3(kWave), 0, st, nr, rows, d_off, d_sz, pool, A->rowptr,
                           A->col, B->rowptr, B->col, d_counts);
        RAMD_HIP(hipGetLastError());
        RAMD_HIP(hipStreamSynchronize(st));  // the host vectors above go out of scope
    }

    // --- bound row pointers and the output arrays at those bounds
    int* d_rpb = tmp.get<int>((size_t)m + 1);
    const int64_t nnz_bound = device_rowptr(ctx, d_counts, m, d_rpb, "spgemm: the product");
    int* d_col = dalloc<int>((size_t)std::max<int64_t>(nnz_bound, 1));
    double* d_val = nullptr;
    int* d_rp = nullptr;
    try {
        d_val = dalloc<double>((size_t)std::max<int64_t>(nnz_bound, 1));

        // --- numeric bins by distinct count (table = 2x, power of two)
        static const int kNumT[] = {32, 64, 128, 256, 512, 1024, 2048};
        const int nnum = 7;
        bin_rows<int>(ctx, m, d_counts, 2, true, kNumT, nnum, d_rows, first, count, tmp);
        int* d_len = tmp.get<int>((size_t)m);
        RAMD_HIP(hipMemsetAsync(d_len, 0, sizeof(int) * std::max(m, 1), st));
        for (int b = 0; b < nnum; b++) {
            const int nr = count[b];
            if (!nr) continue;
            const int* rows = d_rows + first[b];
            switch (kNumT[b]) {
#define NUM_CASE(TT, WW)                                                                                          \
    case TT:                                                                                                      \
        hipLaunchKernelGGL((numeric_lds<TT, WW>), dim3((nr + WW - 1) / WW), dim3(WW * kWave), 0, st, nr, rows,    \
                           A->rowptr, A->col, A->val, B->rowptr, B->col, B->val, d_rpb, d_col, d_val, d_len);     \
        break;
#define NUM_CASE_G(TT, WW, GG)                                                                                    \
    case TT: {                                                                                                    \
        constexpr int groups = kGroups<WW, GG>;                                                                   \
        hipLaunchKernelGGL((numeric_lds<TT, WW, GG>), dim3((nr + groups - 1) / groups), dim3(WW * kWave), 0, st,   \
                           nr, rows, A->rowptr, A->col, A->val, B->rowptr, B->col, B->val, d_rpb, d_col, d_val,   \
                           d_len);                                                                                \
        break;                                                                                                    \
    }
                NUM_CASE_G(32, 4, kSmallGroup)
                NUM_CASE_G(64, 4, kSmallGroup)
                NUM_CASE_G(128, 4, 2 * kSmallGroup)
                NUM_CASE(256, 4)
                NUM_CASE(512, 4)
                NUM_CASE(1024, 2)
                NUM_CASE(2048, 1)
#undef NUM_CASE
#undef NUM_CASE_G
            }
            RAMD_HIP(hipGetLastError());
        }
        if (count[nnum]) {
            const int nr = count[nnum];
            const int* rows = d_rows + first[nnum];
            const std::vector<int64_t> cnt = keys_of<int>(ctx, rows, nr, d_counts, tmp);
            std::vector<int64_t> off(nr);
            std::vector<int> sz(nr);
            int64_t tot = 0;
            for (int r = 0; r < nr; r++) {
                sz[r] = next_pow2(2 * cnt[r]);
                off[r] = tot;
                tot += sz[r];
            }
            int64_t* d_off = tmp.get<int64_t>(nr);
            int* d_sz = tmp.get<int>(nr);
            int* kpool = tmp.get<int>((size_t)tot);
            double* spool = tmp.get<double>((size_t)tot);
            int* lcpool = tmp.get<int>((size_t)tot);
            double* lvpool = tmp.get<double>((size_t)tot);
            RAMD_HIP(hipMemcpyAsync(d_off, off.data(), sizeof(int64_t) * nr, hipMemcpyHostToDevice, st));
            RAMD_HIP(hipMemcpyAsync(d_sz, sz.data(), sizeof(int) * nr, hipMemcpyHostToDevice, st));
            hipLaunchKernelGGL(numeric_global, dim3(nr), dim3(kWave), 0, st, nr, rows, d_off, d_sz, kpool, spool,
                               lcpool, lvpool, A->rowptr, A->col, A->val, B->rowptr, B->col, B->val, d_rpb, d_col,
                               d_val, d_len);
            RAMD_HIP(hipGetLastError());
            RAMD_HIP(hipStreamSynchronize(st));  // the host vectors above go out of scope
        }

        // --- final row pointers, compaction where entries were dropped
        d_rp = dalloc<int>((size_t)m + 1);
        const int64_t nnz = device_rowptr(ctx, d_len, m, d_rp, "spgemm: the product");
        if (nnz != nnz_bound) {
            int* c2 = dalloc<int>((size_t)std::max<int64_t>(nnz, 1));
            double* v2 = nullptr;
            try {
                v2 = dalloc<double>((size_t)std::max<int64_t>(nnz, 1));
            } catch (...) {
                dfree(c2);
                throw;
            }
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
        return csr_wrap(ctx, m, n, nnz, d_rp, d_col, d_val, nullptr);
    } catch (...) {
        dfree(d_col);
        dfree(d_val);
        dfree(d_rp);
        throw;
    }
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

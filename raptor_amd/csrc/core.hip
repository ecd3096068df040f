// Context, device memory, device CSR handles and the C entry points that
// manage them.
#include <atomic>
#include <algorithm>
#include <climits>
#include <cstdarg>
#include <cstring>
#include <new>

#include "common.hpp"

namespace ramd {

static thread_local std::string g_last_error;

void set_error(const char* fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
}

void fail(int code, const char* fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
    throw Failure{code};
}

void* dmalloc(size_t bytes)
{
    void* p = nullptr;
    if (bytes == 0) bytes = 8;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        fail(RAMD_E_ALLOC, "hipMalloc(%zu bytes) failed: %s", bytes, hipGetErrorString(e));
    }
    return p;
}

void dfree(void* p)
{
    if (p) (void)hipFree(p);
}

std::vector<int> make_row_blocks(const int* rp, int n_rows, int cap)
{
    std::vector<int> b;
    b.reserve(n_rows / 64 + 2);
    b.push_back(0);
    int r = 0;
    while (r < n_rows) {
        int start = r;
        if (rp[r + 1] - rp[r] > cap) {
            r++;
        } else {
            int64_t acc = 0;
            while (r < n_rows && r - start < kMaxRowsPerBlock && acc + (rp[r + 1] - rp[r]) <= cap) {
                acc += rp[r + 1] - rp[r];
                r++;
            }
        }
        b.push_back(r);
    }
    return b;
}

ramd_csr_s* csr_wrap(ramd_ctx ctx, int n_rows, int n_cols, int64_t nnz, int* rowptr_d, int* col_d,
                     double* val_d, const int* rowptr_host)
{
    static std::atomic<uint64_t> next_uid{1};
    ramd_csr_s* A = new ramd_csr_s();
    A->uid = next_uid.fetch_add(1);
    A->ctx = ctx;
    A->n_rows = n_rows;
    A->n_cols = n_cols;
    A->nnz = nnz;
    A->rowptr = rowptr_d;
    A->col = col_d;
    A->val = val_d;
    if (rowptr_host) {
        A->h_rowptr.assign(rowptr_host, rowptr_host + n_rows + 1);
    } else {
        A->h_rowptr.resize((size_t)n_rows + 1);
        RAMD_HIP(hipMemcpyAsync(A->h_rowptr.data(), rowptr_d, sizeof(int) * (n_rows + 1),
                                hipMemcpyDeviceToHost, ctx->stream));
        RAMD_HIP(hipStreamSynchronize(ctx->stream));
    }
    std::vector<int> blocks = make_row_blocks(A->h_rowptr.data(), n_rows, kStreamCap);
    A->n_blocks = (int)blocks.size() - 1;
    A->blocks = dalloc<int>(blocks.size());
    RAMD_HIP(hipMemcpyAsync(A->blocks, blocks.data(), sizeof(int) * blocks.size(),
                            hipMemcpyHostToDevice, ctx->stream));
    RAMD_HIP(hipStreamSynchronize(ctx->stream));
    for (auto& rb : A->rb) {
        rb.cap = kStreamCap;
        rb.n = A->n_blocks;
        rb.d = A->blocks;
    }
    return A;
}

// diagonal-indexed planes (plane k of row r at k * n + r) to the slot-grouped
// layout; slots past nd get 0
__global__ void dia_group_kernel(int n, int nd, int nslots, DiaArgs a, const double* __restrict__ planes,
                                 double* __restrict__ out)
{
    const int64_t m = (int64_t)nslots * n;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += stride) {
        const int t = (int)(i / n), r = (int)(i % n);
        out[dia_index<double>(t, r, n)] = t < nd ? planes[(size_t)a.ord[t] * n + r] : 0.0;
    }
}

// FP32 copy of the slot-grouped FP64 values (each value rounded once)
__global__ void dia_fp32_kernel(int n, int nd, int nslots, const double* __restrict__ v64, float* __restrict__ v32)
{
    const int64_t m = (int64_t)nslots * n;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += stride) {
        const int t = (int)(i / n), r = (int)(i % n);
        v32[dia_index<float>(t, r, n)] = t < nd ? (float)v64[dia_index<double>(t, r, n)] : 0.0f;
    }
}

static void dia_make_fp32(ramd_csr_s* A)
{
    Dia* d = A->dia;
    const int n = A->n_rows;
    const int ns = dia_slots<float>(d->nd);
    if (!d->v32) {
        d->v32 = dalloc<float>((size_t)ns * n);
        A->forms++;
    }
    const int grid = (int)std::min<int64_t>(((int64_t)ns * n + 255) / 256, 8192);
    hipLaunchKernelGGL(dia_fp32_kernel, dim3(grid), dim3(256), 0, A->ctx->stream, n, d->nd, ns, d->v64, d->v32);
    RAMD_HIP(hipGetLastError());
}

void dia_free(Dia* d)
{
    if (!d) return;
    dfree(d->mask);
    dfree(d->v64);
    dfree(d->v32);
    delete d;
}

__global__ void fill_int_kernel(int* __restrict__ p, int64_t n, int v)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

template <typename T>
__global__ void fill_zero_kernel(T* __restrict__ p, int64_t n)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = T(0);
}

// Distinct column offsets of the rows, into an open-addressing table of
// kDiaTable ints (INT_MIN = empty); *count counts insertions and *fail is set
// past kDiaMax distinct offsets or a full probe sequence (the form is then
// not built).
constexpr int kDiaTable = 256;
__global__ void dia_offsets_kernel(int n, const int* __restrict__ rp, const int* __restrict__ col,
                                   int* __restrict__ keys, int* __restrict__ count, int* __restrict__ fail)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || __atomic_load_n(fail, __ATOMIC_RELAXED)) return;
    for (int j = rp[i]; j < rp[i + 1]; j++) {
        const int d = col[j] - i;
        int h = (int)(((unsigned)d * 2654435761u) >> 24);
        bool placed = false;
        for (int probe = 0; probe < kDiaTable && !placed; probe++, h = (h + 1) & (kDiaTable - 1)) {
            int k = __atomic_load_n(&keys[h], __ATOMIC_RELAXED);
            if (k == INT_MIN) {
                k = atomicCAS(&keys[h], INT_MIN, d);
                if (k == INT_MIN) {
                    if (atomicAdd(count, 1) >= kDiaMax) atomicOr(fail, 1);
                    placed = true;
                    continue;
                }
            }
            placed = k == d;
        }
        if (!placed) {
            atomicOr(fail, 1);
            return;
        }
    }
}

// Planes (plane k of row r at k * n + r, k = rank of the offset among the
// sorted offs) and row masks; flags: 1 a repeated column, 2 a row not in
// ascending diagonal order, 4 a row not in diagonal-first order, 8 a value
// not exact in FP32.
__global__ void dia_fill_kernel(int n, int nd, int center, DiaArgs offs, const int* __restrict__ rp,
                                const int* __restrict__ col, const double* __restrict__ val,
                                double* __restrict__ planes, unsigned* __restrict__ mask, int* __restrict__ flags)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    unsigned m = 0;
    int f = 0, prev_a = -2, prev_d = -2;
    for (int j = rp[i]; j < rp[i + 1]; j++) {
        const int d = col[j] - i;
        int lo = 0, hi = nd - 1;  // the offset is one of offs.off[0, nd)
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (offs.off[mid] < d) lo = mid + 1;
            else hi = mid;
        }
        const int k = lo;
        if (m & (1u << k)) f |= 1;
        m |= 1u << k;
        const double v = val[j];
        planes[(size_t)k * n + i] = v;
        if ((double)(float)v != v) f |= 8;
        const int rd = k == center ? -1 : k;
        if (k <= prev_a) f |= 2;
        if (rd <= prev_d) f |= 4;
        prev_a = k;
        prev_d = rd;
    }
    mask[i] = m;
    if (f) atomicOr(flags, f);
}

void dia_box_detect(const long long* offs, int nd, int n, DiaBox& b)
{
    b = DiaBox();
    // a decomposition of every offset as dz * P + dy * L + dx, |d.| <= 1
    auto split = [&](long long o, long long L, long long P, int& dz, int& dy, int& dx) {
        for (dz = -1; dz <= 1; dz++)
            for (dy = -1; dy <= 1; dy++) {
                const long long r = o - dz * P - dy * L;
                if (r >= -1 && r <= 1) {
                    dx = (int)r;
                    return true;
                }
            }
        return false;
    };
    std::vector<long long> cand;
    for (int k = 0; k < nd; k++)
        for (int e = -1; e <= 1; e++)
            if (offs[k] + e >= 3) cand.push_back(offs[k] + e);
    std::sort(cand.begin(), cand.end());
    cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
    // the true grid: L the smallest line length, P the smallest plane size
    // that is a multiple of L with at least two lines and divides n
    for (long long L : cand)
        for (long long P : cand) {
            if (P < 2 * L || P % L || n % P || b.L) continue;
            bool ok = true;
            for (int k = 0; k < nd && ok; k++) {
                int dz, dy, dx;
                ok = split(offs[k], L, P, dz, dy, dx);
            }
            if (!ok) continue;
            b.L = (int)L;
            b.P = (int)P;
        }
    if (!b.L) return;
    // tile width: the widest of 64, 32, 16 that divides the line, else 64
    b.tx = b.L % 64 == 0 ? 64 : b.L % 32 == 0 ? 32 : b.L % 16 == 0 ? 16 : 64;
    for (int k = 0; k < kDiaMax; k++) {
        int dz = 0, dy = 0, dx = 0;
        if (k < nd) split(offs[k], b.L, b.P, dz, dy, dx);
        b.args.dz[k] = dz;
        b.args.rel[k] = dy * (b.tx + 2) + dx;
    }
}

// The stencil-diagonal form, built on the device from the matrix's device
// CSR arrays (the host reads back only the distinct offsets and four flags).
bool dia_build(ramd_csr_s* A)
{
    const int n = A->n_rows;
    if (A->n_rows != A->n_cols || n < kDiaMinRows || A->nnz == 0) return false;
    // the kernels gather x through 32-bit byte offsets
    if ((int64_t)n * (int64_t)sizeof(double) >= ((int64_t)1 << 32)) return false;
    hipStream_t s = A->ctx->stream;
    const dim3 rows((unsigned)((n + 255) / 256)), blk(256);
    int* keys = dalloc<int>(kDiaTable + 3);  // table, count, fail, flags
    struct Tmp {
        int* keys;
        double* planes = nullptr;
        unsigned* mask = nullptr;
        ~Tmp()
        {
            dfree(keys);
            dfree(planes);
            dfree(mask);
        }
    } tmp{keys};
    hipLaunchKernelGGL(fill_int_kernel, dim3(1), dim3(256), 0, s, keys, (int64_t)kDiaTable, INT_MIN);
    hipLaunchKernelGGL(fill_int_kernel, dim3(1), dim3(256), 0, s, keys + kDiaTable, (int64_t)3, 0);
    hipLaunchKernelGGL(dia_offsets_kernel, rows, blk, 0, s, n, A->rowptr, A->col, keys, keys + kDiaTable,
                       keys + kDiaTable + 1);
    RAMD_HIP(hipGetLastError());
    std::vector<int> tab(kDiaTable + 3);
    RAMD_HIP(hipMemcpyAsync(tab.data(), keys, sizeof(int) * tab.size(), hipMemcpyDeviceToHost, s));
    RAMD_HIP(hipStreamSynchronize(s));
    if (tab[kDiaTable + 1]) return false;
    std::vector<long long> offs;
    for (int h = 0; h < kDiaTable; h++)
        if (tab[h] != INT_MIN) offs.push_back(tab[h]);
    const int nd = (int)offs.size();
    if (nd == 0 || nd > kDiaMax || (double)nd * n > 1.15 * (double)A->nnz) return false;
    std::sort(offs.begin(), offs.end());
    int center = -1;
    for (int k = 0; k < nd; k++)
        if (offs[k] == 0) center = k;
    DiaArgs oa{};
    for (int k = 0; k < kDiaMax; k++) oa.off[k] = k < nd ? (int)offs[k] : 0;
    tmp.planes = dalloc<double>((size_t)nd * n);
    tmp.mask = dalloc<unsigned>((size_t)n);
    {
        const int g = (int)std::min<int64_t>(((int64_t)nd * n + 255) / 256, 8192);
        hipLaunchKernelGGL(fill_zero_kernel<double>, dim3(g), dim3(256), 0, s, tmp.planes, (int64_t)nd * n);
    }
    hipLaunchKernelGGL(dia_fill_kernel, rows, blk, 0, s, n, nd, center, oa, A->rowptr, A->col, A->val, tmp.planes,
                       tmp.mask, keys + kDiaTable + 2);
    RAMD_HIP(hipGetLastError());
    int flags = 0;
    RAMD_HIP(hipMemcpyAsync(&flags, keys + kDiaTable + 2, sizeof(int), hipMemcpyDeviceToHost, s));
    RAMD_HIP(hipStreamSynchronize(s));
    const bool all_asc = !(flags & 2), all_dfirst = !(flags & 4), exact32 = !(flags & 8);
    if ((flags & 1) || (!all_asc && !all_dfirst)) return false;
    Dia* d = new Dia();
    d->nd = nd;
    d->center = center;
    for (int k = 0; k < kDiaMax; k++) d->args.off[k] = k < nd ? (int)offs[k] : 0;
    int kk = 0;
    if (all_asc) {
        for (int k = 0; k < nd; k++) d->args.ord[kk++] = k;
    } else {
        d->args.ord[kk++] = center;
        for (int k = 0; k < nd; k++)
            if (k != center) d->args.ord[kk++] = k;
    }
    for (; kk < kDiaMax; kk++) d->args.ord[kk] = 0;
    for (int k = 0; k < kDiaMax; k++) d->args.pos[k] = 0;
    for (int t = 0; t < nd; t++) d->args.pos[d->args.ord[t]] = t;
    // runs of consecutive offsets and the LDS window they need
    {
        DiaWindow& w = d->window;
        int tot = 0;
        for (int k = 0; k < nd; k++) {
            if (k == 0 || offs[k] != offs[k - 1] + 1) {
                w.lo[w.nr] = (int)offs[k];
                w.len[w.nr] = 0;
                w.base[w.nr] = tot;
                w.nr++;
            }
            const int j = w.nr - 1;
            w.lofs[k] = w.base[j] + (int)(offs[k] - w.lo[j]);
            w.len[j]++;
            if (k + 1 == nd || offs[k + 1] != offs[k] + 1) tot += 255 + w.len[j];
        }
        w.win = tot;
    }
    d->v32_exact = exact32;
    dia_box_detect(offs.data(), nd, n, d->box);
    try {
        const int ns = dia_slots<double>(nd);
        d->v64 = dalloc<double>((size_t)ns * n);
        const int grid = (int)std::min<int64_t>(((int64_t)ns * n + 255) / 256, 8192);
        hipLaunchKernelGGL(dia_group_kernel, dim3(grid), dim3(256), 0, s, n, nd, ns, d->args, tmp.planes, d->v64);
        RAMD_HIP(hipGetLastError());
        RAMD_HIP(hipStreamSynchronize(s));
        d->mask = tmp.mask;
        tmp.mask = nullptr;
    } catch (...) {
        dia_free(d);
        throw;
    }
    A->dia = d;
    A->forms++;
    return true;
}

bool host_diag_first(int n_rows, const int* rp, const int* col)
{
    for (int i = 0; i < n_rows; i++)
        for (int j = rp[i] + 1; j < rp[i + 1]; j++)
            if (col[j] == i) return false;
    return true;
}

// ---------------------------------------------------------------------------
// Sliced-row form
// ---------------------------------------------------------------------------
// entries of row r, in stored order, to lane r % 64 of slice r / 64; the
// padding keeps what the fill wrote
template <typename T>
__global__ void sliced_scatter_kernel(int n, const int* __restrict__ rp, const int64_t* __restrict__ sptr,
                                      const T* __restrict__ in, T* __restrict__ out)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const int64_t base = sptr[r >> 6] + (r & 63);
    const int s = rp[r], e = rp[r + 1];
    for (int k = s; k < e; k++) out[base + (int64_t)64 * (k - s)] = in[k];
}

// 16-bit offsets col - r of row r's entries, in the sliced positions
__global__ void sliced16_scatter_kernel(int n, const int* __restrict__ rp, const int64_t* __restrict__ sptr,
                                        const int* __restrict__ col, int16_t* __restrict__ out)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const int64_t base = sptr[r >> 6] + (r & 63);
    const int s = rp[r], e = rp[r + 1];
    for (int k = s; k < e; k++) out[base + (int64_t)64 * (k - s)] = (int16_t)(col[k] - r);
}

__global__ void fill_i16_kernel(int16_t* __restrict__ p, int64_t n, int16_t v)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

// largest |col - row| over the matrix, into *out (atomic maximum)
__global__ void col_offset_range_kernel(int n, const int* __restrict__ rp, const int* __restrict__ col,
                                        int* __restrict__ out)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    int m = 0;
    for (int k = rp[r]; k < rp[r + 1]; k++) {
        const int d = col[k] - r;
        m = max(m, d < 0 ? -d : d);
    }
    if (m) atomicMax(out, m);
}

void sliced_free(ramd_csr_s* A)
{
    if (!A->sl) return;
    dfree(A->sl->sptr);
    dfree(A->sl->col);
    dfree(A->sl->col16);
    dfree(A->sl->v64);
    dfree(A->sl->v32);
    delete A->sl;
    A->sl = nullptr;
    A->sl_on[0] = A->sl_on[1] = false;
    A->sl16[0] = A->sl16[1] = false;
    A->forms++;
}

void sliced_build(ramd_csr_s* A)
{
    if (A->sl) return;
    const int n = A->n_rows;
    const int nq = (n + 63) / 64;
    const int* rp = A->h_rowptr.data();
    std::vector<int64_t> sptr((size_t)nq + 1, 0);
    for (int q = 0; q < nq; q++) {
        int w = 0;
        for (int r = 64 * q; r < std::min(n, 64 * q + 64); r++) w = std::max(w, rp[r + 1] - rp[r]);
        sptr[q + 1] = sptr[q] + (int64_t)64 * w;
    }
    hipStream_t s = A->ctx->stream;
    Sliced* S = new Sliced;
    A->forms++;
    S->n_slices = nq;
    S->entries = sptr[nq];
    A->sl = S;
    S->sptr = dalloc<int64_t>(sptr.size());
    S->col = dalloc<int>((size_t)std::max<int64_t>(S->entries, 1));
    RAMD_HIP(hipMemcpyAsync(S->sptr, sptr.data(), sizeof(int64_t) * sptr.size(), hipMemcpyHostToDevice, s));
    const int g = (int)std::max<int64_t>(1, std::min<int64_t>((S->entries + 255) / 256, 8192));
    hipLaunchKernelGGL(fill_int_kernel, dim3(g), dim3(256), 0, s, S->col, S->entries, -1);
    if (n)
        hipLaunchKernelGGL((sliced_scatter_kernel<int>), dim3((n + 255) / 256), dim3(256), 0, s, n, A->rowptr,
                           S->sptr, A->col, S->col);
    RAMD_HIP(hipGetLastError());
    // 16-bit column offsets when every |col - row| fits
    if (n && A->n_rows == A->n_cols) {
        int* range = dalloc<int>(1);
        int m = 0;
        try {
            hipLaunchKernelGGL(fill_int_kernel, dim3(1), dim3(64), 0, s, range, (int64_t)1, 0);
            hipLaunchKernelGGL(col_offset_range_kernel, dim3((n + 255) / 256), dim3(256), 0, s, n, A->rowptr, A->col,
                               range);
            RAMD_HIP(hipGetLastError());
            RAMD_HIP(hipMemcpyAsync(&m, range, sizeof(int), hipMemcpyDeviceToHost, s));
            RAMD_HIP(hipStreamSynchronize(s));
        } catch (...) {
            dfree(range);
            throw;
        }
        dfree(range);
        if (m <= 32767) {
            S->col16 = dalloc<int16_t>((size_t)std::max<int64_t>(S->entries, 1));
            hipLaunchKernelGGL(fill_i16_kernel, dim3(g), dim3(256), 0, s, S->col16, S->entries, kSliced16Pad);
            hipLaunchKernelGGL(sliced16_scatter_kernel, dim3((n + 255) / 256), dim3(256), 0, s, n, A->rowptr, S->sptr,
                               A->col, S->col16);
            RAMD_HIP(hipGetLastError());
        }
    }
    RAMD_HIP(hipStreamSynchronize(s));  // sptr's host copy goes out of scope
}

template <typename T>
static void sliced_copy(ramd_csr_s* A, const T* in, T* out)
{
    hipStream_t s = A->ctx->stream;
    const Sliced* S = A->sl;
    const int g = (int)std::max<int64_t>(1, std::min<int64_t>((S->entries + 255) / 256, 8192));
    hipLaunchKernelGGL(fill_zero_kernel<T>, dim3(g), dim3(256), 0, s, out, S->entries);
    if (A->n_rows)
        hipLaunchKernelGGL((sliced_scatter_kernel<T>), dim3((A->n_rows + 255) / 256), dim3(256), 0, s, A->n_rows,
                           A->rowptr, S->sptr, in, out);
    RAMD_HIP(hipGetLastError());
}

void sliced_values(ramd_csr_s* A, int prec)
{
    RAMD_CHECK(A->sl, RAMD_E_STATE, "sliced form not built");
    Sliced* S = A->sl;
    const size_t m = (size_t)std::max<int64_t>(S->entries, 1);
    A->forms++;
    if (prec == 0) {
        if (!S->v64) S->v64 = dalloc<double>(m);
        sliced_copy<double>(A, A->val, S->v64);
    } else {
        RAMD_CHECK(A->val32, RAMD_E_STATE, "FP32 values of the matrix were not built");
        if (!S->v32) S->v32 = dalloc<float>(m);
        sliced_copy<float>(A, A->val32, S->v32);
    }
}

// One workgroup per row block: the block's (column, position) keys sorted
// in LDS (bitonic, up to kSortedGatherMax entries), then the columns,
// positions and values written in that order.  A block of one row longer
// than the cap is copied in stored order (its kernel path reads it so).
constexpr int kSortedGatherMax = 8192;
template <typename V>
__global__ __launch_bounds__(1024) void sg_build_kernel(const int* __restrict__ rowptr, const int* __restrict__ blocks,
                                                        int cap, const int* __restrict__ col,
                                                        const V* __restrict__ val, int* __restrict__ scol,
                                                        uint16_t* __restrict__ spos, V* __restrict__ sval)
{
    __shared__ unsigned long long key[kSortedGatherMax];
    const int s = rowptr[blocks[blockIdx.x]], e = rowptr[blocks[blockIdx.x + 1]];
    const int len = e - s;
    const int tid = threadIdx.x;
    if (len > cap) {
        for (int j = tid; j < len; j += 1024) {
            scol[s + j] = col[s + j];
            spos[s + j] = 0;
            sval[s + j] = val[s + j];
        }
        return;
    }
    int P = 1;
    while (P < len) P <<= 1;
    for (int j = tid; j < P; j += 1024)
        key[j] = j < len ? ((unsigned long long)(unsigned)col[s + j] << 16) | (unsigned)j : ~0ull;
    __syncthreads();
    for (int k = 2; k <= P; k <<= 1)
        for (int h = k >> 1; h > 0; h >>= 1) {
            for (int i = tid; i < P; i += 1024) {
                const int l = i ^ h;
                if (l > i) {
                    const unsigned long long a = key[i], c = key[l];
                    if ((a > c) == ((i & k) == 0)) {
                        key[i] = c;
                        key[l] = a;
                    }
                }
            }
            __syncthreads();
        }
    for (int j = tid; j < len; j += 1024) {
        const unsigned long long kv = key[j];
        const int p = (int)(kv & 0xFFFFull);
        scol[s + j] = (int)(kv >> 16);
        spos[s + j] = (uint16_t)p;
        sval[s + j] = val[s + p];
    }
}

void sg_free(ramd_csr_s* A, int prec)
{
    SortedGather*& g = A->sg[prec];
    if (!g) return;
    RAMD_HIP(hipStreamSynchronize(A->ctx->stream));
    dfree(g->col);
    dfree(g->pos);
    dfree(g->val);
    delete g;
    g = nullptr;
    A->forms++;
}

void sg_build(ramd_csr_s* A, int prec)
{
    const ramd_csr_s::RowBlocks& rb = A->rb[prec];
    RAMD_CHECK(rb.cap >= 1 && rb.cap <= kSortedGatherMax, RAMD_E_STATE, "sorted gather: row-block cap %d", rb.cap);
    RAMD_CHECK(prec == 0 || A->val32, RAMD_E_STATE, "sorted gather: FP32 values were not built");
    sg_free(A, prec);
    const size_t nz = (size_t)std::max<int64_t>(A->nnz, 1);
    SortedGather* g = new SortedGather();
    try {
        g->cap = rb.cap;
        g->blocks = rb.d;
        g->col = dalloc<int>(nz);
        g->pos = dalloc<uint16_t>(nz);
        g->val = dmalloc(nz * (prec ? sizeof(float) : sizeof(double)));
        if (rb.n > 0) {
            hipStream_t s = A->ctx->stream;
            if (prec)
                hipLaunchKernelGGL(sg_build_kernel<float>, dim3(rb.n), dim3(1024), 0, s, A->rowptr, rb.d, rb.cap,
                                   A->col, A->val32, g->col, g->pos, (float*)g->val);
            else
                hipLaunchKernelGGL(sg_build_kernel<double>, dim3(rb.n), dim3(1024), 0, s, A->rowptr, rb.d, rb.cap,
                                   A->col, A->val, g->col, g->pos, (double*)g->val);
            RAMD_HIP(hipGetLastError());
        }
    } catch (...) {
        dfree(g->col);
        dfree(g->pos);
        dfree(g->val);
        delete g;
        throw;
    }
    A->sg[prec] = g;
    A->forms++;
}

void csr_free(ramd_csr_s* A)
{
    if (!A) return;
    sg_free(A, 0);
    sg_free(A, 1);
    sliced_free(A);
    if (A->trans) csr_free(A->trans);
    dfree(A->rowptr);
    dfree(A->col);
    dfree(A->val);
    for (auto& rb : A->rb)
        if (rb.d != A->blocks) dfree(rb.d);
    dfree(A->blocks);
    dfree(A->val32);
    dia_free(A->dia);
    delete A;
}

// Counting-sort transpose: entries of column c are emitted in ascending
// source-row order, the order CSR_to_CSC (core/matrix.cpp:491-538) produces
// and the order CSR_append_T (util/linalg/spmv.cpp:150-163) adds in.
// Transpose kernels: column counts, a scatter of (row, value) to each
// column's segment by atomic cursors, then every segment put back in row
// order (the order of a counting sort over rows ascending, which is the
// order CSR_to_CSC gives), one thread per column.
__global__ void transpose_count_kernel(int64_t nnz, const int* __restrict__ col, int* __restrict__ cnt)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < nnz; j += stride) atomicAdd(&cnt[col[j]], 1);
}

__global__ void transpose_scatter_kernel(int n, const int* __restrict__ rp, const int* __restrict__ col,
                                         const double* __restrict__ val, int* __restrict__ cursor,
                                         int* __restrict__ trow, double* __restrict__ tval)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int j = rp[i]; j < rp[i + 1]; j++) {
        const int d = atomicAdd(&cursor[col[j]], 1);
        trow[d] = i;
        tval[d] = val[j];
    }
}

__global__ void transpose_sort_kernel(int m, const int* __restrict__ tp, int* __restrict__ trow,
                                      double* __restrict__ tval)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m) return;
    const int s = tp[c], e = tp[c + 1];
    for (int a = s + 1; a < e; a++) {  // rows of one column are distinct
        const int r = trow[a];
        const double v = tval[a];
        int b = a - 1;
        while (b >= s && trow[b] > r) {
            trow[b + 1] = trow[b];
            tval[b + 1] = tval[b];
            b--;
        }
        trow[b + 1] = r;
        tval[b + 1] = v;
    }
}

// ---------------------------------------------------------------------------
// Row pointers from row counts on the device: an exclusive prefix sum in
// int64, 4096 counts per workgroup (per-workgroup sums, one workgroup scans
// the sums, every workgroup then scans its counts from its offset).
// ---------------------------------------------------------------------------
constexpr int kScanPer = 16;
constexpr int kScanItems = kBlockThreads * kScanPer;

__device__ __forceinline__ int64_t block_exclusive_scan64(int64_t v, int64_t* lds, int64_t* total)
{
    const int t = threadIdx.x;
    lds[t] = v;
    __syncthreads();
    for (int off = 1; off < kBlockThreads; off <<= 1) {
        const int64_t a = t >= off ? lds[t - off] : 0;
        __syncthreads();
        lds[t] += a;
        __syncthreads();
    }
    const int64_t incl = lds[t];
    if (total) *total = lds[kBlockThreads - 1];
    __syncthreads();
    return incl - v;
}

__global__ __launch_bounds__(kBlockThreads) void scan_sums_kernel(const int* __restrict__ cnt, int n,
                                                                  int64_t* __restrict__ bsum)
{
    __shared__ int64_t lds[kBlockThreads];
    const int64_t base = (int64_t)blockIdx.x * kScanItems;
    int64_t s = 0;
    for (int j = 0; j < kScanPer; j++) {
        const int64_t i = base + j * kBlockThreads + threadIdx.x;
        if (i < n) s += cnt[i];
    }
    int64_t tot;
    block_exclusive_scan64(s, lds, &tot);
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

// in place: bsum[b] <- sum of bsum[0..b); total <- sum of all
__global__ __launch_bounds__(kBlockThreads) void scan_offsets_kernel(int64_t* __restrict__ bsum, int nb,
                                                                     int64_t* __restrict__ total)
{
    __shared__ int64_t lds[kBlockThreads];
    int64_t carry = 0;
    for (int c0 = 0; c0 < nb; c0 += kBlockThreads) {
        const int b = c0 + threadIdx.x;
        const int64_t v = b < nb ? bsum[b] : 0;
        int64_t tot;
        const int64_t ex = block_exclusive_scan64(v, lds, &tot);
        if (b < nb) bsum[b] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) *total = carry;
}

__global__ __launch_bounds__(kBlockThreads) void scan_write_kernel(const int* __restrict__ cnt, int n,
                                                                   const int64_t* __restrict__ boff,
                                                                   const int64_t* __restrict__ total,
                                                                   int* __restrict__ rp)
{
    __shared__ int vals[kScanItems];
    __shared__ int64_t lds[kBlockThreads];
    const int64_t base = (int64_t)blockIdx.x * kScanItems;
    for (int j = 0; j < kScanPer; j++) {
        const int k = j * kBlockThreads + threadIdx.x;
        vals[k] = base + k < n ? cnt[base + k] : 0;
    }
    __syncthreads();
    int64_t s = 0;
    for (int j = 0; j < kScanPer; j++) s += vals[threadIdx.x * kScanPer + j];
    int64_t run = boff[blockIdx.x] + block_exclusive_scan64(s, lds, nullptr);
    // this thread's counts -> their row pointers, in LDS, then written out
    int* own = vals + threadIdx.x * kScanPer;
    for (int j = 0; j < kScanPer; j++) {
        const int c = own[j];
        own[j] = (int)run;
        run += c;
    }
    __syncthreads();
    for (int j = 0; j < kScanPer; j++) {
        const int k = j * kBlockThreads + threadIdx.x;
        if (base + k < n) rp[base + k] = vals[k];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) rp[n] = (int)*total;
}

int64_t device_rowptr(ramd_ctx ctx, const int* cnt, int n, int* rp, const char* what)
{
    hipStream_t s = ctx->stream;
    const int nb = std::max(1, (int)(((int64_t)n + kScanItems - 1) / kScanItems));
    int64_t* buf = dalloc<int64_t>((size_t)nb + 1);
    int64_t total = 0;
    try {
        hipLaunchKernelGGL(scan_sums_kernel, dim3(nb), dim3(kBlockThreads), 0, s, cnt, n, buf);
        hipLaunchKernelGGL(scan_offsets_kernel, dim3(1), dim3(kBlockThreads), 0, s, buf, nb, buf + nb);
        hipLaunchKernelGGL(scan_write_kernel, dim3(nb), dim3(kBlockThreads), 0, s, cnt, n, buf, buf + nb, rp);
        RAMD_HIP(hipGetLastError());
        RAMD_HIP(hipMemcpyAsync(&total, buf + nb, sizeof(int64_t), hipMemcpyDeviceToHost, s));
        RAMD_HIP(hipStreamSynchronize(s));
    } catch (...) {
        dfree(buf);
        throw;
    }
    dfree(buf);
    RAMD_CHECK(total <= INT32_MAX, RAMD_E_ARG, "%s exceeds 2^31 stored entries", what);
    return total;
}

// longest column the per-thread insertion sort takes (longer: host sort)
constexpr int kTransposeMaxCol = 2048;

ramd_csr_s* csr_transpose(ramd_csr_s* A)
{
    if (A->trans) return A->trans;
    ramd_ctx ctx = A->ctx;
    hipStream_t s = ctx->stream;
    const int n = A->n_rows, m = A->n_cols;
    const int64_t nnz = A->nnz;
    // column counts -> host exclusive sums (the handle keeps them on the host)
    std::vector<int> tp((size_t)m + 1, 0);
    int* d_rp = dalloc<int>((size_t)m + 1);
    int* d_col = nullptr;
    double* d_val = nullptr;
    try {
        int* cnt = d_rp + 1;
        hipLaunchKernelGGL(fill_int_kernel, dim3((unsigned)std::min<int64_t>((m + 256) / 256, 8192)), dim3(256), 0,
                           s, d_rp, (int64_t)m + 1, 0);
        if (nnz)
            hipLaunchKernelGGL(transpose_count_kernel, dim3((unsigned)std::min<int64_t>((nnz + 255) / 256, 8192)),
                               dim3(256), 0, s, nnz, A->col, cnt);
        RAMD_HIP(hipGetLastError());
        RAMD_HIP(hipMemcpyAsync(tp.data(), d_rp, sizeof(int) * (m + 1), hipMemcpyDeviceToHost, s));
        RAMD_HIP(hipStreamSynchronize(s));
        int longest = 0;
        for (int c = 0; c < m; c++) {
            longest = std::max(longest, tp[c + 1]);
            tp[c + 1] += tp[c];
        }
        d_col = dalloc<int>((size_t)std::max<int64_t>(nnz, 1));
        d_val = dalloc<double>((size_t)std::max<int64_t>(nnz, 1));
        RAMD_HIP(hipMemcpyAsync(d_rp, tp.data(), sizeof(int) * (m + 1), hipMemcpyHostToDevice, s));
        if (nnz && longest <= kTransposeMaxCol) {
            int* cursor = dalloc<int>((size_t)std::max(m, 1));
            RAMD_HIP(hipMemcpyAsync(cursor, d_rp, sizeof(int) * m, hipMemcpyDeviceToDevice, s));
            hipLaunchKernelGGL(transpose_scatter_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n,
                               A->rowptr, A->col, A->val, cursor, d_col, d_val);
            hipLaunchKernelGGL(transpose_sort_kernel, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, m, d_rp,
                               d_col, d_val);
            RAMD_HIP(hipGetLastError());
            RAMD_HIP(hipStreamSynchronize(s));
            dfree(cursor);
        } else if (nnz) {
            // a long column: the counting sort on the host
            std::vector<int> col((size_t)nnz), trow((size_t)nnz);
            std::vector<double> val((size_t)nnz), tval((size_t)nnz);
            RAMD_HIP(hipMemcpyAsync(col.data(), A->col, sizeof(int) * nnz, hipMemcpyDeviceToHost, s));
            RAMD_HIP(hipMemcpyAsync(val.data(), A->val, sizeof(double) * nnz, hipMemcpyDeviceToHost, s));
            RAMD_HIP(hipStreamSynchronize(s));
            std::vector<int> pos(tp.begin(), tp.end() - 1);
            const int* rp = A->h_rowptr.data();
            for (int i = 0; i < n; i++)
                for (int j = rp[i]; j < rp[i + 1]; j++) {
                    const int d = pos[col[j]]++;
                    trow[d] = i;
                    tval[d] = val[j];
                }
            RAMD_HIP(hipMemcpyAsync(d_col, trow.data(), sizeof(int) * nnz, hipMemcpyHostToDevice, s));
            RAMD_HIP(hipMemcpyAsync(d_val, tval.data(), sizeof(double) * nnz, hipMemcpyHostToDevice, s));
            RAMD_HIP(hipStreamSynchronize(s));
        }
    } catch (...) {
        dfree(d_rp);
        dfree(d_col);
        dfree(d_val);
        throw;
    }
    A->trans = csr_wrap(ctx, m, n, nnz, d_rp, d_col, d_val, tp.data());
    return A->trans;
}

__global__ void to_fp32_kernel(const double* __restrict__ in, float* __restrict__ out, int64_t n)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) out[i] = (float)in[i];
}

void csr_make_fp32(ramd_csr_s* A)
{
    if (A->val32) return;
    A->val32 = dalloc<float>((size_t)A->nnz);
    A->forms++;
    if (A->nnz) {
        int grid = (int)std::min<int64_t>((A->nnz + 255) / 256, 8192);
        hipLaunchKernelGGL(to_fp32_kernel, dim3(grid), dim3(256), 0, A->ctx->stream, A->val, A->val32, A->nnz);
        RAMD_HIP(hipGetLastError());
    }
    if (A->dia && !A->dia->v32) dia_make_fp32(A);
}

// Rewrites the stencil-diagonal planes from new CSR values (same pattern):
// entry (row, col) lands on the plane of offset col - row.
__global__ void dia_scatter_kernel(int n, DiaArgs a, int nd, const int* __restrict__ rp, const int* __restrict__ col,
                                   const double* __restrict__ val, double* __restrict__ v64)
{
    const int row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= n) return;
    for (int j = rp[row]; j < rp[row + 1]; j++) {
        const int d = col[j] - row;
        int k = 0;
        while (k < nd - 1 && a.off[k] != d) k++;
        v64[dia_index<double>(a.pos[k], row, n)] = val[j];
    }
}

// per-workgroup count of values that FP32 does not hold exactly
__global__ __launch_bounds__(kBlockThreads) void inexact32_kernel(const double* __restrict__ v, int64_t m,
                                                                  double* __restrict__ partials)
{
    double c = 0;
    const int64_t stride = (int64_t)gridDim.x * kBlockThreads;
    for (int64_t i = (int64_t)blockIdx.x * kBlockThreads + threadIdx.x; i < m; i += stride)
        c += ((double)(float)v[i] != v[i]) ? 1.0 : 0.0;
    c = block_sum(c);
    if (threadIdx.x == 0) partials[blockIdx.x] = c;
}

void csr_set_values(ramd_csr_s* A, const double* val_d)
{
    ramd_ctx ctx = A->ctx;
    hipStream_t s = ctx->stream;
    if (A->nnz) RAMD_HIP(hipMemcpyAsync(A->val, val_d, sizeof(double) * A->nnz, hipMemcpyDeviceToDevice, s));
    if (A->trans) {  // stale values: rebuilt on next use
        RAMD_HIP(hipStreamSynchronize(s));
        csr_free(A->trans);
        A->trans = nullptr;
    }
    if (A->val32 && A->nnz) {
        int grid = (int)std::min<int64_t>((A->nnz + 255) / 256, 8192);
        hipLaunchKernelGGL(to_fp32_kernel, dim3(grid), dim3(256), 0, s, A->val, A->val32, A->nnz);
        RAMD_HIP(hipGetLastError());
    }
    if (A->sl) {
        if (A->sl->v64) sliced_values(A, 0);
        if (A->sl->v32) sliced_values(A, 1);
    }
    for (int p = 0; p < 2; p++)  // the sorted values follow (same order)
        if (A->sg[p]) sg_build(A, p);
    if (Dia* d = A->dia) {
        const int n = A->n_rows;
        hipLaunchKernelGGL(dia_scatter_kernel, dim3((n + 255) / 256), dim3(256), 0, s, n, d->args, d->nd, A->rowptr,
                           A->col, A->val, d->v64);
        RAMD_HIP(hipGetLastError());
        const int64_t m = (int64_t)dia_slots<double>(d->nd) * n;  // padding slots hold exact zeros
        const int g = (int)std::max<int64_t>(1, std::min<int64_t>(kRedBlocks, (m + 255) / 256));
        double* part = ctx_partials(ctx, (size_t)g);
        hipLaunchKernelGGL(inexact32_kernel, dim3(g), dim3(kBlockThreads), 0, s, d->v64, m, part);
        RAMD_HIP(hipGetLastError());
        launch_reduce_final(ctx, part, g, ctx->red_result + 8);
        double inexact = 0;
        fetch_scalars(ctx, ctx->red_result + 8, 1, &inexact);
        // the FP32-value operator is chosen only while it is exact
        if (d->v32_exact != (inexact == 0.0)) A->forms++;
        d->v32_exact = inexact == 0.0;
        if (d->v32) dia_make_fp32(A);
    }
    RAMD_HIP(hipStreamSynchronize(s));
}

}  // namespace ramd

void ramd_ctx_s::set_current() const { RAMD_HIP(hipSetDevice(device)); }

using namespace ramd;

extern "C" {

int ramd_last_error(char* buf, size_t len)
{
    if (!buf || len == 0) return RAMD_E_ARG;
    snprintf(buf, len, "%s", g_last_error.c_str());
    return RAMD_OK;
}

int ramd_init(int device, ramd_ctx* out)
{
    return guarded([&] {
        RAMD_CHECK(out, RAMD_E_ARG, "ramd_init: out is NULL");
        int count = 0;
        hipError_t e = hipGetDeviceCount(&count);
        if (e != hipSuccess || count == 0) {
            (void)hipGetLastError();
            fail(RAMD_E_DEVICE, "ramd_init: no HIP device available (%s)",
                 e != hipSuccess ? hipGetErrorString(e) : "0 devices");
        }
        RAMD_CHECK(device >= 0 && device < count, RAMD_E_ARG, "ramd_init: device %d out of range (%d devices)",
                   device, count);
        hipDeviceProp_t prop;
        RAMD_HIP(hipGetDeviceProperties(&prop, device));
        RAMD_CHECK(std::strncmp(prop.gcnArchName, "gfx950", 6) == 0, RAMD_E_DEVICE,
                   "ramd_init: device %d is %s; this library is built for gfx950 only", device,
                   prop.gcnArchName);
        ramd_ctx c = new ramd_ctx_s();
        c->device = device;
        RAMD_HIP(hipSetDevice(device));
        RAMD_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        noalloc_scratch_setup(device);
        c->red_partials = dalloc<double>(4096);
        c->red_result = dalloc<double>(16);
        RAMD_HIP(hipHostMalloc((void**)&c->pinned, 16 * sizeof(double), hipHostMallocDefault));
        RAMD_HIP(hipHostMalloc((void**)&c->step_pin, 8 * sizeof(double), hipHostMallocMapped | hipHostMallocCoherent));
        RAMD_HIP(hipHostGetDevicePointer((void**)&c->step_pin_d, c->step_pin, 0));
        RAMD_HIP(hipEventCreateWithFlags(&c->step_ev, hipEventDisableTiming));
        RAMD_HIP(hipStreamSynchronize(c->stream));
        *out = c;
    });
}

int ramd_finalize(ramd_ctx ctx)
{
    return guarded([&] {
        RAMD_CHECK(ctx, RAMD_E_ARG, "ramd_finalize: NULL context");
        ctx->set_current();
        RAMD_HIP(hipStreamSynchronize(ctx->stream));
        for (const ProfPending& p : ctx->prof_pending) {
            (void)hipEventDestroy(p.beg);
            (void)hipEventDestroy(p.end);
        }
        for (hipEvent_t e : ctx->prof_pool) (void)hipEventDestroy(e);
        dfree(ctx->red_partials);
        dfree(ctx->red_result);
        dfree(ctx->big_partials);
        dfree(ctx->flush_buf);
        if (ctx->pinned) (void)hipHostFree(ctx->pinned);
        if (ctx->step_pin) (void)hipHostFree(ctx->step_pin);
        if (ctx->step_ev) (void)hipEventDestroy(ctx->step_ev);
        (void)hipStreamDestroy(ctx->stream);
        delete ctx;
    });
}

int ramd_sync(ramd_ctx ctx)
{
    return guarded([&] {
        RAMD_CHECK(ctx, RAMD_E_ARG, "ramd_sync: NULL context");
        RAMD_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int ramd_malloc(ramd_ctx ctx, size_t bytes, void** out_d)
{
    return guarded([&] {
        RAMD_CHECK(ctx && out_d, RAMD_E_ARG, "ramd_malloc: NULL argument");
        ctx->set_current();
        *out_d = dmalloc(bytes);
    });
}

int ramd_free(ramd_ctx ctx, void* p_d)
{
    return guarded([&] {
        RAMD_CHECK(ctx, RAMD_E_ARG, "ramd_free: NULL context");
        RAMD_HIP(hipStreamSynchronize(ctx->stream));
        dfree(p_d);
    });
}

int ramd_memcpy_h2d(ramd_ctx ctx, void* dst_d, const void* src, size_t bytes)
{
    return guarded([&] {
        RAMD_CHECK(ctx && (bytes == 0 || (dst_d && src)), RAMD_E_ARG, "ramd_memcpy_h2d: NULL argument");
        if (bytes) RAMD_HIP(hipMemcpyAsync(dst_d, src, bytes, hipMemcpyHostToDevice, ctx->stream));
        RAMD_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int ramd_memcpy_d2h(ramd_ctx ctx, void* dst, const void* src_d, size_t bytes)
{
    return guarded([&] {
        RAMD_CHECK(ctx && (bytes == 0 || (dst && src_d)), RAMD_E_ARG, "ramd_memcpy_d2h: NULL argument");
        if (bytes) RAMD_HIP(hipMemcpyAsync(dst, src_d, bytes, hipMemcpyDeviceToHost, ctx->stream));
        RAMD_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int ramd_memcpy_d2d(ramd_ctx ctx, void* dst_d, const void* src_d, size_t bytes)
{
    return guarded([&] {
        RAMD_CHECK(ctx && (bytes == 0 || (dst_d && src_d)), RAMD_E_ARG, "ramd_memcpy_d2d: NULL argument");
        if (bytes) RAMD_HIP(hipMemcpyAsync(dst_d, src_d, bytes, hipMemcpyDeviceToDevice, ctx->stream));
    });
}

int ramd_memset_zero(ramd_ctx ctx, void* p_d, size_t bytes)
{
    return guarded([&] {
        RAMD_CHECK(ctx && (bytes == 0 || p_d), RAMD_E_ARG, "ramd_memset_zero: NULL argument");
        if (bytes) RAMD_HIP(hipMemsetAsync(p_d, 0, bytes, ctx->stream));
    });
}

int ramd_csr_create(ramd_ctx ctx, int32_t n_rows, int32_t n_cols, int64_t nnz, const int32_t* rowptr,
                    const int32_t* col, const double* val, ramd_csr* out)
{
    return guarded([&] {
        RAMD_CHECK(ctx && out && rowptr, RAMD_E_ARG, "ramd_csr_create: NULL argument");
        RAMD_CHECK(n_rows >= 0 && n_cols >= 0 && nnz >= 0 && nnz <= INT32_MAX, RAMD_E_ARG,
                   "ramd_csr_create: bad sizes n_rows=%d n_cols=%d nnz=%lld", n_rows, n_cols, (long long)nnz);
        RAMD_CHECK(nnz == 0 || (col && val), RAMD_E_ARG, "ramd_csr_create: NULL col/val");
        RAMD_CHECK(rowptr[0] == 0 && rowptr[n_rows] == nnz, RAMD_E_ARG,
                   "ramd_csr_create: rowptr[0]=%d rowptr[n]=%d, expected 0 and %lld", rowptr[0],
                   rowptr[n_rows], (long long)nnz);
        for (int i = 0; i < n_rows; i++)
            RAMD_CHECK(rowptr[i + 1] >= rowptr[i], RAMD_E_ARG, "ramd_csr_create: rowptr decreases at row %d", i);
        for (int64_t j = 0; j < nnz; j++)
            RAMD_CHECK(col[j] >= 0 && col[j] < n_cols, RAMD_E_ARG,
                       "ramd_csr_create: column %d of entry %lld outside [0,%d)", col[j], (long long)j, n_cols);
        ctx->set_current();
        int* d_rp = dalloc<int>((size_t)n_rows + 1);
        int* d_col = dalloc<int>((size_t)nnz);
        double* d_val = dalloc<double>((size_t)nnz);
        RAMD_HIP(hipMemcpyAsync(d_rp, rowptr, sizeof(int) * (n_rows + 1), hipMemcpyHostToDevice, ctx->stream));
        if (nnz) {
            RAMD_HIP(hipMemcpyAsync(d_col, col, sizeof(int) * nnz, hipMemcpyHostToDevice, ctx->stream));
            RAMD_HIP(hipMemcpyAsync(d_val, val, sizeof(double) * nnz, hipMemcpyHostToDevice, ctx->stream));
        }
        *out = csr_wrap(ctx, n_rows, n_cols, nnz, d_rp, d_col, d_val, rowptr);
        (*out)->diag_first = host_diag_first(n_rows, rowptr, col);
        dia_build(*out);
    });
}

int ramd_csr_info(ramd_csr A, int32_t* n_rows, int32_t* n_cols, int64_t* nnz)
{
    return guarded([&] {
        RAMD_CHECK(A, RAMD_E_ARG, "ramd_csr_info: NULL handle");
        if (n_rows) *n_rows = A->n_rows;
        if (n_cols) *n_cols = A->n_cols;
        if (nnz) *nnz = A->nnz;
    });
}

int ramd_csr_download(ramd_csr A, int32_t* rowptr, int32_t* col, double* val)
{
    return guarded([&] {
        RAMD_CHECK(A, RAMD_E_ARG, "ramd_csr_download: NULL handle");
        hipStream_t s = A->ctx->stream;
        if (rowptr)
            RAMD_HIP(hipMemcpyAsync(rowptr, A->rowptr, sizeof(int) * (A->n_rows + 1), hipMemcpyDeviceToHost, s));
        if (col && A->nnz) RAMD_HIP(hipMemcpyAsync(col, A->col, sizeof(int) * A->nnz, hipMemcpyDeviceToHost, s));
        if (val && A->nnz) RAMD_HIP(hipMemcpyAsync(val, A->val, sizeof(double) * A->nnz, hipMemcpyDeviceToHost, s));
        RAMD_HIP(hipStreamSynchronize(s));
    });
}

int ramd_csr_set_values(ramd_csr A, const double* val_d)
{
    return guarded([&] {
        RAMD_CHECK(A && (val_d || A->nnz == 0), RAMD_E_ARG, "ramd_csr_set_values: NULL argument");
        RAMD_CHECK(!A->borrowed, RAMD_E_STATE, "ramd_csr_set_values: handle is owned by a hierarchy");
        A->ctx->set_current();
        csr_set_values(A, val_d);
    });
}

int ramd_csr_set_dia_staging(ramd_csr A, int pp, int mode, int planes)
{
    return guarded([&] {
        RAMD_CHECK(A && pp >= 0 && pp < 3 && mode >= -1 && mode <= 2, RAMD_E_ARG,
                   "ramd_csr_set_dia_staging: bad argument");
        RAMD_CHECK(A->dia, RAMD_E_STATE, "ramd_csr_set_dia_staging: matrix has no stencil-diagonal form");
        RAMD_CHECK(mode != 1 || A->dia->window.win <= kDiaWin, RAMD_E_ARG,
                   "ramd_csr_set_dia_staging: x window of %d values exceeds %d", A->dia->window.win, kDiaWin);
        RAMD_CHECK(mode != 2 || A->dia->box.L > 0, RAMD_E_ARG,
                   "ramd_csr_set_dia_staging: the offsets have no box geometry");
        RAMD_CHECK(planes == 0 || (mode == 2 && planes >= kBoxMinPlanes && planes <= kBoxMaxPlanes), RAMD_E_ARG,
                   "ramd_csr_set_dia_staging: %d planes per workgroup (%d..%d, plane-march only)", planes,
                   kBoxMinPlanes, kBoxMaxPlanes);
        A->dia->lds_x[pp] = mode;
        if (planes) A->dia->box.zc[pp] = planes;
        A->forms++;
    });
}

int ramd_csr_dia_staging(ramd_csr A, int pp, int* mode, int* planes, int* L, int* P)
{
    return guarded([&] {
        RAMD_CHECK(A && pp >= 0 && pp < 3 && mode && planes && L && P, RAMD_E_ARG,
                   "ramd_csr_dia_staging: bad argument");
        RAMD_CHECK(A->dia, RAMD_E_STATE, "ramd_csr_dia_staging: matrix has no stencil-diagonal form");
        *mode = A->dia->lds_x[pp];
        *planes = A->dia->box.zc[pp];
        *L = A->dia->box.L;
        *P = A->dia->box.P;
    });
}

int ramd_csr_set_sliced(ramd_csr A, int prec, int on)
{
    return guarded([&] {
        RAMD_CHECK(A && (prec == 0 || prec == 1) && on >= 0 && on <= 2, RAMD_E_ARG,
                   "ramd_csr_set_sliced: bad argument");
        RAMD_CHECK(!A->dia, RAMD_E_STATE, "ramd_csr_set_sliced: matrix uses its stencil-diagonal form");
        A->ctx->set_current();
        if (on) {
            RAMD_CHECK(prec == 0 || A->val32, RAMD_E_STATE, "ramd_csr_set_sliced: FP32 values were not built");
            RAMD_CHECK(A->nnz > 0, RAMD_E_ARG, "ramd_csr_set_sliced: matrix has no entries");
            sliced_build(A);
            sliced_values(A, prec);
            RAMD_HIP(hipStreamSynchronize(A->ctx->stream));
        }
        RAMD_CHECK(on != 2 || A->sl->col16, RAMD_E_ARG,
                   "ramd_csr_set_sliced: the column offsets do not fit 16 bits (or the matrix is not square)");
        A->sl_on[prec] = on != 0;
        A->sl16[prec] = on == 2;
        if (on) A->lr_on[prec] = A->wr_on[prec] = false;
        A->forms++;
    });
}

int ramd_csr_set_lane_rows(ramd_csr A, int prec, int on)
{
    return guarded([&] {
        RAMD_CHECK(A && (prec == 0 || prec == 1) && (on == 0 || on == 1), RAMD_E_ARG,
                   "ramd_csr_set_lane_rows: bad argument");
        RAMD_CHECK(!A->dia, RAMD_E_STATE, "ramd_csr_set_lane_rows: matrix uses its stencil-diagonal form");
        RAMD_CHECK(prec == 0 || A->val32 || !on, RAMD_E_STATE, "ramd_csr_set_lane_rows: FP32 values were not built");
        A->lr_on[prec] = on != 0;
        if (on) A->sl_on[prec] = A->sl16[prec] = A->wr_on[prec] = false;
        A->forms++;
    });
}

int ramd_csr_set_fp32_outer(ramd_csr A, int on)
{
    return guarded([&] {
        RAMD_CHECK(A && (on == 0 || on == 1), RAMD_E_ARG, "ramd_csr_set_fp32_outer: bad argument");
        A->ctx->set_current();
        RAMD_HIP(hipStreamSynchronize(A->ctx->stream));
        A->v32_outer = on != 0;
        A->forms++;
    });
}

int ramd_csr_set_sorted_gather(ramd_csr A, int prec, int on)
{
    return guarded([&] {
        RAMD_CHECK(A && (prec == 0 || prec == 1) && (on == 0 || on == 1), RAMD_E_ARG,
                   "ramd_csr_set_sorted_gather: bad argument");
        RAMD_CHECK(!A->dia, RAMD_E_STATE, "ramd_csr_set_sorted_gather: matrix uses its stencil-diagonal form");
        A->ctx->set_current();
        if (on) {
            A->sl_on[prec] = A->sl16[prec] = A->lr_on[prec] = A->wr_on[prec] = false;
            sg_build(A, prec);
        } else {
            sg_free(A, prec);
        }
        A->forms++;
    });
}

int ramd_csr_set_wave_rows(ramd_csr A, int prec, int on)
{
    return guarded([&] {
        RAMD_CHECK(A && (prec == 0 || prec == 1) && (on == 0 || on == 1), RAMD_E_ARG,
                   "ramd_csr_set_wave_rows: bad argument");
        RAMD_CHECK(!A->dia, RAMD_E_STATE, "ramd_csr_set_wave_rows: matrix uses its stencil-diagonal form");
        RAMD_CHECK(prec == 0 || A->val32 || !on, RAMD_E_STATE, "ramd_csr_set_wave_rows: FP32 values were not built");
        A->wr_on[prec] = on != 0;
        if (on) A->sl_on[prec] = A->sl16[prec] = A->lr_on[prec] = false;
        A->forms++;
    });
}

int ramd_csr_diagonals(ramd_csr A, int* nd)
{
    return guarded([&] {
        RAMD_CHECK(A && nd, RAMD_E_ARG, "ramd_csr_diagonals: NULL argument");
        *nd = A->dia ? A->dia->nd : 0;
    });
}

int ramd_csr_destroy(ramd_csr A)
{
    return guarded([&] {
        if (!A) return;
        RAMD_CHECK(!A->borrowed, RAMD_E_STATE, "ramd_csr_destroy: handle is owned by a hierarchy");
        (void)hipStreamSynchronize(A->ctx->stream);
        csr_free(A);
    });
}

}  // extern "C"

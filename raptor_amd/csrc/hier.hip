// Device-resident AMG hierarchy: setup driver (host strength / splitting /
// interpolation, device Galerkin products), the V-cycle in FP64 and FP32,
// the coarse dense LU solve, stationary AMG, PCG with the reference's rules
// and the mixed-precision flexible PCG.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>

#include "common.hpp"
#include "host/setup.hpp"

#include "hier.hpp"

namespace ramd {

using host::Csr;

static double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

int EventSpans::open(hipStream_t s, int slot)
{
    hipEvent_t ev[2];
    for (hipEvent_t& e : ev) {
        if (!pool.empty()) {
            e = pool.back();
            pool.pop_back();
        } else {
            RAMD_HIP(hipEventCreate(&e));
        }
    }
    RAMD_HIP(hipEventRecord(ev[0], s));
    pending.push_back(Span{ev[0], ev[1], slot});
    return (int)pending.size() - 1;
}

void EventSpans::close(hipStream_t s, int idx) { RAMD_HIP(hipEventRecord(pending[(size_t)idx].e, s)); }

void EventSpans::resolve(std::vector<double>& acc)
{
    if (pending.empty()) return;
    RAMD_HIP(hipEventSynchronize(pending.back().e));
    for (const Span& sp : pending) {
        float ms = 0;
        RAMD_HIP(hipEventElapsedTime(&ms, sp.b, sp.e));
        if (sp.slot >= 0 && (size_t)sp.slot < acc.size()) acc[(size_t)sp.slot] += 1e-3 * ms;
        pool.push_back(sp.b);
        pool.push_back(sp.e);
    }
    pending.clear();
}

void EventSpans::release()
{
    for (const Span& sp : pending) {
        (void)hipEventDestroy(sp.b);
        (void)hipEventDestroy(sp.e);
    }
    for (hipEvent_t e : pool) (void)hipEventDestroy(e);
    pending.clear();
    pool.clear();
}

// Per-level solve timing: a span on the hierarchy's stream attributed to
// level l while tracking is on (-1: not tracking).  Long runs are resolved
// as they go so the pending list stays short.
static int level_span_open(ramd_hier_s* h, int l)
{
    if (!h->track) return -1;
    if (h->spans.pending.size() >= 4096) h->spans.resolve(h->solve_level_s);
    return h->spans.open(h->ctx->stream, l);
}

static void level_span_close(ramd_hier_s* h, int idx)
{
    if (idx >= 0) h->spans.close(h->ctx->stream, idx);
}

// dia = false skips the stencil-diagonal form (matrices that are only
// multiplied by other sparse matrices).
static ramd_csr_s* upload(ramd_ctx ctx, const Csr& M, bool dia = true)
{
    int* rp = dalloc<int>((size_t)M.n_rows + 1);
    int* c = dalloc<int>((size_t)M.nnz());
    double* v = dalloc<double>((size_t)M.nnz());
    RAMD_HIP(hipMemcpyAsync(rp, M.rowptr.data(), sizeof(int) * (M.n_rows + 1), hipMemcpyHostToDevice, ctx->stream));
    if (M.nnz()) {
        RAMD_HIP(hipMemcpyAsync(c, M.col.data(), sizeof(int) * M.nnz(), hipMemcpyHostToDevice, ctx->stream));
        RAMD_HIP(hipMemcpyAsync(v, M.val.data(), sizeof(double) * M.nnz(), hipMemcpyHostToDevice, ctx->stream));
    }
    ramd_csr_s* D = csr_wrap(ctx, M.n_rows, M.n_cols, M.nnz(), rp, c, v, M.rowptr.data());
    D->diag_first = host_diag_first(M.n_rows, M.rowptr.data(), M.col.data());
    if (dia) dia_build(D);
    return D;
}

static Csr download(ramd_csr_s* A)
{
    Csr M;
    M.n_rows = A->n_rows;
    M.n_cols = A->n_cols;
    M.rowptr = A->h_rowptr;
    M.col.resize((size_t)A->nnz);
    M.val.resize((size_t)A->nnz);
    if (A->nnz) {
        RAMD_HIP(hipMemcpyAsync(M.col.data(), A->col, sizeof(int) * A->nnz, hipMemcpyDeviceToHost, A->ctx->stream));
        RAMD_HIP(hipMemcpyAsync(M.val.data(), A->val, sizeof(double) * A->nnz, hipMemcpyDeviceToHost, A->ctx->stream));
    }
    RAMD_HIP(hipStreamSynchronize(A->ctx->stream));
    return M;
}

// Dense LU with partial pivoting of the row-major coarse matrix read in
// column-major order, i.e. LAPACK dgetrf_ applied as
// multilevel/par_multilevel.hpp:295-331 does.  Returns LAPACK's info.
static int lu_factor_colmajor(int n, std::vector<double>& a, std::vector<int>& piv)
{
    int info = 0;
    piv.assign((size_t)n, 0);
    for (int j = 0; j < n; j++) {
        int p = j;
        double amax = std::fabs(a[j + (size_t)j * n]);
        for (int i = j + 1; i < n; i++) {
            const double t = std::fabs(a[i + (size_t)j * n]);
            if (t > amax) {
                amax = t;
                p = i;
            }
        }
        piv[j] = p;
        if (a[p + (size_t)j * n] != 0.0) {
            if (p != j)
                for (int k = 0; k < n; k++) std::swap(a[j + (size_t)k * n], a[p + (size_t)k * n]);
            const double r = 1.0 / a[j + (size_t)j * n];
            for (int i = j + 1; i < n; i++) a[i + (size_t)j * n] *= r;
        } else if (info == 0) {
            info = j + 1;
        }
        for (int k = j + 1; k < n; k++) {
            const double u = a[j + (size_t)k * n];
            for (int i = j + 1; i < n; i++) a[i + (size_t)k * n] -= a[i + (size_t)j * n] * u;
        }
    }
    return info;
}

constexpr int kCoarseMax = 8192;

// dgetrs_('N') on one right-hand side by one workgroup: row swaps, unit-lower
// then upper substitution in column order (the reference BLAS dtrsm order).
__global__ __launch_bounds__(256) void coarse_solve_kernel(int n, const double* __restrict__ lu,
                                                           const int* __restrict__ piv, const double* __restrict__ b,
                                                           double* __restrict__ x)
{
    extern __shared__ double sb[];
    for (int i = threadIdx.x; i < n; i += blockDim.x) sb[i] = b[i];
    __syncthreads();
    if (threadIdx.x == 0)
        for (int j = 0; j < n; j++) {
            const int p = piv[j];
            if (p != j) {
                const double t = sb[j];
                sb[j] = sb[p];
                sb[p] = t;
            }
        }
    __syncthreads();
    for (int j = 0; j < n; j++) {
        const double bj = sb[j];
        for (int i = j + 1 + threadIdx.x; i < n; i += blockDim.x) sb[i] -= bj * lu[i + (size_t)j * n];
        __syncthreads();
    }
    for (int j = n - 1; j >= 0; j--) {
        if (threadIdx.x == 0) sb[j] /= lu[j + (size_t)j * n];
        __syncthreads();
        const double bj = sb[j];
        for (int i = threadIdx.x; i < j; i += blockDim.x) sb[i] -= bj * lu[i + (size_t)j * n];
        __syncthreads();
    }
    for (int i = threadIdx.x; i < n; i += blockDim.x) x[i] = sb[i];
}

__global__ void extract_diag(int n, const int* __restrict__ rp, const int* __restrict__ col,
                             const double* __restrict__ val, double* __restrict__ d)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double v = 0.0;
    for (int j = rp[i]; j < rp[i + 1]; j++)
        if (col[j] == i) {
            v = val[j];
            break;
        }
    d[i] = v;
}

// First Jacobi sweep from x == 0 (relax.cpp:40 with tmp = 0 and the
// off-diagonal sum of zeros equal to +0): x = (1-w)*0 + w*((b - 0)/d).
// Rows without a usable diagonal (|d| <= 1e-16, or empty rows) keep x = 0.
template <typename X>
__global__ void jacobi_from_zero(int n, const X* __restrict__ d, const X* __restrict__ b, X* __restrict__ x, X omega)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    // d[i] is 0 for a row without a stored diagonal (empty rows included)
    x[i] = jacobi_zero_row(b[i], d[i], omega);
}

template <typename X>
__global__ void add_into(int n, const X* __restrict__ xin, const X* __restrict__ inc, X* __restrict__ y)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = xin[i] + inc[i];
}

__global__ void to_f32(int64_t n, const double* __restrict__ a, float* __restrict__ b)
{
    const int64_t s = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += s) b[i] = (float)a[i];
}

__global__ void to_f64(int64_t n, const float* __restrict__ a, double* __restrict__ b)
{
    const int64_t s = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += s) b[i] = (double)a[i];
}

// ---------------------------------------------------------------------------
// V-cycle
// ---------------------------------------------------------------------------
template <typename TB, typename TX>
__global__ __launch_bounds__(64) void coarse_wave_kernel(int n, const double* __restrict__ lu,
                                                         const int* __restrict__ piv, const TB* __restrict__ b,
                                                         TX* __restrict__ x)
{
    extern __shared__ double ls[];
    coarse_solve_wave<TB, TX>(n, lu, piv, b, x, 1, ls);
}

// ---------------------------------------------------------------------------
// Coarse tail: the restriction into a run of small levels, their whole
// V-cycle (smoothing, residuals, restrictions, the coarse LU solve,
// interpolations) and the interpolation back out, by one workgroup in one
// launch instead of about five launches per level.  Each operation runs
// over its rows with a workgroup barrier after it; a row is computed by one
// thread with the operations of stream_kernel's phase 2 on the CSR arrays
// (products rounded, then added left to right in stored order; Jacobi's
// diagonal rules), so every value is the one the per-level kernels produce.
// ---------------------------------------------------------------------------
constexpr int kTailMax = 8;          // levels in the tail, the coarsest included
constexpr int kTailThreads = 1024;
constexpr int kTailMaxRows = 2048;   // a tail level's rows
constexpr int64_t kTailMaxNnz = 16384;   // a tail level's A (and P) entries
constexpr int kTailMaxOuterRows = 8192;  // rows of the level above the tail
constexpr int64_t kTailMaxOuterNnz = 65536;  // entries of its P

template <typename X>
struct TailLevel {
    int n = 0;
    int jfz = 0;  // the restriction into this level forms its first Jacobi sweep from zero
    int adf = 0;  // A stores every row's diagonal first
    const int *arp = nullptr, *acol = nullptr;
    const X* aval = nullptr;
    const X* diag = nullptr;
    const int *prp = nullptr, *pcol = nullptr;  // P: this level to the next
    const X* pval = nullptr;
    const int *rrp = nullptr, *rcol = nullptr;  // R = P^T
    const X* rval = nullptr;
    X *x = nullptr, *b = nullptr, *t = nullptr, *r = nullptr;
};
template <typename X>
struct TailArgs {
    int nl = 0;  // levels lv[0 .. nl), lv[nl - 1] the coarsest
    int nb = 0;  // rows of the level above the tail
    const int *bprp = nullptr, *bpcol = nullptr;  // its P and R
    const X* bpval = nullptr;
    const int *brrp = nullptr, *brcol = nullptr;
    const X* brval = nullptr;
    const X* rb = nullptr;    // its residual
    const X* curb = nullptr;  // its iterate
    X* outb = nullptr;        // receives curb + P x_tail
    int coarse_n = 0;
    const double* lu = nullptr;
    const int* piv = nullptr;
    X omega = 0;
    int sweeps = 0;
    TailLevel<X> lv[kTailMax];
};

// One operation over rows [0, n) of a CSR matrix (n <= kTailThreads *
// kTailRowsPerThread): OP_SPMV, OP_SPMV_JFZ (y2 = the first Jacobi sweep
// from zero on y, diagonal dvec), OP_RESIDUAL, OP_ADD_INTO (y = b + A x) and
// OP_JACOBI (x old iterate, y new; df: diagonal stored first, dvec its
// values).  The workgroup forms the products of a window of the matrix's
// entries together (independent loads) into prod (LDS, kTailLds entries),
// then each row's thread adds its entries of the window in stored order,
// carrying its sum to the next window: stream_kernel's operations.
constexpr int kTailRowsPerThread = 8;
constexpr int kTailLdsBytes = 65536;
static_assert(kCoarseWave * kCoarseWave * (int)sizeof(double) <= kTailLdsBytes, "coarse LU must fit the tail's LDS");
static_assert(kTailMaxOuterRows <= kTailThreads * kTailRowsPerThread && kTailMaxRows <= kTailMaxOuterRows,
              "tail rows per thread");

template <int OP, typename X>
__device__ void tail_op(X* prod, int n, const int* __restrict__ rp, const int* __restrict__ col,
                        const X* __restrict__ val, const X* x, const X* b, X* y, X omega, const X* dvec, X* y2,
                        int df)
{
    constexpr int CAP = kTailLdsBytes / (int)sizeof(X);
    constexpr int RT = kTailRowsPerThread;
    const int tid = threadIdx.x;
    const int e0 = rp[0], E = rp[n] - e0;
    int lo[RT], hi[RT];
    X acc[RT], diag[RT];
#pragma unroll
    for (int u = 0; u < RT; u++) {
        const int row = tid + u * kTailThreads;
        lo[u] = hi[u] = 0;
        acc[u] = diag[u] = X(0);
        if (row < n) {
            const int rs = rp[row] - e0, re = rp[row + 1] - e0;
            lo[u] = rs;
            hi[u] = re;
            if (OP == OP_RESIDUAL) acc[u] = b[row];
            if (OP == OP_JACOBI && df && rs < re) {
                // the diagonal (first) entry is not added; a row without a
                // usable one keeps its value
                bool use;
                if (dvec) {
                    const X dg = dvec[row];
                    use = fabs(dg) > X(1e-16);
                    if (use) diag[u] = dg;
                } else {
                    use = col[e0 + rs] == row;
                    if (use) diag[u] = val[e0 + rs];
                }
                lo[u] = use ? rs + 1 : re;
            }
        }
    }
    for (int c0 = 0; c0 < E; c0 += CAP) {
        const int cn = min(CAP, E - c0);
        __syncthreads();  // the previous window's products are read
        // eight entries per thread and step, every load issued before the products
        constexpr int U = 8;
        for (int j0 = tid; j0 < cn; j0 += U * kTailThreads) {
            int cc[U];
            X vv[U], xx[U];
#pragma unroll
            for (int v = 0; v < U; v++) {
                const int j = j0 + v * kTailThreads;
                if (j < cn) {
                    cc[v] = col[e0 + c0 + j];
                    vv[v] = val[e0 + c0 + j];
                }
            }
#pragma unroll
            for (int v = 0; v < U; v++)
                if (j0 + v * kTailThreads < cn) xx[v] = x[cc[v]];
#pragma unroll
            for (int v = 0; v < U; v++)
                if (j0 + v * kTailThreads < cn) prod[j0 + v * kTailThreads] = vv[v] * xx[v];
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < RT; u++) {
            const int k0 = max(lo[u], c0), k1 = min(hi[u], c0 + cn);
            const int row = tid + u * kTailThreads;
            if (OP == OP_JACOBI && !df) {
                for (int k = k0; k < k1; k++) {
                    if (col[e0 + k] == row) diag[u] = val[e0 + k];
                    else acc[u] = acc[u] + prod[k - c0];
                }
            } else {
                // eight LDS reads in flight, the additions in order
                int k = k0;
                for (; k + 8 <= k1; k += 8) {
                    X t[8];
#pragma unroll
                    for (int v = 0; v < 8; v++) t[v] = prod[k - c0 + v];
#pragma unroll
                    for (int v = 0; v < 8; v++) acc[u] = (OP == OP_RESIDUAL) ? acc[u] - t[v] : acc[u] + t[v];
                }
                for (; k < k1; k++) acc[u] = (OP == OP_RESIDUAL) ? acc[u] - prod[k - c0] : acc[u] + prod[k - c0];
            }
        }
    }
#pragma unroll
    for (int u = 0; u < RT; u++) {
        const int row = tid + u * kTailThreads;
        if (row >= n) continue;
        if (OP == OP_JACOBI) {
            const X xo = x[row];
            y[row] = (fabs(diag[u]) > X(1e-16)) ? ((X(1.0) - omega) * xo) + (omega * ((b[row] - acc[u]) / diag[u]))
                                                 : xo;
        } else if (OP == OP_ADD_INTO) {
            y[row] = b[row] + acc[u];
        } else {
            y[row] = acc[u];
            if (OP == OP_SPMV_JFZ) y2[row] = jacobi_zero_row(acc[u], dvec[row], omega);
        }
    }
}

// The cycle of cycle64 / cycle32 from the restriction into lv[0] to the
// interpolation out of it, entered with every tail level's x = 0.
template <typename X>
__global__ __launch_bounds__(kTailThreads) void tail_cycle_kernel(TailArgs<X> a)
{
    // LDS: each operation's products, and the coarse LU for the solve
    __shared__ double lds[kTailLdsBytes / sizeof(double)];
    X* prod = reinterpret_cast<X*>(lds);
    double* ls = lds;
    {
        const TailLevel<X>& C = a.lv[0];
        if (C.jfz)
            tail_op<OP_SPMV_JFZ, X>(prod, C.n, a.brrp, a.brcol, a.brval, a.rb, nullptr, C.b, a.omega, C.diag, C.x, 0);
        else tail_op<OP_SPMV, X>(prod, C.n, a.brrp, a.brcol, a.brval, a.rb, nullptr, C.b, X(0), nullptr, nullptr, 0);
        __syncthreads();
    }
    unsigned in_t = 0;  // bit j: level j's iterate is in t after its pre-smoothing
    for (int j = 0; j + 1 < a.nl; j++) {
        const TailLevel<X>& L = a.lv[j];
        X* cur = L.x;
        X* other = L.t;
        if (a.sweeps <= 0) {  // no sweep to start from zero: x = 0
            for (int i = threadIdx.x; i < L.n; i += kTailThreads) cur[i] = X(0);
            __syncthreads();
        }
        for (int it = 0; it < a.sweeps; it++) {
            if (it == 0 && L.jfz) continue;
            if (it == 0) {
                for (int i = threadIdx.x; i < L.n; i += kTailThreads)
                    cur[i] = jacobi_zero_row(L.b[i], L.diag[i], a.omega);
            } else {
                tail_op<OP_JACOBI, X>(prod, L.n, L.arp, L.acol, L.aval, cur, L.b, other, a.omega, L.diag, nullptr,
                                      L.adf);
                X* tmp = cur;
                cur = other;
                other = tmp;
            }
            __syncthreads();
        }
        if (cur != L.x) in_t |= 1u << j;
        tail_op<OP_RESIDUAL, X>(prod, L.n, L.arp, L.acol, L.aval, cur, L.b, L.r, X(0), nullptr, nullptr, 0);
        __syncthreads();
        const TailLevel<X>& C = a.lv[j + 1];
        if (C.jfz)
            tail_op<OP_SPMV_JFZ, X>(prod, C.n, L.rrp, L.rcol, L.rval, L.r, nullptr, C.b, a.omega, C.diag, C.x, 0);
        else tail_op<OP_SPMV, X>(prod, C.n, L.rrp, L.rcol, L.rval, L.r, nullptr, C.b, X(0), nullptr, nullptr, 0);
        __syncthreads();
    }
    {  // the coarsest level: the wavefront LU solve (FP64, rounded once on output)
        const TailLevel<X>& C = a.lv[a.nl - 1];
        const int n = a.coarse_n;
        for (int t = threadIdx.x; t < n * n; t += kTailThreads) ls[t] = a.lu[t];
        __syncthreads();
        if (threadIdx.x < 64) {
            const int lane = threadIdx.x;
            const int pv = lane < n ? a.piv[lane] : 0;
            double v = lane < n ? (double)C.b[lane] : 0.0;
            v = coarse_lu_lane(n, ls, pv, v, lane);
            if (lane < n) C.x[lane] = (X)v;
        }
        __syncthreads();
    }
    for (int j = a.nl - 2; j >= 0; j--) {
        const TailLevel<X>& L = a.lv[j];
        const TailLevel<X>& C = a.lv[j + 1];
        X* cur = ((in_t >> j) & 1u) ? L.t : L.x;
        X* other = ((in_t >> j) & 1u) ? L.x : L.t;
        tail_op<OP_ADD_INTO, X>(prod, L.n, L.prp, L.pcol, L.pval, C.x, cur, other, X(0), nullptr, nullptr, 0);
        X* tmp = cur;
        cur = other;
        other = tmp;
        __syncthreads();
        for (int it = 0; it < a.sweeps; it++) {
            tail_op<OP_JACOBI, X>(prod, L.n, L.arp, L.acol, L.aval, cur, L.b, other, a.omega, L.diag, nullptr,
                                  L.adf);
            tmp = cur;
            cur = other;
            other = tmp;
            __syncthreads();
        }
        if (cur != L.x) {
            for (int i = threadIdx.x; i < L.n; i += kTailThreads) L.x[i] = cur[i];
            __syncthreads();
        }
    }
    tail_op<OP_ADD_INTO, X>(prod, a.nb, a.bprp, a.bpcol, a.bpval, a.lv[0].x, a.curb, a.outb, X(0), nullptr, nullptr, 0);
}

static bool multicolour(const ramd_hier_s* h);
static bool jfz_fused(const ramd_hier_s* h, int l);

// Levels that can start the coarse tail, the largest tail first (none with
// per-level timing on, multicolour smoothing or a coarse system past one
// wavefront).  Tail levels hold at most kTailMaxRows rows and kTailMaxNnz
// entries of A and of P; the level above them at most kTailMaxOuterRows
// rows and kTailMaxOuterNnz entries of P.
static std::vector<int> tail_candidates(const ramd_hier_s* h)
{
    std::vector<int> c;
    const int nlev = (int)h->levels.size();
    if (h->track || multicolour(h) || nlev < 2 || h->coarse_n > kCoarseWave) return c;
    auto small = [&](int l) {
        const Level& L = h->levels[l];
        return L.n <= kTailMaxRows && (!L.A || L.A->nnz <= kTailMaxNnz) && (!L.P || L.P->nnz <= kTailMaxNnz);
    };
    if (!small(nlev - 1)) return c;
    int l0 = nlev - 1;
    while (l0 - 1 >= 1 && nlev - (l0 - 1) <= kTailMax && small(l0 - 1)) l0--;
    for (int l = l0; l < nlev; l++) {
        const Level& B = h->levels[l - 1];
        if (B.n <= kTailMaxOuterRows && B.P->nnz <= kTailMaxOuterNnz) c.push_back(l);
    }
    return c;
}

// First level of the coarse tail for a cycle in precision prec (0 FP64,
// 1 FP32), or -1: the forced level, or in automatic mode the timed choice
// (none before it is made).
static int tail_level(const ramd_hier_s* h, int prec)
{
    if (h->tail_force != -2) return h->tail_force;
    if (h->track || multicolour(h) || h->coarse_n > kCoarseWave || h->tail_mode < 0) return -1;
    if (h->tail_mode > 0) return h->tail_mode;
    return h->tail_pick[prec] >= 0 ? h->tail_pick[prec] : -1;
}

template <typename X>
__global__ void fill_value(int n, X v, X* __restrict__ x)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = v;
}

static void cycle64(ramd_hier_s* h, int l, double* x, const double* b, bool x_zero, bool first_done);
static void cycle32(ramd_hier_s* h, int l, float* x, const float* b, bool x_zero, bool first_done);

// Automatic mode: the first cycle in each precision times whole V-cycles
// without a tail and with every candidate start, back to back and in turn
// (seven rounds: the coarse levels stay cached between cycles as they do in
// a solve), and keeps the candidate with the smallest median (same bits for
// all).  Uses the fine level's own x and b vectors.
void cycle_prepare(ramd_hier_s* h, int prec)
{
    if (prec == 1) enable_fp32(h);
    if (h->tail_mode != 0 || h->tail_pick[prec] != -2 || h->track) return;
    std::vector<int> cand = tail_candidates(h);
    if (cand.empty()) {
        h->tail_pick[prec] = -1;
        return;
    }
    cand.insert(cand.begin(), -1);
    Level& L = h->levels[0];
    hipStream_t s = h->ctx->stream;
    std::function<void()> run;
    if (prec == 1) {
        hipLaunchKernelGGL((fill_value<float>), grid1(L.n), dim3(256), 0, s, L.n, 1.0f, L.b32);
        run = [&] { cycle32(h, 0, L.x32, L.b32, true, false); };
    } else {
        hipLaunchKernelGGL((fill_value<double>), grid1(L.n), dim3(256), 0, s, L.n, 1.0, L.b);
        run = [&] { cycle64(h, 0, L.x, L.b, true, false); };
    }
    RAMD_HIP(hipGetLastError());
    constexpr int kRounds = 7;
    const int nc = (int)cand.size();
    std::vector<hipEvent_t> ev(2 * (size_t)nc * kRounds, nullptr);
    std::vector<float> ms(ev.size() / 2, 0.f);
    try {
        for (hipEvent_t& e : ev) RAMD_HIP(hipEventCreate(&e));
        for (int i = 0; i < nc; i++) {  // first launches: code loading
            h->tail_force = cand[i];
            run();
        }
        for (int r = 0; r < kRounds; r++)
            for (int i = 0; i < nc; i++) {
                const size_t k = (size_t)r * nc + i;
                h->tail_force = cand[i];
                RAMD_HIP(hipEventRecord(ev[2 * k], s));
                run();
                RAMD_HIP(hipEventRecord(ev[2 * k + 1], s));
            }
        h->tail_force = -2;
        RAMD_HIP(hipEventSynchronize(ev.back()));
        for (size_t k = 0; k < ms.size(); k++) RAMD_HIP(hipEventElapsedTime(&ms[k], ev[2 * k], ev[2 * k + 1]));
    } catch (...) {
        h->tail_force = -2;
        for (hipEvent_t e : ev)
            if (e) (void)hipEventDestroy(e);
        throw;
    }
    for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    int best = -1;
    float best_ms = 0;
    for (int i = 0; i < nc; i++) {
        std::vector<float> t(kRounds);
        for (int r = 0; r < kRounds; r++) t[r] = ms[(size_t)r * nc + i];
        std::nth_element(t.begin(), t.begin() + kRounds / 2, t.end());
        if (i == 0 || t[kRounds / 2] < best_ms) {
            best_ms = t[kRounds / 2];
            best = cand[i];
        }
    }
    h->tail_pick[prec] = best;
}

template <typename X>
static const X* vals_of(const ramd_csr_s* M)
{
    if constexpr (sizeof(X) == 8) {
        return M->val;
    } else {
        RAMD_CHECK(M->val32, RAMD_E_STATE, "coarse tail: FP32 values missing");
        return M->val32;
    }
}

// Launches the tail below level lb (restriction of r, the tail levels'
// cycle, out = cur + P_lb x).
template <typename X>
static void launch_tail(ramd_hier_s* h, int lb, const X* r, const X* cur, X* out)
{
    const int nlev = (int)h->levels.size();
    TailArgs<X> a;
    a.nl = nlev - (lb + 1);
    RAMD_CHECK(a.nl >= 1 && a.nl <= kTailMax, RAMD_E_STATE, "coarse tail: %d levels", a.nl);
    const Level& B = h->levels[lb];
    a.nb = B.n;
    a.bprp = B.P->rowptr;
    a.bpcol = B.P->col;
    a.bpval = vals_of<X>(B.P);
    a.brrp = B.R->rowptr;
    a.brcol = B.R->col;
    a.brval = vals_of<X>(B.R);
    a.rb = r;
    a.curb = cur;
    a.outb = out;
    a.coarse_n = h->coarse_n;
    a.lu = h->lu;
    a.piv = h->piv;
    a.omega = (X)h->opts.relax_weight;
    a.sweeps = h->opts.num_smooth_sweeps;
    for (int j = 0; j < a.nl; j++) {
        const int l = lb + 1 + j;
        Level& L = h->levels[l];
        TailLevel<X>& t = a.lv[j];
        t.n = L.n;
        t.jfz = (int)jfz_fused(h, l);
        if constexpr (sizeof(X) == 8) {
            t.x = L.x;
            t.b = L.b;
            t.t = L.t;
            t.r = L.r;
            t.diag = L.diag;
        } else {
            t.x = L.x32;
            t.b = L.b32;
            t.t = L.t32;
            t.r = L.r32;
            t.diag = L.diag32;
        }
        if (l == nlev - 1) continue;
        t.adf = (int)L.A->diag_first;
        t.arp = L.A->rowptr;
        t.acol = L.A->col;
        t.aval = vals_of<X>(L.A);
        t.prp = L.P->rowptr;
        t.pcol = L.P->col;
        t.pval = vals_of<X>(L.P);
        t.rrp = L.R->rowptr;
        t.rcol = L.R->col;
        t.rval = vals_of<X>(L.R);
    }
    RAMD_CHECK(a.lv[a.nl - 1].n == a.coarse_n, RAMD_E_STATE, "coarse tail: coarsest level has %d rows, LU %d",
               a.lv[a.nl - 1].n, a.coarse_n);
    hipLaunchKernelGGL((tail_cycle_kernel<X>), dim3(1), dim3(kTailThreads), 0, h->ctx->stream, a);
    RAMD_HIP(hipGetLastError());
}

static void coarse_solve64(ramd_hier_s* h, const double* b, double* x)
{
    const int n = h->coarse_n;
    if (n <= kCoarseWave)
        hipLaunchKernelGGL((coarse_wave_kernel<double, double>), dim3(1), dim3(64), sizeof(double) * n * n,
                           h->ctx->stream, n, h->lu, h->piv, b, x);
    else
        hipLaunchKernelGGL(coarse_solve_kernel, dim3(1), dim3(256), sizeof(double) * n, h->ctx->stream, n, h->lu,
                           h->piv, b, x);
    RAMD_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Multicolour smoothers (RAMD_MC_SOR / RAMD_MC_SSOR): colours in increasing
// order (then decreasing for the symmetric form), one thread per row of the
// colour with the reference's per-row update on the current x --
// SOR_forward's x_i = (x_i + w (y_i - x_i - sum)) / a_ii going up,
// SOR_backward's x_i = (1 - w) x_i + w ((y_i - sum) / a_ii) going down
// (util/linalg/par_relax.cpp:29-104), the off-diagonal sum in stored order.
// Rows of one colour are never neighbours, so they update independently.
// A row whose first entry is not its diagonal keeps x, as there.
// ---------------------------------------------------------------------------
template <typename V, typename X, bool BACK>
__global__ void mc_sor_kernel(const int* __restrict__ rows, int count, const int* __restrict__ rp,
                              const int* __restrict__ col, const V* __restrict__ val, X* __restrict__ x,
                              const X* __restrict__ y, X omega)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    const int i = rows[t];
    const int s = rp[i], e = rp[i + 1];
    if (s == e || col[s] != i) return;
    const X diag = X(val[s]);
    X row_sum = 0;
    for (int j = s + 1; j < e; j++) row_sum += X(val[j]) * x[col[j]];
    if (BACK) x[i] = ((X(1.0) - omega) * x[i]) + (omega * ((y[i] - row_sum) / diag));
    else x[i] = (x[i] + omega * (y[i] - x[i] - row_sum)) / diag;
}

template <typename X>
__global__ void fill_zero(int n, X* __restrict__ x)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = X(0);
}

static bool multicolour(const ramd_hier_s* h)
{
    return h->opts.relax_type == RAMD_MC_SOR || h->opts.relax_type == RAMD_MC_SSOR;
}

// colouring of every level but the coarsest (built once per hierarchy)
static void build_colours(ramd_hier_s* h)
{
    ramd_ctx ctx = h->ctx;
    for (size_t l = 0; l + 1 < h->levels.size(); l++) {
        Level& L = h->levels[l];
        if (L.color_rows) continue;
        const ramd_csr_s* A = L.A;
        std::vector<int> col((size_t)A->nnz);
        if (A->nnz) {
            RAMD_HIP(hipMemcpyAsync(col.data(), A->col, sizeof(int) * A->nnz, hipMemcpyDeviceToHost, ctx->stream));
            RAMD_HIP(hipStreamSynchronize(ctx->stream));
        }
        std::vector<int> colour;
        const int nc = host::greedy_colors(A->n_rows, A->h_rowptr.data(), col.data(), colour);
        L.color_ptr.assign((size_t)nc + 1, 0);
        for (int c : colour) L.color_ptr[c + 1]++;
        for (int c = 0; c < nc; c++) L.color_ptr[c + 1] += L.color_ptr[c];
        std::vector<int> rows((size_t)A->n_rows), pos(L.color_ptr.begin(), L.color_ptr.end() - 1);
        for (int i = 0; i < A->n_rows; i++) rows[pos[colour[i]]++] = i;
        L.color_rows = dalloc<int>((size_t)std::max(A->n_rows, 1));
        RAMD_HIP(hipMemcpyAsync(L.color_rows, rows.data(), sizeof(int) * rows.size(), hipMemcpyHostToDevice,
                                ctx->stream));
        RAMD_HIP(hipStreamSynchronize(ctx->stream));
    }
}

// num_smooth_sweeps multicolour sweeps in place on x (zeroed first when
// the cycle enters with x = 0)
template <typename V, typename X>
static void mc_relax(ramd_hier_s* h, const Level& L, const V* val, X* x, const X* b, bool x_zero)
{
    hipStream_t s = h->ctx->stream;
    const X w = (X)h->opts.relax_weight;
    if (x_zero) hipLaunchKernelGGL((fill_zero<X>), grid1(L.n), dim3(256), 0, s, L.n, x);
    const int nc = (int)L.color_ptr.size() - 1;
    const bool sym = h->opts.relax_type == RAMD_MC_SSOR;
    for (int it = 0; it < h->opts.num_smooth_sweeps; it++)
        for (int pass = 0; pass < (sym ? 2 : 1); pass++)
            for (int k = 0; k < nc; k++) {
                const int c = pass == 0 ? k : nc - 1 - k;
                const int cnt = L.color_ptr[c + 1] - L.color_ptr[c];
                if (cnt == 0) continue;
                const int* rows = L.color_rows + L.color_ptr[c];
                if (pass == 0)
                    hipLaunchKernelGGL((mc_sor_kernel<V, X, false>), grid1(cnt), dim3(256), 0, s, rows, cnt,
                                       L.A->rowptr, L.A->col, val, x, b, w);
                else
                    hipLaunchKernelGGL((mc_sor_kernel<V, X, true>), grid1(cnt), dim3(256), 0, s, rows, cnt,
                                       L.A->rowptr, L.A->col, val, x, b, w);
            }
    RAMD_HIP(hipGetLastError());
}

// Whether the restriction into level l also forms level l's first Jacobi
// sweep from zero (a smoothed level, not the coarsest, Jacobi smoothing).
static bool jfz_fused(const ramd_hier_s* h, int l)
{
    return l < (int)h->levels.size() - 1 && !multicolour(h) && h->opts.num_smooth_sweeps >= 1;
}

// The off-diagonal Jacobi kernel reads `cur` and writes the other buffer.
// first_done: x already holds the first Jacobi sweep from zero (written by
// the restriction that formed b, with jacobi_from_zero's operations).
static void cycle64(ramd_hier_s* h, int l, double* x, const double* b, bool x_zero, bool first_done = false)
{
    Level& L = h->levels[l];
    hipStream_t s = h->ctx->stream;
    int tk = level_span_open(h, l);
    if (l == (int)h->levels.size() - 1) {
        coarse_solve64(h, b, x);
        level_span_close(h, tk);
        return;
    }
    const double w = h->opts.relax_weight;
    const int sweeps = h->opts.num_smooth_sweeps;
    const bool mc = multicolour(h);
    double* cur = x;
    double* other = L.t;
    // no sweep to start from zero: x = 0 as the reference sets it
    // (multilevel/par_multilevel.hpp:389)
    if (x_zero && !mc && sweeps <= 0) hipLaunchKernelGGL((fill_zero<double>), grid1(L.n), dim3(256), 0, s, L.n, x);
    if (mc) mc_relax<double, double>(h, L, L.A->val, cur, b, x_zero);
    for (int it = 0; it < sweeps && !mc; it++) {
        if (it == 0 && first_done) {
            continue;
        } else if (it == 0 && x_zero) {
            hipLaunchKernelGGL((jacobi_from_zero<double>), grid1(L.n), dim3(256), 0, s, L.n, L.diag, b, cur, w);
            RAMD_HIP(hipGetLastError());
        } else {
            launch_stream_op(L.A, OP_JACOBI, cur, b, other, w, s, L.diag);
            std::swap(cur, other);
        }
    }
    launch_stream_op(L.A, OP_RESIDUAL, cur, b, L.r, 0.0, s);
    if (l + 1 == tail_level(h, 0)) {
        // the restriction, the coarser levels' cycle and x + P e in one launch
        launch_tail<double>(h, l, L.r, cur, other);
    } else {
        Level& C = h->levels[l + 1];
        // the coarse level's first Jacobi sweep from zero comes with its right-hand side
        const bool cj = jfz_fused(h, l + 1);
        if (cj) launch_restrict_jfz(L.R, L.r, C.b, C.x, C.diag, w, s);
        else launch_stream_op(L.R, OP_SPMV, L.r, nullptr, C.b, 0.0, s);
        level_span_close(h, tk);
        cycle64(h, l + 1, C.x, C.b, true, cj);
        tk = level_span_open(h, l);
        // x <- x + P e (CSR_append: row sum first, then added), into the other buffer
        launch_stream_op(L.P, OP_ADD_INTO, C.x, cur, other, 0.0, s);
    }
    std::swap(cur, other);
    if (mc) mc_relax<double, double>(h, L, L.A->val, cur, b, false);
    for (int it = 0; it < sweeps && !mc; it++) {
        launch_stream_op(L.A, OP_JACOBI, cur, b, other, w, s, L.diag);
        std::swap(cur, other);
    }
    if (cur != x) RAMD_HIP(hipMemcpyAsync(x, cur, sizeof(double) * L.n, hipMemcpyDeviceToDevice, s));
    level_span_close(h, tk);
}

// x_zero: x is zero on entry; first_done: x already holds the first Jacobi
// sweep from zero (formed by the caller with jacobi_from_zero's operations)
static void cycle32(ramd_hier_s* h, int l, float* x, const float* b, bool x_zero, bool first_done = false)
{
    Level& L = h->levels[l];
    hipStream_t s = h->ctx->stream;
    int tk = level_span_open(h, l);
    if (l == (int)h->levels.size() - 1) {
        const int n = h->coarse_n;
        if (n <= kCoarseWave) {
            // FP64 solve of the widened right-hand side, rounded once on output
            hipLaunchKernelGGL((coarse_wave_kernel<float, float>), dim3(1), dim3(64), sizeof(double) * n * n, s, n,
                               h->lu, h->piv, b, x);
        } else {
            hipLaunchKernelGGL(to_f64, grid1(n), dim3(256), 0, s, (int64_t)n, b, h->cb);
            coarse_solve64(h, h->cb, h->cx);
            hipLaunchKernelGGL(to_f32, grid1(n), dim3(256), 0, s, (int64_t)n, h->cx, x);
        }
        RAMD_HIP(hipGetLastError());
        level_span_close(h, tk);
        return;
    }
    const float w = (float)h->opts.relax_weight;
    const int sweeps = h->opts.num_smooth_sweeps;
    const bool mc = multicolour(h);
    float* cur = x;
    float* other = L.t32;
    if (x_zero && !mc && sweeps <= 0) hipLaunchKernelGGL((fill_zero<float>), grid1(L.n), dim3(256), 0, s, L.n, x);
    if (mc) mc_relax<float, float>(h, L, L.A->val32, cur, b, x_zero);
    for (int it = 0; it < sweeps && !mc; it++) {
        if (it == 0 && first_done) {
            continue;
        } else if (it == 0 && x_zero) {
            hipLaunchKernelGGL((jacobi_from_zero<float>), grid1(L.n), dim3(256), 0, s, L.n, L.diag32, b, cur, w);
            RAMD_HIP(hipGetLastError());
        } else {
            launch_stream_op32(L.A, OP_JACOBI, cur, b, other, w, s, L.diag32);
            std::swap(cur, other);
        }
    }
    launch_stream_op32(L.A, OP_RESIDUAL, cur, b, L.r32, 0.f, s);
    if (l + 1 == tail_level(h, 1)) {
        launch_tail<float>(h, l, L.r32, cur, other);
    } else {
        Level& C = h->levels[l + 1];
        const bool cj = jfz_fused(h, l + 1);
        if (cj) launch_restrict_jfz32(L.R, L.r32, C.b32, C.x32, C.diag32, w, s);
        else launch_stream_op32(L.R, OP_SPMV, L.r32, nullptr, C.b32, 0.f, s);
        level_span_close(h, tk);
        cycle32(h, l + 1, C.x32, C.b32, true, cj);
        tk = level_span_open(h, l);
        launch_stream_op32(L.P, OP_ADD_INTO, C.x32, cur, other, 0.f, s);
    }
    std::swap(cur, other);
    if (mc) mc_relax<float, float>(h, L, L.A->val32, cur, b, false);
    for (int it = 0; it < sweeps && !mc; it++) {
        launch_stream_op32(L.A, OP_JACOBI, cur, b, other, w, s, L.diag32);
        std::swap(cur, other);
    }
    if (cur != x) RAMD_HIP(hipMemcpyAsync(x, cur, sizeof(float) * L.n, hipMemcpyDeviceToDevice, s));
    level_span_close(h, tk);
}

// ---------------------------------------------------------------------------
// setup
// ---------------------------------------------------------------------------
static void alloc_level_vectors(ramd_hier_s* h, Level& L)
{
    L.n = L.A->n_rows;
    L.x = dalloc<double>((size_t)L.n);
    L.b = dalloc<double>((size_t)L.n);
    L.t = dalloc<double>((size_t)L.n);
    L.r = dalloc<double>((size_t)L.n);
    L.diag = dalloc<double>((size_t)L.n);
    if (L.n) {
        hipLaunchKernelGGL(extract_diag, grid1(L.n), dim3(256), 0, h->ctx->stream, L.n, L.A->rowptr, L.A->col,
                           L.A->val, L.diag);
        RAMD_HIP(hipGetLastError());
    }
}

static void free_level(Level& L)
{
    if (L.A) csr_free(L.A);
    if (L.P) csr_free(L.P);  // frees R (its transpose)
    for (void* p : {(void*)L.diag, (void*)L.x, (void*)L.b, (void*)L.t, (void*)L.r, (void*)L.diag32, (void*)L.x32,
                    (void*)L.b32, (void*)L.t32, (void*)L.r32, (void*)L.color_rows})
        dfree(p);
    L = Level();
}

static void hier_free(ramd_hier_s* h)
{
    if (!h) return;
    if (h->ctx) (void)hipStreamSynchronize(h->ctx->stream);
    block_free(h->blk);
    if (h->fg) {
        for (int i = 0; i < kFpcgGraphSets; i++) fpcg_graphs_free(&h->fg[i]);
        delete[] h->fg;
    }
    dfree(h->fpart);
    h->spans.release();
    for (Level& L : h->levels) free_level(L);
    for (void* p : {(void*)h->lu, (void*)h->piv, (void*)h->cb, (void*)h->cx, (void*)h->pr, (void*)h->pz, (void*)h->pp,
                    (void*)h->pAp, (void*)h->pr32, (void*)h->pz32})
        dfree(p);
    delete h;
}

// ParRugeStubenSolver::extend_hierarchy, ruge_stuben/par_ruge_stuben_solver.hpp:57-162
static Csr rs_level(ramd_hier_s* h, Csr& A, int level, const double* weights, std::vector<int>& states)
{
    const ramd_amg_opts& o = h->opts;
    double t0 = now_s();
    Csr S = o.strength_type == RAMD_SYMMETRIC ? host::symmetric_strength(A, o.strong_threshold)
                                              : host::classical_strength(A, o.strong_threshold);
    double t1 = now_s();
    switch (o.coarsen_type) {
    case RAMD_RS:
        // RS on the first three levels, Falgout below (par_ruge_stuben_solver.hpp:76-86)
        if (level < 3) host::split_rs(S, states);
        else host::split_falgout(S, weights, states);
        break;
    case RAMD_CLJP: host::split_cljp(S, weights, states); break;
    case RAMD_FALGOUT: host::split_falgout(S, weights, states); break;
    case RAMD_PMIS: host::split_pmis(S, weights, states); break;
    case RAMD_HMIS: host::split_hmis(S, weights, states); break;
    default: fail(RAMD_E_ARG, "unknown coarsen_type %d", o.coarsen_type);
    }
    double t2 = now_s();
    Csr P;
    switch (o.interp_type) {
    case RAMD_DIRECT: P = host::direct_interpolation(A, S, states); break;
    case RAMD_MOD_CLASSICAL: P = host::mod_classical_interpolation(A, S, states); break;
    case RAMD_EXTENDED: P = host::extended_interpolation(A, S, states, o.interp_filter); break;
    default: fail(RAMD_E_ARG, "interp_type %d is not supported", o.interp_type);
    }
    double t3 = now_s();
    h->times[0] += t1 - t0;
    h->times[1] += t2 - t1;
    h->times[2] += t3 - t2;
    (void)level;
    return P;
}

// The same level built on the device (classical or symmetric strength;
// PMIS, HMIS, CLJP, Falgout or RS splitting; direct, modified classical or
// extended interpolation; results identical to rs_level's), when the options name
// such a combination.  The sequential Ruge-Stuben passes (RS, and the first
// passes of HMIS and Falgout) run on the host between the device steps.  A
// level whose rows would need more interpolation columns than the device
// kernels hold is interpolated on the host.
static bool device_rs(const ramd_amg_opts& o)
{
    return o.device_setup && o.solver == RAMD_SOLVER_RS &&
           (o.strength_type == RAMD_CLASSICAL || o.strength_type == RAMD_SYMMETRIC) &&
           (o.interp_type == RAMD_EXTENDED || o.interp_type == RAMD_MOD_CLASSICAL || o.interp_type == RAMD_DIRECT);
}

static ramd_csr_s* rs_level_device(ramd_hier_s* h, int level, const double* weights, std::vector<int>& states)
{
    const ramd_amg_opts& o = h->opts;
    ramd_ctx ctx = h->ctx;
    Level& L = h->levels[level];
    const int n = L.n;
    double t0 = now_s();
    ramd_csr_s* S = o.strength_type == RAMD_SYMMETRIC ? gpu_symmetric_strength(L.A, o.strong_threshold)
                                                      : gpu_classical_strength(L.A, o.strong_threshold);
    int* st = nullptr;
    ramd_csr_s* P = nullptr;
    try {
        double t1 = now_s();
        st = dalloc<int>((size_t)std::max(n, 1));
        std::vector<int> prior;
        // RS: Ruge-Stuben on the first three levels, Falgout below
        // (par_ruge_stuben_solver.hpp:76-86)
        const int ct = (o.coarsen_type == RAMD_RS && level >= 3) ? RAMD_FALGOUT : o.coarsen_type;
        switch (ct) {
        case RAMD_PMIS: gpu_split_pmis(S, weights, st); break;
        case RAMD_CLJP: gpu_split_cljp(S, weights, st); break;
        case RAMD_HMIS: {
            Csr Sh = download(S);
            host::hmis_first_pass(Sh, prior);
            gpu_split_pmis(S, weights, st, prior.data());
            break;
        }
        case RAMD_FALGOUT: {
            Csr Sh = download(S);
            host::falgout_first_pass(Sh, prior);
            gpu_split_cljp(S, weights, st, prior.data());
            break;
        }
        case RAMD_RS: {
            Csr Sh = download(S);
            host::split_rs(Sh, prior);
            if (n) RAMD_HIP(hipMemcpyAsync(st, prior.data(), sizeof(int) * n, hipMemcpyHostToDevice, ctx->stream));
            break;
        }
        default: fail(RAMD_E_ARG, "unknown coarsen_type %d", o.coarsen_type);
        }
        states.resize((size_t)n);
        if (n) RAMD_HIP(hipMemcpyAsync(states.data(), st, sizeof(int) * n, hipMemcpyDeviceToHost, ctx->stream));
        RAMD_HIP(hipStreamSynchronize(ctx->stream));
        double t2 = now_s();
        const bool ext = o.interp_type == RAMD_EXTENDED;
        if (o.interp_type == RAMD_DIRECT) P = gpu_direct_interpolation(L.A, S, st, states);
        else if (ext) P = gpu_extended_interpolation(L.A, S, st, states, o.interp_filter);
        else P = gpu_mod_classical_interpolation(L.A, S, st, states);
        if (!P) {
            Csr Ah = download(L.A), Sh = download(S);
            Csr Ph = ext ? host::extended_interpolation(Ah, Sh, states, o.interp_filter)
                         : host::mod_classical_interpolation(Ah, Sh, states);
            P = upload(ctx, Ph, false);
        }
        double t3 = now_s();
        h->times[0] += t1 - t0;
        h->times[1] += t2 - t1;
        h->times[2] += t3 - t2;
    } catch (...) {
        csr_free(S);
        dfree(st);
        throw;
    }
    csr_free(S);
    dfree(st);
    return P;
}

// jacobi_prolongation aggregation/par_prolongation.cpp:7-103: P = T, then
// `steps` times P = P - (omega D^-1 A) P with D the absolute row sums; the
// products run on the device, the row scaling and the subtraction's layout
// pass on the host.
static Csr sa_prolongation(ramd_ctx ctx, const Csr& A, Csr P, double omega, int steps)
{
    if (steps <= 0) return P;
    Csr scaled = host::sa_scaled_operator(A, omega);
    ramd_csr_s* Sd = upload(ctx, scaled, false);
    try {
        for (int s = 0; s < steps; s++) {
            ramd_csr_s* Pd = upload(ctx, P, false);
            ramd_csr_s* Qd = nullptr;
            try {
                Qd = spgemm(Sd, Pd);
            } catch (...) {
                csr_free(Pd);
                throw;
            }
            csr_free(Pd);
            Csr Q = download(Qd);
            csr_free(Qd);
            P = host::sa_subtract(P, Q);
        }
    } catch (...) {
        csr_free(Sd);
        throw;
    }
    csr_free(Sd);
    return P;
}

// ParSmoothedAggregationSolver::extend_hierarchy,
// aggregation/par_smoothed_aggregation_solver.hpp:52-129: strength, MIS-2,
// aggregation, tentative prolongator from the candidate vector B (replaced
// by the aggregate norms for the next level), Jacobi smoothing.
static Csr sa_level(ramd_hier_s* h, Csr& A, const double* weights, std::vector<int>& aggregates,
                    std::vector<double>& B)
{
    const ramd_amg_opts& o = h->opts;
    double t0 = now_s();
    Csr S = o.strength_type == RAMD_SYMMETRIC ? host::symmetric_strength(A, o.strong_threshold)
                                              : host::classical_strength(A, o.strong_threshold);
    double t1 = now_s();
    std::vector<int> states;
    host::mis2(S, weights, states);
    const int n_aggs = host::aggregate(A, S, states, nullptr, aggregates);
    double t2 = now_s();
    std::vector<double> R;
    Csr T = host::fit_candidates(n_aggs, aggregates, B, R);
    Csr P = sa_prolongation(h->ctx, A, std::move(T), o.prolong_weight, o.prolong_smooth_steps);
    B = std::move(R);
    double t3 = now_s();
    h->times[0] += t1 - t0;
    h->times[1] += t2 - t1;
    h->times[2] += t3 - t2;
    return P;
}

// sa_level on the device: strength, MIS-2, aggregation, fit_candidates,
// the scaled operator, its products with the prolongator and the
// subtractions, with sa_level's results.  The candidate vector B (device,
// n) is replaced by the aggregate norms (n_aggs), the next level's B.  The
// host forms only the aggregates' column numbers and T's row pointers.
static bool device_sa(const ramd_amg_opts& o) { return o.device_setup && o.solver == RAMD_SOLVER_SA; }

static ramd_csr_s* sa_level_device(ramd_hier_s* h, int level, const double* weights, std::vector<int>& aggregates,
                                   double*& B)
{
    const ramd_amg_opts& o = h->opts;
    ramd_ctx ctx = h->ctx;
    Level& L = h->levels[level];
    const int n = L.n;
    double t0 = now_s();
    ramd_csr_s* S = o.strength_type == RAMD_SYMMETRIC ? gpu_symmetric_strength(L.A, o.strong_threshold)
                                                      : gpu_classical_strength(L.A, o.strong_threshold);
    int *st = nullptr, *agg = nullptr;
    double* R = nullptr;
    ramd_csr_s *P = nullptr, *M = nullptr, *Q = nullptr;
    try {
        double t1 = now_s();
        st = dalloc<int>((size_t)std::max(n, 1));
        agg = dalloc<int>((size_t)std::max(n, 1));
        gpu_mis2(S, weights, st);
        const int n_aggs = gpu_aggregate(L.A, S, st, agg);
        aggregates.resize((size_t)n);
        if (n) RAMD_HIP(hipMemcpyAsync(aggregates.data(), agg, sizeof(int) * n, hipMemcpyDeviceToHost, ctx->stream));
        RAMD_HIP(hipStreamSynchronize(ctx->stream));
        double t2 = now_s();
        R = dalloc<double>((size_t)std::max(n_aggs, 1));
        P = gpu_fit_candidates(ctx, n, n_aggs, agg, aggregates, B, R);
        if (o.prolong_smooth_steps > 0) {
            M = gpu_sa_scaled_operator(L.A, o.prolong_weight);
            for (int step = 0; step < o.prolong_smooth_steps; step++) {
                Q = spgemm(M, P);
                ramd_csr_s* next = gpu_sa_subtract(P, Q);
                csr_free(Q);
                Q = nullptr;
                csr_free(P);
                P = next;
            }
            csr_free(M);
            M = nullptr;
        }
        dfree(B);
        B = R;
        R = nullptr;
        double t3 = now_s();
        h->times[0] += t1 - t0;
        h->times[1] += t2 - t1;
        h->times[2] += t3 - t2;
    } catch (...) {
        csr_free(S);
        dfree(st);
        dfree(agg);
        dfree(R);
        csr_free(P);
        csr_free(M);
        csr_free(Q);
        throw;
    }
    csr_free(S);
    dfree(st);
    dfree(agg);
    return P;
}

// Device state every hierarchy needs once its levels exist: the coarse LU
// (column-major factors of A_c^T and pivots), the PCG work vectors, the
// kernel tuning of every cycle operator and the ownership marks.
static void finish_hierarchy(ramd_hier_s* h, int n, const std::vector<double>& lu, const std::vector<int>& piv)
{
    ramd_ctx ctx = h->ctx;
    h->coarse_n = n;
    h->lu = dalloc<double>((size_t)n * n);
    h->piv = dalloc<int>((size_t)n);
    h->cb = dalloc<double>((size_t)n);
    h->cx = dalloc<double>((size_t)n);
    if (n) {
        RAMD_HIP(hipMemcpyAsync(h->lu, lu.data(), sizeof(double) * n * n, hipMemcpyHostToDevice, ctx->stream));
        RAMD_HIP(hipMemcpyAsync(h->piv, piv.data(), sizeof(int) * n, hipMemcpyHostToDevice, ctx->stream));
    }
    const int n0 = h->levels[0].n;
    h->pr = dalloc<double>((size_t)n0);
    h->pz = dalloc<double>((size_t)n0);
    h->pp = dalloc<double>((size_t)n0);
    h->pAp = dalloc<double>((size_t)n0);
    // row-block caps of every CSR operator of the cycle and the x staging of
    // the diagonal-form ones, FP64 vectors
    for (Level& L : h->levels) {
        if (&L != &h->levels.back()) {
            csr_autotune(L.A, 0);
            dia_autotune(L.A, 0);
        }
        if (L.P) {
            csr_autotune(L.P, 0);
            csr_autotune(L.R, 0);
        }
    }
    if (multicolour(h)) build_colours(h);
    RAMD_HIP(hipStreamSynchronize(ctx->stream));
    for (Level& L : h->levels) L.A->borrowed = true;
    for (Level& L : h->levels)
        if (L.P) L.P->borrowed = true;
}

static void hier_setup(ramd_hier_s* h, Csr Af)
{
    const ramd_amg_opts& o = h->opts;
    ramd_ctx ctx = h->ctx;
    double tstart = now_s();
    double tlevel = tstart;
    // levels[0]->A = Af->copy(); sort(); move_diag()  (par_multilevel.hpp:141-143)
    host::sort_rows(Af);
    host::move_diag(Af);
    std::vector<double> weights;
    if (o.weights) weights.assign(o.weights, o.weights + Af.n_rows);
    else weights = host::reference_weights(Af.n_rows, 0);

    h->levels.emplace_back();
    h->levels[0].A = upload(ctx, Af);
    alloc_level_vectors(h, h->levels[0]);
    Csr A = std::move(Af);
    bool host_A = true;  // A holds the current level (the device path skips the copy)
    // SA near-nullspace candidate (aggregation/par_smoothed_aggregation_solver.hpp:38-49)
    std::vector<double> B;
    struct DeviceB {
        double* p = nullptr;
        ~DeviceB() { dfree(p); }
    } Bd;
    if (o.solver == RAMD_SOLVER_SA) {
        B.assign((size_t)A.n_rows, 1.0);
        if (device_sa(o)) {
            Bd.p = dalloc<double>((size_t)std::max(A.n_rows, 1));
            if (A.n_rows)
                RAMD_HIP(hipMemcpyAsync(Bd.p, B.data(), sizeof(double) * A.n_rows, hipMemcpyHostToDevice,
                                        ctx->stream));
        }
    }
    while (h->levels.back().n > o.max_coarse &&
           (o.max_levels == -1 || (int)h->levels.size() < o.max_levels)) {
        const int l = (int)h->levels.size() - 1;
        Level& L = h->levels[l];
        ramd_csr_s* Pd = nullptr;
        if (device_rs(o)) {
            Pd = rs_level_device(h, l, weights.data(), L.states);
        } else if (device_sa(o)) {
            Pd = sa_level_device(h, l, weights.data(), L.states, Bd.p);
        } else {
            if (!host_A) A = download(L.A);
            Csr P = o.solver == RAMD_SOLVER_SA ? sa_level(h, A, weights.data(), L.states, B)
                                               : rs_level(h, A, l, weights.data(), L.states);
            Pd = upload(ctx, P);
        }
        L.P = Pd;
        if (Pd->n_cols == 0) fail(RAMD_E_NUMERIC, "level %d produced no coarse points", l);
        double tg = now_s();
        L.R = csr_transpose(L.P);
        ramd_csr_s* Ac = galerkin(L.A, L.P);
        h->times[3] += now_s() - tg;
        h->levels.emplace_back();
        h->levels.back().A = Ac;
        alloc_level_vectors(h, h->levels.back());
        host_A = false;
        RAMD_HIP(hipStreamSynchronize(ctx->stream));
        const double tl = now_s();
        h->setup_level_s.push_back(tl - tlevel);
        tlevel = tl;
    }
    if (!host_A) A = download(h->levels.back().A);
    // duplicate_coarse (par_multilevel.hpp:229-339): dense row-major A_c, dgetrf_
    const int n = A.n_rows;
    RAMD_CHECK(n <= kCoarseMax, RAMD_E_ARG, "coarsest level has %d rows; the dense solve supports <= %d", n,
               kCoarseMax);
    std::vector<double> dense((size_t)n * n, 0.0);
    for (int i = 0; i < n; i++)
        for (int j = A.rowptr[i]; j < A.rowptr[i + 1]; j++) dense[(size_t)i * n + A.col[j]] = A.val[j];
    std::vector<int> piv;
    // as the reference, a zero pivot does not stop the setup (dgetrf_'s info
    // is not inspected, multilevel/par_multilevel.hpp:330): the factors are
    // kept and a cycle then divides by the zero pivot.  The info is kept for
    // ramd_hier_coarse_info.
    h->coarse_info = lu_factor_colmajor(n, dense, piv);
    finish_hierarchy(h, n, dense, piv);
    const double tend = now_s();
    h->setup_level_s.push_back(tend - tlevel);
    double total = tend - tstart;
    h->times[4] = total - h->times[0] - h->times[1] - h->times[2] - h->times[3];
}

void enable_fp32(ramd_hier_s* h)
{
    if (h->fp32) return;
    hipStream_t s = h->ctx->stream;
    for (Level& L : h->levels) {
        csr_make_fp32(L.A);
        if (L.P) {
            csr_make_fp32(L.P);
            csr_make_fp32(L.R);
        }
        if (&L != &h->levels.back()) {
            csr_autotune(L.A, 1);
            dia_autotune(L.A, 1);
        }
        if (L.P) {
            csr_autotune(L.P, 1);
            csr_autotune(L.R, 1);
        }
        L.x32 = dalloc<float>((size_t)L.n);
        L.b32 = dalloc<float>((size_t)L.n);
        L.t32 = dalloc<float>((size_t)L.n);
        L.r32 = dalloc<float>((size_t)L.n);
        L.diag32 = dalloc<float>((size_t)L.n);
        if (L.n) hipLaunchKernelGGL(to_f32, grid1(L.n), dim3(256), 0, s, (int64_t)L.n, L.diag, L.diag32);
        RAMD_HIP(hipGetLastError());
    }
    const int n0 = h->levels[0].n;
    h->pr32 = dalloc<float>((size_t)n0);
    h->pz32 = dalloc<float>((size_t)n0);
    RAMD_HIP(hipStreamSynchronize(s));
    h->fp32 = true;
}

// ---------------------------------------------------------------------------
// PCG vector kernels (reference operand order, one rounding per op)
// ---------------------------------------------------------------------------
__global__ void pcg_update_xr(int n, double alpha, const double* __restrict__ p, const double* __restrict__ Ap,
                              double* __restrict__ x, double* __restrict__ r, int update_r)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    x[i] += p[i] * alpha;
    if (update_r) r[i] += Ap[i] * (-1.0 * alpha);
}

__global__ void pcg_update_p(int n, double beta, const double* __restrict__ z, double* __restrict__ p)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double t = p[i] * beta;
    p[i] = t + z[i] * 1.0;
}

static double dot64(ramd_ctx ctx, int n, const double* a, const double* b)
{
    launch_dot(ctx, n, a, b, ctx->red_result);
    return fetch_scalar(ctx, ctx->red_result);
}

static void vcycle_zero(ramd_hier_s* h, double* z, const double* r) { cycle64(h, 0, z, r, true); }

void vcycle_from_zero(ramd_hier_s* h, double* x, const double* b) { cycle64(h, 0, x, b, true); }

// Flexible-PCG vector kernels.  The reductions run on a fixed grid of
// kRedBlocks workgroups, each summing a fixed strided subset, so every run
// gives the same bits.  (A last-workgroup-finishes variant that would save
// the final reduction's launch needs a device-scope release per workgroup,
// which writes back the XCD's L2: measured slower.)

// r -= alpha Ap; r32 = (float) r when given; partial sums of r^2.  alpha =
// rz / pAp is formed from the device scalars (every thread computes the
// same value; workgroup 0 stores it for the direction update, which also
// takes the step x += alpha p).  With z32 given, also the FP32 cycle's first
// Jacobi sweep from zero on r32 (jacobi_from_zero's operations, diagonal
// d32, weight w), so the cycle starts from it.
__global__ __launch_bounds__(kBlockThreads) void fpcg_update_norm(int n, const double* __restrict__ sc_rz,
                                                                  const double* __restrict__ sc_pap,
                                                                  double* __restrict__ sc_alpha,
                                                                  const double* __restrict__ Ap,
                                                                  double* __restrict__ r, float* __restrict__ r32,
                                                                  double* __restrict__ partials,
                                                                  const float* __restrict__ d32,
                                                                  float* __restrict__ z32, float w)
{
    const double alpha = *sc_rz / *sc_pap;
    if (blockIdx.x == 0 && threadIdx.x == 0) *sc_alpha = alpha;
    double s = 0;
    const int stride = gridDim.x * kBlockThreads;
    // four of the thread's entries per step, their loads issued first (the
    // operations and the order of the sum are the plain loop's)
    constexpr int U = 4;
    for (int i0 = blockIdx.x * kBlockThreads + threadIdx.x; i0 < n; i0 += U * stride) {
        double rv[U], av[U];
        float dv[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int i = i0 + u * stride;
            if (i < n) {
                rv[u] = r[i];
                av[u] = Ap[i];
                if (z32) dv[u] = d32[i];
            }
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int i = i0 + u * stride;
            if (i >= n) break;
            const double ri = rv[u] - alpha * av[u];
            r[i] = ri;
            if (r32) {
                const float rf = (float)ri;
                r32[i] = rf;
                if (z32) {
                    const float di = dv[u];
                    const float zero = 0.0f;
                    float v = zero;
                    if (fabs(di) > float(1e-16)) v = ((1.0f - w) * zero) + (w * ((rf - zero) / di));
                    z32[i] = v;
                }
            }
            s += ri * ri;
        }
    }
    s = block_sum(s);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

// partial sums of r.z (first gridDim.x entries) and z.Ap (next gridDim.x)
template <typename Z>
__global__ __launch_bounds__(kBlockThreads) void fpcg_dots(int n, const double* __restrict__ r,
                                                           const Z* __restrict__ z, const double* __restrict__ Ap,
                                                           double* __restrict__ partials)
{
    double s1 = 0, s2 = 0;
    const int stride = gridDim.x * kBlockThreads;
    // four of the thread's entries per step, their loads issued first (the
    // products and the order of the sums are the plain loop's)
    constexpr int U = 4;
    for (int i0 = blockIdx.x * kBlockThreads + threadIdx.x; i0 < n; i0 += U * stride) {
        double rv[U], zv[U], av[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int i = i0 + u * stride;
            if (i < n) {
                zv[u] = (double)z[i];
                rv[u] = r[i];
                av[u] = Ap[i];
            }
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            if (i0 + u * stride >= n) break;
            s1 += rv[u] * zv[u];
            s2 += zv[u] * av[u];
        }
    }
    s1 = block_sum(s1);
    s2 = block_sum(s2);
    if (threadIdx.x == 0) {
        partials[blockIdx.x] = s1;
        partials[gridDim.x + blockIdx.x] = s2;
    }
}

// x += alpha p (when x is given: the step of this iteration, taken here so
// that p is read once), then p = z + beta p with the flexible
// (Polak-Ribiere) beta = -alpha (z.Ap) / rz_old from the device scalars
template <typename Z>
__global__ void fpcg_dir(int n, const double* __restrict__ sc_alpha, const double* __restrict__ sc_zap,
                         const double* __restrict__ sc_rz_old, const Z* __restrict__ z, double* __restrict__ p,
                         double* __restrict__ x)
{
    const double alpha = *sc_alpha;
    const double beta = -alpha * (*sc_zap) / (*sc_rz_old);
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double pi = p[i];
    if (x) x[i] += alpha * pi;
    p[i] = (double)z[i] + beta * pi;
}

// x += alpha p with alpha from the device scalar (the step not yet taken
// when the iteration stops or checks the true residual)
__global__ void fpcg_step_x(int n, const double* __restrict__ sc_alpha, const double* __restrict__ p,
                            double* __restrict__ x)
{
    const double alpha = *sc_alpha;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] += alpha * p[i];
}

static int fixed_blocks(int n) { return std::max(1, std::min(kRedBlocks, (n + kBlockThreads - 1) / kBlockThreads)); }

// r.z and z.Ap of the flexible-PCG step, in one pass, into device scalars
// (part: 2 * fixed_blocks(n) partial sums)
template <typename Z>
static void fpcg_rz_zap(ramd_ctx ctx, int n, const double* r, const Z* z, const double* Ap, double* rz_d,
                        double* zap_d, double* part)
{
    const int g = fixed_blocks(n);
    hipLaunchKernelGGL((fpcg_dots<Z>), dim3(g), dim3(kBlockThreads), 0, ctx->stream, n, r, z, Ap, part);
    RAMD_HIP(hipGetLastError());
    launch_reduce_final_two(ctx, part, g, rz_d, part + g, g, zap_d);
}

// plain 2-norm (no small-entry skipping, NaN propagates)
static double norm2_64(ramd_ctx ctx, int n, const double* a) { return std::sqrt(dot64(ctx, n, a, a)); }

hipGraphExec_t capture_graph(hipStream_t s, const std::function<void()>& fn)
{
    hipGraph_t g = nullptr;
    RAMD_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    try {
        fn();
    } catch (...) {
        (void)hipStreamEndCapture(s, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
    }
    RAMD_HIP(hipStreamEndCapture(s, &g));
    hipGraphExec_t ex = nullptr;
    const hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    RAMD_HIP(e);
    return ex;
}

void fpcg_graphs_free(FpcgGraphs* g)
{
    if (!g) return;
    for (int i = 0; i < 2; i++) {
        if (g->gA[i]) (void)hipGraphExecDestroy(g->gA[i]);
        if (g->gB[i]) (void)hipGraphExecDestroy(g->gB[i]);
        g->gA[i] = g->gB[i] = nullptr;
    }
    g->a_uid = 0;
}

// The FPCG's own partial-sum buffer (>= count doubles).  Growing it
// releases the captured iterations, which hold the old buffer.
static double* fpcg_partials(ramd_hier_s* h, size_t count)
{
    if (h->fpart_cap < count) {
        RAMD_HIP(hipStreamSynchronize(h->ctx->stream));
        if (h->fg)
            for (int i = 0; i < kFpcgGraphSets; i++) fpcg_graphs_free(&h->fg[i]);
        dfree(h->fpart);
        h->fpart = dalloc<double>(count);
        h->fpart_cap = count;
    }
    return h->fpart;
}

// Flexible PCG in FP64 (true residual, inner products, coarse LU in FP64)
// preconditioned by one V-cycle in `cycle_prec`; stops on the true relative
// residual.  Returns RAMD_OK, RAMD_E_INDEFINITE, RAMD_E_NUMERIC or
// RAMD_E_NOT_CONVERGED; throws on argument and device errors.
//
// An iteration is the operator step (A p with p.Ap, the residual update
// with ||r||^2; the last kernel also writes those 16 bytes to host memory),
// the preconditioner application (V-cycle, r.z and z.Ap) and, after the host
// has read the 16 bytes, the direction update.  While the residual is not
// close to rtol the application is queued before the host waits for the
// operator step, so the device does not idle while the host reads and
// decides.  With the context's graphs on (and no kernel timing or per-level
// tracking active) the operator step and the application are replayed as
// captured HIP graphs, one pair per parity of the alternating r.z slots;
// the application after a true-residual restart runs eagerly.  Replays and
// eager steps launch the same kernels with the same arguments: same bits.
// Partial sums of ramd_fpcg_noalloc: a fixed array in every GPU's memory
// (kNoallocPartials doubles, 512 KiB), so that entry point allocates
// nothing.  Host code uses the address hipGetSymbolAddress gives for the
// device that is current at set-up (noalloc_scratch_setup, from ramd_init);
// the array's own name is not a device address in host code.  Every call
// on a GPU shares that GPU's array: calls that overlap on one GPU race.
static __device__ __align__(256) double scratch[kNoallocPartials];
static std::mutex g_scratch_mu;
static double* g_scratch_ptr[kMaxDevices] = {};

void noalloc_scratch_setup(int device)
{
    RAMD_CHECK(device >= 0 && device < kMaxDevices, RAMD_E_ARG, "device %d past the scratch table (%d)", device,
               kMaxDevices);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return;
    void* ptr = nullptr;
    RAMD_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    g_scratch_ptr[device] = static_cast<double*>(ptr);
}

static double* noalloc_scratch(int device)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    RAMD_CHECK(device >= 0 && device < kMaxDevices && g_scratch_ptr[device], RAMD_E_STATE,
               "ramd_fpcg_noalloc: no scratch address for device %d (the context was not set up by ramd_init)",
               device);
    return g_scratch_ptr[device];
}

// noalloc: ramd_fpcg_noalloc's form -- no device allocation and no graph:
// the FP32 copies, the outer operator's kernel choice and the per-hierarchy
// buffers must exist already; partial sums go to the fixed scratch array.
static int fpcg_run(ramd_hier_s* h, ramd_csr_s* A, double* x_d, const double* b_d, double rtol, int max_iter,
                    int cycle_prec, double* relres_hist, int32_t* iters, double* final_relres, bool noalloc = false)
{
    int rc = RAMD_OK;
    ramd_ctx ctx = h->ctx;
    ctx->set_current();
    hipStream_t s = ctx->stream;
    const int n = A->n_rows;
    RAMD_CHECK(n == h->levels[0].n, RAMD_E_ARG, "ramd_fpcg: A has %d rows, hierarchy %d", n, h->levels[0].n);
    const bool f32 = cycle_prec == RAMD_PREC_FP32;
    // Both forms: the outer FP64 operator reads FP32-stored values where
    // every value is exactly representable in FP32 (same FP64 arithmetic and
    // results, half the value bytes); otherwise the FP64 values
    auto outer = [](const ramd_csr_s* M, int op, const double* x, const double* b, double* y, double w,
                    hipStream_t st) { launch_stream_op_c(M, op, x, b, y, w, st); };
    if (noalloc) {
        RAMD_CHECK(!f32 || h->fp32, RAMD_E_STATE,
                   "ramd_fpcg_noalloc: the hierarchy's FP32 copies are not built (ramd_hier_enable_fp32)");
        RAMD_CHECK(!A->dia || (A->val32 && A->dia->v32 && A->dia->lds_x[A->v32_outer ? 2 : 0] >= 0), RAMD_E_STATE,
                   "ramd_fpcg_noalloc: the operator's FP32 values and kernel choice are made by a first "
                   "ramd_fpcg call with it");
        RAMD_CHECK(!h->track && ctx->prof_min_nnz < 0, RAMD_E_STATE,
                   "ramd_fpcg_noalloc: per-level timing or kernel profiling is on (they create events)");
    } else {
        cycle_prepare(h, f32 ? 1 : 0);
        if (A->dia) {
            csr_make_fp32(A);
            // the outer operator reads FP32 values into FP64 vectors (or,
            // with that turned off, FP64 values)
            dia_autotune(A, A->v32_outer ? 2 : 0);
        }
    }
    double *r = h->pr, *z = h->pz, *p = h->pp, *Ap = h->pAp;
    float *r32 = h->pr32, *z32 = h->pz32;
    const int g = fixed_blocks(n);
    const dim3 gv = grid1(n), bv(256);
    const size_t npart = (size_t)std::max(spmv_dot_partials(A), 2 * g);
    if (noalloc)
        RAMD_CHECK(npart <= (size_t)kNoallocPartials, RAMD_E_ARG,
                   "ramd_fpcg_noalloc: the operator needs %zu partial sums, the scratch holds %d (about %d rows)",
                   npart, kNoallocPartials, kNoallocPartials * kBlockThreads);
    double* part = noalloc ? noalloc_scratch(ctx->device) : fpcg_partials(h, npart);
    // FP32 form: the update pass also forms the cycle's first Jacobi sweep
    // from zero on the fine level (Jacobi smoothing, more than one level)
    const bool fuse_first = f32 && !multicolour(h) && h->opts.num_smooth_sweeps >= 1 && h->levels.size() > 1;
    // device scalars: [0] p.Ap, [1] r.r, [2]/[5] r.z (current / next), [3] z.Ap, [4] alpha
    double* sc = ctx->red_result;
    int cur = 2, nxt = 5;
    // operator step: Ap = A p and p.Ap in one pass over A; then r -= alpha Ap,
    // ||r||^2 (and r32, and in the FP32 form the first sweep z32) in one pass;
    // p.Ap and r.r are also written to host memory, read after step_ev
    auto stepA = [&](int c) {
        launch_spmv_dot(A, true, p, Ap, sc, s, part);
        hipLaunchKernelGGL(fpcg_update_norm, dim3(g), dim3(kBlockThreads), 0, s, n, sc + c, sc, sc + 4, Ap, r,
                           f32 ? r32 : nullptr, part, fuse_first ? h->levels[0].diag32 : nullptr,
                           fuse_first ? z32 : nullptr, (float)h->opts.relax_weight);
        RAMD_HIP(hipGetLastError());
        launch_reduce_final_pub(ctx, part, g, sc + 1, sc, ctx->step_pin_d);
    };
    // preconditioner application: z = M^-1 r (first_done: z32 holds the
    // fused first sweep), r.z into slot nx and z.Ap.  It writes only z, the
    // cycle's level vectors, the partial sums and scalar slots nx and 3.
    auto stepB = [&](int nx, bool first_done) {
        if (f32) cycle32(h, 0, z32, r32, true, first_done);
        else cycle64(h, 0, z, r, true);
        if (f32) fpcg_rz_zap(ctx, n, r, (const float*)z32, Ap, sc + nx, sc + 3, part);
        else fpcg_rz_zap(ctx, n, r, (const double*)z, Ap, sc + nx, sc + 3, part);
    };
    // direction update: the step x += alpha p (unless taken already: xs
    // null), then p = z + beta p with the flexible beta (r_new - r_old =
    // -alpha Ap) and r.z of slot c as the old one
    auto stepD = [&](int c, double* xs) {
        if (f32) hipLaunchKernelGGL((fpcg_dir<float>), gv, bv, 0, s, n, sc + 4, sc + 3, sc + c, z32, p, xs);
        else hipLaunchKernelGGL((fpcg_dir<double>), gv, bv, 0, s, n, sc + 4, sc + 3, sc + c, z, p, xs);
        RAMD_HIP(hipGetLastError());
    };
    const bool use_graphs = !noalloc && ctx->graphs && ctx->prof_min_nnz < 0 && !h->track;
    FpcgGraphs* fg = nullptr;
    if (use_graphs) {
        if (!h->fg) h->fg = new FpcgGraphs[kFpcgGraphSets];
        const uint64_t key = forms_key(h, A);
        for (int i = 0; i < kFpcgGraphSets && !fg; i++) {
            FpcgGraphs& e = h->fg[i];
            if (e.a_uid == A->uid && e.key == key && e.prec == cycle_prec && e.part == part) fg = &e;
        }
        if (!fg) {  // the least recently used entry is captured anew
            fg = &h->fg[0];
            for (int i = 1; i < kFpcgGraphSets; i++)
                if (h->fg[i].used < fg->used) fg = &h->fg[i];
            RAMD_HIP(hipStreamSynchronize(s));
            fpcg_graphs_free(fg);
            fg->a_uid = A->uid;
            fg->key = key;
            fg->prec = cycle_prec;
            fg->part = part;
        }
        fg->used = ++h->fg_tick;
    }
    const double bnorm = norm2_64(ctx, n, b_d);
    const double scale = bnorm > 0 ? bnorm : 1.0;
    outer(A, OP_RESIDUAL, x_d, b_d, r, 0.0, s);
    double relres = norm2_64(ctx, n, r) / scale;
    if (relres_hist) relres_hist[0] = relres;
    int it = 0;
    if (!std::isfinite(bnorm) || !std::isfinite(relres)) {
        rc = RAMD_E_NUMERIC;
    } else if (relres > rtol) {
        if (f32) {
            hipLaunchKernelGGL(to_f32, gv, bv, 0, s, (int64_t)n, r, r32);
            RAMD_HIP(hipGetLastError());
            cycle32(h, 0, z32, r32, true, false);
            hipLaunchKernelGGL(to_f64, gv, bv, 0, s, (int64_t)n, z32, p);
        } else {
            cycle64(h, 0, z, r, true);
            RAMD_HIP(hipMemcpyAsync(p, z, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
        }
        RAMD_HIP(hipGetLastError());
        if (f32) fpcg_rz_zap(ctx, n, r, (const float*)z32, r, sc + cur, sc + 3, part);
        else fpcg_rz_zap(ctx, n, r, (const double*)z, r, sc + cur, sc + 3, part);
        // the step x += alpha p of an iteration is taken by its direction
        // update (one read of p), or here before x is read or returned
        auto step_x = [&] {
            hipLaunchKernelGGL(fpcg_step_x, gv, bv, 0, s, n, sc + 4, p, x_d);
            RAMD_HIP(hipGetLastError());
        };
        double relres_prev = 0.0;  // relative residual one iteration back (0: none yet)
        while (it < max_iter) {
            const int par = cur == 2 ? 0 : 1;
            if (use_graphs) {
                if (!fg->gA[par]) fg->gA[par] = capture_graph(s, [&] { stepA(cur); });
                RAMD_HIP(hipGraphLaunch(fg->gA[par], s));
            } else {
                stepA(cur);
            }
            RAMD_HIP(hipEventRecord(ctx->step_ev, s));
            // Queue the preconditioner application before the host reads
            // this step's scalars, so the device stays busy meanwhile, when
            // the residual is not about to reach rtol (at the last ratio of
            // successive residuals, not within two more iterations).  It
            // changes nothing the iteration keeps when the iteration then
            // stops or restarts from the true residual (it is run again on
            // that residual).
            const double rho = relres_prev > 0.0 ? relres / relres_prev : 0.1;
            const bool ahead = relres * rho * rho > rtol;
            const bool plain_b = fuse_first;  // first_done of a plain iteration's application
            auto run_b = [&](bool first_done) {
                if (use_graphs && first_done == plain_b) {
                    if (!fg->gB[par]) fg->gB[par] = capture_graph(s, [&] { stepB(nxt, plain_b); });
                    RAMD_HIP(hipGraphLaunch(fg->gB[par], s));
                } else {
                    stepB(nxt, first_done);
                }
            };
            if (ahead) run_b(plain_b);
            bool z_first = fuse_first;  // z32 holds the fused first sweep of the current r32
            RAMD_HIP(hipEventSynchronize(ctx->step_ev));
            const volatile double* pin = ctx->step_pin;
            const double pr[2] = {pin[0], pin[1]};
            if (!(pr[0] > 0.0)) {
                // x and r take the step; the error is reported instead of a result
                step_x();
                rc = std::isfinite(pr[0]) ? RAMD_E_INDEFINITE : RAMD_E_NUMERIC;
                break;
            }
            it++;
            relres_prev = relres;
            relres = std::sqrt(pr[1]) / scale;
            bool x_done = false, r_new = false;
            if (relres <= rtol || !std::isfinite(relres) || it == max_iter) {
                step_x();
                x_done = true;
            }
            if (relres <= rtol) {
                // confirm with the true residual; continue from it if the recurrence drifted
                outer(A, OP_RESIDUAL, x_d, b_d, r, 0.0, s);
                relres = norm2_64(ctx, n, r) / scale;
                if (relres_hist) relres_hist[it] = relres;
                if (relres <= rtol) break;
                if (f32) {
                    hipLaunchKernelGGL(to_f32, gv, bv, 0, s, (int64_t)n, r, r32);
                    RAMD_HIP(hipGetLastError());
                }
                z_first = false;  // r32 was replaced
                r_new = true;
            } else if (relres_hist) {
                relres_hist[it] = relres;
            }
            if (!std::isfinite(relres)) {
                rc = RAMD_E_NUMERIC;
                break;
            }
            if (!ahead || r_new) run_b(z_first);
            stepD(cur, x_done ? nullptr : x_d);
            std::swap(cur, nxt);
        }
    }
    RAMD_HIP(hipStreamSynchronize(s));
    *iters = it;
    *final_relres = relres;
    if (rc == RAMD_OK && relres > rtol) rc = RAMD_E_NOT_CONVERGED;
    return rc;
}

static void fpcg_error(const char* who, int rc, int max_iter)
{
    if (rc == RAMD_E_INDEFINITE) set_error("%s: p.Ap <= 0 (matrix or preconditioner not SPD)", who);
    if (rc == RAMD_E_NUMERIC) set_error("%s: non-finite value in the iteration", who);
    if (rc == RAMD_E_NOT_CONVERGED) set_error("%s: %d iterations did not reach rtol", who, max_iter);
}

// dL/dA on A's pattern: g_k = -lambda_row * x_col (adjoint of A x = b)
__global__ void adjoint_outer(int n, const int* __restrict__ rp, const int* __restrict__ col,
                              const double* __restrict__ lam, const double* __restrict__ x, double* __restrict__ g)
{
    const int row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= n) return;
    const double l = lam[row];
    for (int j = rp[row]; j < rp[row + 1]; j++) g[j] = -(l * x[col[j]]);
}

// ---------------------------------------------------------------------------
// hierarchy files: every level's A (stored order), P, splitting/aggregates,
// the coarse LU factors and the options, so a solve can skip the setup
// ---------------------------------------------------------------------------
static const char kHierMagic[8] = {'R', 'A', 'M', 'D', 'H', 'I', 'E', 'R'};
constexpr int32_t kHierVersion = 2;  // 2: relax_type stored after the prolongation steps

struct HierFile {
    FILE* f = nullptr;
    const char* path = "";
    ~HierFile()
    {
        if (f) std::fclose(f);
    }
    void put_bytes(const void* p, size_t n)
    {
        if (n && std::fwrite(p, 1, n, f) != n) fail(RAMD_E_ARG, "%s: write failed", path);
    }
    void get_bytes(void* p, size_t n)
    {
        if (n && std::fread(p, 1, n, f) != n) fail(RAMD_E_ARG, "%s: file ends early", path);
    }
    template <class T>
    void put(T v)
    {
        put_bytes(&v, sizeof v);
    }
    template <class T>
    T get()
    {
        T v;
        get_bytes(&v, sizeof v);
        return v;
    }
    template <class T>
    void put_vec(const std::vector<T>& v)
    {
        put<int64_t>((int64_t)v.size());
        put_bytes(v.data(), v.size() * sizeof(T));
    }
    template <class T>
    std::vector<T> get_vec(int64_t max_n)
    {
        const int64_t n = get<int64_t>();
        RAMD_CHECK(n >= 0 && n <= max_n, RAMD_E_ARG, "%s: corrupt array length %lld", path, (long long)n);
        std::vector<T> v((size_t)n);
        get_bytes(v.data(), (size_t)n * sizeof(T));
        return v;
    }
    void put_csr(const Csr& M)
    {
        put<int32_t>(M.n_rows);
        put<int32_t>(M.n_cols);
        put_vec(M.rowptr);
        put_vec(M.col);
        put_vec(M.val);
    }
    Csr get_csr()
    {
        Csr M;
        M.n_rows = get<int32_t>();
        M.n_cols = get<int32_t>();
        RAMD_CHECK(M.n_rows >= 0 && M.n_cols >= 0, RAMD_E_ARG, "%s: corrupt matrix header", path);
        M.rowptr = get_vec<int>((int64_t)M.n_rows + 1);
        M.col = get_vec<int>(INT64_MAX / 8);
        M.val = get_vec<double>(INT64_MAX / 8);
        RAMD_CHECK((int64_t)M.rowptr.size() == (int64_t)M.n_rows + 1 && M.rowptr[0] == 0 &&
                       M.rowptr[M.n_rows] == (int)M.col.size() && M.col.size() == M.val.size(),
                   RAMD_E_ARG, "%s: inconsistent matrix", path);
        for (int i = 0; i < M.n_rows; i++)
            RAMD_CHECK(M.rowptr[i + 1] >= M.rowptr[i], RAMD_E_ARG, "%s: corrupt row pointers", path);
        for (int c : M.col) RAMD_CHECK(c >= 0 && c < M.n_cols, RAMD_E_ARG, "%s: column out of range", path);
        return M;
    }
};

static void hier_save(ramd_hier_s* h, const char* path)
{
    HierFile w;
    w.path = path;
    w.f = std::fopen(path, "wb");
    RAMD_CHECK(w.f, RAMD_E_ARG, "%s: cannot create", path);
    w.put_bytes(kHierMagic, 8);
    w.put<int32_t>(kHierVersion);
    const ramd_amg_opts& o = h->opts;
    w.put<int32_t>(o.solver);
    w.put<double>(o.strong_threshold);
    w.put<int32_t>(o.coarsen_type);
    w.put<int32_t>(o.interp_type);
    w.put<int32_t>(o.strength_type);
    w.put<double>(o.interp_filter);
    w.put<double>(o.prolong_weight);
    w.put<int32_t>(o.num_smooth_sweeps);
    w.put<double>(o.relax_weight);
    w.put<int32_t>(o.max_coarse);
    w.put<int32_t>(o.max_levels);
    w.put<int32_t>(o.prolong_smooth_steps);
    w.put<int32_t>(o.relax_type);
    for (double t : h->times) w.put<double>(t);
    w.put<int32_t>((int32_t)h->levels.size());
    for (Level& L : h->levels) {
        w.put_csr(download(L.A));
        w.put<int32_t>(L.P ? 1 : 0);
        if (L.P) w.put_csr(download(L.P));
        w.put_vec(L.states);
    }
    const int n = h->coarse_n;
    std::vector<double> lu((size_t)n * n);
    std::vector<int> piv((size_t)n);
    if (n) {
        RAMD_HIP(hipMemcpyAsync(lu.data(), h->lu, sizeof(double) * lu.size(), hipMemcpyDeviceToHost, h->ctx->stream));
        RAMD_HIP(hipMemcpyAsync(piv.data(), h->piv, sizeof(int) * piv.size(), hipMemcpyDeviceToHost, h->ctx->stream));
        RAMD_HIP(hipStreamSynchronize(h->ctx->stream));
    }
    w.put<int32_t>(n);
    w.put_vec(lu);
    w.put_vec(piv);
    RAMD_CHECK(std::fflush(w.f) == 0, RAMD_E_ARG, "%s: write failed", path);
}

static void hier_load(ramd_hier_s* h, const char* path)
{
    HierFile r;
    r.path = path;
    r.f = std::fopen(path, "rb");
    RAMD_CHECK(r.f, RAMD_E_ARG, "%s: cannot open", path);
    char magic[8];
    r.get_bytes(magic, 8);
    RAMD_CHECK(std::memcmp(magic, kHierMagic, 8) == 0, RAMD_E_ARG, "%s: not a hierarchy file", path);
    const int32_t version = r.get<int32_t>();
    RAMD_CHECK(version == 1 || version == kHierVersion, RAMD_E_ARG, "%s: version %d (expected 1 or %d)", path,
               version, kHierVersion);
    ramd_amg_opts& o = h->opts;
    std::memset(&o, 0, sizeof o);
    o.solver = r.get<int32_t>();
    o.strong_threshold = r.get<double>();
    o.coarsen_type = r.get<int32_t>();
    o.interp_type = r.get<int32_t>();
    o.strength_type = r.get<int32_t>();
    o.interp_filter = r.get<double>();
    o.prolong_weight = r.get<double>();
    o.num_smooth_sweeps = r.get<int32_t>();
    o.relax_weight = r.get<double>();
    o.max_coarse = r.get<int32_t>();
    o.max_levels = r.get<int32_t>();
    o.prolong_smooth_steps = r.get<int32_t>();
    o.relax_type = version >= 2 ? r.get<int32_t>() : RAMD_JACOBI;  // version 1 files predate the field
    o.weights = nullptr;
    RAMD_CHECK(o.num_smooth_sweeps >= 0 &&
                   (o.relax_type == RAMD_JACOBI || o.relax_type == RAMD_MC_SOR || o.relax_type == RAMD_MC_SSOR),
               RAMD_E_ARG, "%s: corrupt options", path);
    for (double& t : h->times) t = r.get<double>();
    const int32_t nl = r.get<int32_t>();
    RAMD_CHECK(nl >= 1 && nl <= 1000, RAMD_E_ARG, "%s: %d levels", path, nl);
    for (int l = 0; l < nl; l++) {
        Csr A = r.get_csr();
        RAMD_CHECK(A.n_rows == A.n_cols && (l == 0 || A.n_rows == h->levels.back().P->n_cols), RAMD_E_ARG,
                   "%s: level %d has the wrong size", path, l);
        h->levels.emplace_back();
        Level& L = h->levels.back();
        // only the fine level gets the stencil-diagonal form, as in the setup
        L.A = upload(h->ctx, A, l == 0);
        alloc_level_vectors(h, L);
        if (r.get<int32_t>()) {
            RAMD_CHECK(l + 1 < nl, RAMD_E_ARG, "%s: prolongator below the coarsest level", path);
            Csr P = r.get_csr();
            RAMD_CHECK(P.n_rows == A.n_rows, RAMD_E_ARG, "%s: level %d prolongator has the wrong size", path, l);
            L.P = upload(h->ctx, P, false);
            L.R = csr_transpose(L.P);
        } else {
            RAMD_CHECK(l + 1 == nl, RAMD_E_ARG, "%s: level %d has no prolongator", path, l);
        }
        std::vector<int> st = r.get_vec<int>(A.n_rows);
        L.states = std::move(st);
    }
    const int32_t n = r.get<int32_t>();
    RAMD_CHECK(n == h->levels.back().n && n <= kCoarseMax, RAMD_E_ARG, "%s: coarse size %d", path, n);
    std::vector<double> lu = r.get_vec<double>((int64_t)n * n);
    std::vector<int> piv = r.get_vec<int>(n);
    RAMD_CHECK((int64_t)lu.size() == (int64_t)n * n && (int)piv.size() == n, RAMD_E_ARG, "%s: corrupt coarse LU",
               path);
    for (int p : piv) RAMD_CHECK(p >= 0 && p < n, RAMD_E_ARG, "%s: corrupt pivots", path);
    // dgetrf's info: the first zero pivot of U (1-based), 0 when none
    h->coarse_info = 0;
    for (int j = 0; j < n && !h->coarse_info; j++)
        if (lu[(size_t)j + (size_t)j * n] == 0.0) h->coarse_info = j + 1;
    finish_hierarchy(h, n, lu, piv);
}

}  // namespace ramd

using namespace ramd;

extern "C" {

int ramd_amg_opts_default(int solver, ramd_amg_opts* o)
{
    return guarded([&] {
        RAMD_CHECK(o, RAMD_E_ARG, "ramd_amg_opts_default: NULL opts");
        RAMD_CHECK(solver == RAMD_SOLVER_RS || solver == RAMD_SOLVER_SA, RAMD_E_ARG, "unknown solver %d", solver);
        std::memset(o, 0, sizeof *o);
        o->solver = solver;
        if (solver == RAMD_SOLVER_RS) {
            // ParRugeStubenSolver(0.0, RS, Direct, Classical, SOR) defaults, relax set to Jacobi
            o->strong_threshold = 0.0;
            o->coarsen_type = RAMD_RS;
            o->interp_type = RAMD_DIRECT;
            o->strength_type = RAMD_CLASSICAL;
        } else {
            o->strong_threshold = 0.0;
            o->strength_type = RAMD_SYMMETRIC;
        }
        o->interp_filter = 0.3;
        o->prolong_weight = 4.0 / 3;
        o->num_smooth_sweeps = 1;
        o->relax_weight = 1.0;
        o->max_coarse = 50;
        o->max_levels = 25;
        o->weights = nullptr;
        o->prolong_smooth_steps = 1;
        o->relax_type = RAMD_JACOBI;
        o->device_setup = 1;
    });
}

int ramd_hier_setup(ramd_ctx ctx, int32_t n, int64_t nnz, const int32_t* rowptr, const int32_t* col,
                    const double* val, const ramd_amg_opts* opts, ramd_hier* out)
{
    return guarded([&] {
        RAMD_CHECK(ctx && out && opts && rowptr && (nnz == 0 || (col && val)), RAMD_E_ARG,
                   "ramd_hier_setup: NULL argument");
        RAMD_CHECK(n > 0 && nnz >= 0 && rowptr[0] == 0 && rowptr[n] == nnz, RAMD_E_ARG,
                   "ramd_hier_setup: inconsistent CSR (n=%d nnz=%lld)", n, (long long)nnz);
        RAMD_CHECK(opts->solver == RAMD_SOLVER_RS || opts->solver == RAMD_SOLVER_SA, RAMD_E_ARG,
                   "ramd_hier_setup: solver %d unknown", opts->solver);
        RAMD_CHECK(opts->num_smooth_sweeps >= 0 && opts->max_coarse >= 1 && opts->prolong_smooth_steps >= 0,
                   RAMD_E_ARG, "ramd_hier_setup: bad options");
        RAMD_CHECK(opts->relax_type == RAMD_JACOBI || opts->relax_type == RAMD_MC_SOR ||
                       opts->relax_type == RAMD_MC_SSOR,
                   RAMD_E_ARG,
                   "ramd_hier_setup: relax_type %d is not available on the device (RAMD_JACOBI, RAMD_MC_SOR, "
                   "RAMD_MC_SSOR; the reference's sequential SOR/SSOR is host-only)",
                   opts->relax_type);
        for (int64_t j = 0; j < nnz; j++)
            RAMD_CHECK(col[j] >= 0 && col[j] < n, RAMD_E_ARG, "ramd_hier_setup: column %d out of range", col[j]);
        ctx->set_current();
        Csr A;
        A.n_rows = A.n_cols = n;
        A.rowptr.assign(rowptr, rowptr + n + 1);
        A.col.assign(col, col + nnz);
        A.val.assign(val, val + nnz);
        std::unique_ptr<ramd_hier_s, void (*)(ramd_hier_s*)> h(new ramd_hier_s(), hier_free);
        h->ctx = ctx;
        h->opts = *opts;
        h->opts.weights = nullptr;
        if (opts->weights) {
            h->opts.weights = opts->weights;
        }
        hier_setup(h.get(), std::move(A));
        h->opts.weights = nullptr;
        *out = h.release();
    });
}

int ramd_sa_prolongation(ramd_ctx ctx, int32_t n, const int32_t* rowptr, const int32_t* col, const double* val,
                         int32_t n_aggs, const int32_t* t_rowptr, const int32_t* t_col, const double* t_val,
                         double omega, int steps, int32_t** p_rowptr, int32_t** p_col, double** p_val,
                         int64_t* p_nnz)
{
    return guarded([&] {
        RAMD_CHECK(ctx && rowptr && t_rowptr && p_rowptr && p_col && p_val && p_nnz && n >= 0 && n_aggs >= 0 &&
                       steps >= 0,
                   RAMD_E_ARG, "ramd_sa_prolongation: bad argument");
        auto take = [](int32_t nr, int32_t nc, const int32_t* rp, const int32_t* c, const double* v) {
            RAMD_CHECK(rp[0] == 0, RAMD_E_ARG, "ramd_sa_prolongation: rowptr[0] != 0");
            for (int i = 0; i < nr; i++)
                RAMD_CHECK(rp[i + 1] >= rp[i], RAMD_E_ARG, "ramd_sa_prolongation: rowptr decreases at %d", i);
            const int64_t nnz = rp[nr];
            RAMD_CHECK(nnz == 0 || (c && v), RAMD_E_ARG, "ramd_sa_prolongation: NULL entries");
            for (int64_t j = 0; j < nnz; j++)
                RAMD_CHECK(c[j] >= 0 && c[j] < nc, RAMD_E_ARG, "ramd_sa_prolongation: column %d out of range",
                           c[j]);
            Csr M;
            M.n_rows = nr;
            M.n_cols = nc;
            M.rowptr.assign(rp, rp + nr + 1);
            M.col.assign(c, c + nnz);
            M.val.assign(v, v + nnz);
            return M;
        };
        Csr A = take(n, n, rowptr, col, val);
        Csr T = take(n, n_aggs, t_rowptr, t_col, t_val);
        ctx->set_current();
        Csr P = sa_prolongation(ctx, A, std::move(T), omega, steps);
        auto out = [](const auto& v, auto** dst) {
            using E = typename std::decay_t<decltype(v)>::value_type;
            E* p = (E*)std::malloc(sizeof(E) * (v.size() ? v.size() : 1));
            if (!p) throw std::bad_alloc();
            if (!v.empty()) std::memcpy(p, v.data(), sizeof(E) * v.size());
            *dst = p;
        };
        out(P.rowptr, p_rowptr);
        out(P.col, p_col);
        out(P.val, p_val);
        *p_nnz = P.nnz();
    });
}

int ramd_hier_save(ramd_hier h, const char* path)
{
    return guarded([&] {
        RAMD_CHECK(h && path, RAMD_E_ARG, "ramd_hier_save: NULL argument");
        h->ctx->set_current();
        hier_save(h, path);
    });
}

int ramd_hier_load(ramd_ctx ctx, const char* path, ramd_hier* out)
{
    return guarded([&] {
        RAMD_CHECK(ctx && path && out, RAMD_E_ARG, "ramd_hier_load: NULL argument");
        ctx->set_current();
        std::unique_ptr<ramd_hier_s, void (*)(ramd_hier_s*)> h(new ramd_hier_s(), hier_free);
        h->ctx = ctx;
        hier_load(h.get(), path);
        *out = h.release();
    });
}

int ramd_hier_destroy(ramd_hier h)
{
    return guarded([&] { hier_free(h); });
}

int ramd_hier_info(ramd_hier h, int* num_levels, int32_t* n_rows, int64_t* nnz_A, int64_t* nnz_P)
{
    return guarded([&] {
        RAMD_CHECK(h, RAMD_E_ARG, "ramd_hier_info: NULL handle");
        const int L = (int)h->levels.size();
        if (num_levels) *num_levels = L;
        for (int l = 0; l < L; l++) {
            if (n_rows) n_rows[l] = h->levels[l].n;
            if (nnz_A) nnz_A[l] = h->levels[l].A->nnz;
            if (nnz_P) nnz_P[l] = h->levels[l].P ? h->levels[l].P->nnz : 0;
        }
    });
}

int ramd_hier_level(ramd_hier h, int level, ramd_csr* A, ramd_csr* P)
{
    return guarded([&] {
        RAMD_CHECK(h && level >= 0 && level < (int)h->levels.size(), RAMD_E_ARG, "ramd_hier_level: bad level %d",
                   level);
        if (A) *A = h->levels[level].A;
        if (P) *P = h->levels[level].P;
    });
}

int ramd_hier_restriction(ramd_hier h, int level, ramd_csr* R)
{
    return guarded([&] {
        RAMD_CHECK(h && R && level >= 0 && level < (int)h->levels.size(), RAMD_E_ARG,
                   "ramd_hier_restriction: bad argument");
        *R = h->levels[level].R;
    });
}

int ramd_hier_states(ramd_hier h, int level, int32_t* states)
{
    return guarded([&] {
        RAMD_CHECK(h && states && level >= 0 && level < (int)h->levels.size(), RAMD_E_ARG,
                   "ramd_hier_states: bad argument");
        const std::vector<int>& s = h->levels[level].states;
        RAMD_CHECK(!s.empty() || h->levels[level].n == 0, RAMD_E_STATE, "level %d has no splitting (coarsest)", level);
        std::copy(s.begin(), s.end(), states);
    });
}

int ramd_hier_row_caps(ramd_hier h, int32_t* caps)
{
    return guarded([&] {
        RAMD_CHECK(h && caps, RAMD_E_ARG, "ramd_hier_row_caps: NULL argument");
        for (size_t l = 0; l < h->levels.size(); l++) {
            const Level& L = h->levels[l];
            const ramd_csr_s* M[3] = {L.A, L.P, L.R};
            for (int m = 0; m < 3; m++)
                for (int p = 0; p < 2; p++)
                    caps[6 * l + 2 * m + p] =
                        (M[m] && !M[m]->dia)
                            ? (M[m]->wr_on[p]   ? 2
                               : M[m]->sl_on[p] ? (M[m]->sl16[p] ? -2 : -1)
                               : M[m]->lr_on[p] ? 1
                               : (M[m]->sg[p] && M[m]->sg[p]->blocks == M[m]->rb[p].d) ? -M[m]->rb[p].cap
                                                                                         : M[m]->rb[p].cap)
                            : 0;
        }
    });
}

int ramd_hier_setup_times(ramd_hier h, double* t)
{
    return guarded([&] {
        RAMD_CHECK(h && t, RAMD_E_ARG, "ramd_hier_setup_times: NULL argument");
        for (int i = 0; i < 5; i++) t[i] = h->times[i];
    });
}

int ramd_hier_coarse_info(ramd_hier h, int32_t* n_coarse, int32_t* info)
{
    return guarded([&] {
        RAMD_CHECK(h && info, RAMD_E_ARG, "ramd_hier_coarse_info: NULL argument");
        if (n_coarse) *n_coarse = h->coarse_n;
        *info = h->coarse_info;
    });
}

int ramd_hier_track_times(ramd_hier h, int on)
{
    return guarded([&] {
        RAMD_CHECK(h, RAMD_E_ARG, "ramd_hier_track_times: NULL handle");
        h->ctx->set_current();
        h->spans.resolve(h->solve_level_s);
        h->track = on != 0;
        h->solve_level_s.assign(h->levels.size(), 0.0);
    });
}

int ramd_hier_level_times(ramd_hier h, int phase, double* times)
{
    return guarded([&] {
        RAMD_CHECK(h && times && (phase == 0 || phase == 1), RAMD_E_ARG, "ramd_hier_level_times: bad argument");
        h->ctx->set_current();
        const size_t nl = h->levels.size();
        for (size_t i = 0; i < 5 * nl; i++) times[i] = 0.0;
        if (phase == 0) {
            for (size_t l = 0; l < nl && l < h->setup_level_s.size(); l++) times[5 * l] = h->setup_level_s[l];
        } else {
            h->solve_level_s.resize(nl, 0.0);
            h->spans.resolve(h->solve_level_s);
            for (size_t l = 0; l < nl; l++) times[5 * l] = h->solve_level_s[l];
        }
    });
}

int ramd_set_graphs(ramd_ctx ctx, int on)
{
    return guarded([&] {
        RAMD_CHECK(ctx, RAMD_E_ARG, "ramd_set_graphs: NULL context");
        ctx->graphs = on != 0;
    });
}

int ramd_hier_set_tail(ramd_hier h, int32_t level)
{
    return guarded([&] {
        RAMD_CHECK(h, RAMD_E_ARG, "ramd_hier_set_tail: NULL handle");
        if (level >= 1) {
            const std::vector<int> c = tail_candidates(h);
            RAMD_CHECK(std::find(c.begin(), c.end(), (int)level) != c.end(), RAMD_E_ARG,
                       "ramd_hier_set_tail: level %d cannot start the coarse tail", (int)level);
        }
        RAMD_CHECK(level >= -1, RAMD_E_ARG, "ramd_hier_set_tail: level %d", (int)level);
        h->ctx->set_current();
        RAMD_HIP(hipStreamSynchronize(h->ctx->stream));
        h->tail_mode = level;
    });
}

int ramd_hier_tail_level(ramd_hier h, int32_t prec, int32_t* level, int32_t* first)
{
    return guarded([&] {
        RAMD_CHECK(h && level && first, RAMD_E_ARG, "ramd_hier_tail_level: NULL argument");
        RAMD_CHECK(prec == RAMD_PREC_FP64 || prec == RAMD_PREC_FP32, RAMD_E_ARG, "ramd_hier_tail_level: precision %d",
                   (int)prec);
        const int p = prec == RAMD_PREC_FP32 ? 1 : 0;
        *level = (h->tail_mode == 0 && h->tail_pick[p] == -2) ? -2 : tail_level(h, p);
        const std::vector<int> c = tail_candidates(h);
        *first = c.empty() ? -1 : c[0];
    });
}

int ramd_hier_enable_fp32(ramd_hier h)
{
    return guarded([&] {
        RAMD_CHECK(h, RAMD_E_ARG, "ramd_hier_enable_fp32: NULL handle");
        h->ctx->set_current();
        enable_fp32(h);
    });
}

int ramd_vcycle(ramd_hier h, double* x_d, const double* b_d, int x_is_zero)
{
    return guarded([&] {
        RAMD_CHECK(h && x_d && b_d, RAMD_E_ARG, "ramd_vcycle: NULL argument");
        h->ctx->set_current();
        cycle_prepare(h, 0);
        cycle64(h, 0, x_d, b_d, x_is_zero != 0);
    });
}

int ramd_vcycle_fp32(ramd_hier h, float* x_d, const float* b_d)
{
    return guarded([&] {
        RAMD_CHECK(h && x_d && b_d, RAMD_E_ARG, "ramd_vcycle_fp32: NULL argument");
        h->ctx->set_current();
        cycle_prepare(h, 1);
        cycle32(h, 0, x_d, b_d, true);
    });
}

// ParMultilevel::solve multilevel/par_multilevel.hpp:459-538
int ramd_amg_solve(ramd_hier h, double* x_d, const double* b_d, double tol, int max_iter, double* residuals,
                   int32_t* iters)
{
    return guarded([&] {
        RAMD_CHECK(h && x_d && b_d && iters, RAMD_E_ARG, "ramd_amg_solve: NULL argument");
        ramd_ctx ctx = h->ctx;
        ctx->set_current();
        cycle_prepare(h, 0);
        Level& L0 = h->levels[0];
        auto norm = [&](const double* v) {
            launch_norm2sq(ctx, L0.n, v, ctx->red_result);
            double r = fetch_scalar(ctx, ctx->red_result);
            r = pow(r, 1.0 / 2);
            r = pow(r, 2);
            return pow(r, 1. / 2);
        };
        const double b_norm = norm(b_d);
        launch_stream_op(L0.A, OP_RESIDUAL, x_d, b_d, h->pr, 0.0, ctx->stream);
        double r_norm = std::fabs(b_norm) > 1e-16 ? norm(h->pr) / b_norm : norm(h->pr);
        int iter = 0;
        if (residuals) residuals[0] = r_norm;
        while (r_norm > tol && iter < max_iter) {
            cycle64(h, 0, x_d, b_d, false);
            iter++;
            launch_stream_op(L0.A, OP_RESIDUAL, x_d, b_d, h->pr, 0.0, ctx->stream);
            r_norm = std::fabs(b_norm) > 1e-16 ? norm(h->pr) / b_norm : norm(h->pr);
            if (residuals) residuals[iter] = r_norm;
        }
        *iters = iter;
    });
}

// PCG krylov/par_cg.cpp:121-263.  precond_t / comm_t (may be NULL) get the
// device seconds of the V-cycles and of the inner products added, the
// buckets the reference times with MPI_Wtime (par_cg.cpp:150-228).
static int pcg_run(ramd_hier h, ramd_csr A, double* x_d, const double* b_d, double tol, int max_iter, double* res,
                   int32_t max_res, int32_t* n_res, int32_t* iters, double* precond_t, double* comm_t)
{
    int rc = RAMD_OK;
    int st = guarded([&] {
        RAMD_CHECK(h && A && x_d && b_d && n_res && iters && (res || max_res == 0), RAMD_E_ARG,
                   "ramd_pcg: NULL argument");
        ramd_ctx ctx = h->ctx;
        ctx->set_current();
        cycle_prepare(h, 0);
        hipStream_t s = ctx->stream;
        const int n = A->n_rows;
        RAMD_CHECK(n == h->levels[0].n, RAMD_E_ARG, "ramd_pcg: A has %d rows, hierarchy %d", n, h->levels[0].n);
        double *r = h->pr, *z = h->pz, *p = h->pp, *Ap = h->pAp;
        int nres = 0;
        auto push = [&](double v) {
            if (nres < max_res) res[nres] = v;
            nres++;
        };
        // buckets: [0] preconditioner, [1] inner products
        const bool timed = precond_t || comm_t;
        EventSpans spans;
        std::vector<double> acc(2, 0.0);
        auto precondition = [&](double* zz, const double* rr) {
            const int t = timed ? spans.open(s, 0) : -1;
            vcycle_zero(h, zz, rr);
            if (t >= 0) spans.close(s, t);
        };
        auto inner = [&](const double* a, const double* b) {
            const int t = timed ? spans.open(s, 1) : -1;
            launch_dot(ctx, n, a, b, ctx->red_result);
            if (t >= 0) spans.close(s, t);
            return fetch_scalar(ctx, ctx->red_result);
        };
        try {
            if (max_iter <= 0) max_iter = ((int)(1.3 * n)) + 2;
            precondition(z, b_d);
            const double b_inner = inner(b_d, z);
            const double norm_b = sqrt(b_inner);
            if (norm_b > 1e-16) tol = tol * norm_b;
            launch_stream_op(A, OP_RESIDUAL, x_d, b_d, r, 0.0, s);
            precondition(z, r);
            RAMD_HIP(hipMemcpyAsync(p, z, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
            double rz = inner(r, z);
            push(sqrt(rz));
            int iter = 0;
            const int recompute_r = 8;
            while (iter < max_iter) {
                iter++;
                launch_stream_op(A, OP_SPMV, p, nullptr, Ap, 0.0, s);
                const double App = inner(Ap, p);
                if (App < 0.0) {
                    rc = RAMD_E_INDEFINITE;
                    break;
                }
                const double alpha = rz / App;
                const bool full_r = recompute_r && iter % recompute_r == 0;
                hipLaunchKernelGGL(pcg_update_xr, grid1(n), dim3(256), 0, s, n, alpha, p, Ap, x_d, r,
                                   full_r ? 0 : 1);
                RAMD_HIP(hipGetLastError());
                if (full_r) launch_stream_op(A, OP_RESIDUAL, x_d, b_d, r, 0.0, s);
                precondition(z, r);
                const double next = inner(r, z);
                const double beta = next / rz;
                push(next / b_inner);
                if (next < tol) break;
                if (full_r) {
                    RAMD_HIP(hipMemcpyAsync(p, z, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
                } else {
                    hipLaunchKernelGGL(pcg_update_p, grid1(n), dim3(256), 0, s, n, beta, z, p);
                    RAMD_HIP(hipGetLastError());
                }
                rz = next;
            }
            RAMD_HIP(hipStreamSynchronize(s));
            *n_res = nres;
            *iters = iter;
            if (timed) {
                spans.resolve(acc);
                if (precond_t) *precond_t += acc[0];
                if (comm_t) *comm_t += acc[1];
            }
        } catch (...) {
            spans.release();
            throw;
        }
        spans.release();
    });
    if (st != RAMD_OK) return st;
    if (rc == RAMD_E_INDEFINITE) set_error("ramd_pcg: indefinite matrix detected (p.Ap < 0)");
    return rc;
}

int ramd_pcg(ramd_hier h, ramd_csr A, double* x_d, const double* b_d, double tol, int max_iter, double* res,
             int32_t max_res, int32_t* n_res, int32_t* iters)
{
    return pcg_run(h, A, x_d, b_d, tol, max_iter, res, max_res, n_res, iters, nullptr, nullptr);
}

int ramd_pcg_timed(ramd_hier h, ramd_csr A, double* x_d, const double* b_d, double tol, int max_iter, double* res,
                   int32_t max_res, int32_t* n_res, int32_t* iters, double* precond_t, double* comm_t)
{
    return pcg_run(h, A, x_d, b_d, tol, max_iter, res, max_res, n_res, iters, precond_t, comm_t);
}

int ramd_cg(ramd_csr A, double* x_d, const double* b_d, double tol, int max_iter, double* res, int32_t max_res,
            int32_t* n_res, int32_t* iters)
{
    int rc = RAMD_OK;
    int st = guarded([&] {
        RAMD_CHECK(A && x_d && b_d && n_res && iters && (res || max_res == 0), RAMD_E_ARG, "ramd_cg: NULL argument");
        RAMD_CHECK(A->n_rows == A->n_cols, RAMD_E_ARG, "ramd_cg: A is not square");
        ramd_ctx ctx = A->ctx;
        ctx->set_current();
        hipStream_t s = ctx->stream;
        const int n = A->n_rows;
        struct Bufs {
            double *r = nullptr, *p = nullptr, *Ap = nullptr;
            ~Bufs()
            {
                dfree(r);
                dfree(p);
                dfree(Ap);
            }
        } w;
        w.r = dalloc<double>((size_t)std::max(n, 1));
        w.p = dalloc<double>((size_t)std::max(n, 1));
        w.Ap = dalloc<double>((size_t)std::max(n, 1));
        int nres = 0;
        auto push = [&](double v) {
            if (nres < max_res) res[nres] = v;
            nres++;
        };
        // b.norm(2): entries |v| <= 1e-16 skipped, root / power / root (core/vector.cpp:107-118)
        launch_norm2sq(ctx, n, b_d, ctx->red_result);
        double b_norm = fetch_scalar(ctx, ctx->red_result);
        b_norm = pow(pow(pow(b_norm, 1.0 / 2), 2), 1.0 / 2);
        if (b_norm < 1e-16) b_norm = 1.0;
        if (max_iter <= 0) max_iter = ((int)(1.3 * n)) + 2;
        launch_stream_op(A, OP_RESIDUAL, x_d, b_d, w.r, 0.0, s);
        RAMD_HIP(hipMemcpyAsync(w.p, w.r, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
        double rr = dot64(ctx, n, w.r, w.r);
        double norm_r = sqrt(rr);
        push(norm_r / b_norm);
        if (norm_r != 0.0) tol = tol * norm_r;
        const int recompute_r = 8;
        int iter = 0;
        while (norm_r > tol && iter < max_iter) {
            launch_stream_op(A, OP_SPMV, w.p, nullptr, w.Ap, 0.0, s);
            const double App = dot64(ctx, n, w.Ap, w.p);
            if (App < 0.0) {
                rc = RAMD_E_INDEFINITE;
                break;
            }
            const double alpha = rr / App;
            const bool update_r = (iter % recompute_r) && iter > 0;
            hipLaunchKernelGGL(pcg_update_xr, grid1(n), dim3(256), 0, s, n, alpha, w.p, w.Ap, x_d, w.r,
                               update_r ? 1 : 0);
            RAMD_HIP(hipGetLastError());
            if (!update_r) launch_stream_op(A, OP_RESIDUAL, x_d, b_d, w.r, 0.0, s);
            const double next = dot64(ctx, n, w.r, w.r);
            const double beta = next / rr;
            hipLaunchKernelGGL(pcg_update_p, grid1(n), dim3(256), 0, s, n, beta, w.r, w.p);
            RAMD_HIP(hipGetLastError());
            rr = next;
            norm_r = sqrt(rr);
            push(norm_r / b_norm);
            iter++;
        }
        RAMD_HIP(hipStreamSynchronize(s));
        *n_res = nres;
        *iters = iter;
    });
    if (st != RAMD_OK) return st;
    if (rc == RAMD_E_INDEFINITE) set_error("ramd_cg: indefinite matrix detected (A p . p < 0)");
    return rc;
}

// Flexible PCG (Polak-Ribiere beta = z_{k+1}.(r_{k+1} - r_k) / z_k.r_k) in FP64
// with a V-cycle preconditioner in `cycle_prec`; stops on the true FP64
// relative residual ||b - A x|| / ||b|| <= rtol.
int ramd_fpcg(ramd_hier h, ramd_csr A, double* x_d, const double* b_d, double rtol, int max_iter, int cycle_prec,
              double* relres_hist, int32_t* iters, double* final_relres)
{
    int rc = RAMD_OK;
    int st = guarded([&] {
        RAMD_CHECK(h && A && x_d && b_d && iters && final_relres, RAMD_E_ARG, "ramd_fpcg: NULL argument");
        RAMD_CHECK(cycle_prec == RAMD_PREC_FP64 || cycle_prec == RAMD_PREC_FP32, RAMD_E_ARG,
                   "ramd_fpcg: unknown precision %d", cycle_prec);
        rc = fpcg_run(h, A, x_d, b_d, rtol, max_iter, cycle_prec, relres_hist, iters, final_relres);
    });
    if (st != RAMD_OK) return st;
    fpcg_error("ramd_fpcg", rc, max_iter);
    return rc;
}

int ramd_fpcg_noalloc(ramd_hier h, ramd_csr A, double* x_d, const double* b_d, double rtol, int max_iter,
                      int cycle_prec, double* relres_hist, int32_t* iters, double* final_relres)
{
    int rc = RAMD_OK;
    int st = guarded([&] {
        RAMD_CHECK(h && A && x_d && b_d && iters && final_relres, RAMD_E_ARG, "ramd_fpcg_noalloc: NULL argument");
        RAMD_CHECK(cycle_prec == RAMD_PREC_FP64 || cycle_prec == RAMD_PREC_FP32, RAMD_E_ARG,
                   "ramd_fpcg_noalloc: unknown precision %d", cycle_prec);
        rc = fpcg_run(h, A, x_d, b_d, rtol, max_iter, cycle_prec, relres_hist, iters, final_relres, true);
    });
    if (st != RAMD_OK) return st;
    fpcg_error("ramd_fpcg_noalloc", rc, max_iter);
    return rc;
}

int ramd_adjoint(ramd_hier h, ramd_csr A, const double* x_d, const double* g_d, double rtol, int max_iter,
                 int cycle_prec, double* dA_d, double* db_d, int32_t* iters, double* final_relres)
{
    int rc = RAMD_OK;
    int st = guarded([&] {
        RAMD_CHECK(h && A && x_d && g_d && db_d && iters && final_relres, RAMD_E_ARG, "ramd_adjoint: NULL argument");
        RAMD_CHECK(cycle_prec == RAMD_PREC_FP64 || cycle_prec == RAMD_PREC_FP32, RAMD_E_ARG,
                   "ramd_adjoint: unknown precision %d", cycle_prec);
        RAMD_CHECK(A->n_rows == A->n_cols, RAMD_E_ARG, "ramd_adjoint: A is not square");
        hipStream_t s = h->ctx->stream;
        h->ctx->set_current();
        // A^T lambda = dL/dx; A is symmetric (the SPD operator the hierarchy was built for)
        RAMD_HIP(hipMemsetAsync(db_d, 0, sizeof(double) * A->n_rows, s));
        rc = fpcg_run(h, A, db_d, g_d, rtol, max_iter, cycle_prec, nullptr, iters, final_relres);
        if (rc == RAMD_OK && dA_d && A->n_rows) {
            hipLaunchKernelGGL(adjoint_outer, grid1(A->n_rows), dim3(256), 0, s, A->n_rows, A->rowptr, A->col, db_d,
                               x_d, dA_d);
            RAMD_HIP(hipGetLastError());
        }
        RAMD_HIP(hipStreamSynchronize(s));
    });
    if (st != RAMD_OK) return st;
    fpcg_error("ramd_adjoint", rc, max_iter);
    return rc;
}

}  // extern "C"

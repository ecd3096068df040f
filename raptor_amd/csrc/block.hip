// Multi-vector path: k right-hand sides (k <= 64) held row-interleaved
// (entry (i, c) at i*k + c), one V-cycle for all of them per launch, and a
// lockstep flexible PCG with one set of scalars per right-hand side.
//
// Every row is handled by a group of LPR lanes; lane g of the group owns
// columns g, g+LPR, ...  Each column's sum runs over the row in stored order
// with exactly the single-vector kernels' operation sequence, so column c
// of a block V-cycle is bit-identical to the single-vector V-cycle applied
// to column c.  The matrix entries are read once per row for all k columns.
//
// LPR is picked per matrix by a small autotuner (timed trial launches on
// the real operator, see block_tune) and cached in the hierarchy.
//
// The flexible-PCG iteration is split into two captured HIP graphs (the
// operator step with the convergence test, then the preconditioner step);
// all scalars live on the device, and one 16-byte read per iteration tells
// the host whether to continue.  No memset node is recorded: buffers are
// cleared by kernels.
#include <cmath>
#include <cstring>
#include <functional>

#include "hier.hpp"

namespace ramd {

constexpr int kMaxRhs = 64;

// ---------------------------------------------------------------------------
// SpMM family (CSR and stencil-diagonal forms)
// ---------------------------------------------------------------------------
template <int OP, typename V, typename X, int LPR>
__global__ __launch_bounds__(256) void bcsr_kernel(int n, int k, const int* __restrict__ rp,
                                                   const int* __restrict__ col, const V* __restrict__ val,
                                                   const X* __restrict__ x, const X* __restrict__ b,
                                                   X* __restrict__ y, X omega)
{
    constexpr int RPW = 256 / LPR;
    constexpr int CPL = kMaxRhs / LPR;  // columns per lane (upper bound)
    const int row = xcd_remap(blockIdx.x, gridDim.x) * RPW + threadIdx.x / LPR;
    const int g = threadIdx.x % LPR;
    if (row >= n || g >= k) return;
    const int s = rp[row], e = rp[row + 1];
    const size_t o = (size_t)row * k;
    X acc[CPL], diag = 0;
#pragma unroll
    for (int q = 0; q < CPL; q++) {
        const int c = g + q * LPR;
        acc[q] = (OP == OP_RESIDUAL && c < k) ? b[o + c] : X(0);
    }
    for (int j = s; j < e; j++) {
        const int cj = col[j];
        const X a = X(val[j]);
        const size_t oj = (size_t)cj * k;
        if (OP == OP_JACOBI && cj == row) {
            diag = a;
            continue;
        }
#pragma unroll
        for (int q = 0; q < CPL; q++) {
            const int c = g + q * LPR;
            if (c < k) {
                const X p = a * x[oj + c];
                if (OP == OP_RESIDUAL) acc[q] -= p;
                else acc[q] += p;
            }
        }
    }
#pragma unroll
    for (int q = 0; q < CPL; q++) {
        const int c = g + q * LPR;
        if (c >= k) continue;
        if (OP == OP_SPMV || OP == OP_RESIDUAL) y[o + c] = acc[q];
        else if (OP == OP_APPEND) y[o + c] += acc[q];
        else if (OP == OP_ADD_INTO) y[o + c] = b[o + c] + acc[q];
        else {
            const X xo = x[o + c];
            y[o + c] = (s != e && fabs(diag) > X(1e-16))
                           ? ((X(1.0) - omega) * xo) + (omega * ((b[o + c] - acc[q]) / diag))
                           : xo;
        }
    }
}

template <int OP, typename V, typename X, int LPR, int ND>
__global__ __launch_bounds__(256) void bdia_kernel(int n, int k, DiaArgs a, int nd, int center,
                                                   const unsigned* __restrict__ mask, const V* __restrict__ val,
                                                   const X* __restrict__ x, const X* __restrict__ b,
                                                   X* __restrict__ y, X omega)
{
    constexpr int RPW = 256 / LPR;
    constexpr int CPL = kMaxRhs / LPR;
    const int row = xcd_remap(blockIdx.x, gridDim.x) * RPW + threadIdx.x / LPR;
    const int g = threadIdx.x % LPR;
    if (row >= n || g >= k) return;
    const unsigned m = mask[row];
    const size_t o = (size_t)row * k;
    X acc[CPL], diag = 0;
#pragma unroll
    for (int q = 0; q < CPL; q++) {
        const int c = g + q * LPR;
        acc[q] = (OP == OP_RESIDUAL && c < k) ? b[o + c] : X(0);
    }
#pragma unroll
    for (int kk = 0; kk < ND; kk++) {
        const bool live = (ND == kDiaMax) ? (kk < nd) : true;
        const int d = live ? a.ord[kk] : 0;
        const bool on = live && ((m >> d) & 1u);
        const X v = X(val[dia_index<V>(live ? kk : 0, row, n)]);
        const size_t oj = (size_t)dia_col(row, a.off[d], n) * k;
        if (OP == OP_JACOBI && d == center) {
            diag = on ? v : diag;
            continue;
        }
#pragma unroll
        for (int q = 0; q < CPL; q++) {
            const int c = g + q * LPR;
            if (c < k) {
                const X p = v * x[oj + c];
                if (OP == OP_RESIDUAL) acc[q] = on ? acc[q] - p : acc[q];
                else acc[q] = on ? acc[q] + p : acc[q];
            }
        }
    }
#pragma unroll
    for (int q = 0; q < CPL; q++) {
        const int c = g + q * LPR;
        if (c >= k) continue;
        if (OP == OP_SPMV || OP == OP_RESIDUAL) y[o + c] = acc[q];
        else if (OP == OP_APPEND) y[o + c] += acc[q];
        else if (OP == OP_ADD_INTO) y[o + c] = b[o + c] + acc[q];
        else {
            const X xo = x[o + c];
            y[o + c] = (m != 0u && fabs(diag) > X(1e-16))
                           ? ((X(1.0) - omega) * xo) + (omega * ((b[o + c] - acc[q]) / diag))
                           : xo;
        }
    }
}

// ---------------------------------------------------------------------------
// Vectorized SpMM: the k columns of a row are read and written as 16-byte
// units of VW = 16 / sizeof(X) consecutive columns (k a multiple of VW); lane
// g of a row's LPR lanes owns units g, g + LPR, ... (UPL of them, the last
// possibly past the row's units: it then re-reads unit g and stores
// nothing).  A lane group's load of one x row is one contiguous run of whole
// cache lines, and every load of a step (matrix entries, then x units) is
// issued before the first use.  Every column runs over the row in stored
// order with the scalar kernels' operation sequence, so the bits are the
// same.
// ---------------------------------------------------------------------------
constexpr int kMaxUnits = 2;

template <typename X>
struct Unit16;
template <>
struct Unit16<float> {
    using T = float4;
};
template <>
struct Unit16<double> {
    using T = double2;
};

template <typename X>
__device__ __forceinline__ void unit_load(const X* __restrict__ p, X (&v)[16 / sizeof(X)])
{
    using XV = typename Unit16<X>::T;
    const XV t = *reinterpret_cast<const XV*>(p);
    const X* pt = reinterpret_cast<const X*>(&t);
#pragma unroll
    for (int i = 0; i < (int)(16 / sizeof(X)); i++) v[i] = pt[i];
}

template <typename X>
__device__ __forceinline__ void unit_store(X* __restrict__ p, const X (&v)[16 / sizeof(X)])
{
    using XV = typename Unit16<X>::T;
    XV t;
    X* pt = reinterpret_cast<X*>(&t);
#pragma unroll
    for (int i = 0; i < (int)(16 / sizeof(X)); i++) pt[i] = v[i];
    *reinterpret_cast<XV*>(p) = t;
}

// A lane's units: column offset of each (a valid one for every u) and
// whether it exists.
template <typename X, int LPR, int UPL>
struct Units {
    static constexpr int VW = 16 / sizeof(X);
    int off[UPL];
    bool on[UPL];
    __device__ __forceinline__ Units(int g, int units)
    {
#pragma unroll
        for (int u = 0; u < UPL; u++) {
            const int un = g + u * LPR;
            on[u] = un < units;
            off[u] = (on[u] ? un : g) * VW;
        }
    }
};

// accumulators <- b (residual) or 0
template <int OP, typename X, int LPR, int UPL>
__device__ __forceinline__ void bvec_init(X (&acc)[UPL][16 / sizeof(X)], const Units<X, LPR, UPL>& U,
                                          const X* __restrict__ brow)
{
    constexpr int VW = 16 / sizeof(X);
#pragma unroll
    for (int u = 0; u < UPL; u++) {
        if (OP == OP_RESIDUAL) {
            unit_load<X>(brow + U.off[u], acc[u]);
        } else {
#pragma unroll
            for (int t = 0; t < VW; t++) acc[u][t] = X(0);
        }
    }
}

// y row from the accumulators (Jacobi: usable = the row has a usable diagonal)
template <int OP, typename X, int LPR, int UPL>
__device__ __forceinline__ void bvec_finish(const X (&acc)[UPL][16 / sizeof(X)], const Units<X, LPR, UPL>& U,
                                            const X* __restrict__ xrow, const X* __restrict__ brow,
                                            X* __restrict__ yrow, X diag, bool usable, X omega)
{
    constexpr int VW = 16 / sizeof(X);
#pragma unroll
    for (int u = 0; u < UPL; u++) {
        if (!U.on[u]) continue;
        X out[VW];
        if (OP == OP_SPMV || OP == OP_RESIDUAL) {
#pragma unroll
            for (int t = 0; t < VW; t++) out[t] = acc[u][t];
        } else if (OP == OP_ADD_INTO) {
            X bv[VW];
            unit_load<X>(brow + U.off[u], bv);
#pragma unroll
            for (int t = 0; t < VW; t++) out[t] = bv[t] + acc[u][t];
        } else {
            X xo[VW], bv[VW];
            unit_load<X>(xrow + U.off[u], xo);
            unit_load<X>(brow + U.off[u], bv);
#pragma unroll
            for (int t = 0; t < VW; t++)
                out[t] = usable ? ((X(1.0) - omega) * xo[t]) + (omega * ((bv[t] - acc[u][t]) / diag)) : xo[t];
        }
        unit_store<X>(yrow + U.off[u], out);
    }
}

template <int OP, typename V, typename X, int LPR, int UPL>
__global__ __launch_bounds__(256) void bcsr_vec_kernel(int n, int k, const int* __restrict__ rp,
                                                       const int* __restrict__ col, const V* __restrict__ val,
                                                       const X* __restrict__ x, const X* __restrict__ b,
                                                       X* __restrict__ y, X omega)
{
    constexpr int VW = 16 / sizeof(X);
    constexpr int RPW = 256 / LPR;
    constexpr int EB = 8 / UPL;  // entries per step
    const int row = xcd_remap(blockIdx.x, gridDim.x) * RPW + threadIdx.x / LPR;
    const int g = threadIdx.x % LPR;
    const int units = k / VW;
    if (row >= n || g >= units) return;
    const Units<X, LPR, UPL> U(g, units);
    const int s = rp[row], e = rp[row + 1];
    const size_t o = (size_t)row * k;
    X acc[UPL][VW], diag = 0;
    bvec_init<OP, X, LPR, UPL>(acc, U, b + o);
    // entries of the first step; each step then issues its x loads and the
    // next step's entry loads before adding (entries past the row read
    // column 0 and are skipped)
    int cj[EB];
    X a[EB];
#pragma unroll
    for (int t = 0; t < EB; t++) {
        const bool in = s + t < e;
        cj[t] = in ? col[s + t] : 0;
        a[t] = in ? X(val[s + t]) : X(0);
    }
    for (int j0 = s; j0 < e; j0 += EB) {
        X xv[EB][UPL][VW];
#pragma unroll
        for (int t = 0; t < EB; t++)
#pragma unroll
            for (int u = 0; u < UPL; u++) unit_load<X>(x + (size_t)cj[t] * k + U.off[u], xv[t][u]);
        int cn[EB];
        X an[EB];
#pragma unroll
        for (int t = 0; t < EB; t++) {
            const bool in = j0 + EB + t < e;
            cn[t] = in ? col[j0 + EB + t] : 0;
            an[t] = in ? X(val[j0 + EB + t]) : X(0);
        }
#pragma unroll
        for (int t = 0; t < EB; t++) {
            if (j0 + t >= e) break;
            if (OP == OP_JACOBI && cj[t] == row) {
                diag = a[t];
                continue;
            }
#pragma unroll
            for (int u = 0; u < UPL; u++)
#pragma unroll
                for (int v = 0; v < VW; v++) {
                    const X p = a[t] * xv[t][u][v];
                    if (OP == OP_RESIDUAL) acc[u][v] -= p;
                    else acc[u][v] += p;
                }
        }
#pragma unroll
        for (int t = 0; t < EB; t++) {
            cj[t] = cn[t];
            a[t] = an[t];
        }
    }
    bvec_finish<OP, X, LPR, UPL>(acc, U, x + o, b + o, y + o, diag, s != e && fabs(diag) > X(1e-16), omega);
}

template <int OP, typename V, typename X, int LPR, int UPL, int ND>
__global__ __launch_bounds__(256) void bdia_vec_kernel(int n, int k, DiaArgs a, int nd, int center,
                                                       const unsigned* __restrict__ mask, const V* __restrict__ val,
                                                       const X* __restrict__ x, const X* __restrict__ b,
                                                       X* __restrict__ y, X omega)
{
    constexpr int VW = 16 / sizeof(X);
    constexpr int RPW = 256 / LPR;
    constexpr int DCH = 9;  // diagonals per step
    const int row = xcd_remap(blockIdx.x, gridDim.x) * RPW + threadIdx.x / LPR;
    const int g = threadIdx.x % LPR;
    const int units = k / VW;
    if (row >= n || g >= units) return;
    const Units<X, LPR, UPL> U(g, units);
    const unsigned m = mask[row];
    const size_t o = (size_t)row * k;
    X acc[UPL][VW], diag = 0;
    bvec_init<OP, X, LPR, UPL>(acc, U, b + o);
#pragma unroll
    for (int c0 = 0; c0 < ND; c0 += DCH) {
        X vv[DCH];
        int dd[DCH];
        X xv[DCH][UPL][VW];
#pragma unroll
        for (int t = 0; t < DCH; t++) {
            const int kk = c0 + t;
            if (kk < ND) {
                const bool live = (ND == kDiaMax) ? (kk < nd) : true;
                dd[t] = live ? a.ord[kk] : -1;
                vv[t] = X(val[dia_index<V>(live ? kk : 0, row, n)]);
                const X* xr = x + (size_t)dia_col(row, a.off[live ? dd[t] : 0], n) * k;
#pragma unroll
                for (int u = 0; u < UPL; u++) unit_load<X>(xr + U.off[u], xv[t][u]);
            }
        }
#pragma unroll
        for (int t = 0; t < DCH; t++) {
            if (c0 + t >= ND) break;
            const int d = dd[t];
            const bool on = d >= 0 && ((m >> d) & 1u);
            if (OP == OP_JACOBI && d == center) {
                diag = on ? vv[t] : diag;
                continue;
            }
#pragma unroll
            for (int u = 0; u < UPL; u++)
#pragma unroll
                for (int v = 0; v < VW; v++) {
                    const X p = vv[t] * xv[t][u][v];
                    if (OP == OP_RESIDUAL) acc[u][v] = on ? acc[u][v] - p : acc[u][v];
                    else acc[u][v] = on ? acc[u][v] + p : acc[u][v];
                }
        }
    }
    bvec_finish<OP, X, LPR, UPL>(acc, U, x + o, b + o, y + o, diag, m != 0u && fabs(diag) > X(1e-16), omega);
}

template <int OP, typename V, typename X, int LPR, int UPL>
static void blaunch_vec_op(const ramd_csr_s* A, const V* val, const V* dval, int k, const X* x, const X* b, X* y,
                           X omega, hipStream_t s)
{
    const int n = A->n_rows;
    if (n == 0) return;
    const dim3 grid((unsigned)((n + 256 / LPR - 1) / (256 / LPR))), block(256);
    if (dval) {
        const Dia* d = A->dia;
#define BDIAV(NDV)                                                                                            \
    hipLaunchKernelGGL((bdia_vec_kernel<OP, V, X, LPR, UPL, NDV>), grid, block, 0, s, n, k, d->args, d->nd, \
                       d->center, d->mask, dval, x, b, y, omega)
        if (d->nd == 27) BDIAV(27);
        else if (d->nd == 9) BDIAV(9);
        else if (d->nd == 7) BDIAV(7);
        else BDIAV(kDiaMax);
#undef BDIAV
    } else {
        hipLaunchKernelGGL((bcsr_vec_kernel<OP, V, X, LPR, UPL>), grid, block, 0, s, n, k, A->rowptr, A->col, val,
                           x, b, y, omega);
    }
    RAMD_HIP(hipGetLastError());
}

template <typename V, typename X, int LPR>
static void blaunch_vec_lpr(const ramd_csr_s* A, const V* val, const V* dval, int op, int k, const X* x,
                            const X* b, X* y, X omega, hipStream_t s)
{
    constexpr int VW = 16 / sizeof(X);
    const int upl = (k / VW + LPR - 1) / LPR;
#define BVEC(UPLV)                                                                                            \
    switch (op) {                                                                                            \
    case OP_SPMV: blaunch_vec_op<OP_SPMV, V, X, LPR, UPLV>(A, val, dval, k, x, b, y, omega, s); break;       \
    case OP_RESIDUAL: blaunch_vec_op<OP_RESIDUAL, V, X, LPR, UPLV>(A, val, dval, k, x, b, y, omega, s); break; \
    case OP_JACOBI: blaunch_vec_op<OP_JACOBI, V, X, LPR, UPLV>(A, val, dval, k, x, b, y, omega, s); break;   \
    case OP_ADD_INTO: blaunch_vec_op<OP_ADD_INTO, V, X, LPR, UPLV>(A, val, dval, k, x, b, y, omega, s); break; \
    default: fail(RAMD_E_ARG, "block op %d not supported", op);                                              \
    }
    if (upl == 1) {
        BVEC(1)
    } else if (upl == 2) {
        BVEC(2)
    } else {
        fail(RAMD_E_ARG, "k = %d needs %d units per lane at %d lanes per row", k, upl, LPR);
    }
#undef BVEC
}

template <int OP, typename V, typename X, int LPR>
static void blaunch_op(const ramd_csr_s* A, const V* val, const V* dval, int k, const X* x, const X* b, X* y,
                       X omega, hipStream_t s)
{
    const int n = A->n_rows;
    if (n == 0) return;
    const dim3 grid((unsigned)((n + 256 / LPR - 1) / (256 / LPR))), block(256);
    if (dval) {
        const Dia* d = A->dia;
#define BDIA(NDV)                                                                                                  \
    hipLaunchKernelGGL((bdia_kernel<OP, V, X, LPR, NDV>), grid, block, 0, s, n, k, d->args, d->nd, d->center, \
                       d->mask, dval, x, b, y, omega)
        if (d->nd == 27) BDIA(27);
        else if (d->nd == 9) BDIA(9);
        else if (d->nd == 7) BDIA(7);
        else BDIA(kDiaMax);
#undef BDIA
    } else {
        hipLaunchKernelGGL((bcsr_kernel<OP, V, X, LPR>), grid, block, 0, s, n, k, A->rowptr, A->col, val, x, b, y,
                           omega);
    }
    RAMD_HIP(hipGetLastError());
}

template <typename V, typename X, int LPR>
static void blaunch_lpr(const ramd_csr_s* A, const V* val, const V* dval, int op, int k, const X* x, const X* b,
                        X* y, X omega, hipStream_t s)
{
    switch (op) {
    case OP_SPMV: blaunch_op<OP_SPMV, V, X, LPR>(A, val, dval, k, x, b, y, omega, s); break;
    case OP_RESIDUAL: blaunch_op<OP_RESIDUAL, V, X, LPR>(A, val, dval, k, x, b, y, omega, s); break;
    case OP_JACOBI: blaunch_op<OP_JACOBI, V, X, LPR>(A, val, dval, k, x, b, y, omega, s); break;
    case OP_ADD_INTO: blaunch_op<OP_ADD_INTO, V, X, LPR>(A, val, dval, k, x, b, y, omega, s); break;
    default: fail(RAMD_E_ARG, "block op %d not supported", op);
    }
}

// lanes per row: 8, 16, 32 or 64 (at least k/8, since a lane holds <= 8 columns)
static bool lpr_ok(int lpr, int k) { return lpr * 8 >= k && (lpr <= 8 || lpr / 2 < k); }

// vectorized kernel with lpr lanes per row for k columns of X: whole 16-byte
// units, at most kMaxUnits per lane, no lane without a unit
template <typename X>
static bool vec_ok(int lpr, int k)
{
    constexpr int VW = 16 / sizeof(X);
    if (k % VW) return false;
    const int units = k / VW;
    return lpr <= units && (units + lpr - 1) / lpr <= kMaxUnits;
}

static bool aligned16(const void* p) { return ((uintptr_t)p & 15u) == 0; }

// code > 0: scalar kernel with `code` lanes per row; code < 0: vectorized
// kernel with -code lanes per row (falls back to 32 scalar lanes when the
// vectors are not 16-byte aligned)
template <typename X>
static void blaunch(const ramd_csr_s* A, int code, int op, int k, const X* x, const X* b, X* y, X omega,
                    hipStream_t s)
{
    if (code < 0 && !(aligned16(x) && aligned16(y) && (!b || aligned16(b)))) code = lpr_ok(32, k) ? 32 : 64;
    // FP64 vectors read FP64 values; FP32 vectors read the FP32 copies
    using V = X;
    const V* val;
    const V* dval;
    if constexpr (sizeof(X) == 8) {
        val = (const V*)A->val;
        dval = A->dia ? (const V*)A->dia->v64 : nullptr;
    } else {
        RAMD_CHECK(A->val32 && (!A->dia || A->dia->v32), RAMD_E_STATE, "FP32 copies of the matrix were not built");
        val = (const V*)A->val32;
        dval = A->dia ? (const V*)A->dia->v32 : nullptr;
    }
    switch (code) {
    case 8: blaunch_lpr<V, X, 8>(A, val, dval, op, k, x, b, y, omega, s); break;
    case 16: blaunch_lpr<V, X, 16>(A, val, dval, op, k, x, b, y, omega, s); break;
    case 32: blaunch_lpr<V, X, 32>(A, val, dval, op, k, x, b, y, omega, s); break;
    case 64: blaunch_lpr<V, X, 64>(A, val, dval, op, k, x, b, y, omega, s); break;
    case -4: blaunch_vec_lpr<V, X, 4>(A, val, dval, op, k, x, b, y, omega, s); break;
    case -8: blaunch_vec_lpr<V, X, 8>(A, val, dval, op, k, x, b, y, omega, s); break;
    case -16: blaunch_vec_lpr<V, X, 16>(A, val, dval, op, k, x, b, y, omega, s); break;
    default: fail(RAMD_E_ARG, "block kernel code %d", code);
    }
}

// ---------------------------------------------------------------------------
// elementwise and per-column kernels.  Layout of the per-column kernels:
// with kp = the power of two >= k, thread t of a workgroup owns column
// t % kp and row offset t / kp; rows advance by (256 / kp) * gridDim.x.
// Sums are combined in a fixed order.
// ---------------------------------------------------------------------------
constexpr int kColBlocks = 4096;

// first Jacobi sweep from X = 0 (the single-vector jacobi_from_zero per column)
// The per-column kernels take kColUnroll rows per thread per step (rows
// i, i + stride, ...) so that many independent loads are in flight.
constexpr int kColUnroll = 4;

template <typename X>
__global__ __launch_bounds__(256) void bjacobi_from_zero(int n, int k, int kp, const X* __restrict__ d,
                                                         const X* __restrict__ b, X* __restrict__ x, X omega)
{
    const int c = threadIdx.x & (kp - 1);
    const int rpb = 256 / kp;
    if (c >= k) return;
    const X zero = X(0.0);
    const int64_t stride = (int64_t)gridDim.x * rpb;
    for (int64_t i0 = (int64_t)blockIdx.x * rpb + threadIdx.x / kp; i0 < n; i0 += kColUnroll * stride) {
        X di[kColUnroll], bi[kColUnroll];
#pragma unroll
        for (int u = 0; u < kColUnroll; u++) {
            const int64_t i = i0 + u * stride;
            di[u] = i < n ? d[i] : X(0);  // 0 for a row without a stored diagonal
            bi[u] = i < n ? b[(size_t)i * k + c] : X(0);
        }
#pragma unroll
        for (int u = 0; u < kColUnroll; u++) {
            const int64_t i = i0 + u * stride;
            if (i >= n) break;
            X v = zero;
            if (fabs(di[u]) > X(1e-16)) v = ((X(1.0) - omega) * zero) + (omega * ((bi[u] - zero) / di[u]));
            x[(size_t)i * k + c] = v;
        }
    }
}

template <typename A, typename B>
__global__ void bconvert(int64_t m, const A* __restrict__ a, B* __restrict__ b)
{
    const int64_t s = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += s) b[i] = (B)a[i];
}

__global__ void bfill(int64_t m, double v, double* __restrict__ a)
{
    const int64_t s = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += s) a[i] = v;
}

__global__ void bfill32(int64_t m, float v, float* __restrict__ a)
{
    const int64_t s = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += s) a[i] = v;
}

// dgetrs_('N') for column blockIdx.x of a row-interleaved block (see coarse_solve_kernel)
// Column c of the block by one wavefront (n <= kCoarseWave).
template <typename X>
__global__ __launch_bounds__(64) void bcoarse_wave(int n, int k, const double* __restrict__ lu,
                                                   const int* __restrict__ piv, const X* __restrict__ b,
                                                   X* __restrict__ x)
{
    extern __shared__ double ls[];
    const int c = blockIdx.x;
    coarse_solve_wave<X, X>(n, lu, piv, b + c, x + c, (size_t)k, ls);
}

template <typename X>
__global__ __launch_bounds__(256) void bcoarse_solve(int n, int k, const double* __restrict__ lu,
                                                     const int* __restrict__ piv, const X* __restrict__ b,
                                                     X* __restrict__ x)
{
    extern __shared__ double sb[];
    const int c = blockIdx.x;
    for (int i = threadIdx.x; i < n; i += blockDim.x) sb[i] = (double)b[(size_t)i * k + c];
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
    for (int i = threadIdx.x; i < n; i += blockDim.x) x[(size_t)i * k + c] = (X)sb[i];
}

static int pow2_at_least(int k)
{
    int p = 1;
    while (p < k) p <<= 1;
    return p;
}

// workgroups of the per-column kernels for n rows
static int col_blocks(int n, int kp) { return std::max(1, std::min(kColBlocks, (n + 256 / kp - 1) / (256 / kp))); }

// Combines the 256 / kp row-offset sums of each column in LDS (in offset
// order) and stores the workgroup's partial for column c at
// partials[blockIdx.x * k + c].
__device__ __forceinline__ void col_partial(double v, int k, int kp, double* __restrict__ partials)
{
    __shared__ double sh[256];
    sh[threadIdx.x] = v;
    __syncthreads();
    if (threadIdx.x < k) {
        double t = sh[threadIdx.x];
        for (int o = kp; o < 256; o += kp) t += sh[threadIdx.x + o];
        partials[(size_t)blockIdx.x * k + threadIdx.x] = t;
    }
    __syncthreads();
}

// per-column a.b (MODE 0) or r.z and z.w (MODE 1: two partial arrays)
template <int MODE, typename Z>
__global__ __launch_bounds__(256) void bcol_dots(int n, int k, int kp, const double* __restrict__ a,
                                                 const Z* __restrict__ z, const double* __restrict__ w,
                                                 double* __restrict__ partials)
{
    const int c = threadIdx.x & (kp - 1);
    const int rpb = 256 / kp;
    double s1 = 0, s2 = 0;
    const int64_t stride = (int64_t)gridDim.x * rpb;
    if (c < k)
        for (int64_t i0 = (int64_t)blockIdx.x * rpb + threadIdx.x / kp; i0 < n; i0 += kColUnroll * stride) {
            double av[kColUnroll], zv[kColUnroll], wv[kColUnroll];
#pragma unroll
            for (int u = 0; u < kColUnroll; u++) {
                const int64_t i = i0 + u * stride;
                const size_t e = (size_t)(i < n ? i : 0) * k + c;
                av[u] = i < n ? a[e] : 0.0;
                zv[u] = i < n ? (double)z[e] : 0.0;
                wv[u] = (MODE == 1 && i < n) ? w[e] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < kColUnroll; u++) {
                if (i0 + u * stride >= n) break;
                s1 += av[u] * zv[u];
                if (MODE == 1) s2 += zv[u] * wv[u];
            }
        }
    col_partial(s1, k, kp, partials);
    if (MODE == 1) col_partial(s2, k, kp, partials + (size_t)gridDim.x * k);
}

// r -= alpha Ap; r32 = (float) r; per-column sums of r^2.  The step
// x += alpha p of the iteration is taken by the direction update (bdir),
// which reads p anyway, or by bstep_x when the iteration ends there.
__global__ __launch_bounds__(256) void bupdate(int n, int k, int kp, const double* __restrict__ alpha,
                                               const double* __restrict__ Ap, double* __restrict__ r,
                                               float* __restrict__ r32, double* __restrict__ partials)
{
    const int c = threadIdx.x & (kp - 1);
    const int rpb = 256 / kp;
    double s = 0;
    const int64_t stride = (int64_t)gridDim.x * rpb;
    if (c < k) {
        const double al = alpha[c];
        for (int64_t i0 = (int64_t)blockIdx.x * rpb + threadIdx.x / kp; i0 < n; i0 += kColUnroll * stride) {
            double rv[kColUnroll], av[kColUnroll];
#pragma unroll
            for (int u = 0; u < kColUnroll; u++) {
                const int64_t i = i0 + u * stride;
                const size_t e = (size_t)(i < n ? i : 0) * k + c;
                rv[u] = i < n ? r[e] : 0.0;
                av[u] = i < n ? Ap[e] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < kColUnroll; u++) {
                const int64_t i = i0 + u * stride;
                if (i >= n) break;
                const size_t e = (size_t)i * k + c;
                const double ri = rv[u] - al * av[u];
                r[e] = ri;
                if (r32) r32[e] = (float)ri;
                s += ri * ri;
            }
        }
    }
    col_partial(s, k, kp, partials);
}

// p = z + beta p
template <typename Z>
__global__ __launch_bounds__(256) void bdir(int n, int k, int kp, const double* __restrict__ beta,
                                            const Z* __restrict__ z, double* __restrict__ p,
                                            const double* __restrict__ alpha, double* __restrict__ x)
{
    const int c = threadIdx.x & (kp - 1);
    const int rpb = 256 / kp;
    if (c >= k) return;
    const double be = beta[c];
    const double al = x ? alpha[c] : 0.0;
    const int64_t stride = (int64_t)gridDim.x * rpb;
    for (int64_t i0 = (int64_t)blockIdx.x * rpb + threadIdx.x / kp; i0 < n; i0 += kColUnroll * stride) {
        double zv[kColUnroll], pv[kColUnroll], xv[kColUnroll];
#pragma unroll
        for (int u = 0; u < kColUnroll; u++) {
            const int64_t i = i0 + u * stride;
            const size_t e = (size_t)(i < n ? i : 0) * k + c;
            zv[u] = i < n ? (double)z[e] : 0.0;
            pv[u] = i < n ? p[e] : 0.0;
            if (x) xv[u] = i < n ? x[e] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < kColUnroll; u++) {
            const int64_t i = i0 + u * stride;
            if (i >= n) break;
            const size_t e = (size_t)i * k + c;
            if (x) x[e] = xv[u] + al * pv[u];
            p[e] = zv[u] + be * pv[u];
        }
    }
}

// x += alpha p (the iteration's step, when no direction update follows)
__global__ __launch_bounds__(256) void bstep_x(int n, int k, int kp, const double* __restrict__ alpha,
                                               const double* __restrict__ p, double* __restrict__ x)
{
    const int c = threadIdx.x & (kp - 1);
    const int rpb = 256 / kp;
    if (c >= k) return;
    const double al = alpha[c];
    const int64_t stride = (int64_t)gridDim.x * rpb;
    for (int64_t i = (int64_t)blockIdx.x * rpb + threadIdx.x / kp; i < n; i += stride) {
        const size_t e = (size_t)i * k + c;
        x[e] = x[e] + al * p[e];
    }
}

// out[c] = sum over g of partials[g * k + c] (workgroup c): thread t adds
// g = t, t + 256, ... in order, then the 256 sums combine in a fixed tree
__global__ __launch_bounds__(256) void bcol_final(int np, int k, const double* __restrict__ partials,
                                                  double* __restrict__ out)
{
    __shared__ double sh[256];
    const int c = blockIdx.x;
    const double s = strided_sum(partials + c, 0, np, (size_t)k);
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[c] = sh[0];
}

// Device scalars of the lockstep flexible PCG, one entry per column.
struct BScal {
    double *bb, *rr, *rz, *rzn, *zap, *pap, *alpha, *beta, *scale, *relres, *active, *iters;
    double* flags;  // [0] active count, [1] error (1 indefinite, 2 non-finite)
};

__global__ void bscal_init(int k, double rtol, BScal S)
{
    const int c = threadIdx.x;
    if (c < k) {
        const double bn = sqrt(S.bb[c]);
        S.scale[c] = bn > 0 ? bn : 1.0;
        S.relres[c] = sqrt(S.rr[c]) / S.scale[c];
        S.active[c] = S.relres[c] > rtol ? 1.0 : 0.0;
        S.iters[c] = 0;
        if (!isfinite(bn) || !isfinite(S.relres[c])) S.flags[1] = 2;
    }
    __syncthreads();
    if (c == 0) {
        double a = 0;
        for (int i = 0; i < k; i++) a += S.active[i];
        S.flags[0] = a;
    }
}

__global__ void bscal_alpha(int k, BScal S)
{
    const int c = threadIdx.x;
    if (c >= k) return;
    double al = 0;
    if (S.active[c] != 0) {
        if (!(S.pap[c] > 0)) S.flags[1] = isfinite(S.pap[c]) ? 1 : 2;
        else al = S.rz[c] / S.pap[c];
    }
    S.alpha[c] = al;
}

__global__ void bscal_relres(int k, double rtol, BScal S)
{
    const int c = threadIdx.x;
    if (c < k && S.active[c] != 0) {
        S.iters[c] += 1;
        const double rel = sqrt(S.rr[c]) / S.scale[c];
        S.relres[c] = rel;
        if (!isfinite(rel)) S.flags[1] = 2;
        if (rel <= rtol) S.active[c] = 0;
    }
    __syncthreads();
    if (c == 0) {
        double a = 0;
        for (int i = 0; i < k; i++) a += S.active[i];
        S.flags[0] = a;
    }
}

__global__ void bscal_beta(int k, BScal S)
{
    const int c = threadIdx.x;
    if (c >= k) return;
    double be = 0;
    if (S.active[c] != 0) {
        be = -S.alpha[c] * S.zap[c] / S.rz[c];
        S.rz[c] = S.rzn[c];
    }
    S.beta[c] = be;
}

// after the true-residual check: reactivate columns that are not below rtol
__global__ void bscal_recheck(int k, double rtol, BScal S)
{
    const int c = threadIdx.x;
    if (c < k) {
        const double rel = sqrt(S.rr[c]) / S.scale[c];
        S.relres[c] = rel;
        S.active[c] = rel > rtol ? 1.0 : 0.0;
    }
    __syncthreads();
    if (c == 0) {
        double a = 0;
        for (int i = 0; i < k; i++) a += S.active[i];
        S.flags[0] = a;
    }
}

// ---------------------------------------------------------------------------
// work space
// ---------------------------------------------------------------------------
struct BlockLevel {
    double *x = nullptr, *b = nullptr, *t = nullptr, *r = nullptr;
    float *x32 = nullptr, *b32 = nullptr, *t32 = nullptr, *r32 = nullptr;
    // kernel codes (see blaunch) per vector precision, [0] FP64, [1] FP32
    int code_A[2] = {0, 0}, code_P[2] = {0, 0}, code_R[2] = {0, 0};
};

struct BlockState {
    int k = 0;
    std::vector<BlockLevel> lv;
    double *X = nullptr, *B = nullptr, *R = nullptr, *Z = nullptr, *P = nullptr, *AP = nullptr;
    float *R32 = nullptr, *Z32 = nullptr;
    double* scal = nullptr;     // 16 arrays of kMaxRhs
    double* partials = nullptr;  // 2 * kColBlocks * kMaxRhs
    double* pinned = nullptr;    // host, 2 + 2 * kMaxRhs
    BScal S{};
    // captured iteration (operator step, preconditioner step) for one (A,
    // precision, rtol); the graphs hold raw pointers to A's buffers and to
    // every level operator's, so they are keyed on A's identity and on the
    // buffer counts of all of them (a handle freed and another created at the
    // same address has a new uid)
    hipGraphExec_t gA = nullptr, gB = nullptr;
    uint64_t g_uid = 0;
    uint64_t g_forms = 0;
    int g_prec = -1;
    double g_rtol = -1;
};


static void free_graphs(BlockState* b)
{
    if (b->gA) (void)hipGraphExecDestroy(b->gA);
    if (b->gB) (void)hipGraphExecDestroy(b->gB);
    b->gA = b->gB = nullptr;
    b->g_uid = 0;
}

void block_free(BlockState* b)
{
    if (!b) return;
    free_graphs(b);
    for (BlockLevel& L : b->lv)
        for (void* p : {(void*)L.x, (void*)L.b, (void*)L.t, (void*)L.r, (void*)L.x32, (void*)L.b32, (void*)L.t32,
                        (void*)L.r32})
            dfree(p);
    for (void* p : {(void*)b->X, (void*)b->B, (void*)b->R, (void*)b->Z, (void*)b->P, (void*)b->AP, (void*)b->R32,
                    (void*)b->Z32, (void*)b->scal, (void*)b->partials})
        dfree(p);
    if (b->pinned) (void)hipHostFree(b->pinned);
    delete b;
}

static double time_ms(hipStream_t s, const std::function<void()>& fn, int reps)
{
    hipEvent_t e0, e1;
    RAMD_HIP(hipEventCreate(&e0));
    RAMD_HIP(hipEventCreate(&e1));
    fn();  // warm
    RAMD_HIP(hipEventRecord(e0, s));
    for (int i = 0; i < reps; i++) fn();
    RAMD_HIP(hipEventRecord(e1, s));
    RAMD_HIP(hipEventSynchronize(e1));
    float ms = 0;
    RAMD_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return ms / reps;
}

// Autotuner: kernel and lanes per row for one operator, k and vector
// precision, by timing the residual (square operators) or the product on
// the level's own block vectors.
template <typename X>
static int block_tune(ramd_hier_s* h, const ramd_csr_s* M, int k, const X* x, const X* b, X* y)
{
    hipStream_t s = h->ctx->stream;
    int best = 0;
    double best_ms = 1e30;
    std::vector<int> codes;
    for (int lpr : {8, 16, 32, 64})
        if (lpr_ok(lpr, k)) codes.push_back(lpr);
    if (aligned16(x) && aligned16(b) && aligned16(y))
        for (int lpr : {4, 8, 16})
            if (vec_ok<X>(lpr, k)) codes.push_back(-lpr);
    // small operators: twenty timed launches (microsecond launches rank unreliably over three)
    const int reps = M->nnz * (int64_t)k < ((int64_t)4 << 20) ? 20 : 3;
    for (int code : codes) {
        const double ms = time_ms(
            s, [&] { blaunch<X>(M, code, M->n_rows == M->n_cols ? OP_RESIDUAL : OP_SPMV, k, x, b, y, X(0), s); },
            reps);
        if (ms < best_ms) {
            best_ms = ms;
            best = code;
        }
    }
    return best;
}

static BlockState* block_state(ramd_hier_s* h, int k, bool fp32)
{
    RAMD_CHECK(h->opts.relax_type == RAMD_JACOBI, RAMD_E_ARG,
               "the multi-vector cycle smooths with Jacobi only (hierarchy relax_type %d)", h->opts.relax_type);
    BlockState*& b = h->blk;
    if (b && b->k != k) {
        RAMD_HIP(hipStreamSynchronize(h->ctx->stream));
        block_free(b);
        b = nullptr;
    }
    if (!b) {
        b = new BlockState();
        b->k = k;
        b->lv.resize(h->levels.size());
        for (size_t l = 0; l < h->levels.size(); l++) {
            const size_t m = (size_t)h->levels[l].n * k;
            BlockLevel& L = b->lv[l];
            L.x = dalloc<double>(m);
            L.b = dalloc<double>(m);
            L.t = dalloc<double>(m);
            L.r = dalloc<double>(m);
        }
        const size_t m0 = (size_t)h->levels[0].n * k;
        b->X = dalloc<double>(m0);
        b->B = dalloc<double>(m0);
        b->R = dalloc<double>(m0);
        b->Z = dalloc<double>(m0);
        b->P = dalloc<double>(m0);
        b->AP = dalloc<double>(m0);
        b->scal = dalloc<double>(16 * kMaxRhs);
        b->partials = dalloc<double>(2 * (size_t)kColBlocks * kMaxRhs);
        RAMD_HIP(hipHostMalloc((void**)&b->pinned, sizeof(double) * (2 + 2 * kMaxRhs), hipHostMallocDefault));
        double** f[] = {&b->S.bb, &b->S.rr, &b->S.rz, &b->S.rzn, &b->S.zap, &b->S.pap, &b->S.alpha, &b->S.beta,
                        &b->S.scale, &b->S.relres, &b->S.active, &b->S.iters, &b->S.flags};
        for (int i = 0; i < 13; i++) *f[i] = b->scal + i * kMaxRhs;
        hipLaunchKernelGGL(bfill, dim3(4), dim3(256), 0, h->ctx->stream, (int64_t)16 * kMaxRhs, 0.0, b->scal);
        RAMD_HIP(hipGetLastError());
        // tune every level's operators in FP64 on the level vectors
        for (size_t l = 0; l < h->levels.size(); l++) {
            const int64_t m = (int64_t)h->levels[l].n * k;
            for (double* v : {b->lv[l].x, b->lv[l].b, b->lv[l].t, b->lv[l].r})
                hipLaunchKernelGGL(bfill, dim3(256), dim3(256), 0, h->ctx->stream, m, 1.0, v);
            RAMD_HIP(hipGetLastError());
        }
        for (size_t l = 0; l + 1 < h->levels.size(); l++) {
            Level& Lv = h->levels[l];
            BlockLevel& L = b->lv[l];
            L.code_A[0] = block_tune<double>(h, Lv.A, k, L.x, L.b, L.r);
            L.code_P[0] = block_tune<double>(h, Lv.P, k, b->lv[l + 1].x, L.b, L.r);
            L.code_R[0] = block_tune<double>(h, Lv.R, k, L.x, b->lv[l + 1].b, b->lv[l + 1].r);
        }
    }
    if (fp32 && !b->R32) {
        enable_fp32(h);
        for (size_t l = 0; l < h->levels.size(); l++) {
            const size_t m = (size_t)h->levels[l].n * k;
            BlockLevel& L = b->lv[l];
            L.x32 = dalloc<float>(m);
            L.b32 = dalloc<float>(m);
            L.t32 = dalloc<float>(m);
            L.r32 = dalloc<float>(m);
        }
        const size_t m0 = (size_t)h->levels[0].n * k;
        b->R32 = dalloc<float>(m0);
        b->Z32 = dalloc<float>(m0);
        // tune every level's operators for FP32 vectors on the level vectors
        for (size_t l = 0; l < h->levels.size(); l++) {
            const int64_t m = (int64_t)h->levels[l].n * k;
            for (float* v : {b->lv[l].x32, b->lv[l].b32, b->lv[l].t32, b->lv[l].r32})
                hipLaunchKernelGGL(bfill32, dim3(256), dim3(256), 0, h->ctx->stream, m, 1.0f, v);
            RAMD_HIP(hipGetLastError());
        }
        for (size_t l = 0; l + 1 < h->levels.size(); l++) {
            Level& Lv = h->levels[l];
            BlockLevel& L = b->lv[l];
            BlockLevel& C = b->lv[l + 1];
            L.code_A[1] = block_tune<float>(h, Lv.A, k, L.x32, L.b32, L.r32);
            L.code_P[1] = block_tune<float>(h, Lv.P, k, C.x32, L.b32, L.r32);
            L.code_R[1] = block_tune<float>(h, Lv.R, k, L.x32, C.b32, C.r32);
        }
    }
    return b;
}

// ---------------------------------------------------------------------------
// block V-cycle (x = 0 on entry at every level, as inside PCG)
// ---------------------------------------------------------------------------
template <typename X>
static void bvec(BlockLevel& L, X*& x, X*& b, X*& t, X*& r)
{
    if constexpr (sizeof(X) == 8) {
        x = L.x;
        b = L.b;
        t = L.t;
        r = L.r;
    } else {
        x = L.x32;
        b = L.b32;
        t = L.t32;
        r = L.r32;
    }
}

template <typename X>
static void bcycle(ramd_hier_s* h, BlockState* bs, int l, X* x, const X* b)
{
    Level& L = h->levels[l];
    BlockLevel& BL = bs->lv[l];
    const int k = bs->k;
    hipStream_t s = h->ctx->stream;
    if (l == (int)h->levels.size() - 1) {
        const int n = h->coarse_n;
        if (n && n <= kCoarseWave)
            hipLaunchKernelGGL((bcoarse_wave<X>), dim3(k), dim3(64), sizeof(double) * n * n, s, n, k, h->lu, h->piv,
                               b, x);
        else if (n)
            hipLaunchKernelGGL((bcoarse_solve<X>), dim3(k), dim3(256), sizeof(double) * n, s, n, k, h->lu, h->piv, b,
                               x);
        RAMD_HIP(hipGetLastError());
        return;
    }
    const X w = (X)h->opts.relax_weight;
    const int sweeps = h->opts.num_smooth_sweeps;
    X *lx, *lb, *lt, *lr;
    bvec<X>(BL, lx, lb, lt, lr);
    (void)lx;
    (void)lb;
    X* cur = x;
    X* other = lt;
    const X* diag;
    if constexpr (sizeof(X) == 8) diag = L.diag;
    else diag = L.diag32;
    if (sweeps <= 0) {  // no sweep to start from zero: X = 0 as the reference sets it
        const int64_t m = (int64_t)L.n * k;
        if constexpr (sizeof(X) == 8) hipLaunchKernelGGL(bfill, dim3(256), dim3(256), 0, s, m, 0.0, x);
        else hipLaunchKernelGGL(bfill32, dim3(256), dim3(256), 0, s, m, 0.0f, x);
        RAMD_HIP(hipGetLastError());
    }
    for (int it = 0; it < sweeps; it++) {
        if (it == 0) {
            const int kp = pow2_at_least(k);
            hipLaunchKernelGGL((bjacobi_from_zero<X>), dim3(col_blocks(L.n, kp)), dim3(256), 0, s, L.n, k, kp, diag,
                               b, cur, w);
            RAMD_HIP(hipGetLastError());
        } else {
            blaunch<X>(L.A, BL.code_A[(sizeof(X) == 4 ? 1 : 0)], OP_JACOBI, k, cur, b, other, w, s);
            std::swap(cur, other);
        }
    }
    blaunch<X>(L.A, BL.code_A[(sizeof(X) == 4 ? 1 : 0)], OP_RESIDUAL, k, cur, b, lr, X(0), s);
    BlockLevel& BC = bs->lv[l + 1];
    X *cx, *cb, *ct, *cr;
    bvec<X>(BC, cx, cb, ct, cr);
    (void)ct;
    (void)cr;
    blaunch<X>(L.R, BL.code_R[(sizeof(X) == 4 ? 1 : 0)], OP_SPMV, k, lr, (const X*)nullptr, cb, X(0), s);
    bcycle<X>(h, bs, l + 1, cx, cb);
    blaunch<X>(L.P, BL.code_P[(sizeof(X) == 4 ? 1 : 0)], OP_ADD_INTO, k, cx, cur, other, X(0), s);
    std::swap(cur, other);
    for (int it = 0; it < sweeps; it++) {
        blaunch<X>(L.A, BL.code_A[(sizeof(X) == 4 ? 1 : 0)], OP_JACOBI, k, cur, b, other, w, s);
        std::swap(cur, other);
    }
    if (cur != x) RAMD_HIP(hipMemcpyAsync(x, cur, sizeof(X) * (size_t)L.n * k, hipMemcpyDeviceToDevice, s));
}


static void col_dots(ramd_ctx ctx, BlockState* bs, int n, int mode, const double* a, const void* z, bool z32,
                     const double* w, double* out1, double* out2)
{
    hipStream_t s = ctx->stream;
    const int k = bs->k;
    const int kp = pow2_at_least(k);
    const int g = col_blocks(n, kp);
    if (mode == 0) {
        if (z32) hipLaunchKernelGGL((bcol_dots<0, float>), dim3(g), dim3(256), 0, s, n, k, kp, a, (const float*)z, w,
                                    bs->partials);
        else hipLaunchKernelGGL((bcol_dots<0, double>), dim3(g), dim3(256), 0, s, n, k, kp, a, (const double*)z, w,
                                bs->partials);
    } else {
        if (z32) hipLaunchKernelGGL((bcol_dots<1, float>), dim3(g), dim3(256), 0, s, n, k, kp, a, (const float*)z, w,
                                    bs->partials);
        else hipLaunchKernelGGL((bcol_dots<1, double>), dim3(g), dim3(256), 0, s, n, k, kp, a, (const double*)z, w,
                                bs->partials);
    }
    RAMD_HIP(hipGetLastError());
    hipLaunchKernelGGL(bcol_final, dim3(k), dim3(256), 0, s, g, k, bs->partials, out1);
    if (mode == 1) hipLaunchKernelGGL(bcol_final, dim3(k), dim3(256), 0, s, g, k, bs->partials + (size_t)g * k, out2);
    RAMD_HIP(hipGetLastError());
}

// outer operator of the block solve: FP64 values (the caller's matrix has
// the fine level's pattern, so the fine level's tuning applies)
static void bouter(ramd_hier_s* h, BlockState* bs, const ramd_csr_s* A, int op, const double* x, const double* b,
                   double* y)
{
    const int code = bs->lv.size() > 1 ? bs->lv[0].code_A[0] : (bs->k <= 8 ? 8 : 32);
    blaunch<double>(A, code, op, bs->k, x, b, y, 0.0, h->ctx->stream);
}

static void bprecondition(ramd_hier_s* h, BlockState* bs, bool f32)
{
    if (f32) bcycle<float>(h, bs, 0, bs->Z32, bs->R32);
    else bcycle<double>(h, bs, 0, bs->Z, bs->R);
}

static int fpcg_block_run(ramd_hier_s* h, ramd_csr_s* A, int k, double* X_d, const double* B_d, double rtol,
                          int max_iter, int prec, int use_graph, int32_t* iters_col, double* relres_col,
                          int32_t* iters_total)
{
    ramd_ctx ctx = h->ctx;
    ctx->set_current();
    hipStream_t s = ctx->stream;
    const int n = A->n_rows;
    RAMD_CHECK(n == h->levels[0].n, RAMD_E_ARG, "ramd_fpcg_block: A has %d rows, hierarchy %d", n, h->levels[0].n);
    const bool f32 = prec == RAMD_PREC_FP32;
    BlockState* bs = block_state(h, k, f32);
    const int64_t m = (int64_t)n * k;
    const BScal S = bs->S;
    const dim3 ge(std::max<unsigned>(1, (unsigned)std::min<int64_t>((m + 255) / 256, 65536)));
    const int kp = pow2_at_least(k);
    const int gc = col_blocks(n, kp);
    RAMD_HIP(hipMemcpyAsync(bs->X, X_d, sizeof(double) * m, hipMemcpyDeviceToDevice, s));
    RAMD_HIP(hipMemcpyAsync(bs->B, B_d, sizeof(double) * m, hipMemcpyDeviceToDevice, s));
    hipLaunchKernelGGL(bfill, dim3(4), dim3(256), 0, s, (int64_t)16 * kMaxRhs, 0.0, bs->scal);
    // R = B - A X; per-column ||B||, ||R||
    bouter(h, bs, A, OP_RESIDUAL, bs->X, bs->B, bs->R);
    col_dots(ctx, bs, n, 0, bs->B, bs->B, false, nullptr, S.bb, nullptr);
    col_dots(ctx, bs, n, 0, bs->R, bs->R, false, nullptr, S.rr, nullptr);
    hipLaunchKernelGGL(bscal_init, dim3(1), dim3(64), 0, s, k, rtol, S);
    RAMD_HIP(hipGetLastError());
    auto read_flags = [&](double f[2]) {
        RAMD_HIP(hipMemcpyAsync(bs->pinned, S.flags, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
        RAMD_HIP(hipStreamSynchronize(s));
        f[0] = bs->pinned[0];
        f[1] = bs->pinned[1];
    };
    double fl[2];
    read_flags(fl);
    int rc = RAMD_OK;
    int it = 0;
    if (fl[1] != 0) rc = RAMD_E_NUMERIC;
    if (rc == RAMD_OK && fl[0] > 0) {
        // first direction: Z = M^-1 R, P = Z, rz = R.Z
        if (f32) hipLaunchKernelGGL((bconvert<double, float>), ge, dim3(256), 0, s, m, bs->R, bs->R32);
        bprecondition(h, bs, f32);
        if (f32) hipLaunchKernelGGL((bconvert<float, double>), ge, dim3(256), 0, s, m, bs->Z32, bs->P);
        else RAMD_HIP(hipMemcpyAsync(bs->P, bs->Z, sizeof(double) * m, hipMemcpyDeviceToDevice, s));
        col_dots(ctx, bs, n, 0, bs->R, f32 ? (const void*)bs->Z32 : (const void*)bs->Z, f32, nullptr, S.rz, nullptr);

        auto stepA = [&] {
            bouter(h, bs, A, OP_SPMV, bs->P, nullptr, bs->AP);
            col_dots(ctx, bs, n, 0, bs->P, bs->AP, false, nullptr, S.pap, nullptr);
            hipLaunchKernelGGL(bscal_alpha, dim3(1), dim3(64), 0, s, k, S);
            hipLaunchKernelGGL(bupdate, dim3(gc), dim3(256), 0, s, n, k, kp, S.alpha, bs->AP, bs->R,
                               f32 ? bs->R32 : nullptr, bs->partials);
            hipLaunchKernelGGL(bcol_final, dim3(k), dim3(256), 0, s, gc, k, bs->partials, S.rr);
            hipLaunchKernelGGL(bscal_relres, dim3(1), dim3(64), 0, s, k, rtol, S);
            RAMD_HIP(hipGetLastError());
        };
        // preconditioner step; with step_x the direction update also takes
        // the iteration's step x += alpha p (one read of p)
        auto stepB = [&](bool step_x) {
            bprecondition(h, bs, f32);
            col_dots(ctx, bs, n, 1, bs->R, f32 ? (const void*)bs->Z32 : (const void*)bs->Z, f32, bs->AP, S.rzn,
                     S.zap);
            hipLaunchKernelGGL(bscal_beta, dim3(1), dim3(64), 0, s, k, S);
            double* xs = step_x ? bs->X : nullptr;
            if (f32)
                hipLaunchKernelGGL((bdir<float>), dim3(gc), dim3(256), 0, s, n, k, kp, S.beta, bs->Z32, bs->P, S.alpha,
                                   xs);
            else
                hipLaunchKernelGGL((bdir<double>), dim3(gc), dim3(256), 0, s, n, k, kp, S.beta, bs->Z, bs->P, S.alpha,
                                   xs);
            RAMD_HIP(hipGetLastError());
        };
        auto step_x = [&] {
            hipLaunchKernelGGL(bstep_x, dim3(gc), dim3(256), 0, s, n, k, kp, S.alpha, bs->P, bs->X);
            RAMD_HIP(hipGetLastError());
        };
        const uint64_t fk = forms_key(h, A);
        if (use_graph && (bs->g_uid != A->uid || bs->g_forms != fk || bs->g_prec != prec || bs->g_rtol != rtol ||
                          !bs->gA)) {
            free_graphs(bs);
            bs->gA = capture_graph(s, stepA);
            bs->gB = capture_graph(s, [&] { stepB(true); });
            bs->g_uid = A->uid;
            bs->g_forms = fk;
            bs->g_prec = prec;
            bs->g_rtol = rtol;
        }
        while (it < max_iter) {
            if (use_graph) RAMD_HIP(hipGraphLaunch(bs->gA, s));
            else stepA();
            it++;
            read_flags(fl);
            if (fl[1] != 0) {
                step_x();
                rc = fl[1] == 1 ? RAMD_E_INDEFINITE : RAMD_E_NUMERIC;
                break;
            }
            bool x_done = false;
            if (fl[0] == 0) {
                // confirm every column with its true residual
                step_x();
                x_done = true;
                bouter(h, bs, A, OP_RESIDUAL, bs->X, bs->B, bs->R);
                col_dots(ctx, bs, n, 0, bs->R, bs->R, false, nullptr, S.rr, nullptr);
                hipLaunchKernelGGL(bscal_recheck, dim3(1), dim3(64), 0, s, k, rtol, S);
                RAMD_HIP(hipGetLastError());
                read_flags(fl);
                if (fl[0] == 0) break;
                if (f32) hipLaunchKernelGGL((bconvert<double, float>), ge, dim3(256), 0, s, m, bs->R, bs->R32);
            }
            if (use_graph && !x_done) RAMD_HIP(hipGraphLaunch(bs->gB, s));
            else stepB(!x_done);
        }
    }
    RAMD_HIP(hipMemcpyAsync(X_d, bs->X, sizeof(double) * m, hipMemcpyDeviceToDevice, s));
    RAMD_HIP(hipMemcpyAsync(bs->pinned + 2, S.iters, sizeof(double) * k, hipMemcpyDeviceToHost, s));
    RAMD_HIP(hipMemcpyAsync(bs->pinned + 2 + k, S.relres, sizeof(double) * k, hipMemcpyDeviceToHost, s));
    RAMD_HIP(hipStreamSynchronize(s));
    bool all = true;
    for (int c = 0; c < k; c++) {
        if (iters_col) iters_col[c] = (int32_t)bs->pinned[2 + c];
        if (relres_col) relres_col[c] = bs->pinned[2 + k + c];
        if (!(bs->pinned[2 + k + c] <= rtol)) all = false;
    }
    *iters_total = it;
    if (rc == RAMD_OK && !all) rc = RAMD_E_NOT_CONVERGED;
    return rc;
}

}  // namespace ramd

using namespace ramd;

extern "C" {

int ramd_fpcg_block(ramd_hier h, ramd_csr A, int32_t k, double* X_d, const double* B_d, double rtol, int max_iter,
                    int cycle_prec, int use_graph, int32_t* iters_col, double* relres_col, int32_t* iters_total)
{
    int rc = RAMD_OK;
    int st = guarded([&] {
        RAMD_CHECK(h && A && X_d && B_d && iters_total, RAMD_E_ARG, "ramd_fpcg_block: NULL argument");
        RAMD_CHECK(k >= 1 && k <= kMaxRhs, RAMD_E_ARG, "ramd_fpcg_block: k = %d (1..%d)", k, kMaxRhs);
        RAMD_CHECK(cycle_prec == RAMD_PREC_FP64 || cycle_prec == RAMD_PREC_FP32, RAMD_E_ARG,
                   "ramd_fpcg_block: unknown precision %d", cycle_prec);
        rc = fpcg_block_run(h, A, k, X_d, B_d, rtol, max_iter, cycle_prec, use_graph, iters_col, relres_col,
                            iters_total);
    });
    if (st != RAMD_OK) return st;
    if (rc == RAMD_E_INDEFINITE) set_error("ramd_fpcg_block: p.Ap <= 0 in a column (not SPD)");
    if (rc == RAMD_E_NUMERIC) set_error("ramd_fpcg_block: non-finite value in the iteration");
    if (rc == RAMD_E_NOT_CONVERGED) set_error("ramd_fpcg_block: %d iterations did not reach rtol", max_iter);
    return rc;
}

int ramd_vcycle_block(ramd_hier h, int32_t k, int prec, void* X_d, const void* B_d)
{
    return guarded([&] {
        RAMD_CHECK(h && X_d && B_d, RAMD_E_ARG, "ramd_vcycle_block: NULL argument");
        RAMD_CHECK(k >= 1 && k <= kMaxRhs, RAMD_E_ARG, "ramd_vcycle_block: k = %d (1..%d)", k, kMaxRhs);
        RAMD_CHECK(prec == RAMD_PREC_FP64 || prec == RAMD_PREC_FP32, RAMD_E_ARG, "ramd_vcycle_block: precision %d",
                   prec);
        h->ctx->set_current();
        BlockState* bs = block_state(h, k, prec == RAMD_PREC_FP32);
        if (prec == RAMD_PREC_FP32) bcycle<float>(h, bs, 0, (float*)X_d, (const float*)B_d);
        else bcycle<double>(h, bs, 0, (double*)X_d, (const double*)B_d);
        RAMD_HIP(hipStreamSynchronize(h->ctx->stream));
    });
}

int ramd_block_tuning(ramd_hier h, int32_t k, int32_t prec, int32_t* codes)
{
    return guarded([&] {
        RAMD_CHECK(h && codes, RAMD_E_ARG, "ramd_block_tuning: NULL argument");
        RAMD_CHECK(k >= 1 && k <= kMaxRhs, RAMD_E_ARG, "ramd_block_tuning: k = %d", k);
        RAMD_CHECK(prec == RAMD_PREC_FP64 || prec == RAMD_PREC_FP32, RAMD_E_ARG, "ramd_block_tuning: precision %d",
                   prec);
        h->ctx->set_current();
        const int p = prec == RAMD_PREC_FP32 ? 1 : 0;
        BlockState* bs = block_state(h, k, p == 1);
        RAMD_HIP(hipStreamSynchronize(h->ctx->stream));
        for (size_t l = 0; l < h->levels.size(); l++) {
            codes[3 * l] = bs->lv[l].code_A[p];
            codes[3 * l + 1] = bs->lv[l].code_P[p];
            codes[3 * l + 2] = bs->lv[l].code_R[p];
        }
    });
}

int ramd_block_set_tuning(ramd_hier h, int32_t k, int32_t prec, const int32_t* codes)
{
    return guarded([&] {
        RAMD_CHECK(h && codes, RAMD_E_ARG, "ramd_block_set_tuning: NULL argument");
        RAMD_CHECK(k >= 1 && k <= kMaxRhs, RAMD_E_ARG, "ramd_block_set_tuning: k = %d", k);
        RAMD_CHECK(prec == RAMD_PREC_FP64 || prec == RAMD_PREC_FP32, RAMD_E_ARG,
                   "ramd_block_set_tuning: precision %d", prec);
        h->ctx->set_current();
        const int p = prec == RAMD_PREC_FP32 ? 1 : 0;
        const size_t nl = h->levels.size();
        for (size_t l = 0; l + 1 < nl; l++)
            for (int m = 0; m < 3; m++) {
                const int c = codes[3 * l + m];
                const bool ok = c > 0 ? lpr_ok(c, k)
                                      : (c < 0 && (p ? vec_ok<float>(-c, k) : vec_ok<double>(-c, k)) &&
                                         (-c == 4 || -c == 8 || -c == 16));
                RAMD_CHECK(ok && (c < 0 || c == 8 || c == 16 || c == 32 || c == 64), RAMD_E_ARG,
                           "ramd_block_set_tuning: code %d for k = %d at level %zu", c, k, l);
            }
        BlockState* bs = block_state(h, k, p == 1);
        RAMD_HIP(hipStreamSynchronize(h->ctx->stream));
        free_graphs(bs);  // captured launches used the old kernels
        for (size_t l = 0; l + 1 < nl; l++) {
            bs->lv[l].code_A[p] = codes[3 * l];
            bs->lv[l].code_P[p] = codes[3 * l + 1];
            bs->lv[l].code_R[p] = codes[3 * l + 2];
        }
    });
}

}  // extern "C"

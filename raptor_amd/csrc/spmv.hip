// Row-block ("CSR-stream") kernels for the SpMV family and the
// deterministic vector reductions.
//
// Each workgroup owns a block of consecutive rows whose nonzeros fit in
// kStreamCap products.  Phase 1 streams the block's values and column
// indices with fully coalesced loads and stages a_ij * x_j in LDS; phase 2
// gives every row to one lane, which adds its products left to right.  The
// per-row order of operations is therefore exactly the reference's
// sequential loop (util/linalg/spmv.cpp:52-103, util/linalg/relax.cpp:23-41),
// so FP64 results are bit-identical to it (the file is compiled with
// -ffp-contract=off).  A row longer than kStreamCap is processed in chunks
// by its own workgroup.
#include <algorithm>
#include <cmath>
#include <type_traits>

#include "common.hpp"

namespace ramd {

// acc (+ or -) prod[k0] (+ or -) ... prod[k1 - 1], left to right; the LDS
// reads are issued eight at a time so their latency is paid once per eight
// entries, the additions keep their order
template <bool SUB, typename X>
__device__ __forceinline__ X lds_row_sum(const X* prod, int k0, int k1, X acc)
{
    int k = k0;
    for (; k + 8 <= k1; k += 8) {
        X t[8];
#pragma unroll
        for (int u = 0; u < 8; u++) t[u] = prod[k + u];
#pragma unroll
        for (int u = 0; u < 8; u++) acc = SUB ? acc - t[u] : acc + t[u];
    }
    for (; k < k1; k++) acc = SUB ? acc - prod[k] : acc + prod[k];
    return acc;
}

// SGA: col and val are a SortedGather's arrays (SpMV and its first-sweep
// form only), spos the entries' positions in their block.
template <int OP, typename V, typename X, int CAP, bool SGA = false>
__global__ __launch_bounds__(kBlockThreads) void stream_kernel(
    const int* __restrict__ rowptr, const int* __restrict__ col, const V* __restrict__ val,
    const int* __restrict__ blocks, int nb, const X* __restrict__ x, const X* __restrict__ b,
    X* __restrict__ y, X omega, int diag_first, double* __restrict__ dotp, const X* __restrict__ dvec,
    X* __restrict__ y2, const uint16_t* __restrict__ spos)
{
    static_assert(!SGA || OP == OP_SPMV || OP == OP_SPMV_JFZ, "sorted gathers serve the SpMV forms only");
    __shared__ X prod[CAP];
    const int blk = xcd_remap(blockIdx.x, nb);
    const int r0 = blocks[blk], r1 = blocks[blk + 1];
    const int s = rowptr[r0], e = rowptr[r1];
    const int tid = threadIdx.x;
    // the lane's row bounds and per-row operands, loaded together with the
    // block's bounds instead of after phase 1
    const int row = r0 + tid;
    const bool has_row = row < r1;
    int rs = 0, re = 0;
    X pre = X(0), bj = X(0), dgv = X(0);  // b / y / x of the row, b and a_ii for Jacobi
    if (has_row) {
        rs = rowptr[row];
        re = rowptr[row + 1];
        if (OP == OP_RESIDUAL || OP == OP_ADD_INTO) pre = b[row];
        else if (OP == OP_APPEND) pre = y[row];
        else if (OP == OP_SPMV && dotp) pre = x[row];
        else if (OP == OP_SPMV_JFZ) dgv = dvec[row];
        else if (OP == OP_JACOBI) {
            pre = x[row];
            bj = b[row];
            if (dvec) dgv = dvec[row];
        }
    }

    if (e - s > CAP) {
        // one long row, processed in LDS-sized chunks, summed by lane 0
        X acc = (OP == OP_RESIDUAL) ? b[r0] : X(0);
        X diag = 0;
        for (int c0 = s; c0 < e; c0 += CAP) {
            const int cn = min(CAP, e - c0);
            for (int j = tid; j < cn; j += kBlockThreads) prod[j] = X(val[c0 + j]) * x[col[c0 + j]];
            __syncthreads();
            if (tid == 0) {
                for (int k = 0; k < cn; k++) {
                    if (OP == OP_JACOBI) {
                        if (col[c0 + k] == r0) diag = X(val[c0 + k]);
                        else acc += prod[k];
                    } else if (OP == OP_RESIDUAL) {
                        acc -= prod[k];
                    } else {
                        acc += prod[k];
                    }
                }
            }
            __syncthreads();
        }
        if (OP == OP_SPMV && dotp) {
            const double v = block_sum(tid == 0 ? (double)x[r0] * (double)acc : 0.0);
            if (tid == 0) dotp[blk] = v;
        }
        if (tid == 0) {
            if (OP == OP_SPMV || OP == OP_RESIDUAL) y[r0] = acc;
            else if (OP == OP_SPMV_JFZ) {
                y[r0] = acc;
                y2[r0] = jacobi_zero_row(acc, dvec[r0], omega);
            } else if (OP == OP_APPEND) y[r0] += acc;
            else if (OP == OP_ADD_INTO) y[r0] = b[r0] + acc;
            else {
                const X xo = x[r0];
                y[r0] = (fabs(diag) > X(1e-16)) ? ((X(1.0) - omega) * xo) + (omega * ((b[r0] - acc) / diag)) : xo;
            }
        }
        return;
    }

    if constexpr (SGA) {
        // phase 1 (sorted gathers): lane j of the block takes the block's j-th
        // entry in column order, so a wavefront's 64 gathers fall on a few
        // adjacent cache lines; the product goes to its stored position.
        // Eight entries per lane and step, their loads issued first.
        constexpr int U = CAP / kBlockThreads, UB = 8;
        const int len = e - s;
    #pragma unroll
        for (int u0 = 0; u0 < U; u0 += UB) {
            if (u0 * kBlockThreads >= len) break;
            int cc[UB], pp[UB];
            X vv[UB], xx[UB];
    #pragma unroll
            for (int t = 0; t < UB; t++) {
                const int j = tid + (u0 + t) * kBlockThreads;
                if (j < len) {  // streamed once: past the caches
                    cc[t] = __builtin_nontemporal_load(&col[s + j]);
                    vv[t] = X(__builtin_nontemporal_load(&val[s + j]));
                    pp[t] = __builtin_nontemporal_load(&spos[s + j]);
                }
            }
    #pragma unroll
            for (int t = 0; t < UB; t++)
                if (tid + (u0 + t) * kBlockThreads < len) xx[t] = x[cc[t]];
    #pragma unroll
            for (int t = 0; t < UB; t++)
                if (tid + (u0 + t) * kBlockThreads < len) prod[pp[t]] = vv[t] * xx[t];
        }
    } else if constexpr (sizeof(X) == 4) {
        // phase 1 (FP32 vectors): the block's entries [s, e) are read in aligned
        // groups of four -- one 16-byte column load and one 16-byte value load per
        // group and lane instead of four of each -- then the gathers, then the
        // products; all of a step's loads are in flight before the first multiply.
        // Group entries outside [s, e) are neither gathered nor stored.  The last
        // block's final partial group, which may end past the arrays, is read
        // entry by entry.  (FP64 vectors keep one entry per load: their wide
        // loads measured slower.)
        typedef int I4 __attribute__((ext_vector_type(4)));
        typedef V V4 __attribute__((ext_vector_type(4)));
        constexpr int QPL = (CAP / 4 + 1 + kBlockThreads - 1) / kBlockThreads;  // groups per lane, at most
        constexpr int QC = 4;                                                      // groups per step
        const int q0 = s >> 2, nq = ((e + 3) >> 2) - q0;
        const bool tail_scalar = blk == nb - 1 && (e & 3) != 0;
    #pragma unroll
        for (int u0 = 0; u0 < QPL; u0 += QC) {
            int cc[QC][4];
            V vv[QC][4];
            X xx[QC][4];
    #pragma unroll
            for (int t = 0; t < QC; t++) {
                const int q = tid + (u0 + t) * kBlockThreads;
                if (u0 + t < QPL && q < nq) {
                    const int g = q0 + q;
                    if (tail_scalar && q == nq - 1) {
    #pragma unroll
                        for (int k = 0; k < 4; k++) {
                            const int idx = 4 * g + k;
                            cc[t][k] = idx < e ? col[idx] : 0;
                            vv[t][k] = idx < e ? val[idx] : V(0);
                        }
                    } else {  // streamed once: past the caches
                        const I4 c4 = __builtin_nontemporal_load(reinterpret_cast<const I4*>(col) + g);
                        const V4 v4 = __builtin_nontemporal_load(reinterpret_cast<const V4*>(val) + g);
    #pragma unroll
                        for (int k = 0; k < 4; k++) {
                            cc[t][k] = c4[k];
                            vv[t][k] = v4[k];
                        }
                    }
                }
            }
    #pragma unroll
            for (int t = 0; t < QC; t++) {
                const int q = tid + (u0 + t) * kBlockThreads;
    #pragma unroll
                for (int k = 0; k < 4; k++) {
                    const int idx = 4 * (q0 + q) + k;
                    if (u0 + t < QPL && q < nq && idx >= s && idx < e) xx[t][k] = x[cc[t][k]];
                }
            }
    #pragma unroll
            for (int t = 0; t < QC; t++) {
                const int q = tid + (u0 + t) * kBlockThreads;
    #pragma unroll
                for (int k = 0; k < 4; k++) {
                    const int idx = 4 * (q0 + q) + k;
                    if (u0 + t < QPL && q < nq && idx >= s && idx < e) prod[idx - s] = X(vv[t][k]) * xx[t][k];
                }
            }
        }
    } else {
        // phase 1 (FP64 vectors): every lane issues all of its column/value
        // loads, then all of its x gathers, before the first multiply (16
        // independent loads in flight per lane instead of one dependent chain)
        constexpr int U = CAP / kBlockThreads;
        const int len = e - s;
        int cc[U];
        X vv[U], xx[U];
    #pragma unroll
        for (int u = 0; u < U; u++) {
            const int j = tid + u * kBlockThreads;
            if (j < len) {  // streamed once: past the caches
                cc[u] = __builtin_nontemporal_load(&col[s + j]);
                vv[u] = X(__builtin_nontemporal_load(&val[s + j]));
            }
        }
    #pragma unroll
        for (int u = 0; u < U; u++) {
            const int j = tid + u * kBlockThreads;
            if (j < len) xx[u] = x[cc[u]];
        }
    #pragma unroll
        for (int u = 0; u < U; u++) {
            const int j = tid + u * kBlockThreads;
            if (j < len) prod[j] = vv[u] * xx[u];
        }
    }
    __syncthreads();

    // phase 2: one lane per row, products added left to right
    const int k0 = rs - s, k1 = re - s;
    if (OP == OP_SPMV || OP == OP_SPMV_JFZ) {
        X acc = 0;
        if (has_row) {
            acc = lds_row_sum<false>(prod, k0, k1, acc);
            y[row] = acc;
            if (OP == OP_SPMV_JFZ) y2[row] = jacobi_zero_row(acc, dgv, omega);
        }
        if (OP == OP_SPMV && dotp) {
            const double v = block_sum(has_row ? (double)pre * (double)acc : 0.0);
            if (tid == 0) dotp[blk] = v;
        }
        return;
    }
    if (!has_row) return;
    if (OP == OP_APPEND || OP == OP_ADD_INTO) {
        // CSR_append: the row sum is formed first, then added (spmv.cpp:96-101)
        const X acc = lds_row_sum<false>(prod, k0, k1, X(0));
        y[row] = pre + acc;
    } else if (OP == OP_RESIDUAL) {
        y[row] = lds_row_sum<true>(prod, k0, k1, pre);
    } else {
        // relax.cpp:23-41: x holds the old iterate (tmp), y receives the new
        const X xo = pre;
        if (k0 == k1) {
            y[row] = xo;
            return;
        }
        X acc = 0, diag = 0;
        if (diag_first && dvec) {
            // the level's diagonal array holds a_ii (0 where absent); a row
            // whose |a_ii| > 1e-16 stores it first, any other row keeps its
            // value -- the row's first entry is not read again
            const X dg = dgv;
            if (fabs(dg) > X(1e-16)) {
                diag = dg;
                acc = lds_row_sum<false>(prod, k0 + 1, k1, acc);
            }
        } else if (diag_first) {
            // diagonal stored first, or absent (then the row keeps its value)
            if (col[s + k0] == row) {
                diag = X(val[s + k0]);
                acc = lds_row_sum<false>(prod, k0 + 1, k1, acc);
            }
        } else {
            for (int k = k0; k < k1; k++) {
                if (col[s + k] == row) diag = X(val[s + k]);
                else acc += prod[k];
            }
        }
        y[row] = (fabs(diag) > X(1e-16)) ? ((X(1.0) - omega) * xo) + (omega * ((bj - acc) / diag)) : xo;
    }
}

// One-lane-per-row kernels, over three layouts.  LAYOUT 0, the sliced form:
// every wavefront reads its slice's columns and values as contiguous runs
// and padding (column -1) is loaded and dropped by a select.  LAYOUT 2, the
// sliced form with 16-bit column offsets (column = row + offset, padding
// kSliced16Pad).  LAYOUT 1, the CSR arrays themselves (short rows, no LDS
// staging): a lane walks its own row, entries past the row's end are not
// loaded and count as padding.  A
// lane adds its row's entries in stored order with the same operations as
// phase 2 of stream_kernel, so the bits are the same; eight entries' loads
// are in flight per lane.
template <int OP, typename V, typename X, int LAYOUT>
__global__ __launch_bounds__(kBlockThreads) void sliced_kernel(
    int n, const int64_t* __restrict__ sptr, const int* __restrict__ rowptr, const int* __restrict__ col,
    const int16_t* __restrict__ col16, const V* __restrict__ val, const X* __restrict__ x, const X* __restrict__ b,
    X* __restrict__ y, X omega, int diag_first, double* __restrict__ dotp, const X* __restrict__ dvec,
    X* __restrict__ y2)
{
    constexpr int B = 8;
    constexpr int64_t STEP = LAYOUT == 1 ? 1 : 64;  // CSR rows: consecutive entries; sliced: 64 apart
    const int blk = xcd_remap(blockIdx.x, gridDim.x);
    const int row = blk * kBlockThreads + threadIdx.x;
    const bool valid = row < n;
    const int lane = threadIdx.x & 63;
    // lanes past n belong to no slice when n is not a multiple of 64 * 4
    int64_t base = 0;
    int w = 0;
    if (valid) {
        if (LAYOUT != 1) {
            base = sptr[row >> 6] + lane;
            w = (int)((sptr[(row >> 6) + 1] - sptr[row >> 6]) >> 6);
        } else {
            base = rowptr[row];
            w = rowptr[row + 1] - (int)base;
        }
    }
    (void)lane;
    X acc = (OP == OP_RESIDUAL && valid) ? b[row] : X(0);
    X diag = 0;
    // Jacobi: first entry of the row to add (w: none -- the row keeps x)
    int k_first = 0;
    if (OP == OP_JACOBI && valid && diag_first) {
        if (dvec) {
            const X dg = dvec[row];
            if (fabs(dg) > X(1e-16)) {
                diag = dg;
                k_first = 1;
            } else {
                k_first = w;
            }
        } else if (w > 0 && (LAYOUT == 2 ? col16[base] == 0 : col[base] == row)) {
            diag = X(val[base]);
            k_first = 1;
        } else {
            k_first = w;
        }
    }
    for (int k0 = 0; k0 < w; k0 += B) {
        int cc[B];
        V vv[B];
        X xx[B];
#pragma unroll
        for (int t = 0; t < B; t++) {
            const bool in = k0 + t < w;
            // entries are streamed once: past the caches
            if (LAYOUT == 2) {
                const int d = in ? (int)__builtin_nontemporal_load(&col16[base + STEP * (k0 + t)]) : kSliced16Pad;
                cc[t] = d == kSliced16Pad ? -1 : row + d;
            } else {
                cc[t] = in ? __builtin_nontemporal_load(&col[base + STEP * (k0 + t)]) : -1;
            }
            vv[t] = in ? __builtin_nontemporal_load(&val[base + STEP * (k0 + t)]) : V(0);
        }
#pragma unroll
        for (int t = 0; t < B; t++) xx[t] = x[cc[t] < 0 ? 0 : cc[t]];
#pragma unroll
        for (int t = 0; t < B; t++) {
            const X p = X(vv[t]) * xx[t];
            const bool use = cc[t] >= 0 && k0 + t >= k_first;
            if (OP == OP_JACOBI && !diag_first) {
                if (use && cc[t] == row) diag = X(vv[t]);
                else acc = use ? acc + p : acc;
            } else if (OP == OP_RESIDUAL) {
                acc = use ? acc - p : acc;
            } else {
                acc = use ? acc + p : acc;
            }
        }
    }
    if (OP == OP_SPMV && dotp) {
        const double v = block_sum(valid ? (double)x[row] * (double)acc : 0.0);
        if (threadIdx.x == 0) dotp[blk] = v;
    }
    if (!valid) return;
    if (OP == OP_SPMV || OP == OP_RESIDUAL) {
        y[row] = acc;
    } else if (OP == OP_SPMV_JFZ) {
        y[row] = acc;
        y2[row] = jacobi_zero_row(acc, dvec[row], omega);
    } else if (OP == OP_APPEND) {
        y[row] += acc;
    } else if (OP == OP_ADD_INTO) {
        y[row] = b[row] + acc;
    } else {
        const X xo = x[row];
        y[row] = (fabs(diag) > X(1e-16)) ? ((X(1.0) - omega) * xo) + (omega * ((b[row] - acc) / diag)) : xo;
    }
}

// Wavefront-level barrier: LDS written by lanes of one wavefront is visible
// to its other lanes after it (no workgroup barrier: the wavefronts of a
// workgroup run different trip counts).
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One wavefront per row: the form for operators with few, long rows (the
// coarse Galerkin operators of aggregation hierarchies), where one lane per
// row leaves most of the GPU idle and walks a long dependent chain of
// loads.  The wavefront streams the row's columns and values with coalesced
// loads (64 consecutive entries per step), gathers x and stages the products
// in LDS, kWaveChunk at a time; lane 0 then adds them left to right with the
// operations of stream_kernel's phase 2 (the first-entry diagonal rule of a
// diagonal-first row, or the entry whose column is the row itself), so the
// bits are the same.
constexpr int kWaveRows = kBlockThreads / 64;  // rows per workgroup
constexpr int kWaveChunk = 512;                // products staged per wavefront and pass

template <int OP, typename V, typename X>
__global__ __launch_bounds__(kBlockThreads) void wave_kernel(int n, const int* __restrict__ rowptr,
                                                            const int* __restrict__ col, const V* __restrict__ val,
                                                            const X* __restrict__ x, const X* __restrict__ b,
                                                            X* __restrict__ y, X omega, int diag_first,
                                                            double* __restrict__ dotp, const X* __restrict__ dvec,
                                                            X* __restrict__ y2)
{
    __shared__ X prod[kWaveRows * kWaveChunk];
    __shared__ int dpos[kWaveRows];
    const int blk = xcd_remap(blockIdx.x, gridDim.x);
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int row = blk * kWaveRows + wv;
    const bool valid = row < n;
    X* pw = prod + wv * kWaveChunk;
    int s = 0, e = 0;
    if (valid) {
        s = rowptr[row];
        e = rowptr[row + 1];
    }
    const int len = e - s;
    // entries [first, len) are added, except entry `skip` (Jacobi's diagonal
    // when it is not stored first); a Jacobi row without a usable diagonal
    // keeps its value
    X acc = X(0), diag = X(0);
    int first = 0;
    if (lane == 0) dpos[wv] = -1;
    if (OP == OP_RESIDUAL && valid) acc = b[row];
    if (OP == OP_JACOBI && valid && diag_first) {
        if (dvec) {
            const X dg = dvec[row];
            if (fabs(dg) > X(1e-16)) {
                diag = dg;
                first = 1;
            } else {
                first = len;
            }
        } else if (len > 0 && col[s] == row) {
            diag = X(val[s]);
            first = 1;
        } else {
            first = len;
        }
    }
    wave_sync();
    for (int c0 = 0; c0 < len; c0 += kWaveChunk) {
        const int cn = min(kWaveChunk, len - c0);
        constexpr int U = kWaveChunk / 64;
        int cc[U];
        V vv[U];
        X xx[U];
#pragma unroll
        for (int u = 0; u < U; u++) {  // streamed once: past the caches
            const int j = lane + 64 * u;
            if (j < cn) {
                cc[u] = __builtin_nontemporal_load(&col[s + c0 + j]);
                vv[u] = __builtin_nontemporal_load(&val[s + c0 + j]);
            }
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int j = lane + 64 * u;
            if (j < cn) xx[u] = x[cc[u]];
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int j = lane + 64 * u;
            if (j < cn) {
                pw[j] = X(vv[u]) * xx[u];
                if (OP == OP_JACOBI && !diag_first && cc[u] == row) dpos[wv] = c0 + j;
            }
        }
        wave_sync();
        if (lane == 0) {
            const int k0 = max(first - c0, 0);
            if (OP == OP_JACOBI && !diag_first) {
                const int dp = dpos[wv];
                if (dp >= c0 && dp < c0 + cn) {
                    diag = X(val[s + dp]);
                    acc = lds_row_sum<false>(pw, k0, dp - c0, acc);
                    acc = lds_row_sum<false>(pw, dp - c0 + 1, cn, acc);
                } else {
                    acc = lds_row_sum<false>(pw, k0, cn, acc);
                }
            } else {
                acc = lds_row_sum<OP == OP_RESIDUAL>(pw, min(k0, cn), cn, acc);
            }
        }
        wave_sync();
    }
    if (OP == OP_SPMV && dotp) {
        const double v = block_sum((valid && lane == 0) ? (double)x[row] * (double)acc : 0.0);
        if (threadIdx.x == 0) dotp[blk] = v;
    }
    if (!valid || lane != 0) return;
    if (OP == OP_SPMV || OP == OP_RESIDUAL) {
        y[row] = acc;
    } else if (OP == OP_SPMV_JFZ) {
        y[row] = acc;
        y2[row] = jacobi_zero_row(acc, dvec[row], omega);
    } else if (OP == OP_APPEND) {
        y[row] += acc;
    } else if (OP == OP_ADD_INTO) {
        y[row] = b[row] + acc;
    } else {
        const X xo = x[row];
        y[row] = (len > 0 && fabs(diag) > X(1e-16)) ? ((X(1.0) - omega) * xo) + (omega * ((b[row] - acc) / diag))
                                                    : xo;
    }
}

// Stencil-diagonal kernels: one lane per row, diagonals visited in the
// row's stored order (slot t holds diagonal a.ord[t]).  A lane reads its
// row's values a 16-byte group of slots at a time (dia_index layout).  Every
// load is issued unconditionally and without waiting for the row's mask --
// a diagonal absent from the row reads its own zero and x at the clamped
// column (dia_col) -- and the accumulation skips it with a select, so the
// operation sequence of each row is exactly the CSR row loop's and no
// out-of-range x is touched.

// slots per chunk: all of a 27-point row's FP32 values, ten FP64 values;
// x is gathered and added in sub-chunks of dia_xchunk slots (nine FP64
// values against FP32 matrix values, to hold fewer registers)
template <typename V>
constexpr int dia_chunk() { return sizeof(V) == 4 ? 28 : 10; }
template <typename V, typename X>
constexpr int dia_xchunk() { return (sizeof(V) == 4 && sizeof(X) == 8) ? 9 : dia_chunk<V>(); }

// x[c] for 0 <= c < n through a 32-bit byte offset from the uniform base,
// so the load takes the scalar-base + vector-offset form (one address VGPR
// per gather); the diagonal form is built only where n * 8 < 2^32.
template <typename X>
__device__ __forceinline__ X dia_gather(const X* __restrict__ x, int c)
{
    return *reinterpret_cast<const X*>(reinterpret_cast<const char*>(x) + (unsigned)c * (unsigned)sizeof(X));
}

// Values of slots [c0, c0 + CH) of `row`; groups past the stored slots (nsl,
// a multiple of the group size) read nothing and give 0.
template <typename V, int CH, int ND>
__device__ __forceinline__ void dia_load(const V* __restrict__ val, int c0, int row, int n, int nsl, V (&vv)[CH])
{
    constexpr int G = 16 / (int)sizeof(V);
    constexpr int NSL = (ND + G - 1) / G * G;
    static_assert(CH % G == 0, "chunk must hold whole groups");
    typedef V VG __attribute__((ext_vector_type(G)));
    const VG* vg = reinterpret_cast<const VG*>(val);
#pragma unroll
    for (int g = 0; g < CH / G; g++) {
        const int s0 = c0 + g * G;
        const bool ok = (ND == kDiaMax) ? s0 < nsl : s0 < NSL;
        VG v = VG(V(0));
        // values are read once per operation: streamed past the caches
        if (ok) v = __builtin_nontemporal_load(&vg[(size_t)(s0 / G) * n + row]);
#pragma unroll
        for (int t = 0; t < G; t++) vv[g * G + t] = v[t];
    }
}

// y_row of the plain product (OP_SPMV) in the stencil-diagonal form: the
// same load and select sequence as dia_kernel's body.
template <int OP, typename V, typename X, int ND>
__device__ __forceinline__ X dia_row(int row, int n, const DiaArgs& a, int nd, int center,
                                     const unsigned* __restrict__ mask, const V* __restrict__ val,
                                     const X* __restrict__ x, const X* __restrict__ b)
{
    const unsigned m = mask[row];
    X acc = X(0);
    constexpr int CH = dia_chunk<V>();
    const int nsl = dia_slots<V>(nd);
    constexpr int XC = dia_xchunk<V, X>();
#pragma unroll
    for (int c0 = 0; c0 < ND; c0 += CH) {
        V vv[CH];
        dia_load<V, CH, ND>(val, c0, row, n, nsl, vv);
#pragma unroll
        for (int x0 = 0; x0 < CH; x0 += XC) {
            X xx[XC];
            // the gathers first: their addresses do not wait for the mask
#pragma unroll
            for (int t = 0; t < XC; t++) {
                const int kk = c0 + x0 + t;
                if (x0 + t < CH && kk < ND) {
                    const bool live = (ND == kDiaMax) ? (kk < nd) : true;
                    xx[t] = dia_gather(x, dia_col(row, a.off[live ? a.ord[kk] : 0], n));
                }
            }
#pragma unroll
            for (int t = 0; t < XC; t++) {
                const int kk = c0 + x0 + t;
                if (x0 + t < CH && kk < ND) {
                    const bool live = (ND == kDiaMax) ? (kk < nd) : true;
                    const bool on = live && ((m >> (live ? a.ord[kk] : 0)) & 1u);
                    acc = on ? acc + X(vv[x0 + t]) * xx[t] : acc;
                }
            }
        }
    }
    (void)b;
    (void)center;
    return acc;
}

template <int OP, typename V, typename X, int ND>
__global__ __launch_bounds__(kBlockThreads) void dia_kernel(int n, DiaArgs a, int nd, int center,
                                                           const unsigned* __restrict__ mask, const V* __restrict__ val,
                                                           const X* __restrict__ x, const X* __restrict__ b,
                                                           X* __restrict__ y, X omega, double* __restrict__ dotp)
{
    const int blk = xcd_remap(blockIdx.x, gridDim.x);
    const int row = blk * kBlockThreads + threadIdx.x;
    if (OP == OP_SPMV && dotp) {
        // all lanes take part in the workgroup sum; rows past n add zero
        const X acc = row < n ? dia_row<OP, V, X, ND>(row, n, a, nd, center, mask, val, x, b) : X(0);
        if (row < n) y[row] = acc;
        const double v = block_sum(row < n ? (double)x[row] * (double)acc : 0.0);
        if (threadIdx.x == 0) dotp[blk] = v;
        return;
    }
    if (row >= n) return;
    const unsigned m = mask[row];
    X acc = (OP == OP_RESIDUAL) ? b[row] : X(0);
    X diag = 0;
    constexpr int CH = dia_chunk<V>();
    const int nsl = dia_slots<V>(nd);
    constexpr int XC = dia_xchunk<V, X>();
#pragma unroll
    for (int c0 = 0; c0 < ND; c0 += CH) {
        V vv[CH];
        dia_load<V, CH, ND>(val, c0, row, n, nsl, vv);
#pragma unroll
        for (int x0 = 0; x0 < CH; x0 += XC) {
            X xx[XC];
            bool on[XC];
            int kd[XC];
            // the gathers first: their addresses do not wait for the mask
#pragma unroll
            for (int t = 0; t < XC; t++) {
                const int kk = c0 + x0 + t;
                if (x0 + t < CH && kk < ND) {
                    const bool live = (ND == kDiaMax) ? (kk < nd) : true;
                    kd[t] = live ? a.ord[kk] : 0;
                    xx[t] = dia_gather(x, dia_col(row, a.off[kd[t]], n));
                }
            }
#pragma unroll
            for (int t = 0; t < XC; t++) {
                const int kk = c0 + x0 + t;
                if (x0 + t < CH && kk < ND) on[t] = ((ND == kDiaMax) ? (kk < nd) : true) && ((m >> kd[t]) & 1u);
            }
#pragma unroll
            for (int t = 0; t < XC; t++) {
                if (x0 + t < CH && c0 + x0 + t < ND) {
                    const X v = X(vv[x0 + t]);
                    const X p = v * xx[t];
                    if (OP == OP_JACOBI) {
                        if (kd[t] == center) diag = on[t] ? v : diag;
                        else acc = on[t] ? acc + p : acc;
                    } else if (OP == OP_RESIDUAL) {
                        acc = on[t] ? acc - p : acc;
                    } else {
                        acc = on[t] ? acc + p : acc;
                    }
                }
            }
        }
    }
    if (OP == OP_SPMV || OP == OP_RESIDUAL) {
        y[row] = acc;
    } else if (OP == OP_APPEND) {
        y[row] += acc;
    } else if (OP == OP_ADD_INTO) {
        y[row] = b[row] + acc;
    } else {
        const X xo = x[row];
        y[row] = (m != 0u && fabs(diag) > X(1e-16)) ? ((X(1.0) - omega) * xo) + (omega * ((b[row] - acc) / diag))
                                                   : xo;
    }
}

// The same operations with the workgroup's x window staged in LDS: the
// 256 rows of a workgroup read x only at row + offset, i.e. inside a few
// runs of consecutive addresses, which are loaded once with coalesced loads
// (indices clamped into [0, n): a clamped value is only ever read for an
// absent diagonal, whose product the select drops).  Same operation
// sequence per row as dia_kernel, so the same bits.
template <int OP, typename V, typename X, int ND>
__global__ __launch_bounds__(kBlockThreads) void dia_lds_kernel(int n, DiaArgs a, DiaWindow w, int nd, int center,
                                                               const unsigned* __restrict__ mask,
                                                               const V* __restrict__ val, const X* __restrict__ x,
                                                               const X* __restrict__ b, X* __restrict__ y, X omega,
                                                               double* __restrict__ dotp)
{
    __shared__ X xs[kDiaWin];
    const int blk = xcd_remap(blockIdx.x, gridDim.x);
    const int r0 = blk * kBlockThreads;
    const int tid = threadIdx.x;
    for (int j = 0; j < w.nr; j++) {
        const int len = kBlockThreads - 1 + w.len[j];
        for (int t = tid; t < len; t += kBlockThreads) {
            int g = r0 + w.lo[j] + t;
            g = g < 0 ? 0 : (g >= n ? n - 1 : g);
            xs[w.base[j] + t] = x[g];
        }
    }
    __syncthreads();
    const int row = r0 + tid;
    const bool valid = row < n;
    X acc = (OP == OP_RESIDUAL && valid) ? b[row] : X(0);
    X diag = 0;
    const unsigned m = valid ? mask[row] : 0u;
    const int vrow = valid ? row : 0;
    constexpr int CH = dia_chunk<V>();
    const int nsl = dia_slots<V>(nd);
#pragma unroll
    for (int c0 = 0; c0 < ND; c0 += CH) {
        V vv[CH];
        bool on[CH];
        int kd[CH];
        dia_load<V, CH, ND>(val, c0, vrow, n, nsl, vv);
#pragma unroll
        for (int t = 0; t < CH; t++) {
            const int kk = c0 + t;
            if (kk < ND) {
                const bool live = (ND == kDiaMax) ? (kk < nd) : true;
                const int k = live ? a.ord[kk] : 0;
                kd[t] = k;
                on[t] = live && ((m >> k) & 1u);
            }
        }
#pragma unroll
        for (int t = 0; t < CH; t++) {
            if (c0 + t < ND) {
                const X p = X(vv[t]) * xs[w.lofs[kd[t]] + tid];
                if (OP == OP_JACOBI) {
                    if (kd[t] == center) diag = on[t] ? X(vv[t]) : diag;
                    else acc = on[t] ? acc + p : acc;
                } else if (OP == OP_RESIDUAL) {
                    acc = on[t] ? acc - p : acc;
                } else {
                    acc = on[t] ? acc + p : acc;
                }
            }
        }
    }
    if (OP == OP_SPMV && dotp) {
        const double v = block_sum(valid ? (double)x[row] * (double)acc : 0.0);
        if (tid == 0) dotp[blk] = v;
    }
    if (!valid) return;
    if (OP == OP_SPMV || OP == OP_RESIDUAL) {
        y[row] = acc;
    } else if (OP == OP_APPEND) {
        y[row] += acc;
    } else if (OP == OP_ADD_INTO) {
        y[row] = b[row] + acc;
    } else {
        const X xo = x[row];
        y[row] = (m != 0u && fabs(diag) > X(1e-16)) ? ((X(1.0) - omega) * xo) + (omega * ((b[row] - acc) / diag))
                                                   : xo;
    }
}

// Plane-march form for a box geometry (DiaBox): a workgroup owns a TX x TY
// tile of (x, y) positions and walks zc consecutive planes.  For every plane
// it stages the tile's x values with a one-value halo, (TX + 2) x (TY + 2),
// in a ring of four LDS planes, so the 27 reads of a row come from LDS and
// each x value is loaded from memory about (1 + 2 / TX + 2 / TY) times per
// operation instead of once per diagonal.  The staged value of tile position
// (i, j) on plane z' is x at clamp(z' P + (y0 + j - 1) L + x0 + i - 1) into
// [0, n), which is x at dia_col(row, off) for every row of the tile and every
// box offset, so each row reads exactly what dia_kernel gathers and runs its
// operation sequence: the same bits.  One barrier per plane: plane z + 1 is
// written into slot (z + 1) & 3 while slots z - 1 .. z + 1 are read; the
// next plane's loads are in flight during the current plane's rows.
template <int OP, typename V, typename X, int ND, int TX>
__global__ __launch_bounds__(kBlockThreads) void dia_box_kernel(int n, int L, int P, int nxt, int nyt, int zc,
                                                               DiaArgs a, BoxArgs bx, int nd, int center,
                                                               const unsigned* __restrict__ mask,
                                                               const V* __restrict__ val, const X* __restrict__ x,
                                                               const X* __restrict__ b, X* __restrict__ y, X omega,
                                                               double* __restrict__ dotp)
{
    constexpr int TY = kBlockThreads / TX;
    constexpr int RW = TX + 2;
    constexpr int TILE = RW * (TY + 2);
    constexpr int PL = (TILE + 7) / 8 * 8;
    constexpr int NLD = (TILE + kBlockThreads - 1) / kBlockThreads;
    __shared__ X xs[4 * PL];
    const int blk = xcd_remap(blockIdx.x, gridDim.x);
    const int xt = blk % nxt, yt = (blk / nxt) % nyt, zt = blk / (nxt * nyt);
    const int x0 = xt * TX, y0 = yt * TY, z0 = zt * zc;
    const int nz = n / P, ny = P / L;
    const int z1 = min(z0 + zc, nz);
    const int tid = threadIdx.x;
    const int tx = tid % TX, ty = tid / TX;
    const bool col_ok = x0 + tx < L && y0 + ty < ny;
    const int lb = (ty + 1) * RW + tx + 1;  // the lane's own position in a plane
    // tile positions this thread stages, as offsets from the plane's start
    int goff[NLD];
#pragma unroll
    for (int i = 0; i < NLD; i++) {
        const int e = tid + i * kBlockThreads;
        goff[i] = (y0 + e / RW - 1) * L + x0 + e % RW - 1;
    }
    auto load_plane = [&](int zp, X (&r)[NLD]) {
#pragma unroll
        for (int i = 0; i < NLD; i++) {
            if (tid + i * kBlockThreads < TILE) {
                int g = zp * P + goff[i];
                g = g < 0 ? 0 : (g >= n ? n - 1 : g);
                r[i] = dia_gather(x, g);
            }
        }
    };
    auto store_plane = [&](int zp, const X (&r)[NLD]) {
        X* dst = xs + (zp & 3) * PL;
#pragma unroll
        for (int i = 0; i < NLD; i++)
            if (tid + i * kBlockThreads < TILE) dst[tid + i * kBlockThreads] = r[i];
    };
    X nx[NLD];
    load_plane(z0 - 1, nx);
    store_plane(z0 - 1, nx);
    load_plane(z0, nx);
    store_plane(z0, nx);
    load_plane(z0 + 1, nx);
    double dsum = 0.0;
    constexpr int CH = dia_chunk<V>();
    const int nsl = dia_slots<V>(nd);
    // the next plane's row mask and first value chunk (all of a row's FP32
    // values, ten FP64 ones) are loaded one plane ahead, so they arrive
    // while this plane's rows run
    constexpr bool PF = true;
    V vpf[PF ? CH : 1];
    unsigned mpf = 0u;
    if constexpr (PF) {
        const int r0 = col_ok ? z0 * P + (y0 + ty) * L + x0 + tx : 0;
        mpf = col_ok ? mask[r0] : 0u;
        dia_load<V, CH, ND>(val, 0, r0, n, nsl, vpf);
    }
    for (int z = z0; z < z1; z++) {
        store_plane(z + 1, nx);
        if (z + 2 <= z1) load_plane(z + 2, nx);
        const int row = z * P + (y0 + ty) * L + x0 + tx;
        const int vrow = col_ok ? row : 0;
        // the row's mask and first value chunk are in flight across the barrier
        unsigned m;
        V vv0[CH];
        if constexpr (PF) {
            m = mpf;
#pragma unroll
            for (int t = 0; t < CH; t++) vv0[t] = vpf[t];
            if (z + 1 < z1) {
                const int rn = col_ok ? row + P : 0;
                mpf = col_ok ? mask[rn] : 0u;
                dia_load<V, CH, ND>(val, 0, rn, n, nsl, vpf);
            }
        } else {
            m = col_ok ? mask[vrow] : 0u;
            dia_load<V, CH, ND>(val, 0, vrow, n, nsl, vv0);
        }
        __syncthreads();
        X acc = (OP == OP_RESIDUAL && col_ok) ? b[vrow] : X(0);
        X diag = 0;
#pragma unroll
        for (int c0 = 0; c0 < ND; c0 += CH) {
            V vv[CH];
            if (c0 == 0) {
#pragma unroll
                for (int t = 0; t < CH; t++) vv[t] = vv0[t];
            } else {
                dia_load<V, CH, ND>(val, c0, vrow, n, nsl, vv);
            }
#pragma unroll
            for (int t = 0; t < CH; t++) {
                const int kk = c0 + t;
                if (kk < ND) {
                    const bool live = (ND == kDiaMax) ? (kk < nd) : true;
                    const int k = live ? a.ord[kk] : 0;
                    const bool on = live && ((m >> k) & 1u);
                    const X v = X(vv[t]);
                    // uniform part of the LDS address: the plane's slot and the offset
                    const int so = ((z + bx.dz[k] + 4) & 3) * PL + bx.rel[k];
                    const X p = v * xs[lb + so];
                    if (OP == OP_JACOBI) {
                        if (k == center) diag = on ? v : diag;
                        else acc = on ? acc + p : acc;
                    } else if (OP == OP_RESIDUAL) {
                        acc = on ? acc - p : acc;
                    } else {
                        acc = on ? acc + p : acc;
                    }
                }
            }
        }
        if (col_ok) {
            if (OP == OP_SPMV || OP == OP_RESIDUAL) {
                y[row] = acc;
            } else if (OP == OP_APPEND) {
                y[row] += acc;
            } else if (OP == OP_ADD_INTO) {
                y[row] = b[row] + acc;
            } else {
                const X xo = xs[(z & 3) * PL + lb];
                y[row] = (m != 0u && fabs(diag) > X(1e-16)) ? ((X(1.0) - omega) * xo) + (omega * ((b[row] - acc) / diag))
                                                           : xo;
            }
            if (OP == OP_SPMV && dotp) dsum += (double)xs[(z & 3) * PL + lb] * (double)acc;
        }
    }
    if (OP == OP_SPMV && dotp) {
        const double v = block_sum(dsum);
        if (tid == 0) dotp[blk] = v;
    }
}

int dia_grid(const ramd_csr_s* A, int pp)
{
    const Dia* d = A->dia;
    const int n = A->n_rows;
    if (d->lds_x[pp] == 2 && d->box.L > 0) {
        const DiaBox& bx = d->box;
        const int ty = kBlockThreads / bx.tx;
        const int nxt = (bx.L + bx.tx - 1) / bx.tx, nyt = (bx.P / bx.L + ty - 1) / ty;
        const int nzt = (n / bx.P + bx.zc[pp] - 1) / bx.zc[pp];
        return std::max(1, nxt * nyt * nzt);
    }
    return std::max(1, (n + kBlockThreads - 1) / kBlockThreads);
}

int dia_grid_bound(const ramd_csr_s* A)
{
    const Dia* d = A->dia;
    const int n = A->n_rows;
    int np = (n + kBlockThreads - 1) / kBlockThreads;
    if (d->box.L > 0) {
        const DiaBox& bx = d->box;
        const int ty = kBlockThreads / bx.tx;
        const int nxt = (bx.L + bx.tx - 1) / bx.tx, nyt = (bx.P / bx.L + ty - 1) / ty;
        np = std::max(np, nxt * nyt * ((n / bx.P + kBoxMinPlanes - 1) / kBoxMinPlanes));
    }
    return std::max(np, 1);
}

template <int OP, typename V, typename X>
static void launch_dia_op(const ramd_csr_s* A, const V* val, const X* x, const X* b, X* y, X omega, hipStream_t s,
                          double* dotp)
{
    const Dia* d = A->dia;
    const int n = A->n_rows;
    dim3 grid((n + kBlockThreads - 1) / kBlockThreads), block(kBlockThreads);
    const int pp = sizeof(X) == 4 ? 1 : (sizeof(V) == 4 ? 2 : 0);
    if (d->lds_x[pp] == 2 && d->box.L > 0) {
        const DiaBox& bx = d->box;
        const int ty = kBlockThreads / bx.tx;
        const int nxt = (bx.L + bx.tx - 1) / bx.tx, nyt = (bx.P / bx.L + ty - 1) / ty;
        const int zc = bx.zc[pp];
        const dim3 bgrid((unsigned)dia_grid(A, pp));
#define DIA_BOX(NDV, TXV)                                                                                            \
    hipLaunchKernelGGL((dia_box_kernel<OP, V, X, NDV, TXV>), bgrid, block, 0, s, n, bx.L, bx.P, nxt, nyt, zc, d->args, \
                       bx.args, d->nd, d->center, d->mask, val, x, b, y, omega, dotp)
#define DIA_BOX_TX(NDV)              \
    if (bx.tx == 64) DIA_BOX(NDV, 64); \
    else if (bx.tx == 32) DIA_BOX(NDV, 32); \
    else DIA_BOX(NDV, 16)
        if (d->nd == 27) {
            DIA_BOX_TX(27);
        } else if (d->nd == 7) {
            DIA_BOX_TX(7);
        } else {
            DIA_BOX_TX(kDiaMax);
        }
#undef DIA_BOX_TX
#undef DIA_BOX
    } else if (d->lds_x[pp] == 1 && d->window.win <= kDiaWin) {
#define DIA_LDS(NDV)                                                                                            \
    hipLaunchKernelGGL((dia_lds_kernel<OP, V, X, NDV>), grid, block, 0, s, n, d->args, d->window, d->nd, d->center, \
                       d->mask, val, x, b, y, omega, dotp)
        if (d->nd == 27) DIA_LDS(27);
        else if (d->nd == 9) DIA_LDS(9);
        else if (d->nd == 7) DIA_LDS(7);
        else DIA_LDS(kDiaMax);
#undef DIA_LDS
    } else {
#define DIA_LAUNCH(NDV)                                                                                         \
    hipLaunchKernelGGL((dia_kernel<OP, V, X, NDV>), grid, block, 0, s, n, d->args, d->nd, d->center, d->mask, val, \
                       x, b, y, omega, dotp)
        if (d->nd == 27) DIA_LAUNCH(27);
        else if (d->nd == 9) DIA_LAUNCH(9);
        else if (d->nd == 7) DIA_LAUNCH(7);
        else DIA_LAUNCH(kDiaMax);
#undef DIA_LAUNCH
    }
    RAMD_HIP(hipGetLastError());
}

template <typename V, typename X>
static void launch_dia(const ramd_csr_s* A, const V* val, int op, const X* x, const X* b, X* y, X omega,
                       hipStream_t s, double* dotp)
{
    switch (op) {
    case OP_SPMV: launch_dia_op<OP_SPMV>(A, val, x, b, y, omega, s, dotp); break;
    case OP_APPEND: launch_dia_op<OP_APPEND>(A, val, x, b, y, omega, s, nullptr); break;
    case OP_RESIDUAL: launch_dia_op<OP_RESIDUAL>(A, val, x, b, y, omega, s, nullptr); break;
    case OP_JACOBI: launch_dia_op<OP_JACOBI>(A, val, x, b, y, omega, s, nullptr); break;
    case OP_ADD_INTO: launch_dia_op<OP_ADD_INTO>(A, val, x, b, y, omega, s, nullptr); break;
    default: fail(RAMD_E_ARG, "unknown stream op %d", op);
    }
}

static hipEvent_t prof_event(ramd_ctx ctx)
{
    if (!ctx->prof_pool.empty()) {
        hipEvent_t e = ctx->prof_pool.back();
        ctx->prof_pool.pop_back();
        return e;
    }
    hipEvent_t e;
    RAMD_HIP(hipEventCreate(&e));
    return e;
}

// Algorithmic bytes of one launch: the matrix streamed once, x gathered with
// ideal reuse (n_cols entries), the output written once; residual, Jacobi
// and append read one more n-vector (b, or y for append).
// In the stencil-diagonal form the matrix costs nd * n values (padding
// included) plus a 4-byte mask per row, and no column indices.
static double op_bytes(const ramd_csr_s* A, int op, int sv, int sx, bool dia)
{
    double b;
    if (dia) b = (double)A->dia->nd * A->n_rows * sv + 4.0 * A->n_rows;
    else b = (double)A->nnz * (sv + 4.0) + 4.0 * (A->n_rows + 1);
    b += (double)sx * A->n_cols + (double)sx * A->n_rows;
    if (op == OP_SPMV_JFZ) b += 2.0 * sx * A->n_rows;  // the diagonal read, the sweep written
    else if (op != OP_SPMV) b += (double)sx * A->n_rows;
    return b;
}

template <typename V, typename X, int CAP>
static void launch_stream_cap(const ramd_csr_s* A, const ramd_csr_s::RowBlocks& rb, const V* val, int op,
                              const X* x, const X* b, X* y, X omega, hipStream_t s, double* dotp, const X* dvec,
                              X* y2)
{
    const int nb = rb.n;
    const dim3 grid(nb), block(kBlockThreads);
    const int df = (int)A->diag_first;
    // the sorted-gather order, when built for these row blocks and values
    // (it holds the CSR values of the vector precision: FP64 or FP32)
    const SortedGather* g = A->sg[sizeof(X) == 4 ? 1 : 0];
    if (g && g->blocks == rb.d && (op == OP_SPMV || op == OP_SPMV_JFZ) && sizeof(V) == sizeof(X) &&
        (sizeof(V) == 4 ? (const void*)val == (const void*)A->val32 : (const void*)val == (const void*)A->val)) {
        const V* gv = static_cast<const V*>(g->val);
        if (op == OP_SPMV)
            hipLaunchKernelGGL((stream_kernel<OP_SPMV, V, X, CAP, true>), grid, block, 0, s, A->rowptr, g->col, gv,
                               rb.d, nb, x, b, y, omega, df, dotp, nullptr, nullptr, g->pos);
        else
            hipLaunchKernelGGL((stream_kernel<OP_SPMV_JFZ, V, X, CAP, true>), grid, block, 0, s, A->rowptr, g->col,
                               gv, rb.d, nb, x, b, y, omega, df, nullptr, dvec, y2, g->pos);
        return;
    }
    switch (op) {
    case OP_SPMV:
        hipLaunchKernelGGL((stream_kernel<OP_SPMV, V, X, CAP>), grid, block, 0, s, A->rowptr, A->col, val, rb.d, nb,
                           x, b, y, omega, df, dotp, nullptr, nullptr, nullptr);
        break;
    case OP_SPMV_JFZ:
        hipLaunchKernelGGL((stream_kernel<OP_SPMV_JFZ, V, X, CAP>), grid, block, 0, s, A->rowptr, A->col, val, rb.d,
                           nb, x, b, y, omega, df, nullptr, dvec, y2, nullptr);
        break;
    case OP_APPEND:
        hipLaunchKernelGGL((stream_kernel<OP_APPEND, V, X, CAP>), grid, block, 0, s, A->rowptr, A->col, val, rb.d,
                           nb, x, b, y, omega, df, nullptr, nullptr, nullptr, nullptr);
        break;
    case OP_RESIDUAL:
        hipLaunchKernelGGL((stream_kernel<OP_RESIDUAL, V, X, CAP>), grid, block, 0, s, A->rowptr, A->col, val, rb.d,
                           nb, x, b, y, omega, df, nullptr, nullptr, nullptr, nullptr);
        break;
    case OP_JACOBI:
        hipLaunchKernelGGL((stream_kernel<OP_JACOBI, V, X, CAP>), grid, block, 0, s, A->rowptr, A->col, val, rb.d,
                           nb, x, b, y, omega, df, nullptr, dvec, nullptr, nullptr);
        break;
    case OP_ADD_INTO:
        hipLaunchKernelGGL((stream_kernel<OP_ADD_INTO, V, X, CAP>), grid, block, 0, s, A->rowptr, A->col, val, rb.d,
                           nb, x, b, y, omega, df, nullptr, nullptr, nullptr, nullptr);
        break;
    default:
        fail(RAMD_E_ARG, "unknown stream op %d", op);
    }
}

// LAYOUT 0: the sliced form; 2: the same with 16-bit column offsets; 1: lanes
// over the CSR rows (val: CSR values)
template <typename V, typename X, int LAYOUT>
static void launch_sliced(const ramd_csr_s* A, const V* csr_val, int op, const X* x, const X* b, X* y, X omega,
                          hipStream_t s, double* dotp, const X* dvec, X* y2)
{
    const Sliced* S = A->sl;
    const V* val = csr_val;
    const int64_t* sptr = nullptr;
    const int* col = A->col;
    const int16_t* col16 = nullptr;
    if (LAYOUT != 1) {
        if constexpr (sizeof(V) == 4) val = S->v32;
        else val = S->v64;
        sptr = S->sptr;
        col = S->col;
        col16 = S->col16;
        RAMD_CHECK(LAYOUT == 0 || col16, RAMD_E_STATE, "16-bit column offsets of the sliced form were not built");
    }
    RAMD_CHECK(val, RAMD_E_STATE, "values of this precision were not built for the lane-per-row kernel");
    const int n = A->n_rows;
    const dim3 grid((n + kBlockThreads - 1) / kBlockThreads), block(kBlockThreads);
    const int df = (int)A->diag_first;
#define SLICED(OPV, DOT, DV, Y2)                                                                                \
    hipLaunchKernelGGL((sliced_kernel<OPV, V, X, LAYOUT>), grid, block, 0, s, n, sptr, A->rowptr, col, col16, val, \
                       x, b, y, omega, df, DOT, DV, Y2)
    switch (op) {
    case OP_SPMV: SLICED(OP_SPMV, dotp, nullptr, nullptr); break;
    case OP_SPMV_JFZ: SLICED(OP_SPMV_JFZ, nullptr, dvec, y2); break;
    case OP_APPEND: SLICED(OP_APPEND, nullptr, nullptr, nullptr); break;
    case OP_RESIDUAL: SLICED(OP_RESIDUAL, nullptr, nullptr, nullptr); break;
    case OP_JACOBI: SLICED(OP_JACOBI, nullptr, dvec, nullptr); break;
    case OP_ADD_INTO: SLICED(OP_ADD_INTO, nullptr, nullptr, nullptr); break;
    default: fail(RAMD_E_ARG, "unknown stream op %d", op);
    }
#undef SLICED
    RAMD_HIP(hipGetLastError());
}

template <typename V, typename X>
static void launch_wave(const ramd_csr_s* A, const V* val, int op, const X* x, const X* b, X* y, X omega,
                        hipStream_t s, double* dotp, const X* dvec, X* y2)
{
    const int n = A->n_rows;
    const dim3 grid((n + kWaveRows - 1) / kWaveRows), block(kBlockThreads);
    const int df = (int)A->diag_first;
#define WAVE(OPV, DOT, DV, Y2)                                                                                     \
    hipLaunchKernelGGL((wave_kernel<OPV, V, X>), grid, block, 0, s, n, A->rowptr, A->col, val, x, b, y, omega, df, \
                       DOT, DV, Y2)
    switch (op) {
    case OP_SPMV: WAVE(OP_SPMV, dotp, nullptr, nullptr); break;
    case OP_SPMV_JFZ: WAVE(OP_SPMV_JFZ, nullptr, dvec, y2); break;
    case OP_APPEND: WAVE(OP_APPEND, nullptr, nullptr, nullptr); break;
    case OP_RESIDUAL: WAVE(OP_RESIDUAL, nullptr, nullptr, nullptr); break;
    case OP_JACOBI: WAVE(OP_JACOBI, nullptr, dvec, nullptr); break;
    case OP_ADD_INTO: WAVE(OP_ADD_INTO, nullptr, nullptr, nullptr); break;
    default: fail(RAMD_E_ARG, "unknown stream op %d", op);
    }
#undef WAVE
    RAMD_HIP(hipGetLastError());
}

// val: CSR values; dval: stencil-diagonal values (nullptr -> CSR kernel)
template <typename V, typename X>
static void launch_stream(const ramd_csr_s* A, const V* val, const V* dval, int op, const X* x, const X* b, X* y,
                          X omega, hipStream_t s, double* dotp = nullptr, const X* dvec = nullptr,
                          X* y2 = nullptr)
{
    if (A->n_rows == 0) return;
    RAMD_CHECK(op != OP_SPMV_JFZ || (!dval && dvec && y2), RAMD_E_STATE,
               "restriction with the first sweep needs a CSR operator, a diagonal and a second output");
    ramd_ctx ctx = A->ctx;
    const bool prof = ctx->prof_min_nnz >= 0 && A->nnz >= ctx->prof_min_nnz;
    ProfPending pend{};
    if (prof) {
        pend.beg = prof_event(ctx);
        pend.end = prof_event(ctx);
        pend.slot = (op == OP_SPMV_JFZ ? OP_SPMV : op) * 2 + (sizeof(X) == 4 ? 1 : 0);
        RAMD_HIP(hipEventRecord(pend.beg, s));
        ctx->prof[pend.slot].bytes += op_bytes(A, op, (int)sizeof(V), (int)sizeof(X), dval != nullptr);
        ctx->prof[pend.slot].launches++;
        if (dval) ctx->prof[pend.slot].dia_launches++;
    }
    if (dval) {
        launch_dia(A, dval, op, x, b, y, omega, s, dotp);
        if (prof) {
            RAMD_HIP(hipEventRecord(pend.end, s));
            ctx->prof_pending.push_back(pend);
        }
        return;
    }
    if (A->wr_on[sizeof(X) == 4 ? 1 : 0]) {
        launch_wave<V, X>(A, val, op, x, b, y, omega, s, dotp, dvec, y2);
        if (prof) {
            RAMD_HIP(hipEventRecord(pend.end, s));
            ctx->prof_pending.push_back(pend);
        }
        return;
    }
    if (A->lr_on[sizeof(X) == 4 ? 1 : 0] || (A->sl && A->sl_on[sizeof(X) == 4 ? 1 : 0])) {
        if (A->lr_on[sizeof(X) == 4 ? 1 : 0]) launch_sliced<V, X, 1>(A, val, op, x, b, y, omega, s, dotp, dvec, y2);
        else if (A->sl16[sizeof(X) == 4 ? 1 : 0])
            launch_sliced<V, X, 2>(A, val, op, x, b, y, omega, s, dotp, dvec, y2);
        else launch_sliced<V, X, 0>(A, val, op, x, b, y, omega, s, dotp, dvec, y2);
        if (prof) {
            RAMD_HIP(hipEventRecord(pend.end, s));
            ctx->prof_pending.push_back(pend);
        }
        return;
    }
    const ramd_csr_s::RowBlocks& rb = A->rb[sizeof(X) == 4 ? 1 : 0];
    switch (rb.cap) {
    case 1024: launch_stream_cap<V, X, 1024>(A, rb, val, op, x, b, y, omega, s, dotp, dvec, y2); break;
    case 2048: launch_stream_cap<V, X, 2048>(A, rb, val, op, x, b, y, omega, s, dotp, dvec, y2); break;
    case 8192:
        // 8192 FP64 products would fill the whole 64 KiB of LDS a workgroup may use
        if constexpr (sizeof(X) == 4)
            launch_stream_cap<V, X, 8192>(A, rb, val, op, x, b, y, omega, s, dotp, dvec, y2);
        else fail(RAMD_E_STATE, "row-block cap 8192 is for FP32 vectors only");
        break;
    default: launch_stream_cap<V, X, 4096>(A, rb, val, op, x, b, y, omega, s, dotp, dvec, y2); break;
    }
    RAMD_HIP(hipGetLastError());
    if (prof) {
        RAMD_HIP(hipEventRecord(pend.end, s));
        ctx->prof_pending.push_back(pend);
    }
}

// Resolves recorded event pairs into the per-slot totals.
static void prof_resolve(ramd_ctx ctx)
{
    if (ctx->prof_pending.empty()) return;
    RAMD_HIP(hipStreamSynchronize(ctx->stream));
    for (const ProfPending& p : ctx->prof_pending) {
        float ms = 0;
        RAMD_HIP(hipEventElapsedTime(&ms, p.beg, p.end));
        ctx->prof[p.slot].ms += ms;
        ctx->prof_pool.push_back(p.beg);
        ctx->prof_pool.push_back(p.end);
    }
    ctx->prof_pending.clear();
}

void launch_stream_op(const ramd_csr_s* A, int op, const double* x, const double* b, double* y, double omega,
                      hipStream_t s, const double* dvec)
{
    launch_stream<double, double>(A, A->val, A->dia ? A->dia->v64 : nullptr, op, x, b, y, omega, s, nullptr, dvec);
}

void launch_stream_op_c(const ramd_csr_s* A, int op, const double* x, const double* b, double* y, double omega,
                        hipStream_t s)
{
    if (A->dia && A->dia->v32 && A->dia->v32_exact && A->v32_outer)
        launch_stream<float, double>(A, nullptr, A->dia->v32, op, x, b, y, omega, s);
    else
        launch_stream_op(A, op, x, b, y, omega, s);
}

void launch_restrict_jfz(const ramd_csr_s* R, const double* x, double* y, double* xc, const double* dc,
                         double omega, hipStream_t s)
{
    launch_stream<double, double>(R, R->val, nullptr, OP_SPMV_JFZ, x, nullptr, y, omega, s, nullptr, dc, xc);
}

void launch_restrict_jfz32(const ramd_csr_s* R, const float* x, float* y, float* xc, const float* dc, float omega,
                           hipStream_t s)
{
    RAMD_CHECK(R->val32, RAMD_E_STATE, "FP32 values of the matrix were not built");
    launch_stream<float, float>(R, R->val32, nullptr, OP_SPMV_JFZ, x, nullptr, y, omega, s, nullptr, dc, xc);
}

void launch_stream_op32(const ramd_csr_s* A, int op, const float* x, const float* b, float* y, float omega,
                        hipStream_t s, const float* dvec)
{
    RAMD_CHECK(A->val32, RAMD_E_STATE, "FP32 values of the matrix were not built");
    RAMD_CHECK(!A->dia || A->dia->v32, RAMD_E_STATE, "FP32 diagonal values of the matrix were not built");
    launch_stream<float, float>(A, A->val32, A->dia ? A->dia->v32 : nullptr, op, x, b, y, omega, s, nullptr, dvec);
}

__global__ void flush_kernel(uint4* __restrict__ p, int64_t n)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = make_uint4(0, 0, 0, 0);
}

float cold_median_ms(ramd_ctx ctx, const std::function<void()>& run, int reps, size_t op_bytes)
{
    hipStream_t s = ctx->stream;
    const bool flush = op_bytes < kFlushBytes;
    if (flush && !ctx->flush_buf) ctx->flush_buf = dmalloc(kFlushBytes);
    std::vector<hipEvent_t> ev(2 * (size_t)reps, nullptr);
    std::vector<float> ms((size_t)reps, 0.f);
    try {
        for (hipEvent_t& e : ev) RAMD_HIP(hipEventCreate(&e));
        run();  // first launch: code loading
        for (int r = 0; r < reps; r++) {
            if (flush) {
                hipLaunchKernelGGL(flush_kernel, dim3(4096), dim3(256), 0, s, (uint4*)ctx->flush_buf,
                                   (int64_t)(kFlushBytes / sizeof(uint4)));
                RAMD_HIP(hipGetLastError());
            }
            RAMD_HIP(hipEventRecord(ev[2 * r], s));
            run();
            RAMD_HIP(hipEventRecord(ev[2 * r + 1], s));
        }
        RAMD_HIP(hipEventSynchronize(ev.back()));
        for (int r = 0; r < reps; r++) RAMD_HIP(hipEventElapsedTime(&ms[r], ev[2 * r], ev[2 * r + 1]));
    } catch (...) {
        for (hipEvent_t e : ev)
            if (e) (void)hipEventDestroy(e);
        throw;
    }
    for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    std::nth_element(ms.begin(), ms.begin() + reps / 2, ms.end());
    return ms[reps / 2];
}

void csr_autotune(ramd_csr_s* A, int prec)
{
    if (A->n_rows == 0 || A->nnz == 0 || A->dia) return;
    RAMD_CHECK(prec == 0 || A->val32, RAMD_E_STATE, "csr_autotune: FP32 values were not built");
    ramd_ctx ctx = A->ctx;
    hipStream_t s = ctx->stream;
    RAMD_HIP(hipStreamSynchronize(s));
    A->forms++;  // kernel choice and row blocks change
    if (A->rb[prec].d != A->blocks) {  // tuned before: start from the default blocks
        dfree(A->rb[prec].d);
        A->rb[prec] = {kStreamCap, A->n_blocks, A->blocks};
    }
    A->sl_on[prec] = false;
    A->sl16[prec] = false;
    A->lr_on[prec] = false;
    A->wr_on[prec] = false;
    sg_free(A, prec);
    const size_t w = prec ? sizeof(float) : sizeof(double);
    struct Scratch {
        void *x = nullptr, *b = nullptr, *y = nullptr;
        std::vector<int*> blocks;
        ~Scratch()
        {
            dfree(x);
            dfree(b);
            dfree(y);
            for (int* p : blocks) dfree(p);
        }
    } t;
    t.x = dmalloc(w * std::max(A->n_cols, 1));
    t.b = dmalloc(w * std::max(A->n_rows, 1));
    t.y = dmalloc(w * std::max(A->n_rows, 1));
    RAMD_HIP(hipMemsetAsync(t.x, 0, w * std::max(A->n_cols, 1), s));
    RAMD_HIP(hipMemsetAsync(t.b, 0, w * std::max(A->n_rows, 1), s));
    const int op = A->n_rows == A->n_cols ? OP_RESIDUAL : OP_SPMV;
    ramd_csr_s::RowBlocks best = A->rb[prec];
    float best_ms = 1e30f;
    // the median of seven launches for small operators, whose single
    // launches are too short to rank reliably, three above
    const int reps = A->nnz < (int64_t)4 << 20 ? 7 : 3;
    const size_t op_b = (size_t)A->nnz * (4 + w) + ((size_t)A->n_rows * 2 + (size_t)A->n_cols) * w;
    auto time_ms = [&] {
        return cold_median_ms(ctx, [&] {
            if (prec) launch_stream<float, float>(A, A->val32, nullptr, op, (const float*)t.x, (const float*)t.b,
                                                  (float*)t.y, 0.f, s);
            else launch_stream<double, double>(A, A->val, nullptr, op, (const double*)t.x, (const double*)t.b,
                                               (double*)t.y, 0.0, s);
        }, reps, op_b);
    };
    for (int cap : kStreamCaps) {
        if (prec == 0 && cap > 4096) continue;
        ramd_csr_s::RowBlocks rb;
        rb.cap = cap;
        if (cap == kStreamCap) {
            rb.n = A->n_blocks;
            rb.d = A->blocks;
        } else {
            std::vector<int> blk = make_row_blocks(A->h_rowptr.data(), A->n_rows, cap);
            rb.n = (int)blk.size() - 1;
            rb.d = dalloc<int>(blk.size());
            t.blocks.push_back(rb.d);
            RAMD_HIP(hipMemcpyAsync(rb.d, blk.data(), sizeof(int) * blk.size(), hipMemcpyHostToDevice, s));
        }
        A->rb[prec] = rb;
        const float ms = time_ms();
        if (ms < best_ms) {
            best_ms = ms;
            best = rb;
        }
    }
    // keep the winner's blocks, release the others
    for (int*& p : t.blocks)
        if (p == best.d) p = nullptr;
    A->rb[prec] = best;
    // restrictions (more columns than rows: rows gather the finer vector at
    // scattered columns): the same blocks with the entries in sorted-gather
    // order
    if (A->n_cols > A->n_rows) {
        sg_build(A, prec);
        const float ms = time_ms();
        if (ms < best_ms) best_ms = ms;
        else sg_free(A, prec);
    }
    // lanes straight over the CSR rows (no extra storage), for short rows:
    // a lane adds its whole row alone
    if ((double)A->nnz <= kLaneRowsMaxLen * (double)A->n_rows) {
        A->lr_on[prec] = true;
        const float ms = time_ms();
        A->lr_on[prec] = ms < best_ms;
        if (ms < best_ms) best_ms = ms;
    }
    // square operators with long rows: the sliced form against the best so far
    if (A->n_rows == A->n_cols && (double)A->nnz >= kSlicedMinRowLen * (double)A->n_rows) {
        sliced_build(A);
        if ((double)A->sl->entries <= kSlicedMaxPad * (double)A->nnz) {
            sliced_values(A, prec);
            const bool lr = A->lr_on[prec];
            A->lr_on[prec] = false;
            A->sl_on[prec] = true;
            float ms = time_ms();
            if (A->sl->col16) {  // and with 16-bit column offsets
                A->sl16[prec] = true;
                const float ms16 = time_ms();
                A->sl16[prec] = ms16 < ms;
                ms = std::min(ms, ms16);
            }
            A->sl_on[prec] = ms < best_ms;
            if (!A->sl_on[prec]) A->sl16[prec] = false;
            A->lr_on[prec] = lr && !A->sl_on[prec];
        }
        if (!A->sl_on[prec]) {  // release what this precision built
            RAMD_HIP(hipStreamSynchronize(s));
            if (prec) {
                dfree(A->sl->v32);
                A->sl->v32 = nullptr;
            } else {
                dfree(A->sl->v64);
                A->sl->v64 = nullptr;
            }
            if (!A->sl_on[0] && !A->sl_on[1]) sliced_free(A);
        }
    }
    // few long rows: one wavefront per row against the best so far
    if ((double)A->nnz >= kWaveRowsMinLen * (double)A->n_rows) {
        const bool lr = A->lr_on[prec], sl = A->sl_on[prec];
        A->wr_on[prec] = true;
        const float ms = time_ms();
        if (ms < best_ms) {
            best_ms = ms;
            A->lr_on[prec] = false;
            if (sl) {
                A->sl_on[prec] = false;
                A->sl16[prec] = false;
                RAMD_HIP(hipStreamSynchronize(s));
                if (prec) {
                    dfree(A->sl->v32);
                    A->sl->v32 = nullptr;
                } else {
                    dfree(A->sl->v64);
                    A->sl->v64 = nullptr;
                }
                if (!A->sl_on[0] && !A->sl_on[1]) sliced_free(A);
            }
        } else {
            A->wr_on[prec] = false;
            A->lr_on[prec] = lr;
        }
    }
    if (A->lr_on[prec] || A->sl_on[prec] || A->wr_on[prec]) sg_free(A, prec);
}

void dia_autotune(ramd_csr_s* A, int pp)
{
    Dia* d = A->dia;
    if (!d || A->n_rows == 0 || d->lds_x[pp] >= 0) return;
    RAMD_CHECK(pp == 0 || d->v32, RAMD_E_STATE, "dia_autotune: FP32 values were not built");
    ramd_ctx ctx = A->ctx;
    hipStream_t s = ctx->stream;
    const size_t w = pp == 1 ? sizeof(float) : sizeof(double);
    const size_t bytes = w * (size_t)A->n_rows;
    struct Scratch {
        void *x = nullptr, *b = nullptr, *y = nullptr;
        double* part = nullptr;
        ~Scratch()
        {
            dfree(x);
            dfree(b);
            dfree(y);
            dfree(part);
        }
    } t;
    t.x = dmalloc(bytes);
    t.b = dmalloc(bytes);
    t.y = dmalloc(bytes);
    t.part = dalloc<double>((size_t)dia_grid_bound(A));
    RAMD_HIP(hipMemsetAsync(t.x, 0, bytes, s));
    RAMD_HIP(hipMemsetAsync(t.b, 0, bytes, s));
    // bytes an operation moves: values, mask and three vectors per row
    const size_t op_b = (size_t)A->n_rows * ((size_t)d->nd * (pp == 0 ? 8 : 4) + 4 + 3 * w);
    // candidates: (mode, planes per workgroup)
    std::vector<std::pair<int, int>> cands = {{0, 0}};
    if (d->window.win <= kDiaWin) cands.push_back({1, 0});
    // the plane-march kernel with four and with eight planes per workgroup
    // (fewer planes: more workgroups in flight; more: fewer halo planes of x
    // loaded, which matters most for FP64 x); 16 and 32 planes were slower
    // on the whole 256^3 FP32-cycle solve (24.55, 24.62, 24.77 ms for 4, 8,
    // 16 planes on every operator)
    if (d->box.L > 0) {
        cands.push_back({2, kBoxPlanes});
        cands.push_back({2, 2 * kBoxPlanes});
    }
    float best = 1e30f;
    std::pair<int, int> pick = cands[0];
    const int zc0 = d->box.zc[pp];
    for (const auto& c : cands) {
        d->lds_x[pp] = c.first;
        if (c.first == 2) d->box.zc[pp] = c.second;
        // the operations each pair runs: the cycle's residual and Jacobi
        // sweep, the outer product with its fused dot
        auto run = [&](int op) {
            if (pp == 1)
                launch_dia<float, float>(A, d->v32, op, (const float*)t.x, (const float*)t.b, (float*)t.y, 0.5f, s,
                                         nullptr);
            else if (pp == 2)
                launch_dia<float, double>(A, d->v32, OP_SPMV, (const double*)t.x, (const double*)t.b, (double*)t.y,
                                          0.0, s, t.part);
            else
                launch_dia<double, double>(A, d->v64, op, (const double*)t.x, (const double*)t.b, (double*)t.y, 0.5,
                                           s, nullptr);
        };
        // medians of seven launches: the candidates differ by a few percent
        float ms = cold_median_ms(ctx, [&] { run(pp == 2 ? OP_SPMV : OP_RESIDUAL); }, 7, op_b);
        if (pp != 2) ms += cold_median_ms(ctx, [&] { run(OP_JACOBI); }, 7, op_b);
        if (ms < best) {
            best = ms;
            pick = c;
        }
    }
    d->lds_x[pp] = pick.first;
    d->box.zc[pp] = pick.first == 2 ? pick.second : zc0;
    A->forms++;
}

double* ctx_partials(ramd_ctx ctx, size_t count)
{
    if (count > ctx->big_cap) {
        RAMD_HIP(hipStreamSynchronize(ctx->stream));
        dfree(ctx->big_partials);
        ctx->big_partials = nullptr;
        ctx->big_cap = 0;
        ctx->big_partials = dalloc<double>(count);
        ctx->big_cap = count;
    }
    return ctx->big_partials;
}

int spmv_dot_partials(const ramd_csr_s* A)
{
    if (A->dia) return dia_grid_bound(A);  // every kernel choice of either value precision
    const int np = A->wr_on[0] ? (A->n_rows + kWaveRows - 1) / kWaveRows
                   : (A->lr_on[0] || (A->sl && A->sl_on[0])) ? (A->n_rows + kBlockThreads - 1) / kBlockThreads
                                                             : A->rb[0].n;
    return std::max(np, 1);
}

void launch_spmv_dot(const ramd_csr_s* A, bool values_fp32, const double* x, double* y, double* out_d,
                     hipStream_t s, double* partials)
{
    ramd_ctx ctx = A->ctx;
    const bool dia32 = values_fp32 && A->dia && A->dia->v32 && A->dia->v32_exact && A->v32_outer;
    const int np = A->dia ? dia_grid(A, dia32 ? 2 : 0) : spmv_dot_partials(A);
    double* part = partials ? partials : ctx_partials(ctx, (size_t)np);
    if (dia32)
        launch_stream<float, double>(A, nullptr, A->dia->v32, OP_SPMV, x, nullptr, y, 0.0, s, part);
    else
        launch_stream<double, double>(A, A->val, A->dia ? A->dia->v64 : nullptr, OP_SPMV, x, nullptr, y, 0.0, s,
                                      part);
    launch_reduce_final(ctx, part, np, out_d);
}

// ---------------------------------------------------------------------------
// Deterministic reductions: a fixed grid of workgroups, each summing a fixed
// strided subset in a fixed order, then one workgroup combining the partials
// in index order.  Same inputs -> same bits, every run.
// ---------------------------------------------------------------------------
template <int MODE>  // 0: x.y   1: sum x^2 over |x| > 1e-16
__global__ __launch_bounds__(kBlockThreads) void reduce_partial(const double* __restrict__ x,
                                                                const double* __restrict__ y, int64_t n,
                                                                double* __restrict__ partials)
{
    double s = 0;
    const int64_t stride = (int64_t)gridDim.x * kBlockThreads;
    // four of the thread's entries per step, their loads issued first (the
    // order of the sum is the plain loop's)
    constexpr int U = 4;
    for (int64_t i0 = (int64_t)blockIdx.x * kBlockThreads + threadIdx.x; i0 < n; i0 += U * stride) {
        double xv[U], yv[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int64_t i = i0 + u * stride;
            if (i < n) {
                xv[u] = x[i];
                if (MODE == 0) yv[u] = y[i];
            }
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            if (i0 + u * stride >= n) break;
            if (MODE == 0) {
                s += xv[u] * yv[u];
            } else {
                const double v = xv[u];
                if (fabs(v) > 1e-16) s += v * v;
            }
        }
    }
    s = block_sum(s);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

__global__ __launch_bounds__(kBlockThreads) void reduce_final(const double* __restrict__ partials, int np,
                                                              double* __restrict__ out)
{
    double s = strided_sum(partials, 0, np);
    s = block_sum(s);
    if (threadIdx.x == 0) *out = s;
}

// Two reduce_final sums in one launch (same operations, same bits):
// *o1 = the sum of p1[0, n1), *o2 = the sum of p2[0, n2).
__global__ __launch_bounds__(kBlockThreads) void reduce_final_two(const double* __restrict__ p1, int n1,
                                                                  double* __restrict__ o1,
                                                                  const double* __restrict__ p2, int n2,
                                                                  double* __restrict__ o2)
{
    double s1 = strided_sum(p1, 0, n1);
    s1 = block_sum(s1);
    double s2 = strided_sum(p2, 0, n2);
    s2 = block_sum(s2);
    if (threadIdx.x == 0) {
        *o1 = s1;
        *o2 = s2;
    }
}

void launch_reduce_final_two(ramd_ctx ctx, const double* p1, int n1, double* o1, const double* p2, int n2,
                             double* o2)
{
    RAMD_CHECK(n1 >= 1 && n1 <= 4096 && n2 >= 1 && n2 <= 4096, RAMD_E_ARG, "reduce_final_two: %d and %d partial sums",
               n1, n2);
    hipLaunchKernelGGL(reduce_final_two, dim3(1), dim3(kBlockThreads), 0, ctx->stream, p1, n1, o1, p2, n2, o2);
    RAMD_HIP(hipGetLastError());
}

// reduce_final (same operations, same bits) that also publishes two scalars
// to device-writable host memory: host[0] = *also (a scalar an earlier
// kernel of the stream wrote), host[1] = the sum.  The host then waits on an
// event recorded after it instead of copying the scalars back.
__global__ __launch_bounds__(kBlockThreads) void reduce_final_pub(const double* __restrict__ partials, int np,
                                                                  double* __restrict__ out,
                                                                  const double* __restrict__ also, double* host)
{
    double s = strided_sum(partials, 0, np);
    s = block_sum(s);
    if (threadIdx.x == 0) {
        *out = s;
        host[0] = *also;
        host[1] = s;
    }
}

void launch_reduce_final_pub(ramd_ctx ctx, const double* partials, int np, double* out_d, const double* also_d,
                             double* host)
{
    RAMD_CHECK(np >= 1 && np <= 4096, RAMD_E_ARG, "reduce_final_pub: %d partial sums", np);
    hipLaunchKernelGGL(reduce_final_pub, dim3(1), dim3(kBlockThreads), 0, ctx->stream, partials, np, out_d, also_d,
                       host);
    RAMD_HIP(hipGetLastError());
}

static int red_blocks(int64_t n)
{
    int64_t b = (n + kBlockThreads - 1) / kBlockThreads;
    return (int)std::max<int64_t>(1, std::min<int64_t>(b, kRedBlocks));
}

void launch_dot(ramd_ctx ctx, int64_t n, const double* x, const double* y, double* out_d)
{
    const int nb = red_blocks(n);
    hipLaunchKernelGGL(reduce_partial<0>, dim3(nb), dim3(kBlockThreads), 0, ctx->stream, x, y, n,
                       ctx->red_partials);
    hipLaunchKernelGGL(reduce_final, dim3(1), dim3(kBlockThreads), 0, ctx->stream, ctx->red_partials, nb, out_d);
    RAMD_HIP(hipGetLastError());
}

void launch_norm2sq(ramd_ctx ctx, int64_t n, const double* x, double* out_d)
{
    const int nb = red_blocks(n);
    hipLaunchKernelGGL(reduce_partial<1>, dim3(nb), dim3(kBlockThreads), 0, ctx->stream, x, x, n,
                       ctx->red_partials);
    hipLaunchKernelGGL(reduce_final, dim3(1), dim3(kBlockThreads), 0, ctx->stream, ctx->red_partials, nb, out_d);
    RAMD_HIP(hipGetLastError());
}

// First stage for long partial arrays: workgroup g sums the contiguous
// chunk [g*chunk, (g+1)*chunk) in a fixed order.
__global__ __launch_bounds__(kBlockThreads) void reduce_chunks(const double* __restrict__ partials, int np, int chunk,
                                                               double* __restrict__ out)
{
    const int b0 = blockIdx.x * chunk, b1 = min(np, b0 + chunk);
    double s = strided_sum(partials, b0, b1);
    s = block_sum(s);
    if (threadIdx.x == 0) out[blockIdx.x] = s;
}

void launch_reduce_final(ramd_ctx ctx, const double* partials, int np, double* out_d)
{
    constexpr int kChunks = 256;
    if (np > 4096) {
        // one workgroup would read every partial serially: split in two stages
        const int chunk = (np + kChunks - 1) / kChunks;
        double* mid = ctx->red_partials + 3072;  // 256 doubles past the grid-stride partials
        hipLaunchKernelGGL(reduce_chunks, dim3(kChunks), dim3(kBlockThreads), 0, ctx->stream, partials, np, chunk, mid);
        hipLaunchKernelGGL(reduce_final, dim3(1), dim3(kBlockThreads), 0, ctx->stream, mid, kChunks, out_d);
    } else {
        hipLaunchKernelGGL(reduce_final, dim3(1), dim3(kBlockThreads), 0, ctx->stream, partials, np, out_d);
    }
    RAMD_HIP(hipGetLastError());
}

void fetch_scalars(ramd_ctx ctx, const double* d, int k, double* out)
{
    RAMD_CHECK(k >= 1 && k <= 16, RAMD_E_ARG, "fetch_scalars: %d scalars", k);
    RAMD_HIP(hipMemcpyAsync(ctx->pinned, d, sizeof(double) * k, hipMemcpyDeviceToHost, ctx->stream));
    RAMD_HIP(hipStreamSynchronize(ctx->stream));
    for (int i = 0; i < k; i++) out[i] = ctx->pinned[i];
}

double fetch_scalar(ramd_ctx ctx, const double* d)
{
    RAMD_HIP(hipMemcpyAsync(ctx->pinned, d, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    RAMD_HIP(hipStreamSynchronize(ctx->stream));
    return ctx->pinned[0];
}

__global__ void axpy_kernel(int64_t n, double alpha, const double* __restrict__ x, double* __restrict__ y)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) y[i] += x[i] * alpha;
}

}  // namespace ramd

using namespace ramd;

static void check_vec_op(ramd_csr A, const void* a, const void* b, const char* who)
{
    RAMD_CHECK(A, RAMD_E_ARG, "%s: NULL matrix", who);
    RAMD_CHECK((a && b) || (A->n_rows == 0), RAMD_E_ARG, "%s: NULL vector", who);
    A->ctx->set_current();
}

extern "C" {

int ramd_spmv(ramd_csr A, const double* x_d, double* y_d)
{
    return guarded([&] {
        check_vec_op(A, x_d, y_d, "ramd_spmv");
        launch_stream_op(A, OP_SPMV, x_d, nullptr, y_d, 0.0, A->ctx->stream);
    });
}

int ramd_spmv_append(ramd_csr A, const double* x_d, double* y_d)
{
    return guarded([&] {
        check_vec_op(A, x_d, y_d, "ramd_spmv_append");
        launch_stream_op(A, OP_APPEND, x_d, nullptr, y_d, 0.0, A->ctx->stream);
    });
}

int ramd_residual(ramd_csr A, const double* x_d, const double* b_d, double* r_d)
{
    return guarded([&] {
        check_vec_op(A, x_d, r_d, "ramd_residual");
        RAMD_CHECK(b_d || A->n_rows == 0, RAMD_E_ARG, "ramd_residual: NULL b");
        launch_stream_op(A, OP_RESIDUAL, x_d, b_d, r_d, 0.0, A->ctx->stream);
    });
}

int ramd_spmv_t(ramd_csr A, const double* x_d, double* y_d)
{
    return guarded([&] {
        check_vec_op(A, x_d, y_d, "ramd_spmv_t");
        ramd_csr_s* T = csr_transpose(A);
        launch_stream_op(T, OP_SPMV, x_d, nullptr, y_d, 0.0, A->ctx->stream);
    });
}

int ramd_jacobi(ramd_csr A, double* x_d, const double* b_d, double* tmp_d, int sweeps, double omega)
{
    return guarded([&] {
        check_vec_op(A, x_d, b_d, "ramd_jacobi");
        RAMD_CHECK(tmp_d || A->n_rows == 0, RAMD_E_ARG, "ramd_jacobi: NULL tmp");
        RAMD_CHECK(A->n_rows == A->n_cols, RAMD_E_ARG, "ramd_jacobi: matrix is %dx%d, not square", A->n_rows,
                   A->n_cols);
        for (int it = 0; it < sweeps; it++) {
            RAMD_HIP(hipMemcpyAsync(tmp_d, x_d, sizeof(double) * A->n_rows, hipMemcpyDeviceToDevice,
                                    A->ctx->stream));
            launch_stream_op(A, OP_JACOBI, tmp_d, b_d, x_d, omega, A->ctx->stream);
        }
    });
}

int ramd_dot(ramd_ctx ctx, int64_t n, const double* x_d, const double* y_d, double* out)
{
    return guarded([&] {
        RAMD_CHECK(ctx && out && n >= 0 && (n == 0 || (x_d && y_d)), RAMD_E_ARG, "ramd_dot: bad argument");
        ctx->set_current();
        launch_dot(ctx, n, x_d, y_d, ctx->red_result);
        *out = fetch_scalar(ctx, ctx->red_result);
    });
}

int ramd_norm2(ramd_ctx ctx, int64_t n, const double* x_d, double* out)
{
    return guarded([&] {
        RAMD_CHECK(ctx && out && n >= 0 && (n == 0 || x_d), RAMD_E_ARG, "ramd_norm2: bad argument");
        ctx->set_current();
        launch_norm2sq(ctx, n, x_d, ctx->red_result);
        double r = fetch_scalar(ctx, ctx->red_result);
        // core/vector.cpp:117 and core/par_vector.cpp:94-97: root, undo, root
        r = pow(r, 1.0 / 2);
        r = pow(r, 2);
        *out = pow(r, 1. / 2);
    });
}

int ramd_prof_enable(ramd_ctx ctx, int64_t min_nnz)
{
    return guarded([&] {
        RAMD_CHECK(ctx, RAMD_E_ARG, "ramd_prof_enable: NULL context");
        prof_resolve(ctx);
        ctx->prof_min_nnz = min_nnz;
        for (ProfSlot& s : ctx->prof) s = ProfSlot();
    });
}

int ramd_prof_read(ramd_ctx ctx, int op, int prec, int64_t* launches, int64_t* dia_launches, double* ms,
                   double* bytes)
{
    return guarded([&] {
        RAMD_CHECK(ctx && op >= 0 && op < 5 && (prec == RAMD_PREC_FP64 || prec == RAMD_PREC_FP32), RAMD_E_ARG,
                   "ramd_prof_read: bad argument");
        prof_resolve(ctx);
        const ProfSlot& s = ctx->prof[op * 2 + prec];
        if (launches) *launches = s.launches;
        if (dia_launches) *dia_launches = s.dia_launches;
        if (ms) *ms = s.ms;
        if (bytes) *bytes = s.bytes;
    });
}

int ramd_axpy(ramd_ctx ctx, int64_t n, double alpha, const double* x_d, double* y_d)
{
    return guarded([&] {
        RAMD_CHECK(ctx && n >= 0 && (n == 0 || (x_d && y_d)), RAMD_E_ARG, "ramd_axpy: bad argument");
        ctx->set_current();
        if (n == 0) return;
        int grid = (int)std::min<int64_t>((n + 255) / 256, 16384);
        hipLaunchKernelGGL(axpy_kernel, dim3(grid), dim3(256), 0, ctx->stream, n, alpha, x_d, y_d);
        RAMD_HIP(hipGetLastError());
    });
}

}  // extern "C"

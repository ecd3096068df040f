// Shared definitions of libraptor_amd: error plumbing, the context and the
// device-resident CSR handle.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <functional>
#include <string>
#include <vector>

#include "raptor_amd.h"
#include "status.hpp"

namespace ramd {

#define RAMD_HIP(call)                                                                    \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess)                                                             \
            ::ramd::fail(RAMD_E_DEVICE, "%s failed: %s (%s:%d)", #call,                   \
                         hipGetErrorString(e_), __FILE__, __LINE__);                      \
    } while (0)

// Device allocation that reports RAMD_E_ALLOC with the size.
void* dmalloc(size_t bytes);
void dfree(void* p);

template <class T>
T* dalloc(size_t n)
{
    return static_cast<T*>(dmalloc(n * sizeof(T)));
}

// Launch geometry of the row-block ("CSR-stream") kernels.
constexpr int kBlockThreads = 256;   // one workgroup = 4 wavefronts of 64
// the sliced-row form is tried for square operators averaging at least this
// many entries per row, when padding adds at most half the entries again
constexpr double kSlicedMinRowLen = 12.0;
constexpr double kSlicedMaxPad = 1.5;
// the lane-per-row CSR kernel is tried for operators averaging at most this
// many entries per row
constexpr double kLaneRowsMaxLen = 8.0;
// the wavefront-per-row kernel is tried for operators averaging at least
// this many entries per row
constexpr double kWaveRowsMinLen = 24.0;
constexpr int kStreamCap = 4096;     // products staged in LDS per workgroup
constexpr int kMaxRowsPerBlock = 256;

// Row blocks for the LDS-staged kernels: consecutive rows whose nonzeros fit
// in kStreamCap products and number at most kMaxRowsPerBlock; a row longer
// than kStreamCap forms a block of its own.
std::vector<int> make_row_blocks(const int* rowptr_host, int n_rows, int cap);

// Keeps XCD-local row ranges together: workgroup ids are dealt round-robin
// over the 8 XCDs, so remap id -> (id % 8) chunk so each XCD sweeps one
// contiguous slice of rows (one L2 per XCD keeps its x halo).
__device__ __forceinline__ int xcd_remap(int bid, int nb)
{
    int q = nb >> 3, r = nb & 7, x = bid & 7, k = bid >> 3;
    return x * q + (x < r ? x : r) + k;
}


// Thread t's share of a one-workgroup reduction: p[i0 + t], p[i0 + t + T],
// ... below i1 (T = threads per workgroup, `stride` apart in memory), added
// in that order; eight loads are issued before their additions, so a
// thread waits for one memory round trip per eight entries instead of per
// entry (the additions and their order are the plain loop's).
__device__ __forceinline__ double strided_sum(const double* __restrict__ p, int i0, int i1, size_t stride = 1)
{
    constexpr int T = kBlockThreads;
    double s = 0;
    int i = i0 + (int)threadIdx.x;
    for (; i + 7 * T < i1; i += 8 * T) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; u++) v[u] = p[(size_t)(i + u * T) * stride];
#pragma unroll
        for (int u = 0; u < 8; u++) s += v[u];
    }
    for (; i < i1; i += T) s += p[(size_t)i * stride];
    return s;
}

// Workgroup sum in a fixed order (wavefront shuffle tree, then the four
// wavefront sums in index order); the result is valid in thread 0.  Every
// thread of the workgroup must call it.
__device__ __forceinline__ double block_sum(double v)
{
    __shared__ double wsum[kBlockThreads / 64];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) wsum[w] = v;
    __syncthreads();
    double t = 0;
    if (threadIdx.x == 0)
        for (int i = 0; i < kBlockThreads / 64; i++) t += wsum[i];
    __syncthreads();
    return t;
}

}  // namespace ramd

// Live kernel timing: hipEvent pairs around the row-block kernels of large
// matrices, accumulated per (operation, precision).
struct ProfSlot {
    int64_t launches = 0;
    int64_t dia_launches = 0;  // of which in the stencil-diagonal form
    double bytes = 0;  // algorithmic bytes, summed over launches
    double ms = 0;     // resolved event time
};
struct ProfPending {
    hipEvent_t beg, end;
    int slot;
};

struct ramd_ctx_s {
    int device = 0;
    hipStream_t stream = nullptr;
    int64_t prof_min_nnz = -1;  // < 0: off
    ProfSlot prof[10];
    std::vector<ProfPending> prof_pending;
    std::vector<hipEvent_t> prof_pool;
    // scratch for reductions: partial sums
    double* red_partials = nullptr;
    double* red_result = nullptr;  // device scalars (16 slots)
    double* pinned = nullptr;      // pinned host staging for scalars (16 slots)
    // coherent pinned host memory that kernels write (8 slots; the solvers'
    // per-iteration scalars) and the event recorded after such a write
    double* step_pin = nullptr;
    double* step_pin_d = nullptr;  // its device address
    hipEvent_t step_ev = nullptr;
    // per-workgroup partial sums of fused kernels (grown on demand)
    double* big_partials = nullptr;
    size_t big_cap = 0;
    // a buffer larger than the L2 and infinity caches together, written
    // before each autotuner launch (allocated at the first autotune)
    void* flush_buf = nullptr;
    // solvers replay captured HIP graphs of their iterations (ramd_set_graphs;
    // off by default: eager launches measured faster)
    bool graphs = false;
    void set_current() const;
};

// Stencil-diagonal form of a matrix whose rows use few distinct column
// offsets (col - row): values stored per diagonal (val[k * n + row], zero
// where absent), a per-row presence mask, and the order in which every row
// stores its diagonals (ascending, or diagonal first then ascending).  The
// kernels visit a row's entries in that order, so sums are bit-identical to
// the CSR ones; column indices are never read.
constexpr int kDiaMax = 32;
struct DiaArgs {
    int off[kDiaMax];  // offset of diagonal k (ascending in k)
    int ord[kDiaMax];  // visiting order: ord[kk] = diagonal index
    int pos[kDiaMax];  // slot of diagonal k: ord[pos[k]] == k
};
// Stencil-diagonal values are stored by slot (position in the rows' common
// stored order) in 16-byte groups: slot t of row r sits at
// ((t / G) * n + r) * G + t % G with G = 16 / sizeof(V), so one lane reads
// G consecutive slots of its row with one 16-byte load and a wavefront's
// load is one contiguous run.  Slots past nd (up to a multiple of G) hold 0.
template <typename V>
__host__ __device__ __forceinline__ size_t dia_index(int slot, int row, int n)
{
    constexpr int G = 16 / (int)sizeof(V);
    return ((size_t)(slot / G) * (size_t)n + (size_t)row) * G + (size_t)(slot % G);
}
// Column read for diagonal offset off of row: row + off clamped into
// [0, n).  The address does not wait for the row's mask; a clamped (or
// absent-diagonal) value is only ever multiplied into a product that the
// mask's select then drops.
__device__ __forceinline__ int dia_col(int row, int off, int n)
{
    const int c = row + off;
    return c < 0 ? 0 : (c >= n ? n - 1 : c);
}
template <typename V>
__host__ __device__ inline int dia_slots(int nd)
{
    constexpr int G = 16 / (int)sizeof(V);
    return (nd + G - 1) / G * G;
}
// x window of a 256-row workgroup for the LDS-staged kernels: the offsets
// form runs of consecutive values; run j needs x[r0 + lo[j] .. r0 + 255 +
// lo[j] + len[j] - 1], stored at base[j] in LDS, and diagonal k is read at
// lofs[k] + (row - r0).
constexpr int kDiaWin = 2560;  // LDS values per workgroup (20 KiB in FP64)
struct DiaWindow {
    int nr = 0;
    int win = 0;            // total values
    int lo[kDiaMax];
    int len[kDiaMax];
    int base[kDiaMax];
    int lofs[kDiaMax];
};
// Box geometry of a stencil-diagonal form: every offset is
// dz * P + dy * L + dx with |dx|, |dy|, |dz| <= 1, P a multiple of L and n a
// multiple of P (rows of an L x (P / L) x (n / P) grid in x-fastest order).
// The plane-march kernel tiles rows TX x TY per workgroup and walks zc
// planes; diagonal k reads x from the LDS plane dz[k] at lane offset rel[k].
struct BoxArgs {
    int dz[kDiaMax];   // plane step of diagonal k: -1, 0 or 1
    int rel[kDiaMax];  // dy * (TX + 2) + dx of diagonal k
};
// Planes per workgroup of the plane-march kernel: at least this many.
constexpr int kBoxMinPlanes = 4;
constexpr int kBoxPlanes = 4;  // the autotuner's choice
constexpr int kBoxMaxPlanes = 256;
struct DiaBox {
    int L = 0, P = 0;  // 0: the offsets have no box decomposition
    int tx = 64;       // tile width (TY = kBlockThreads / tx)
    BoxArgs args{};
    int zc[3] = {kBoxPlanes, kBoxPlanes, kBoxPlanes};  // planes per workgroup, per precision pair
};
struct Dia {
    int nd = 0;
    int center = -1;            // diagonal index of offset 0 (-1 if none)
    DiaArgs args{};
    DiaWindow window{};
    DiaBox box{};
    // per (values, vectors) precision ([0] FP64/FP64, [1] FP32/FP32,
    // [2] FP32/FP64): 2 = plane-march with the x planes in LDS (box
    // geometry only), 1 = stage the x window in LDS, 0 = gather, -1 = not
    // timed yet (gather).  Chosen by timing; the result bits are the same.
    int lds_x[3] = {-1, -1, -1};
    unsigned* mask = nullptr;   // device, n
    double* v64 = nullptr;      // device, dia_slots<double>(nd) * n, dia_index layout
    float* v32 = nullptr;       // device, dia_slots<float>(nd) * n (built with the FP32 copy)
    bool v32_exact = false;     // every value is exactly representable in FP32
};

namespace ramd {
// Sliced rows: 64 consecutive rows per slice (one wavefront, one lane per
// row), the slice's entries stored column-major (entry k of lane l at
// sptr[q] + 64k + l) and padded to its longest row with column -1.  The
// one-lane-per-row form of operators with long rows: every load of a
// wavefront is one contiguous run.
// A square operator whose columns all lie within 32767 of their row also
// gets 16-bit column offsets (col - row; padding kSliced16Pad), two bytes per
// entry less to stream.
constexpr int16_t kSliced16Pad = -32768;
struct Sliced {
    int n_slices = 0;
    int64_t entries = 0;
    int64_t* sptr = nullptr;  // device, n_slices + 1
    int* col = nullptr;       // device, entries
    int16_t* col16 = nullptr; // device, entries (when the offsets fit)
    double* v64 = nullptr;    // device, entries (built on demand)
    float* v32 = nullptr;
};
// Sorted-gather order for the row-block kernels' products (restrictions,
// whose rows gather the finer vector at scattered columns): within each row
// block of one partition, the entries in ascending column order (ties in
// stored order) with each entry's position in the block's stored order.  A
// wavefront's gathers then touch a few cache lines instead of dozens; the
// products land in LDS at their stored positions and each row adds them in
// stored order as before, so the results are the same bits.  The arrays run
// parallel to the CSR arrays (the same block bounds); a block holding one
// row longer than the cap keeps its stored order.
struct SortedGather {
    int cap = 0;
    const int* blocks = nullptr;  // the partition it was built for
    int* col = nullptr;           // device, nnz
    uint16_t* pos = nullptr;      // device, nnz: position within the block
    void* val = nullptr;          // device, nnz: FP64 (vectors FP64) or FP32 values
};
}  // namespace ramd

struct ramd_csr_s {
    ramd_ctx ctx = nullptr;
    Dia* dia = nullptr;         // stencil-diagonal form (fine-level operators)
    ramd::Sliced* sl = nullptr; // sliced-row form (long-row operators)
    bool sl_on[2] = {false, false};  // per vector precision: use the sliced form
    bool sl16[2] = {false, false};   // ... with its 16-bit column offsets
    bool lr_on[2] = {false, false};  // per vector precision: one lane per CSR row
    bool wr_on[2] = {false, false};  // per vector precision: one wavefront per CSR row
    // per vector precision: the row-block kernels' SpMV in sorted-gather order
    // (used while built for that precision's row blocks)
    ramd::SortedGather* sg[2] = {nullptr, nullptr};
    int32_t n_rows = 0, n_cols = 0;
    int64_t nnz = 0;
    int* rowptr = nullptr;   // device, n_rows+1
    int* col = nullptr;      // device, nnz
    double* val = nullptr;   // device, nnz
    // row blocks of the LDS-staged kernels (device, n_blocks+1) for the
    // default cap of kStreamCap products per workgroup
    int* blocks = nullptr;
    int n_blocks = 0;
    // blocks used per vector precision ([0] FP64, [1] FP32 vectors): the
    // default ones, or those of the cap the autotuner measured fastest
    struct RowBlocks {
        int cap = 0;
        int n = 0;
        int* d = nullptr;  // == blocks for the default cap
    } rb[2];
    // lazily built transpose (gather form of A^T x), same drop-free pattern
    ramd_csr_s* trans = nullptr;
    // FP32 copy of the values (mixed-precision cycle)
    float* val32 = nullptr;
    // host copy of rowptr (kept: used to build blocks and by setup)
    std::vector<int> h_rowptr;
    bool borrowed = false;   // owned by a hierarchy; ramd_csr_destroy refuses it
    // every row stores its diagonal first or has none (move_diag layout,
    // core/matrix.cpp:729-787); lets Jacobi read one column per row
    bool diag_first = false;
    // the flexible PCG's FP64 products with this matrix may read its FP32
    // copy of the values when every value is exact in FP32 (the same
    // products; on by default; ramd_csr_set_fp32_outer)
    bool v32_outer = true;
    // identity of the handle (never reused within a process) and a count of
    // the changes to its device forms and kernel choices after creation (FP32
    // copies, the stencil-diagonal and sliced forms, row blocks, the kernel
    // picked per precision, the FP32-value operator's exactness): captured
    // HIP graphs hold raw pointers to these buffers and the kernels chosen
    // at capture, and are keyed on both
    uint64_t uid = 0;
    uint32_t forms = 0;
};

namespace ramd {
// Creates a handle around device arrays the library already owns.
ramd_csr_s* csr_wrap(ramd_ctx ctx, int n_rows, int n_cols, int64_t nnz, int* rowptr_d, int* col_d,
                     double* val_d, const int* rowptr_host);
void csr_free(ramd_csr_s* A);
// Sliced-row form: pattern (sliced_build), values of one precision copied in
// from A->val / A->val32 (sliced_values), release (sliced_free).
void sliced_build(ramd_csr_s* A);
void sliced_values(ramd_csr_s* A, int prec);
void sliced_free(ramd_csr_s* A);
// Sorted-gather order for vector precision prec over A->rb[prec] (values
// from A->val or A->val32), and its release.
void sg_build(ramd_csr_s* A, int prec);
void sg_free(ramd_csr_s* A, int prec);
// True when every row has its diagonal first or no diagonal at all.
bool host_diag_first(int n_rows, const int* rowptr, const int* col);
// Builds A->dia from host arrays when the matrix qualifies (square, at least
// kDiaMinRows rows, <= kDiaMax distinct offsets, padding <= 15%, every row in
// one common diagonal order).  Returns whether it was built.
constexpr int kDiaMinRows = 4096;
bool dia_build(ramd_csr_s* A);
// Box geometry of the offsets (sorted ascending) for n rows: fills b.L, b.P,
// b.tx and b.args, or leaves b.L = 0 when there is none.
void dia_box_detect(const long long* offs, int nd, int n, DiaBox& b);
// Workgroups of the stencil-diagonal kernels of A for precision pair pp
// (the number of per-workgroup partial sums of a fused product).
int dia_grid(const ramd_csr_s* A, int pp);
// An upper bound of dia_grid over every kernel choice and precision pair.
int dia_grid_bound(const ramd_csr_s* A);
void dia_free(Dia* d);
// Fixed scratch of ramd_fpcg_noalloc (partial sums), per GPU, and the set-up
// of its per-device address (called by ramd_init with the device current).
constexpr int kNoallocPartials = 65536;
constexpr int kMaxDevices = 64;
void noalloc_scratch_setup(int device);
// Row pointers rp[0..n] from row counts cnt[0..n) on the device (exclusive
// prefix sum in int64; rp[n] is the total), one 8-byte read of the total;
// throws RAMD_E_ARG naming `what` when it exceeds 2^31 - 1.
int64_t device_rowptr(ramd_ctx ctx, const int* cnt, int n, int* rp, const char* what);
// Builds (once) A->trans, the CSR of A^T with rows in ascending source-row order.
ramd_csr_s* csr_transpose(ramd_csr_s* A);
void csr_make_fp32(ramd_csr_s* A);
// Row-block caps the LDS-staged kernels can use (products per workgroup).
constexpr int kStreamCaps[4] = {1024, 2048, 4096, 8192};
// Times the CSR kernels of A with every cap for vectors of precision prec
// (0 FP64, 1 FP32; the FP32 values must exist) and keeps the fastest row
// blocks.  The cap changes only which workgroup owns a row, never the order
// of a row's sum, so results are identical for every choice.
void csr_autotune(ramd_csr_s* A, int prec);
// Median time (ms) of `reps` launches of run().  For an operation that
// moves fewer than kFlushBytes (more than the L2 and infinity caches hold)
// each launch follows a pass that writes kFlushBytes, so it finds its
// operands in HBM, as inside a cycle after the other levels' kernels,
// rather than left in the caches by its own previous launch; a larger one
// cannot keep its operands cached and is timed back to back, without the
// flush's write-backs in its time.  The autotuners rank kernel choices by it.
constexpr size_t kFlushBytes = (size_t)320 << 20;
float cold_median_ms(ramd_ctx ctx, const std::function<void()>& run, int reps, size_t op_bytes);
// For a matrix with the stencil-diagonal form: times the gather kernel, the
// LDS x-window kernel (when the window fits kDiaWin) and the plane-march
// kernel (box geometry; several planes per workgroup) for the precision pair
// `pp` (0 FP64 values/vectors, 1 FP32/FP32, 2 FP32 values/FP64 vectors) and
// keeps the fastest (same result bits for every choice).  Timed once per pair.
void dia_autotune(ramd_csr_s* A, int pp);
// New values for the same pattern (device array); refreshes every derived
// copy (FP32 values, stencil-diagonal planes, the cached transpose).
void csr_set_values(ramd_csr_s* A, const double* val_d);

// Kernel launchers (stream-ordered), FP64 and FP32.
// OP_ADD_INTO: y = b + A x with b a separate input (x <- x + P e written to a
// second buffer); same arithmetic as OP_APPEND.
// OP_SPMV_JFZ: y = A x, and y2 = the first Jacobi sweep from zero of the
// system with right-hand side y (jacobi_zero_row with the diagonal dvec):
// the restriction writes the coarse level's first smoothing sweep with it
enum StreamOp { OP_SPMV = 0, OP_APPEND = 1, OP_RESIDUAL = 2, OP_JACOBI = 3, OP_ADD_INTO = 4, OP_SPMV_JFZ = 5 };

// One row of a Jacobi sweep from x = 0 (util/linalg/relax.cpp:23-41 with
// x = 0, the same operations): 0 for a row whose |a_ii| <= 1e-16.
template <typename X>
__device__ __forceinline__ X jacobi_zero_row(X b, X d, X omega)
{
    const X zero = X(0.0);
    X v = zero;
    if (fabs(d) > X(1e-16)) v = ((X(1.0) - omega) * zero) + (omega * ((b - zero) / d));
    return v;
}
// dvec (Jacobi only, may be NULL): the matrix's diagonal (0 where absent);
// on a diagonal-first matrix it replaces re-reading each row's first entry.
void launch_stream_op(const ramd_csr_s* A, int op, const double* x, const double* b, double* y,
                      double omega, hipStream_t s, const double* dvec = nullptr);
// Restriction y = R x fused with the coarse level's first Jacobi sweep from
// zero: x_c = jacobi_zero_row(y, d_c) (OP_SPMV_JFZ), FP64 and FP32.
void launch_restrict_jfz(const ramd_csr_s* R, const double* x, double* y, double* xc, const double* dc,
                         double omega, hipStream_t s);
void launch_restrict_jfz32(const ramd_csr_s* R, const float* x, float* y, float* xc, const float* dc, float omega,
                           hipStream_t s);
void launch_stream_op32(const ramd_csr_s* A, int op, const float* x, const float* b, float* y,
                        float omega, hipStream_t s, const float* dvec = nullptr);
// FP64 arithmetic reading the FP32 copy of the values where every value is
// exactly representable in FP32 (same results as launch_stream_op, half the
// value bytes); falls back to launch_stream_op otherwise.
void launch_stream_op_c(const ramd_csr_s* A, int op, const double* x, const double* b, double* y,
                        double omega, hipStream_t s);

// y = A x (FP64 vectors) with x.y summed per workgroup in the same pass and
// reduced into *out_d (fixed order: deterministic).  values_fp32 selects the
// lossless FP32-value form of launch_stream_op_c.
void launch_spmv_dot(const ramd_csr_s* A, bool values_fp32, const double* x, double* y, double* out_d,
                     hipStream_t s, double* partials = nullptr);
// per-workgroup partials launch_spmv_dot writes for A
int spmv_dot_partials(const ramd_csr_s* A);

// Deterministic device reductions; result left in *out_d (device scalar).
void launch_dot(ramd_ctx ctx, int64_t n, const double* x, const double* y, double* out_d);
void launch_norm2sq(ramd_ctx ctx, int64_t n, const double* x, double* out_d);
// Sums partials[0..np) into *out_d in a fixed order (one workgroup).
void launch_reduce_final(ramd_ctx ctx, const double* partials, int np, double* out_d);
// two launch_reduce_final sums (n1, n2 <= 4096) in one launch
void launch_reduce_final_two(ramd_ctx ctx, const double* p1, int n1, double* o1, const double* p2, int n2,
                             double* o2);
// launch_reduce_final for np <= 4096 that also writes host[0] = *also_d and
// host[1] = the sum (host: the device address of device-writable host
// memory such as ctx->step_pin_d)
void launch_reduce_final_pub(ramd_ctx ctx, const double* partials, int np, double* out_d, const double* also_d,
                             double* host);
// Device scratch of at least `count` doubles for per-workgroup partials.
double* ctx_partials(ramd_ctx ctx, size_t count);
double fetch_scalar(ramd_ctx ctx, const double* d);
// Copies k (<= 16) device scalars to out after the stream drains.
void fetch_scalars(ramd_ctx ctx, const double* d, int k, double* out);
// Fixed grid of the deterministic grid-stride reductions.
constexpr int kRedBlocks = 2048;
}  // namespace ramd

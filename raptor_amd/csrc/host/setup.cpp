// Host AMG setup: layout helpers, the reference's random stream, strength
// of connection, C/F splitting, interpolation and stencil generators.
//
// All routines are the one-rank specialisation of the reference's parallel
// code: every off_proc block, halo exchange and transpose exchange is empty
// at one rank (core/comm_pkg.hpp: num_msgs == 0), so only the on_proc
// arithmetic remains, restated here in the reference's visiting order.
#include "setup.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <numeric>
#include <stdexcept>
#include <thread>

namespace ramd {
namespace host {

// ---------------------------------------------------------------------------
// row-parallel drivers: rows are independent in every routine below except
// the splittings, so contiguous row ranges go to threads with private
// scratch, and the pieces are concatenated in row order (same result as one
// sequential pass).  RAMD_HOST_THREADS caps the thread count (default: the
// hardware threads, at most 16).
// ---------------------------------------------------------------------------
static int host_threads(int64_t work)
{
    if (work < (1 << 16)) return 1;
    int t = (int)std::thread::hardware_concurrency();
    if (const char* e = std::getenv("RAMD_HOST_THREADS")) t = std::atoi(e);
    return std::max(1, std::min(t, 16));
}

// runs fn(t, r0, r1) on T contiguous ranges of [0, n)
template <class Fn>
static void parallel_ranges(int n, int T, Fn fn)
{
    if (T <= 1) {
        fn(0, 0, n);
        return;
    }
    std::vector<std::thread> th;
    std::vector<std::exception_ptr> err((size_t)T);
    for (int t = 0; t < T; t++) {
        const int r0 = (int)((int64_t)n * t / T), r1 = (int)((int64_t)n * (t + 1) / T);
        th.emplace_back([&, t, r0, r1] {
            try {
                fn(t, r0, r1);
            } catch (...) {
                err[t] = std::current_exception();
            }
        });
    }
    for (auto& x : th) x.join();
    for (auto& e : err)
        if (e) std::rethrow_exception(e);
}

struct RowChunk {
    std::vector<int> len, col;
    std::vector<double> val;
};

// Concatenates per-range row pieces into out (rowptr/col/val).
static void assemble(int n, std::vector<RowChunk>& ch, Csr& out)
{
    const int T = (int)ch.size();
    out.rowptr.assign((size_t)n + 1, 0);
    std::vector<int64_t> base((size_t)T + 1, 0);
    int r = 0;
    for (int t = 0; t < T; t++) {
        for (int l : ch[t].len) {
            out.rowptr[r + 1] = out.rowptr[r] + l;
            r++;
        }
        base[t + 1] = base[t] + (int64_t)ch[t].col.size();
    }
    if (base[T] > INT32_MAX) throw std::length_error("matrix exceeds 2^31 stored entries");
    out.col.resize((size_t)base[T]);
    out.val.resize((size_t)base[T]);
    parallel_ranges(T, T, [&](int, int t0, int t1) {
        for (int t = t0; t < t1; t++) {
            if (ch[t].col.empty()) continue;
            std::memcpy(out.col.data() + base[t], ch[t].col.data(), sizeof(int) * ch[t].col.size());
            std::memcpy(out.val.data() + base[t], ch[t].val.data(), sizeof(double) * ch[t].val.size());
            std::vector<int>().swap(ch[t].col);
            std::vector<double>().swap(ch[t].val);
        }
    });
}

// ---------------------------------------------------------------------------
// layout
// ---------------------------------------------------------------------------
void sort_rows(Csr& A)
{
    parallel_ranges(A.n_rows, host_threads(A.nnz()), [&](int, int r0, int r1) {
    std::vector<int> perm;
    std::vector<int> c2;
    std::vector<double> v2;
    const bool has_val = !A.val.empty();
    for (int i = r0; i < r1; i++) {
        const int s = A.rowptr[i], e = A.rowptr[i + 1];
        bool sorted = true;
        for (int j = s + 1; j < e && sorted; j++) sorted = A.col[j - 1] <= A.col[j];
        if (sorted) continue;
        perm.resize(e - s);
        std::iota(perm.begin(), perm.end(), 0);
        std::stable_sort(perm.begin(), perm.end(), [&](int a, int b) { return A.col[s + a] < A.col[s + b]; });
        c2.resize(e - s);
        v2.resize(e - s);
        for (int k = 0; k < e - s; k++) {
            c2[k] = A.col[s + perm[k]];
            if (has_val) v2[k] = A.val[s + perm[k]];
        }
        std::copy(c2.begin(), c2.end(), A.col.begin() + s);
        if (has_val) std::copy(v2.begin(), v2.end(), A.val.begin() + s);
    }
    });
}

void move_diag(Csr& A)
{
    const bool has_val = !A.val.empty();
    parallel_ranges(A.n_rows, host_threads(A.nnz()), [&](int, int r0, int r1) {
    for (int i = r0; i < r1; i++) {
        const int s = A.rowptr[i], e = A.rowptr[i + 1];
        for (int j = s; j < e; j++) {
            if (A.col[j] == i) {
                const double t = has_val ? A.val[j] : 0.0;
                for (int k = j; k > s; k--) {
                    A.col[k] = A.col[k - 1];
                    if (has_val) A.val[k] = A.val[k - 1];
                }
                A.col[s] = i;
                if (has_val) A.val[s] = t;
                break;
            }
        }
    }
    });
}

// ---------------------------------------------------------------------------
// glibc rand(): srandom_r fills 31 words with 16807*x mod (2^31-1) (Schrage's
// method), the front pointer starts 3 words ahead, 310 outputs are discarded,
// and each output is (state[f] += state[r]) >> 1 over uint32 arithmetic.
// ---------------------------------------------------------------------------
GlibcRand::GlibcRand(unsigned seed)
{
    if (seed == 0) seed = 1;
    r_[0] = (int32_t)seed;
    int32_t word = (int32_t)seed;
    for (int i = 1; i < 31; i++) {
        long hi = word / 127773, lo = word % 127773;
        word = (int32_t)(16807 * lo - 2836 * hi);
        if (word < 0) word += 2147483647;
        r_[i] = word;
    }
    f_ = 3;
    b_ = 0;
    for (int i = 0; i < 310; i++) next();
}

int GlibcRand::next()
{
    uint32_t v = (uint32_t)r_[f_] + (uint32_t)r_[b_];
    r_[f_] = (int32_t)v;
    f_ = (f_ + 1) % 31;
    b_ = (b_ + 1) % 31;
    return (int)(v >> 1);
}

std::vector<double> reference_weights(int n, int first_row)
{
    std::vector<double> w((size_t)n);
    GlibcRand g((unsigned)(2448422 + first_row));
    for (int i = 0; i < n; i++) w[i] = double(g.next()) / GlibcRand::kRandMax;
    return w;
}

// ---------------------------------------------------------------------------
// strength, par_strength.cpp
// ---------------------------------------------------------------------------
static void sort_and_diag_first(Csr& A)
{
    sort_rows(A);
    move_diag(A);
}

// classical_strength par_strength.cpp:8-340 (num_variables == 1)
Csr classical_strength(Csr& A, double theta)
{
    sort_and_diag_first(A);
    Csr S;
    S.n_rows = A.n_rows;
    S.n_cols = A.n_cols;
    const int T = host_threads(A.nnz());
    std::vector<RowChunk> ch((size_t)T);
    parallel_ranges(A.n_rows, T, [&](int t, int r0, int r1) {
        RowChunk& c = ch[t];
        c.len.resize((size_t)(r1 - r0));
        c.col.reserve((size_t)(A.rowptr[r1] - A.rowptr[r0]));
        c.val.reserve((size_t)(A.rowptr[r1] - A.rowptr[r0]));
        for (int i = r0; i < r1; i++) {
            const size_t before = c.col.size();
            int s = A.rowptr[i];
            const int e = A.rowptr[i + 1];
            if (e - s) {
                double diag;
                if (A.col[s] == i) {
                    diag = A.val[s];
                    s++;
                } else {
                    diag = 0.0;
                }
                double row_scale;
                if (diag < 0.0) {
                    row_scale = -RAND_MAX;
                    for (int j = s; j < e; j++)
                        if (A.val[j] > row_scale) row_scale = A.val[j];
                } else {
                    row_scale = RAND_MAX;
                    for (int j = s; j < e; j++)
                        if (A.val[j] < row_scale) row_scale = A.val[j];
                }
                const double threshold = row_scale * theta;
                c.col.push_back(i);
                c.val.push_back(diag);
                if (diag < 0) {
                    for (int j = s; j < e; j++)
                        if (A.val[j] > threshold) {
                            c.col.push_back(A.col[j]);
                            c.val.push_back(A.val[j]);
                        }
                } else {
                    for (int j = s; j < e; j++)
                        if (A.val[j] < threshold) {
                            c.col.push_back(A.col[j]);
                            c.val.push_back(A.val[j]);
                        }
                }
            }
            c.len[i - r0] = (int)(c.col.size() - before);
        }
    });
    assemble(A.n_rows, ch, S);
    return S;
}

// symmetric_strength par_strength.cpp:342-534
Csr symmetric_strength(Csr& A, double theta)
{
    sort_and_diag_first(A);
    const int n = A.n_rows;
    std::vector<int> neg_diags((size_t)n, 0);
    std::vector<double> row_scales((size_t)n, 0.0);
    for (int i = 0; i < n; i++) {
        int s = A.rowptr[i];
        const int e = A.rowptr[i + 1];
        if (!(e - s)) continue;
        double diag;
        if (A.col[s] == i) {
            diag = A.val[s];
            s++;
        } else {
            diag = 0.0;
        }
        double row_scale;
        if (diag < 0.0) {
            neg_diags[i] = 1;
            row_scale = -RAND_MAX;
            for (int j = s; j < e; j++)
                if (A.val[j] > row_scale) row_scale = A.val[j];
        } else {
            neg_diags[i] = 0;
            row_scale = RAND_MAX;
            for (int j = s; j < e; j++)
                if (A.val[j] < row_scale) row_scale = A.val[j];
        }
        row_scales[i] = row_scale * theta;
    }
    Csr S;
    S.n_rows = n;
    S.n_cols = A.n_cols;
    S.rowptr.assign((size_t)n + 1, 0);
    S.col.reserve(A.col.size());
    S.val.reserve(A.col.size());
    for (int i = 0; i < n; i++) {
        int s = A.rowptr[i];
        const int e = A.rowptr[i + 1];
        if (e - s) {
            const bool neg_diag = neg_diags[i];
            const double threshold = row_scales[i];
            // the first stored entry is taken as the diagonal (:466-468)
            S.col.push_back(i);
            S.val.push_back(A.val[s++]);
            for (int j = s; j < e; j++) {
                const double v = A.val[j];
                const int c = A.col[j];
                if ((neg_diag && v > threshold) || (!neg_diag && v < threshold) ||
                    (neg_diags[c] && v > row_scales[c]) || (!neg_diags[c] && v < row_scales[c])) {
                    S.col.push_back(c);
                    S.val.push_back(v);
                }
            }
        }
        S.rowptr[i + 1] = (int)S.col.size();
    }
    return S;
}

// ---------------------------------------------------------------------------
// C/F splitting, ruge_stuben/par_cf_splitting.cpp
// ---------------------------------------------------------------------------

// set_initial_states :114-131
static void set_initial_states(const Csr& S, std::vector<int>& states)
{
    states.assign((size_t)S.n_rows, kUnassigned);
    for (int i = 0; i < S.n_rows; i++)
        states[i] = (S.rowptr[i + 1] - S.rowptr[i] > 1) ? kUnassigned : kNoNeighbors;
}

// transpose :155-228 -- CSC of S without each row's leading diagonal, rows
// ascending within a column.
static void transpose_s(const Csr& S, std::vector<int>& col_ptr, std::vector<int>& col_idx)
{
    const int n = S.n_cols;
    col_ptr.assign((size_t)n + 1, 0);
    std::vector<int> sizes((size_t)n, 0);
    for (int i = 0; i < S.n_rows; i++) {
        int s = S.rowptr[i];
        const int e = S.rowptr[i + 1];
        if (s < e && S.col[s] == i) s++;
        for (int j = s; j < e; j++) sizes[S.col[j]]++;
    }
    for (int i = 0; i < n; i++) {
        col_ptr[i + 1] = col_ptr[i] + sizes[i];
        sizes[i] = 0;
    }
    col_idx.assign((size_t)col_ptr[n], 0);
    for (int i = 0; i < S.n_rows; i++) {
        int s = S.rowptr[i];
        const int e = S.rowptr[i + 1];
        if (s < e && S.col[s] == i) s++;
        for (int j = s; j < e; j++) {
            const int c = S.col[j];
            col_idx[col_ptr[c] + sizes[c]++] = i;
        }
    }
}

// initial_weights :232-289 with rand_vals given: w_i = rand_i + #{rows that
// strongly depend on i}.
static void initial_weights(const Csr& S, const double* rand_vals, std::vector<double>& w)
{
    for (int i = 0; i < S.n_cols; i++) w[i] = rand_vals[i];
    for (int i = 0; i < S.n_rows; i++) {
        int s = S.rowptr[i];
        const int e = S.rowptr[i + 1];
        if (s < e && S.col[s] == i) s++;
        for (int j = s; j < e; j++) w[S.col[j]] += 1;
    }
}

// select_independent_set :403-495 (one rank: max_off_weights are all 0)
static int select_independent_set(const Csr& S, int remaining, const std::vector<int>& unassigned,
                                  const std::vector<double>& w, const std::vector<int>& col_ptr,
                                  const std::vector<int>& col_idx, std::vector<int>& states,
                                  std::vector<int>& new_coarse)
{
    int num = 0;
    for (int i = 0; i < remaining; i++) {
        const int u = unassigned[i];
        const double weight = w[u];
        if (0.0 > weight) continue;
        int s = S.rowptr[u];
        const int e = S.rowptr[u + 1];
        if (s < e && S.col[s] == u) s++;
        int j;
        for (j = s; j < e; j++)
            if (w[S.col[j]] > weight) break;
        if (j != e) continue;
        for (j = col_ptr[u]; j < col_ptr[u + 1]; j++)
            if (w[col_idx[j]] > weight) break;
        if (j != col_ptr[u + 1]) continue;
        states[u] = kNewSelection;
        new_coarse[num++] = u;
    }
    return num;
}

// update_states :1200-1227
static int update_states(std::vector<double>& w, std::vector<int>& states, int remaining,
                         std::vector<int>& unassigned)
{
    int ctr = 0;
    for (int i = 0; i < remaining; i++) {
        const int u = unassigned[i];
        if (states[u] == kNewSelection) {
            w[u] = 0.0;
            states[u] = kSelected;
        } else if (w[u] < 1.0 || states[u] == kNewUnselection) {
            w[u] = 0.0;
            states[u] = kUnselected;
        } else {
            unassigned[ctr++] = unassigned[i];
        }
    }
    return ctr;
}

// update_row_weights :497-536 (on_proc part)
static void update_row_weights(const Csr& S, int num_new, const std::vector<int>& new_coarse,
                               std::vector<char>& edgemark, const std::vector<int>& states,
                               std::vector<double>& w)
{
    for (int i = 0; i < num_new; i++) {
        const int c = new_coarse[i];
        int s = S.rowptr[c];
        const int e = S.rowptr[c + 1];
        if (s < e && S.col[s] == c) s++;
        for (int j = s; j < e; j++) {
            const int idx = S.col[j];
            if (states[idx] == kUnassigned && edgemark[j]) {
                edgemark[j] = 0;
                w[idx]--;
            }
        }
    }
}

// update_local_dist2_weights :538-640 (on_proc new coarse points)
static void update_local_dist2_weights(const Csr& S, int num_new, const std::vector<int>& new_coarse,
                                       const std::vector<int>& col_ptr, const std::vector<int>& col_idx,
                                       std::vector<char>& edgemark, const std::vector<int>& states,
                                       std::vector<double>& w, std::vector<int>& c_dep_cache)
{
    for (int i = 0; i < num_new; i++) {
        const int c = new_coarse[i];
        const int s = col_ptr[c], e = col_ptr[c + 1];
        for (int j = s; j < e; j++) {
            const int idx = col_idx[j];
            if (states[idx] == kUnassigned) c_dep_cache[idx] = c;
        }
        for (int j = s; j < e; j++) {
            const int idx = col_idx[j];
            if (states[idx] == kSelected) continue;
            int is = S.rowptr[idx];
            const int ie = S.rowptr[idx + 1];
            if (is < ie && S.col[is] == idx) is++;
            for (int k = is; k < ie; k++) {
                const int ik = S.col[k];
                if (states[ik] == kUnassigned && edgemark[k] && c_dep_cache[ik] == c) {
                    edgemark[k] = 0;
                    w[ik]--;
                }
            }
        }
    }
}

// cljp_main_loop :1383-1607
static void cljp_main_loop(Csr& S, std::vector<int>& states, const double* rand_vals)
{
    const int n = S.n_rows;
    std::vector<double> w((size_t)n, 0.0);
    std::vector<int> new_coarse((size_t)n), unassigned((size_t)n), c_dep((size_t)n, kUnassigned);
    std::vector<char> edgemark(S.col.size(), 1);
    std::vector<int> col_ptr, col_idx;
    initial_weights(S, rand_vals, w);
    transpose_s(S, col_ptr, col_idx);
    int remaining = 0, num_new = 0;
    for (int i = 0; i < S.n_cols; i++) {
        if (states[i] == kUnassigned) {
            unassigned[remaining++] = i;
        } else {
            w[i] = 0.0;
            if (states[i] == kSelected) new_coarse[num_new++] = i;
        }
    }
    if (num_new) {
        update_row_weights(S, num_new, new_coarse, edgemark, states, w);
        update_local_dist2_weights(S, num_new, new_coarse, col_ptr, col_idx, edgemark, states, w, c_dep);
    }
    bool first_pass = true;
    while (remaining || first_pass) {
        num_new = select_independent_set(S, remaining, unassigned, w, col_ptr, col_idx, states, new_coarse);
        update_row_weights(S, num_new, new_coarse, edgemark, states, w);
        update_local_dist2_weights(S, num_new, new_coarse, col_ptr, col_idx, edgemark, states, w, c_dep);
        remaining = update_states(w, states, remaining, unassigned);
        first_pass = false;
    }
}

// pmis_main_loop :1229-1381
static void pmis_main_loop(Csr& S, std::vector<int>& states, const double* rand_vals)
{
    const int n = S.n_rows;
    std::vector<double> w((size_t)n, 0.0);
    std::vector<int> new_coarse((size_t)n), unassigned((size_t)n);
    std::vector<int> col_ptr, col_idx;
    transpose_s(S, col_ptr, col_idx);
    initial_weights(S, rand_vals, w);
    for (int i = 0; i < S.n_cols; i++) {
        if (states[i] == kSelected) {
            for (int j = col_ptr[i]; j < col_ptr[i + 1]; j++) {
                const int row = col_idx[j];
                if (states[row] == kUnassigned) states[row] = kUnselected;
            }
        }
    }
    int remaining = 0;
    for (int i = 0; i < n; i++) {
        if (states[i] == kUnassigned && w[i] < 1) {
            states[i] = kUnselected;
            w[i] = 0.0;
        } else if (states[i] == kUnassigned) {
            unassigned[remaining++] = i;
        } else {
            w[i] = 0.0;
        }
    }
    bool first_pass = true;
    while (remaining || first_pass) {
        const int num_new =
            select_independent_set(S, remaining, unassigned, w, col_ptr, col_idx, states, new_coarse);
        for (int i = 0; i < num_new; i++) {
            const int idx = new_coarse[i];
            for (int j = col_ptr[idx]; j < col_ptr[idx + 1]; j++) {
                const int row = col_idx[j];
                if (states[row] == kUnassigned) states[row] = kNewUnselection;
            }
        }
        remaining = update_states(w, states, remaining, unassigned);
        first_pass = false;
    }
}

// ---------------------------------------------------------------------------
// serial Ruge-Stuben splitting, ruge_stuben/cf_splitting.cpp
// ---------------------------------------------------------------------------

// rs_first_pass cf_splitting.cpp:69-209: columns kept in buckets by weight
// (number of rows that strongly depend on them); positions are visited from
// the heaviest end, and each new C point moves the affected columns between
// adjacent buckets by swapping them to the bucket boundary.
static void rs_first_pass(const Csr& S, const std::vector<int>& col_ptr, const std::vector<int>& col_idx,
                          std::vector<int>& w, std::vector<int>& states)
{
    const int nc = S.n_cols;
    std::vector<int> wptr((size_t)nc + 1, 0), wsize((size_t)nc, 0), pos_col((size_t)nc), col_pos((size_t)nc);
    for (int i = 0; i < nc; i++) wsize[w[i]]++;
    for (int i = 0; i < nc; i++) {
        wptr[i + 1] = wptr[i] + wsize[i];
        wsize[i] = 0;
    }
    for (int i = 0; i < nc; i++) {
        const int idx = wptr[w[i]] + wsize[w[i]]++;
        pos_col[idx] = i;
        col_pos[i] = idx;
    }
    auto swap_pos = [&](int a, int b) {
        col_pos[pos_col[a]] = b;
        col_pos[pos_col[b]] = a;
        std::swap(pos_col[a], pos_col[b]);
    };
    for (int i = nc - 1; i >= 0; i--) {
        const int c = pos_col[i];
        wsize[w[c]]--;
        if (states[c] != kUnassigned) continue;
        states[c] = kSelected;
        for (int j = col_ptr[c]; j < col_ptr[c + 1]; j++) {
            const int idx = col_idx[j];
            if (states[idx] != kUnassigned) continue;
            states[idx] = kUnselected;
            int ks = S.rowptr[idx];
            const int ke = S.rowptr[idx + 1];
            if (ks < ke && S.col[ks] == idx) ks++;
            for (int k = ks; k < ke; k++) {
                const int ik = S.col[k];
                if (states[ik] != kUnassigned) continue;
                const int wk = w[ik];
                if (wk >= nc - 1) continue;
                const int old_pos = col_pos[ik];
                const int new_pos = wptr[wk] + wsize[wk] - 1;
                swap_pos(old_pos, new_pos);
                wsize[wk] -= 1;
                wsize[wk + 1] += 1;
                wptr[wk + 1] = new_pos;
                w[ik]++;
            }
        }
        int s = S.rowptr[c];
        const int e = S.rowptr[c + 1];
        if (s < e && S.col[s] == c) s++;
        for (int j = s; j < e; j++) {
            const int idx = S.col[j];
            if (states[idx] != kUnassigned) continue;
            const int wi = w[idx];
            if (wi == 0) continue;
            const int old_pos = col_pos[idx];
            const int new_pos = wptr[wi];
            swap_pos(old_pos, new_pos);
            wsize[wi] -= 1;
            wsize[wi - 1] += 1;
            wptr[wi] += 1;
            wptr[wi - 1] = wptr[wi] - wsize[wi - 1];
            w[idx]--;
        }
    }
}

// rs_second_pass cf_splitting.cpp:211-261: every F point whose strong F
// neighbours share no C point with it promotes them (rows include their
// diagonal here, so an F point without strong C neighbours promotes itself).
static void rs_second_pass(const Csr& S, std::vector<int>& states)
{
    std::vector<int> row_coarse((size_t)S.n_rows, -1);
    for (int i = 0; i < S.n_rows; i++) {
        if (states[i] == kSelected) continue;
        const int s = S.rowptr[i], e = S.rowptr[i + 1];
        for (int j = s; j < e; j++)
            if (states[S.col[j]] == kSelected) row_coarse[S.col[j]] = i;
        for (int j = s; j < e; j++) {
            const int c = S.col[j];
            if (states[c] != kUnselected) continue;
            const int ks = S.rowptr[c], ke = S.rowptr[c + 1];
            if (ks == ke) continue;
            bool connection = false;
            for (int k = ks; k < ke; k++)
                if (row_coarse[S.col[k]] == i) {
                    connection = true;
                    break;
                }
            if (!connection) {
                row_coarse[c] = i;
                states[c] = kSelected;
            }
        }
    }
}

// split_rs(CSRMatrix*) cf_splitting.cpp:277-322
void split_rs_serial(Csr& S, std::vector<int>& states, bool prev_states, bool second_pass)
{
    move_diag(S);
    if (!prev_states) states.assign((size_t)S.n_rows, kUnassigned);
    std::vector<int> col_ptr, col_idx;
    transpose_s(S, col_ptr, col_idx);
    std::vector<int> w((size_t)S.n_rows, 0);
    for (int i = 0; i < S.n_cols; i++) w[i] = col_ptr[i + 1] - col_ptr[i];
    rs_first_pass(S, col_ptr, col_idx, w, states);
    if (second_pass) rs_second_pass(S, states);
}

// split_rs(ParCSRMatrix*) par_cf_splitting.cpp:7-23
void split_rs(Csr& S, std::vector<int>& states)
{
    set_initial_states(S, states);
    split_rs_serial(S, states, true, true);
}

// split_falgout par_cf_splitting.cpp:50-73 (reset_boundaries marks nothing
// at one rank: there are no off_proc columns or send indices)
void split_falgout(Csr& S, const double* weights, std::vector<int>& states)
{
    move_diag(S);
    set_initial_states(S, states);
    split_rs_serial(S, states, true, true);
    cljp_main_loop(S, states, weights);
}

// split_hmis par_cf_splitting.cpp:91-112
void split_hmis(Csr& S, const double* weights, std::vector<int>& states)
{
    move_diag(S);
    set_initial_states(S, states);
    split_rs_serial(S, states, true, false);
    pmis_main_loop(S, states, weights);
}

// the Ruge-Stuben part of split_hmis (states before its PMIS loop)
void hmis_first_pass(Csr& S, std::vector<int>& states)
{
    move_diag(S);
    set_initial_states(S, states);
    split_rs_serial(S, states, true, false);
}

// the Ruge-Stuben part of split_falgout (states before its CLJP loop)
void falgout_first_pass(Csr& S, std::vector<int>& states)
{
    move_diag(S);
    set_initial_states(S, states);
    split_rs_serial(S, states, true, true);
}

// split_cljp :32-48
void split_cljp(Csr& S, const double* weights, std::vector<int>& states)
{
    move_diag(S);
    set_initial_states(S, states);
    cljp_main_loop(S, states, weights);
}

// split_pmis :75-89
void split_pmis(Csr& S, const double* weights, std::vector<int>& states)
{
    move_diag(S);
    set_initial_states(S, states);
    pmis_main_loop(S, states, weights);
}

// ---------------------------------------------------------------------------
// interpolation, ruge_stuben/par_interpolation.cpp
// ---------------------------------------------------------------------------
static std::vector<int> coarse_index(const std::vector<int>& states, bool any_nonzero, int& n_coarse)
{
    std::vector<int> cmap(states.size(), -1);
    n_coarse = 0;
    for (size_t i = 0; i < states.size(); i++)
        if (any_nonzero ? (states[i] != 0) : (states[i] == kSelected)) cmap[i] = n_coarse++;
    return cmap;
}

// direct_interpolation :1453-1754.  Every state != 0 (C points and points
// without neighbours) gets a column (:1527-1533), as in the reference.
Csr direct_interpolation(Csr& A, Csr& S, const std::vector<int>& states)
{
    sort_and_diag_first(A);
    sort_and_diag_first(S);
    const int n = A.n_rows;
    // values of A on S's pattern (:1470-1488)
    std::vector<double> sa(S.col.size());
    for (int i = 0; i < n; i++) {
        int ctr = A.rowptr[i];
        for (int j = S.rowptr[i]; j < S.rowptr[i + 1]; j++) {
            const int c = S.col[j];
            while (A.col[ctr] != c) ctr++;
            sa[j] = A.val[ctr];
        }
    }
    int nc;
    std::vector<int> cmap = coarse_index(states, true, nc);
    Csr P;
    P.n_rows = n;
    P.n_cols = nc;
    P.rowptr.assign((size_t)n + 1, 0);
    for (int i = 0; i < n; i++) {
        if (states[i] == 1) {
            P.col.push_back(cmap[i]);
            P.val.push_back(1);
        } else {
            double sum_strong_pos = 0, sum_strong_neg = 0, sum_all_pos = 0, sum_all_neg = 0;
            int s = S.rowptr[i];
            const int e = S.rowptr[i + 1];
            if (S.col[s] == i) s++;
            for (int j = s; j < e; j++) {
                if (states[S.col[j]] == 1) {
                    const double v = sa[j];
                    if (v < 0) sum_strong_neg += v;
                    else sum_strong_pos += v;
                }
            }
            int as = A.rowptr[i];
            const int ae = A.rowptr[i + 1];
            double diag = A.val[as];
            as++;
            for (int j = as; j < ae; j++) {
                const double v = A.val[j];
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
            for (int j = s; j < e; j++) {
                const int c = S.col[j];
                if (states[c] == 1) {
                    const double v = sa[j];
                    P.col.push_back(cmap[c]);
                    P.val.push_back(v < 0 ? neg_coeff * v : pos_coeff * v);
                }
            }
        }
        P.rowptr[i + 1] = (int)P.col.size();
    }
    return P;
}

// mod_classical_interpolation :991-1451 (num_variables == 1)
Csr mod_classical_interpolation(Csr& A, Csr& S, const std::vector<int>& states)
{
    sort_and_diag_first(A);
    sort_and_diag_first(S);
    const int n = A.n_rows;
    int nc;
    std::vector<int> cmap = coarse_index(states, false, nc);
    Csr P;
    P.n_rows = n;
    P.n_cols = nc;
    const int T = host_threads(A.nnz());
    std::vector<RowChunk> ch((size_t)T);
    parallel_ranges(n, T, [&](int t, int r0, int r1) {
    RowChunk& Pc = ch[t];
    Pc.len.resize((size_t)(r1 - r0));
    std::vector<int> pos((size_t)A.n_cols, -1);
    for (int i = r0; i < r1; i++) {
        const int row_first = (int)Pc.col.size();
        if (states[i] == kSelected) {
            Pc.col.push_back(cmap[i]);
            Pc.val.push_back(1);
            Pc.len[i - r0] = 1;
            continue;
        }
        const int row_start = (int)Pc.col.size();
        for (int j = S.rowptr[i] + 1; j < S.rowptr[i + 1]; j++) {
            const int c = S.col[j];
            if (states[c] == kSelected) {
                pos[c] = (int)Pc.col.size();
                Pc.col.push_back(cmap[c]);
                Pc.val.push_back(S.val[j]);
            }
        }
        int start = A.rowptr[i];
        const int end = A.rowptr[i + 1];
        int ctr = S.rowptr[i] + 1;
        const int end_S = S.rowptr[i + 1];
        double weak_sum = A.val[start++];
        for (int j = start; j < end; j++) {
            const int c = A.col[j];
            const double v = A.val[j];
            if (ctr < end_S && S.col[ctr] == c) {
                ctr++;
                if (states[c] == kSelected) continue;
                double coarse_sum = 0;
                int sk = A.rowptr[c];
                const int ek = A.rowptr[c + 1];
                const double diag = A.val[sk++];
                const int sign = diag > 0 ? 1 : -1;
                for (int k = sk; k < ek; k++) {
                    const int ck = A.col[k];
                    if (states[ck] == kSelected) {
                        const double vk = A.val[k];
                        if (vk * sign < 0 && pos[ck] >= row_start) coarse_sum += vk;
                    }
                }
                if (fabs(coarse_sum) < kZeroTol) weak_sum += v;
                else coarse_sum = v / coarse_sum;
                if (coarse_sum) {
                    for (int k = A.rowptr[c] + 1; k < ek; k++) {
                        const int ck = A.col[k];
                        if (states[ck] == kSelected) {
                            const double vk = A.val[k];
                            const int idx = pos[ck];
                            if (vk * sign < 0 && idx >= row_start) Pc.val[idx] += (coarse_sum * vk);
                        }
                    }
                }
            } else if (states[c] != kNoNeighbors) {
                weak_sum += v;
            }
        }
        const int row_end = (int)Pc.col.size();
        for (int j = row_first; j < row_end; j++) Pc.val[j] /= -weak_sum;
        Pc.len[i - r0] = row_end - row_first;
    }
    });
    assemble(n, ch, P);
    return P;
}

// filter_interp :175-278 over rows given by their lengths, compacting in
// place: entries below filter * max|p_ij| are dropped and the survivors are
// rescaled to keep the row sum.
static void filter_rows(std::vector<int>& col, std::vector<double>& val, std::vector<int>& len,
                        const double filter_threshold)
{
    if (!(filter_threshold > kZeroTol && filter_threshold <= 1)) return;
    int row_start = 0, ctr = 0;
    for (size_t i = 0; i < len.size(); i++) {
        const int prev_ctr = ctr;
        const int row_end = row_start + len[i];
        double row_max = 0;
        for (int j = row_start; j < row_end; j++) {
            const double a = fabs(val[j]);
            if (a > row_max) row_max = a;
        }
        row_max *= filter_threshold;
        double row_sum = 0, remain_sum = 0;
        for (int j = row_start; j < row_end; j++) {
            const double v = val[j];
            row_sum += v;
            if (fabs(v) >= row_max) {
                col[ctr] = col[j];
                val[ctr] = v;
                ctr++;
                remain_sum += v;
            }
        }
        if (fabs(remain_sum) > kZeroTol && fabs(row_sum - remain_sum) > kZeroTol) {
            const double row_scale = row_sum / remain_sum;
            for (int j = prev_ctr; j < ctr; j++) val[j] *= row_scale;
        }
        len[i] = ctr - prev_ctr;
        row_start = row_end;
    }
    col.resize(ctr);
    val.resize(ctr);
}

void filter_interp(Csr& P, const double filter_threshold)
{
    std::vector<int> len((size_t)P.n_rows);
    for (int i = 0; i < P.n_rows; i++) len[i] = P.rowptr[i + 1] - P.rowptr[i];
    filter_rows(P.col, P.val, len, filter_threshold);
    for (int i = 0; i < P.n_rows; i++) P.rowptr[i + 1] = P.rowptr[i] + len[i];
}

// extended_interpolation :280-989 (num_variables == 1).  Row entries are
// kept in discovery order, which later sets the order of every sum over P.
Csr extended_interpolation(Csr& A, Csr& S, const std::vector<int>& states, double filter)
{
    sort_and_diag_first(A);
    sort_and_diag_first(S);
    const int n = A.n_rows;
    int nc;
    std::vector<int> cmap = coarse_index(states, false, nc);
    Csr P;
    P.n_rows = n;
    P.n_cols = nc;
    const int T = host_threads(A.nnz());
    std::vector<RowChunk> ch((size_t)T);
    parallel_ranges(n, T, [&](int t, int r0, int r1) {
    RowChunk& Pc = ch[t];
    Pc.len.resize((size_t)(r1 - r0));
    std::vector<int> pos((size_t)A.n_cols, -1);
    // upper bound of this range's entries (:531-566)
    int64_t bound = 0;
    for (int i = r0; i < r1; i++) {
        if (states[i] == kUnselected) {
            bound += S.rowptr[i + 1] - S.rowptr[i];
            for (int j = S.rowptr[i] + 1; j < S.rowptr[i + 1]; j++) {
                const int c = S.col[j];
                if (states[c] == kUnselected) bound += S.rowptr[c + 1] - S.rowptr[c];
            }
        } else {
            bound++;
        }
    }
    Pc.col.reserve((size_t)bound);
    Pc.val.reserve((size_t)bound);
    for (int i = r0; i < r1; i++) {
        if (states[i] != kUnselected) {
            if (states[i] == kSelected) {
                Pc.col.push_back(cmap[i]);
                Pc.val.push_back(1);
                Pc.len[i - r0] = 1;
            } else {
                Pc.len[i - r0] = 0;
            }
            continue;
        }
        const int row_start = (int)Pc.col.size();
        for (int j = S.rowptr[i] + 1; j < S.rowptr[i + 1]; j++) {
            const int c = S.col[j];
            if (states[c] == kSelected) {
                if (pos[c] < row_start) {
                    pos[c] = (int)Pc.col.size();
                    Pc.col.push_back(cmap[c]);
                    Pc.val.push_back(0.0);
                }
            } else if (states[c] == kUnselected) {
                for (int k = S.rowptr[c] + 1; k < S.rowptr[c + 1]; k++) {
                    const int ck = S.col[k];
                    if (states[ck] == kSelected && pos[ck] < row_start) {
                        pos[ck] = (int)Pc.col.size();
                        Pc.col.push_back(cmap[ck]);
                        Pc.val.push_back(0.0);
                    }
                }
            }
        }
        pos[i] = (int)Pc.col.size();
        const int row_end = (int)Pc.col.size();
        int start = A.rowptr[i];
        const int end = A.rowptr[i + 1];
        double weak_sum = A.val[start++];
        int ctr = S.rowptr[i] + 1;
        const int end_S = S.rowptr[i + 1];
        for (int j = start; j < end; j++) {
            const int c = A.col[j];
            const double v = A.val[j];
            const int idx = pos[c];
            if (idx >= row_start) {
                Pc.val[idx] += v;
                if (ctr < end_S && S.col[ctr] == c) ctr++;
            } else if (ctr < end_S && S.col[ctr] == c) {
                ctr++;
                if (states[c] != kUnselected) continue;
                double col_sum = 0;
                int sk = A.rowptr[c];
                const int ek = A.rowptr[c + 1];
                const double diag = A.val[sk++];
                const int sign = diag > 0 ? 1 : -1;
                for (int k = sk; k < ek; k++) {
                    const double vk = A.val[k];
                    if (vk * sign < 0 && pos[A.col[k]] >= row_start) col_sum += vk;
                }
                if (fabs(col_sum) < kZeroTol) {
                    weak_sum += v;
                } else {
                    col_sum = v / col_sum;
                    for (int k = A.rowptr[c] + 1; k < ek; k++) {
                        const int ck = A.col[k];
                        const double vk = A.val[k];
                        const int pk = pos[ck];
                        if (vk * sign < 0 && pk >= row_start) {
                            if (ck == i) weak_sum += (col_sum * vk);
                            else Pc.val[pk] += (col_sum * vk);
                        }
                    }
                }
            } else {
                if (states[c] != kNoNeighbors) weak_sum += v;
            }
        }
        if (fabs(weak_sum) > kZeroTol)
            for (int j = row_start; j < row_end; j++) Pc.val[j] /= -weak_sum;
        pos[i] = -1;
        Pc.len[i - r0] = row_end - row_start;
    }
    // the filter runs after every row of the range exists (row-local, like
    // the reference's post-pass), so no stored position is invalidated early
    filter_rows(Pc.col, Pc.val, Pc.len, filter);
    });
    assemble(n, ch, P);
    return P;
}

// ---------------------------------------------------------------------------
// smoothed aggregation, aggregation/par_*.cpp
// ---------------------------------------------------------------------------
constexpr int kTmpSelection = 4;  // core/types.hpp:31

// mis2 aggregation/par_mis.cpp:204-645 at one rank.  Distance-two maximal
// independent set on S, driven by the random values r = weights.  The
// reference's quirks are kept: a vertex marked TmpSelection keeps that mark
// until it is selected or unselected, and the reset of the "touched" list
// sets next[] to 0 instead of Unassigned, so a vertex touched in two
// iterations keeps C[w] == Selected from then on (:587-593).
void mis2(const Csr& S, const double* r, std::vector<int>& states)
{
    const int n = S.n_rows;
    states.assign((size_t)n, kUnassigned);
    if (!n) return;
    std::vector<int> V((size_t)n);
    std::iota(V.begin(), V.end(), 0);
    std::vector<int> C((size_t)n, kUnselected), next((size_t)n, kUnassigned);
    std::vector<double> row_max((size_t)n);

    // directed graph: entries of S toward smaller random values (:326-357)
    std::vector<int> d_ptr((size_t)n + 1, 0), d_col;
    d_col.reserve(S.col.size() / 2);
    for (int i = 0; i < n; i++) {
        for (int j = S.rowptr[i]; j < S.rowptr[i + 1]; j++)
            if (r[i] > r[S.col[j]]) d_col.push_back(S.col[j]);
        d_ptr[i + 1] = (int)d_col.size();
    }
    // columns of S (CSR_to_CSC, core/matrix.cpp:492-538)
    std::vector<int> c_ptr((size_t)n + 1, 0), c_row(S.col.size());
    for (int c : S.col) c_ptr[c + 1]++;
    for (int i = 0; i < n; i++) c_ptr[i + 1] += c_ptr[i];
    {
        std::vector<int> pos(c_ptr.begin(), c_ptr.end() - 1);
        for (int i = 0; i < n; i++)
            for (int j = S.rowptr[i]; j < S.rowptr[i + 1]; j++) c_row[pos[S.col[j]]++] = i;
    }

    int remaining = n;
    bool first_pass = true;
    int64_t guard = 0;
    while (remaining || first_pass) {
        if (++guard > (int64_t)n + 2) throw std::runtime_error("mis2: independent-set loop does not terminate");
        // tentative selection (:378-415)
        for (int i = 0; i < remaining; i++) {
            const int v = V[i];
            bool found = false;
            for (int j = d_ptr[v]; j < d_ptr[v + 1]; j++) {
                const int w = d_col[j];
                if (states[w] == kUnassigned || states[w] > kSelected) {
                    found = true;
                    break;
                }
            }
            if (!found) states[v] = kTmpSelection;
        }
        // largest tentative random value in each row (:420-446)
        for (int i = 0; i < n; i++) {
            double max_val = -1.0;
            for (int j = S.rowptr[i]; j < S.rowptr[i + 1]; j++) {
                const int u = S.col[j];
                if (states[u] > kSelected && r[u] > max_val) max_val = r[u];
            }
            row_max[i] = max_val;
        }
        // conflicts at distance two (:449-473)
        for (int i = 0; i < remaining; i++) {
            const int v = V[i];
            if (states[v] != kTmpSelection) continue;
            bool found = false;
            for (int j = S.rowptr[v]; j < S.rowptr[v + 1]; j++)
                if (row_max[S.col[j]] > r[v]) {
                    found = true;
                    break;
                }
            if (!found) states[v] = kNewSelection;
        }
        // mark the neighbours of new selections (:505-526)
        int head = kTmpSelection, length = 0;
        for (int i = 0; i < remaining; i++) {
            const int v = V[i];
            if (states[v] != kNewSelection) continue;
            for (int j = c_ptr[v]; j < c_ptr[v + 1]; j++) {
                const int w = c_row[j];
                C[w] = kSelected;
                if (next[w] == kUnassigned) {
                    next[w] = head;
                    head = w;
                    length++;
                }
            }
        }
        // unselect everything within distance two of a new selection (:555-586)
        for (int i = 0; i < remaining; i++) {
            const int v = V[i];
            if (states[v] == kNewSelection) continue;
            bool found = false;
            for (int j = S.rowptr[v]; j < S.rowptr[v + 1]; j++) {
                const int w = S.col[j];
                if (states[w] == kNewSelection || C[w] == kSelected) {
                    found = true;
                    break;
                }
            }
            if (found) states[v] = kNewUnselection;
        }
        for (int i = 0; i < length; i++) {
            const int t = head;
            head = next[head];
            C[t] = kUnselected;
            next[t] = 0;
        }
        // settle this iteration's decisions (:595-612)
        int ctr = 0;
        for (int i = 0; i < remaining; i++) {
            const int v = V[i];
            if (states[v] == kNewSelection)
                states[v] = kSelected;
            else if (states[v] == kNewUnselection)
                states[v] = kUnselected;
            else
                V[ctr++] = v;
        }
        remaining = ctr;
        first_pass = false;
    }
}

// aggregate aggregation/par_aggregate.cpp:5-185 at one rank.  S and A are
// sorted with the diagonal first in place (:14-17).  aggregates[i] is the
// root row of i's aggregate, or -1; returns the number of aggregates.
// rand_vals (may be NULL: zeros, as the solver calls it) break ties in the
// second pass.
int aggregate(Csr& A, Csr& S, const std::vector<int>& states, const double* rand_vals,
              std::vector<int>& aggregates)
{
    sort_and_diag_first(S);
    sort_and_diag_first(A);
    const int n = S.n_rows;
    const int N = n;  // global rows at one rank
    aggregates.assign((size_t)n, -1);
    std::vector<double> r((size_t)n, 0.0);
    if (rand_vals) std::copy(rand_vals, rand_vals + n, r.begin());
    int n_aggs = 0;
    for (int i = 0; i < n; i++) {
        if (S.rowptr[i + 1] - S.rowptr[i] <= 1)
            aggregates[i] = -N;
        else if (states[i] == kSelected) {
            aggregates[i] = i;
            n_aggs++;
        }
    }
    // pass 1: join the first selected strong neighbour (:96-126)
    for (int i = 0; i < n; i++) {
        if (states[i] == kSelected) continue;
        for (int j = S.rowptr[i]; j < S.rowptr[i + 1]; j++) {
            const int c = S.col[j];
            if (states[c] == kSelected && aggregates[c] >= 0) {
                aggregates[i] = aggregates[c];
                break;
            }
        }
    }
    // pass 2: join the aggregate of the strongest aggregated neighbour,
    // encoded as -(agg+1) until every row is done (:134-172)
    for (int i = 0; i < n; i++) {
        if (aggregates[i] != -1) continue;
        int ctr = A.rowptr[i];
        const int a_end = A.rowptr[i + 1];
        double max_val = 0.0;
        int max_agg = -N;
        for (int j = S.rowptr[i]; j < S.rowptr[i + 1]; j++) {
            const int c = S.col[j];
            while (ctr < a_end && A.col[ctr] != c) ctr++;
            if (ctr == a_end) throw std::runtime_error("aggregate: strength entry missing from A");
            const double val = std::fabs(A.val[ctr]) + r[c];
            if (val > max_val && aggregates[c] >= 0) {
                max_val = val;
                max_agg = aggregates[c];
            }
        }
        aggregates[i] = -(max_agg + 1);
    }
    for (int i = 0; i < n; i++) {
        if (aggregates[i] <= -N)
            aggregates[i] = -1;
        else if (aggregates[i] < 0)
            aggregates[i] = -(aggregates[i] + 1);
    }
    return n_aggs;
}

// fit_candidates aggregation/par_candidates.cpp:6-216 (one candidate) at one
// rank.  Aggregates become columns in increasing root order; column k holds
// B over its rows, scaled by 1/sqrt(sum of B^2 in ascending row order); R
// gets the norms (the next level's B).
Csr fit_candidates(int n_aggs, const std::vector<int>& aggregates, const std::vector<double>& B,
                   std::vector<double>& R)
{
    const int n = (int)aggregates.size();
    std::vector<int> colmap((size_t)n, 0);
    for (int a : aggregates)
        if (a >= 0) colmap[a] = 1;
    int nc = 0;
    for (int i = 0; i < n; i++)
        if (colmap[i]) colmap[i] = nc++;
    if (nc != n_aggs)
        throw std::runtime_error("fit_candidates: an aggregate has no selected root (the reference's aggregation "
                                 "leaves such a point undefined)");
    R.assign((size_t)n_aggs, 0.0);
    for (int i = 0; i < n; i++) {
        const int a = aggregates[i];
        if (a < 0) continue;
        R[colmap[a]] += B[i] * B[i];
    }
    std::vector<double> scale((size_t)n_aggs);
    for (int k = 0; k < n_aggs; k++) {
        R[k] = std::sqrt(R[k]);
        scale[k] = 1.0 / R[k];
    }
    Csr T;
    T.n_rows = n;
    T.n_cols = n_aggs;
    T.rowptr.assign((size_t)n + 1, 0);
    for (int i = 0; i < n; i++) {
        const int a = aggregates[i];
        if (a >= 0) {
            const int k = colmap[a];
            T.col.push_back(k);
            T.val.push_back(B[i] * scale[k]);
        }
        T.rowptr[i + 1] = (int)T.col.size();
    }
    return T;
}

// The row scaling of jacobi_prolongation aggregation/par_prolongation.cpp:
// 7-103: row i of A times omega / sum_j |a_ij| (the absolute row sum in
// stored order; rows summing to zero become zero).  A keeps its layout.
Csr sa_scaled_operator(const Csr& A, double omega)
{
    Csr M = A;
    parallel_ranges(A.n_rows, host_threads(A.nnz()), [&](int, int r0, int r1) {
        for (int i = r0; i < r1; i++) {
            double row_sum = 0.0;
            for (int j = A.rowptr[i]; j < A.rowptr[i + 1]; j++) row_sum += std::fabs(A.val[j]);
            double inv = 0.0;
            if (row_sum != 0.0) inv = (1.0 / std::fabs(row_sum)) * omega;
            for (int j = A.rowptr[i]; j < A.rowptr[i + 1]; j++) M.val[j] = A.val[j] * inv;
        }
    });
    return M;
}

// ParCSRMatrix::subtract util/linalg/par_add.cpp:184-310 at one rank: the
// rows of P followed by the negated rows of Q, sorted by column, equal
// columns summed in order, |v| < 1e-16 dropped (core/matrix.cpp:922-986),
// then the diagonal (column == row) moved first.  A row of P has at most
// one entry and the rows of Q have distinct columns, so no column appears
// more than twice and the sum does not depend on how the sort orders ties.
Csr sa_subtract(const Csr& P, const Csr& Q)
{
    const int n = P.n_rows;
    Csr C;
    C.n_rows = n;
    C.n_cols = P.n_cols;
    std::vector<RowChunk> ch((size_t)host_threads(P.nnz() + Q.nnz()));
    parallel_ranges(n, (int)ch.size(), [&](int t, int r0, int r1) {
        RowChunk& c = ch[t];
        std::vector<std::pair<int, double>> row;
        for (int i = r0; i < r1; i++) {
            row.clear();
            for (int j = P.rowptr[i]; j < P.rowptr[i + 1]; j++) row.emplace_back(P.col[j], P.val[j]);
            for (int j = Q.rowptr[i]; j < Q.rowptr[i + 1]; j++) row.emplace_back(Q.col[j], -Q.val[j]);
            std::stable_sort(row.begin(), row.end(),
                             [](const std::pair<int, double>& a, const std::pair<int, double>& b) {
                                 return a.first < b.first;
                             });
            // remove_duplicates
            const size_t base = c.col.size();
            for (size_t k = 0; k < row.size(); k++) {
                if (k > 0 && row[k].first == row[k - 1].first) {
                    c.val.back() += row[k].second;
                    continue;
                }
                if (c.col.size() > base && std::fabs(c.val.back()) < kZeroTol) {
                    c.col.pop_back();
                    c.val.pop_back();
                }
                c.col.push_back(row[k].first);
                c.val.push_back(row[k].second);
            }
            if (c.col.size() > base && std::fabs(c.val.back()) < kZeroTol) {
                c.col.pop_back();
                c.val.pop_back();
            }
            // move_diag
            for (size_t k = base; k < c.col.size(); k++) {
                if (c.col[k] == i) {
                    const double v = c.val[k];
                    for (size_t m = k; m > base; m--) {
                        c.col[m] = c.col[m - 1];
                        c.val[m] = c.val[m - 1];
                    }
                    c.col[base] = i;
                    c.val[base] = v;
                    break;
                }
            }
            c.len.push_back((int)(c.col.size() - base));
        }
    });
    assemble(n, ch, C);
    return C;
}

// ---------------------------------------------------------------------------
// stencil generators
// ---------------------------------------------------------------------------

// par_stencil_grid gallery/par_stencil.hpp:16-236 at one rank.  Row r takes,
// for its d-th nonzero stencil offset (ascending), the value of nonzero
// stencil entry N_s-1-d (:220), dropped when the neighbour leaves the grid
// in any dimension (the boundary zeroing of :137-209 for the symmetric
// stencils the gallery provides).
Csr stencil_grid(const double* stencil, const int* grid, int dim)
{
    int len = 1;
    for (int i = 0; i < dim; i++) len *= 3;
    int64_t nv = 1;
    for (int i = 0; i < dim; i++) nv *= grid[i];
    if (nv > INT32_MAX) throw std::length_error("stencil grid exceeds int32 rows");
    // extents and strides: dimension k (0 = fastest) has stride prod grid[dim-1..dim-k]
    std::vector<int64_t> stride(dim), extent(dim);
    stride[0] = 1;
    for (int k = 0; k < dim; k++) {
        extent[k] = grid[dim - 1 - k];
        if (k + 1 < dim) stride[k + 1] = stride[k] * extent[k];
    }
    std::vector<std::vector<int>> off;  // per nonzero entry: offset digit per dimension k
    std::vector<double> nz;
    std::vector<int64_t> diag;
    for (int i = 0; i < len; i++) {
        if (fabs(stencil[i]) > kZeroTol) {
            std::vector<int> o(dim);
            int64_t d = 0;
            int t = i;
            for (int k = 0; k < dim; k++) {
                o[k] = (t % 3) - 1;
                t /= 3;
                d += stride[k] * o[k];
            }
            off.push_back(o);
            nz.push_back(stencil[i]);
            diag.push_back(d);
        }
    }
    const int ns = (int)nz.size();
    Csr A;
    A.n_rows = A.n_cols = (int)nv;
    A.rowptr.assign((size_t)nv + 1, 0);
    A.col.reserve((size_t)nv * ns);
    A.val.reserve((size_t)nv * ns);
    std::vector<int64_t> coord(dim);
    for (int64_t r = 0; r < nv; r++) {
        int64_t t = r;
        for (int k = 0; k < dim; k++) {
            coord[k] = t % extent[k];
            t /= extent[k];
        }
        for (int d = 0; d < ns; d++) {
            bool inside = true;
            for (int k = 0; k < dim && inside; k++) {
                const int64_t c = coord[k] + off[d][k];
                inside = c >= 0 && c < extent[k];
            }
            const double v = nz[ns - 1 - d];
            if (inside && fabs(v) > kZeroTol) {
                A.col.push_back((int)(r + diag[d]));
                A.val.push_back(v);
            }
        }
        A.rowptr[r + 1] = (int)A.col.size();
    }
    return A;
}

// laplace_stencil_27pt gallery/laplacian27pt.hpp:27-39
std::vector<double> laplace_stencil_27pt()
{
    std::vector<double> s(27, -1.0);
    s[13] = 26;
    return s;
}

// diffusion_stencil_2d gallery/diffusion.hpp:58-85
std::vector<double> diffusion_stencil_2d(double eps, double theta)
{
    const double C = cos(theta), S = sin(theta);
    const double CS = C * S, CC = C * C, SS = S * S;
    const double v1 = ((-1 * eps - 1) * CC + (-1 * eps - 1) * SS + (3 * eps - 3) * CS) / 6.0;
    const double v2 = ((2 * eps - 4) * CC + (-4 * eps + 2) * SS) / 6.0;
    const double v3 = ((-1 * eps - 1) * CC + (-1 * eps - 1) * SS + (-3 * eps + 3) * CS) / 6.0;
    const double v4 = ((-4 * eps + 2) * CC + (2 * eps - 4) * SS) / 6.0;
    const double v5 = ((8 * eps + 8) * CC + (8 * eps + 8) * SS) / 6.0;
    return {v1, v2, v3, v4, v5, v4, v3, v2, v1};
}

// SplitMix64 -> uniform (u >> 11) * 2^-53 -> Box-Muller.
static uint64_t splitmix64(uint64_t& s)
{
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static double uniform01(uint64_t& s) { return (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0); }

static double gaussian(uint64_t& s)
{
    double u1 = uniform01(s), u2 = uniform01(s);
    if (u1 < 1e-300) u1 = 1e-300;
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

Csr fv7_variable(int m, uint64_t seed)
{
    const int64_t n = (int64_t)m * m * m;
    if (n > INT32_MAX) throw std::length_error("fv7 grid exceeds int32 rows");
    std::vector<double> k((size_t)n);
    uint64_t s = seed;
    for (int64_t i = 0; i < n; i++) k[i] = exp(gaussian(s));
    auto face = [&](int64_t p, int64_t q) { return 2.0 * k[p] * k[q] / (k[p] + k[q]); };
    Csr A;
    A.n_rows = A.n_cols = (int)n;
    A.rowptr.assign((size_t)n + 1, 0);
    A.col.reserve((size_t)n * 7);
    A.val.reserve((size_t)n * 7);
    const int64_t st[3] = {1, m, (int64_t)m * m};
    for (int64_t p = 0; p < n; p++) {
        const int64_t c[3] = {p % m, (p / m) % m, p / ((int64_t)m * m)};
        double off_vals[6];
        int64_t off_cols[6];
        double diag = 0;
        // neighbours in ascending column order: -z, -y, -x, +x, +y, +z
        const int dims[6] = {2, 1, 0, 0, 1, 2};
        const int sgn[6] = {-1, -1, -1, 1, 1, 1};
        for (int f = 0; f < 6; f++) {
            const int d = dims[f];
            const int64_t cc = c[d] + sgn[f];
            if (cc < 0 || cc >= m) {
                diag += 2.0 * k[p];  // Dirichlet wall at half-cell distance
                off_cols[f] = -1;
            } else {
                const int64_t q = p + sgn[f] * st[d];
                const double w = face(p, q);
                diag += w;
                off_cols[f] = q;
                off_vals[f] = -w;
            }
        }
        for (int f = 0; f < 3; f++)
            if (off_cols[f] >= 0) {
                A.col.push_back((int)off_cols[f]);
                A.val.push_back(off_vals[f]);
            }
        A.col.push_back((int)p);
        A.val.push_back(diag);
        for (int f = 3; f < 6; f++)
            if (off_cols[f] >= 0) {
                A.col.push_back((int)off_cols[f]);
                A.val.push_back(off_vals[f]);
            }
        A.rowptr[p + 1] = (int)A.col.size();
    }
    return A;
}

// ---------------------------------------------------------------------------
// Greedy colouring of A + A^T for the multicolour smoothers
// ---------------------------------------------------------------------------
int greedy_colors(int n, const int* rp, const int* col, std::vector<int>& color)
{
    // transpose pattern, so that a row sees every row that references it
    std::vector<int> tp((size_t)n + 1, 0);
    for (int i = 0; i < n; i++)
        for (int j = rp[i]; j < rp[i + 1]; j++)
            if (col[j] != i && col[j] >= 0 && col[j] < n) tp[col[j] + 1]++;
    for (int i = 0; i < n; i++) tp[i + 1] += tp[i];
    std::vector<int> trow((size_t)tp[n]);
    std::vector<int> pos(tp.begin(), tp.end() - 1);
    for (int i = 0; i < n; i++)
        for (int j = rp[i]; j < rp[i + 1]; j++)
            if (col[j] != i && col[j] >= 0 && col[j] < n) trow[pos[col[j]]++] = i;
    color.assign((size_t)n, -1);
    std::vector<int> seen;  // seen[c] == i + 1: colour c taken by a neighbour of row i
    int ncolors = 0;
    for (int i = 0; i < n; i++) {
        auto mark = [&](int j) {
            const int c = color[j];
            if (c >= 0) seen[c] = i + 1;
        };
        if ((int)seen.size() < ncolors + 1) seen.resize((size_t)ncolors + 1, 0);
        for (int j = rp[i]; j < rp[i + 1]; j++)
            if (col[j] != i && col[j] >= 0 && col[j] < n) mark(col[j]);
        for (int j = tp[i]; j < tp[i + 1]; j++) mark(trow[j]);
        int c = 0;
        while (c < ncolors && seen[c] == i + 1) c++;
        color[i] = c;
        if (c == ncolors) ncolors++;
    }
    return ncolors;
}

}  // namespace host
}  // namespace ramd

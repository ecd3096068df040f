// Device-resident AMG hierarchy shared by the single-vector path
// (hier.hip) and the multi-vector path (block.hip).
#pragma once

#include <vector>

#include "common.hpp"

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
};

struct BlockState;  // multi-vector work space (block.hip)
void block_free(BlockState* b);

}  // namespace ramd

struct ramd_hier_s {
    ramd_ctx ctx = nullptr;
    ramd_amg_opts opts{};
    std::vector<ramd::Level> levels;
    int coarse_n = 0;
    double* lu = nullptr;  // column-major LU of the coarse matrix read column-major (A_c^T)
    int* piv = nullptr;
    double* cb = nullptr;  // FP64 staging for the FP32 cycle's coarse solve
    double* cx = nullptr;
    double times[5] = {0, 0, 0, 0, 0};
    bool fp32 = false;
    // PCG work vectors (fine level)
    double *pr = nullptr, *pz = nullptr, *pp = nullptr, *pAp = nullptr;
    float *pr32 = nullptr, *pz32 = nullptr;
    ramd::BlockState* blk = nullptr;
};

namespace ramd {
// Builds the FP32 copies of every level (values, diagonal, level vectors).
void enable_fp32(ramd_hier_s* h);
inline dim3 grid1(int64_t n, int bs = 256)
{
    int64_t g = (n + bs - 1) / bs;
    return dim3((unsigned)(g > 0 ? g : 1));
}
}  // namespace ramd

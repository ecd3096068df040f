// BiCGStab and AMG-preconditioned BiCGStab (krylov/par_bicgstab.cpp) at one
// rank on the device.  The control flow and every scalar follow the
// reference: res[] receives ||r_i||_2 (ParVector::norm: entries with
// |v| <= 1e-16 skipped, root re-taken) from ||r_0|| on, the loop runs while
// ||r|| > tol ||r_0|| and iter < max_iter, and each vector update is the
// reference's sequence of copy / axpy / scale passes fused per element into
// one pass (the same operations in the same order, so the same bits).  The
// inner products are the library's deterministic reductions (fixed grid,
// fixed order): their rounding differs from the reference's serial sums by
// a few ulps, so histories agree to round-off, not bit for bit.
#include <cmath>

#include "hier.hpp"

namespace ramd {

// s = r + (-alpha) Ap          (s.copy(r); s.axpy(Ap, -1.0 * alpha))
__global__ void bicg_s_kernel(int n, double m_alpha, const double* __restrict__ r, const double* __restrict__ Ap,
                              double* __restrict__ s)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) s[i] = r[i] + m_alpha * Ap[i];
}

// x = (x + alpha pd) + omega sd;  r = s + (-omega) As
__global__ void bicg_xr_kernel(int n, double alpha, double omega, double m_omega, const double* __restrict__ pd,
                               const double* __restrict__ sd, const double* __restrict__ s,
                               const double* __restrict__ As, double* __restrict__ x, double* __restrict__ r)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double t = x[i] + alpha * pd[i];
    x[i] = t + omega * sd[i];
    r[i] = s[i] + m_omega * As[i];
}

// p = ((p * beta) + 1.0 r) + (-beta omega) Ap   (p.scale; p.axpy(r, 1.0); p.axpy(Ap, -1.0*beta*omega))
__global__ void bicg_p_kernel(int n, double beta, double m_beta_omega, const double* __restrict__ r,
                              const double* __restrict__ Ap, double* __restrict__ p)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double t = p[i] * beta;
    const double u = t + 1.0 * r[i];
    p[i] = u + m_beta_omega * Ap[i];
}

namespace {

dim3 grid_n(int n) { return dim3((unsigned)std::max(1, (n + 255) / 256)); }

// ParVector::norm(2) from the sum of squares over |v| > 1e-16
double reference_norm(double ss) { return pow(pow(pow(ss, 1.0 / 2), 2), 1.0 / 2); }

int bicgstab_run(ramd_hier_s* h, ramd_csr_s* A, double* x, const double* b, double tol, int max_iter,
                 double* res, int max_res, int32_t* n_res, int32_t* iters)
{
    ramd_ctx ctx = A->ctx;
    hipStream_t s = ctx->stream;
    const int n = A->n_rows;
    struct Bufs {
        double* v[8] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
        ~Bufs()
        {
            for (double* p : v) dfree(p);
        }
    } w;
    for (double*& p : w.v) p = dalloc<double>((size_t)std::max(n, 1));
    double *r = w.v[0], *rs = w.v[1], *sv = w.v[2], *p = w.v[3], *Ap = w.v[4], *As = w.v[5], *ph = w.v[6],
           *sh = w.v[7];
    int nres = 0;
    auto push = [&](double v) {
        if (nres < max_res) res[nres] = v;
        nres++;
    };
    double* red = ctx->red_result;
    auto dots2 = [&](const double* a1, const double* b1, const double* a2, const double* b2, double out[2]) {
        launch_dot(ctx, n, a1, b1, red);
        launch_dot(ctx, n, a2, b2, red + 1);
        fetch_scalars(ctx, red, 2, out);
    };
    if (max_iter <= 0) max_iter = ((int)(1.3 * n)) + 2;
    // r0 = b - A x0; r* = r0; p0 = r0
    launch_stream_op(A, OP_RESIDUAL, x, b, r, 0.0, s);
    RAMD_HIP(hipMemcpyAsync(rs, r, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    RAMD_HIP(hipMemcpyAsync(p, r, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    launch_dot(ctx, n, r, rs, red);
    launch_norm2sq(ctx, n, r, red + 1);
    double sc[2];
    fetch_scalars(ctx, red, 2, sc);
    double rr_inner = sc[0];
    double norm_r = reference_norm(sc[1]);
    push(norm_r);
    if (norm_r != 0.0) tol = tol * norm_r;
    int iter = 0;
    while (norm_r > tol && iter < max_iter) {
        // p_hat = M^-1 p (or p itself); alpha = (r, r*) / (A p_hat, r*)
        const double* pd = p;
        if (h) {
            vcycle_from_zero(h, ph, p);
            pd = ph;
        }
        launch_stream_op(A, OP_SPMV, pd, nullptr, Ap, 0.0, s);
        launch_dot(ctx, n, Ap, rs, red);
        const double Apr_inner = fetch_scalar(ctx, red);
        const double alpha = rr_inner / Apr_inner;
        hipLaunchKernelGGL(bicg_s_kernel, grid_n(n), dim3(256), 0, s, n, -1.0 * alpha, r, Ap, sv);
        RAMD_HIP(hipGetLastError());
        // s_hat = M^-1 s; omega = (A s_hat, s) / (A s_hat, A s_hat)
        const double* sd = sv;
        if (h) {
            vcycle_from_zero(h, sh, sv);
            sd = sh;
        }
        launch_stream_op(A, OP_SPMV, sd, nullptr, As, 0.0, s);
        dots2(As, sv, As, As, sc);
        const double omega = sc[0] / sc[1];
        hipLaunchKernelGGL(bicg_xr_kernel, grid_n(n), dim3(256), 0, s, n, alpha, omega, -1.0 * omega, pd, sd, sv,
                           As, x, r);
        RAMD_HIP(hipGetLastError());
        // beta = (r_new, r*) / (r, r*) * alpha / omega; p = r + beta (p - omega Ap)
        launch_dot(ctx, n, r, rs, red);
        launch_norm2sq(ctx, n, r, red + 1);
        fetch_scalars(ctx, red, 2, sc);
        const double next_inner = sc[0];
        const double beta = (next_inner / rr_inner) * (alpha / omega);
        hipLaunchKernelGGL(bicg_p_kernel, grid_n(n), dim3(256), 0, s, n, beta, -1.0 * beta * omega, r, Ap, p);
        RAMD_HIP(hipGetLastError());
        rr_inner = next_inner;
        norm_r = reference_norm(sc[1]);
        push(norm_r);
        iter++;
    }
    RAMD_HIP(hipStreamSynchronize(s));
    *n_res = nres;
    *iters = iter;
    return std::isfinite(norm_r) ? RAMD_OK : RAMD_E_NUMERIC;
}

}  // namespace
}  // namespace ramd

using namespace ramd;

extern "C" {

int ramd_bicgstab(ramd_csr A, double* x_d, const double* b_d, double tol, int32_t max_iter, double* res,
                  int32_t max_res, int32_t* n_res, int32_t* iters)
{
    int rc = RAMD_OK;
    int st = guarded([&] {
        RAMD_CHECK(A && x_d && b_d && n_res && iters && (res || max_res == 0), RAMD_E_ARG,
                   "ramd_bicgstab: NULL argument");
        RAMD_CHECK(A->n_rows == A->n_cols, RAMD_E_ARG, "ramd_bicgstab: A is not square");
        A->ctx->set_current();
        rc = bicgstab_run(nullptr, A, x_d, b_d, tol, max_iter, res, max_res, n_res, iters);
    });
    if (st != RAMD_OK) return st;
    if (rc != RAMD_OK) set_error("ramd_bicgstab: breakdown (the residual norm is not finite)");
    return rc;
}

int ramd_pre_bicgstab(ramd_hier h, ramd_csr A, double* x_d, const double* b_d, double tol, int32_t max_iter,
                      double* res, int32_t max_res, int32_t* n_res, int32_t* iters)
{
    int rc = RAMD_OK;
    int st = guarded([&] {
        RAMD_CHECK(h && A && x_d && b_d && n_res && iters && (res || max_res == 0), RAMD_E_ARG,
                   "ramd_pre_bicgstab: NULL argument");
        RAMD_CHECK(A->n_rows == A->n_cols && A->n_rows == h->levels[0].n, RAMD_E_ARG,
                   "ramd_pre_bicgstab: A has %d rows, the hierarchy %d", A->n_rows, h->levels[0].n);
        RAMD_CHECK(A->ctx == h->ctx, RAMD_E_ARG, "ramd_pre_bicgstab: A and the hierarchy use different contexts");
        h->ctx->set_current();
        cycle_prepare(h, 0);
        rc = bicgstab_run(h, A, x_d, b_d, tol, max_iter, res, max_res, n_res, iters);
    });
    if (st != RAMD_OK) return st;
    if (rc != RAMD_OK) set_error("ramd_pre_bicgstab: breakdown (the residual norm is not finite)");
    return rc;
}

}  // extern "C"

// libtorch autograd Function over the C ABI: x = A(vals)^-1 b solved by the
// device AMG-preconditioned flexible PCG; the backward pass is the adjoint
// (ramd_adjoint): A^T lambda = dL/dx through the same hierarchy, then
// dL/db = lambda and dL/dA_k = -lambda_row(k) x_col(k) on A's pattern.
//
// The reference has no differentiation; this wraps the solve path for
// callers that train through it.  Tensors are FP64 device tensors on the
// context's GPU; the library runs on its own stream, so torch's queued work
// is drained before each call and every library call returns synchronised.
#include <torch/cuda.h>
#include <torch/extension.h>

#include <string>

#include "raptor_amd.h"

namespace {

std::string last_error()
{
    char buf[1024];
    ramd_last_error(buf, sizeof buf);
    return buf;
}

void check_input(const torch::Tensor& t, const char* name)
{
    TORCH_CHECK(t.is_cuda(), name, " must be a GPU tensor");
    TORCH_CHECK(t.scalar_type() == torch::kFloat64, name, " must be float64");
    TORCH_CHECK(t.is_contiguous() && t.dim() == 1, name, " must be a contiguous vector");
}

struct AmgSolve : public torch::autograd::Function<AmgSolve> {
    static torch::Tensor forward(torch::autograd::AutogradContext* ctx, torch::Tensor vals, torch::Tensor b,
                                 int64_t hier, int64_t mat, double rtol, int64_t max_iter, int64_t prec)
    {
        check_input(vals, "vals");
        check_input(b, "b");
        auto A = reinterpret_cast<ramd_csr>(mat);
        auto h = reinterpret_cast<ramd_hier>(hier);
        int32_t n = 0, nc = 0;
        int64_t nnz = 0;
        TORCH_CHECK(ramd_csr_info(A, &n, &nc, &nnz) == RAMD_OK, last_error());
        TORCH_CHECK(vals.numel() == nnz && b.numel() == n, "vals has ", vals.numel(), " entries (A has ", nnz,
                    "), b has ", b.numel(), " (A has ", n, " rows)");
        torch::Tensor x = torch::zeros_like(b);
        torch::cuda::synchronize(b.device().index());
        TORCH_CHECK(ramd_csr_set_values(A, vals.data_ptr<double>()) == RAMD_OK, last_error());
        int32_t it = 0;
        double rel = 0;
        const int rc = ramd_fpcg(h, A, x.data_ptr<double>(), b.data_ptr<double>(), rtol, (int)max_iter, (int)prec,
                                 nullptr, &it, &rel);
        TORCH_CHECK(rc == RAMD_OK, "amg_solve forward: ", last_error());
        ctx->save_for_backward({x});
        ctx->saved_data["hier"] = hier;
        ctx->saved_data["mat"] = mat;
        ctx->saved_data["rtol"] = rtol;
        ctx->saved_data["max_iter"] = max_iter;
        ctx->saved_data["prec"] = prec;
        return x;
    }

    static torch::autograd::tensor_list backward(torch::autograd::AutogradContext* ctx,
                                                 torch::autograd::tensor_list grad_out)
    {
        torch::Tensor x = ctx->get_saved_variables()[0];
        torch::Tensor g = grad_out[0].contiguous();
        auto A = reinterpret_cast<ramd_csr>(ctx->saved_data["mat"].toInt());
        auto h = reinterpret_cast<ramd_hier>(ctx->saved_data["hier"].toInt());
        int32_t n = 0, nc = 0;
        int64_t nnz = 0;
        TORCH_CHECK(ramd_csr_info(A, &n, &nc, &nnz) == RAMD_OK, last_error());
        torch::Tensor dA = torch::empty({nnz}, x.options());
        torch::Tensor db = torch::empty_like(x);
        torch::cuda::synchronize(x.device().index());
        int32_t it = 0;
        double rel = 0;
        const int rc = ramd_adjoint(h, A, x.data_ptr<double>(), g.data_ptr<double>(), ctx->saved_data["rtol"].toDouble(),
                                    (int)ctx->saved_data["max_iter"].toInt(), (int)ctx->saved_data["prec"].toInt(),
                                    dA.data_ptr<double>(), db.data_ptr<double>(), &it, &rel);
        TORCH_CHECK(rc == RAMD_OK, "amg_solve backward: ", last_error());
        return {dA, db, torch::Tensor(), torch::Tensor(), torch::Tensor(), torch::Tensor(), torch::Tensor()};
    }
};

torch::Tensor amg_solve(torch::Tensor vals, torch::Tensor b, int64_t hier, int64_t mat, double rtol, int64_t max_iter,
                        int64_t prec)
{
    return AmgSolve::apply(vals, b, hier, mat, rtol, max_iter, prec);
}

}  // namespace

PYBIND11_MODULE(TORCH_EXTENSION_NAME, m)
{
    m.def("amg_solve", &amg_solve,
          "x = A(vals)^-1 b by device AMG-PCG; differentiable in vals and b (adjoint backward)");
}

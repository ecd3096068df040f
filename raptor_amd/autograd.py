"""Differentiable AMG solve for PyTorch: x = A(vals)^-1 b.

The forward pass is the device flexible PCG preconditioned by a hierarchy
built once for A's pattern and values; the backward pass is the adjoint
(A^T lambda = dL/dx through the same hierarchy, dL/db = lambda,
dL/dA_k = -lambda_row(k) x_col(k) on A's stored pattern).  Both run in the
libtorch extension raptor_amd/_torch_ext.so over the C ABI; there is no
Python or CPU fallback.

    ml = raptor_amd.ParSmoothedAggregationSolver(ctx, 0.08).setup(rp, col, val)
    A = raptor_amd.ParCSRMatrix.from_csr(ctx, rp, col, val)
    vals = torch.tensor(val, device="cuda", requires_grad=True)
    x = raptor_amd.autograd.amg_solve(vals, b, ml, A)
    loss = 0.5 * ((x - x_star) ** 2).sum(); loss.backward()   # vals.grad, b.grad

Import torch before the first raptor_amd call: the library then binds to
torch's HIP runtime (one process cannot use two HIP runtimes).

The hierarchy stays that of the original values: when vals moves far from
them, the solve still converges to rtol (the hierarchy is only the
preconditioner) but needs more iterations.
"""
import importlib.util
import os

from . import _lib
from ._lib import PREC_FP64

_ext = None


def _hip_runtimes():
    """Paths of the HIP runtime libraries mapped into this process."""
    paths = set()
    try:
        with open("/proc/self/maps") as f:
            for line in f:
                p = line.split()[-1] if len(line.split()) >= 6 else ""
                if "libamdhip64" in os.path.basename(p):
                    paths.add(os.path.realpath(p))
    except OSError:
        pass
    return paths


def _load():
    global _ext
    if _ext is None:
        import torch  # noqa: F401  (libtorch and its HIP runtime before the library)
        _lib.lib()  # libraptor_amd.so first, so the extension binds to the same instance
        if len(_hip_runtimes()) > 1:
            raise RuntimeError("two HIP runtimes are loaded (libraptor_amd.so was loaded before torch): "
                               "import torch before the first raptor_amd call when using raptor_amd.autograd")
        path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_torch_ext.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is missing: build it with __graft_entry__.build() "
                               "(make -C raptor_amd/csrc torch_ext ...)")
        spec = importlib.util.spec_from_file_location("raptor_amd._torch_ext", path)
        mod = importlib.util.module_from_spec(spec)
        spec.loader.exec_module(mod)
        _ext = mod
    return _ext


def amg_solve(vals, b, ml, A, rtol=1e-10, max_iter=500, cycle_prec=PREC_FP64):
    """x = A(vals)^-1 b; vals (nnz, A's stored order) and b are float64 GPU
    tensors.  A's device values are replaced by vals on every call; A must
    not be a hierarchy's own level matrix."""
    return _load().amg_solve(vals, b, int(ml.h.value), int(A.h.value), float(rtol), int(max_iter),
                             int(cycle_prec))

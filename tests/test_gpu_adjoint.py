"""Adjoint of the device AMG solve (ramd_adjoint) and the libtorch autograd
Function over it: gradients of L = 0.5 ||x - x*||^2, x = A^-1 b, with respect
to b and to the stored values of A.

Checks: central finite differences (h = 1e-6, FP64) computed with dense
LAPACK solves of the perturbed systems on the CPU (the closed form the
gradient must match, independent of the formula under test); the FP64
oracle's PCG on the same hierarchy for lambda; the FP32-cycle form; and the
torch autograd path against the direct C ABI call."""
import numpy as np
import pytest

import checkers as K
import raptor_amd as ra

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def problem(ctx):
    # the 7-point finite-volume operator with log-normal coefficients (C5
    # shape) on a 10^3 grid, smoothed-aggregation hierarchy, theta 0.08
    rp, col, val = ra.gallery_fv7(10, seed=3)
    n = rp.size - 1
    ml = ra.ParSmoothedAggregationSolver(ctx, 0.08)
    ml.relax_weight = 2.0 / 3.0
    ml.max_coarse = 20
    ml.setup(rp, col, val)
    A = ra.ParCSRMatrix.from_csr(ctx, rp, col, val)
    rng = np.random.default_rng(4)
    b = rng.standard_normal(n)
    xs = rng.standard_normal(n)
    return dict(rp=rp, col=col, val=val, n=n, ml=ml, A=A, b=b, xs=xs)


def dense(rp, col, val, n):
    D = np.zeros((n, n))
    for i in range(n):
        D[i, col[rp[i]:rp[i + 1]]] = val[rp[i]:rp[i + 1]]
    return D


def loss_dense(D, b, xs):
    x = np.linalg.solve(D, b)
    return 0.5 * np.sum((x - xs) ** 2)


def device_grad(problem, ctx, prec=ra.PREC_FP64):
    A, ml = problem["A"], problem["ml"]
    x = ra.ParVector.from_numpy(ctx, np.zeros(problem["n"]))
    _, it, rel = ra.FPCG(A, ml, x, ra.ParVector.from_numpy(ctx, problem["b"]), 1e-13, 500, prec)
    xh = x.numpy()
    g = ra.ParVector.from_numpy(ctx, xh - problem["xs"])
    dA, db, it = ra.adjoint(A, ml, x, g, rtol=1e-13, cycle_prec=prec)
    return xh, dA.numpy(), db.numpy()


def test_adjoint_matches_finite_differences(problem, ctx):
    rp, col, val, n = problem["rp"], problem["col"], problem["val"], problem["n"]
    b, xs = problem["b"], problem["xs"]
    x, dA, db = device_grad(problem, ctx)
    D = dense(rp, col, val, n)
    assert np.allclose(x, np.linalg.solve(D, b), rtol=0, atol=1e-9 * np.abs(x).max())
    h = 1e-6
    rng = np.random.default_rng(5)
    rows = np.repeat(np.arange(n), np.diff(rp))
    for k in rng.choice(col.size, 32, replace=False):
        i, j = rows[k], col[k]
        Dp, Dm = D.copy(), D.copy()
        Dp[i, j] += h
        Dm[i, j] -= h
        fd = (loss_dense(Dp, b, xs) - loss_dense(Dm, b, xs)) / (2 * h)
        assert abs(dA[k] - fd) <= 1e-5 * max(abs(fd), 1e-3 * np.abs(dA).max()), (k, i, j, dA[k], fd)
    for i in rng.choice(n, 8, replace=False):
        bp, bm = b.copy(), b.copy()
        bp[i] += h
        bm[i] -= h
        fd = (loss_dense(D, bp, xs) - loss_dense(D, bm, xs)) / (2 * h)
        assert abs(db[i] - fd) <= 1e-5 * max(abs(fd), 1e-3 * np.abs(db).max()), (i, db[i], fd)


def test_adjoint_matches_oracle_gradient(problem, ctx):
    """lambda from the oracle's FP64 PCG (reference rules, Jacobi) on the
    product's own hierarchy; dL/dA = -lambda_i x_j formed on the host."""
    x, dA, db = device_grad(problem, ctx)
    ml = problem["ml"]
    As, Ps = [], []
    for l in range(ml.num_levels):
        Al, Pl = ml.level(l)
        As.append(K.Csr(*Al.to_csr(), Al.n_cols))
        if Pl is not None:
            Ps.append(K.Csr(*Pl.to_csr(), Pl.n_cols))
    H = K.Hier(As, Ps, relax_type=0, sweeps=1, omega=2.0 / 3.0)
    A = K.Csr(problem["rp"], problem["col"], problem["val"])
    g = x - problem["xs"]
    # the reference rule stops on r.z < tol * sqrt(g.M g): ask for r.z ~ 1e-26 g.M g
    gmg = float(np.dot(g, H.cycle(np.zeros(problem["n"]), g)))
    lam, _, it = H.pcg(A, np.zeros(problem["n"]), g, tol=1e-26 * np.sqrt(gmg), max_iter=300)
    assert it > 0
    assert np.allclose(db, lam, rtol=1e-8, atol=1e-10 * np.abs(lam).max())
    rows = np.repeat(np.arange(problem["n"]), np.diff(problem["rp"]))
    ref = -(lam[rows] * x[problem["col"]])
    assert np.allclose(dA, ref, rtol=1e-7, atol=1e-9 * np.abs(ref).max())


def test_adjoint_fp32_cycle_same_gradient(problem, ctx):
    _, dA64, db64 = device_grad(problem, ctx, ra.PREC_FP64)
    _, dA32, db32 = device_grad(problem, ctx, ra.PREC_FP32)
    assert np.allclose(db32, db64, rtol=1e-9, atol=1e-11 * np.abs(db64).max())
    assert np.allclose(dA32, dA64, rtol=1e-9, atol=1e-11 * np.abs(dA64).max())


_TORCH_SCRIPT = r"""
import sys
import numpy as np
import torch                      # first: the library then binds to torch's HIP runtime
sys.path.insert(0, sys.argv[1])
import raptor_amd as ra
from raptor_amd import autograd
ctx = ra.Context(0)
rp, col, val = ra.gallery_fv7(10, seed=3)
n = rp.size - 1
ml = ra.ParSmoothedAggregationSolver(ctx, 0.08)
ml.relax_weight = 2.0 / 3.0
ml.max_coarse = 20
ml.setup(rp, col, val)
rng = np.random.default_rng(4)
b_h, xs_h = rng.standard_normal(n), rng.standard_normal(n)
dev = torch.device("cuda", 0)
vals = torch.tensor(val, dtype=torch.float64, device=dev, requires_grad=True)
b = torch.tensor(b_h, dtype=torch.float64, device=dev, requires_grad=True)
xs = torch.tensor(xs_h, dtype=torch.float64, device=dev)
Aop = ra.ParCSRMatrix.from_csr(ctx, rp, col, val)   # the autograd path rewrites its values
x = autograd.amg_solve(vals, b, ml, Aop, rtol=1e-13)
loss = 0.5 * ((x - xs) ** 2).sum()
loss.backward()
# the same gradient through the C ABI directly
A = ra.ParCSRMatrix.from_csr(ctx, rp, col, val)
xd = ra.ParVector.from_numpy(ctx, np.zeros(n))
ra.FPCG(A, ml, xd, ra.ParVector.from_numpy(ctx, b_h), 1e-13, 500)
xh = xd.numpy()
dA, db, _ = ra.adjoint(A, ml, xd, ra.ParVector.from_numpy(ctx, xh - xs_h), rtol=1e-13)
assert np.array_equal(x.detach().cpu().numpy(), xh), "x"
assert np.array_equal(vals.grad.cpu().numpy(), dA.numpy()), "dA"
assert np.array_equal(b.grad.cpu().numpy(), db.numpy()), "db"
# a second forward with moved values converges and differs
with torch.no_grad():
    v2 = vals * 1.01
x2 = autograd.amg_solve(v2, b.detach(), ml, Aop, rtol=1e-13)
assert torch.allclose(x2 * 1.01, x.detach(), rtol=1e-6, atol=0), "scaled A"
ctx.close()
print("TORCH_AUTOGRAD_OK")
"""


def test_torch_autograd_function():
    """The libtorch autograd Function gives exactly the C ABI's solution and
    adjoint gradient.  Runs in its own interpreter, which imports torch
    before the library so that both use torch's HIP runtime (two HIP
    runtimes cannot share one process)."""
    pytest.importorskip("torch")
    import os
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, "-c", _TORCH_SCRIPT, root], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0 and "TORCH_AUTOGRAD_OK" in r.stdout, r.stdout[-2000:] + r.stderr[-4000:]


def test_set_values_refreshes_operator(ctx):
    """ramd_csr_set_values on a matrix with the stencil-diagonal form: the
    product with new values equals the oracle's, bit for bit."""
    rp, col, val = ra.gallery_fv7(18, seed=9)
    A = ra.ParCSRMatrix.from_csr(ctx, rp, col, val)
    assert A.diagonals() == 7
    v2 = val * (1.0 + 0.25 * np.random.default_rng(1).random(val.size))
    A.set_values(ra.ParVector.from_numpy(ctx, v2))
    xh = np.random.default_rng(2).standard_normal(rp.size - 1)
    y = ra.ParVector(ctx, rp.size - 1)
    A.mult(ra.ParVector.from_numpy(ctx, xh), y)
    assert np.array_equal(y.numpy(), K.spmv(K.Csr(rp, col, v2), xh))

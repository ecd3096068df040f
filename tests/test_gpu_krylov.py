"""GPU BiCGStab and AMG-preconditioned BiCGStab (krylov/par_bicgstab.cpp)
against the oracle, which is itself bit-identical to the compiled reference,
and the reference's golden history (krylov/tests/test_par_bicgstab.cpp)."""
import math

import numpy as np
import pytest

import checkers as K
import raptor_amd as ra

pytestmark = pytest.mark.gpu

OMEGA = 2.0 / 3.0


def aniso(m):
    rp, col, val = ra.gallery_aniso2d(m, 0.001, math.pi / 8)
    return K.Csr(rp, col, val)


def dev(ctx, a):
    return ra.ParVector.from_numpy(ctx, np.asarray(a, dtype=np.float64))


def test_bicgstab_golden_and_oracle(ctx, golden):
    """50 x 50 anisotropic, b = A*1, x0 = 0 (the reference test): the first 30
    residuals match the golden file to 1e-6 (its test's bound) and the
    oracle's to round-off.  Unpreconditioned BiCGStab on this problem
    amplifies round-off (the golden file, the reference and this build take
    103, 101 and about 105 iterations), so the tail is held to the stop
    rule, not to the oracle's history."""
    A = aniso(50)
    n = A.n_rows
    b = K.spmv(A, np.ones(n))
    D = ra.ParCSRMatrix.from_csr(ctx, A.rowptr, A.col, A.val)
    x = dev(ctx, np.zeros(n))
    res, it = ra.BiCGStab(D, x, dev(ctx, b))
    assert len(res) == it + 1
    assert np.allclose(res[:30], golden["bicgstab_res"][:30], rtol=0, atol=1e-6)
    xo, reso, ito = K.bicgstab(A, np.zeros(n), b)
    # round-off differences start at 1e-15 and grow about tenfold every
    # three iterations on this problem (1e-14 at step 10, 1.5e-6 at step 30)
    rel = np.abs(res[:30] - reso[:30]) / reso[:30]
    assert rel[:10].max() < 1e-12 and rel.max() < 1e-4, rel
    assert abs(it - ito) <= 0.15 * ito, (it, ito)
    # the reference's stop rule holds for the true residual up to round-off
    true = np.linalg.norm(b - K.spmv(A, x.numpy()))
    assert res[-1] <= 1e-5 * res[0] and true <= 1.001e-5 * res[0]


def test_bicgstab_well_conditioned_matches_oracle(ctx):
    """27-point Laplacian 12^3 shifted to strong diagonal dominance: the whole
    history matches the oracle to 1e-10 relative and the iteration count is
    the same (only the inner products' summation order differs)."""
    rp, col, val = ra.gallery_laplace27(12)
    val = val.copy()
    for i in range(rp.size - 1):
        for j in range(rp[i], rp[i + 1]):
            if col[j] == i:
                val[j] += 10.0
    A = K.Csr(rp, col, val)
    n = A.n_rows
    b = np.random.default_rng(9).standard_normal(n)
    D = ra.ParCSRMatrix.from_csr(ctx, rp, col, val)
    x = dev(ctx, np.zeros(n))
    res, it = ra.BiCGStab(D, x, dev(ctx, b), tol=1e-12)
    xo, reso, ito = K.bicgstab(A, np.zeros(n), b, tol=1e-12)
    assert it == ito and len(res) == len(reso)
    assert np.allclose(res, reso, rtol=1e-10, atol=0)
    assert np.allclose(x.numpy(), xo, rtol=0, atol=1e-12 * np.abs(xo).max())


def test_bicgstab_zero_rhs(ctx):
    """b = 0 and x0 = 0: ||r0|| = 0, no iteration, x stays 0."""
    A = aniso(20)
    D = ra.ParCSRMatrix.from_csr(ctx, A.rowptr, A.col, A.val)
    x = dev(ctx, np.zeros(A.n_rows))
    res, it = ra.BiCGStab(D, x, dev(ctx, np.zeros(A.n_rows)))
    assert it == 0 and list(res) == [0.0]
    assert not np.any(x.numpy())


@pytest.mark.parametrize("kind", ["rs", "sa"])
def test_pre_bicgstab_matches_oracle(ctx, kind):
    """Pre_BiCGStab with the product's hierarchy (FP64 V-cycle, bit-identical
    to the oracle's) against orc_pre_bicgstab on the same levels: same
    iteration count within one, histories equal to round-off at the start
    (the inner products' summation order differs) and to 1e-4 relative
    while above 1e-6 of the first residual, and a converged true residual."""
    rp, col, val = ra.gallery_aniso2d(80, 0.001, math.pi / 8)
    A = K.Csr(rp, col, val)
    n = A.n_rows
    if kind == "rs":
        ml = ra.ParRugeStubenSolver(ctx, 0.25, ra.CLJP, ra.ModClassical, ra.Classical, ra.Jacobi)
    else:
        ml = ra.ParSmoothedAggregationSolver(ctx, 0.0)
    ml.relax_weight = OMEGA
    ml.setup(rp, col, val)
    As, Ps = [], []
    for l in range(ml.num_levels):
        Al, Pl = ml.level(l)
        As.append(K.Csr(*Al.to_csr(), Al.n_cols))
        if Pl is not None:
            Ps.append(K.Csr(*Pl.to_csr(), Pl.n_cols))
    H = K.Hier(As, Ps, relax_type=0, sweeps=1, omega=OMEGA)
    b = K.spmv(A, np.ones(n))
    D = ra.ParCSRMatrix.from_csr(ctx, rp, col, val)
    x = dev(ctx, np.zeros(n))
    res, it = ra.Pre_BiCGStab(D, x, dev(ctx, b), ml, tol=1e-8)
    xo, reso, ito = K.bicgstab(A, np.zeros(n), b, tol=1e-8, hier=H)
    assert abs(it - ito) <= 1, (it, ito)
    k = min(len(res), len(reso))
    big = reso[:k] > 1e-6 * reso[0]
    rel = np.abs(res[:k] - reso[:k]) / reso[:k]
    assert rel[:3].max() < 1e-10 and rel[big].max() < 1e-4, rel
    true = np.linalg.norm(b - K.spmv(A, x.numpy()))
    assert true <= 1.001e-8 * res[0]
    # far fewer iterations than without the preconditioner
    x2 = dev(ctx, np.zeros(n))
    _, it_plain = ra.BiCGStab(D, x2, dev(ctx, b), tol=1e-8)
    assert it < it_plain / 3

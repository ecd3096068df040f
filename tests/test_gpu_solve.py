"""GPU parity of the hierarchy, the V-cycle, stationary AMG and PCG
(multilevel/par_multilevel.hpp, krylov/par_cg.cpp) against the oracle on the
product's own hierarchy and against the reference compiled from source."""
import ctypes as C
import math

import numpy as np
import pytest

import checkers as K
import raptor_amd as ra

pytestmark = pytest.mark.gpu
needs_ref = pytest.mark.skipif(not K.have_ref(), reason="oracle/_ref not built")

OMEGA = 2.0 / 3.0


def build(ctx, rp, col, val, coarsen, interp, theta=0.25, max_coarse=50):
    ml = ra.ParRugeStubenSolver(ctx, theta, coarsen, interp, ra.Classical, ra.Jacobi)
    ml.relax_weight = OMEGA
    ml.max_coarse = max_coarse
    ml.setup(rp, col, val)
    return ml


def levels(ml):
    As, Ps = [], []
    for l in range(ml.num_levels):
        A, P = ml.level(l)
        As.append(K.Csr(*A.to_csr(), A.n_cols))
        if P is not None:
            Ps.append(K.Csr(*P.to_csr(), P.n_cols))
    return As, Ps


CONFIGS = [
    # C1: aniso 100^2, CLJP + ModClassical
    ("c1_aniso100", lambda: ra.gallery_aniso2d(100, 0.001, math.pi / 8), ra.CLJP, ra.ModClassical),
    # C2 shape at reduced size: 27-pt, PMIS + Extended
    ("lap27_16_pmis_ext", lambda: ra.gallery_laplace27(16), ra.PMIS, ra.Extended),
    ("aniso60_pmis_direct", lambda: ra.gallery_aniso2d(60, 0.001, math.pi / 8), ra.PMIS, ra.Direct),
]


@pytest.fixture(scope="module", params=CONFIGS, ids=[c[0] for c in CONFIGS])
def problem(request, ctx):
    name, gen, coarsen, interp = request.param
    rp, col, val = gen()
    ml = build(ctx, rp, col, val, coarsen, interp)
    As, Ps = levels(ml)
    return dict(name=name, rp=rp, col=col, val=val, ml=ml, As=As, Ps=Ps, coarsen=coarsen, interp=interp)


def test_hierarchy_galerkin_bit_exact(problem):
    """Every coarse operator equals the oracle's P^T (A P) of the product's own
    A_l and P_l, bit for bit; level 0 is the caller's A sorted, diagonal first."""
    As, Ps = problem["As"], problem["Ps"]
    A0 = K.move_diag(K.sort_rows(K.Csr(problem["rp"], problem["col"], problem["val"])))
    assert np.array_equal(As[0].col, A0.col) and np.array_equal(As[0].val, A0.val)
    for l in range(len(Ps)):
        O = K.galerkin(As[l], Ps[l])
        assert np.array_equal(As[l + 1].rowptr, O.rowptr), l
        assert np.array_equal(As[l + 1].col, O.col), l
        assert np.array_equal(As[l + 1].val, O.val), l
    assert As[-1].n_rows <= 50


@needs_ref
def test_hierarchy_matches_reference(problem):
    """Same number of levels, same C/F splittings, same P and A_l patterns
    (exact) and values (1e-12 relative) as the reference's setup."""
    rp, col, val = problem["rp"], problem["col"], problem["val"]
    L = K.ref()
    rpa, cola, vala = K.i32(rp), K.i32(col), K.f64(val)
    nl = L.ref_hier_setup(rpa.size - 1, K.ip(rpa), K.ip(cola), K.dp(vala), C.c_double(0.25), problem["coarsen"],
                          problem["interp"], 0, 0, C.c_double(OMEGA), 50, 25, None)
    ml, As, Ps = problem["ml"], problem["As"], problem["Ps"]
    assert nl == ml.num_levels
    for l in range(nl):
        rpr, colr, valr, nc = K.ref_fetch(L.ref_hier_level(l, 0))
        assert np.array_equal(As[l].rowptr, rpr) and np.array_equal(As[l].col, colr), l
        assert np.allclose(As[l].val, valr, rtol=1e-12, atol=1e-300), l
        if l < nl - 1:
            rpr, colr, valr, nc = K.ref_fetch(L.ref_hier_level(l, 1))
            assert nc == Ps[l].n_cols
            assert np.array_equal(Ps[l].rowptr, rpr) and np.array_equal(Ps[l].col, colr), l
            assert np.allclose(Ps[l].val, valr, rtol=1e-12, atol=1e-300), l
            st = np.empty(As[l].n_rows, np.int32)
            L.ref_hier_states(l, K.ip(st))
            assert np.array_equal(ml.states(l), st), l


def test_vcycle_bit_exact(problem, ctx):
    ml, As, Ps = problem["ml"], problem["As"], problem["Ps"]
    H = K.Hier(As, Ps, relax_type=0, sweeps=1, omega=OMEGA)
    n = As[0].n_rows
    rng = np.random.default_rng(21)
    b = rng.standard_normal(n)
    # x == 0 on entry (as PCG calls it) and a nonzero initial guess
    x = ra.ParVector.from_numpy(ctx, np.zeros(n))
    ml.cycle(x, ra.ParVector.from_numpy(ctx, b), x_is_zero=True)
    xo = H.cycle(np.zeros(n), b)
    assert np.array_equal(x.numpy(), xo)
    x0 = rng.standard_normal(n)
    x = ra.ParVector.from_numpy(ctx, x0)
    ml.cycle(x, ra.ParVector.from_numpy(ctx, b))
    assert np.array_equal(x.numpy(), H.cycle(x0, b))


def test_amg_solve_matches_oracle(problem, ctx):
    ml, As, Ps = problem["ml"], problem["As"], problem["Ps"]
    H = K.Hier(As, Ps, relax_type=0, sweeps=1, omega=OMEGA)
    A = K.Csr(problem["rp"], problem["col"], problem["val"])
    n = A.n_rows
    b = K.spmv(A, np.ones(n))
    x = ra.ParVector.from_numpy(ctx, np.zeros(n))
    it = ml.solve(x, ra.ParVector.from_numpy(ctx, b))
    xo, reso, ito = H.solve(np.zeros(n), b)
    assert it == ito
    assert np.allclose(ml.residuals, reso, rtol=1e-10, atol=0)


def test_pcg_matches_oracle(problem, ctx):
    """PCG with the reference's rules on the same hierarchy: identical
    iteration count; residual history equal to 1e-8 (the inner products are
    summed in a different, fixed order)."""
    ml, As, Ps = problem["ml"], problem["As"], problem["Ps"]
    H = K.Hier(As, Ps, relax_type=0, sweeps=1, omega=OMEGA)
    A = K.Csr(problem["rp"], problem["col"], problem["val"])
    n = A.n_rows
    b = K.spmv(A, np.ones(n))
    Ad = ra.ParCSRMatrix.from_csr(ml.ctx, A.rowptr, A.col, A.val)
    x = ra.ParVector.from_numpy(ctx, np.zeros(n))
    res, it = ra.PCG(Ad, ml, x, ra.ParVector.from_numpy(ctx, b), tol=1e-8)
    xo, reso, ito = H.pcg(A, np.zeros(n), b, tol=1e-8)
    assert it == ito
    assert len(res) == len(reso)
    assert np.allclose(res, reso, rtol=1e-8, atol=0)
    assert np.allclose(x.numpy(), xo, rtol=1e-9, atol=1e-12)


@pytest.mark.parametrize("prec", [ra.PREC_FP64, ra.PREC_FP32])
def test_fpcg_reaches_true_residual(problem, ctx, prec):
    ml = problem["ml"]
    A = K.Csr(problem["rp"], problem["col"], problem["val"])
    n = A.n_rows
    rng = np.random.default_rng(22)
    xt = rng.uniform(-1, 1, n)
    b = K.spmv(A, xt)
    Ad = ra.ParCSRMatrix.from_csr(ctx, A.rowptr, A.col, A.val)
    x = ra.ParVector.from_numpy(ctx, np.zeros(n))
    hist, it, rel = ra.FPCG(Ad, ml, x, ra.ParVector.from_numpy(ctx, b), 1e-8, 300, prec)
    xs = x.numpy()
    true_rel = np.linalg.norm(b - K.spmv(A, xs)) / np.linalg.norm(b)
    assert rel <= 1e-8 and true_rel <= 1.0000001e-8
    # error against the FP64 oracle solution of the same system
    H = K.Hier(problem["As"], problem["Ps"], relax_type=0, sweeps=1, omega=OMEGA)
    xo, _, _ = H.pcg(A, np.zeros(n), b, tol=1e-14, max_iter=500)
    assert np.linalg.norm(xs - xt) / np.linalg.norm(xt) < 1e-5


def test_pcg_indefinite_reports_error(ctx):
    rp, col, val = ra.gallery_aniso2d(30, 1.0, 0.0)
    ml = build(ctx, rp, col, val, ra.CLJP, ra.ModClassical)
    A = ra.ParCSRMatrix.from_csr(ctx, rp, col, -np.asarray(val))
    n = rp.size - 1
    x = ra.ParVector.from_numpy(ctx, np.zeros(n))
    with pytest.raises(ra.RamdError) as e:
        ra.PCG(A, ml, x, ra.ParVector.from_numpy(ctx, np.ones(n)), tol=1e-8)
    assert e.value.code == -4


def test_fpcg_nan_rhs_reports_error(ctx):
    rp, col, val = ra.gallery_aniso2d(30, 1.0, 0.0)
    ml = build(ctx, rp, col, val, ra.CLJP, ra.ModClassical)
    A = ra.ParCSRMatrix.from_csr(ctx, rp, col, val)
    n = rp.size - 1
    b = np.ones(n)
    b[7] = np.nan
    x = ra.ParVector.from_numpy(ctx, np.zeros(n))
    with pytest.raises(ra.RamdError) as e:
        ra.FPCG(A, ml, x, ra.ParVector.from_numpy(ctx, b), 1e-8, 50)
    assert e.value.code in (-7, -5)

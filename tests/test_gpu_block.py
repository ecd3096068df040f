"""Multi-vector path (k right-hand sides, row-interleaved n x k blocks): the
block V-cycle is bit-identical column by column to the single-vector
V-cycle (FP64 and FP32), the lockstep flexible PCG reaches the true
relative residual in every column with the single-vector iteration counts,
and the captured-graph replay gives exactly the eager result."""
import math

import numpy as np
import pytest

import checkers as K
import raptor_amd as ra

pytestmark = pytest.mark.gpu

OMEGA = 2.0 / 3.0

PROBLEMS = [
    # 9-point anisotropic (stencil-diagonal fine level), C4 shape at reduced size
    ("aniso80_cljp_modclass", lambda: ra.gallery_aniso2d(80, 0.001, math.pi / 8), "rs", ra.CLJP, ra.ModClassical),
    # 27-point, extended interpolation (long coarse rows)
    ("lap27_16_pmis_ext", lambda: ra.gallery_laplace27(16), "rs", ra.PMIS, ra.Extended),
    # smoothed aggregation on the 7-point FV operator
    ("fv7_16_sa", lambda: ra.gallery_fv7(16, seed=3), "sa", None, None),
]


@pytest.fixture(scope="module", params=PROBLEMS, ids=[p[0] for p in PROBLEMS])
def problem(request, ctx):
    name, gen, kind, coarsen, interp = request.param
    rp, col, val = gen()
    if kind == "sa":
        ml = ra.ParSmoothedAggregationSolver(ctx, 0.08)
    else:
        ml = ra.ParRugeStubenSolver(ctx, 0.25, coarsen, interp, ra.Classical, ra.Jacobi)
    ml.relax_weight = OMEGA
    ml.setup(rp, col, val)
    ml.enable_fp32()
    A = ra.ParCSRMatrix.from_csr(ctx, rp, col, val)
    return dict(rp=rp, col=col, val=val, n=rp.size - 1, ml=ml, A=A)


def oracle_hier(ml):
    As, Ps = [], []
    for l in range(ml.num_levels):
        Al, Pl = ml.level(l)
        As.append(K.Csr(*Al.to_csr(), Al.n_cols))
        if Pl is not None:
            Ps.append(K.Csr(*Pl.to_csr(), Pl.n_cols))
    return K.Hier(As, Ps, relax_type=0, sweeps=1, omega=OMEGA)


@pytest.mark.parametrize("k", [4, 32])
def test_block_vcycle_columns_equal_oracle(problem, ctx, k):
    """Every column of the block V-cycle (FP64 and FP32) equals the oracle's
    cycle (multilevel/par_multilevel.hpp:341-457; its single-precision
    sequence for FP32) applied to that column on the same hierarchy, bit for
    bit."""
    ml, n = problem["ml"], problem["n"]
    H = oracle_hier(ml)
    B = np.random.default_rng(100 + k).standard_normal((n, k))
    X = ra.ParVector(ctx, n * k)
    ra.vcycle_block(ml, X, ra.ParVector.from_numpy(ctx, B.ravel()), k, ra.PREC_FP64)
    Xh = X.numpy().reshape(n, k)
    B32 = B.astype(np.float32)
    X32 = ra.ParVector(ctx, n * k, np.float32)
    ra.vcycle_block(ml, X32, ra.ParVector.from_numpy(ctx, B32.ravel()), k, ra.PREC_FP32)
    X32h = X32.numpy().reshape(n, k)
    for c in range(k):
        assert np.array_equal(Xh[:, c], H.cycle(np.zeros(n), B[:, c])), ("fp64", c)
        assert np.array_equal(X32h[:, c], H.cycle_fp32(np.zeros(n, np.float32), B32[:, c])), ("fp32", c)


@pytest.mark.parametrize("sweeps", [0, 2])
def test_block_vcycle_sweep_counts(ctx, sweeps):
    """Block V-cycles with no smoothing (each level's iterate starts from the
    zero the reference sets) and with two sweeps: every column equals the
    oracle's cycle, FP64 and FP32, bit for bit."""
    rp, col, val = ra.gallery_laplace27(12)
    ml = ra.ParRugeStubenSolver(ctx, 0.25, ra.PMIS, ra.Extended, ra.Classical, ra.Jacobi)
    ml.relax_weight = OMEGA
    ml.num_smooth_sweeps = sweeps
    ml.setup(rp, col, val)
    ml.enable_fp32()
    n, k = rp.size - 1, 3
    As, Ps = [], []
    for l in range(ml.num_levels):
        Al, Pl = ml.level(l)
        As.append(K.Csr(*Al.to_csr(), Al.n_cols))
        if Pl is not None:
            Ps.append(K.Csr(*Pl.to_csr(), Pl.n_cols))
    H = K.Hier(As, Ps, relax_type=0, sweeps=sweeps, omega=OMEGA)
    B = np.random.default_rng(40 + sweeps).standard_normal((n, k))
    X = ra.ParVector.from_numpy(ctx, np.full(n * k, 5.0))
    ra.vcycle_block(ml, X, ra.ParVector.from_numpy(ctx, B.ravel()), k, ra.PREC_FP64)
    B32 = B.astype(np.float32)
    X32 = ra.ParVector.from_numpy(ctx, np.full(n * k, 5.0, np.float32))
    ra.vcycle_block(ml, X32, ra.ParVector.from_numpy(ctx, B32.ravel()), k, ra.PREC_FP32)
    Xh, X32h = X.numpy().reshape(n, k), X32.numpy().reshape(n, k)
    for c in range(k):
        assert np.array_equal(Xh[:, c], H.cycle(np.zeros(n), B[:, c])), ("fp64", c)
        assert np.array_equal(X32h[:, c], H.cycle_fp32(np.zeros(n, np.float32), B32[:, c])), ("fp32", c)


@pytest.mark.parametrize("k", [1, 5, 32])
def test_block_vcycle_columns_bit_exact(problem, ctx, k):
    ml, n = problem["ml"], problem["n"]
    B = np.random.default_rng(k).standard_normal((n, k))
    # FP64
    X = ra.ParVector(ctx, n * k)
    ra.vcycle_block(ml, X, ra.ParVector.from_numpy(ctx, B.ravel()), k, ra.PREC_FP64)
    Xh = X.numpy().reshape(n, k)
    for c in range(k):
        x = ra.ParVector.from_numpy(ctx, np.zeros(n))
        ml.cycle(x, ra.ParVector.from_numpy(ctx, B[:, c].copy()), x_is_zero=True)
        assert np.array_equal(Xh[:, c], x.numpy()), ("fp64", c)
    # FP32
    B32 = B.astype(np.float32)
    X32 = ra.ParVector(ctx, n * k, np.float32)
    ra.vcycle_block(ml, X32, ra.ParVector.from_numpy(ctx, B32.ravel()), k, ra.PREC_FP32)
    X32h = X32.numpy().reshape(n, k)
    from raptor_amd import _lib
    for c in range(k):
        x = ra.ParVector(ctx, n, np.float32)
        b = ra.ParVector.from_numpy(ctx, B32[:, c].copy())
        ml.cycle_fp32(x, b)
        assert np.array_equal(X32h[:, c], x.numpy()), ("fp32", c)


@pytest.mark.parametrize("kind", ["scalar", "units"])
def test_block_vcycle_kernel_variants_bit_exact(problem, ctx, kind):
    """k = 32 with every operator forced onto the one-column-per-lane-step
    kernel or onto the 16-byte-unit kernel: columns still bit-identical to
    the single-vector cycles (FP64 and FP32).  The tuned choice is restored."""
    ml, n = problem["ml"], problem["n"]
    k = 32
    B = np.random.default_rng(41).standard_normal((n, k))
    from raptor_amd import _lib
    for prec, units_lpr in ((ra.PREC_FP64, -8), (ra.PREC_FP32, -4)):
        tuned = ra.block_tuning(ml, k, prec)
        codes = np.zeros_like(tuned)
        codes[:-1] = 32 if kind == "scalar" else units_lpr
        ra.block_set_tuning(ml, k, codes, prec)
        try:
            if prec == ra.PREC_FP64:
                X = ra.ParVector(ctx, n * k)
                ra.vcycle_block(ml, X, ra.ParVector.from_numpy(ctx, B.ravel()), k, prec)
                Xh = X.numpy().reshape(n, k)
                for c in (0, 13, 31):
                    x = ra.ParVector.from_numpy(ctx, np.zeros(n))
                    bc = ra.ParVector.from_numpy(ctx, B[:, c].copy())
                    ml.cycle(x, bc, x_is_zero=True)
                    assert np.array_equal(Xh[:, c], x.numpy()), (kind, "fp64", c)
            else:
                B32 = B.astype(np.float32)
                X32 = ra.ParVector(ctx, n * k, np.float32)
                ra.vcycle_block(ml, X32, ra.ParVector.from_numpy(ctx, B32.ravel()), k, prec)
                X32h = X32.numpy().reshape(n, k)
                for c in (0, 13, 31):
                    x = ra.ParVector(ctx, n, np.float32)
                    bc = ra.ParVector.from_numpy(ctx, B32[:, c].copy())
                    ml.cycle_fp32(x, bc)
                    assert np.array_equal(X32h[:, c], x.numpy()), (kind, "fp32", c)
        finally:
            ra.block_set_tuning(ml, k, tuned, prec)
        assert np.array_equal(ra.block_tuning(ml, k, prec), tuned)


@pytest.mark.parametrize("prec", [ra.PREC_FP64, ra.PREC_FP32])
def test_block_tuning_choices(problem, prec):
    """k = 32: every operator got a kernel the launcher knows (lanes per row of
    the scalar kernel, or minus those of the 16-byte-unit kernel)."""
    t = ra.block_tuning(problem["ml"], 32, prec)
    assert t.shape == (problem["ml"].num_levels, 3)
    vec = {-4, -8, -16}
    assert set(np.unique(t[:-1])) <= {8, 16, 32} | vec
    assert np.all(t[-1] == 0)


@pytest.mark.parametrize("prec", [ra.PREC_FP64, ra.PREC_FP32])
def test_block_fpcg_all_columns_converge(problem, ctx, prec):
    ml, A, n = problem["ml"], problem["A"], problem["n"]
    k = 8
    Ah = K.Csr(problem["rp"], problem["col"], problem["val"])
    B = np.random.default_rng(100).standard_normal((n, k))
    X = ra.ParVector.from_numpy(ctx, np.zeros(n * k))
    its, rel, tot = ra.FPCG_block(A, ml, X, ra.ParVector.from_numpy(ctx, B.ravel()), k, 1e-8, 300, prec)
    Xh = X.numpy().reshape(n, k)
    assert tot == its.max()
    for c in range(k):
        true = np.linalg.norm(B[:, c] - K.spmv(Ah, Xh[:, c])) / np.linalg.norm(B[:, c])
        assert rel[c] <= 1e-8 and true <= 1.0000001e-8, (c, rel[c], true)
        x = ra.ParVector.from_numpy(ctx, np.zeros(n))
        _, it1, _ = ra.FPCG(A, ml, x, ra.ParVector.from_numpy(ctx, B[:, c].copy()), 1e-8, 300, prec)
        assert abs(int(its[c]) - it1) <= 1, (c, its[c], it1)


def test_block_fpcg_graph_equals_eager(problem, ctx):
    ml, A, n = problem["ml"], problem["A"], problem["n"]
    k = 4
    B = np.random.default_rng(7).standard_normal((n, k)).ravel()
    out = []
    for use_graph in (False, True, True):  # the second graph call replays the cached graphs
        X = ra.ParVector.from_numpy(ctx, np.zeros(n * k))
        its, rel, tot = ra.FPCG_block(A, ml, X, ra.ParVector.from_numpy(ctx, B), k, 1e-8, 300, ra.PREC_FP32,
                                      use_graph=use_graph)
        out.append((X.numpy(), its, rel))
    for X, its, rel in out[1:]:
        assert np.array_equal(X, out[0][0]) and np.array_equal(its, out[0][1]) and np.array_equal(rel, out[0][2])

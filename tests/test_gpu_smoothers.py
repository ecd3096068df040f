"""Multicolour SOR / SSOR smoothers (this build's labelled non-reference
smoothers; row updates of util/linalg/par_relax.cpp:29-104 in greedy colour
order) on the device: V-cycles bit-identical to the oracle's multicolour
cycle, FPCG convergence, save/load and the multi-vector guard."""
import math

import numpy as np
import pytest

import checkers as K
import raptor_amd as ra

pytestmark = pytest.mark.gpu

PROBLEMS = [
    ("aniso100_cljp_modclass", lambda: ra.gallery_aniso2d(100, 0.001, math.pi / 8), "rs", ra.CLJP, ra.ModClassical),
    ("lap27_14_pmis_ext", lambda: ra.gallery_laplace27(14), "rs", ra.PMIS, ra.Extended),
    ("fv7_16_sa", lambda: ra.gallery_fv7(16, seed=3), "sa", None, None),
]


def build(ctx, gen, kind, coarsen, interp, relax, omega):
    rp, col, val = gen()
    if kind == "sa":
        ml = ra.ParSmoothedAggregationSolver(ctx, 0.08, relax_type=relax)
    else:
        ml = ra.ParRugeStubenSolver(ctx, 0.25, coarsen, interp, ra.Classical, relax)
    ml.relax_weight = omega
    ml.setup(rp, col, val)
    As, Ps = [], []
    for l in range(ml.num_levels):
        A, P = ml.level(l)
        As.append(K.Csr(*A.to_csr(), A.n_cols))
        if P is not None:
            Ps.append(K.Csr(*P.to_csr(), P.n_cols))
    return rp, col, val, ml, As, Ps


@pytest.mark.parametrize("relax,omega", [(ra.MC_SOR, 1.0), (ra.MC_SSOR, 1.0), (ra.MC_SOR, 0.8)])
@pytest.mark.parametrize("name,gen,kind,coarsen,interp", PROBLEMS, ids=[p[0] for p in PROBLEMS])
def test_vcycle_multicolour_bit_exact(ctx, name, gen, kind, coarsen, interp, relax, omega):
    rp, col, val, ml, As, Ps = build(ctx, gen, kind, coarsen, interp, relax, omega)
    H = K.Hier(As, Ps, relax_type=relax, sweeps=1, omega=omega)
    n = As[0].n_rows
    rng = np.random.default_rng(51)
    b = rng.standard_normal(n)
    x = ra.ParVector.from_numpy(ctx, np.zeros(n))
    bd = ra.ParVector.from_numpy(ctx, b)
    ml.cycle(x, bd, x_is_zero=True)
    assert np.array_equal(x.numpy(), H.cycle(np.zeros(n), b))
    x0 = rng.standard_normal(n)
    x = ra.ParVector.from_numpy(ctx, x0)
    ml.cycle(x, bd)
    assert np.array_equal(x.numpy(), H.cycle(x0, b))


@pytest.mark.parametrize("prec", [ra.PREC_FP64, ra.PREC_FP32])
def test_fpcg_with_multicolour_ssor(ctx, prec):
    """MC-SSOR with w = 1 (a symmetric smoother) preconditions FPCG to the true
    1e-8 residual, in fewer iterations than Jacobi 2/3 on the same problem."""
    gen = lambda: ra.gallery_aniso2d(100, 0.001, math.pi / 8)
    rp, col, val, ml, As, Ps = build(ctx, gen, "rs", ra.CLJP, ra.ModClassical, ra.MC_SSOR, 1.0)
    _, _, _, mlj, _, _ = build(ctx, gen, "rs", ra.CLJP, ra.ModClassical, ra.Jacobi, 2.0 / 3.0)
    A = ra.ParCSRMatrix.from_csr(ctx, rp, col, val)
    n = rp.size - 1
    b = np.random.default_rng(52).standard_normal(n)
    bd = ra.ParVector.from_numpy(ctx, b)
    its = []
    for m in (ml, mlj):
        x = ra.ParVector.from_numpy(ctx, np.zeros(n))
        _, it, rel = ra.FPCG(A, m, x, bd, 1e-8, 300, prec)
        true = np.linalg.norm(b - K.spmv(K.Csr(rp, col, val), x.numpy())) / np.linalg.norm(b)
        assert rel <= 1e-8 and true <= 1.000001e-8
        its.append(it)
    assert its[0] < its[1], its


def test_multicolour_save_load_and_block_guard(ctx, tmp_path):
    gen = lambda: ra.gallery_aniso2d(60, 0.001, math.pi / 8)
    rp, col, val, ml, As, Ps = build(ctx, gen, "rs", ra.PMIS, ra.Direct, ra.MC_SSOR, 1.0)
    path = tmp_path / "mc.bin"
    ml.save(path)
    ml2 = type(ml).load(ctx, path)
    n = As[0].n_rows
    b = ra.ParVector.from_numpy(ctx, np.random.default_rng(53).standard_normal(n))
    x1 = ra.ParVector.from_numpy(ctx, np.zeros(n))
    x2 = ra.ParVector.from_numpy(ctx, np.zeros(n))
    ml.cycle(x1, b, x_is_zero=True)
    ml2.cycle(x2, b, x_is_zero=True)
    assert np.array_equal(x1.numpy(), x2.numpy())
    X = ra.ParVector(ctx, n * 4)
    with pytest.raises(ra.RamdError):
        ra.vcycle_block(ml, X, ra.ParVector.from_numpy(ctx, np.ones(n * 4)), 4)

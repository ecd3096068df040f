"""The batching workload at its full size: rotated anisotropic 9-point
diffusion (eps 0.001, theta pi/8) on 2048^2 (4.19M rows), RS(CLJP,
modified classical, theta 0.25), Jacobi 2/3, 32 right-hand sides
b_c ~ N(0,1) from seeds 100..131, lockstep block flexible PCG with the
captured HIP graphs.

Checked against the reference's per-column algorithm (PCG,
krylov/par_cg.cpp:121-263, one column at a time; the cycle
multilevel/par_multilevel.hpp:341-457):
  - every column's true relative residual, recomputed on the host by the
    oracle's SpMV, is <= 1e-8 (FP32 and FP64 cycles);
  - the graph replay gives exactly the eager result (X, per-column
    iteration counts and residuals);
  - each column's iteration count is within 1 of the single-vector FPCG's;
  - columns of one block V-cycle equal the oracle's cycle (FP64 and its
    single-precision sequence) on the same hierarchy, bit for bit."""
import math

import numpy as np
import pytest

import checkers as K
import raptor_amd as ra
from raptor_amd import synth

pytestmark = pytest.mark.gpu

OMEGA = 2.0 / 3.0
M, KRHS = 2048, 32


@pytest.fixture(scope="module")
def c4(ctx):
    rp, col, val = ra.gallery_aniso2d(M, 0.001, math.pi / 8)
    ml = ra.ParRugeStubenSolver(ctx, 0.25, ra.CLJP, ra.ModClassical, ra.Classical, ra.Jacobi)
    ml.relax_weight = OMEGA
    ml.setup(rp, col, val)
    ml.enable_fp32()
    A = ra.ParCSRMatrix.from_csr(ctx, rp, col, val)
    n = rp.size - 1
    B = np.stack([synth.normal(100 + c, n) for c in range(KRHS)], axis=1)
    return dict(ml=ml, A=A, Ah=K.Csr(rp, col, val), n=n, B=B, Bd=ra.ParVector.from_numpy(ctx, B.ravel()))


def block_solve(ctx, c4, prec, use_graph):
    n = c4["n"]
    X = ra.ParVector(ctx, n * KRHS)
    X.set_const_value(0.0)
    its, rel, tot = ra.FPCG_block(c4["A"], c4["ml"], X, c4["Bd"], KRHS, 1e-8, 500, prec, use_graph=use_graph)
    return X.numpy().reshape(n, KRHS), its, rel, tot


@pytest.mark.parametrize("prec", [ra.PREC_FP32, ra.PREC_FP64], ids=["fp32", "fp64"])
def test_c4_block_fpcg_true_residuals(ctx, c4, prec):
    X, its, rel, tot = block_solve(ctx, c4, prec, True)
    assert tot == its.max()
    B, Ah = c4["B"], c4["Ah"]
    for c in range(KRHS):
        true = np.linalg.norm(B[:, c] - K.spmv(Ah, X[:, c])) / np.linalg.norm(B[:, c])
        assert rel[c] <= 1e-8 and true <= 1.0000001e-8, (c, rel[c], true)


def test_c4_graph_replay_equals_eager(ctx, c4):
    """Eager, first graph call (capture + replay) and second graph call
    (cached graphs): identical bits."""
    outs = [block_solve(ctx, c4, ra.PREC_FP32, g) for g in (False, True, True)]
    X0, its0, rel0, tot0 = outs[0]
    for X, its, rel, tot in outs[1:]:
        assert tot == tot0
        assert np.array_equal(its, its0) and np.array_equal(rel, rel0)
        assert np.array_equal(X, X0)


def test_c4_iterations_match_single_vector(ctx, c4):
    _, its, _, _ = block_solve(ctx, c4, ra.PREC_FP32, True)
    n, B = c4["n"], c4["B"]
    x = ra.ParVector(ctx, n)
    for c in range(KRHS):
        x.set_const_value(0.0)
        _, it1, rel1 = ra.FPCG(c4["A"], c4["ml"], x, ra.ParVector.from_numpy(ctx, B[:, c].copy()), 1e-8, 500,
                               ra.PREC_FP32)
        assert rel1 <= 1e-8
        assert abs(int(its[c]) - it1) <= 1, (c, its[c], it1)


def test_c4_block_vcycle_columns_equal_oracle(ctx, c4):
    ml, n, B = c4["ml"], c4["n"], c4["B"]
    As, Ps = [], []
    for l in range(ml.num_levels):
        Al, Pl = ml.level(l)
        As.append(K.Csr(*Al.to_csr(), Al.n_cols))
        if Pl is not None:
            Ps.append(K.Csr(*Pl.to_csr(), Pl.n_cols))
    H = K.Hier(As, Ps, relax_type=0, sweeps=1, omega=OMEGA)
    X = ra.ParVector(ctx, n * KRHS)
    ra.vcycle_block(ml, X, c4["Bd"], KRHS, ra.PREC_FP64)
    Xh = X.numpy().reshape(n, KRHS)
    B32 = B.astype(np.float32)
    X32 = ra.ParVector(ctx, n * KRHS, np.float32)
    ra.vcycle_block(ml, X32, ra.ParVector.from_numpy(ctx, B32.ravel()), KRHS, ra.PREC_FP32)
    X32h = X32.numpy().reshape(n, KRHS)
    for c in (0, 17, KRHS - 1):
        assert np.array_equal(Xh[:, c], H.cycle(np.zeros(n), B[:, c])), ("fp64", c)
        assert np.array_equal(X32h[:, c], H.cycle_fp32(np.zeros(n, np.float32), np.ascontiguousarray(B32[:, c]))), \
            ("fp32", c)

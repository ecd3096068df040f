"""Parity at the benchmark sizes.

27-point Laplacian 128^3 (2.1M rows, RS(PMIS, extended), theta 0.25): the
device-built hierarchy equals the host-built one level by level, and the
device V-cycle (FP64 and FP32) is bit-identical to the oracle's cycle on
that hierarchy.

27-point Laplacian 256^3 (16.8M rows): properties that do not depend on an
oracle run at this size -- the FP32-cycle and FP64-cycle flexible PCG both
reach a true FP64 relative residual of 1e-8 in the same number of
iterations, their solutions agree to the solver tolerance, and the V-cycle
is exactly linear under scaling by 2 (every operation of the cycle scales
exactly by a power of two)."""
import numpy as np
import pytest

import checkers as K
import raptor_amd as ra
from raptor_amd import synth

pytestmark = pytest.mark.gpu

OMEGA = 2.0 / 3.0


def rs_pmis_ext(ctx, rp, col, val, device_setup=True):
    ml = ra.ParRugeStubenSolver(ctx, 0.25, ra.PMIS, ra.Extended, ra.Classical, ra.Jacobi)
    ml.relax_weight = OMEGA
    ml.device_setup = device_setup
    ml.setup(rp, col, val)
    return ml


@pytest.fixture(scope="module")
def c2(ctx):
    rp, col, val = ra.gallery_laplace27(128)
    ml = rs_pmis_ext(ctx, rp, col, val)
    return dict(rp=rp, col=col, val=val, ml=ml)


def test_c2_device_hierarchy_equals_host(ctx, c2):
    """Every level's A and P and every splitting at 128^3, device vs host setup."""
    a = c2["ml"]
    b = rs_pmis_ext(ctx, c2["rp"], c2["col"], c2["val"], device_setup=False)
    assert a.num_levels == b.num_levels >= 4
    for l in range(a.num_levels):
        for M1, M2 in zip(a.level(l), b.level(l)):
            if M1 is None:
                assert M2 is None
                continue
            assert all(np.array_equal(u, v) for u, v in zip(M1.to_csr(), M2.to_csr())), l
        if l < a.num_levels - 1:
            assert np.array_equal(a.states(l), b.states(l)), l


def test_c2_vcycle_bit_exact(ctx, c2):
    """FP64 and FP32 V-cycles at 128^3 equal the oracle's cycles on the same levels."""
    ml = c2["ml"]
    As, Ps = [], []
    for l in range(ml.num_levels):
        A, P = ml.level(l)
        As.append(K.Csr(*A.to_csr(), A.n_cols))
        if P is not None:
            Ps.append(K.Csr(*P.to_csr(), P.n_cols))
    H = K.Hier(As, Ps, relax_type=0, sweeps=1, omega=OMEGA)
    n = As[0].n_rows
    b = synth.normal(7, n)
    x = ra.ParVector(ctx, n)
    ml.cycle(x, ra.ParVector.from_numpy(ctx, b), x_is_zero=True)
    assert np.array_equal(x.numpy(), H.cycle(np.zeros(n), b))
    ml.enable_fp32()
    b32 = b.astype(np.float32)
    x32 = ra.ParVector(ctx, n, np.float32)
    ml.cycle_fp32(x32, ra.ParVector.from_numpy(ctx, b32))
    assert np.array_equal(x32.numpy(), H.cycle_fp32(np.zeros(n, np.float32), b32))


@pytest.fixture(scope="module")
def c3(ctx):
    rp, col, val = ra.gallery_laplace27(256)
    ml = rs_pmis_ext(ctx, rp, col, val)
    A = ra.ParCSRMatrix.from_csr(ctx, rp, col, val)
    del rp, col, val
    return dict(ml=ml, A=A, n=A.n_rows)


def true_relres(ctx, A, x, b_h):
    r = ra.ParVector(ctx, A.n_rows)
    A.residual(x, ra.ParVector.from_numpy(ctx, b_h), r)
    return np.linalg.norm(r.numpy()) / np.linalg.norm(b_h)


def test_c3_fpcg_fp32_and_fp64_cycles(ctx, c3):
    """256^3, b ~ N(0,1) seed 2: both forms reach 1e-8 (true FP64 residual,
    recomputed here) in the same iteration count; solutions agree."""
    ml, A, n = c3["ml"], c3["A"], c3["n"]
    b_h = synth.normal(2, n)
    b = ra.ParVector.from_numpy(ctx, b_h)
    xs, its = [], []
    for prec in (ra.PREC_FP32, ra.PREC_FP64):
        x = ra.ParVector(ctx, n)
        _, it, rel = ra.FPCG(A, ml, x, b, 1e-8, 100, prec)
        assert rel <= 1e-8
        assert true_relres(ctx, A, x, b_h) <= 1e-8
        xs.append(x.numpy())
        its.append(it)
    assert its[0] == its[1] and its[0] <= 20
    assert np.linalg.norm(xs[0] - xs[1]) <= 1e-6 * np.linalg.norm(xs[1])


def test_c3_vcycle_scales_exactly(ctx, c3):
    """cycle(2 b) == 2 cycle(b) bit for bit, FP64 and FP32, at 256^3."""
    ml, n = c3["ml"], c3["n"]
    b_h = synth.normal(11, n)
    y1, y2 = ra.ParVector(ctx, n), ra.ParVector(ctx, n)
    ml.cycle(y1, ra.ParVector.from_numpy(ctx, b_h), x_is_zero=True)
    ml.cycle(y2, ra.ParVector.from_numpy(ctx, 2.0 * b_h), x_is_zero=True)
    assert np.array_equal(2.0 * y1.numpy(), y2.numpy())
    ml.enable_fp32()
    b32 = b_h.astype(np.float32)
    z1, z2 = ra.ParVector(ctx, n, np.float32), ra.ParVector(ctx, n, np.float32)
    ml.cycle_fp32(z1, ra.ParVector.from_numpy(ctx, b32))
    ml.cycle_fp32(z2, ra.ParVector.from_numpy(ctx, np.float32(2.0) * b32))
    assert np.array_equal(np.float32(2.0) * z1.numpy(), z2.numpy())

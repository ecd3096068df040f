"""ramd_fpcg_noalloc (the flexible PCG that allocates no device memory)
against ramd_fpcg, the entry point it is a variant of, through the C ABI:
two calls in a row, the second on a smaller system, each compared with
ramd_fpcg's result on the same hierarchy and operator bit for bit (same
kernels, same arguments, fixed-order reductions).  Also its refusals: an
FP32 cycle before the hierarchy's FP32 copies exist, and an operator with
the stencil-diagonal form that no ramd_fpcg call has prepared."""
import numpy as np
import pytest

import raptor_amd as ra
from raptor_amd import synth

pytestmark = pytest.mark.gpu


def system(ctx, m, seed):
    rp, col, val = ra.gallery_laplace27(m)
    ml = ra.ParRugeStubenSolver(ctx, 0.25, ra.PMIS, ra.Extended, ra.Classical, ra.Jacobi)
    ml.relax_weight = 2.0 / 3.0
    ml.setup(rp, col, val)
    A = ra.ParCSRMatrix.from_csr(ctx, rp, col, val)
    b = ra.ParVector.from_numpy(ctx, synth.normal(seed, rp.size - 1))
    return ml, A, b


def test_noalloc_equals_fpcg_two_calls(ctx):
    cases = [(40, 21, ra.PREC_FP32), (20, 22, ra.PREC_FP64)]  # the second system is smaller
    systems, refs = [], []
    for m, seed, prec in cases:
        ml, A, b = system(ctx, m, seed)
        if prec == ra.PREC_FP32:
            ml.enable_fp32()
        x = ra.ParVector(ctx, A.n_rows)
        hist, it, rel = ra.FPCG(A, ml, x, b, 1e-8, 200, prec)
        assert rel <= 1e-8
        systems.append((ml, A, b, prec))
        refs.append((x.numpy(), hist, it, rel))
    outs = []
    for ml, A, b, prec in systems:  # two no-allocation calls in a row
        x = ra.ParVector(ctx, A.n_rows)
        hist, it, rel = ra.FPCG_noalloc(A, ml, x, b, 1e-8, 200, prec)
        outs.append((x.numpy(), hist, it, rel))
    for (xr, hr, itr, relr), (xn, hn, itn, reln) in zip(refs, outs):
        assert itn == itr and reln == relr
        assert np.array_equal(hn, hr)
        assert np.array_equal(xn, xr)


def test_noalloc_refuses_unprepared(ctx):
    ml, A, b = system(ctx, 16, 23)
    x = ra.ParVector(ctx, A.n_rows)
    with pytest.raises(ra.RamdError):  # FP32 copies of the hierarchy not built
        ra.FPCG_noalloc(A, ml, x, b, 1e-8, 50, ra.PREC_FP32)
    if A.diagonals():
        with pytest.raises(ra.RamdError):  # the operator was never prepared by ramd_fpcg
            ra.FPCG_noalloc(A, ml, x, b, 1e-8, 50, ra.PREC_FP64)
    ra.FPCG(A, ml, x, b, 1e-8, 50, ra.PREC_FP64)
    x.set_const_value(0.0)
    _, it, rel = ra.FPCG_noalloc(A, ml, x, b, 1e-8, 50, ra.PREC_FP64)
    assert rel <= 1e-8 and it > 0

"""Parity at the benchmark sizes.

27-point Laplacian 128^3 (2.1M rows, RS(PMIS or HMIS, extended), theta
0.25): the device-built hierarchy equals the host-built one level by level
and matches the fingerprints of the reference's own setup of the same
problem (tests/golden/ref_hier_27pt128.json: pattern, splitting and value
hashes per level), and the device V-cycle (FP64 and FP32) is bit-identical
to the oracle's cycle on that hierarchy.

27-point Laplacian 256^3 (16.8M rows): properties that do not depend on an
oracle run at this size -- the FP32-cycle and FP64-cycle flexible PCG both
reach a true FP64 relative residual of 1e-8 in the same number of
iterations, their solutions agree to the solver tolerance, and the V-cycle
is exactly linear under scaling by 2 (every operation of the cycle scales
exactly by a power of two)."""
import hashlib
import json
import os

import numpy as np
import pytest

import checkers as K
import raptor_amd as ra
from raptor_amd import synth

pytestmark = pytest.mark.gpu

OMEGA = 2.0 / 3.0


def rs_pmis_ext(ctx, rp, col, val, device_setup=True, coarsen=ra.PMIS):
    ml = ra.ParRugeStubenSolver(ctx, 0.25, coarsen, ra.Extended, ra.Classical, ra.Jacobi)
    ml.relax_weight = OMEGA
    ml.device_setup = device_setup
    ml.setup(rp, col, val)
    return ml


@pytest.fixture(scope="module")
def c2(ctx):
    rp, col, val = ra.gallery_laplace27(128)
    ml = rs_pmis_ext(ctx, rp, col, val)
    return dict(rp=rp, col=col, val=val, ml=ml)


@pytest.fixture(scope="module")
def c2_hmis(ctx, c2):
    return rs_pmis_ext(ctx, c2["rp"], c2["col"], c2["val"], coarsen=ra.HMIS)


def fingerprint(rp, col, val):
    """The fingerprint tests/golden/make_ref_hier_fingerprints.py records."""
    h = hashlib.sha256()
    h.update(np.asarray(rp, np.int32).astype("<i4").tobytes())
    h.update(np.asarray(col, np.int32).astype("<i4").tobytes())
    val = np.asarray(val, np.float64)
    w = (np.arange(val.size) % 97 + 1).astype(np.float64)
    return dict(n_rows=int(len(rp) - 1), nnz=int(len(col)), pattern_sha256=h.hexdigest(),
                sum=float(np.sum(val)), abs_sum=float(np.sum(np.abs(val))), weighted_sum=float(np.dot(w, val)))


def assert_matches_reference(ml, ref):
    """Levels, patterns and splittings exact; value checksums to 1e-12."""
    assert ml.num_levels == len(ref["levels"])
    for l, lev in enumerate(ref["levels"]):
        A, P = ml.level(l)
        mats = [(A, lev["A"])] + ([(P, lev["P"])] if "P" in lev else [])
        assert (P is None) == ("P" not in lev), l
        for M, want in mats:
            got = fingerprint(*M.to_csr())
            for key in ("n_rows", "nnz", "pattern_sha256"):
                assert got[key] == want[key], (l, key)
            for key in ("sum", "abs_sum", "weighted_sum"):
                assert abs(got[key] - want[key]) <= 1e-12 * max(abs(want[key]), want["abs_sum"]), (l, key)
        if "P" in lev:
            assert P.n_cols == lev["P"]["n_cols"]
            st = ml.states(l)
            assert hashlib.sha256(st.astype("<i4").tobytes()).hexdigest() == lev["states_sha256"], l


REF_FP = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_hier_27pt128.json")


@pytest.mark.parametrize("name", ["pmis", "hmis"])
def test_c2_hierarchy_matches_reference_fingerprints(ctx, c2, c2_hmis, name):
    """128^3 RS(PMIS / HMIS, extended) hierarchies against the reference's
    own setup (ruge_stuben/par_ruge_stuben_solver.hpp:57-162)."""
    with open(REF_FP) as f:
        ref = json.load(f)
    assert ref["n"] == c2["rp"].size - 1
    assert_matches_reference(c2["ml"] if name == "pmis" else c2_hmis, ref["hierarchies"][name])


@pytest.mark.parametrize("coarsen", [ra.PMIS, ra.HMIS], ids=["pmis", "hmis"])
def test_c2_device_hierarchy_equals_host(ctx, c2, c2_hmis, coarsen):
    """Every level's A and P and every splitting at 128^3, device vs host setup."""
    a = c2["ml"] if coarsen == ra.PMIS else c2_hmis
    b = rs_pmis_ext(ctx, c2["rp"], c2["col"], c2["val"], device_setup=False, coarsen=coarsen)
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

"""The differentiation workload at its full size: 7-point finite-volume
diffusion on 96^3 (884,736 rows) with log-normal cell coefficients (seed 3),
smoothed aggregation (symmetric strength 0.08, MIS-2, Jacobi prolongation
4/3), Jacobi 2/3, L = 0.5 ||x - x*||^2 with x* ~ N(0,1) (seed 4), b ~ N(0,1)
(seed 6).

  - dL/dA on A's pattern against central differences (h = 1e-6, FP64) on 32
    random stored entries (seed 5); each perturbed solve is done by
    refinement with the unperturbed hierarchy, so dL follows from the
    perturbation delta as (x - x*).delta + |delta|^2 / 2;
  - lambda = dL/db against the oracle's FP64 PCG (krylov/par_cg.cpp:121-263)
    for A^T lambda = x - x* on the same hierarchy;
  - the FP32-cycle form gives the same gradient to 1e-8."""
import math

import numpy as np
import pytest

import checkers as K
import raptor_amd as ra
from raptor_amd import synth

pytestmark = pytest.mark.gpu

M = 96


@pytest.fixture(scope="module")
def c5(ctx):
    rp, col, val = ra.gallery_fv7(M, seed=3)
    ml = ra.ParSmoothedAggregationSolver(ctx, 0.08)
    ml.relax_weight = 2.0 / 3.0
    ml.setup(rp, col, val)
    A = ra.ParCSRMatrix.from_csr(ctx, rp, col, val)
    n = rp.size - 1
    b_h = synth.normal(6, n)
    xs_h = synth.normal(4, n)
    out = {}
    for prec in (ra.PREC_FP64, ra.PREC_FP32):
        x = ra.ParVector(ctx, n)
        _, itf, rel = ra.FPCG(A, ml, x, ra.ParVector.from_numpy(ctx, b_h), 1e-10, 500, prec)
        g = ra.ParVector.from_numpy(ctx, x.numpy() - xs_h)
        dA, db, ita = ra.adjoint(A, ml, x, g, rtol=1e-10, cycle_prec=prec)
        out[prec] = dict(x=x.numpy(), dA=dA.numpy(), db=db.numpy(), itf=itf, ita=ita, rel=rel)
    return dict(rp=rp, col=col, val=val, ml=ml, A=A, n=n, b=b_h, xs=xs_h, out=out)


def test_c5_gradient_central_differences(ctx, c5):
    rp, col, val, n = c5["rp"], c5["col"], c5["val"], c5["n"]
    ml, A = c5["ml"], c5["A"]
    o = c5["out"][ra.PREC_FP64]
    xh, gA = o["x"], o["dA"]
    assert o["rel"] <= 1e-10
    h = 1e-6
    rng = np.random.default_rng(5)
    rows = np.repeat(np.arange(n), np.diff(rp))
    Ap = ra.ParCSRMatrix.from_csr(ctx, rp, col, val)
    vdev = ra.ParVector(ctx, col.size)
    r, z, d = ra.ParVector(ctx, n), ra.ParVector(ctx, n), ra.ParVector(ctx, n)
    e = xh - c5["xs"]
    worst = 0.0
    for kk in rng.choice(col.size, 32, replace=False):
        i, j = int(rows[kk]), int(col[kk])
        deltas = []
        for sgn in (1.0, -1.0):
            v = val.copy()
            v[kk] += sgn * h
            vdev.copy_from(v)
            Ap.set_values(vdev)
            rhs_h = np.zeros(n)
            rhs_h[i] = -sgn * h * xh[j]
            rhs = ra.ParVector.from_numpy(ctx, rhs_h)
            d.set_const_value(0.0)
            for _ in range(3):
                Ap.residual(d, rhs, r)
                z.set_const_value(0.0)
                try:
                    ra.FPCG(A, ml, z, r, 1e-12, 500, ra.PREC_FP64)
                except ra.RamdError as err:
                    if err.code != -5:
                        raise
                d.axpy(z, 1.0)
            deltas.append(d.numpy())
        dp_, dm_ = deltas
        fd = (float(e @ (dp_ - dm_)) + 0.5 * float(dp_ @ dp_ - dm_ @ dm_)) / (2 * h)
        worst = max(worst, abs(gA[kk] - fd) / max(abs(fd), 1e-300))
    assert worst < 1e-6, worst


def test_c5_lambda_matches_oracle_pcg(ctx, c5):
    ml, n = c5["ml"], c5["n"]
    o = c5["out"][ra.PREC_FP64]
    As, Ps = [], []
    for l in range(ml.num_levels):
        Al, Pl = ml.level(l)
        As.append(K.Csr(*Al.to_csr(), Al.n_cols))
        if Pl is not None:
            Ps.append(K.Csr(*Pl.to_csr(), Pl.n_cols))
    H = K.Hier(As, Ps, relax_type=0, sweeps=1, omega=2.0 / 3.0)
    Ah = K.Csr(c5["rp"], c5["col"], c5["val"])
    gh = o["x"] - c5["xs"]
    gmg = float(gh @ H.cycle(np.zeros(n), gh))
    lam, _, _ = H.pcg(Ah, np.zeros(n), gh, tol=1e-20 * math.sqrt(gmg), max_iter=300)
    assert np.linalg.norm(lam - o["db"]) <= 1e-9 * np.linalg.norm(lam)
    # dL/dA = -lambda_i x_j on A's pattern, from the oracle's lambda
    rows = np.repeat(np.arange(n), np.diff(c5["rp"]))
    ref = -lam[rows] * o["x"][c5["col"]]
    assert np.linalg.norm(ref - o["dA"]) <= 1e-9 * np.linalg.norm(ref)


def test_c5_fp32_cycles_same_gradient(c5):
    a, b = c5["out"][ra.PREC_FP64], c5["out"][ra.PREC_FP32]
    assert b["rel"] <= 1e-10
    assert np.linalg.norm(b["dA"] - a["dA"]) <= 1e-8 * np.linalg.norm(a["dA"])
    assert np.linalg.norm(b["db"] - a["db"]) <= 1e-8 * np.linalg.norm(a["db"])


def test_c5_sa_setup_device_equals_host(ctx, c5):
    """The smoothed-aggregation hierarchy of the full-size workload built with
    the device setup steps (strength, MIS-2, aggregation, fit_candidates,
    Jacobi prolongation) equals the one built with the host steps: every
    level's A and P and the aggregates, bit for bit."""
    ml = ra.ParSmoothedAggregationSolver(ctx, 0.08)
    ml.relax_weight = 2.0 / 3.0
    ml.device_setup = False
    ml.setup(c5["rp"], c5["col"], c5["val"])
    a = c5["ml"]
    assert a.num_levels == ml.num_levels >= 4
    for l in range(a.num_levels):
        for M1, M2 in zip(a.level(l), ml.level(l)):
            if M1 is None:
                assert M2 is None
                continue
            assert M1.n_cols == M2.n_cols
            assert all(np.array_equal(u, v) for u, v in zip(M1.to_csr(), M2.to_csr())), l
        if l < a.num_levels - 1:
            assert np.array_equal(a.states(l), ml.states(l)), l

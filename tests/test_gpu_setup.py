"""Setup steps on the device (classical strength, PMIS, extended
interpolation with filter) against the host restatement, which is pinned to
the reference's fixtures and the compiled reference (test_host_setup.py):
S patterns, C/F states and P bit-identical."""
import math

import numpy as np
import pytest

import checkers as K
import raptor_amd as ra

pytestmark = pytest.mark.gpu


def random_mmatrix(n, seed):
    """Symmetric M-matrix with a random pattern (row sums >= 0)."""
    rng = np.random.default_rng(seed)
    rows = [dict() for _ in range(n)]
    for i in range(n):
        for j in rng.choice(n, size=rng.integers(1, 9), replace=False):
            if j == i:
                continue
            v = -rng.uniform(0.1, 2.0)
            rows[i][j] = v
            rows[j][i] = v
    rp, col, val = [0], [], []
    for i in range(n):
        d = -sum(rows[i].values()) + rng.uniform(0.0, 0.5)
        items = sorted(rows[i].items())
        col += [i] + [j for j, _ in items]
        val += [d] + [v for _, v in items]
        rp.append(len(col))
    return np.array(rp, np.int32), np.array(col, np.int32), np.array(val)


CASES = [
    ("lap27_12", lambda: ra.gallery_laplace27(12)),
    ("aniso50", lambda: ra.gallery_aniso2d(50, 0.001, math.pi / 8)),
    ("fv7_10", lambda: ra.gallery_fv7(10, seed=4)),
    ("random_mm_700", lambda: random_mmatrix(700, 5)),
]


def diag_first(rp, col, val):
    A = K.move_diag(K.sort_rows(K.Csr(rp, col, val)))
    return A


@pytest.mark.parametrize("theta", [0.0, 0.25, 0.5])
@pytest.mark.parametrize("name,gen", CASES, ids=[c[0] for c in CASES])
def test_strength_pmis_extended_match_host(ctx, name, gen, theta):
    A = diag_first(*gen())
    n = A.n_rows
    D = ra.ParCSRMatrix.from_csr(ctx, A.rowptr, A.col, A.val)
    S = ra.strength_gpu(D, theta)
    srp, scol, sval = S.to_csr()
    hrp, hcol = ra.strength(A.rowptr, A.col, A.val, ra.Classical, theta)
    assert np.array_equal(srp, hrp) and np.array_equal(scol, hcol)
    # S values are A's kept values (the diagonal slot first)
    assert np.all(scol[srp[:-1]] == np.arange(n))
    w = np.random.default_rng(6).random(n)
    st = ra.split_pmis_gpu(S, w)
    hst = ra.split(hrp, hcol, ra.PMIS, w)
    assert np.array_equal(st, hst)
    assert np.array_equal(ra.split_hmis_gpu(S, w), ra.split(hrp, hcol, ra.HMIS, w))
    assert np.array_equal(ra.split_cljp_gpu(S, w), ra.split(hrp, hcol, ra.CLJP, w))
    Pm = ra.interp_mod_classical_gpu(D, S, hst)
    mrp, mcol, mval = Pm.to_csr()
    hmrp, hmcol, hmval, hmnc = ra.interpolation(A.rowptr, A.col, A.val, hrp, hcol, hst, ra.ModClassical, 0.0)
    assert Pm.n_cols == hmnc
    assert np.array_equal(mrp, hmrp) and np.array_equal(mcol, hmcol) and np.array_equal(mval, hmval)
    Pd = ra.interp_direct_gpu(D, S, hst)
    hdrp, hdcol, hdval, hdnc = ra.interpolation(A.rowptr, A.col, A.val, hrp, hcol, hst, ra.Direct, 0.0)
    assert Pd.n_cols == hdnc
    assert all(np.array_equal(u, v) for u, v in zip(Pd.to_csr(), (hdrp, hdcol, hdval)))
    ssrp, sscol, _ = ra.strength_gpu(D, theta, ra.Symmetric).to_csr()
    hsrp, hscol = ra.strength(A.rowptr, A.col, A.val, ra.Symmetric, theta)
    assert np.array_equal(ssrp, hsrp) and np.array_equal(sscol, hscol)
    assert set(np.unique(st)) <= {-2, 0, 1}
    for filt in (0.3, 0.0):
        P = ra.interp_extended_gpu(D, S, st, filt)
        prp, pcol, pval = P.to_csr()
        hprp, hpcol, hpval, hnc = ra.interpolation(A.rowptr, A.col, A.val, hrp, hcol, hst, ra.Extended, filt)
        assert P.n_cols == hnc
        assert np.array_equal(prp, hprp) and np.array_equal(pcol, hpcol)
        assert np.array_equal(pval, hpval)


def test_pmis_reference_weights(ctx):
    """Without weights both sides draw the reference's srand(2448422) stream."""
    A = diag_first(*ra.gallery_laplace27(10))
    D = ra.ParCSRMatrix.from_csr(ctx, A.rowptr, A.col, A.val)
    S = ra.strength_gpu(D, 0.25)
    hrp, hcol = ra.strength(A.rowptr, A.col, A.val, ra.Classical, 0.25)
    assert np.array_equal(ra.split_pmis_gpu(S), ra.split(hrp, hcol, ra.PMIS))


@pytest.mark.parametrize("coarsen,stype,interp", [(ra.PMIS, ra.Classical, ra.Extended),
                                                  (ra.HMIS, ra.Classical, ra.Extended),
                                                  (ra.PMIS, ra.Symmetric, ra.Extended),
                                                  (ra.CLJP, ra.Classical, ra.ModClassical),
                                                  (ra.Falgout, ra.Classical, ra.ModClassical),
                                                  (ra.RS, ra.Classical, ra.Extended),
                                                  (ra.CLJP, ra.Classical, ra.Direct)],
                         ids=["pmis", "hmis", "pmis_sym", "cljp_modc", "falgout_modc", "rs_ext", "cljp_direct"])
@pytest.mark.parametrize("name,gen", [("lap27_16", lambda: ra.gallery_laplace27(16)),
                                      ("aniso80", lambda: ra.gallery_aniso2d(80, 0.001, math.pi / 8))])
def test_hierarchy_device_setup_equals_host(ctx, name, gen, coarsen, stype, interp):
    """Ruge-Stuben hierarchies (every splitting, extended or modified
    classical interpolation, classical or symmetric strength) built with the
    device steps and with the host steps: every level's A and P, and the C/F
    states, identical."""
    rp, col, val = gen()
    mls = []
    for dev in (True, False):
        ml = ra.ParRugeStubenSolver(ctx, 0.25, coarsen, interp, stype, ra.Jacobi)
        ml.device_setup = dev
        # a coarse matrix with a zero pivot does not stop the setup (the
        # reference ignores dgetrf_'s info): every level is compared
        ml.setup(rp, col, val)
        mls.append(ml)
    a, b = mls
    assert a.coarse_info() == b.coarse_info()
    assert a.num_levels == b.num_levels >= 3
    for l in range(a.num_levels):
        for M1, M2 in zip(a.level(l), b.level(l)):
            if M1 is None:
                assert M2 is None
                continue
            assert all(np.array_equal(u, v) for u, v in zip(M1.to_csr(), M2.to_csr()))
        if l < a.num_levels - 1:
            assert np.array_equal(a.states(l), b.states(l))


@pytest.mark.parametrize("name", ["aniso", "laplacian"])
def test_device_steps_golden(ctx, golden, name):
    """The reference's own fixtures (tests/test_par_strength.cpp,
    ruge_stuben/tests/test_par_interpolation.cpp): classical strength
    (theta 0.25) and extended interpolation (filter 0) on the device."""
    A = K.move_diag(K.fixture_csr(golden, name))
    D = ra.ParCSRMatrix.from_csr(ctx, A.rowptr, A.col, A.val)
    for stype, suffix in ((ra.Symmetric, "_SS"), (ra.Classical, "_S")):
        srp, scol, _ = ra.strength_gpu(D, 0.25, stype).to_csr()
        E = K.move_diag(K.fixture_csr(golden, name + suffix))
        S = K.move_diag(K.sort_rows(K.Csr(srp, scol, np.ones(scol.size))))
        assert np.array_equal(S.rowptr, E.rowptr) and np.array_equal(S.col, E.col)
    Sf = ra.ParCSRMatrix.from_csr(ctx, E.rowptr, E.col, E.val)
    states = golden[name + "_split"].astype(np.int32)
    for P, suffix in ((ra.interp_extended_gpu(D, Sf, states, 0.0), "extend"),
                      (ra.interp_mod_classical_gpu(D, Sf, states), "mod_class"),
                      (ra.interp_direct_gpu(D, Sf, states), "direct")):
        prp, pcol, pval = P.to_csr()
        F = K.fixture_csr(golden, f"{name}_P_{suffix}")
        assert P.n_cols == F.n_cols
        M = K.sort_rows(K.Csr(prp, pcol, pval, F.n_cols))
        assert np.array_equal(M.rowptr, F.rowptr) and np.array_equal(M.col, F.col)
        assert np.allclose(M.val, F.val, rtol=0, atol=1e-6)  # the fixture's own bound (par_compare.hpp)


@pytest.mark.parametrize("level", [0, 1])
def test_device_pmis_cljp_golden(ctx, golden, level):
    """ruge_stuben/tests/test_par_splitting.cpp: PMIS and CLJP on rss_S{0,1}
    with weights.txt give rss_cf{0,1}_pmis and rss_cf{0,1} on the device."""
    S = K.move_diag(K.fixture_csr(golden, f"rss_S{level}"))
    D = ra.ParCSRMatrix.from_csr(ctx, S.rowptr, S.col, S.val)
    w = golden["weights"][:S.n_rows]
    assert np.array_equal(ra.split_pmis_gpu(D, w), golden[f"rss_cf{level}_pmis"])
    assert np.array_equal(ra.split_cljp_gpu(D, w), golden[f"rss_cf{level}"])


def hub_mmatrix(n, hub_deg, seed):
    """A sparse M-matrix (a chain) with one hub row that depends strongly on
    hub_deg rows while they depend on it only weakly: the hub stays F with
    more than 256 strong C neighbours, so the device interpolation hands that
    level to the host."""
    rng = np.random.default_rng(seed)
    rows = [dict() for _ in range(n)]
    for i in range(n - 1):
        rows[i][i + 1] = -1.0
        rows[i + 1][i] = -1.0
    for j in rng.choice(np.arange(2, n), size=hub_deg, replace=False):
        rows[0][int(j)] = -1.0
        rows[int(j)][0] = -0.01
    rp, col, val = [0], [], []
    for i in range(n):
        items = sorted(rows[i].items())
        col += [i] + [j for j, _ in items]
        val += [-sum(v for _, v in items) + 0.1] + [v for _, v in items]
        rp.append(len(col))
    return np.array(rp, np.int32), np.array(col, np.int32), np.array(val)


@pytest.mark.parametrize("interp", [ra.Extended, ra.ModClassical], ids=["extended", "modclass"])
def test_device_setup_row_overflow_falls_back(ctx, interp):
    """A level with a row past the device kernels' 256 columns is built by
    the host restatement; the hierarchy still equals the all-host one (and
    the device entry point reports the overflow)."""
    rp, col, val = hub_mmatrix(3000, 1200, 5)
    A = K.move_diag(K.sort_rows(K.Csr(rp, col, val)))
    D = ra.ParCSRMatrix.from_csr(ctx, A.rowptr, A.col, A.val)
    S = ra.strength_gpu(D, 0.25)
    st = ra.split_pmis_gpu(S, np.random.default_rng(1).random(A.n_rows))
    fn = ra.interp_extended_gpu if interp == ra.Extended else ra.interp_mod_classical_gpu
    args = (D, S, st, 0.3) if interp == ra.Extended else (D, S, st)
    overflow = False
    try:
        fn(*args)
    except ra.RamdError:
        overflow = True
    mls = []
    for dev in (True, False):
        ml = ra.ParRugeStubenSolver(ctx, 0.25, ra.PMIS, interp, ra.Classical, ra.Jacobi)
        ml.device_setup = dev
        ml.setup(rp, col, val)
        mls.append(ml)
    a, b = mls
    assert a.num_levels == b.num_levels
    for l in range(a.num_levels):
        for M1, M2 in zip(a.level(l), b.level(l)):
            if M1 is None:
                assert M2 is None
                continue
            assert all(np.array_equal(u, v) for u, v in zip(M1.to_csr(), M2.to_csr())), l
    assert overflow  # the hub row needs more than 256 columns


# ---------------------------------------------------------------------------
# smoothed aggregation (aggregation/tests/test_par_smoothed_aggregation.cpp)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("level", [0, 1])
def test_device_sa_steps_golden(ctx, golden, level):
    """Symmetric strength (0.25) -> MIS-2 with weights.txt -> aggregation on
    the device against sas_S/mis/agg of the reference's test (aggregate
    numbers there index the selected rows in increasing order)."""
    A = K.move_diag(K.sort_rows(K.fixture_csr(golden, f"sas_A{level}", sort=False)))
    n = A.n_rows
    w = golden["weights"][:n]
    D = ra.ParCSRMatrix.from_csr(ctx, A.rowptr, A.col, A.val)
    S = ra.strength_gpu(D, 0.25, ra.Symmetric)
    srp, scol, _ = S.to_csr()
    E = K.move_diag(K.fixture_csr(golden, f"sas_S{level}"))
    assert np.array_equal(srp, E.rowptr) and np.array_equal(scol, E.col)
    st = ra.mis2_gpu(S, w)
    assert np.array_equal(st, golden[f"sas_mis{level}"])
    agg, na = ra.aggregate_gpu(D, S, st, w)
    roots = np.flatnonzero(st > 0)
    assert na == roots.size
    assert np.array_equal(roots[golden[f"sas_agg{level}"]], agg)


@pytest.mark.parametrize("stype", [ra.Symmetric, ra.Classical], ids=["symmetric", "classical"])
@pytest.mark.parametrize("theta", [0.0, 0.08, 0.25])
@pytest.mark.parametrize("name,gen", CASES, ids=[c[0] for c in CASES])
def test_device_mis2_aggregate_match_host(ctx, name, gen, theta, stype):
    """MIS-2 and aggregation (with and without random values) on the device
    equal the host restatement's on the same S, which the compiled reference
    pins (test_host_setup.py)."""
    A = diag_first(*gen())
    n = A.n_rows
    D = ra.ParCSRMatrix.from_csr(ctx, A.rowptr, A.col, A.val)
    S = ra.strength_gpu(D, theta, stype)
    srp, scol, _ = S.to_csr()
    w = np.random.default_rng(7).random(n)
    st = ra.mis2_gpu(S, w)
    assert np.array_equal(st, ra.mis2(srp, scol, w))
    assert np.array_equal(ra.mis2_gpu(S), ra.mis2(srp, scol))  # the srand(2448422) stream
    for rv in (None, w):
        agg, na = ra.aggregate_gpu(D, S, st, rv)
        agg_h, na_h = ra.aggregate(A.rowptr, A.col, A.val, srp, scol, st, rv)
        assert na == na_h and np.array_equal(agg, agg_h)


@pytest.mark.parametrize("frac", [0.0, 0.002, 0.05])
def test_device_aggregate_row_order_pass(ctx, frac):
    """States that leave rows with no aggregated neighbour (none or very few
    selected rows): the second pass then depends on the row order (such a
    row's code n - 1 reads as an aggregate to the rows after it), and the
    device reruns it in row order; the result equals the host's."""
    A = diag_first(*random_mmatrix(900, 11))
    n = A.n_rows
    D = ra.ParCSRMatrix.from_csr(ctx, A.rowptr, A.col, A.val)
    S = ra.strength_gpu(D, 0.0, ra.Symmetric)
    srp, scol, _ = S.to_csr()
    st = (np.random.default_rng(12).random(n) < frac).astype(np.int32)
    agg, na = ra.aggregate_gpu(D, S, st)
    agg_h, na_h = ra.aggregate(A.rowptr, A.col, A.val, srp, scol, st)
    assert na == na_h and np.array_equal(agg, agg_h)
    if frac == 0.0:
        assert na == 0 and (agg == n - 1).any()  # the code n - 1 survives decoding


def test_device_aggregate_missing_entry_raises(ctx):
    """An S entry that A does not store stops the aggregation (host and
    device alike)."""
    A = diag_first(*ra.gallery_aniso2d(12, 0.001, math.pi / 8))
    n = A.n_rows
    # S: each row's diagonal and a column that A's row does not hold
    srp = np.arange(0, 2 * n + 1, 2, dtype=np.int32)
    far = (np.arange(n) + n // 2) % n
    scol = np.empty(2 * n, np.int32)
    scol[0::2] = np.arange(n)
    scol[1::2] = far
    assert not any(far[i] in A.col[A.rowptr[i]:A.rowptr[i + 1]] for i in range(n))
    D = ra.ParCSRMatrix.from_csr(ctx, A.rowptr, A.col, A.val)
    S = ra.ParCSRMatrix.from_csr(ctx, srp, scol, np.ones(2 * n))
    st = np.zeros(n, np.int32)
    with pytest.raises(ra.RamdError):
        ra.aggregate(A.rowptr, A.col, A.val, srp, scol, st)
    with pytest.raises(ra.RamdError):
        ra.aggregate_gpu(D, S, st)


SA_CASES = [
    ("fv7_14_s3", lambda: ra.gallery_fv7(14, seed=3)),
    ("fv7_20_s9", lambda: ra.gallery_fv7(20, seed=9)),
    ("aniso60", lambda: ra.gallery_aniso2d(60, 0.001, math.pi / 8)),
    ("lap27_12", lambda: ra.gallery_laplace27(12)),
    ("random_mm_700", lambda: random_mmatrix(700, 5)),
]


@pytest.mark.parametrize("steps", [0, 1, 2])
@pytest.mark.parametrize("theta,stype", [(0.0, ra.Symmetric), (0.08, ra.Symmetric), (0.25, ra.Classical)],
                         ids=["sym0", "sym008", "cls025"])
@pytest.mark.parametrize("name,gen", SA_CASES, ids=[c[0] for c in SA_CASES])
def test_sa_hierarchy_device_setup_equals_host(ctx, name, gen, theta, stype, steps):
    """Smoothed-aggregation hierarchies built with the device steps
    (strength, MIS-2, aggregation, the prolongation smoothing) and with the
    host steps: every level's A and P and the aggregates identical."""
    rp, col, val = gen()
    mls, errs = [], []
    for dev in (True, False):
        ml = ra.ParSmoothedAggregationSolver(ctx, theta, stype, ra.Jacobi, prolong_smooth_steps=steps)
        ml.device_setup = dev
        ml.max_coarse = 20
        try:
            ml.setup(rp, col, val)
            errs.append(None)
        except ra.RamdError as e:  # e.g. an aggregate without a root: both builds stop alike
            errs.append(str(e))
        mls.append(ml)
    assert errs[0] == errs[1]
    if errs[0] is not None:
        pytest.skip(f"both builds stop alike: {errs[0]}")
    a, b = mls
    assert a.num_levels == b.num_levels >= 2
    for l in range(a.num_levels):
        for M1, M2 in zip(a.level(l), b.level(l)):
            if M1 is None:
                assert M2 is None
                continue
            assert M1.n_cols == M2.n_cols
            assert all(np.array_equal(u, v) for u, v in zip(M1.to_csr(), M2.to_csr())), l
        if l < a.num_levels - 1:
            assert np.array_equal(a.states(l), b.states(l)), l


def candidate(kind, n, seed):
    rng = np.random.default_rng(seed)
    if kind == "ones":
        return np.ones(n)
    if kind == "uniform":
        return rng.random(n) + 0.5
    return np.exp(rng.normal(0.0, 20.0, n)) * rng.choice([-1.0, 1.0], n)  # magnitudes over ~1e-26..1e26


@pytest.mark.parametrize("bkind", ["ones", "uniform", "wide"])
@pytest.mark.parametrize("name,gen", CASES, ids=[c[0] for c in CASES])
def test_device_fit_candidates_match_host(ctx, name, gen, bkind):
    """fit_candidates on the device (aggregate norms in ascending row order,
    their square roots and reciprocals) equals the host restatement bit for
    bit, on aggregates from the device MIS-2 and aggregation."""
    A = diag_first(*gen())
    n = A.n_rows
    D = ra.ParCSRMatrix.from_csr(ctx, A.rowptr, A.col, A.val)
    S = ra.strength_gpu(D, 0.08, ra.Symmetric)
    agg, na = ra.aggregate_gpu(D, S, ra.mis2_gpu(S, np.random.default_rng(8).random(n)))
    B = candidate(bkind, n, 9)
    T, R = ra.fit_candidates_gpu(ctx, agg, na, B)
    trp, tcol, tval, Rh = ra.fit_candidates(agg, na, B)
    drp, dcol, dval = T.to_csr()
    assert T.n_cols == na
    assert np.array_equal(drp, trp) and np.array_equal(dcol, tcol) and np.array_equal(dval, tval)
    assert np.array_equal(R, Rh)


def test_device_fit_candidates_many_aggregates(ctx):
    """200000 rows in 50000 aggregates of random membership with candidate
    values over ~52 decades: every norm (a device square root) and every
    scaled entry equals the host's."""
    n, m = 200000, 50000
    rng = np.random.default_rng(10)
    roots = np.sort(rng.choice(n, m, replace=False))
    agg = roots[rng.integers(0, m, n)].astype(np.int32)
    agg[roots] = roots
    agg[rng.choice(n, 1000, replace=False)] = -1  # unaggregated rows
    agg[roots] = roots
    B = candidate("wide", n, 11)
    T, R = ra.fit_candidates_gpu(ctx, agg, m, B)
    trp, tcol, tval, Rh = ra.fit_candidates(agg, m, B)
    drp, dcol, dval = T.to_csr()
    assert np.array_equal(drp, trp) and np.array_equal(dcol, tcol) and np.array_equal(dval, tval)
    assert np.array_equal(R, Rh)


def test_device_aggregate_rejects_unsorted_rows(ctx):
    """The device aggregation walks A's row alongside S's, so it takes both
    stored sorted with the diagonal first; other orders are refused (the
    host restatement sorts its copies instead)."""
    A = diag_first(*ra.gallery_aniso2d(12, 0.001, math.pi / 8))
    D = ra.ParCSRMatrix.from_csr(ctx, A.rowptr, A.col, A.val)
    S = ra.strength_gpu(D, 0.25, ra.Symmetric)
    st = ra.mis2_gpu(S, np.random.default_rng(3).random(A.n_rows))
    ra.aggregate_gpu(D, S, st)  # sorted: accepted
    col = A.col.copy()
    s, e = A.rowptr[40] + 1, A.rowptr[41]
    col[s:e] = col[s:e][::-1]  # row 40's off-diagonal entries descending
    val = A.val.copy()
    val[s:e] = val[s:e][::-1]
    U = ra.ParCSRMatrix.from_csr(ctx, A.rowptr, col, val)
    with pytest.raises(ra.RamdError):
        ra.aggregate_gpu(U, S, st)
    rp = A.rowptr
    col = A.col.copy()
    col[rp[7]], col[rp[7] + 1] = col[rp[7] + 1], col[rp[7]]  # row 7's diagonal second
    val = A.val.copy()
    val[rp[7]], val[rp[7] + 1] = val[rp[7] + 1], val[rp[7]]
    U = ra.ParCSRMatrix.from_csr(ctx, rp, col, val)
    with pytest.raises(ra.RamdError):
        ra.aggregate_gpu(U, S, st)

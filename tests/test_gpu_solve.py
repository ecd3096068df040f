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


def build(ctx, rp, col, val, cfg, max_coarse=50):
    if cfg["solver"] == "sa":
        ml = ra.ParSmoothedAggregationSolver(ctx, cfg["theta"], ra.Symmetric, ra.Jacobi)
    else:
        ml = ra.ParRugeStubenSolver(ctx, cfg.get("theta", 0.25), cfg["coarsen"], cfg["interp"], ra.Classical,
                                    ra.Jacobi)
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


def rs(coarsen, interp):
    return dict(solver="rs", coarsen=coarsen, interp=interp)


def sa(theta):
    return dict(solver="sa", theta=theta)


CONFIGS = [
    # C1: aniso 100^2, CLJP + ModClassical
    ("c1_aniso100", lambda: ra.gallery_aniso2d(100, 0.001, math.pi / 8), rs(ra.CLJP, ra.ModClassical)),
    # C2 shape at reduced size: 27-pt, PMIS + Extended and HMIS + Extended
    ("lap27_16_pmis_ext", lambda: ra.gallery_laplace27(16), rs(ra.PMIS, ra.Extended)),
    ("lap27_12_hmis_ext", lambda: ra.gallery_laplace27(12), rs(ra.HMIS, ra.Extended)),
    ("aniso60_pmis_direct", lambda: ra.gallery_aniso2d(60, 0.001, math.pi / 8), rs(ra.PMIS, ra.Direct)),
    ("aniso50_rs_modclass", lambda: ra.gallery_aniso2d(50, 0.001, math.pi / 8), rs(ra.RS, ra.ModClassical)),
    # C5 shape at reduced size: 7-point FV, smoothed aggregation theta 0.08
    ("c5_fv7_20_sa", lambda: ra.gallery_fv7(20, seed=3), sa(0.08)),
    ("aniso60_sa", lambda: ra.gallery_aniso2d(60, 0.001, math.pi / 8), sa(0.0)),
]


@pytest.fixture(scope="module", params=CONFIGS, ids=[c[0] for c in CONFIGS])
def problem(request, ctx):
    name, gen, cfg = request.param
    rp, col, val = gen()
    ml = build(ctx, rp, col, val, cfg)
    As, Ps = levels(ml)
    return dict(name=name, rp=rp, col=col, val=val, ml=ml, As=As, Ps=Ps, cfg=cfg)


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
    """Same number of levels, same C/F splittings (or SA aggregates), same P
    and A_l patterns (exact) and values (1e-12 relative) as the reference's
    setup."""
    rp, col, val = problem["rp"], problem["col"], problem["val"]
    L = K.ref()
    rpa, cola, vala = K.i32(rp), K.i32(col), K.f64(val)
    cfg = problem["cfg"]
    if cfg["solver"] == "sa":
        nl = L.ref_hier_setup_sa(rpa.size - 1, K.ip(rpa), K.ip(cola), K.dp(vala), C.c_double(cfg["theta"]), 1, 0,
                                 C.c_double(OMEGA), 50, 25, None, C.c_double(4.0 / 3), 1)
    else:
        nl = L.ref_hier_setup(rpa.size - 1, K.ip(rpa), K.ip(cola), K.dp(vala), C.c_double(0.25), cfg["coarsen"],
                              cfg["interp"], 0, 0, C.c_double(OMEGA), 50, 25, None)
    ml, As, Ps = problem["ml"], problem["As"], problem["Ps"]
    assert nl == ml.num_levels
    for l in range(nl):
        R = K.Csr(*K.ref_fetch(L.ref_hier_level(l, 0)))
        M = As[l]
        if cfg["solver"] == "sa" and l == nl - 1:
            # the reference leaves the coarsest SA operator in the product's
            # ascending order (only the next level's strength would move its
            # diagonal first); the dense coarse solve does not see the order
            R, M = K.move_diag(K.sort_rows(R)), K.move_diag(K.sort_rows(M))
        assert np.array_equal(M.rowptr, R.rowptr) and np.array_equal(M.col, R.col), l
        assert np.allclose(M.val, R.val, rtol=1e-12, atol=1e-300), l
        if l < nl - 1:
            rpr, colr, valr, nc = K.ref_fetch(L.ref_hier_level(l, 1))
            assert nc == Ps[l].n_cols
            assert np.array_equal(Ps[l].rowptr, rpr) and np.array_equal(Ps[l].col, colr), l
            assert np.allclose(Ps[l].val, valr, rtol=1e-12, atol=1e-300), l
            st = np.empty(As[l].n_rows, np.int32)
            L.ref_hier_states(l, K.ip(st))
            assert np.array_equal(ml.states(l), st), l


def set_kernel(M, p, code):
    """Selects the CSR kernel of row_caps' encoding (-1/-2 sliced, 1 lane
    rows, 2 wavefront rows, a cap: the tuned row blocks, minus a cap: those
    blocks in sorted-gather order) for precision p."""
    M.set_wave_rows(p, False)
    M.set_lane_rows(p, False)
    M.set_sliced(p, False)
    M.set_sorted_gather(p, False)
    if code in (-1, -2):
        M.set_sliced(p, int(-code))
    elif code == 1:
        M.set_lane_rows(p, True)
    elif code == 2:
        M.set_wave_rows(p, True)
    elif code < -2:
        M.set_sorted_gather(p, True)


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


def test_vcycle_fp32_bit_exact(problem, ctx):
    """FP32 cycle (the FPCG preconditioner at single precision) against the
    oracle's single-precision cycle on the same hierarchy: bit-identical."""
    ml, As, Ps = problem["ml"], problem["As"], problem["Ps"]
    H = K.Hier(As, Ps, relax_type=0, sweeps=1, omega=OMEGA)
    n = As[0].n_rows
    b = np.random.default_rng(22).standard_normal(n).astype(np.float32)
    x = ra.ParVector(ctx, n, np.float32)
    bd = ra.ParVector.from_numpy(ctx, b)
    ml.cycle_fp32(x, bd)
    assert np.array_equal(x.numpy(), H.cycle_fp32(np.zeros(n, np.float32), b))


def test_vcycle_coarse_tail(problem, ctx):
    """The small coarsest levels can run as one workgroup kernel (the coarse
    tail).  With the tail forced to start at every level that can start it,
    FP64 and FP32 cycles equal the one-launch-per-operation cycles (which
    equal the oracle's) bit for bit; the automatic mode keeps one of these
    choices."""
    ml, As = problem["ml"], problem["As"]
    ml.enable_fp32()
    _, first = ml.tail_level()
    assert 1 <= first <= ml.num_levels - 1
    n = As[0].n_rows
    b = np.random.default_rng(26).standard_normal(n)
    out = {}
    try:
        for mode in [-1] + list(range(first, ml.num_levels)):
            ml.set_tail(mode)
            assert ml.tail_level(ra.PREC_FP64)[0] == mode and ml.tail_level(ra.PREC_FP32)[0] == mode
            x = ra.ParVector.from_numpy(ctx, np.zeros(n))
            ml.cycle(x, ra.ParVector.from_numpy(ctx, b), x_is_zero=True)
            x32 = ra.ParVector(ctx, n, np.float32)
            ml.cycle_fp32(x32, ra.ParVector.from_numpy(ctx, b.astype(np.float32)))
            out[mode] = (x.numpy(), x32.numpy())
        ml.set_tail(0)
        x = ra.ParVector.from_numpy(ctx, np.zeros(n))
        ml.cycle(x, ra.ParVector.from_numpy(ctx, b), x_is_zero=True)
        assert ml.tail_level(ra.PREC_FP64)[0] in out
    finally:
        ml.set_tail(0)
    for mode in out:
        assert np.array_equal(out[mode][0], out[-1][0]), mode
        assert np.array_equal(out[mode][1], out[-1][1]), mode
    assert np.array_equal(x.numpy(), out[-1][0])


@pytest.mark.parametrize("sweeps", [0, 2, 3])
@pytest.mark.parametrize("cfg", ["lap27_16_pmis_ext", "c5_fv7_20_sa"])
def test_vcycle_sweep_counts(ctx, cfg, sweeps):
    """Other sweep counts (0: no smoothing, the coarse iterates start from
    the zero the reference sets; 2, 3: sweeps after the first from zero),
    with the largest coarse tail and without: FP64 and FP32 cycles
    bit-identical to the oracle's."""
    name, gen, c = [t for t in CONFIGS if t[0] == cfg][0]
    rp, col, val = gen()
    if c["solver"] == "sa":
        ml = ra.ParSmoothedAggregationSolver(ctx, c["theta"], ra.Symmetric, ra.Jacobi)
    else:
        ml = ra.ParRugeStubenSolver(ctx, 0.25, c["coarsen"], c["interp"], ra.Classical, ra.Jacobi)
    ml.relax_weight = OMEGA
    ml.max_coarse = 50
    ml.num_smooth_sweeps = sweeps
    ml.setup(rp, col, val)
    ml.enable_fp32()
    As, Ps = levels(ml)
    H = K.Hier(As, Ps, relax_type=0, sweeps=sweeps, omega=OMEGA)
    n = As[0].n_rows
    b = np.random.default_rng(27).standard_normal(n)
    b32 = b.astype(np.float32)
    xo, xo32 = H.cycle(np.zeros(n), b), H.cycle_fp32(np.zeros(n, np.float32), b32)
    _, first = ml.tail_level()
    assert first >= 1
    for mode in (first, -1):
        ml.set_tail(mode)
        # a nonzero buffer: the cycle from zero must not read it
        x = ra.ParVector.from_numpy(ctx, np.full(n, 7.0))
        ml.cycle(x, ra.ParVector.from_numpy(ctx, b), x_is_zero=True)
        assert np.array_equal(x.numpy(), xo), mode
        x32 = ra.ParVector.from_numpy(ctx, np.full(n, 7.0, np.float32))
        ml.cycle_fp32(x32, ra.ParVector.from_numpy(ctx, b32))
        assert np.array_equal(x32.numpy(), xo32), mode


@pytest.mark.parametrize("planes", [4, 7])
@pytest.mark.parametrize("cfg", ["lap27_16_pmis_ext", "c5_fv7_20_sa"])
def test_vcycle_plane_march_bit_exact(ctx, cfg, planes):
    """The fine level's residual and Jacobi sweeps on the plane-march kernel
    (FP64 values with FP64 vectors, FP32 values with FP32 vectors, the
    latter loading the next plane's values one plane ahead), whatever the
    tuner picked: FP64 and FP32 cycles bit-identical to the oracle's."""
    name, gen, c = [t for t in CONFIGS if t[0] == cfg][0]
    rp, col, val = gen()
    ml = build(ctx, rp, col, val, c)
    ml.enable_fp32()
    A0 = ml.level(0)[0]
    assert A0.diagonals() > 0 and A0.dia_staging(0)[2] > 0  # a box geometry
    before = [A0.dia_staging(pp)[:2] for pp in (0, 1)]
    try:
        for pp in (0, 1):
            A0.set_dia_staging(pp, 2, planes)
        As, Ps = levels(ml)
        H = K.Hier(As, Ps, relax_type=0, sweeps=1, omega=OMEGA)
        n = As[0].n_rows
        b = np.random.default_rng(29).standard_normal(n)
        x = ra.ParVector.from_numpy(ctx, np.zeros(n))
        ml.cycle(x, ra.ParVector.from_numpy(ctx, b), x_is_zero=True)
        assert np.array_equal(x.numpy(), H.cycle(np.zeros(n), b))
        b32 = b.astype(np.float32)
        x32 = ra.ParVector(ctx, n, np.float32)
        ml.cycle_fp32(x32, ra.ParVector.from_numpy(ctx, b32))
        assert np.array_equal(x32.numpy(), H.cycle_fp32(np.zeros(n, np.float32), b32))
    finally:
        for pp in (0, 1):
            mode, pl = before[pp]
            A0.set_dia_staging(pp, mode, pl if mode == 2 else 0)


@pytest.mark.parametrize("max_coarse", [400])
def test_vcycle_large_coarse_system(ctx, max_coarse):
    """A coarsest system larger than one wavefront takes the workgroup
    solve; FP64 and FP32 cycles stay bit-identical to the oracle."""
    rp, col, val = ra.gallery_aniso2d(60, 0.001, math.pi / 8)
    ml = build(ctx, rp, col, val, rs(ra.PMIS, ra.Direct), max_coarse=max_coarse)
    As, Ps = levels(ml)
    assert As[-1].n_rows > 64
    H = K.Hier(As, Ps, relax_type=0, sweeps=1, omega=OMEGA)
    n = As[0].n_rows
    b = np.random.default_rng(23).standard_normal(n)
    x = ra.ParVector.from_numpy(ctx, np.zeros(n))
    ml.cycle(x, ra.ParVector.from_numpy(ctx, b), x_is_zero=True)
    assert np.array_equal(x.numpy(), H.cycle(np.zeros(n), b))
    b32 = b.astype(np.float32)
    x32 = ra.ParVector(ctx, n, np.float32)
    b32d = ra.ParVector.from_numpy(ctx, b32)
    ml.cycle_fp32(x32, b32d)
    assert np.array_equal(x32.numpy(), H.cycle_fp32(np.zeros(n, np.float32), b32))


@pytest.mark.parametrize("mode", [1, 2], ids=["cols32", "offsets16"])
def test_vcycle_sliced_bit_exact(problem, ctx, mode):
    """Every CSR level operator switched to the sliced-row kernel, with 32-bit
    columns or 16-bit column offsets (the Jacobi sweeps then read the level's
    diagonal array): FP64 and FP32 cycles still bit-identical to the oracle.
    The tuned choices are restored."""
    ml, As, Ps = problem["ml"], problem["As"], problem["Ps"]
    ml.enable_fp32()
    before = ml.row_caps()
    mats = []
    for l in range(ml.num_levels):
        A = ml.level(l)[0]
        if A.diagonals() == 0 and A.nnz > 0:
            mats.append((l, A))
            for p in (0, 1):
                A.set_sliced(p, mode)
    if not mats:
        pytest.skip("no CSR level operator")
    assert np.all(ml.row_caps()[[l for l, _ in mats], 0, :] == -mode)
    ml.set_tail(-1)  # every level through the switched kernels
    try:
        H = K.Hier(As, Ps, relax_type=0, sweeps=1, omega=OMEGA)
        n = As[0].n_rows
        b = np.random.default_rng(24).standard_normal(n)
        x = ra.ParVector.from_numpy(ctx, np.zeros(n))
        ml.cycle(x, ra.ParVector.from_numpy(ctx, b), x_is_zero=True)
        assert np.array_equal(x.numpy(), H.cycle(np.zeros(n), b))
        b32 = b.astype(np.float32)
        x32 = ra.ParVector(ctx, n, np.float32)
        b32d = ra.ParVector.from_numpy(ctx, b32)
        ml.cycle_fp32(x32, b32d)
        assert np.array_equal(x32.numpy(), H.cycle_fp32(np.zeros(n, np.float32), b32))
    finally:
        for l, A in mats:
            for p in (0, 1):
                set_kernel(A, p, before[l, 0, p])
        ml.set_tail(0)
    assert np.array_equal(ml.row_caps(), before)


def test_vcycle_lane_rows_bit_exact(problem, ctx):
    """Every CSR level operator and interpolation switched to the
    lane-per-row kernel over the CSR arrays: FP64 and FP32 cycles still
    bit-identical to the oracle.  The tuned choices are restored."""
    ml, As, Ps = problem["ml"], problem["As"], problem["Ps"]
    ml.enable_fp32()
    before = ml.row_caps()
    mats = []
    for l in range(ml.num_levels):
        A, P = ml.level(l)
        for m, M in ((0, A), (1, P)):
            if M is not None and M.diagonals() == 0 and M.nnz > 0:
                mats.append((l, m, M))
                for p in (0, 1):
                    M.set_lane_rows(p, True)
    if not mats:
        pytest.skip("no CSR operator")
    caps = ml.row_caps()
    assert all(np.all(caps[l, m, :] == 1) for l, m, _ in mats)
    ml.set_tail(-1)  # every level through the switched kernels
    try:
        H = K.Hier(As, Ps, relax_type=0, sweeps=1, omega=OMEGA)
        n = As[0].n_rows
        b = np.random.default_rng(25).standard_normal(n)
        x = ra.ParVector.from_numpy(ctx, np.zeros(n))
        ml.cycle(x, ra.ParVector.from_numpy(ctx, b), x_is_zero=True)
        assert np.array_equal(x.numpy(), H.cycle(np.zeros(n), b))
        b32 = b.astype(np.float32)
        x32 = ra.ParVector(ctx, n, np.float32)
        b32d = ra.ParVector.from_numpy(ctx, b32)
        ml.cycle_fp32(x32, b32d)
        assert np.array_equal(x32.numpy(), H.cycle_fp32(np.zeros(n, np.float32), b32))
    finally:
        for l, m, M in mats:
            for p in (0, 1):
                set_kernel(M, p, before[l, m, p])
        ml.set_tail(0)
    assert np.array_equal(ml.row_caps(), before)


def test_vcycle_sorted_gather_bit_exact(problem, ctx):
    """Every restriction R = P^T switched to the row-block kernel with its
    entries read in sorted-gather order: FP64 and FP32 cycles still
    bit-identical to the oracle.  The tuned choices are restored."""
    ml, As, Ps = problem["ml"], problem["As"], problem["Ps"]
    ml.enable_fp32()
    before = ml.row_caps()
    mats = []
    for l in range(ml.num_levels - 1):
        R = ml.restriction(l)
        if R is not None and R.nnz > 0 and R.diagonals() == 0:
            mats.append((l, 2, R))
    if not mats:
        pytest.skip("no CSR restriction")
    for _, _, M in mats:
        for p in (0, 1):
            set_kernel(M, p, 0)
            M.set_sorted_gather(p, True)
    caps = ml.row_caps()
    assert all(np.all(caps[l, m, :] < -2) for l, m, _ in mats)
    ml.set_tail(-1)  # every level through the switched kernels
    try:
        H = K.Hier(As, Ps, relax_type=0, sweeps=1, omega=OMEGA)
        n = As[0].n_rows
        b = np.random.default_rng(28).standard_normal(n)
        x = ra.ParVector.from_numpy(ctx, np.zeros(n))
        ml.cycle(x, ra.ParVector.from_numpy(ctx, b), x_is_zero=True)
        assert np.array_equal(x.numpy(), H.cycle(np.zeros(n), b))
        b32 = b.astype(np.float32)
        x32 = ra.ParVector(ctx, n, np.float32)
        ml.cycle_fp32(x32, ra.ParVector.from_numpy(ctx, b32))
        assert np.array_equal(x32.numpy(), H.cycle_fp32(np.zeros(n, np.float32), b32))
    finally:
        for l, m, M in mats:
            for p in (0, 1):
                set_kernel(M, p, before[l, m, p])
        ml.set_tail(0)
    assert np.array_equal(ml.row_caps(), before)


def test_vcycle_wave_rows_bit_exact(problem, ctx):
    """Every CSR level operator and interpolation switched to the
    wavefront-per-row kernel: FP64 and FP32 cycles still bit-identical to
    the oracle.  The tuned choices are restored."""
    ml, As, Ps = problem["ml"], problem["As"], problem["Ps"]
    ml.enable_fp32()
    before = ml.row_caps()
    mats = []
    for l in range(ml.num_levels):
        A, P = ml.level(l)
        if A.diagonals() == 0 and A.nnz > 0:
            mats.append((l, 0, A))
        if P is not None and P.nnz > 0:
            mats.append((l, 1, P))
    if not mats:
        pytest.skip("no CSR operator")
    for _, _, M in mats:
        for p in (0, 1):
            M.set_wave_rows(p, True)
    assert all(np.all(ml.row_caps()[l, m, :] == 2) for l, m, _ in mats)
    ml.set_tail(-1)  # every level through the switched kernels
    try:
        H = K.Hier(As, Ps, relax_type=0, sweeps=1, omega=OMEGA)
        n = As[0].n_rows
        b = np.random.default_rng(26).standard_normal(n)
        x = ra.ParVector.from_numpy(ctx, np.zeros(n))
        ml.cycle(x, ra.ParVector.from_numpy(ctx, b), x_is_zero=True)
        assert np.array_equal(x.numpy(), H.cycle(np.zeros(n), b))
        b32 = b.astype(np.float32)
        x32 = ra.ParVector(ctx, n, np.float32)
        ml.cycle_fp32(x32, ra.ParVector.from_numpy(ctx, b32))
        assert np.array_equal(x32.numpy(), H.cycle_fp32(np.zeros(n, np.float32), b32))
    finally:
        for l, m, M in mats:
            for p in (0, 1):
                set_kernel(M, p, before[l, m, p])
        ml.set_tail(0)
    assert np.array_equal(ml.row_caps(), before)


def test_fpcg_sliced_outer_operator(ctx):
    """FPCG on a CSR operator (a coarse Galerkin operator with long rows)
    gives the same iterates with the outer product on the sliced kernel
    (fused p.Ap partials per workgroup) as on the row-block kernel."""
    rp, col, val = ra.gallery_laplace27(14)
    ml0 = build(ctx, rp, col, val, rs(ra.PMIS, ra.Extended))
    A1 = ml0.level(1)[0]
    r1, c1, v1 = A1.to_csr()
    ml = build(ctx, r1, c1, v1, rs(ra.PMIS, ra.Extended))
    A = ra.ParCSRMatrix.from_csr(ctx, r1, c1, v1)
    assert A.diagonals() == 0
    n = r1.size - 1
    b = np.random.default_rng(25).standard_normal(n)
    out = []
    for sliced in (False, True):
        A.set_sliced(0, sliced)
        x = ra.ParVector.from_numpy(ctx, np.zeros(n))
        hist, it, rel = ra.FPCG(A, ml, x, ra.ParVector.from_numpy(ctx, b), 1e-8, 200, ra.PREC_FP64)
        out.append((x.numpy(), it, rel))
    assert out[0][1] == out[1][1] and out[0][2] == out[1][2]
    assert np.array_equal(out[0][0], out[1][0])


@pytest.mark.parametrize("prec", [ra.PREC_FP64, ra.PREC_FP32])
def test_fpcg_fp32_outer_switch(ctx, prec):
    """FPCG on a stencil operator whose values are exact in FP32, with its FP64
    products reading the FP32 value copy (the default) or the FP64 values:
    with both launches on the same workgroup shape the iterates are
    bit-identical (the products are the same and p.Ap is summed over the
    same workgroups); with different shapes (4 and 8 planes per workgroup)
    only p.Ap's rounding differs: same iteration count, results equal to
    rounding."""
    rp, col, val = ra.gallery_laplace27(24)
    ml = build(ctx, rp, col, val, rs(ra.PMIS, ra.Extended))
    A = ra.ParCSRMatrix.from_csr(ctx, rp, col, val)
    assert A.diagonals() > 0
    n = rp.size - 1
    b = np.random.default_rng(27).standard_normal(n)

    def run(on, pl64, pl32):
        A.set_fp32_outer(on)
        A.set_dia_staging(0, 2, pl64)
        A.set_dia_staging(2, 2, pl32)
        x = ra.ParVector.from_numpy(ctx, np.zeros(n))
        hist, it, rel = ra.FPCG(A, ml, x, ra.ParVector.from_numpy(ctx, b), 1e-8, 200, prec)
        assert A.dia_staging(0)[:2] == (2, pl64) and A.dia_staging(2)[:2] == (2, pl32)
        return x.numpy(), it, rel

    try:
        ref = run(True, 4, 4)
        for o in (run(False, 4, 4), run(True, 4, 4)):
            assert o[1] == ref[1] and o[2] == ref[2]
            assert np.array_equal(o[0], ref[0])
        o = run(False, 8, 4)
        assert o[1] == ref[1] and abs(o[2] - ref[2]) <= 1e-6 * ref[2]
        assert np.allclose(o[0], ref[0], rtol=0, atol=1e-10 * np.abs(ref[0]).max())
    finally:
        A.set_fp32_outer(True)
        A.set_dia_staging(0, -1)
        A.set_dia_staging(2, -1)
    assert ra._lib.lib().ramd_csr_set_fp32_outer(A.h, 2) != 0  # only 0 or 1


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
    # error against an FP64 oracle solve of the same system to 1e-12 (the
    # oracle's V-cycle, bit-identical to the reference's, in the oracle's
    # flexible PCG): the oracle solution is that accurate, and the device
    # solution is within the solver tolerance's reach of it
    H = K.Hier(problem["As"], problem["Ps"], relax_type=0, sweeps=1, omega=OMEGA)
    xo, ito, relo = K.MtHier(H, 4).fpcg(A, np.zeros(n), b, 1e-12, 500)
    assert relo <= 1e-12
    assert np.linalg.norm(b - K.spmv(A, xo)) <= 1.0000001e-12 * np.linalg.norm(b)
    assert np.linalg.norm(xo - xt) <= 1e-7 * np.linalg.norm(xt)
    assert np.linalg.norm(xs - xo) <= 1e-5 * np.linalg.norm(xo)
    assert np.linalg.norm(xs - xt) / np.linalg.norm(xt) < 1e-5
    if prec == ra.PREC_FP64:
        # the same algorithm in FP64: the oracle's iteration count to 1e-8
        _, it8, _ = K.MtHier(H, 4).fpcg(A, np.zeros(n), b, 1e-8, 500)
        assert abs(it8 - it) <= 1, (it8, it)


@pytest.mark.parametrize("prec", [ra.PREC_FP64, ra.PREC_FP32])
def test_fpcg_graph_replay_equals_eager(problem, ctx, prec):
    """The FPCG iteration replayed from captured HIP graphs (first call
    captures, second replays the cached graphs) gives exactly the eager
    iterates: solution, iteration count and residual history."""
    A = K.Csr(problem["rp"], problem["col"], problem["val"])
    n = A.n_rows
    b = np.random.default_rng(27).standard_normal(n)
    Ad = ra.ParCSRMatrix.from_csr(ctx, A.rowptr, A.col, A.val)
    out = []
    x = ra.ParVector(ctx, n)
    bd = ra.ParVector.from_numpy(ctx, b)
    try:
        for graphs in (False, True, True):
            ctx.set_graphs(graphs)
            x.set_const_value(0.0)
            hist, it, rel = ra.FPCG(Ad, problem["ml"], x, bd, 1e-10, 300, prec)
            out.append((x.numpy(), it, rel, np.array(hist)))
    finally:
        ctx.set_graphs(False)
    for x, it, rel, hist in out[1:]:
        assert it == out[0][1] and rel == out[0][2]
        assert np.array_equal(hist, out[0][3]) and np.array_equal(x, out[0][0])


def test_pcg_indefinite_reports_error(ctx):
    rp, col, val = ra.gallery_aniso2d(30, 1.0, 0.0)
    ml = build(ctx, rp, col, val, rs(ra.CLJP, ra.ModClassical))
    A = ra.ParCSRMatrix.from_csr(ctx, rp, col, -np.asarray(val))
    n = rp.size - 1
    x = ra.ParVector.from_numpy(ctx, np.zeros(n))
    with pytest.raises(ra.RamdError) as e:
        ra.PCG(A, ml, x, ra.ParVector.from_numpy(ctx, np.ones(n)), tol=1e-8)
    assert e.value.code == -4


def test_fpcg_nan_rhs_reports_error(ctx):
    rp, col, val = ra.gallery_aniso2d(30, 1.0, 0.0)
    ml = build(ctx, rp, col, val, rs(ra.CLJP, ra.ModClassical))
    A = ra.ParCSRMatrix.from_csr(ctx, rp, col, val)
    n = rp.size - 1
    b = np.ones(n)
    b[7] = np.nan
    x = ra.ParVector.from_numpy(ctx, np.zeros(n))
    with pytest.raises(ra.RamdError) as e:
        ra.FPCG(A, ml, x, ra.ParVector.from_numpy(ctx, b), 1e-8, 50)
    assert e.value.code in (-7, -5)


# ---------------------------------------------------------------------------
# smoothed-aggregation prolongation (device products) in isolation
# ---------------------------------------------------------------------------
def test_sa_prolongation_golden(golden, ctx):
    """aggregation/tests/test_par_smoothed_aggregation.cpp: jacobi_prolongation
    of sas_A0 with sas_T0 (omega 4/3, one step) gives sas_P0, and of sas_A1
    with sas_T1 gives sas_P1 (pattern exact, values 1e-10)."""
    for a, t, p in (("sas_A0", "sas_T0", "sas_P0"), ("sas_A1", "sas_T1", "sas_P1")):
        A = K.move_diag(K.fixture_csr(golden, a))
        T = K.fixture_csr(golden, t)
        prp, pcol, pval = ra.jacobi_prolongation(ctx, A.rowptr, A.col, A.val, T.rowptr, T.col, T.val, T.n_cols)
        E = K.fixture_csr(golden, p)
        M = K.sort_rows(K.Csr(prp, pcol, pval, T.n_cols))
        assert np.array_equal(M.rowptr, E.rowptr) and np.array_equal(M.col, E.col), a
        assert np.allclose(M.val, E.val, rtol=0, atol=1e-10), a


@needs_ref
@pytest.mark.parametrize("steps", [1, 2])
def test_sa_prolongation_bit_exact_vs_ref(ctx, steps):
    L = K.ref()
    for rp, col, val in (ra.gallery_fv7(14, seed=5), ra.gallery_aniso2d(40, 0.001, math.pi / 8)):
        A = K.move_diag(K.sort_rows(K.Csr(rp, col, val)))
        n = A.n_rows
        srp, scol = ra.strength(A.rowptr, A.col, A.val, ra.Symmetric, 0.08)
        st = ra.mis2(srp, scol, np.random.default_rng(1).random(n))
        agg, na = ra.aggregate(A.rowptr, A.col, A.val, srp, scol, st)
        trp, tcol, tval, _ = ra.fit_candidates(agg, na, np.ones(n))
        prp, pcol, pval = ra.jacobi_prolongation(ctx, A.rowptr, A.col, A.val, trp, tcol, tval, na, 4.0 / 3, steps)
        nnz = L.ref_jacobi_prolongation(n, K.ip(A.rowptr), K.ip(A.col), K.dp(A.val), na, K.ip(trp), K.ip(tcol),
                                        K.dp(tval), C.c_double(4.0 / 3), steps)
        E = K.ref_fetch(nnz)
        assert np.array_equal(prp, E[0]) and np.array_equal(pcol, E[1]) and np.array_equal(pval, E[2])


def test_cg_golden_history(golden, ctx):
    """krylov/tests/test_par_cg.cpp: CG on the 50x50 anisotropic operator,
    b = A*1, x0 = 0: the residual history of the reference (cg_res.txt,
    161 entries, scaled by ||b|| there) and the oracle's, iteration for
    iteration."""
    rp, col, val = ra.gallery_aniso2d(50, 0.001, math.pi / 8)
    A = ra.ParCSRMatrix.from_csr(ctx, rp, col, val)
    Ah = K.Csr(rp, col, val)
    n = Ah.n_rows
    b = K.spmv(Ah, np.ones(n))
    x = ra.ParVector.from_numpy(ctx, np.zeros(n))
    res, it = ra.CG(A, x, ra.ParVector.from_numpy(ctx, b))
    ref = golden["cg_res"]
    assert len(res) == len(ref)
    assert np.allclose(res * K.norm2(b), ref, rtol=0, atol=1e-6)
    xo, reso, ito = K.cg(Ah, np.zeros(n), b)
    # CG amplifies the rounding of the (differently ordered) inner products
    assert it == ito and np.allclose(res * K.norm2(b), reso * K.norm2(b), rtol=0, atol=1e-6)


def test_row_cap_autotuner_choices(problem):
    """Every CSR operator of the cycle got one of the caps the kernels are
    built for (FP32-only cap 8192 never for FP64 vectors); the bit-exact
    V-cycle tests above ran with these choices."""
    ml = problem["ml"]
    ml.enable_fp32()
    caps = ml.row_caps()
    assert caps.shape == (ml.num_levels, 3, 2)
    # -1/-2 sliced, 1 lane, 2 wave rows, minus a cap: sorted-gather order
    assert set(np.unique(caps)) <= {-8192, -4096, -2048, -1024, -2, -1, 0, 1, 2, 1024, 2048, 4096, 8192}
    assert not np.any(np.abs(caps[:, :, 0]) == 8192)
    assert np.all(caps[:, :2, :] >= -2)  # sorted gathers only for restrictions
    assert np.all(caps[-1, 1:] == 0)  # no P, R at the coarsest level


def test_hierarchy_save_load_round_trip(problem, ctx, tmp_path):
    """A saved and reloaded hierarchy has the same levels, splittings and
    coarse factors, and cycles bit-identically."""
    ml = problem["ml"]
    path = tmp_path / "h.bin"
    ml.save(path)
    ml2 = type(ml).load(ctx, path)
    assert ml2.num_levels == ml.num_levels
    for l in range(ml.num_levels):
        A1, P1 = ml.level(l)
        A2, P2 = ml2.level(l)
        for M1, M2 in ((A1, A2), (P1, P2)):
            if M1 is None:
                assert M2 is None
                continue
            a, b = M1.to_csr(), M2.to_csr()
            assert all(np.array_equal(u, v) for u, v in zip(a, b))
        if l < ml.num_levels - 1:
            assert np.array_equal(ml.states(l), ml2.states(l))
    n = problem["As"][0].n_rows
    b = np.random.default_rng(31).standard_normal(n)
    x1 = ra.ParVector.from_numpy(ctx, np.zeros(n))
    x2 = ra.ParVector.from_numpy(ctx, np.zeros(n))
    ml.cycle(x1, ra.ParVector.from_numpy(ctx, b), x_is_zero=True)
    ml2.cycle(x2, ra.ParVector.from_numpy(ctx, b), x_is_zero=True)
    assert np.array_equal(x1.numpy(), x2.numpy())
    ml2.free()

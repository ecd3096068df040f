"""GPU parity of the SpMV family, the reductions and the Galerkin SpGEMM,
through the C ABI, against the oracle (bit-exact wherever the reference's
per-row order is reproduced) and the reference's golden data
(util/tests/test_par_spmv_*.cpp, ruge_stuben/tests/test_par_rap.cpp)."""
import math

import numpy as np
import pytest

import checkers as K
import raptor_amd as ra

pytestmark = pytest.mark.gpu


def dev(ctx, a):
    return ra.ParVector.from_numpy(ctx, np.asarray(a, dtype=np.float64))


def to_dev(ctx, A):
    return ra.ParCSRMatrix.from_csr(ctx, A.rowptr, A.col, A.val, A.n_cols)


def random_csr(rng, n_rows, n_cols, avg, long_rows=()):
    rp, col, val = [0], [], []
    for i in range(n_rows):
        k = min(n_cols, rng.poisson(avg) if i not in long_rows else n_cols)
        c = np.sort(rng.choice(n_cols, size=k, replace=False))
        col += list(c)
        val += list(rng.standard_normal(k))
        rp.append(len(col))
    return K.Csr(rp, col, val, n_cols)


def cases():
    rng = np.random.default_rng(11)
    out = [
        ("aniso64", K.move_diag(K.Csr(*ra.gallery_aniso2d(64, 0.001, math.pi / 8)))),
        ("lap27_20", K.Csr(*ra.gallery_laplace27(20))),
        ("rect", random_csr(rng, 700, 333, 9)),
        ("long_rows", random_csr(rng, 300, 9000, 4, long_rows=(5, 77, 299))),
        ("empty_rows", K.Csr([0, 0, 2, 2, 3], [1, 3, 0], [1.5, -2.0, 4.0], 4)),
    ]
    return out


@pytest.mark.parametrize("name,A", cases(), ids=[c[0] for c in cases()])
def test_spmv_family_bit_exact(ctx, name, A):
    rng = np.random.default_rng(12)
    x = rng.standard_normal(A.n_cols)
    xr = rng.standard_normal(A.n_rows)
    b = rng.standard_normal(A.n_rows)
    D = to_dev(ctx, A)
    y = ra.ParVector(ctx, A.n_rows)
    D.mult(dev(ctx, x), y)
    assert np.array_equal(y.numpy(), K.spmv(A, x))
    y2 = dev(ctx, b)
    D.mult_append(dev(ctx, x), y2)
    assert np.array_equal(y2.numpy(), K.spmv_append(A, x, b))
    r = ra.ParVector(ctx, A.n_rows)
    D.residual(dev(ctx, x), dev(ctx, b), r)
    assert np.array_equal(r.numpy(), K.residual(A, x, b))
    t = ra.ParVector(ctx, A.n_cols)
    D.mult_T(dev(ctx, xr), t)
    assert np.array_equal(t.numpy(), K.spmv_t(A, xr))


def sliced_cases():
    rng = np.random.default_rng(31)
    sq = random_csr(rng, 900, 900, 20)
    # every row gets a diagonal; a few are (near) zero, one row is emptied
    rp, col, val = [0], [], []
    for i in range(sq.n_rows):
        c = list(sq.col[sq.rowptr[i]:sq.rowptr[i + 1]])
        v = list(sq.val[sq.rowptr[i]:sq.rowptr[i + 1]])
        if i == 17:
            c, v = [], []
        elif i not in c:
            c.append(i)
            v.append(0.0 if i % 97 == 5 else 30.0 + i % 7)
        o = np.argsort(c)
        col += [c[j] for j in o]
        val += [v[j] for j in o]
        rp.append(len(col))
    S = K.Csr(rp, col, val, 900)
    # rows longer than the wavefront kernel's 512-entry LDS chunk, each with
    # its diagonal somewhere past the first chunk
    L = random_csr(rng, 1500, 1500, 40, long_rows=(3, 700, 1499))
    rp, col, val = [0], [], []
    for i in range(L.n_rows):
        c = list(L.col[L.rowptr[i]:L.rowptr[i + 1]])
        v = list(L.val[L.rowptr[i]:L.rowptr[i + 1]])
        if i not in c:
            c.append(i)
            v.append(50.0)
        o = np.argsort(c)
        col += [c[j] for j in o]
        val += [v[j] for j in o]
        rp.append(len(col))
    SL = K.Csr(rp, col, val, 1500)
    return [("square_sorted", S), ("square_diag_first", K.move_diag(S)),
            ("square_long_rows", SL), ("square_long_rows_diag_first", K.move_diag(SL)),
            ("rect", random_csr(rng, 700, 333, 9)),
            ("long_rows", random_csr(rng, 300, 9000, 4, long_rows=(5, 77, 299))),
            ("empty_rows", K.Csr([0, 0, 2, 2, 3], [1, 3, 0], [1.5, -2.0, 4.0], 4))]


@pytest.mark.parametrize("layout", ["sliced", "sliced16", "lane_rows", "wave_rows"])
@pytest.mark.parametrize("name,A", sliced_cases(), ids=[c[0] for c in sliced_cases()])
def test_sliced_kernel_bit_exact(ctx, name, A, layout):
    """The one-lane-per-row kernels -- over the sliced form (64 rows per
    slice, padding dropped by a select) and straight over the CSR rows --
    and the wavefront-per-row kernel (rows longer than its 512-entry LDS
    chunk included) give the oracle's bits for every operation."""
    rng = np.random.default_rng(32)
    x, b = rng.standard_normal(A.n_cols), rng.standard_normal(A.n_rows)
    D = to_dev(ctx, A)
    if layout == "sliced":
        D.set_sliced(0, True)
    elif layout == "sliced16":
        if A.n_rows != A.n_cols:  # offsets col - row need a square operator
            with pytest.raises(ra.RamdError):
                D.set_sliced(0, 2)
            return
        D.set_sliced(0, 2)
    elif layout == "lane_rows":
        D.set_lane_rows(0, True)
    else:
        D.set_wave_rows(0, True)
    y = ra.ParVector(ctx, A.n_rows)
    D.mult(dev(ctx, x), y)
    assert np.array_equal(y.numpy(), K.spmv(A, x))
    y2 = dev(ctx, b)
    D.mult_append(dev(ctx, x), y2)
    assert np.array_equal(y2.numpy(), K.spmv_append(A, x, b))
    r = ra.ParVector(ctx, A.n_rows)
    D.residual(dev(ctx, x), dev(ctx, b), r)
    assert np.array_equal(r.numpy(), K.residual(A, x, b))
    if A.n_rows == A.n_cols:
        xj = dev(ctx, x)
        D.jacobi(xj, dev(ctx, b), ra.ParVector(ctx, A.n_rows), 2, 2.0 / 3.0)
        assert np.array_equal(xj.numpy(), K.jacobi(A, x, b, 2, 2.0 / 3.0))
    D.set_sliced(0, False)
    D.set_lane_rows(0, False)
    D.set_wave_rows(0, False)
    y3 = ra.ParVector(ctx, A.n_rows)
    D.mult(dev(ctx, x), y3)
    assert np.array_equal(y3.numpy(), y.numpy())


@pytest.mark.parametrize("name,A", cases(), ids=[c[0] for c in cases()])
def test_sorted_gather_bit_exact(ctx, name, A):
    """The row-block SpMV with each block's entries read in column order
    (sorted-gather order; rows longer than the block cap -- the long_rows
    case -- keep their stored order) gives the oracle's bits, before and
    after the values change, and the stored-order kernel again once it is
    released."""
    rng = np.random.default_rng(33)
    x = rng.standard_normal(A.n_cols)
    D = to_dev(ctx, A)
    if D.diagonals():  # the stencil-diagonal form has no row blocks to reorder
        with pytest.raises(ra.RamdError):
            D.set_sorted_gather(0, True)
        return
    D.set_sorted_gather(0, True)
    y = ra.ParVector(ctx, A.n_rows)
    D.mult(dev(ctx, x), y)
    assert np.array_equal(y.numpy(), K.spmv(A, x))
    # new values: the sorted copy follows
    v2 = rng.standard_normal(A.col.size)
    D.set_values(dev(ctx, v2))
    A2 = K.Csr(A.rowptr, A.col, v2, A.n_cols)
    D.mult(dev(ctx, x), y)
    assert np.array_equal(y.numpy(), K.spmv(A2, x))
    D.set_sorted_gather(0, False)
    y2 = ra.ParVector(ctx, A.n_rows)
    D.mult(dev(ctx, x), y2)
    assert np.array_equal(y2.numpy(), y.numpy())


@pytest.mark.parametrize("mode,planes", [(0, 0), (1, 0), (2, 4), (2, 7)])
@pytest.mark.parametrize("name,gen", [("aniso64", lambda: ra.gallery_aniso2d(64, 0.001, math.pi / 8)),
                                      ("lap27_20", lambda: ra.gallery_laplace27(20)),
                                      ("lap27_16", lambda: ra.gallery_laplace27(16)),
                                      ("fv7_18", lambda: ra.gallery_fv7(18, seed=2)),
                                      ("fv7_32", lambda: ra.gallery_fv7(32, seed=2))])
def test_dia_staging_variants_bit_exact(ctx, name, gen, mode, planes):
    """Every kernel of the stencil-diagonal form (x gathered, x window staged
    in LDS, plane-march over the box grid with 4 or 7 planes per workgroup),
    with FP64 values and with the exact FP32 copy, gives the CSR oracle's
    bits for every operation.  The 2-D operator has no box geometry, so the
    plane-march kernel is refused for it."""
    A = K.move_diag(K.Csr(*gen()))
    D = to_dev(ctx, A)
    assert D.diagonals() > 0
    if mode == 2:
        _, _, L, P = D.dia_staging(0)
        if name.startswith("aniso"):
            assert L == 0
            with pytest.raises(ra.RamdError):
                D.set_dia_staging(0, 2, planes)
            return
        m = round(A.n_rows ** (1 / 3))
        assert (L, P) == (m, m * m)
    for pp in (0, 2):
        D.set_dia_staging(pp, mode, planes)
        assert D.dia_staging(pp)[0] == mode
    rng = np.random.default_rng(14)
    x, b = rng.standard_normal(A.n_rows), rng.standard_normal(A.n_rows)
    y = ra.ParVector(ctx, A.n_rows)
    D.mult(dev(ctx, x), y)
    assert np.array_equal(y.numpy(), K.spmv(A, x))
    r = ra.ParVector(ctx, A.n_rows)
    D.residual(dev(ctx, x), dev(ctx, b), r)
    assert np.array_equal(r.numpy(), K.residual(A, x, b))
    y2 = dev(ctx, b)
    D.mult_append(dev(ctx, x), y2)
    assert np.array_equal(y2.numpy(), K.spmv_append(A, x, b))
    xj = dev(ctx, x)
    D.jacobi(xj, dev(ctx, b), ra.ParVector(ctx, A.n_rows), 2, 2.0 / 3.0)
    assert np.array_equal(xj.numpy(), K.jacobi(A, x, b, 2, 2.0 / 3.0))


@pytest.mark.parametrize("name,A", cases()[:2], ids=[c[0] for c in cases()[:2]])
def test_jacobi_bit_exact(ctx, name, A):
    rng = np.random.default_rng(13)
    x0, b = rng.standard_normal(A.n_rows), rng.standard_normal(A.n_rows)
    D = to_dev(ctx, A)
    x = dev(ctx, x0)
    tmp = ra.ParVector(ctx, A.n_rows)
    D.jacobi(x, dev(ctx, b), tmp, 2, 2.0 / 3.0)
    assert np.array_equal(x.numpy(), K.jacobi(A, x0, b, 2, 2.0 / 3.0))


@pytest.mark.parametrize("name", ["laplacian27", "aniso", "random"])
def test_spmv_golden(ctx, golden, name):
    A = K.fixture_csr(golden, name)
    D = to_dev(ctx, A)
    y = ra.ParVector(ctx, A.n_rows)
    D.mult(dev(ctx, np.arange(A.n_cols, dtype=float)), y)
    assert np.allclose(y.numpy(), golden[name + "_inc_b"], rtol=1e-12, atol=1e-10)
    t = ra.ParVector(ctx, A.n_cols)
    D.mult_T(dev(ctx, np.ones(A.n_rows)), t)
    assert np.allclose(t.numpy(), golden[name + "_ones_b_T"], rtol=1e-12, atol=1e-10)


def test_reductions_deterministic(ctx):
    rng = np.random.default_rng(14)
    for n in (1, 255, 256, 100_000, 3_000_001):
        a, b = rng.standard_normal(n), rng.standard_normal(n)
        da, db = dev(ctx, a), dev(ctx, b)
        d1, d2 = da.inner_product(db), da.inner_product(db)
        assert d1 == d2
        exact = math.fsum(a * b)
        assert abs(d1 - exact) <= 1e-14 * n * np.abs(a * b).sum() / max(n, 1) + 1e-13 * np.abs(a * b).sum()
        nr = da.norm(2)
        assert math.isclose(nr, math.sqrt(math.fsum(a * a)), rel_tol=1e-13)
    # integers: exact in any order
    x = np.arange(1, 1001, dtype=float)
    assert dev(ctx, x).inner_product(dev(ctx, x)) == 333833500.0


@pytest.mark.parametrize("level", [0, 1])
def test_galerkin_golden_and_bit_exact(ctx, golden, level):
    A = K.move_diag(K.fixture_csr(golden, f"rss_A{level}"))
    P = K.fixture_csr(golden, f"rss_P{level}")
    Ac = to_dev(ctx, A).galerkin(to_dev(ctx, P))
    rp, col, val = Ac.to_csr()
    O = K.galerkin(A, P)
    assert np.array_equal(rp, O.rowptr) and np.array_equal(col, O.col) and np.array_equal(val, O.val)
    E = K.move_diag(K.fixture_csr(golden, f"rss_A{level + 1}"))
    assert np.array_equal(rp, E.rowptr) and np.array_equal(col, E.col)
    assert np.allclose(val, E.val, rtol=0, atol=1e-12)


def test_spgemm_bit_exact_bins(ctx):
    """Rows spanning every hash-table bin, including the global-memory path."""
    rng = np.random.default_rng(15)
    A = random_csr(rng, 400, 3000, 6, long_rows=(3, 200))
    B = random_csr(rng, 3000, 5000, 5, long_rows=(17,))
    C = to_dev(ctx, A).mult(to_dev(ctx, B))
    rp, col, val = C.to_csr()
    O = K.spgemm(A, B)
    assert np.array_equal(rp, O.rowptr) and np.array_equal(col, O.col) and np.array_equal(val, O.val)


def test_spgemm_tn_bit_exact(ctx):
    rng = np.random.default_rng(16)
    P = random_csr(rng, 2000, 300, 4)
    B = random_csr(rng, 2000, 700, 12)
    C = to_dev(ctx, B).mult_T(to_dev(ctx, P))
    rp, col, val = C.to_csr()
    O = K.spgemm_tn(P, B)
    assert np.array_equal(rp, O.rowptr) and np.array_equal(col, O.col) and np.array_equal(val, O.val)


def test_spgemm_cancellation_dropped(ctx):
    A = K.Csr([0, 2], [0, 1], [1.0, 1.0], 2)
    B = K.Csr([0, 2, 4], [0, 1, 0, 1], [1.0, 1e-16, -1.0, 0.0], 2)
    C = to_dev(ctx, A).mult(to_dev(ctx, B))
    assert C.nnz == 0

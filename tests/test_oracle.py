"""Pins the CPU oracle (oracle/oracle.c) before it is trusted as the checker.

Each oracle function is checked against (a) the reference's own golden
vectors from test_data, (b) the reference compiled from source
(oracle/_ref/libraptor_ref.so) where it is available, and (c) closed forms /
brute force chosen so that a dropped term, a wrong sign or a transposed
operand fails.  Mirrors the reference tests util/tests/test_par_spmv_*.cpp,
krylov/tests/test_par_cg.cpp and ruge_stuben/tests/test_par_rap.cpp.
"""
import ctypes as C
import math

import numpy as np
import pytest

import checkers as K


def stencil_grid_np(stencil, grid):
    """Independent numpy statement of par_stencil_grid for symmetric
    stencils (gallery/par_stencil.hpp:16-236): neighbours inside the grid,
    columns ascending, value of the reversed nonzero-stencil entry."""
    dim = len(grid)
    ext = list(reversed(grid))
    strides = [1]
    for k in range(1, dim):
        strides.append(strides[-1] * ext[k - 1])
    offs, nz = [], []
    for i in range(3 ** dim):
        if abs(stencil[i]) > 1e-16:
            digits = [(i // 3 ** k) % 3 - 1 for k in range(dim)]
            offs.append(digits)
            nz.append(stencil[i])
    n = int(np.prod(grid))
    rp, col, val = [0], [], []
    for r in range(n):
        coord = [(r // strides[k]) % ext[k] for k in range(dim)]
        for d, o in enumerate(offs):
            if all(0 <= coord[k] + o[k] < ext[k] for k in range(dim)):
                col.append(r + sum(strides[k] * o[k] for k in range(dim)))
                val.append(nz[len(nz) - 1 - d])
        rp.append(len(col))
    return K.Csr(rp, col, val)


def lap27():
    s = [-1.0] * 27
    s[13] = 26.0
    return s


def diffusion_2d(eps, theta):
    C_, S_ = math.cos(theta), math.sin(theta)
    CS, CC, SS = C_ * S_, C_ * C_, S_ * S_
    v1 = ((-1 * eps - 1) * CC + (-1 * eps - 1) * SS + (3 * eps - 3) * CS) / 6.0
    v2 = ((2 * eps - 4) * CC + (-4 * eps + 2) * SS) / 6.0
    v3 = ((-1 * eps - 1) * CC + (-1 * eps - 1) * SS + (-3 * eps + 3) * CS) / 6.0
    v4 = ((-4 * eps + 2) * CC + (2 * eps - 4) * SS) / 6.0
    v5 = ((8 * eps + 8) * CC + (8 * eps + 8) * SS) / 6.0
    return [v1, v2, v3, v4, v5, v4, v3, v2, v1]


def rand_int_csr(rng, n_rows, n_cols, density=0.3, diag=False):
    d = (rng.random((n_rows, n_cols)) < density) * rng.integers(-4, 5, (n_rows, n_cols))
    if diag:
        for i in range(min(n_rows, n_cols)):
            d[i, i] = 10 + abs(d[i]).sum()
    rp, col, val = [0], [], []
    for i in range(n_rows):
        for j in range(n_cols):
            if d[i, j] != 0:
                col.append(j)
                val.append(float(d[i, j]))
        rp.append(len(col))
    return K.Csr(rp, col, val, n_cols), d.astype(float)


# ---------------------------------------------------------------------------
# (a) golden vectors
# ---------------------------------------------------------------------------
def test_stencil_generator_matches_fixtures(golden):
    A = K.fixture_csr(golden, "laplacian27")
    B = stencil_grid_np(lap27(), [10, 10, 10])
    assert np.array_equal(A.rowptr, B.rowptr) and np.array_equal(A.col, B.col)
    assert np.allclose(A.val, B.val, rtol=0, atol=1e-14)
    A = K.fixture_csr(golden, "aniso")
    B = stencil_grid_np(diffusion_2d(0.001, math.pi / 8), [25, 25])
    assert np.array_equal(A.rowptr, B.rowptr) and np.array_equal(A.col, B.col)
    assert np.allclose(A.val, B.val, rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("name", ["laplacian27", "aniso", "random"])
def test_spmv_golden(golden, name):
    """util/tests/test_par_spmv_{laplacian,aniso,random}.cpp"""
    A = K.fixture_csr(golden, name)
    ones_r, ones_c = np.ones(A.n_cols), np.ones(A.n_rows)
    inc_c, inc_r = np.arange(A.n_cols, dtype=float), np.arange(A.n_rows, dtype=float)
    tol = dict(rtol=1e-12, atol=1e-10)
    assert np.allclose(K.spmv(A, ones_r), golden[name + "_ones_b"], **tol)
    assert np.allclose(K.spmv_t(A, ones_c), golden[name + "_ones_b_T"], **tol)
    assert np.allclose(K.spmv(A, inc_c), golden[name + "_inc_b"], **tol)
    assert np.allclose(K.spmv_t(A, inc_r), golden[name + "_inc_b_T"], **tol)


def test_cg_residual_history_golden(golden):
    """krylov/tests/test_par_cg.cpp: 50x50 aniso, b = A*1, x0 = 0, CG(A,x,b,res)."""
    A = stencil_grid_np(diffusion_2d(0.001, math.pi / 8), [50, 50])
    b = K.spmv(A, np.ones(A.n_rows))
    b_norm = K.norm2(b)
    x, res, it = K.cg(A, np.zeros(A.n_rows), b)
    ref = golden["cg_res"]
    assert len(res) == len(ref)
    assert np.allclose(res * b_norm, ref, rtol=0, atol=1e-6)


@pytest.mark.parametrize("level", [0, 1])
def test_galerkin_golden(golden, level):
    """ruge_stuben/tests/test_par_rap.cpp: rss_A{l} and rss_P{l} -> rss_A{l+1}."""
    A = K.move_diag(K.fixture_csr(golden, f"rss_A{level}"))
    P = K.fixture_csr(golden, f"rss_P{level}")
    Ac = K.galerkin(A, P)
    E = K.move_diag(K.fixture_csr(golden, f"rss_A{level + 1}"))
    assert np.array_equal(Ac.rowptr, E.rowptr)
    assert np.array_equal(Ac.col, E.col)
    assert np.allclose(Ac.val, E.val, rtol=0, atol=1e-12)


# ---------------------------------------------------------------------------
# (c) closed forms and brute force
# ---------------------------------------------------------------------------
def test_spmv_family_bruteforce():
    rng = np.random.default_rng(1)
    A, D = rand_int_csr(rng, 37, 29)
    x = rng.integers(-5, 6, 29).astype(float)
    xr = rng.integers(-5, 6, 37).astype(float)
    b = rng.integers(-5, 6, 37).astype(float)
    assert np.array_equal(K.spmv(A, x), D @ x)
    assert np.array_equal(K.spmv_t(A, xr), D.T @ xr)
    assert np.array_equal(K.residual(A, x, b), b - D @ x)
    assert np.array_equal(K.spmv_append(A, x, b), b + D @ x)


def test_jacobi_closed_form():
    # 3x3: x1 = (1-w) x0 + w (b - (A - D) x0) / d, written out by hand
    A = K.Csr([0, 2, 5, 7], [0, 1, 0, 1, 2, 1, 2], [4.0, -1.0, -1.0, 4.0, -2.0, -1.0, 5.0])
    x0 = np.array([1.0, 2.0, 3.0])
    b = np.array([2.0, 1.0, 0.5])
    w = 2.0 / 3.0
    e = np.array([(1 - w) * 1.0 + w * ((2.0 - (-1.0 * 2.0)) / 4.0),
                  (1 - w) * 2.0 + w * ((1.0 - (-1.0 * 1.0 + -2.0 * 3.0)) / 4.0),
                  (1 - w) * 3.0 + w * ((0.5 - (-1.0 * 2.0)) / 5.0)])
    assert np.allclose(K.jacobi(A, x0, b, 1, w), e, rtol=1e-15, atol=0)
    # a zero / missing diagonal leaves the row unchanged (relax.cpp:39)
    Z = K.Csr([0, 1, 2], [1, 1], [3.0, 2.0])
    assert K.jacobi(Z, [7.0, 1.0], [1.0, 1.0], 1, 1.0)[0] == 7.0


def test_dot_norm_closed_form():
    x = np.arange(1, 101, dtype=float)
    assert K.dot(x, x) == 338350.0                     # sum k^2
    assert K.dot(x, np.ones(100)) == 5050.0
    v = np.array([3.0, 4.0, 1e-17, -1e-20])           # entries <= 1e-16 skipped
    assert K.norm2(v) == 5.0


def test_lu_matches_library_solve():
    rng = np.random.default_rng(2)
    for n in (1, 5, 40):
        Ad = rng.standard_normal((n, n)) + n * np.eye(n)
        b = rng.standard_normal(n)
        lu, piv, info = K.lu_factor(Ad)
        assert info == 0
        x = K.lu_solve(lu, piv, b)
        # row-major data read column-major: the factor is of A^T (F4)
        assert np.allclose(x, np.linalg.solve(Ad.T, b), rtol=1e-12, atol=1e-12)
    # singular matrix -> info > 0
    _, _, info = K.lu_factor(np.zeros((3, 3)))
    assert info == 1


def test_galerkin_dense_triple_product():
    rng = np.random.default_rng(3)
    A, Ad = rand_int_csr(rng, 30, 30, 0.2, diag=True)
    P, Pd = rand_int_csr(rng, 30, 9, 0.25)
    Ac = K.galerkin(K.move_diag(A), P)
    assert np.array_equal(Ac.dense(), Pd.T @ Ad @ Pd)
    # diagonal first in every row that has one
    for i in range(Ac.n_rows):
        s, e = Ac.rowptr[i], Ac.rowptr[i + 1]
        if i in Ac.col[s:e]:
            assert Ac.col[s] == i
            assert np.all(np.diff(Ac.col[s + 1:e]) > 0)


def test_spgemm_drop_rule():
    # exact cancellation and a |v| == 1e-16 entry are dropped (> 1e-16 kept)
    A = K.Csr([0, 2], [0, 1], [1.0, 1.0], 2)
    B = K.Csr([0, 2, 4], [0, 1, 0, 1], [1.0, 1e-16, -1.0, 0.0], 2)
    C = K.spgemm(A, B)
    assert C.nnz == 0
    B2 = K.Csr([0, 2, 4], [0, 1, 0, 1], [1.0, 3e-16, -0.5, 0.0], 2)
    C2 = K.spgemm(A, B2)
    assert list(C2.col) == [0, 1] and C2.val[0] == 0.5


def test_pcg_and_cycle_quirks_small():
    """2-level hierarchy built by hand: PCG's first entry is sqrt(r0.z0) and
    the cycle is linear in b (x0 = 0)."""
    A = stencil_grid_np(diffusion_2d(1.0, 0.0), [8, 8])
    A = K.move_diag(A)
    n = A.n_rows
    # aggregation-like P: pairs of rows
    rp = np.arange(n + 1)
    P = K.Csr(rp, np.arange(n) // 2, np.ones(n), n // 2)
    Ac = K.galerkin(A, P)
    H = K.Hier([A, Ac], [P], relax_type=0, sweeps=1, omega=2.0 / 3.0)
    rng = np.random.default_rng(4)
    b1, b2 = rng.standard_normal(n), rng.standard_normal(n)
    z1, z2, z12 = H.cycle(np.zeros(n), b1), H.cycle(np.zeros(n), b2), H.cycle(np.zeros(n), b1 + 2 * b2)
    assert np.allclose(z12, z1 + 2 * z2, rtol=1e-12, atol=1e-12)
    x, res, it = H.pcg(A, np.zeros(n), b1, tol=1e-8)
    z0 = H.cycle(np.zeros(n), b1)
    b_inner = K.dot(b1, z0)
    assert math.isclose(res[0], math.sqrt(b_inner), rel_tol=1e-14)
    # stop rule of krylov/par_cg.cpp:155-161,232: r.z < tol * sqrt(b.z0), res[k] = r.z / b.z0
    assert it > 0 and res[-1] * b_inner < 1e-8 * math.sqrt(b_inner)
    assert res[-2] * b_inner >= 1e-8 * math.sqrt(b_inner)
    assert np.linalg.norm(b1 - A.dense() @ x) / np.linalg.norm(b1) < 1e-4


# ---------------------------------------------------------------------------
# (b) the reference compiled from source
# ---------------------------------------------------------------------------
needs_ref = pytest.mark.skipif(not K.have_ref(), reason="oracle/_ref not built (reference sources absent)")


@needs_ref
def test_ref_serial_jacobi_bit_exact():
    A = K.move_diag(stencil_grid_np(diffusion_2d(0.001, math.pi / 8), [12, 12]))
    rng = np.random.default_rng(5)
    x0, b = rng.standard_normal(A.n_rows), rng.standard_normal(A.n_rows)
    xr = x0.copy()
    K.ref().ref_serial_jacobi(A.n_rows, K.ip(A.rowptr), K.ip(A.col), K.dp(A.val), K.dp(xr), K.dp(b), 2,
                              C.c_double(2.0 / 3.0))
    assert np.array_equal(K.jacobi(A, x0, b, 2, 2.0 / 3.0), xr)


@needs_ref
def test_ref_spgemm_and_rap_bit_exact(golden):
    A = K.fixture_csr(golden, "rss_A0")
    P = K.fixture_csr(golden, "rss_P0")
    L = K.ref()
    nnz = L.ref_spgemm(A.n_rows, A.n_cols, K.ip(A.rowptr), K.ip(A.col), K.dp(A.val), P.n_rows, P.n_cols,
                       K.ip(P.rowptr), K.ip(P.col), K.dp(P.val))
    rp, col, val, _ = K.ref_fetch(nnz)
    C_ = K.spgemm(A, P)
    assert np.array_equal(C_.rowptr, rp) and np.array_equal(C_.col, col) and np.array_equal(C_.val, val)
    nnz = L.ref_rap(A.n_rows, K.ip(A.rowptr), K.ip(A.col), K.dp(A.val), P.n_cols, K.ip(P.rowptr), K.ip(P.col),
                    K.dp(P.val), 1)
    rp, col, val, _ = K.ref_fetch(nnz)
    Ac = K.galerkin(K.move_diag(A), P)
    assert np.array_equal(Ac.rowptr, rp) and np.array_equal(Ac.col, col) and np.array_equal(Ac.val, val)


def _ref_hierarchy(A, theta, coarsen, interp, relax, omega, max_coarse=50):
    L = K.ref()
    nl = L.ref_hier_setup(A.n_rows, K.ip(A.rowptr), K.ip(A.col), K.dp(A.val), C.c_double(theta), coarsen, interp,
                          0, relax, C.c_double(omega), max_coarse, 25, None)
    As, Ps = [], []
    for l in range(nl):
        rp, col, val, nc = K.ref_fetch(L.ref_hier_level(l, 0))
        As.append(K.Csr(rp, col, val, nc))
        if l < nl - 1:
            rp, col, val, nc = K.ref_fetch(L.ref_hier_level(l, 1))
            Ps.append(K.Csr(rp, col, val, nc))
    return As, Ps


@needs_ref
def test_ref_cycle_and_pcg_sor():
    """Oracle V-cycle / stationary solve / PCG driven with the reference's
    default SOR smoother on the reference's own hierarchy (C1 problem):
    same iteration counts, histories equal up to the coarse-LU rounding."""
    A = stencil_grid_np(diffusion_2d(0.001, math.pi / 8), [100, 100])
    As, Ps = _ref_hierarchy(A, 0.25, 1, 1, 1, 1.0)
    H = K.Hier(As, Ps, relax_type=1, sweeps=1, omega=1.0)
    n = A.n_rows
    b = K.spmv(A, np.ones(n))
    L = K.ref()
    xr = np.zeros(n)
    L.ref_hier_cycle(K.dp(xr), K.dp(b))
    xo = H.cycle(np.zeros(n), b)
    assert np.allclose(xo, xr, rtol=1e-10, atol=1e-12 * np.abs(xr).max())
    res_r = np.zeros(101)
    xr = np.zeros(n)
    it_r = L.ref_hier_solve(K.dp(xr), K.dp(b), C.c_double(1e-7), 100, K.dp(res_r))
    _, res_o, it_o = H.solve(np.zeros(n), b)
    assert it_o == it_r
    assert np.allclose(res_o, res_r[:it_r + 1], rtol=1e-8, atol=0)
    res_r = np.zeros(1000)
    xr = np.zeros(n)
    nr = L.ref_pcg(n, K.ip(A.rowptr), K.ip(A.col), K.dp(A.val), K.dp(xr), K.dp(b), C.c_double(1e-8), -1,
                   K.dp(res_r), 1000)
    xo, res_o, it_o = H.pcg(A, np.zeros(n), b, tol=1e-8)
    assert len(res_o) == nr
    assert np.allclose(res_o, res_r[:nr], rtol=1e-6, atol=0)
    assert np.allclose(xo, xr, rtol=1e-8, atol=1e-10)


@needs_ref
@pytest.mark.parametrize("omega", [2.0 / 3.0, 1.0])
def test_cycle_fp32_tracks_fp64(omega):
    """The single-precision cycle follows the FP64 cycle (pinned above) to
    FP32 accuracy on the reference's own C1 hierarchy: a dropped or
    sign-flipped term would leave an O(1) difference, not 1e-6."""
    A = stencil_grid_np(diffusion_2d(0.001, math.pi / 8), [60, 60])
    As, Ps = _ref_hierarchy(A, 0.25, 1, 1, 0, omega)
    assert len(As) >= 3
    H = K.Hier(As, Ps, relax_type=0, sweeps=1, omega=omega)
    b = np.random.default_rng(5).standard_normal(A.n_rows).astype(np.float32)
    x64 = H.cycle(np.zeros(A.n_rows), b.astype(np.float64))
    x32 = H.cycle_fp32(np.zeros(A.n_rows, np.float32), b)
    assert x32.dtype == np.float32
    err = np.linalg.norm(x32 - x64) / np.linalg.norm(x64)
    assert 0 < err < 2e-6, err
    # a nonzero float32 input on the coarsest level alone is solved in FP64
    # and rounded once
    Hc = K.Hier(As[-1:], [], relax_type=0, sweeps=1, omega=omega)
    bc = np.random.default_rng(6).standard_normal(As[-1].n_rows).astype(np.float32)
    xc = Hc.cycle_fp32(np.zeros_like(bc), bc)
    assert np.array_equal(xc, Hc.cycle(np.zeros(bc.size), bc.astype(np.float64)).astype(np.float32))


@needs_ref
def test_ref_coarse_lu_matches_oracle():
    A = stencil_grid_np(diffusion_2d(0.001, math.pi / 8), [40, 40])
    As, Ps = _ref_hierarchy(A, 0.25, 1, 1, 1, 1.0)
    L = K.ref()
    n = L.ref_hier_coarse(None, None)
    lu_r = np.zeros(n * n)
    piv_r = np.zeros(n, np.int32)
    L.ref_hier_coarse(K.dp(lu_r), K.ip(piv_r))
    lu_o, piv_o, info = K.lu_factor(As[-1].dense())
    assert info == 0
    assert np.array_equal(piv_o, piv_r - 1)
    assert np.allclose(lu_o, lu_r, rtol=1e-12, atol=1e-14)


def test_bicgstab_residual_history_golden(golden):
    """krylov/tests/test_bicgstab.cpp / test_par_bicgstab.cpp: 50x50 aniso,
    b = A*1, x0 = 0, BiCGStab(A,x,b,res); the reference checks the first 30
    residuals to 1e-6."""
    A = stencil_grid_np(diffusion_2d(0.001, math.pi / 8), [50, 50])
    b = K.spmv(A, np.ones(A.n_rows))
    x, res, it = K.bicgstab(A, np.zeros(A.n_rows), b)
    ref = golden["bicgstab_res"]
    assert np.allclose(res[:30], ref[:30], rtol=0, atol=1e-6)
    assert len(res) == it + 1


@needs_ref
def test_ref_bicgstab_bit_exact():
    """orc_bicgstab against the reference's own BiCGStab compiled in _ref:
    every residual and the solution bit for bit (same passes, same order)."""
    import ctypes as C
    A = stencil_grid_np(diffusion_2d(0.001, math.pi / 8), [50, 50])
    b = K.spmv(A, np.ones(A.n_rows))
    x, res, it = K.bicgstab(A, np.zeros(A.n_rows), b)
    L = K.ref()
    xr = np.zeros(A.n_rows)
    rr = np.zeros(5000)
    nr = L.ref_bicgstab(A.n_rows, K.ip(A.rowptr), K.ip(A.col), K.dp(A.val), K.dp(xr), K.dp(b), C.c_double(1e-5),
                        -1, K.dp(rr), 5000)
    assert nr == len(res)
    assert np.array_equal(rr[:nr], res)
    assert np.array_equal(xr, x)


@needs_ref
@pytest.mark.parametrize("omega", [2.0 / 3.0])
def test_pre_bicgstab_oracle_invariants(omega):
    """orc_pre_bicgstab (the reference's routine cannot be run: it reads an
    uninitialised counter) on the reference's own C1 hierarchy: the
    recursively updated residual equals b - A x at the end to round-off
    (a dropped or mis-signed term in the x or r update breaks this), the
    stop rule holds, and the AMG preconditioner cuts the iterations."""
    A = stencil_grid_np(diffusion_2d(0.001, math.pi / 8), [60, 60])
    As, Ps = _ref_hierarchy(A, 0.25, 1, 1, 0, omega)
    H = K.Hier(As, Ps, relax_type=0, sweeps=1, omega=omega)
    b = K.spmv(A, np.ones(A.n_rows))
    x, res, it = K.bicgstab(A, np.zeros(A.n_rows), b, tol=1e-8, hier=H)
    true = np.linalg.norm(b - K.spmv(A, x))
    assert res[-1] <= 1e-8 * res[0]
    assert abs(true - res[-1]) <= 1e-6 * res[-1] + 1e-13 * res[0]
    assert np.allclose(x, 1.0, atol=1e-5)
    _, _, it_plain = K.bicgstab(A, np.zeros(A.n_rows), b, tol=1e-8)
    assert it < it_plain / 3


def _colors_oracle(A):
    import ctypes as C
    col = np.empty(A.n_rows, np.int32)
    nc = K.orc().orc_greedy_colors(A.n_rows, K.ip(A.rowptr), K.ip(A.col), K.ip(col))
    return col, nc


def _adjacent_same_colour(A, colour):
    for i in range(A.n_rows):
        for j in A.col[A.rowptr[i]:A.rowptr[i + 1]]:
            if j != i and colour[j] == colour[i]:
                return True
    return False


def test_greedy_colouring_oracle():
    """Greedy colouring of A + A^T: a proper colouring (no two neighbours
    share a colour), 2 colours for the 5-point Laplacian (red-black), 4 for
    the 9-point stencil, 8 for the 27-point one, and a pattern that is only
    one-sided (row 0 -> 3) still separates rows 0 and 3."""
    A5 = stencil_grid_np(np.array([0, -1.0, 0, -1, 4, -1, 0, -1, 0]), [9, 7])
    A9 = stencil_grid_np(diffusion_2d(0.001, math.pi / 8), [12, 12])
    for A, want in ((A5, 2), (A9, 4)):
        c, nc = _colors_oracle(A)
        assert nc == want and not _adjacent_same_colour(A, c)
    c5, _ = _colors_oracle(A5)
    assert np.array_equal(c5, (np.arange(63) % 9 + np.arange(63) // 9) % 2)  # checkerboard
    one_sided = K.Csr([0, 2, 3, 4, 5], [0, 3, 1, 2, 3], [1.0] * 5, 4)
    c, nc = _colors_oracle(one_sided)
    assert nc == 2 and c[0] != c[3]


@pytest.mark.parametrize("symmetric", [0, 1])
def test_mc_sor_oracle_invariants(symmetric):
    """Multicolour SOR with w = 1 is Gauss-Seidel in colour order: after a
    forward sweep every row of the last colour satisfies its own equation
    (its neighbours are all of earlier colours, already final), so its
    residual is round-off; a symmetric sweep leaves the first colour so.
    A dropped term, a wrong sign or a stale neighbour value breaks this.
    Both forms then reduce the error of a stationary iteration on an SPD
    M-matrix."""
    import ctypes as C
    A = K.move_diag(stencil_grid_np(diffusion_2d(0.001, math.pi / 8), [12, 12]))
    n = A.n_rows
    c, nc = _colors_oracle(A)
    b = np.random.default_rng(3).standard_normal(n)
    x = np.random.default_rng(4).standard_normal(n)
    K.orc().orc_mc_sor(n, K.ip(A.rowptr), K.ip(A.col), K.dp(A.val), K.dp(b), K.dp(x), 1, C.c_double(1.0),
                       symmetric)
    r = b - K.spmv(A, x)
    last = c == (0 if symmetric else nc - 1)
    assert np.abs(r[last]).max() < 1e-12 * np.abs(b).max()
    assert np.abs(r[~last]).max() > 1e-3
    # stationary iteration x <- S(x) converges for w = 1 and w = 2/3
    for w in (1.0, 2.0 / 3.0):
        x = np.zeros(n)
        r0 = np.linalg.norm(b)
        for _ in range(30):
            K.orc().orc_mc_sor(n, K.ip(A.rowptr), K.ip(A.col), K.dp(A.val), K.dp(b), K.dp(x), 1, C.c_double(w),
                               symmetric)
        assert np.linalg.norm(b - K.spmv(A, x)) < 0.5 * r0


@needs_ref
@pytest.mark.parametrize("threads", [1, 3, 8])
def test_mt_cycle_and_fpcg(threads):
    """oracle_mt.c (the CPU baseline): its row-parallel cycle equals
    orc_cycle bit for bit at any thread count (the restriction as a gather
    over P^T keeps each entry's term order); its flexible PCG reaches the
    true relative residual it reports and gives the same iterates for every
    thread count (fixed-chunk inner products)."""
    A = stencil_grid_np(diffusion_2d(0.001, math.pi / 8), [60, 60])
    As, Ps = _ref_hierarchy(A, 0.25, 1, 1, 0, 2.0 / 3.0)
    H = K.Hier(As, Ps, relax_type=0, sweeps=1, omega=2.0 / 3.0)
    M = K.MtHier(H, threads)
    n = A.n_rows
    b = np.random.default_rng(9).standard_normal(n)
    assert np.array_equal(M.cycle(np.zeros(n), b), H.cycle(np.zeros(n), b))
    x0 = np.random.default_rng(10).standard_normal(n)
    assert np.array_equal(M.cycle(x0, b), H.cycle(x0, b))
    x, it, rel = M.fpcg(A, np.zeros(n), b, 1e-8, 200)
    assert 0 < it < 200 and rel <= 1e-8
    assert np.linalg.norm(b - K.spmv(A, x)) / np.linalg.norm(b) == pytest.approx(rel, rel=1e-12)
    x1, it1, rel1 = K.MtHier(H, 1).fpcg(A, np.zeros(n), b, 1e-8, 200)
    assert it1 == it and rel1 == rel and np.array_equal(x1, x)

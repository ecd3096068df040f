"""Test-side access to the two CPU checkers built under oracle/_ref:

  orc  -- liboracle.so, the plain-C restatement (oracle/oracle.c)
  ref  -- libraptor_ref.so, the reference compiled from its own sources plus
          oracle/ref_harness.cpp (present only where the reference was
          available at build time)

Only tests, the smoke check and bench.py's cpu_baseline leg use this module.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_DIR = os.path.join(ROOT, "oracle", "_ref")

PI = C.POINTER(C.c_int)
PD = C.POINTER(C.c_double)


def ip(a):
    return a.ctypes.data_as(PI)


def dp(a):
    return a.ctypes.data_as(PD)


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class OrcCsr(C.Structure):
    _fields_ = [("n_rows", C.c_int), ("n_cols", C.c_int), ("rowptr", PI), ("col", PI), ("val", PD)]


class OrcHier(C.Structure):
    _fields_ = [("num_levels", C.c_int), ("A", C.POINTER(OrcCsr)), ("P", C.POINTER(OrcCsr)),
                ("coarse_lu", PD), ("coarse_piv", PI), ("coarse_n", C.c_int), ("relax_type", C.c_int),
                ("num_sweeps", C.c_int), ("relax_weight", C.c_double)]


_orc = None
_ref = None


def orc():
    global _orc
    if _orc is None:
        path = os.path.join(REF_DIR, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} missing: run make -C oracle port")
        L = C.CDLL(path)
        L.orc_dot.restype = C.c_double
        L.orc_norm2.restype = C.c_double
        L.orc_spgemm.restype = C.c_int64
        L.orc_spgemm_tn.restype = C.c_int64
        L.orc_mt_prepare.restype = C.c_void_p
        L.orc_mt_prepare.argtypes = [C.c_void_p]
        L.orc_mt_free.argtypes = [C.c_void_p]
        L.orc_mt_cycle.argtypes = [C.c_void_p, PD, PD]
        L.orc_mt_fpcg.argtypes = [C.c_void_p, C.c_void_p, PD, PD, C.c_double, C.c_int, PD]
        _orc = L
    return _orc


def have_ref():
    return os.path.exists(os.path.join(REF_DIR, "libraptor_ref.so"))


def ref():
    global _ref
    if _ref is None:
        L = C.CDLL(os.path.join(REF_DIR, "libraptor_ref.so"))
        for f in ("ref_gallery", "ref_strength", "ref_interp", "ref_spgemm", "ref_rap", "ref_hier_level",
                  "ref_fit_candidates", "ref_jacobi_prolongation"):
            getattr(L, f).restype = C.c_long
        L.ref_init()
        _ref = L
    return _ref


def ref_fetch(nnz):
    L = ref()
    n = L.ref_last_n()
    rp = np.empty(n + 1, np.int32)
    col = np.empty(max(nnz, 1), np.int32)
    val = np.empty(max(nnz, 1), np.float64)
    L.ref_fetch(ip(rp), ip(col), dp(val))
    return rp, col[:nnz], val[:nnz], L.ref_last_ncols()


# ---------------------------------------------------------------------------
# oracle wrappers
# ---------------------------------------------------------------------------
class Csr:
    """Host CSR bundle that keeps its arrays alive for ctypes."""

    def __init__(self, rowptr, col, val, n_cols=None):
        self.rowptr, self.col, self.val = i32(rowptr), i32(col), f64(val)
        self.n_rows = self.rowptr.size - 1
        self.n_cols = self.n_rows if n_cols is None else int(n_cols)

    def orc(self):
        return OrcCsr(self.n_rows, self.n_cols, ip(self.rowptr), ip(self.col), dp(self.val))

    @property
    def nnz(self):
        return self.col.size

    def dense(self):
        d = np.zeros((self.n_rows, self.n_cols))
        for i in range(self.n_rows):
            for j in range(self.rowptr[i], self.rowptr[i + 1]):
                d[i, self.col[j]] += self.val[j]
        return d


def spmv(A, x):
    y = np.empty(A.n_rows)
    orc().orc_spmv(A.n_rows, ip(A.rowptr), ip(A.col), dp(A.val), dp(f64(x)), dp(y))
    return y


def spmv_t(A, x):
    y = np.empty(A.n_cols)
    orc().orc_spmv_t(A.n_rows, A.n_cols, ip(A.rowptr), ip(A.col), dp(A.val), dp(f64(x)), dp(y))
    return y


def spmv_append(A, x, y):
    y = f64(y).copy()
    orc().orc_spmv_append(A.n_rows, ip(A.rowptr), ip(A.col), dp(A.val), dp(f64(x)), dp(y))
    return y


def residual(A, x, b):
    r = np.empty(A.n_rows)
    orc().orc_residual(A.n_rows, ip(A.rowptr), ip(A.col), dp(A.val), dp(f64(x)), dp(f64(b)), dp(r))
    return r


def jacobi(A, x, b, sweeps, omega):
    x = f64(x).copy()
    tmp = np.empty(A.n_rows)
    orc().orc_jacobi(A.n_rows, ip(A.rowptr), ip(A.col), dp(A.val), dp(f64(b)), dp(x), dp(tmp), sweeps,
                     C.c_double(omega))
    return x


def sor(A, x, b, sweeps, omega):
    x = f64(x).copy()
    orc().orc_sor(A.n_rows, ip(A.rowptr), ip(A.col), dp(A.val), dp(f64(b)), dp(x), sweeps, C.c_double(omega))
    return x


def dot(x, y):
    x, y = f64(x), f64(y)
    return orc().orc_dot(x.size, dp(x), dp(y))


def norm2(x):
    x = f64(x)
    return orc().orc_norm2(x.size, dp(x))


def _take_csr(n_rows, n_cols, rp, pcol, pval, nnz):
    L = orc()
    col = np.ctypeslib.as_array(pcol, shape=(max(nnz, 1),))[:nnz].copy() if nnz else np.zeros(0, np.int32)
    val = np.ctypeslib.as_array(pval, shape=(max(nnz, 1),))[:nnz].copy() if nnz else np.zeros(0)
    L.orc_free(C.cast(pcol, C.c_void_p))
    L.orc_free(C.cast(pval, C.c_void_p))
    return Csr(rp, col, val, n_cols)


def spgemm(A, B):
    rp = np.empty(A.n_rows + 1, np.int32)
    pcol, pval = PI(), PD()
    nnz = orc().orc_spgemm(A.n_rows, ip(A.rowptr), ip(A.col), dp(A.val), B.n_rows, B.n_cols, ip(B.rowptr),
                           ip(B.col), dp(B.val), ip(rp), C.byref(pcol), C.byref(pval))
    assert nnz >= 0
    return _take_csr(A.n_rows, B.n_cols, rp, pcol, pval, nnz)


def spgemm_tn(P, B):
    rp = np.empty(P.n_cols + 1, np.int32)
    pcol, pval = PI(), PD()
    nnz = orc().orc_spgemm_tn(P.n_rows, P.n_cols, ip(P.rowptr), ip(P.col), dp(P.val), B.n_cols, ip(B.rowptr),
                              ip(B.col), dp(B.val), ip(rp), C.byref(pcol), C.byref(pval))
    assert nnz >= 0
    return _take_csr(P.n_cols, B.n_cols, rp, pcol, pval, nnz)


def move_diag(A):
    col, val = A.col.copy(), A.val.copy()
    orc().orc_move_diag(A.n_rows, ip(A.rowptr), ip(col), dp(val))
    return Csr(A.rowptr, col, val, A.n_cols)


def sort_rows(A):
    col, val = A.col.copy(), A.val.copy()
    orc().orc_sort_rows(A.n_rows, ip(A.rowptr), ip(col), dp(val))
    return Csr(A.rowptr, col, val, A.n_cols)


def galerkin(A, P):
    """Ac = P^T (A P), diagonal first, as ruge_stuben/par_ruge_stuben_solver.hpp:139-143."""
    return move_diag(spgemm_tn(P, spgemm(A, P)))


def lu_factor(dense_rowmajor):
    n = dense_rowmajor.shape[0]
    a = np.ascontiguousarray(dense_rowmajor, dtype=np.float64).ravel().copy()
    piv = np.empty(n, np.int32)
    info = orc().orc_lu_factor(n, dp(a), ip(piv))
    return a, piv, info


def lu_solve(lu, piv, b):
    b = f64(b).copy()
    orc().orc_lu_solve(b.size, dp(lu), ip(piv), dp(b))
    return b


class Hier:
    """Oracle hierarchy over given level matrices (A_l diag-first, P_l)."""

    def __init__(self, As, Ps, relax_type=0, sweeps=1, omega=1.0):
        self.As, self.Ps = As, Ps
        Ac = As[-1]
        self.lu, self.piv, info = lu_factor(Ac.dense())
        assert info == 0
        self.cA = (OrcCsr * len(As))(*[a.orc() for a in As])
        self.cP = (OrcCsr * max(1, len(Ps)))(*[p.orc() for p in Ps])
        self.h = OrcHier(len(As), self.cA, self.cP, dp(self.lu), ip(self.piv), Ac.n_rows, relax_type, sweeps,
                         C.c_double(omega))

    def cycle(self, x, b):
        x = f64(x).copy()
        orc().orc_cycle(C.byref(self.h), dp(x), dp(f64(b)))
        return x

    def cycle_fp32(self, x, b):
        x = np.ascontiguousarray(x, dtype=np.float32).copy()
        b = np.ascontiguousarray(b, dtype=np.float32)
        orc().orc_cycle_fp32(C.byref(self.h), x.ctypes.data_as(C.POINTER(C.c_float)),
                             b.ctypes.data_as(C.POINTER(C.c_float)))
        return x

    def solve(self, x, b, tol=1e-7, max_iter=100):
        x = f64(x).copy()
        res = np.zeros(max_iter + 1)
        it = orc().orc_amg_solve(C.byref(self.h), dp(x), dp(f64(b)), C.c_double(tol), max_iter, dp(res))
        return x, res[:it + 1], it

    def pcg(self, A, x, b, tol=1e-5, max_iter=-1, max_res=5000):
        x = f64(x).copy()
        res = np.zeros(max_res)
        nres = C.c_int()
        cA = A.orc()
        it = orc().orc_pcg(C.byref(cA), C.byref(self.h), dp(x), dp(f64(b)), C.c_double(tol), max_iter, dp(res),
                           max_res, C.byref(nres))
        return x, res[:min(nres.value, max_res)], it


class MtHier:
    """oracle_mt.c over a Hier: the cycle and a flexible PCG on `threads`
    host threads (OpenMP).  The cycle is bit-identical to Hier.cycle."""

    def __init__(self, hier, threads):
        L = orc()
        L.orc_mt_set_threads(int(threads))
        self.threads = L.orc_mt_threads()
        self.hier = hier
        self.p = L.orc_mt_prepare(C.byref(hier.h))

    def cycle(self, x, b):
        x = f64(x).copy()
        assert orc().orc_mt_cycle(self.p, dp(x), dp(f64(b))) == 0
        return x

    def fpcg(self, A, x, b, rtol=1e-8, max_iter=500):
        """(x, iterations, final true relative residual)"""
        x = f64(x).copy()
        rel = C.c_double()
        cA = A.orc()
        it = orc().orc_mt_fpcg(C.byref(cA), self.p, dp(x), dp(f64(b)), C.c_double(rtol), int(max_iter),
                               C.byref(rel))
        return x, it, rel.value

    def free(self):
        if self.p:
            orc().orc_mt_free(self.p)
            self.p = None

    def __del__(self):
        self.free()


def cg(A, x, b, tol=1e-5, max_iter=-1, max_res=5000):
    x = f64(x).copy()
    res = np.zeros(max_res)
    nres = C.c_int()
    cA = A.orc()
    it = orc().orc_cg(C.byref(cA), dp(x), dp(f64(b)), C.c_double(tol), max_iter, dp(res), max_res, C.byref(nres))
    return x, res[:min(nres.value, max_res)], it


def bicgstab(A, x, b, tol=1e-5, max_iter=-1, max_res=5000, hier=None):
    """orc_bicgstab, or orc_pre_bicgstab with hier (a Hier) as preconditioner."""
    x = f64(x).copy()
    res = np.zeros(max_res)
    nres = C.c_int()
    cA = A.orc()
    if hier is None:
        it = orc().orc_bicgstab(C.byref(cA), dp(x), dp(f64(b)), C.c_double(tol), max_iter, dp(res), max_res,
                                C.byref(nres))
    else:
        it = orc().orc_pre_bicgstab(C.byref(cA), C.byref(hier.h), dp(x), dp(f64(b)), C.c_double(tol), max_iter,
                                    dp(res), max_res, C.byref(nres))
    return x, res[:min(nres.value, max_res)], it


# ---------------------------------------------------------------------------
# fixtures
# ---------------------------------------------------------------------------
def fixture_csr(golden, name, sort=True):
    """A matrix of the reference's test_data, as readParMatrix + finalize()
    leaves it (rows sorted by column)."""
    rp, col, val = golden[name + "/rowptr"], golden[name + "/col"], golden[name + "/val"]
    shape = golden[name + "/shape"]
    A = Csr(rp, col, val, int(shape[1]))
    return sort_rows(A) if sort else A

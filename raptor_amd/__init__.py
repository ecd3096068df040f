"""raptor_amd -- MI355X (gfx950) AMG solve path behind RAPtor's interfaces.

Python mirror of the reference's one-rank API, over the C ABI of
libraptor_amd.so (include/raptor_amd.h):

    ParCSRMatrix      device CSR; mult / mult_append / residual / mult_T and
                      the Galerkin products mult(P) / mult_T(AP)
                      (core/par_matrix.hpp, util/linalg/par_spmv.cpp,
                      util/linalg/par_matmult.cpp)
    ParVector         device FP64 vector (core/par_vector.hpp)
    ParRugeStubenSolver, ParSmoothedAggregationSolver
                      ParMultilevel with RS or SA setup: setup / cycle /
                      solve (multilevel/par_multilevel.hpp,
                      ruge_stuben/par_ruge_stuben_solver.hpp,
                      aggregation/par_smoothed_aggregation_solver.hpp)
    PCG               krylov/par_cg.cpp:121-263
    BiCGStab, Pre_BiCGStab
                      krylov/par_bicgstab.cpp:11-110, 240-360
    strength / split / interpolation / mis2 / aggregate /
    fit_candidates / jacobi_prolongation / gallery
                      setup pieces (par_strength.cpp,
                      ruge_stuben/par_cf_splitting.cpp,
                      ruge_stuben/par_interpolation.cpp, aggregation/,
                      gallery/)

Every compute call runs the HIP library; nothing falls back to a CPU path.
"""
import ctypes as C
import os
import weakref

import numpy as np

from . import _lib
from ._lib import (CLJP, HMIS, MC_SOR, MC_SSOR, PMIS, RS, Classical, Direct, Extended, Falgout, Jacobi,
                   ModClassical, PREC_FP32, PREC_FP64, RamdError, Symmetric, check, lib)

__all__ = [
    "Context", "ParVector", "ParCSRMatrix", "ParMultilevel", "ParRugeStubenSolver",
    "ParSmoothedAggregationSolver", "PCG", "CG", "BiCGStab", "Pre_BiCGStab", "FPCG", "FPCG_block", "vcycle_block", "block_tuning", "block_set_tuning", "adjoint", "strength", "split", "interpolation", "mis2",
    "aggregate", "fit_candidates", "strength_gpu", "split_pmis_gpu", "interp_extended_gpu", "mis2_gpu", "aggregate_gpu", "fit_candidates_gpu", "jacobi_prolongation", "read_petsc", "write_petsc", "gallery_laplace27", "gallery_aniso2d", "gallery_fv7", "RamdError",
    "RS", "CLJP", "Falgout", "PMIS", "HMIS", "Direct", "ModClassical", "Extended", "Classical",
    "Symmetric", "Jacobi", "MC_SOR", "MC_SSOR", "greedy_colors", "PREC_FP64", "PREC_FP32",
]


def _ptr(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class Context:
    """One GPU and its stream (ramd_init).  Closing it first releases every
    live object created on it, so nothing is freed after its context."""

    # whether new contexts replay solver iterations from captured HIP graphs
    # (the library's default is eager launches)
    graphs_default = False

    def __init__(self, device=0):
        self.h = C.c_void_p()
        self._children = weakref.WeakSet()
        check(lib().ramd_init(device, C.byref(self.h)))
        if Context.graphs_default:
            self.set_graphs(True)

    def _adopt(self, obj):
        self._children.add(obj)

    def set_graphs(self, on=True):
        """Replay the solvers' iterations as captured HIP graphs or launch
        every kernel eagerly (the default; ramd_set_graphs); same results."""
        check(lib().ramd_set_graphs(self.h, 1 if on else 0))

    def sync(self):
        check(lib().ramd_sync(self.h))

    def close(self):
        if self.h:
            # hierarchies own level matrices: free them after borrowed views
            for obj in sorted(list(self._children), key=lambda o: isinstance(o, ParRugeStubenSolver)):
                obj.free()
            lib().ramd_finalize(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ParVector:
    """Device FP64 vector of length n (core/par_vector.hpp:44-176), zero on
    construction like the reference's (core/vector.hpp:54-77: the values are
    a resized std::vector)."""

    def __init__(self, ctx, n, dtype=np.float64):
        self.ctx, self.n, self.dtype = ctx, int(n), np.dtype(dtype)
        self.ptr = C.c_void_p()
        check(lib().ramd_malloc(ctx.h, max(1, self.n) * self.dtype.itemsize, C.byref(self.ptr)))
        ctx._adopt(self)
        check(lib().ramd_memset_zero(ctx.h, self.ptr, max(1, self.n) * self.dtype.itemsize))

    @classmethod
    def from_numpy(cls, ctx, a):
        a = np.ascontiguousarray(a)
        v = cls(ctx, a.size, a.dtype)
        v.copy_from(a)
        return v

    def copy_from(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.size == self.n
        check(lib().ramd_memcpy_h2d(self.ctx.h, self.ptr, a.ctypes.data_as(C.c_void_p), a.nbytes))

    def numpy(self):
        out = np.empty(self.n, dtype=self.dtype)
        check(lib().ramd_memcpy_d2h(self.ctx.h, out.ctypes.data_as(C.c_void_p), self.ptr, out.nbytes))
        return out

    def set_const_value(self, v):
        if v == 0:
            check(lib().ramd_memset_zero(self.ctx.h, self.ptr, self.n * self.dtype.itemsize))
        else:
            self.copy_from(np.full(self.n, v, dtype=self.dtype))

    def copy(self, other):
        check(lib().ramd_memcpy_d2d(self.ctx.h, self.ptr, other.ptr, self.n * self.dtype.itemsize))

    def inner_product(self, other):
        out = C.c_double()
        check(lib().ramd_dot(self.ctx.h, self.n, self.ptr, other.ptr, C.byref(out)))
        return out.value

    def norm(self, p=2):
        assert p == 2
        out = C.c_double()
        check(lib().ramd_norm2(self.ctx.h, self.n, self.ptr, C.byref(out)))
        return out.value

    def axpy(self, x, alpha):
        check(lib().ramd_axpy(self.ctx.h, self.n, float(alpha), x.ptr, self.ptr))

    def free(self):
        if self.ptr and self.ctx.h:
            lib().ramd_free(self.ctx.h, self.ptr)
        self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class ParCSRMatrix:
    """Device CSR matrix: a ParCSRMatrix at one rank (on_proc block only)."""

    def __init__(self, ctx, handle, owned=True):
        self.ctx, self.h, self.owned = ctx, handle, owned
        ctx._adopt(self)
        nr, nc, nnz = C.c_int32(), C.c_int32(), C.c_int64()
        check(lib().ramd_csr_info(self.h, C.byref(nr), C.byref(nc), C.byref(nnz)))
        self.n_rows, self.n_cols, self.nnz = nr.value, nc.value, nnz.value

    @classmethod
    def from_csr(cls, ctx, rowptr, col, val, n_cols=None):
        rowptr, col, val = _i32(rowptr), _i32(col), _f64(val)
        n = rowptr.size - 1
        n_cols = n if n_cols is None else int(n_cols)
        h = C.c_void_p()
        check(lib().ramd_csr_create(ctx.h, n, n_cols, col.size, _ptr(rowptr, C.c_int32), _ptr(col, C.c_int32),
                                    _ptr(val, C.c_double), C.byref(h)))
        return cls(ctx, h)

    def diagonals(self):
        """Stored diagonals of the stencil-diagonal form (0: CSR only)."""
        nd = C.c_int()
        check(lib().ramd_csr_diagonals(self.h, C.byref(nd)))
        return nd.value

    def to_csr(self):
        rp = np.empty(self.n_rows + 1, np.int32)
        col = np.empty(max(1, self.nnz), np.int32)
        val = np.empty(max(1, self.nnz), np.float64)
        check(lib().ramd_csr_download(self.h, _ptr(rp, C.c_int32), _ptr(col, C.c_int32), _ptr(val, C.c_double)))
        return rp, col[:self.nnz], val[:self.nnz]

    # --- SpMV family (util/linalg/par_spmv.cpp)
    def mult(self, x, b=None):
        if isinstance(x, ParCSRMatrix):
            h = C.c_void_p()
            check(lib().ramd_spgemm(self.h, x.h, C.byref(h)))
            return ParCSRMatrix(self.ctx, h)
        check(lib().ramd_spmv(self.h, x.ptr, b.ptr))

    def mult_append(self, x, b):
        check(lib().ramd_spmv_append(self.h, x.ptr, b.ptr))

    def residual(self, x, b, r):
        check(lib().ramd_residual(self.h, x.ptr, b.ptr, r.ptr))

    def mult_T(self, x, b=None):
        """mult_T(vector, out): b = A^T x.   mult_T(P): returns P^T * self."""
        if isinstance(x, ParCSRMatrix):
            h = C.c_void_p()
            check(lib().ramd_spgemm_tn(x.h, self.h, C.byref(h)))
            return ParCSRMatrix(self.ctx, h)
        check(lib().ramd_spmv_t(self.h, x.ptr, b.ptr))

    def galerkin(self, P):
        h = C.c_void_p()
        check(lib().ramd_galerkin(self.h, P.h, C.byref(h)))
        return ParCSRMatrix(self.ctx, h)

    def set_dia_staging(self, pp, mode, planes=0):
        """Stencil-diagonal kernel variant for precision pair pp (0 FP64, 1 FP32,
        2 FP32 values with FP64 vectors): 0 gather, 1 LDS-staged x window,
        2 plane-march over a box grid (`planes` per workgroup, 0 keeps the
        current count), -1 autotune.  Every variant gives the same bits."""
        check(lib().ramd_csr_set_dia_staging(self.h, int(pp), int(mode), int(planes)))

    def dia_staging(self, pp):
        """(mode, planes per workgroup, L, P) of the stencil-diagonal kernel
        for precision pair pp; L = P = 0 when the offsets have no box geometry."""
        v = [C.c_int() for _ in range(4)]
        check(lib().ramd_csr_dia_staging(self.h, int(pp), *[C.byref(t) for t in v]))
        return tuple(t.value for t in v)

    def set_sliced(self, prec, on):
        """CSR kernel for FP64 (prec 0) or FP32 (prec 1) vectors: True (or 1)
        selects the sliced-row form, 2 the sliced-row form with 16-bit column
        offsets, False the row-block kernel (identical bits)."""
        check(lib().ramd_csr_set_sliced(self.h, int(prec), int(on)))

    def set_lane_rows(self, prec, on):
        """CSR kernel for FP64 (prec 0) or FP32 (prec 1) vectors: True selects
        one lane per row over the CSR arrays, False the row-block kernel
        (identical bits)."""
        check(lib().ramd_csr_set_lane_rows(self.h, int(prec), int(bool(on))))

    def set_wave_rows(self, prec, on):
        """CSR kernel for FP64 (prec 0) or FP32 (prec 1) vectors: True selects
        one wavefront per row, False the row-block kernel (identical bits)."""
        check(lib().ramd_csr_set_wave_rows(self.h, int(prec), int(bool(on))))

    def set_fp32_outer(self, on):
        """FP64 products with this matrix read its exact FP32 value copy (True,
        default) or the FP64 values (False): the same products; FPCG's p.Ap
        follows each launch's workgroup shape in its last bits."""
        check(lib().ramd_csr_set_fp32_outer(self.h, int(bool(on))))

    def set_sorted_gather(self, prec, on):
        """Row-block kernel's SpMV for FP64 (prec 0) or FP32 (prec 1) vectors
        with the entries read in column order within each block (True) or in
        stored order (False); identical bits."""
        check(lib().ramd_csr_set_sorted_gather(self.h, int(prec), int(bool(on))))

    def set_values(self, vals):
        """New values (a device ParVector of nnz entries, stored order) for the same pattern."""
        check(lib().ramd_csr_set_values(self.h, vals.ptr))

    def jacobi(self, x, b, tmp, num_sweeps=1, omega=1.0):
        check(lib().ramd_jacobi(self.h, x.ptr, b.ptr, tmp.ptr, num_sweeps, float(omega)))

    def free(self):
        if self.h and self.owned and self.ctx.h:
            lib().ramd_csr_destroy(self.h)
        self.h = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def default_opts(solver=_lib.SOLVER_RS):
    o = _lib.AmgOpts()
    check(lib().ramd_amg_opts_default(solver, C.byref(o)))
    return o


class ParMultilevel:
    """Fields and solve path shared by both solvers (ParMultilevel,
    multilevel/par_multilevel.hpp:79-100,626-658)."""

    _solver = _lib.SOLVER_RS

    def _init_common(self, ctx, strong_threshold, strength_type, relax_type):
        if relax_type not in (Jacobi, MC_SOR, MC_SSOR):
            raise ValueError("the device V-cycle smooths with Jacobi or the multicolour MC_SOR / MC_SSOR "
                             "(the reference's sequential SOR is host-only)")
        self.ctx = ctx
        self.strong_threshold = strong_threshold
        self.strength_type = strength_type
        self.relax_type = relax_type
        o = default_opts(self._solver)
        self.num_smooth_sweeps = o.num_smooth_sweeps
        self.relax_weight = o.relax_weight
        self.max_coarse = o.max_coarse
        self.max_levels = o.max_levels
        self.solve_tol = 1e-7
        self.max_iterations = 100
        self.weights = None
        self.device_setup = True
        self.h = None

    def _opts(self):
        o = default_opts(self._solver)
        o.strong_threshold = self.strong_threshold
        o.strength_type = self.strength_type
        o.num_smooth_sweeps = self.num_smooth_sweeps
        o.relax_weight = self.relax_weight
        o.max_coarse = self.max_coarse
        o.max_levels = self.max_levels
        o.relax_type = self.relax_type
        o.device_setup = 1 if self.device_setup else 0
        return o

    def setup(self, rowptr, col, val):
        """ParMultilevel::setup(Af): A is given as host CSR arrays."""
        rowptr, col, val = _i32(rowptr), _i32(col), _f64(val)
        o = self._opts()
        w = None
        if self.weights is not None:
            w = _f64(self.weights)
            o.weights = _ptr(w, C.c_double)
        h = C.c_void_p()
        check(lib().ramd_hier_setup(self.ctx.h, rowptr.size - 1, col.size, _ptr(rowptr, C.c_int32),
                                    _ptr(col, C.c_int32), _ptr(val, C.c_double), C.byref(o), C.byref(h)))
        self.h = h
        self.ctx._adopt(self)
        nl = C.c_int()
        check(lib().ramd_hier_info(self.h, C.byref(nl), None, None, None))
        self.num_levels = nl.value
        return self

    def level_sizes(self):
        L = self.num_levels
        n = np.zeros(L, np.int32)
        na = np.zeros(L, np.int64)
        npp = np.zeros(L, np.int64)
        check(lib().ramd_hier_info(self.h, None, _ptr(n, C.c_int32), _ptr(na, C.c_int64), _ptr(npp, C.c_int64)))
        return n, na, npp

    def level(self, l):
        a, p = C.c_void_p(), C.c_void_p()
        check(lib().ramd_hier_level(self.h, l, C.byref(a), C.byref(p)))
        A = ParCSRMatrix(self.ctx, a, owned=False)
        P = ParCSRMatrix(self.ctx, p, owned=False) if p.value else None
        return A, P

    def restriction(self, l):
        """Level l's restriction operator R = P_l^T (None on the coarsest level)."""
        r = C.c_void_p()
        check(lib().ramd_hier_restriction(self.h, l, C.byref(r)))
        return ParCSRMatrix(self.ctx, r, owned=False) if r.value else None

    def states(self, l):
        n, _, _ = self.level_sizes()
        s = np.empty(int(n[l]), np.int32)
        check(lib().ramd_hier_states(self.h, l, _ptr(s, C.c_int32)))
        return s

    def save(self, path):
        """Writes the hierarchy to `path` (ramd_hier_save)."""
        check(lib().ramd_hier_save(self.h, os.fsencode(path)))

    @classmethod
    def load(cls, ctx, path):
        """A hierarchy read back from `path` (ramd_hier_load); cycles and
        solves exactly as the one that was saved."""
        ml = cls.__new__(cls)
        ml._init_common(ctx, 0.0, Classical, Jacobi)
        h = C.c_void_p()
        check(lib().ramd_hier_load(ctx.h, os.fsencode(path), C.byref(h)))
        ml.h = h
        ctx._adopt(ml)
        nl = C.c_int()
        check(lib().ramd_hier_info(ml.h, C.byref(nl), None, None, None))
        ml.num_levels = nl.value
        return ml

    def row_caps(self):
        """Row-block caps chosen by the autotuner: array (levels, 3, 2) for
        A, P, R with FP64 and FP32 vectors (0: stencil-diagonal or absent)."""
        c = np.zeros(6 * self.num_levels, np.int32)
        check(lib().ramd_hier_row_caps(self.h, _ptr(c, C.c_int32)))
        return c.reshape(-1, 3, 2)

    def setup_times(self):
        t = np.zeros(5)
        check(lib().ramd_hier_setup_times(self.h, _ptr(t, C.c_double)))
        return t

    def enable_fp32(self):
        check(lib().ramd_hier_enable_fp32(self.h))

    def set_tail(self, level=0):
        """Coarse tail (ramd_hier_set_tail; same bits either way): -1 off, 0
        automatic (timed at the first cycle of each precision), >= 1 forced
        to start at that level."""
        check(lib().ramd_hier_set_tail(self.h, int(level)))

    def tail_level(self, prec=PREC_FP64):
        """(level the next cycle in prec starts the tail at: -1 none, -2 not
        timed yet; smallest level that can start one, -1 none)."""
        v, f = C.c_int32(), C.c_int32()
        check(lib().ramd_hier_tail_level(self.h, int(prec), C.byref(v), C.byref(f)))
        return v.value, f.value

    def set_track_times(self, on=True):
        """ParMultilevel::track_times: per-level device time of the cycle
        work from now on (clears the solve buckets)."""
        check(lib().ramd_hier_track_times(self.h, 1 if on else 0))

    def level_times(self, phase):
        """setup_times (phase 0) / solve_times (phase 1) in the reference's
        layout: array (levels, 5), column 0 the level's total seconds."""
        t = np.zeros(5 * self.num_levels)
        check(lib().ramd_hier_level_times(self.h, phase, _ptr(t, C.c_double)))
        return t.reshape(-1, 5)

    def coarse_info(self):
        """(rows of the dense coarse system, dgetrf info: > 0 for a zero pivot)."""
        n, info = C.c_int32(), C.c_int32()
        check(lib().ramd_hier_coarse_info(self.h, C.byref(n), C.byref(info)))
        return n.value, info.value

    def cycle(self, x, b, x_is_zero=False):
        check(lib().ramd_vcycle(self.h, x.ptr, b.ptr, 1 if x_is_zero else 0))

    def cycle_fp32(self, x, b):
        """One V-cycle from x = 0 on the FP32 copies (float32 ParVectors)."""
        assert x.dtype == np.float32 and b.dtype == np.float32
        check(lib().ramd_vcycle_fp32(self.h, x.ptr, b.ptr))

    def solve(self, x, b):
        res = np.zeros(self.max_iterations + 1)
        it = C.c_int32()
        check(lib().ramd_amg_solve(self.h, x.ptr, b.ptr, self.solve_tol, self.max_iterations,
                                   _ptr(res, C.c_double), C.byref(it)))
        self.residuals = res[:it.value + 1]
        return it.value

    def free(self):
        if self.h and self.ctx.h:
            lib().ramd_hier_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass



class ParRugeStubenSolver(ParMultilevel):
    """ParRugeStubenSolver(strong_threshold, coarsen, interp, strength, relax)
    (ruge_stuben/par_ruge_stuben_solver.hpp:15-25)."""

    _solver = _lib.SOLVER_RS

    def __init__(self, ctx, strong_threshold=0.0, coarsen_type=RS, interp_type=Direct,
                 strength_type=Classical, relax_type=Jacobi):
        self._init_common(ctx, strong_threshold, strength_type, relax_type)
        self.coarsen_type = coarsen_type
        self.interp_type = interp_type
        self.interp_filter = default_opts(self._solver).interp_filter

    def _opts(self):
        o = super()._opts()
        o.coarsen_type = self.coarsen_type
        o.interp_type = self.interp_type
        o.interp_filter = self.interp_filter
        return o


class ParSmoothedAggregationSolver(ParMultilevel):
    """ParSmoothedAggregationSolver(strong_threshold, MIS, JacobiProlongation,
    strength, relax, prolong_smooth_steps, prolong_weight)
    (aggregation/par_smoothed_aggregation_solver.hpp:17-32)."""

    _solver = _lib.SOLVER_SA

    def __init__(self, ctx, strong_threshold=0.0, strength_type=Symmetric, relax_type=Jacobi,
                 prolong_smooth_steps=1, prolong_weight=4.0 / 3):
        self._init_common(ctx, strong_threshold, strength_type, relax_type)
        self.prolong_smooth_steps = prolong_smooth_steps
        self.prolong_weight = prolong_weight

    def _opts(self):
        o = super()._opts()
        o.prolong_smooth_steps = self.prolong_smooth_steps
        o.prolong_weight = self.prolong_weight
        return o

def PCG(A, ml, x, b, tol=1e-5, max_iter=-1, max_res=10000, times=None):
    """PCG(A, ml, x, b, res, tol, max_iter, precond_t, comm_t),
    krylov/par_cg.cpp:121-263.  Returns (residual history, iteration count).
    times, when given, is a float64 array of two entries to which the
    device seconds of the V-cycles (precond_t) and of the inner products
    (comm_t) are added."""
    res = np.zeros(max_res)
    nres, it = C.c_int32(), C.c_int32()
    if times is None:
        check(lib().ramd_pcg(ml.h, A.h, x.ptr, b.ptr, float(tol), int(max_iter), _ptr(res, C.c_double), max_res,
                             C.byref(nres), C.byref(it)))
    else:
        assert times.dtype == np.float64 and times.size >= 2 and times.flags["C_CONTIGUOUS"]
        pt, ct = C.c_double(times[0]), C.c_double(times[1])
        check(lib().ramd_pcg_timed(ml.h, A.h, x.ptr, b.ptr, float(tol), int(max_iter), _ptr(res, C.c_double),
                                   max_res, C.byref(nres), C.byref(it), C.byref(pt), C.byref(ct)))
        times[0], times[1] = pt.value, ct.value
    return res[:min(nres.value, max_res)], it.value


def CG(A, x, b, tol=1e-5, max_iter=-1, max_res=100000):
    """CG(A, x, b, res, tol, max_iter), krylov/par_cg.cpp:8-118.  Returns
    (residual history, iteration count)."""
    res = np.zeros(max_res)
    nres, it = C.c_int32(), C.c_int32()
    check(lib().ramd_cg(A.h, x.ptr, b.ptr, float(tol), int(max_iter), _ptr(res, C.c_double), max_res,
                        C.byref(nres), C.byref(it)))
    return res[:min(nres.value, max_res)], it.value


def greedy_colors(rowptr, col):
    """Greedy colouring of A + A^T (ramd_host_greedy_colors): (colour per row, number of colours)."""
    rowptr, col = _i32(rowptr), _i32(col)
    n = rowptr.size - 1
    out = np.empty(max(n, 1), np.int32)
    nc = C.c_int32()
    check(lib().ramd_host_greedy_colors(n, _ptr(rowptr, C.c_int32), _ptr(col, C.c_int32),
                                        _ptr(out, C.c_int32), C.byref(nc)))
    return out[:n], nc.value


def BiCGStab(A, x, b, tol=1e-5, max_iter=-1, max_res=100000):
    """BiCGStab(A, x, b, res, tol, max_iter), krylov/par_bicgstab.cpp:11-110.
    Returns (residual 2-norm history from ||r_0||, iteration count)."""
    res = np.zeros(max_res)
    nres, it = C.c_int32(), C.c_int32()
    check(lib().ramd_bicgstab(A.h, x.ptr, b.ptr, float(tol), int(max_iter), _ptr(res, C.c_double), max_res,
                              C.byref(nres), C.byref(it)))
    return res[:min(nres.value, max_res)], it.value


def Pre_BiCGStab(A, x, b, ml, tol=1e-5, max_iter=-1, max_res=100000):
    """Pre_BiCGStab(A, x, b, ml, res, tol, max_iter), krylov/par_bicgstab.cpp:
    240-360, with one V-cycle of ml as the preconditioner.  Returns
    (residual history, iteration count)."""
    res = np.zeros(max_res)
    nres, it = C.c_int32(), C.c_int32()
    check(lib().ramd_pre_bicgstab(ml.h, A.h, x.ptr, b.ptr, float(tol), int(max_iter), _ptr(res, C.c_double),
                                  max_res, C.byref(nres), C.byref(it)))
    return res[:min(nres.value, max_res)], it.value


def FPCG(A, ml, x, b, rtol=1e-8, max_iter=200, cycle_prec=PREC_FP64):
    """Flexible PCG to a true FP64 relative residual; returns (history, iters, final relres)."""
    hist = np.zeros(max_iter + 1)
    it, fin = C.c_int32(), C.c_double()
    check(lib().ramd_fpcg(ml.h, A.h, x.ptr, b.ptr, float(rtol), int(max_iter), int(cycle_prec),
                          _ptr(hist, C.c_double), C.byref(it), C.byref(fin)))
    return hist[:it.value + 1], it.value, fin.value


def FPCG_noalloc(A, ml, x, b, rtol=1e-8, max_iter=200, cycle_prec=PREC_FP64):
    """FPCG without any device allocation (ramd_fpcg_noalloc): partial sums
    in the library's fixed per-GPU scratch, eager launches; the same results
    as FPCG once FPCG has run with the same hierarchy and operator."""
    hist = np.zeros(max_iter + 1)
    it, fin = C.c_int32(), C.c_double()
    check(lib().ramd_fpcg_noalloc(ml.h, A.h, x.ptr, b.ptr, float(rtol), int(max_iter), int(cycle_prec),
                                  _ptr(hist, C.c_double), C.byref(it), C.byref(fin)))
    return hist[:it.value + 1], it.value, fin.value


def FPCG_block(A, ml, X, B, k, rtol=1e-8, max_iter=200, cycle_prec=PREC_FP64, use_graph=True):
    """k right-hand sides (row-interleaved n x k device blocks) solved in
    lockstep (ramd_fpcg_block).  Returns (per-column iterations, per-column
    final relative residuals, lockstep iterations)."""
    its = np.zeros(k, np.int32)
    rel = np.zeros(k)
    tot = C.c_int32()
    check(lib().ramd_fpcg_block(ml.h, A.h, int(k), X.ptr, B.ptr, float(rtol), int(max_iter), int(cycle_prec),
                                1 if use_graph else 0, _ptr(its, C.c_int32), _ptr(rel, C.c_double),
                                C.byref(tot)))
    return its, rel, tot.value


def vcycle_block(ml, X, B, k, prec=PREC_FP64):
    """One V-cycle from X = 0 on k row-interleaved columns (ramd_vcycle_block)."""
    check(lib().ramd_vcycle_block(ml.h, int(k), int(prec), X.ptr, B.ptr))


def block_tuning(ml, k, prec=PREC_FP64):
    """Multi-vector kernels picked by the autotuner for k columns and vector
    precision prec: array (levels, 3) for A, P, R -- lanes per row, negative
    for the kernel that reads 16-byte units of consecutive columns."""
    out = np.zeros(3 * ml.num_levels, np.int32)
    check(lib().ramd_block_tuning(ml.h, int(k), int(prec), _ptr(out, C.c_int32)))
    return out.reshape(-1, 3)


def block_set_tuning(ml, k, codes, prec=PREC_FP64):
    """Overrides the multi-vector kernel choices (array (levels, 3), same
    encoding as block_tuning)."""
    codes = np.ascontiguousarray(codes, dtype=np.int32).ravel()
    assert codes.size == 3 * ml.num_levels
    check(lib().ramd_block_set_tuning(ml.h, int(k), int(prec), _ptr(codes, C.c_int32)))


def adjoint(A, ml, x, g, rtol=1e-10, max_iter=500, cycle_prec=PREC_FP64, with_dA=True, dA=None, db=None):
    """Gradients of L(x) for x = A^-1 b from g = dL/dx (ramd_adjoint): solves
    A^T lambda = g through ml's hierarchy; returns (dL/dA on A's pattern as a
    device ParVector of nnz entries or None, dL/db = lambda, iterations).
    dA / db: device vectors to write into (nnz and n entries) instead of new
    ones, for loops that solve repeatedly."""
    nnz = A.nnz
    if db is None:
        db = ParVector(A.ctx, x.n)
    if dA is None and with_dA:
        dA = ParVector(A.ctx, nnz)
    assert db.n == x.n and (dA is None or dA.n == nnz)
    it, rel = C.c_int32(), C.c_double()
    check(lib().ramd_adjoint(ml.h, A.h, x.ptr, g.ptr, float(rtol), int(max_iter), int(cycle_prec),
                             dA.ptr if dA is not None else None, db.ptr, C.byref(it), C.byref(rel)))
    return dA, db, it.value


# ---------------------------------------------------------------------------
# host setup pieces
# ---------------------------------------------------------------------------
def _take(ptr, n, dtype):
    if n == 0:
        lib().ramd_host_free(C.cast(ptr, C.c_void_p))
        return np.zeros(0, dtype)
    a = np.ctypeslib.as_array(ptr, shape=(n,)).copy().astype(dtype)
    lib().ramd_host_free(C.cast(ptr, C.c_void_p))
    return a


def strength(rowptr, col, val, strength_type=Classical, theta=0.0):
    rowptr, col, val = _i32(rowptr), _i32(col), _f64(val)
    n = rowptr.size - 1
    srp, scol, snnz = C.POINTER(C.c_int32)(), C.POINTER(C.c_int32)(), C.c_int64()
    check(lib().ramd_host_strength(n, _ptr(rowptr, C.c_int32), _ptr(col, C.c_int32), _ptr(val, C.c_double),
                                   strength_type, float(theta), C.byref(srp), C.byref(scol), C.byref(snnz)))
    return _take(srp, n + 1, np.int32), _take(scol, snnz.value, np.int32)


def split(s_rowptr, s_col, coarsen_type, weights=None):
    s_rowptr, s_col = _i32(s_rowptr), _i32(s_col)
    n = s_rowptr.size - 1
    states = np.empty(n, np.int32)
    w = None if weights is None else _f64(weights)
    check(lib().ramd_host_split(n, _ptr(s_rowptr, C.c_int32), _ptr(s_col, C.c_int32), coarsen_type,
                                None if w is None else _ptr(w, C.c_double), _ptr(states, C.c_int32)))
    return states


def strength_gpu(A, theta, strength_type=Classical):
    """Classical (ramd_strength_gpu) or symmetric (ramd_symmetric_strength_gpu)
    strength of a device matrix on the device."""
    h = C.c_void_p()
    fn = lib().ramd_symmetric_strength_gpu if strength_type == Symmetric else lib().ramd_strength_gpu
    check(fn(A.h, float(theta), C.byref(h)))
    return ParCSRMatrix(A.ctx, h)


def split_pmis_gpu(S, weights=None):
    """PMIS C/F states of a device strength matrix (ramd_split_pmis_gpu)."""
    states = np.empty(S.n_rows, np.int32)
    w = None if weights is None else _f64(weights)
    check(lib().ramd_split_pmis_gpu(S.h, None if w is None else _ptr(w, C.c_double), _ptr(states, C.c_int32)))
    return states



def split_hmis_gpu(S, weights=None):
    """HMIS C/F states of a device strength matrix (ramd_split_hmis_gpu): the
    Ruge-Stuben first pass on the host, the PMIS loop on the device."""
    states = np.empty(S.n_rows, np.int32)
    w = None if weights is None else _f64(weights)
    check(lib().ramd_split_hmis_gpu(S.h, None if w is None else _ptr(w, C.c_double), _ptr(states, C.c_int32)))
    return states


def interp_extended_gpu(A, S, states, filter_threshold=0.3):
    """Extended interpolation (with filter) on the device (ramd_interp_extended_gpu)."""
    states = _i32(states)
    h = C.c_void_p()
    check(lib().ramd_interp_extended_gpu(A.h, S.h, _ptr(states, C.c_int32), float(filter_threshold), C.byref(h)))
    return ParCSRMatrix(A.ctx, h)


def split_cljp_gpu(S, weights=None):
    """CLJP C/F states of a device strength matrix (ramd_split_cljp_gpu)."""
    states = np.empty(S.n_rows, np.int32)
    w = None if weights is None else _f64(weights)
    check(lib().ramd_split_cljp_gpu(S.h, None if w is None else _ptr(w, C.c_double), _ptr(states, C.c_int32)))
    return states


def interp_mod_classical_gpu(A, S, states):
    """Modified classical interpolation on the device (ramd_interp_mod_classical_gpu)."""
    states = _i32(states)
    h = C.c_void_p()
    check(lib().ramd_interp_mod_classical_gpu(A.h, S.h, _ptr(states, C.c_int32), C.byref(h)))
    return ParCSRMatrix(A.ctx, h)


def interp_direct_gpu(A, S, states):
    """Direct interpolation on the device (ramd_interp_direct_gpu)."""
    states = _i32(states)
    h = C.c_void_p()
    check(lib().ramd_interp_direct_gpu(A.h, S.h, _ptr(states, C.c_int32), C.byref(h)))
    return ParCSRMatrix(A.ctx, h)


def mis2_gpu(S, weights=None):
    """MIS-2 states (1 selected, 0 not) of a device strength matrix on the
    device (ramd_mis2_gpu)."""
    states = np.empty(S.n_rows, np.int32)
    w = None if weights is None else _f64(weights)
    check(lib().ramd_mis2_gpu(S.h, None if w is None else _ptr(w, C.c_double), _ptr(states, C.c_int32)))
    return states


def aggregate_gpu(A, S, states, rand_vals=None):
    """Aggregation of device A and S on the device (ramd_aggregate_gpu).
    Returns (aggregates: root row per row or -1, n_aggs)."""
    states = _i32(states)
    agg = np.empty(A.n_rows, np.int32)
    na = C.c_int32()
    r = None if rand_vals is None else _f64(rand_vals)
    check(lib().ramd_aggregate_gpu(A.h, S.h, _ptr(states, C.c_int32), None if r is None else _ptr(r, C.c_double),
                                   _ptr(agg, C.c_int32), C.byref(na)))
    return agg, na.value


def fit_candidates_gpu(ctx, aggregates, n_aggs, B):
    """fit_candidates on the device (ramd_fit_candidates_gpu): returns (T, a
    device matrix, and the aggregate norms R)."""
    agg, B = _i32(aggregates), _f64(B)
    R = np.empty(n_aggs)
    h = C.c_void_p()
    check(lib().ramd_fit_candidates_gpu(ctx.h, agg.size, _ptr(agg, C.c_int32), int(n_aggs), _ptr(B, C.c_double),
                                        _ptr(R, C.c_double), C.byref(h)))
    return ParCSRMatrix(ctx, h), R


def split_rs_serial(s_rowptr, s_col, second_pass=True):
    """split_rs(CSRMatrix*, states, false, second_pass) of the reference."""
    s_rowptr, s_col = _i32(s_rowptr), _i32(s_col)
    n = s_rowptr.size - 1
    states = np.empty(n, np.int32)
    check(lib().ramd_host_split_rs_serial(n, _ptr(s_rowptr, C.c_int32), _ptr(s_col, C.c_int32),
                                          1 if second_pass else 0, _ptr(states, C.c_int32)))
    return states


def interpolation(rowptr, col, val, s_rowptr, s_col, states, interp_type, filter_threshold=0.0):
    rowptr, col, val = _i32(rowptr), _i32(col), _f64(val)
    s_rowptr, s_col, states = _i32(s_rowptr), _i32(s_col), _i32(states)
    n = rowptr.size - 1
    nc = C.c_int32()
    prp, pcol, pval, pnnz = C.POINTER(C.c_int32)(), C.POINTER(C.c_int32)(), C.POINTER(C.c_double)(), C.c_int64()
    check(lib().ramd_host_interp(n, _ptr(rowptr, C.c_int32), _ptr(col, C.c_int32), _ptr(val, C.c_double),
                                 _ptr(s_rowptr, C.c_int32), _ptr(s_col, C.c_int32), _ptr(states, C.c_int32),
                                 interp_type, float(filter_threshold), C.byref(nc), C.byref(prp), C.byref(pcol),
                                 C.byref(pval), C.byref(pnnz)))
    return (_take(prp, n + 1, np.int32), _take(pcol, pnnz.value, np.int32), _take(pval, pnnz.value, np.float64),
            nc.value)


def _gallery(kind, m, eps=1.0, theta=0.0, seed=0):
    n, nnz = C.c_int32(), C.c_int64()
    rp, col, val = C.POINTER(C.c_int32)(), C.POINTER(C.c_int32)(), C.POINTER(C.c_double)()
    check(lib().ramd_host_gallery(kind, int(m), float(eps), float(theta), int(seed), C.byref(n), C.byref(rp),
                                  C.byref(col), C.byref(val), C.byref(nnz)))
    return (_take(rp, n.value + 1, np.int32), _take(col, nnz.value, np.int32), _take(val, nnz.value, np.float64))


def gallery_laplace27(m):
    """par_stencil_grid(laplace_stencil_27pt(), {m,m,m}, 3)."""
    return _gallery(0, m)


def gallery_aniso2d(m, eps=0.001, theta=np.pi / 8):
    """par_stencil_grid(diffusion_stencil_2d(eps, theta), {m,m}, 2)."""
    return _gallery(1, m, eps, theta)


def gallery_fv7(m, seed=3):
    """7-point finite-volume diffusion, log-normal cell coefficients (seeded)."""
    return _gallery(2, m, seed=seed)


def mis2(s_rowptr, s_col, weights=None):
    """mis2(S, states, off_proc_states, false, weights), aggregation/par_mis.cpp:204-645."""
    s_rowptr, s_col = _i32(s_rowptr), _i32(s_col)
    n = s_rowptr.size - 1
    states = np.empty(n, np.int32)
    w = None if weights is None else _f64(weights)
    check(lib().ramd_host_mis2(n, _ptr(s_rowptr, C.c_int32), _ptr(s_col, C.c_int32),
                               None if w is None else _ptr(w, C.c_double), _ptr(states, C.c_int32)))
    return states


def aggregate(rowptr, col, val, s_rowptr, s_col, states, rand_vals=None):
    """aggregate(A, S, states, ...), aggregation/par_aggregate.cpp:5-185.
    Returns (aggregates: root row per row or -1, n_aggs)."""
    rowptr, col, val = _i32(rowptr), _i32(col), _f64(val)
    s_rowptr, s_col, states = _i32(s_rowptr), _i32(s_col), _i32(states)
    n = rowptr.size - 1
    agg = np.empty(n, np.int32)
    na = C.c_int32()
    r = None if rand_vals is None else _f64(rand_vals)
    check(lib().ramd_host_aggregate(n, _ptr(rowptr, C.c_int32), _ptr(col, C.c_int32), _ptr(val, C.c_double),
                                    _ptr(s_rowptr, C.c_int32), _ptr(s_col, C.c_int32), _ptr(states, C.c_int32),
                                    None if r is None else _ptr(r, C.c_double), _ptr(agg, C.c_int32),
                                    C.byref(na)))
    return agg, na.value


def fit_candidates(aggregates, n_aggs, B):
    """fit_candidates(A, n_aggs, aggregates, B, R, 1, ...), aggregation/par_candidates.cpp:6-216.
    Returns (T rowptr, col, val, R)."""
    aggregates, B = _i32(aggregates), _f64(B)
    n = aggregates.size
    R = np.zeros(max(n_aggs, 1))
    trp, tcol, tval, tnnz = C.POINTER(C.c_int32)(), C.POINTER(C.c_int32)(), C.POINTER(C.c_double)(), C.c_int64()
    check(lib().ramd_host_fit_candidates(n, _ptr(aggregates, C.c_int32), int(n_aggs), _ptr(B, C.c_double),
                                         _ptr(R, C.c_double), C.byref(trp), C.byref(tcol), C.byref(tval),
                                         C.byref(tnnz)))
    return (_take(trp, n + 1, np.int32), _take(tcol, tnnz.value, np.int32), _take(tval, tnnz.value, np.float64),
            R[:n_aggs])


def jacobi_prolongation(ctx, rowptr, col, val, t_rowptr, t_col, t_val, n_aggs, omega=4.0 / 3, steps=1):
    """jacobi_prolongation(A, T, false, omega, steps), aggregation/par_prolongation.cpp:7-103
    (the sparse products on the device of ctx).  Returns (P rowptr, col, val)."""
    rowptr, col, val = _i32(rowptr), _i32(col), _f64(val)
    t_rowptr, t_col, t_val = _i32(t_rowptr), _i32(t_col), _f64(t_val)
    n = rowptr.size - 1
    prp, pcol, pval, pnnz = C.POINTER(C.c_int32)(), C.POINTER(C.c_int32)(), C.POINTER(C.c_double)(), C.c_int64()
    check(lib().ramd_sa_prolongation(ctx.h, n, _ptr(rowptr, C.c_int32), _ptr(col, C.c_int32),
                                     _ptr(val, C.c_double), int(n_aggs), _ptr(t_rowptr, C.c_int32),
                                     _ptr(t_col, C.c_int32), _ptr(t_val, C.c_double), float(omega), int(steps),
                                     C.byref(prp), C.byref(pcol), C.byref(pval), C.byref(pnnz)))
    return _take(prp, n + 1, np.int32), _take(pcol, pnnz.value, np.int32), _take(pval, pnnz.value, np.float64)


def read_petsc(path):
    """readMatrix(filename) of gallery/matrix_IO.cpp: (rowptr, col, val, n_cols)."""
    nr, nc, nnz = C.c_int32(), C.c_int32(), C.c_int64()
    rp, col, val = C.POINTER(C.c_int32)(), C.POINTER(C.c_int32)(), C.POINTER(C.c_double)()
    check(lib().ramd_host_read_petsc(os.fsencode(path), C.byref(nr), C.byref(nc), C.byref(rp), C.byref(col),
                                     C.byref(val), C.byref(nnz)))
    return (_take(rp, nr.value + 1, np.int32), _take(col, nnz.value, np.int32), _take(val, nnz.value, np.float64),
            nc.value)


def write_petsc(path, rowptr, col, val, n_cols=None):
    """Big-endian PETSc binary matrix file."""
    rowptr, col, val = _i32(rowptr), _i32(col), _f64(val)
    n = rowptr.size - 1
    check(lib().ramd_host_write_petsc(os.fsencode(path), n, n if n_cols is None else int(n_cols),
                                      _ptr(rowptr, C.c_int32), _ptr(col, C.c_int32), _ptr(val, C.c_double)))

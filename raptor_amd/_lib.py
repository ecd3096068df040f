"""ctypes binding of libraptor_amd.so (the C ABI declared in include/raptor_amd.h).

The library is loaded from this package directory only; there is no fallback
implementation.  If the shared object is missing, build it with
``python -c "import __graft_entry__ as g; g.build()"`` (or
``make -C raptor_amd/csrc``).
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libraptor_amd.so")

RAMD_OK = 0
RAMD_E_ARG = -1
RAMD_E_DEVICE = -2
RAMD_E_ALLOC = -3
RAMD_E_INDEFINITE = -4
RAMD_E_NOT_CONVERGED = -5
RAMD_E_STATE = -6
RAMD_E_NUMERIC = -7

# enums of include/raptor_amd.h (core/types.hpp:158-164 of the reference)
RS, CLJP, Falgout, PMIS, HMIS = 0, 1, 2, 3, 4
Direct, ModClassical, Extended = 0, 1, 2
Classical, Symmetric = 0, 1
Jacobi = 0
SOR, SSOR, MC_SOR, MC_SSOR = 1, 2, 3, 4
PREC_FP64, PREC_FP32 = 0, 1
SOLVER_RS, SOLVER_SA = 0, 1


class RamdError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"raptor_amd error {code}: {msg}")
        self.code = code


class AmgOpts(C.Structure):
    _fields_ = [
        ("solver", C.c_int),
        ("strong_threshold", C.c_double),
        ("coarsen_type", C.c_int),
        ("interp_type", C.c_int),
        ("strength_type", C.c_int),
        ("interp_filter", C.c_double),
        ("prolong_weight", C.c_double),
        ("num_smooth_sweeps", C.c_int),
        ("relax_weight", C.c_double),
        ("max_coarse", C.c_int),
        ("max_levels", C.c_int),
        ("weights", C.POINTER(C.c_double)),
        ("prolong_smooth_steps", C.c_int),
        ("relax_type", C.c_int),
        ("device_setup", C.c_int),
    ]


_lib = None

P = C.c_void_p
I32 = C.c_int32
I64 = C.c_int64
D = C.c_double
PI32 = C.POINTER(C.c_int32)
PI64 = C.POINTER(C.c_int64)
PD = C.POINTER(C.c_double)
PPI32 = C.POINTER(PI32)
PPD = C.POINTER(PD)

_SIGS = {
    "ramd_last_error": (C.c_char_p, C.c_size_t),
    "ramd_init": (C.c_int, C.POINTER(P)),
    "ramd_finalize": (P,),
    "ramd_sync": (P,),
    "ramd_malloc": (P, C.c_size_t, C.POINTER(P)),
    "ramd_free": (P, P),
    "ramd_memcpy_h2d": (P, P, P, C.c_size_t),
    "ramd_memcpy_d2h": (P, P, P, C.c_size_t),
    "ramd_memcpy_d2d": (P, P, P, C.c_size_t),
    "ramd_memset_zero": (P, P, C.c_size_t),
    "ramd_csr_create": (P, I32, I32, I64, PI32, PI32, PD, C.POINTER(P)),
    "ramd_csr_info": (P, PI32, PI32, PI64),
    "ramd_csr_download": (P, PI32, PI32, PD),
    "ramd_csr_destroy": (P,),
    "ramd_csr_diagonals": (P, C.POINTER(C.c_int)),
    "ramd_csr_set_dia_staging": (P, C.c_int, C.c_int, C.c_int),
    "ramd_csr_dia_staging": (P, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                             C.POINTER(C.c_int)),
    "ramd_csr_set_sliced": (P, C.c_int, C.c_int),
    "ramd_csr_set_lane_rows": (P, C.c_int, C.c_int),
    "ramd_csr_set_wave_rows": (P, C.c_int, C.c_int),
    "ramd_csr_set_sorted_gather": (P, C.c_int, C.c_int),
    "ramd_csr_set_fp32_outer": (P, C.c_int),
    "ramd_spmv": (P, P, P),
    "ramd_spmv_append": (P, P, P),
    "ramd_residual": (P, P, P, P),
    "ramd_spmv_t": (P, P, P),
    "ramd_jacobi": (P, P, P, P, C.c_int, D),
    "ramd_dot": (P, I64, P, P, PD),
    "ramd_norm2": (P, I64, P, PD),
    "ramd_axpy": (P, I64, D, P, P),
    "ramd_prof_enable": (P, I64),
    "ramd_prof_read": (P, C.c_int, C.c_int, PI64, PI64, PD, PD),
    "ramd_spgemm": (P, P, C.POINTER(P)),
    "ramd_spgemm_tn": (P, P, C.POINTER(P)),
    "ramd_galerkin": (P, P, C.POINTER(P)),
    "ramd_amg_opts_default": (C.c_int, C.POINTER(AmgOpts)),
    "ramd_hier_setup": (P, I32, I64, PI32, PI32, PD, C.POINTER(AmgOpts), C.POINTER(P)),
    "ramd_hier_destroy": (P,),
    "ramd_hier_info": (P, C.POINTER(C.c_int), PI32, PI64, PI64),
    "ramd_hier_level": (P, C.c_int, C.POINTER(P), C.POINTER(P)),
    "ramd_hier_restriction": (P, C.c_int, C.POINTER(C.c_void_p)),
    "ramd_hier_states": (P, C.c_int, PI32),
    "ramd_hier_setup_times": (P, PD),
    "ramd_hier_row_caps": (P, PI32),
    "ramd_hier_save": (P, C.c_char_p),
    "ramd_hier_load": (P, C.c_char_p, C.POINTER(P)),
    "ramd_host_read_petsc": (C.c_char_p, PI32, PI32, PPI32, PPI32, PPD, PI64),
    "ramd_host_write_petsc": (C.c_char_p, I32, I32, PI32, PI32, PD),
    "ramd_hier_enable_fp32": (P,),
    "ramd_hier_set_tail": (P, C.c_int32),
    "ramd_hier_tail_level": (P, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int32)),
    "ramd_set_graphs": (P, C.c_int),
    "ramd_vcycle": (P, P, P, C.c_int),
    "ramd_vcycle_fp32": (P, P, P),
    "ramd_amg_solve": (P, P, P, D, C.c_int, PD, PI32),
    "ramd_pcg": (P, P, P, P, D, C.c_int, PD, I32, PI32, PI32),
    "ramd_pcg_timed": (P, P, P, P, D, C.c_int, PD, I32, PI32, PI32, PD, PD),
    "ramd_hier_level_times": (P, C.c_int, PD),
    "ramd_hier_track_times": (P, C.c_int),
    "ramd_hier_coarse_info": (P, PI32, PI32),
    "ramd_fpcg": (P, P, P, P, D, C.c_int, C.c_int, PD, PI32, PD),
    "ramd_cg": (P, P, P, D, C.c_int, PD, I32, PI32, PI32),
    "ramd_bicgstab": (P, P, P, D, I32, PD, I32, PI32, PI32),
    "ramd_host_greedy_colors": (I32, PI32, PI32, PI32, PI32),
    "ramd_strength_gpu": (P, D, C.POINTER(P)),
    "ramd_symmetric_strength_gpu": (P, D, C.POINTER(P)),
    "ramd_split_pmis_gpu": (P, PD, PI32),
    "ramd_split_cljp_gpu": (P, PD, PI32),
    "ramd_split_hmis_gpu": (P, PD, PI32),
    "ramd_interp_extended_gpu": (P, P, PI32, D, C.POINTER(P)),
    "ramd_interp_mod_classical_gpu": (P, P, PI32, C.POINTER(P)),
    "ramd_interp_direct_gpu": (P, P, PI32, C.POINTER(P)),
    "ramd_mis2_gpu": (P, PD, PI32),
    "ramd_aggregate_gpu": (P, P, PI32, PD, PI32, PI32),
    "ramd_fit_candidates_gpu": (P, I32, PI32, I32, PD, PD, C.POINTER(P)),
    "ramd_pre_bicgstab": (P, P, P, P, D, I32, PD, I32, PI32, PI32),
    "ramd_adjoint": (P, P, P, P, D, C.c_int, C.c_int, P, P, PI32, PD),
    "ramd_fpcg_noalloc": (P, P, P, P, D, C.c_int, C.c_int, PD, PI32, PD),
    "ramd_fpcg_block": (P, P, I32, P, P, D, C.c_int, C.c_int, C.c_int, PI32, PD, PI32),
    "ramd_vcycle_block": (P, I32, C.c_int, P, P),
    "ramd_block_tuning": (P, I32, I32, PI32),
    "ramd_block_set_tuning": (P, I32, I32, PI32),
    "ramd_csr_set_values": (P, P),
    "ramd_host_free": None,
    "ramd_host_strength": (I32, PI32, PI32, PD, C.c_int, D, PPI32, PPI32, PI64),
    "ramd_host_split": (I32, PI32, PI32, C.c_int, PD, PI32),
    "ramd_host_split_rs_serial": (I32, PI32, PI32, C.c_int, PI32),
    "ramd_host_interp": (I32, PI32, PI32, PD, PI32, PI32, PI32, C.c_int, D, PI32, PPI32, PPI32, PPD, PI64),
    "ramd_host_gallery": (C.c_int, I32, D, D, C.c_uint64, PI32, PPI32, PPI32, PPD, PI64),
    "ramd_host_mis2": (I32, PI32, PI32, PD, PI32),
    "ramd_host_aggregate": (I32, PI32, PI32, PD, PI32, PI32, PI32, PD, PI32, PI32),
    "ramd_host_fit_candidates": (I32, PI32, I32, PD, PD, PPI32, PPI32, PPD, PI64),
    "ramd_sa_prolongation": (P, I32, PI32, PI32, PD, I32, PI32, PI32, PD, D, C.c_int, PPI32, PPI32, PPD, PI64),
}


def lib():
    """Load libraptor_amd.so (once).  Raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: the HIP library must be built (make -C raptor_amd/csrc); "
            "raptor_amd has no CPU fallback")
    L = C.CDLL(LIB_PATH)
    for name, args in _SIGS.items():
        fn = getattr(L, name, None)
        if fn is None:
            continue
        if name == "ramd_host_free":
            fn.argtypes = [P]
            fn.restype = None
        else:
            fn.argtypes = list(args)
            fn.restype = C.c_int
    _lib = L
    return L


def last_error():
    buf = C.create_string_buffer(2048)
    lib().ramd_last_error(buf, len(buf))
    return buf.value.decode(errors="replace")


def check(code):
    if code != RAMD_OK:
        raise RamdError(code, last_error())
    return code


def declared_symbols():
    """Names of every function declared in include/raptor_amd.h."""
    import re
    hdr = os.path.join(os.path.dirname(_HERE), "include", "raptor_amd.h")
    txt = open(hdr).read()
    return sorted(set(re.findall(r"^(?:int|void)\s+(ramd_[a-z0-9_]+)\s*\(", txt, re.M)))

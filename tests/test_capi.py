"""The C-ABI library loads, exports every symbol include/raptor_amd.h
declares, and fails cleanly (no CPU fallback) where no GPU is usable."""
import ctypes as C

import pytest

import raptor_amd as ra
from raptor_amd import _lib


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    declared = _lib.declared_symbols()
    assert len(declared) > 30
    missing = [s for s in declared if not hasattr(L, s)]
    assert not missing, missing


def test_error_codes_without_device_or_arguments():
    L = _lib.lib()
    assert L.ramd_init(0, None) == _lib.RAMD_E_ARG
    assert "NULL" in _lib.last_error()
    assert L.ramd_csr_info(None, None, None, None) == _lib.RAMD_E_ARG


def test_opts_defaults_match_reference():
    o = ra.default_opts()
    # multilevel/par_multilevel.hpp:87-99, ruge_stuben/par_ruge_stuben_solver.hpp:24
    assert o.num_smooth_sweeps == 1 and o.relax_weight == 1.0
    assert o.max_coarse == 50 and o.max_levels == 25 and o.interp_filter == 0.3


def test_host_entry_points_reject_bad_input():
    with pytest.raises(ra.RamdError):
        ra.strength([0, 1], [5], [1.0], ra.Classical, 0.25)  # column out of range
    with pytest.raises(ra.RamdError):
        ra.split([0, 1], [0], 99, None)                       # unknown coarsening


def test_torch_extension_loads():
    """The libtorch autograd wrapper links against this library, binds to
    torch's HIP runtime when torch is imported first, and exports amg_solve
    (own interpreter: this one has already loaded the library)."""
    pytest.importorskip("torch")
    import os
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    script = ("import sys, torch; sys.path.insert(0, sys.argv[1]); from raptor_amd import autograd; "
              "m = autograd._load(); assert hasattr(m, 'amg_solve'); assert len(autograd._hip_runtimes()) == 1; "
              "print('EXT_OK')")
    r = subprocess.run([sys.executable, "-c", script, root], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and "EXT_OK" in r.stdout, r.stderr[-3000:]


@pytest.mark.parametrize("name", ["aniso", "random", "laplacian_P_extend"])
def test_petsc_reference_files_byte_for_byte(tmp_path, golden, name):
    """The reference's own fixture files (test_data/<name>.pm, copied
    unchanged to tests/golden/pm/): read_petsc gives the matrix the
    reference's reader gives (gallery/matrix_IO.cpp:24-110: rows in file
    order, columns and values as stored), and write_petsc reproduces the
    file byte for byte."""
    import numpy as np
    import os
    src = os.path.join(os.path.dirname(__file__), "golden", "pm", name + ".pm")
    rp, col, val, nc = ra.read_petsc(src)
    assert np.array_equal(rp, golden[name + "/rowptr"]) and np.array_equal(col, golden[name + "/col"])
    assert np.array_equal(val, golden[name + "/val"]) and nc == int(golden[name + "/shape"][1])
    out = tmp_path / (name + ".pm")
    ra.write_petsc(out, rp, col, val, nc)
    assert out.read_bytes() == open(src, "rb").read()


def test_petsc_matrix_round_trip(tmp_path, golden):
    """write_petsc writes PETSc's big-endian binary layout (the format of the
    reference's test_data/*.pm, gallery/matrix_IO.cpp) and read_petsc reads
    it back, also from a little-endian file."""
    import numpy as np
    import checkers as K
    A = K.fixture_csr(golden, "aniso", sort=False)
    p = tmp_path / "a.pm"
    ra.write_petsc(p, A.rowptr, A.col, A.val, A.n_cols)
    raw = p.read_bytes()
    head = np.frombuffer(raw, dtype=">i4", count=4)
    assert list(head) == [1211216, A.n_rows, A.n_cols, A.col.size]
    rp, col, val, nc = ra.read_petsc(p)
    assert nc == A.n_cols and np.array_equal(rp, A.rowptr) and np.array_equal(col, A.col)
    assert np.array_equal(val, A.val)
    # the same matrix stored little-endian
    lens = np.diff(A.rowptr).astype("<i4")
    q = tmp_path / "le.pm"
    q.write_bytes(np.array([1211216, A.n_rows, A.n_cols, A.col.size], "<i4").tobytes() + lens.tobytes()
                  + A.col.astype("<i4").tobytes() + A.val.astype("<f8").tobytes())
    rp2, col2, val2, nc2 = ra.read_petsc(q)
    assert np.array_equal(rp2, rp) and np.array_equal(col2, col) and np.array_equal(val2, val)
    with pytest.raises(ra.RamdError):
        ra.read_petsc(tmp_path / "missing.pm")
    q.write_bytes(raw[:40])
    with pytest.raises(ra.RamdError):
        ra.read_petsc(q)

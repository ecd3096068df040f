"""RAPtor's own C++ interface with the device path behind it: a driver
compiled against the reference's headers and library (tests/dropin/
dropin_c1.cpp) builds ParAMDRugeStubenSolver(0.25, CLJP, ModClassical,
Classical, Jacobi) with relax_weight = 2/3 and calls the reference's
unchanged PCG(A, ml, x, b, res, 1e-8, -1, &precond_t, &comm_t) on the 100x100
anisotropic problem (b = A*1).  The iteration count, residual history and
solution equal the oracle's PCG on the same hierarchy with omega = 2/3
(krylov/par_cg.cpp:121-263, multilevel/par_multilevel.hpp:341-457); the
mirrored host levels equal the device ones; the timing buckets are filled."""
import math
import os
import subprocess

import numpy as np
import pytest

import checkers as K
import raptor_amd as ra

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "oracle", "_ref", "dropin_c1")
OMEGA = 2.0 / 3.0


def parse(stdout):
    out = {}
    for line in stdout.splitlines():
        if line.startswith("RAMD_"):
            key, *vals = line.split()
            out[key] = vals
    return out


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="drop-in driver not built (needs the reference sources)")
def test_dropin_pcg_matches_oracle(ctx, tmp_path):
    m = 100
    xfile = tmp_path / "x.bin"
    r = subprocess.run([DRIVER, str(m), str(xfile)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    o = parse(r.stdout)
    assert o["RAMD_MIRROR"] == ["equal"]
    its = int(o["RAMD_ITERS"][0])
    res = np.array([float(v) for v in o["RAMD_RES"]])
    x = np.fromfile(xfile, dtype=np.float64)
    # the same hierarchy through the package (same options, same setup)
    rp, col, val = ra.gallery_aniso2d(m, 0.001, math.pi / 8)
    ml = ra.ParRugeStubenSolver(ctx, 0.25, ra.CLJP, ra.ModClassical, ra.Classical, ra.Jacobi)
    ml.relax_weight = OMEGA
    ml.setup(rp, col, val)
    assert [int(v) for v in o["RAMD_LEVELS"][1:]] == [int(v) for v in ml.level_sizes()[0]]
    As, Ps = [], []
    for l in range(ml.num_levels):
        Al, Pl = ml.level(l)
        As.append(K.Csr(*Al.to_csr(), Al.n_cols))
        if Pl is not None:
            Ps.append(K.Csr(*Pl.to_csr(), Pl.n_cols))
    A = K.Csr(rp, col, val)
    n = A.n_rows
    b = K.spmv(A, np.ones(n))
    H = K.Hier(As, Ps, relax_type=0, sweeps=1, omega=OMEGA)
    xo, reso, ito = H.pcg(A, np.zeros(n), b, tol=1e-8)
    assert its == ito
    assert res.size == reso.size and np.allclose(res, reso, rtol=1e-8, atol=0)
    assert np.allclose(x, xo, rtol=1e-9, atol=1e-12)
    # relax_weight reached the device: omega = 1 (the default) gives another history
    _, res1, it1 = K.Hier(As, Ps, relax_type=0, sweeps=1, omega=1.0).pcg(A, np.zeros(n), b, tol=1e-8)
    assert it1 != its or not np.allclose(res1, res, rtol=1e-6, atol=0)
    # precond_t, comm_t and the per-level buckets
    pt, ct = (float(v) for v in o["RAMD_TIMES"])
    assert pt > 0 and ct > 0
    st = [float(v) for v in o["RAMD_SETUP_TIMES"]]
    so = [float(v) for v in o["RAMD_SOLVE_TIMES"]]
    assert len(st) == len(so) == ml.num_levels and all(t > 0 for t in st) and all(t > 0 for t in so)
    # ParMultilevel::solve through the adapter: the oracle's stationary iteration count
    _, reso2, ito2 = H.solve(np.zeros(n), b)
    it2, rel2 = int(o["RAMD_SOLVE"][0]), float(o["RAMD_SOLVE"][1])
    assert it2 == ito2 and math.isclose(rel2, reso2[-1], rel_tol=1e-8)

"""Writes tests/golden/ref_hier_27pt128.json: fingerprints of the hierarchy
the reference itself builds for the 27-point Laplacian on 128^3 (2,097,152
rows) with ParRugeStubenSolver(0.25, PMIS or HMIS, Extended, Classical)
(ruge_stuben/par_ruge_stuben_solver.hpp:57-162, setup_helper
multilevel/par_multilevel.hpp:126-212), run from oracle/_ref/libraptor_ref.so
(the reference compiled from its own sources by oracle/Makefile).

Per level: sizes, SHA-256 of the CSR pattern (rowptr then columns, int32
little-endian) of A_l and P_l, SHA-256 of the C/F states, SHA-256 of the
FP64 values, and three value checksums (sum, sum of |v|, sum of v times a
position weight) that a test compares to 1e-12 relative.  The hierarchy is
too large to commit; these fingerprints pin it.

Run in the container that holds the reference sources:
    python tests/golden/make_ref_hier_fingerprints.py
"""
import ctypes as C
import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import checkers as K  # noqa: E402

OUT = os.path.join(HERE, "ref_hier_27pt128.json")
M = 128
OMEGA = 2.0 / 3.0


def fingerprint(rp, col, val):
    rp, col, val = np.asarray(rp, np.int32), np.asarray(col, np.int32), np.asarray(val, np.float64)
    h = hashlib.sha256()
    h.update(rp.astype("<i4").tobytes())
    h.update(col.astype("<i4").tobytes())
    w = (np.arange(val.size) % 97 + 1).astype(np.float64)
    return dict(n_rows=int(rp.size - 1), nnz=int(col.size), pattern_sha256=h.hexdigest(),
                values_sha256=hashlib.sha256(val.astype("<f8").tobytes()).hexdigest(),
                sum=float(np.sum(val)), abs_sum=float(np.sum(np.abs(val))), weighted_sum=float(np.dot(w, val)))


def main():
    L = K.ref()
    nnz = L.ref_gallery(0, M, C.c_double(0.0), C.c_double(0.0))
    rp, col, val, _ = K.ref_fetch(nnz)
    n = rp.size - 1
    out = dict(workload=f"27-point Laplacian {M}^3, RS(theta 0.25, Extended, Classical strength), "
                        "the reference's setup", n=int(n), nnz=int(nnz), hierarchies={})
    for name, coarsen in (("pmis", 3), ("hmis", 4)):
        t = time.time()
        nl = L.ref_hier_setup(n, K.ip(rp), K.ip(col), K.dp(val), C.c_double(0.25), coarsen, 2, 0, 0,
                              C.c_double(OMEGA), 50, 25, None)
        levels = []
        for l in range(nl):
            A = K.ref_fetch(L.ref_hier_level(l, 0))
            lev = dict(A=fingerprint(*A[:3]))
            if l < nl - 1:
                P = K.ref_fetch(L.ref_hier_level(l, 1))
                lev["P"] = fingerprint(*P[:3])
                lev["P"]["n_cols"] = int(P[3])
                st = np.empty(A[0].size - 1, np.int32)
                L.ref_hier_states(l, K.ip(st))
                lev["states_sha256"] = hashlib.sha256(st.astype("<i4").tobytes()).hexdigest()
                lev["n_coarse"] = int(np.sum(st == 1))
            levels.append(lev)
        out["hierarchies"][name] = dict(levels=levels)
        print(f"{name}: {nl} levels {[lv['A']['n_rows'] for lv in levels]} in {time.time() - t:.1f}s", flush=True)
    with open(OUT, "w") as f:
        json.dump(out, f, indent=1)
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()

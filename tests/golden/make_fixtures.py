"""Convert the reference's golden test data into one compressed .npz.

Reads the pyamg-generated fixtures that RAPtor's own tests use
(test_data/*.pm in PETSc binary, big-endian: int32 code 1211216, n_rows,
n_cols, nnz, int32 row lengths, int32 columns, float64 values -- the layout
read by gallery/matrix_IO.cpp:24-110 -- and test_data/*.txt, one value per
line) and stores them as plain arrays:

    <name>/rowptr, <name>/col, <name>/val, <name>/shape   for matrices
    <name>                                                 for vectors

It also copies three of the .pm files unchanged into pm/ (PM_FILES), so the
library's PETSc reader and writer are checked against the reference's own
files byte for byte.

Run in the container that holds the reference:
    python tests/golden/make_fixtures.py [/root/reference/test_data]
"""
import os
import shutil
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "reference_test_data.npz")

MATRICES = [
    # SpMV fixtures (util/tests/test_par_spmv_*.cpp)
    "laplacian27", "aniso", "random",
    # strength (tests/test_par_strength.cpp)
    "aniso_S", "aniso_SS", "laplacian", "laplacian_S", "laplacian_SS",
    # interpolation (ruge_stuben/tests/test_par_interpolation.cpp)
    "aniso_P_direct", "aniso_P_mod_class", "aniso_P_extend",
    "laplacian_P_direct", "laplacian_P_mod_class", "laplacian_P_extend",
    # RS hierarchy pieces (ruge_stuben/tests/test_par_{rap,splitting,ruge_stuben}.cpp)
    "rss_A0", "rss_A1", "rss_A2", "rss_P0", "rss_P1", "rss_S0", "rss_S1",
    "rss_P0_mc", "rss_P1_mc", "rss_P0_extend", "rss_P1_extend",
    # smoothed aggregation (aggregation/tests/test_par_smoothed_aggregation.cpp)
    "sas_A0", "sas_S0", "sas_T0", "sas_P0", "sas_A1", "sas_S1", "sas_T1", "sas_P1",
]
# kept as files: the byte-for-byte reader / writer check
PM_FILES = ["aniso", "random", "laplacian_P_extend"]
INT_VECTORS = [
    "rss_cf0", "rss_cf1", "rss_cf0_pmis", "rss_cf1_pmis", "rss_cf0_rs", "rss_cf1_rs",
    "aniso_split", "laplacian_split", "sas_mis0", "sas_mis1", "sas_agg0", "sas_agg1",
    "aniso_mis", "laplacian_mis",
]
FLOAT_VECTORS = [
    "laplacian27_ones_b", "laplacian27_ones_b_T", "laplacian27_inc_b", "laplacian27_inc_b_T",
    "aniso_ones_b", "aniso_ones_b_T", "aniso_inc_b", "aniso_inc_b_T",
    "random_ones_b", "random_ones_b_T", "random_inc_b", "random_inc_b_T",
    "weights", "cg_res",
    # BiCGStab on the 50 x 50 rotated anisotropic problem (krylov/tests/test_bicgstab.cpp,
    # krylov/tests/test_par_bicgstab.cpp)
    "bicgstab_res", "bicgstab_x",
]


def read_pm(path):
    raw = open(path, "rb").read()
    head = np.frombuffer(raw, dtype=">i4", count=4)
    code, n_rows, n_cols, nnz = (int(v) for v in head)
    assert code == 1211216, path
    off = 16
    lens = np.frombuffer(raw, dtype=">i4", count=n_rows, offset=off).astype(np.int64)
    off += 4 * n_rows
    col = np.frombuffer(raw, dtype=">i4", count=nnz, offset=off).astype(np.int32)
    off += 4 * nnz
    val = np.frombuffer(raw, dtype=">f8", count=nnz, offset=off).astype(np.float64)
    rowptr = np.zeros(n_rows + 1, dtype=np.int32)
    rowptr[1:] = np.cumsum(lens)
    return rowptr, col, val, np.array([n_rows, n_cols], dtype=np.int64)


def main(src):
    arrays = {}
    for name in MATRICES:
        rowptr, col, val, shape = read_pm(os.path.join(src, name + ".pm"))
        arrays[name + "/rowptr"] = rowptr
        arrays[name + "/col"] = col
        arrays[name + "/val"] = val
        arrays[name + "/shape"] = shape
    for name in INT_VECTORS:
        arrays[name] = np.loadtxt(os.path.join(src, name + ".txt"), dtype=np.int64).astype(np.int32)
    for name in FLOAT_VECTORS:
        arrays[name] = np.loadtxt(os.path.join(src, name + ".txt"), dtype=np.float64)
    np.savez_compressed(OUT, **arrays)
    print("wrote", OUT, os.path.getsize(OUT), "bytes,", len(arrays), "arrays")
    os.makedirs(os.path.join(HERE, "pm"), exist_ok=True)
    for name in PM_FILES:
        shutil.copyfile(os.path.join(src, name + ".pm"), os.path.join(HERE, "pm", name + ".pm"))
    print("copied", len(PM_FILES), "PETSc binary files to", os.path.join(HERE, "pm"))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/test_data")

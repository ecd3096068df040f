"""The drop-in adapter include/raptor_amd_raptor.hpp on the CPU: it compiles
against the reference's own headers (multilevel/par_multilevel.hpp,
krylov/par_cg.hpp), and the driver built with it refuses the reference's
default SOR smoother instead of silently running another one."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference/raptor"
DRIVER = os.path.join(ROOT, "oracle", "_ref", "dropin_c1")


@pytest.mark.skipif(not os.path.exists(os.path.join(REF, "raptor.hpp")), reason="reference headers absent")
def test_adapter_compiles_against_reference_headers():
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "tests", "dropin"), "check"], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="drop-in driver not built (needs the reference sources)")
def test_adapter_refuses_sequential_sor(tmp_path):
    """ParRugeStubenSolver's default relax_type is SOR (a sequential sweep):
    setup throws raptor_amd::Error before touching a device."""
    r = subprocess.run([DRIVER, "20", str(tmp_path / "x.bin"), "sor"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "RAMD_REFUSED -1" in r.stdout and "SOR" in r.stdout

"""The Fortran side of the boundary: include/cylpic.f90 compiles with the
image's Fortran compiler, and a Fortran host (tests/fortran/driver.f90, the
reference main loop's call order) links against libcylpic.so and drives
it.  Here, without a GPU, the library must refuse cleanly through the
interface (nonzero status + message), not crash."""
import os
import subprocess
import sys

import pytest

from fortran_io import DRIVER_LIB, write_input

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def driver():
    from cylindrical_epoch_amd import build
    if not os.path.exists(build.LIB):
        build.build()
    subprocess.check_call(["bash", os.path.join(HERE, "fortran", "build.sh")], stdout=subprocess.DEVNULL)
    return DRIVER_LIB


def _input(tmp_path):
    from cylindrical_epoch_amd import deck
    p = deck.uniform_plasma(16, 8, 1e-7, 1e-7, 2, 1e25, 100.0, seed=1)
    fin = str(tmp_path / "in.bin")
    write_input(fin, 16, 8, 2, 3, 0, (1, 1, 2, 10), (1, 1, 2, 9), (8, 4), 1e-7, 1e-7,
                deck.set_dt(1e-7, 1e-7, 0.95), 0.0, -deck.Q0, deck.M0, p)
    return fin


@pytest.mark.skipif(os.path.exists("/dev/kfd") and os.access("/dev/kfd", os.R_OK | os.W_OK),
                    reason="a GPU is present: the GPU test runs the driver instead")
def test_fortran_driver_reports_missing_gpu(driver, tmp_path):
    # run in a child process: the Fortran runtime writes to the process's stdout
    fin = _input(tmp_path)
    code = ("import sys; sys.path.insert(0, %r); from fortran_io import run_driver; "
            "sys.exit(run_driver(%r, %r))" % (HERE, fin, str(tmp_path / "out.bin")))
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2, (r.stdout, r.stderr)
    assert "cylpic_init failed with 2" in r.stdout
    assert "HIP" in r.stdout

"""Fortran module layer (iso_c_binding over the C-ABI) and the sp2run driver.

The driver mirrors examples/sp2run/src/main.F90: it reads a Matrix Market
Hamiltonian, calls prg_sp2_alg2 / prg_sp2_alg1 / prg_sp2_basic through the
`prg_sp2_mod` module and writes rho.  CPU tests check that the module layer
was built and that a missing device is reported as a library error; the GPU
test compares the written rho with the oracle on the reference's own
384-orbital Hamiltonian (examples/implicitfermi/hamiltonian_ortho.mtx).
"""
import subprocess
from pathlib import Path

import numpy as np
import pytest

import progress_oracle as O

ROOT = Path(__file__).resolve().parent.parent
SP2RUN = ROOT / "qmd-progress_amd" / "bin" / "sp2run"
FLANG = Path("/opt/rocm/bin/amdflang")


def _need_driver():
    if not SP2RUN.exists():
        if FLANG.exists():
            pytest.fail("amdflang is present but the sp2run driver was not built")
        pytest.skip("no Fortran compiler: module layer not built")


def test_driver_usage():
    _need_driver()
    r = subprocess.run([str(SP2RUN)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2
    assert "usage: sp2run" in r.stdout


def test_driver_reports_missing_device(golden):
    _need_driver()
    import torch
    if torch.cuda.device_count() > 0:
        pytest.skip("a HIP device is present")
    r = subprocess.run([str(SP2RUN), str(golden / "poly.16.mtx"), "/dev/null"],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 1
    assert "library error 6" in r.stdout  # ERR_NO_DEVICE


@pytest.mark.gpu
@pytest.mark.parametrize("method", ["alg2", "alg1", "basic"])
def test_driver_matches_oracle(golden, tmp_path, method):
    _need_driver()
    print(f"[sp2run] method={method}", flush=True)
    out = tmp_path / "rho.mtx"
    r = subprocess.run([str(SP2RUN), str(golden / "hamiltonian_ortho_384.mtx"), str(out), method,
                        "0.5", "Abs", "1e-10"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    h = O.read_mtx(str(golden / "hamiltonian_ortho_384.mtx"))
    rho = O.read_mtx(str(out))
    f = {"alg2": O.sp2_alg2, "alg1": O.sp2_alg1, "basic": O.sp2_basic}[method]
    ref = f(h, 0.5, 10, 200, "Abs", 1e-10)
    ref = ref[0] if isinstance(ref, tuple) else ref
    # rho is written with 17 significant digits
    assert np.linalg.norm(rho - ref) < 1e-9
    assert "sp2run: norb=384" in r.stdout

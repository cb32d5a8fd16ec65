"""Fortran module layer (iso_c_binding over the C-ABI) and the sp2run driver.

The driver mirrors examples/sp2run/src/main.F90: it reads a Matrix Market
Hamiltonian, calls prg_sp2_alg2 / prg_sp2_alg1 / prg_sp2_basic through the
`prg_sp2_mod` module and writes rho.  CPU tests check that the module layer
was built and that a missing device is reported as a library error; the GPU
test compares the written rho with the oracle on the reference's own
384-orbital Hamiltonian (examples/implicitfermi/hamiltonian_ortho.mtx).
"""
import re
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


RESPSOLVE = ROOT / "qmd-progress_amd" / "bin" / "respsolve"


def _write_mtx(path, a):
    n = a.shape[0]
    with open(path, "w") as f:
        f.write("%%MatrixMarket matrix coordinate real general\n")
        f.write(f"{n} {n} {n * n}\n")
        for j in range(n):
            for i in range(n):
                f.write(f"{i + 1} {j + 1} {a[i, j]:.17e}\n")


@pytest.mark.gpu
def test_respsolve_driver_matches_oracle(tmp_path):
    """The Fortran modules' prg_compute_response_RS / _FD, prg_pcg and
    prg_newtonschulz (fortran/respsolve.F90) against the oracle."""
    if not RESPSOLVE.exists():
        _need_driver()
        pytest.fail("amdflang is present but the respsolve driver was not built")
    print("[respsolve] n=96", flush=True)
    n = 96
    rng = np.random.default_rng(14)
    q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    ev = np.concatenate([np.sort(rng.uniform(-3.0, -1.0, n // 2)),
                         np.sort(rng.uniform(1.0, 3.0, n // 2))])
    h = (q * ev) @ q.T
    h = 0.5 * (h + h.T)
    h1 = rng.standard_normal((n, n))
    h1 = 0.5 * (h1 + h1.T)
    a = (q * rng.uniform(1.0, 3.0, n)) @ q.T
    a = 0.5 * (a + a.T)
    for name, m in (("h", h), ("h1", h1), ("a", a)):
        _write_mtx(tmp_path / f"{name}.mtx", m)
    pre = tmp_path / "out"
    r = subprocess.run([str(RESPSOLVE), str(tmp_path / "h.mtx"), str(tmp_path / "h1.mtx"),
                        str(tmp_path / "a.mtx"), str(pre)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    # the files hold what the library computed from the 17-digit inputs
    h, h1, a = (O.read_mtx(str(tmp_path / f"{x}.mtx")) for x in ("h", "h1", "a"))
    rs = O.read_mtx(f"{pre}_rs.mtx")
    fd = O.read_mtx(f"{pre}_fd.mtx")
    x = O.read_mtx(f"{pre}_pcg.mtx")
    ai = O.read_mtx(f"{pre}_ns.mtx")
    assert np.max(np.abs(rs - O.compute_response_rs(h, h1, 0.5))) < 1e-9
    assert np.max(np.abs(fd - O.compute_response_fd(h, h1, 1e-4, 0.5))) < 1e-8
    x_o, _, _ = O.pcg(a, np.zeros((n, n)), np.eye(n), 1e-16)
    assert np.max(np.abs(x - x_o)) < 1e-9
    ai_o, it_o = O.newtonschulz(a, 0.25 * np.eye(n), 1e-10)
    assert np.max(np.abs(ai - ai_o)) < 1e-11
    assert f"iterations={it_o}" in r.stdout
    # prg_implicit_fermi_zero and prg_implicit_fermi_response (order 2, H2 =
    # H1 / 2) at the reference's Fortran signatures
    ifz = O.read_mtx(f"{pre}_ifz.mtx")
    assert np.linalg.norm(ifz - O.implicit_fermi_zero(h, 6, 0.0, 0, 1e-18)) < 1e-8
    p0_o, ps_o, mu0_o, mu_o = O.implicit_fermi_response(h, [h1, 0.5 * h1], 6, 0.0, 10.0,
                                                        0.5 * n, 1e-8, 2, lin_tol=1e-10)
    assert np.linalg.norm(O.read_mtx(f"{pre}_ifr0.mtx") - p0_o) < 1e-8
    for k in (1, 2):
        pk = O.read_mtx(f"{pre}_ifr{k}.mtx")
        assert np.linalg.norm(pk - ps_o[k - 1]) < 1e-7 * max(1.0, np.linalg.norm(ps_o[k - 1]))
    line = [x for x in r.stdout.splitlines() if "mu0=" in x][0]
    # Fortran es24.16 fields are right-aligned: "mu0=  1.0E+00 mu1= ..."
    vals = dict(re.findall(r"(\w+)=\s*(\S+)", line))
    assert abs(float(vals["mu0"]) - mu0_o) < 1e-8
    assert abs(float(vals["mu1"]) - mu_o[0]) < 1e-8 and abs(float(vals["mu2"]) - mu_o[1]) < 1e-8


GPTEST = ROOT / "qmd-progress_amd" / "bin" / "gptest"


def test_gptest_usage():
    if not GPTEST.exists():
        _need_driver()
        pytest.fail("amdflang is present but the gptest driver was not built")
    r = subprocess.run([str(GPTEST)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "usage: gptest" in r.stdout


@pytest.mark.gpu
def test_gptest_driver_matches_oracle(golden, tmp_path):
    """The gptest-shaped Fortran caller (prg_graph_mod, prg_subgraphloop_mod,
    prg_homolumo_mod and the bml accessors, in the call order of
    examples/gptest/gptest.F90:57-151) on the reference's poly.16.mtx with 16
    core nodes per part, against the same flow through the oracle."""
    if not GPTEST.exists():
        _need_driver()
        pytest.fail("amdflang is present but the gptest driver was not built")
    print("[gptest] poly.16 nodesPerPart=16", flush=True)
    out = tmp_path / "rho.mtx"
    r = subprocess.run([str(GPTEST), str(golden / "poly.16.mtx"), "16", "1e-5", "1e-12", "0.5",
                        str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    got = dict(l.split(" = ") for l in r.stdout.splitlines() if " = " in l)
    h = O.read_mtx(str(golden / "poly.16.mtx"))
    n = h.shape[0]
    rho_full, pp, ic, vv = O.sp2_alg2_genseq(h, 0.5, 25, 100, "Rel", 1e-10)
    assert int(got["icount"]) == ic
    assert abs(float(got["trace_hp_sp2"]) - np.sum(h * rho_full)) < 1e-8
    lo, hi = O.gershgorin(h)
    eh, el, eg = O.homolumogap(vv, ic, pp, lo, hi)
    # The last vv (6.6e-11) is ||X - X^2||_F at the rounding level of X's
    # entries, so it carries ~1e-4 relative rounding; the gap estimator then
    # has two admissible answers 4.8e-8 (relative) apart, and any change of
    # the product's summation order (the GPU's K slices against numpy) may
    # pick either.
    assert abs(float(got["egap_sp2"]) - eg) < 1e-7 * max(1.0, abs(eg))
    g = np.where(np.abs(rho_full) > 1e-5, rho_full, 0.0)
    parts = O.equal_partition(16, n)
    seq, imax = O.sp2sequence(lo, hi, eh, el, 1e-12)
    assert int(got["parts"]) == len(parts) and int(got["maxIter"]) == imax
    rho_o, vv_o = O.subgraph_sp2_loop(h, g, parts, seq, imax, lo, hi)
    rho = O.read_mtx(str(out))
    assert np.linalg.norm(rho - rho_o) < 1e-9
    assert abs(float(got["trace_hp_graph"]) - np.sum(h * rho_o)) < 1e-8
    assert abs(float(got["frob"]) - np.linalg.norm(rho_full - rho_o)) < 1e-8
    _, _, eg_g = O.homolumogap(vv_o, imax, seq, lo, hi)
    assert abs(float(got["egap_graph"]) - eg_g) < 1e-6 * max(1.0, abs(eg_g))

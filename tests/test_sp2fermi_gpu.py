"""GPU parity of the truncated (finite-temperature) SP2 family
(prg_sp2_fermi_init, prg_sp2_fermi_init_norecs, prg_sp2_fermi) against the
oracle: same sign list, chemical-potential shift and beta to 1e-9 relative,
X to 1e-9 (Frobenius), on the reference's 384-orbital Hamiltonian (bndfil 2/3,
the reference test's settings, tests/src/main.F90:740-800) and the model
Hamiltonian."""
import numpy as np
import pytest

import progress_oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def P():
    import progress_amd
    progress_amd.set_precision(progress_amd.PREC_FP64)
    return progress_amd


@pytest.mark.parametrize("case", ["ortho384", "model256"])
def test_sp2_fermi_init_and_solve_match_oracle(P, golden, case):
    print(f"[sp2_fermi] {case}", flush=True)
    if case == "ortho384":
        h = O.read_mtx(str(golden / "hamiltonian_ortho_384.mtx"))
        nocc = 2.0 / 3.0 * h.shape[0]
    else:
        h = O.model_hamiltonian(256, 1)
        nocc = 0.5 * h.shape[0]
    ref = O.sp2_fermi_init(h, 20, nocc, 1.0, 1e-9, 1e-12)
    got = P.sp2_fermi_init(h, 20, nocc, 1.0, 1e-9, 1e-12)
    assert got["sgnlist"] == ref["sgnlist"]
    for k in ("mu", "h1", "hN"):
        assert abs(got[k] - ref[k]) < 1e-9 * max(1.0, abs(ref[k]))
    assert abs(got["beta"] - ref["beta"]) < 1e-7 * abs(ref["beta"])
    assert np.linalg.norm(got["x"] - ref["x"]) < 1e-9
    xo, muo = O.sp2_fermi(h, 0, 20, nocc, ref["mu"], ref["beta"], ref["h1"], ref["hN"],
                          ref["sgnlist"], 1e-4, 1e-12)
    xg, mug = P.sp2_fermi(h, 0, 20, nocc, ref["mu"], ref["beta"], ref["h1"], ref["hN"],
                          ref["sgnlist"], 1e-4, 1e-12)
    assert abs(mug - muo) < 1e-9 * max(1.0, abs(muo))
    assert np.linalg.norm(xg - xo) < 1e-9


def test_sp2_fermi_init_norecs_matches_oracle(P, golden):
    print("[sp2_fermi_init_norecs]", flush=True)
    h = O.read_mtx(str(golden / "hamiltonian_ortho_384.mtx"))
    nocc = 2.0 / 3.0 * h.shape[0]
    ref = O.sp2_fermi_init(h, 20, nocc, 1.0, 1e-9, 1e-12, norecs=True, beta=20.1)
    got = P.sp2_fermi_init(h, 20, nocc, 1.0, 1e-9, 1e-12, norecs=True, beta=20.1)
    assert got["nsteps"] == ref["nsteps"] < 20
    assert got["sgnlist"] == ref["sgnlist"]
    assert abs(got["mu"] - ref["mu"]) < 1e-9
    assert np.linalg.norm(got["x"] - ref["x"]) < 1e-9


def test_sp2_entropy_ts_matches_oracle(P):
    print("[sp2_entropy_ts]", flush=True)
    h = O.model_hamiltonian(128, 2)
    lo, hi = O.gershgorin(h)
    sg = [1, -1, 1, -1, -1, 1, 1, -1]
    gg, ee = O.sp2_entropy_function(0.1, lo, hi, sg)
    x, _ = O.sp2_fermi(h, 1, len(sg), 64.0, 0.1, 5.0, lo, hi, sg, 1e-4, 1e-12)
    assert abs(P.sp2_entropy_ts(x, gg, ee) - O.sp2_entropy_ts(x, gg, ee)) < 1e-9


def test_orthogonalize_deorthogonalize_match_oracle(P):
    print("[nonortho]", flush=True)
    rng = np.random.default_rng(5)
    n = 300
    a = rng.standard_normal((n, n))
    a = a + a.T
    z = rng.standard_normal((n, n)) / np.sqrt(n)
    assert np.linalg.norm(P.orthogonalize(a, z) - O.orthogonalize(a, z)) < 1e-11 * n
    assert np.linalg.norm(P.deorthogonalize(a, z) - O.deorthogonalize(a, z)) < 1e-11 * n

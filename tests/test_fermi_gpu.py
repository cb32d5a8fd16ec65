"""GPU parity of the eigensolver densities, the implicit Fermi family and the
density-matrix responses against the oracle (FP64; tolerances in each test)."""
import numpy as np
import pytest

import progress_oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def P():
    import progress_amd
    progress_amd.set_precision(progress_amd.PREC_FP64)
    return progress_amd


@pytest.fixture(scope="module")
def h256():
    return O.model_hamiltonian(256, 0)


def test_eigensolver_densities(P, h256):
    print("[eigen] n=256", flush=True)
    h = h256
    rho, w = P.build_density_T0(h, 0.5)
    rho_o, w_o = O.build_density_T0(h, 0.5)
    assert np.allclose(w, w_o, rtol=0, atol=1e-11)
    assert np.linalg.norm(rho - rho_o) < 1e-9
    rho3, _ = P.build_density_T0(h, 0.4987)  # fractional occupation rule
    rho3_o, _ = O.build_density_T0(h, 0.4987)
    assert np.linalg.norm(rho3 - rho3_o) < 1e-9
    rf, dr = P.build_density_T_fermi(h, 0.05, 0.6)
    rf_o, dr_o = O.build_density_T_fermi(h, 0.05, 0.6)
    assert np.linalg.norm(rf - rf_o) < 1e-9 and abs(dr - dr_o) < 1e-9 * max(1, abs(dr_o))
    assert np.allclose(P.get_eigenvalues(h), w_o, atol=1e-11)
    assert abs(P.check_idempotency(0.5 * rho_o) - O.check_idempotency(0.5 * rho_o)) < 1e-10
    pd = P.get_density_matrix(h, 20.0, 0.6)
    assert np.linalg.norm(pd - O.get_density_matrix(h, 20.0, 0.6)) < 1e-9
    pt, mu = P.test_density_matrix(h, 20.0, 0.6, 128.0, 0, 1e-9)
    pt_o, mu_o = O.test_density_matrix(h, 20.0, 0.6, 128.0, 0, 1e-9)
    assert abs(mu - mu_o) < 1e-9 and np.linalg.norm(pt - pt_o) < 1e-8


@pytest.mark.parametrize("method", [0, 1])
def test_implicit_fermi(P, h256, method):
    print(f"[implicit_fermi] method={method}", flush=True)
    h = h256
    # method 0: CG to ||R||_F^2 <= 1e-16; method 1: Newton-Schulz to ||I - Y Ai||_F <= 1e-10
    args = (8, 2, 128.0, 0.6, 20.0, method, 3, 1e-9, 1e-16 if method == 0 else 1e-10)
    p, mu = P.implicit_fermi(h, *args)
    p_o, mu_o = O.implicit_fermi(h, *args)
    assert abs(mu - mu_o) < 1e-9
    assert np.linalg.norm(p - p_o) < 1e-8
    if method == 1:  # the Newton-Schulz variant solves the same recursion as CG
        p0, _ = O.implicit_fermi(h, 8, 2, 128.0, 0.6, 20.0, 0, 3, 1e-9, 1e-16)
        assert np.linalg.norm(p - p0) < 1e-6


def test_implicit_fermi_reference_tolerance(P, golden):
    """tests/src/main.F90:202-227 on the shipped 384-orbital Hamiltonian:
    nsteps 10, k 2, nocc 10, mu 0.2, beta 4, CG, one mu step;
    ||P - P_diag||_F <= 0.1 against prg_test_density_matrix."""
    print("[implicit_fermi ref test] n=384", flush=True)
    h = O.read_mtx(str(golden / "hamiltonian_ortho_384.mtx"))
    h = 0.5 * (h + h.T)
    p, _ = P.implicit_fermi(h, 10, 2, 10.0, 0.2, 4.0, 0, 1, 1.0, 10e-8)
    pt, _ = P.test_density_matrix(h, 4.0, 0.2, 10.0, 1, 1.0)
    assert np.linalg.norm(p - pt) <= 0.1


def test_save_inverse_and_first_order_response(P, h256):
    print("[save_inverse + first order response] n=256", flush=True)
    h = h256
    n, nsteps = 256, 8
    inv = [P.Matrix.from_numpy(np.eye(n)) for _ in range(nsteps)]
    p, mu, oc, tn = P.implicit_fermi_save_inverse(inv, h, nsteps, 128.0, 0.6, 20.0, 1e-8, 1e-10, 1)
    inv_o = [np.eye(n)] * nsteps
    p_o, mu_o, inv_o, it_o, ns_o = O.implicit_fermi_save_inverse(inv_o, h, nsteps, 128.0, 0.6,
                                                                  20.0, 1e-8, 1e-10, 1)
    assert abs(mu - mu_o) < 1e-9
    assert np.linalg.norm(p - p_o) < 1e-8
    assert oc == it_o and tn == ns_o
    for a, b in zip(inv, inv_o):
        assert np.linalg.norm(a.numpy() - b) < 1e-8
    # the reference's tolerance check (tests/src/main.F90:229-258)
    pt, _ = P.test_density_matrix(h, 20.0, mu, 128.0, 1, 1e-4)
    assert np.linalg.norm(0.5 * p - pt) <= 0.1
    h1 = O.banded_perturbation(n, 8, seed=3)
    p0, p1 = P.implicit_fermi_first_order_response(h, h1, inv, nsteps, mu, 20.0, 128.0)
    p0_o, p1_o = O.implicit_fermi_first_order_response(h, h1, [a.numpy() for a in inv], nsteps,
                                                       mu, 20.0)
    assert np.linalg.norm(p0 - p0_o) < 1e-9
    assert np.linalg.norm(p1 - p1_o) < 1e-9 * max(1.0, np.linalg.norm(p1_o))


def test_implicit_fermi_zero(P, h256):
    print("[implicit_fermi_zero]", flush=True)
    for method in (0, 1):
        p = P.implicit_fermi_zero(h256, 6, 0.6, method, 1e-16)
        p_o = O.implicit_fermi_zero(h256, 6, 0.6, method, 1e-16)
        # CG stops at ||R||_F^2 <= tol: agreement to ~sqrt(tol) * condition
        assert np.linalg.norm(p - p_o) < 1e-6


@pytest.mark.parametrize("order", [1, 2, 3])
def test_implicit_fermi_response_orders(P, order):
    print(f"[implicit_fermi_response] order={order}", flush=True)
    n = 128
    h0 = O.model_hamiltonian(n, 4)
    hs = [O.banded_perturbation(n, 4, seed=10 + j) * 0.1 for j in range(3)]
    p0, ps, mu0, mu = P.implicit_fermi_response(h0, hs, 6, 0.6, 10.0, 64.0, 1e-6, order,
                                                lin_tol=1e-10)
    p0_o, ps_o, mu0_o, mu_o = O.implicit_fermi_response(h0, hs, 6, 0.6, 10.0, 64.0, 1e-6, order,
                                                        lin_tol=1e-10)
    assert abs(mu0 - mu0_o) < 1e-8
    assert np.allclose(mu, mu_o, atol=1e-8)
    assert np.linalg.norm(p0 - p0_o) < 1e-8
    for a, b in zip(ps, ps_o):
        assert np.linalg.norm(a - b) < 1e-7 * max(1.0, np.linalg.norm(b))


def test_response_sp2(P):
    print("[response_sp2] n=192", flush=True)
    n = 192
    h = O.model_hamiltonian(n, 1)
    h1 = O.banded_perturbation(n, 8, seed=7)
    rho, rho1 = P.compute_response_sp2(h, h1, 0.5, 10, 200, "Abs", 1e-10)
    rho_o, rho1_o, it, pp = O.compute_response_sp2(h, h1, 0.5, 10, 200, "Abs", 1e-10)
    assert np.linalg.norm(rho - rho_o) < 1e-9
    assert np.linalg.norm(rho1 - rho1_o) < 1e-9 * max(1.0, np.linalg.norm(rho1_o))


def test_canon_response(P):
    print("[canon_response] n=160", flush=True)
    n = 160
    h = O.model_hamiltonian(n, 2)
    w, q = np.linalg.eigh(h)
    h1e = q.T @ O.banded_perturbation(n, 6, seed=1) @ q
    h1e = 0.5 * (h1e + h1e.T)
    p1 = P.canon_response(h1e, 80.0, 10.0, w, 0.6, 6)
    p1_o = O.canon_response(h1e, 80.0, 10.0, w, 0.6, 6)
    assert np.linalg.norm(p1 - p1_o) < 1e-10 * max(1.0, np.linalg.norm(p1_o))


def test_canon_response_variants_match_oracle(P, h256):
    """prg_canon_response_orig / _p1_dpdmu / _vector against the oracle (the
    eigenbasis recursions are element-wise: FP64 rounding only)."""
    print("[canon response variants] n=256", flush=True)
    e = np.linalg.eigvalsh(h256)
    rng = np.random.default_rng(8)
    h1 = rng.standard_normal((256, 256))
    h1 = 0.5 * (h1 + h1.T)
    beta, mu0, m = 20.0, 0.6, 8
    p = P.canon_response_orig(h1, 128.0, beta, e, mu0, m)
    assert np.max(np.abs(p - O.canon_response(h1, 128.0, beta, e, mu0, m))) < 1e-12
    p1, dp = P.canon_response_p1_dpdmu(h1, beta, e, mu0, m)
    p1_o, dp_o = O.canon_response_p1_dpdmu(h1, beta, e, mu0, m)
    # dPdmu up to beta / 4 = 5: a few ulps of the 8-step scalar recursion
    assert np.max(np.abs(p1 - p1_o)) < 1e-12 and np.max(np.abs(dp - dp_o)) < 1e-12
    pv = P.canon_response_vector(h1, beta, e, mu0)
    pv_o = O.canon_response_vector(h1, beta, e, mu0)
    assert np.max(np.abs(pv - pv_o)) < 1e-10 * max(1.0, np.max(np.abs(pv_o)))
    assert abs(np.trace(pv)) < 1e-9


def test_response_rs_and_fd_match_oracle(P):
    """prg_compute_response_RS and prg_compute_response_FD on a gapped
    Hamiltonian (half filling at the gap): each against the oracle, and RS
    against FD (central difference, delta 1e-4: O(delta^2) apart)."""
    print("[response RS / FD] n=160", flush=True)
    n = 160
    rng = np.random.default_rng(9)
    q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    ev = np.concatenate([np.sort(rng.uniform(-3.0, -1.0, n // 2)),
                         np.sort(rng.uniform(1.0, 3.0, n // 2))])
    h = (q * ev) @ q.T
    h = 0.5 * (h + h.T)
    h1 = rng.standard_normal((n, n))
    h1 = 0.5 * (h1 + h1.T)
    rs = P.compute_response_rs(h, h1, 0.5)
    rs_o = O.compute_response_rs(h, h1, 0.5)
    assert np.max(np.abs(rs - rs_o)) < 1e-9
    fd = P.compute_response_fd(h, h1, 1e-4, 0.5)
    fd_o = O.compute_response_fd(h, h1, 1e-4, 0.5)
    assert np.max(np.abs(fd - fd_o)) < 1e-8
    assert np.max(np.abs(rs - fd)) < 1e-5


@pytest.mark.parametrize("method", [0, 1])
def test_implicit_fermi_order_four_matches_oracle(P, h256, method):
    """Expansion order k = 4 (prg_setup_linsys: A = X^4 + (I - X)^4, B = X^4)."""
    print(f"[implicit_fermi k=4] method={method}", flush=True)
    args = (6, 4, 128.0, 0.6, 20.0, method, 3, 1e-9, 1e-16 if method == 0 else 1e-10)
    p, mu = P.implicit_fermi(h256, *args)
    p_o, mu_o = O.implicit_fermi(h256, *args)
    assert abs(mu - mu_o) < 1e-9
    assert np.linalg.norm(p - p_o) < 1e-8


def test_implicit_fermi_rejects_odd_order(P, h256):
    with pytest.raises(P.ProgressError):
        P.implicit_fermi(h256, 6, 3, 128.0, 0.6, 20.0, 0, 1, 1e-9, 1e-16)


def test_linear_solvers_match_oracle(P):
    """prg_pcg (Jacobi-preconditioned), prg_conjgrad and prg_newtonschulz
    through the C ABI against the oracle on an SPD system with a diagonal
    spread over three decades; B = I (the inverse, as in the recursion)."""
    print("[pcg / conjgrad / newtonschulz] n=120", flush=True)
    n = 120
    rng = np.random.default_rng(11)
    q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    core = (q * rng.uniform(1.0, 2.0, n)) @ q.T
    dsc = np.diag(np.logspace(0, 1.5, n))
    a = dsc @ core @ dsc
    a = 0.5 * (a + a.T)
    eye = np.eye(n)
    x, r = P.pcg(a, np.zeros((n, n)), eye, 1e-16)
    x_o, r_o, _ = O.pcg(a, np.zeros((n, n)), eye, 1e-16)
    assert np.max(np.abs(a @ x - eye)) < 1e-7
    assert np.max(np.abs(x - x_o)) < 1e-9 * np.max(np.abs(x_o))
    assert np.max(np.abs(r - (eye - a @ x))) < 1e-9
    spd = (q * rng.uniform(1.0, 3.0, n)) @ q.T
    spd = 0.5 * (spd + spd.T)
    xc = P.conjgrad(spd, np.zeros((n, n)), eye, 1e-20)
    assert np.max(np.abs(spd @ xc - eye)) < 1e-9
    ai, it = P.newtonschulz(spd, 0.25 * eye, 1e-10)
    ai_o, it_o = O.newtonschulz(spd, 0.25 * eye, 1e-10)
    assert it == it_o and np.max(np.abs(ai - ai_o)) < 1e-11


def test_finite_diff_matches_oracle(P, h256):
    """prg_finite_diff (orders 1-3) against the oracle: eigensolver densities
    on both sides, differences scaled by 1/h^k (h = 1e-3), so the bar grows
    with the order."""
    print("[finite diff] n=256", flush=True)
    rng = np.random.default_rng(13)
    hs = []
    for _ in range(3):
        x = rng.standard_normal((256, 256)) * 0.01
        hs.append(0.5 * (x + x.T))
    mus = [0.02, -0.01, 0.005]
    got = P.finite_diff(h256, hs, 0.6, mus, 20.0, 3, 0.1, 1e-3)
    ref = O.finite_diff(h256, hs, 0.6, mus, 20.0, 3, 0.1, 1e-3)
    for k, (g, r) in enumerate(zip(got, ref)):
        assert np.max(np.abs(g - r)) < 1e-11 / 1e-3 ** (k + 1), k


@pytest.mark.parametrize("order", [1, 3])
def test_implicit_fermi_response_reference_arity(P, order):
    """prg_implicit_fermi_response at the arity of src/prg_progress_mod.h:240-258
    (separate H1..H3 / P1..P3 handles, NULL above `order`; mu0 and mu by
    value, read back with prg_implicit_fermi_response_mu) gives the oracle's
    result, as the array form does."""
    print(f"[implicit_fermi_response C arity] order={order}", flush=True)
    n = 128
    h0 = O.model_hamiltonian(n, 4)
    hs = [O.banded_perturbation(n, 4, seed=10 + j) * 0.1 for j in range(3)]
    p0, ps, mu0, mu = P.implicit_fermi_response_c(h0, hs, 6, 0.6, 10.0, 64.0, 1e-6, order,
                                                  lin_tol=1e-10)
    p0_o, ps_o, mu0_o, mu_o = O.implicit_fermi_response(h0, hs, 6, 0.6, 10.0, 64.0, 1e-6, order,
                                                        lin_tol=1e-10)
    assert abs(mu0 - mu0_o) < 1e-8
    assert np.allclose(mu, mu_o, atol=1e-8)
    assert np.linalg.norm(p0 - p0_o) < 1e-8
    for a, b in zip(ps, ps_o):
        assert np.linalg.norm(a - b) < 1e-7 * max(1.0, np.linalg.norm(b))

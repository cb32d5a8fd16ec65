"""Pins of the CPU oracle (oracle/progress_oracle.py) -- CPU only.

Every check here is independent of the oracle's own formulas: eigenvalue
closed forms (the SP2 and implicit-Fermi recursions are polynomials /
rational maps of H), the T=0 eigendecomposition density, the reference's
known-answer tests and shipped fixtures, divided differences and finite
differences.  A planted-mistake test shows the closed-form check catches a
plausible error.
"""
import math

import numpy as np
import pytest

import progress_oracle as O


def eig(h):
    return np.linalg.eigh(h)


@pytest.fixture(scope="module")
def model256():
    return O.model_hamiltonian(256, seed=0)


@pytest.fixture(scope="module")
def poly16(golden):
    h = O.read_mtx(str(golden / "poly.16.mtx"))
    return 0.5 * (h + h.T)


@pytest.fixture(scope="module")
def ortho384(golden):
    h = O.read_mtx(str(golden / "hamiltonian_ortho_384.mtx"))
    return 0.5 * (h + h.T)


def test_model_hamiltonian_definition():
    h = O.model_hamiltonian(200, seed=3)
    assert np.array_equal(h, h.T)
    i, j = np.triu_indices(200, 65)
    assert np.all(h[i, j] == 0.0)
    d = np.diag(h)
    assert d.min() >= -1.0 and d.max() <= 1.0
    # first neighbour hopping ~ -exp(-1/8) (1 + 0.1 z)
    nn = np.diag(h, 1)
    assert abs(nn.mean() + math.exp(-1 / 8)) < 0.02


def test_fixture_reader(golden):
    h = O.read_mtx(str(golden / "poly.16.mtx"))
    assert h.shape == (192, 192)
    assert h[0, 0] == -13.294
    t = O.read_mtx(str(golden / "hamiltonian-twolevel-ref.mtx"))
    assert t.shape == (10, 10) and t[0, 0] == -2.0 and t[0, 1] == -1.0


@pytest.mark.parametrize("which", ["model256", "poly16", "ortho384"])
def test_sp2_alg2_closed_form(which, request):
    h = request.getfixturevalue(which)
    n = h.shape[0]
    w, q = eig(h)
    lo, hi = O.gershgorin(h)
    rho, it, pp = O.sp2_alg2(h, 0.5, 10, 200, "Abs", 1e-10)
    x, ite, ppe = O.sp2_alg2_eigen(w, lo, hi, 0.5 * n, 10, 200, "Abs", 1e-10)
    assert it == ite and pp == ppe
    assert np.linalg.norm(rho - (q * (2 * x)) @ q.T) < 1e-10
    d0, _ = O.build_density_T0(h, 0.5)
    assert np.linalg.norm(rho - d0) < 1e-8
    assert abs(np.trace(rho) - n) < 1e-9


def test_gershgorin_bounds_contain_spectrum(model256):
    w, _ = eig(model256)
    lo, hi = O.gershgorin(model256)
    assert lo <= w.min() and w.max() <= hi
    # a diagonal matrix's Gershgorin interval is its exact spectrum
    d = np.diag([3.0, -2.0, 1.0])
    assert O.gershgorin(d) == (-2.0, 3.0)


@pytest.mark.parametrize("fn", ["alg1", "basic", "alg2_genseq", "alg1_genseq"])
def test_sp2_flavours_reach_T0_density(fn, model256):
    h = model256
    d0, _ = O.build_density_T0(h, 0.5)
    if fn == "alg1":
        rho = O.sp2_alg1(h, 0.5, 10, 200, "Abs", 1e-10)[0]
    elif fn == "basic":
        rho = O.sp2_basic(h, 0.5, 10, 200, "Abs", 1e-10)[0]
    elif fn == "alg2_genseq":
        rho = O.sp2_alg2_genseq(h, 0.5, 10, 200, "Abs", 1e-10)[0]
    else:
        rho = O.sp2_alg1_genseq(h, 0.5, 10, 200, "Abs", 1e-10)[0]
    assert np.linalg.norm(rho - d0) < 1e-8


def test_rel_stop_and_idempotency_of_reference_settings(ortho384):
    # the reference tests run "Rel" with minsp2iter 25 and check idempotency
    # of rho/2 (tests/src/main.F90:86-91,260-335).  This shipped 384-orbital
    # fixture has its spectrum clustered at 0 and 1, so half filling is the
    # gapped occupation (bndfil 2/3 lands inside the upper cluster).
    rho, it, pp = O.sp2_alg2(ortho384, 0.5, 10, 100, "Rel", 1e-10)
    assert O.check_idempotency(0.5 * rho) < 1e-8
    rho1, it1, _, _ = O.sp2_alg1(ortho384, 0.5, 10, 100, "Rel", 1e-10)
    assert O.check_idempotency(0.5 * rho1) < 1e-8


def test_replay_variants_reproduce_genseq(model256):
    h = model256
    rho, pp, ic, vv = O.sp2_alg2_genseq(h, 0.5, 10, 200, "Abs", 1e-10)
    r2, vv2 = O.sp2_alg2_seq(h, pp, ic)
    assert np.array_equal(r2, rho) and vv2 == vv[:ic]
    lo, hi = O.gershgorin(h)
    r3, vv3 = O.sp2_alg2_seq_inplace(h, pp, ic, lo, hi)
    assert np.abs(r3 - rho).max() < 1e-13
    r4, vv4 = O.sp2_alg1_seq(h, pp, ic)
    assert np.abs(r4 - rho).max() < 1e-10
    assert np.allclose(vv4, vv2, rtol=1e-6, atol=1e-12)
    # submatrix with every row core = alg1 replay, vv = squared Frobenius norms
    r5, vv5 = O.sp2_submatrix(h, pp, ic, [0.0] * ic, lo, hi, h.shape[0])
    assert np.abs(r5 - r4).max() < 1e-12
    assert np.allclose(np.sqrt(vv5), vv4, rtol=1e-9, atol=1e-14)


def test_homolumogap_recovers_gap(model256):
    h = model256
    n = h.shape[0]
    w, _ = eig(h)
    lo, hi = O.gershgorin(h)
    rho, pp, ic, vv = O.sp2_alg2_genseq(h, 0.5, 10, 200, "Abs", 1e-10)
    eh, el, eg = O.homolumogap(vv, ic, pp, lo, hi)
    assert abs(eh - w[n // 2 - 1]) < 1e-3 * (hi - lo)
    assert abs(el - w[n // 2]) < 1e-3 * (hi - lo)
    assert abs(eg - (w[n // 2] - w[n // 2 - 1])) / (w[n // 2] - w[n // 2 - 1]) < 0.05


def test_sp2sequence_converges_homo_lumo(model256):
    h = model256
    n = h.shape[0]
    w, _ = eig(h)
    lo, hi = O.gershgorin(h)
    ehomo, elumo = w[n // 2 - 1], w[n // 2]
    pp, imax = O.sp2sequence(lo, hi, ehomo, elumo, 1e-12)
    assert imax < 100
    x = O.sp2_poly_eigen(np.array([ehomo, elumo]), lo, hi, pp)
    # homo pushed to 1, lumo to 0 within the error limit
    assert abs(1 - x[0]) + abs(x[1]) <= 1e-12
    # and the sequence purifies the whole spectrum to the T=0 occupations
    xs = O.sp2_poly_eigen(w, lo, hi, pp)
    assert np.abs(xs - np.r_[np.ones(n // 2), np.zeros(n // 2)]).max() < 1e-6


def test_partition_kats(golden):
    # tests/src/main.F90:807-843
    assert len(O.equal_partition(6, 72)) == 12
    assert len(O.equal_partition(7, 72)) == 11
    parts7 = O.equal_partition(7, 72)
    assert [len(p) for p in parts7] == [7] * 10 + [2]
    name, total, parts = O.read_part(str(golden / "test.part"))
    assert len(parts) == 104 and total == 673
    assert sorted(sum(parts, [])) == list(range(673))


def test_subgraph_loop_block_diagonal_is_exact():
    # independent blocks with zero halo: the graph loop equals full SP2 with
    # the same sequence (special case of src/prg_subgraphloop_mod.F90:36-126)
    rng = np.random.default_rng(4)
    blocks = [O.model_hamiltonian(48, seed=s) for s in range(3)]
    n = 144
    h = np.zeros((n, n))
    for b, blk in enumerate(blocks):
        h[48 * b:48 * b + 48, 48 * b:48 * b + 48] = blk
    lo, hi = O.gershgorin(h)
    rho_full, pp, ic, _ = O.sp2_alg2_genseq(h, 0.5, 10, 200, "Abs", 1e-10)
    g = (np.abs(rho_full) > 1e-12).astype(float)
    parts = O.equal_partition(24, n)
    rho, vv = O.subgraph_sp2_loop(h, g, parts, pp, ic, lo, hi)
    ref, vvref = O.sp2_alg2_seq(h, pp, ic)
    assert np.abs(rho - ref).max() < 1e-10
    # core-row vv sums add up to the whole-matrix Frobenius norms
    assert np.allclose(vv, vvref, rtol=1e-8, atol=1e-12)
    del rng


def test_implicit_fermi_closed_form(model256):
    h = model256
    n = h.shape[0]
    w, q = eig(h)
    beta, nsteps, mu = 20.0, 8, 0.6
    p, mu_new = O.implicit_fermi(h, nsteps, 2, 0.5 * n, mu, beta, 0, 1, 1e-9, 1e-18)
    x = O.implicit_fermi_eigen(w, mu, beta, nsteps)
    assert np.linalg.norm(p - (q * x) @ q.T) < 1e-7
    # the recursion approximates the Fermi function
    f = O.fermi(w, mu, beta)
    assert np.abs(x - f).max() < 2e-2
    x12 = O.implicit_fermi_eigen(w, mu, beta, 12)
    assert np.abs(x12 - f).max() < np.abs(x - f).max()


def test_newtonschulz_inverse():
    rng = np.random.default_rng(0)
    a = rng.standard_normal((40, 40))
    a = a @ a.T / 40 + np.eye(40)
    ai0 = np.eye(40) / np.linalg.norm(a, 2)
    ai, it = O.newtonschulz(a, ai0, 1e-12)
    assert np.linalg.norm(ai @ a - np.eye(40)) < 1e-11
    assert it > 3


def test_implicit_fermi_save_inverse_matches_fermi(model256):
    h = model256
    n = h.shape[0]
    w, q = eig(h)
    beta, nsteps = 20.0, 8
    inv = [np.eye(n)] * nsteps
    p, mu, inv, it, ns = O.implicit_fermi_save_inverse(inv, h, nsteps, 0.5 * n, 0.6, beta,
                                                      1e-8, 1e-10, 1)
    assert abs(np.trace(p) / 2 - 0.5 * n) < 1e-6
    x = O.implicit_fermi_eigen(w, mu, beta, nsteps)
    assert np.linalg.norm(p - 2 * (q * x) @ q.T) < 1e-7


def _exact_inverses(h, mu, beta, nsteps):
    n = h.shape[0]
    cnst = beta / 2.0 ** (nsteps + 2)
    p = O.normalize_implicit_fermi(h, cnst, mu)
    inv = []
    for _ in range(nsteps):
        p2 = p @ p
        y = 2 * (p2 - p) + np.eye(n)
        yi = np.linalg.inv(y)
        inv.append(yi)
        p = yi @ p2
    return inv


def test_first_order_response_divided_differences():
    h = O.model_hamiltonian(96, seed=2)
    h1 = O.banded_perturbation(96, 8, seed=5)
    w, q = eig(h)
    beta, nsteps, mu = 10.0, 6, 0.6
    inv = _exact_inverses(h, mu, beta, nsteps)
    p0, p1 = O.implicit_fermi_first_order_response(h, h1, inv, nsteps, mu, beta)
    # Daleckii-Krein response of the scalar map (before the trace correction)
    f = lambda e: O.implicit_fermi_eigen(e, mu, beta, nsteps)
    eps = 1e-6
    df = lambda e: (f(e + eps) - f(e - eps)) / (2 * eps)
    k = O.divided_differences(f, df, w, tol=1e-7)
    h1e = q.T @ h1 @ q
    dk = k * h1e
    x0 = f(w)
    dpdmu = beta * x0 * (1 - x0)
    mu1 = -np.trace(dk) / dpdmu.sum()
    ref = q @ (dk + np.diag(mu1 * dpdmu)) @ q.T
    assert np.linalg.norm(p0 - (q * x0) @ q.T) < 1e-9
    assert np.linalg.norm(p1 - ref) / np.linalg.norm(ref) < 1e-6
    # canonical eigenbasis routine gives the same response
    c1 = O.canon_response(h1e, 48, beta, w, mu, nsteps)
    # prg_canon_response starts from p0 = 0.5 + c(h - mu) (reflected); the
    # recursion is reflection symmetric, so P1 changes sign except for the
    # trace correction which is recomputed: compare through the traceless part
    assert np.allclose(np.abs(c1), np.abs(q.T @ p1 @ q), rtol=1e-6, atol=1e-9)


def test_implicit_fermi_response_orders_are_taylor_coefficients():
    """prg_implicit_fermi_response orders 1-3: along H(l) = H0 + l H1 + l^2 H2
    + l^3 H3 at the returned mu(l) = mu0 + l mu1 + l^2 mu2 + l^3 mu3, the
    density is the scalar recursion applied to the eigenvalues of H(l), so
    P1..P3 are the Taylor coefficients of Q(l) f(w(l)) Q(l)^T.  They are
    read off a degree-6 fit through l = -3s..3s (s = 0.005; fit error 7e-7
    relative at order 3), with non-commuting perturbations.  Dropping the
    cross terms of the higher orders (the response to H2, H3 alone) moves P2
    and P3 by about 80 %, far outside the 1e-5 bar."""
    n, nsteps, beta = 48, 6, 10.0
    h0 = O.model_hamiltonian(n, 3)
    hs = [0.25 * O.banded_perturbation(n, 4, seed=s) for s in (11, 12, 13)]
    p0, ps, mu0, mu = O.implicit_fermi_response(h0, hs, nsteps, 0.6, beta, 0.5 * n, 1e-12, 3,
                                                1e-13)
    for p in ps:
        assert abs(np.trace(p)) < 1e-10

    def density(lam):
        hl = h0 + lam * hs[0] + lam ** 2 * hs[1] + lam ** 3 * hs[2]
        ml = mu0 + lam * mu[0] + lam ** 2 * mu[1] + lam ** 3 * mu[2]
        w, q = eig(hl)
        return (q * O.implicit_fermi_eigen(w, ml, beta, nsteps)) @ q.T

    lams = 0.005 * np.arange(-3, 4)
    vals = np.stack([density(lam) for lam in lams]).reshape(7, -1)
    coef = np.linalg.solve(np.vander(lams, 7, increasing=True), vals).reshape(7, n, n)
    assert np.linalg.norm(coef[0] - p0) < 1e-12
    for j in range(3):
        assert np.linalg.norm(coef[j + 1] - ps[j]) / np.linalg.norm(ps[j]) < 1e-5
    _, ps_x, _, _ = O.implicit_fermi_response(h0, [0.0 * hs[0], hs[1], hs[2]], nsteps, 0.6, beta,
                                              0.5 * n, 1e-12, 3, 1e-13)
    for j in (1, 2):
        assert np.linalg.norm(coef[j + 1] - ps_x[j]) / np.linalg.norm(ps[j]) > 0.5


def test_response_sp2_divided_differences():
    h = O.model_hamiltonian(96, seed=1)
    h1 = O.banded_perturbation(96, 8, seed=7)
    n = h.shape[0]
    w, q = eig(h)
    lo, hi = O.gershgorin(h)
    rho, rho1, it, pp = O.compute_response_sp2(h, h1, 0.5, 10, 200, "Abs", 1e-10)
    # composite map e -> 2 p((hi - e)/(hi - lo)) with the executed sequence
    f = lambda e: 2 * O.sp2_poly_eigen(e, lo, hi, pp)
    eps = 1e-7
    df = lambda e: (f(e + eps) - f(e - eps)) / (2 * eps)
    k = O.divided_differences(f, df, w, tol=1e-9)
    ref = q @ (k * (q.T @ h1 @ q)) @ q.T
    assert np.linalg.norm(rho1 - ref) / np.linalg.norm(ref) < 1e-6
    # finite differences with frozen sequence and bounds
    hstep = 1e-6
    rp, _ = O.sp2_alg2_seq_inplace(h + hstep * h1, pp, len(pp), lo, hi)
    rm, _ = O.sp2_alg2_seq_inplace(h - hstep * h1, pp, len(pp), lo, hi)
    fd = (rp - rm) / (2 * hstep)
    assert np.linalg.norm(rho1 - fd) / np.linalg.norm(fd) < 1e-4
    assert abs(np.trace(rho) - n) < 1e-8


def test_planted_normalisation_sign_error_is_caught(model256, monkeypatch):
    """A plausible mistake (normalising with (H - lo I) instead of (hi I - H))
    must fail the closed-form pin."""
    h = model256
    n = h.shape[0]
    w, q = eig(h)
    lo, hi = O.gershgorin(h)

    def wrong_normalize(hh):
        a, b = O.gershgorin(hh)
        return (hh - a * np.eye(hh.shape[0])) / (b - a)

    monkeypatch.setattr(O, "normalize", wrong_normalize)
    rho, it, pp = O.sp2_alg2(h, 0.5, 10, 200, "Abs", 1e-10)
    d0, _ = O.build_density_T0(h, 0.5)
    assert np.linalg.norm(rho - d0) > 1.0


def test_planted_branch_swap_is_caught(model256):
    """Swapping the SP2 branches (X^2 <-> 2X - X^2) must change the sequence
    check against the eigenvalue closed form."""
    h = model256
    w, q = eig(h)
    lo, hi = O.gershgorin(h)
    rho, ic, pp = O.sp2_alg2(h, 0.5, 10, 200, "Abs", 1e-10)
    swapped = [1 - s for s in pp]
    x = O.sp2_poly_eigen(w, lo, hi, swapped)
    assert np.linalg.norm(rho - (q * (2 * x)) @ q.T) > 1.0


def _fermi_closed_form(h, kbt, ef):
    w, q = eig(h)
    f = 1.0 / (1.0 + np.exp((w - ef) / kbt))
    df = f * f * np.exp((w - ef) / kbt) / kbt
    return (q * (2.0 * f)) @ q.T, float(np.sum(2.0 * df))


def test_gp_t0_single_part_is_the_fermi_density():
    """prg_build_densityGP_T0 with one part holding every node: the density is
    2 f(H) at kT = 0.1 and the Fermi-level step follows the reference's rule
    (error <= 1: ef - error/dN; nocc below target flips the gradient sign)."""
    n = 120
    h = O.model_hamiltonian(n, 2)
    g = np.where(np.abs(h) > 0, 1.0, 0.0)
    w = np.linalg.eigvalsh(h)
    branches = set()
    for shift in (0.0, 0.02, -0.3, 0.4):
        ef = 0.5 * (w[n // 2 - 1] + w[n // 2]) + shift
        rho, ef1 = O.build_density_gp_t0(h, g, 0.5, ef, [list(range(n))])
        d, dn = _fermi_closed_form(h, 0.1, ef)
        assert np.linalg.norm(rho - d) < 1e-10
        nocc = np.trace(d)
        err = abs(nocc - n)
        signed = dn if nocc >= n else -dn
        step = (nocc if err > 1.0 else err) / signed
        branches.add(err > 1.0)
        assert abs(ef1 - (ef - step)) < 1e-12 * max(1.0, abs(ef1))
        if err <= 1.0:
            # the small-error step moves Ef toward the target occupation
            assert (ef1 - ef) * (n - nocc) > 0
    assert branches == {True, False}


def test_gp_t0_block_diagonal_parts_are_exact():
    """Block-diagonal H whose blocks are the parts: the halos are empty and the
    assembled density equals 2 f(H) of the whole matrix."""
    rng = np.random.default_rng(7)
    sizes = [30, 45, 25]
    n = sum(sizes)
    h = np.zeros((n, n))
    parts, o = [], 0
    for s in sizes:
        a = rng.standard_normal((s, s))
        h[o:o + s, o:o + s] = 0.5 * (a + a.T)
        parts.append(list(range(o, o + s)))
        o += s
    g = np.where(h != 0, 1.0, 0.0)
    rho, _ = O.build_density_gp_t0(h, g, 0.5, 0.0, parts)
    d, _ = _fermi_closed_form(h, 0.1, 0.0)
    assert np.linalg.norm(rho - d) < 1e-10
    # coupled parts: every row comes from the part that owns the node as core
    # (its core+halo solution), with zeros outside that part's index list
    h2 = h.copy()
    h2[29, 30] = h2[30, 29] = 0.3
    g2 = np.where(h2 != 0, 1.0, 0.0)
    rho2, _ = O.build_density_gp_t0(h2, g2, 0.5, 0.0, parts)
    for p, core in enumerate(parts):
        idx, nc = O.matrix2submatrix_index(g2, core, True)
        dp, _ = _fermi_closed_form(h2[np.ix_(idx, idx)], 0.1, 0.0)
        assert np.allclose(rho2[np.ix_(idx[:nc], idx)], dp[:nc], atol=1e-12)
        others = np.setdiff1d(np.arange(n), idx)
        assert np.all(rho2[np.ix_(idx[:nc], others)] == 0.0)



def test_bfs_partition_kats():
    """Path graph of 10 nodes into 3 parts and a 4 x 4 grid into 4 parts (hand
    traced: BFS from the far end, proportional cuts, left part first)."""
    n = 10
    g = np.zeros((n, n))
    for i in range(n - 1):
        g[i, i + 1] = g[i + 1, i] = 1.0
    assert O.bfs_partition(g, 3) == [[7, 8, 9], [3, 4, 5, 6], [0, 1, 2]]
    assert O.bfs_partition(g, 1) == [list(range(10))]
    assert O.bfs_partition(g, 2) == [[5, 6, 7, 8, 9], [0, 1, 2, 3, 4]]


def _sp2_fermi_poly(x0, sgnlist):
    x = np.asarray(x0, dtype=float).copy()
    for s in sgnlist:
        x = 2.0 * x - x * x if s == 1 else x * x
    return x


def test_sp2_fermi_matches_eigenvalue_recursion(golden):
    """prg_sp2_fermi_init + prg_sp2_fermi on the reference's 384-orbital
    Hamiltonian against the reference test's bar (tests/src/main.F90:740-800;
    its 600-orbital file is not shipped): ||2 X - 2 f_FD||_F <= 0.01 with f_FD
    the Fermi function at the init's beta and the Fermi level that holds nocc
    electrons.  The routine's mu is the shift of the normalise_fermi map, not
    an energy: the polynomial's half-occupation energy sp2_inverse(0.5) is
    the Fermi level, and mu itself is far from it on this spectrum."""
    h = O.read_mtx(str(golden / "hamiltonian_ortho_384.mtx"))
    n = h.shape[0]
    nocc = 2.0 / 3.0 * n
    ini = O.sp2_fermi_init(h, 20, nocc, 1.0, 1e-9, 1e-12)
    x, mu = O.sp2_fermi(h, 0, 20, nocc, ini["mu"], ini["beta"], ini["h1"], ini["hN"],
                        ini["sgnlist"], 1e-4, 1e-12)
    w, q = eig(h)
    xi = ini["x"]
    wi = np.linalg.eigvalsh(xi)
    assert np.all(wi > -1e-9) and np.all(wi < 1 + 1e-9)
    from scipy.optimize import brentq
    beta = ini["beta"]
    ef = brentq(lambda m: np.sum(1.0 / (1.0 + np.exp(beta * (w - m)))) - nocc, w[0], w[-1])
    f = 1.0 / (1.0 + np.exp(beta * (w - ef)))
    assert np.linalg.norm(2.0 * x - (q * (2.0 * f)) @ q.T) <= 0.01
    # the SP2 polynomial's own half-occupation energy is that Fermi level
    e_half = O.sp2_inverse(0.5, mu, ini["h1"], ini["hN"], ini["sgnlist"])
    assert abs(e_half - ef) < 1e-4
    # the shift itself is not the Fermi level on this spectrum
    f_mu = 1.0 / (1.0 + np.exp(ini["beta"] * (w - mu)))
    assert np.linalg.norm(2.0 * x - (q * (2.0 * f_mu)) @ q.T) > 1.0
    assert abs(np.trace(x) - nocc) < 1e-3


def test_sp2_fermi_replay_is_the_polynomial(golden):
    """With osteps = 1 the matrix recursion is the sign-list polynomial of the
    normalise_fermi map applied to the eigenvalues, plus the occupation
    correction lambda * DX (DX = -beta X (I - X))."""
    h = O.model_hamiltonian(96, 3)
    n = h.shape[0]
    lo, hi = O.gershgorin(h)
    sg = [1, -1, -1, 1, -1, 1, 1, -1]
    mu0, beta = 0.3, 4.0
    x, mu1 = O.sp2_fermi(h, 1, len(sg), 0.5 * n, mu0, beta, lo, hi, sg, 1e-4, 1e-12)
    w, q = eig(h)
    p = _sp2_fermi_poly((hi - mu0 - w) / (hi - lo), sg)
    dx = -beta * p * (1.0 - p)
    lam = (0.5 * n - p.sum()) / dx.sum()
    assert abs(mu1 - (mu0 + lam)) < 1e-10
    assert np.linalg.norm(x - (q * (p + lam * dx)) @ q.T) < 1e-9


def test_sp2_inverse_round_trip_and_entropy_quadrature():
    """sp2_inverse inverts the sign-list polynomial of the normalise_fermi map
    (p(x0(sp2_inverse(f))) = f), and the entropy function integrates it: GG
    agrees with an independent adaptive quadrature at mesh points."""
    from scipy.integrate import quad
    sg = [1, -1, 1, -1, -1, 1]
    mu, h1, hn = 0.2, -3.0, 2.5
    for f in (0.05, 0.3, 0.77, 0.999):
        e = O.sp2_inverse(f, mu, h1, hn, sg)
        x0 = (hn - mu - e) / (hn - h1)
        assert abs(_sp2_fermi_poly([x0], sg)[0] - f) < 1e-10
    gg, ee = O.sp2_entropy_function(mu, h1, hn, sg)
    total = quad(lambda t: O.sp2_inverse(t, mu, h1, hn, sg), 0.0, 1.0, limit=200)[0]
    for k in (1000, 5000, 9000):
        part = quad(lambda t: O.sp2_inverse(t, mu, h1, hn, sg), 0.0, ee[k], limit=200)[0]
        # the integrand has square-root endpoints: two-point Gauss per 1e-4
        # cell agrees with adaptive quadrature to ~1e-6
        assert abs(gg[k] - (part - total * ee[k])) < 2e-6
    assert abs(gg[0]) < 1e-15 and abs(gg[-1]) < 1e-12


@pytest.mark.parametrize("nmod", [8, 11, 14])
def test_crt_square_reproduces_x_squared(nmod):
    """The int8 Chinese-remainder square: for X with spectrum in [0, 1] the
    reconstruction equals X^2 up to the rounding of X to 2^-a (a = 32, 44, 52:
    the largest with 2^(2a - 1 + 0.001) <= M, since the residues stand for
    C - 2^a A whose entries are at most 2^(2a) / 4), i.e. |error| <=
    2^-a (||X_i||_1 + ||X_j||_1) / 2 + FP64 noise."""
    rng = np.random.default_rng(nmod)
    n = 200
    q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    x = (q * rng.uniform(0.0, 1.0, n)) @ q.T
    x = 0.5 * (x + x.T)
    ps, qs, a, m = O.crt_constants(nmod)
    assert a == {8: 32, 11: 44, 14: 52}[nmod]
    assert 4 ** a / 2 * 2.0 ** 0.001 <= m
    assert a == 52 or 4 ** (a + 1) / 2 * 2.0 ** 0.001 > m
    # a general product keeps the |C| <= 2^(2a) bound: one bit less
    assert O.crt_constants(nmod, square=False)[2] == {8: 31, 11: 43, 14: 52}[nmod]
    got = O.crt_square(x, nmod)
    ref = x @ x
    l1 = np.abs(x).sum(axis=1)
    bound = 2.0 ** -(a + 1) * (l1[:, None] + l1[None, :]) + 1e-13
    assert np.all(np.abs(got - ref) <= bound)


def test_crt_square_planted_mistakes_are_caught():
    """A wrong CRT weight (q_k = 1) or a dropped modulus breaks the product."""
    rng = np.random.default_rng(1)
    n = 64
    q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    x = (q * rng.uniform(0.0, 1.0, n)) @ q.T
    ref = x @ x
    good = O.crt_square(x, 11)
    assert np.max(np.abs(good - ref)) < 1e-10
    orig = O.crt_constants

    def wrong_q(nmod):
        ps, qs, a, m = orig(nmod)
        return ps, [1] * len(qs), a, m

    O.crt_constants = wrong_q
    try:
        assert np.max(np.abs(O.crt_square(x, 11) - ref)) > 1e-3
    finally:
        O.crt_constants = orig

    def dropped(nmod):
        ps, qs, a, m = orig(nmod)
        return ps[:-1], qs[:-1], a, m

    O.crt_constants = dropped
    try:
        assert np.max(np.abs(O.crt_square(x, 11) - ref)) > 1e-3
    finally:
        O.crt_constants = orig


def _commuting_pair(n, seed, lo, hi):
    rng = np.random.default_rng(seed)
    q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    x = (q * rng.uniform(0.0, 1.0, n)) @ q.T
    y = (q * rng.uniform(lo, hi, n)) @ q.T
    return 0.5 * (x + x.T), 0.5 * (y + y.T)


@pytest.mark.parametrize("nmod", [8, 11, 14])
def test_crt_product_reproduces_commuting_product(nmod):
    """The int8 CRT general product of commuting symmetric X (spectrum in
    [0, 1]) and Y (spectrum in [2, 6], scale 8): the reconstruction equals X Y
    up to the rounding of the operands, |error| <= 2^-(a+1) (sX ||Y_j||_1 +
    sY ||X_i||_1) + 4^-(a+1) sX sY n + FP64 noise."""
    n = 160
    x, y = _commuting_pair(n, nmod, 2.0, 6.0)
    sx, sy = O.crt_scale(x), O.crt_scale(y)
    assert sx == 1.0 and sy == 8.0
    assert np.max(np.sqrt(np.sum(y * y, axis=0))) <= sy
    ps, qs, a, m = O.crt_constants(nmod, square=False)
    got = O.crt_product(x, y, nmod)
    assert np.array_equal(got, got.T)
    ref = x @ y
    lx, ly = np.abs(x).sum(axis=1), np.abs(y).sum(axis=1)
    bound = (2.0 ** -(a + 1) * (sx * ly[None, :] + sy * lx[:, None]) +
             4.0 ** -(a + 1) * sx * sy * n + 1e-13 * sx * sy)
    assert np.all(np.abs(got - ref) <= bound)
    # the square is the X = Y case of the same arithmetic (on a grid one bit
    # finer): both within their roundings of X^2
    assert np.max(np.abs(O.crt_product(x, x, nmod) - O.crt_square(x, nmod))) <= \
        2.0 ** -a * 2.0 * np.max(lx) + 1e-13


def test_crt_product_planted_scale_mistake_is_caught():
    """Scaling Y by 1 instead of the power of two above its row norms lets the
    integer products pass M / 2 and wrap: the reconstruction breaks."""
    x, y = _commuting_pair(96, 5, 2.0, 6.0)
    ref = x @ y
    assert np.max(np.abs(O.crt_product(x, y, 11) - ref)) < 1e-9
    orig = O.crt_scale
    O.crt_scale = lambda z: 1.0
    try:
        assert np.max(np.abs(O.crt_product(x, y, 11) - ref)) > 1e-3
    finally:
        O.crt_scale = orig


def test_orthogonalize_round_trip_with_lowdin_factor():
    """Z = S^(-1/2) (Loewdin): Z^T S Z = I, and deorthogonalize inverts
    orthogonalize for Z^T Z = S^-1 ... i.e. Z (Z^T A Z) Z^T = S^-1 A S^-1."""
    rng = np.random.default_rng(3)
    n = 50
    b = rng.standard_normal((n, n)) * 0.05
    s = np.eye(n) + 0.5 * (b + b.T)
    w, q = np.linalg.eigh(s)
    z = (q / np.sqrt(w)) @ q.T
    assert np.allclose(O.orthogonalize(s, z), np.eye(n), atol=1e-12)
    a = rng.standard_normal((n, n))
    a = a + a.T
    sinv = np.linalg.inv(s)
    assert np.allclose(O.deorthogonalize(O.orthogonalize(a, z), z), sinv @ a @ sinv, atol=1e-10)


@pytest.mark.parametrize("nmod", [8, 10, 11, 13])
def test_crt_range_bound_is_tight(nmod):
    """The largest shifted product the bound admits: X = e_1 e_1^T / 2 (the
    eigenvalue 1/2 where x^2 - x = -1/4, so C_11 - 2^a A_11 = -2^(2a) / 4)
    reconstructs to FP64 rounding of the fraction sum.  Planted mistakes: one
    more bit of scale (a + 1) makes |C_11 - 2^a A_11| exceed M / 2 and the
    reconstruction wraps by M / 4^(a+1); and without the shift (residues
    read as C itself, whose entries reach 2^(2a)) X = e_1 e_1^T wraps at
    this a."""
    n = 8
    x = np.zeros((n, n))
    x[0, 0] = 0.5
    sq = np.zeros((n, n))
    sq[0, 0] = 0.25
    assert np.max(np.abs(O.crt_square(x, nmod) - sq)) < 1e-12
    orig = O.crt_constants
    ps, qs, a, m = orig(nmod)

    def one_more_bit(k, square=True):
        ps, qs, a, m = orig(k, square)
        return ps, qs, a + 1, m
    O.crt_constants = one_more_bit
    try:
        assert abs(O.crt_square(x, nmod)[0, 0] - 0.25) > 0.1
    finally:
        O.crt_constants = orig
    # the unshifted reconstruction (S mod 1 times M / 4^a) of e_1 e_1^T
    e = np.zeros((n, n))
    e[0, 0] = 1.0
    av = np.rint(e * 2.0 ** a)
    s = 0.0
    for p, q in zip(ps, qs):
        v = float((int(av[0, 0]) % p) ** 2 * q % p)
        s += (v - p * round(v / p)) / p
    unshifted = (s - round(s)) * (float(m) / 4.0 ** a)
    assert abs(unshifted - 1.0) > 0.1
    assert abs(O.crt_square(e, nmod)[0, 0] - 1.0) < 1e-12


def test_sp2_basic_tcore_pins():
    """prg_sp2_basic_tcore runs the basic branch rule for exactly maxit + 1
    iterations.  Pins: the same rule applied to the eigenvalues of X (fresh
    trace = sum of eigenvalues) takes the same branches until the trace has
    converged (after that |Tr X - occ| is rounding noise and both branches
    leave the idempotent X unchanged), the density equals Q 2x(e) Q^T, and it
    agrees with prg_sp2_basic run to convergence."""
    n = 96
    h = O.model_hamiltonian(n, 2)
    maxit = 60
    rho, pp = O.sp2_basic_tcore(h, 0.5, maxit)
    assert len(pp) == maxit + 1
    w, q = np.linalg.eigh(h)
    lo, hi = O.gershgorin(h)
    x = (hi - w) / (hi - lo)
    pp_e, conv = [], None
    for k in range(maxit + 1):
        if x.sum() - 0.5 * n <= 0.0:
            x = 2.0 * x - x * x
            pp_e.append(0)
        else:
            x = x * x
            pp_e.append(1)
        if conv is None and abs(x.sum() - 0.5 * n) < 1e-8:
            conv = k
    assert conv is not None and conv < maxit - 5
    assert pp_e[:conv] == pp[:conv]
    assert np.linalg.norm(rho - (q * (2.0 * x)) @ q.T) < 1e-9
    rb, it = O.sp2_basic(h, 0.5, 10, 200, "Abs", 1e-10)
    assert np.linalg.norm(rb - rho) < 1e-9
    # a planted mistake: the branch sign flipped (X^2 while Tr X <= occ) drives
    # the trace away from occ, and the eigenvalue check above catches it
    x = (hi - w) / (hi - lo)
    for _ in range(maxit + 1):
        x = x * x if x.sum() - 0.5 * n <= 0.0 else 2.0 * x - x * x
    assert np.linalg.norm(rho - (q * (2.0 * x)) @ q.T) > 1.0


def _if_scalar_and_derivative(nsteps, mu, beta):
    """The exact-inverse implicit-Fermi recursion on one eigenvalue and its
    derivative by the chain rule: x0 = 1/2 - c (e - mu), x <- x^2 / (2x^2 - 2x + 1),
    dx_{k+1}/dx_k = 2 x (1 - x) / (2x^2 - 2x + 1)^2."""
    cnst = beta / (2.0 ** (nsteps + 2))

    def f(e):
        return O.implicit_fermi_eigen(e, mu, beta, nsteps)

    def df(e):
        x = 0.5 - cnst * (np.asarray(e) - mu)
        d = -cnst * np.ones_like(x)
        for _ in range(nsteps):
            q = 2.0 * x * x - 2.0 * x + 1.0
            d = d * 2.0 * x * (1.0 - x) / (q * q)
            x = x * x / q
        return d

    return f, df


@pytest.mark.parametrize("canonical", [False, True])
def test_implicit_fermi_adjoint_daleckii_krein(canonical):
    """The oracle's reverse sweep equals the closed form: P = f(H) for the
    scalar recursion f, so dTr(P A)/dH = Q (K o Q^T A Q) Q^T with K the
    divided differences of f (Daleckii-Krein); the canonical variant is the
    same with A - kappa I."""
    n, nsteps, mu, beta = 60, 6, 0.3, 8.0
    h = O.model_hamiltonian(n, 4)
    rng = np.random.default_rng(2)
    a = rng.standard_normal((n, n))
    a = 0.5 * (a + a.T)
    g = O.implicit_fermi_adjoint(h, a, nsteps, mu, beta, canonical)
    assert np.array_equal(g, g.T)
    e, q = np.linalg.eigh(h)
    f, df = _if_scalar_and_derivative(nsteps, mu, beta)
    k = O.divided_differences(f, df, e)
    aa = a
    if canonical:
        p = (q * f(e)) @ q.T
        cm = beta * (p - p @ p)
        aa = a - (np.sum(a * cm) / np.trace(cm)) * np.eye(n)
    ref = q @ (k * (q.T @ aa @ q)) @ q.T
    assert np.max(np.abs(g - ref)) < 1e-10 * max(1.0, np.max(np.abs(ref)))


def test_implicit_fermi_adjoint_canonical_matches_reference_response():
    """Tr(A P1[H1]) for the reference's first-order response (with its mu
    correction) equals Tr(H1 G) for the canonical adjoint G, for any H1."""
    n, nsteps, mu, beta = 50, 6, 0.3, 8.0
    h = O.model_hamiltonian(n, 5)
    rng = np.random.default_rng(3)
    a = rng.standard_normal((n, n))
    a = 0.5 * (a + a.T)
    cnst = beta / (2.0 ** (nsteps + 2))
    xs = [O.normalize_implicit_fermi(h, cnst, mu)]
    inv = []
    for _ in range(nsteps):
        x = xs[-1]
        inv.append(np.linalg.inv(2.0 * (x @ x - x) + np.eye(n)))
        xs.append(inv[-1] @ (x @ x))
    g = O.implicit_fermi_adjoint(h, a, nsteps, mu, beta, canonical=True)
    for s in range(3):
        h1 = rng.standard_normal((n, n))
        h1 = 0.5 * (h1 + h1.T)
        _, p1 = O.implicit_fermi_first_order_response(h, h1, inv, nsteps, mu, beta)
        assert abs(np.sum(a * p1) - np.sum(h1 * g)) < 1e-10 * max(1.0, abs(np.sum(a * p1)))


def test_implicit_fermi_adjoint_planted_mistake_is_caught():
    """Dropping the W F + (W F)^T term of the reverse step breaks the closed form."""
    n, nsteps, mu, beta = 40, 5, 0.3, 8.0
    h = O.model_hamiltonian(n, 6)
    a = np.eye(n)
    good = O.implicit_fermi_adjoint(h, a, nsteps, mu, beta)
    e, q = np.linalg.eigh(h)
    f, df = _if_scalar_and_derivative(nsteps, mu, beta)
    ref = q @ (O.divided_differences(f, df, e) * (q.T @ a @ q)) @ q.T
    assert np.max(np.abs(good - ref)) < 1e-10

    def broken(h, a, nsteps, mu, beta, canonical=False):
        n = h.shape[0]
        cnst = beta / (2.0 ** (nsteps + 2))
        xs = [O.normalize_implicit_fermi(h, cnst, mu)]
        invs = []
        for _ in range(nsteps):
            x = xs[-1]
            invs.append(np.linalg.inv(2.0 * (x @ x - x) + np.eye(n)))
            xs.append(invs[-1] @ (x @ x))
        g = a.copy()
        for i in reversed(range(nsteps)):
            w = invs[i] @ g
            u = w - 2.0 * (w @ xs[i + 1])
            v = xs[i] @ (0.5 * (u + u.T))
            g = v + v.T
        return -cnst * g

    assert np.max(np.abs(broken(h, a, nsteps, mu, beta) - ref)) > 1e-6


def test_canon_response_vector_is_the_fermi_derivative():
    """Off the diagonal, the vector response equals the central difference of
    the eigenbasis Fermi density f(L + lambda H1) (Daleckii-Krein); the
    diagonal correction makes the trace vanish (canonical)."""
    n, beta, mu0 = 40, 6.0, 0.1
    rng = np.random.default_rng(4)
    e = np.sort(rng.uniform(-1.0, 1.0, n))
    h1 = rng.standard_normal((n, n))
    h1 = 0.5 * (h1 + h1.T)
    p1 = O.canon_response_vector(h1, beta, e, mu0)
    assert abs(np.trace(p1)) < 1e-12

    def dens(lam):
        w, q = np.linalg.eigh(np.diag(e) + lam * h1)
        return (q / (1.0 + np.exp(beta * (w - mu0)))) @ q.T

    eps = 1e-5
    fd = (dens(eps) - dens(-eps)) / (2.0 * eps)
    off = ~np.eye(n, dtype=bool)
    # central difference error O(eps^2) on the closest eigenvalue pairs: 2e-8
    assert np.max(np.abs(p1[off] - fd[off])) < 1e-7
    # the diagonal is the derivative shifted by -dmu f'(e), dmu = Tr(dP) / Tr f'
    t = np.exp(beta * (e - mu0))
    df = -beta * t / (1.0 + t) ** 2
    ref = np.diag(fd) - (np.trace(fd) / np.sum(df)) * df
    assert np.max(np.abs(np.diag(p1) - ref)) < 1e-7


def test_canon_response_p1_dpdmu_plus_correction_is_canon_response():
    n, beta, mu0, m = 30, 10.0, 0.2, 8
    rng = np.random.default_rng(5)
    e = np.sort(rng.uniform(-1.0, 1.0, n))
    h1 = rng.standard_normal((n, n))
    h1 = 0.5 * (h1 + h1.T)
    p1, dpdmu = O.canon_response_p1_dpdmu(h1, beta, e, mu0, m)
    ref = O.canon_response(h1, n / 2, beta, e, mu0, m)
    corr = p1.copy()
    corr[np.diag_indices(n)] += (-np.trace(p1) / np.sum(dpdmu)) * dpdmu
    assert np.max(np.abs(corr - ref)) < 1e-14


def test_response_rs_matches_finite_differences():
    """Rayleigh-Schroedinger first-order response of the T = 0 density equals
    the central difference of prg_build_density_T0 on a gapped system, and
    keeps the trace (fixed occupation)."""
    n = 60
    rng = np.random.default_rng(6)
    q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    e = np.concatenate([np.sort(rng.uniform(-3.0, -1.0, n // 2)), np.sort(rng.uniform(1.0, 3.0, n // 2))])
    h = (q * e) @ q.T
    h = 0.5 * (h + h.T)
    h1 = rng.standard_normal((n, n))
    h1 = 0.5 * (h1 + h1.T)
    rs = O.compute_response_rs(h, h1, 0.5)
    fd = O.compute_response_fd(h, h1, 1e-5, 0.5)
    assert np.max(np.abs(rs - fd)) < 1e-7
    assert abs(np.trace(rs)) < 1e-10
    # a planted sign error (e_i - e_j) is caught
    e0, c = np.linalg.eigh(h)
    v = c.T @ h1 @ c
    de = e0[:, None] - e0[None, :]
    vt = np.where(np.eye(n, dtype=bool), 0.0, v / np.where(np.eye(n, dtype=bool), 1.0, de))
    occ = O.rs_occupations(n, 0.5)
    c1 = c @ vt
    bad = (c * occ) @ c1.T + c1 @ (occ[:, None] * c.T)
    assert np.max(np.abs(bad - fd)) > 1e-3


@pytest.mark.parametrize("k", [4, 6])
def test_implicit_fermi_even_order_is_the_scalar_map(k):
    """Order k > 2 (prg_setup_linsys): with converged CG solves the level map
    on every eigenvalue is x <- x^k / (x^k + (1 - x)^k) from
    x0 = 1/2 - beta / (4 k^nsteps) (e - mu); a mistake in the system (the
    k = 2 matrix A for k = 4) changes the result."""
    n, nsteps, mu, beta = 48, 5, 0.1, 6.0
    h = O.model_hamiltonian(n, 3)
    p, _ = O.implicit_fermi(h, nsteps, k, n / 2.0, mu, beta, 0, 1, 1e30, 1e-24)
    e, q = np.linalg.eigh(h)
    x = 0.5 - beta / (4.0 * k ** nsteps) * (e - mu)
    for _ in range(nsteps):
        x = x ** k / (x ** k + (1.0 - x) ** k)
    assert np.max(np.abs(p - (q * x) @ q.T)) < 1e-10
    xs = np.array([0.2, 0.4, 0.7, 0.9])
    a, b = O.setup_linsys(np.diag(xs), k)
    assert np.allclose(np.diag(b), xs ** k) and np.allclose(np.diag(a), xs ** k + (1 - xs) ** k)
    wrong_a = 2.0 * (np.diag(xs ** 2) - np.diag(xs)) + np.eye(4)
    assert not np.allclose(wrong_a, a)


def test_pcg_solves_and_beats_plain_cg_on_a_badly_scaled_system():
    """prg_pcg on the system the implicit-Fermi recursion solves (B = I: the
    inverse) with a diagonal spread over three decades: it reaches A^-1, the
    Jacobi preconditioner needs far fewer iterations than plain CG, and the
    returned residual is B - A X."""
    n = 80
    rng = np.random.default_rng(11)
    q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    core = (q * rng.uniform(1.0, 2.0, n)) @ q.T
    dsc = np.diag(np.logspace(0, 1.5, n))
    a = dsc @ core @ dsc
    a = 0.5 * (a + a.T)
    b = np.eye(n)
    x, r, k = O.pcg(a, np.zeros((n, n)), b, 1e-18)
    assert np.max(np.abs(a @ x - b)) < 1e-8
    assert np.allclose(r, b - a @ x, atol=1e-9)
    _, k_cg = O.conjgrad(a, np.zeros((n, n)), b, 1e-18, maxit=1000)
    assert k < k_cg / 4


def test_finite_diff_matches_fermi_taylor_coefficients():
    """prg_finite_diff on commuting (diagonal) H0, H1..H3: the density is the
    Fermi function of x(l) = e + l a + l^2 b + l^3 c - mu(l) element-wise, so
    the three differences approach its Taylor coefficients f' x', (f'' x'^2 +
    f' x'')/2 and (f''' x'^3 + 3 f'' x' x'' + f' x''')/6 at l = 0, with errors
    O(h) (forward) and O(h^2) (central)."""
    rng = np.random.default_rng(12)
    n, beta, mu0 = 12, 4.0, 0.1
    e = rng.uniform(-1.0, 1.0, n)
    a, b, c = (rng.uniform(-1.0, 1.0, n) for _ in range(3))
    mus = [0.2, -0.1, 0.05]
    h = 1e-3
    o = O.finite_diff(np.diag(e), [np.diag(a), np.diag(b), np.diag(c)], mu0, mus, beta, 3, 0.0, h)
    f = 1.0 / (1.0 + np.exp(beta * (e - mu0)))
    f1 = -beta * f * (1 - f)
    f2 = beta ** 2 * f * (1 - f) * (1 - 2 * f)
    f3 = -beta ** 3 * f * (1 - f) * (1 - 6 * f + 6 * f * f)
    x1, x2, x3 = a - mus[0], 2 * (b - mus[1]), 6 * (c - mus[2])
    t1 = f1 * x1
    t2 = (f2 * x1 ** 2 + f1 * x2) / 2
    t3 = (f3 * x1 ** 3 + 3 * f2 * x1 * x2 + f1 * x3) / 6
    assert np.max(np.abs(np.diag(o[0]) - t1)) < 20 * h
    assert np.max(np.abs(np.diag(o[1]) - t2)) < 100 * h * h
    assert np.max(np.abs(np.diag(o[2]) - t3)) < 1000 * h * h
    assert np.max(np.abs(o[0] - np.diag(np.diag(o[0])))) == 0.0


# --------------------------------------------------------------------------
# core+halo sets (bml_matrix2submatrix_index), halo order, and the remaining
# oracle functions: implicit_fermi_zero, sp2_entropy_ts, test_density_matrix
# --------------------------------------------------------------------------


def _random_sparse_sym(n, density, rng):
    m = (rng.random((n, n)) < density) * rng.normal(size=(n, n))
    m = m + m.T
    np.fill_diagonal(m, rng.normal(size=n))
    return m


def _halo_cases():
    rng = np.random.default_rng(21)
    n = 60
    g = _random_sparse_sym(n, 0.03, rng)
    h = _random_sparse_sym(n, 0.04, rng)
    cores = [list(range(0, 6)), [7, 3, 3, 41], [59], list(rng.permutation(n)[:10])]
    return g, h, cores


@pytest.mark.parametrize("double_jump", [True, False])
@pytest.mark.parametrize("with_h", [True, False])
def test_halo_set_is_the_symbolic_product_pattern(double_jump, with_h):
    """The set matrix2submatrix_index lists equals the nonzero columns of the
    core rows of (I + |G| + |H|)(I + |G|) (double jump) or (I + |G| + |H|)
    (single jump), computed here by integer matrix products.  The list holds
    the cores first (repeats dropped) and no node twice."""
    g, h, cores = _halo_cases()
    hh = h if with_h else None
    for core in cores:
        idx, nc = O.matrix2submatrix_index(g, core, double_jump, hh)
        uniq = list(dict.fromkeys(int(c) for c in core))
        assert idx[:nc] == uniq
        assert len(set(idx)) == len(idx)
        assert set(idx) == O.halo_set_bruteforce(g, core, double_jump, hh)


def test_halo_set_planted_mistakes_are_caught():
    """A single-jump extraction and one that ignores H both fail the
    double-jump symbolic pattern with H on these graphs."""
    g, h, cores = _halo_cases()
    ref = [O.halo_set_bruteforce(g, c, True, h) for c in cores]
    single = [set(O.matrix2submatrix_index(g, c, False, h)[0]) for c in cores]
    no_h = [set(O.matrix2submatrix_index(g, c, True, None)[0]) for c in cores]
    assert any(a != b for a, b in zip(single, ref))
    assert any(a != b for a, b in zip(no_h, ref))


def test_halo_order_does_not_move_the_core_rows():
    """The order of the halo nodes inside a part is a similarity permutation of
    the block, which commutes with the SP2 polynomial: shuffling the halo
    moves the core rows of the part's density (and its vv) by rounding only."""
    rng = np.random.default_rng(5)
    n = 96
    h = O.model_hamiltonian(n, seed=2)
    lo, hi = O.gershgorin(h)
    _, pp, ic, _ = O.sp2_alg2_genseq(h, 0.5, 10, 200, "Abs", 1e-10)
    g = (np.abs(h) > 0.05).astype(float)
    core = list(range(20, 36))
    idx, nc = O.matrix2submatrix_index(g, core, True, h)
    halo = idx[nc:]
    perm = idx[:nc] + list(rng.permutation(halo))
    d1, vv1 = O.sp2_submatrix(h[np.ix_(idx, idx)], pp, ic, [0.0] * ic, lo, hi, nc)
    d2, vv2 = O.sp2_submatrix(h[np.ix_(perm, perm)], pp, ic, [0.0] * ic, lo, hi, nc)
    pos = {k: i for i, k in enumerate(perm)}
    back = [pos[k] for k in idx]
    assert np.max(np.abs(d1[:nc, :] - d2[:nc, :][:, back])) < 1e-12
    assert np.allclose(vv1, vv2, rtol=1e-12, atol=1e-14)


def _gapped_hamiltonian(n, seed, gap=2.0):
    """Symmetric H with eigenvalues in [-3, -gap/2] u [gap/2, 3] (half each)."""
    rng = np.random.default_rng(seed)
    e = np.concatenate([rng.uniform(-3.0, -gap / 2, n // 2), rng.uniform(gap / 2, 3.0, n - n // 2)])
    q, _ = np.linalg.qr(rng.normal(size=(n, n)))
    return (q * e) @ q.T


def test_implicit_fermi_zero_closed_form():
    """prg_implicit_fermi_zero, method 0 (CG solves): with converged solves the
    recursion is the scalar map x <- x^2/(2x^2 - 2x + 1) on the eigenvalues of
    X0 = (0.5 + c mu) I - c H, c = 0.5 min(1/(mu - lo), 1/(hi - mu)) from the
    Gershgorin interval (src/prg_implicit_fermi_mod.F90:379-382), i.e.
    implicit_fermi_eigen at beta = c 2^(nsteps + 2).  On a gapped H it tends
    to the T = 0 projector onto the states below mu."""
    n, mu = 40, 0.1
    h = _gapped_hamiltonian(n, 3)
    w, q = eig(h)
    lo, hi = O.gershgorin(h)
    for nsteps in (4, 10):
        c = 0.5 * min(1.0 / (mu - lo), 1.0 / (hi - mu))
        x = O.implicit_fermi_eigen(w, mu, c * 2.0 ** (nsteps + 2), nsteps)
        p = O.implicit_fermi_zero(h, nsteps, mu, 0, 1e-24)
        assert np.linalg.norm(p - (q * x) @ q.T) < 1e-9
    proj = (q[:, w < mu]) @ q[:, w < mu].T
    assert np.linalg.norm(O.implicit_fermi_zero(h, 14, mu, 0, 1e-24) - proj) < 1e-8


def test_implicit_fermi_zero_planted_mistakes_are_caught(monkeypatch):
    """A constant without the 0.5 factor, or the reflected map (X0 = 0.5 +
    c (H - mu)), must fail the closed form (model H, states at mu: the map's
    slope there is what the constant sets)."""
    n, mu, nsteps = 40, 0.0, 3
    h = O.model_hamiltonian(n, seed=3)
    w, q = eig(h)
    lo, hi = O.gershgorin(h)
    c = 0.5 * min(1.0 / (mu - lo), 1.0 / (hi - mu))
    ref = (q * O.implicit_fermi_eigen(w, mu, c * 2.0 ** (nsteps + 2), nsteps)) @ q.T
    real = O.normalize_implicit_fermi
    monkeypatch.setattr(O, "normalize_implicit_fermi", lambda hh, cn, m: real(hh, 2.0 * cn, m))
    assert np.linalg.norm(O.implicit_fermi_zero(h, nsteps, mu, 0, 1e-24) - ref) > 1e-3
    monkeypatch.setattr(O, "normalize_implicit_fermi", lambda hh, cn, m: real(hh, -cn, m))
    assert np.linalg.norm(O.implicit_fermi_zero(h, nsteps, mu, 0, 1e-24) - ref) > 1.0


def test_implicit_fermi_zero_method1_is_a_function_of_h():
    """Method 1 reuses one CG inverse (tolerance 0.9) of the first level's
    Y = 2(X^2 - X) + I; that inverse is a polynomial in Y, so the result is a
    spectral function of H: it commutes with H and equals the same recursion
    run on the diagonal matrix of eigenvalues, rotated back."""
    n, mu = 40, 0.1
    h = _gapped_hamiltonian(n, 4)
    w, q = eig(h)
    p = O.implicit_fermi_zero(h, 6, mu, 1, 1e-12)
    pd = O.implicit_fermi_zero(np.diag(w), 6, mu, 1, 1e-12)
    # the Gershgorin interval of diag(w) differs from H's: run both with the
    # same interval by comparing only what the interval does not change
    assert np.linalg.norm(p @ h - h @ p) < 1e-9
    lo, hi = O.gershgorin(h)
    lod, hid = O.gershgorin(np.diag(w))
    if (lo, hi) == (lod, hid):
        assert np.linalg.norm(p - (q * np.diag(pd)) @ q.T) < 1e-9


def _entropy_G(x, mu, h1, hn, sg, total):
    from scipy.integrate import quad
    return quad(lambda t: O.sp2_inverse(t, mu, h1, hn, sg), 0.0, x, limit=200)[0] - total * x


def test_sp2_entropy_ts_is_the_quadrature_of_the_inverse():
    """sp2_entropy_ts on D0 = Q diag(d) Q^T: the sum over |d_i| of the entropy
    function G(x) = int_0^x sp2_inverse - x int_0^1 sp2_inverse, evaluated
    here by adaptive quadrature (independent of the 1e-4 mesh); eigenvalue
    invariance through the rotation Q."""
    from scipy.integrate import quad
    sg = [1, -1, 1, -1, -1, 1]
    mu, h1, hn = 0.2, -3.0, 2.5
    gg, ee = O.sp2_entropy_function(mu, h1, hn, sg)
    total = quad(lambda t: O.sp2_inverse(t, mu, h1, hn, sg), 0.0, 1.0, limit=200)[0]
    rng = np.random.default_rng(8)
    d = np.concatenate([rng.uniform(0.02, 0.98, 10), [-0.3, 0.51234567]])
    qm, _ = np.linalg.qr(rng.normal(size=(12, 12)))
    d0 = (qm * d) @ qm.T
    ts = O.sp2_entropy_ts(d0, gg, ee)
    gx = [_entropy_G(abs(x), mu, h1, hn, sg, total) for x in d]
    assert abs(ts - sum(gx)) < 5e-5
    # one eigenvalue at a time (errors of opposite sign cannot cancel)
    err = [abs(O.sp2_entropy_ts(np.array([[x]]), gg, ee) - g) for x, g in zip(d, gx)]
    assert max(err) < 1e-5
    # planted mistake: the mesh read one cell off (GG(j) for GG(j+1))
    off = [abs(O.sp2_entropy_ts(np.array([[x]]), np.roll(gg, 1), ee) - g) for x, g in zip(d, gx)]
    assert max(off) > 5e-5


def test_test_density_matrix_converges_to_the_occupation_root():
    """prg_test_density_matrix (osteps = 0): the returned mu satisfies
    |sum_i f(e_i; mu) - nocc| <= occErrLimit, P = Q f Q^T at that mu, and mu
    is within the Newton bound of the root found by bracketing."""
    from scipy.optimize import brentq
    n, beta = 50, 4.0
    h = O.model_hamiltonian(n, seed=7)
    w, q = eig(h)
    nocc = 0.5 * n + 0.3
    limit = 1e-9
    # undamped Newton, as the reference: start near the root (mu = 0.73)
    p, mu = O.test_density_matrix(h, beta, 0.5, nocc, 0, limit)
    occ = np.sum(1.0 / (1.0 + np.exp(beta * (w - mu))))
    assert abs(occ - nocc) <= limit
    f = 1.0 / (1.0 + np.exp(beta * (w - mu)))
    assert np.linalg.norm(p - (q * f) @ q.T) < 1e-12
    root = brentq(lambda m: np.sum(1.0 / (1.0 + np.exp(beta * (w - m)))) - nocc, -10, 10,
                  xtol=1e-15)
    dndmu = beta * np.sum(f * (1 - f))
    assert abs(mu - root) <= 2 * limit / dndmu
    # osteps > 0: exactly osteps Newton steps; from a start near the root
    # three steps already converge quadratically
    _, mu3 = O.test_density_matrix(h, beta, root + 0.05, nocc, 3, limit)
    assert abs(mu3 - root) < 1e-10


def test_test_density_matrix_planted_mistake_is_caught():
    """A Newton step that drops the sum of squares (dN/dmu = beta Tr P instead
    of beta (Tr P - sum P_ij^2)) does not reach the root in three steps."""
    from scipy.optimize import brentq
    n, beta = 50, 4.0
    h = O.model_hamiltonian(n, seed=7)
    w, q = eig(h)
    nocc = 0.5 * n + 0.3
    root = brentq(lambda m: np.sum(1.0 / (1.0 + np.exp(beta * (w - m)))) - nocc, -10, 10,
                  xtol=1e-15)

    def broken(h, beta, mu, nocc, osteps):
        for _ in range(osteps):
            f = 1.0 / (1.0 + np.exp(beta * (w - mu)))
            p = (q * f) @ q.T
            mu = mu + (nocc - np.trace(p)) / (beta * np.trace(p))
        return mu

    assert abs(broken(h, beta, root + 0.05, nocc, 3) - root) > 1e-6


# prg_build_zmatGP (graph-partitioned congruence transform)

def _spd_overlap(n, blocks, seed):
    """Overlap S = I + 0.1 B with B symmetric, nonzero only inside the given
    diagonal blocks (dense within each), and its graph (the blocks' pattern)."""
    rng = np.random.default_rng(seed)
    s = np.eye(n)
    g = np.zeros((n, n))
    start = 0
    for b in blocks:
        x = rng.standard_normal((b, b))
        s[start:start + b, start:start + b] += 0.05 * (x + x.T)
        g[start:start + b, start:start + b] = 1.0
        start += b
    return s, g


def test_zmat_gp_single_part_is_the_loewdin_factor():
    """One part holding every node: Z = S^-1/2, i.e. Z S Z = I, Z = Z^T."""
    s, g = _spd_overlap(40, [40], 1)
    z = O.build_zmat_gp(s, g, [list(range(40))])
    assert np.allclose(z @ s @ z, np.eye(40), atol=1e-12)
    assert np.allclose(z, z.T, atol=1e-14)


def test_zmat_gp_block_diagonal_overlap_is_exact():
    """Block-diagonal S whose graph is its pattern (each block a clique): the
    core+halo of any core covers its blocks, so every core row of Z_p is the
    row of S^-1/2 whatever the partition (BFS parts and an interleaved one)."""
    s, g = _spd_overlap(60, [12, 20, 8, 20], 2)
    exact = O.buildzdiag(s)
    for parts in (O.bfs_partition(g, 5), [list(range(k, 60, 7)) for k in range(7)]):
        z = O.build_zmat_gp(s, g, parts)
        assert np.max(np.abs(z - exact)) < 1e-12


def test_zmat_gp_planted_mistakes_are_caught():
    """Planted mistakes: S^+1/2 in place of S^-1/2, and parts without their
    halo (core blocks only), both move Z far from S^-1/2 on the block case."""
    s, g = _spd_overlap(60, [12, 20, 8, 20], 2)
    exact = O.buildzdiag(s)
    parts = [list(range(k, 60, 7)) for k in range(7)]
    orig = O.buildzdiag
    try:
        O.buildzdiag = lambda m: (lambda w, u: (u * np.sqrt(w)) @ u.T)(*np.linalg.eigh(m))
        assert np.max(np.abs(O.build_zmat_gp(s, g, parts) - exact)) > 1e-2
    finally:
        O.buildzdiag = orig
    z = np.zeros((60, 60))
    for core in parts:
        zp = O.buildzdiag(s[np.ix_(core, core)])
        z[np.ix_(core, core)] = zp
    assert np.max(np.abs(z - exact)) > 1e-3


def test_crt_symmetric_residue_without_correction():
    """The int8 CRT kernels' residue shortcut: for an odd modulus p <= 255 and
    an integer-valued double |v| / p < 2^38, v - p rint(v (1/p)) is already
    the symmetric residue in [-(p-1)/2, (p-1)/2] (the FP64 quotient is within
    2^-14 of v / p, which stays at least 1/(2p) from a half-integer); for p =
    256 only the tie +128 needs folding to -128.  Checked in IEEE FP64 on
    random values up to 2^45 and on every residue class near the top."""
    rng = np.random.default_rng(11)
    for p in O.CRT_MODULI:
        inv = 1.0 / p
        top = 2.0 ** 45
        v = np.concatenate([np.rint(rng.uniform(-top, top, 20000)),
                            np.floor(top / p) * p - np.arange(2 * p, dtype=float),
                            -np.floor(top / p) * p + np.arange(2 * p, dtype=float)])
        m = v - p * np.rint(v * inv)
        if p == 256:
            m = np.where(m > 127, m - p, m)
        exact = np.array([((int(x) + p // 2) % p) - p // 2 for x in v], dtype=float)
        if p == 256:
            exact = np.where(exact > 127, exact - p, exact)
        assert np.array_equal(m, exact), p
        assert m.min() >= -128 and m.max() <= 127


def test_crt_product_epilogue_magic_read():
    """The int8 product epilogue's residue of v = a q (a the int32 MFMA sum,
    |a| <= 2^31, q < p): t = rint(a fl(q fl(1/p))), then w = (a q + 1.5 2^52)
    - p t, exact in FP64, whose low 32 bits are the residue as an int32.  Its
    low byte equals the byte of the symmetric residue for every modulus
    (including the p = 256 tie +-128, byte 0x80), on random sums and on the
    residue classes next to the extremes.  A quotient taken as a plain
    truncation (int(a q / p)) instead of rint fails it for every odd p."""
    rng = np.random.default_rng(12)
    magic = 6755399441055744.0
    ps, qs, _, _ = O.crt_constants(11, square=False)
    for p, qk in zip(ps, qs):
        q = float(qk)
        qmi = q * (1.0 / p)
        top = 2.0 ** 31
        a = np.concatenate([np.rint(rng.uniform(-top, top, 20000)),
                            np.arange(-2 * p, 2 * p, dtype=float),
                            top - 1 - np.arange(2 * p, dtype=float), -top + np.arange(2 * p, dtype=float)])
        w = (a * q + magic) - p * np.rint(a * qmi)
        lo = (w.view(np.uint64) & np.uint64(0xffffffff)).astype(np.int64)
        got = lo & 0xff
        exact = np.array([(((int(x) * int(q)) + p // 2) % p) - p // 2 for x in a])
        assert np.array_equal(got, exact & 0xff), p
        if p == 256:
            continue  # any integer quotient leaves the same low byte for p = 256
        # planted mistake: truncating the quotient gives a different byte
        bad = (a * q + magic) - p * np.trunc(a * qmi)
        badlo = (bad.view(np.uint64) & np.uint64(0xffffffff)).astype(np.int64) & 0xff
        assert not np.array_equal(badlo, exact & 0xff), p


def test_crt_exact_int_product_equals_int64_matmul():
    """The oracle's FP64 route to the exact residue products equals numpy's
    int64 matmul on residue-range matrices (including all +-128 entries,
    the largest partial sums)."""
    rng = np.random.default_rng(4)
    a = rng.integers(-128, 128, size=(150, 300))
    b = rng.integers(-128, 128, size=(120, 300))
    assert np.array_equal(O._exact_int_product(a, b), a @ b.T)
    full = np.full((8, 1000), -128, dtype=np.int64)
    assert np.array_equal(O._exact_int_product(full, full), full @ full.T)

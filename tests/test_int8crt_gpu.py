"""GPU parity of the int8 Chinese-remainder SP2 products (PRG_PREC_INT8CRT).

The products are exact integer products of X's residues, so the only error is
the rounding of X to 2^-a (a = 44 with the default 11 moduli).  Bars (stated
per test): the same branch sequence as the FP64 oracle, the density within
1e-7 (Frobenius) of the oracle's, and the error shrinking as moduli are
added.  Replays of a fixed branch sequence isolate the product accuracy from
the branch decisions.
"""
import numpy as np
import pytest

import progress_oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture()
def P():
    import progress_amd
    progress_amd.set_precision(progress_amd.PREC_INT8CRT)
    progress_amd.set_int8_moduli(11)
    yield progress_amd
    progress_amd.set_int8_moduli(11)
    progress_amd.set_precision(progress_amd.PREC_FP64)


@pytest.mark.parametrize("n,seed", [(256, 0), (300, 1), (1024, 0)])
def test_crt_alg2_matches_oracle(P, n, seed):
    print(f"[int8crt alg2] n={n} seed={seed}", flush=True)
    h = O.model_hamiltonian(n, seed)
    rho_o, pp_o, ic_o, _ = O.sp2_alg2_genseq(h, 0.5, 10, 200, "Abs", 1e-10)
    rho_g, pp_g, ic_g, _ = P.sp2_alg2_genseq(h, 0.5, 10, 200, "Abs", 1e-10)
    assert ic_g == ic_o and list(pp_g) == list(pp_o)
    # rounding X to 2^-43 per entry, amplified by the inverse gap: 1e-7 bar
    assert np.linalg.norm(rho_g - rho_o) < 1e-7
    assert np.array_equal(rho_g, rho_g.T)
    assert abs(np.trace(rho_g) - n) < 1e-8


def test_crt_replay_isolates_products(P):
    """A fixed branch sequence: the int8 path and the FP64 oracle apply the
    same polynomial, so the difference is the product error alone."""
    n = 512
    print(f"[int8crt replay] n={n}", flush=True)
    h = O.model_hamiltonian(n, 3)
    _, pp, ic, _ = O.sp2_alg2_genseq(h, 0.5, 10, 200, "Abs", 1e-10)
    rho_o = O.sp2_alg2_seq(h, pp, ic)
    rho_o = rho_o[0] if isinstance(rho_o, tuple) else rho_o
    rho_g, _ = P.sp2_alg2_seq(h, pp, ic)
    assert np.linalg.norm(rho_g - rho_o) < 1e-7


def test_crt_error_falls_with_moduli(P):
    """The rounding of X is 2^-(a+1) with a = 31, 39, 46, 52 for 8, 10, 12, 14
    moduli: each step of two moduli must cut the error, and 14 moduli reach
    the FP64 path's own agreement with the oracle on this input (1e-9, the
    bar of the FP64 parity tests: summation order alone moves the density by
    ~1e-10 here)."""
    n = 384
    print(f"[int8crt moduli] n={n}", flush=True)
    h = O.model_hamiltonian(n, 5)
    d0, _ = O.build_density_T0(h, 0.5)
    rho_o, _, pp_o = O.sp2_alg2(h, 0.5, 10, 200, "Abs", 1e-10)
    errs = {}
    for npm in (8, 10, 12, 14):
        P.set_int8_moduli(npm)
        assert P.get_int8_moduli() == npm
        _, pp, ic, _ = O.sp2_alg2_genseq(h, 0.5, 10, 200, "Abs", 1e-10)
        rho_g, _ = P.sp2_alg2_seq(h, pp, ic)
        errs[npm] = float(np.linalg.norm(rho_g - rho_o))
    print("errors vs FP64 oracle:", errs, flush=True)
    assert errs[14] < 1e-9
    assert errs[12] < 1e-8
    assert errs[8] > errs[10] > errs[12]
    assert float(np.linalg.norm(rho_o - d0)) < 1e-9


def test_crt_rejects_bad_moduli_count(P):
    with pytest.raises(P.ProgressError):
        P.set_int8_moduli(5)
    with pytest.raises(P.ProgressError):
        P.set_int8_moduli(15)


@pytest.mark.parametrize("n", [2000])
def test_crt_device_api_accuracy(P, n):
    """Device-pointer entry point (the benchmark's path) at a size where the
    oracle's eigendecomposition still runs in seconds; n = 2000 is padded to
    ld = 2048 and the padding must stay zero."""
    import torch
    print(f"[int8crt device api] n={n}", flush=True)
    h = O.model_hamiltonian(n, 1)
    ld = P.padded_ld(n)
    hd = torch.zeros((ld, ld), dtype=torch.float64, device="cuda")
    hd[:n, :n] = torch.from_numpy(h)
    rd = torch.zeros_like(hd)
    it, pp = P.sp2_alg2_device(hd, rd, n, 0.5, 1, 200, "Abs", 1e-10, precision=P.PREC_INT8CRT)
    rho = rd[:n, :n].cpu().numpy()
    d0, _ = O.build_density_T0(h, 0.5)
    rho_o, it_o, pp_o = O.sp2_alg2(h, 0.5, 1, 200, "Abs", 1e-10)
    assert it == it_o and list(pp[:it]) == list(pp_o)
    assert np.linalg.norm(rho - rho_o) < 1e-7
    assert np.linalg.norm(rho - d0) < 1e-6
    assert float(rd[n:, :].abs().max()) == 0.0 and float(rd[:, n:].abs().max()) == 0.0


@pytest.mark.parametrize("npm,n", [(11, 300), (8, 256), (14, 200)])
def test_crt_single_product_matches_oracle_arithmetic(P, npm, n):
    """One replayed step (pp = [1]: X1 = X0^2) isolates the kernel's product:
    the same integer residue arithmetic in the oracle (crt_square) must agree
    to FP64 rounding of the fraction sum, far below the 2^-a rounding of X."""
    print(f"[int8crt one product] moduli={npm} n={n}", flush=True)
    P.set_int8_moduli(npm)
    h = O.model_hamiltonian(n, 7)
    rho_g, _ = P.sp2_alg2_seq(h, [1], 1)
    # the GPU's own X0 (an empty replay returns 2 X0): up to 11 moduli X is
    # carried by its residues, i.e. on the 2^-a grid; with more, as FP64
    _, _, a, _ = O.crt_constants(npm)
    rho0, _ = P.sp2_alg2_seq(h, [1], 0)
    x0 = 0.5 * rho0
    assert np.max(np.abs(x0 - O.normalize(h))) <= (2.0 ** -(a + 1) if a <= 45 else 1e-15)
    ref = 2.0 * O.crt_square(x0, npm)
    # the two sides sum the npm fractions (each <= 1/2) in different orders:
    # FP64 rounding of the sum times M / 2^(2a), times 2 for rho = 2 X^2; and
    # where X is carried by its residues, X1 itself is rounded to 2^-a
    _, _, a, m = O.crt_constants(npm)
    noise = max(1e-13, 8.0 * npm * 2.0 ** -52 * (float(m) / 4.0 ** a))
    noise += 2.0 ** -a if a <= 45 else 0.0
    assert np.max(np.abs(rho_g - ref)) < noise
    # and X^2 itself to the 2^-a rounding
    bound = max(2.0 ** -(a - 8), noise)
    assert np.max(np.abs(rho_g - 2.0 * (x0 @ x0))) < bound


def _commuting_pair(n, seed, lo, hi):
    rng = np.random.default_rng(seed)
    q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    x = (q * rng.uniform(0.0, 1.0, n)) @ q.T
    y = (q * rng.uniform(lo, hi, n)) @ q.T
    return 0.5 * (x + x.T), 0.5 * (y + y.T)


def _device(P, z, n):
    import torch
    ld = P.padded_ld(n)
    t = torch.zeros((ld, ld), dtype=torch.float64, device="cuda")
    if z is not None:
        t[:n, :n] = torch.from_numpy(z)
    return t


@pytest.mark.parametrize("npm,n", [(10, 200), (11, 256), (14, 300)])
def test_crt_general_product_matches_oracle(P, npm, n):
    """C = X Y for commuting symmetric X (spectrum [0, 1]) and Y (spectrum
    [2, 6], scaled by 8) through the device entry point in int8 CRT product
    precision: the oracle's integer arithmetic (crt_product) to the FP64
    rounding of the fraction sum, exactly symmetric, padding untouched."""
    print(f"[int8crt general product] moduli={npm} n={n}", flush=True)
    x, y = _commuting_pair(n, npm, 2.0, 6.0)
    a, b, c = _device(P, x, n), _device(P, y, n), _device(P, None, n)
    P.set_int8_moduli(npm)
    P.set_product_precision(P.PREC_INT8CRT)
    try:
        P.gemm_counters(reset=True)
        P.symmetric_product_device(a, b, c, n)
        nprod, _, exe, int8_ops = P.gemm_counters(reset=True, int8=True)
    finally:
        P.set_product_precision(P.PREC_FP64)
    assert nprod == 1 and exe == 0.0 and int8_ops > 0.0
    got = c[:n, :n].cpu().numpy()
    ref = O.crt_product(x, y, npm)
    _, _, abits, m = O.crt_constants(npm, square=False)
    sxy = O.crt_scale(x) * O.crt_scale(y)
    noise = max(1e-13, 8.0 * npm * 2.0 ** -52 * (float(m) / 4.0 ** abits)) * sxy
    assert np.max(np.abs(got - ref)) < noise
    assert np.array_equal(got, got.T)
    ld = P.padded_ld(n)
    if ld > n:
        assert float(c[n:, :].abs().max()) == 0.0 and float(c[:, n:].abs().max()) == 0.0


def test_symmetric_product_fp64_matches_numpy(P):
    """The same entry point in FP64 product precision: the FP64 MFMA kernel."""
    n = 200
    x, y = _commuting_pair(n, 3, -1.0, 1.0)
    a, b, c = _device(P, x, n), _device(P, y, n), _device(P, None, n)
    assert P.get_product_precision() == P.PREC_FP64
    P.symmetric_product_device(a, b, c, n)
    got = c[:n, :n].cpu().numpy()
    assert np.max(np.abs(got - x @ y)) < 1e-13
    assert np.array_equal(got, got.T)


@pytest.mark.parametrize("method", [0, 1])
def test_implicit_fermi_int8_products_match_oracle(P, method):
    """prg_implicit_fermi with its symmetric products (X^2, the CG and
    Newton-Schulz products, Inv X^2) as int8 CRT products, 12 moduli (operands
    rounded to 2^-46 of their scale).  Against the oracle's FP64 recursion:
    the FP64 path's own bars (P to 1e-8 Frobenius, mu to 1e-9: the CG / NS
    stopping points differ by rounding); against the GPU's FP64 products:
    1e-9 (measured 4e-11), which isolates the product arithmetic."""
    print(f"[int8crt implicit fermi] method={method}", flush=True)
    h = O.model_hamiltonian(256, 0)
    args = (8, 2, 128.0, 0.6, 20.0, method, 3, 1e-9, 1e-16 if method == 0 else 1e-10)
    p_o, mu_o = O.implicit_fermi(h, *args)
    assert P.get_product_precision() == P.PREC_FP64
    p64, _ = P.implicit_fermi(h, *args)
    P.set_int8_moduli(12)
    P.set_product_precision(P.PREC_INT8CRT)
    try:
        P.gemm_counters(reset=True)
        p, mu = P.implicit_fermi(h, *args)
        nprod, _, exe, int8_ops = P.gemm_counters(reset=True, int8=True)
    finally:
        P.set_product_precision(P.PREC_FP64)
    assert nprod > 0 and exe == 0.0 and int8_ops > 0.0
    assert np.array_equal(p, p.T)
    assert abs(mu - mu_o) < 1e-9
    assert np.linalg.norm(p - p_o) < 1e-8
    assert np.linalg.norm(p - p64) < 1e-9


def test_full_size_n8192_fp64_and_int8_properties(P):
    """The benchmark's workload itself (N = 8192 model Hamiltonian, seed 0,
    occupation N/2), where the oracle's eigendecomposition would take minutes
    on the host: size-independent properties instead.  Both precisions give
    the same branch sequence (66 steps on this input); each density is exactly
    symmetric, keeps Tr D = N to 1e-8 (measured 2e-11) and is within 1e-6 (Frobenius) of the
    GPU eigendecomposition density 2 Q_occ Q_occ^T (measured 4e-12 FP64,
    1.7e-7 with 10 moduli); the two agree to 1e-6; D^2 / 2 = D to 1e-6."""
    import torch
    n = 8192
    print(f"[int8crt full size] n={n}", flush=True)
    h = torch.from_numpy(O.model_hamiltonian(n, 0)).cuda()
    ld = P.padded_ld(n)
    assert ld == n
    evals, evecs = torch.linalg.eigh(h)
    q = evecs[:, : n // 2]
    ddiag = 2.0 * (q @ q.T)
    del evecs, q
    out = {}
    for prec, nmod in ((P.PREC_FP64, 11), (P.PREC_INT8CRT, 10)):
        P.set_int8_moduli(nmod)
        rho = torch.zeros_like(h)
        it, pp = P.sp2_alg2_device(h, rho, n, 0.5, 1, 200, "Abs", 1e-10, precision=prec)
        err = float(torch.linalg.norm(rho - ddiag))
        print(f"  precision {prec}: {it} steps, ||D - D_diag||_F = {err:.3e}", flush=True)
        assert err < 1e-6
        assert bool(torch.equal(rho, rho.T))
        assert abs(float(torch.trace(rho)) - n) < 1e-8
        assert float(torch.linalg.norm(rho @ rho / 2 - rho)) < 1e-6
        out[prec] = (it, list(pp[:it]), rho)
    (it64, pp64, r64), (it8, pp8, r8) = out[P.PREC_FP64], out[P.PREC_INT8CRT]
    assert it8 == it64 == 66 and pp8 == pp64
    assert float(torch.linalg.norm(r8 - r64)) < 1e-6


def test_full_size_n8192_implicit_fermi_properties(P):
    """prg_implicit_fermi at fixed mu (8 levels, beta 20, Newton-Schulz to
    1e-10) on the N = 8192 model Hamiltonian, with FP64 products and with
    12-moduli int8 CRT products.  With exact inverses the recursion maps each
    eigenvalue through the oracle's scalar map, so P must equal
    Q f(e) Q^T built from a GPU eigendecomposition and the oracle's
    implicit_fermi_eigen on the 8192 eigenvalues, to 1e-10 (Frobenius) with
    FP64 products and 1e-8 with int8 products (measured 1.3e-12 and 3.5e-10:
    the Newton-Schulz stopping point and the operands' rounding to 2^-46 of
    their scale are the only approximations), exactly symmetric, and the two
    product precisions within 1e-8 of each other."""
    import torch
    n = 8192
    print(f"[implicit fermi full size] n={n}", flush=True)
    h = torch.from_numpy(O.model_hamiltonian(n, 0)).cuda()
    assert P.padded_ld(n) == n
    e, q = torch.linalg.eigh(h)
    mu = 0.5 * float(e[n // 2 - 1] + e[n // 2])
    nsteps, beta = 8, 20.0
    f = torch.from_numpy(O.implicit_fermi_eigen(e.cpu().numpy(), mu, beta, nsteps)).cuda()
    ref = (q * f) @ q.T
    del q
    out = []
    for prec in (P.PREC_FP64, P.PREC_INT8CRT):
        P.set_int8_moduli(12)
        P.set_product_precision(prec)
        try:
            p = torch.zeros_like(h)
            P.implicit_fermi_device(h, p, n, nsteps, mu, beta, 1e-10)
        finally:
            P.set_product_precision(P.PREC_FP64)
        err = float(torch.linalg.norm(p - ref))
        print(f"  product precision {prec}: ||P - Q f(e) Q^T||_F = {err:.3e}", flush=True)
        assert err < (1e-10 if prec == P.PREC_FP64 else 1e-8)
        assert bool(torch.equal(p, p.T))
        out.append(p)
    assert float(torch.linalg.norm(out[0] - out[1])) < 1e-8


def test_crt_after_shutdown_and_init(P):
    """prg_progress_shutdown frees every workspace, the int8 product's item
    list and per-modulus constants included; the next int8 solve must upload
    them again (a stale cache would run every item as modulus 0, tile 0).
    Same density and branch sequence before and after, and the oracle's."""
    n = 300
    print(f"[int8crt shutdown/init] n={n}", flush=True)
    h = O.model_hamiltonian(n, 1)
    rho_o, pp_o, ic_o, _ = O.sp2_alg2_genseq(h, 0.5, 10, 200, "Abs", 1e-10)
    before, pp_b, ic_b, _ = P.sp2_alg2_genseq(h, 0.5, 10, 200, "Abs", 1e-10)
    P.lib().prg_progress_shutdown()
    P.lib().prg_progress_init()
    after, pp_a, ic_a, _ = P.sp2_alg2_genseq(h, 0.5, 10, 200, "Abs", 1e-10)
    assert ic_a == ic_b == ic_o and list(pp_a) == list(pp_b) == list(pp_o)
    assert np.array_equal(after, before)
    assert np.linalg.norm(after - rho_o) < 1e-7


@pytest.mark.parametrize("config", [803, 804, 805, 904, 944])
@pytest.mark.parametrize("n", [1000, 2000])
def test_crt_product_configs_match_oracle(P, config, n):
    """Every int8 CRT product kernel configuration (one item per workgroup
    with 8 waves x 3/4/5 ring stages; the persistent 8-wave kernel with a
    4-stage ring, one workgroup per CU walking its XCD's items), forced: one
    replayed step (X1 = X0^2) agrees with the oracle's integer
    arithmetic (crt_square) to FP64 rounding of the fraction sum, at n = 1000
    (110 items) and n = 2000 (396 items: the persistent workgroups take one
    or two items each, diagonal and off-diagonal)."""
    print(f"[int8crt config] {config} n={n}", flush=True)
    npm = 11
    P.set_int8_moduli(npm)
    P.set_int8_config(config)
    try:
        h = O.model_hamiltonian(n, 9)
        rho_g, _ = P.sp2_alg2_seq(h, [1], 1)
        used = P.int8_waves()
        rho0, _ = P.sp2_alg2_seq(h, [1], 0)
    finally:
        P.set_int8_config(0)
    assert used == config
    _, _, a, m = O.crt_constants(npm)
    ref = 2.0 * O.crt_square(0.5 * rho0, npm)
    noise = max(1e-13, 8.0 * npm * 2.0 ** -52 * (float(m) / 4.0 ** a)) + 2.0 ** -a
    assert np.max(np.abs(rho_g - ref)) < noise

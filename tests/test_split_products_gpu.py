"""GPU checks of the BF16 hi+lo split products of the iterative solvers
(prg_set_product_precision(PRG_PREC_BF16SPLIT)).

Each operand x is carried as hi = bf16(x) and lo = bf16((x - hi) 2^8), which
holds x to 2^-18 of its magnitude, and a product is hi*hi + 2^-8 (hi*lo +
lo*hi) with FP32 accumulation.  Bar of one product (stated per test):
|C - X Y| <= 2^-14 (|X| |Y|) elementwise.  It covers the representation error
(2^-18 per operand, plus the dropped lo*lo term, 2^-18) and FP32 accumulation
over K <= 384 terms (K 2^-24 = 2^-15.4 in the worst case), while a dropped
cross term (2^-9) or a wrong lo scale fails it.  The closed form X Y (numpy,
FP64) is the reference.
"""
import numpy as np
import pytest

import progress_oracle as O


@pytest.fixture()
def P():
    import progress_amd
    yield progress_amd
    progress_amd.set_product_precision(progress_amd.PREC_FP64)


def _commuting_pair(n, seed, lo, hi):
    rng = np.random.default_rng(seed)
    q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    x = (q * rng.uniform(0.0, 1.0, n)) @ q.T
    y = (q * rng.uniform(lo, hi, n)) @ q.T
    return 0.5 * (x + x.T), 0.5 * (y + y.T)


def _device(z, n, ld):
    import torch
    t = torch.zeros((ld, ld), dtype=torch.float64, device="cuda")
    if z is not None:
        t[:n, :n] = torch.from_numpy(z)
    return t


@pytest.mark.gpu
@pytest.mark.parametrize("n,lo,hi", [(200, 2.0, 6.0), (256, -1.0, 1.0), (300, -40.0, 25.0)])
def test_bf16_split_product_matches_closed_form(P, n, lo, hi):
    """C = X Y for commuting symmetric X (spectrum [0, 1]) and Y (spectrum
    [lo, hi]) through the device entry point: within 2^-14 (|X| |Y|) of the
    FP64 product, exactly symmetric, padding untouched, and counted as one
    product with no FP64 or int8 MFMA work."""
    print(f"[bf16 split product] n={n} spectrum=[{lo}, {hi}]", flush=True)
    x, y = _commuting_pair(n, n, lo, hi)
    ld = P.padded_ld(n)
    a, b, c = _device(x, n, ld), _device(y, n, ld), _device(None, n, ld)
    P.set_product_precision(P.PREC_BF16SPLIT)
    try:
        P.gemm_counters(reset=True)
        P.symmetric_product_device(a, b, c, n)
        nprod, alg, exe, int8_ops = P.gemm_counters(reset=True, int8=True)
    finally:
        P.set_product_precision(P.PREC_FP64)
    assert nprod == 1 and exe == 0.0 and int8_ops == 0.0 and alg == 2.0 * n ** 3
    got = c[:n, :n].cpu().numpy()
    bound = 2.0 ** -14 * (np.abs(x) @ np.abs(y))
    assert np.all(np.abs(got - x @ y) <= bound)
    assert np.array_equal(got, got.T)
    if ld > n:
        assert float(c[n:, :].abs().max()) == 0.0 and float(c[:, n:].abs().max()) == 0.0


def test_bf16_split_bar_rejects_a_dropped_cross_term():
    """The bar above, applied on the CPU to the split product with one cross
    term left out (and to one with the lo plane scaled by 2^-11 instead of
    2^-8), fails: the tolerance pins both terms."""
    n = 128
    x, y = _commuting_pair(n, 7, -1.0, 1.0)

    def bf16(v):
        f = v.astype(np.float32).view(np.uint32).astype(np.uint64)
        r = ((f + 0x7FFF + ((f >> 16) & 1)) >> 16) << 16
        return r.astype(np.uint32).view(np.float32).astype(np.float64)

    def split(v, s):
        h = bf16(v)
        return h, bf16((v - h) * s)

    bound = 2.0 ** -14 * (np.abs(x) @ np.abs(y))
    (xh, xl), (yh, yl) = split(x, 256.0), split(y, 256.0)
    full = xh @ yh + (xh @ yl + xl @ yh) / 256.0
    assert np.all(np.abs(full - x @ y) <= bound)
    dropped = xh @ yh + (xh @ yl) / 256.0
    assert not np.all(np.abs(dropped - x @ y) <= bound)
    (xh, xl), (yh, yl) = split(x, 2048.0), split(y, 2048.0)
    wrong = xh @ yh + (xh @ yl + xl @ yh) / 256.0
    assert not np.all(np.abs(wrong - x @ y) <= bound)


@pytest.mark.gpu
def test_product_precision_rejects_fp16_split(P):
    """Only FP64, int8 CRT and BF16 split are product precisions."""
    with pytest.raises(RuntimeError):
        P.set_product_precision(P.PREC_F16SPLIT)
    assert P.get_product_precision() == P.PREC_FP64


@pytest.mark.gpu
@pytest.mark.parametrize("method,tol", [(1, 1e-4), (0, 1e-6)])
def test_implicit_fermi_bf16_products(P, method, tol):
    """prg_implicit_fermi with its symmetric products as BF16 split products
    (Newton-Schulz to ||I - Y Inv||_F <= 1e-4, or CG to ||R||_F^2 <= 1e-6,
    both above the products' noise floor).  Against the oracle's FP64
    recursion at the same tolerance: P within 1e-3 relative (Frobenius),
    exactly symmetric, and mu within 2e-4: a relative trace error of 1e-4
    (the products' 16-bit noise carried through 8 levels) moves the Newton
    step on mu by 1e-4 N / (beta sum f (1 - f)) = 2e-4 for this H (beta sum
    f (1 - f) = 130)."""
    print(f"[bf16 implicit fermi] method={method} tol={tol}", flush=True)
    h = O.model_hamiltonian(256, 0)
    args = (8, 2, 128.0, 0.6, 20.0, method, 3, 1e-9, tol)
    p_o, mu_o = O.implicit_fermi(h, *args)
    P.set_product_precision(P.PREC_BF16SPLIT)
    try:
        P.gemm_counters(reset=True)
        p, mu = P.implicit_fermi(h, *args)
        nprod, _, exe, int8_ops = P.gemm_counters(reset=True, int8=True)
    finally:
        P.set_product_precision(P.PREC_FP64)
    print(f"  products={nprod:.0f} |dP|/|P|={np.linalg.norm(p - p_o) / np.linalg.norm(p_o):.3e} "
          f"dmu={abs(mu - mu_o):.3e}", flush=True)
    assert nprod > 0 and exe == 0.0 and int8_ops == 0.0
    assert np.array_equal(p, p.T)
    assert abs(mu - mu_o) < 2e-4
    assert np.linalg.norm(p - p_o) < 1e-3 * np.linalg.norm(p_o)

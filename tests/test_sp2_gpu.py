"""GPU parity of the SP2 family (HIP path through the C-ABI) against the oracle.

Bar: the same branch sequence and iteration count as the oracle, and the FP64
density within 1e-9 (Frobenius) of the oracle's on the same input; the
reference's own tolerance tests (idempotency of rho/2) on its shipped
fixtures.
"""
import numpy as np
import pytest

import progress_oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def P():
    import progress_amd
    progress_amd.set_precision(progress_amd.PREC_FP64)
    return progress_amd


@pytest.mark.parametrize("n,seed", [(256, 0), (300, 1), (1024, 0)])
def test_sp2_alg2_matches_oracle(P, n, seed):
    print(f"[sp2_alg2] n={n} seed={seed}", flush=True)
    h = O.model_hamiltonian(n, seed)
    rho_o, it_o, pp_o = O.sp2_alg2(h, 0.5, 10, 200, "Abs", 1e-10)
    rho_g = P.sp2_alg2(h, 0.5, 10, 200, "Abs", 1e-10)
    assert np.linalg.norm(rho_g - rho_o) < 1e-9 * max(1.0, n / 256)
    d0, _ = O.build_density_T0(h, 0.5)
    assert np.linalg.norm(rho_g - d0) < 1e-8 * max(1.0, n / 256)
    assert np.array_equal(rho_g, rho_g.T)


@pytest.mark.parametrize("n", [200, 256])
def test_genseq_sequence_and_vv(P, n):
    print(f"[genseq] n={n}", flush=True)
    h = O.model_hamiltonian(n, 2)
    rho_o, pp_o, ic_o, vv_o = O.sp2_alg2_genseq(h, 0.5, 10, 200, "Abs", 1e-10)
    rho_g, pp_g, ic_g, vv_g = P.sp2_alg2_genseq(h, 0.5, 10, 200, "Abs", 1e-10)
    assert ic_g == ic_o and pp_g == pp_o
    assert np.allclose(vv_g[:len(vv_o)], vv_o, rtol=1e-9, atol=1e-11)
    assert np.linalg.norm(rho_g - rho_o) < 1e-9
    rho1_o, pp1_o, ic1_o, vv1_o = O.sp2_alg1_genseq(h, 0.5, 10, 200, "Abs", 1e-10)
    rho1_g, pp1_g, ic1_g, vv1_g = P.sp2_alg1_genseq(h, 0.5, 10, 200, "Abs", 1e-10)
    assert ic1_g == ic1_o and pp1_g == pp1_o
    assert np.allclose(vv1_g[:ic1_o], vv1_o[:ic1_o], rtol=1e-9, atol=1e-11)
    assert np.linalg.norm(rho1_g - rho1_o) < 1e-9


def test_alg1_basic_match_oracle(P):
    print("[alg1/basic] n=256", flush=True)
    h = O.model_hamiltonian(256, 0)
    r1o = O.sp2_alg1(h, 0.5, 10, 200, "Abs", 1e-10)[0]
    r1g = P.sp2_alg1(h, 0.5, 10, 200, "Abs", 1e-10)
    assert np.linalg.norm(r1g - r1o) < 1e-9
    rbo = O.sp2_basic(h, 0.5, 10, 200, "Abs", 1e-10)[0]
    rbg = P.sp2_basic(h, 0.5, 10, 200, "Abs", 1e-10)
    assert np.linalg.norm(rbg - rbo) < 1e-9


def test_rel_convergence_matches_oracle(P, golden):
    print("[rel] ortho384", flush=True)
    h = O.read_mtx(str(golden / "hamiltonian_ortho_384.mtx"))
    h = 0.5 * (h + h.T)
    for fo, fg in [(O.sp2_alg2, P.sp2_alg2), (O.sp2_alg1, P.sp2_alg1)]:
        ro = fo(h, 0.5, 10, 100, "Rel", 1e-10)[0]
        rg = fg(h, 0.5, 10, 100, "Rel", 1e-10)
        assert np.linalg.norm(rg - ro) < 1e-9
        assert O.check_idempotency(0.5 * rg) < 1e-8


def test_replay_and_inplace_variants(P):
    print("[replay] n=192", flush=True)
    h = O.model_hamiltonian(192, 5)
    lo, hi = O.gershgorin(h)
    glo, ghi = P.gershgorin(h)
    assert (glo, ghi) == pytest.approx((lo, hi), rel=1e-14, abs=1e-14)
    rho, pp, ic, vv = O.sp2_alg2_genseq(h, 0.5, 10, 200, "Abs", 1e-10)
    for fo, fg in [(O.sp2_alg2_seq, P.sp2_alg2_seq), (O.sp2_alg1_seq, P.sp2_alg1_seq)]:
        ro, vo = fo(h, pp, ic)
        rg, vg = fg(h, pp, ic)
        assert np.linalg.norm(rg - ro) < 1e-9
        assert np.allclose(vg, vo, rtol=1e-9, atol=1e-11)
    for fo, fg in [(O.sp2_alg2_seq_inplace, P.sp2_alg2_seq_inplace),
                   (O.sp2_alg1_seq_inplace, P.sp2_alg1_seq_inplace)]:
        ro, vo = fo(h, pp, ic, lo, hi)
        rg, vg = fg(h, pp, ic, lo, hi)
        assert np.linalg.norm(rg - ro) < 1e-9
        assert np.allclose(vg, vo, rtol=1e-9, atol=1e-11)


@pytest.mark.parametrize("core", [0, 37, 192])
def test_submatrix_core_rows(P, core):
    print(f"[submatrix] core={core}", flush=True)
    h = O.model_hamiltonian(192, 6)
    lo, hi = O.gershgorin(h)
    rho, pp, ic, vv = O.sp2_alg2_genseq(h, 0.5, 10, 200, "Abs", 1e-10)
    vv0 = [0.25] * ic  # accumulation onto existing values
    ro, vo = O.sp2_submatrix(h, pp, ic, vv0, lo, hi, core)
    rg, vg = P.sp2_submatrix(h, pp, ic, vv0, lo, hi, core)
    assert np.linalg.norm(rg - ro) < 1e-9
    assert np.allclose(vg, vo, rtol=1e-9, atol=1e-12)


def test_reference_poly512_idempotency(P, golden):
    """tests/src/main.F90:336-394: poly.512.mtx, bndfil 0.5, Rel, minsp2iter 25,
    idempotency of rho/2 <= 1e-6 for alg1 and alg2."""
    print("[poly512] n=6144", flush=True)
    h = O.read_mtx(str(golden / "poly.512.mtx.gz"))
    for fg in (P.sp2_alg2, P.sp2_alg1):
        rho = fg(h, 0.5, 25, 100, "Rel", 1e-10)
        x = 0.5 * rho
        assert np.linalg.norm(x @ x - x) < 1e-6
        assert abs(np.trace(rho) - 6144) < 1e-6


def test_reference_poly512_ellpack_thresholds(P, golden):
    """tests/src/main.F90:336-424, the ellpack cases: poly.512.mtx in
    ellpack handles, threshold 1e-9 (alg1 and alg2, idempotency of rho/2 <=
    1e-6) and 1e-5 (alg2, idempotency <= 1e-2).  The library applies the
    sparse formats' thresholding (|a| <= threshold dropped from every product
    and sum); at 1e-5 the density gains exact zeros the dense solve does not
    have."""
    print("[poly512 ellpack] n=6144", flush=True)
    h = O.read_mtx(str(golden / "poly.512.mtx.gz"))
    dense = P.sp2_alg2(h, 0.5, 25, 100, "Rel", 1e-10)
    for fg, thr, bar in ((P.sp2_alg1, 1e-9, 1e-6), (P.sp2_alg2, 1e-9, 1e-6),
                         (P.sp2_alg2, 1e-5, 1e-2)):
        rho = fg(h, 0.5, 25, 100, "Rel", 1e-10, threshold=thr, matrix_type=P.ELLPACK)
        x = 0.5 * rho
        idem = np.linalg.norm(x @ x - x)
        print(f"  threshold {thr:.0e}: idempotency {idem:.2e}, zeros {np.sum(rho == 0)} "
              f"(dense {np.sum(dense == 0)})", flush=True)
        assert idem <= bar
        if thr == 1e-5:
            assert np.sum(rho == 0) > np.sum(dense == 0)


def test_ellpack_threshold_matches_oracle(P, golden):
    """prg_sp2_alg2 in ellpack handles, threshold 1e-5, on the reference's
    384-orbital Hamiltonian, against the oracle's thresholded alg2 (product
    dropped before 2X - X^2, the sum dropped again): the same iteration
    count and branch sequence, the density to 1e-9.  A dense handle ignores
    the threshold (BML dense does): bit-equal to threshold 0."""
    h = O.read_mtx(str(golden / "hamiltonian_ortho_384.mtx"))
    rho = P.sp2_alg2(h, 0.5, 10, 100, "Abs", 1e-10, threshold=1e-5, matrix_type=P.ELLPACK)
    ro, it, pp = O.sp2_alg2(h, 0.5, 10, 100, "Abs", 1e-10, threshold=1e-5)
    r0, _, _ = O.sp2_alg2(h, 0.5, 10, 100, "Abs", 1e-10)
    print(f"[ellpack 384] |rho - oracle| {np.linalg.norm(rho - ro):.2e}, threshold moves the "
          f"density by {np.linalg.norm(ro - r0):.2e}", flush=True)
    assert np.linalg.norm(ro - r0) > 1e-6
    assert np.linalg.norm(rho - ro) < 1e-9
    d1 = P.sp2_alg2(h, 0.5, 10, 100, "Abs", 1e-10, threshold=1e-5)
    d0 = P.sp2_alg2(h, 0.5, 10, 100, "Abs", 1e-10)
    assert np.array_equal(d1, d0)


def test_device_pointer_entry_point(P):
    import torch
    print("[device api] n=512", flush=True)
    n = 512
    h = O.model_hamiltonian(n, 0)
    ld = P.padded_ld(n)
    hd = torch.zeros((ld, ld), dtype=torch.float64, device="cuda")
    hd[:n, :n] = torch.from_numpy(h)
    rd = torch.zeros_like(hd)
    it, pp = P.sp2_alg2_device(hd, rd, n, 0.5, 10, 200, "Abs", 1e-10)
    rho_o, it_o, pp_o = O.sp2_alg2(h, 0.5, 10, 200, "Abs", 1e-10)
    assert it == it_o and pp == pp_o
    r = rd.cpu().numpy()
    assert np.linalg.norm(r[:n, :n] - rho_o) < 1e-9
    assert np.all(r[n:, :] == 0) and np.all(r[:, n:] == 0)


def test_bad_inputs_are_errors(P):
    print("[errors]", flush=True)
    with pytest.raises(P.ProgressError):
        P.sp2_alg2(np.zeros((4, 4)), 0.5, 1, 10, "Abs", 1e-10)  # zero-width interval
    with pytest.raises(P.ProgressError):
        # maxsp2iter too small: the reference calls stop here
        P.sp2_alg2(O.model_hamiltonian(128, 0), 0.5, 1, 5, "Abs", 1e-10)


@pytest.mark.parametrize("n", [256, 1024])
def test_f16split_mode_tracks_fp64(P, n):
    """FP16 hi+lo split products with FP32 accumulation: same iteration count
    (+-2) and an FP32-class density (||D - D_oracle||_F within 1e-3 at these
    sizes; the measured value is reported by bench.py at N=8192)."""
    print(f"[f16split] n={n}", flush=True)
    h = O.model_hamiltonian(n, 0)
    rho_o, it_o, pp_o = O.sp2_alg2(h, 0.5, 10, 200, "Abs", 1e-10)
    P.set_precision(P.PREC_F16SPLIT)
    try:
        rho_g = P.sp2_alg2(h, 0.5, 10, 200, "Abs", 1e-10)
    finally:
        P.set_precision(P.PREC_FP64)
    err = np.linalg.norm(rho_g - rho_o)
    print(f"  ||D_f16split - D_oracle||_F = {err:.3e}")
    assert err < 1e-3
    assert abs(np.trace(rho_g) - n) < 1e-3
    assert np.linalg.norm(rho_g @ rho_g / 2 - rho_g) < 1e-3


@pytest.mark.parametrize("n", [256, 1024])
def test_bf16split_mode_tracks_fp64(P, n):
    """BF16 hi+lo split products (about 16 significant bits) with FP32
    accumulation, the same noise-floor stop and FP64 polishing as the FP16
    split: the density within 2e-2 (Frobenius) of the oracle's (measured
    4.2e-3 at n = 256, 9.2e-3 at 1024), trace and idempotency within 1e-3
    after the polishing steps."""
    print(f"[bf16split] n={n}", flush=True)
    h = O.model_hamiltonian(n, 0)
    rho_o, it_o, pp_o = O.sp2_alg2(h, 0.5, 10, 200, "Abs", 1e-10)
    P.set_precision(P.PREC_BF16SPLIT)
    try:
        rho_g = P.sp2_alg2(h, 0.5, 10, 200, "Abs", 1e-10)
    finally:
        P.set_precision(P.PREC_FP64)
    err = np.linalg.norm(rho_g - rho_o)
    print(f"  ||D_bf16split - D_oracle||_F = {err:.3e}", flush=True)
    assert err < 2e-2
    assert abs(np.trace(rho_g) - n) < 1e-3
    assert np.linalg.norm(rho_g @ rho_g / 2 - rho_g) < 1e-3
    assert np.array_equal(rho_g, rho_g.T)


def test_sp2_basic_tcore_matches_oracle():
    """prg_sp2_basic_tcore: maxit + 1 iterations of the basic rule, no early
    exit; the density matches the oracle (and prg_sp2_basic once converged)."""
    import progress_amd as P
    P.set_precision(P.PREC_FP64)
    n = 300
    print(f"[sp2_basic_tcore] n={n}", flush=True)
    h = O.model_hamiltonian(n, 4)
    rho_o, _ = O.sp2_basic_tcore(h, 0.5, 80)
    rho_full = O.sp2_basic(h, 0.5, 10, 200, "Abs", 1e-10)[0]
    rho_g = P.sp2_basic_tcore(h, 0.5, 10, 80, "Abs", 1e-10, rhofull=rho_full, verbose=2)
    # after convergence (about iteration 30) the sign of Tr X - occ is rounding
    # noise, so each side keeps taking noise-chosen branches on an idempotent X
    # and wanders by ~1e-9 in trace: 5e-8 bar (the converged densities agree
    # to 1e-12 otherwise)
    assert np.linalg.norm(rho_g - rho_o) < 5e-8
    assert np.linalg.norm(rho_g - rho_full) < 5e-8
    assert abs(np.trace(rho_g) - n) < 1e-7


def test_sp2_tensorcore_entry():
    """prg_sp2_tensorcore: float32 H in, D = X out (not 2X), FP16 split
    products with the noise-floor stop and FP64 refinement.  Bars (the split
    products' accuracy on this near-gapless model input, as for the f16split
    mode of prg_sp2_alg2): D within 5e-3 (Frobenius) of the FP64 oracle's X
    and Tr D within 1e-3 of int(bndfil N); D symmetric."""
    import progress_amd as P
    n = 512
    print(f"[sp2_tensorcore] n={n}", flush=True)
    h = O.model_hamiltonian(n, 2)
    d = P.sp2_tensorcore(h, 0.5, 1, 100)
    x_o = 0.5 * O.build_density_T0(h.astype(np.float32).astype(np.float64), 0.5)[0]
    err, tr = float(np.linalg.norm(d - x_o)), abs(float(np.trace(d)) - n // 2)
    print(f"  ||D - X_oracle||_F = {err:.3e}, |Tr D - Nocc| = {tr:.3e}", flush=True)
    assert err < 5e-3
    assert tr < 1e-3
    assert np.array_equal(d, d.T)


@pytest.mark.parametrize("prec,moduli", [("int8crt", 10), ("int8crt", 11), ("f16split", 0),
                                         ("bf16split", 0)])
@pytest.mark.parametrize("core", [37, 120])
def test_submatrix_core_rows_low_precision(P, prec, moduli, core):
    """prg_sp2_submatrix under the other product precisions: vv accumulates the
    core rows only, as with FP64 products (src/prg_sp2_mod.F90:1224).  Bars:
    int8 CRT (X rounded to 2^-39 / 2^-43) density within 1e-7 and vv within
    1e-6 relative; the FP32-accumulated split products track the oracle to
    their own accuracy (vv within 3 % of its largest entry for FP16, 20 % for
    BF16) -- both far from the whole-matrix sums, which are about n / core
    times larger."""
    print(f"[submatrix low precision] {prec} moduli={moduli} core={core}", flush=True)
    n = 192
    h = O.model_hamiltonian(n, 6)
    lo, hi = O.gershgorin(h)
    _, pp, ic, _ = O.sp2_alg2_genseq(h, 0.5, 10, 200, "Abs", 1e-10)
    vv0 = [0.0] * ic
    ro, vo = O.sp2_submatrix(h, pp, ic, vv0, lo, hi, core)
    _, vwhole = O.sp2_submatrix(h, pp, ic, vv0, lo, hi, n)
    mode = {"int8crt": P.PREC_INT8CRT, "f16split": P.PREC_F16SPLIT,
            "bf16split": P.PREC_BF16SPLIT}[prec]
    P.set_precision(mode)
    if moduli:
        P.set_int8_moduli(moduli)
    try:
        rg, vg = P.sp2_submatrix(h, pp, ic, vv0, lo, hi, core)
    finally:
        P.set_int8_moduli(11)
        P.set_precision(P.PREC_FP64)
    vo, vg, vwhole = np.asarray(vo), np.asarray(vg[:ic]), np.asarray(vwhole)
    if prec == "int8crt":
        assert np.linalg.norm(rg - ro) < 1e-7
        assert np.allclose(vg, vo, rtol=1e-6, atol=1e-12)
    else:
        tol = (0.03 if prec == "f16split" else 0.2) * vo.max()
        assert np.max(np.abs(vg - vo)) < tol
    assert np.max(np.abs(vg - vwhole)) > 0.5 * vo.max()


def test_homolumogap_and_sequence_from_gpu_genseq(P):
    """prg_homolumogap and prg_sp2sequence fed by the GPU's genseq, as
    examples/gptest/gptest.F90:57-115 chains them: the same gap estimate and
    the same generated sequence as the oracle's routines on the oracle's
    genseq (pp identical, vv within 1e-9 relative), and the estimate brackets
    the true HOMO/LUMO of an eigendecomposition to 1e-3 of the interval."""
    print("[homolumogap/sp2sequence from GPU genseq] n=256", flush=True)
    n = 256
    h = O.model_hamiltonian(n, 0)
    lo, hi = P.gershgorin(h)
    lo_o, hi_o = O.gershgorin(h)
    assert abs(lo - lo_o) < 1e-12 and abs(hi - hi_o) < 1e-12
    _, pp_g, ic_g, vv_g = P.sp2_alg2_genseq(h, 0.5, 10, 200, "Abs", 1e-10)
    _, pp_o, ic_o, vv_o = O.sp2_alg2_genseq(h, 0.5, 10, 200, "Abs", 1e-10)
    assert ic_g == ic_o and list(pp_g) == list(pp_o)
    ours = P.homolumogap(vv_g, ic_g, pp_g, lo, hi)
    ref = O.homolumogap(vv_o, ic_o, pp_o, lo_o, hi_o)
    assert np.allclose(ours, ref, rtol=1e-9, atol=1e-12)
    w = np.linalg.eigvalsh(h)
    assert abs(ours[0] - w[n // 2 - 1]) < 1e-3 * (hi - lo)
    assert abs(ours[1] - w[n // 2]) < 1e-3 * (hi - lo)
    pps, imax = P.sp2sequence(lo, hi, ours[0], ours[1], 1e-12)
    ppo, imo = O.sp2sequence(lo_o, hi_o, ref[0], ref[1], 1e-12)
    assert imax == imo and list(pps) == list(ppo)

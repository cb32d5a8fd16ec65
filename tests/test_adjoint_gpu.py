"""Adjoint of the SP2 recursion (dL/dH for L = Tr(D A)) on the GPU.

Checks: (1) against the first-order response prg_compute_response_SP2 with
H1 = A (the Frechet derivative of a matrix function of a symmetric matrix is
self-adjoint); (2) against central finite differences in FP64 with the branch
sequence and Gershgorin bounds frozen (<= 1e-5 relative); (3) checkpoint
spacing does not change the result (bitwise).
"""
import numpy as np
import pytest

import progress_oracle as O

pytestmark = pytest.mark.gpu


def test_sp2_adjoint_vs_response_and_fd():
    import torch
    import progress_amd as P
    from progress_amd.autograd import sp2_density
    P.set_precision(P.PREC_FP64)
    n = 256
    print(f"[adjoint] n={n}", flush=True)
    h = O.model_hamiltonian(n, 1)
    a = O.banded_perturbation(n, 32, seed=2)
    ht = torch.tensor(h, device="cuda", requires_grad=True)
    at = torch.tensor(a, device="cuda")
    d = sp2_density(ht, 0.5, 1e-10, 8)
    loss = torch.sum(d * at)
    loss.backward()
    g = ht.grad.detach().cpu().numpy()
    # (1) response with H1 = A
    rho, rho1 = P.compute_response_sp2(h, a, 0.5, 1, 500, "Abs", 1e-10)
    assert np.linalg.norm(g - rho1) <= 1e-9 * np.linalg.norm(rho1)
    # (2) central finite differences, frozen sequence and bounds
    ld = P.padded_ld(n)
    hp = torch.zeros((ld, ld), dtype=torch.float64, device="cuda")
    hp[:n, :n] = torch.from_numpy(h)
    rd = torch.zeros_like(hp)
    lo, hi = P.gershgorin_device(hp, n)
    it, pp = P.sp2_alg2_device(hp, rd, n, 0.5, 1, 500, "Abs", 1e-10)
    rng = np.random.default_rng(0)
    step = 1e-5
    errs = []
    for _ in range(8):
        i, j = rng.integers(0, n, 2)
        e = np.zeros((n, n))
        e[i, j] += 1.0
        e[j, i] += 1.0
        vals = []
        for s in (+step, -step):
            hp[:n, :n] = torch.from_numpy(h + s * e)
            P.sp2_seq_device(hp, rd, n, pp, lo, hi)
            vals.append(float(torch.sum(rd[:n, :n] * at)))
        fd = (vals[0] - vals[1]) / (2 * step)
        an = g[i, j] + g[j, i] if i != j else g[i, i] * 2
        errs.append(abs(fd - an) / max(abs(an), 1e-3))
    assert max(errs) <= 1e-5, errs
    # (3) checkpoint spacing
    go = torch.tensor(a, device="cuda")
    gp = torch.zeros_like(hp)
    gp[:n, :n] = go
    hp[:n, :n] = torch.from_numpy(h)
    outs = []
    for ck in (1, 8, 500):
        out = torch.zeros_like(hp)
        P.sp2_adjoint_device(hp, gp, out, n, pp, lo, hi, ck)
        outs.append(out.cpu().numpy())
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[1], outs[2])


def test_sp2_adjoint_nonsymmetric_seed():
    """A non-symmetric dL/dD takes full products (the symmetric seed takes
    upper-triangle tiles only).  The adjoint map is linear and commutes with
    transposition (S(G^T) = S(G)^T for symmetric X_k), so for any B:
    adj(B) + adj(B)^T = adj(B + B^T), the right side on the symmetric path."""
    import torch
    import progress_amd as P
    n = 320
    print(f"[adjoint non-symmetric seed] n={n}", flush=True)
    h = O.model_hamiltonian(n, 4)
    ld = P.padded_ld(n)
    hp = torch.zeros((ld, ld), dtype=torch.float64, device="cuda")
    hp[:n, :n] = torch.from_numpy(h)
    rd = torch.zeros_like(hp)
    lo, hi = P.gershgorin_device(hp, n)
    it, pp = P.sp2_alg2_device(hp, rd, n, 0.5, 1, 500, "Abs", 1e-10)
    b = np.random.default_rng(5).standard_normal((n, n))
    outs = []
    for seed in (b, b + b.T):
        gp = torch.zeros_like(hp)
        gp[:n, :n] = torch.from_numpy(seed)
        out = torch.zeros_like(hp)
        P.sp2_adjoint_device(hp, gp, out, n, pp, lo, hi, 8)
        outs.append(out[:n, :n].cpu().numpy())
    g, gs = outs
    assert np.linalg.norm(g - g.T) > 1e-3 * np.linalg.norm(g)  # genuinely non-symmetric
    assert np.array_equal(gs, gs.T)                            # symmetric path is exact
    assert np.linalg.norm(g + g.T - gs) <= 1e-12 * np.linalg.norm(gs)


def test_sp2_adjoint_c4_configuration():
    """The differentiation workload as specified: N = 2048 model Hamiltonian
    (seed 0), A symmetric banded (bandwidth 32, N(0,1) entries), L = Tr(D A),
    checkpoints every 8 steps; dL/dH against central finite differences with
    step 1e-4 on 64 random (i, j) pairs (h_ij and h_ji perturbed together,
    compared with G_ij + G_ji, sequence and Gershgorin bounds frozen, FP64),
    and against the first-order response prg_compute_response_SP2 with H1 = A.
    Bars: response 1e-9 relative (Frobenius); central finite differences 1e-5
    relative per pair (measured: 6e-7 at most, median 3e-8).  A Richardson
    step (1e-4 and 5e-5) is reported beside them: it cancels the O(h^2) term
    but doubles the rounding noise, which dominates at these steps."""
    import torch
    import progress_amd as P
    from progress_amd.autograd import sp2_density
    P.set_precision(P.PREC_FP64)
    n = 2048
    print(f"[adjoint C4] n={n}", flush=True)
    h = O.model_hamiltonian(n, 0)
    a = O.banded_perturbation(n, 32, seed=1)
    ht = torch.tensor(h, device="cuda", requires_grad=True)
    at = torch.tensor(a, device="cuda")
    d = sp2_density(ht, 0.5, 1e-10, 8)
    torch.sum(d * at).backward()
    g = ht.grad.detach().cpu().numpy()
    _, rho1 = P.compute_response_sp2(h, a, 0.5, 1, 500, "Abs", 1e-10)
    assert np.linalg.norm(g - rho1) <= 1e-9 * np.linalg.norm(rho1)
    ld = P.padded_ld(n)
    hp = torch.zeros((ld, ld), dtype=torch.float64, device="cuda")
    hp[:n, :n] = torch.from_numpy(h)
    rd = torch.zeros_like(hp)
    lo, hi = P.gershgorin_device(hp, n)
    it, pp = P.sp2_alg2_device(hp, rd, n, 0.5, 1, 500, "Abs", 1e-10)
    rng = np.random.default_rng(64)

    def loss_at(i, j, s):
        hp[i, j] += s
        if i != j:
            hp[j, i] += s
        P.sp2_seq_device(hp, rd, n, pp, lo, hi)
        v = float(torch.sum(rd[:n, :n] * at))
        hp[:n, :n] = torch.from_numpy(h)
        return v

    plain, rich = [], []
    for _ in range(64):
        i, j = (int(v) for v in rng.integers(0, n, 2))
        an = g[i, j] + g[j, i] if i != j else 2.0 * g[i, i]
        d1 = (loss_at(i, j, 1e-4) - loss_at(i, j, -1e-4)) / 2e-4
        d2 = (loss_at(i, j, 5e-5) - loss_at(i, j, -5e-5)) / 1e-4
        scale = max(abs(an), 1e-3)
        plain.append(abs(d1 - an) / scale)
        rich.append(abs((4.0 * d2 - d1) / 3.0 - an) / scale)
    print(f"  FD rel. error: central max {max(plain):.2e} median {np.median(plain):.2e}; "
          f"Richardson max {max(rich):.2e}", flush=True)
    assert max(plain) <= 1e-5


@pytest.mark.parametrize("canonical", [False, True])
def test_implicit_fermi_adjoint_vs_oracle_and_fd(canonical):
    """Adjoint of the fixed-mu implicit-Fermi map (8 levels, beta 20, Newton-
    Schulz inverses to 1e-10) through the torch.autograd binding: (1) the
    oracle's reverse sweep with exact inverses, itself pinned by the
    Daleckii-Krein closed form, to 1e-10 relative (measured 3.5e-14); (2) at
    fixed mu, central finite differences in FP64 (step 1e-4) along A, along
    the gradient and along a random direction, to 1e-5 of the gradient norm
    (measured 1e-8); (3) the forward P equals the oracle's exact-inverse
    recursion to 1e-11."""
    import torch
    import progress_amd as P
    from progress_amd.autograd import implicit_fermi_density
    n, nsteps, mu, beta = 256, 8, 0.6, 20.0
    print(f"[implicit fermi adjoint] n={n} canonical={canonical}", flush=True)
    assert P.get_product_precision() == P.PREC_FP64
    h = O.model_hamiltonian(n, 0)
    a = O.banded_perturbation(n, 32, seed=3)
    ht = torch.tensor(h, device="cuda", requires_grad=True)
    at = torch.tensor(a, device="cuda")
    p = implicit_fermi_density(ht, nsteps, mu, beta, 1e-10, canonical)
    torch.sum(p * at).backward()
    g = ht.grad.detach().cpu().numpy()
    assert np.array_equal(g, g.T)
    g_o = O.implicit_fermi_adjoint(h, a, nsteps, mu, beta, canonical)
    assert np.linalg.norm(g - g_o) <= 1e-10 * np.linalg.norm(g_o)
    cnst = beta / 2.0 ** (nsteps + 2)
    x = O.normalize_implicit_fermi(h, cnst, mu)
    for _ in range(nsteps):
        x2 = x @ x
        x = np.linalg.inv(2.0 * (x2 - x) + np.eye(n)) @ x2
    assert np.linalg.norm(p.detach().cpu().numpy() - x) < 1e-11
    if canonical:
        return
    ld = P.padded_ld(n)

    def p_of(hm):
        hp = torch.zeros((ld, ld), dtype=torch.float64, device="cuda")
        hp[:n, :n] = torch.from_numpy(hm)
        pp = torch.zeros_like(hp)
        P.implicit_fermi_device(hp, pp, n, nsteps, mu, beta, 1e-10)
        return pp[:n, :n].cpu().numpy()

    rng = np.random.default_rng(5)
    r = rng.standard_normal((n, n))
    gn = np.linalg.norm(g)
    eps = 1e-4
    for e in (a, g, 0.5 * (r + r.T)):
        e = e / np.linalg.norm(e)
        fd = (np.sum(p_of(h + eps * e) * a) - np.sum(p_of(h - eps * e) * a)) / (2.0 * eps)
        assert abs(fd - np.sum(g * e)) <= 1e-5 * gn


def test_sp2_adjoint_fp32_input_c4():
    """Config C4 names an FP32 Hamiltonian: H given as float32 (N = 2048,
    seed 0; A banded, bandwidth 32) is promoted exactly to FP64, and D and
    dL/dH come back as float32.  They equal the FP64 path on the promoted
    input rounded to float32 (relative 1e-6, the float32 rounding), and the
    gradient is symmetric."""
    import torch
    import progress_amd as P
    from progress_amd.autograd import sp2_density
    P.set_precision(P.PREC_FP64)
    n = 2048
    print(f"[adjoint C4 fp32 input] n={n}", flush=True)
    h32 = torch.tensor(O.model_hamiltonian(n, 0), dtype=torch.float32, device="cuda")
    a = torch.tensor(O.banded_perturbation(n, 32, seed=1), device="cuda")
    x32 = h32.clone().requires_grad_(True)
    d32 = sp2_density(x32, 0.5, 1e-10, 8)
    assert d32.dtype == torch.float32
    torch.sum(d32.to(torch.float64) * a).backward()
    g32 = x32.grad
    assert g32.dtype == torch.float32
    x64 = h32.to(torch.float64).clone().requires_grad_(True)
    d64 = sp2_density(x64, 0.5, 1e-10, 8)
    torch.sum(d64 * a).backward()
    g64 = x64.grad
    assert torch.linalg.norm(d32.double() - d64) <= 1e-6 * torch.linalg.norm(d64)
    assert torch.linalg.norm(g32.double() - g64) <= 1e-6 * torch.linalg.norm(g64)
    assert torch.linalg.norm(g32 - g32.T) <= 1e-6 * torch.linalg.norm(g32)


def test_sp2_adjoint_matches_oracle_response():
    """dTr(D A)/dH of the SP2 recursion from the autograd binding against the
    oracle's prg_compute_response_SP2 with H1 = A (the adjoint of the
    self-adjoint first-order response), FP64, N = 256: 1e-9 relative."""
    import torch
    import progress_amd as P
    from progress_amd.autograd import sp2_density
    P.set_precision(P.PREC_FP64)
    n = 256
    print(f"[adjoint vs oracle response] n={n}", flush=True)
    h = O.model_hamiltonian(n, 2)
    a = O.banded_perturbation(n, 32, seed=5)
    ht = torch.tensor(h, device="cuda", requires_grad=True)
    d = sp2_density(ht, 0.5, 1e-10, 8)
    torch.sum(d * torch.tensor(a, device="cuda")).backward()
    g = ht.grad.detach().cpu().numpy()
    _, rho1, _, _ = O.compute_response_sp2(h, a, 0.5, 1, 500, "Abs", 1e-10)
    assert np.linalg.norm(g - rho1) <= 1e-9 * np.linalg.norm(rho1)

"""CPU oracle for the PROGRESS density-matrix solvers (TEST INFRASTRUCTURE ONLY).

This module is a plain numpy FP64 restatement of the reference algorithms.  It is
the oracle the HIP library is tested against, never part of the product: only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it.  Every function cites the reference routine it restates
(paths relative to the lanl/qmd-progress tree).

Matrices are dense, symmetric FP64 ``numpy.ndarray`` objects.  BML (the
reference's matrix library) is absent; the BML calls used by the reference are
restated with the dense semantics inferred from their call sites:

* ``bml_multiply_x2(X, X2, thr, trace)`` -> ``X2 = X @ X``, ``trace = (Tr X, Tr X2)``
* ``bml_multiply(A, B, C, a, b)``        -> ``C = a A @ B + b C``
* ``bml_add_deprecated(a, A, b, B)``     -> ``A = a A + b B``
* ``bml_add(A, B, a, b)``                -> ``A = a A + b B``
* ``bml_scale_add_identity(A, a, b)``    -> ``A = a A + b I``
* ``bml_gershgorin``                     -> textbook Gershgorin interval
* ``bml_normalize(A, lo, hi)``           -> ``A = (hi I - A) / (hi - lo)``
* thresholds are ignored (dense matrices).

Parity pins (checked in ``tests/test_oracle.py``):

* SP2 family, normalisation, sequence/gap helpers: the SP2 iteration is a
  polynomial in H, so running the same control flow on the eigenvalue vector
  (``sp2_*_eigen``) must reproduce the matrix oracle's branch sequence and
  ``Q p(x) Q^T``; the converged result must equal the T=0 eigendecomposition
  density (``build_density_T0``); the reference's own tolerance tests on the
  shipped ``poly.512.mtx`` fixture (idempotency <= 1e-6) and the partition
  KATs (72/6 -> 12 parts, 72/7 -> 11, ``test.part`` -> 104 parts).
* implicit Fermi: with exact inverses the recursion is the scalar map
  x <- x^2 / (2x^2 - 2x + 1) on the eigenvalues, whose limit is the Fermi
  function 1/(1+exp(beta (e-mu))) as the number of steps grows.
* responses: Daleckii-Krein divided differences of the same scalar maps and
  central finite differences of the forward solvers; the order-1..3 implicit
  Fermi responses as the Taylor coefficients of the scalar map applied to the
  eigenvalues of H0 + l H1 + l^2 H2 + l^3 H3 (a polynomial fit in l).
* int8 Chinese-remainder products (``crt_square``, ``crt_product``; no
  reference counterpart): against the FP64 products within the operand
  rounding bound, the scale bound at its edge, and planted mistakes (wrong CRT
  weights, a dropped modulus, an operand scaled by 1) that must break them.

Functions whose values only BML could pin (``matrix2submatrix_index`` halo
ordering) are "parity unpinned"; the subgraph loop's result is pinned through
the block-diagonal special case (parts independent) and the full-halo limit.
"""
from __future__ import annotations

import math
import numpy as np

# --------------------------------------------------------------------------
# inputs
# --------------------------------------------------------------------------


def model_hamiltonian(n: int, seed: int = 0) -> np.ndarray:
    """Synthetic dense tight-binding Hamiltonian used by the benchmark.

    Off-diagonal h_ij = -exp(-|i-j|/8) (1 + 0.1 z_ij), z ~ N(0,1) drawn for the
    full n x n array first, zeroed for |i-j| > 64, symmetrised as
    (off + off^T)/2; diagonal ~ U[-1, 1] drawn second.  numpy PCG64 via
    ``default_rng(seed)``.
    """
    rng = np.random.default_rng(seed)
    z = rng.standard_normal((n, n))
    idx = np.arange(n)
    dist = np.abs(idx[:, None] - idx[None, :])
    off = -np.exp(-dist / 8.0) * (1.0 + 0.1 * z)
    off[dist > 64] = 0.0
    np.fill_diagonal(off, 0.0)
    h = 0.5 * (off + off.T)
    h[idx, idx] = rng.uniform(-1.0, 1.0, n)
    return h


def banded_perturbation(n: int, bandwidth: int = 32, seed: int = 1) -> np.ndarray:
    """Random symmetric banded matrix (|i-j| <= bandwidth, N(0,1) entries)."""
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((n, n))
    idx = np.arange(n)
    a[np.abs(idx[:, None] - idx[None, :]) > bandwidth] = 0.0
    return np.triu(a) + np.triu(a, 1).T


def read_mtx(path: str) -> np.ndarray:
    """Matrix Market ``coordinate real general`` reader (1-based indices).

    Accepts the ``%%%MatrixMarket`` header variant found in
    tests/tests_data/hamiltonian-twolevel-ref.mtx.  Missing entries are zero.
    """
    import gzip
    opener = gzip.open if path.endswith(".gz") else open
    with opener(path, "rt") as fh:
        header = fh.readline()
        if "MatrixMarket" not in header:
            raise ValueError(f"{path}: not a Matrix Market file")
        line = fh.readline()
        while line.startswith("%"):
            line = fh.readline()
        nr, nc, nnz = (int(v) for v in line.split())
        data = np.loadtxt(fh, ndmin=2)
    if data.shape[0] != nnz:
        raise ValueError(f"{path}: expected {nnz} entries, got {data.shape[0]}")
    a = np.zeros((nr, nc))
    a[data[:, 0].astype(int) - 1, data[:, 1].astype(int) - 1] = data[:, 2]
    return a


# --------------------------------------------------------------------------
# normalisation (src/prg_normalize_mod.F90)
# --------------------------------------------------------------------------


def gershgorin(h: np.ndarray) -> tuple[float, float]:
    """Gershgorin interval (bml_gershgorin as used at src/prg_normalize_mod.F90:41)."""
    d = np.diag(h)
    r = np.abs(h).sum(axis=1) - np.abs(d)
    return float((d - r).min()), float((d + r).max())


def bml_normalize(x: np.ndarray, mineval: float, maxeval: float) -> np.ndarray:
    """X = (maxeval I - X)/(maxeval - mineval) (src/prg_normalize_mod.F90:25-29;
    explicit form at src/prg_response_mod.F90:694-699)."""
    n = x.shape[0]
    return (maxeval * np.eye(n) - x) / (maxeval - mineval)


def normalize(h: np.ndarray) -> np.ndarray:
    """prg_normalize (src/prg_normalize_mod.F90:32-46)."""
    lo, hi = gershgorin(h)
    return bml_normalize(h, lo, hi)


def normalize_implicit_fermi(h, cnst, mu):
    """prg_normalize_implicit_fermi (src/prg_normalize_mod.F90:83-94):
    X0 = (0.5 + cnst mu) I - cnst H."""
    n = h.shape[0]
    return -cnst * h + (0.5 + cnst * mu) * np.eye(n)


# --------------------------------------------------------------------------
# SP2 family (src/prg_sp2_mod.F90)
# --------------------------------------------------------------------------


class NotConverging(RuntimeError):
    """The reference calls ``stop`` / ``error stop`` here."""


def _is_rel(sp2conv: str) -> bool:
    # Fortran character comparison pads the shorter operand with blanks.
    return sp2conv.rstrip() == "Rel"


def sp2_basic(h, bndfil, minsp2iter, maxsp2iter, sp2conv, idemtol):
    """prg_sp2_basic (src/prg_sp2_mod.F90:49-135). Returns (rho, iterations)."""
    n = h.shape[0]
    occ = bndfil * n
    x = normalize(h)
    for it in range(0, maxsp2iter + 1):
        trx = np.trace(x)
        x2 = x @ x
        trx2 = np.trace(x2)
        if trx - occ <= 0.0:
            x = 2.0 * x - x2
            trx = 2.0 * trx - trx2
        else:
            x = x2
            trx = trx2
        if abs(occ - trx) < idemtol and it > minsp2iter:
            return 2.0 * x, it
        if it == maxsp2iter:
            raise NotConverging("sp2 purification is not converging")
    raise NotConverging("unreachable")


def sp2_basic_tcore(h, bndfil, maxsp2iter):
    """prg_sp2_basic_tcore (src/prg_sp2_mod.F90:138-246): the prg_sp2_basic
    branch rule (fresh trace of X each iteration, :177, :198-212) for
    iterations 0..maxsp2iter with the convergence exit commented out (:214);
    D = 2X (:241).  Returns (rho, pp) with pp the branch sequence taken."""
    n = h.shape[0]
    occ = bndfil * n
    x = normalize(h)
    pp = []
    for _ in range(0, maxsp2iter + 1):
        trx = np.trace(x)
        x2 = x @ x
        if trx - occ <= 0.0:
            x = 2.0 * x - x2
            pp.append(0)
        else:
            x = x2
            pp.append(1)
    return 2.0 * x, pp


def _drop(a, threshold):
    """A sparse BML format's thresholding: entries with |a| <= threshold are
    dropped (lanl/bml, absent here: parity unpinned beyond this rule)."""
    return np.where(np.abs(a) > threshold, a, 0.0) if threshold > 0.0 else a


def _sp2_alg2_core(x, occ, minsp2iter, maxsp2iter, sp2conv, idemtol, fresh_trace, want_vv,
                   threshold=0.0):
    """Shared body of prg_sp2_alg2 (src/prg_sp2_mod.F90:263-380) and
    prg_sp2_alg2_genseq (:395-533).  ``fresh_trace`` selects genseq's
    ``trx = trace(1)`` (:452) instead of the recursed trace (:334,341).
    threshold > 0 is the sparse formats' arithmetic: bml_multiply_x2 drops
    the product's entries at or below it (its trace taken before, as the
    product is formed) and bml_add_deprecated drops those of 2X - X2
    (:310, :328)."""
    trx = np.trace(x)
    it = 0
    pp, vv = [], []
    e0 = e1 = e2 = 0.0
    while it < maxsp2iter:
        it += 1
        x2 = x @ x
        if fresh_trace:
            trx = np.trace(x)
        trx2 = np.trace(x2)
        x2 = _drop(x2, threshold)
        if want_vv:
            vv.append(math.sqrt(float(np.sum((x - x2) ** 2))))
        tr2xx2 = 2.0 * trx - trx2
        trx_old = trx
        limdiff = abs(trx2 - occ) - abs(tr2xx2 - occ)
        brk = False
        if limdiff >= idemtol:
            x = _drop(2.0 * x - x2, threshold)
            trx = 2.0 * trx - trx2
            pp.append(0)
        elif limdiff < -idemtol:
            x = x2
            trx = trx2
            pp.append(1)
        else:
            it -= 1
            trx = trx_old
            brk = True
        e2, e1 = e1, e0
        e0 = abs(trx - trx_old)
        if _is_rel(sp2conv) and it >= minsp2iter and (e0 >= e2 or e0 < idemtol):
            brk = True
        if it == maxsp2iter:
            raise NotConverging("SP2 purification is not converging")
        if brk:
            break
    return x, it, pp, vv


def sp2_alg2(h, bndfil, minsp2iter, maxsp2iter, sp2conv, idemtol, threshold=0.0):
    """prg_sp2_alg2 (src/prg_sp2_mod.F90:263-380). Returns (rho, iterations, pp).
    threshold > 0: the arithmetic of a sparse BML format (see _sp2_alg2_core)."""
    occ = bndfil * h.shape[0]
    x, it, pp, _ = _sp2_alg2_core(normalize(h), occ, minsp2iter, maxsp2iter, sp2conv,
                                  idemtol, False, False, threshold)
    return 2.0 * x, it, pp


def sp2_alg2_genseq(h, bndfil, minsp2iter, maxsp2iter, sp2conv, idemtol):
    """prg_sp2_alg2_genseq (src/prg_sp2_mod.F90:395-533).
    Returns (rho, pp, icount, vv); vv has one entry per executed product
    (including the discarded last product when the loop ends in the
    limdiff branch, as the reference writes vv(iter) before undoing iter)."""
    occ = bndfil * h.shape[0]
    x, it, pp, vv = _sp2_alg2_core(normalize(h), occ, minsp2iter, maxsp2iter, sp2conv,
                                   idemtol, True, True)
    return 2.0 * x, pp, it, vv


def _sp2_alg2_replay(x, pp, icount):
    """Loop body of prg_sp2_alg2_seq (:543-627) and the in-place variant (:638-722)."""
    vv = []
    for it in range(icount):
        x2 = x @ x
        vv.append(math.sqrt(float(np.sum((x - x2) ** 2))))
        x = 2.0 * x - x2 if pp[it] == 0 else x2
    return 2.0 * x, vv


def sp2_alg2_seq(h, pp, icount):
    """prg_sp2_alg2_seq (src/prg_sp2_mod.F90:543-627). Returns (rho, vv)."""
    return _sp2_alg2_replay(normalize(h), pp, icount)


def sp2_alg2_seq_inplace(x, pp, icount, mineval, maxeval):
    """prg_prg_sp2_alg2_seq_inplace (src/prg_sp2_mod.F90:638-722)."""
    return _sp2_alg2_replay(bml_normalize(x, mineval, maxeval), pp, icount)


def sp2_alg1(h, bndfil, minsp2iter, maxsp2iter, sp2conv, idemtol, genseq=False):
    """prg_sp2_alg1 (src/prg_sp2_mod.F90:734-864) and, with genseq=True,
    prg_sp2_alg1_genseq (:879-1000).  Returns (rho, iterations, pp, vv)."""
    n = h.shape[0]
    occ = bndfil * n
    x = normalize(h)
    trx = np.trace(x)
    it = 0
    pp, vv = [], []
    e0 = e1 = e2 = 0.0
    while it < maxsp2iter:
        it += 1
        x2 = x - x @ x
        if genseq:
            vv.append(math.sqrt(float(np.sum(x2 ** 2))))
        trx2 = np.trace(x2)
        limdiff = abs(trx - trx2 - occ) - abs(trx + trx2 - occ)
        brk = False
        if limdiff >= idemtol:
            x = x + x2
            trx = trx + trx2
            pp.append(0)
        elif limdiff < -idemtol:
            x = x - x2
            trx = trx - trx2
            pp.append(1)
        elif (not genseq) and limdiff == 0.0 and it == 1:
            x = x - x2
            trx = trx - trx2
            pp.append(1)
        else:
            it -= 1
            brk = True
        e2, e1 = e1, e0
        e0 = abs(trx2)
        if _is_rel(sp2conv) and it >= minsp2iter and (e2 <= e0 or e0 < idemtol):
            brk = True
        if genseq and sp2conv.rstrip() == "Abs" and abs(limdiff) < idemtol:
            brk = True
        if it == maxsp2iter:
            raise NotConverging("SP2 purification is not converging")
        if brk:
            break
    return 2.0 * x, it, pp, vv


def sp2_alg1_genseq(h, bndfil, minsp2iter, maxsp2iter, sp2conv, idemtol):
    """prg_sp2_alg1_genseq (src/prg_sp2_mod.F90:879-1000). Returns (rho, pp, icount, vv)."""
    rho, it, pp, vv = sp2_alg1(h, bndfil, minsp2iter, maxsp2iter, sp2conv, idemtol, genseq=True)
    return rho, pp, it, vv


def _sp2_alg1_replay(x, pp, icount):
    """Loop body of prg_sp2_alg1_seq (:1010-1089) and the in-place variant (:1100-1177)."""
    vv = []
    for it in range(icount):
        x2 = x - x @ x
        vv.append(math.sqrt(float(np.sum(x2 ** 2))))
        x = x + x2 if pp[it] == 0 else x - x2
    return 2.0 * x, vv


def sp2_alg1_seq(h, pp, icount):
    """prg_sp2_alg1_seq (src/prg_sp2_mod.F90:1010-1089)."""
    return _sp2_alg1_replay(normalize(h), pp, icount)


def sp2_alg1_seq_inplace(x, pp, icount, mineval, maxeval):
    """prg_prg_sp2_alg1_seq_inplace (src/prg_sp2_mod.F90:1100-1177)."""
    return _sp2_alg1_replay(bml_normalize(x, mineval, maxeval), pp, icount)


def sp2_submatrix(x, pp, icount, vv, mineval, maxeval, core_size):
    """prg_sp2_submatrix[_inplace] (src/prg_sp2_mod.F90:1190-1310).
    vv(iter) += sum over the first core_size rows of (X - X^2)^2 (no sqrt).
    Returns (rho, vv) with vv updated in a copy."""
    vv = list(vv)
    x = bml_normalize(x, mineval, maxeval)
    for it in range(icount):
        x2 = x - x @ x
        vv[it] += float(np.sum(x2[:core_size, :] ** 2))
        factor = 1.0 - 2.0 * pp[it]
        x = x + factor * x2
    return 2.0 * x, vv


# --------------------------------------------------------------------------
# sequence / gap helpers (src/prg_homolumo_mod.F90)
# --------------------------------------------------------------------------


def homolumogap(vv, imax, pp, mineval, maxeval):
    """prg_homolumogap (src/prg_homolumo_mod.F90:22-89). vv/pp are 0-based lists
    holding the reference's vv(1..), pp(1..).  Returns (ehomo, elumo, egap)."""
    x_a, x_b = 0.0, 1.0
    i = imax
    while vv[i - 1] <= 0.0:
        i -= 1
    hgamma = 6.0 - 4.0 * math.sqrt(2.0)
    hgamma = hgamma * (1.0 - hgamma)
    while vv[i - 1] < hgamma:
        precomp = math.sqrt(1.0 - 4.0 * vv[i - 1])
        y_a = 0.5 * (1.0 + precomp)
        y_b = 0.5 * (1.0 - precomp)
        for j in range(i - 1, 0, -1):
            if pp[j - 1] > 0:
                y_a = math.sqrt(y_a)
                y_b = math.sqrt(y_b)
            else:
                y_a = 1.0 - math.sqrt(1.0 - y_a)
                y_b = 1.0 - math.sqrt(1.0 - y_b)
        x_a = max(x_a, y_a)
        x_b = min(x_b, y_b)
        i -= 1
        if i < 1:
            break
    ehomo = maxeval - x_a * (maxeval - mineval)
    elumo = maxeval - x_b * (maxeval - mineval)
    return ehomo, elumo, elumo - ehomo


def sp2sequence(mineval, maxeval, ehomo, elumo, errlimit):
    """prg_sp2sequence (src/prg_homolumo_mod.F90:97-147). Returns (pp, imax).
    Keeps the reference's parenthesisation of the second comparison operand."""
    eh = (maxeval - ehomo) / (maxeval - mineval)
    el = (maxeval - elumo) / (maxeval - mineval)
    error = 1.0
    it = 0
    pp = []
    while error > errlimit:
        it += 1
        if (abs(1.0 - eh * eh) + abs(el * el)) < abs(1.0 - (2.0 * eh - eh * eh) + abs(2.0 * el - el * el)):
            pp.append(1)
        else:
            pp.append(0)
        sgm = 1.0 - 2.0 * pp[-1]
        eh = eh + sgm * (eh - eh * eh)
        el = el + sgm * (el - el * el)
        error = abs(1.0 - eh) + abs(el)
        if it >= 100:
            error = 0.0
    return pp, it


# --------------------------------------------------------------------------
# eigendecomposition densities (src/prg_densitymatrix_mod.F90)
# --------------------------------------------------------------------------


def t0_occupations(n, bndfil):
    """Occupation rule of prg_build_density_T0 (src/prg_densitymatrix_mod.F90:73-84):
    2 for state i (1-based) when i - nocc < 1e-4, else 0; state int(nocc)+1
    gets 1 when nocc has a fractional part above 0.01."""
    nocc = n * bndfil
    occ = np.zeros(n)
    for i in range(1, n + 1):
        if i - nocc < 0.0001 or abs(i - nocc) < 0.0001:
            occ[i - 1] = 2.0
    if abs(nocc - int(nocc)) > 0.01:
        occ[int(nocc)] = 1.0
    return occ


def build_density_T0(h, bndfil):
    """prg_build_density_T0 (src/prg_densitymatrix_mod.F90:36-101)."""
    n = h.shape[0]
    w, q = np.linalg.eigh(h)
    occ = t0_occupations(n, bndfil)
    return (q * occ) @ q.T, w


def fermi_dm(e, ef, kbt):
    """fermi of src/prg_densitymatrix_mod.F90:886-896 (0 beyond (e-ef)/kbt > 100)."""
    e = np.asarray(e, dtype=float)
    x = (e - ef) / kbt
    out = np.zeros_like(e)
    ok = x <= 100.0
    out[ok] = 1.0 / (1.0 + np.exp(x[ok]))
    return out


def dfermi_dm(e, ef, kbt):
    """dfermi of src/prg_densitymatrix_mod.F90:902-916 (with its -1000 sentinel)."""
    e = np.asarray(e, dtype=float)
    x = (e - ef) / kbt
    out = np.zeros_like(e)
    big = x > 100.0
    out[big & (np.abs(e - ef) < 0.001)] = -1000.0
    ok = ~big
    f = 1.0 / (1.0 + np.exp(x[ok]))
    out[ok] = f * f * np.exp(x[ok]) / kbt
    return out


def build_density_T_fermi(h, kbt, ef):
    """prg_build_density_T_Fermi (src/prg_densitymatrix_mod.F90:509-582):
    returns (rho, drho) with rho = Q 2 f(e) Q^T and drho = Tr(Q 2 f'(e) Q^T)."""
    w, q = np.linalg.eigh(h)
    rho = (q * (2.0 * fermi_dm(w, ef, kbt))) @ q.T
    drho = float(np.sum(2.0 * dfermi_dm(w, ef, kbt)))
    return rho, drho


def fermi(e, mu, beta):
    """fermi (src/prg_implicit_fermi_mod.F90:1192-1198)."""
    return 1.0 / (1.0 + np.exp(beta * (e - mu)))


def get_density_matrix(h, beta, mu):
    """prg_get_density_matrix (src/prg_implicit_fermi_mod.F90:1056-1100): no factor 2."""
    w, q = np.linalg.eigh(h)
    return (q * fermi(w, mu, beta)) @ q.T


def test_density_matrix(h, beta, mu, nocc, osteps, occ_err_limit):
    """prg_test_density_matrix (src/prg_implicit_fermi_mod.F90:1112-1186).
    Returns (P, mu)."""
    occ_err = 1000.0
    it = 0
    w, q = np.linalg.eigh(h)
    p = None
    while (osteps == 0 and occ_err > occ_err_limit) or (osteps > 0 and it < osteps):
        it += 1
        p = (q * fermi(w, mu, beta)) @ q.T
        trp = np.trace(p)
        dpdmu = beta * (trp - np.sum(p * p))
        occ_err = abs(trp - nocc)
        if occ_err > occ_err_limit:
            mu = mu + (nocc - trp) / dpdmu
    return p, mu


def check_idempotency(a):
    """prg_check_idempotency (src/prg_densitymatrix_mod.F90:862-880): ||A^2 - A||_F."""
    return float(np.linalg.norm(a @ a - a))


# --------------------------------------------------------------------------
# implicit Fermi (src/prg_implicit_fermi_mod.F90)
# --------------------------------------------------------------------------


def conjgrad(a, p, b, cg_tol, maxit=50):
    """prg_conjgrad (src/prg_implicit_fermi_mod.F90:1006-1048): solve A X = B,
    stop when the squared Frobenius norm of the residual <= cg_tol."""
    p = p.copy()
    r = b - a @ p
    rn = np.sum(r * r)
    k = 0
    d = None
    rn_old = rn
    while rn > cg_tol:
        k += 1
        if k == 1:
            d = r.copy()
        else:
            d = (rn / rn_old) * d + r
        w = a @ d
        alpha = rn / np.sum(d * w)
        p = p + alpha * d
        r = r - alpha * w
        rn_old = rn
        rn = np.sum(r * r)
        if k > maxit:
            raise NotConverging("Conjugate gradient is not converging")
    return p, k


def finite_diff(h0, hs, mu0, mus, beta, order, lam, h):
    """prg_finite_diff (src/prg_implicit_fermi_mod.F90:726-831): the order-1..3
    finite differences of get_density_matrix along H0 + sum l^i H_i,
    mu0 + sum l^i mu_i at l = lam (order 1 forward, 2 and 3 central).
    Returns the list of difference matrices (the reference prints them)."""
    def dens(l):
        hh = np.array(h0, dtype=np.float64)
        mu = mu0
        for i in range(order):
            hh = hh + l ** (i + 1) * hs[i]
            mu += l ** (i + 1) * mus[i]
        return get_density_matrix(hh, beta, mu)

    d0, d1p = dens(lam), dens(lam + h)
    out = [(d1p - d0) / h]
    if order > 1:
        d1m = dens(lam - h)
        out.append((d1p - 2.0 * d0 + d1m) / (2.0 * h * h))
    if order > 2:
        d2m, d2p = dens(lam - 2 * h), dens(lam + 2 * h)
        out.append((d2p - d2m) / (12.0 * h ** 3) - (d1p - d1m) / (6.0 * h ** 3))
    return out


def pcg(a, p, b, cg_tol, maxit=100):
    """prg_pcg (src/prg_implicit_fermi_mod.F90:934-996): conjugate gradient for
    A X = B preconditioned by M = diag(A)^-1, run while ||R||_F^2 > cg_tol; the
    inner products are bml_trace_mult, Tr(X Y).  Returns (X, R, iterations)
    (the reference leaves R in the right-hand side's matrix)."""
    def trm(x, y):
        return float(np.sum(x * y.T))

    minv = 1.0 / np.diag(a)
    p = p.copy()
    r = b - a @ p
    z = minv[:, None] * r
    r_new = trm(z, r)
    d = z.copy()
    k = 0
    r_old = r_new
    while np.sum(r * r) > cg_tol:
        k += 1
        if k != 1:
            d = (r_new / r_old) * d + z
        w = a @ d
        alpha = trm(r, z) / trm(d, w)
        p = p + alpha * d
        r = r - alpha * w
        z = minv[:, None] * r
        r_old = r_new
        r_new = trm(r, z)
        if k > maxit:
            raise NotConverging("PCG is not converging")
    return p, r, k


def newtonschulz(a, ai, tol, max_iter=200):
    """prg_newtonschulz (src/prg_implicit_fermi_mod.F90:888-922):
    R = I - A Ai, err = ||R||_F, Ai <- Ai + Ai R until err <= tol; falls back
    to CG when the error grows by more than 0.01."""
    n = a.shape[0]
    err = 100000.0
    i = 0
    while err > tol:
        tmp = ai.copy()
        r = np.eye(n) - a @ ai
        prev = err
        err = float(np.linalg.norm(r))
        if prev + 0.01 < err:
            ai, _ = conjgrad(a, ai, np.eye(n), 0.00001)
        else:
            ai = tmp + tmp @ r
        i += 1
        if i > max_iter:
            raise NotConverging("Newton-Schulz is not converging")
    return ai, i


def setup_linsys(x, k):
    """prg_setup_linsys (src/prg_implicit_fermi_mod.F90:843-876): (A, B) of one
    level, A X_next = B, for even k: k = 2 gives B = X^2, A = 2(X^2 - X) + I;
    otherwise B = X^k, A = (I - X)^k + X^k by (k - 2)/2 products with X^2 and
    (I - X)^2."""
    n = x.shape[0]
    x2 = x @ x
    if k == 2:
        return 2.0 * (x2 - x) + np.eye(n), x2
    y = x2 - 2.0 * x + np.eye(n)
    b, a = x2.copy(), y.copy()
    for _ in range((k - 2) // 2):
        b = b @ x2
        a = a @ y
    return a + b, b


def implicit_fermi(h, nsteps, k, nocc, mu, beta, method, osteps, occ_err_limit, tol):
    """prg_implicit_fermi (src/prg_implicit_fermi_mod.F90:209-336), even k
    (setup_linsys).
    method 0: CG solve per level (:262-277).  method 1: a level-1 CG inverse at
    tol 0.9 refined by Newton-Schulz (to `tol`) at every level (:278-298).  The
    reference has that Newton-Schulz call commented out, which reuses the
    level-1 inverse for all levels and makes the recursion diverge (||P|| ~ 1e30
    on the model Hamiltonian); this restates the intended iteration.
    Output P is not scaled by 2.  Returns (P, mu)."""
    if k < 2 or k % 2:
        raise ValueError("the expansion order k must be even")
    n = h.shape[0]
    cnst = beta / (4 * k ** nsteps)
    occ_err = 10000.0
    it = 0
    p = None
    ai = np.eye(n)
    trp0 = 0.0
    while (osteps == 0 and occ_err > occ_err_limit) or (osteps > 0 and it < osteps):
        it += 1
        p = normalize_implicit_fermi(h, cnst, mu)
        for i in range(nsteps):
            y, p2 = setup_linsys(p, k)
            if method == 0:
                p, _ = conjgrad(y, p, p2, tol)
            else:
                if i == 0:
                    ai, _ = conjgrad(y, ai, np.eye(n), 0.9)
                ai, _ = newtonschulz(y, ai, tol)
                p = ai @ p2
        trp0 = np.trace(p)
        dpdmu = beta * (trp0 - np.sum(p * p))
        occ_err = abs(trp0 - nocc)
        if occ_err > occ_err_limit:
            mu = mu + (nocc - trp0) / dpdmu
        if it > 1000:
            raise NotConverging("chemical potential")
    return p, mu


def implicit_fermi_adjoint(h, a, nsteps, mu, beta, canonical=False):
    """Reverse sweep of the fixed-mu implicit-Fermi recursion (k = 2, exact
    inverses; the forward of src/prg_implicit_fermi_mod.F90:253-298): the
    gradient dTr(P A)/dH for symmetric A (no reference routine: the adjoint
    the GPU library adds).  Level i maps X to F = Y^-1 X^2, Y = 2 X^2 - 2X + I;
    with W = Y^-1 Gbar and U = W - 2 W F the adjoint of X is
    X Us + Us X + W F + (W F)^T, Us = (U + U^T) / 2, and X0 = (1/2 + c mu) I - c H
    gives dH = -c Gbar0 (c = beta / 2^(nsteps+2)).  canonical: the
    derivative at fixed occupation with the reference's dP/dmu ~ beta P (I - P)
    (src/prg_implicit_fermi_mod.F90:512-521): A is replaced by A - kappa I,
    kappa = Tr(A beta P (I - P)) / Tr(beta P (I - P))."""
    n = h.shape[0]
    eye = np.eye(n)
    cnst = beta / (2.0 ** (nsteps + 2))
    xs = [normalize_implicit_fermi(h, cnst, mu)]
    invs = []
    for _ in range(nsteps):
        x = xs[-1]
        x2 = x @ x
        yi = np.linalg.inv(2.0 * (x2 - x) + eye)
        invs.append(yi)
        xs.append(yi @ x2)
    g = np.array(a, dtype=np.float64)
    if canonical:
        p = xs[-1]
        cm = beta * (p - p @ p)
        tr = np.trace(cm)
        if abs(tr) > 1e-8:
            g = g - (np.sum(g * cm) / tr) * eye
    for i in reversed(range(nsteps)):
        x, f, yi = xs[i], xs[i + 1], invs[i]
        w = yi @ g
        wf = w @ f
        u = w - 2.0 * wf
        v = x @ (0.5 * (u + u.T))
        g = (v + v.T) + (wf + wf.T)
    return -cnst * g


def implicit_fermi_save_inverse(inv, h, nsteps, nocc, mu, beta, occ_err_limit, tol,
                                scf_it, maxiter=30):
    """prg_implicit_fermi_save_inverse (src/prg_implicit_fermi_mod.F90:46-193).
    ``inv`` is a list of nsteps matrices (updated copies are returned).
    Returns (P (scaled by 2), mu, inv, occiter_increment, ns_iterations)."""
    n = h.shape[0]
    eye = np.eye(n)
    inv = [m.copy() for m in inv]
    occ_err = 10.0
    alpha = 1.0
    prev = 1
    it = 0
    cnst = beta / (2.0 ** (nsteps + 2))
    if scf_it == 1:
        alpha = 4.0
        p = normalize_implicit_fermi(h, cnst, mu)
        p2 = p @ p
        y = 2.0 * (p2 - p) + eye
        ai, _ = conjgrad(y, eye, eye, 0.0001)
        inv = [eye.copy() for _ in range(nsteps)]
    totns = 0
    muadj = 0
    p = None
    while (occ_err > occ_err_limit or muadj == 1) and it < maxiter:
        it += 1
        muadj = 0
        p = normalize_implicit_fermi(h, cnst, mu)
        for i in range(nsteps):
            p2 = p @ p
            y = 2.0 * (p2 - p) + eye
            if it == 1:
                inv[i], _ = conjgrad(y, inv[i], eye, 0.001)
            inv[i], nsit = newtonschulz(y, inv[i], tol)
            totns += nsit
            p = inv[i] @ p2
        trp0 = np.trace(p)
        dpdmu = beta * (trp0 - np.sum(p * p))
        occ_err = abs(trp0 - nocc)
        if occ_err > 1.0:
            if nocc - trp0 < 0.0 and prev == -1:
                prev = -1
            elif nocc - trp0 > 0.0 and prev == 1:
                prev = 1
            elif nocc - trp0 > 0.0 and prev == -1:
                prev = 1
                alpha = alpha / 2
            else:
                prev = -1
                alpha = alpha / 2
            mu = mu + prev * alpha
            muadj = 1
        elif occ_err > occ_err_limit:
            mu = mu + (nocc - trp0) / dpdmu
            muadj = 1
    return 2.0 * p, mu, inv, it, totns


def implicit_fermi_zero(h, nsteps, mu, method, tol):
    """prg_implicit_fermi_zero (src/prg_implicit_fermi_mod.F90:347-425) with
    cnst = 0.5 min(1/(mu - lo), 1/(hi - mu)) from the Gershgorin interval.
    Method 1 is restated as intended (P2 = P^2 before the solve); the
    reference's method-1 branch multiplies by a P2 it never computed."""
    n = h.shape[0]
    lo, hi = gershgorin(h)
    cnst = 0.5 * min(1.0 / (mu - lo), 1.0 / (hi - mu))
    p = normalize_implicit_fermi(h, cnst, mu)
    ai = np.eye(n)
    for i in range(nsteps):
        p2 = p @ p
        y = 2.0 * (p2 - p) + np.eye(n)
        if method == 0:
            p, _ = conjgrad(y, p, p2, tol)
        else:
            if i == 0:
                ai, _ = conjgrad(y, ai, np.eye(n), 0.9)
            p = ai @ p2
    return p


def implicit_fermi_first_order_response(h0, h1, inv, nsteps, mu0, beta):
    """prg_implicit_fermi_first_order_response (src/prg_implicit_fermi_mod.F90:439-529).
    Returns (P0, P1) (neither scaled by 2)."""
    n = h0.shape[0]
    cnst = beta / (2.0 ** (2 + nsteps))
    p0 = normalize_implicit_fermi(h0, cnst, mu0)
    p1 = -cnst * h1
    for i in range(nsteps):
        c0 = p0 @ p0
        c = p0 @ p1 + p1 @ p0
        b = 2.0 * p1 - 2.0 * c
        p0 = inv[i] @ c0
        c = c + b @ p0
        p1 = inv[i] @ c
    bm = np.eye(n) - p0
    cm = beta * (p0 @ bm)
    dpdmu_trace = np.trace(cm)
    mu1 = -np.trace(p1) / dpdmu_trace
    if abs(dpdmu_trace) > 1e-8:
        p1 = p1 + mu1 * cm
    return p0, p1


def implicit_fermi_response(h0, hs, nsteps, mu0, beta, nocc, occ_tol, order, lin_tol=1e-10,
                            max_k=50):
    """prg_implicit_fermi_response (src/prg_implicit_fermi_mod.F90:547-712),
    orders 1..3, with its own chemical-potential loop.  Per level the inverse
    T^-1 is a tol-0.01 CG solve at level 1 (from the previous inverse, starting
    at zero) refined by Newton-Schulz to lin_tol at every level.  The reference
    has the Newton-Schulz call commented out (it reuses the level-1 inverse
    for all levels, and the recursion diverges); this restates the intended
    iteration.  Returns (P0, [P1..Porder], mu0, mu)."""
    n = h0.shape[0]
    eye = np.eye(n)
    cnst = beta / (2.0 ** (2 + nsteps))
    mu = [0.0] * order
    occ_err = 10000.0
    k = 0
    ti = np.zeros((n, n))
    p0 = None
    ps = []
    while occ_err > occ_tol:
        k += 1
        p0 = normalize_implicit_fermi(h0, cnst, mu0)
        ps = [normalize_implicit_fermi(hs[j], cnst, mu[j]) - 0.5 * eye for j in range(order)]
        for i in range(nsteps):
            c0 = p0 @ p0
            cs = [None] * order
            bs = [None] * order
            cs[0] = p0 @ ps[0] + ps[0] @ p0
            bs[0] = 2.0 * ps[0] - 2.0 * cs[0]
            if order > 1:
                cs[1] = ps[0] @ ps[0] + p0 @ ps[1] + ps[1] @ p0
                bs[1] = 2.0 * ps[1] - 2.0 * cs[1]
            if order > 2:
                cs[2] = ps[0] @ ps[1] + ps[1] @ ps[0] + p0 @ ps[2] + ps[2] @ p0
                bs[2] = 2.0 * ps[2] - 2.0 * cs[2]
            t = c0 - p0
            t = 2.0 * t + eye
            if i == 0:
                ti, _ = conjgrad(t, ti, eye, 0.01)
            ti, _ = newtonschulz(t, ti, lin_tol)
            p0 = ti @ c0
            cs[0] = cs[0] + bs[0] @ p0
            ps[0] = ti @ cs[0]
            if order > 1:
                cs[1] = cs[1] + bs[1] @ p0 + bs[0] @ ps[0]
                ps[1] = ti @ cs[1]
            if order > 2:
                cs[2] = cs[2] + bs[2] @ p0 + bs[1] @ ps[0] + bs[0] @ ps[1]
                ps[2] = ti @ cs[2]
        pmu = beta * (p0 @ (eye - p0))
        pmu_trace = np.trace(pmu)
        p0_trace = np.trace(p0)
        occ_err = abs(p0_trace - nocc)
        mu0 = mu0 + (nocc - p0_trace) / pmu_trace
        for j in range(order):
            ptr = np.trace(ps[j])
            mu[j] = mu[j] - ptr / pmu_trace
            occ_err += abs(ptr)
        if k > max_k:
            break
    return p0, ps, mu0, mu


# --------------------------------------------------------------------------
# density-matrix response (src/prg_response_mod.F90)
# --------------------------------------------------------------------------


def compute_response_sp2(h, h1, bndfil, minsp2iter, maxsp2iter, sp2conv, idemtol):
    """prg_compute_response_SP2 (src/prg_response_mod.F90:653-782).
    Returns (rho, rho1, iterations, pp)."""
    n = h.shape[0]
    occ = bndfil * n
    lo, hi = gershgorin(h)
    mm = hi - lo
    alpha = -1.0 / mm
    beta = hi / mm
    x = alpha * h + beta * np.eye(n)
    d = alpha * h1
    trx = np.trace(x)
    it = 0
    pp = []
    e0 = e1 = e2 = 0.0
    while it < maxsp2iter:
        it += 1
        x2 = x @ x
        a = x @ d + d @ x
        trx2 = np.trace(x2)
        tr2xx2 = 2.0 * trx - trx2
        trx_old = trx
        limdiff = abs(trx2 - occ) - abs(tr2xx2 - occ)
        brk = False
        if limdiff >= idemtol:
            x = 2.0 * x - x2
            d = 2.0 * d - a
            trx = 2.0 * trx - trx2
            pp.append(0)
        elif limdiff < -idemtol:
            x = x2
            d = a
            trx = trx2
            pp.append(1)
        else:
            trx = trx_old
            brk = True
        e2, e1 = e1, e0
        e0 = abs(trx - trx_old)
        if _is_rel(sp2conv) and it >= minsp2iter and (e0 >= e2 or e0 < idemtol):
            brk = True
        if it == maxsp2iter:
            raise NotConverging("SP2 purification is not converging")
        if brk:
            break
    return 2.0 * x, 2.0 * d, it, pp


def canon_response(h1_eig, nocc, beta, evals, mu0, m):
    """prg_canon_response (src/prg_response_mod.F90:947-1022): first-order
    response of the implicit-Fermi recursion in the eigenbasis of H0, with the
    canonical trace correction.  ``h1_eig`` is H1 in that eigenbasis."""
    cnst = beta / (2.0 ** (m + 2))
    p0 = 0.5 + cnst * (evals - mu0)
    p1 = -cnst * h1_eig
    for _ in range(m):
        p02 = p0 * p0
        dx1 = p0[:, None] * p1 + p1 * p0[None, :]
        id0 = 1.0 / (2.0 * (p02 - p0) + 1.0)
        p0 = id0 * p02
        p1 = id0[:, None] * (dx1 + 2.0 * (p1 - dx1) * p0[None, :])
    dpdmu = beta * p0 * (1.0 - p0)
    mu1 = -np.trace(p1) / np.sum(dpdmu)
    p1 = p1.copy()
    p1[np.diag_indices_from(p1)] += mu1 * dpdmu
    return p1


def canon_response_p1_dpdmu(h1_eig, beta, evals, mu0, m):
    """prg_canon_response_p1_dpdmu (src/prg_response_mod.F90:1112-1183): the
    recursion of canon_response without the trace correction.  Returns
    (P1, dPdmu) with dPdmu = beta p0 (1 - p0) after the m steps.
    (prg_canon_response_orig, :1038-1105, is canon_response itself.)"""
    cnst = beta / (2.0 ** (m + 2))
    p0 = 0.5 + cnst * (np.asarray(evals) - mu0)
    p1 = -cnst * np.asarray(h1_eig, dtype=np.float64)
    for _ in range(m):
        p02 = p0 * p0
        dx1 = (p0[:, None] + p0[None, :]) * p1
        id0 = 1.0 / (2.0 * (p02 - p0) + 1.0)
        p0 = id0 * p02
        p1 = id0[:, None] * (dx1 + 2.0 * (p1 - dx1) * p0[None, :])
    return p1, beta * p0 * (1.0 - p0)


def canon_response_vector(h1_eig, beta, evals, mu0):
    """prg_canon_response_vector (src/prg_response_mod.F90:849-931): the exact
    Fermi response in the eigenbasis, P1 = K o H1 with K_ji = (f_i - f_j) /
    (e_i - e_j) off the diagonal and K_ii = df/de = -beta t / (1 + t)^2,
    t = exp(beta (e_i - mu0)), then the diagonal shifted by -dmu K_ii with
    dmu = Tr P1 / Tr K (0 when |Tr K| <= 1e-12).  Degenerate eigenvalue pairs
    divide by zero, as in the reference."""
    e = np.asarray(evals, dtype=np.float64)
    n = e.shape[0]
    f = 1.0 / (1.0 + np.exp(beta * (e - mu0)))
    x = e[None, :] - e[:, None]
    y = f[None, :] - f[:, None]
    x[np.diag_indices(n)] += 1.0
    with np.errstate(divide="ignore", invalid="ignore"):
        k = y / x
    t = np.exp((e - mu0) * beta)
    k[np.diag_indices(n)] += -t / ((1.0 + t) * (1.0 + t)) * beta
    p1 = k * np.asarray(h1_eig, dtype=np.float64)
    trk = float(np.sum(np.diag(k)))
    dmu = float(np.sum(np.diag(p1))) / trk if abs(trk) > 1e-12 else 0.0
    p1[np.diag_indices(n)] -= dmu * np.diag(k)
    return p1


def rs_occupations(n, bndfil):
    """Occupations of prg_compute_response_RS (src/prg_response_mod.F90:319-329):
    2 for state i (1-based) with i <= nocc, else 0; state int(nocc)+1 gets 1
    when nocc has a fractional part above 0.01."""
    nocc = n * bndfil
    occ = np.array([2.0 if i <= nocc else 0.0 for i in range(1, n + 1)])
    if abs(nocc - int(nocc)) > 0.01:
        occ[int(nocc)] = 1.0
    return occ


def compute_response_rs(h, h1, bndfil):
    """prg_compute_response_RS (src/prg_response_mod.F90:250-361): first-order
    Rayleigh-Schroedinger response at T = 0.  V = C^T H1 C, V~_ij = V_ij /
    (e_j - e_i) (0 on the diagonal), C1 = C V~, rho1 = C f C1^T + C1 f C^T."""
    e, c = np.linalg.eigh(h)
    v = c.T @ h1 @ c
    n = h.shape[0]
    de = e[None, :] - e[:, None]
    with np.errstate(divide="ignore", invalid="ignore"):
        vt = np.where(np.eye(n, dtype=bool), 0.0, v / np.where(np.eye(n, dtype=bool), 1.0, de))
    c1 = c @ vt
    occ = rs_occupations(n, bndfil)
    return (c * occ) @ c1.T + c1 @ (occ[:, None] * c.T)


def compute_response_fd(h, h1, delta, bndfil):
    """prg_compute_response_FD (src/prg_response_mod.F90:380-424):
    (rho(H + delta H1) - rho(H - delta H1)) / (2 delta), rho = prg_build_density_T0."""
    rp, _ = build_density_T0(h + delta * h1, bndfil)
    rm, _ = build_density_T0(h - delta * h1, bndfil)
    return (rp - rm) / (2.0 * delta)


# --------------------------------------------------------------------------
# graph partitioning and the subgraph SP2 loop
# (src/prg_graph_mod.F90, src/prg_subgraphloop_mod.F90)
# --------------------------------------------------------------------------


def equal_partition(nodes_per_part, nnodes):
    """prg_equalPartition (src/prg_graph_mod.F90:359-399): list of 0-based core
    node lists."""
    if nnodes <= nodes_per_part:
        np_ = 1
    else:
        np_ = math.ceil(nnodes / nodes_per_part)
    parts = []
    for i in range(1, np_ + 1):
        psize = nodes_per_part if i * nodes_per_part <= nnodes else nnodes - nodes_per_part * (i - 1)
        parts.append([(i - 1) * nodes_per_part + j for j in range(psize)])
    return parts


def read_part(path):
    """prg_readPart (src/prg_graph_mod.F90:888-929): returns (name, total_nodes,
    parts) with parts as 0-based node lists (the file is 0-based; the reference
    converts to 1-based in memory)."""
    with open(path) as fh:
        tokens = fh.read().split()
    name = tokens[0]
    total_nodes, total_parts = int(tokens[1]), int(tokens[2])
    sizes = [int(t) for t in tokens[3:3 + total_parts]]
    pos = 3 + total_parts
    parts = []
    for i in range(total_parts):
        pos += 1  # part id
        parts.append([int(t) for t in tokens[pos:pos + sizes[i]]])
        pos += sizes[i]
    return name, total_nodes, parts


def matrix2submatrix_index(g, core, double_jump=True, h=None):
    """Core+halo index list for one part: bml_matrix2submatrix_index as the
    reference calls it (graph only at src/prg_subgraphloop_mod.F90:355-358 and
    src/prg_graphsolver_mod.F90:112-115; graph plus the Hamiltonian as the
    trailing optional matrix at src/prg_subgraphloop_mod.F90:60-63).  BML
    (lanl/bml, master, commit unpinned) is absent; this restates its published
    dense algorithm:

    1. the core nodes in the given order (repeats dropped);
    2. for each core node in that order, its unseen neighbours in G
       (nonzero off-diagonal entries of its row) in ascending index order;
    3. when h is given, the same walk over the rows of H (H's sparsity
       adds halo nodes the graph misses);
    4. with double_jump, the unseen G-neighbours of every node listed so far,
       in list order (a symbolic X^2 over the graph).

    So the set is S1 u N_G(S1) with S1 = C u N_G(C) u N_H(C), i.e. the
    nonzero columns of the core rows of (I + |G| + |H|)(I + |G|)
    (``halo_set_bruteforce``).  Returns (index_list, core_size)."""
    seen = np.zeros(g.shape[0], dtype=bool)
    idx = []
    for c in core:
        c = int(c)
        if not seen[c]:
            seen[c] = True
            idx.append(c)
    ncore = len(idx)

    def walk(m, rows):
        # each row's unseen nonzero columns in ascending order, rows in order
        for c in rows:
            nz = np.nonzero(m[c])[0]
            new = nz[~seen[nz]]
            seen[new] = True
            idx.extend(int(k) for k in new)

    walk(g, list(idx))
    if h is not None:
        walk(h, idx[:ncore])
    if double_jump:
        walk(g, idx[ncore:])
    return idx, ncore


def halo_set_bruteforce(g, core, double_jump=True, h=None):
    """The core+halo set of ``matrix2submatrix_index`` by symbolic products:
    the nonzero columns of the core rows of (I + |G| + |H|) (I + |G|) with
    double jump, of (I + |G| + |H|) without.  Ordering-free check of the set."""
    n = g.shape[0]
    pg = (np.abs(g) > 0).astype(np.int64)
    first = np.eye(n, dtype=np.int64) + pg
    if h is not None:
        first = first + (np.abs(h) > 0).astype(np.int64)
    row = first[list(core)].sum(axis=0) > 0
    if double_jump:
        row = (row.astype(np.int64) @ (np.eye(n, dtype=np.int64) + pg)) > 0
    return set(int(k) for k in np.nonzero(row)[0])


def subgraph_sp2_loop(h, g, parts, pp, max_iter, mineval, maxeval, double_jump=True,
                      halo_from_h=True):
    """prg_subgraphSP2Loop (src/prg_subgraphloop_mod.F90:36-126) on one rank:
    per part extract the core+halo block of H (halos from G and, as the
    reference passes h_bml to bml_matrix2submatrix_index, from H), run
    prg_sp2_submatrix_inplace, scatter the core rows (and the matching
    columns, the density being symmetric) back, then prg_fnormGraph (sqrt of
    the accumulated vv).  Returns (rho, vv)."""
    n = h.shape[0]
    rho = np.zeros((n, n))
    vv = [0.0] * max_iter
    for core in parts:
        idx, llsize = matrix2submatrix_index(g, core, double_jump, h if halo_from_h else None)
        sub = h[np.ix_(idx, idx)]
        d, vv = sp2_submatrix(sub, pp, max_iter, vv, mineval, maxeval, llsize)
        cidx = idx[:llsize]
        rho[np.ix_(cidx, idx)] = d[:llsize, :]
    vv = [math.sqrt(v) for v in vv]
    return rho, vv


def normalize_fermi(h, h1, hn, mu):
    """prg_normalize_fermi (src/prg_normalize_mod.F90:59-72):
    X = ((hN - mu) I - H) / (hN - h1)."""
    n = h.shape[0]
    return ((hn - mu) * np.eye(n) - h) / (hn - h1)


def _absmaxderivative(f, de):
    """absmaxderivative (src/prg_sp2_fermi_mod.F90:754-768)."""
    return float(np.max(np.abs(np.diff(f)) / de)) if len(f) > 1 else -10000.0


def sp2_fermi_init(h, nsteps, nocc, tscale, occ_err_limit, trace_limit, norecs=False,
                   beta=0.0, max_outer=200):
    """prg_sp2_fermi_init (src/prg_sp2_fermi_mod.F90:43-185) and, with norecs,
    prg_sp2_fermi_init_norecs (:203-387).  Returns
    dict(x, mu, beta, h1, hN, sgnlist, nsteps)."""
    lo, hi = gershgorin(h)
    beta0 = beta
    mu = 0.5 * (hi + lo)
    h1, hn = tscale * lo, tscale * hi
    n = h.shape[0]
    sgn = [0] * nsteps
    first = True
    occ_err = 1.0
    outer = 0
    while occ_err > occ_err_limit:
        outer += 1
        if outer > max_outer:
            raise NotConverging("prg_sp2_fermi_init: occupation loop did not converge")
        x = normalize_fermi(h, h1, hn, mu)
        probe = 1.0 - 0.001 * np.arange(1, 1001)
        x1 = -np.eye(n) / (hn - h1)
        for i in range(nsteps):
            x2 = x @ x
            tr0, tr2 = np.trace(x), np.trace(x2)
            if first:
                sgn[i] = -1 if abs(tr2 - nocc) < abs(2.0 * tr0 - tr2 - nocc) else 1
            t = x @ x1 + x1 @ x
            if sgn[i] == 1:
                x1 = 2.0 * x1 - t
                x = 2.0 * x - x2
                probe = 2.0 * probe - probe * probe
            else:
                x1 = t
                x = x2
                probe = probe * probe
            if norecs:
                beta = 4.0 * _absmaxderivative(probe, 0.001) / (hi - lo)
                if beta > beta0:
                    nsteps = i + 1
                    break
        first = False
        tr0, tr1 = np.trace(x), np.trace(x1)
        occ_err = abs(nocc - tr0)
        lam = (nocc - tr0) / tr1 if abs(tr1) > trace_limit else 0.0
        mu += lam
    trx = np.trace(x @ (np.eye(n) - x))
    beta = -np.trace(x1) / trx if abs(trx) > trace_limit else -1000.0
    return {"x": x, "mu": mu, "beta": beta, "h1": h1, "hN": hn, "sgnlist": sgn[:nsteps],
            "nsteps": nsteps}


def sp2_fermi(h, osteps, nsteps, nocc, mu, beta, h1, hn, sgnlist, eps, trace_limit,
              max_outer=200):
    """prg_sp2_fermi (src/prg_sp2_fermi_mod.F90:405-470).  Returns (x, mu)."""
    n = h.shape[0]
    occ_err = 1.0 + eps
    it = 0
    lam = 0.0
    dx = np.zeros((n, n))
    x = None
    while (osteps == 0 and occ_err > eps) or (osteps > 0 and it < osteps):
        it += 1
        if osteps == 0 and it > max_outer:
            raise NotConverging("prg_sp2_fermi: occupation loop did not converge")
        x = normalize_fermi(h, h1, hn, mu)
        for i in range(nsteps):
            x2 = x @ x
            x = 2.0 * x - x2 if sgnlist[i] == 1 else x2
        tr0 = np.trace(x)
        occ_err = abs(nocc - tr0)
        dx = -beta * (x @ (np.eye(n) - x))
        trdx = np.trace(dx)
        lam = (nocc - tr0) / trdx if abs(trdx) > trace_limit else 0.0
        mu += lam
    return x + lam * dx, mu


def sp2_inverse(f, mu, h1, hn, sgnlist):
    """sp2_inverse (src/prg_sp2_fermi_mod.F90:720-745)."""
    e = f
    for s in reversed(sgnlist):
        c = 0.5 * (1.0 + s)
        e = c - s * math.sqrt(abs(c - s * e))
    return hn - mu - e * (hn - h1)


def sp2_entropy_function(mu, h1, hn, sgnlist):
    """prg_sp2_entropy_function (src/prg_sp2_fermi_mod.F90:480-531): GG on the
    mesh ee = 0, 1e-4, ..., 1 (two-point Gauss per cell), minus its linear
    trend.  Returns (GG, ee)."""
    n = 10001
    ee = np.arange(n) * 0.0001
    gg = np.zeros(n)
    x1 = 0.5773502691896257
    acc = 0.0
    for i in range(n - 1):
        a, b = ee[i], ee[i + 1]
        c = 0.5 * (b - a)
        acc += c * sp2_inverse(0.5 * ((b - a) * x1 + (b + a)), mu, h1, hn, sgnlist)
        acc += c * sp2_inverse(0.5 * ((b - a) * -x1 + (b + a)), mu, h1, hn, sgnlist)
        gg[i + 1] = acc
    return gg - gg[-1] * ee, ee


def sp2_entropy_ts(d0, gg, ee):
    """sp2_entropy_ts (src/prg_sp2_fermi_mod.F90:541-587): linear interpolation
    of GG at |eigenvalue| of D0 (mesh index 1..10000)."""
    ts = 0.0
    for x in np.linalg.eigvalsh(d0):
        hs = abs(x)
        j = int(math.floor(hs / 0.0001)) + 1
        if 0 < j <= 10000:
            ts += ((hs - ee[j - 1]) * gg[j] + (ee[j] - hs) * gg[j - 1]) / (ee[j] - ee[j - 1])
    return ts


CRT_MODULI = (256, 255, 253, 251, 247, 241, 239, 233, 229, 227, 223, 217, 211, 199)


def crt_constants(nmod, square=True):
    """Moduli, q_k = (M/p_k)^-1 mod p_k and the scaling exponent a of the int8
    Chinese-remainder products (no reference counterpart: the build's
    FP64-accurate mixed-precision products).

    square=True (the SP2 square of an iterate with spectrum in [0, 1]): the
    residues stand for C - 2^a A = 2^(2a) (X~^2 - X~), whose entries lie in
    [-2^(2a)/4, 2^(2a)/4] (x^2 - x in [-1/4, 0] on [0, 1]; an entry of a
    symmetric matrix is at most its spectral norm), so the largest a with
    2^(2a - 1 + 0.001) <= M.  square=False (general product of operands
    scaled to row 2-norm <= 1): |C| <= 2^(2a) (1 + 1e-10), so 2^(2a + 1.001)
    <= M.  a <= 52 either way."""
    ps = CRT_MODULI[:nmod]
    m = 1
    for p in ps:
        m *= p
    qs = [pow(m // p % p, -1, p) for p in ps]
    slack = -1.001 if square else 1.001
    a = min(52, int(math.floor((math.log2(m) - slack) / 2)))
    return ps, qs, a, m


def _exact_int_product(a, b):
    """A B^T of integer matrices with entries in [-128, 128]: every partial sum
    is an integer of magnitude <= K 2^14 < 2^53 (K < 2^17 columns), so an
    FP64 BLAS product is exact whatever its summation order; returned as
    int64 (numpy's int64 matmul has no BLAS path and takes minutes at
    K = 2000)."""
    c = np.asarray(a, dtype=np.float64) @ np.asarray(b, dtype=np.float64).T
    return np.rint(c).astype(np.int64)


def crt_square(x, nmod):
    """X^2 the way the int8 CRT SP2 step forms it: A = rint(2^a X), per modulus
    the exact integer product of the symmetric residues, the fraction sum
    S = sum_k ((C mod p) q mod p) / p = C / M (mod 1); then
    (C - 2^a A) / M = S - A 2^a / M (mod 1), which lies in (-1/2, 1/2), and
    X~^2 = (C - 2^a A) / 2^(2a) + A 2^-a.  Integer products in int64 (exact:
    |sum| < K 128^2)."""
    ps, qs, a, m = crt_constants(nmod)
    av = np.rint(np.asarray(x) * 2.0 ** a)
    s = np.zeros_like(av)
    for p, q in zip(ps, qs):
        r = av - p * np.rint(av / p)
        r = np.where(r > 127, r - p, r)
        r = np.where(r < -128, r + p, r).astype(np.int64)
        c = _exact_int_product(r, r)
        v = (c * q).astype(np.float64)
        md = v - p * np.rint(v / p)
        s += md / p
    xa = av * 2.0 ** -a
    t = s - xa * (4.0 ** a / float(m))
    return (t - np.rint(t)) * (float(m) / 4.0 ** a) + xa


def crt_scale(x):
    """Power of two above the largest row 2-norm of a symmetric matrix (1 for a
    zero matrix): the operand scale of the int8 CRT general product."""
    norm = math.sqrt(float(np.max(np.sum(np.asarray(x) ** 2, axis=0))))
    if norm == 0.0:
        return 1.0
    return math.ldexp(1.0, math.frexp(norm)[1])


def crt_product(x, y, nmod):
    """X Y for symmetric, commuting X and Y the way the int8 CRT general
    product forms it: A = rint(2^a X / sX), B = rint(2^a Y / sY) with sX, sY
    from crt_scale, per modulus the exact integer product A B^T of the
    symmetric residues, the fraction sum of ((C mod p) q mod p) / p reduced
    mod 1, times M sX sY / 2^(2a); the upper triangle mirrored onto the lower.
    Same integer arithmetic as crt_square, which is the case X = Y with both
    scales 1."""
    ps, qs, a, m = crt_constants(nmod, square=False)
    sx, sy = crt_scale(x), crt_scale(y)
    xv = np.rint(np.asarray(x) * (2.0 ** a / sx))
    yv = np.rint(np.asarray(y) * (2.0 ** a / sy))

    def residues(v, p):
        r = v - p * np.rint(v / p)
        r = np.where(r > 127, r - p, r)
        return np.where(r < -128, r + p, r).astype(np.int64)

    s = np.zeros_like(xv)
    for p, q in zip(ps, qs):
        c = _exact_int_product(residues(xv, p), residues(yv, p))
        v = (c * q).astype(np.float64)
        md = v - p * np.rint(v / p)
        s += md / p
    f = (s - np.rint(s)) * (float(m) / 4.0 ** a) * sx * sy
    return np.triu(f) + np.triu(f, 1).T


def orthogonalize(a, z):
    """prg_orthogonalize (src/prg_nonortho_mod.F90:35-68): Z^T A Z."""
    return (z.T @ a) @ z


def deorthogonalize(a_ortho, z):
    """prg_deorthogonalize (src/prg_nonortho_mod.F90:81-120): Z A Z^T."""
    return z @ (a_ortho @ z.T)


def bfs_partition(g, nparts):
    """The build's replacement for METIS in prg_build_densityGP_T0
    (prg_metisPartition, src/prg_partition_mod.F90:216-311): recursive
    bisection of breadth-first orderings.  A bisection orders the subset by
    BFS (ascending neighbours) from the last node that a BFS from the
    component's lowest node reaches, components in ascending order of their
    lowest node, and cuts where the two sides get node counts proportional
    to their part counts (left = floor(k/2) parts, rounded half up).
    Returns the list of cores (ascending node ids)."""
    n = g.shape[0]
    adj = [[int(j) for j in np.nonzero((g[i] != 0) | (g[:, i] != 0))[0] if j != i]
           for i in range(n)]
    part = [-1] * n

    def bfs(start, inset, mark):
        out = [start]
        mark[start] = True
        head = 0
        while head < len(out):
            v = out[head]
            head += 1
            for nb in adj[v]:
                if nb in inset and not mark.get(nb, False):
                    mark[nb] = True
                    out.append(nb)
        return out

    def bisect(nodes, k, first):
        if k == 1:
            for v in nodes:
                part[v] = first
            return
        inset = set(nodes)
        mark = {}
        order = []
        for v in nodes:
            if mark.get(v, False):
                continue
            probe = bfs(v, inset, {})
            order.extend(bfs(probe[-1], inset, mark))
        kl = k // 2
        nl = int(math.floor(len(nodes) * kl / k + 0.5))
        bisect(sorted(order[:nl]), kl, first)
        bisect(sorted(order[nl:]), k - kl, first + kl)

    bisect(list(range(n)), nparts, 0)
    return [[v for v in range(n) if part[v] == p] for p in range(nparts)]


def build_density_gp_t0(h, g, bndfil, ef, parts, kbt=0.1):
    """prg_build_densityGP_T0 (src/prg_graphsolver_mod.F90:39-191) on one
    rank for a given partition: core+halo with the double jump (:122-126),
    per part prg_build_density_T_Fermi at kT = 0.1 (:147), core rows of each
    part written into rho (:155), then the Fermi-level step (:162-177) from
    the trace error and the summed dN/dEf of the parts.  Returns (rho, ef)."""
    n = h.shape[0]
    rho = np.zeros((n, n))
    dnocc = 0.0
    for core in parts:
        idx, llsize = matrix2submatrix_index(g, core, True)
        d, dn = build_density_T_fermi(h[np.ix_(idx, idx)], kbt, ef)
        dnocc += dn
        rho[np.ix_(idx[:llsize], idx)] = d[:llsize, :]
    nocc = float(np.trace(rho))
    error = abs(nocc - 2.0 * bndfil * n)
    if nocc < 2.0 * bndfil * n:
        dnocc = -dnocc
    if error > 1e-6 and abs(dnocc) > 1e-6:
        ef = ef - (nocc / dnocc if error > 1.0 else error / dnocc)
    return rho, ef


def buildzdiag(s):
    """prg_buildZdiag (src/prg_genz_mod.F90:176-276): Z = U s^-1/2 U^T from the
    eigen-decomposition of the overlap S (the Loewdin factor; a negative
    eigenvalue is the reference's 'not positive definite' stop)."""
    w, u = np.linalg.eigh(s)
    if np.min(w) <= 0.0:
        raise ValueError("overlap block is not positive definite")
    return (u / np.sqrt(w)) @ u.T


def build_zmat_gp(s, g, parts, z0=None):
    """prg_build_zmatGP (src/prg_graphsolver_mod.F90:202-331) on one rank for
    a given partition: the double-jump core+halo lists of the graph alone
    (:282-285, no Hamiltonian), per part prg_buildZdiag of the overlap's
    core+halo block (:306), its core rows written into zmat
    (bml_submatrix2matrix, :308); entries not written keep z0's value."""
    n = s.shape[0]
    z = np.zeros((n, n)) if z0 is None else np.array(z0, dtype=float)
    for core in parts:
        idx, llsize = matrix2submatrix_index(g, core, True)
        zp = buildzdiag(s[np.ix_(idx, idx)])
        z[np.ix_(idx[:llsize], idx)] = zp[:llsize, :]
    return z


# --------------------------------------------------------------------------
# closed forms on the eigenvalues (the pins of the matrix oracle)
# --------------------------------------------------------------------------


def sp2_alg2_eigen(evals, mineval, maxeval, occ, minsp2iter, maxsp2iter, sp2conv, idemtol):
    """The prg_sp2_alg2 control flow applied to the eigenvalue vector: X0 has
    eigenvalues (maxeval - e)/(maxeval - mineval); traces are sums.  Returns
    (x_final, iterations, pp) with x_final the eigenvalues of the final X."""
    x = (maxeval - np.asarray(evals)) / (maxeval - mineval)
    trx = float(np.sum(x))
    it = 0
    pp = []
    e0 = e1 = e2 = 0.0
    while it < maxsp2iter:
        it += 1
        x2 = x * x
        trx2 = float(np.sum(x2))
        tr2xx2 = 2.0 * trx - trx2
        trx_old = trx
        limdiff = abs(trx2 - occ) - abs(tr2xx2 - occ)
        brk = False
        if limdiff >= idemtol:
            x = 2.0 * x - x2
            trx = 2.0 * trx - trx2
            pp.append(0)
        elif limdiff < -idemtol:
            x = x2
            trx = trx2
            pp.append(1)
        else:
            it -= 1
            trx = trx_old
            brk = True
        e2, e1 = e1, e0
        e0 = abs(trx - trx_old)
        if _is_rel(sp2conv) and it >= minsp2iter and (e0 >= e2 or e0 < idemtol):
            brk = True
        if brk:
            break
    return x, it, pp


def sp2_poly_eigen(evals, mineval, maxeval, pp):
    """Replay a branch sequence on the eigenvalues: eigenvalues of the final X."""
    x = (maxeval - np.asarray(evals)) / (maxeval - mineval)
    for s in pp:
        x = x * x if s == 1 else 2.0 * x - x * x
    return x


def implicit_fermi_eigen(evals, mu, beta, nsteps):
    """Exact-inverse implicit-Fermi recursion on the eigenvalues:
    x0 = 0.5 - cnst (e - mu), cnst = beta/2^(nsteps+2);
    x <- x^2 / (2 x^2 - 2 x + 1), nsteps times."""
    cnst = beta / (2.0 ** (nsteps + 2))
    x = 0.5 - cnst * (np.asarray(evals) - mu)
    for _ in range(nsteps):
        x = x * x / (2.0 * x * x - 2.0 * x + 1.0)
    return x


def divided_differences(f_of_evals, df_of_evals, evals, tol=1e-9):
    """Daleckii-Krein kernel K_ij = (f(e_i) - f(e_j))/(e_i - e_j), with the
    derivative on (near-)degenerate pairs."""
    e = np.asarray(evals)
    f = f_of_evals(e)
    de = e[:, None] - e[None, :]
    df = f[:, None] - f[None, :]
    with np.errstate(divide="ignore", invalid="ignore"):
        k = np.where(np.abs(de) > tol, df / np.where(np.abs(de) > tol, de, 1.0), 0.0)
    deg = np.abs(de) <= tol
    d = df_of_evals(e)
    k[deg] = (0.5 * (d[:, None] + d[None, :]))[deg]
    return k

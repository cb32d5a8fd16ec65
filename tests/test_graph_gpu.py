"""GPU parity of the graph-partitioned SP2 loop (batched kernels) against the
oracle, and the reference's own subgraph test on its shipped fixture."""
import numpy as np
import pytest

import progress_oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def P():
    import progress_amd
    progress_amd.set_precision(progress_amd.PREC_FP64)
    return progress_amd


def _sequence(h, P):
    lo, hi = O.gershgorin(h)
    rho, pp, ic, vv = O.sp2_alg2_genseq(h, 0.5, 10, 200, "Abs", 1e-10)
    eh, el, eg = O.homolumogap(vv, ic, pp, lo, hi)
    seq, imax = O.sp2sequence(lo, hi, eh, el, 1e-12)
    return rho, lo, hi, seq, imax


@pytest.mark.parametrize("n,per_part,gthr", [(384, 48, 1e-4), (500, 37, 1e-3)])
def test_subgraph_loop_matches_oracle(P, n, per_part, gthr):
    print(f"[subgraph] n={n} per_part={per_part}", flush=True)
    h = O.model_hamiltonian(n, 3)
    rho_full, lo, hi, seq, imax = _sequence(h, P)
    g = np.where(np.abs(rho_full) >= gthr, rho_full, 0.0)
    parts = O.equal_partition(per_part, n)
    rho_o, vv_o = O.subgraph_sp2_loop(h, g, parts, seq, imax, lo, hi)
    gp = P.equal_partition(per_part, n)
    rho_g, vv_g = P.subgraph_sp2_loop(h, g, gp, seq, imax, lo, hi)
    # the loop passes H to the core+halo extraction, as the reference does
    # (src/prg_subgraphloop_mod.F90:60-63): H's neighbours of the core join
    # the halo before the double jump; the set is pinned by symbolic products
    halos_o = [O.matrix2submatrix_index(g, c, True, h)[0] for c in parts]
    assert gp.halos() == halos_o
    for c, lst in zip(parts, gp.halos()):
        assert set(lst) == O.halo_set_bruteforce(g, c, True, h)
    assert np.linalg.norm(rho_g - rho_o) < 1e-9
    assert np.allclose(vv_g, vv_o, rtol=1e-9, atol=1e-12)


@pytest.fixture(scope="module")
def wide_case():
    """Parts of 254-426 core+halo rows (padded 256-448): 128 x 128,
    128 x 64 and 64 x 64 tiles with 128-wide panels."""
    n = 1200
    h = O.model_hamiltonian(n, 5)
    rho_full, lo, hi, seq, imax = _sequence(h, None)
    g = np.where(np.abs(rho_full) >= 1e-3, rho_full, 0.0)
    rho_o, vv_o = O.subgraph_sp2_loop(h, g, O.equal_partition(150, n), seq, imax, lo, hi)
    return dict(n=n, h=h, g=g, lo=lo, hi=hi, seq=seq, imax=imax, rho=rho_o, vv=vv_o)


@pytest.mark.parametrize("panel,depth,kslab", [(64, 1, 16), (64, 2, 16), (64, 1, 32),
                                               (128, 1, 16)])
def test_subgraph_loop_tile_configs_match_oracle(P, wide_case, panel, depth, kslab, monkeypatch):
    """Every compiled configuration of the batched SP2 kernel (panel width,
    slabs of prefetch, slab depth) against the oracle."""
    print(f"[subgraph tiles] panel={panel} depth={depth} kslab={kslab}", flush=True)
    c = wide_case
    monkeypatch.setenv("PRG_BATCH_TILE", str(panel))
    monkeypatch.setenv("PRG_BATCH_DEPTH", str(depth))
    monkeypatch.setenv("PRG_BATCH_KSLAB", str(kslab))
    gp = P.equal_partition(150, c["n"])
    rho_g, vv_g = P.subgraph_sp2_loop(c["h"], c["g"], gp, c["seq"], c["imax"], c["lo"], c["hi"])
    assert P.subgraph_config()[:3] == (panel, depth, kslab)
    assert max(gp.lsize[i] for i in range(gp.totalParts)) > 384
    assert np.linalg.norm(rho_g - c["rho"]) < 1e-9
    assert np.allclose(vv_g, c["vv"], rtol=1e-9, atol=1e-12)


def test_subgraph_loop_graph_replays_equal_eager(P, wide_case, monkeypatch):
    """The iteration launches captured in a hipGraph and replayed by later
    calls give the eager launches' result bit for bit, on every replay, with
    other library work and synchronisations between the calls."""
    print("[subgraph hipGraph replays]", flush=True)
    c = wide_case
    mh, mg = P.Matrix.from_numpy(c["h"]), P.Matrix.from_numpy(c["g"])
    gp = P.equal_partition(150, c["n"])
    args = (c["seq"], c["imax"], c["lo"], c["hi"])
    monkeypatch.setenv("PRG_BATCH_GRAPH", "0")
    ref, vv_ref = P.subgraph_sp2_loop(mh, mg, gp, *args)
    assert P.subgraph_config()[3] == 0
    monkeypatch.delenv("PRG_BATCH_GRAPH")
    modes = []
    for r in range(4):
        rho, vv = P.subgraph_sp2_loop(mh, mg, gp, *args)
        modes.append(P.subgraph_config()[3])
        assert np.array_equal(rho, ref), f"call {r}"
        assert vv == vv_ref, f"call {r}"
        P.gershgorin(c["h"])
    assert modes[0] in (1, 2) and modes[1:] == [2, 2, 2]
    assert np.linalg.norm(ref - c["rho"]) < 1e-9


def test_subgraph_loop_block_diagonal_is_exact(P):
    print("[subgraph block diagonal]", flush=True)
    blocks = [O.model_hamiltonian(80, seed=s) for s in range(3)]
    n = 240
    h = np.zeros((n, n))
    for b, blk in enumerate(blocks):
        h[80 * b:80 * b + 80, 80 * b:80 * b + 80] = blk
    rho_full, lo, hi, seq, imax = _sequence(h, P)
    g = (np.abs(h) > 0).astype(float)
    gp = P.equal_partition(40, n)
    rho_g, _ = P.subgraph_sp2_loop(h, g, gp, seq, imax, lo, hi)
    ref, _ = O.sp2_alg2_seq_inplace(h, seq, imax, lo, hi)
    assert np.linalg.norm(rho_g - ref) < 1e-9


def test_reference_subgraphsp2_equal(P, golden):
    """tests/src/main.F90:847-883 + tests/src/test_subgraphloop.F90:59-152:
    poly.512.mtx, genseq (Rel, minsp2iter 25), graph = density thresholded at
    1e-3, equal parts of 48 nodes, sequence with errlimit 1e-12;
    idempotency of rho/2 <= 1e-2."""
    print("[subgraph reference test] n=6144", flush=True)
    h = O.read_mtx(str(golden / "poly.512.mtx.gz"))
    n = h.shape[0]
    rho, pp, ic, vv = P.sp2_alg2_genseq(h, 0.5, 25, 100, "Rel", 1e-10)
    lo, hi = P.gershgorin(h)
    eh, el, eg = P.homolumogap(vv, ic, pp, lo, hi)
    seq, imax = P.sp2sequence(lo, hi, eh, el, 1e-12)
    g = np.where(np.abs(rho) >= 1e-3, rho, 0.0)
    gp = P.equal_partition(48, n)
    rho_g, vvg = P.subgraph_sp2_loop(h, g, gp, seq, imax, lo, hi)
    x = 0.5 * rho_g
    assert np.linalg.norm(x @ x - x) <= 1e-2
    lsz = [gp.lsize[i] for i in range(gp.totalParts)]
    print(f"  parts={gp.totalParts} lsize range=({min(lsz)}, {max(lsz)}) iters={imax}")


@pytest.mark.parametrize("n,nparts,gthr", [(400, 6, 0.05), (700, 9, 0.02)])
def test_build_density_gp_t0_matches_oracle(P, n, nparts, gthr):
    """prg_build_densityGP_T0: the partition (the build's BFS partitioner, in
    place of METIS), the per-part kT = 0.1 densities from batched eigensolves,
    the core-row assembly and the Fermi-level step, against the oracle."""
    print(f"[gp_t0] n={n} nparts={nparts}", flush=True)
    h = O.model_hamiltonian(n, 4)
    g = np.where(np.abs(h) >= gthr, h, 0.0)
    parts = O.bfs_partition(g, nparts)
    assert P.bfs_partition(g, nparts).parts() == parts
    w = np.linalg.eigvalsh(h)
    ef = 0.5 * (w[n // 2 - 1] + w[n // 2])
    rho_o, ef_o = O.build_density_gp_t0(h, g, 0.5, ef, parts)
    rho_g, ef_g = P.build_density_gp_t0(h, g, 0.5, ef, nparts)
    assert np.linalg.norm(rho_g - rho_o) < 1e-9
    assert abs(ef_g - ef_o) < 1e-9 * max(1.0, abs(ef_o))


def test_build_density_gp_t0_one_part_is_exact(P):
    """One part holding every node: the result is the full kT = 0.1 density."""
    n = 300
    print(f"[gp_t0 one part] n={n}", flush=True)
    h = O.model_hamiltonian(n, 1)
    g = np.where(h != 0, 1.0, 0.0)
    rho_g, _ = P.build_density_gp_t0(h, g, 0.5, 0.2, 1)
    d, _ = O.build_density_T_fermi(h, 0.1, 0.2)
    assert np.linalg.norm(rho_g - d) < 1e-9


@pytest.mark.parametrize("panel,depth", [(64, 1), (128, 1), (128, 4)])
@pytest.mark.parametrize("mode", ["f16split", "bf16split"])
def test_subgraph_loop_split_precision(P, wide_case, mode, panel, depth, monkeypatch):
    """The batched loop with FP16 / BF16 hi+lo split products (prg_set_precision
    F16SPLIT / BF16SPLIT; three 16-bit MFMA products with FP32 accumulation
    per tile, FP64 update and core-row sums) on the n = 1200, 8-part case,
    with each of the three split kernels (PRG_BATCH_TILE / PRG_BATCH_DEPTH:
    64 x 64 tiles; 128-wide panels giving 128 x 128, 128 x 64 and 64 x 64
    tiles, slabs staged through registers; the same with a 4-slot LDS ring):
    the same core+halo lists as the FP64 loop, and the density and vv of the
    oracle's FP64 loop within the split products' accuracy (relative
    Frobenius bars 1e-3 FP16, 3e-2 BF16; vv within 1e-3 / 3e-2 of its
    largest entry)."""
    print(f"[subgraph split] {mode} panel={panel} depth={depth}", flush=True)
    c = wide_case
    gp64 = P.equal_partition(150, c["n"])
    P.subgraph_sp2_loop(c["h"], c["g"], gp64, c["seq"], c["imax"], c["lo"], c["hi"])
    monkeypatch.setenv("PRG_BATCH_TILE", str(panel))
    monkeypatch.setenv("PRG_BATCH_DEPTH", str(depth))
    prec = {"f16split": P.PREC_F16SPLIT, "bf16split": P.PREC_BF16SPLIT}[mode]
    P.set_precision(prec)
    try:
        gp = P.equal_partition(150, c["n"])
        rho_g, vv_g = P.subgraph_sp2_loop(c["h"], c["g"], gp, c["seq"], c["imax"], c["lo"],
                                          c["hi"])
        cfg = P.subgraph_config()
    finally:
        P.set_precision(P.PREC_FP64)
    assert cfg[0] == panel and cfg[1] == depth
    assert gp.halos() == gp64.halos()
    rel = np.linalg.norm(rho_g - c["rho"]) / np.linalg.norm(c["rho"])
    vv_o = np.asarray(c["vv"])
    dvv = np.max(np.abs(np.asarray(vv_g) - vv_o)) / vv_o.max()
    print(f"  rel. Frobenius {rel:.2e}, vv {dvv:.2e}", flush=True)
    bar = 1e-3 if mode == "f16split" else 3e-2
    assert rel < bar and dvv < bar


@pytest.mark.parametrize("nparts", [1, 6, 24])
def test_zmat_gp_matches_oracle(nparts):
    """prg_build_zmatGP against the oracle on a banded overlap S = I + 0.1 B
    (B symmetric, bandwidth 6) with its pattern as the graph, the library's
    BFS parts: Z to 1e-10 (batched rocSOLVER eigensolves vs LAPACK); one part
    gives S^-1/2 (Z S Z = I).  Entries outside the parts' rows keep the
    output handle's values."""
    import progress_amd as P
    n = 300
    rng = np.random.default_rng(nparts)
    b = np.triu(rng.standard_normal((n, n)), 0)
    b = np.triu(b) - np.triu(b, 7)
    s = np.eye(n) + 0.05 * (b + b.T)
    g = (np.abs(s) > 0).astype(float)
    print(f"[zmatGP] n={n} parts={nparts}", flush=True)
    z = P.build_zmat_gp(s, g, nparts)
    parts = P.bfs_partition(g, nparts).parts()
    ref = O.build_zmat_gp(s, g, parts)
    assert np.max(np.abs(z - ref)) < 1e-10
    if nparts == 1:
        assert np.allclose(z @ s @ z, np.eye(n), atol=1e-10)
    # an output handle with prior contents: unwritten entries are kept
    z0 = np.full((n, n), 3.0)
    out = P.Matrix.from_numpy(z0)
    got = P.build_zmat_gp(s, g, nparts, out=out).numpy()
    assert np.max(np.abs(got - O.build_zmat_gp(s, g, parts, z0=z0))) < 1e-10

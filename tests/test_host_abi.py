"""Host-side logic behind the C-ABI (no GPU compute): partition KATs of the
reference's tests, the HOMO-LUMO estimate and the SP2 sequence generator,
checked against the oracle."""
import numpy as np
import pytest

import progress_oracle as O

from test_abi import LIB

pytestmark = pytest.mark.skipif(not LIB.exists(), reason="library not built")


def test_equal_and_file_partition_kats(golden):
    import progress_amd as P
    # tests/src/main.F90:802-843
    assert P.equal_partition(6, 72).totalParts == 12
    gp = P.equal_partition(7, 72)
    assert gp.totalParts == 11 and gp.parts() == O.equal_partition(7, 72)
    fp = P.file_partition(golden / "test.part")
    assert fp.totalParts == 104 and fp.totalNodes == 673
    assert fp.parts() == O.read_part(str(golden / "test.part"))[2]


def test_homolumogap_and_sequence_match_oracle():
    import progress_amd as P
    h = O.model_hamiltonian(256, 0)
    lo, hi = O.gershgorin(h)
    rho, pp, ic, vv = O.sp2_alg2_genseq(h, 0.5, 10, 200, "Abs", 1e-10)
    ours = P.homolumogap(vv, ic, pp, lo, hi)
    ref = O.homolumogap(vv, ic, pp, lo, hi)
    assert np.allclose(ours, ref, rtol=0, atol=1e-14)
    pps, imax = P.sp2sequence(lo, hi, ref[0], ref[1], 1e-12)
    ppo, imo = O.sp2sequence(lo, hi, ref[0], ref[1], 1e-12)
    assert imax == imo and pps == ppo


def _grid(m):
    n = m * m
    g = np.zeros((n, n))
    for r in range(m):
        for c in range(m):
            i = r * m + c
            if c + 1 < m:
                g[i, i + 1] = g[i + 1, i] = 1.0
            if r + 1 < m:
                g[i, i + m] = g[i + m, i] = 1.0
    return g


def _geometric(n, cutoff, seed):
    rng = np.random.default_rng(seed)
    x = rng.uniform(0.0, 1.0, (n, 3))
    d = np.linalg.norm(x[:, None, :] - x[None, :, :], axis=-1)
    g = np.where((d < cutoff) & (d > 0), np.exp(-d), 0.0)
    return g


@pytest.mark.parametrize("case", ["path", "grid", "geometric", "components"])
def test_bfs_partition_matches_oracle(case):
    """The partitioner that replaces METIS: the C-ABI and the oracle produce the
    same parts; every node lands in exactly one part; part sizes follow the
    proportional cuts (within one node per bisection level)."""
    import progress_amd as P
    if case == "path":
        n = 10
        g = np.zeros((n, n))
        for i in range(n - 1):
            g[i, i + 1] = g[i + 1, i] = 1.0
        nparts, expect = 3, [[7, 8, 9], [3, 4, 5, 6], [0, 1, 2]]
    elif case == "grid":
        g, nparts, expect = _grid(12), 8, None
    elif case == "geometric":
        g, nparts, expect = _geometric(300, 0.22, 4), 7, None
    else:
        g = np.zeros((40, 40))
        g[:25, :25] = _grid(5)
        g[25:, 25:] = np.eye(15, k=1) + np.eye(15, k=-1)
        nparts, expect = 5, None
    ours = P.bfs_partition(g, nparts).parts()
    ref = O.bfs_partition(g, nparts)
    assert ours == ref
    if expect is not None:
        assert ref == expect
    allnodes = sorted(v for p in ref for v in p)
    assert allnodes == list(range(g.shape[0]))
    sizes = [len(p) for p in ref]
    assert max(sizes) - min(sizes) <= int(np.ceil(np.log2(nparts))) + 1


def test_bfs_partition_rejects_bad_count():
    import progress_amd as P
    with pytest.raises(P.ProgressError):
        P.bfs_partition(np.eye(4), 5)


def test_sp2_inverse_and_entropy_function_match_oracle():
    """Host-side helpers of the truncated SP2 module through the C-ABI."""
    import progress_amd as P
    sg = [1, -1, 1, 1, -1, 1, -1]
    mu, h1, hn = 0.15, -2.0, 3.0
    for f in (0.01, 0.4, 0.5, 0.93):
        assert abs(P.sp2_inverse(f, mu, h1, hn, sg) - O.sp2_inverse(f, mu, h1, hn, sg)) < 1e-14
    gg, ee = P.sp2_entropy_function(mu, h1, hn, sg)
    go, eo = O.sp2_entropy_function(mu, h1, hn, sg)
    assert np.array_equal(ee, eo)
    assert np.max(np.abs(gg - go)) < 1e-12


def test_sparse_import_keeps_entries_above_threshold():
    """bml_import_from_dense with a sparse type (ellpack) keeps |a| > threshold,
    the dense type keeps every entry (host-side, no GPU compute)."""
    import progress_amd as P
    a = np.array([[1.0, 1e-6, -2e-5], [1e-6, 2.0, 3e-4], [-2e-5, 3e-4, -1.0]])
    ell = P.Matrix.from_numpy(a, matrix_type=2, threshold=1e-5).numpy()
    assert np.array_equal(ell, np.where(np.abs(a) > 1e-5, a, 0.0))
    assert np.array_equal(P.Matrix.from_numpy(a, threshold=1e-5).numpy(), a)

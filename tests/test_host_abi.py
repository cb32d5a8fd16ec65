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

"""prg_gpu_sp2_alg2_device_noalloc: the FP64 device-pointer SP2 without any
allocation (static device scratch, N <= 1024), called through the C-ABI and
checked against prg_gpu_sp2_alg2_device on the same inputs: two calls in a
row (N = 1024, then the smaller N = 600 on the same scratch), each equal to
the allocating entry point's density to 1e-12 (the same kernels in the same
order; the bar allows for nothing else) with the same iteration count and
branch sequence, and the density within 1e-9 of the oracle at N = 600.  An
N above the scratch limit is rejected with status 1 before any launch."""
import ctypes as C

import numpy as np
import pytest

import progress_oracle as O

pytestmark = pytest.mark.gpu


def _noalloc(L, dh, dr, n, maxit=200):
    pp = (C.c_int * maxit)()
    it = C.c_int(0)
    st = L.prg_gpu_sp2_alg2_device_noalloc(
        C.c_int(n), C.c_void_p(dh.data_ptr()), C.c_void_p(dr.data_ptr()), C.c_double(0.5),
        C.c_int(1), C.c_int(maxit), b"Abs", C.c_double(1e-10), C.byref(it), pp)
    return st, it.value, list(pp[:it.value])


def test_noalloc_sp2_matches_allocating_entry_point():
    import torch
    import progress_amd as P
    L = P.lib()
    L.prg_gpu_sp2_alg2_device_noalloc.restype = C.c_int
    L.prg_progress_init()
    for n in (1024, 600):
        print(f"[noalloc SP2] n={n}", flush=True)
        h = O.model_hamiltonian(n, 3)
        ld = P.padded_ld(n)
        dh = torch.zeros((ld, ld), dtype=torch.float64, device="cuda")
        dh[:n, :n] = torch.from_numpy(h)
        r_ref = torch.zeros_like(dh)
        it_ref, pp_ref = P.sp2_alg2_device(dh, r_ref, n, 0.5, 1, 200, "Abs", 1e-10)
        # garbage in the block the call must overwrite; padding zero, as the
        # allocating entry point expects
        r_na = torch.zeros_like(dh)
        r_na[:n, :n] = 7.0
        torch.cuda.synchronize()
        st, it, pp = _noalloc(L, dh, r_na, n)
        assert st == 0, L.prg_last_error().decode()
        assert it == it_ref and pp == pp_ref
        diff = torch.max(torch.abs(r_na - r_ref)).item()
        print(f"  iterations {it}, max |rho_noalloc - rho| = {diff:.2e}", flush=True)
        assert diff <= 1e-12
        if n == 600:
            ref, it_o, pp_o = O.sp2_alg2(h, 0.5, 1, 200, "Abs", 1e-10)
            assert it == it_o and pp == list(pp_o)
            assert np.linalg.norm(r_na[:n, :n].cpu().numpy() - ref) < 1e-9


def test_noalloc_sp2_rejects_n_above_the_scratch():
    import torch
    import progress_amd as P
    L = P.lib()
    L.prg_gpu_sp2_alg2_device_noalloc.restype = C.c_int
    dh = torch.zeros((128, 128), dtype=torch.float64, device="cuda")
    st, _, _ = _noalloc(L, dh, dh, 1100)
    assert st == 1
    assert "1024" in L.prg_last_error().decode()

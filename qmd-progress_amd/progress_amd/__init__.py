"""Python binding of libprogress_gpu.so (MI355X density-matrix solvers).

Thin ctypes layer over the C-ABI declared in include/progress_gpu.h.  The
functions below mirror the reference's routines (names, argument meaning and
order of src/prg_sp2_mod.F90 & co.) on numpy arrays: each call imports the
Hamiltonian into a library matrix handle, runs the HIP path and exports the
result.  There is no CPU fallback: if the shared library is missing, or no
GPU is visible, the calls raise.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
LIB_PATH = _HERE.parent / "lib" / "libprogress_gpu.so"

PREC_FP64 = 0
PREC_F16SPLIT = 1
PREC_INT8CRT = 2
PREC_BF16SPLIT = 3

_dense, _double_real, _sequential, _col_major = 1, 2, 0, 1


class ProgressError(RuntimeError):
    """A library call recorded a non-zero status."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"[status {code}] {msg}")
        self.code = code


_lib = None


def lib():
    """Load the shared library (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(f"{LIB_PATH} is missing: run `make -C qmd-progress_amd` (or "
                          "__graft_entry__.build()) first; there is no CPU fallback")
    # One HIP runtime per process: when PyTorch is installed, load it first so
    # that this library's libamdhip64.so.7 dependency resolves to the runtime
    # torch already mapped (torch bundles its own copy; two copies in one
    # process cannot share the device).
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(str(LIB_PATH))
    vp, i, d, ip, dp, cp = C.c_void_p, C.c_int, C.c_double, C.POINTER(C.c_int), \
        C.POINTER(C.c_double), C.c_char_p
    sig = {
        "bml_zero_matrix": (vp, [i, i, i, i, i]),
        "bml_import_from_dense": (vp, [i, i, i, i, i, vp, d, i]),
        "bml_export_to_dense": (vp, [vp, i]),
        "bml_deallocate": (None, [C.POINTER(vp)]),
        "bml_get_N": (i, [vp]),
        "bml_trace": (d, [vp]),
        "bml_read_bml_matrix": (None, [vp, cp]),
        "prg_last_error_code": (i, []),
        "prg_last_error": (cp, []),
        "prg_progress_init": (None, []),
        "prg_progress_shutdown": (None, []),
        "prg_set_precision": (None, [i]),
        "prg_get_precision": (i, []),
        "prg_set_int8_moduli": (i, [i]),
        "prg_get_int8_moduli": (i, []),
        "prg_gpu_int8_waves": (i, []),
        "prg_gpu_set_int8_config": (i, [i]),
        "prg_gershgorin": (None, [vp, dp]),
        "prg_normalize": (None, [vp]),
        "prg_normalize_fermi": (None, [vp, d, d, d]),
        "prg_normalize_implicit_fermi": (None, [vp, d, d]),
        "prg_sp2_basic": (None, [vp, vp, d, d, i, i, cp, d, i]),
        "prg_sp2_alg2": (None, [vp, vp, d, d, i, i, cp, d, i]),
        "prg_sp2_basic_tcore": (None, [vp, vp, vp, d, d, i, i, cp, d, i]),
        "prg_sp2_tensorcore": (None, [i, C.POINTER(C.c_float), dp, C.c_float, C.c_float, i, i,
                                      C.c_char, C.c_float, i]),
        "prg_sp2_alg1": (None, [vp, vp, d, d, i, i, cp, d, i]),
        "prg_sp2_alg2_genseq": (None, [vp, vp, d, d, i, i, cp, d, ip, ip, dp, i]),
        "prg_sp2_alg1_genseq": (None, [vp, vp, d, d, i, i, cp, d, ip, ip, dp]),
        "prg_sp2_alg2_seq": (None, [vp, vp, d, ip, ip, dp, i]),
        "prg_sp2_alg1_seq": (None, [vp, vp, d, ip, ip, dp]),
        "prg_prg_sp2_alg2_seq_inplace": (None, [vp, d, i, ip, ip, dp, d, d]),
        "prg_prg_sp2_alg1_seq_inplace": (None, [vp, d, i, ip, ip, dp, d, d]),
        "prg_sp2_submatrix": (None, [vp, vp, d, ip, ip, dp, d, d, i]),
        "prg_sp2_submatrix_inplace": (None, [vp, d, ip, ip, dp, d, d, i]),
        "prg_homolumogap": (None, [dp, i, ip, d, d, dp, dp, dp, i]),
        "prg_sp2sequence": (None, [ip, ip, d, d, d, d, d, i]),
        "prg_gpu_sp2_alg2_device": (i, [i, vp, vp, d, i, i, cp, d, i, ip, ip]),
        "prg_gpu_padded_ld": (i, [i]),
        "prg_gpu_sp2_split_slices": (i, [i]),
        "prg_gpu_timing": (None, [i]),
        "prg_equalPartition": (None, [vp, i, i]),
        "prg_filePartition": (None, [vp, cp]),
        "prg_setPartition": (None, [vp, i, i, ip, ip]),
        "prg_destroyGraphPartitioning": (None, [vp]),
        "prg_getPartitionHalosFromGraph": (None, [vp, vp, i]),
        "prg_fnormGraph": (None, [vp]),
        "prg_subgraphSP2Loop": (None, [vp, vp, vp, vp, d]),
        "prg_bfsPartition": (None, [vp, vp, i]),
        "prg_sp2_fermi_init": (None, [vp, i, d, d, d, d, d, vp, dp, dp, dp, dp, ip, i]),
        "prg_sp2_fermi_init_norecs": (None, [vp, ip, d, d, d, d, d, vp, dp, dp, dp, dp, ip, i]),
        "prg_sp2_fermi": (None, [vp, i, i, d, dp, dp, dp, dp, ip, d, d, d, vp]),
        "prg_sp2_entropy_function": (None, [d, d, d, i, ip, dp, dp]),
        "prg_sp2_inverse": (d, [d, d, d, d, i, ip]),
        "prg_sp2_entropy_ts": (d, [vp, dp, dp]),
        "prg_orthogonalize": (None, [vp, vp, vp, d, cp, i]),
        "prg_deorthogonalize": (None, [vp, vp, vp, d, cp, i]),
        "prg_build_densityGP_T0": (None, [vp, vp, vp, d, d, d, i, i]),
        "prg_build_densityGP_T0_ef": (d, []),
        "prg_build_zmatGP": (None, [vp, vp, vp, d, i, i]),
        "prg_gpu_subgraph_timing": (None, [dp, dp, dp]),
        "prg_gpu_subgraph_config": (None, [ip, ip, ip, ip]),
        "prg_gpu_gershgorin_device": (i, [i, vp, dp]),
        "prg_gpu_sp2_seq_device": (i, [i, vp, vp, ip, i, d, d]),
        "prg_gpu_sp2_adjoint_device": (i, [i, vp, vp, vp, ip, i, d, d, i]),
        "prg_gpu_implicit_fermi_device": (i, [i, vp, vp, i, d, d, d]),
        "prg_gpu_implicit_fermi_adjoint_device": (i, [i, vp, vp, vp, i, d, d, d, i, vp]),
        "prg_build_density_T0": (None, [i, vp, vp, d, d, dp]),
        "prg_build_density_T_fermi": (None, [vp, vp, d, d, d, i, dp]),
        "prg_get_eigenvalues": (None, [i, vp, dp, i]),
        "prg_check_idempotency": (None, [vp, d, dp]),
        "prg_get_density_matrix": (None, [vp, vp, d, d, d]),
        "prg_test_density_matrix": (None, [vp, vp, d, d, d, i, d, d]),
        "prg_test_density_matrix_mu": (d, []),
        "prg_implicit_fermi": (None, [vp, vp, i, i, d, d, d, i, i, d, d, d]),
        "prg_implicit_fermi_last_mu": (d, []),
        "prg_implicit_fermi_save_inverse": (None, [C.POINTER(vp), vp, vp, i, d, dp, d, d, d, d,
                                                   i, i, i]),
        "prg_implicit_fermi_counters": (None, [ip, ip]),
        "prg_implicit_fermi_zero": (None, [vp, vp, i, d, i, d, d]),
        "prg_implicit_fermi_first_order_response": (None, [vp, vp, vp, vp, C.POINTER(vp), i, d,
                                                           d, d, d]),
        "prg_implicit_fermi_response": (None, [vp, vp, vp, vp, vp, vp, vp, vp, i, d, d, d, d,
                                               d, d, i, d]),
        "prg_implicit_fermi_response_mu": (None, [dp, dp]),
        "prg_gpu_implicit_fermi_response": (None, [vp, C.POINTER(vp), vp, C.POINTER(vp), i, dp,
                                                   dp, d, d, d, d, i, d]),
        "prg_compute_response_SP2": (None, [vp, vp, vp, vp, d, d, i, i, cp, d, d, i]),
        "prg_canon_response": (None, [vp, vp, d, d, dp, d, i, i]),
        "prg_canon_response_orig": (None, [vp, vp, d, d, dp, d, i, d, i]),
        "prg_pcg": (None, [vp, vp, vp, vp, vp, d, d]),
        "prg_finite_diff": (None, [vp, vp, d, d, d, i, d, d, d]),
        "prg_gpu_finite_diff": (None, [vp, C.POINTER(vp), d, dp, d, i, d, d, d, C.POINTER(vp)]),
        "prg_conjgrad": (None, [vp, vp, vp, vp, vp, vp, d, d]),
        "prg_newtonschulz": (None, [vp, vp, vp, vp, vp, vp, d, d, ip]),
        "prg_canon_response_p1_dpdmu": (None, [vp, dp, vp, i, d, vp, dp, d, i, i]),
        "prg_canon_response_vector": (None, [vp, vp, d, d, dp, d, i, d, i]),
        "prg_compute_response_RS": (None, [vp, vp, vp, d, d, d, i]),
        "prg_compute_response_FD": (None, [vp, vp, vp, d, d, d, i]),
        "prg_gpu_step_timing": (i, [dp, ip]),
        "prg_gpu_int8_product_timing": (i, [dp, ip]),
        "prg_gpu_gemm_counters": (None, [i, dp, dp, dp, dp]),
        "prg_set_product_precision": (None, [i]),
        "prg_gpu_symmetric_product_device": (i, [i, vp, vp, vp]),
        "prg_get_product_precision": (i, []),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _check():
    L = lib()
    code = L.prg_last_error_code()
    if code != 0:
        raise ProgressError(code, L.prg_last_error().decode())


def _free(ptr):
    libc = C.CDLL(None)
    libc.free.argtypes = [C.c_void_p]
    libc.free(ptr)


ELLPACK = 2  # bml_matrix_type_t ellpack (a sparse BML format)


class Matrix:
    """Owning wrapper of a library matrix handle (dense FP64)."""

    def __init__(self, handle):
        if not handle:
            _check()
            raise ProgressError(-1, "null matrix handle")
        self.h = C.c_void_p(handle)

    @classmethod
    def zeros(cls, n: int, matrix_type: int = None) -> "Matrix":
        h = lib().bml_zero_matrix(_dense if matrix_type is None else matrix_type, _double_real,
                                  n, n, _sequential)
        _check()
        return cls(h)

    @classmethod
    def from_numpy(cls, a: np.ndarray, matrix_type: int = None,
                   threshold: float = 0.0) -> "Matrix":
        """bml_import_from_dense; a sparse matrix_type (ellpack = 2, ...) keeps
        only the entries above `threshold`, as BML's sparse formats do."""
        a = np.asfortranarray(np.asarray(a, dtype=np.float64))
        n = a.shape[0]
        if a.shape != (n, n):
            raise ValueError("square matrix expected")
        h = lib().bml_import_from_dense(_dense if matrix_type is None else matrix_type,
                                        _double_real, _col_major, n, n,
                                        a.ctypes.data_as(C.c_void_p), threshold, _sequential)
        _check()
        return cls(h)

    @property
    def n(self) -> int:
        return lib().bml_get_N(self.h)

    def numpy(self) -> np.ndarray:
        n = self.n
        p = lib().bml_export_to_dense(self.h, _col_major)
        _check()
        buf = (C.c_double * (n * n)).from_address(p)
        out = np.frombuffer(buf, dtype=np.float64).copy().reshape((n, n), order="F")
        _free(p)
        return out

    def close(self):
        if self.h:
            lib().bml_deallocate(C.byref(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def set_precision(mode: int):
    lib().prg_set_precision(mode)
    _check()


def get_precision() -> int:
    return int(lib().prg_get_precision())


def set_int8_moduli(count: int):
    """Number of moduli of the int8 Chinese-remainder products (6..14)."""
    lib().prg_set_int8_moduli(count)
    _check()


def get_int8_moduli() -> int:
    return int(lib().prg_get_int8_moduli())


def set_int8_config(config: int):
    """Force the int8 CRT product kernel configuration (803, 804, 805, 904, 944;
    0 = time them per size)."""
    st = lib().prg_gpu_set_int8_config(config)
    if st != 0:
        raise ProgressError(st, lib().prg_last_error().decode())


def sp2_split_slices(n: int) -> int:
    """K slices of the FP64 SP2 step at size n (1: the fused one-kernel step)."""
    return int(lib().prg_gpu_sp2_split_slices(n))


def int8_waves() -> int:
    """Workgroup layout (4 or 8 waves) of the last int8 CRT product."""
    return int(lib().prg_gpu_int8_waves())


def _ivec(n, init=None):
    a = np.zeros(max(n, 1), dtype=np.int32)
    if init is not None:
        a[:len(init)] = init
    return a


def _dvec(n, init=None):
    a = np.zeros(max(n, 1), dtype=np.float64)
    if init is not None:
        a[:len(init)] = init
    return a


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def gershgorin(h: np.ndarray):
    m = Matrix.from_numpy(h)
    g = _dvec(2)
    lib().prg_gershgorin(m.h, _dp(g))
    _check()
    return float(g[0]), float(g[1])


def _h_rho(fn, h, *args, matrix_type=None):
    """H in, rho out; matrix_type (ellpack = 2, ...) stores both handles in a
    sparse BML format, whose thresholding the SP2 routines then apply."""
    mh = Matrix.from_numpy(h, matrix_type)
    mr = Matrix.zeros(h.shape[0], matrix_type)
    fn(mh.h, mr.h, *args)
    _check()
    return mr.numpy()


def sp2_basic(h, bndfil, minsp2iter, maxsp2iter, sp2conv, idemtol, threshold=0.0, verbose=0):
    """prg_sp2_basic (src/prg_sp2_mod.F90:49-135)."""
    return _h_rho(lib().prg_sp2_basic, h, threshold, bndfil, minsp2iter, maxsp2iter,
                  sp2conv.encode(), idemtol, verbose)


def sp2_basic_tcore(h, bndfil, minsp2iter, maxsp2iter, sp2conv, idemtol, rhofull=None,
                    threshold=0.0, verbose=0):
    """prg_sp2_basic_tcore (src/prg_sp2_mod.F90:138-246): maxsp2iter + 1
    iterations of the basic branch rule, no convergence exit."""
    n = h.shape[0]
    mh = Matrix.from_numpy(h)
    rho = Matrix.zeros(n)
    mf = Matrix.from_numpy(rhofull) if rhofull is not None else None
    lib().prg_sp2_basic_tcore(mh.h, rho.h, mf.h if mf is not None else None, threshold, bndfil,
                              minsp2iter, maxsp2iter, sp2conv.encode(), idemtol, verbose)
    _check()
    return rho.numpy()


def sp2_tensorcore(h, bndfil, minsp2iter=1, maxsp2iter=100, sp2conv=b"R", idemtol=1e-5,
                   eps=1e-5, verbose=0):
    """prg_sp2_tensorcore (src/sp2_tensorcore/src/prg_sp2_tensorcore.cu:131-141):
    float32 H in, float64 D = X (not 2X) out, FP16 split products."""
    hf = np.asfortranarray(np.asarray(h, dtype=np.float32))
    n = hf.shape[0]
    out = np.zeros((n, n), dtype=np.float64, order="F")
    lib().prg_sp2_tensorcore(n, hf.ctypes.data_as(C.POINTER(C.c_float)),
                             out.ctypes.data_as(C.POINTER(C.c_double)), eps, bndfil, minsp2iter,
                             maxsp2iter, sp2conv, idemtol, verbose)
    _check()
    return out


def sp2_alg2(h, bndfil, minsp2iter, maxsp2iter, sp2conv, idemtol, threshold=0.0, verbose=0,
             matrix_type=None):
    """prg_sp2_alg2 (src/prg_sp2_mod.F90:263-380)."""
    return _h_rho(lib().prg_sp2_alg2, h, threshold, bndfil, minsp2iter, maxsp2iter,
                  sp2conv.encode(), idemtol, verbose, matrix_type=matrix_type)


def sp2_alg1(h, bndfil, minsp2iter, maxsp2iter, sp2conv, idemtol, threshold=0.0, verbose=0,
             matrix_type=None):
    """prg_sp2_alg1 (src/prg_sp2_mod.F90:734-864)."""
    return _h_rho(lib().prg_sp2_alg1, h, threshold, bndfil, minsp2iter, maxsp2iter,
                  sp2conv.encode(), idemtol, verbose, matrix_type=matrix_type)


def sp2_alg2_genseq(h, bndfil, minsp2iter, maxsp2iter, sp2conv, idemtol, threshold=0.0,
                    verbose=0):
    """prg_sp2_alg2_genseq: returns (rho, pp, icount, vv) with vv[:icount] and,
    when the loop ended on a discarded product, vv[icount]."""
    pp, vv, ic = _ivec(maxsp2iter), _dvec(maxsp2iter), _ivec(1)
    rho = _h_rho(lib().prg_sp2_alg2_genseq, h, threshold, bndfil, minsp2iter, maxsp2iter,
                 sp2conv.encode(), idemtol, _ip(pp), _ip(ic), _dp(vv), verbose)
    icount = int(ic[0])
    return rho, pp[:icount].tolist(), icount, vv


def sp2_alg1_genseq(h, bndfil, minsp2iter, maxsp2iter, sp2conv, idemtol, threshold=0.0):
    """prg_sp2_alg1_genseq: returns (rho, pp, icount, vv)."""
    pp, vv, ic = _ivec(maxsp2iter), _dvec(maxsp2iter), _ivec(1)
    rho = _h_rho(lib().prg_sp2_alg1_genseq, h, threshold, bndfil, minsp2iter, maxsp2iter,
                 sp2conv.encode(), idemtol, _ip(pp), _ip(ic), _dp(vv))
    icount = int(ic[0])
    return rho, pp[:icount].tolist(), icount, vv


def sp2_alg2_seq(h, pp, icount, threshold=0.0):
    """prg_sp2_alg2_seq: returns (rho, vv)."""
    ppa, vv, ic = _ivec(icount, pp), _dvec(icount), _ivec(1, [icount])
    rho = _h_rho(lib().prg_sp2_alg2_seq, h, threshold, _ip(ppa), _ip(ic), _dp(vv), 0)
    return rho, vv[:icount].tolist()


def sp2_alg1_seq(h, pp, icount, threshold=0.0):
    """prg_sp2_alg1_seq: returns (rho, vv)."""
    ppa, vv, ic = _ivec(icount, pp), _dvec(icount), _ivec(1, [icount])
    rho = _h_rho(lib().prg_sp2_alg1_seq, h, threshold, _ip(ppa), _ip(ic), _dp(vv))
    return rho, vv[:icount].tolist()


def _inplace(fn, x, pp, icount, mineval, maxeval, threshold):
    m = Matrix.from_numpy(x)
    ppa, vv, ic = _ivec(icount, pp), _dvec(icount), _ivec(1, [icount])
    fn(m.h, threshold, icount, _ip(ppa), _ip(ic), _dp(vv), mineval, maxeval)
    _check()
    return m.numpy(), vv[:icount].tolist()


def sp2_alg2_seq_inplace(x, pp, icount, mineval, maxeval, threshold=0.0):
    """prg_prg_sp2_alg2_seq_inplace: returns (rho, vv)."""
    return _inplace(lib().prg_prg_sp2_alg2_seq_inplace, x, pp, icount, mineval, maxeval,
                    threshold)


def sp2_alg1_seq_inplace(x, pp, icount, mineval, maxeval, threshold=0.0):
    """prg_prg_sp2_alg1_seq_inplace: returns (rho, vv)."""
    return _inplace(lib().prg_prg_sp2_alg1_seq_inplace, x, pp, icount, mineval, maxeval,
                    threshold)


def sp2_submatrix(x, pp, icount, vv, mineval, maxeval, core_size, threshold=0.0):
    """prg_sp2_submatrix_inplace: returns (rho, vv) with vv accumulated."""
    m = Matrix.from_numpy(x)
    ppa, vva, ic = _ivec(icount, pp), _dvec(icount, vv), _ivec(1, [icount])
    lib().prg_sp2_submatrix_inplace(m.h, threshold, _ip(ppa), _ip(ic), _dp(vva), mineval,
                                    maxeval, core_size)
    _check()
    return m.numpy(), vva[:icount].tolist()


def homolumogap(vv, imax, pp, mineval, maxeval):
    """prg_homolumogap: returns (ehomo, elumo, egap)."""
    vva, ppa = _dvec(len(vv), vv), _ivec(len(pp), pp)
    out = _dvec(3)
    o = [C.c_double() for _ in range(3)]
    lib().prg_homolumogap(_dp(vva), imax, _ip(ppa), mineval, maxeval, C.byref(o[0]),
                          C.byref(o[1]), C.byref(o[2]), 0)
    _check()
    del out
    return o[0].value, o[1].value, o[2].value


def sp2sequence(mineval, maxeval, ehomo, elumo, errlimit):
    """prg_sp2sequence: returns (pp, imax)."""
    pp, im = _ivec(100), _ivec(1)
    lib().prg_sp2sequence(_ip(pp), _ip(im), mineval, maxeval, ehomo, elumo, errlimit, 0)
    _check()
    return pp[:int(im[0])].tolist(), int(im[0])


def padded_ld(n: int) -> int:
    return lib().prg_gpu_padded_ld(n)


def sp2_alg2_device(dH, dRho, n, bndfil, minsp2iter, maxsp2iter, sp2conv, idemtol,
                    precision=PREC_FP64):
    """Device-pointer prg_sp2_alg2: dH, dRho are torch CUDA tensors of shape
    (ld, ld), ld = padded_ld(n), zero outside the n x n block.  Returns
    (iterations, pp).  Synchronises torch's stream before the call."""
    import torch
    torch.cuda.synchronize()
    pp = _ivec(maxsp2iter)
    it = _ivec(1)
    st = lib().prg_gpu_sp2_alg2_device(n, C.c_void_p(dH.data_ptr()), C.c_void_p(dRho.data_ptr()),
                                       bndfil, minsp2iter, maxsp2iter, sp2conv.encode(), idemtol,
                                       precision, _ip(it), _ip(pp))
    if st != 0:
        raise ProgressError(st, lib().prg_last_error().decode())
    return int(it[0]), pp[:int(it[0])].tolist()


def step_timing(enable=None):
    """Enable/reset (enable=True/False) or read (enable=None) the product-kernel
    timers: returns (ms_sum, count)."""
    if enable is not None:
        lib().prg_gpu_timing(1 if enable else 0)
        return None
    ms = C.c_double()
    cnt = C.c_int()
    lib().prg_gpu_step_timing(C.byref(ms), C.byref(cnt))
    return ms.value, cnt.value


def gemm_counters(reset=False, int8=False):
    """(products, algorithmic flops, executed FP64 MFMA flops) of the products
    the iterative solvers issued since the last reset; with int8=True also the
    int8 MFMA operations of their int8 CRT products."""
    v = [C.c_double(), C.c_double(), C.c_double(), C.c_double()]
    lib().prg_gpu_gemm_counters(1 if reset else 0, *[C.byref(x) for x in v])
    out = tuple(x.value for x in v)
    return out if int8 else out[:3]


def set_product_precision(mode):
    """Arithmetic of the symmetric products of commuting symmetric matrices in
    implicit Fermi, SP2-Fermi and the adjoint: PREC_FP64, PREC_INT8CRT or
    PREC_BF16SPLIT (about 16 significant bits: solver tolerances must sit
    above that noise floor)."""
    lib().prg_set_product_precision(int(mode))
    _check()


def get_product_precision():
    return int(lib().prg_get_product_precision())


def symmetric_product_device(a, b, c, n):
    """c = a b for commuting symmetric a, b (torch float64 cuda tensors in the
    padded layout, padded_ld(n) square), in the current product precision."""
    for t in (a, b, c):
        if not (t.is_cuda and t.dtype.is_floating_point and t.element_size() == 8 and
                t.is_contiguous() and t.shape == (padded_ld(n), padded_ld(n))):
            raise ValueError("symmetric_product_device: padded contiguous float64 cuda tensors")
    lib().prg_gpu_symmetric_product_device(n, C.c_void_p(a.data_ptr()), C.c_void_p(b.data_ptr()),
                                           C.c_void_p(c.data_ptr()))
    _check()


def int8_product_timing():
    """(ms_sum, count) of the int8 CRT product kernel alone, for the steps
    timed since step_timing(True)."""
    ms = C.c_double()
    cnt = C.c_int()
    lib().prg_gpu_int8_product_timing(C.byref(ms), C.byref(cnt))
    return ms.value, cnt.value


# ---------------------------------------------------------------- eigensolver densities

def build_density_T0(h, bndfil):
    """prg_build_density_T0: returns (rho, eigenvalues)."""
    n = h.shape[0]
    w = _dvec(n)
    rho = _h_rho(lambda a, b: lib().prg_build_density_T0(n, a, b, 0.0, bndfil, _dp(w)), h)
    return rho, w


def build_density_T_fermi(h, kbt, ef):
    """prg_build_density_T_fermi: returns (rho, drho)."""
    dr = C.c_double()
    rho = _h_rho(lambda a, b: lib().prg_build_density_T_fermi(a, b, 0.0, kbt, ef, 0,
                                                              C.byref(dr)), h)
    return rho, dr.value


def get_eigenvalues(h):
    n = h.shape[0]
    m = Matrix.from_numpy(h)
    w = _dvec(n)
    lib().prg_get_eigenvalues(n, m.h, _dp(w), 0)
    _check()
    return w


def check_idempotency(a):
    m = Matrix.from_numpy(a)
    v = C.c_double()
    lib().prg_check_idempotency(m.h, 0.0, C.byref(v))
    _check()
    return v.value


def get_density_matrix(h, beta, mu):
    return _h_rho(lambda a, b: lib().prg_get_density_matrix(a, b, beta, mu, 0.0), h)


def test_density_matrix(h, beta, mu, nocc, osteps, occ_err_limit):
    """prg_test_density_matrix: returns (P, mu)."""
    p = _h_rho(lambda a, b: lib().prg_test_density_matrix(a, b, beta, mu, nocc, osteps,
                                                          occ_err_limit, 0.0), h)
    return p, lib().prg_test_density_matrix_mu()


# ---------------------------------------------------------------- implicit Fermi

def implicit_fermi(h, nsteps, k, nocc, mu, beta, method, osteps, occ_err_limit, tol):
    """prg_implicit_fermi: returns (P, mu)."""
    p = _h_rho(lambda a, b: lib().prg_implicit_fermi(a, b, nsteps, k, nocc, mu, beta, method,
                                                     osteps, occ_err_limit, 0.0, tol), h)
    return p, lib().prg_implicit_fermi_last_mu()


def _handles(mats):
    arr = (C.c_void_p * len(mats))(*[m.h.value for m in mats])
    return arr


def implicit_fermi_save_inverse(inv, h, nsteps, nocc, mu, beta, occ_err_limit, tol, scf_it,
                                occiter=0, totns=0):
    """prg_implicit_fermi_save_inverse.  `inv` is a list of nsteps Matrix
    handles (updated in place on the device).  Returns (P, mu, occiter, totns)."""
    mu_c = C.c_double(mu)
    hh = _handles(inv)
    p = _h_rho(lambda a, b: lib().prg_implicit_fermi_save_inverse(
        hh, a, b, nsteps, nocc, C.byref(mu_c), beta, occ_err_limit, 0.0, tol, scf_it, occiter,
        totns), h)
    oc, tn = C.c_int(), C.c_int()
    lib().prg_implicit_fermi_counters(C.byref(oc), C.byref(tn))
    return p, mu_c.value, oc.value, tn.value


def implicit_fermi_zero(h, nsteps, mu, method, tol):
    return _h_rho(lambda a, b: lib().prg_implicit_fermi_zero(a, b, nsteps, mu, method, 0.0, tol),
                  h)


def implicit_fermi_first_order_response(h0, h1, inv, nsteps, mu0, beta, nocc):
    """Returns (P0, P1); `inv` is the list of Matrix handles of the inverses."""
    n = h0.shape[0]
    m0, m1 = Matrix.from_numpy(h0), Matrix.from_numpy(h1)
    p0, p1 = Matrix.zeros(n), Matrix.zeros(n)
    lib().prg_implicit_fermi_first_order_response(m0.h, m1.h, p0.h, p1.h, _handles(inv), nsteps,
                                                  mu0, beta, nocc, 0.0)
    _check()
    return p0.numpy(), p1.numpy()


def implicit_fermi_response(h0, hs, nsteps, mu0, beta, nocc, occ_tol, order, lin_tol=1e-10):
    """Returns (P0, [P1..], mu0, [mu1..])."""
    n = h0.shape[0]
    m0 = Matrix.from_numpy(h0)
    mh = [Matrix.from_numpy(x) for x in hs[:order]]
    p0 = Matrix.zeros(n)
    ps = [Matrix.zeros(n) for _ in range(order)]
    mu0_c = C.c_double(mu0)
    mu = _dvec(order)
    lib().prg_gpu_implicit_fermi_response(m0.h, _handles(mh), p0.h, _handles(ps), nsteps,
                                          C.byref(mu0_c), _dp(mu), beta, nocc, occ_tol, lin_tol,
                                          order, 0.0)
    _check()
    return p0.numpy(), [p.numpy() for p in ps], mu0_c.value, mu.tolist()


def implicit_fermi_response_c(h0, hs, nsteps, mu0, beta, nocc, occ_tol, order, lin_tol=1e-10):
    """The same through the reference header's arity (prg_implicit_fermi_response
    with separate H1..H3 / P1..P3 handles and mu0 / mu by value), the updated
    chemical potentials read back with prg_implicit_fermi_response_mu.
    Returns (P0, [P1..], mu0, [mu1..])."""
    n = h0.shape[0]
    m0 = Matrix.from_numpy(h0)
    mh = [Matrix.from_numpy(x) for x in hs[:order]]
    p0 = Matrix.zeros(n)
    ps = [Matrix.zeros(n) for _ in range(order)]
    hv = [m.h for m in mh] + [None] * (3 - order)
    pv = [m.h for m in ps] + [None] * (3 - order)
    lib().prg_implicit_fermi_response(m0.h, hv[0], hv[1], hv[2], p0.h, pv[0], pv[1], pv[2],
                                      nsteps, mu0, 0.0, beta, nocc, occ_tol, lin_tol, order, 0.0)
    _check()
    m0_out = C.c_double()
    mu = _dvec(3)
    lib().prg_implicit_fermi_response_mu(C.byref(m0_out), _dp(mu))
    return p0.numpy(), [p.numpy() for p in ps], m0_out.value, mu[:order].tolist()


# ---------------------------------------------------------------- response

def compute_response_sp2(h, h1, bndfil, minsp2iter, maxsp2iter, sp2conv, idemtol):
    """prg_compute_response_SP2: returns (rho, rho1)."""
    n = h.shape[0]
    mh, m1 = Matrix.from_numpy(h), Matrix.from_numpy(h1)
    rsp, rho = Matrix.zeros(n), Matrix.zeros(n)
    lib().prg_compute_response_SP2(mh.h, m1.h, rsp.h, rho.h, 0.0, bndfil, minsp2iter,
                                   maxsp2iter, sp2conv.encode(), idemtol, 0.0, 0)
    _check()
    return rho.numpy(), rsp.numpy()


def canon_response(h1_eig, nocc, beta, evals, mu0, m):
    n = h1_eig.shape[0]
    mh = Matrix.from_numpy(h1_eig)
    p1 = Matrix.zeros(n)
    e = _dvec(n, evals)
    lib().prg_canon_response(p1.h, mh.h, nocc, beta, _dp(e), mu0, m, n)
    _check()
    return p1.numpy()


def finite_diff(h0, hs, mu0, mus, beta, order, lam, h):
    """prg_finite_diff: returns the list of the order 1..`order` difference
    matrices (the reference prints them)."""
    n = h0.shape[0]
    m0 = Matrix.from_numpy(h0)
    mh = [Matrix.from_numpy(x) for x in hs[:order]]
    outs = [Matrix.zeros(n) for _ in range(order)]
    mu = _dvec(order, mus[:order])
    lib().prg_gpu_finite_diff(m0.h, _handles(mh), mu0, _dp(mu), beta, order, lam, h, 0.0,
                              _handles(outs))
    _check()
    return [o.numpy() for o in outs]


def pcg(a, x0, b, cg_tol):
    """prg_pcg: returns (X, R) with R the residual left in B's matrix."""
    ma, mx, mb = Matrix.from_numpy(a), Matrix.from_numpy(x0), Matrix.from_numpy(b)
    lib().prg_pcg(ma.h, mx.h, mb.h, None, None, cg_tol, 0.0)
    _check()
    return mx.numpy(), mb.numpy()


def conjgrad(a, x0, b, cg_tol):
    """prg_conjgrad: returns X."""
    ma, mx, mb = Matrix.from_numpy(a), Matrix.from_numpy(x0), Matrix.from_numpy(b)
    lib().prg_conjgrad(ma.h, mx.h, mb.h, None, None, None, cg_tol, 0.0)
    _check()
    return mx.numpy()


def newtonschulz(a, ai0, tol):
    """prg_newtonschulz: returns (Ai, iterations)."""
    ma, mi = Matrix.from_numpy(a), Matrix.from_numpy(ai0)
    it = C.c_int()
    lib().prg_newtonschulz(ma.h, mi.h, None, None, None, None, tol, 0.0, C.byref(it))
    _check()
    return mi.numpy(), it.value


def canon_response_orig(h1_eig, nocc, beta, evals, mu0, m):
    n = h1_eig.shape[0]
    mh, p1 = Matrix.from_numpy(h1_eig), Matrix.zeros(n)
    e = _dvec(n, evals)
    lib().prg_canon_response_orig(p1.h, mh.h, nocc, beta, _dp(e), mu0, m, 0.0, n)
    _check()
    return p1.numpy()


def canon_response_p1_dpdmu(h1_eig, beta, evals, mu0, m):
    """Returns (P1, dPdmu)."""
    n = h1_eig.shape[0]
    mh, p1 = Matrix.from_numpy(h1_eig), Matrix.zeros(n)
    e, dp = _dvec(n, evals), _dvec(n)
    lib().prg_canon_response_p1_dpdmu(p1.h, _dp(dp), mh.h, 0, beta, None, _dp(e), mu0, m, n)
    _check()
    return p1.numpy(), np.array(dp[:n])


def canon_response_vector(h1_eig, beta, evals, mu0):
    n = h1_eig.shape[0]
    mh, p1 = Matrix.from_numpy(h1_eig), Matrix.zeros(n)
    e = _dvec(n, evals)
    lib().prg_canon_response_vector(p1.h, mh.h, 0.0, beta, _dp(e), mu0, 0, 0.0, n)
    _check()
    return p1.numpy()


def compute_response_rs(h, h1, bndfil):
    n = h.shape[0]
    m0, m1, r = Matrix.from_numpy(h), Matrix.from_numpy(h1), Matrix.zeros(n)
    lib().prg_compute_response_RS(m0.h, m1.h, r.h, 0.0, bndfil, 0.0, 0)
    _check()
    return r.numpy()


def compute_response_fd(h, h1, delta, bndfil):
    n = h.shape[0]
    m0, m1, r = Matrix.from_numpy(h), Matrix.from_numpy(h1), Matrix.zeros(n)
    lib().prg_compute_response_FD(m0.h, m1.h, r.h, delta, bndfil, 0.0, 0)
    _check()
    return r.numpy()


# ---------------------------------------------------------------- device-pointer helpers

def _st(code):
    if code != 0:
        raise ProgressError(code, lib().prg_last_error().decode())


def gershgorin_device(dH, n):
    import torch
    torch.cuda.synchronize()
    b = _dvec(2)
    _st(lib().prg_gpu_gershgorin_device(n, C.c_void_p(dH.data_ptr()), _dp(b)))
    return float(b[0]), float(b[1])


def sp2_seq_device(dH, dRho, n, pp, mineval, maxeval):
    import torch
    torch.cuda.synchronize()
    ppa = _ivec(len(pp), pp)
    _st(lib().prg_gpu_sp2_seq_device(n, C.c_void_p(dH.data_ptr()), C.c_void_p(dRho.data_ptr()),
                                     _ip(ppa), len(pp), mineval, maxeval))


def sp2_adjoint_device(dH, dGout, dGrad, n, pp, mineval, maxeval, ckpt=8):
    import torch
    torch.cuda.synchronize()
    ppa = _ivec(len(pp), pp)
    _st(lib().prg_gpu_sp2_adjoint_device(n, C.c_void_p(dH.data_ptr()),
                                         C.c_void_p(dGout.data_ptr()),
                                         C.c_void_p(dGrad.data_ptr()), _ip(ppa), len(pp),
                                         mineval, maxeval, ckpt))


def implicit_fermi_device(dH, dP, n, nsteps, mu, beta, tol=1e-10):
    """prg_implicit_fermi at fixed mu (method 1) on padded float64 cuda tensors:
    dP = P (not scaled by 2)."""
    import torch
    torch.cuda.synchronize()
    _st(lib().prg_gpu_implicit_fermi_device(n, C.c_void_p(dH.data_ptr()),
                                            C.c_void_p(dP.data_ptr()), nsteps, mu, beta, tol))


def implicit_fermi_adjoint_device(dH, dA, dG, n, nsteps, mu, beta, tol=1e-10, canonical=False,
                                  dP=None):
    """dG = dTr(P A)/dH of the fixed-mu implicit-Fermi map (canonical: at fixed
    occupation), padded float64 cuda tensors; dP (optional) receives P."""
    import torch
    torch.cuda.synchronize()
    _st(lib().prg_gpu_implicit_fermi_adjoint_device(
        n, C.c_void_p(dH.data_ptr()), C.c_void_p(dA.data_ptr()), C.c_void_p(dG.data_ptr()),
        nsteps, mu, beta, tol, 1 if canonical else 0,
        C.c_void_p(dP.data_ptr()) if dP is not None else None))


# ---------------------------------------------------------------- graph partitioning

class GraphPartitioning(C.Structure):
    """prg_graph_partitioning_t (include/progress_gpu.h)."""
    _fields_ = [("totalParts", C.c_int), ("totalNodes", C.c_int),
                ("nnodesInPart", C.POINTER(C.c_int)),
                ("nodeInPart", C.POINTER(C.POINTER(C.c_int))),
                ("lsize", C.POINTER(C.c_int)), ("llsize", C.POINTER(C.c_int)),
                ("core_halo_index", C.POINTER(C.POINTER(C.c_int))),
                ("pp", C.c_int * 100), ("maxIter", C.c_int),
                ("ehomo", C.c_double), ("elumo", C.c_double),
                ("mineval", C.c_double), ("maxeval", C.c_double),
                ("vv", C.c_double * 100)]

    def __init__(self):
        super().__init__()

    def parts(self):
        return [[self.nodeInPart[i][j] for j in range(self.nnodesInPart[i])]
                for i in range(self.totalParts)]

    def halos(self):
        return [[self.core_halo_index[i][j] for j in range(self.lsize[i])]
                for i in range(self.totalParts)]

    def destroy(self):
        lib().prg_destroyGraphPartitioning(C.byref(self))

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def equal_partition(nodes_per_part, nnodes):
    gp = GraphPartitioning()
    lib().prg_equalPartition(C.byref(gp), nodes_per_part, nnodes)
    _check()
    return gp


def file_partition(path):
    gp = GraphPartitioning()
    lib().prg_filePartition(C.byref(gp), str(path).encode())
    _check()
    return gp


def set_partition(parts, nnodes):
    gp = GraphPartitioning()
    sizes = np.array([len(p) for p in parts], dtype=np.int32)
    nodes = np.concatenate([np.asarray(p, dtype=np.int32) for p in parts])
    lib().prg_setPartition(C.byref(gp), len(parts), nnodes, _ip(sizes), _ip(nodes))
    _check()
    return gp


def subgraph_sp2_loop(h, g, gp, pp, max_iter, mineval, maxeval, vv=None, out=None):
    """prg_subgraphSP2Loop: returns (rho, vv[:max_iter]); gp is filled with
    the core+halo lists.  h and g are numpy arrays or Matrix handles (whose
    device copies stay resident across calls); with `out` (a Matrix) the
    density stays in that handle and is returned as the handle."""
    mh = h if isinstance(h, Matrix) else Matrix.from_numpy(h)
    mg = g if isinstance(g, Matrix) else Matrix.from_numpy(g)
    rho = out if out is not None else Matrix.zeros(mh.n)
    gp.maxIter = max_iter
    for k, v in enumerate(pp):
        gp.pp[k] = int(v)
    for k in range(100):
        gp.vv[k] = 0.0 if vv is None or k >= len(vv) else float(vv[k])
    gp.mineval, gp.maxeval = mineval, maxeval
    lib().prg_subgraphSP2Loop(mh.h, mg.h, rho.h, C.byref(gp), 0.0)
    _check()
    return (rho if out is not None else rho.numpy()), [gp.vv[k] for k in range(max_iter)]


def sp2_fermi_init(h, nsteps, nocc, tscale, occ_err_limit, trace_limit, norecs=False,
                   beta=0.0, threshold=0.0):
    """prg_sp2_fermi_init (norecs=False) or prg_sp2_fermi_init_norecs (norecs=True,
    beta = target inverse temperature).  Returns
    dict(x, mu, beta, h1, hN, sgnlist, nsteps)."""
    n = h.shape[0]
    mh = Matrix.from_numpy(h)
    mx = Matrix.zeros(n)
    mu, be, h1, hn = _dvec(1), _dvec(1, [beta]), _dvec(1), _dvec(1)
    sg = _ivec(max(nsteps, 1))
    if norecs:
        ns = _ivec(1, [nsteps])
        lib().prg_sp2_fermi_init_norecs(mh.h, _ip(ns), nocc, tscale, threshold, occ_err_limit,
                                        trace_limit, mx.h, _dp(mu), _dp(be), _dp(h1), _dp(hn),
                                        _ip(sg), 0)
        nsteps = int(ns[0])
    else:
        lib().prg_sp2_fermi_init(mh.h, nsteps, nocc, tscale, threshold, occ_err_limit,
                                 trace_limit, mx.h, _dp(mu), _dp(be), _dp(h1), _dp(hn), _ip(sg), 0)
    _check()
    return {"x": mx.numpy(), "mu": float(mu[0]), "beta": float(be[0]), "h1": float(h1[0]),
            "hN": float(hn[0]), "sgnlist": sg[:nsteps].tolist(), "nsteps": nsteps}


def sp2_fermi(h, osteps, nsteps, nocc, mu, beta, h1, hN, sgnlist, eps, trace_limit,
              threshold=0.0):
    """prg_sp2_fermi: returns (x, mu)."""
    n = h.shape[0]
    mh = Matrix.from_numpy(h)
    mx = Matrix.zeros(n)
    m, be, a, b = _dvec(1, [mu]), _dvec(1, [beta]), _dvec(1, [h1]), _dvec(1, [hN])
    sg = _ivec(max(nsteps, 1), list(sgnlist[:nsteps]) + [0] * max(0, 1 - nsteps))
    lib().prg_sp2_fermi(mh.h, osteps, nsteps, nocc, _dp(m), _dp(be), _dp(a), _dp(b), _ip(sg),
                        threshold, eps, trace_limit, mx.h)
    _check()
    return mx.numpy(), float(m[0])


def sp2_entropy_function(mu, h1, hN, sgnlist):
    """prg_sp2_entropy_function: returns (GG, ee), 10001 points each."""
    sg = _ivec(max(len(sgnlist), 1), list(sgnlist) or [0])
    gg, ee = _dvec(10001), _dvec(10001)
    lib().prg_sp2_entropy_function(mu, h1, hN, len(sgnlist), _ip(sg), _dp(gg), _dp(ee))
    _check()
    return gg, ee


def sp2_inverse(f, mu, h1, hN, sgnlist):
    sg = _ivec(max(len(sgnlist), 1), list(sgnlist) or [0])
    return float(lib().prg_sp2_inverse(f, mu, h1, hN, len(sgnlist), _ip(sg)))


def sp2_entropy_ts(d0, gg, ee):
    md = Matrix.from_numpy(d0)
    ts = float(lib().prg_sp2_entropy_ts(md.h, _dp(np.ascontiguousarray(gg, dtype=np.float64)),
                                        _dp(np.ascontiguousarray(ee, dtype=np.float64))))
    _check()
    return ts


def orthogonalize(a, z):
    """prg_orthogonalize: Z^T A Z."""
    ma, mz = Matrix.from_numpy(a), Matrix.from_numpy(z)
    mo = Matrix.zeros(a.shape[0])
    lib().prg_orthogonalize(ma.h, mz.h, mo.h, 0.0, b"dense", 0)
    _check()
    return mo.numpy()


def deorthogonalize(a_ortho, z):
    """prg_deorthogonalize: Z A_ortho Z^T."""
    mo, mz = Matrix.from_numpy(a_ortho), Matrix.from_numpy(z)
    ma = Matrix.zeros(a_ortho.shape[0])
    lib().prg_deorthogonalize(mo.h, mz.h, ma.h, 0.0, b"dense", 0)
    _check()
    return ma.numpy()


def bfs_partition(g, nparts):
    """The build's graph partitioner (in place of METIS): returns a
    GraphPartitioning with nparts cores (ascending node ids)."""
    gp = GraphPartitioning()
    mg = Matrix.from_numpy(g)
    lib().prg_bfsPartition(C.byref(gp), mg.h, nparts)
    _check()
    return gp


def build_density_gp_t0(h, g, bndfil, ef, nparts, threshold=0.0, verbose=0, out=None):
    """prg_build_densityGP_T0: returns (rho, updated Ef).  h and g are numpy
    arrays or Matrix handles; with `out` (a Matrix) rho stays in that handle."""
    mh = h if isinstance(h, Matrix) else Matrix.from_numpy(h)
    mg = g if isinstance(g, Matrix) else Matrix.from_numpy(g)
    rho = out if out is not None else Matrix.zeros(mh.n)
    lib().prg_build_densityGP_T0(mh.h, mg.h, rho.h, threshold, bndfil, ef, nparts, verbose)
    _check()
    return (rho if out is not None else rho.numpy()), float(lib().prg_build_densityGP_T0_ef())


def build_zmat_gp(s, g, nparts, threshold=0.0, verbose=0, out=None):
    """prg_build_zmatGP: the graph-partitioned congruence transform Z ~ S^-1/2.
    s and g are numpy arrays or Matrix handles; with `out` (a Matrix, whose
    entries outside the parts' core rows are kept) Z stays in that handle."""
    ms = s if isinstance(s, Matrix) else Matrix.from_numpy(s)
    mg = g if isinstance(g, Matrix) else Matrix.from_numpy(g)
    z = out if out is not None else Matrix.zeros(ms.n)
    lib().prg_build_zmatGP(ms.h, mg.h, z.h, threshold, nparts, verbose)
    _check()
    return z if out is not None else z.numpy()


def subgraph_timing():
    a, b, c = C.c_double(), C.c_double(), C.c_double()
    lib().prg_gpu_subgraph_timing(C.byref(a), C.byref(b), C.byref(c))
    return a.value, b.value, c.value


def subgraph_config():
    """(panel width, prefetch slabs, slab depth, launch) of the last subgraph
    loop: launch 0 eager, 1 captured into a hipGraph, 2 replayed graph."""
    a, b, c, d = C.c_int(), C.c_int(), C.c_int(), C.c_int()
    lib().prg_gpu_subgraph_config(C.byref(a), C.byref(b), C.byref(c), C.byref(d))
    return a.value, b.value, c.value, d.value

"""The C-ABI library loads and exports every function include/*.h declares
(CPU only: no compute call is made)."""
import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
LIB = ROOT / "qmd-progress_amd" / "lib" / "libprogress_gpu.so"


def declared_functions():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        for m in re.finditer(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", text):
            name = m.group(1)
            if name.startswith(("prg_", "bml_")):
                names.add(name)
    return sorted(names)


def test_header_declares_reference_entry_points():
    names = declared_functions()
    for must in ["prg_sp2_alg2", "prg_sp2_alg1", "prg_sp2_basic", "prg_sp2_alg2_genseq",
                 "prg_sp2_submatrix_inplace", "prg_normalize", "prg_homolumogap",
                 "prg_sp2sequence", "bml_zero_matrix", "bml_deallocate"]:
        assert must in names


@pytest.mark.skipif(not LIB.exists(), reason="library not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(LIB))
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.skipif(not LIB.exists(), reason="library not built")
def test_no_device_is_a_loud_error_not_a_fallback(monkeypatch):
    """Without a GPU the library refuses to compute (no CPU fallback)."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    import progress_amd as P
    import numpy as np
    with pytest.raises(P.ProgressError):
        P.sp2_alg2(np.eye(4), 0.5, 1, 10, "Abs", 1e-10)

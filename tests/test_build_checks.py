"""Checks on the gfx950 assembly of the library's own LDS-ring kernels.

The int8 CRT product kernels and the ring form of the batched split-precision
SP2 kernel copy operand stages global -> LDS with global_load_lds_dwordx4
written as inline assembly that sets M0 itself (so that the compiler adds no
conservative vmcnt waits before LDS reads).  LLVM treats M0 as reserved and
does not honour a clobber of it, so correctness needs every M0 access in
those kernels to be the helper's own: an `s_mov_b32 m0, sN` followed by
`s_nop 0` and the DMA.  This test compiles csrc/prg_gemm_crt.hip and
csrc/prg_batched.hip for gfx950 (device code only, to assembly) and fails if
any other instruction of those kernels reads or writes m0.
"""
import re
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
CSRC = ROOT / "qmd-progress_amd" / "csrc"
# (source file, mangled-name pattern of its LDS-ring kernels)
RING_KERNELS = [("prg_gemm_crt.hip", r"_Z\S*crt_product\S*kernel\S*"),
                ("prg_batched.hip", r"_Z\S*batched_split_ring_kernel\S*")]
HIPCC = Path("/opt/rocm/bin/hipcc")


def _kernel_bodies(asm, pattern=r"_Z\S*crt_product\S*kernel\S*"):
    """{symbol: [instruction lines]} of the kernels whose symbol matches."""
    bodies = {}
    cur = None
    for line in asm.splitlines():
        m = re.match(r"^(" + pattern + r"):", line)
        if m:
            cur = m.group(1)
            bodies[cur] = []
            continue
        if cur is not None:
            if "-- End function" in line:
                cur = None
                continue
            s = line.split(";")[0].strip()
            if s and not s.startswith(".") and not s.endswith(":"):
                bodies[cur].append(s)
    return bodies


@pytest.mark.skipif(not HIPCC.exists() or shutil.which("bash") is None, reason="hipcc not found")
@pytest.mark.parametrize("src,pattern", RING_KERNELS)
def test_m0_is_only_touched_by_the_lds_dma_helper(tmp_path, src, pattern):
    out = tmp_path / "ring.s"
    r = subprocess.run([str(HIPCC), "-O3", "-std=c++17", "--offload-arch=gfx950",
                        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                        "--offload-device-only", "-S", str(CSRC / src), "-o", str(out)],
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-2000:]
    bodies = _kernel_bodies(out.read_text(), pattern)
    assert bodies, "no LDS-ring kernel found in the assembly"
    for name, ins in bodies.items():
        dmas = 0
        for i, s in enumerate(ins):
            if re.search(r"\bm0\b", s) is None:
                continue
            # the helper: s_mov_b32 m0, sN / s_nop 0 / global_load_lds_dwordx4
            assert re.fullmatch(r"s_mov_b32\s+m0,\s*s\d+", s), (name, s)
            assert ins[i + 1].startswith("s_nop"), (name, ins[i + 1])
            assert ins[i + 2].startswith("global_load_lds_dwordx4"), (name, ins[i + 2])
            dmas += 1
        n_dma = sum(1 for s in ins if s.startswith("global_load_lds_dwordx4"))
        assert dmas == n_dma > 0, (name, dmas, n_dma)


def test_m0_checker_rejects_a_foreign_m0_use():
    """The m0 scan itself: a body with one more m0 access (an indexed move
    through m0) is rejected; the helper's own pattern is accepted."""
    good = ("_ZN3prg12_GLOBAL__N_118crt_product_kernelILi8ELi4EEEvNS0_11ProductArgsE:\n"
            "  s_mov_b32 m0, s4\n  s_nop 0\n  global_load_lds_dwordx4 v[0:1], off\n"
            "  s_endpgm\n; -- End function\n")
    bad = good.replace("  s_endpgm\n", "  s_mov_b32 m0, 12\n  v_movrels_b32 v1, v2\n  s_endpgm\n")
    bodies = _kernel_bodies(good)
    ins = next(iter(bodies.values()))
    assert sum(1 for s in ins if re.search(r"\bm0\b", s)) == 1
    ins_bad = next(iter(_kernel_bodies(bad).values()))
    foreign = [s for s in ins_bad if re.search(r"\bm0\b", s)
               and not re.fullmatch(r"s_mov_b32\s+m0,\s*s\d+", s)]
    assert foreign == ["s_mov_b32 m0, 12"]

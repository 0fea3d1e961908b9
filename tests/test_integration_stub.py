"""The reference-side binding in INTEGRATION.md compiles against the
reference's own headers: fmm/evaluate.h (preproc, run, fmm/evaluate.h:18,23),
node.h / input.h (AllNodes, NodeTree), trans.h (Trans_matrix with its
real_t** UE2UC / DE2DC, fmm/trans.h:23-27) and util.h.  g++ -fsyntax-only
with the reference's gcc flag -Drestrict=__restrict__
(arch/x86_64-Xeon-Linux-gcc:3); the one FFTW header that fmm/node.h:5
includes is not in this image, so a typedef-only stand-in for its two types
is written into a temporary folder.  Nothing is linked or run.  Skipped
where the reference tree is absent."""
import re
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
REF = Path("/root/reference/fmm")

pytestmark = pytest.mark.skipif(not (REF / "evaluate.h").exists() or shutil.which("g++") is None,
                                reason="reference headers or g++ not available")


def stub_source():
    text = (ROOT / "INTEGRATION.md").read_text()
    sec = text[text.index("## 1. Drop-in"):text.index("## 2.")]
    blocks = re.findall(r"```c\+\+\n(.*?)```", sec, flags=re.S)
    assert len(blocks) == 1
    return blocks[0]


def test_binding_stub_compiles_against_the_reference_headers(tmp_path):
    src = stub_source()
    for name in ("preproc(AllNodes* A)", "run(FMMWrapper_t* F)", "kifmm_gpu_create", "kifmm_gpu_evaluate",
                 "kifmm_gpu_output", "compute_UE2UC_mat", "compute_DE2DC_mat"):
        assert name in src, name
    (tmp_path / "evaluate--hip.cc").write_text(src)
    (tmp_path / "fftw3.h").write_text("typedef double fftw_complex[2];\ntypedef struct fftw_plan_s* fftw_plan;\n"
                                      "typedef float fftwf_complex[2];\ntypedef struct fftwf_plan_s* fftwf_plan;\n")
    r = subprocess.run(["g++", "-fsyntax-only", "-Drestrict=__restrict__", f"-I{tmp_path}", f"-I{REF}",
                        f"-I{ROOT / 'include'}", str(tmp_path / "evaluate--hip.cc")],
                       capture_output=True, text=True, timeout=120)
    errors = [ln for ln in r.stderr.splitlines() if "error" in ln]
    assert r.returncode == 0 and not errors, r.stderr[-4000:]


def test_binding_packs_the_eight_octant_matrices_and_sums_leaf_counts():
    """The two defects of the earlier stub: counts taken from the root
    (0 for any tree deeper than its root, fmm/input.cc:285-292,440-446) and
    the real_t** octant matrices passed as if contiguous."""
    src = stub_source()
    assert "nv[0].srcNum" not in src and "nv[0].trgNum" not in src
    assert "n_src += d.srcNum" in src and "n_trg += d.trgNum" in src
    assert re.search(r"memcpy\(&ue2uc\[\(size_t\)o \* nuc \* ne\], TM->UE2UC\[o\]", src)
    assert re.search(r"memcpy\(&de2dc\[\(size_t\)o \* ne \* ne\], TM->DE2DC\[o\]", src)

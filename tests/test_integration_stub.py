"""The reference-side binding stub of INTEGRATION.md section 2 compiles.

The stub is taken verbatim from INTEGRATION.md and syntax-checked with
hipcc against the reference's own headers, as a module file inside the
reference tree would be (flags of the reference's ARCH=HIP build,
Makefile:56-68), with this build's include/ on the include path.  The
reference tree is only read: the stub goes into a temporary directory
whose src/ is a symbolic link to the reference's src/.  Skips where the
reference tree or hipcc is absent.
"""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_SRC = "/root/reference/src"
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


def stub_text():
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    sec = text.split("## 2. Binding the C-ABI from the reference's own tree", 1)[1]
    sec = sec.split("\n## 3.", 1)[0]
    blocks = re.findall(r"```cpp\n(.*?)```", sec, re.S)
    assert len(blocks) == 1, "section 2 holds one C++ block"
    return blocks[0]


def test_stub_block_present():
    s = stub_text()
    assert "class mi355x_mhd_t : public mhd_t" in s
    # the corrected bases of the host arrays (dat_3d_t::at takes an index triple)
    assert "h.hyd.u.at(o)" in s and "h.fld.bf.at(g)" in s
    assert "g(-2, -2, -2)" in s


@pytest.mark.skipif(not os.path.isdir(REF_SRC) or not os.path.exists(HIPCC),
                    reason="reference tree or hipcc absent")
def test_stub_compiles_against_reference_headers(tmp_path):
    os.symlink(REF_SRC, tmp_path / "src")
    mod = tmp_path / "usr_ext" / "mi355x"
    mod.mkdir(parents=True)
    (mod / "mi355x_mhd.h").write_text(stub_text())
    (mod / "check.cpp").write_text(
        '#include "mi355x_mhd.h"\n'
        "#include <memory>\n"
        "std::shared_ptr<mesh::mod_base_t> make_module() {\n"
        "    return std::make_shared<dynamics::mi355x_mhd_t>();\n"
        "}\n")
    cmd = [HIPCC, "-std=c++17", "-fsyntax-only", "-x", "hip", "--offload-arch=gfx950",
           "-include", "hip/hip_runtime.h", "-Wno-return-type",
           "-Wno-unused-command-line-argument", "-Wno-inconsistent-missing-override",
           "-I", os.path.join(ROOT, "include"), str(mod / "check.cpp")]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600, cwd=str(mod))
    assert r.returncode == 0, r.stderr[-4000:]

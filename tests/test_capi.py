"""CPU checks of the C-ABI boundary (no kernel launches).

* libkratos_amd.so loads and exports every entry point include/kratos_amd.h
  declares, and nothing the header does not declare under the kr_ prefix.
* The ctypes structs of the Python host layer have the size and field
  offsets the C compiler gives the header's structs (a probe compiled with
  gcc from the header itself).
* Argument errors come back as status codes with a message, never as a
  crash, also without a GPU (checks that run before any device call).
"""
import ctypes as C
import os
import re
import subprocess

import pytest

import kratos_amd
from kratos_amd import _capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "kratos_amd.h")


def declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:[A-Za-z_][\w\s\*]*?)\b(kr_\w+)\s*\(", text, re.M)))


def test_header_declares_entry_points():
    names = declared()
    assert "kr_create" in names and "kr_advance" in names
    assert set(names) == set(_capi.EXPORTS), "ctypes table and header disagree"


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(kratos_amd.LIB_PATH)
    for name in declared():
        assert hasattr(lib, name), f"{name} missing from {kratos_amd.LIB_PATH}"
    out = subprocess.run(["nm", "-D", "--defined-only", kratos_amd.LIB_PATH],
                         capture_output=True, text=True, check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    kr = {s for s in exported if s.startswith("kr_")}
    assert kr == set(declared()), f"undeclared kr_ exports: {sorted(kr - set(declared()))}"


PROBE = r"""
#include <stddef.h>
#include <stdio.h>
#include "kratos_amd.h"
#define F(T, m) printf(#T "." #m " %zu\n", offsetof(T, m))
int main(void) {
    printf("kr_config %zu\nkr_block_desc %zu\nkr_stats %zu\n", sizeof(kr_config),
           sizeof(kr_block_desc), sizeof(kr_stats));
    F(kr_config, gamma); F(kr_config, bnd); F(kr_config, t_0); F(kr_config, top_up);
    F(kr_config, n_cell_block); F(kr_config, kernel_path); F(kr_config, tile);
    F(kr_block_desc, xf0); F(kr_block_desc, dx0); F(kr_block_desc, nbr);
    F(kr_stats, n_cells); F(kr_stats, kernel_ms); F(kr_stats, kernel_calls);
    F(kr_stats, alg_bytes_per_cell_update); F(kr_stats, sweep_tile); F(kr_stats, tune_ms);
    F(kr_block_desc, group); F(kr_stats, tune_pending); F(kr_stats, graph_ready);
    F(kr_stats, sweep_kernel);
    return 0;
}
"""


def test_struct_layouts_match_header(tmp_path):
    src = tmp_path / "probe.c"
    src.write_text(PROBE)
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.dirname(HEADER), str(src), "-o", str(exe)],
                   check=True)
    lines = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    got = dict(ln.split() for ln in lines if ln.strip())
    structs = {"kr_config": _capi.KrConfig, "kr_block_desc": _capi.KrBlockDesc,
               "kr_stats": _capi.KrStats}
    for key, val in got.items():
        if "." in key:
            s, m = key.split(".")
            assert getattr(structs[s], m).offset == int(val), key
        else:
            assert C.sizeof(structs[key]) == int(val), key


def test_create_rejects_bad_config_without_device():
    lib = _capi.load()
    cfg = _capi.KrConfig()
    cfg.model = 7                               # no such model
    cfg.gamma = 5.0 / 3.0
    cfg.n_cell_block[:] = [8, 8, 8]
    h = C.c_void_p()
    rc = lib.kr_create(C.byref(cfg), C.byref(h))
    assert rc == -1 and not h.value
    # the reason comes back through kr_last_error(NULL)
    assert lib.kr_last_error(None).decode() == "kr_create: unknown model"
    cfg.model = 1
    cfg.cfl = 0.0
    assert lib.kr_create(C.byref(cfg), C.byref(h)) == -1
    assert "cfl" in lib.kr_last_error(None).decode()
    cfg.cfl = 0.3
    cfg.bnd[:] = [0, 1, 0, 0, 0, 0]
    assert lib.kr_create(C.byref(cfg), C.byref(h)) == -1
    assert "periodic" in lib.kr_last_error(None).decode()
    with pytest.raises(_capi.KratosError, match="periodic must be set on both sides"):
        _capi.check(-1, None, "kr_create")
    assert lib.kr_create(None, C.byref(h)) == -1
    assert "null" in lib.kr_last_error(None).decode()
    lib.kr_destroy(None)                        # no-op on a null handle


def test_python_layer_fails_loudly_without_library(monkeypatch, tmp_path):
    monkeypatch.setattr(_capi, "_lib", None)
    monkeypatch.setattr(_capi, "LIB_PATH", str(tmp_path / "missing.so"))
    with pytest.raises(_capi.KratosError):
        _capi.load()

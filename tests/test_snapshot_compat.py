"""Snapshots of the host layer read by the reference's own reader.

oracle/_ref/snapshot_reader is the reference's binary container reader
(io/binary/binary_io_base.cpp) and tree deserializer (mesh/tree/tree.cpp
:22-31) compiled from the reference's sources by oracle/Makefile.ref; the
tests skip where it is not built (the reference tree is absent).  The host
layer's generators run against a stand-in of the C-ABI (no device: the
blocks keep their initial state), write a snapshot, and the reference's
reader must find in it:

* every tag with the bytes this build wrote (hashes agree with an
  independent parse of the container);
* the parameters, with [mesh] dist_mode = 0 (the block numbering the
  snapshot uses);
* a mesh_tree whose nodes are the tree this build's Python builder
  (RefinedMeshSpec, or the base level of a same-level mesh) derives from
  the parameter file: same keys in the reference's order, masked parents
  with id -1, leaves numbered 0..n-1 in the tree's order, and each leaf's
  block_<id>|level and |i_logic tags naming that leaf.
"""
import os
import struct
import subprocess

import pytest

from kratos_amd.mesh import RefinedMeshSpec
from test_host_api import HOST, MOCK_ABI, ROOT, parse_snapshot

READER = os.path.join(ROOT, "oracle", "_ref", "snapshot_reader")
pytestmark = pytest.mark.skipif(not os.access(READER, os.X_OK),
                                reason="reference reader not built (reference tree absent)")


def fnv1a(b):
    h = 1469598103934665603
    for c in b:
        h ^= c
        h = (h * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return h


def mock_pgen(name, tmp_path):
    mock = tmp_path / "mock_abi.cpp"
    mock.write_text(MOCK_ABI)
    exe = tmp_path / f"{name}_mock"
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", HOST, "-I", os.path.join(ROOT, "include"),
                    os.path.join(HOST, "usr_example", name, "pgen.cpp"),
                    os.path.join(HOST, "kratos", "kratos.cpp"), str(mock), "-o", str(exe),
                    "-lpthread"], check=True)
    return exe


def snapshot(name, tmp_path, over=""):
    exe = mock_pgen(name, tmp_path)
    # one parameter file: a later file would reset the [refine_region_*]
    # section list (input.cpp:83); later keys override earlier ones
    text = open(os.path.join(HOST, "usr_example", name, f"{name}.par")).read()
    o = tmp_path / "run.par"
    o.write_text(text + f"\n[cycle]\nn_cycle_lim = 1\nfinal_output = 1\n"
                 f"prefix_output = {tmp_path}/S\n" + over)
    subprocess.run([str(exe), str(o)], check=True, capture_output=True, cwd=tmp_path)
    return tmp_path / "S_00001.bin"


def read_ref(path):
    out = subprocess.run([READER, str(path)], check=True, capture_output=True,
                         text=True).stdout.splitlines()
    rec = {"tag": {}, "hash": {}, "arg": {}, "tree": None, "node": []}
    for ln in out:
        f = ln.split(" ")
        if f[0] == "tag":
            rec["tag"][f[1]] = (int(f[2]), int(f[3]))
        elif f[0] == "hash":
            rec["hash"][f[1]] = int(f[2], 16)
        elif f[0] == "arg":
            rec["arg"][f[1]] = " ".join(f[2:])
        elif f[0] == "tree":
            rec["tree"] = [int(x) for x in f[1:]]
        elif f[0] == "node":
            rec["node"].append(tuple(int(x) for x in f[1:]))
    return rec


def check_common(path):
    ref = read_ref(path)
    mine = parse_snapshot(path)
    assert set(ref["tag"]) == set(mine)
    for k, (unit, data) in mine.items():
        assert ref["tag"][k] == (len(data), unit), k
        assert ref["hash"][k] == fnv1a(data), k
    assert ref["arg"]["mesh|dist_mode"] == "0"
    return ref, mine


def leaf_tags(mine, ids):
    out = {}
    for i in ids:
        lvl = struct.unpack("<i", mine[f"block_{i}|level"][1])[0]
        reg = struct.unpack("<3i", mine[f"block_{i}|i_logic"][1])
        out[i] = (lvl,) + reg
    return out


def test_same_level_tree(tmp_path):
    """vortex: 32^3 in 8 blocks of 16^3, periodic, the base level only."""
    ref, mine = check_common(snapshot("vortex", tmp_path))
    n_dim, rx, ry, rz, px, py, pz, size, eff = ref["tree"]
    assert (n_dim, rx, ry, rz, px, py, pz, size, eff) == (3, 2, 2, 2, 1, 1, 1, 8, 8)
    keys = [(lv, x, y, z) for lv, x, y, z, *_ in ref["node"]]
    assert keys == sorted((0, x, y, z) for x in range(2) for y in range(2) for z in range(2))
    for i, (lv, x, y, z, id_l, id_g, rank, mask) in enumerate(ref["node"]):
        assert (id_l, id_g, rank, mask) == (i, i, 0, 0)
    assert leaf_tags(mine, range(8)) == {i: k for i, k in enumerate(keys)}


def test_same_level_tree_2d(tmp_path):
    """vortex on a 64 x 64 x 1 mesh of 16 x 32 x 1 blocks: reg_max 0 past
    n_dim, as the reference leaves it."""
    over = "[mesh]\nn_cell_global = 64 64 1\nn_cell_block = 16 32 1\n"
    ref, mine = check_common(snapshot("vortex", tmp_path, over))
    assert ref["tree"][:4] == [2, 4, 2, 0]
    assert ref["tree"][7:] == [8, 8]
    keys = [tuple(n[:4]) for n in ref["node"]]
    assert keys == sorted((0, x, y, 0) for x in range(4) for y in range(2))
    assert leaf_tags(mine, range(8)) == {i: k for i, k in enumerate(keys)}


def test_refined_tree(tmp_path):
    """blast: 3 levels from [refine_region_*]; the nodes are the tree the
    Python builder derives from the same regions, masked parents included."""
    ref, mine = check_common(snapshot("blast", tmp_path))
    spec = RefinedMeshSpec((32, 32, 32), (8, 8, 8),
                           refine=[(1, (-0.25,) * 3, (0.2,) * 3), (2, (-0.05,) * 3, (0.05,) * 3)],
                           x_min=(-0.5,) * 3, x_max=(0.5,) * 3)
    want = sorted(spec.tree.items())
    assert ref["tree"] == [3, 4, 4, 4, 1, 1, 1, len(want), spec.n_blocks]
    assert [tuple(n[:4]) for n in ref["node"]] == [k for k, _ in want]
    nxt = 0
    for (lv, x, y, z, id_l, id_g, rank, mask), (k, masked) in zip(ref["node"], want):
        assert mask == int(masked)
        if masked:
            assert (id_l, id_g, rank) == (-1, -1, -1)
        else:
            assert (id_l, id_g, rank) == (nxt, nxt, 0)
            nxt += 1
    assert leaf_tags(mine, range(spec.n_blocks)) == dict(enumerate(spec.leaves))

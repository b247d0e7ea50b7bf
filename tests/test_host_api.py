"""CPU checks of the C++ host API (kratos_amd/host): the reference-style
problem generators compile against the mirror headers and the C-ABI, the
parameter-file reader follows the reference's rules (sections, '#'
comments, 3-vectors, sequences, missing keys), and a generator fails
loudly when no device is present.
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "kratos_amd", "host")


def test_examples_build():
    subprocess.run(["make", "-s", "-C", HOST], check=True)
    for name in ("sod", "vortex"):
        assert os.access(os.path.join(HOST, "bin", name), os.X_OK)


PARSER = r"""
#include <cstdio>
#include "kratos/input.h"
int main(int argc, char **argv) {
    input args;
    args.set_file(argv[1]);
    if (!args.read()) return 3;
    idx_t n = args.get<idx_t>("mesh", "n_cell_global");
    coord2_t x = args.get<coord2_t>("mesh", "x_min");
    std::vector<std::string> kinds;
    args("boundary", "kinds", kinds);
    double g = args.get<double>("dynamics", "gamma", 1.5);
    double c = args.get<double>("dynamics", "cfl", 0.3);
    bool top = args.get<bool>("cycle", "t_lim_top_up", false);
    int missing_ok = 0;
    try { args.get<int>("cycle", "nope"); } catch (const std::runtime_error &) { missing_ok = 1; }
    int untouched = 42;
    bool f = args("cycle", "nope", untouched);
    printf("%d %d %d|%g %g %g|%zu %s %s|%g %g %d|%d %d %d\n", n[0], n[1], n[2], x[0], x[1], x[2],
           kinds.size(), kinds[0].c_str(), kinds[5].c_str(), g, c, (int)top, missing_ok, (int)f,
           untouched);
    return 0;
}
"""

PAR = """# comment line
[mesh]
n_cell_global = 512 2 1   # trailing comment
x_min = 0 -1 -1.5
[boundary]
kinds = out out per per fre ref
[dynamics]
gamma=1.4
[cycle]
t_lim_top_up = 1
"""


def test_parameter_file_rules(tmp_path):
    src = tmp_path / "p.cpp"
    src.write_text(PARSER)
    exe = tmp_path / "p"
    subprocess.run(["g++", "-std=c++17", "-I", HOST, str(src), "-o", str(exe)], check=True)
    par = tmp_path / "a.par"
    par.write_text(PAR)
    out = subprocess.run([str(exe), str(par)], capture_output=True, text=True, check=True).stdout
    line = out.strip().splitlines()[-1]
    assert line == "512 2 1|0 -1 -1.5|6 out ref|1.4 0.3 1|1 0 42", line


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is present; the GPU tests run it")
def test_generator_fails_loudly_without_device(tmp_path):
    subprocess.run(["make", "-s", "-C", HOST], check=True)
    par = os.path.join(HOST, "usr_example", "sod", "sod.par")
    r = subprocess.run([os.path.join(HOST, "bin", "sod"), par], capture_output=True, text=True,
                       cwd=tmp_path)
    assert r.returncode != 0
    assert "kr_create" in r.stderr


SNAP = r"""
#include <cstdio>
#include "kratos/binary_io.h"
int main(int argc, char **argv) {
    binary_io::base_t w;
    std::map<std::string, std::string> args{{"mesh|n_cell_global", "32 32 32"}, {"cycle|t_lim", "0.1"}};
    w.write_map(args, "input_args");
    const double t = 0.125; const int cyc = 7; const int v3[3] = {1, 2, 3};
    double f[6] = {1, 2, 3, 4, 5, 6};
    w.write_value(t, "time"); w.write_value(cyc, "cycle");
    w.write(v3, 3, sizeof(int), "block_0|i_logic");
    w.write(f, 6, sizeof(double), "block_0|hydro_cons_field");
    w.save(argv[1]);
    binary_io::base_t r;
    r.load(argv[1]);
    double t2; int c2; int v[3]; double g[6];
    r.read(&t2, 1, "time"); r.read(&c2, 1, "cycle"); r.read(v, 3, "block_0|i_logic");
    r.read(g, 6, "block_0|hydro_cons_field");
    auto m = r.read_map("input_args");
    bool ok = t2 == t && c2 == cyc && v[2] == 3 && g[5] == 6 && m == args;
    printf("%s\n", ok ? "roundtrip ok" : "roundtrip FAILED");
    return ok ? 0 : 1;
}
"""


def parse_snapshot(path):
    """Independent reader of the snapshot container: byte 0 = little-endian
    flag | sizeof(size_t); size_t header size; size_t count; per tag: size_t
    key length, key, size_t size, 1-byte unit, size_t offset; data at header
    size + offset."""
    import struct
    raw = open(path, "rb").read()
    assert raw[0] == (1 | 8)
    hsz = struct.unpack_from("<Q", raw, 1)[0]
    pos = 9
    n = struct.unpack_from("<Q", raw, pos)[0]
    pos += 8
    tags = {}
    for _ in range(n):
        klen = struct.unpack_from("<Q", raw, pos)[0]
        pos += 8
        key = raw[pos:pos + klen].decode("ascii")
        pos += klen
        size, = struct.unpack_from("<Q", raw, pos)
        unit = raw[pos + 8]
        off, = struct.unpack_from("<Q", raw, pos + 9)
        pos += 17
        tags[key] = (unit, raw[hsz + off:hsz + off + size])
    assert pos == hsz
    return tags


def parse_args_map(blob):
    import struct
    n, = struct.unpack_from("<Q", blob, 0)
    pos, out = 8, {}
    for _ in range(n):
        kv = []
        for _ in range(2):
            ln, = struct.unpack_from("<Q", blob, pos)
            kv.append(blob[pos + 8:pos + 8 + ln].decode("ascii"))
            pos += 8 + ln
        out[kv[0]] = kv[1]
    return out


def test_snapshot_container(tmp_path):
    import numpy as np
    src = tmp_path / "s.cpp"
    src.write_text(SNAP)
    exe = tmp_path / "s"
    subprocess.run(["g++", "-std=c++17", "-I", HOST, str(src), "-o", str(exe)], check=True)
    snap = tmp_path / "x.bin"
    out = subprocess.run([str(exe), str(snap)], capture_output=True, text=True)
    assert out.returncode == 0 and "roundtrip ok" in out.stdout, out.stdout
    tags = parse_snapshot(snap)
    assert sorted(tags) == sorted(["input_args", "time", "cycle", "block_0|i_logic",
                                   "block_0|hydro_cons_field"])
    assert tags["time"][0] == 8 and np.frombuffer(tags["time"][1], "<f8")[0] == 0.125
    assert tags["cycle"][0] == 4 and np.frombuffer(tags["cycle"][1], "<i4")[0] == 7
    assert list(np.frombuffer(tags["block_0|i_logic"][1], "<i4")) == [1, 2, 3]
    assert list(np.frombuffer(tags["block_0|hydro_cons_field"][1], "<f8")) == [1, 2, 3, 4, 5, 6]
    assert parse_args_map(tags["input_args"][1]) == {"mesh|n_cell_global": "32 32 32",
                                                    "cycle|t_lim": "0.1"}

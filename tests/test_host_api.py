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


MOCK_ABI = r"""
// A stand-in for libkratos_amd.so with no device: every cycle sleeps 20 ms
// and advances t by 1e-3, so that the host layer's evolve loop can be
// driven (and interrupted) on a machine without a GPU.
#include <chrono>
#include <cstring>
#include <thread>
#include "kratos_amd.h"
struct kr_ctx { double t = 0; int cycles = 0; };
static kr_ctx g_ctx;
extern "C" {
int kr_create(const kr_config *, kr_ctx **out) { *out = &g_ctx; return KR_OK; }
int kr_add_blocks(kr_ctx *, int, const kr_block_desc *) { return KR_OK; }
int kr_upload(kr_ctx *, int, const double *, const double *) { return KR_OK; }
int kr_download(kr_ctx *, int, double *, double *) { return KR_OK; }
int kr_step(kr_ctx *c, double *dt) {
    std::this_thread::sleep_for(std::chrono::milliseconds(20));
    c->t += 1e-3; ++c->cycles; if (dt) *dt = 1e-3; return KR_OK;
}
int kr_advance(kr_ctx *c, int n, int *done) {
    for (int i = 0; i < n; ++i) kr_step(c, nullptr);
    if (done) *done = n; return KR_OK;
}
int kr_get_time(kr_ctx *c, double *t, double *dt_last, double *dt_seed) {
    if (t) *t = c->t; if (dt_last) *dt_last = 1e-3; if (dt_seed) *dt_seed = 1e-3; return KR_OK;
}
int kr_divb_max(kr_ctx *, double *d) { *d = 0; return KR_OK; }
int kr_sync(kr_ctx *) { return KR_OK; }
const char *kr_last_error(kr_ctx *) { return ""; }
void kr_destroy(kr_ctx *) {}
}
"""


def test_sigint_saves_final_snapshot(tmp_path):
    """SIGINT during evolve (the reference's handler, cycle.cpp:32 and
    :169-181): the run stops between cycles and, with final_output set,
    writes its final snapshot.  The host layer runs against a stand-in of
    the C-ABI here (no device); the GPU test of the same behaviour drives
    the real library."""
    import signal
    import time as _time
    mock = tmp_path / "mock_abi.cpp"
    mock.write_text(MOCK_ABI)
    exe = tmp_path / "vortex_mock"
    inc = os.path.join(ROOT, "include")
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", HOST, "-I", inc,
                    os.path.join(HOST, "usr_example", "vortex", "pgen.cpp"),
                    os.path.join(HOST, "kratos", "kratos.cpp"), str(mock), "-o", str(exe),
                    "-lpthread"], check=True)
    over = tmp_path / "over.par"
    over.write_text(f"[cycle]\nn_cycle_lim = 1000000\nn_display_freq = 1\nfinal_output = 1\n"
                    f"prefix_output = {tmp_path}/S\n")
    par = os.path.join(HOST, "usr_example", "vortex", "vortex.par")
    p = subprocess.Popen([str(exe), par, str(over)], stdout=subprocess.PIPE,
                         stderr=subprocess.PIPE, text=True, cwd=tmp_path)
    seen = 0
    t_end = _time.time() + 60
    while seen < 3 and _time.time() < t_end:
        line = p.stdout.readline()
        if not line:
            break
        seen += line.startswith("cycle =")
    assert seen == 3, "the run did not reach its third cycle"
    p.send_signal(signal.SIGINT)
    out, err = p.communicate(timeout=60)
    assert p.returncode == 0, err
    assert "Signal : 2" in err
    snaps = sorted(x.name for x in tmp_path.iterdir() if x.name.endswith(".bin"))
    assert snaps[0] == "S_00000.bin" and len(snaps) == 2, snaps   # initial + final
    assert "Making output" in out.splitlines()[-1] or "Making output" in out

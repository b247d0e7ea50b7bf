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

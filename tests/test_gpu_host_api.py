"""GPU: the reference-style C++ problem generators (kratos_amd/host), run as
executables with parameter files, against the CPU oracle started from the
initial state each generator wrote.

* sod: 512 x 2 x 1 hydro (PLM + HLLC + RK2, outflow, cfl 0.4) to t = 0.2
  with the t_lim top-up; the initial state must equal the Python Sod
  generator's and the final state the oracle's, bit for bit.
* blast: MHD blast on a 3-level [refine_region_*] mesh; same tree as the
  Python builder, final state bit-exact with the oracle.
* vortex: 3-D Orszag-Tang-type MHD, 32^3 as 8 blocks of 16^3; n_cycle_lim
  = 9 runs 10 cycles (the reference's loop rule); final state bit-exact
  with the oracle, div B at roundoff; the same generator on a 2-D mesh
  (64 x 64 x 1) set in the override file.
"""
import os
import subprocess

import numpy as np
import pytest

import kratos_amd
import oracle
from kratos_amd import problems
from kratos_amd.mesh import bf_host_size

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "kratos_amd", "host")


def run_pgen(name, par, tmp_path, extra=""):
    exe = os.path.join(HOST, "bin", name)
    if not os.access(exe, os.X_OK):
        raise RuntimeError(f"{exe} is not built: run __graft_entry__.build() first")
    dump = tmp_path / f"{name}.raw"
    over = tmp_path / "over.par"
    over.write_text(f"[prob]\ndump = {dump}\n{extra}")
    print(f"\n[case] host pgen {name} {par}", flush=True)
    r = subprocess.run(["timeout", "-k", "10", "120", exe, par, str(over)], capture_output=True,
                       text=True, cwd=tmp_path)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0, r.stderr
    return np.fromfile(dump, dtype=np.float64), r.stdout


def host_order(spec):
    """MeshSpec block of each host-API block id: the host layer numbers the
    blocks of a same-level mesh in the tree's order, x slowest (the
    reference's dist_mode = 0), MeshSpec x fastest."""
    nx, ny, nz = spec.nblk
    return [(b % nz) * nx * ny + ((b // nz) % ny) * nx + b // (ny * nz)
            for b in range(spec.n_blocks)]


def split_dump(raw, spec, mhd):
    n = spec.n_cell_block
    nu = n[0] * n[1] * n[2] * (8 if mhd else 5)
    nf = bf_host_size(n) if mhd else 0
    per = nu + nf
    assert raw.size == 2 * spec.n_blocks * per
    order = host_order(spec) if isinstance(spec, kratos_amd.MeshSpec) else list(
        range(spec.n_blocks))
    states = []
    for s in range(2):
        blocks = [None] * spec.n_blocks
        for b in range(spec.n_blocks):
            o = (s * spec.n_blocks + b) * per
            blocks[order[b]] = (raw[o:o + nu], raw[o + nu:o + per] if mhd else None)
        states.append(spec.gather(blocks, mhd))
    return states


def test_sod_generator(gpu_lib, tmp_path):
    par = os.path.join(HOST, "usr_example", "sod", "sod.par")
    spec = kratos_amd.MeshSpec((512, 2, 1), (512, 2, 1), (0, -1, -1), (1, 1, 1),
                               ("out", "out", "out", "out", "out", "out"))
    raw, out = run_pgen("sod", par, tmp_path)
    (u0, _), (u1, _) = split_dump(raw, spec, False)
    np.testing.assert_array_equal(u0, problems.sod())
    phys = kratos_amd.Physics(mhd=False, gamma=1.4, cfl=0.4, t_lim=0.2, top_up=True)
    ref = oracle.OracleMesh(spec, phys)
    ref.upload(u0, None)
    ref.advance(100000)
    ou, _ = ref.download()
    assert ref.time()[0] == 0.2
    np.testing.assert_array_equal(u1, ou)
    assert "Speed =" in out


def test_vortex_generator(gpu_lib, tmp_path):
    par = os.path.join(HOST, "usr_example", "vortex", "vortex.par")
    spec = kratos_amd.MeshSpec((32, 32, 32), (16, 16, 16))
    raw, out = run_pgen("vortex", par, tmp_path)
    (u0, b0), (u1, b1) = split_dump(raw, spec, True)
    phys = kratos_amd.Physics(mhd=True, gamma=1.6666666666666667, cfl=0.3)
    ref = oracle.OracleMesh(spec, phys)
    ref.upload(u0, b0)
    assert ref.advance(10) == 10
    ou, ob = ref.download()
    np.testing.assert_array_equal(u1, ou)
    for x, y in zip(b1, ob):
        np.testing.assert_array_equal(x, y)
    divb = float(out.split("max |div B| =")[1].split()[0])
    assert divb < 1e-12


def test_vortex_generator_2d(gpu_lib, tmp_path):
    """The same generator on a 2-D mesh (thin z) set in the override file:
    64 x 64 x 1 as 4 blocks of 32 x 32 x 1."""
    par = os.path.join(HOST, "usr_example", "vortex", "vortex.par")
    spec = kratos_amd.MeshSpec((64, 64, 1), (32, 32, 1))
    raw, out = run_pgen("vortex", par, tmp_path,
                        "[mesh]\nn_cell_global = 64 64 1\nn_cell_block = 32 32 1\n")
    (u0, b0), (u1, b1) = split_dump(raw, spec, True)
    phys = kratos_amd.Physics(mhd=True, gamma=1.6666666666666667, cfl=0.3)
    ref = oracle.OracleMesh(spec, phys)
    ref.upload(u0, b0)
    assert ref.advance(10) == 10
    ou, ob = ref.download()
    np.testing.assert_array_equal(u1, ou)
    for x, y in zip(b1, ob):
        np.testing.assert_array_equal(x, y)
    assert np.abs(u1[3] - u0[3]).max() > 1e-3
    divb = float(out.split("max |div B| =")[1].split()[0])
    assert divb < 1e-12


def _snap_fields(path):
    from test_host_api import parse_snapshot
    tags = parse_snapshot(path)
    f = {k: v[1] for k, v in tags.items()}
    return f


def test_vortex_snapshot_restart(gpu_lib, tmp_path):
    """Snapshots in the reference's container: a run stopped after 5 cycles
    and restarted from its snapshot (the snapshot as the parameter file)
    ends bit-identical to an uninterrupted run of the same 13 cycles.  (The
    evolve loop's rule: a run stopped by n_cycle_lim = 4 has taken 5 cycles
    and stores cycle = 4; the restart to n_cycle_lim = 11 takes 8 more.)"""
    exe = os.path.join(HOST, "bin", "vortex")
    par = os.path.join(HOST, "usr_example", "vortex", "vortex.par")

    def run(args, over):
        o = tmp_path / f"over_{len(list(tmp_path.iterdir()))}.par"
        o.write_text(over)
        print(f"\n[case] host pgen vortex {args} {over!r}", flush=True)
        r = subprocess.run(["timeout", "-k", "10", "120", exe] + args + [str(o)],
                           capture_output=True, text=True, cwd=tmp_path)
        print(r.stdout[-1500:], r.stderr[-1500:], flush=True)
        assert r.returncode == 0, r.stderr

    run([par], f"[cycle]\nn_cycle_lim = 12\nfinal_output = 1\nprefix_output = {tmp_path}/A\n")
    run([par], f"[cycle]\nn_cycle_lim = 4\nfinal_output = 1\nprefix_output = {tmp_path}/B\n")
    b1 = tmp_path / "B_00001.bin"
    run([str(b1)], f"[cycle]\nn_cycle_lim = 11\nprefix_output = {tmp_path}/C\n")
    a, c = _snap_fields(tmp_path / "A_00001.bin"), _snap_fields(tmp_path / "C_00002.bin")
    assert np.frombuffer(a["cycle"], "<i4")[0] == 12
    assert np.frombuffer(c["cycle"], "<i4")[0] == 11
    assert a["time"] == c["time"] and a["dt"] == c["dt"]
    blocks = sorted(k for k in a if k.endswith("|hydro_cons_field"))
    assert len(blocks) == 8
    for k in blocks + [k.replace("hydro_cons", "field_bf") for k in blocks]:
        np.testing.assert_array_equal(np.frombuffer(a[k], "<f8"), np.frombuffer(c[k], "<f8"))
    # the initial snapshot holds the initial condition and the parameters
    s0 = _snap_fields(tmp_path / "A_00000.bin")
    assert np.frombuffer(s0["time"], "<f8")[0] == 0.0
    from test_host_api import parse_args_map
    args0 = parse_args_map(s0["input_args"])
    assert args0["mesh|n_cell_global"].split() == ["32", "32", "32"]


def test_blast_generator_refined(gpu_lib, tmp_path):
    """The blast generator on its [refine_region_*] mesh: the host API builds
    the same tree as RefinedMeshSpec (block count, order and geometry: the
    initial state matches the Python blast on that spec), and the final state
    equals the oracle's on that mesh bit for bit (n_cycle_lim = 5 runs 6
    cycles)."""
    from kratos_amd.mesh import RefinedMeshSpec
    # one parameter file: like the reference's input::read (input.cpp:83),
    # each file read resets the section list, so [refine_region_*] must sit
    # in the last file given
    par = tmp_path / "blast_dump.par"
    par.write_text(open(os.path.join(HOST, "usr_example", "blast", "blast.par")).read()
                   + f"\n[prob]\ndump = {tmp_path / 'blast.raw'}\n")
    spec = RefinedMeshSpec((32, 32, 32), (8, 8, 8),
                           refine=[(1, (-0.25,) * 3, (0.2,) * 3), (2, (-0.05,) * 3, (0.05,) * 3)],
                           x_min=(-0.5,) * 3, x_max=(0.5,) * 3)
    exe = os.path.join(HOST, "bin", "blast")
    print(f"\n[case] host pgen blast {par}", flush=True)
    r = subprocess.run(["timeout", "-k", "10", "120", exe, str(par)], capture_output=True,
                       text=True, cwd=tmp_path)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0, r.stderr
    raw, out = np.fromfile(tmp_path / "blast.raw", dtype=np.float64), r.stdout
    assert f"blocks = {spec.n_blocks}" in out
    n = spec.n_cell_block
    nu, nf = n[0] * n[1] * n[2] * 8, (n[0] + 5) * (n[1] + 5) * (n[2] + 5) * 3
    assert raw.size == 2 * spec.n_blocks * (nu + nf)
    states = [[(raw[(s * spec.n_blocks + b) * (nu + nf):][:nu],
                raw[(s * spec.n_blocks + b) * (nu + nf) + nu:][:nf])
               for b in range(spec.n_blocks)] for s in range(2)]
    py = problems.blast_blocks(spec)
    for b in range(spec.n_blocks):
        np.testing.assert_allclose(states[0][b][0], py[b][0], rtol=1e-14, atol=1e-15)
    phys = kratos_amd.Physics(mhd=True, gamma=1.6666666666666667, cfl=0.3)
    ref = oracle.OracleMesh(spec, phys, n_threads=16)
    ref.upload_blocks(states[0])
    assert ref.advance(6) == 6
    fin = ref.download_blocks()
    for b in range(spec.n_blocks):
        np.testing.assert_array_equal(states[1][b][0].reshape(-1, 8)[:, :5],
                                      fin[b][0].reshape(-1, 8)[:, :5])
        f1 = states[1][b][1].reshape(n[2] + 5, n[1] + 5, n[0] + 5, 3)
        f2 = fin[b][1].reshape(n[2] + 5, n[1] + 5, n[0] + 5, 3)
        for c, sl in enumerate(((slice(2, n[2] + 2), slice(2, n[1] + 2), slice(2, n[0] + 3)),
                                (slice(2, n[2] + 2), slice(2, n[1] + 3), slice(2, n[0] + 2)),
                                (slice(2, n[2] + 3), slice(2, n[1] + 2), slice(2, n[0] + 2)))):
            np.testing.assert_array_equal(f1[sl + (c,)], f2[sl + (c,)])
    assert "max |div B|" in out


def test_blast_refined_snapshot_restart(gpu_lib, tmp_path):
    """A refined run restarted from its snapshot takes its blocks (level,
    i_logic) from the file, as the reference reads its tree back, and ends
    bit-identical to the uninterrupted run (the loop rule: n_cycle_lim = 9
    runs 10 cycles; 4 runs 5 and stores cycle 4; the restart to 8 runs 5
    more)."""
    exe = os.path.join(HOST, "bin", "blast")
    base = open(os.path.join(HOST, "usr_example", "blast", "blast.par")).read()

    def run(args, text=None):
        if text is not None:
            o = tmp_path / f"par_{len(list(tmp_path.iterdir()))}.par"
            o.write_text(text)
            args = args + [str(o)]
        print(f"\n[case] host pgen blast {args}", flush=True)
        r = subprocess.run(["timeout", "-k", "10", "120", exe] + args, capture_output=True,
                           text=True, cwd=tmp_path)
        print(r.stdout[-1500:], r.stderr[-1500:], flush=True)
        assert r.returncode == 0, r.stderr

    run([], base + f"\n[cycle]\nn_cycle_lim = 9\nfinal_output = 1\nprefix_output = {tmp_path}/A\n")
    run([], base + f"\n[cycle]\nn_cycle_lim = 4\nfinal_output = 1\nprefix_output = {tmp_path}/B\n")
    run([str(tmp_path / "B_00001.bin")], f"[cycle]\nn_cycle_lim = 8\nprefix_output = {tmp_path}/C\n")
    a, c = _snap_fields(tmp_path / "A_00001.bin"), _snap_fields(tmp_path / "C_00002.bin")
    assert a["time"] == c["time"]
    blocks = sorted(k for k in a if k.endswith("|hydro_cons_field"))
    assert len(blocks) > 64
    levels = {np.frombuffer(a[k.replace("hydro_cons_field", "level")], "<i4")[0] for k in blocks}
    assert levels == {0, 1, 2}
    for k in blocks + [k.replace("hydro_cons", "field_bf") for k in blocks]:
        np.testing.assert_array_equal(np.frombuffer(a[k], "<f8"), np.frombuffer(c[k], "<f8"))


def test_vortex_sigint_final_snapshot(gpu_lib, tmp_path):
    """SIGINT during evolve (cycle.cpp:32, :169-181): the run stops between
    cycles, writes its final snapshot (final_output = 1), and that snapshot
    holds the state of an uninterrupted run of the same number of cycles,
    bit for bit."""
    import signal
    import time as _time
    exe = os.path.join(HOST, "bin", "vortex")
    par = os.path.join(HOST, "usr_example", "vortex", "vortex.par")
    over = tmp_path / "over_s.par"
    over.write_text(f"[cycle]\nn_cycle_lim = 1000000\nn_display_freq = 1\nfinal_output = 1\n"
                    f"prefix_output = {tmp_path}/S\n")
    print("\n[case] host pgen vortex, SIGINT after 4 cycles", flush=True)
    p = subprocess.Popen([exe, par, str(over)], stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                         text=True, cwd=tmp_path)
    try:
        seen = 0
        t_end = _time.time() + 90
        while seen < 4 and _time.time() < t_end:
            line = p.stdout.readline()
            if not line:
                break
            seen += line.startswith("cycle =")
        assert seen == 4, "the run did not reach its fourth cycle"
        p.send_signal(signal.SIGINT)
        out, err = p.communicate(timeout=90)
    finally:
        if p.poll() is None:
            p.kill()
            p.wait()
    print(out[-1500:], err[-1500:], flush=True)
    assert p.returncode == 0, err
    assert "Signal : 2" in err
    s = _snap_fields(tmp_path / "S_00001.bin")
    k = int(np.frombuffer(s["cycle"], "<i4")[0])          # cycles taken
    assert k >= 4
    # the reference's loop rule: n_cycle_lim = k - 1 takes k cycles
    over2 = tmp_path / "over_r.par"
    over2.write_text(f"[cycle]\nn_cycle_lim = {k - 1}\nfinal_output = 1\n"
                     f"prefix_output = {tmp_path}/R\n")
    r = subprocess.run(["timeout", "-k", "10", "120", exe, par, str(over2)], capture_output=True,
                       text=True, cwd=tmp_path)
    assert r.returncode == 0, r.stderr
    a = _snap_fields(tmp_path / "R_00001.bin")
    assert a["time"] == s["time"]
    blocks = sorted(x for x in a if x.endswith("|hydro_cons_field"))
    assert len(blocks) == 8
    for x in blocks + [x.replace("hydro_cons", "field_bf") for x in blocks]:
        np.testing.assert_array_equal(np.frombuffer(a[x], "<f8"), np.frombuffer(s[x], "<f8"))

"""GPU parity of 2-D and 1-D (thin-axis) MHD: the staged HIP kernels
through the C-ABI against the CPU oracle.  FP64 form bit for bit; mixed
form within the relative L1 bar of 1e-5 on rho and B.  Cases: the
Orszag-Tang vortex (periodic, with a z component so the thin face family
moves), a 2-D blast between outflow / reflecting / free sides, the Brio-Wu
shock tube in 1-D with outflow ends, and div B of a long 2-D run.
"""
import math

import numpy as np
import pytest

import kratos_amd
import oracle
from kratos_amd import problems
from kratos_amd.mesh import divb_rel

pytestmark = pytest.mark.gpu

G = 5.0 / 3.0


def rel_l1(a, b):
    d = np.abs(a - b).sum()
    s = np.abs(b).sum()
    return d / s if s > 0 else d


def _ot(N):
    u, bfs = problems.orszag_tang(N)
    xc = (np.arange(N[0]) + 0.5) / N[0]
    yc = (np.arange(N[1]) + 0.5) / N[1]
    tp = 2 * math.pi
    vz = 0.1 * np.sin(tp * (xc[None, :] + yc[:, None]))
    bz = 0.05 * np.cos(tp * (xc[None, :] - 2 * yc[:, None]))
    u[4, 0] = u[0, 0] * vz
    u[1, 0] += 0.5 * u[0, 0] * vz * vz + 0.5 * bz * bz
    return u, (bfs[0], bfs[1], np.stack([bz, bz]))


def run_both(spec, phys, u, b, n_cycles, precision="fp64"):
    print(f"\n[case] thin N={spec.n_cell_global} blk={spec.n_cell_block} bnd={spec.bnd} "
          f"prec={precision} cycles={n_cycles}", flush=True)
    ctx = kratos_amd.Context(spec, phys, precision, path="auto")
    ctx.upload(u, b)
    done = ctx.advance(n_cycles)
    gu, gb = ctx.download()
    gt = ctx.time()
    divb = divb_rel(gb, tuple((spec.x_max[a] - spec.x_min[a]) / spec.n_cell_global[a]
                              for a in range(3)))
    ctx.close()
    ref = oracle.OracleMesh(spec, phys)
    ref.upload(u, b)
    rdone = ref.advance(n_cycles)
    ou, ob = ref.download()
    assert done == rdone
    return (gu, gb, gt, divb), (ou, ob, ref.time())


def _equal(g, o):
    (gu, gb, gt, _), (ou, ob, ot) = g, o
    assert gt[0] == ot[0]
    np.testing.assert_array_equal(gu, ou)
    for a in range(3):
        np.testing.assert_array_equal(gb[a], ob[a])


@pytest.mark.parametrize("blk", [(64, 64, 1), (32, 16, 1)])
def test_orszag_tang_fp64(gpu_lib, blk):
    N = (64, 64, 1)
    spec = kratos_amd.MeshSpec(N, blk)
    phys = kratos_amd.Physics(mhd=True, gamma=G, cfl=0.3)
    u, b = _ot(N)
    g, o = run_both(spec, phys, u, b, 20)
    _equal(g, o)
    assert np.abs(o[1][2][0] - b[2][0]).max() > 1e-3     # the thin face family moved


def test_orszag_tang_mixed(gpu_lib):
    N = (64, 64, 1)
    spec = kratos_amd.MeshSpec(N, (32, 32, 1))
    phys = kratos_amd.Physics(mhd=True, gamma=G, cfl=0.3)
    u, b = _ot(N)
    (gu, gb, gt, divb), (ou, ob, ot) = run_both(spec, phys, u, b, 20, "mixed")
    assert abs(gt[0] - ot[0]) <= 1e-5 * ot[0]
    assert rel_l1(gu[0], ou[0]) <= 1e-5
    for a in range(3):
        assert rel_l1(gb[a], ob[a]) <= 1e-5, a
    assert divb <= 1e-12


@pytest.mark.parametrize("bnd", [("out", "out", "ref", "ref", "per", "per"),
                                 ("fre", "ref", "out", "fre", "out", "out")])
def test_blast_2d_physical_sides_fp64(gpu_lib, bnd):
    N = (48, 32, 1)
    spec = kratos_amd.MeshSpec(N, (16, 16, 1), (-0.75, -0.5, 0.0), (0.75, 0.5, 1.0), bnd)
    phys = kratos_amd.Physics(mhd=True, gamma=G, cfl=0.3)
    u, b = problems.blast(N, b0=(0.6, 0.5, 0.3), x_min=(-0.75, -0.5, -0.5),
                          x_max=(0.75, 0.5, 0.5))
    g, o = run_both(spec, phys, u, b, 15)
    _equal(g, o)


@pytest.mark.parametrize("blk", [(256, 1, 1), (64, 1, 1)])
def test_brio_wu_1d_fp64(gpu_lib, blk):
    N = (256, 1, 1)
    spec = kratos_amd.MeshSpec(N, blk, bnd=("out", "out", "per", "per", "per", "per"))
    phys = kratos_amd.Physics(mhd=True, gamma=2.0, cfl=0.4, t_lim=0.1, top_up=True)
    u, b = problems.brio_wu(N)
    g, o = run_both(spec, phys, u, b, 100000)
    assert o[2][0] == pytest.approx(0.1, abs=0)
    _equal(g, o)


def test_orszag_tang_divb_long(gpu_lib):
    """128^2 to t = 0.2 in the FP64 form: div B at round-off."""
    N = (128, 128, 1)
    spec = kratos_amd.MeshSpec(N, (64, 64, 1))
    phys = kratos_amd.Physics(mhd=True, gamma=G, cfl=0.3, t_lim=0.2, top_up=True)
    u, b = _ot(N)
    print(f"\n[case] thin OT divb N={N}", flush=True)
    ctx = kratos_amd.Context(spec, phys, "fp64")
    ctx.upload(u, b)
    ctx.advance(100000)
    assert ctx.time()[0] == pytest.approx(0.2, abs=0)
    _, gb = ctx.download()
    ctx.close()
    assert divb_rel(gb, (1.0 / N[0], 1.0 / N[1], 1.0)) <= 1e-12


def test_thin_layout_rejected(gpu_lib):
    phys = kratos_amd.Physics(mhd=True)
    with pytest.raises(RuntimeError):
        kratos_amd.Context(kratos_amd.MeshSpec((16, 1, 16), (16, 1, 16)), phys)
    with pytest.raises(RuntimeError):
        kratos_amd.Context(kratos_amd.MeshSpec((16, 1, 16), (16, 1, 16)),
                           kratos_amd.Physics(mhd=False))
    u, b = problems.orszag_tang((16, 16, 2))
    ctx = kratos_amd.Context(kratos_amd.MeshSpec((16, 16, 2), (16, 16, 1)), phys)
    with pytest.raises(RuntimeError, match="thin axis"):
        ctx.upload(u, b)
    ctx.close()

"""GPU parity of the fused hydro stage kernel (k_sweep_hyd: PLM + HLLC
fluxes along x, y, z and the Heun update in one launch per substep) through
the C-ABI against the CPU oracle.  FP64 form bit for bit on periodic,
reflecting / free / outflow and fixed sides and with every tile the
autotuner may pick; mixed form within the relative L1 bar of 1e-5, with
mass, energy and momentum conserved to round-off across tile seams (a seam
face that two tiles solved with different rounding would break that).
"""
import numpy as np
import pytest

import kratos_amd
import oracle
from kratos_amd import problems

pytestmark = pytest.mark.gpu

G = 5.0 / 3.0


def rel_l1(a, b):
    d = np.abs(a - b).sum()
    s = np.abs(b).sum()
    return d / s if s > 0 else d


def run_both(spec, phys, u, n_cycles, precision="fp64", path="fused", tile=(0, 0, 0)):
    print(f"\n[case] hydro N={spec.n_cell_global} blk={spec.n_cell_block} bnd={spec.bnd} "
          f"prec={precision} path={path} tile={tile} cycles={n_cycles}", flush=True)
    ctx = kratos_amd.Context(spec, phys, precision, path=path, tile=tile)
    ctx.upload(u, None)
    done = ctx.advance(n_cycles)
    gu, _ = ctx.download()
    gt = ctx.time()
    st = ctx.stats()
    ctx.close()
    ref = oracle.OracleMesh(spec, phys)
    ref.upload(u, None)
    assert ref.advance(n_cycles) == done
    ou, _ = ref.download()
    return gu, gt, st, ou, ref.time()


@pytest.mark.parametrize("path", ["staged", "fused"])
@pytest.mark.parametrize("N,blk", [((32, 32, 32), (16, 16, 16)), ((64, 64, 64), (32, 32, 32)),
                                   ((32, 48, 16), (16, 16, 8))])
def test_hydro_turbulence_fp64(gpu_lib, N, blk, path):
    spec = kratos_amd.MeshSpec(N, blk)
    phys = kratos_amd.Physics(mhd=False, gamma=G, cfl=0.3)
    u, _, _ = problems.turbulence(N, seed=21)
    gu, gt, st, ou, ot = run_both(spec, phys, u, 10, path=path)
    assert (st["sweep_tile"][0] > 0) == (path == "fused")     # the kernel that ran
    assert gt[0] == ot[0]
    np.testing.assert_array_equal(gu, ou)


@pytest.mark.parametrize("tile", [(32, 8, 1), (16, 16, 1), (16, 8, 1), (8, 8, 1)])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_hydro_tile_variants(gpu_lib, tile, precision):
    N, blk = (64, 64, 64), (32, 32, 32)
    spec = kratos_amd.MeshSpec(N, blk)
    phys = kratos_amd.Physics(mhd=False, gamma=G, cfl=0.3)
    u, _, _ = problems.turbulence(N, seed=22)
    gu, gt, st, ou, ot = run_both(spec, phys, u, 4, precision, tile=tile)
    assert st["sweep_tile"][0:2] == tile[0:2]
    if precision == "fp64":
        np.testing.assert_array_equal(gu, ou)
    else:
        assert rel_l1(gu[0], ou[0]) <= 1e-5
        assert rel_l1(gu[1], ou[1]) <= 1e-5


@pytest.mark.parametrize("bnd", [("ref", "ref", "fre", "out", "per", "per"),
                                 ("out", "fre", "ref", "ref", "out", "ref"),
                                 ("fixd", "out", "per", "per", "fixd", "fixd")])
def test_hydro_physical_sides_fp64(gpu_lib, bnd):
    """Ghost primitives of physical sides in halo blocks (k_phys_fill with
    the hydro boundary functors), read by the fused kernel like a
    neighbour's cells."""
    spec = kratos_amd.MeshSpec((32, 32, 32), (16, 16, 16), bnd=bnd)
    phys = kratos_amd.Physics(mhd=False, gamma=1.4, cfl=0.3, fix_w=(1.1, 0.7, 0.3, -0.2, 0.1))
    u, _, _ = problems.turbulence((32, 32, 32), seed=23)
    gu, gt, st, ou, ot = run_both(spec, phys, u, 10)
    assert st["sweep_tile"][0] > 0
    assert gt[0] == ot[0]
    np.testing.assert_array_equal(gu, ou)


def test_hydro_mixed_conservation(gpu_lib):
    """Mixed form, 64^3 periodic over 8 blocks: within the L1 bar of the
    FP64 oracle, and the totals of the conserved variables kept to
    round-off (every face flux enters its two cells with one value)."""
    N = (64, 64, 64)
    spec = kratos_amd.MeshSpec(N, (32, 32, 32))
    phys = kratos_amd.Physics(mhd=False, gamma=G, cfl=0.3)
    u, _, _ = problems.turbulence(N, seed=24)
    gu, gt, st, ou, ot = run_both(spec, phys, u, 12, "mixed")
    assert abs(gt[0] - ot[0]) <= 1e-5 * ot[0]
    for v in (0, 1):
        assert rel_l1(gu[v], ou[v]) <= 1e-5, v
    for v in range(5):
        assert abs(gu[v].sum() - u[v].sum()) <= 1e-12 * np.abs(u[v]).sum(), v


def test_hydro_mixed_fused_equals_staged(gpu_lib):
    """Mixed form: the fused kernel solves every face with the FP32
    functions the staged kernels use and sums the update in their order, so
    the two paths agree bit for bit (which also lets both kinds of block
    share faces on a refined mesh)."""
    N = (32, 32, 32)
    spec = kratos_amd.MeshSpec(N, (16, 16, 16), bnd=("ref", "out", "per", "per", "fre", "fre"))
    phys = kratos_amd.Physics(mhd=False, gamma=G, cfl=0.3)
    u, _, _ = problems.turbulence(N, seed=25)
    res = {}
    for path in ("staged", "fused"):
        ctx = kratos_amd.Context(spec, phys, "mixed", path=path)
        ctx.upload(u, None)
        assert ctx.advance(8) == 8
        res[path] = ctx.download()[0]
        ctx.close()
    np.testing.assert_array_equal(res["fused"], res["staged"])


@pytest.mark.parametrize("path", ["staged", "fused"])
def test_hydro_ensemble_per_box_dt(gpu_lib, path):
    """Independent periodic hydro boxes in one context, each its own time
    group (own dt, own stop at t_lim; a finished box's workgroups return at
    once): each equals an oracle run of that box alone, bit for bit."""
    N = (16, 16, 16)
    spec = kratos_amd.MeshSpec(N, N)
    phys = kratos_amd.Physics(mhd=False, gamma=G, cfl=0.3, t_lim=0.03, top_up=True)
    ics = [problems.turbulence(N, seed=s, mach=mc)[0] for s, mc in enumerate((0.2, 0.5, 1.0, 2.0))]
    print(f"\n[case] hydro ensemble {len(ics)} x {N} path={path}", flush=True)
    ctx = kratos_amd.Context(spec, phys, "fp64", path=path, members=len(ics))
    for k, u in enumerate(ics):
        ctx.upload(u, None, member=k)
    ctx.advance(100000)
    cycles = []
    for k, u in enumerate(ics):
        gu, _ = ctx.download(member=k)
        t, dt, cyc = ctx.time(member=k)
        cycles.append(cyc)
        ref = oracle.OracleMesh(spec, phys)
        ref.upload(u, None)
        ref.advance(100000)
        ou, _ = ref.download()
        assert t == ref.time()[0] == 0.03
        np.testing.assert_array_equal(gu, ou)
    ctx.close()
    assert len(set(cycles)) > 1, "boxes should need different cycle counts"


def test_hydro_fused_needs_3d_tile(gpu_lib):
    """Blocks no tile divides, 2-D meshes and the waves-per-SIMD variants
    (an MHD-only form) cannot take the fused hydro path: the context
    refuses them when the mesh is committed."""
    phys = kratos_amd.Physics(mhd=False)
    for N, blk, tile in (((24, 16, 16), (12, 8, 8), (0, 0, 0)),
                         ((512, 2, 1), (512, 2, 1), (0, 0, 0)),
                         ((32, 32, 32), (16, 16, 16), (16, 8, 4))):
        ctx = kratos_amd.Context(kratos_amd.MeshSpec(N, blk), phys, path="fused", tile=tile)
        with pytest.raises(RuntimeError, match="no fused tile"):
            ctx.upload(np.ones((5, N[2], N[1], N[0])))
        ctx.close()

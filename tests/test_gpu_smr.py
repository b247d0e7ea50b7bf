"""GPU parity on statically refined meshes: the staged HIP path through the
C-ABI against the CPU oracle, same per-block initial states.  FP64 form:
every conserved value and face field bit-identical; mixed form: density L1
<= 1e-5 relative.  Cases cover hydro and MHD, two and three levels, a
physical (outflow / reflecting) hydro box, and a refined mesh whose blocks
all sit on level 0 (the same-level path fed by the refined-mesh builder).
Blocks of two sizes: a mesh whose groups of eight same-level leaves are
given as blocks of twice the size gives the leaves' states bit for bit.
MHD runs the fused kernel on blocks with no cross-level direction and its
producer form (fluxes and EMFs into HBM) plus the staged update on the
others (and everything staged when asked); 8^3 blocks take the 8x8
producer tile, 16^3 blocks the 16x8 one."""
import numpy as np
import pytest

import kratos_amd
import oracle
from kratos_amd import problems
from kratos_amd.mesh import MergedMeshSpec, Physics, RefinedMeshSpec, divb_rel_blocks

pytestmark = pytest.mark.gpu
G = 5.0 / 3.0
TWO = [(1, (0.3, 0.3, 0.3), (0.6, 0.6, 0.6))]
THREE = [(1, (0.25, 0.25, 0.25), (0.7, 0.7, 0.7)), (2, (0.45, 0.45, 0.45), (0.55, 0.55, 0.55))]
BLAST3 = [(1, (-0.25,) * 3, (0.2,) * 3), (2, (-0.05,) * 3, (0.05,) * 3)]
WALL = [(1, (0.0, 0.3, 0.3), (0.35, 0.6, 0.6))]        # level 1 reaching the x- side


def _faces(spec, ba):
    n = spec.n_cell_block
    f = np.asarray(ba).reshape(n[2] + 5, n[1] + 5, n[0] + 5, 3)
    return (f[2:n[2] + 2, 2:n[1] + 2, 2:n[0] + 3, 0], f[2:n[2] + 2, 2:n[1] + 3, 2:n[0] + 2, 1],
            f[2:n[2] + 3, 2:n[1] + 2, 2:n[0] + 2, 2])


def _run(spec, phys, blocks, cycles, precision="fp64", path="auto"):
    print(f"\n[case] smr blocks={spec.n_blocks} levels={spec.n_levels} mhd={phys.mhd} "
          f"prec={precision} path={path} cycles={cycles}", flush=True)
    ctx = kratos_amd.Context(spec, phys, precision, path=path)
    ctx.upload_blocks(blocks)
    done = ctx.advance(cycles)
    g = ctx.download_blocks()
    gt = ctx.time()
    # div B bar, both forms: max |div B| dx / max |B| <= 1e-12
    divb = divb_rel_blocks(spec, g) if phys.mhd else 0.0
    ctx.close()
    o = oracle.OracleMesh(spec, phys, n_threads=16)
    o.upload_blocks(blocks)
    assert o.advance(cycles) == done
    return g, o.download_blocks(), gt, o.time(), divb


def _assert_equal(spec, g, o, mhd):
    for b in range(spec.n_blocks):
        gu = np.asarray(g[b][0]).reshape(-1, 8 if mhd else 5)[:, :5]
        ou = np.asarray(o[b][0]).reshape(-1, 8 if mhd else 5)[:, :5]
        np.testing.assert_array_equal(gu, ou, err_msg=f"block {b}")
        if mhd:
            for x, y in zip(_faces(spec, g[b][1]), _faces(spec, o[b][1])):
                np.testing.assert_array_equal(x, y, err_msg=f"block {b} faces")


@pytest.mark.parametrize("mhd,path", [(False, "auto"), (False, "staged"), (True, "auto"),
                                      (True, "staged")])
def test_smr_two_levels_fp64(gpu_lib, mhd, path):
    """path auto: blocks away from level changes run the fused kernel (MHD
    or hydro), the others the staged kernels (MHD: the fused producer form
    for fluxes and EMFs); staged: every block staged."""
    spec = RefinedMeshSpec((32, 32, 32), (8, 8, 8), refine=TWO)
    phys = Physics(mhd=mhd, gamma=G, cfl=0.3)
    g, o, gt, ot, divb = _run(spec, phys, problems.smooth_blocks(spec, seed=5, mhd=mhd), 6,
                              path=path)
    assert gt[0] == ot[0]
    _assert_equal(spec, g, o, mhd)
    if mhd:
        assert divb <= 1e-12


def test_smr_16_blocks_mhd_fp64(gpu_lib):
    """16^3 blocks: the 16x8 producer tile on the blocks at level changes."""
    spec = RefinedMeshSpec((64, 64, 64), (16, 16, 16), refine=TWO)
    phys = Physics(mhd=True, gamma=G, cfl=0.3)
    g, o, gt, ot, divb = _run(spec, phys, problems.smooth_blocks(spec, seed=8, mhd=True), 4)
    assert gt[0] == ot[0]
    _assert_equal(spec, g, o, True)
    assert divb <= 1e-12


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_smr_32_blocks_mhd(gpu_lib, precision):
    """32^3 blocks (the block size of the refined blast's outer levels):
    one root block of eight refined, so every level-0 block meets the
    level-1 octant across a face, an edge or a corner (periodic wrap
    included).  FP64 bit-identical; mixed: density L1 <= 1e-5 relative."""
    spec = RefinedMeshSpec((64, 64, 64), (32, 32, 32),
                           refine=[(1, (0.3, 0.3, 0.3), (0.45, 0.45, 0.45))])
    assert spec.n_blocks == 15
    phys = Physics(mhd=True, gamma=G, cfl=0.3)
    g, o, gt, ot, divb = _run(spec, phys, problems.smooth_blocks(spec, seed=12, mhd=True), 3,
                              precision)
    assert divb <= 1e-12
    if precision == "fp64":
        assert gt[0] == ot[0]
        _assert_equal(spec, g, o, True)
        return
    num = den = 0.0
    for b in range(spec.n_blocks):
        _, dx = spec.geometry(b)
        v = dx[0] * dx[1] * dx[2]
        gr = np.asarray(g[b][0]).reshape(-1, 8)[:, 0]
        orr = np.asarray(o[b][0]).reshape(-1, 8)[:, 0]
        num += np.abs(gr - orr).sum() * v
        den += np.abs(orr).sum() * v
    assert num / den <= 1e-5


def _merged_run(m, phys, blocks, cycles, precision="fp64", path="auto"):
    print(f"\n[case] merged blocks={m.n_blocks} (leaves {m.base.n_blocks}) mhd={phys.mhd} "
          f"prec={precision} path={path} cycles={cycles}", flush=True)
    ctx = kratos_amd.Context(m, phys, precision, path=path)
    ctx.upload_blocks(m.merge_blocks(blocks))
    done = ctx.advance(cycles)
    g = m.split_blocks(ctx.download_blocks())
    t = ctx.time()
    ctx.close()
    return g, t, done


@pytest.mark.parametrize("mhd,path", [(True, "auto"), (True, "staged"), (False, "auto")])
def test_merged_blocks_fp64(gpu_lib, mhd, path):
    """Blocks of two sizes in one mesh: the two-level mesh with its groups
    of eight same-level leaves given as blocks of 16^3 beside blocks of 8^3.
    FP64 bit for bit against the oracle on the leaves (the same cells)."""
    base = RefinedMeshSpec((32, 32, 32), (8, 8, 8), refine=TWO)
    m = MergedMeshSpec(base)
    assert 0 < sum(m.merged(b) for b in range(m.n_blocks)) < m.n_blocks
    phys = Physics(mhd=mhd, gamma=G, cfl=0.3)
    # one value per shared face, as a merged block holds it
    blocks = m.split_blocks(m.merge_blocks(problems.smooth_blocks(base, seed=13, mhd=mhd)))
    g, gt, done = _merged_run(m, phys, blocks, 4, path=path)
    o = oracle.OracleMesh(base, phys, n_threads=16)
    o.upload_blocks(blocks)
    assert o.advance(4) == done == 4
    assert gt[0] == o.time()[0]
    _assert_equal(base, g, o.download_blocks(), mhd)
    if mhd:
        assert divb_rel_blocks(base, g) <= 1e-12


def test_merged_blast_mixed_equals_leaves(gpu_lib):
    """The 3-level blast in the mixed form with its full groups of eight
    given as blocks of 16^3: the same state, bit for bit, as the mesh of
    8^3 leaves."""
    base = RefinedMeshSpec((32, 32, 32), (8, 8, 8), refine=BLAST3, x_min=(-0.5,) * 3,
                           x_max=(0.5,) * 3)
    m = MergedMeshSpec(base)
    phys = Physics(mhd=True, gamma=G, cfl=0.3)
    blocks = m.split_blocks(m.merge_blocks(problems.smooth_blocks(base, seed=14, mhd=True)))
    g, gt, done = _merged_run(m, phys, blocks, 5, "mixed")
    print(f"[case] leaves blocks={base.n_blocks} mixed cycles=5", flush=True)
    ctx = kratos_amd.Context(base, phys, "mixed")
    ctx.upload_blocks(blocks)
    assert ctx.advance(5) == done
    ref = ctx.download_blocks()
    assert ctx.time() == gt
    ctx.close()
    _assert_equal(base, g, ref, True)


def test_block_size_neither_n_nor_twice_rejected(gpu_lib):
    base = RefinedMeshSpec((32, 32, 32), (8, 8, 8), refine=TWO)

    class Bad(MergedMeshSpec):
        def block_descs(self, members=1):
            d = super().block_descs(members)
            b = next(b for b in range(self.n_blocks) if self.merged(b))
            d[b].n_ceff[2] = 8                   # 16 x 16 x 8: neither size
            return d

    with pytest.raises(kratos_amd.KratosError, match="neither n_cell_block nor twice"):
        kratos_amd.Context(Bad(base), Physics(mhd=True, gamma=G, cfl=0.3))


def test_smr_three_levels_mhd_fp64(gpu_lib):
    spec = RefinedMeshSpec((32, 32, 32), (8, 8, 8), refine=THREE)
    assert spec.n_levels == 3
    phys = Physics(mhd=True, gamma=G, cfl=0.3)
    g, o, gt, ot, divb = _run(spec, phys, problems.smooth_blocks(spec, seed=7), 4)
    assert gt[0] == ot[0]
    _assert_equal(spec, g, o, True)
    assert divb <= 1e-12


def test_smr_blast_fp64(gpu_lib):
    spec = RefinedMeshSpec((32, 32, 32), (8, 8, 8), refine=BLAST3, x_min=(-0.5,) * 3,
                           x_max=(0.5,) * 3)
    phys = Physics(mhd=True, gamma=G, cfl=0.3)
    g, o, gt, ot, divb = _run(spec, phys, problems.blast_blocks(spec), 5)
    _assert_equal(spec, g, o, True)
    assert divb <= 1e-12


def test_smr_hydro_physical_sides_fp64(gpu_lib):
    spec = RefinedMeshSpec((32, 32, 32), (8, 8, 8), refine=TWO,
                           bnd=("out", "out", "ref", "ref", "fre", "fre"))
    phys = Physics(mhd=False, gamma=1.4, cfl=0.4)
    g, o, gt, ot, _ = _run(spec, phys, problems.smooth_blocks(spec, seed=2, mhd=False), 6)
    assert gt[0] == ot[0]
    _assert_equal(spec, g, o, False)


@pytest.mark.parametrize("mhd", [False, True])
def test_single_level_refined_spec_fp64(gpu_lib, mhd):
    spec = RefinedMeshSpec((16, 16, 16), (8, 8, 8))
    phys = Physics(mhd=mhd, gamma=G, cfl=0.3)
    g, o, gt, ot, _ = _run(spec, phys, problems.smooth_blocks(spec, seed=3, mhd=mhd), 4)
    _assert_equal(spec, g, o, mhd)


def test_smr_mixed_l1(gpu_lib):
    spec = RefinedMeshSpec((32, 32, 32), (8, 8, 8), refine=THREE)
    phys = Physics(mhd=True, gamma=G, cfl=0.3)
    g, o, _, _, divb = _run(spec, phys, problems.smooth_blocks(spec, seed=11), 6, "mixed")
    num = den = 0.0
    for b in range(spec.n_blocks):
        _, dx = spec.geometry(b)
        v = dx[0] * dx[1] * dx[2]
        gr = np.asarray(g[b][0]).reshape(-1, 8)[:, 0]
        orr = np.asarray(o[b][0]).reshape(-1, 8)[:, 0]
        num += np.abs(gr - orr).sum() * v
        den += np.abs(orr).sum() * v
    assert num / den <= 1e-5
    assert divb <= 1e-12


def test_smr_hydro_mixed_fused_equals_staged(gpu_lib):
    """Hydro, three levels, mixed form.  path auto runs the fused kernel on
    the blocks away from level changes and the staged kernels with the
    cross-level corrections on the others; it must equal the all-staged run
    bit for bit (both solve every face with the same FP32 functions), and
    the density L1 against the FP64 oracle stays within 1e-5.  The totals of
    the periodic box drift only by the FP32 rounding of the restricted
    fluxes at level changes (about 1e-12 here; FP64 keeps them to 1e-15)."""
    spec = RefinedMeshSpec((32, 32, 32), (8, 8, 8), refine=THREE)
    phys = Physics(mhd=False, gamma=G, cfl=0.3)
    blocks = problems.smooth_blocks(spec, seed=12, mhd=False)
    g, o, _, _, _ = _run(spec, phys, blocks, 6, "mixed")
    ctx = kratos_amd.Context(spec, phys, "mixed", path="staged")
    ctx.upload_blocks(blocks)
    ctx.advance(6)
    gs = ctx.download_blocks()
    ctx.close()
    num = den = 0.0
    tot0 = np.zeros(5)
    tot1 = np.zeros(5)
    for b in range(spec.n_blocks):
        np.testing.assert_array_equal(g[b][0], gs[b][0], err_msg=f"block {b}")
        _, dx = spec.geometry(b)
        v = dx[0] * dx[1] * dx[2]
        gu = np.asarray(g[b][0]).reshape(-1, 5)
        ou = np.asarray(o[b][0]).reshape(-1, 5)
        num += np.abs(gu[:, 0] - ou[:, 0]).sum() * v
        den += np.abs(ou[:, 0]).sum() * v
        tot0 += np.asarray(blocks[b][0]).reshape(-1, 5).sum(0) * v
        tot1 += gu.sum(0) * v
    assert num / den <= 1e-5
    for k in range(5):
        assert abs(tot1[k] - tot0[k]) <= 1e-10 * max(1.0, abs(tot0[k])), k


def test_unbalanced_mesh_rejected(gpu_lib):
    """kr_add_blocks / the first upload refuses a block list whose levels
    differ by two across a ghost region, with a message (no launch)."""
    from test_oracle_smr import _unbalanced_spec
    spec = _unbalanced_spec()
    phys = Physics(mhd=True, gamma=G, cfl=0.3)
    print("\n[case] unbalanced refined mesh (host-side rejection)", flush=True)
    ctx = kratos_amd.Context(spec, phys, "fp64")
    with pytest.raises(kratos_amd.KratosError, match="balanced|within one level"):
        ctx.upload_blocks(problems.smooth_blocks(spec, seed=1))
    ctx.close()


@pytest.mark.parametrize("bnd", [("out", "fre", "ref", "ref", "fre", "out"), ("ref",) * 6])
def test_smr_mhd_physical_sides_fp64(gpu_lib, bnd):
    """Refined MHD meshes with physical sides (staged kernels: the side
    functors applied where cells and face fields are read, the boundary-plane
    EMF, the cross-level halo and corrections): FP64 bit-exact with the
    oracle, div B at round-off."""
    spec = RefinedMeshSpec((32, 32, 32), (8, 8, 8), refine=WALL, bnd=bnd)
    assert spec.n_levels == 2 and min(x for L, x, y, z in spec.leaves if L == 1) == 0
    phys = Physics(mhd=True, gamma=G, cfl=0.3)
    g, o, gt, ot, divb = _run(spec, phys, problems.smooth_blocks(spec, seed=7), 6)
    assert gt[0] == ot[0]
    _assert_equal(spec, g, o, True)
    assert divb <= 1e-12


def test_merged_blocks_physical_sides_fp64(gpu_lib):
    """Blocks of two sizes on a refined MHD mesh with physical sides (the
    staged kernels throughout): the groups of eight leaves as blocks of 16^3,
    some of them on the sides; FP64 bit for bit against the oracle."""
    bnd = ("out", "fre", "ref", "ref", "fre", "out")
    base = RefinedMeshSpec((32, 32, 32), (8, 8, 8), refine=WALL, bnd=bnd)
    m = MergedMeshSpec(base)
    on_side = [b for b in range(m.n_blocks) if m.merged(b) and
               any(base.reg(m.parts[b][0])[a] == 0 for a in range(3))]
    assert on_side
    phys = Physics(mhd=True, gamma=G, cfl=0.3)
    blocks = m.split_blocks(m.merge_blocks(problems.smooth_blocks(base, seed=15)))
    g, gt, done = _merged_run(m, phys, blocks, 4)
    o = oracle.OracleMesh(base, phys, n_threads=16)
    o.upload_blocks(blocks)
    assert o.advance(4) == done == 4
    assert gt[0] == o.time()[0]
    _assert_equal(base, g, o.download_blocks(), True)
    assert divb_rel_blocks(base, g) <= 1e-12


def test_smr_mhd_physical_sides_mixed(gpu_lib):
    spec = RefinedMeshSpec((32, 32, 32), (8, 8, 8), refine=WALL,
                           bnd=("out", "fre", "ref", "ref", "fre", "out"))
    phys = Physics(mhd=True, gamma=G, cfl=0.3)
    g, o, _, _, divb = _run(spec, phys, problems.smooth_blocks(spec, seed=7), 6, "mixed")
    num = den = 0.0
    for b in range(spec.n_blocks):
        _, dx = spec.geometry(b)
        v = dx[0] * dx[1] * dx[2]
        gr = np.asarray(g[b][0]).reshape(-1, 8)[:, 0]
        orr = np.asarray(o[b][0]).reshape(-1, 8)[:, 0]
        num += np.abs(gr - orr).sum() * v
        den += np.abs(orr).sum() * v
    assert num / den <= 1e-5
    assert divb <= 1e-12

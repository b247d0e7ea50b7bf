"""Blocks of two sizes on the host side: a refined mesh's groups of eight
same-level leaves at an even index become blocks of twice the block size
(kratos_amd.MergedMeshSpec); the per-block arrays split back into the
leaves' arrays exactly (every cell, every face of the leaves)."""
import numpy as np

from kratos_amd import problems
from kratos_amd.mesh import MergedMeshSpec, RefinedMeshSpec

TWO = [(1, (0.3, 0.3, 0.3), (0.6, 0.6, 0.6))]
BLAST3 = [(1, (-0.25,) * 3, (0.2,) * 3), (2, (-0.05,) * 3, (0.05,) * 3)]


def _faces(n, ba):
    f = np.asarray(ba).reshape(n[2] + 5, n[1] + 5, n[0] + 5, 3)
    return (f[2:n[2] + 2, 2:n[1] + 2, 2:n[0] + 3, 0], f[2:n[2] + 2, 2:n[1] + 3, 2:n[0] + 2, 1],
            f[2:n[2] + 3, 2:n[1] + 2, 2:n[0] + 2, 2])


def test_merge_groups_and_descriptors():
    base = RefinedMeshSpec((32, 32, 32), (8, 8, 8), refine=BLAST3, x_min=(-0.5,) * 3,
                           x_max=(0.5,) * 3)
    m = MergedMeshSpec(base)
    merged = [b for b in range(m.n_blocks) if m.merged(b)]
    assert 0 < len(merged) < m.n_blocks
    assert m.n_blocks + 7 * len(merged) == base.n_blocks
    assert sorted(i for p in m.parts for i in p) == list(range(base.n_blocks))
    descs = m.block_descs()
    for b in merged:
        lo = m.parts[b][0]
        L, r = base.level(lo), base.reg(lo)
        assert all(v % 2 == 0 for v in r)
        for o, leaf in enumerate(m.parts[b]):     # octant o: bits x, y, z
            assert base.leaves[leaf] == (L,) + tuple(r[a] + ((o >> a) & 1) for a in range(3))
        d = descs[b]
        assert tuple(d.n_ceff) == (16, 16, 16) and d.level == L and tuple(d.reg) == r
        assert tuple(d.xf0) == tuple(base.geometry(lo)[0])


def test_merge_split_round_trip():
    base = RefinedMeshSpec((32, 32, 32), (8, 8, 8), refine=TWO)
    m = MergedMeshSpec(base)
    n = base.n_cell_block
    blocks = problems.smooth_blocks(base, seed=3, mhd=True)
    per_block = m.merge_blocks(blocks)
    for b in range(m.n_blocks):
        nb = m.block_cells(b)
        assert np.asarray(per_block[b][0]).size == nb[0] * nb[1] * nb[2] * 8
        assert np.asarray(per_block[b][1]).size == (nb[0] + 5) * (nb[1] + 5) * (nb[2] + 5) * 3
    back = m.split_blocks(per_block)
    for a, b in zip(blocks, back):
        np.testing.assert_array_equal(np.asarray(a[0]), np.asarray(b[0]))
        for x, y in zip(_faces(n, a[1]), _faces(n, b[1])):
            np.testing.assert_array_equal(x, y)

"""Pins the oracle's MHD physical sides on statically refined meshes (the
MHD boundary functors, mhd/boundary/phys.h:144-251, with the cross-level
halo of hydro/boundary/comm_functions.h:329-452 and the side kernels run
again after the prolongation, mhd/boundary/keeper.h:288-293):

* A reflecting wall is a mirror, also where the mesh is refined at the
  wall: a 1-D magnetised flow between walls on a mesh refined next to the
  left wall equals the left half of the periodic box twice as long, refined
  at the mirrored place, holding the state and its mirror image (1e-12).
* A reflecting refined box conserves mass to round-off and keeps div B at
  round-off.
* Free sides keep a uniform moving magnetised state uniform across level
  changes and at the sides (1e-13).
* One level through the refined-mesh code equals the uniform path bit for
  bit with out / fre / ref sides (the second side pass changes nothing
  there).
"""
import math

import numpy as np
import pytest

import oracle
from kratos_amd import problems
from kratos_amd.mesh import MeshSpec, Physics, RefinedMeshSpec, divb_rel_blocks

from test_oracle_mhd_bnd import _slab_blocks
from test_oracle_smr import _faces

G = 5.0 / 3.0


def _mirror(spec_w, spec_t):
    """Leaf of the twin holding each wall-mesh leaf (same level and logical
    index), and the twin's mirror symmetry about x = 1."""
    where = {leaf: i for i, leaf in enumerate(spec_t.leaves)}
    for L, x, y, z in spec_t.leaves:
        nx = spec_t.nblk[0] << L
        assert (L, nx - 1 - x, y, z) in where, "twin tree not mirror symmetric"
    return [where[leaf] for leaf in spec_w.leaves]


def test_refined_reflecting_wall_equals_mirrored_periodic_box():
    wall = RefinedMeshSpec((32, 16, 16), (8, 8, 8), refine=[(1, (0.0, 0.0, 0.0), (0.2, 1.0, 1.0))],
                           x_max=(1.0, 1.0, 1.0), bnd=("ref", "ref", "per", "per", "per", "per"))
    twin = RefinedMeshSpec((64, 16, 16), (8, 8, 8),
                           refine=[(1, (0.0, 0.0, 0.0), (0.2, 1.0, 1.0)),
                                   (1, (1.8, 0.0, 0.0), (2.0, 1.0, 1.0))],
                           x_max=(2.0, 1.0, 1.0))
    assert wall.n_levels == 2 and twin.n_levels == 2
    idx = _mirror(wall, twin)
    phys = Physics(mhd=True, gamma=G, cfl=0.3)
    ow = oracle.OracleMesh(wall, phys, n_threads=8)
    ow.upload_blocks(_slab_blocks(wall))
    ot = oracle.OracleMesh(twin, phys, n_threads=8)
    ot.upload_blocks(_slab_blocks(twin, mirror_from=1.0))
    assert ow.advance(16) == 16 and ot.advance(16) == 16
    assert ow.time()[0] == pytest.approx(ot.time()[0], rel=1e-14)
    gw, gt = ow.download_blocks(), ot.download_blocks()
    vmax = 0.0
    for b, tb in enumerate(idx):
        uw = np.asarray(gw[b][0]).reshape(-1, 8)[:, :5]
        ut = np.asarray(gt[tb][0]).reshape(-1, 8)[:, :5]
        np.testing.assert_allclose(uw, ut, rtol=0, atol=1e-12, err_msg=f"block {b}")
        vmax = max(vmax, float(np.abs(uw[:, 2]).max()))
    assert vmax > 0.05                  # the flow reached the walls


def test_refined_reflecting_box_conserves_mass_and_divb():
    spec = RefinedMeshSpec((16, 16, 16), (8, 8, 8), bnd=("ref",) * 6,
                           refine=[(1, (0.25, 0.25, 0.25), (0.7, 0.7, 0.7))])
    assert spec.n_levels == 2
    phys = Physics(mhd=True, gamma=G, cfl=0.3)
    blocks = problems.smooth_blocks(spec, seed=6)
    o = oracle.OracleMesh(spec, phys, n_threads=8)
    o.upload_blocks(blocks)

    def mass(bs):
        tot = 0.0
        for b, (ua, _) in enumerate(bs):
            _, dx = spec.geometry(b)
            tot += np.asarray(ua).reshape(-1, 8)[:, 0].sum() * dx[0] * dx[1] * dx[2]
        return tot

    m0 = mass(blocks)
    assert o.advance(10) == 10
    out = o.download_blocks()
    assert abs(mass(out) - m0) <= 1e-13 * m0
    assert divb_rel_blocks(spec, out) <= 1e-12


def test_refined_free_sides_keep_a_uniform_moving_state():
    spec = RefinedMeshSpec((16, 16, 16), (8, 8, 8), bnd=("fre",) * 6,
                           refine=[(1, (0.0, 0.3, 0.3), (0.4, 0.6, 0.6))])
    assert spec.n_levels == 2
    phys = Physics(mhd=True, gamma=G, cfl=0.3)
    blocks = problems.wave_blocks(spec, amp=0.0, vel=(0.3, -0.2, 0.1), mhd=True,
                                  b0=(0.5, 0.25, -0.4))
    o = oracle.OracleMesh(spec, phys, n_threads=8)
    o.upload_blocks(blocks)
    assert o.advance(5) == 5
    for (ua, _), (ub, _) in zip(o.download_blocks(), blocks):
        np.testing.assert_allclose(ua, ub, rtol=1e-13, atol=1e-14)


def test_single_level_with_sides_matches_uniform_path():
    bnd = ("out", "fre", "ref", "ref", "fre", "out")
    rs = RefinedMeshSpec((16, 16, 16), (8, 8, 8), bnd=bnd)
    us = MeshSpec((16, 16, 16), (8, 8, 8), bnd=bnd)
    phys = Physics(mhd=True, gamma=G, cfl=0.3)
    blocks = problems.smooth_blocks(rs, seed=5)
    perm = [us.bid(rs.reg(b)) for b in range(rs.n_blocks)]
    o1 = oracle.OracleMesh(rs, phys, n_threads=4)
    o1.upload_blocks(blocks)
    assert o1.advance(4) == 4
    o2 = oracle.OracleMesh(us, phys, n_threads=4)
    ub = [None] * rs.n_blocks
    for b in range(rs.n_blocks):
        ub[perm[b]] = blocks[b]
    o2.upload_blocks(ub)
    assert o2.advance(4) == 4
    r1, r2 = o1.download_blocks(), o2.download_blocks()
    for b in range(rs.n_blocks):
        np.testing.assert_array_equal(r1[b][0], r2[perm[b]][0])
        for x, y in zip(_faces(rs, r1[b][1]), _faces(rs, r2[perm[b]][1])):
            np.testing.assert_array_equal(x, y)
    assert o1.time() == o2.time()

"""Pins the oracle's statically refined (SMR) mesh against what the
mathematics and the tree rules fix.

* The tree builder (tree.cpp insert / fill_base_level): leaves tile the
  domain exactly and every pair of touching leaves differs by at most one
  level (faces, edges and corners).
* One level through the SMR code path is bit-identical to the uniform
  path (the cross-level code only acts where levels meet).
* Periodic SMR boxes: mass, momentum and energy conserved to round-off
  (the coarse face takes the area-weighted fine flux and the skipped face
  is added back); div B stays at round-off in every block (edge EMF
  restriction and write-back keep CT telescoping).
* The coarse face B on a coarse/fine interface stays the area mean of the
  four fine faces (the shared edges carry one EMF).
* A density wave advected through the refined region: the SMR error lies
  between the uniform coarse and the uniform fine errors -- a wrong
  prolongation slope or restriction weight puts an O(1) error at the
  interface and fails it.
* A uniform state stays uniform to round-off on a refined mesh.
"""
import math

import numpy as np
import pytest

import oracle
from kratos_amd import problems
from kratos_amd.mesh import MeshSpec, Physics, RefinedMeshSpec

G = 5.0 / 3.0
REFINE = [(1, (0.3, 0.3, 0.3), (0.6, 0.6, 0.6))]


def _vol(spec, b):
    _, dx = spec.geometry(b)
    return dx[0] * dx[1] * dx[2]


def _totals(spec, blocks, mhd):
    nv = 8 if mhd else 5
    return sum(np.asarray(ua).reshape(-1, nv)[:, :5].sum(0) * _vol(spec, b)
               for b, (ua, _) in enumerate(blocks))


def _faces(spec, ba):
    n = spec.n_cell_block
    f = np.asarray(ba).reshape(n[2] + 5, n[1] + 5, n[0] + 5, 3)
    return (f[2:n[2] + 2, 2:n[1] + 2, 2:n[0] + 3, 0], f[2:n[2] + 2, 2:n[1] + 3, 2:n[0] + 2, 1],
            f[2:n[2] + 3, 2:n[1] + 2, 2:n[0] + 2, 2])


def _divb(spec, blocks):
    m = 0.0
    for b, (_, ba) in enumerate(blocks):
        _, dx = spec.geometry(b)
        bx, by, bz = _faces(spec, ba)
        d = ((bx[:, :, 1:] - bx[:, :, :-1]) / dx[0] + (by[:, 1:] - by[:, :-1]) / dx[1]
             + (bz[1:] - bz[:-1]) / dx[2])
        m = max(m, float(np.abs(d).max()))
    return m


def test_tree_balanced_and_tiling():
    spec = RefinedMeshSpec((64, 64, 64), (8, 8, 8),
                           refine=[(1, (0.3,) * 3, (0.6,) * 3), (2, (0.45,) * 3, (0.55,) * 3)])
    assert spec.n_levels == 3
    # exact tiling: leaf volumes (in level-2 cells) add up to the domain
    tot = sum(8 ** (2 - spec.level(b)) for b in range(spec.n_blocks))
    assert tot == (8 * 4) ** 3
    # 2:1 balance over faces, edges and corners (periodic box)
    lv = {}
    for b in range(spec.n_blocks):
        L, r = spec.level(b), spec.reg(b)
        for z in range(r[2] << (2 - L), (r[2] + 1) << (2 - L)):
            for y in range(r[1] << (2 - L), (r[1] + 1) << (2 - L)):
                for x in range(r[0] << (2 - L), (r[0] + 1) << (2 - L)):
                    lv[(x, y, z)] = L
    assert len(lv) == 32 ** 3
    for (x, y, z), L in lv.items():
        for d in ((1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 0), (1, 0, 1), (0, 1, 1), (1, 1, 1),
                  (1, -1, 0), (1, 0, -1), (0, 1, -1), (1, 1, -1), (1, -1, 1), (-1, 1, 1)):
            q = ((x + d[0]) % 32, (y + d[1]) % 32, (z + d[2]) % 32)
            assert abs(lv[q] - L) <= 1


@pytest.mark.parametrize("mhd", [False, True])
def test_single_level_matches_uniform_path(mhd):
    rs = RefinedMeshSpec((16, 16, 16), (8, 8, 8))
    us = MeshSpec((16, 16, 16), (8, 8, 8))
    phys = Physics(mhd=mhd, gamma=G, cfl=0.3)
    blocks = problems.smooth_blocks(rs, seed=3, mhd=mhd)
    perm = [us.bid(rs.reg(b)) for b in range(rs.n_blocks)]
    o1 = oracle.OracleMesh(rs, phys, n_threads=4)
    o1.upload_blocks(blocks)
    o1.advance(3)
    o2 = oracle.OracleMesh(us, phys, n_threads=4)
    ub = [None] * rs.n_blocks
    for b in range(rs.n_blocks):
        ub[perm[b]] = blocks[b]
    o2.upload_blocks(ub)
    o2.advance(3)
    r1, r2 = o1.download_blocks(), o2.download_blocks()
    for b in range(rs.n_blocks):
        np.testing.assert_array_equal(r1[b][0], r2[perm[b]][0])
        if mhd:          # faces of the block (ghost layers are exchange scratch)
            for x, y in zip(_faces(rs, r1[b][1]), _faces(rs, r2[perm[b]][1])):
                np.testing.assert_array_equal(x, y)
    assert o1.time() == o2.time()


@pytest.mark.parametrize("mhd", [False, True])
def test_conservation_and_divb(mhd):
    spec = RefinedMeshSpec((32, 32, 32), (8, 8, 8), refine=REFINE)
    assert spec.n_levels == 2
    phys = Physics(mhd=mhd, gamma=G, cfl=0.3)
    blocks = problems.smooth_blocks(spec, seed=5, mhd=mhd)
    o = oracle.OracleMesh(spec, phys, n_threads=8)
    o.upload_blocks(blocks)
    t0 = _totals(spec, blocks, mhd)
    assert o.advance(6) == 6
    out = o.download_blocks()
    t1 = _totals(spec, out, mhd)
    scale = np.array([abs(t0[0]), abs(t0[1]), abs(t0[0]), abs(t0[0]), abs(t0[0])])
    assert np.all(np.abs(t1 - t0) <= 1e-13 * scale), (t1 - t0) / scale
    if mhd:
        assert _divb(spec, out) < 1e-11


def _coarse_fine_face_gap(spec, blocks):
    """Max |B_n(coarse face) - mean of the 4 fine faces| over all interfaces."""
    where = {(spec.level(b),) + tuple(spec.reg(b)): b for b in range(spec.n_blocks)}
    n = spec.n_cell_block[0]
    gap = 0.0
    for b in range(spec.n_blocks):
        L, r = spec.level(b), spec.reg(b)
        fc = _faces(spec, blocks[b][1])
        for a in range(3):
            for s in (0, 1):
                d = [0, 0, 0]
                d[a] = 1 if s else -1
                q = [r[x] + d[x] for x in range(3)]
                q = [q[x] % (4 << L) for x in range(3)]
                kids = [(L + 1,) + tuple(2 * q[x] + ((c >> x) & 1) * (x != a) + (0 if s else 1) * (x == a)
                                         for x in range(3)) for c in range(8)]
                kids = [k for k in set(kids) if k in where]
                if not kids:
                    continue
                for k in kids:
                    f = _faces(spec, blocks[where[k]][1])[a]
                    f = np.moveaxis(f, 2 - a, 0)        # face-normal index first
                    fine = f[0 if s else n]
                    cf = np.moveaxis(fc[a], 2 - a, 0)[n if s else 0]
                    # coarse faces covered by this fine block
                    o = [(k[1 + x] % 2) * (n // 2) for x in range(3)]
                    t = [x for x in (2, 1, 0) if x != a]   # array axes order (z, y, x) minus a
                    sl = tuple(slice(o[x], o[x] + n // 2) for x in t)
                    cs = cf[sl]
                    mean = 0.25 * (fine[0::2, 0::2] + fine[1::2, 0::2] + fine[0::2, 1::2]
                                   + fine[1::2, 1::2])
                    gap = max(gap, float(np.abs(cs - mean).max()))
    return gap


def test_coarse_face_stays_mean_of_fine_faces():
    spec = RefinedMeshSpec((32, 32, 32), (8, 8, 8), refine=REFINE)
    phys = Physics(mhd=True, gamma=G, cfl=0.3)
    blocks = problems.smooth_blocks(spec, seed=9, mhd=True, b_amp=0.0)   # uniform B: consistent
    assert _coarse_fine_face_gap(spec, blocks) == 0.0
    o = oracle.OracleMesh(spec, phys, n_threads=8)
    o.upload_blocks(blocks)
    o.advance(6)
    out = o.download_blocks()
    moved = max(float(np.abs(_faces(spec, out[b][1])[0] - _faces(spec, blocks[b][1])[0]).max())
                for b in range(spec.n_blocks))
    assert moved > 1e-4                      # the field did evolve
    assert _coarse_fine_face_gap(spec, out) < 1e-13


def _wave_error(spec, t_end, n_threads=8):
    phys = Physics(mhd=False, gamma=G, cfl=0.3, t_lim=t_end, top_up=True)
    o = oracle.OracleMesh(spec, phys, n_threads=n_threads)
    o.upload_blocks(problems.wave_blocks(spec))
    o.advance(100000)
    assert o.time()[0] == t_end
    out = o.download_blocks()
    err, vol = 0.0, 0.0
    for b, (ua, _) in enumerate(out):
        rho = np.asarray(ua).reshape(-1, 5)[:, 0].reshape(spec.n_cell_block[::-1])
        v = _vol(spec, b)
        err += np.abs(rho - problems.wave_exact(spec, b, t_end)).sum() * v
        vol += rho.size * v
    return err / vol


def test_wave_through_refined_region():
    t_end = 0.25
    coarse = _wave_error(RefinedMeshSpec((32, 32, 32), (8, 8, 8)), t_end)
    fine = _wave_error(RefinedMeshSpec((64, 64, 64), (8, 8, 8)), t_end)
    smr = _wave_error(RefinedMeshSpec((32, 32, 32), (8, 8, 8), refine=REFINE), t_end)
    assert fine < 0.5 * coarse
    assert fine < smr < coarse, (fine, smr, coarse)


def test_uniform_state_on_refined_mesh():
    spec = RefinedMeshSpec((32, 32, 32), (8, 8, 8), refine=REFINE)
    phys = Physics(mhd=True, gamma=G, cfl=0.3)
    blocks = problems.wave_blocks(spec, amp=0.0, vel=(0.3, -0.2, 0.1), mhd=True,
                                  b0=(0.5, 0.25, -0.4))
    o = oracle.OracleMesh(spec, phys, n_threads=8)
    o.upload_blocks(blocks)
    o.advance(4)
    out = o.download_blocks()
    for b in range(spec.n_blocks):
        np.testing.assert_allclose(out[b][0], blocks[b][0], rtol=1e-13, atol=1e-14)
        for x, y in zip(_faces(spec, out[b][1]), _faces(spec, blocks[b][1])):
            np.testing.assert_allclose(x, y, rtol=1e-13, atol=1e-14)


def _unbalanced_spec():
    """Root 2^3 blocks of 8^3; root block (0,0,0) refined to level 1 and its
    child (1,1,1) again to level 2: those level-2 blocks touch level-0
    blocks across the root's faces (tree_t::insert would have refined the
    neighbours; a hand-made list need not)."""
    spec = RefinedMeshSpec((16, 16, 16), (8, 8, 8))
    leaves = [(0, x, y, z) for x in range(2) for y in range(2) for z in range(2) if (x, y, z) != (0, 0, 0)]
    leaves += [(1, x, y, z) for x in range(2) for y in range(2) for z in range(2) if (x, y, z) != (1, 1, 1)]
    leaves += [(2, 2 + x, 2 + y, 2 + z) for x in range(2) for y in range(2) for z in range(2)]
    spec.leaves = sorted(leaves)
    spec.n_levels = 3
    return spec


def test_unbalanced_mesh_is_rejected():
    with pytest.raises(RuntimeError, match="balanced|within one level"):
        oracle.OracleMesh(_unbalanced_spec(), Physics(mhd=True, gamma=G), n_threads=1)

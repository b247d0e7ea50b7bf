"""Pins the oracle's MHD physical boundaries (mhd/boundary/phys.h,
phys_functors.h) with what the mathematics fixes:

* A reflecting wall is a mirror: a magnetised 1-D flow between walls equals
  the left half of the periodic box twice as long holding the state and its
  mirror image, to 1e-12 (a wrong sign or mirror index is off by 1e-2); a 3-D
  reflecting box conserves mass to round-off.
* Outflow keeps an outflowing normal velocity and zeroes an inflowing one.
* Free (zero-gradient) sides keep a uniform moving magnetised state
  exactly uniform, including the high-side boundary-plane edge EMF
  (rho = 1, so its v = w_mom / rho reading is the velocity).
* One block or eight blocks with the same physical sides: bit-identical.

The reference's reflect functor mirrors the tangential face field at -i
(phys_functors.h:118-132) where the cell-centred values mirror at -i-1;
the oracle follows that index as written.  A mirror twin with a
tangential field separates the two: the geometric -i-1 would equal the
twin, -i does not, and the test asserts the difference (so -i-1 is
excluded).  Nothing fixes -i itself, so that detail stays parity
unpinned.
"""
import math

import numpy as np
import pytest

import oracle
from kratos_amd import problems
from kratos_amd.mesh import MeshSpec, Physics, RefinedMeshSpec

G = 5.0 / 3.0


def _sym_blocks(spec):
    """Per-block state mirror-symmetric about x = 0.5 (box [0,1]^3)."""
    n = spec.n_cell_block
    out = []
    tp = 2 * math.pi
    for b in range(spec.n_blocks):
        cc, fc, dx = problems._block_axes(spec, b)
        X, Y, Z = problems._grid(cc[0], cc[1], cc[2])
        sx = X - 0.5
        rho = 1.0 + 0.2 * np.cos(tp * sx) * (1 + 0.1 * np.sin(tp * Y))
        pre = 0.6 + 0.1 * np.cos(2 * tp * sx) * np.cos(tp * Z)
        vel = [0.3 * np.sin(tp * sx) * (1 + 0.2 * np.cos(tp * Y)),
               0.2 * np.cos(tp * sx) * np.sin(tp * Z),
               0.1 * np.cos(2 * tp * sx) * np.sin(tp * Y)]
        a = 0.05

        def Ay(x, y, z):
            return a * np.cos(tp * (x - 0.5)) * np.cos(tp * z)

        def Az(x, y, z):
            return a * np.cos(2 * tp * (x - 0.5)) * np.cos(tp * y)

        ay = Ay(*problems._grid(fc[0], cc[1], fc[2]))
        az = Az(*problems._grid(fc[0], fc[1], cc[2]))
        bx = (az[:, 1:, :] - az[:, :-1, :]) / dx[1] - (ay[1:, :, :] - ay[:-1, :, :]) / dx[2] + 0.4
        by = -(az[:, :, 1:] - az[:, :, :-1]) / dx[0]
        bz = (ay[:, :, 1:] - ay[:, :, :-1]) / dx[0]
        out.append(problems._pack_block(n, rho, pre, vel, (bx, by, bz), G))
    return out


def _cells(spec, blocks):
    """Global (var, z, y, x) cell array of a single-level spec (x-slowest order)."""
    n = spec.n_cell_block
    N = spec.n_cell_global
    u = np.zeros((5, N[2], N[1], N[0]))
    for b, (ua, _) in enumerate(blocks):
        r = spec.reg(b)
        ub = np.moveaxis(np.asarray(ua).reshape(n[2], n[1], n[0], 8), -1, 0)[:5]
        u[:, r[2] * n[2]:(r[2] + 1) * n[2], r[1] * n[1]:(r[1] + 1) * n[1],
          r[0] * n[0]:(r[0] + 1) * n[0]] = ub
    return u


def _slab_blocks(spec, mirror_from=None, b_t=0.0):
    """1-D state along x (uniform in y, z), B = (0.6, B_y(x), 0), v = (v_x,
    0, 0).  With b_t = 0 the face field keeps no tangential component, so
    the flow is a magnetised 1-D gas; b_t scales a tangential B_y(x), which
    the mirror image negates (B is an axial vector)."""
    n = spec.n_cell_block
    tp = 2 * math.pi
    out = []
    for b in range(spec.n_blocks):
        cc, _, _ = problems._block_axes(spec, b)
        x = cc[0] if mirror_from is None else np.where(cc[0] < 1.0, cc[0], 2.0 - cc[0])
        sgn = 1.0 if mirror_from is None else np.where(cc[0] < 1.0, 1.0, -1.0)
        rho1 = 1.0 + 0.3 * np.exp(-((x - 0.3) / 0.1) ** 2)
        pre1 = 0.6 + 0.4 * np.exp(-((x - 0.6) / 0.08) ** 2)
        vx1 = sgn * 0.4 * np.sin(tp * x) * np.exp(-((x - 0.45) / 0.2) ** 2)
        sh = (n[2], n[1], n[0])
        rho = np.broadcast_to(rho1[None, None, :], sh).copy()
        pre = np.broadcast_to(pre1[None, None, :], sh).copy()
        vel = [np.broadcast_to(vx1[None, None, :], sh).copy(), np.zeros(sh), np.zeros(sh)]
        by1 = sgn * b_t * (0.5 + np.cos(tp * x))
        bfs = (np.full((n[2], n[1], n[0] + 1), 0.6),
               np.broadcast_to(by1[None, None, :], (n[2], n[1] + 1, n[0])).copy(),
               np.zeros((n[2] + 1, n[1], n[0])))
        out.append(problems._pack_block(n, rho, pre, vel, bfs, G))
    return out


def test_reflecting_wall_equals_mirrored_periodic_box():
    """A reflecting wall is a mirror: [0, 1] with walls equals the left half
    of the periodic [0, 2] holding the state and its mirror image (v_x odd)."""
    wall = RefinedMeshSpec((32, 4, 4), (32, 4, 4), x_max=(1.0, 1.0, 1.0),
                           bnd=("ref", "ref", "per", "per", "per", "per"))
    twin = RefinedMeshSpec((64, 4, 4), (64, 4, 4), x_max=(2.0, 1.0, 1.0))
    phys = Physics(mhd=True, gamma=G, cfl=0.3)
    ow = oracle.OracleMesh(wall, phys, n_threads=1)
    ow.upload_blocks(_slab_blocks(wall))
    ot = oracle.OracleMesh(twin, phys, n_threads=1)
    ot.upload_blocks(_slab_blocks(twin, mirror_from=1.0))
    assert ow.advance(40) == 40 and ot.advance(40) == 40
    assert ow.time()[0] == pytest.approx(ot.time()[0], rel=1e-14)
    uw = np.asarray(ow.download_blocks()[0][0]).reshape(4, 4, 32, 8)
    ut = np.asarray(ot.download_blocks()[0][0]).reshape(4, 4, 64, 8)
    # equal up to the round-off by which the periodic twin's right half
    # drifts from an exact mirror of its left half
    np.testing.assert_allclose(uw[..., :5], ut[:, :, :32, :5], rtol=0, atol=1e-12)
    assert np.abs(uw[..., 2]).max() > 0.05


def test_reflect_tangential_face_mirror_is_not_geometric():
    """With a tangential B_y(x) and v_x != 0 at the wall, a wall built on the
    geometric mirror of the tangential face field (ghost cell -1 <- cell 0,
    index -i-1) would equal the mirrored periodic twin as the test above
    does.  The reference's functor reads index -i (ghost -1 <- cell 1,
    phys_functors.h:118-132), and so does the oracle: the two runs then
    differ near the wall by far more than round-off.  This excludes -i-1;
    no invariant fixes the -i choice itself."""
    wall = RefinedMeshSpec((32, 4, 4), (32, 4, 4), x_max=(1.0, 1.0, 1.0),
                           bnd=("ref", "ref", "per", "per", "per", "per"))
    twin = RefinedMeshSpec((64, 4, 4), (64, 4, 4), x_max=(2.0, 1.0, 1.0))
    phys = Physics(mhd=True, gamma=G, cfl=0.3)
    ow = oracle.OracleMesh(wall, phys, n_threads=1)
    ow.upload_blocks(_slab_blocks(wall, b_t=0.3))
    ot = oracle.OracleMesh(twin, phys, n_threads=1)
    ot.upload_blocks(_slab_blocks(twin, mirror_from=1.0, b_t=0.3))
    assert ow.advance(10) == 10 and ot.advance(10) == 10
    uw = np.asarray(ow.download_blocks()[0][0]).reshape(4, 4, 32, 8)
    ut = np.asarray(ot.download_blocks()[0][0]).reshape(4, 4, 64, 8)
    d = np.abs(uw[..., :5] - ut[:, :, :32, :5]).max(axis=(0, 1, 3))
    assert d.max() > 1e-6, d
    # the difference enters at the walls (with -i-1 the runs agree to 5e-16)
    assert d[[0, 1, 30, 31]].min() > 10 * d[4:28].max(), d


def test_reflecting_box_conserves_mass():
    spec = RefinedMeshSpec((16, 16, 16), (8, 8, 8), bnd=("ref",) * 6)
    phys = Physics(mhd=True, gamma=G, cfl=0.3)
    blocks = problems.smooth_blocks(spec, seed=6)
    o = oracle.OracleMesh(spec, phys, n_threads=4)
    o.upload_blocks(blocks)
    m0 = sum(np.asarray(ua).reshape(-1, 8)[:, 0].sum() for ua, _ in blocks)
    o.advance(10)
    m1 = sum(np.asarray(ua).reshape(-1, 8)[:, 0].sum() for ua, _ in o.download_blocks())
    assert abs(m1 - m0) <= 1e-13 * m0


def test_free_sides_keep_a_uniform_moving_state():
    spec = RefinedMeshSpec((16, 16, 16), (8, 8, 8), bnd=("fre",) * 6)
    phys = Physics(mhd=True, gamma=G, cfl=0.3)
    blocks = problems.wave_blocks(spec, amp=0.0, vel=(0.3, -0.2, 0.1), mhd=True,
                                  b0=(0.5, 0.25, -0.4))
    o = oracle.OracleMesh(spec, phys, n_threads=4)
    o.upload_blocks(blocks)
    o.advance(5)
    for (ua, ba), (ub, bb) in zip(o.download_blocks(), blocks):
        np.testing.assert_allclose(ua, ub, rtol=1e-13, atol=1e-14)


def test_outflow_passes_outflow_and_blocks_inflow():
    """Uniform flow along +x with a free left side and an outflow right side
    stays uniform (an outflowing normal velocity is kept); reversed, the
    outflow side zeroes the inflowing normal velocity of its ghosts and the
    state changes."""
    for vx, keeps in ((0.3, True), (-0.3, False)):
        spec = RefinedMeshSpec((16, 8, 8), (8, 8, 8), bnd=("fre", "out", "per", "per", "per", "per"))
        phys = Physics(mhd=True, gamma=G, cfl=0.3)
        blocks = problems.wave_blocks(spec, amp=0.0, vel=(vx, 0.1, -0.2), mhd=True,
                                      b0=(0.5, 0.25, -0.4))
        o = oracle.OracleMesh(spec, phys, n_threads=2)
        o.upload_blocks(blocks)
        o.advance(4)
        diff = max(float(np.abs(ua - ub).max()) for (ua, _), (ub, _) in zip(o.download_blocks(), blocks))
        assert (diff < 1e-13) == keeps, (vx, diff)


def test_block_decomposition_with_physical_sides():
    bnd = ("out", "fre", "ref", "ref", "fre", "out")
    one = MeshSpec((16, 16, 16), (16, 16, 16), bnd=bnd)
    many = MeshSpec((16, 16, 16), (8, 8, 8), bnd=bnd)
    phys = Physics(mhd=True, gamma=G, cfl=0.3)
    u, b, _ = problems.turbulence((16, 16, 16), seed=4, beta=2.0)
    res = []
    for spec in (one, many):
        o = oracle.OracleMesh(spec, phys, n_threads=4)
        o.upload(u, b)
        assert o.advance(6) == 6
        res.append(o.download())
    np.testing.assert_array_equal(res[0][0], res[1][0])
    for x, y in zip(res[0][1], res[1][1]):
        np.testing.assert_array_equal(x, y)


def test_fixed_sides_hold_the_ambient_state():
    """A "fixd" (Dirichlet) side holds the ambient state in its ghosts: a
    uniform state equal to that ambient medium stays exactly uniform (hydro
    and MHD), and a different interior state is driven by the ghosts (the
    ghosts are the fixed state, not copies of the interior)."""
    for mhd in (False, True):
        spec = RefinedMeshSpec((16, 8, 8), (8, 8, 8), bnd=("fixd", "fixd", "per", "per", "fixd", "fre"))
        amb_v, amb_b = (0.2, -0.1, 0.05), (0.3, 0.2, -0.1)
        phys = Physics(mhd=mhd, gamma=G, cfl=0.3, fix_w=(1.0, 0.6, 0.2, -0.1, 0.05),
                       fix_b=amb_b if mhd else (0.0, 0.0, 0.0))
        blocks = problems.wave_blocks(spec, amp=0.0, vel=amb_v, pre=0.6, mhd=mhd, b0=amb_b)
        o = oracle.OracleMesh(spec, phys, n_threads=2)
        o.upload_blocks(blocks)
        o.advance(4)
        for (ua, _), (ub, _) in zip(o.download_blocks(), blocks):
            np.testing.assert_allclose(ua, ub, rtol=1e-13, atol=1e-14)
        other = problems.wave_blocks(spec, amp=0.0, vel=amb_v, pre=0.9, mhd=mhd, b0=amb_b)
        o = oracle.OracleMesh(spec, phys, n_threads=2)
        o.upload_blocks(other)
        o.advance(4)
        diff = max(float(np.abs(ua - ub).max()) for (ua, _), (ub, _) in zip(o.download_blocks(), other))
        assert diff > 1e-3

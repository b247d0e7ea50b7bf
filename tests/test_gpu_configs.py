"""GPU parity on the benchmark configurations' own shapes and on the
corner-upwind EMF case, through the C-ABI against the CPU oracle.

* Field-loop advection (Gardiner & Stone 2005), extruded in 3-D, in each
  of the three planes: the case whose result depends on the corner-upwind
  EMF selection (the oracle tests pin it there).  FP64 bit-exact on the
  staged and fused kernels; mixed form within 1e-5.
* The ensemble (64^3 single-block periodic boxes, the turbulence recipe,
  seeds 0.., beta ~ U[0.1, 10] from the same SplitMix64 stream, per-box dt
  to t = 0.05) in the mixed form: each box's density L1 against its own
  FP64 oracle run <= 1e-5.
* The refined blast (three static levels, 16^3 blocks, periodic box) in
  both forms: FP64 bit-exact, mixed density L1 <= 1e-5.
* The 256^3 measurement's block shape at 128^3 (64 blocks of 32^3, the
  autotuned fused kernel), beta = 1 and 0.01, mixed form, 20 cycles:
  density, energy and B L1 <= 1e-5.
Every MHD case checks the div B bar max |div B| dx / max |B| <= 1e-12.
"""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import kratos_amd
import oracle
from kratos_amd import problems
from kratos_amd.mesh import MeshSpec, Physics, RefinedMeshSpec, divb_rel, divb_rel_blocks

from test_gpu_smr import _assert_equal as assert_blocks_equal
from test_oracle_emf_upwind import loop_case

pytestmark = pytest.mark.gpu
G = 5.0 / 3.0


def rel_l1(a, b):
    return float(np.abs(a - b).sum() / np.abs(b).sum())


def cat(bs):
    return np.concatenate([x.ravel() for x in bs])


_LOOP = {}


def _loop_oracle(orient):
    if orient not in _LOOP:
        u, b, kw = loop_case((64, 32, 16), orient, vz=1.0)
        N = (u.shape[3], u.shape[2], u.shape[1])
        spec = MeshSpec(N, (16, 16, 16), **kw)
        phys = Physics(mhd=True, gamma=G, cfl=0.3, t_lim=0.25, top_up=True)
        o = oracle.OracleMesh(spec, phys, n_threads=16)
        o.upload(u, b)
        o.advance(100000)
        _LOOP[orient] = (spec, phys, u, b, o.download(), o.time())
    return _LOOP[orient]


@pytest.mark.parametrize("orient", [0, 1, 2])
@pytest.mark.parametrize("path", ["staged", "fused"])
def test_field_loop_fp64(gpu_lib, orient, path):
    spec, phys, u, b, (ou, ob), ot = _loop_oracle(orient)
    print(f"\n[case] field loop orient={orient} N={spec.n_cell_global} path={path}", flush=True)
    ctx = kratos_amd.Context(spec, phys, "fp64", path=path)
    ctx.upload(u, b)
    ctx.advance(100000)
    gu, gb = ctx.download()
    gt = ctx.time()
    ctx.close()
    assert gt[0] == ot[0] == 0.25
    np.testing.assert_array_equal(gu, ou)
    for x, y in zip(gb, ob):
        np.testing.assert_array_equal(x, y)
    dx = [(spec.x_max[a] - spec.x_min[a]) / spec.n_cell_global[a] for a in range(3)]
    assert divb_rel(gb, dx) <= 1e-12


@pytest.mark.parametrize("orient", [0, 1, 2])
def test_field_loop_mixed(gpu_lib, orient):
    spec, phys, u, b, (ou, ob), ot = _loop_oracle(orient)
    print(f"\n[case] field loop mixed orient={orient}", flush=True)
    ctx = kratos_amd.Context(spec, phys, "mixed")
    ctx.upload(u, b)
    ctx.advance(100000)
    gu, gb = ctx.download()
    ctx.close()
    assert rel_l1(gu[0], ou[0]) <= 1e-5
    assert rel_l1(cat(gb), cat(ob)) <= 1e-5
    e_g, e_o = problems.magnetic_energy(gb), problems.magnetic_energy(ob)
    assert abs(e_g - e_o) <= 1e-5 * e_o
    dx = [(spec.x_max[a] - spec.x_min[a]) / spec.n_cell_global[a] for a in range(3)]
    assert divb_rel(gb, dx) <= 1e-12


def test_ensemble_mixed_per_box(gpu_lib):
    """Eight 64^3 single-block boxes of the ensemble recipe, mixed form,
    per-box dt to t = 0.05, each box against its own FP64 oracle run."""
    N = (64, 64, 64)
    n_box = 8
    spec = MeshSpec(N, N)
    phys = Physics(mhd=True, gamma=G, cfl=0.3, t_lim=0.05, top_up=True)
    ics = [problems.turbulence(N, seed=s, beta_range=(0.1, 10.0)) for s in range(n_box)]
    betas = [info["beta"] for _, _, info in ics]
    assert min(betas) < 2.0 and max(betas) > 5.0, betas
    print(f"\n[case] ensemble mixed {n_box} x {N}", flush=True)
    ctx = kratos_amd.Context(spec, phys, "mixed", members=n_box)
    for m, (u, b, _) in enumerate(ics):
        ctx.upload(u, b, member=m)
    ctx.advance(100000)
    got = [ctx.download(member=m) + (ctx.time(member=m),) for m in range(n_box)]
    ctx.close()

    def ref(m):
        u, b, _ = ics[m]
        o = oracle.OracleMesh(spec, phys, n_threads=1)
        o.upload(u, b)
        o.advance(100000)
        return o.download() + (o.time(),)

    with ThreadPoolExecutor(max_workers=n_box) as ex:
        refs = list(ex.map(ref, range(n_box)))
    cycles = []
    for m in range(n_box):
        gu, gb, gt = got[m]
        ou, ob, ot = refs[m]
        assert gt[0] == ot[0] == 0.05
        cycles.append(gt[2])
        assert rel_l1(gu[0], ou[0]) <= 1e-5, (m, betas[m])
        assert rel_l1(gu[1], ou[1]) <= 1e-5, (m, betas[m])
        assert rel_l1(cat(gb), cat(ob)) <= 1e-5, (m, betas[m])
        assert divb_rel(gb, (1.0 / 64,) * 3) <= 1e-12
    assert len(set(cycles)) > 1, "boxes of different beta need different cycle counts"


def test_ensemble_mixed_equals_boxes_alone(gpu_lib):
    """The mixed ensemble's stage kernel walks a device work list of the
    boxes still running, splitting their columns into z-ranges as fewer
    remain; a box run alone takes the grid launch.  Boxes are independent,
    so every box of the ensemble must equal its own run bit for bit."""
    N = (32, 32, 32)
    n_box = 6
    spec = MeshSpec(N, N)
    phys = Physics(mhd=True, gamma=G, cfl=0.3, t_lim=0.02, top_up=True)
    ics = [problems.turbulence(N, seed=s, beta_range=(0.1, 10.0)) for s in range(n_box)]
    print(f"\n[case] ensemble mixed {n_box} x {N} vs boxes alone", flush=True)
    ctx = kratos_amd.Context(spec, phys, "mixed", members=n_box)
    for m, (u, b, _) in enumerate(ics):
        ctx.upload(u, b, member=m)
    ctx.advance(100000)
    got = [ctx.download(member=m) + (ctx.time(member=m),) for m in range(n_box)]
    st = ctx.stats()
    ctx.close()
    assert st["sweep_kernel"] == 1, st
    cycles = []
    for m, (u, b, _) in enumerate(ics):
        one = kratos_amd.Context(spec, phys, "mixed", tile=tuple(st["sweep_tile"][0:2]) + (1,))
        one.upload(u, b)
        done = one.advance(100000)
        ou, ob = one.download()
        ot = one.time()
        one.close()
        gu, gb, gt = got[m]
        assert gt[0] == ot[0] == 0.02 and gt[2] == done, (m, gt, ot, done)
        cycles.append(gt[2])
        np.testing.assert_array_equal(gu, ou, err_msg=f"box {m}")
        for x, y in zip(gb, ob):
            np.testing.assert_array_equal(x, y, err_msg=f"box {m} faces")
    assert len(set(cycles)) > 1, "boxes of different beta need different cycle counts"


_BLAST = {}


def _blast_spec():
    return RefinedMeshSpec((64, 64, 64), (16, 16, 16),
                           refine=[(1, (-0.25,) * 3, (0.25,) * 3), (2, (-0.125,) * 3, (0.125,) * 3)],
                           x_min=(-0.5,) * 3, x_max=(0.5,) * 3)


def _blast_oracle(cycles):
    if cycles not in _BLAST:
        spec = _blast_spec()
        phys = Physics(mhd=True, gamma=G, cfl=0.3)
        blocks = problems.blast_blocks(spec)
        o = oracle.OracleMesh(spec, phys, n_threads=16)
        o.upload_blocks(blocks)
        assert o.advance(cycles) == cycles
        _BLAST[cycles] = (spec, phys, blocks, o.download_blocks(), o.time())
    return _BLAST[cycles]


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_refined_blast(gpu_lib, precision):
    spec, phys, blocks, o, ot = _blast_oracle(10)
    assert spec.n_levels == 3
    print(f"\n[case] refined blast {spec.n_blocks} blocks {precision}", flush=True)
    ctx = kratos_amd.Context(spec, phys, precision)
    ctx.upload_blocks(blocks)
    assert ctx.advance(10) == 10
    g = ctx.download_blocks()
    gt = ctx.time()
    ctx.close()
    assert divb_rel_blocks(spec, g) <= 1e-12
    if precision == "fp64":
        # conserved variables and the face field bit for bit (u's Bcc slots
        # are the face average the next c2p would write: the library returns
        # it from the final face field, the oracle keeps the last c2p's)
        assert gt[0] == ot[0]
        assert_blocks_equal(spec, g, o, True)
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


@pytest.mark.parametrize("beta", [1.0, 0.01])
def test_turbulence_128_in_32_blocks_mixed(gpu_lib, beta):
    N, blk = (128, 128, 128), (32, 32, 32)
    spec = MeshSpec(N, blk)
    phys = Physics(mhd=True, gamma=G, cfl=0.3)
    u, b, _ = problems.turbulence(N, seed=0, beta=beta)
    print(f"\n[case] 128^3 in 32^3 blocks mixed beta={beta}", flush=True)
    ctx = kratos_amd.Context(spec, phys, "mixed")
    ctx.upload(u, b)
    assert ctx.advance(20) == 20
    gu, gb = ctx.download()
    ctx.close()
    o = oracle.OracleMesh(spec, phys, n_threads=16)
    o.upload(u, b)
    assert o.advance(20) == 20
    ou, ob = o.download()
    for v in (0, 1):
        assert rel_l1(gu[v], ou[v]) <= 1e-5, v
    assert rel_l1(cat(gb), cat(ob)) <= 1e-5
    assert divb_rel(gb, (1.0 / 128,) * 3) <= 1e-12

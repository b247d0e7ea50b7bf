"""GPU: kr_divb_max_noalloc, the allocation-free variant of kr_divb_max.

The face fields of a context's blocks (downloaded in kr_upload's AoS
layout after a few MHD cycles, so that div B is a non-zero round-off
value) are put in device memory by the test (hipMalloc / hipMemcpy of the
HIP runtime the library links, called through ctypes: plumbing) and
handed to kr_divb_max_noalloc; its result must equal kr_divb_max on the
same state bit for bit.  Two calls in a row, the second on a smaller mesh
(fewer blocks, smaller blocks), each against its own context; and a
request beyond the scratch limit is refused with a message and no launch.
"""
import ctypes as C

import numpy as np
import pytest

import kratos_amd
from kratos_amd import _capi, problems
from kratos_amd.mesh import MeshSpec, Physics

pytestmark = pytest.mark.gpu
G = 5.0 / 3.0


def _case(N, blk, seed, precision):
    spec = MeshSpec(N, blk)
    phys = Physics(mhd=True, gamma=G, cfl=0.3)
    u, b, _ = problems.turbulence(N, seed=seed)
    ctx = kratos_amd.Context(spec, phys, precision)
    ctx.upload(u, b)
    ctx.advance(3)
    want = ctx.divb_max()
    blocks = ctx.download_blocks()
    dx = []
    for k in range(spec.n_blocks):
        _, d = spec.geometry(k)
        dx.extend(d)
    ctx.close()
    bf = np.concatenate([np.asarray(x[1], dtype=np.float64) for x in blocks])
    return spec, bf, np.asarray(dx, dtype=np.float64), want


class _Hip:
    """The few HIP runtime calls the test needs, from the runtime
    libkratos_amd.so links (one runtime in the process)."""

    def __init__(self):
        self.h = C.CDLL("libamdhip64.so.7")
        self.h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.h.hipFree.argtypes = [C.c_void_p]
        self.h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.h.hipDeviceSynchronize.argtypes = []

    def ok(self, rc, what):
        assert rc == 0, f"{what} returned {rc}"

    def put(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        p = C.c_void_p()
        self.ok(self.h.hipMalloc(C.byref(p), a.nbytes), "hipMalloc")
        self.ok(self.h.hipMemcpy(p, a.ctypes.data, a.nbytes, 1), "hipMemcpy H2D")
        return p

    def get(self, p, n):
        a = np.zeros(n, dtype=np.float64)
        self.ok(self.h.hipMemcpy(a.ctypes.data, p, a.nbytes, 2), "hipMemcpy D2H")
        return a


def _noalloc(lib, bf, dx, spec):
    hip = _Hip()
    bf_d, dx_d = hip.put(bf), hip.put(dx)
    out = hip.put(np.array([-1.0]))
    n = (C.c_int * 3)(*spec.n_cell_block)
    try:
        rc = lib.kr_divb_max_noalloc(bf_d, spec.n_blocks, n, dx_d, out, None)
        _capi.check(rc, None, "kr_divb_max_noalloc")
        hip.ok(hip.h.hipDeviceSynchronize(), "hipDeviceSynchronize")
        return float(hip.get(out, 1)[0])
    finally:
        for p in (bf_d, dx_d, out):
            hip.h.hipFree(p)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_noalloc_divb_equals_kr_divb_max(gpu_lib, precision):
    big = _case((64, 64, 64), (16, 16, 16), 1, precision)       # 64 blocks
    small = _case((32, 32, 16), (16, 8, 8), 2, precision)       # 16 smaller blocks
    for spec, bf, dx, want in (big, small):
        print(f"\n[case] noalloc divb {spec.n_blocks} blocks of {spec.n_cell_block} "
              f"{precision}", flush=True)
        got = _noalloc(gpu_lib, bf, dx, spec)
        assert want > 0
        assert got == want, (got, want)


def test_noalloc_rejects_beyond_scratch(gpu_lib):
    n = (C.c_int * 3)(4, 4, 4)
    # refused on the host before any launch: the pointers are never read
    rc = gpu_lib.kr_divb_max_noalloc(C.c_void_p(16), 65537, n, C.c_void_p(16), C.c_void_p(16),
                                     None)
    assert rc == -1
    assert b"scratch limit" in gpu_lib.kr_last_error(None)

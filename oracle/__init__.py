"""TEST INFRASTRUCTURE ONLY: the CPU oracle.

ctypes wrapper of ``oracle/libkratos_oracle.so``, the C++17 FP64
restatement of the reference's Godunov hydro/MHD stage (PRECISION=2
semantics).  Only tests/, ``__graft_entry__.smoke()`` and the cpu_baseline
leg of ``bench.py`` may import this package, and only as the checker; the
kratos_amd product never does.

Parity status of each oracle function (how it is pinned) is listed in the
header comment of each test module under tests/.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libkratos_oracle.so")

_BND = {"per": 0, "out": 1, "fre": 2, "ref": 3, "fixd": 4}


class KroConfig(C.Structure):
    _fields_ = [
        ("mhd", C.c_int),
        ("gamma", C.c_double), ("cfl", C.c_double),
        ("rho_floor", C.c_double), ("tmp_floor", C.c_double),
        ("n_cell_global", C.c_int * 3),
        ("n_cell_block", C.c_int * 3),
        ("x_min", C.c_double * 3), ("x_max", C.c_double * 3),
        ("bnd", C.c_int * 6),
        ("dt_init", C.c_double), ("dt_expand", C.c_double), ("t_lim", C.c_double),
        ("top_up", C.c_int),
        ("n_threads", C.c_int),
        ("fix_w", C.c_double * 5),
        ("fix_b", C.c_double * 3),
    ]


_lib = None


def _bf_size(n):
    """Doubles in a block's face-field array: n + (n > 1) faces plus two
    ghost layers each side, per axis, times 3 components
    (mhd/types/fld_data.cpp:18-26)."""
    s = 3
    for v in n:
        s *= v + (1 if v > 1 else 0) + 4
    return s


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    lib = C.CDLL(LIB_PATH)
    D = C.POINTER(C.c_double)
    lib.kro_create.restype = C.c_void_p
    lib.kro_create.argtypes = [C.POINTER(KroConfig)]
    lib.kro_create_blocks.restype = C.c_void_p
    lib.kro_create_blocks.argtypes = [C.POINTER(KroConfig), C.c_int, C.POINTER(C.c_int),
                                      C.POINTER(C.c_int)]
    lib.kro_destroy.argtypes = [C.c_void_p]
    lib.kro_last_error.restype = C.c_char_p
    lib.kro_n_blocks.argtypes = [C.c_void_p]
    lib.kro_block_geom.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                   D, D]
    lib.kro_set_block.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
    lib.kro_get_block.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
    lib.kro_advance.argtypes = [C.c_void_p, C.c_int]
    for n in ("kro_time", "kro_dt", "kro_dt_next"):
        getattr(lib, n).restype = C.c_double
        getattr(lib, n).argtypes = [C.c_void_p]
    lib.kro_plm.argtypes = [D, D, D]
    lib.kro_hllc.argtypes = [D, D, C.c_double, D]
    lib.kro_hlld.argtypes = [D, D, C.c_double, D]
    lib.kro_face_flux.argtypes = [C.c_int, C.c_int, D, D, C.c_double, D]
    lib.kro_c2p_hydro.restype = C.c_double
    lib.kro_c2p_hydro.argtypes = [D, D, C.c_double, C.c_double, C.c_double]
    _lib = lib
    return lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class OracleMesh:
    """CPU oracle over the same uniform block mesh as kratos_amd.mesh.Context."""

    def __init__(self, spec, phys, n_threads=0):
        lib = load()
        self.lib, self.spec, self.phys, self.mhd = lib, spec, phys, phys.mhd
        c = KroConfig()
        c.mhd = int(phys.mhd)
        c.gamma, c.cfl = phys.gamma, phys.cfl
        c.rho_floor, c.tmp_floor = phys.rho_floor, phys.tmp_floor
        for a in range(3):
            c.n_cell_global[a] = spec.n_cell_global[a]
            c.n_cell_block[a] = spec.n_cell_block[a]
            c.x_min[a] = spec.x_min[a]
            c.x_max[a] = spec.x_max[a]
        for s in range(6):
            c.bnd[s] = _BND[spec.bnd[s]]
        c.dt_init, c.dt_expand, c.t_lim = phys.dt_init, phys.dt_expand, phys.t_lim
        c.top_up = int(phys.top_up)
        c.n_threads = int(n_threads)
        for k in range(5):
            c.fix_w[k] = float(getattr(phys, "fix_w", (1.0, 1.0, 0.0, 0.0, 0.0))[k])
        for k in range(3):
            c.fix_b[k] = float(getattr(phys, "fix_b", (0.0, 0.0, 0.0))[k])
        if phys.t_0 != 0.0:
            raise ValueError("oracle starts at t = 0")
        if hasattr(spec, "leaves"):             # statically refined mesh
            nb = spec.n_blocks
            lv = (C.c_int * nb)(*[spec.level(b) for b in range(nb)])
            rg = (C.c_int * (3 * nb))(*[v for b in range(nb) for v in spec.reg(b)])
            h = lib.kro_create_blocks(C.byref(c), nb, lv, rg)
        else:
            h = lib.kro_create(C.byref(c))
        if not h:
            raise RuntimeError("kro_create: " + lib.kro_last_error().decode())
        self.h = h

    def close(self):
        if self.h:
            self.lib.kro_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def block_geom(self, b):
        reg, n = (C.c_int * 3)(), (C.c_int * 3)()
        xf0, dx0 = np.zeros(3), np.zeros(3)
        self.lib.kro_block_geom(self.h, b, reg, n, _dp(xf0), _dp(dx0))
        return tuple(reg), tuple(n), xf0, dx0

    def upload(self, u, bfs=None):
        for b, (ua, ba) in enumerate(self.spec.split(u, bfs if self.mhd else None)):
            ua = np.ascontiguousarray(ua)
            bp = None
            if ba is not None:
                ba = np.ascontiguousarray(ba)
                bp = ba.ctypes.data
            if self.lib.kro_set_block(self.h, b, ua.ctypes.data, bp) != 0:
                raise RuntimeError("kro_set_block failed")

    def download(self):
        n = self.spec.n_cell_block
        nvar = 8 if self.mhd else 5
        blocks = []
        for b in range(self.spec.n_blocks):
            ua = np.zeros(n[0] * n[1] * n[2] * nvar)
            ba = np.zeros(_bf_size(n)) if self.mhd else None
            self.lib.kro_get_block(self.h, b, ua.ctypes.data,
                                   ba.ctypes.data if ba is not None else None)
            blocks.append((ua, ba))
        return self.spec.gather(blocks, self.mhd)

    def upload_blocks(self, blocks):
        for b, (ua, ba) in enumerate(blocks):
            ua = np.ascontiguousarray(ua, dtype=np.float64)
            bp = None
            if self.mhd:
                ba = np.ascontiguousarray(ba, dtype=np.float64)
                bp = ba.ctypes.data
            if self.lib.kro_set_block(self.h, b, ua.ctypes.data, bp) != 0:
                raise RuntimeError("kro_set_block failed")

    def download_blocks(self):
        n = self.spec.n_cell_block
        nvar = 8 if self.mhd else 5
        out = []
        for b in range(self.spec.n_blocks):
            ua = np.zeros(n[0] * n[1] * n[2] * nvar)
            ba = np.zeros(_bf_size(n)) if self.mhd else None
            self.lib.kro_get_block(self.h, b, ua.ctypes.data,
                                   ba.ctypes.data if ba is not None else None)
            out.append((ua, ba))
        return out

    def advance(self, n_cycles):
        r = self.lib.kro_advance(self.h, int(n_cycles))
        if r < 0:
            raise RuntimeError("kro_advance: " + self.lib.kro_last_error().decode())
        return r

    def time(self):
        return (self.lib.kro_time(self.h), self.lib.kro_dt(self.h),
                self.lib.kro_dt_next(self.h))


def plm(q):
    q = np.ascontiguousarray(q, dtype=np.float64)
    ql, qr = C.c_double(), C.c_double()
    load().kro_plm(_dp(q), C.byref(ql), C.byref(qr))
    return ql.value, qr.value


def hllc(wl, wr, gamma):
    wl = np.ascontiguousarray(wl, dtype=np.float64)
    wr = np.ascontiguousarray(wr, dtype=np.float64)
    f = np.zeros(5)
    load().kro_hllc(_dp(wl), _dp(wr), gamma, _dp(f))
    return f


def hlld(wl, wr, gamma):
    wl = np.ascontiguousarray(wl, dtype=np.float64)
    wr = np.ascontiguousarray(wr, dtype=np.float64)
    f = np.zeros(8)
    load().kro_hlld(_dp(wl), _dp(wr), gamma, _dp(f))
    return f


def face_flux(mhd, axis, wl, wr, gamma):
    wl = np.ascontiguousarray(wl, dtype=np.float64)
    wr = np.ascontiguousarray(wr, dtype=np.float64)
    f = np.zeros(8 if mhd else 5)
    load().kro_face_flux(int(mhd), axis, _dp(wl), _dp(wr), gamma, _dp(f))
    return f


def c2p_hydro(u, gamma, rho_floor=1e-9, tmp_floor=1e-9):
    u = np.array(u, dtype=np.float64)
    w = np.zeros(5)
    cs2 = load().kro_c2p_hydro(_dp(u), _dp(w), gamma, rho_floor, tmp_floor)
    return u, w, cs2

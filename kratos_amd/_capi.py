"""ctypes binding of the C-ABI declared in include/kratos_amd.h.

The shared library is built in-tree (``kratos_amd/libkratos_amd.so``) by
``__graft_entry__.build()``.  There is no fallback: if the library is absent
or cannot be loaded, every entry point raises.
"""
import ctypes as C
import os

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libkratos_amd.so")

KR_HYDRO, KR_MHD = 0, 1
KR_FP64, KR_MIXED = 0, 1
KR_PER, KR_OUT, KR_FRE, KR_REF, KR_FIX = 0, 1, 2, 3, 4
KR_PATH_AUTO, KR_PATH_STAGED, KR_PATH_FUSED = 0, 1, 2
PATHS = {"auto": KR_PATH_AUTO, "staged": KR_PATH_STAGED, "fused": KR_PATH_FUSED}
BND_CODES = {"per": KR_PER, "out": KR_OUT, "fre": KR_FRE, "ref": KR_REF, "fixd": KR_FIX}

KERNEL_KINDS = ["c2p", "dt_finalize", "flux_x", "flux_y", "flux_z", "emf", "update",
                "cycle_begin"]  # the fused stage kernel reports under "update"


class KrConfig(C.Structure):
    _fields_ = [
        ("model", C.c_int),
        ("precision", C.c_int),
        ("gamma", C.c_double), ("cfl", C.c_double),
        ("rho_floor", C.c_double), ("tmp_floor", C.c_double),
        ("bnd", C.c_int * 6),
        ("t_0", C.c_double), ("t_lim", C.c_double),
        ("dt_init", C.c_double), ("dt_expand", C.c_double),
        ("top_up", C.c_int),
        ("n_cell_block", C.c_int * 3),
        ("kernel_path", C.c_int),
        ("tile", C.c_int * 3),
        ("fix_w", C.c_double * 5),
        ("fix_b", C.c_double * 3),
    ]


class KrBlockDesc(C.Structure):
    _fields_ = [
        ("id", C.c_int),
        ("level", C.c_int),
        ("group", C.c_int),
        ("reg", C.c_int * 3),
        ("n_ceff", C.c_int * 3),
        ("xf0", C.c_double * 3),
        ("dx0", C.c_double * 3),
        ("nbr", C.c_int * 27),
    ]


class KrStats(C.Structure):
    _fields_ = [
        ("gpu_ms", C.c_double),
        ("cycles", C.c_int),
        ("n_blocks", C.c_int),
        ("n_cells", C.c_longlong),
        ("kernel_ms", C.c_double * 8),
        ("kernel_calls", C.c_longlong * 8),
        ("alg_bytes_per_cell_update", C.c_double),
        ("sweep_tile", C.c_int * 4),
        ("tune_ms", C.c_float * 16),
        ("tune_pending", C.c_int),
        ("graph_ready", C.c_int),
        ("sweep_kernel", C.c_int),
    ]


EXPORTS = {
    "kr_create": (C.c_int, [C.POINTER(KrConfig), C.POINTER(C.c_void_p)]),
    "kr_add_blocks": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(KrBlockDesc)]),
    "kr_upload": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]),
    "kr_download": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]),
    "kr_step": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    "kr_advance": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int)]),
    "kr_get_time": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double),
                              C.POINTER(C.c_double)]),
    "kr_set_time": (C.c_int, [C.c_void_p, C.c_double, C.c_double]),
    "kr_get_group_time": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_double),
                                    C.POINTER(C.c_double), C.POINTER(C.c_int)]),
    "kr_divb_max": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    "kr_divb_max_noalloc": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.c_void_p,
                                      C.c_void_p, C.c_void_p]),
    "kr_profile": (C.c_int, [C.c_void_p, C.c_int]),
    "kr_get_stats": (C.c_int, [C.c_void_p, C.POINTER(KrStats)]),
    "kr_sync": (C.c_int, [C.c_void_p]),
    "kr_last_error": (C.c_char_p, [C.c_void_p]),
    "kr_destroy": (None, [C.c_void_p]),
}

_lib = None


class KratosError(RuntimeError):
    pass


def load():
    """Load libkratos_amd.so (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise KratosError(f"{LIB_PATH} is missing: run __graft_entry__.build() first")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in EXPORTS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc, ctx=None, what=""):
    if rc != 0:
        # with no context, kr_last_error(NULL) gives the reason of the last
        # kr_create on this thread
        m = load().kr_last_error(ctx if ctx else None)
        msg = m.decode() if m else ""
        raise KratosError(f"{what} failed ({rc}): {msg}")

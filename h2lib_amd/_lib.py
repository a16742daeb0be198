"""Loader of the in-tree shared libraries.

The HIP library (libh2hip.so) is the only compute path.  If it has not been
built, importing the package still works, but every operation raises
``RuntimeError``: there is no CPU fallback.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
HIP_SO = os.path.join(_HERE, "libh2hip.so")

_lib = None


class H2HipError(RuntimeError):
    pass


def _sig(lib, name, res, args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = args


def hip():
    """The loaded libh2hip.so (raises if it is missing)."""
    global _lib
    if _lib is None:
        if not os.path.exists(HIP_SO):
            raise H2HipError(
                "libh2hip.so is not built; run __graft_entry__.build() or "
                "make -C h2lib_amd")
        lib = C.CDLL(HIP_SO, mode=C.RTLD_LOCAL)
        vp, u32, i32, sz, dp = (C.c_void_p, C.c_uint32, C.c_int, C.c_size_t,
                                C.c_void_p)
        _sig(lib, "h2hip_last_error", C.c_char_p, [])
        _sig(lib, "h2hip_device_count", i32, [])
        _sig(lib, "h2hip_surface_create", i32,
             [C.POINTER(vp), u32, dp, u32, dp, dp, dp])
        _sig(lib, "h2hip_surface_destroy", None, [vp])
        _sig(lib, "h2hip_surface_geometry", i32, [vp, dp, dp])
        _sig(lib, "h2hip_op_create", i32,
             [C.POINTER(vp), vp, i32, i32, i32, i32, u32, u32, dp, dp])
        _sig(lib, "h2hip_op_destroy", None, [vp])
        _sig(lib, "h2hip_op_set_precision", i32, [vp, i32])
        _sig(lib, "h2hip_nearfield", i32,
             [vp, i32, dp, u32, dp, u32, i32, dp, sz])
        _sig(lib, "h2hip_nfplan_create", i32,
             [C.POINTER(vp), vp, i32, i32, dp, u32, dp, u32, u32, dp])
        _sig(lib, "h2hip_nfplan_destroy", None, [vp])
        _sig(lib, "h2hip_nfplan_entries", C.c_uint64, [vp])
        _sig(lib, "h2hip_nfplan_offsets", i32, [vp, dp])
        _sig(lib, "h2hip_nfplan_work", i32, [vp, vp, i32, dp])
        _sig(lib, "h2hip_nfplan_set_schedule", i32, [vp, i32, i32])
        _sig(lib, "h2hip_nearfield_execute", i32, [vp, vp, i32, dp, vp])
        _sig(lib, "h2hip_stream_sync", i32, [vp])
        _sig(lib, "h2hip_malloc", i32, [C.POINTER(vp), sz])
        _sig(lib, "h2hip_free", i32, [vp])
        _sig(lib, "h2hip_memcpy", i32, [vp, vp, sz, i32])
        _sig(lib, "h2hip_memset", i32, [vp, i32, sz])
        _sig(lib, "h2hip_memset_async", i32, [vp, i32, sz, vp])
        _sig(lib, "h2hip_stream_create", i32, [C.POINTER(vp)])
        _sig(lib, "h2hip_stream_destroy", i32, [vp])
        _sig(lib, "h2hip_event_create", i32, [C.POINTER(vp)])
        _sig(lib, "h2hip_event_destroy", i32, [vp])
        _sig(lib, "h2hip_event_record", i32, [vp, vp])
        _sig(lib, "h2hip_event_elapsed", i32,
             [vp, vp, C.POINTER(C.c_float)])
        _lib = lib
    return _lib


def check(rc, what):
    if rc != 0:
        msg = hip().h2hip_last_error()
        raise H2HipError("%s failed (%d): %s" % (
            what, rc, msg.decode() if msg else ""))


def set_reproducible(on=True):
    """Reproducible mode for the process: H2 products and Krylov inner
    products add in a fixed order (the same bits on every run, and each
    vector of a batched product gets the bits of the single product)
    instead of with floating-point atomics (default, faster)."""
    f = hip().h2hip_set_reproducible
    f.restype, f.argtypes = C.c_int, [C.c_int]
    check(f(int(bool(on))), "h2hip_set_reproducible")


def get_reproducible():
    f = hip().h2hip_get_reproducible
    f.restype, f.argtypes = C.c_int, []
    return bool(f())

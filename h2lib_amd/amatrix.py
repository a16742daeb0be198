"""Dense matrix products on the device: addeval_amatrix_avector and
addevaltrans_amatrix_avector (Library/amatrix.c:862-925) through
h2hip_addeval_amatrix."""
import ctypes as C

import numpy as np

from ._lib import check, hip

_done = False


def _sigs():
    global _done
    if not _done:
        f = hip().h2hip_addeval_amatrix
        f.restype = C.c_int
        f.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_uint32, C.c_uint32,
                      C.c_void_p, C.c_uint64, C.c_void_p, C.c_void_p,
                      C.c_void_p]
        _done = True


def addeval_amatrix(alpha, d_A, rows, cols, ld, d_x, d_y, trans=False,
                    stream=None):
    """y += alpha A x (trans: y += alpha A^* x) for the column-major device
    matrix d_A (rows x cols, leading dimension ld) and device vectors; the
    dtype of d_A (float64 or complex128) selects real or complex."""
    _sigs()
    cplx = d_A.dtype == np.complex128
    if d_x.dtype != d_A.dtype or d_y.dtype != d_A.dtype:
        raise TypeError("addeval_amatrix: A, x and y need one dtype")
    if ld < rows or d_A.n < (cols - 1) * ld + rows or \
            d_x.n < (rows if trans else cols) or \
            d_y.n < (cols if trans else rows):
        raise ValueError("addeval_amatrix: sizes do not fit")
    a = np.array([np.real(alpha), np.imag(alpha)], dtype=np.float64)
    check(hip().h2hip_addeval_amatrix(
        a.ctypes.data_as(C.c_void_p), int(cplx), int(bool(trans)), rows,
        cols, C.c_void_p(d_A.ptr), ld, C.c_void_p(d_x.ptr),
        C.c_void_p(d_y.ptr), C.c_void_p(stream.handle if stream else 0)),
        "h2hip_addeval_amatrix")

"""ctypes view of the reference's 1998 C codec (oracle/_ref/libmars1.so).

Test infrastructure only: `make -C oracle ref` compiles reference/mars1's
encoder sources unchanged into a shared library, and the Fisher tests call
its classification (index_func.c `newclass`, `variance_class`,
coding_func.c `flips`), its `FisherIndexing` and read its `ordering` and
`mapping` tables (globals.h) to check the oracle's restatement.  The
library exists only where the reference sources do; tests skip otherwise.
"""

from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent.parent
LIB = ROOT / "oracle" / "_ref" / "libmars1.so"
SRC = Path("/root/reference/reference/mars1")


class Node(C.Structure):
    pass


# globals.h:97-104 struct c (a Fisher bucket node).
Node._fields_ = [("ptr_x", C.c_short), ("ptr_y", C.c_short), ("sum", C.c_double),
                 ("sum2", C.c_double), ("iso", C.c_short), ("next", C.POINTER(Node))]

_lib = None


def available() -> bool:
    if LIB.exists():
        return True
    if not SRC.exists():
        return False
    return subprocess.run(["make", "-C", str(ROOT / "oracle"), "ref"],
                          capture_output=True).returncode == 0 and LIB.exists()


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not available():
            raise FileNotFoundError(LIB)
        L = C.CDLL(str(LIB))
        DD = C.POINTER(C.POINTER(C.c_double))
        L.newclass.argtypes = [C.c_int, DD, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.newclass.restype = None
        L.variance_class.argtypes = [C.c_int, DD]
        L.variance_class.restype = C.c_int
        L.flips.argtypes = [C.c_int, DD, DD, C.c_int]
        L.flips.restype = None
        L.FisherIndexing.argtypes = [C.c_int, C.c_int]
        L.FisherIndexing.restype = None
        _lib = L
    return _lib


class Rows:
    """A row-major float64 array seen as C `double **` (rows kept alive)."""

    def __init__(self, a: np.ndarray):
        self.a = np.ascontiguousarray(a, np.float64)
        n = self.a.shape[0]
        self.ptrs = (C.POINTER(C.c_double) * n)(
            *[self.a[i].ctypes.data_as(C.POINTER(C.c_double)) for i in range(n)])

    @property
    def pp(self):
        return C.cast(self.ptrs, C.POINTER(C.POINTER(C.c_double)))


def tables():
    L = lib()
    o = np.ctypeslib.as_array((C.c_int * 144).in_dll(L, "ordering")).reshape(24, 6).copy()
    m = np.ctypeslib.as_array((C.c_int * 64).in_dll(L, "mapping")).reshape(8, 8).copy()
    return o, m


def newclass(block: np.ndarray):
    b = Rows(block)
    iso, clas = C.c_int(), C.c_int()
    lib().newclass(b.a.shape[0], b.pp, C.byref(iso), C.byref(clas))
    return iso.value, clas.value


def variance_class(block: np.ndarray) -> int:
    b = Rows(block)
    return lib().variance_class(b.a.shape[0], b.pp)


def flips(block: np.ndarray, iso: int) -> np.ndarray:
    b = Rows(block)
    out = Rows(np.zeros_like(b.a))
    lib().flips(b.a.shape[0], b.pp, out.pp, iso)
    return out.a.copy()


def fisher_index(contracted: np.ndarray, w: int, h: int, size: int, shift: int):
    """FisherIndexing's 72 buckets as sets of (dom_row, dom_col, iso).

    `contracted` is the plane the search classifies (any positive scale of
    the 2x2 sums: the classification only compares sums and variances).
    """
    L = lib()
    tip = int(round(np.log2(size)))
    rows = Rows(contracted)
    C.c_int.in_dll(L, "image_width").value = w
    C.c_int.in_dll(L, "image_height").value = h
    C.c_int.in_dll(L, "SHIFT").value = shift
    C.c_void_p.in_dll(L, "contract").value = C.cast(rows.ptrs, C.c_void_p).value
    heads = (C.POINTER(Node) * (8 * 3 * 24)).in_dll(L, "class_fisher")
    for b in range(72):
        heads[tip * 72 + b] = C.POINTER(Node)()
    L.FisherIndexing(size, tip)
    out = []
    for b in range(72):
        s = []
        p = heads[tip * 72 + b]
        while p:
            n = p.contents
            s.append((n.ptr_x, n.ptr_y, n.iso))
            p = n.next
        out.append(s)
    return out

"""CPU oracle for the betweenness path — TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this module, and only as the checker (or the timed CPU baseline); the
product path never routes through it.  See oracle.c for what each function
restates and which reference lines it follows.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
_LIB = None

_I32P = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_I64P = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_U32P = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_U64P = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
_F64P = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")


def build() -> Path:
    lib = _HERE / "liboracle.so"
    subprocess.run(["gcc", "-O2", "-std=c11", "-fPIC", "-shared", "-fopenmp", "-Wall",
                    str(_HERE / "oracle.c"), "-o", str(lib), "-lm"], check=True)
    return lib


def _lib():
    global _LIB
    if _LIB is None:
        path = _HERE / "liboracle.so"
        if not path.exists():
            build()
        L = C.CDLL(str(path))
        L.orc_sssp.restype = C.c_int64
        L.orc_sssp.argtypes = [C.c_int32, _I64P, _I32P, _I32P, C.c_int32, _I64P, _U64P, _U32P,
                               C.c_void_p, C.POINTER(C.c_int32)]
        L.orc_brandes.restype = C.c_int
        L.orc_brandes.argtypes = [C.c_int32, _I64P, _I32P, _I32P, _I32P, C.c_int32, _F64P,
                                  C.POINTER(C.c_int32)]
        L.orc_brandes_mt.restype = C.c_int
        L.orc_brandes_mt.argtypes = [C.c_int32, _I64P, _I32P, _I32P, _I32P, C.c_int32, _F64P,
                                     C.POINTER(C.c_int32), C.c_int32]
        L.orc_label_bc.restype = C.c_int
        L.orc_label_bc.argtypes = [C.c_int32, _I32P, _I64P, _I32P, _U64P, _U32P, _I64P, _I32P, _I32P, _I32P,
                                   C.c_int32, _F64P, C.c_int32]
        L.orc_rq_pairs.restype = C.c_int
        L.orc_rq_pairs.argtypes = [C.c_int32, _I64P, _I32P, _I32P, _I32P, C.c_int32, _F64P]
        _LIB = L
    return _LIB


class CSR:
    """Symmetric CSR (0-based) from arc arrays, arcs of a vertex in input order."""

    def __init__(self, n: int, u, v, w):
        u = np.asarray(u, np.int64)
        order = np.argsort(u, kind="stable")
        self.n = int(n)
        self.row = np.zeros(n + 1, np.int64)
        np.add.at(self.row, u + 1, 1)
        self.row = np.cumsum(self.row).astype(np.int64)
        self.col = np.ascontiguousarray(np.asarray(v, np.int32)[order])
        self.w = np.ascontiguousarray(np.asarray(w, np.int32)[order])

    @classmethod
    def from_arcgraph(cls, g) -> "CSR":
        return cls(g.n, g.u, g.v, g.w)


def sssp(g: CSR, s: int):
    """(d int64 [-1 unreachable], sigma uint64, hop uint32, overflow bool)."""
    d = np.zeros(g.n, np.int64)
    sg = np.zeros(g.n, np.uint64)
    hp = np.zeros(g.n, np.uint32)
    ovf = C.c_int32(0)
    r = _lib().orc_sssp(g.n, g.row, g.col, g.w, int(s), d, sg, hp, None, C.byref(ovf))
    if r < 0:
        raise MemoryError("oracle sssp")
    return d, sg, hp, bool(ovf.value)


def default_threads() -> int:
    """Worker threads for the oracle: the host's cores, at most 16."""
    import os
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def brandes(g: CSR, sources=None, threads: int = 1):
    """BC (0.5 * sum of dependencies over `sources`, default all) and overflow.
    threads > 1 splits the sources over OpenMP threads (same values within
    rounding; bitwise the same for a given thread count)."""
    src = np.arange(g.n, dtype=np.int32) if sources is None else np.ascontiguousarray(sources, np.int32)
    bc = np.zeros(g.n, np.float64)
    ovf = C.c_int32(0)
    if threads > 1:
        rc = _lib().orc_brandes_mt(g.n, g.row, g.col, g.w, src, len(src), bc, C.byref(ovf), int(threads))
    else:
        rc = _lib().orc_brandes(g.n, g.row, g.col, g.w, src, len(src), bc, C.byref(ovf))
    if rc != 0:
        raise MemoryError("oracle brandes")
    return bc, bool(ovf.value)


def label_bc(g: CSR, tree: dict, sources=None, threads: int = 1):
    """Reference-method BC (label queries + layered Brandes) over `sources`
    from an index's host structure `tree` (parent, label_off, dis, cnt, hop
    by vertex id, as bcx's Index.export() or the reference's TreeDecomp
    give it)."""
    src = np.arange(g.n, dtype=np.int32) if sources is None else np.ascontiguousarray(sources, np.int32)
    bc = np.zeros(g.n, np.float64)
    t = {k: np.ascontiguousarray(tree[k], dt) for k, dt in (
        ("parent", np.int32), ("label_off", np.int64), ("dis", np.int32), ("cnt", np.uint64),
        ("hop", np.uint32))}
    if _lib().orc_label_bc(g.n, t["parent"], t["label_off"], t["dis"], t["cnt"], t["hop"], g.row, g.col, g.w,
                           src, len(src), bc, int(threads)) != 0:
        raise MemoryError("oracle label_bc (allocation failed or not one tree)")
    return bc


def rq_pairs(g: CSR, points):
    """RQ pair-dependency formula over all unordered pairs (tiny graphs)."""
    p = np.ascontiguousarray(points, np.int32)
    out = np.zeros(len(p), np.float64)
    if _lib().orc_rq_pairs(g.n, g.row, g.col, g.w, p, len(p), out) != 0:
        raise MemoryError("oracle rq_pairs")
    return out

"""bcx — MI355X-native exact betweenness centrality (QBC / MHL-QBC / RQ path).

Python mirror of the C ABI in include/bcx.h (ctypes, numpy arrays).  The
native library ``libbcx.so`` next to this file holds the HIP kernels and the
host index builder; importing this package never falls back to anything else:
a missing or stale library raises ``BcxError``.

Reference entry points mirrored (CSU-WRZ/GPU_BetweennessCentrality):
  Graph.read            Graph::ReadGraph (src/Graph.cpp:4)
  Index.build_mhl       TreeDecomp Reduce/MakeTree/MakeIndex (MHLQBC.cu:497-501)
  Index.build_order     CHconsorderMT/makeTree/makeIndex (QBC.cu:421-423)
  Context.full_bc       Graph::brandesBCIndexGPU_5 (MHLQBC.cu:495, QBC.cu:419)
  Context.subset_bc     Graph::brandesBCIndexGPU_PointBatch (RQ.cu:465)
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

__all__ = ["BcxError", "Graph", "Index", "Context", "RQServer", "lib", "MODE_PLAIN", "MODE_BATCHED",
           "MODE_SWEEP"]

MODE_PLAIN = 0
MODE_BATCHED = 1
MODE_SWEEP = 2
SWEEP_FORM_AUTO = 0
SWEEP_FORM_ONE_PER_CU = 1
SWEEP_FORM_TWO_PER_CU = 2
INDEX_DEVICE_LABELS = 1
_PKG = Path(__file__).resolve().parent
_LIB = None


class BcxError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"bcx error {code}: {msg}")
        self.code = code


class IndexStats(C.Structure):
    _fields_ = [("n", C.c_int32), ("tree_height", C.c_int32), ("tree_width", C.c_int32),
                ("count_overflow", C.c_int32), ("label_entries", C.c_int64),
                ("mean_depth", C.c_double), ("build_seconds", C.c_double),
                ("bag_entries", C.c_int64)]


class Tuning(C.Structure):
    _fields_ = [("tile", C.c_int32), ("tchunk", C.c_int32), ("bp_threads", C.c_int32),
                ("batch", C.c_int32), ("sweep_batch", C.c_int32), ("sweep_form", C.c_int32),
                ("unpack_wide", C.c_int32), ("sweep_mem_limit", C.c_int64)]


class Timing(C.Structure):
    _fields_ = [("scan_ms", C.c_double), ("dep_ms", C.c_double), ("reduce_ms", C.c_double),
                ("scan_launches", C.c_int64), ("dep_launches", C.c_int64), ("sources", C.c_int64),
                ("sweep_fwd_ms", C.c_double), ("sweep_bwd_ms", C.c_double),
                ("sweep_asc_ms", C.c_double)]


_P = C.c_void_p
_I32P = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_I64P = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_U32P = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_U64P = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
_F64P = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")

_SIGS = {
    "bcx_last_error": (C.c_char_p, []),
    "bcx_version": (C.c_char_p, []),
    "bcx_graph_read": (C.c_int, [C.c_char_p, C.POINTER(_P)]),
    "bcx_graph_from_arcs": (C.c_int, [C.c_int32, C.c_int64, _I32P, _I32P, _I32P, C.POINTER(_P)]),
    "bcx_graph_info": (C.c_int, [_P, C.POINTER(C.c_int32), C.POINTER(C.c_int64)]),
    "bcx_graph_connected": (C.c_int, [_P]),
    "bcx_graph_free": (None, [_P]),
    "bcx_index_build_mhl": (C.c_int, [_P, C.POINTER(_P)]),
    "bcx_index_build_order": (C.c_int, [_P, C.c_char_p, C.POINTER(_P)]),
    "bcx_index_build_flags": (C.c_int, [_P, C.c_char_p, C.c_void_p, C.c_uint32, C.POINTER(_P)]),
    "bcx_index_build_with_order": (C.c_int, [_P, _I32P, C.POINTER(_P)]),
    "bcx_index_stats_get": (C.c_int, [_P, C.POINTER(IndexStats)]),
    "bcx_index_query": (C.c_int, [_P, C.c_int32, C.c_int32, C.POINTER(C.c_int32),
                                  C.POINTER(C.c_uint64), C.POINTER(C.c_uint32)]),
    "bcx_index_mean_scan": (C.c_int, [_P, C.c_int64, C.c_uint64, C.POINTER(C.c_double)]),
    "bcx_index_build_h2h": (C.c_int, [_P]),
    "bcx_min_degree_order": (C.c_int, [_P, _I32P, C.POINTER(C.c_int64), _I32P]),
    "bcx_index_h2h_scan_stats": (C.c_int, [_P, C.c_int64, C.c_uint64, C.POINTER(C.c_double),
                                           C.POINTER(C.c_double)]),
    "bcx_index_h2h_query": (C.c_int, [_P, C.c_int32, C.c_int32, C.POINTER(C.c_int32)]),
    "bcx_index_tree_order": (C.c_int, [_P, _I32P]),
    "bcx_index_tree": (C.c_int, [_P, _I32P, _I32P]),
    "bcx_index_save": (C.c_int, [_P, C.c_char_p]),
    "bcx_index_load": (C.c_int, [C.c_char_p, C.POINTER(_P)]),
    "bcx_index_free": (None, [_P]),
    "bcx_index_from_arrays": (C.c_int, [C.c_int32, _I32P, _I64P, _I32P, _U64P, _U32P, _I64P, _I32P, _I32P,
                                        _U64P, _U32P, _P, C.POINTER(_P)]),
    "bcx_index_export": (C.c_int, [_P, _I32P, _I64P, _I32P, _U64P, _U32P, _I64P, _I32P, _I32P, _U64P, _U32P]),
    "bcx_index_h2h_from_arrays": (C.c_int, [_P, _I32P, _I64P, _I32P, _I64P, _I32P, C.c_int32, _I32P, _I64P,
                                            _I32P]),
    "bcx_index_h2h_sizes": (C.c_int, [_P, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int32),
                                      C.POINTER(C.c_int64)]),
    "bcx_index_h2h_export": (C.c_int, [_P, _I64P, _I32P, _I64P, _I32P, _I32P, _I64P, _I32P]),
    "bcx_ctx_create": (C.c_int, [C.c_int, _P, _P, C.POINTER(_P)]),
    "bcx_ctx_export_labels": (C.c_int, [_P, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "bcx_ctx_free": (None, [_P]),
    "bcx_ctx_set_tuning": (C.c_int, [_P, C.POINTER(Tuning)]),
    "bcx_ctx_get_tuning": (C.c_int, [_P, C.POINTER(Tuning)]),
    "bcx_ctx_autotune": (C.c_int, [_P, _I32P, C.c_int32, C.POINTER(Tuning)]),
    "bcx_ctx_sweep_capacity": (C.c_int, [_P, C.POINTER(C.c_int32)]),
    "bcx_ctx_sweep_slots": (C.c_int, [_P, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "bcx_ctx_workspace_bytes": (C.c_int, [_P, C.c_int32, C.POINTER(C.c_size_t)]),
    "bcx_ctx_mem_info": (C.c_int, [_P, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
    "bcx_ctx_timing": (C.c_int, [_P, C.POINTER(Timing)]),
    "bcx_ctx_timing_reset": (None, [_P]),
    "bcx_ctx_set_profiling": (C.c_int, [_P, C.c_int]),
    "bcx_ctx_copy_probe": (C.c_int, [_P, C.c_size_t, C.POINTER(C.c_double)]),
    "bcx_ctx_read_probe": (C.c_int, [_P, C.c_size_t, C.POINTER(C.c_double)]),
    "bcx_full_bc": (C.c_int, [_P, C.c_int, _F64P, C.POINTER(C.c_uint32)]),
    "bcx_bc_sources": (C.c_int, [_P, C.c_int, _I32P, C.c_int32, _F64P, C.POINTER(C.c_uint32)]),
    "bcx_query_dist_count": (C.c_int, [_P, C.c_int32, _I32P, _U64P, _U32P]),
    "bcx_query_dist_count_noalloc": (C.c_int, [_P, C.c_int32, _I32P, _U64P, _U32P]),
    "bcx_query_dist_count_mode": (C.c_int, [_P, C.c_int, _I32P, C.c_int32, _I32P, _U64P, C.c_void_p]),
    "bcx_subset_bc": (C.c_int, [_P, _I32P, C.c_int32, _F64P]),
    "bcx_h2h_distances": (C.c_int, [_P, _I32P, C.c_int32, _I32P]),
    "bcx_rq_create": (C.c_int, [_P, C.c_size_t, C.POINTER(_P)]),
    "bcx_rq_create_mode": (C.c_int, [_P, C.c_int, C.c_size_t, C.POINTER(_P)]),
    "bcx_rq_free": (None, [_P]),
    "bcx_rq_submit": (C.c_int, [_P, _I32P, C.c_int32, C.POINTER(C.c_int64)]),
    "bcx_rq_cancel": (C.c_int, [_P, C.c_int64]),
    "bcx_rq_step": (C.c_int, [_P, C.POINTER(C.c_int32)]),
    "bcx_rq_poll": (C.c_int, [_P, C.c_int64, _F64P]),
    "bcx_rq_bytes_needed": (C.c_int, [_P, C.c_int32, C.POINTER(C.c_size_t)]),
    "bcx_rq_memory": (C.c_int, [_P, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
}


def lib() -> C.CDLL:
    """Load libbcx.so (in-tree build) and declare every C ABI signature."""
    global _LIB
    if _LIB is None:
        path = Path(os.environ.get("BCX_LIBRARY", _PKG / "libbcx.so"))
        if not path.exists():
            raise BcxError(-5, f"{path} not built; run `python -m gpu_betweennesscentrality_amd.build`")
        L = C.CDLL(str(path))
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def _check(rc: int) -> int:
    if rc < 0:
        raise BcxError(rc, lib().bcx_last_error().decode(errors="replace"))
    return rc


class Graph:
    """Parsed graph (host memory)."""

    def __init__(self, handle: int):
        self._h = _P(handle)

    @classmethod
    def read(cls, path: str | os.PathLike) -> "Graph":
        h = _P()
        _check(lib().bcx_graph_read(str(path).encode(), C.byref(h)))
        return cls(h.value)

    @classmethod
    def from_arcs(cls, n: int, u, v, w) -> "Graph":
        """0-based arc arrays; every edge listed both ways with one weight >= 1."""
        u = np.ascontiguousarray(u, dtype=np.int32)
        v = np.ascontiguousarray(v, dtype=np.int32)
        w = np.ascontiguousarray(w, dtype=np.int32)
        h = _P()
        _check(lib().bcx_graph_from_arcs(int(n), len(u), u, v, w, C.byref(h)))
        return cls(h.value)

    @property
    def n(self) -> int:
        n = C.c_int32()
        _check(lib().bcx_graph_info(self._h, C.byref(n), None))
        return n.value

    @property
    def arcs(self) -> int:
        m = C.c_int64()
        _check(lib().bcx_graph_info(self._h, None, C.byref(m)))
        return m.value

    def connected(self) -> bool:
        return bool(_check(lib().bcx_graph_connected(self._h)))

    def min_degree_order(self) -> np.ndarray:
        """order[v] = contraction rank (Graph::GraphConstract,
        buildIndex.cpp:759-925); the order file QBC reads."""
        order = np.zeros(self.n, np.int32)
        widths = np.zeros(self.n // 100000 + 1, np.int32)
        arcs = C.c_int64()
        _check(lib().bcx_min_degree_order(self._h, order, C.byref(arcs), widths))
        return order

    def close(self) -> None:
        if self._h:
            lib().bcx_graph_free(self._h)
            self._h = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Index:
    """Tree-decomposition label index (distance, uint64 count, max-hop)."""

    def __init__(self, handle: int):
        self._h = _P(handle)

    @classmethod
    def build_mhl(cls, g: Graph, device_labels: bool = False) -> "Index":
        """device_labels: tree and bags only, the label rows computed by the
        context on the device (no host label rows)."""
        h = _P()
        if device_labels:
            _check(lib().bcx_index_build_flags(g._h, None, None, INDEX_DEVICE_LABELS, C.byref(h)))
        else:
            _check(lib().bcx_index_build_mhl(g._h, C.byref(h)))
        return cls(h.value)

    @classmethod
    def build_order(cls, g: Graph, order, device_labels: bool = False) -> "Index":
        """order: path of an order file, or an array order[v] = contraction rank.
        device_labels: as for build_mhl."""
        h = _P()
        if device_labels:
            if isinstance(order, (str, os.PathLike)):
                _check(lib().bcx_index_build_flags(g._h, str(order).encode(), None, INDEX_DEVICE_LABELS,
                                                   C.byref(h)))
            else:
                o = np.ascontiguousarray(order, dtype=np.int32)
                _check(lib().bcx_index_build_flags(g._h, None, o.ctypes.data_as(C.c_void_p),
                                                   INDEX_DEVICE_LABELS, C.byref(h)))
        elif isinstance(order, (str, os.PathLike)):
            _check(lib().bcx_index_build_order(g._h, str(order).encode(), C.byref(h)))
        else:
            o = np.ascontiguousarray(order, dtype=np.int32)
            _check(lib().bcx_index_build_with_order(g._h, o, C.byref(h)))
        return cls(h.value)

    @classmethod
    def load(cls, path) -> "Index":
        h = _P()
        _check(lib().bcx_index_load(str(path).encode(), C.byref(h)))
        return cls(h.value)

    @classmethod
    def from_arrays(cls, a: dict, graph: "Graph | None" = None) -> "Index":
        """Index from the reference's host structure (keys as export())."""
        n = len(a["parent"])
        args = [np.ascontiguousarray(a[k], dt) for k, dt in (
            ("parent", np.int32), ("label_off", np.int64), ("dis", np.int32), ("cnt", np.uint64),
            ("hop", np.uint32), ("bag_off", np.int64), ("bag_v", np.int32), ("bag_len", np.int32),
            ("bag_cnt", np.uint64), ("bag_hop", np.uint32))]
        h = _P()
        _check(lib().bcx_index_from_arrays(n, *args, graph._h if graph is not None else None, C.byref(h)))
        return cls(h.value)

    def export(self) -> dict:
        """Tree, label rows over ancestors and bags, indexed by vertex id."""
        st = self.stats()
        n, E, B = st["n"], st["label_entries"] - st["n"], st["bag_entries"]
        a = {"parent": np.zeros(n, np.int32), "label_off": np.zeros(n + 1, np.int64),
             "dis": np.zeros(E, np.int32), "cnt": np.zeros(E, np.uint64), "hop": np.zeros(E, np.uint32),
             "bag_off": np.zeros(n + 1, np.int64), "bag_v": np.zeros(B, np.int32),
             "bag_len": np.zeros(B, np.int32), "bag_cnt": np.zeros(B, np.uint64),
             "bag_hop": np.zeros(B, np.uint32)}
        _check(lib().bcx_index_export(self._h, a["parent"], a["label_off"], a["dis"], a["cnt"], a["hop"],
                                      a["bag_off"], a["bag_v"], a["bag_len"], a["bag_cnt"], a["bag_hop"]))
        return a

    def save(self, path) -> None:
        _check(lib().bcx_index_save(self._h, str(path).encode()))

    def stats(self) -> dict:
        st = IndexStats()
        _check(lib().bcx_index_stats_get(self._h, C.byref(st)))
        return {k: getattr(st, k) for k, _ in IndexStats._fields_}

    def query(self, s: int, t: int) -> tuple[int, int, int, bool]:
        d, c, h = C.c_int32(), C.c_uint64(), C.c_uint32()
        rc = _check(lib().bcx_index_query(self._h, s, t, C.byref(d), C.byref(c), C.byref(h)))
        return d.value, c.value, h.value, rc == 1

    def tree_order(self) -> np.ndarray:
        """Vertex ids in DFS preorder of the decomposition tree."""
        n = self.stats()["n"]
        out = np.zeros(n, np.int32)
        _check(lib().bcx_index_tree_order(self._h, out))
        return out

    def tree(self) -> tuple[np.ndarray, np.ndarray]:
        """(parent vertex id or -1, height = depth + 1) per vertex."""
        n = self.stats()["n"]
        par = np.zeros(n, np.int32)
        hgt = np.zeros(n, np.int32)
        _check(lib().bcx_index_tree(self._h, par, hgt))
        return par, hgt

    def build_h2h(self) -> None:
        """QBC's H2H distance labels and separators (buildIndex.cpp:222-278,
        313-386); needed before creating a Context for h2h_distances."""
        _check(lib().bcx_index_build_h2h(self._h))

    def h2h_from_arrays(self, a: dict) -> None:
        """QBC's H2H index from the reference's host structure (keys as
        h2h_export(), plus "parent"), installed on this index's tree."""
        n = self.stats()["n"]
        br = np.ascontiguousarray(a.get("branch", np.zeros(0, np.int32)), np.int32)
        args = [np.ascontiguousarray(a[k], dt) for k, dt in (
            ("parent", np.int32), ("gdis_off", np.int64), ("gdis", np.int32), ("sep_off", np.int64),
            ("sep", np.int32))]
        p2h_off = np.ascontiguousarray(a.get("p2h_off", np.zeros(len(br) * n + 1, np.int64)), np.int64)
        p2h = np.ascontiguousarray(a.get("p2h", np.zeros(0, np.int32)), np.int32)
        _check(lib().bcx_index_h2h_from_arrays(self._h, *args, len(br), br, p2h_off, p2h))

    def h2h_export(self) -> dict:
        """The H2H distance rows (no own entry), separators and P2H tables,
        indexed by vertex id (the layout h2h_from_arrays takes)."""
        n = self.stats()["n"]
        g, s, b, p = C.c_int64(), C.c_int64(), C.c_int32(), C.c_int64()
        _check(lib().bcx_index_h2h_sizes(self._h, C.byref(g), C.byref(s), C.byref(b), C.byref(p)))
        a = {"gdis_off": np.zeros(n + 1, np.int64), "gdis": np.zeros(g.value, np.int32),
             "sep_off": np.zeros(n + 1, np.int64), "sep": np.zeros(s.value, np.int32),
             "branch": np.zeros(b.value, np.int32), "p2h_off": np.zeros(b.value * n + 1, np.int64),
             "p2h": np.zeros(p.value, np.int32)}
        _check(lib().bcx_index_h2h_export(self._h, a["gdis_off"], a["gdis"], a["sep_off"], a["sep"], a["branch"],
                                          a["p2h_off"], a["p2h"]))
        return a

    def h2h_query(self, s: int, t: int) -> int:
        d = C.c_int32()
        _check(lib().bcx_index_h2h_query(self._h, s, t, C.byref(d)))
        return d.value

    def h2h_scan_stats(self, pairs: int = 1 << 18, seed: int = 1) -> tuple[float, float]:
        """Mean H2H scan length: optimizeVS separators only, and with P2H."""
        a, b = C.c_double(), C.c_double()
        _check(lib().bcx_index_h2h_scan_stats(self._h, pairs, seed, C.byref(a), C.byref(b)))
        return a.value, b.value

    def mean_scan(self, pairs: int = 1 << 18, seed: int = 1) -> float:
        m = C.c_double()
        _check(lib().bcx_index_mean_scan(self._h, pairs, seed, C.byref(m)))
        return m.value

    def close(self) -> None:
        if self._h:
            lib().bcx_index_free(self._h)
            self._h = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Context:
    """Device-resident graph + index on one GPU."""

    def __init__(self, g: Graph, ix: Index, device: int = 0):
        h = _P()
        _check(lib().bcx_ctx_create(device, g._h, ix._h, C.byref(h)))
        self._h = h
        self.n = g.n

    def set_tuning(self, **kw) -> None:
        t = self.tuning()
        t.update(kw)
        tn = Tuning(**t)
        _check(lib().bcx_ctx_set_tuning(self._h, C.byref(tn)))

    def tuning(self) -> dict:
        tn = Tuning()
        _check(lib().bcx_ctx_get_tuning(self._h, C.byref(tn)))
        return {k: getattr(tn, k) for k, _ in Tuning._fields_}

    def autotune(self, sources) -> dict:
        """Pick target chunk and dependency-group width on a sample batch."""
        src = np.ascontiguousarray(sources, dtype=np.int32)
        tn = Tuning()
        _check(lib().bcx_ctx_autotune(self._h, src, len(src), C.byref(tn)))
        return {k: getattr(tn, k) for k, _ in Tuning._fields_}

    def set_profiling(self, on: bool) -> None:
        _check(lib().bcx_ctx_set_profiling(self._h, int(on)))

    def copy_probe(self, nbytes: int) -> float:
        """Milliseconds for one device-to-device copy of nbytes."""
        ms = C.c_double()
        _check(lib().bcx_ctx_copy_probe(self._h, nbytes, C.byref(ms)))
        return ms.value

    def read_probe(self, nbytes: int) -> float:
        """Milliseconds for one streaming read of nbytes (multiple of 16)."""
        ms = C.c_double()
        _check(lib().bcx_ctx_read_probe(self._h, nbytes, C.byref(ms)))
        return ms.value

    def timing(self) -> dict:
        t = Timing()
        _check(lib().bcx_ctx_timing(self._h, C.byref(t)))
        return {k: getattr(t, k) for k, _ in Timing._fields_}

    def reset_timing(self) -> None:
        lib().bcx_ctx_timing_reset(self._h)

    def sweep_capacity(self) -> int:
        """Sources one bag-sweep launch takes after the device-memory cap."""
        k = C.c_int32()
        _check(lib().bcx_ctx_sweep_capacity(self._h, C.byref(k)))
        return k.value

    def export_labels(self) -> dict:
        """The context's label rows (as Index.export lays them out), from the device."""
        n = self.n
        off = np.zeros(n + 1, np.int64)
        ptr = lambda x: x.ctypes.data_as(C.c_void_p)
        _check(lib().bcx_ctx_export_labels(self._h, ptr(off), None, None, None))
        E = int(off[-1])
        a = {"label_off": off, "dis": np.zeros(E, np.int32), "cnt": np.zeros(E, np.uint64),
             "hop": np.zeros(E, np.uint32)}
        _check(lib().bcx_ctx_export_labels(self._h, ptr(a["label_off"]), ptr(a["dis"]), ptr(a["cnt"]),
                                           ptr(a["hop"])))
        return a

    def sweep_slots(self) -> tuple[int, int]:
        """(bag-sweep state slots allocated, compute units of the device)."""
        slots, cus = C.c_int32(0), C.c_int32(0)
        _check(lib().bcx_ctx_sweep_slots(self._h, C.byref(slots), C.byref(cus)))
        return slots.value, cus.value

    def workspace_bytes(self, batch: int) -> int:
        b = C.c_size_t()
        _check(lib().bcx_ctx_workspace_bytes(self._h, batch, C.byref(b)))
        return b.value

    def mem_info(self) -> tuple[int, int]:
        f, t = C.c_size_t(), C.c_size_t()
        _check(lib().bcx_ctx_mem_info(self._h, C.byref(f), C.byref(t)))
        return f.value, t.value

    def full_bc(self, mode: int = MODE_BATCHED) -> tuple[np.ndarray, bool]:
        out = np.zeros(self.n, np.float64)
        ovf = C.c_uint32()
        _check(lib().bcx_full_bc(self._h, mode, out, C.byref(ovf)))
        return out, bool(ovf.value)

    def bc_sources(self, sources, mode: int = MODE_BATCHED, bc=None) -> tuple[np.ndarray, bool]:
        src = np.ascontiguousarray(sources, dtype=np.int32)
        out = np.zeros(self.n, np.float64) if bc is None else np.ascontiguousarray(bc, np.float64)
        ovf = C.c_uint32()
        _check(lib().bcx_bc_sources(self._h, mode, src, len(src), out, C.byref(ovf)))
        return out, bool(ovf.value)

    def query(self, source: int) -> tuple[np.ndarray, np.ndarray, np.ndarray, bool]:
        d = np.zeros(self.n, np.int32)
        s = np.zeros(self.n, np.uint64)
        h = np.zeros(self.n, np.uint32)
        rc = _check(lib().bcx_query_dist_count(self._h, source, d, s, h))
        return d, s, h, rc == 1

    def query_noalloc(self, source: int) -> tuple[np.ndarray, np.ndarray, np.ndarray, bool]:
        """query() through the allocation-free entry point (library-held
        staging, graphs up to 1,572,864 vertices)."""
        d = np.zeros(self.n, np.int32)
        s = np.zeros(self.n, np.uint64)
        h = np.zeros(self.n, np.uint32)
        rc = _check(lib().bcx_query_dist_count_noalloc(self._h, source, d, s, h))
        return d, s, h, rc == 1

    def query_rows(self, sources, mode: int = MODE_SWEEP):
        """(d, sigma, hop, overflow) rows [len(sources), n] of one form's
        distance/count stage; hop is None for the bag sweep (no layers)."""
        src = np.ascontiguousarray(sources, np.int32)
        d = np.zeros((len(src), self.n), np.int32)
        s = np.zeros((len(src), self.n), np.uint64)
        h = None if mode == MODE_SWEEP else np.zeros((len(src), self.n), np.uint32)
        rc = _check(lib().bcx_query_dist_count_mode(
            self._h, mode, src, len(src), d, s, None if h is None else h.ctypes.data_as(C.c_void_p)))
        return d, s, h, rc == 1

    def h2h_distances(self, sources) -> np.ndarray:
        """QBC distance-only scan (QBC.cu:243-331): [len(sources), n] int32."""
        src = np.ascontiguousarray(sources, np.int32)
        out = np.zeros((len(src), self.n), np.int32)
        _check(lib().bcx_h2h_distances(self._h, src, len(src), out))
        return out

    def subset_bc(self, points) -> np.ndarray:
        p = np.ascontiguousarray(points, dtype=np.int32)
        out = np.zeros(len(p), np.float64)
        _check(lib().bcx_subset_bc(self._h, p, len(p), out))
        return out

    def close(self) -> None:
        if self._h:
            lib().bcx_ctx_free(self._h)
            self._h = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class RQServer:
    """Continuous-batching subset-BC engine over one Context (bcx_rq_*)."""

    def __init__(self, ctx: Context, hbm_budget_bytes: int, mode: int = MODE_SWEEP):
        h = _P()
        _check(lib().bcx_rq_create_mode(ctx._h, mode, hbm_budget_bytes, C.byref(h)))
        self._h = h
        self._ctx = ctx
        self._q: dict[int, int] = {}

    def bytes_needed(self, q: int) -> int:
        b = C.c_size_t()
        _check(lib().bcx_rq_bytes_needed(self._h, q, C.byref(b)))
        return b.value

    def memory(self) -> tuple[int, int]:
        """(bytes held now, budget)."""
        u, b = C.c_size_t(), C.c_size_t()
        _check(lib().bcx_rq_memory(self._h, C.byref(u), C.byref(b)))
        return u.value, b.value

    def submit(self, points) -> int | None:
        """Request id, or None when it cannot be admitted now (BCX_EBUSY)."""
        p = np.ascontiguousarray(points, dtype=np.int32)
        rid = C.c_int64()
        rc = lib().bcx_rq_submit(self._h, p, len(p), C.byref(rid))
        if rc == -6:
            return None
        _check(rc)
        self._q[rid.value] = len(p)
        return rid.value

    def cancel(self, rid: int) -> None:
        _check(lib().bcx_rq_cancel(self._h, rid))
        self._q.pop(rid, None)

    def step(self) -> int:
        a = C.c_int32()
        _check(lib().bcx_rq_step(self._h, C.byref(a)))
        return a.value

    def poll(self, rid: int) -> np.ndarray | None:
        out = np.zeros(self._q[rid], np.float64)
        done = _check(lib().bcx_rq_poll(self._h, rid, out))
        if done:
            self._q.pop(rid, None)
            return out
        return None

    def close(self) -> None:
        if self._h:
            lib().bcx_rq_free(self._h)
            self._h = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

"""ctypes bindings for the two in-tree shared libraries.

``libptl_host.so``  host front end (source -> DepthTree arrays) and the seeded
                    program generator; plain C++.
``libptl_hip.so``   the code generator itself: HIP kernels for gfx950 behind the
                    extern "C" interface of ``include/ptl.h``.

Both are built in-tree by ``python -m ptilopsis_amd.build`` (or
``__graft_entry__.build()``).  Loading fails loudly when a library is missing:
there is no fallback path.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
HOST_SO = os.path.join(HERE, "libptl_host.so")
HIP_SO = os.path.join(HERE, "libptl_hip.so")

_host = None
_hip = None


class HostTree(C.Structure):
    _fields_ = [
        ("node_types", C.POINTER(C.c_uint8)),
        ("result_types", C.POINTER(C.c_uint8)),
        ("parents", C.POINTER(C.c_int32)),
        ("depth", C.POINTER(C.c_int32)),
        ("child_idx", C.POINTER(C.c_int32)),
        ("node_data", C.POINTER(C.c_uint32)),
        ("n_nodes", C.c_uint64),
        ("max_depth", C.c_uint32),
        ("n_functions", C.c_uint32),
    ]


class GenParams(C.Structure):
    _fields_ = [
        ("seed", C.c_uint64),
        ("n_functions", C.c_uint32),
        ("min_args", C.c_uint32), ("max_args", C.c_uint32),
        ("min_stmts", C.c_uint32), ("max_stmts", C.c_uint32),
        ("nested_min", C.c_uint32), ("nested_max", C.c_uint32),
        ("max_nest", C.c_uint32),
        ("p_assign", C.c_double), ("p_if", C.c_double), ("p_while", C.c_double), ("p_return", C.c_double),
        ("p_var_leaf", C.c_double),
        ("expr_p", C.c_double),
        ("p_leaf_operand", C.c_double),
        ("p_arith", C.c_double), ("p_compare", C.c_double),
        ("float_frac", C.c_double),
        ("balanced", C.c_uint32),
        ("balanced_min", C.c_uint32), ("balanced_max", C.c_uint32),
        ("chain_min", C.c_uint32), ("chain_max", C.c_uint32),
        ("p_chain", C.c_double),
        ("total_target", C.c_uint64),
    ]


def host_lib():
    global _host
    if _host is None:
        if not os.path.exists(HOST_SO):
            raise RuntimeError(f"{HOST_SO} is not built; run python -m ptilopsis_amd.build")
        lib = C.CDLL(HOST_SO)
        lib.ptl_parse_source.argtypes = [C.c_char_p, C.c_uint64, C.POINTER(HostTree), C.c_char_p, C.c_uint64]
        lib.ptl_parse_source.restype = C.c_int
        lib.ptl_free_host_tree.argtypes = [C.POINTER(HostTree)]
        lib.ptl_parse_arena.argtypes = [C.c_char_p, C.c_uint64, C.POINTER(HostArena), C.c_char_p, C.c_uint64]
        lib.ptl_parse_arena.restype = C.c_int
        lib.ptl_free_host_arena.argtypes = [C.POINTER(HostArena)]
        lib.ptl_flatten_records.argtypes = [C.POINTER(HostArena), C.POINTER(HostTree)]
        lib.ptl_flatten_records.restype = C.c_int
        lib.ptl_generate_source.argtypes = [C.POINTER(GenParams), C.POINTER(C.c_uint32), C.POINTER(C.c_char_p),
                                            C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
        lib.ptl_generate_source.restype = C.c_int
        lib.ptl_free_source.argtypes = [C.c_void_p]
        _host = lib
    return _host


def hip_lib():
    global _hip
    if _hip is None:
        if not os.path.exists(HIP_SO):
            raise RuntimeError(f"{HIP_SO} is not built; run python -m ptilopsis_amd.build")
        from . import abi
        _hip = abi.bind(C.CDLL(HIP_SO))
    return _hip


@dataclass
class Tree:
    """The six DepthTree arrays (pareas-ast include/codegen/depthtree.hpp:12-58)."""
    node_types: np.ndarray   # u8, pareas NodeType
    result_types: np.ndarray  # u8, pareas DataType
    parents: np.ndarray      # i32, root -1
    depth: np.ndarray        # i32
    child_idx: np.ndarray    # i32, root -1
    node_data: np.ndarray    # u32
    max_depth: int
    n_functions: int = -1

    @property
    def n(self) -> int:
        return int(self.node_types.shape[0])

    def arrays(self):
        return (self.node_types, self.result_types, self.parents, self.depth, self.child_idx, self.node_data)

    def save(self, path: str) -> None:
        np.savez(path, node_types=self.node_types, result_types=self.result_types, parents=self.parents,
                 depth=self.depth, child_idx=self.child_idx, node_data=self.node_data,
                 max_depth=np.array([self.max_depth], np.int64), n_functions=np.array([self.n_functions], np.int64))

    @staticmethod
    def load(path: str) -> "Tree":
        z = np.load(path)
        return Tree(z["node_types"], z["result_types"], z["parents"], z["depth"], z["child_idx"], z["node_data"],
                    int(z["max_depth"][0]), int(z["n_functions"][0]))


def parse_source(src: str | bytes) -> Tree:
    """Front end: Pareas-language source -> DepthTree arrays (raises ValueError on parse errors)."""
    lib = host_lib()
    b = src.encode() if isinstance(src, str) else src
    t = HostTree()
    err = C.create_string_buffer(512)
    if lib.ptl_parse_source(b, len(b), C.byref(t), err, 512) != 0:
        raise ValueError("parse error: " + err.value.decode(errors="replace"))
    n = t.n_nodes
    try:
        def arr(ptr, dt):
            return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dt, copy=True) if n else np.zeros(0, dt)
        tree = Tree(arr(t.node_types, np.uint8), arr(t.result_types, np.uint8), arr(t.parents, np.int32),
                    arr(t.depth, np.int32), arr(t.child_idx, np.int32), arr(t.node_data, np.uint32),
                    int(t.max_depth), int(t.n_functions))
    finally:
        lib.ptl_free_host_tree(C.byref(t))
    return tree


class HostArena(C.Structure):
    _fields_ = [
        ("node_types", C.POINTER(C.c_uint8)),
        ("result_types", C.POINTER(C.c_uint8)),
        ("node_data", C.POINTER(C.c_uint32)),
        ("first_child", C.POINTER(C.c_uint32)),
        ("n_children", C.POINTER(C.c_uint32)),
        ("kids", C.POINTER(C.c_uint32)),
        ("n_nodes", C.c_uint64),
        ("n_kids", C.c_uint64),
        ("root", C.c_uint32),
        ("n_functions", C.c_uint32),
    ]


@dataclass
class Arena:
    """Parser node records before flattening: node v has the children
    kids[first_child[v] : first_child[v] + n_children[v]]; roots are the
    programs' roots in order (one for a parsed program, several for a forest
    made by concat)."""
    node_types: np.ndarray    # u8
    result_types: np.ndarray  # u8
    node_data: np.ndarray     # u32
    first_child: np.ndarray   # u32
    n_children: np.ndarray    # u32
    kids: np.ndarray          # u32
    roots: np.ndarray         # u32
    n_functions: int = -1

    @property
    def n(self) -> int:
        return int(self.node_types.shape[0])

    @staticmethod
    def concat(arenas) -> "Arena":
        """One forest of the programs, in order: node and kid indices shifted."""
        nb = np.cumsum([0] + [a.n for a in arenas])
        kb = np.cumsum([0] + [int(a.kids.shape[0]) for a in arenas])
        cat = lambda f: np.concatenate([f(a, i) for i, a in enumerate(arenas)])
        return Arena(cat(lambda a, i: a.node_types), cat(lambda a, i: a.result_types),
                     cat(lambda a, i: a.node_data),
                     cat(lambda a, i: (a.first_child + np.uint32(kb[i])).astype(np.uint32)),
                     cat(lambda a, i: a.n_children),
                     cat(lambda a, i: (a.kids + np.uint32(nb[i])).astype(np.uint32)),
                     cat(lambda a, i: (a.roots + np.uint32(nb[i])).astype(np.uint32)),
                     sum(a.n_functions for a in arenas))


def parse_arena(src: str | bytes) -> Arena:
    """Front end stopped before flattening (ptl_parse_arena): the typed node
    records, which flatten to exactly what parse_source returns."""
    lib = host_lib()
    b = src.encode() if isinstance(src, str) else src
    a = HostArena()
    err = C.create_string_buffer(512)
    if lib.ptl_parse_arena(b, len(b), C.byref(a), err, 512) != 0:
        raise ValueError("parse error: " + err.value.decode(errors="replace"))
    n, k = a.n_nodes, a.n_kids
    try:
        def arr(ptr, m, dt):
            return np.ctypeslib.as_array(ptr, shape=(m,)).astype(dt, copy=True) if m else np.zeros(0, dt)
        out = Arena(arr(a.node_types, n, np.uint8), arr(a.result_types, n, np.uint8),
                    arr(a.node_data, n, np.uint32), arr(a.first_child, n, np.uint32),
                    arr(a.n_children, n, np.uint32), arr(a.kids, k, np.uint32),
                    np.array([a.root], np.uint32), int(a.n_functions))
    finally:
        lib.ptl_free_host_arena(C.byref(a))
    return out


def _tree_from_host(lib, t: HostTree) -> Tree:
    n = t.n_nodes
    try:
        def arr(ptr, dt):
            return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dt, copy=True) if n else np.zeros(0, dt)
        return Tree(arr(t.node_types, np.uint8), arr(t.result_types, np.uint8), arr(t.parents, np.int32),
                    arr(t.depth, np.int32), arr(t.child_idx, np.int32), arr(t.node_data, np.uint32),
                    int(t.max_depth), int(t.n_functions))
    finally:
        lib.ptl_free_host_tree(C.byref(t))


def flatten_records(ar: Arena, timing: list | None = None) -> Tree:
    """The host front end's flattening of one program's records
    (ptl_flatten_records, one thread); appends its seconds to `timing`."""
    import time
    lib = host_lib()
    keep = [np.ascontiguousarray(x) for x in (ar.node_types, ar.result_types, ar.node_data, ar.first_child,
                                                ar.n_children, ar.kids)]
    a = HostArena(keep[0].ctypes.data_as(C.POINTER(C.c_uint8)), keep[1].ctypes.data_as(C.POINTER(C.c_uint8)),
                  *[k.ctypes.data_as(C.POINTER(C.c_uint32)) for k in keep[2:]], ar.n, int(ar.kids.shape[0]),
                  int(ar.roots[0]), max(0, ar.n_functions))
    t = HostTree()
    t0 = time.perf_counter()
    if lib.ptl_flatten_records(C.byref(a), C.byref(t)) != 0:
        raise ValueError("bad records")
    if timing is not None:
        timing.append(time.perf_counter() - t0)
    return _tree_from_host(lib, t)


def generate_source(params: dict, targets=None) -> tuple[str, int]:
    """Runs the seeded program generator; returns (source, expected node count)."""
    lib = host_lib()
    p = GenParams()
    for k, v in params.items():
        setattr(p, k, v)
    out = C.c_char_p()
    ln = C.c_uint64()
    nodes = C.c_uint64()
    tp = None
    if targets is not None:
        ta = np.ascontiguousarray(targets, dtype=np.uint32)
        tp = ta.ctypes.data_as(C.POINTER(C.c_uint32))
    if lib.ptl_generate_source(C.byref(p), tp, C.byref(out), C.byref(ln), C.byref(nodes)) != 0:
        raise MemoryError("generator failed")
    try:
        s = C.string_at(out, ln.value).decode()
    finally:
        lib.ptl_free_source(out)
    return s, int(nodes.value)

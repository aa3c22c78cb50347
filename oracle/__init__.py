"""CPU oracle bindings -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import
this package, and only as the checker (or the timed CPU baseline), never as the
path that produces results.  ``liboracle.so`` is built by ``oracle/Makefile``
from ``ptl_oracle.c``, a restatement of the reference's rv_generator_avx at one
thread (src/codegen/rv_generator_avx.cpp).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SO = os.path.join(HERE, "liboracle.so")
REF_DUMP = os.path.join(HERE, "_ref", "pareas_dump")

STAGES = ["preprocess", "isn_cnt", "isn_gen", "optimize", "regalloc", "fix_jumps", "postprocess"]
FLAG_OOB, FLAG_CALL_PATH, FLAG_EMPTY, FLAG_HEAPSORT, FLAG_TIE = 1, 2, 4, 8, 16

_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(SO):
            build()
        L = C.CDLL(SO)
        L.ptlo_create.argtypes = [C.c_void_p] * 6 + [C.c_int64, C.c_uint32]
        L.ptlo_create.restype = C.c_void_p
        L.ptlo_destroy.argtypes = [C.c_void_p]
        L.ptlo_run.argtypes = [C.c_void_p, C.c_int]
        L.ptlo_run.restype = C.c_uint32
        L.ptlo_get.argtypes = [C.c_void_p, C.c_char_p, C.POINTER(C.c_int64)]
        L.ptlo_get.restype = C.POINTER(C.c_uint32)
        L.ptlo_num_functions.argtypes = [C.c_void_p]
        L.ptlo_num_functions.restype = C.c_uint32
        L.ptlo_opt_iters.argtypes = [C.c_void_p]
        L.ptlo_opt_iters.restype = C.c_uint32
        L.ptlo_ties.argtypes = [C.c_void_p]
        L.ptlo_ties.restype = C.c_uint64
        L.ptlo_exclusive_scan.argtypes = [C.c_void_p, C.c_void_p, C.c_int64]
        L.ptlo_introsort.argtypes = [C.c_void_p, C.c_int64, C.c_void_p]
        L.ptlo_asm.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.c_int64]
        L.ptlo_asm.restype = C.c_int64
        L.ptlo_run_scalar.argtypes = [C.c_void_p, C.c_int, C.c_uint32]
        L.ptlo_run_scalar.restype = C.c_uint32
        L.ptlo_run_stage.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_uint32]
        L.ptlo_run_stage.restype = C.c_uint32
        L.ptlo_listing.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.c_int64, C.c_void_p, C.c_int64]
        L.ptlo_listing.restype = C.c_int64
        _lib = L
    return _lib


class Oracle:
    """One oracle run over a tree; stage-boundary arrays via get()."""

    def __init__(self, tree):
        L = lib()
        self._keep = [np.ascontiguousarray(a) for a in tree.arrays()]
        ptrs = [a.ctypes.data for a in self._keep]
        self.h = L.ptlo_create(*ptrs, tree.n, tree.max_depth)
        self.flags = 0

    def run(self, upto: int = 7) -> int:
        self.flags = lib().ptlo_run(self.h, upto)
        return self.flags

    def run_scalar(self, upto: int = 7, avx: int = 0) -> int:
        """The scalar generator's semantics instead (rv_generator_st,
        src/codegen/rv_generator.cpp:227-1736), with the AVX2 behaviours named
        by the SC_* bits in `avx`; a context runs its stages once."""
        self.flags = lib().ptlo_run_scalar(self.h, upto, avx)
        return self.flags

    def run_stage(self, stage: int, scalar: bool = False, avx: int = 0) -> int:
        """The next stage alone, of either semantics (see ptlo_run_stage)."""
        self.flags = lib().ptlo_run_stage(self.h, stage, 1 if scalar else 0, avx)
        return self.flags

    def get(self, name: str) -> np.ndarray:
        n = C.c_int64()
        p = lib().ptlo_get(self.h, name.encode(), C.byref(n))
        if not p or n.value == 0:
            return np.zeros(0, np.uint32)
        return np.ctypeslib.as_array(p, shape=(n.value,)).copy()

    @property
    def n_functions(self) -> int:
        return lib().ptlo_num_functions(self.h)

    @property
    def opt_iters(self) -> int:
        return lib().ptlo_opt_iters(self.h)

    @property
    def ties(self) -> int:
        return lib().ptlo_ties(self.h)

    def result(self) -> dict:
        """Final outputs: words and the function table."""
        return dict(words=self.get("instr"), func_ids=self.get("func_ids"), func_starts=self.get("func_starts"),
                    func_ends=self.get("func_ends"), func_sizes=self.get("func_sizes"))

    def close(self):
        if self.h:
            lib().ptlo_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# SC_* bits of oracle/ptl_scalar.inc: divergences switched to the AVX2 behaviour
SC = {"D1": 1 << 1, "D2": 1 << 2, "D3": 1 << 3, "D4": 1 << 4, "D6": 1 << 6, "D9": 1 << 9, "D10": 1 << 10,
      "D11": 1 << 11, "D12": 1 << 12, "D13": 1 << 13, "D14": 1 << 14, "D15": 1 << 15, "D17": 1 << 17}
SC_ALL = sum(SC.values())


def compile_tree(tree, scalar: bool = False, avx: int = 0) -> dict:
    """The oracle's words and function table for a tree: rv_generator_avx at one
    thread, or with scalar=True rv_generator_st (with the AVX2 behaviours in `avx`)."""
    o = Oracle(tree)
    flags = o.run_scalar(7, avx) if scalar else o.run(7)
    r = o.result()
    r["flags"] = flags
    o.close()
    return r


def exclusive_scan(x: np.ndarray) -> np.ndarray:
    a = np.ascontiguousarray(x, dtype=np.uint32)
    out = np.zeros_like(a)
    lib().ptlo_exclusive_scan(a.ctypes.data, out.ctypes.data, a.shape[0])
    return out


def introsort(keys: np.ndarray) -> np.ndarray:
    k = np.ascontiguousarray(keys, dtype=np.int32)
    idx = np.arange(k.shape[0], dtype=np.uint32)
    lib().ptlo_introsort(idx.ctypes.data, k.shape[0], k.ctypes.data)
    return idx


def to_asm(words) -> bytes:
    """rv_generator::to_asm text of the words (src/codegen/rv_generator.cpp:114-120)."""
    w = np.ascontiguousarray(words, dtype=np.uint32)
    out = np.zeros(max(1, 64 * w.shape[0]), np.uint8)
    n = lib().ptlo_asm(w.ctypes.data, w.shape[0], out.ctypes.data, out.shape[0])
    if n < 0:
        raise RuntimeError("ptlo_asm: output buffer too small")
    return out[:n].tobytes()


def ref_front_end(src_path: str, out_path: str):
    """Runs the reference front end (oracle/_ref/pareas_dump) when it was built."""
    if not os.path.exists(REF_DUMP):
        return None
    r = subprocess.run([REF_DUMP, src_path, out_path], capture_output=True)
    if r.returncode != 0:
        raise ValueError(r.stderr.decode())
    raw = open(out_path, "rb").read()
    n, md = np.frombuffer(raw[:16], np.uint64)
    n = int(n)
    o = 16
    def take(dt, cnt):
        nonlocal o
        a = np.frombuffer(raw[o:o + np.dtype(dt).itemsize * cnt], dt).copy()
        o += np.dtype(dt).itemsize * cnt
        return a
    from ptilopsis_amd.native import Tree
    return Tree(take(np.uint8, n), take(np.uint8, n), take(np.int32, n), take(np.int32, n), take(np.int32, n),
                take(np.uint32, n), int(md))


def listing(words, func_starts) -> bytes:
    """rvdisasm::disassemble(os, words, 0, func_starts) text (src/disassembler.cpp:411-451),
    the listing of rv_generator::print(os, true)."""
    w = np.ascontiguousarray(words, dtype=np.uint32)
    fs = np.ascontiguousarray(func_starts, dtype=np.uint32)
    out = np.zeros(400 * (w.shape[0] + 1) + 256, np.uint8)
    n = lib().ptlo_listing(w.ctypes.data, w.shape[0], fs.ctypes.data, fs.shape[0], out.ctypes.data, out.shape[0])
    if n < 0:
        raise RuntimeError("ptlo_listing: output buffer too small")
    return out[:n].tobytes()

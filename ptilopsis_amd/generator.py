"""Host-side mirror of the reference's rv_generator interface over the HIP library.

The reference's interface (include/codegen/rv_generator.hpp:19-74):
``rv_generator(const DepthTree&)``, ``process(bool profile)``,
``get_instructions()``, ``to_binary``/``to_asm``/``print``, and the protected
function table.  :class:`RvGeneratorHip` keeps those names and their meaning;
errors are raised as ``RuntimeError(ptl_strerror(status))`` the way the
reference throws ``std::runtime_error``.  Every call goes through
``libptl_hip.so``; there is no other code path.
"""
from __future__ import annotations

import ctypes as C
import sys

import numpy as np

from . import abi, tune
from .native import Arena, Tree, hip_lib


class PtlError(RuntimeError):
    def __init__(self, status: int):
        self.status = status
        super().__init__(hip_lib().ptl_strerror(status).decode())


class Context:
    """One ptl_ctx: a HIP stream plus a pooled device arena on one device."""

    def __init__(self, device: int = 0, arena_bytes: int = 0):
        self.lib = hip_lib()
        tune.apply_defaults(self.lib)
        h = C.c_void_p()
        s = self.lib.ptl_create(C.byref(h), device, arena_bytes)
        if s != abi.PTL_OK:
            raise PtlError(s)
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.ptl_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _ast(tree: Tree):
        keep = [np.ascontiguousarray(a) for a in tree.arrays()]
        a = abi.PtlAst(*[k.ctypes.data for k in keep], tree.n, tree.max_depth)
        return a, keep

    def compile(self, tree: Tree) -> dict:
        a, keep = self._ast(tree)
        r = abi.PtlResult()
        s = self.lib.ptl_compile(self.h, C.byref(a), C.byref(r))
        if s != abi.PTL_OK:
            raise PtlError(s)
        return result_dict(r)

    def upload(self, slot: int, tree: Tree) -> None:
        a, keep = self._ast(tree)
        s = self.lib.ptl_upload(self.h, slot, C.byref(a))
        if s != abi.PTL_OK:
            raise PtlError(s)
        self._slot_programs = getattr(self, "_slot_programs", {})
        self._slot_programs[slot] = 1

    def compile_resident(self, slot: int, copy_out: bool = True) -> dict:
        r = abi.PtlResult()
        s = self.lib.ptl_compile_resident(self.h, slot, 1 if copy_out else 0, C.byref(r))
        if s != abi.PTL_OK:
            raise PtlError(s)
        return result_dict(r, copy_out)

    @staticmethod
    def _batch(trees):
        """Concatenated arrays (parents stay local to each program), the i64
        program node offsets and the per-program max depths."""
        arrs = [np.concatenate([np.ascontiguousarray(t.arrays()[k]) for t in trees]) for k in range(6)]
        offs = np.zeros(len(trees) + 1, np.int64)
        offs[1:] = np.cumsum([t.n for t in trees])
        md = np.array([t.max_depth for t in trees], np.uint32)
        a = abi.PtlAst(*[x.ctypes.data for x in arrs], int(offs[-1]), int(md.max()))
        return a, (arrs, offs, md)

    def compile_batch(self, trees, copy_out: bool = True) -> list[dict]:
        """ptl_compile_batch: every program through one pipeline; one result per program."""
        a, keep = self._batch(trees)
        _, offs, md = keep
        out = (abi.PtlResult * len(trees))()
        s = self.lib.ptl_compile_batch(self.h, C.byref(a), offs.ctypes.data, md.ctypes.data, len(trees), out)
        if s != abi.PTL_OK:
            raise PtlError(s)
        return [result_dict(out[p], copy_out) for p in range(len(trees))]

    def upload_batch(self, slot: int, trees) -> None:
        a, keep = self._batch(trees)
        _, offs, md = keep
        s = self.lib.ptl_upload_batch(self.h, slot, C.byref(a), offs.ctypes.data, md.ctypes.data, len(trees))
        if s != abi.PTL_OK:
            raise PtlError(s)
        self._slot_programs = getattr(self, "_slot_programs", {})
        self._slot_programs[slot] = len(trees)

    def compile_resident_batch(self, slot: int, copy_out: bool = True, results: bool = True):
        """ptl_compile_resident_batch; one result dict per program, or (results
        False) none: the library fills its result array either way."""
        n = getattr(self, "_slot_programs", {}).get(slot, 1)
        out = (abi.PtlResult * n)()
        s = self.lib.ptl_compile_resident_batch(self.h, slot, 1 if copy_out else 0, out)
        if s != abi.PTL_OK:
            raise PtlError(s)
        return [result_dict(out[p], copy_out) for p in range(n)] if results else None

    @staticmethod
    def _arena(ar: Arena):
        keep = [np.ascontiguousarray(x) for x in (ar.node_types, ar.result_types, ar.node_data, ar.first_child,
                                                    ar.n_children, ar.kids, ar.roots)]
        a = abi.PtlArena(keep[0].ctypes.data, keep[1].ctypes.data, keep[2].ctypes.data, keep[3].ctypes.data,
                         keep[4].ctypes.data, keep[5].ctypes.data, ar.n, int(ar.kids.shape[0]),
                         keep[6].ctypes.data, int(ar.roots.shape[0]))
        return a, keep

    def tree_build(self, ar: Arena) -> list[Tree]:
        """ptl_tree_build: the DepthTree of every program of the forest, built
        on the device and copied back."""
        a, keep = self._arena(ar)
        n, P = ar.n, int(ar.roots.shape[0])
        t = np.empty(n, np.uint8); rt = np.empty(n, np.uint8); par = np.empty(n, np.int32)
        dep = np.empty(n, np.int32); ci = np.empty(n, np.int32); nd = np.empty(n, np.uint32)
        offs = np.empty(P + 1, np.int64)
        md = C.c_uint32()
        s = self.lib.ptl_tree_build(self.h, C.byref(a), t.ctypes.data, rt.ctypes.data, par.ctypes.data,
                                    dep.ctypes.data, ci.ctypes.data, nd.ctypes.data, offs.ctypes.data, C.byref(md))
        if s != abi.PTL_OK:
            raise PtlError(s)
        out = []
        for p in range(P):
            b, e = int(offs[p]), int(offs[p + 1])
            out.append(Tree(t[b:e].copy(), rt[b:e].copy(), par[b:e].copy(), dep[b:e].copy(), ci[b:e].copy(),
                            nd[b:e].copy(), int(dep[b:e].max()), -1))
        return out

    def upload_arena(self, slot: int, ar: Arena) -> None:
        """ptl_upload_arena: the trees built on the device into a resident slot."""
        a, keep = self._arena(ar)
        s = self.lib.ptl_upload_arena(self.h, slot, C.byref(a))
        if s != abi.PTL_OK:
            raise PtlError(s)
        self._slot_programs = getattr(self, "_slot_programs", {})
        self._slot_programs[slot] = int(ar.roots.shape[0])

    def compile_arena(self, ar: Arena, copy_out: bool = True) -> list[dict]:
        """ptl_compile_arena: build and compile; one result per program."""
        a, keep = self._arena(ar)
        P = int(ar.roots.shape[0])
        out = (abi.PtlResult * P)()
        s = self.lib.ptl_compile_arena(self.h, C.byref(a), out)
        if s != abi.PTL_OK:
            raise PtlError(s)
        return [result_dict(out[p], copy_out) for p in range(P)]

    def profile(self, on: bool = True) -> None:
        self.lib.ptl_profile_enable(self.h, 1 if on else 0)

    def kernel_times(self) -> dict:
        """{kernel name: (total ns, launches)} of the last compile with profiling on."""
        out, i = {}, 0
        name, ns, cnt = C.c_char_p(), C.c_uint64(), C.c_uint32()
        while self.lib.ptl_profile_get(self.h, i, C.byref(name), C.byref(ns), C.byref(cnt)) == 0:
            out[name.value.decode()] = (int(ns.value), int(cnt.value))
            i += 1
        return out

    def stop_after(self, stage: int) -> None:
        self.lib.ptl_debug_stop_after(self.h, stage)

    def _text(self, s, t, n) -> bytes:
        if s != abi.PTL_OK:
            raise PtlError(s)
        return C.string_at(t.value, n.value) if n.value else b""

    def result_asm(self, program: int = 0) -> bytes:
        """to_asm text of program `program` of the last compile, formatted on the device."""
        t, n = C.c_void_p(), C.c_uint64()
        return self._text(self.lib.ptl_result_asm(self.h, program, C.byref(t), C.byref(n)), t, n)

    def result_listing(self, program: int = 0) -> bytes:
        """The listing of rv_generator::print(os, true) for program `program` of
        the last compile (ptl_result_listing), formatted on the device."""
        t, n = C.c_void_p(), C.c_uint64()
        return self._text(self.lib.ptl_result_listing(self.h, program, C.byref(t), C.byref(n)), t, n)

    def disassemble(self, words) -> bytes:
        """to_asm text of host words, formatted on the device."""
        w = np.ascontiguousarray(words, dtype=np.uint32)
        t, n = C.c_void_p(), C.c_uint64()
        return self._text(self.lib.ptl_disassemble(self.h, w.ctypes.data, w.shape[0], C.byref(t), C.byref(n)), t, n)

    def tune_set(self, knob: str, bucket: int, value: int) -> None:
        """One launch parameter of this context (include/ptl.h, ptl_tune_set)."""
        s = self.lib.ptl_tune_set(self.h, knob.encode(), bucket, value)
        if s != abi.PTL_OK:
            raise PtlError(s)

    def tune_get(self, knob: str, bucket: int) -> int:
        v = C.c_uint32()
        s = self.lib.ptl_tune_get(self.h, knob.encode(), bucket, C.byref(v))
        if s != abi.PTL_OK:
            raise PtlError(s)
        return int(v.value)

    def set_overlap(self, on) -> None:
        """Walk and lockstep at the same time (True, default) or one after the
        other (False); 2 forces the lockstep's bounded wait to give up and the
        allocation to be redone one kernel after the other."""
        self.lib.ptl_debug_overlap(self.h, int(on))

    def copy_bandwidth(self, nbytes: int = 1 << 30, reps: int = 5) -> dict:
        """Measured device copy bandwidth (ptl_debug_copy_bandwidth): GB/s moved
        (2 x nbytes per copy) by the fastest of `reps` copies, and the device's
        compute units and HBM size."""
        ns, cus, hbm = C.c_uint64(), C.c_uint32(), C.c_uint64()
        s = self.lib.ptl_debug_copy_bandwidth(self.h, nbytes, reps, C.byref(ns), C.byref(cus), C.byref(hbm))
        if s != abi.PTL_OK:
            raise PtlError(s)
        return {"bytes": nbytes, "best_ns": int(ns.value), "GBps": 2 * nbytes / max(1, ns.value),
                "compute_units": int(cus.value), "hbm_bytes": int(hbm.value)}

    def exclusive_scan(self, x: np.ndarray) -> np.ndarray:
        a = np.ascontiguousarray(x, dtype=np.uint32)
        out = np.zeros_like(a)
        s = self.lib.ptl_debug_exclusive_scan(self.h, a.ctypes.data, out.ctypes.data, a.shape[0])
        if s != abi.PTL_OK:
            raise PtlError(s)
        return out

    def depth_sort(self, keys: np.ndarray, bounds) -> np.ndarray:
        """The device depth sort of each program range [bounds[p], bounds[p+1]) of
        `keys`: node ids by position (ptl_debug_depth_sort)."""
        k = np.ascontiguousarray(keys, dtype=np.int32)
        b = np.ascontiguousarray(bounds, dtype=np.uint32)
        out = np.zeros(k.shape[0], np.uint32)
        s = self.lib.ptl_debug_depth_sort(self.h, k.ctypes.data, b.ctypes.data, b.shape[0] - 1, out.ctypes.data)
        if s != abi.PTL_OK:
            raise PtlError(s)
        return out

    def dump(self, name: str) -> np.ndarray:
        n = self.lib.ptl_debug_dump(self.h, name.encode(), None, 0)
        if n < 0:
            raise PtlError(int(n))
        out = np.zeros(max(n, 1), np.uint32)
        self.lib.ptl_debug_dump(self.h, name.encode(), out.ctypes.data, n)
        return out[:n]


class Server:
    """ptl_server: the request loop over the batched pipeline (continuous
    batching, admission by estimated HBM footprint)."""

    def __init__(self, device: int = 0, hbm_budget: int = 0, max_batch: int = 4096, workers: int = 2):
        self.lib = hip_lib()
        tune.apply_defaults(self.lib)
        h = C.c_void_p()
        s = self.lib.ptl_server_create(C.byref(h), device, hbm_budget, max_batch, workers)
        if s != abi.PTL_OK:
            raise PtlError(s)
        self.h = h

    def submit(self, tree: Tree) -> int:
        keep = [np.ascontiguousarray(a) for a in tree.arrays()]
        a = abi.PtlAst(*[k.ctypes.data for k in keep], tree.n, tree.max_depth)
        t = C.c_uint64()
        s = self.lib.ptl_server_submit(self.h, C.byref(a), C.byref(t))
        if s != abi.PTL_OK:
            raise PtlError(s)
        return int(t.value)

    def submit_records(self, ar: Arena) -> int:
        """ptl_server_submit_arena: one program's parser records; its tree is
        built on the device inside its batch."""
        a, keep = Context._arena(ar)
        t = C.c_uint64()
        s = self.lib.ptl_server_submit_arena(self.h, C.byref(a), C.byref(t))
        if s != abi.PTL_OK:
            raise PtlError(s)
        return int(t.value)

    def wait(self, ticket: int) -> tuple[dict, dict]:
        """(result, times) of a submitted request; its storage is released."""
        r = abi.PtlResult()
        tm = abi.PtlServeTimes()
        s = self.lib.ptl_server_wait(self.h, ticket, C.byref(r), C.byref(tm))
        try:
            if s != abi.PTL_OK:
                raise PtlError(s)
            out = result_dict(r)
        finally:
            self.lib.ptl_server_release(self.h, ticket)
        return out, {k: int(getattr(tm, k)) for k, _ in abi.PtlServeTimes._fields_}

    def stats(self) -> dict:
        v = [C.c_uint64() for _ in range(4)]
        self.lib.ptl_server_stats(self.h, *[C.byref(x) for x in v])
        return dict(zip(("batches", "requests_done", "nodes_done", "largest_batch"), (int(x.value) for x in v)))

    def close(self):
        if getattr(self, "h", None):
            self.lib.ptl_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# rvdisasm::color (include/disassembler.hpp:9-17)
_WHITE, _INSTR, _EXTRA = "\033[37m", "\033[38;5;186m", "\033[38;5;114m"


def stage_table(names, ns) -> str:
    """The stage table rv_generator.cpp:164-202 writes to stderr."""
    total = sum(ns)
    name_len = 1 + max(len(n) for n in names)
    times = [f"{t}ns" for t in ns]
    total_s = f"{total}ns"
    time_len = 1 + max(len(total_s), max(len(t) for t in times))
    pcts = [f"{(t * 100.0) / total:.2f}%" if total else "nan%" for t in ns]
    pct_len = max(12, 1 + max(len(p) for p in pcts))
    out = [f"{'Stage':<{name_len}} {'Duration':<{time_len}} % of total\n", "-" * (name_len + time_len + 12) + "\n"]
    for n, t, p in zip(names, times, pcts):
        out.append(f"{_INSTR}{n:<{name_len}}{_EXTRA}{t:>{time_len}}{p:>{pct_len}}{_WHITE}\n")
    out.append("-" * (name_len + time_len + 12) + "\n")
    out.append(f"{_INSTR}{'Total':<{name_len}}{_EXTRA}{total_s:>{time_len}}{_WHITE}\n")
    return "".join(out)


_EMPTY = np.zeros(0, np.uint32)
_EMPTY.setflags(write=False)


def result_dict(r: abi.PtlResult, copy_out: bool = True) -> dict:
    def arr(p, n):
        if not copy_out or n == 0:
            return _EMPTY
        return np.ctypeslib.as_array(p, shape=(n,)).copy()
    nf = r.n_funcs
    return dict(words=arr(r.words, r.n_words), func_ids=arr(r.func_ids, nf), func_starts=arr(r.func_starts, nf),
                func_ends=arr(r.func_ends, nf), func_sizes=arr(r.func_sizes, nf), n_words=int(r.n_words),
                n_funcs=int(nf), stage_ns=[int(x) for x in r.stage_ns], W0=int(r.n_words_gen),
                W1=int(r.n_words_alloc), opt_iterations=int(r.opt_iterations), tie_slots=int(r.tie_slots))


# pareas DataType names (pareas-ast src/codegen/datatype.cpp:5-12)
_TYPE_NAMES = ("invalid", "void", "int", "float", "int_ref", "float_ref")


class RvGeneratorHip:
    """Drop-in for rv_generator_avx (rv_generator.hpp:19-74; rv_generator_avx.cpp:104-120)."""

    def __init__(self, tree: Tree, ctx: Context | None = None):
        self.tree = tree
        self.ctx = ctx or Context()
        self._res = None

    def process(self, profile: bool = False) -> None:
        """Runs the seven stages; with profile, prints the two-line CSV to stdout,
        then the stage table to stderr, as rv_generator.cpp:150-202 does (stage
        names, times as C++ chrono prints nanoseconds, shares with two decimals,
        the disassembler's colour codes)."""
        self._res = self.ctx.compile(self.tree)
        ns = self._res["stage_ns"]
        if profile:
            print(",".join(abi.STAGES) + "," + abi.STAGES[-1])
            print(",".join(str(x) for x in ns) + "," + str(ns[-1]), flush=True)
        sys.stderr.write(stage_table(abi.STAGES, ns))

    def get_instructions(self) -> np.ndarray:
        return self._res["words"]

    @property
    def function_ids(self): return self._res["func_ids"]

    @property
    def func_starts(self): return self._res["func_starts"]

    @property
    def func_ends(self): return self._res["func_ends"]

    @property
    def function_sizes(self): return self._res["func_sizes"]

    def print(self, stream=None, disassemble: bool = False) -> None:
        """rv_generator::print(os, disassemble) (rv_generator.cpp:75-107): one line
        per node of the generator's arrays after process() -- the rv node type,
        result type name, parent, depth, child index, instruction count, node
        data and location, each right-aligned as there -- then, with
        `disassemble` and at least one word, the listing
        (rvdisasm::disassemble, src/disassembler.cpp:411-451).  The arrays are
        read back from the device after compiling the tree again (the context's
        last compile may be another call's)."""
        stream = stream or sys.stdout
        ctx = self.ctx
        self._res = ctx.compile(self.tree)
        t = self.tree
        n = t.n
        ty, rt, nd = ctx.dump("types"), ctx.dump("rtypes"), ctx.dump("node_data")
        sz, loc = ctx.dump("sizes"), ctx.dump("locs")
        # digits as the reference computes them (1 + log10 of the array's
        # largest value, at least 1; log10 of the node count + 1)
        dig = lambda a: len(str(max(1, int(np.max(a).astype(np.uint32)) if a.size else 1)))
        nd_d, dep_d, sz_d, dat_d, loc_d = len(str(n)), dig(t.depth), dig(sz), dig(nd), dig(loc)
        lines = []
        for i in range(n):
            lines.append(f"Node {i:>{nd_d}}, type = {int(ty[i]):>3}, result = {_TYPE_NAMES[int(rt[i])]:>9}, "
                         f"parent = {int(t.parents[i]):>{nd_d}}, depth = {int(t.depth[i]):>{dep_d}}, "
                         f"child_idx = {int(t.child_idx[i]):>{nd_d + 1}}, size = {int(sz[i]):>{sz_d}}, "
                         f"data = {int(nd[i]):>{dat_d}}, loc = {int(loc[i]):>{loc_d}}\n")
        text = "".join(lines)
        if disassemble and self._res["n_words"]:
            text += ctx.result_listing(0).decode("ascii")
        stream.write(text)

    def to_binary(self, stream) -> None:
        """Raw little-endian words (rv_generator.cpp:116-119)."""
        stream.write(self.get_instructions().astype("<u4").tobytes())

    def to_asm(self, stream) -> None:
        """One line per word, rvdisasm::instruction(word) and a newline
        (rv_generator.cpp:114-120), formatted on the device.  Text streams get
        str, binary streams bytes."""
        txt = self.ctx.disassemble(self.get_instructions())
        try:
            stream.write(txt)
        except TypeError:
            stream.write(txt.decode("ascii"))

"""ctypes declarations of the C ABI in include/ptl.h."""
from __future__ import annotations

import ctypes as C

STAGES = ("preprocess", "isn_cnt", "isn_gen", "optimize", "regalloc", "fix_jumps", "postprocess")

PTL_OK = 0
ERRORS = {
    -1: "PTL_ERR_ARG", -2: "PTL_ERR_INVALID_AST", -3: "PTL_ERR_EMPTY", -4: "PTL_ERR_NOMEM",
    -5: "PTL_ERR_HIP", -6: "PTL_ERR_UNSUPPORTED", -7: "PTL_ERR_ADMISSION", -8: "PTL_ERR_TOO_LARGE",
}

# every symbol include/ptl.h declares
EXPORTS = ("ptl_create", "ptl_destroy", "ptl_strerror", "ptl_compile", "ptl_upload", "ptl_compile_resident",
           "ptl_result_device", "ptl_compile_batch", "ptl_upload_batch", "ptl_compile_resident_batch",
           "ptl_server_create", "ptl_server_destroy", "ptl_server_submit", "ptl_server_wait",
           "ptl_server_release", "ptl_server_stats",
           "ptl_result_asm", "ptl_disassemble", "ptl_tune_set", "ptl_tune_get",
           "ptl_debug_dump", "ptl_debug_stop_after", "ptl_debug_overlap",
           "ptl_debug_exclusive_scan", "ptl_debug_depth_sort", "ptl_profile_enable", "ptl_profile_get",
           "ptl_compile_noalloc", "ptl_tree_build", "ptl_upload_arena", "ptl_compile_arena",
           "ptl_server_submit_arena", "ptl_result_listing", "ptl_debug_copy_bandwidth")


class PtlAst(C.Structure):
    _fields_ = [
        ("node_types", C.c_void_p), ("result_types", C.c_void_p), ("parents", C.c_void_p),
        ("depth", C.c_void_p), ("child_idx", C.c_void_p), ("node_data", C.c_void_p),
        ("n_nodes", C.c_uint64), ("max_depth", C.c_uint32),
    ]


class PtlArena(C.Structure):
    _fields_ = [
        ("node_types", C.c_void_p), ("result_types", C.c_void_p), ("node_data", C.c_void_p),
        ("first_child", C.c_void_p), ("n_children", C.c_void_p), ("kids", C.c_void_p),
        ("n_nodes", C.c_uint64), ("n_kids", C.c_uint64), ("roots", C.c_void_p), ("n_roots", C.c_uint32),
    ]


class PtlResult(C.Structure):
    _fields_ = [
        ("words", C.POINTER(C.c_uint32)), ("n_words", C.c_uint64),
        ("func_ids", C.POINTER(C.c_uint32)), ("func_starts", C.POINTER(C.c_uint32)),
        ("func_ends", C.POINTER(C.c_uint32)), ("func_sizes", C.POINTER(C.c_uint32)),
        ("n_funcs", C.c_uint32),
        ("stage_ns", C.c_uint64 * 7),
        ("n_words_gen", C.c_uint64), ("n_words_alloc", C.c_uint64),
        ("opt_iterations", C.c_uint32), ("tie_slots", C.c_uint32),
    ]


class PtlServeTimes(C.Structure):
    _fields_ = [("queue_ns", C.c_uint64), ("total_ns", C.c_uint64), ("batch_programs", C.c_uint32),
                ("worker", C.c_uint32), ("batch_id", C.c_uint64), ("pack_ns", C.c_uint64), ("run_ns", C.c_uint64),
                ("device_ns", C.c_uint64), ("copyout_ns", C.c_uint64)]


def bind(lib):
    lib.ptl_create.argtypes = [C.POINTER(C.c_void_p), C.c_int, C.c_uint64]
    lib.ptl_create.restype = C.c_int
    lib.ptl_destroy.argtypes = [C.c_void_p]
    lib.ptl_strerror.argtypes = [C.c_int]
    lib.ptl_strerror.restype = C.c_char_p
    lib.ptl_compile.argtypes = [C.c_void_p, C.POINTER(PtlAst), C.POINTER(PtlResult)]
    lib.ptl_compile.restype = C.c_int
    lib.ptl_upload.argtypes = [C.c_void_p, C.c_int, C.POINTER(PtlAst)]
    lib.ptl_upload.restype = C.c_int
    lib.ptl_compile_resident.argtypes = [C.c_void_p, C.c_int, C.c_int, C.POINTER(PtlResult)]
    lib.ptl_compile_resident.restype = C.c_int
    lib.ptl_result_device.argtypes = [C.c_void_p]
    lib.ptl_result_device.restype = C.c_void_p
    lib.ptl_compile_batch.argtypes = [C.c_void_p, C.POINTER(PtlAst), C.c_void_p, C.c_void_p, C.c_uint32,
                                      C.POINTER(PtlResult)]
    lib.ptl_compile_batch.restype = C.c_int
    lib.ptl_upload_batch.argtypes = [C.c_void_p, C.c_int, C.POINTER(PtlAst), C.c_void_p, C.c_void_p, C.c_uint32]
    lib.ptl_upload_batch.restype = C.c_int
    lib.ptl_compile_resident_batch.argtypes = [C.c_void_p, C.c_int, C.c_int, C.POINTER(PtlResult)]
    lib.ptl_compile_resident_batch.restype = C.c_int
    lib.ptl_server_create.argtypes = [C.POINTER(C.c_void_p), C.c_int, C.c_uint64, C.c_uint32, C.c_uint32]
    lib.ptl_server_create.restype = C.c_int
    lib.ptl_server_destroy.argtypes = [C.c_void_p]
    lib.ptl_server_submit.argtypes = [C.c_void_p, C.POINTER(PtlAst), C.POINTER(C.c_uint64)]
    lib.ptl_server_submit.restype = C.c_int
    lib.ptl_server_wait.argtypes = [C.c_void_p, C.c_uint64, C.POINTER(PtlResult), C.POINTER(PtlServeTimes)]
    lib.ptl_server_wait.restype = C.c_int
    lib.ptl_server_release.argtypes = [C.c_void_p, C.c_uint64]
    lib.ptl_server_release.restype = C.c_int
    lib.ptl_server_stats.argtypes = [C.c_void_p] + [C.POINTER(C.c_uint64)] * 4
    lib.ptl_server_stats.restype = C.c_int
    lib.ptl_debug_dump.argtypes = [C.c_void_p, C.c_char_p, C.c_void_p, C.c_uint64]
    lib.ptl_debug_dump.restype = C.c_int64
    lib.ptl_debug_stop_after.argtypes = [C.c_void_p, C.c_int]
    lib.ptl_debug_stop_after.restype = C.c_int
    lib.ptl_result_asm.argtypes = [C.c_void_p, C.c_uint32, C.POINTER(C.c_void_p), C.POINTER(C.c_uint64)]
    lib.ptl_result_asm.restype = C.c_int
    lib.ptl_disassemble.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64, C.POINTER(C.c_void_p), C.POINTER(C.c_uint64)]
    lib.ptl_disassemble.restype = C.c_int
    lib.ptl_tune_set.argtypes = [C.c_void_p, C.c_char_p, C.c_uint32, C.c_uint32]
    lib.ptl_tune_set.restype = C.c_int
    lib.ptl_tune_get.argtypes = [C.c_void_p, C.c_char_p, C.c_uint32, C.POINTER(C.c_uint32)]
    lib.ptl_tune_get.restype = C.c_int
    lib.ptl_debug_overlap.argtypes = [C.c_void_p, C.c_int]
    lib.ptl_debug_overlap.restype = C.c_int
    lib.ptl_debug_exclusive_scan.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64]
    lib.ptl_debug_exclusive_scan.restype = C.c_int
    lib.ptl_debug_depth_sort.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p]
    lib.ptl_debug_depth_sort.restype = C.c_int
    lib.ptl_profile_enable.argtypes = [C.c_void_p, C.c_int]
    lib.ptl_profile_enable.restype = C.c_int
    lib.ptl_profile_get.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_char_p), C.POINTER(C.c_uint64),
                                    C.POINTER(C.c_uint32)]
    lib.ptl_profile_get.restype = C.c_int
    lib.ptl_compile_noalloc.argtypes = [C.c_void_p, C.POINTER(PtlAst), C.c_void_p, C.c_uint64] + \
        [C.c_void_p] * 4 + [C.c_uint32, C.POINTER(PtlResult)]
    lib.ptl_compile_noalloc.restype = C.c_int
    lib.ptl_tree_build.argtypes = [C.c_void_p, C.POINTER(PtlArena)] + [C.c_void_p] * 8
    lib.ptl_tree_build.restype = C.c_int
    lib.ptl_upload_arena.argtypes = [C.c_void_p, C.c_int, C.POINTER(PtlArena)]
    lib.ptl_upload_arena.restype = C.c_int
    lib.ptl_compile_arena.argtypes = [C.c_void_p, C.POINTER(PtlArena), C.POINTER(PtlResult)]
    lib.ptl_compile_arena.restype = C.c_int
    lib.ptl_server_submit_arena.argtypes = [C.c_void_p, C.POINTER(PtlArena), C.POINTER(C.c_uint64)]
    lib.ptl_server_submit_arena.restype = C.c_int
    lib.ptl_result_listing.argtypes = [C.c_void_p, C.c_uint32, C.POINTER(C.c_void_p), C.POINTER(C.c_uint64)]
    lib.ptl_result_listing.restype = C.c_int
    lib.ptl_debug_copy_bandwidth.argtypes = [C.c_void_p, C.c_uint64, C.c_uint32, C.POINTER(C.c_uint64),
                                             C.POINTER(C.c_uint32), C.POINTER(C.c_uint64)]
    lib.ptl_debug_copy_bandwidth.restype = C.c_int
    return lib

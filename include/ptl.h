/* ptl.h -- C ABI of the MI355X code generator (libptl_hip.so).
 *
 * Drop-in boundary for the reference's rv_generator
 * (include/codegen/rv_generator.hpp:19-74): a pareas-ast DepthTree goes in,
 * RISC-V instruction words and the function table come out, bit-exact to
 * rv_generator_avx run with one thread.  Plain pointers and sizes only.
 *
 * Status codes: 0 = PTL_OK; negative values are errors, ptl_strerror() names them.
 * A context owns one HIP stream and a pooled device arena; it is not
 * thread-safe, but separate contexts may be used from separate threads.
 */
#ifndef PTL_H
#define PTL_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

enum {
    PTL_OK = 0,
    PTL_ERR_ARG = -1,          /* null pointer or inconsistent sizes */
    PTL_ERR_INVALID_AST = -2,  /* arrays violate the DepthTree contract */
    PTL_ERR_EMPTY = -3,        /* no functions: the reference takes max() of an empty range */
    PTL_ERR_NOMEM = -4,        /* arena or host allocation failed */
    PTL_ERR_HIP = -5,          /* a HIP runtime call failed */
    PTL_ERR_UNSUPPORTED = -6,  /* input reaches a path the reference leaves undefined */
    PTL_ERR_ADMISSION = -7,    /* serving: request larger than the arena can ever hold */
    PTL_ERR_TOO_LARGE = -8,    /* ptl_compile_noalloc: input above its node limit */
};

/* Input: exactly the six DepthTree arrays (pareas-ast include/codegen/depthtree.hpp:12-58,
 * built by depthtree.cpp:11-100), host memory. Root: parent -1, child_idx -1. */
typedef struct {
    const uint8_t *node_types;    /* pareas NodeType (astnode.hpp:11-56) */
    const uint8_t *result_types;  /* pareas DataType (datatype.hpp:7-14) */
    const int32_t *parents;
    const int32_t *depth;
    const int32_t *child_idx;
    const uint32_t *node_data;
    uint64_t n_nodes;
    uint32_t max_depth;           /* DepthTree::maxDepth() */
} ptl_ast;

/* Output, owned by the context and valid until the next call on it or
 * ptl_destroy.  words == rv_generator::get_instructions()
 * (rv_generator_avx.cpp:82-84); the function table is the protected
 * function_ids / func_starts / func_ends / function_sizes members
 * (rv_generator.hpp:46-51) in post-order function order. */
typedef struct {
    const uint32_t *words;
    uint64_t n_words;
    const uint32_t *func_ids, *func_starts, *func_ends, *func_sizes;
    uint32_t n_funcs;
    uint64_t stage_ns[7];         /* preprocess, isn_cnt, isn_gen, optimize, regalloc, fix_jumps, postprocess */
    uint64_t n_words_gen;         /* W0: words after isn_gen */
    uint64_t n_words_alloc;       /* W1: words after regalloc */
    uint32_t opt_iterations;      /* optimize fixpoint launches */
    uint32_t tie_slots;           /* word slots written by two nodes of equal depth; 0xFFFFFFFF: not
                                     counted (inputs of 4M nodes or more always take the depth sort) */
} ptl_result;

typedef struct ptl_ctx ptl_ctx;

/* Creates a context on HIP device `hip_device` with a device arena of
 * `arena_bytes` (0: grow on demand). */
int ptl_create(ptl_ctx **out, int hip_device, uint64_t arena_bytes);
void ptl_destroy(ptl_ctx *ctx);
const char *ptl_strerror(int status);

/* Replaces rv_generator_avx(tree, 1, 0, false, 0) + process() + the getters
 * (src/main.cpp:121-157): host arrays in, host result out, synchronous. */
int ptl_compile(ptl_ctx *ctx, const ptl_ast *in, ptl_result *out);

/* Device-resident variants for callers that keep trees in HBM.
 * ptl_upload copies a tree into the context's resident slot `slot`
 * (0 <= slot < 64); ptl_compile_resident runs the pipeline on it.  With
 * copy_out == 0 the result's host pointers are NULL and only the sizes are
 * filled; ptl_result_device then yields the device pointer of the words. */
int ptl_upload(ptl_ctx *ctx, int slot, const ptl_ast *in);
int ptl_compile_resident(ptl_ctx *ctx, int slot, int copy_out, ptl_result *out);
const uint32_t *ptl_result_device(ptl_ctx *ctx);

/* Batched compile (no counterpart in the reference, which compiles one
 * DepthTree per rv_generator): n_programs trees concatenated node-wise; program
 * p covers nodes [prog_node_offsets[p], prog_node_offsets[p+1]) (offsets[0] = 0,
 * offsets[n_programs] = concat->n_nodes), its parents are local to that range
 * and max_depths[p] is its DepthTree::maxDepth().  All programs run through one
 * pipeline; out[p] (n_programs entries) is exactly what ptl_compile returns for
 * program p alone, its words and function table pointing into the context's
 * result storage.  stage_ns, opt_iterations and tie_slots are the whole batch's.
 * On the device the programs' words lie back to back in program order. */
int ptl_compile_batch(ptl_ctx *ctx, const ptl_ast *concat, const int64_t *prog_node_offsets,
                      const uint32_t *max_depths, uint32_t n_programs, ptl_result *out);

/* Resident batch: ptl_upload_batch copies a batch into slot `slot` as
 * ptl_upload does a tree; ptl_compile_resident_batch compiles the batch (or a
 * single tree) held there into out[0..n_programs).  ptl_compile_resident
 * refuses a slot holding more than one program (PTL_ERR_ARG). */
int ptl_upload_batch(ptl_ctx *ctx, int slot, const ptl_ast *concat, const int64_t *prog_node_offsets,
                     const uint32_t *max_depths, uint32_t n_programs);
int ptl_compile_resident_batch(ptl_ctx *ctx, int slot, int copy_out, ptl_result *out);

/* Parser node records, the input of the device-side tree construction that
 * replaces DepthTree::DepthTree (pareas-ast src/codegen/depthtree.cpp:11-100):
 * the tree before flattening, host memory.  Node v (records in any order) has
 * the children kids[first_child[v] .. first_child[v] + n_children[v]) in child
 * order; roots[0..n_roots) are the programs' roots (STATEMENT_LIST), in
 * program order.  Each program becomes exactly the DepthTree that
 * depthtree.cpp builds from its tree (post-order with children in order, root
 * depth 0, root parent and child_idx -1), the programs back to back.  Every
 * record must lie in exactly one root's tree, listed once as a child:
 * otherwise PTL_ERR_INVALID_AST. */
typedef struct {
    const uint8_t *node_types;    /* pareas NodeType */
    const uint8_t *result_types;  /* pareas DataType */
    const uint32_t *node_data;
    const uint32_t *first_child;
    const uint32_t *n_children;
    const uint32_t *kids;
    uint64_t n_nodes;
    uint64_t n_kids;
    const uint32_t *roots;
    uint32_t n_roots;
} ptl_arena;

/* Builds the DepthTree arrays on the device and copies them into the caller's
 * arrays (n_nodes entries each, the programs back to back, each program's
 * parents local to its own node range as ptl_compile_batch takes them);
 * prog_node_offsets gets n_roots + 1 entries and max_depth the largest depth. */
int ptl_tree_build(ptl_ctx *ctx, const ptl_arena *in, uint8_t *node_types, uint8_t *result_types, int32_t *parents,
                   int32_t *depth, int32_t *child_idx, uint32_t *node_data, int64_t *prog_node_offsets,
                   uint32_t *max_depth);
/* Builds the trees into resident slot `slot`, as ptl_upload (one root) or
 * ptl_upload_batch (several) would hold them flattened; no flattened copy
 * passes through the host. */
int ptl_upload_arena(ptl_ctx *ctx, int slot, const ptl_arena *in);
/* Build and compile in one call: out[0..n_roots) as ptl_compile_batch returns
 * them for the flattened programs (for one root, as ptl_compile). */
int ptl_compile_arena(ptl_ctx *ctx, const ptl_arena *in, ptl_result *out);

/* Serving (no counterpart in the reference, whose only caller compiles one
 * file per process, src/main.cpp:121-157): a request loop over the batched
 * pipeline.  `workers` threads each own a context (HIP stream + pooled arena)
 * and repeatedly take every queued request that fits the HBM budget still free
 * (at most max_batch), compile them as one batch and return the results, so
 * requests arriving while batches run join the next batch and up to `workers`
 * batches are on the device at once.  A request's device footprint is
 * estimated as 1 MiB + 384 B per node; one larger than the whole budget is
 * refused (PTL_ERR_ADMISSION).  hbm_budget_bytes 0 means 64 GiB.  Requests of
 * 256k nodes or more queue apart from smaller ones (even workers take small
 * requests first, odd workers large ones), so small requests do not wait
 * behind a large one's batch.
 * ptl_server_submit copies the arrays and returns at once; ptl_server_wait
 * blocks until the request is done and fills out (valid until
 * ptl_server_release of that ticket) and, if given, its times.  Each result is
 * exactly what ptl_compile returns for that request alone.  Requests still
 * queued at ptl_server_destroy are dropped; no wait may be pending then. */
typedef struct ptl_server ptl_server;
typedef struct {
    uint64_t queue_ns;            /* submit to the start of its batch */
    uint64_t total_ns;            /* submit to results ready */
    uint32_t batch_programs;      /* programs in the batch it ran in */
    uint32_t worker;              /* worker (context) that ran the batch */
    uint64_t batch_id;            /* 1-based sequence number of the batch in this server */
    uint64_t pack_ns;             /* host: the batch's arrays concatenated */
    uint64_t run_ns;              /* upload, the seven stages and their host synchronisations */
    uint64_t device_ns;           /* of run_ns: the stages' device time (sum of stage_ns) */
    uint64_t copyout_ns;          /* results copied back and split per request */
} ptl_serve_times;
int ptl_server_create(ptl_server **out, int hip_device, uint64_t hbm_budget_bytes, uint32_t max_batch,
                      uint32_t workers);
void ptl_server_destroy(ptl_server *s);
int ptl_server_submit(ptl_server *s, const ptl_ast *in, uint64_t *ticket);
/* A request given as parser records of one program (n_roots == 1; see
 * ptl_arena): its tree is built on the device inside its batch, so no
 * flattened copy is made on the host.  Malformed records fail that request
 * alone (PTL_ERR_INVALID_AST at ptl_server_wait).  The result is the one
 * ptl_compile gives for the flattened tree. */
int ptl_server_submit_arena(ptl_server *s, const ptl_arena *in, uint64_t *ticket);
int ptl_server_wait(ptl_server *s, uint64_t ticket, ptl_result *out, ptl_serve_times *times);
int ptl_server_release(ptl_server *s, uint64_t ticket);
int ptl_server_stats(ptl_server *s, uint64_t *batches, uint64_t *requests_done, uint64_t *nodes_done,
                     uint64_t *largest_batch);

/* Text output: replaces rv_generator::to_asm (src/codegen/rv_generator.cpp:114-120),
 * which prints every word as rvdisasm::instruction(word) (src/disassembler.cpp:453-462)
 * followed by '\n': the mnemonic, one space, the operands, exactly as the
 * reference prints them (its switch fall-throughs included).  The lines are
 * formatted on the device.  ptl_result_asm formats program `program` (0 for a
 * single tree) of the last compile, whose words are still in HBM;
 * ptl_disassemble formats host words (and, like a compile, replaces the
 * context's previous results).  The text (not NUL-terminated) is owned by the
 * context and valid until the next call on it.  PTL_ERR_ARG when there is no
 * complete compile or program is out of range.  Raw words (to_binary,
 * rv_generator.cpp:109-112) are ptl_result.words as they are. */
int ptl_result_asm(ptl_ctx *ctx, uint32_t program, const char **text, uint64_t *n_bytes);
int ptl_disassemble(ptl_ctx *ctx, const uint32_t *words, uint64_t n_words, const char **text, uint64_t *n_bytes);
/* The listing of rv_generator::print(os, true) (src/codegen/rv_generator.cpp:100-103),
 * i.e. rvdisasm::disassemble(os, words, 0, func_starts) (src/disassembler.cpp:411-451)
 * of program `program` of the last compile: the header line, then per word its
 * index, byte address, bytes, mnemonic and coloured operands (branch targets
 * absolute), with a "func<k>:" line before each function's first word; the
 * reference's ANSI colours included.  Empty text for a program without words
 * (the reference prints no listing then).  Text ownership and errors as
 * ptl_result_asm. */
int ptl_result_listing(ptl_ctx *ctx, uint32_t program, const char **text, uint64_t *n_bytes);

/* Host-side stage dumps for differential testing (stage 1..7 boundary arrays
 * of the last compile): name is one of "types", "rtypes", "node_data", "locs",
 * "func_ids", "func_starts", "func_ends", "func_sizes", "instr", "rd", "rs1",
 * "rs2", "jt", "used", "sym_reg", "sym_sw", "sorted", and per program
 * "ra_rows" (steps of the allocation lockstep) and "ra_maxsize" (its longest
 * function, in words).  Copies up to cap elements (u32) into dst and returns
 * the element count, or a negative status. */
int64_t ptl_debug_dump(ptl_ctx *ctx, const char *name, uint32_t *dst, uint64_t cap);

/* Runs the pipeline only up to and including `stage` (1..7) on the next
 * ptl_compile (0 = all stages); for stage-boundary parity tests. */
int ptl_debug_stop_after(ptl_ctx *ctx, int stage);

/* Launch parameters chosen by the autotuner (ptilopsis_amd/tune.py writes
 * ptilopsis_amd/tuning.json; the Python Context applies it): per size bucket of
 * a compile's node count (0: < 64Ki nodes, 1: < 1Mi, 2: < 8Mi, 3: larger), the
 * block size of each stage's per-node and per-word kernels ("block.preprocess",
 * "block.isn_cnt", "block.isn_gen", "block.optimize", "block.emit",
 * "block.fix_jumps": 64, 128 or 256) and the register-allocation walks' ring
 * ("walk_ring": 4 or 8 entries).  Output never depends on them.  No counterpart
 * in the reference.  PTL_ERR_ARG for unknown knobs, buckets or values. */
/* With ctx NULL, ptl_tune_set / ptl_tune_get address the process-wide defaults
 * that each context (the serving workers' included) copies at creation. */
int ptl_tune_set(ptl_ctx *ctx, const char *knob, uint32_t bucket, uint32_t value);
int ptl_tune_get(ptl_ctx *ctx, const char *knob, uint32_t bucket, uint32_t *value);

/* Register allocation: lockstep launches of few programs run at the same time
 * as the walk (on = 1, the default) or every lockstep after it (on = 0); the
 * output is the same.  on = 2 overlaps but has the first overlapped program's
 * lockstep wait for one more walk workgroup than there is: its bounded wait
 * gives up (0.25 s) and the library redoes the allocation one kernel after the
 * other, as it does when the device serializes the kernels.  PTL_ERR_ARG for
 * other values. */
int ptl_debug_overlap(ptl_ctx *ctx, int on);

/* Per-kernel timing: with profiling on, every kernel launch of the following
 * calls is bracketed by HIP events on the context's stream.  ptl_profile_get
 * returns, for entry idx (0-based, by first launch), the kernel name, the
 * summed duration over all its launches in the last compile, and the number
 * of launches; PTL_ERR_ARG past the last entry. */
int ptl_profile_enable(ptl_ctx *ctx, int on);
int ptl_profile_get(ptl_ctx *ctx, int idx, const char **name, uint64_t *total_ns, uint32_t *launches);

/* The device exclusive-scan primitive on its own (host arrays in and out),
 * for the known-answer test of benchmarks/prefix_sum.cpp:35-40,215-226. */
int ptl_debug_exclusive_scan(ptl_ctx *ctx, const uint32_t *in, uint32_t *out, uint64_t n);
/* Measured streaming bandwidth of the device: the fastest of `reps` plain copies
 * of `bytes` bytes (read + write = 2 x bytes moved) in *best_ns, with the device's
 * compute-unit count and HBM size; for reporting kernels against a measured
 * peak beside the published one. */
int ptl_debug_copy_bandwidth(ptl_ctx *ctx, uint64_t bytes, uint32_t reps, uint64_t *best_ns, uint32_t *cus,
                              uint64_t *hbm_bytes);

/* The device depth sort on its own: each program's node range
 * [bounds[p], bounds[p+1]) of `keys` is ordered as std::sort orders node ids
 * by depth in rv_generator_avx.cpp:539-541 (libstdc++ introsort), and
 * out[position] = node id.  bounds[0] = 0, bounds[n_programs] = n; every
 * program non-empty (PTL_ERR_ARG otherwise).  For tests of the sort on chosen
 * key patterns; a compile runs it itself. */
int ptl_debug_depth_sort(ptl_ctx *ctx, const int32_t *keys, const uint32_t *bounds, uint32_t n_programs,
                         uint32_t *out);

/* Compile without allocating.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (a tight serving loop, a harness
 * or code region that forbids hipMalloc).  Same result as ptl_compile of the
 * same tree, but the inputs are device arrays (a ptl_ast whose pointers are
 * device addresses), the outputs go into the caller's device buffers (words:
 * words_cap entries; the four function-table arrays: funcs_cap entries each),
 * and every scratch buffer of the pipeline lives in a fixed 128 MiB device
 * array inside the library, one per GPU, held until the process ends.  The
 * call allocates no device or pinned host memory.  Inputs above 262144 nodes
 * are refused (PTL_ERR_TOO_LARGE); if the output does not fit, the call
 * returns PTL_ERR_ARG with out->n_words / out->n_funcs set to the sizes
 * needed.  Inputs are checked on the device against the DepthTree contract
 * (PTL_ERR_INVALID_AST, PTL_ERR_EMPTY as ptl_compile).  On success out holds
 * the sizes and stage times, and its pointer fields are the caller's device
 * buffers.  Calls on one GPU share the array: the library serializes them
 * (per-device lock), and each call leaves its scratch contents behind.
 * Synchronous; the context's results of an earlier call are dropped. */
int ptl_compile_noalloc(ptl_ctx *ctx, const ptl_ast *dev_in, uint32_t *words, uint64_t words_cap,
                        uint32_t *func_ids, uint32_t *func_starts, uint32_t *func_ends, uint32_t *func_sizes,
                        uint32_t funcs_cap, ptl_result *out);

#ifdef __cplusplus
}
#endif
#endif

// MI355X (gfx950) implementation of the reference's data-parallel RISC-V code
// generator: rv_generator_avx with one thread (src/codegen/rv_generator_avx.cpp),
// stage for stage, bit-exact.
//
// Every stage is integer gather/scatter/scan work bound by HBM bandwidth; no
// stage is a dense contraction.  Arrays are structure-of-arrays in HBM; u8 for
// node and data types, u32 elsewhere.  Stages and where their semantics come from:
//
//   preprocess   k_pre12 (passes 1+2, :126-266), k_pre3 (pass 3, :270-305)
//   isn_cnt      k_scan<CountOp> (count :315-369 + exclusive scan :374-392),
//                k_cnt_fix (:398-447), function table (:451-515)
//   isn_gen      k_gen_regs (parent-slot propagation) then k_gen_owner /
//                k_gen_write (:584-886).  The reference walks depth levels
//                bottom-up, but every parent slot has one writer and is read
//                only by its parent, so both passes run over all nodes at once;
//                word slots written by more than one node go to the writer the
//                reference writes last: the shallowest node, then the later one
//                in the reference's depth sort (its libstdc++ std::sort order,
//                reproduced exactly by the introsort kernels below).
//   optimize     k_opt_init + k_opt_iter (:891-1016), a chaotic in-place
//                iteration of the same monotone operator; it reaches the same
//                greatest fixpoint as the reference's two-copy iteration.
//   regalloc     k_ra_walk (functions not in lane 0 of an 8-function group:
//                independent), k_ra_lane0 (lane-0 functions in step lockstep:
//                they share the first 64-entry region of the reference's
//                register_state snapshot, :1290-1354), k_ra_serial (exact
//                in-order walk for programs where two functions name one
//                register), then counts, scan, emit (:1422-1697), function
//                table fix and frame words (:1714-1811).
//   fix_jumps    size + scan + fused move/encode/postprocess (:1853-1993, 1995-2011).
//
// Batches: P programs concatenated node-wise run through the same kernels.
// Each element finds its program by binary search in the program tables (node,
// word and function bases); isn_gen computes in the program's own word indices
// (register numbers are ORed with them) and then offsets virtual registers and
// jump targets by the program's word base, so every program's words are
// exactly those of its own compile, back to back.
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <chrono>
#include <string>
#include <vector>
#include <algorithm>
#include <future>
#include <mutex>
#include <type_traits>

#include "rv_spec.h"
#include "../../include/ptl.h"

#define DEV __device__ __forceinline__

// ---------------------------------------------------------------- constants
__constant__ uint32_t c_word[RV_NTYPES * 4 * 6];
__constant__ uint32_t c_bits[RV_NTYPES * 4 * 6];
__constant__ uint8_t c_dense[256];
__constant__ uint8_t c_pareas[44];

DEV uint32_t spec_i(uint32_t ty, uint32_t o, uint32_t rt) { return ((uint32_t)c_dense[ty] * 4u + o) * 6u + rt; }
DEV uint32_t size_of(uint32_t ty, uint32_t rt) { return (c_bits[spec_i(ty, 0, rt)] >> 19) & 7u; }

// error flags raised by kernels (bit set in ctx->d_flags)
// EF_XNAMED: some function names a virtual register of a word outside itself;
// EF_XCONFLICT: such a register is also named by another function (see xref_note).
// EF_WAIT: the allocation lockstep gave up waiting for the concurrent walk's rows.
enum : uint32_t { EF_OOB = 1, EF_CALLPATH = 4, EF_XNAMED = 16, EF_XCONFLICT = 32, EF_WAIT = 64 };

// ------------------------------------------------------------------ helpers
// Workgroup barrier that orders LDS only: global loads and stores stay in
// flight across it.  Correct where every global location a thread reads was
// last written by that same thread (program order covers it) or before the
// kernel started, and where global atomics are read only by later kernels.
DEV void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Program of element i given the programs' bases (last p with base[p] <= i).
DEV uint32_t prog_of(const uint32_t *base, uint32_t P, uint32_t i) {
    uint32_t lo = 0, hi = P;
    while (hi - lo > 1) { const uint32_t m = (lo + hi) >> 1; if (base[m] <= i) lo = m; else hi = m; }
    return lo;
}

// The same where neighbouring lanes look up neighbouring elements (node- or
// word-parallel kernels): one search for the wave's first active lane, on a
// wave-uniform index, then each lane steps from there to its own program (the
// programs are long against a wave, so mostly no step).
DEV uint32_t prog_of_w(const uint32_t *base, uint32_t P, uint32_t i) {
    const uint32_t i0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)i);
    uint32_t p = prog_of(base, P, i0);
    while (p > 0 && base[p] > i) --p;
    while (p + 1 < P && base[p + 1] <= i) ++p;
    return p;
}

DEV uint32_t ffz_gcc(uint64_t v) {
    // include/utils.hpp:25-45 as compiled by GCC: __builtin_ffs(int) of ~val sees 32 bits.
    uint32_t low = (uint32_t)(~v);
    return low == 0 ? 64u : (uint32_t)__builtin_ctz(low);
}

template <typename T>
__global__ void k_fill(T *p, uint64_t n, T v) {
    uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

__global__ void k_set_u8(uint8_t *p, uint8_t v) { *p = v; }

// A plain streaming copy (16 B per thread and step, grid-stride), for the
// measured HBM bandwidth the stages are reported against.
__global__ __launch_bounds__(256) void k_copy16(const uint4 *a, uint4 *b, uint64_t n) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += stride) b[i] = a[i];
}

static inline uint32_t grid_for(uint64_t n, uint32_t block) {
    uint64_t g = (n + block - 1) / block;
    if (g == 0) g = 1;
    return (uint32_t)std::min<uint64_t>(g, 0x7FFFFFFFull);
}

// ------------------------------------------------------- reduce-then-scan
// Exclusive u32 scan over n elements in three launches with no inter-block
// waiting: per-tile sums, one workgroup scanning the tile sums, then each tile
// rescans its elements with its offset.  A tile is 256 threads x 16 items.
// Elements are produced striped (element k*256 + t by thread t, so every load
// of Op::value is coalesced) and, for the rescan, transposed through LDS so
// that each thread owns 16 consecutive elements, which a wave64 shuffle scan
// and LDS combine.  Op::value(i) produces element i, Op::emit(i, exclusive,
// value) consumes the result.
constexpr int SCAN_BLOCK = 256;
constexpr int SCAN_ITEMS = 16;
constexpr uint64_t SCAN_TILE = SCAN_BLOCK * SCAN_ITEMS;

DEV uint32_t wave_incl_scan(uint32_t x, uint32_t lane) {
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t y = __shfl_up(x, d, 64);
        if (lane >= (uint32_t)d) x += y;
    }
    return x;
}

DEV uint32_t wave_sum(uint32_t x) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    return x;
}

template <class Op>
__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_reduce(Op op, uint64_t n, uint32_t *tile_sums) {
    __shared__ uint32_t s_w[SCAN_BLOCK / 64];
    const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE + threadIdx.x;
    uint32_t sum = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        uint64_t i = base + (uint64_t)k * SCAN_BLOCK;
        if (i < n) sum += op.value(i);
    }
    sum = wave_sum(sum);
    if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = sum;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
#pragma unroll
        for (int w = 0; w < SCAN_BLOCK / 64; ++w) t += s_w[w];
        tile_sums[blockIdx.x] = t;
    }
}

// One workgroup: exclusive scan of the tile sums in place; the grand total to *total.
__global__ __launch_bounds__(1024) void k_scan_tiles(uint32_t *sums, uint32_t ntiles, uint32_t *total) {
    __shared__ uint32_t s_w[16];
    __shared__ uint32_t s_carry;
    const uint32_t lane = threadIdx.x & 63u, wid = threadIdx.x >> 6;
    if (threadIdx.x == 0) s_carry = 0;
    __syncthreads();
    for (uint32_t b = 0; b < ntiles; b += 1024) {
        const uint32_t i = b + threadIdx.x;
        const uint32_t v = i < ntiles ? sums[i] : 0u;
        const uint32_t x = wave_incl_scan(v, lane);
        if (lane == 63) s_w[wid] = x;
        __syncthreads();
        uint32_t pre = s_carry, all = 0;
        for (uint32_t w = 0; w < 16; ++w) { if (w < wid) pre += s_w[w]; all += s_w[w]; }
        if (i < ntiles) sums[i] = pre + x - v;
        __syncthreads();
        if (threadIdx.x == 0) s_carry += all;
        __syncthreads();
    }
    if (threadIdx.x == 0 && total) *total = s_carry;
}

// The rescan of one tile: v[k] = element base + k of this thread (base = tile
// start + 16 * thread), and the exclusive prefix of its first element.
template <class Op>
DEV uint32_t scan_tile(const Op &op, uint64_t n, const uint32_t *tile_pre, uint32_t *s_w, uint32_t *s_v,
                       uint32_t (&v)[SCAN_ITEMS]) {
    const uint64_t tb = (uint64_t)blockIdx.x * SCAN_TILE;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        const uint32_t e = (uint32_t)k * SCAN_BLOCK + threadIdx.x;
        s_v[e + e / 16u] = tb + e < n ? op.value(tb + e) : 0u;  // element e at e + e/16 (one pad word per 16)
    }
    __syncthreads();
    uint32_t sum = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        v[k] = s_v[threadIdx.x * (SCAN_ITEMS + 1) + k];  // element t*16 + k
        sum += v[k];
    }
    const uint32_t lane = threadIdx.x & 63u, wid = threadIdx.x >> 6;
    const uint32_t x = wave_incl_scan(sum, lane);
    if (lane == 63) s_w[wid] = x;
    __syncthreads();
    uint32_t run = tile_pre[blockIdx.x] + x - sum;
#pragma unroll
    for (int w = 0; w < SCAN_BLOCK / 64; ++w) if ((uint32_t)w < wid) run += s_w[w];
    return run;
}

template <class Op>
__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_down(Op op, uint64_t n, const uint32_t *tile_pre) {
    __shared__ uint32_t s_w[SCAN_BLOCK / 64];
    __shared__ uint32_t s_v[SCAN_TILE + SCAN_TILE / 16];
    const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE + (uint64_t)threadIdx.x * SCAN_ITEMS;
    uint32_t v[SCAN_ITEMS];
    uint32_t run = scan_tile(op, n, tile_pre, s_w, s_v, v);
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        uint64_t i = base + k;
        if (i < n) op.emit(i, run, v[k]);
        run += v[k];
    }
}

// ------------------------------------------------------- single-pass scan
// The same exclusive scan in one launch with decoupled look-back: each element
// is produced once.  A workgroup takes the next tile from a ticket counter, so
// tiles start in order and every tile it waits on is already running; it
// scans its tile as above, publishes the tile's aggregate, and wave 0 looks
// back over the published words of its predecessors 64 at a time, summing
// aggregates back to the nearest inclusive prefix (tile -1 counts as an
// inclusive 0) and waiting while a nearer word is unpublished.  It then
// publishes its own inclusive prefix and the tile emits.  Status word of a
// tile: value | flag << 32 (1 aggregate, 2 inclusive), all zero before the
// launch; the ticket counter is the word after the last tile's.  A status
// word carries its own value, so no other data is ordered by it: the stores
// and polls are relaxed device-scope atomics.
constexpr uint64_t SCAN_AGG = 1ull << 32, SCAN_INCL = 2ull << 32;

template <class Op>
__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_dlb(Op op, uint64_t n, unsigned long long *status,
                                                         uint32_t ntiles, uint32_t *total) {
    __shared__ uint32_t s_w[SCAN_BLOCK / 64];
    __shared__ uint32_t s_v[SCAN_TILE + SCAN_TILE / 16];
    __shared__ uint32_t s_tile, s_pre;
    const uint32_t lane = threadIdx.x & 63u, wid = threadIdx.x >> 6;
    if (threadIdx.x == 0) s_tile = atomicAdd(reinterpret_cast<uint32_t *>(status + ntiles), 1u);
    __syncthreads();
    const uint32_t tile = s_tile;
    const uint64_t tb = (uint64_t)tile * SCAN_TILE;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        const uint32_t e = (uint32_t)k * SCAN_BLOCK + threadIdx.x;
        s_v[e + e / 16u] = tb + e < n ? op.value(tb + e) : 0u;  // element e at e + e/16 (one pad word per 16)
    }
    __syncthreads();
    uint32_t v[SCAN_ITEMS], sum = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        v[k] = s_v[threadIdx.x * (SCAN_ITEMS + 1) + k];  // element t*16 + k
        sum += v[k];
    }
    const uint32_t x = wave_incl_scan(sum, lane);
    if (lane == 63) s_w[wid] = x;
    __syncthreads();
    uint32_t agg = 0, run = x - sum;
#pragma unroll
    for (int w = 0; w < SCAN_BLOCK / 64; ++w) {
        agg += s_w[w];
        if ((uint32_t)w < wid) run += s_w[w];
    }
    if (wid == 0) {
        if (lane == 0)
            __hip_atomic_store(status + tile, (unsigned long long)agg | (tile == 0 ? SCAN_INCL : SCAN_AGG),
                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        uint32_t pre = 0;
        int64_t j = (int64_t)tile - 1;  // nearest predecessor not yet summed (lane 0 reads it)
        while (j >= 0) {
            const int64_t idx = j - (int64_t)lane;
            const unsigned long long st =
                idx >= 0 ? __hip_atomic_load(status + idx, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : SCAN_INCL;
            const uint64_t incl = __ballot((st >> 32) == 2u), none = __ballot((st >> 32) == 0u);
            const uint32_t stop = incl ? (uint32_t)__builtin_ctzll(incl) : 64u;
            if (none && (uint32_t)__builtin_ctzll(none) < stop) {  // a nearer tile has not published yet
                __builtin_amdgcn_s_sleep(1);
                continue;
            }
            pre += wave_sum(lane <= stop ? (uint32_t)st : 0u);
            if (incl) break;
            j -= 64;
        }
        if (lane == 0) {
            if (tile > 0)
                __hip_atomic_store(status + tile, (unsigned long long)(pre + agg) | SCAN_INCL, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
            s_pre = pre;
            if (total && tile == ntiles - 1) *total = pre + agg;
        }
    }
    __syncthreads();
    run += s_pre;
    const uint64_t base = tb + (uint64_t)threadIdx.x * SCAN_ITEMS;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        const uint64_t i = base + k;
        if (i < n) op.emit(i, run, v[k]);
        run += v[k];
    }
}

// ----------------------------------------------------------- device state
struct Dev {
    // inputs: P programs concatenated node-wise, parents global
    const uint8_t *pt, *prt;
    const int32_t *par, *depth, *cidx;
    const uint32_t *ndin;
    uint64_t n;
    uint32_t maxd;           // largest max depth of the programs
    uint32_t P;
    const uint32_t *pn;      // node base of each program (P+1)
    uint32_t *pw;            // word base of each program after isn_gen (P+1)
    uint32_t *pf;            // function base of each program (P+1)
    // preprocess / isn_cnt
    uint8_t *ty8, *rt8;
    uint32_t *nd;
    uint32_t *loc;
    uint32_t *decl;          // func_decl node indices, F
    uint32_t *fids, *fstarts, *fsizes, *fends;
    uint32_t F;
    // isn_gen
    uint32_t *regs;          // 3n
    uint32_t *rank;          // n (sorted position of each node)
    uint32_t *rinv;          // n (node at each sorted position), with rank
    uint64_t *owner;         // W0
    uint32_t W0;
    uint32_t *w0_instr, *w0_rd, *w0_rs1, *w0_rs2, *w0_jt;
    // optimize
    uint8_t *used, *sstore, *usedi;
    // regalloc
    uint4 *w0_f4;            // per word: rd, rs1, rs2, instruction (one 16-byte load)
    uint16_t *sym;           // per virtual register: assigned register | swapped << 8
    uint32_t *xref;          // per virtual register named from outside its word's function: namer + 1
    uint32_t *cflag;         // per program: a register is named by two of its functions
    uint32_t *ra_off;
    uint32_t W1;
    uint32_t *w1_instr, *w1_rd, *w1_rs1, *w1_rs2, *w1_jt;
    // fix_jumps
    uint32_t *fj_off;
    uint32_t W2;
    uint32_t *words;
    uint32_t *flags;
    uint32_t *counter;
};

// ================================================================ preprocess
DEV void classify(uint32_t &ty, uint32_t rt, uint32_t ci, uint32_t &nd) {
    if ((ty & 0xF0u) != 0xF0u) return;
    if (rt == DT_FLOAT || rt == DT_FLOAT_REF) {
        bool comp = (int32_t)nd > 7;
        uint32_t reg = ci - 8u;
        bool stk = (int32_t)reg > 7 && comp;
        bool asint = !stk && comp;
        if (stk) ty |= 2u;
        if (asint) ty |= 1u;
        if (stk) reg -= 8u;
        if (comp) nd = reg;
    } else {
        uint32_t id = nd, fa = ci - id;
        uint32_t reg = id + (fa - ((int32_t)fa > 7 ? 8u : 0u));
        uint32_t osm = ((int32_t)reg > 7 ? 0xFFFFFFFFu : 0u) & id;  // ANDed with the value (D9)
        ty |= osm & 2u;
        reg -= 8u & osm;
        nd = reg;
    }
}

// Passes 1 and 2.  Pass 2's stores are gated by "the 8-node vector holds a
// call-argument list" but not masked by it (:236-262): 8-lane groups of the
// wave64 ballot are exactly the reference's 8-node vectors.
__global__ __launch_bounds__(256) void k_pre12(Dev d) {
    uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    uint32_t ty = 0, rt = 0, ci = 0, nd = 0;
    bool in = i < d.n;
    if (in) {
        ty = c_pareas[d.pt[i]]; rt = d.prt[i]; ci = (uint32_t)d.cidx[i]; nd = d.ndin[i];
        classify(ty, rt, ci, nd);
    }
    if (!in) return;
    // does the node's 8-node vector (counted from its program's first node)
    // hold a call argument list?
    const uint32_t prg = prog_of_w(d.pn, d.P, (uint32_t)i);
    const uint32_t v0 = d.pn[prg] + (((uint32_t)i - d.pn[prg]) & ~7u), v1 = min(v0 + 8u, d.pn[prg + 1]);
    bool block_has_list = false;
    for (uint32_t j = v0; j < v1; ++j) block_has_list |= c_pareas[d.pt[j]] == RV_CALL_ARGS;
    if (block_has_list && (uint32_t)i > d.pn[prg]) {
        uint64_t j = i - 1;
        uint32_t pty = c_pareas[d.pt[j]], prt = d.prt[j], pci = (uint32_t)d.cidx[j], pnd = d.ndin[j];
        classify(pty, prt, pci, pnd);
        if (pty == RV_CARG_STK) nd = pnd + 1u;
        else if (pty == RV_CARG && prt == DT_FLOAT) {
            int32_t ia = (int32_t)(pci - pnd) - 8;
            nd = (uint32_t)(ia > 0 ? ia : 0);
        }
    }
    d.ty8[i] = (uint8_t)ty;
    d.rt8[i] = (uint8_t)rt;
    d.nd[i] = nd;
}

// Pass 3 (:270-305): a FLOAT node marks its comparison-like parent FLOAT; the
// parent passes it on only if it sits in a later 8-node vector than the child
// (within one vector the scatter lands after the vector's load).
__global__ __launch_bounds__(256) void k_pre3(Dev d) {
    uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i >= d.n || d.prt[i] != DT_FLOAT) return;
    const uint32_t b = d.pn[prog_of_w(d.pn, d.P, (uint32_t)i)];  // vectors count from the program's first node
    uint64_t cur = i;
    for (;;) {
        int32_t p = d.par[cur];
        if (p < 0) break;
        if ((d.ty8[p] & 0xE0u) != 0xE0u) break;
        d.rt8[p] = DT_FLOAT;
        if (((cur - b) >> 3) < (((uint64_t)p - b) >> 3)) cur = (uint64_t)p;
        else break;
    }
}

// ================================================================== isn_cnt
struct CountOp {
    Dev d;
    DEV uint32_t value(uint64_t i) const {
        uint32_t ty = d.ty8[i], rt = d.rt8[i];
        uint32_t s = size_of(ty, rt);
        if (ty == RV_CALL_ARGS) {
            s += 1;
            if (i > 0 && (d.ty8[i - 1] & 0xF4u) == 0xF4u) s += (uint32_t)d.cidx[i - 1] + 1u;
        }
        int32_t p = d.par[i];
        uint32_t pty = d.ty8[p != -1 ? (uint32_t)p : d.pn[prog_of(d.pn, d.P, (uint32_t)i)]];  // root: the program's node 0
        int32_t ci = d.cidx[i];
        bool ife = pty == RV_IF_ELSE;
        if (ci == 0 && (ife || pty == RV_IF)) s += 1;
        if (ci == 1 && (ife || pty == RV_WHILE)) s += 1;
        return s;
    }
    DEV void emit(uint64_t i, uint32_t excl, uint32_t) const { d.loc[i] = excl; }
};

// The counts alone (the generator's node_sizes member), for the debug print.
__global__ void k_node_sizes(Dev d, uint32_t *out) {
    const uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i < d.n) out[i] = CountOp{d}.value(i);
}

// Location fix (:398-447).  A condition child moves its control-flow parent
// right after its own base words; call arguments then move to their slot in the
// argument list, which always wins over the first rule at one thread.
__global__ __launch_bounds__(256) void k_cnt_fix(Dev d) {
    uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i >= d.n) return;
    int32_t p = d.par[i];
    uint32_t ty = d.ty8[i];
    uint32_t ci = (uint32_t)d.cidx[i];
    if (p > -1) {
        uint32_t pty = d.ty8[p];
        bool cond = ((pty & 0xD0u) == 0xD0u) && ci == (pty & 1u);
        if (cond && (pty & 0xF4u) != 0xF4u) d.loc[p] = d.loc[i] + size_of(ty, d.rt8[i]);
    }
    if ((ty & 0xF4u) == 0xF4u && p > -1) d.loc[i] = d.loc[p] + ci + 1u;
}

struct DeclOp {
    Dev d;
    DEV uint32_t value(uint64_t i) const { return d.ty8[i] == RV_FDECL ? 1u : 0u; }
    DEV void emit(uint64_t i, uint32_t excl, uint32_t v) const { if (v) d.decl[excl] = (uint32_t)i; }
};

// Every program's first position starts an insertion-sort leaf.
__global__ void k_prog_bounds(uint8_t *bound, const uint32_t *pn, uint32_t P) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < P) bound[pn[p]] = 1;
}

// Batched input: parents arrive local to their program's node range.
__global__ void k_globalize_parents(int32_t *par, uint64_t n, const uint32_t *pn, uint32_t P) {
    const uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t p = par[i];
    if (p >= 0) par[i] = p + (int32_t)pn[prog_of_w(pn, P, (uint32_t)i)];
}

// Word base of every program: the exclusive count scan at its first node (the
// location fix does not move a program's first node); total at P.
__global__ void k_prog_words(Dev d, const uint32_t *total) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p > d.P) return;
    d.pw[p] = p < d.P ? d.loc[d.pn[p]] : *total;
}

// Function base of every program: functions (decl nodes, in node order) before
// its first node.
__global__ void k_prog_funcs(Dev d) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p > d.P) return;
    const uint32_t b = p < d.P ? d.pn[p] : (uint32_t)d.n;
    uint32_t lo = 0, hi = d.F;  // first decl index >= b
    while (lo < hi) { const uint32_t m = (lo + hi) >> 1; if (d.decl[m] < b) lo = m + 1; else hi = m; }
    d.pf[p] = lo;
}

// Function table (:451-515): ids = node data, starts = previous decl location + 6.
__global__ void k_functab(Dev d) {
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= d.F) return;
    uint32_t dn = d.decl[k];
    uint32_t off = 6u + d.loc[dn];
    const uint32_t prg = prog_of(d.pf, d.P, k);
    uint32_t prev = k > d.pf[prg] ? 6u + d.loc[d.decl[k - 1]] : d.pw[prg];
    d.fids[k] = d.nd[dn];
    d.fstarts[k] = prev;
    d.fsizes[k] = off - prev;
    d.fends[k] = off;
}

// =================================================================== isn_gen
struct Lane {
    bool has;
    int32_t pai;
    uint32_t dpv, loc, word, rd, rs1, rs2, jt, wb;
};

// The inputs of one node that all four of its lanes share, loaded once: its
// program (word base, function base and count), types, parent, parent type,
// child index, location and data.
struct GenNode {
    uint32_t nn, prg, wb, fb, Fp, wn, ty, rt, ci, loc, ndv, pty;
    int32_t par;
};

DEV GenNode gen_node(const Dev &d, uint32_t nn) {
    GenNode N;
    N.nn = nn;
    N.prg = prog_of_w(d.pn, d.P, nn);
    N.wb = d.pw[N.prg]; N.wn = d.pw[N.prg + 1] - N.wb;
    N.fb = d.pf[N.prg]; N.Fp = d.pf[N.prg + 1] - N.fb;
    N.ty = d.ty8[nn]; N.rt = d.rt8[nn];
    N.par = d.par[nn];
    N.pty = N.par != -1 ? d.ty8[N.par] : 0u;
    N.ci = (uint32_t)d.cidx[nn];
    N.loc = d.loc[nn];
    N.ndv = d.nd[nn];
    return N;
}

// One (node, offset) lane of the level loop body (:585-884), computed in the
// word indices of the node's own program (register numbers are ORed with word
// indices, :676-697, :842); L.loc is the global word slot.
DEV void gen_lane(const Dev &d, const GenNode &N, uint32_t o, bool full, Lane &L) {
    const uint32_t nn = N.nn, wb = N.wb, fb = N.fb, Fp = N.Fp, ty = N.ty, rt = N.rt;
    const uint32_t si = spec_i(ty, o, rt);
    const uint32_t b = c_bits[si];
    const bool has = b & 1u;
    const bool prop = has || o == 0;
    const int32_t par = prop ? N.par : 0;
    const bool nopar = par == -1;
    const bool lfp = !nopar && prop;
    const uint32_t pty = lfp ? N.pty : 0u;
    const uint32_t ci = lfp ? N.ci : 0u;
    const bool noncond = ((pty & 0xD0u) == 0xD0u) && !(ci == (pty & 1u));
    bool call = noncond;
    if ((nopar || !noncond) && prop) {
        uint32_t ct = (b >> 2) & 3u;
        if (ct == 1) call = true;
        if (ct == 2 && (int32_t)rt > 1) call = true;
    }
    L.has = has;
    L.pai = call ? (int32_t)((uint32_t)par * 3u + ci) : -1;
    const uint32_t iib = o + (prop ? N.loc - wb : 0u);
    uint32_t rd = 0;
    const uint32_t ndv = N.ndv;
    if (has) {
        uint32_t ct = (b >> 4) & 7u;
        if (ct == 3) rd |= 32u;
        if (ct == 4) rd |= 10u;
        if (ct == 1) rd |= ndv + 10u;
        if (ct == 2) rd |= ndv + 42u;
        bool hov = call || (ct == 0 && ((b >> 1) & 1u));
        if (hov) rd |= iib + 64u;
    }
    const bool use_rd = nopar || (!noncond && lfp);
    const uint32_t instr_no = iib + (noncond ? ((b >> 19) & 7u) : 0u);
    L.dpv = (use_rd ? rd : 0u) | (noncond ? instr_no : 0u);
    if (!full || !has) { L.rd = rd; return; }
    uint32_t iloc = iib;
    if (o == 1 && ty == RV_IF_ELSE) iloc = d.regs[(uint64_t)nn * 3 + 1];
    if (o == 1 && ty == RV_WHILE) iloc = d.regs[(uint64_t)nn * 3 + 2];
    const bool add2 = ty == RV_FDECL || (o > 1 && ty == RV_FDECL_DUMMY);
    const bool cal1 = o == 1 && ty == RV_CALL_ARGS;
    if (add2 || cal1) iloc += 2u;
    if (cal1 && nn > 0) {
        uint32_t ppt = d.ty8[nn - 1];
        if ((ppt & 0xF4u) == 0xF4u) iloc += (uint32_t)d.cidx[nn - 1];
    }
    uint32_t word = c_word[si];
    const uint32_t cc = (b >> 13) & 7u;
    const uint32_t nd4 = ndv * 4u;
    if (cc == 1) word |= ndv & 0xFFFFF000u;
    if (cc == 2) word |= (ndv & 0xFFFu) << 20;
    if (cc == 3) word |= (0u - (nd4 + 8u)) << 20;
    if (cc == 4) word |= nd4 << 20;
    if (cc == 5) word |= (0u - nd4) << 20;
    const uint32_t jc = (b >> 16) & 7u;
    uint32_t jt = 0;
    if (jc >= 1 && jc <= 5) jt = d.regs[(uint64_t)nn * 3 + 1 + (jc > 2 ? 1 : 0)] + ((jc & 1u) == 0 ? 1u : 0u);
    if (jc == 6) { if (ndv < Fp) jt |= d.fends[fb + ndv] - wb - 6u; else atomicOr(d.flags, EF_OOB); }
    if (jc == 7) { if (ndv < Fp) jt |= d.fstarts[fb + ndv] - wb; else atomicOr(d.flags, EF_OOB); }
    uint32_t rs[2];
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        uint32_t oc = (b >> (7 + 3 * s)) & 7u;
        uint32_t arg = 0;
        if (oc == 1 || oc == 4 || oc == 5 || oc == 6) {
            uint64_t ri = (uint64_t)nn * 3;
            if (oc == 1) ri += s;
            if (oc == 4 || oc == 5) ri += 1;
            if (oc == 5) ri -= s;
            arg = d.regs[ri];
        }
        if (oc == 2) arg |= ndv + 10u;
        if (oc == 3) arg |= ndv + 42u;
        if (oc == 7) arg |= iib + 63u;
        rs[s] = arg;
    }
    L.loc = iloc + wb; L.word = word; L.rd = rd; L.rs1 = rs[0]; L.rs2 = rs[1]; L.jt = jt;
    L.wb = wb;
}

// Parent-slot propagation: the last propagating lane of a node wins the slot.
__global__ __launch_bounds__(256) void k_gen_regs(Dev d) {
    uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i >= d.n) return;
    int32_t slot = -1; uint32_t val = 0;
    const GenNode N = gen_node(d, (uint32_t)i);
    for (uint32_t o = 0; o < 4; ++o) {
        Lane L;
        gen_lane(d, N, o, false, L);
        if (L.pai >= 0) { slot = L.pai; val = L.dpv; }
    }
    if (slot >= 0) {
        if ((uint64_t)slot < d.n * 3) d.regs[slot] = val;
        else atomicOr(d.flags, EF_OOB);
    }
}

DEV uint64_t owner_key(const Dev &d, uint32_t nn, uint32_t o, bool use_rank) {
    uint64_t level = (uint64_t)(d.maxd - (uint32_t)d.depth[nn]) + 1u;
    uint64_t r = use_rank ? d.rank[nn] : nn;
    return (level << 34) | (r << 2) | o;
}

__global__ __launch_bounds__(256) void k_gen_owner(Dev d, int use_rank) {
    uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i >= d.n) return;
    const GenNode N = gen_node(d, (uint32_t)i);
    const uint64_t k0 = owner_key(d, (uint32_t)i, 0, use_rank != 0);
    for (uint32_t o = 0; o < 4; ++o) {
        Lane L;
        gen_lane(d, N, o, true, L);
        if (!L.has) continue;
        if (L.loc >= d.W0) { atomicOr(d.flags, EF_OOB); continue; }
        atomicMax((unsigned long long *)&d.owner[L.loc], (unsigned long long)(k0 | o));
    }
}

// Counts slots won by one node but also written by another node of equal depth.
__global__ __launch_bounds__(256) void k_gen_ties(Dev d) {
    uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i >= d.n) return;
    const GenNode N = gen_node(d, (uint32_t)i);
    const uint64_t k0 = owner_key(d, (uint32_t)i, 0, false);
    for (uint32_t o = 0; o < 4; ++o) {
        Lane L;
        gen_lane(d, N, o, true, L);
        if (!L.has || L.loc >= d.W0) continue;
        const uint64_t k = k0 | o;
        uint64_t w = d.owner[L.loc];
        if ((w >> 34) == (k >> 34) && ((w >> 2) & 0xFFFFFFFFull) != i) atomicAdd(d.counter, 1u);
    }
}

// The words of isn_gen, one thread per word slot (W0 + 1 of them, the last a
// zero pad): the slot's owner key names its writer, a node (through rinv when the
// keys hold sorted positions) and one of its lanes, whose word is computed
// again; a slot nobody writes is zero.  The five arrays are written in slot
// order, each slot once, so they need no clearing beforehand.
__global__ __launch_bounds__(256) void k_gen_write(Dev d, int use_rank) {
    const uint64_t w = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (w > d.W0) return;
    const uint64_t key = w < d.W0 ? d.owner[w] : 0ull;  // keys are >= 1 << 34
    uint32_t instr = 0, rd = 0, rs1 = 0, rs2 = 0, jt = 0;
    if (key != 0ull) {
        const uint32_t o = (uint32_t)key & 3u, r = (uint32_t)(key >> 2);
        const uint32_t nn = use_rank ? d.rinv[r] : r;
        const GenNode N = gen_node(d, nn);
        Lane L;
        gen_lane(d, N, o, true, L);
        if (!L.has || L.loc != w) atomicOr(d.flags, EF_OOB);
        const uint32_t wn = N.wn;
        if (d.P > 1 && (((int32_t)L.rd > 63 && L.rd - 64u >= wn) || ((int32_t)L.rs1 > 63 && L.rs1 - 64u >= wn) ||
                        ((int32_t)L.rs2 > 63 && L.rs2 - 64u >= wn)))
            atomicOr(d.flags, EF_OOB);
        instr = L.word;
        rd = (int32_t)L.rd > 63 ? L.rd + L.wb : L.rd;
        rs1 = (int32_t)L.rs1 > 63 ? L.rs1 + L.wb : L.rs1;
        rs2 = (int32_t)L.rs2 > 63 ? L.rs2 + L.wb : L.rs2;
        jt = L.jt + L.wb;
    }
    d.w0_instr[w] = instr; d.w0_rd[w] = rd; d.w0_rs1[w] = rs1; d.w0_rs2[w] = rs2; d.w0_jt[w] = jt;
}

// ------------------------------------------ libstdc++ introsort, reproduced
// std::sort over node ids with key depth (rv_generator_avx.cpp:539-541) is
// introsort: depth limit 2*floor(log2 n), median of (first+1, mid, last-1)
// moved to first, unguarded Hoare partition around it, ranges <= 16 left for a
// final insertion sort, heapsort where the limit runs out.  The k-th left stop
// (key >= pivot, scanning up from first+1) swaps with the k-th right stop
// (key <= pivot, scanning down from last-1) while they have not crossed, so one
// partition is a stream compaction of the stops and a search; all ranges of one
// recursion depth are partitioned together.  Each level works on the list of
// its ranges only (their positions concatenated: the level's active list,
// offsets in Seg::off, ranges in position order); ranges that are done leave it.
struct Seg { uint32_t first, last, limit, off; };

__global__ void k_sort_init(uint32_t *sidx, int32_t *skey, const int32_t *depth, uint64_t n) {
    uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    sidx[i] = (uint32_t)i; skey[i] = depth[i];
}

DEV void swap_pos(uint32_t *sidx, int32_t *skey, uint32_t a, uint32_t b) {
    uint32_t ti = sidx[a]; sidx[a] = sidx[b]; sidx[b] = ti;
    int32_t tk = skey[a]; skey[a] = skey[b]; skey[b] = tk;
}

// std::__move_median_to_first(first, first+1, mid, last-1); returns the pivot key.
DEV int32_t median_to_first(uint32_t first, uint32_t last, uint32_t *sidx, int32_t *skey) {
    uint32_t a = first + 1, b = first + (last - first) / 2, c = last - 1;
    int32_t ka = skey[a], kb = skey[b], kc = skey[c];
    uint32_t pick;
    if (ka < kb) { if (kb < kc) pick = b; else if (ka < kc) pick = c; else pick = a; }
    else if (ka < kc) pick = a;
    else if (kb < kc) pick = c;
    else pick = b;
    swap_pos(sidx, skey, first, pick);
    return skey[first];
}

// The first level: one range per program of more than 16 nodes.
__global__ void k_sort_root(const Seg *segs, uint32_t m, uint32_t *sidx, int32_t *skey, int32_t *pivot) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < m) pivot[s] = median_to_first(segs[s].first, segs[s].last, sidx, skey);
}

// Last range whose active offset is <= a.
DEV uint32_t seg_find(const Seg *segs, uint32_t m, uint32_t a) {
    uint32_t lo = 0, hi = m;
    while (hi - lo > 1) { const uint32_t mid = (lo + hi) >> 1; if (segs[mid].off <= a) lo = mid; else hi = mid; }
    return lo;
}

// Walks SCAN_ITEMS consecutive active indices from a0, tracking their range.
// Last range whose active offset is <= a, among ranges [lo, hi) (the answer
// is known to lie there).
DEV uint32_t seg_find_in(const Seg *segs, uint32_t lo, uint32_t hi, uint32_t a) {
    while (hi - lo > 1) { const uint32_t mid = (lo + hi) >> 1; if (segs[mid].off <= a) lo = mid; else hi = mid; }
    return lo;
}

// The ranges holding a workgroup's items [a_lo, a_hi]: one search per
// workgroup, so each thread then searches only this window.
// Called by every thread of the workgroup; items [a_lo, a_lo + span) clipped to A.
struct SegWindow { uint32_t lo, hi; };
DEV SegWindow seg_window(const Seg *segs, uint32_t m, uint32_t A, uint32_t a_lo, uint32_t span) {
    __shared__ uint32_t w[2];
    if (a_lo < A) {
        const uint32_t a_hi = min(A, a_lo + span) - 1;
        if (threadIdx.x == 0) w[0] = seg_find(segs, m, a_lo);
        if (threadIdx.x == 64) w[1] = seg_find(segs, m, a_hi);
    }
    __syncthreads();
    const SegWindow r{w[0], w[1] + 1};
    __syncthreads();
    return r;
}

struct SegWalk {
    const Seg *segs; uint32_t m, s, nxt; Seg g;
    DEV SegWalk(const Seg *sg, uint32_t mm, uint32_t a0) : segs(sg), m(mm) {
        s = seg_find(sg, mm, a0); g = sg[s]; nxt = s + 1 < mm ? sg[s + 1].off : 0xFFFFFFFFu;
    }
    DEV SegWalk(const Seg *sg, uint32_t mm, SegWindow wd, uint32_t a0) : segs(sg), m(mm) {
        s = seg_find_in(sg, wd.lo, wd.hi, a0); g = sg[s]; nxt = s + 1 < mm ? sg[s + 1].off : 0xFFFFFFFFu;
    }
    DEV void to(uint32_t a) {
        while (a >= nxt) { ++s; g = segs[s]; nxt = s + 1 < m ? segs[s + 1].off : 0xFFFFFFFFu; }
    }
    DEV uint32_t pos(uint32_t a) const { return g.first + (a - g.off); }
};

// Left/right stop flags of the element at pos of range g (pivot key p): the
// pivot itself, at first, takes part in neither.
DEV uint32_t stop_bits(const Seg &g, int32_t p, const int32_t *skey, uint32_t i) {
    if (i == g.first) return 0;
    const int32_t k = skey[i];
    return (!(k < p) ? 1u : 0u) | (!(p < k) ? 2u : 0u);
}

// Per-tile counts of left and right stops over the active list.
// The level kernels read the level's counts from the device (lv[0] active
// elements, lv[1] ranges) and are launched on grids sized for an upper bound
// of them; workgroups past the counts exit.
__global__ __launch_bounds__(SCAN_BLOCK) void k_sortlvl_reduce(const Seg *segs, const uint32_t *lv,
                                                               const int32_t *pivot, const int32_t *skey,
                                                               uint2 *tile_sums) {
    __shared__ uint32_t sL[SCAN_BLOCK / 64], sR[SCAN_BLOCK / 64];
    const uint32_t A = lv[0], m = lv[1];
    if (blockIdx.x * (uint32_t)SCAN_TILE >= A) return;
    const SegWindow wd = seg_window(segs, m, A, blockIdx.x * (uint32_t)SCAN_TILE, (uint32_t)SCAN_TILE);
    const uint32_t a0 = blockIdx.x * (uint32_t)SCAN_TILE + threadIdx.x * SCAN_ITEMS;
    uint32_t cl = 0, cr = 0;
    if (a0 < A) {
        SegWalk w(segs, m, wd, a0);
        int32_t p = pivot[w.s];
        for (int k = 0; k < SCAN_ITEMS; ++k) {
            const uint32_t a = a0 + k;
            if (a >= A) break;
            const uint32_t s0 = w.s;
            w.to(a);
            if (w.s != s0) p = pivot[w.s];
            const uint32_t b = stop_bits(w.g, p, skey, w.pos(a));
            cl += b & 1u; cr += b >> 1;
        }
    }
    cl = wave_sum(cl); cr = wave_sum(cr);
    if ((threadIdx.x & 63) == 0) { sL[threadIdx.x >> 6] = cl; sR[threadIdx.x >> 6] = cr; }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t tl = 0, tr = 0;
#pragma unroll
        for (int w = 0; w < SCAN_BLOCK / 64; ++w) { tl += sL[w]; tr += sR[w]; }
        tile_sums[blockIdx.x] = make_uint2(tl, tr);
    }
}

// One workgroup: exclusive scan of (left, right) tile sums in place; totals to *total.
__global__ __launch_bounds__(1024) void k_scan_tiles2(uint2 *sums, const uint32_t *lv, uint2 *total) {
    __shared__ uint32_t wl[16], wr[16];
    __shared__ uint32_t cl, cr;
    const uint32_t ntiles = (uint32_t)((lv[0] + SCAN_TILE - 1) / SCAN_TILE);
    const uint32_t lane = threadIdx.x & 63u, wid = threadIdx.x >> 6;
    if (threadIdx.x == 0) { cl = 0; cr = 0; }
    __syncthreads();
    for (uint32_t b = 0; b < ntiles; b += 1024) {
        const uint32_t i = b + threadIdx.x;
        const uint2 v = i < ntiles ? sums[i] : make_uint2(0, 0);
        const uint32_t xl = wave_incl_scan(v.x, lane), xr = wave_incl_scan(v.y, lane);
        if (lane == 63) { wl[wid] = xl; wr[wid] = xr; }
        __syncthreads();
        uint32_t pl = cl, pr = cr, al = 0, ar = 0;
        for (uint32_t w = 0; w < 16; ++w) {
            if (w < wid) { pl += wl[w]; pr += wr[w]; }
            al += wl[w]; ar += wr[w];
        }
        if (i < ntiles) sums[i] = make_uint2(pl + xl - v.x, pr + xr - v.y);
        __syncthreads();
        if (threadIdx.x == 0) { cl += al; cr += ar; }
        __syncthreads();
    }
    if (threadIdx.x == 0) *total = make_uint2(cl, cr);
}

// Compaction of the stops: Lpos[k] / Rpos[k] = position of the k-th left / right
// stop of the level; base[s] = counts before range s.
__global__ __launch_bounds__(SCAN_BLOCK) void k_sortlvl_down(const Seg *segs, const uint32_t *lv,
                                                             const int32_t *pivot, const int32_t *skey,
                                                             const uint2 *tile_pre, uint32_t *Lpos, uint32_t *Rpos,
                                                             uint2 *base) {
    __shared__ uint32_t sL[SCAN_BLOCK / 64], sR[SCAN_BLOCK / 64];
    const uint32_t A = lv[0], m = lv[1];
    if (blockIdx.x * (uint32_t)SCAN_TILE >= A) return;
    // the tile's stops, staged: its left and right stops are each one
    // contiguous range of Lpos / Rpos, written out coalesced at the end
    __shared__ uint32_t stL[SCAN_TILE], stR[SCAN_TILE];
    const SegWindow wd = seg_window(segs, m, A, blockIdx.x * (uint32_t)SCAN_TILE, (uint32_t)SCAN_TILE);
    const uint32_t a0 = blockIdx.x * (uint32_t)SCAN_TILE + threadIdx.x * SCAN_ITEMS;
    uint32_t bits = 0, cl = 0, cr = 0;  // 2 bits per item
    if (a0 < A) {
        SegWalk w(segs, m, wd, a0);
        int32_t p = pivot[w.s];
        for (int k = 0; k < SCAN_ITEMS; ++k) {
            const uint32_t a = a0 + k;
            if (a >= A) break;
            const uint32_t s0 = w.s;
            w.to(a);
            if (w.s != s0) p = pivot[w.s];
            const uint32_t b = stop_bits(w.g, p, skey, w.pos(a));
            bits |= b << (2 * k);
            cl += b & 1u; cr += b >> 1;
        }
    }
    const uint32_t lane = threadIdx.x & 63u, wid = threadIdx.x >> 6;
    const uint32_t xl = wave_incl_scan(cl, lane), xr = wave_incl_scan(cr, lane);
    if (lane == 63) { sL[wid] = xl; sR[wid] = xr; }
    __syncthreads();
    const uint32_t tL = tile_pre[blockIdx.x].x, tR = tile_pre[blockIdx.x].y;
    uint32_t rl = xl - cl, rr = xr - cr;  // offsets within the tile's ranges
#pragma unroll
    for (int w = 0; w < SCAN_BLOCK / 64; ++w) if ((uint32_t)w < wid) { rl += sL[w]; rr += sR[w]; }
    if (a0 < A) {
        SegWalk w(segs, m, wd, a0);
        for (int k = 0; k < SCAN_ITEMS; ++k) {
            const uint32_t a = a0 + k;
            if (a >= A) break;
            w.to(a);
            const uint32_t i = w.pos(a);
            if (a == w.g.off) base[w.s] = make_uint2(tL + rl, tR + rr);
            const uint32_t b = (bits >> (2 * k)) & 3u;
            if (b & 1u) stL[rl++] = i;
            if (b & 2u) stR[rr++] = i;
        }
    }
    uint32_t nL = 0, nR = 0;
#pragma unroll
    for (int w = 0; w < SCAN_BLOCK / 64; ++w) { nL += sL[w]; nR += sR[w]; }
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < nL; k += SCAN_BLOCK) Lpos[tL + k] = stL[k];
    for (uint32_t k = threadIdx.x; k < nR; k += SCAN_BLOCK) Rpos[tR + k] = stR[k];
}

struct SegCut { uint32_t baseL, cntL, baseR, cntR, m, cut, uni, pad; };

// Ranges whose keys are all equal.  For them the introsort's moves do not
// depend on the keys: median-of-three picks the middle element (all three
// comparisons false: iter_swap(first, first + len/2)), the unguarded partition
// stops at every element and swaps first+1+k with last-1-k for k < (len-1)/2,
// and the cut falls at first+1+(len-1)/2; the final insertion sort moves none of
// them.  So each element's final place follows from its place alone.
// uniform_depth: partition levels until every subrange is <= 16 (the left
// child, 1+(len-1)/2, is never the smaller one).
DEV uint32_t uniform_depth(uint32_t len) {
    uint32_t dpt = 0;
    while (len > 16) { len = 1u + (len - 1u) / 2u; ++dpt; }
    return dpt;
}

// Final place of the element at pos of an all-equal range [a, b), whose first
// median swap has been made already.
DEV uint32_t uniform_final(uint32_t a, uint32_t b, uint32_t pos) {
    bool first = true;
    while (b - a > 16) {
        const uint32_t len = b - a;
        if (!first) { const uint32_t mid = a + len / 2; if (pos == a) pos = mid; else if (pos == mid) pos = a; }
        first = false;
        const uint32_t K = (len - 1) / 2;
        if ((pos >= a + 1 && pos < a + 1 + K) || (pos >= b - K && pos < b)) pos = a + b - pos;
        const uint32_t cut = a + 1 + K;
        if (pos >= cut) a = cut; else b = cut;
    }
    return pos;
}

__global__ void k_sort_search(const Seg *segs, const uint32_t *lv, const uint2 *base, const uint2 *total,
                              const uint32_t *Lpos, const uint32_t *Rpos, SegCut *cuts) {
    const uint32_t m = lv[1];
    uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= m) return;
    uint32_t first = segs[s].first, last = segs[s].last;
    const uint2 b0 = base[s], b1 = s + 1 < m ? base[s + 1] : *total;
    uint32_t bL = b0.x, bR = b0.y, cL = b1.x - b0.x, cR = b1.y - b0.y;
    // m = largest k with L_k < R_k (1-based), L_k = Lpos[bL+k-1], R_k = Rpos[bR+cR-k]
    uint32_t lo = 0, hi = cL < cR ? cL : cR;
    while (lo < hi) {
        uint32_t mid = (lo + hi + 1) / 2;
        if (Lpos[bL + mid - 1] < Rpos[bR + cR - mid]) lo = mid; else hi = mid - 1;
    }
    uint32_t mm = lo;
    uint32_t cut = 0xFFFFFFFFu;
    if (mm + 1 <= cL) cut = Lpos[bL + mm];
    if (mm >= 1) { uint32_t r = Rpos[bR + cR - mm]; if (r < cut) cut = r; }
    if (cut <= first || cut >= last) cut = first + 1;  // unreachable after a median-of-three pivot
    // every element stopped on both sides: all keys equal the pivot
    const uint32_t len = last - first;
    if (cL == len - 1 && cR == len - 1 && uniform_depth(len) <= segs[s].limit) {
        cuts[s] = SegCut{bL, cL, bR, cR, 0, last, 1, 0};
        return;
    }
    cuts[s] = SegCut{bL, cL, bR, cR, mm, cut, 0, 0};
}

// The partition's swaps: pair k of a range (1-based) swaps its k-th left and
// k-th-from-the-end right stop; the index within the range is the pair number.
// A workgroup covers 256*SCAN_ITEMS active indices, consecutive lanes taking
// consecutive pairs (coalesced stop lists).
__global__ __launch_bounds__(256) void k_sort_swap(const Seg *segs, const uint32_t *lv, const SegCut *cuts,
                                                   const uint32_t *Lpos, const uint32_t *Rpos, uint32_t *sidx,
                                                   int32_t *skey) {
    const uint32_t A = lv[0], m = lv[1];
    if (blockIdx.x * 256u * SCAN_ITEMS >= A) return;
    const SegWindow wd = seg_window(segs, m, A, blockIdx.x * 256u * SCAN_ITEMS, 256u * SCAN_ITEMS);
    const uint32_t a0 = blockIdx.x * 256u * SCAN_ITEMS + threadIdx.x;
    if (a0 >= A) return;
    SegWalk w(segs, m, wd, a0);
    SegCut c = cuts[w.s];
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        const uint32_t a = a0 + 256u * k;
        if (a >= A) break;
        const uint32_t s0 = w.s;
        w.to(a);
        if (w.s != s0) c = cuts[w.s];
        const uint32_t kk = a - w.g.off + 1;
        if (kk <= c.m) swap_pos(sidx, skey, Lpos[c.baseL + kk - 1], Rpos[c.baseR + c.cntR - kk]);
    }
}

// Level counters.  qa[l & 1] = ranges << 32 | active elements of level l
// (each level's plan resets and fills the next one's); ue = all-equal ranges
// << 32 | their elements (accumulated); one 64-bit atomic per pair keeps a
// list's slots and offsets in the same order.
// nloc: ranges listed for k_sort_local; bad: a listed range of a size it does
// not take (never, by construction; the host reports it as an internal error).
struct SortCnt { unsigned long long qa[2], ue; uint32_t nheap, nloc, bad, pad; };

// Ranges of at most SORT_LOCAL elements finish in one workgroup's LDS
// (k_sort_local): k_sort_plan lists them (median already moved to first)
// instead of the next level.
constexpr uint32_t SORT_LOCAL = 2048;

DEV uint32_t block_excl_scan(uint32_t v, uint32_t *tmp /* 16 */, uint32_t &total) {
    const uint32_t lane = threadIdx.x & 63u, wid = threadIdx.x >> 6;
    const uint32_t x = wave_incl_scan(v, lane);
    if (lane == 63) tmp[wid] = x;
    __syncthreads();
    uint32_t pre = 0, all = 0;
    for (uint32_t w = 0; w < blockDim.x / 64; ++w) { if (w < wid) pre += tmp[w]; all += tmp[w]; }
    __syncthreads();
    total = all;
    return pre + x - v;
}

// The children of every partitioned range (1024 ranges per workgroup): next
// level (with their active offsets), heapsort list, or leaves (marked in
// `bound`); all-equal ranges to the closed-form list (offsets in its own
// element order); then median-of-three for every new range of the next level.
// Workgroups take their slots of each list with one atomic, so the lists are
// in workgroup-arrival order; every consumer needs only that a list's slots and
// offsets agree.
__global__ __launch_bounds__(1024) void k_sort_plan(const Seg *segs, const SegCut *cuts, const uint32_t *lv,
                                                    Seg *next, int32_t *next_pivot, Seg *heap, Seg *uni,
                                                    Seg *loc, int32_t *loc_pivot,
                                                    unsigned long long *qa_next, SortCnt *cnt, uint8_t *bound,
                                                    uint32_t *sidx, int32_t *skey) {
    __shared__ uint32_t tmp[16];
    __shared__ uint32_t base[6];
    const uint32_t m = lv[1];
    if (blockIdx.x * blockDim.x >= m) return;
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    Seg g{0, 0, 0, 0};
    SegCut ct{};
    if (s < m) { g = segs[s]; ct = cuts[s]; }
    const bool v = s < m, un = v && ct.uni;
    const uint32_t lim = g.limit - 1;
    const uint32_t c0 = ct.cut - g.first, c1 = g.last - ct.cut;
    const bool r0 = v && !un && c0 > 16 && lim > 0, r1 = v && !un && c1 > 16 && lim > 0;
    const bool n0 = r0 && c0 > SORT_LOCAL, n1 = r1 && c1 > SORT_LOCAL;    // next level
    const bool l0 = r0 && c0 <= SORT_LOCAL, l1 = r1 && c1 <= SORT_LOCAL;  // finished in LDS
    const bool h0b = v && !un && c0 > 16 && lim == 0, h1b = v && !un && c1 > 16 && lim == 0;
    uint32_t tq, ta, th, tu, te, tl;
    const uint32_t xl = block_excl_scan((l0 ? 1u : 0u) + (l1 ? 1u : 0u), tmp, tl);
    const uint32_t xq = block_excl_scan((n0 ? 1u : 0u) + (n1 ? 1u : 0u), tmp, tq);
    const uint32_t xa = block_excl_scan((n0 ? c0 : 0u) + (n1 ? c1 : 0u), tmp, ta);
    const uint32_t xh = block_excl_scan((h0b ? 1u : 0u) + (h1b ? 1u : 0u), tmp, th);
    const uint32_t xu = block_excl_scan(un ? 1u : 0u, tmp, tu);
    const uint32_t xe = block_excl_scan(un ? g.last - g.first : 0u, tmp, te);
    if (threadIdx.x == 0) {
        const unsigned long long qa = (tq | ta) ? atomicAdd(qa_next, ((unsigned long long)tq << 32) | ta) : 0ull;
        const unsigned long long ue = (tu | te) ? atomicAdd(&cnt->ue, ((unsigned long long)tu << 32) | te) : 0ull;
        base[0] = (uint32_t)(qa >> 32); base[1] = (uint32_t)qa;
        base[2] = th ? atomicAdd(&cnt->nheap, th) : 0u;
        base[3] = (uint32_t)(ue >> 32); base[4] = (uint32_t)ue;
        base[5] = tl ? atomicAdd(&cnt->nloc, tl) : 0u;
    }
    __syncthreads();
    if (!v) return;
    if (un) {
        uni[base[3] + xu] = Seg{g.first, g.last, g.limit, base[4] + xe};
        bound[g.first] = 2;
        return;
    }
    bound[ct.cut] = 1;
    uint32_t q = base[0] + xq, a = base[1] + xa, h = base[2] + xh;
    if (n0) { next[q] = Seg{g.first, ct.cut, lim, a}; next_pivot[q] = median_to_first(g.first, ct.cut, sidx, skey); ++q; a += c0; }
    if (n1) { next[q] = Seg{ct.cut, g.last, lim, a}; next_pivot[q] = median_to_first(ct.cut, g.last, sidx, skey); }
    uint32_t lq = base[5] + xl;
    if (l0) { loc[lq] = Seg{g.first, ct.cut, lim, 0}; loc_pivot[lq] = median_to_first(g.first, ct.cut, sidx, skey); ++lq; }
    if (l1) { loc[lq] = Seg{ct.cut, g.last, lim, 0}; loc_pivot[lq] = median_to_first(ct.cut, g.last, sidx, skey); }
    if (h0b) heap[h++] = Seg{g.first, ct.cut, 0, 0};
    if (h1b) heap[h++] = Seg{ct.cut, g.last, 0, 0};
}

// One listed range per workgroup, all its remaining levels in LDS: the same
// partitions as the level kernels above (stops, pairs, cut, all-equal ranges
// in closed form when their recursion fits the limit), the same outputs
// (boundaries in `bound`, entries of the heap and all-equal lists), its
// subranges kept in position order.  Thread t holds positions [8t, 8t + 8).
constexpr uint32_t SORT_LOCAL_SUBS = SORT_LOCAL / 17 + 2;
__global__ __launch_bounds__(256) void k_sort_local(const Seg *loc, const int32_t *loc_pivot, const SortCnt *cntr,
                                                    uint32_t *sidx, int32_t *skey, uint8_t *bound, Seg *heap,
                                                    Seg *uni, SortCnt *cnt) {
    static_assert(SORT_LOCAL == 256 * 8, "8 positions per thread");
    __shared__ int32_t K[SORT_LOCAL];
    __shared__ uint32_t X[SORT_LOCAL];
    __shared__ uint16_t LP[SORT_LOCAL], RP[SORT_LOCAL];
    __shared__ uint16_t sf[2][SORT_LOCAL_SUBS], sl[2][SORT_LOCAL_SUBS];
    __shared__ uint8_t slim[2][SORT_LOCAL_SUBS];
    __shared__ int32_t spiv[2][SORT_LOCAL_SUBS];
    __shared__ uint16_t bL[SORT_LOCAL_SUBS], bR[SORT_LOCAL_SUBS], cLs[SORT_LOCAL_SUBS], cRs[SORT_LOCAL_SUBS];
    __shared__ uint16_t mms[SORT_LOCAL_SUBS], cts[SORT_LOCAL_SUBS];
    __shared__ uint8_t unis[SORT_LOCAL_SUBS];
    __shared__ uint32_t tmp[16];
    if (blockIdx.x >= cntr->nloc) return;
    const uint32_t tid = threadIdx.x;
    const Seg g = loc[blockIdx.x];
    const uint32_t base = g.first, len = g.last - g.first;
    if (len > SORT_LOCAL || len < 17) { if (tid == 0) atomicOr(&cnt->bad, 1u); return; }
    for (uint32_t p = tid; p < len; p += 256) { K[p] = skey[base + p]; X[p] = sidx[base + p]; }
    if (tid == 0) { sf[0][0] = 0; sl[0][0] = (uint16_t)len; slim[0][0] = (uint8_t)g.limit; spiv[0][0] = loc_pivot[blockIdx.x]; }
    uint32_t cur = 0, nsub = 1;
    __syncthreads();
    const uint32_t p0 = tid * 8u;
    while (nsub > 0) {
        // stops of this thread's positions: subrange walk from the last one starting at or before p0
        uint32_t s0 = 0;
        {
            uint32_t lo = 0, hi = nsub;
            while (hi - lo > 1) { const uint32_t mid = (lo + hi) >> 1; if (sf[cur][mid] <= p0) lo = mid; else hi = mid; }
            s0 = lo;
        }
        uint32_t bits = 0, cl = 0, cr = 0;
        {
            uint32_t s = s0;
            for (uint32_t k = 0; k < 8; ++k) {
                const uint32_t p = p0 + k;
                if (p >= len) break;
                while (s + 1 < nsub && sf[cur][s + 1] <= p) ++s;
                if (p < sf[cur][s] || p >= sl[cur][s] || p == sf[cur][s]) continue;
                const int32_t kk = K[p], pv = spiv[cur][s];
                const uint32_t b = (!(kk < pv) ? 1u : 0u) | (!(pv < kk) ? 2u : 0u);
                bits |= b << (2 * k);
                cl += b & 1u; cr += b >> 1;
            }
        }
        // both counts in one scan (each at most SORT_LOCAL, so 16 bits apiece)
        uint32_t tot2;
        const uint32_t x2 = block_excl_scan(cl | (cr << 16), tmp, tot2);
        const uint32_t totL = tot2 & 0xFFFFu, totR = tot2 >> 16;
        uint32_t rl = x2 & 0xFFFFu, rr = x2 >> 16;
        const uint32_t rl0 = rl;
        {
            uint32_t s = s0;
            for (uint32_t k = 0; k < 8; ++k) {
                const uint32_t p = p0 + k;
                if (p >= len) break;
                while (s + 1 < nsub && sf[cur][s + 1] <= p) ++s;
                if (p == sf[cur][s]) { bL[s] = (uint16_t)rl; bR[s] = (uint16_t)rr; }
                const uint32_t b = (bits >> (2 * k)) & 3u;
                if (b & 1u) LP[rl++] = (uint16_t)p;
                if (b & 2u) RP[rr++] = (uint16_t)p;
            }
        }
        __syncthreads();
        // per subrange: stop counts, pairs, cut (k_sort_search)
        if (tid < nsub) {
            const uint32_t s = tid, f = sf[cur][s], l = sl[cur][s];
            const uint32_t eL = s + 1 < nsub ? bL[s + 1] : totL, eR = s + 1 < nsub ? bR[s + 1] : totR;
            const uint32_t b0 = bL[s], b1 = bR[s], cL = eL - b0, cR = eR - b1;
            uint32_t lo = 0, hi = cL < cR ? cL : cR;
            while (lo < hi) {
                const uint32_t mid = (lo + hi + 1) / 2;
                if (LP[b0 + mid - 1] < RP[b1 + cR - mid]) lo = mid; else hi = mid - 1;
            }
            uint32_t cut = 0xFFFFFFFFu;
            if (lo + 1 <= cL) cut = LP[b0 + lo];
            if (lo >= 1) { const uint32_t r = RP[b1 + cR - lo]; if (r < cut) cut = r; }
            if (cut <= f || cut >= l) cut = f + 1;
            const uint32_t n = l - f;
            const bool un = cL == n - 1 && cR == n - 1 && uniform_depth(n) <= slim[cur][s];
            cLs[s] = (uint16_t)cL; cRs[s] = (uint16_t)cR;
            mms[s] = (uint16_t)(un ? 0u : lo); cts[s] = (uint16_t)(un ? l : cut); unis[s] = un ? 1 : 0;
        }
        __syncthreads();
        // the pairs' swaps, each by the thread holding its left stop
        {
            uint32_t s = s0, r = rl0;
            for (uint32_t k = 0; k < 8; ++k) {
                const uint32_t p = p0 + k;
                if (p >= len) break;
                while (s + 1 < nsub && sf[cur][s + 1] <= p) ++s;
                if (!((bits >> (2 * k)) & 1u)) continue;
                const uint32_t kk = r++ - bL[s];
                if (kk < mms[s]) {
                    const uint32_t q = RP[bR[s] + cRs[s] - 1u - kk];
                    const int32_t tk = K[p]; K[p] = K[q]; K[q] = tk;
                    const uint32_t tx = X[p]; X[p] = X[q]; X[q] = tx;
                }
            }
        }
        __syncthreads();
        // children (k_sort_plan): the next subranges in position order
        const uint32_t nx = cur ^ 1u;
        uint32_t nc = 0, f = 0, l = 0, cut = 0, lim = 0;
        bool c0n = false, c1n = false;
        if (tid < nsub) {
            f = sf[cur][tid]; l = sl[cur][tid]; cut = cts[tid];
            const uint32_t limit = slim[cur][tid];
            lim = limit - 1u;
            if (unis[tid]) {
                const unsigned long long ue = atomicAdd(&cnt->ue, (1ull << 32) | (unsigned long long)(l - f));
                uni[(uint32_t)(ue >> 32)] = Seg{base + f, base + l, limit, (uint32_t)ue};
                bound[base + f] = 2;
            } else {
                bound[base + cut] = 1;
                const uint32_t c0 = cut - f, c1 = l - cut;
                c0n = c0 > 16 && lim > 0; c1n = c1 > 16 && lim > 0;
                if (c0 > 16 && lim == 0) heap[atomicAdd(&cnt->nheap, 1u)] = Seg{base + f, base + cut, 0, 0};
                if (c1 > 16 && lim == 0) heap[atomicAdd(&cnt->nheap, 1u)] = Seg{base + cut, base + l, 0, 0};
                nc = (c0n ? 1u : 0u) + (c1n ? 1u : 0u);
            }
        }
        uint32_t tot;
        uint32_t q = block_excl_scan(nc, tmp, tot);
        if (c0n) { sf[nx][q] = (uint16_t)f; sl[nx][q] = (uint16_t)cut; slim[nx][q] = (uint8_t)lim; ++q; }
        if (c1n) { sf[nx][q] = (uint16_t)cut; sl[nx][q] = (uint16_t)l; slim[nx][q] = (uint8_t)lim; }
        __syncthreads();
        if (tid < tot) spiv[nx][tid] = median_to_first(sf[nx][tid], sl[nx][tid], X, K);
        __syncthreads();
        cur = nx;
        nsub = tot;
    }
    for (uint32_t p = tid; p < len; p += 256) { skey[base + p] = K[p]; sidx[base + p] = X[p]; }
}

DEV bool KLESS(const int32_t *skey, uint32_t a, uint32_t b) { return skey[a] < skey[b]; }

// libstdc++ heap routines on one range, one thread (std::__make_heap,
// std::__adjust_heap, std::__push_heap, std::__sort_heap).
DEV void heap_adjust(uint32_t *si, int32_t *sk, uint32_t base, int64_t hole, int64_t len, uint32_t vi, int32_t vk) {
    const int64_t top = hole;
    int64_t child = hole;
    while (child < (len - 1) / 2) {
        child = 2 * (child + 1);
        if (sk[base + child] < sk[base + child - 1]) child--;
        si[base + hole] = si[base + child]; sk[base + hole] = sk[base + child];
        hole = child;
    }
    if ((len & 1) == 0 && child == (len - 2) / 2) {
        child = 2 * (child + 1);
        si[base + hole] = si[base + child - 1]; sk[base + hole] = sk[base + child - 1];
        hole = child - 1;
    }
    int64_t parent = (hole - 1) / 2;
    while (hole > top && sk[base + parent] < vk) {
        si[base + hole] = si[base + parent]; sk[base + hole] = sk[base + parent];
        hole = parent;
        parent = (hole - 1) / 2;
    }
    si[base + hole] = vi; sk[base + hole] = vk;
}

__global__ void k_sort_heap(const Seg *heap, const uint32_t *heap_cnt, uint32_t *si, int32_t *sk) {
    uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= *heap_cnt) return;
    uint32_t base = heap[s].first;
    int64_t len = heap[s].last - heap[s].first;
    if (len >= 2) {
        for (int64_t parent = (len - 2) / 2;; --parent) {
            heap_adjust(si, sk, base, parent, len, si[base + parent], sk[base + parent]);
            if (parent == 0) break;
        }
    }
    for (int64_t last = len; last > 1;) {
        --last;
        uint32_t vi = si[base + last]; int32_t vk = sk[base + last];
        si[base + last] = si[base]; sk[base + last] = sk[base];
        heap_adjust(si, sk, base, 0, last, vi, vk);
    }
}

// Leaves are the ranges between consecutive boundaries; boundaries are marked in
// `bound` (1 at each range start).  Each boundary thread insertion-sorts its range.
__global__ void k_sort_insertion(const uint8_t *bound, uint64_t n, uint32_t *si, int32_t *sk) {
    uint64_t b = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (b >= n || bound[b] != 1) return;  // 2: an all-equal range, nothing moves
    uint64_t e = b + 1;
    while (e < n && !bound[e]) ++e;
    for (uint64_t i = b + 1; i < e; ++i) {
        uint32_t vi = si[i]; int32_t vk = sk[i];
        uint64_t j = i;
        while (j > b && vk < sk[j - 1]) { si[j] = si[j - 1]; sk[j] = sk[j - 1]; --j; }
        si[j] = vi; sk[j] = vk;
    }
}

__global__ void k_sort_rank(const uint32_t *si, uint64_t n, uint32_t *rank, uint32_t *rinv) {
    uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t v = si[i];
    rank[v] = (uint32_t)i;
    rinv[i] = v;
}

// Elements of all-equal ranges go to their closed-form places (after the
// plain ranks, which they overwrite).
__global__ void k_sort_rank_uniform(const Seg *uni, uint32_t nuni, uint32_t total, const uint32_t *si,
                                    uint32_t *rank, uint32_t *rinv) {
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= total) return;
    // the range of the wave's first active element by one search on a
    // wave-uniform index, then each lane steps on to its own (a range has more
    // than 16 elements, so a wave spans at most five)
    uint32_t si0 = seg_find(uni, nuni, (uint32_t)__builtin_amdgcn_readfirstlane((int)e));
    while (si0 + 1 < nuni && uni[si0 + 1].off <= e) ++si0;
    const Seg g = uni[si0];
    const uint32_t p = g.first + (e - g.off);
    const uint32_t v = si[p], q = uniform_final(g.first, g.last, p);
    rank[v] = q;
    rinv[q] = v;  // (positions of an all-equal range are a permutation of it)
}

// ================================================================= optimize
__global__ __launch_bounds__(256) void k_opt_init(Dev d) {
    uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i >= d.W0) return;
    int32_t a = (int32_t)(d.w0_rs1[i] - 64u), b = (int32_t)(d.w0_rs2[i] - 64u);
    uint32_t w = d.w0_instr[i] & 0x707Fu;
    bool st = w == 0x2023u || w == 0x2027u;
    if (a >= 0) { if ((uint32_t)a < d.W0) { d.used[a] = 1; if (st) d.sstore[a] = 1; } else atomicOr(d.flags, EF_OOB); }
    if (b >= 0) { if ((uint32_t)b < d.W0) { d.used[b] = 1; if (st) d.sstore[b] = 1; } else atomicOr(d.flags, EF_OOB); }
}

// One sweep: a word whose destination register is unused clears its sources
// (except sources of stores).  Repeated until a sweep changes nothing.  A
// word's sources are mostly the words just before it, so each workgroup
// repeats its own range of OPT_RANGE words until no word of the range lost its
// use; only chains crossing ranges need further sweeps.  `changed` is set by a
// clear of a word outside the clearing workgroup's range, by a word whose
// destination lies outside its range (its flag could change under it), or by a
// range that did not settle within its 64 rounds: a launch with none of these
// has every range settled on values no other workgroup changed, which is the
// fixpoint.  (Generated code names a word's own register as its destination,
// so the second case does not arise there.)
constexpr uint32_t OPT_RANGE = 256 * 8;

// The range's fields and its used/store flags are staged in LDS once; the
// rounds read them there (flags of other ranges from HBM).  A clear is written
// to both (flags only go from 1 to 0, so HBM never takes back a clear made by
// another workgroup); a clear another workgroup makes in this range during the
// launch is a cross clear, which asks for the next launch, where it is staged.
//
// The round's "some word of the range lost its use" flag is double-buffered by
// round parity.  Thread 0 clears local[(round + 1) & 1] at the top of the next
// round while slower waves may still be reading local[round & 1] after the
// round's last barrier; it can only come back to local[round & 1] (two rounds
// later) after every wave has passed the next round's first barrier, i.e.
// after every wave has read it.  With a single flag, a wave still to read it
// could see the next round's clear, leave the loop and drop its words from
// the later rounds, and the launch could then report a fixpoint early.
// Thread t sweeps its own 8 consecutive words, highest first, so a chain of
// words naming the word before (the usual shape) settles inside one sweep of
// its thread rather than one link per round; the range is kept in LDS
// transposed (word k at (k % 8) * 256 + k / 8), so that the threads' k-th words
// sit in consecutive banks.
DEV uint32_t opt_tp(uint32_t k) { return (k & 7u) * 256u + (k >> 3); }

__global__ __launch_bounds__(256) void k_opt_iter(Dev d, uint32_t *changed) {
    static_assert(OPT_RANGE == 256 * 8, "8 words per thread");
    __shared__ int local[2];
    __shared__ uint32_t s_rd[OPT_RANGE], s_r1[OPT_RANGE], s_r2[OPT_RANGE];
    __shared__ uint8_t s_used[OPT_RANGE], s_st[OPT_RANGE];
    const uint64_t b0 = (uint64_t)blockIdx.x * OPT_RANGE;
    const uint64_t b1 = b0 + OPT_RANGE < d.W0 ? b0 + OPT_RANGE : d.W0;
    const uint32_t nr = (uint32_t)(b1 - b0);
    const uint32_t t = threadIdx.x, k0 = 8u * t;
    if (k0 + 8u <= nr) {
        // this thread's 8 words with two 16-byte loads per field (b0 and k0 are multiples of 8)
        const uint4 *rd4 = reinterpret_cast<const uint4 *>(d.w0_rd + b0 + k0);
        const uint4 *r14 = reinterpret_cast<const uint4 *>(d.w0_rs1 + b0 + k0);
        const uint4 *r24 = reinterpret_cast<const uint4 *>(d.w0_rs2 + b0 + k0);
        const uint4 x0 = rd4[0], x1 = rd4[1], y0 = r14[0], y1 = r14[1], z0 = r24[0], z1 = r24[1];
        const uint2 u8 = *reinterpret_cast<const uint2 *>(d.used + b0 + k0);
        const uint2 s8 = *reinterpret_cast<const uint2 *>(d.sstore + b0 + k0);
        const uint32_t xr[8] = {x0.x, x0.y, x0.z, x0.w, x1.x, x1.y, x1.z, x1.w};
        const uint32_t yr[8] = {y0.x, y0.y, y0.z, y0.w, y1.x, y1.y, y1.z, y1.w};
        const uint32_t zr[8] = {z0.x, z0.y, z0.z, z0.w, z1.x, z1.y, z1.z, z1.w};
#pragma unroll
        for (uint32_t j = 0; j < 8; ++j) {
            const uint32_t e = j * 256u + t;
            s_rd[e] = xr[j]; s_r1[e] = yr[j]; s_r2[e] = zr[j];
            s_used[e] = (uint8_t)((j < 4 ? u8.x : u8.y) >> (8u * (j & 3u)));
            s_st[e] = (uint8_t)((j < 4 ? s8.x : s8.y) >> (8u * (j & 3u)));
        }
    } else {
        for (uint32_t j = 0; j < 8; ++j) {
            const uint32_t k = k0 + j;
            if (k >= nr) break;
            const uint32_t e = j * 256u + t;
            s_rd[e] = d.w0_rd[b0 + k]; s_r1[e] = d.w0_rs1[b0 + k]; s_r2[e] = d.w0_rs2[b0 + k];
            s_used[e] = d.used[b0 + k]; s_st[e] = d.sstore[b0 + k];
        }
    }
    auto used_of = [&](uint32_t x) -> bool { return (uint64_t)x - b0 < nr ? s_used[opt_tp(x - (uint32_t)b0)] != 0 : d.used[x] != 0; };
    auto store_of = [&](uint32_t x) -> bool { return (uint64_t)x - b0 < nr ? s_st[opt_tp(x - (uint32_t)b0)] != 0 : d.sstore[x] != 0; };
    bool cross = false;  // a clear outside the range, or a destination outside it
    bool settled = false;
    for (int round = 0; round < 64; ++round) {
        if (threadIdx.x == 0) local[round & 1] = 0;
        __syncthreads();
        bool loc = false;
        for (int32_t j = 7; j >= 0; --j) {
            const uint32_t k = k0 + (uint32_t)j;
            if (k >= nr) continue;
            const uint32_t e = (uint32_t)j * 256u + t;
            const uint32_t rd = s_rd[e];
            if (!((int32_t)rd > 63)) continue;
            const uint32_t v = rd - 64u;
            if (v >= d.W0) continue;
            cross |= (uint64_t)v - b0 >= nr;
            if (used_of(v)) continue;
            const int32_t a = (int32_t)(s_r1[e] - 64u), b = (int32_t)(s_r2[e] - 64u);
            if (a >= 0 && (uint32_t)a < d.W0 && !store_of((uint32_t)a) && used_of((uint32_t)a)) {
                const bool in = (uint64_t)(uint32_t)a - b0 < nr;
                if (in) s_used[opt_tp((uint32_t)a - (uint32_t)b0)] = 0;
                d.used[a] = 0;
                loc |= in; cross |= !in;
            }
            if (b >= 0 && (uint32_t)b < d.W0 && !store_of((uint32_t)b) && used_of((uint32_t)b)) {
                const bool in = (uint64_t)(uint32_t)b - b0 < nr;
                if (in) s_used[opt_tp((uint32_t)b - (uint32_t)b0)] = 0;
                d.used[b] = 0;
                loc |= in; cross |= !in;
            }
        }
        if (loc) local[round & 1] = 1;
        __syncthreads();
        if (!local[round & 1]) { settled = true; break; }
    }
    if (cross || (!settled && threadIdx.x == 0)) *changed = 1;
}

__global__ __launch_bounds__(256) void k_opt_final(Dev d) {
    uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i >= d.W0) return;
    const uint32_t rd = d.w0_rd[i];
    d.usedi[i] = (!((int32_t)rd > 63) || d.used[i]) ? 1 : 0;
    d.w0_f4[i] = make_uint4(rd, d.w0_rs1[i], d.w0_rs2[i], d.w0_instr[i]);  // regalloc's view of the word
}

// ================================================================= regalloc
struct RaStep {
    uint32_t ord, rs1, rs2, r1r, r2r, rdr;
    bool sw1, sw2, swd;
    uint64_t newlife;
};

// One allocation step of one function (rv_generator_avx.cpp:1062-1397), given
// the word's fields, the operands' gathered registers (raw1/raw2: 0 when not
// virtual) and their swapped states.
DEV RaStep ra_eval_raw(uint32_t ins, uint32_t ord, uint32_t rs1, uint32_t rs2, uint32_t raw1, uint32_t raw2,
                       bool s1, bool s2, uint64_t life) {
    RaStep r;
    r.rs1 = rs1; r.rs2 = rs2; r.ord = ord;
    const bool v1 = (int32_t)rs1 > 63, v2 = (int32_t)rs2 > 63;
    r.r1r = v1 ? raw1 : 0u;
    r.r2r = v2 ? raw2 : 0u;
    if (v1 && s1) r.r1r = (ins == 0xC0007053u || (ins & 0x7Fu) == 0x53u) ? 37u : 5u;
    if (v2 && s2) r.r2r = ((ins & 0x7Fu) == 0x53u) ? 38u : 6u;
    const uint64_t clr = ((1ull << r.r1r) | (1ull << r.r2r)) & ~1ull;
    const uint64_t cl = life & ~clr;
    uint32_t rdr = ord;
    if ((int32_t)ord > 63) {
        const bool rdf = ins == 0xD0007053u || (ins & 0x7Fu) == 0x53u;
        rdr = ffz_gcc(cl | (rdf ? 0xFFFFFFFFull : 0xFFFFFFFF00000000ull));
        if (rdr == 64) rdr = rdf ? 37u : 5u;
    }
    r.rdr = rdr;
    r.sw1 = r.r1r != 0 && ((life >> r.r1r) & 1ull);
    r.sw2 = r.r2r != 0 && ((life >> r.r2r) & 1ull);
    r.swd = rdr != 0 && ((cl >> rdr) & 1ull);
    r.newlife = cl | (1ull << rdr);
    return r;
}

DEV bool virt_ok(const Dev &d, uint32_t v) {
    if ((int32_t)v <= 63) return false;
    if (v - 64u >= d.W0) { atomicOr(d.flags, EF_OOB); return false; }
    return true;
}

DEV RaStep ra_eval(const Dev &d, uint32_t off, uint64_t life, bool s1, bool s2) {
    const uint4 w = d.w0_f4[off];
    const uint32_t raw1 = virt_ok(d, w.y) ? (d.sym[w.y - 64u] & 0xFFu) : 0u;
    const uint32_t raw2 = virt_ok(d, w.z) ? (d.sym[w.z - 64u] & 0xFFu) : 0u;
    return ra_eval_raw(w.w, w.x, w.y, w.z, raw1, raw2, s1, s2, life);
}

DEV bool sym_swapped(const Dev &d, uint32_t v) { return (d.sym[v] >> 8) != 0; }

// Register allocation keeps each virtual register's state with the one thread
// that walks the function naming it, which is exact when no register is named
// by two functions.  isn_gen ORs a word's own virtual register into physical
// argument registers (rv_generator_avx.cpp:676-697 for rd; operands, which
// read such a result from a parent slot, :842), and for calls with many
// arguments the result can name a word of another function.  Such
// names are noted here; the host then checks that no other function names the
// same register (k_xref_check), and reallocates every program where one does
// in the reference's exact step and group order (k_ra_serial).
DEV void xref_note(const Dev &d, uint32_t v, uint32_t f) {
    const uint32_t e = v - 64u;
    if (e >= d.W0) return;  // out of range: flagged by the caller's bounds checks
    const uint32_t prev = atomicCAS(&d.xref[e], 0u, f + 1u);
    if (prev != 0u && prev != f + 1u) { atomicOr(d.flags, EF_XCONFLICT); d.cflag[prog_of(d.pf, d.P, f)] = 1; }
    atomicOr(d.flags, EF_XNAMED);
}

DEV void xref_check_fields(const Dev &d, uint32_t f, uint32_t start, uint32_t size, uint32_t ord, uint32_t rs1,
                           uint32_t rs2) {
    if ((int32_t)ord > 63 && ord - 64u - start >= size) xref_note(d, ord, f);
    if ((int32_t)rs1 > 63 && rs1 - 64u - start >= size) xref_note(d, rs1, f);
    if ((int32_t)rs2 > 63 && rs2 - 64u - start >= size) xref_note(d, rs2, f);
}
DEV void sym_mark(const Dev &d, uint32_t v) { reinterpret_cast<uint8_t *>(d.sym)[2u * v + 1u] = 1; }

// register_state updates of one step (rv_generator_avx.cpp:1356-1397): each
// named virtual register takes its register and, as the lane-0 lockstep
// commits, a cleared swapped flag.  The swapped flags of functions outside
// lane 0 are never set, so writing the whole entry is the same as writing the
// register alone there.
DEV void ra_commit_syms(const Dev &d, const RaStep &r) {
    if (((int32_t)r.ord > 63 && r.ord - 64u >= d.W0) || ((int32_t)r.rs1 > 63 && r.rs1 - 64u >= d.W0) ||
        ((int32_t)r.rs2 > 63 && r.rs2 - 64u >= d.W0)) { atomicOr(d.flags, EF_OOB); return; }
    if ((int32_t)r.ord > 63) d.sym[r.ord - 64u] = (uint16_t)r.rdr;
    if ((int32_t)r.rs1 > 63) d.sym[r.rs1 - 64u] = (uint16_t)r.r1r;
    if ((int32_t)r.rs2 > 63) d.sym[r.rs2 - 64u] = (uint16_t)r.r2r;
}

// Functions outside lane 0 of their group: their swapped flags are never set
// (only lane-0 registers are ever looked up), so each walks alone.  Per step,
// the registers each group looks up are recorded as min/max group per register.
// permw lists these functions by size, largest first, so that the 64 walks of
// a workgroup take about the same number of steps.  A function's virtual
// registers are named by that function alone (k_xref_check), so each walk keeps
// them in an LDS cache of WALK_CACHE entries (slot = word offset in the
// function mod WALK_CACHE; entry = offset << 10 | valid << 9 | state), reads
// HBM only on a miss, and writes every change through to HBM (stores that are
// never waited for).  The words themselves are staged
// into LDS WALK_CHUNK steps at a time, loaded cooperatively (consecutive
// threads, consecutive words of one function).
constexpr uint32_t WALK_CACHE = 64;
// The walks' register ring of the newest definitions holds RING + 1 entries
// (RING is the index mask: 3 or 7, chosen by the autotuner per size bucket;
// 2 and 16 entries measured slower on every workload).
constexpr uint32_t WALK_CHUNK = 32;
// Workgroups spread their per-step lookup records over up to WALK_COPIES
// copies of the (step, register) tables (copy = workgroup mod copies), which
// k_walk_merge folds: one table shared by every workgroup serializes its
// atomics on the same few addresses.
constexpr uint32_t WALK_COPIES = 16;
// Walk and lockstep may run at the same time: for each program whose lockstep
// does, each walk workgroup holding functions of it publishes the program's
// table rows in chunks of WALK_PUB rows, adding 1 to done[cbase + chunk]
// (release, agent scope) once it has posted every row of the chunk it takes
// part in (the lowest lane of the program in the workgroup publishes).  The
// lockstep reads a row after the counter reaches need[cbase + chunk], the
// number of walk workgroups holding a function of the program that still runs
// at the chunk's first row.  Walk workgroups never wait, so when both kernels are resident
// the lockstep's wait always ends.  Where dispatches are serialized (a profiler
// collecting counters), the walk cannot start before the lockstep ends: a wait
// that sees no progress for RA0_WAIT_TICKS of the 100 MHz wall clock gives up
// for the rest of the kernel (EF_WAIT), and the host redoes the allocation with
// the two kernels one after the other.
constexpr uint32_t WALK_PUB = 32;
constexpr uint32_t RA0_MAXCOPIES = 8;  // table copies the lockstep folds itself
constexpr uint64_t RA0_WAIT_TICKS = 25000000ull;  // 0.25 s

// One program's allocation: functions from fbase (nf of them) in groups of 8
// (L groups), maxsize = its longest function; its slices of the lane-0 group
// order (perm + poff) and of the (step, register) lookup tables (+ gbase).  The
// tables hold `rows` steps: every step while a lane-0 function is running, then
// (when some other function runs longer) one last row that gathers all later
// steps.  Past its lane-0 functions the snapshot no longer changes and no owner
// reads, so a later lookup only marks, whatever its step.
// cbase: its first row chunk in the walk's publish counters (RA0_NOPUB when its
// lockstep does not run at the same time as the walk).
struct Ra0Prog { uint32_t fbase, L, maxsize, poff; uint64_t gbase; uint32_t nf, rows, cbase, pad; };
constexpr uint32_t RA0_NOPUB = 0xFFFFFFFFu;

// OR of v over the 8 lanes of each group of a wave, every lane active (DPP:
// swap neighbours, swap pairs, then mirror each half row of 8).
DEV uint32_t sm_or8(uint32_t v) {
    v |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0xB1, 0xF, 0xF, false);   // quad_perm [1,0,3,2]
    v |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x4E, 0xF, 0xF, false);   // quad_perm [2,3,0,1]
    v |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x141, 0xF, 0xF, false);  // row_half_mirror
    return v;
}
DEV uint64_t sm_or8_64(uint64_t v) { return (uint64_t)sm_or8((uint32_t)v) | ((uint64_t)sm_or8((uint32_t)(v >> 32)) << 32); }

// The walk's cache of a function's states, kept by one thread: ring slot
// x & WALK_RING holds the newest definitions (tag x << 10 | valid << 9 |
// state); the LDS cache c[x % WALK_CACHE][lane] holds older ones; HBM (sym)
// all of them (written through).  The ring is named registers selected by
// index bits (an array would be indexed through scratch memory); the walks use
// four of its eight.
struct WcRing { uint32_t r0, r1, r2, r3, r4, r5, r6, r7; };

DEV uint32_t wc_ring_at(const WcRing &g, uint32_t k) {
    const uint32_t a = (k & 1u) ? g.r1 : g.r0, b = (k & 1u) ? g.r3 : g.r2;
    const uint32_t c = (k & 1u) ? g.r5 : g.r4, e = (k & 1u) ? g.r7 : g.r6;
    const uint32_t lo = (k & 2u) ? b : a, hi = (k & 2u) ? e : c;
    return (k & 4u) ? hi : lo;
}

DEV void wc_ring_set(WcRing &g, uint32_t k, uint32_t v) {
    g.r0 = k == 0 ? v : g.r0; g.r1 = k == 1 ? v : g.r1; g.r2 = k == 2 ? v : g.r2; g.r3 = k == 3 ? v : g.r3;
    g.r4 = k == 4 ? v : g.r4; g.r5 = k == 5 ? v : g.r5; g.r6 = k == 6 ? v : g.r6; g.r7 = k == 7 ? v : g.r7;
}

// Write-through variants: every state change is stored to HBM as it is made,
// so HBM always holds each register's current state and a cache slot may be
// overwritten without a look at what it held.  Invariant: a register held in
// the ring is current there; one held in its LDS slot and not in the ring is
// current there; a register leaving the ring overwrites its LDS slot.
//
// A step reads its two operands' states without branching on where they are:
// both ring slots and both cache slots are read for every lane, and the lanes
// whose operand is in neither (or is another function's register) load it from
// HBM inside one wave-uniform branch that waits for its own loads.  raw1 / raw2
// are 0 for an operand that is not a virtual register (one past the register
// space is flagged, as virt_ok does).
template <uint32_t WALK_RING>
DEV void walk_gets(const Dev &d, uint32_t (*c)[64], const WcRing &ring, uint32_t lane, uint32_t start,
                   uint32_t size, bool con, bool on, uint32_t y, uint32_t z, uint32_t &raw1, uint32_t &raw2) {
    const bool n1 = on && (int32_t)y > 63, n2 = on && (int32_t)z > 63;
    if ((n1 && y - 64u >= d.W0) || (n2 && z - 64u >= d.W0)) atomicOr(d.flags, EF_OOB);
    const bool v1 = n1 && y - 64u < d.W0, v2 = n2 && z - 64u < d.W0;
    const uint32_t x1 = y - 64u - start, x2 = z - 64u - start;
    const bool l1 = v1 && con && x1 < size, l2 = v2 && con && x2 < size;
    const uint32_t e1 = wc_ring_at(ring, x1 & WALK_RING), e2 = wc_ring_at(ring, x2 & WALK_RING);
    const uint32_t c1 = c[x1 % WALK_CACHE][lane], c2 = c[x2 % WALK_CACHE][lane];
    const bool r1 = e1 != 0u && (e1 >> 10) == x1, r2 = e2 != 0u && (e2 >> 10) == x2;
    const bool m1 = v1 && !(l1 && (r1 || (c1 != 0u && (c1 >> 10) == x1)));
    const bool m2 = v2 && !(l2 && (r2 || (c2 != 0u && (c2 >> 10) == x2)));
    raw1 = v1 ? ((r1 ? e1 : c1) & 0x1FFu) : 0u;
    raw2 = v2 ? ((r2 ? e2 : c2) & 0x1FFu) : 0u;
    if (__ballot(m1 || m2)) {
        // every lane loads (entry 0 where it hit), so that both loads issue
        // back to back under one wait
        const uint32_t g1 = d.sym[m1 ? y - 64u : 0u];
        const uint32_t g2 = d.sym[m2 ? z - 64u : 0u];
        __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): no load is outstanding past this branch
        if (m1) { raw1 = g1; if (l1) c[x1 % WALK_CACHE][lane] = (x1 << 10) | 0x200u | g1; }
        if (m2) { raw2 = g2; if (l2) c[x2 % WALK_CACHE][lane] = (x2 << 10) | 0x200u | g2; }
    }
}

// Stages the words of steps i .. i+WALK_CHUNK-1 of the 64 walked functions in
// LDS: thread t loads word i + t % 32 of functions t / 32, t / 32 + 2, ... (two
// batches of 16 loads, consecutive threads on consecutive words of one
// function).  Each load's function start and size come from the lanes that hold
// them (readlane), and past a function's end the load reads word 0 and stages
// zeros, so that the loads issue back to back.
DEV void walk_stage(const Dev &d, uint4 (*words)[64], uint32_t i, uint32_t lane, uint32_t start, uint32_t size) {
    const uint32_t kk = lane % WALK_CHUNK;
    const bool odd = lane >= WALK_CHUNK;
#pragma unroll
    for (uint32_t h = 0; h < 2; ++h) {
        uint4 tmp[16];
#pragma unroll
        for (uint32_t r = 0; r < 16; ++r) {
            const uint32_t g = 2u * (16u * h + r);
            const uint32_t st0 = (uint32_t)__builtin_amdgcn_readlane((int)start, (int)g);
            const uint32_t st1 = (uint32_t)__builtin_amdgcn_readlane((int)start, (int)g + 1);
            const uint32_t sz0 = (uint32_t)__builtin_amdgcn_readlane((int)size, (int)g);
            const uint32_t sz1 = (uint32_t)__builtin_amdgcn_readlane((int)size, (int)g + 1);
            const uint32_t st = odd ? st1 : st0, sz = odd ? sz1 : sz0;
            const bool in = i + kk < sz;
            const uint4 wv = d.w0_f4[in ? st + i + kk : 0u];
            tmp[r] = in ? wv : make_uint4(0, 0, 0, 0);
        }
#pragma unroll
        for (uint32_t r = 0; r < 16; ++r) words[kk][(odd ? 1u : 0u) + 2u * (16u * h + r)] = tmp[r];
    }
}


template <uint32_t WALK_RING>
__global__ __launch_bounds__(64) void k_ra_walk(Dev d, uint32_t *gmin_all, uint32_t *gmax_all, uint64_t stride,
                                                uint32_t copies, const Ra0Prog *progs, const uint32_t *permw,
                                                uint32_t nw, uint32_t *done) {
    __shared__ uint32_t smin[64], smax[64];
    __shared__ uint32_t cache[WALK_CACHE][64];
    __shared__ uint4 words[WALK_CHUNK][64];
    const uint32_t lane = threadIdx.x;
    const uint32_t t = blockIdx.x * 64u + lane;
    const uint32_t f = t < nw ? permw[t] : d.F;
    const bool act = f < d.F;
    const uint32_t start = act ? d.fstarts[f] : 0u;
    const uint32_t size = act ? d.fsizes[f] : 0u;
    // group = function index within its program / 8; lookup table of the program
    const uint32_t prg = act ? prog_of(d.pf, d.P, f) : 0xFFFFFFFFu;
    const uint32_t grp = act ? (f - d.pf[prg]) >> 3 : 0u;
    uint32_t pmin = prg, pmax = act ? prg : 0u;
    for (int s = 32; s > 0; s >>= 1) { pmin = min(pmin, (uint32_t)__shfl_xor((int)pmin, s, 64)); pmax = max(pmax, (uint32_t)__shfl_xor((int)pmax, s, 64)); }
    const bool one = pmin == pmax;  // the workgroup's functions all belong to one program
    uint32_t *gmin = gmin_all + (blockIdx.x % copies) * stride;
    uint32_t *gmax = gmax_all + (blockIdx.x % copies) * stride;
    const uint64_t tb = one ? progs[pmin].gbase : 0u;  // table base of that program
    const uint32_t tlast = one ? progs[pmin].rows - 1u : 0u;
    const uint64_t mbase = act ? progs[prg].gbase : 0u;  // own program's table and its last row
    const uint32_t mlast = act ? progs[prg].rows - 1u : 0u;
    // the lowest lane of each program in the workgroup publishes that program's
    // rows; pw = the size of the program's longest function here
    uint32_t pw = 0;
    bool lead = false;
    if (done) {
        uint64_t rem = __ballot(act);
        while (rem) {
            const uint32_t l = (uint32_t)__builtin_ctzll(rem);
            const uint32_t p = (uint32_t)__shfl((int)prg, (int)l, 64);
            const bool mine = act && prg == p;
            const uint64_t same = __ballot(mine);
            uint32_t m = mine ? size : 0u;
            for (int s = 32; s > 0; s >>= 1) m = max(m, (uint32_t)__shfl_xor((int)m, s, 64));
            if (mine) { pw = m; lead = lane == l; }
            rem &= ~same;
        }
    }
    const uint32_t cbase = act ? progs[prg].cbase : RA0_NOPUB;
    const bool publishes = done && lead && cbase != RA0_NOPUB;
    const bool con = size < (1u << 22);  // cache on (tags hold 22 bits of word offset)
    WcRing ring{0, 0, 0, 0, 0, 0, 0, 0};
    // update of virtual register v (inside [64, 64 + W0))
    auto put = [&](uint32_t v, uint32_t val, bool own) {  // own: v is defined by the word being committed
        const uint32_t x = v - 64u - start;
        d.sym[v - 64u] = (uint16_t)val;  // write-through
        if (!con || x >= size) return;
        const uint32_t e = wc_ring_at(ring, x & WALK_RING);
        if (e && (e >> 10) == x) { wc_ring_set(ring, x & WALK_RING, (x << 10) | 0x200u | val); return; }
        if (!own) { cache[x % WALK_CACHE][lane] = (x << 10) | 0x200u | val; return; }
        if (e) cache[(e >> 10) % WALK_CACHE][lane] = e;
        wc_ring_set(ring, x & WALK_RING, (x << 10) | 0x200u | val);
    };
    for (uint32_t k = 0; k < WALK_CACHE; ++k) cache[k][lane] = 0;
    smin[lane] = 0xFFFFFFFFu; smax[lane] = 0;
    uint32_t wmax = size;
    for (int s = 32; s > 0; s >>= 1) { uint32_t o = __shfl_xor(wmax, s, 64); wmax = wmax > o ? wmax : o; }
    uint64_t life = 0x1Full;
    uint4 wn = make_uint4(0, 0, 0, 0);  // the word of the next step
    __syncthreads();
    for (uint32_t i = 0; i < wmax; ++i) {
        if (i % WALK_CHUNK == 0) {
            walk_stage(d, words, i, lane, start, size);
            __syncthreads();
            wn = words[0][lane];
        }
        const uint4 w = wn;
        if ((i + 1) % WALK_CHUNK != 0) wn = words[(i + 1) % WALK_CHUNK][lane];  // a step ahead
        const bool on = (int32_t)i < (int32_t)size;
        uint32_t raw1, raw2;
        walk_gets<WALK_RING>(d, cache, ring, lane, start, size, con, on, w.y, w.z, raw1, raw2);
        if (on) {
            raw1 &= 0xFFu; raw2 &= 0xFFu;
            const RaStep r = ra_eval_raw(w.w, w.x, w.y, w.z, raw1, raw2, false, false, life);
            xref_check_fields(d, f, start, size, r.ord, r.rs1, r.rs2);
            if (one) {
                if (r.swd) { atomicMin(&smin[r.rdr], grp); atomicMax(&smax[r.rdr], grp + 1u); }
                if (r.sw1) { atomicMin(&smin[r.r1r], grp); atomicMax(&smax[r.r1r], grp + 1u); }
                if (r.sw2) { atomicMin(&smin[r.r2r], grp); atomicMax(&smax[r.r2r], grp + 1u); }
            } else {  // functions of several programs: post straight to each program's table
                const uint64_t row = mbase + (uint64_t)min(i, mlast) * 64u;
                if (r.swd) { atomicMin(&gmin[row + r.rdr], grp); atomicMax(&gmax[row + r.rdr], grp + 1u); }
                if (r.sw1) { atomicMin(&gmin[row + r.r1r], grp); atomicMax(&gmax[row + r.r1r], grp + 1u); }
                if (r.sw2) { atomicMin(&gmin[row + r.r2r], grp); atomicMax(&gmax[row + r.r2r], grp + 1u); }
            }
            life = r.newlife;
            // register_state updates (ra_commit_syms): destination, then rs1, then rs2
            if (((int32_t)r.ord > 63 && r.ord - 64u >= d.W0) || ((int32_t)r.rs1 > 63 && r.rs1 - 64u >= d.W0) ||
                ((int32_t)r.rs2 > 63 && r.rs2 - 64u >= d.W0)) {
                atomicOr(d.flags, EF_OOB);
            } else {
                // only entries whose value changes are stored; a destination is
                // skipped when an operand names it (the operand's store is later)
                const uint32_t self = 64u + start + i;
                const bool vo = (int32_t)r.ord > 63, v1 = (int32_t)r.rs1 > 63, v2 = (int32_t)r.rs2 > 63;
                if (vo && !(v1 && r.rs1 == r.ord) && !(v2 && r.rs2 == r.ord)) put(r.ord, r.rdr, r.ord == self);
                if (v1 && !(v2 && r.rs2 == r.rs1) && r.r1r != raw1) put(r.rs1, r.r1r, false);
                if (v2 && r.r2r != raw2) put(r.rs2, r.r2r, false);
            }
        }
        lds_barrier();
        uint32_t mx = smax[lane];
        if (mx) {
            atomicMin(&gmin[tb + (uint64_t)min(i, tlast) * 64u + lane], smin[lane]);
            atomicMax(&gmax[tb + (uint64_t)min(i, tlast) * 64u + lane], mx);
            smin[lane] = 0xFFFFFFFFu; smax[lane] = 0;
        }
        lds_barrier();
        if (done) {
            // a program's functions here leave a chunk of its rows (steps past
            // its last row all post to that row): publish it after every
            // lane's posts are visible
            bool pub = false;
            uint32_t chunk = 0;
            if (publishes && i < pw) {
                const uint32_t r = min(i, mlast), rn = min(i + 1u, mlast);
                pub = i + 1u == pw || r / WALK_PUB != rn / WALK_PUB;
                chunk = cbase + r / WALK_PUB;
            }
            if (__ballot(pub)) {
                __threadfence();
                __syncthreads();
                if (pub) __hip_atomic_fetch_add(&done[chunk], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
    }
}

// gmin/gmax[e] = min/max over the copies (copy c at c * stride).
__global__ void k_walk_merge(uint32_t *gmin, uint32_t *gmax, uint64_t stride, uint32_t copies) {
    const uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (e >= stride) return;
    uint32_t mn = gmin[e], mx = gmax[e];
    for (uint32_t c = 1; c < copies; ++c) {
        mn = min(mn, gmin[c * stride + e]);
        mx = max(mx, gmax[c * stride + e]);
    }
    gmin[e] = mn; gmax[e] = mx;
}

// Programs of at most 8 functions (one group): the reference's vector loop
// runs them as one 8-lane vector, and only its lane 0 writes region 0 of
// register_state.  One wave holds 8 such programs, lane 8s + k walking function
// k of program slot s (the walk's register ring, LDS cache and word staging).
// Per step the 8 lanes' swap lookups are ORed across the group; lane 0 then
// marks the snapshot entries looked up (not one its own current word names:
// its reset lands after a same-group mark) and updates region 0 with its own
// write and resets.  Every lookup and mark of the step comes after every
// operand read of the step, as in the vector loop.  A mark on a register outside
// lane 0's own function cannot be kept in its cache: the program then takes the
// exact serial path.  The first `solo` programs of the list (the longest) have
// a wave each, in slot 0, so that no other program's rarer paths (a cache
// miss, a mark) hold up their steps.
template <uint32_t WALK_RING>
__global__ __launch_bounds__(64) void k_ra_group(Dev d, const uint32_t *plist, uint32_t np, uint32_t solo) {
    __shared__ uint32_t cache[WALK_CACHE][64];
    __shared__ uint4 words[WALK_CHUNK][64];
    __shared__ int32_t Sv[8][64];
    const uint32_t lane = threadIdx.x, slot = lane >> 3, k = lane & 7u;
    const uint32_t pi = blockIdx.x < solo ? (slot == 0 ? blockIdx.x : np) : solo + (blockIdx.x - solo) * 8u + slot;
    const uint32_t prg = pi < np ? plist[pi] : 0xFFFFFFFFu;
    const uint32_t fb = prg != 0xFFFFFFFFu ? d.pf[prg] : 0u;
    const uint32_t nf = prg != 0xFFFFFFFFu ? d.pf[prg + 1] - fb : 0u;
    const bool act = k < nf && nf <= 8u;
    if (prg != 0xFFFFFFFFu && nf > 8u) atomicOr(d.flags, EF_OOB);
    const uint32_t f = fb + k;
    const uint32_t start = act ? d.fstarts[f] : 0u;
    const uint32_t size = act ? d.fsizes[f] : 0u;
    const bool con = size < (1u << 22);
    WcRing ring{0, 0, 0, 0, 0, 0, 0, 0};
    auto put = [&](uint32_t v, uint32_t val, bool own) {  // own: v is defined by the word being committed
        const uint32_t x = v - 64u - start;
        d.sym[v - 64u] = (uint16_t)val;  // write-through
        if (!con || x >= size) return;
        const uint32_t e = wc_ring_at(ring, x & WALK_RING);
        if (e && (e >> 10) == x) { wc_ring_set(ring, x & WALK_RING, (x << 10) | 0x200u | val); return; }
        if (!own) { cache[x % WALK_CACHE][lane] = (x << 10) | 0x200u | val; return; }
        if (e) cache[(e >> 10) % WALK_CACHE][lane] = e;
        wc_ring_set(ring, x & WALK_RING, (x << 10) | 0x200u | val);
    };
    for (uint32_t q = 0; q < WALK_CACHE; ++q) cache[q][lane] = 0;
    Sv[slot][k] = -1; Sv[slot][k + 8] = -1; Sv[slot][k + 16] = -1; Sv[slot][k + 24] = -1;
    Sv[slot][k + 32] = -1; Sv[slot][k + 40] = -1; Sv[slot][k + 48] = -1; Sv[slot][k + 56] = -1;
    uint32_t wmax = size;
    for (int s = 32; s > 0; s >>= 1) { uint32_t o = __shfl_xor(wmax, s, 64); wmax = wmax > o ? wmax : o; }
    // a mark sets the swapped flag in place: in the ring, in the LDS cache, or
    // as a byte store to HBM (the register bits stay)
    auto mark = [&](uint32_t v) {
        const uint32_t x = v - 64u - start;
        sym_mark(d, v - 64u);  // write-through
        const uint32_t e = wc_ring_at(ring, x & WALK_RING);
        if (e && (e >> 10) == x) { wc_ring_set(ring, x & WALK_RING, e | 0x100u); return; }
        const uint32_t ce = cache[x % WALK_CACHE][lane];
        if (ce && (ce >> 10) == x) cache[x % WALK_CACHE][lane] = ce | 0x100u;
    };
    uint64_t life = 0x1Full;
    uint64_t valid = 0;  // lane 0 of a slot: region-0 entries holding a virtual register
    uint64_t mkd = 0;    // lane 0: entries whose register is known to be marked already
    uint4 wn = make_uint4(0, 0, 0, 0);  // the word of the next step
    __syncthreads();
    for (uint32_t i = 0; i < wmax; ++i) {
        if (i % WALK_CHUNK == 0) {
            walk_stage(d, words, i, lane, start, size);
            __syncthreads();
            wn = words[0][lane];
        }
        const bool on = (int32_t)i < (int32_t)size;
        const uint4 w = wn;
        if ((i + 1) % WALK_CHUNK != 0) wn = words[(i + 1) % WALK_CHUNK][lane];  // a step ahead
        RaStep r{};
        uint64_t look = 0;
        uint32_t raw1, raw2;
        walk_gets<WALK_RING>(d, cache, ring, lane, start, size, con, on, w.y, w.z, raw1, raw2);
        if (on) {
            // this step's commit may clear a set swapped flag (an operand's, or a
            // destination that is not the word's own new register): forget what
            // is marked
            if (((raw1 | raw2) & 0x100u) || ((int32_t)w.x > 63 && w.x != 64u + start + i)) mkd = 0;
            r = ra_eval_raw(w.w, w.x, w.y, w.z, raw1 & 0xFFu, raw2 & 0xFFu, (raw1 >> 8) & 1u, (raw2 >> 8) & 1u, life);
            xref_check_fields(d, f, start, size, r.ord, r.rs1, r.rs2);
            if (r.swd) look |= 1ull << r.rdr;
            if (r.sw1) look |= 1ull << r.r1r;
            if (r.sw2) look |= 1ull << r.r2r;
            life = r.newlife;
        }
        // the group's lookups of this step, at every lane of the group
        const uint64_t glook = sm_or8_64(look);
        if (k == 0) {
            // marks on the step-start snapshot, before any commit of the step
            uint64_t m = glook & valid & ~mkd;
            while (m) {
                const uint32_t reg = (uint32_t)__builtin_ctzll(m);
                m &= m - 1;
                const uint32_t v = (uint32_t)Sv[slot][reg];
                if (on && (v == r.ord || v == r.rs1 || v == r.rs2)) continue;
                if (v - 64u - start < size && con) { mark(v); mkd |= 1ull << reg; }
                else { atomicOr(d.flags, EF_XCONFLICT); d.cflag[prg] = 1; }
            }
        }
        if (on) {
            // register_state updates (ra_commit_syms): destination, then rs1, then rs2
            if (((int32_t)r.ord > 63 && r.ord - 64u >= d.W0) || ((int32_t)r.rs1 > 63 && r.rs1 - 64u >= d.W0) ||
                ((int32_t)r.rs2 > 63 && r.rs2 - 64u >= d.W0)) {
                atomicOr(d.flags, EF_OOB);
            } else {
                // only entries whose value changes are stored; a destination is
                // skipped when an operand names it (the operand's store is later)
                const uint32_t self = 64u + start + i;
                const bool vo = (int32_t)r.ord > 63, v1 = (int32_t)r.rs1 > 63, v2 = (int32_t)r.rs2 > 63;
                if (vo && !(v1 && r.rs1 == r.ord) && !(v2 && r.rs2 == r.ord)) put(r.ord, r.rdr, r.ord == self);
                if (v1 && !(v2 && r.rs2 == r.rs1) && r.r1r != raw1) put(r.rs1, r.r1r, false);
                if (v2 && r.r2r != raw2) put(r.rs2, r.r2r, false);
            }
            if (k == 0) {  // region 0: the write, then the resets (a reset of the written register is skipped)
                if (r.ord) {
                    Sv[slot][r.rdr] = (int32_t)r.ord;
                    valid = (int32_t)r.ord >= 64 ? valid | (1ull << r.rdr) : valid & ~(1ull << r.rdr);
                    mkd &= ~(1ull << r.rdr);
                }
                if (r.r1r != r.rdr) { Sv[slot][r.r1r] = -1; valid &= ~(1ull << r.r1r); }
                if (r.r2r != r.rdr) { Sv[slot][r.r2r] = -1; valid &= ~(1ull << r.r2r); }
            }
        }
    }
}

// ============================================================ small programs
// Programs of at most 64 functions (at most 8 groups of 8) whose words fit
// the LDS budget are allocated by one wave: lane = function (its index in the
// program, so lanes 8g .. 8g+7 are group g), each lane keeping the states of
// its function's virtual registers in LDS, one byte per word (register |
// touched << 6 | swapped << 7), and its lifetime in registers.  A step is the
// reference's step (rv_generator_avx.cpp:1041-1418 at one thread) for every
// function of the program at once:
//   * a function's state reads see its own group's body's view: the only
//     effect of the lower groups' bodies in the same step is a mark on a
//     region-0 entry (only lane-0 functions write region 0, and every lookup
//     reads it), which an owner sees as a swapped operand (a flip) when it
//     reads that register.  Owners that read an entry they own are settled in
//     group order, each from the final lookups of the groups below it, so one
//     pass reaches the in-order result;
//   * a lookup of register r marks the virtual that region 0 held at r at the
//     start of the step.  The owner's commit resets the swapped flag of the
//     registers it names, so a mark stands if a higher group made it, or if
//     the owner does not name that register in this step;
//   * region 0 of the next step: each lane-0 function's rd write, then its rs1
//     and rs2 frees, groups in order (the last one wins).
// Several programs can share the wave (program k on lanes seg[k] .. seg[k] +
// nf - 1, seg[k] a multiple of 8).  Lanes exchange values with DPP inside a
// group's 8 lanes and with readlane across groups; region 0 of each program
// is an LDS table read by lane = register, double-buffered by step parity.
// A register named outside its own function is noted for k_xref_check as in
// the other kernels; a mark on one, which the owner's bytes cannot hold,
// sends the program to the exact serial path.
constexpr uint32_t SM_CHUNK = 32;        // steps per staged chunk of words
constexpr uint32_t SM_PROGS = 8;         // programs per wave at most
constexpr uint32_t SM_STATE_OFF = 40960; // LDS bytes before the register states
constexpr uint32_t SM_STATE_MAX = 40960;   // LDS bytes of register states per wave at most (80 KB of
                                           // LDS in all: two waves per compute unit)
constexpr uint32_t SM_STATE_SMALL = 40960;
constexpr uint32_t SM_MIN_FUNCS = 9;       // programs of one group keep k_ra_group

struct SmWave { uint32_t np, state_bytes, prog[SM_PROGS], seg[SM_PROGS]; };

DEV uint32_t sm_rl(uint32_t v, uint32_t l) { return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)l); }
DEV uint64_t sm_rl64(uint64_t v, uint32_t l) { return (uint64_t)sm_rl((uint32_t)v, l) | ((uint64_t)sm_rl((uint32_t)(v >> 32), l) << 32); }

DEV uint64_t sm_look(const RaStep &r) {
    uint64_t m = 0;
    if (r.swd) m |= 1ull << r.rdr;
    if (r.sw1) m |= 1ull << r.r1r;
    if (r.sw2) m |= 1ull << r.r2r;
    return m;
}

__global__ __launch_bounds__(64) void k_ra_small(Dev d, const SmWave *waves) {
    extern __shared__ __align__(16) uint8_t sm_lds[];
    uint4 (*words)[64] = reinterpret_cast<uint4 (*)[64]>(sm_lds);                        // [SM_CHUNK][64]
    int32_t (*Sv)[SM_PROGS][64] = reinterpret_cast<int32_t (*)[SM_PROGS][64]>(sm_lds + 32768);  // [2]
    int8_t (*So)[SM_PROGS][64] = reinterpret_cast<int8_t (*)[SM_PROGS][64]>(sm_lds + 36864);    // [2]
    uint32_t *slot_base = reinterpret_cast<uint32_t *>(sm_lds + 37888);  // per group slot: lane-0 function's
    uint32_t *slot_start = slot_base + 8;                                // state base and first word
    uint32_t *slot_size = slot_base + 16;
    uint32_t *prog_slots = slot_base + 24;  // per program: first group slot | end slot << 8
    uint8_t *st = sm_lds + SM_STATE_OFF;
    __shared__ uint32_t fstart[64], fsize[64];
    const uint32_t lane = threadIdx.x;
    const SmWave wv = waves[blockIdx.x];
    // this lane's program, function and group slot
    uint32_t pk = 0;
    for (uint32_t k = 1; k < wv.np; ++k) if (wv.seg[k] <= lane) pk = k;
    const uint32_t prg = wv.prog[pk], seg = wv.seg[pk];
    const uint32_t fb = d.pf[prg], nf = d.pf[prg + 1] - fb;
    const uint32_t li = lane - seg;
    const bool has = li < nf;
    const uint32_t f = fb + li;
    const uint32_t start = has ? d.fstarts[f] : 0u;
    const uint32_t size = has ? d.fsizes[f] : 0u;
    const uint32_t slot = lane >> 3, k8 = lane & 7u;
    // state bytes: exclusive scan of the sizes rounded up to 4
    const uint32_t sz4 = (size + 3u) & ~3u;
    uint32_t incl = sz4;
    for (uint32_t o = 1; o < 64; o <<= 1) {
        const uint32_t u = (uint32_t)__shfl_up((int)incl, o, 64);
        if (lane >= o) incl += u;
    }
    const uint32_t base = incl - sz4;
    uint8_t *my = st + base;
    for (uint32_t x = 0; x < sz4; x += 4) *reinterpret_cast<uint32_t *>(my + x) = 0u;
    fstart[lane] = start; fsize[lane] = size;
    if (k8 == 0) { slot_base[slot] = base; slot_start[slot] = start; slot_size[slot] = size; }
    if (lane == seg) prog_slots[pk] = (seg >> 3) | (((seg + nf + 7u) >> 3) << 8);
    for (uint32_t p = 0; p < wv.np; ++p) { Sv[0][p][lane] = -1; So[0][p][lane] = -1; }
    const uint32_t np = wv.np;
    uint32_t wmax = size;
    for (int s = 32; s > 0; s >>= 1) { const uint32_t o = (uint32_t)__shfl_xor((int)wmax, s, 64); wmax = max(wmax, o); }
    uint64_t life = 0x1Full;
    uint64_t own = 0;  // lane-0 functions: the region-0 entries this group owns
    bool bad = false;  // a word names a register past the program's words: this lane stops (EF_OOB)
    __syncthreads();
    // state of virtual register v as read by this function: (register, swapped);
    // v is within the words (checked when its chunk was staged)
    auto rd_state = [&](uint32_t v, uint32_t &reg, bool &sw) {
        reg = 0; sw = false;
        if (!((int32_t)v > 63)) return;
        const uint32_t x = v - 64u - start;
        if (x < size) { const uint32_t b = my[x]; reg = b & 0x3Fu; sw = (b >> 7) != 0; }
        else { const uint32_t g = d.sym[v - 64u]; reg = g & 0xFFu; sw = (g >> 8) != 0; }
    };
    auto wr_state = [&](uint32_t v, uint32_t reg) {
        const uint32_t x = v - 64u - start;
        if (x < size) my[x] = (uint8_t)(reg | 0x40u);
        else d.sym[v - 64u] = (uint16_t)reg;
    };
    for (uint32_t i = 0; i < wmax; ++i) {
        if (i % SM_CHUNK == 0) {
            // thread t stages word i + t % 32 of functions t / 32, t / 32 + 2, ...
            const uint32_t kk = lane % SM_CHUNK, g0 = lane / SM_CHUNK;
#pragma unroll
            for (uint32_t h = 0; h < 2; ++h) {
                uint4 tmp[16];
#pragma unroll
                for (uint32_t r = 0; r < 16; ++r) {
                    const uint32_t g = g0 + 2u * (16u * h + r);
                    tmp[r] = i + kk < fsize[g] ? d.w0_f4[fstart[g] + i + kk] : make_uint4(0, 0, 0, 0);
                }
#pragma unroll
                for (uint32_t r = 0; r < 16; ++r) words[kk][g0 + 2u * (16u * h + r)] = tmp[r];
            }
            __syncthreads();
            // this function's words of the chunk: registers past the words, and
            // names outside the function (noted for k_xref_check)
            for (uint32_t k = 0; k < SM_CHUNK && i + k < size; ++k) {
                const uint4 wk = words[k][lane];
                if (((int32_t)wk.x > 63 && wk.x - 64u >= d.W0) || ((int32_t)wk.y > 63 && wk.y - 64u >= d.W0) ||
                    ((int32_t)wk.z > 63 && wk.z - 64u >= d.W0)) {
                    atomicOr(d.flags, EF_OOB);
                    bad = true;
                }
                xref_check_fields(d, f, start, size, wk.x, wk.y, wk.z);
            }
        }
        const uint32_t c = i & 1u, nx = c ^ 1u;
        const bool act = (int32_t)i < (int32_t)size && !bad;
        const uint4 w = act ? words[i % SM_CHUNK][lane] : make_uint4(0, 0, 0, 0);
        uint32_t raw1, raw2;
        bool s1, s2;
        rd_state(w.y, raw1, s1);
        rd_state(w.z, raw2, s2);
        RaStep r = ra_eval_raw(w.w, w.x, w.y, w.z, raw1, raw2, s1, s2, life);
        uint64_t look = act ? sm_look(r) : 0ull;
        uint64_t gl = sm_or8_64(look);
        // owners reading an entry they own: settled in group order
        uint64_t or1 = 0, or2 = 0;
        if (k8 == 0 && act && own) {
            uint64_t m = own;
            while (m) {
                const uint32_t reg = (uint32_t)__builtin_ctzll(m);
                m &= m - 1;
                const int32_t v = Sv[c][pk][reg];
                if ((int32_t)w.y > 63 && v == (int32_t)w.y) or1 |= 1ull << reg;
                if ((int32_t)w.z > 63 && v == (int32_t)w.z) or2 |= 1ull << reg;
            }
        }
        uint64_t cand = __ballot(k8 == 0 && act && (or1 | or2) != 0);
        while (cand) {
            const uint32_t L0 = (uint32_t)__builtin_ctzll(cand);
            cand &= cand - 1;
            const uint32_t s = L0 >> 3, first = sm_rl(seg, L0) >> 3;
            uint64_t lower = 0;
            for (uint32_t t = first; t < s; ++t) lower |= sm_rl64(gl, 8u * t);
            const bool f1 = lane == L0 && (or1 & lower) != 0, f2 = lane == L0 && (or2 & lower) != 0;
            if (__ballot(f1 || f2)) {
                if (f1 || f2) { r = ra_eval_raw(w.w, w.x, w.y, w.z, raw1, raw2, s1 || f1, s2 || f2, life); look = sm_look(r); }
                gl = sm_or8_64(look);
            }
        }
        // commit: rd, then rs1, then rs2 (a later one wins); the lifetime
        if (act) {
            if ((int32_t)r.ord > 63) wr_state(r.ord, r.rdr);
            if ((int32_t)r.rs1 > 63) wr_state(r.rs1, r.r1r);
            if ((int32_t)r.rs2 > 63) wr_state(r.rs2, r.r2r);
            life = r.newlife;
        }
        // per program: marks on region 0's virtuals (entry `lane`), then region 0
        // of the next step.  Entry `lane` needs bit `lane` of the lookups of the
        // groups below its owner, of its owner's group and of the groups above it.
        for (uint32_t p = 0; p < np; ++p) {
            const uint32_t ps = prog_slots[p];
            const uint32_t s0 = ps & 0xFFu, s1e = ps >> 8;
            const int32_t v = Sv[c][p][lane];
            const int32_t o = So[c][p][lane];
            int32_t nv = v, no = o;
            bool lowb = false, ownb = false, highb = false, named = false;
            for (uint32_t t = s0; t < s1e; ++t) {
                const uint32_t L0 = 8u * t;
                const bool bt = ((sm_rl64(gl, L0) >> lane) & 1ull) != 0;
                lowb |= (int32_t)t < o && bt;
                highb |= (int32_t)t > o && bt;
                const bool at = sm_rl((uint32_t)act, L0) != 0;
                const uint32_t ot = sm_rl(w.x, L0), o1 = sm_rl(w.y, L0), o2 = sm_rl(w.z, L0);
                if ((int32_t)t == o) {
                    ownb = bt;
                    named = at && ((uint32_t)v == ot || (uint32_t)v == o1 || (uint32_t)v == o2);
                }
                // group t's lane-0 function acts on region 0: rd write, rs1 and rs2 frees
                if (at) {
                    const uint32_t rdt = sm_rl(r.rdr, L0), r1t = sm_rl(r.r1r, L0), r2t = sm_rl(r.r2r, L0);
                    if (ot && rdt == lane) { nv = (int32_t)ot; no = (int32_t)t; }
                    if (r1t != rdt && r1t == lane) { nv = -1; no = -1; }
                    if (r2t != rdt && r2t == lane) { nv = -1; no = -1; }
                }
            }
            if (o >= 0 && v >= 64 && (highb || (!named && (lowb || ownb)))) {
                const uint32_t x = (uint32_t)v - 64u - slot_start[o];
                if (x < slot_size[o]) st[slot_base[o] + x] |= 0x80u;
                else { atomicOr(d.flags, EF_XCONFLICT); d.cflag[waves[blockIdx.x].prog[p]] = 1; }
            }
            Sv[nx][p][lane] = nv;
            So[nx][p][lane] = (int8_t)no;
            for (uint32_t t = s0; t < s1e; ++t) {
                const uint64_t m = __ballot(no == (int32_t)t);
                if (lane == 8u * t) own = m;
            }
        }
    }
    __syncthreads();
    // final states to HBM (the bytes a function wrote or a mark set), the wave
    // storing each function's range with consecutive lanes
    for (uint32_t l = 0; l < 64; ++l) {
        const uint32_t bl = sm_rl(base, l), sl = sm_rl(start, l), zl = sm_rl(size, l);
        for (uint32_t x = lane; x < zl; x += 64u) {
            const uint32_t b = st[bl + x];
            if (b & 0xC0u) d.sym[sl + x] = (uint16_t)((b & 0x3Fu) | ((b >> 7) << 8));
        }
    }
}

// Lane-0 functions in step lockstep, one workgroup.  Region 0 of the
// reference's register_state holds the last lane-0 write per physical register;
// every lookup of any function at step i reads its step-start snapshot S and
// marks the virtual register found there as swapped.  A lookup by a lower group
// at step i is seen by the owner's operand read at step i; one by a higher group
// lands after the owner's own reset.  Intra-step effects between lane-0
// functions are settled by re-evaluation until stable (lower groups settle
// first, so this converges to the reference's in-order result).
constexpr int RA0_THREADS = 1024;

struct Ra0Rec { uint32_t ord, rs1, rs2; uint32_t regs; /* r1r | r2r<<8 | rdr<<16 */ };

DEV void ra0_flips(const int32_t *So, const uint32_t *fbits, uint32_t j, bool &f1, bool &f2) {
    f1 = false; f2 = false;
#pragma unroll 8
    for (int r = 0; r < 64; ++r)
        if (So[r] == (int32_t)j) { f1 |= (fbits[r] & 1u) != 0; f2 |= (fbits[r] & 2u) != 0; }
}

DEV RaStep ra0_eval(const Dev &d, uint32_t off, uint64_t life, bool f1, bool f2) {
    const uint32_t rs1 = d.w0_f4[off].y, rs2 = d.w0_f4[off].z;
    bool s1 = ((int32_t)rs1 > 63 && rs1 - 64u < d.W0 && sym_swapped(d, rs1 - 64u)) || f1;
    bool s2 = ((int32_t)rs2 > 63 && rs2 - 64u < d.W0 && sym_swapped(d, rs2 - 64u)) || f2;
    return ra_eval(d, off, life, s1, s2);
}


__global__ __launch_bounds__(RA0_THREADS) void k_ra_lane0_generic(Dev d, const Ra0Prog *progs, const uint32_t *gmin_all,
                                                         const uint32_t *gmax_all, uint64_t *life0, Ra0Rec *rec) {
    __shared__ int32_t Rv[64], Ro[64], Sv[64], So[64];
    __shared__ uint32_t Lmin[64], Lmax[64], fbits[64], wkey[64];
    __shared__ int changed;
    const uint32_t tid = threadIdx.x;
    const Ra0Prog pg = progs[blockIdx.x];
    const uint32_t L = pg.L, maxsize = pg.rows, fbase = pg.fbase;
    const uint32_t *gmin = gmin_all + pg.gbase, *gmax = gmax_all + pg.gbase;
    life0 += pg.poff; rec += pg.poff;
    if (tid < 64) { Rv[tid] = -1; Ro[tid] = -1; }
    for (uint32_t j = tid; j < L; j += RA0_THREADS) life0[j] = 0x1Full;
    __syncthreads();
    for (uint32_t i = 0; i < maxsize; ++i) {
        if (tid < 64) { Sv[tid] = Rv[tid]; So[tid] = Ro[tid]; fbits[tid] = 0; wkey[tid] = 0; }
        __syncthreads();
        for (uint32_t round = 0;; ++round) {
            if (tid < 64) { Lmin[tid] = gmin[(uint64_t)i * 64u + tid]; Lmax[tid] = gmax[(uint64_t)i * 64u + tid]; }
            if (tid == 0) changed = 0;
            __syncthreads();
            for (uint32_t j = tid; j < L; j += RA0_THREADS) {
                const uint32_t f = fbase + 8u * j;
                if (!((int32_t)i < (int32_t)d.fsizes[f])) continue;
                bool f1, f2;
                ra0_flips(So, fbits, j, f1, f2);
                RaStep r = ra0_eval(d, d.fstarts[f] + i, life0[j], f1, f2);
                if (r.swd) { atomicMin(&Lmin[r.rdr], j); atomicMax(&Lmax[r.rdr], j + 1u); }
                if (r.sw1) { atomicMin(&Lmin[r.r1r], j); atomicMax(&Lmax[r.r1r], j + 1u); }
                if (r.sw2) { atomicMin(&Lmin[r.r2r], j); atomicMax(&Lmax[r.r2r], j + 1u); }
            }
            __syncthreads();
            if (tid < 64) {
                uint32_t nb = 0;
                if (Sv[tid] >= 64 && So[tid] >= 0) {
                    const uint32_t o = (uint32_t)So[tid];
                    const uint32_t f = fbase + 8u * o;
                    if ((int32_t)i < (int32_t)d.fsizes[f] && Lmin[tid] < o) {
                        const uint32_t off = d.fstarts[f] + i;
                        if (d.w0_f4[off].y == (uint32_t)Sv[tid]) nb |= 1u;
                        if (d.w0_f4[off].z == (uint32_t)Sv[tid]) nb |= 2u;
                    }
                }
                if (nb != fbits[tid]) changed = 1;
                Lmin[tid] = nb;  // staged; fbits is updated after the barrier
            }
            __syncthreads();
            int again = changed;
            if (again && round > L + 1) { again = 0; if (tid == 0) atomicOr(d.flags, EF_OOB); }
            if (tid < 64 && again) fbits[tid] = Lmin[tid];
            __syncthreads();
            if (!again) break;
        }
        // commit: lifetime, records, region-0 write keys (group order: last writer wins)
        for (uint32_t j = tid; j < L; j += RA0_THREADS) {
            const uint32_t f = fbase + 8u * j;
            if (!((int32_t)i < (int32_t)d.fsizes[f])) continue;
            bool f1, f2;
            ra0_flips(So, fbits, j, f1, f2);
            RaStep r = ra0_eval(d, d.fstarts[f] + i, life0[j], f1, f2);
            life0[j] = r.newlife;
            rec[j] = Ra0Rec{r.ord, r.rs1, r.rs2, r.r1r | (r.r2r << 8) | (r.rdr << 16)};
            if (r.ord) atomicMax(&wkey[r.rdr], ((j + 1u) << 2) | 1u);
            if (r.r1r != r.rdr) atomicMax(&wkey[r.r1r], ((j + 1u) << 2) | 2u);
            if (r.r2r != r.rdr) atomicMax(&wkey[r.r2r], ((j + 1u) << 2) | 3u);
        }
        __syncthreads();
        for (uint32_t j = tid; j < L; j += RA0_THREADS) {
            const uint32_t f = fbase + 8u * j;
            if (!((int32_t)i < (int32_t)d.fsizes[f])) continue;
            const Ra0Rec q = rec[j];
            RaStep r;
            r.ord = q.ord; r.rs1 = q.rs1; r.rs2 = q.rs2;
            r.r1r = q.regs & 0xFFu; r.r2r = (q.regs >> 8) & 0xFFu; r.rdr = (q.regs >> 16) & 0xFFu;
            if (r.ord && wkey[r.rdr] == (((j + 1u) << 2) | 1u)) { Rv[r.rdr] = (int32_t)r.ord; Ro[r.rdr] = (int32_t)j; }
            if (r.r1r != r.rdr && wkey[r.r1r] == (((j + 1u) << 2) | 2u)) { Rv[r.r1r] = -1; Ro[r.r1r] = -1; }
            if (r.r2r != r.rdr && wkey[r.r2r] == (((j + 1u) << 2) | 3u)) { Rv[r.r2r] = -1; Ro[r.r2r] = -1; }
            ra_commit_syms(d, r);
        }
        __syncthreads();
        // marks on the snapshot's virtual registers
        if (tid < 64 && Sv[tid] >= 64 && Lmax[tid] != 0) {
            const uint32_t v = (uint32_t)Sv[tid];
            bool mark = true;
            if (So[tid] >= 0) {
                const uint32_t o = (uint32_t)So[tid];
                const uint32_t f = fbase + 8u * o;
                if ((int32_t)i < (int32_t)d.fsizes[f]) {
                    const Ra0Rec q = rec[o];
                    bool reset = (q.ord == v) || (q.rs1 == v) || (q.rs2 == v);
                    if (reset) mark = Lmax[tid] > o + 1u;
                }
            }
            if (mark) {
                if (v - 64u < d.W0) sym_mark(d, v - 64u);
                else atomicOr(d.flags, EF_OOB);
            }
        }
        __syncthreads();
    }
}


// Register-resident lockstep: each thread owns PER lane-0 functions (slots
// t = tid + q*blockDim; slot t runs group perm[t], the groups ordered by
// function size, largest first, so that the functions still running at a step
// fill the first waves and finished waves skip each phase) and keeps their
// state in registers: the fields of step i+1 and the packed operand state of
// step i.  The fields of step i sit in LDS (cur_*), where the snapshot checks
// read them too.  A function's virtual registers are read and written only by
// its own thread (marks by other groups reach the owner through LDS), so the
// barriers of a step order LDS alone.  Each step ends with the loads for the
// next: the operands' symbol states (waited for), the walk's row two steps
// ahead (by the loading waves) and the fields two steps ahead (waited for a
// step later).  Per-step LDS tables are double-buffered by step parity so that
// no barrier is needed between the end of a step and the start of the next.
//
// Snapshot owners: only a group owning an entry of the region-0 snapshot (a
// virtual register it wrote there) can see a swap made earlier in the same
// step, so only owners take part in the in-order settle; every other group's
// step result is final after its first evaluation.  Each function keeps the
// set of snapshot registers it may own (`own`, a superset refreshed against
// the snapshot every step), so owners find their flips and marks by register
// without any table.
//
// Of an instruction word, register allocation looks only at whether its major
// opcode is OP-FP (0x53): the other word tests of ra_eval_raw (0xC0007053,
// 0xD0007053) are OP-FP words themselves, so the lockstep keeps that bit alone.
struct Ra0Fn {
    uint32_t j, start, size;              // group, first word, word count
    uint64_t life;
    uint64_t own;                         // snapshot registers this function may own
    uint32_t ord, rs1, rs2;               // step i (kept, or reloaded from LDS in every phase)
    uint4 nw;                             // step i+1: rd, rs1, rs2, instruction (zeros past the end)
    uint32_t rr;                          // step i's result: r1r | r2r << 8 | rdr << 16
    uint32_t pk;                          // step i: raw1 | raw2 << 8 | s1 << 16 | s2 << 17 | flips << 18 |
                                          // won << 20 | OP-FP << 31 (won: 0x40 | the snapshot entry its
                                          // step i-1 commit won, 0 for none)
};

DEV RaStep ra0_step(const Ra0Fn &f) {
    const uint32_t fl = (f.pk >> 18) & 3u;
    return ra_eval_raw((f.pk >> 31) ? 0x53u : 0u, f.ord, f.rs1, f.rs2, f.pk & 0xFFu, (f.pk >> 8) & 0xFFu,
                       ((f.pk >> 16) & 1u) || (fl & 1u), ((f.pk >> 17) & 1u) || (fl & 2u), f.life);
}

// Word `step` of the function (zeros past its end).
DEV uint4 ra0_word(const Dev &d, const Ra0Fn &f, uint32_t step) {
    return (int32_t)step < (int32_t)f.size ? d.w0_f4[f.start + step] : make_uint4(0, 0, 0, 0);
}

DEV uint32_t ra0_fpbit(uint32_t ins) { return (ins & 0x7Fu) == 0x53u ? 1u : 0u; }

// sym entries of a word's operands (0 for physical operands).
DEV void ra0_gather(const Dev &d, uint32_t rs1, uint32_t rs2, uint32_t &g1, uint32_t &g2) {
    g1 = virt_ok(d, rs1) ? d.sym[rs1 - 64u] : 0u;
    g2 = virt_ok(d, rs2) ? d.sym[rs2 - 64u] : 0u;
}

// (bit 31 of g1, the OP-FP bit where the caller sets it, is kept)
DEV uint32_t ra0_pack(uint32_t g1, uint32_t g2) {
    return (g1 & 0x800000FFu) | ((g2 & 0xFFu) << 8) | (((g1 >> 8) & 1u) << 16) | (((g2 >> 8) & 1u) << 17);
}

// The commit of ra_commit_syms, storing only entries whose value changes:
// old1/old2 are the operands' entries before the commit; a destination entry is
// always stored unless an operand names the same entry (the later store wins).
DEV void ra0_commit(const Dev &d, const RaStep &r, uint32_t old1, uint32_t old2) {
    const bool vo = (int32_t)r.ord > 63, v1 = (int32_t)r.rs1 > 63, v2 = (int32_t)r.rs2 > 63;
    if ((vo && r.ord - 64u >= d.W0) || (v1 && r.rs1 - 64u >= d.W0) || (v2 && r.rs2 - 64u >= d.W0)) {
        atomicOr(d.flags, EF_OOB);
        return;
    }
    if (vo && !(v1 && r.rs1 == r.ord) && !(v2 && r.rs2 == r.ord)) d.sym[r.ord - 64u] = (uint16_t)r.rdr;
    if (v1 && !(v2 && r.rs2 == r.rs1) && r.r1r != old1) d.sym[r.rs1 - 64u] = (uint16_t)r.r1r;
    if (v2 && r.r2r != old2) d.sym[r.rs2 - 64u] = (uint16_t)r.r2r;
}

// The flip bits of the snapshot entries of `own` that group j owns, each
// entry's owner, register and flip bits read together.
DEV uint32_t ra0_owned_flips(uint64_t own, const int32_t *Sv, const int32_t *So, const uint32_t *flip, uint32_t j) {
    uint32_t f = 0;
    while (own) {
        const uint32_t r = (uint32_t)__builtin_ctzll(own);
        own &= own - 1;
        const int32_t so = So[r], sv = Sv[r];
        const uint32_t fl = flip[r];
        f |= fl & (0u - (uint32_t)((so == (int32_t)j) & (sv >= 64)));
    }
    return f;
}

constexpr uint32_t RA0_LDS_FNS = 8192;  // lane-0 functions whose current fields fit in LDS

// A function's fields of step i, from LDS where its thread does not keep them
// in registers across the step (KEEP: at most three functions per thread).
template <bool KEEP>
DEV void ra0_cur(Ra0Fn &f, const uint32_t *crd, const uint32_t *crs1, const uint32_t *crs2, uint32_t j) {
    if (!KEEP) { f.ord = crd[j]; f.rs1 = crs1[j]; f.rs2 = crs2[j]; }
}

DEV void ra0_lookups(uint32_t *mn, uint32_t *mx, const RaStep &r, uint32_t j) {
    if (r.swd) { atomicMin(&mn[r.rdr], j); atomicMax(&mx[r.rdr], j + 1u); }
    if (r.sw1) { atomicMin(&mn[r.r1r], j); atomicMax(&mx[r.r1r], j + 1u); }
    if (r.sw2) { atomicMin(&mn[r.r2r], j); atomicMax(&mx[r.r2r], j + 1u); }
}

DEV void ra0_keys(uint32_t *wkey, const RaStep &r, uint32_t j) {
    if (r.ord) atomicMax(&wkey[r.rdr], ((j + 1u) << 2) | 1u);
    if (r.r1r != r.rdr) atomicMax(&wkey[r.r1r], ((j + 1u) << 2) | 2u);
    if (r.r2r != r.rdr) atomicMax(&wkey[r.r2r], ((j + 1u) << 2) | 3u);
}

// Lookups of table row r by the walked functions (entry `tid`): min and max
// over the table's `copies` copies, `stride` apart.  Every load is issued
// before the first is used.
DEV void ra0_row(const uint32_t *gmin, const uint32_t *gmax, uint64_t stride, uint32_t copies, uint32_t r,
                 uint32_t tid, uint32_t &mn, uint32_t &mx) {
    const uint64_t e = (uint64_t)r * 64u + tid;
    uint32_t a[RA0_MAXCOPIES], b[RA0_MAXCOPIES];
#pragma unroll
    for (uint32_t c = 0; c < RA0_MAXCOPIES; ++c) {
        const uint64_t o = (c < copies ? c : 0u) * stride + e;
        a[c] = gmin[o]; b[c] = gmax[o];
    }
    mn = a[0]; mx = b[0];
#pragma unroll
    for (uint32_t c = 1; c < RA0_MAXCOPIES; ++c) { mn = min(mn, a[c]); mx = max(mx, b[c]); }
}

// Waits (every lane of the wave) until the walk has published row chunk c;
// `ok` turns false, for good, when a wait sees no progress for RA0_WAIT_TICKS.
DEV void ra0_wait(const Dev &d, const uint32_t *done, const uint32_t *need, uint32_t c, bool &ok) {
    if (!ok) return;
    const uint32_t want = need[c];
    const uint64_t t0 = wall_clock64();
    while (__hip_atomic_load(&done[c], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < want) {
        if (wall_clock64() - t0 > RA0_WAIT_TICKS) { ok = false; atomicOr(d.flags, EF_WAIT); return; }
        __builtin_amdgcn_s_sleep(2);
    }
}

// Folds row r of the walk's tables into the lockstep's lookup row of its step
// (mn / mx, cleared beforehand; min / max over the copies by LDS atomics):
// copy k is loaded by wave nwv-1-(k mod nwv), the last waves holding the
// shortest functions (idle through most steps), so that no wave with work
// waits for rows coming from HBM.  A loading wave waits for the row's chunk to
// be published (`first`: for the chunk of a first row).
DEV void ra0_stage(const Dev &d, const uint32_t *gmin, const uint32_t *gmax, uint64_t stride, uint32_t copies,
                   uint32_t r, const uint32_t *pdone, const uint32_t *pneed, bool first, bool &ok,
                   uint32_t *mn, uint32_t *mx, uint32_t wid, uint32_t nwv, uint32_t lane) {
    const uint32_t k0 = nwv - 1u - wid;
    if (k0 >= copies) return;
    if (pdone && (first || r % WALK_PUB == 0)) ra0_wait(d, pdone, pneed, r / WALK_PUB, ok);
    for (uint32_t k = k0; k < copies; k += nwv) {
        const uint64_t o = (uint64_t)k * stride + (uint64_t)r * 64u + lane;
        atomicMin(&mn[lane], gmin[o]);
        atomicMax(&mx[lane], gmax[o]);
    }
}

template <int PER, uint32_t NL>
__global__ __launch_bounds__(1024) void k_ra_lane0(Dev d, const Ra0Prog *progs, const uint32_t *gmin_all,
                                                   const uint32_t *gmax_all, const uint32_t *perm_all,
                                                   const uint32_t *done, const uint32_t *need, uint32_t copies,
                                                   uint64_t stride) {
    __shared__ uint32_t cur_rd[NL], cur_rs1[NL], cur_rs2[NL];
    __shared__ int32_t Sv[2][64], So[2][64];                  // region-0 snapshot of step i / i+1
    // lookups by all groups but the owners, and by the walk: the walk's row of
    // a step is folded in two steps ahead, into the table cleared after its last
    // read
    __shared__ uint32_t Lmin[2][64], Lmax[2][64];
    __shared__ uint32_t Omin[2][64], Omax[2][64];             // lookups by the (possible) owners
    __shared__ uint32_t wkey[2][64];
    __shared__ uint32_t flip[64];                             // flip bits per snapshot entry
    __shared__ uint32_t fmk[2];                               // snapshot entries with flip bits
    __shared__ uint32_t mk[2][2];                             // snapshot entries marked at step i
    __shared__ uint32_t keyed[2][2];                          // snapshot entries rewritten in the commit of step i
    __shared__ int changed;
    const uint32_t tid = threadIdx.x, nth = blockDim.x;
    const uint32_t wid = tid >> 6, nwv = nth >> 6, lane = tid & 63u;
    constexpr bool KEEP = PER <= 3;  // a function's fields of step i stay in registers
    // with one table copy per loading wave, a loading wave holds its copy's
    // row of step i+2 in registers, loaded a step before it is staged
    const uint32_t k0 = nwv - 1u - wid;
    const bool reg_rows = nwv >= copies;
    uint32_t rmn = 0xFFFFFFFFu, rmx = 0;
    // this workgroup's program: groups of 8 functions from fbase, L of them
    const Ra0Prog pg = progs[blockIdx.x];
    const uint32_t L = pg.L, maxsize = pg.rows, fbase = pg.fbase;  // steps: the table's rows
    const uint32_t *perm = perm_all + pg.poff;
    const uint32_t *gmin = gmin_all + pg.gbase, *gmax = gmax_all + pg.gbase;
    const uint32_t cb = nth > 64 ? 64u : 0u;  // first thread of the table-clearing threads
    // this program's publish counters (none: the walk has finished)
    const uint32_t *pdone = done && pg.cbase != RA0_NOPUB ? done + pg.cbase : nullptr;
    const uint32_t *pneed = pdone ? need + pg.cbase : nullptr;
    Ra0Fn fn[PER];
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const uint32_t t = tid + q * nth;
        Ra0Fn &f = fn[q];
        f.j = 0; f.start = 0; f.size = 0; f.life = 0x1Full; f.own = 0; f.rr = 0;
        if (t < L) {
            f.j = perm[t];
            if (f.j < L) { f.start = d.fstarts[fbase + 8u * f.j]; f.size = d.fsizes[fbase + 8u * f.j]; }
            else { atomicOr(d.flags, EF_OOB); f.j = 0; }
        }
        const uint4 w0 = ra0_word(d, f, 0);
        f.ord = w0.x; f.rs1 = w0.y; f.rs2 = w0.z;
        f.nw = ra0_word(d, f, 1);
        uint32_t g1 = 0, g2 = 0;
        if ((int32_t)f.size > 0) ra0_gather(d, f.rs1, f.rs2, g1, g2);
        f.pk = ra0_pack(g1, g2) | (ra0_fpbit(w0.w) << 31);
        if (t < L) {
            const bool act = (int32_t)f.size > 0;
            cur_rd[f.j] = act ? f.ord : 0u; cur_rs1[f.j] = act ? f.rs1 : 0u; cur_rs2[f.j] = act ? f.rs2 : 0u;
        }
    }
    bool wok = true;  // the walk's rows still arrive (waves that wait for them)
    if (tid < 64) {
        Sv[0][tid] = -1; So[0][tid] = -1; wkey[0][tid] = 0;
        if (tid < 2) keyed[1][tid] = 0;
        uint32_t m0 = 0xFFFFFFFFu, x0 = 0u;
        if (maxsize) {
            if (pdone) ra0_wait(d, pdone, pneed, 0, wok);
            ra0_row(gmin, gmax, stride, copies, 0, tid, m0, x0);
        }
        Lmin[0][tid] = m0; Lmax[0][tid] = x0;
        Lmin[1][tid] = 0xFFFFFFFFu; Lmax[1][tid] = 0u;
        Omin[0][tid] = Omin[1][tid] = 0xFFFFFFFFu; Omax[0][tid] = Omax[1][tid] = 0;
    }
    __syncthreads();
    if (maxsize > 1) ra0_stage(d, gmin, gmax, stride, copies, 1, pdone, pneed, true, wok, Lmin[1], Lmax[1], wid, nwv, lane);
    if (reg_rows && k0 < copies && maxsize > 2) {
        if (pdone && 2u % WALK_PUB == 0) ra0_wait(d, pdone, pneed, 2u / WALK_PUB, wok);
        const uint64_t o = (uint64_t)k0 * stride + 2u * 64u + lane;
        rmn = gmin[o]; rmx = gmax[o];
    }
    __syncthreads();
    uint32_t fb = 0;   // wave 0: flip bits of snapshot register tid
    uint32_t om = 0xFFFFFFFFu, ox = 0;  // wave 0: owners' lookups of register tid at step i
    for (uint32_t i = 0; i < maxsize; ++i) {
        const uint32_t c = i & 1u, n = c ^ 1u;
        // P1: every group evaluates with no flips; a group that cannot own a
        // snapshot entry finalizes its step, a possible owner keeps its
        // lookups apart
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            Ra0Fn &f = fn[q];
            const uint32_t j = f.j;
            if ((int32_t)i < (int32_t)f.size) {
                ra0_cur<KEEP>(f, cur_rd, cur_rs1, cur_rs2, j);
                xref_check_fields(d, fbase + 8u * j, f.start, f.size, f.ord, f.rs1, f.rs2);
                const RaStep r = ra0_step(f);
                if (f.own) {
                    ra0_lookups(Omin[c], Omax[c], r, j);
                    f.rr = r.r1r | (r.r2r << 8) | (r.rdr << 16);  // final unless a settle round re-evaluates
                } else {
                    ra0_lookups(Lmin[c], Lmax[c], r, j);
                    f.rr = r.r1r | (r.r2r << 8) | (r.rdr << 16);
                    f.life = r.newlife;
                    ra0_keys(wkey[c], r, j);
                }
            }
        }
        lds_barrier();
        // P2: an owner of a snapshot entry reading it as an operand sees the
        // swap when a lower group looked the register up in this step
        if (tid < 64) {
            om = Omin[c][tid]; ox = Omax[c][tid];
            Omin[c][tid] = 0xFFFFFFFFu; Omax[c][tid] = 0;  // cleared for the settle rounds and step i+2
            uint32_t nb = 0;
            const int32_t o = So[c][tid], v = Sv[c][tid];
            const uint32_t lm = Lmin[c][tid];  // read with the entry, not behind its test
            if ((v >= 64) & (o >= 0) & (min(lm, om) < (uint32_t)o)) {
                if (cur_rs1[o] == (uint32_t)v) nb |= 1u;
                if (cur_rs2[o] == (uint32_t)v) nb |= 2u;
            }
            flip[tid] = nb;
            fb = nb;
            const uint64_t any = __ballot(nb != 0);
            if (tid == 0) { changed = any != 0; fmk[0] = (uint32_t)any; fmk[1] = (uint32_t)(any >> 32); }
        }
        if (tid >= cb)
            for (uint32_t k = tid - cb; k < 64; k += nth - cb) wkey[n][k] = 0;
        lds_barrier();
        // (the flipped entries read with the flag, for the first round)
        const uint64_t fm0 = ((uint64_t)fmk[1] << 32) | fmk[0];
        if (changed) {
            // settle intra-step flips among the owners: re-evaluate until the
            // flip bits are stable (lower groups settle first, so this reaches
            // the in-order result); O[c] was cleared when wave 0 last read it
            for (uint32_t round = 0;; ++round) {
                // only entries with flip bits can give an owner flips
                const uint64_t fm = round == 0 ? fm0 : ((uint64_t)fmk[1] << 32) | fmk[0];
#pragma unroll
                for (int q = 0; q < PER; ++q) {
                    const uint32_t j = fn[q].j;
                    if (fn[q].own && (int32_t)i < (int32_t)fn[q].size) {
                        fn[q].pk = (fn[q].pk & ~(3u << 18)) |
                                   (ra0_owned_flips(fn[q].own & fm, Sv[c], So[c], flip, j) << 18);
                        ra0_cur<KEEP>(fn[q], cur_rd, cur_rs1, cur_rs2, j);
                        const RaStep r = ra0_step(fn[q]);
                        ra0_lookups(Omin[c], Omax[c], r, j);
                        fn[q].rr = r.r1r | (r.r2r << 8) | (r.rdr << 16);
                    }
                }
                lds_barrier();
                if (tid < 64) {
                    om = Omin[c][tid]; ox = Omax[c][tid];
                    Omin[c][tid] = 0xFFFFFFFFu; Omax[c][tid] = 0;
                    uint32_t nb = 0;
                    const int32_t o = So[c][tid], v = Sv[c][tid];
                    const uint32_t lm = Lmin[c][tid];
                    if ((v >= 64) & (o >= 0) & (min(lm, om) < (uint32_t)o)) {
                        if (cur_rs1[o] == (uint32_t)v) nb |= 1u;
                        if (cur_rs2[o] == (uint32_t)v) nb |= 2u;
                    }
                    flip[tid] = nb;
                    const uint64_t diff = __ballot(nb != fb), any = __ballot(nb != 0);
                    fb = nb;
                    if (tid == 0) { changed = diff != 0; fmk[0] = (uint32_t)any; fmk[1] = (uint32_t)(any >> 32); }
                }
                lds_barrier();
                if (!changed) break;
                if (round > L + 1) { if (tid == 0) atomicOr(d.flags, EF_OOB); break; }
            }
        }
        // P3: possible owners finalize, from the registers of their last
        // evaluation (P1's, or the last settle round's: the same inputs, so the
        // same result) and the lifetime update of ra_eval_raw; marks on the
        // snapshot's virtual registers; the snapshot of step i+1 starts as a copy
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const uint32_t j = fn[q].j;
            if (fn[q].own && (int32_t)i < (int32_t)fn[q].size) {
                const uint32_t rr = fn[q].rr;
                RaStep r;
                r.ord = KEEP ? fn[q].ord : cur_rd[j];
                r.r1r = rr & 0xFFu; r.r2r = (rr >> 8) & 0xFFu; r.rdr = (rr >> 16) & 0xFFu;
                const uint64_t clr = ((1ull << r.r1r) | (1ull << r.r2r)) & ~1ull;
                fn[q].life = (fn[q].life & ~clr) | (1ull << r.rdr);
                ra0_keys(wkey[c], r, j);
            }
        }
        if (tid < 64) {
            const int32_t v = Sv[c][tid], o = So[c][tid];
            Sv[n][tid] = v; So[n][tid] = o;
            const uint32_t mx = max(Lmax[c][tid], ox);
            Lmin[c][tid] = 0xFFFFFFFFu; Lmax[c][tid] = 0u;  // last read: cleared for the walk's row of step i+2
            bool mark = false;
            if (v >= 64 && o >= 0 && mx != 0) {
                mark = true;
                const uint32_t uv = (uint32_t)v;
                const uint32_t a = cur_rd[o], b = cur_rs1[o], e = cur_rs2[o];  // read together
                if ((a == uv) | (b == uv) | (e == uv)) mark = mx > (uint32_t)o + 1u;
            }
            const uint64_t m = __ballot(mark);
            if (tid == 0) { mk[c][0] = (uint32_t)m; mk[c][1] = (uint32_t)(m >> 32); }
        }
        lds_barrier();
        // P4: snapshot of step i+1, commit of the symbol state, marks; then the
        // fields of step i+1 into LDS and one batch of loads for the next step.
        // The fields of step i+1, loaded a step ago, are in: one wait for every
        // vector load (vmcnt 0) here leaves none outstanding across the
        // branches below, so the gathers issue back to back.
        __builtin_amdgcn_s_waitcnt(0x0F70);
        const uint64_t marked = ((uint64_t)mk[c][1] << 32) | mk[c][0];
        // entries whose region-0 state the commit of step i-1 rewrote; every
        // other entry a group owned at the start of step i-1 it still owns
        const uint64_t kprev = ((uint64_t)keyed[n][1] << 32) | keyed[n][0];
        if (tid < 64) {
            const uint64_t kb = __ballot(wkey[c][tid] != 0);
            if (tid == 0) { keyed[c][0] = (uint32_t)kb; keyed[c][1] = (uint32_t)(kb >> 32); }
        }
        uint32_t wb[PER];  // the entry each function's commit wins: 0x40 | register, 0 for none
        // the operand states of step i+1 (sym entries), loaded right after each
        // function's commit (the commit's stores come first in program order);
        // a mark of one of them that follows, and the OP-FP bit, are added at
        // packing (pm: bit 0 / 1 operand marked, bit 2 OP-FP)
        uint32_t g1[PER], g2[PER], pm[PER];
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const uint32_t t = tid + q * nth;
            Ra0Fn &f = fn[q];
            const uint32_t j = f.j;
            wb[q] = 0; g1[q] = 0; g2[q] = 0; pm[q] = 0;
            if (t >= L) continue;
            uint64_t won = 0;
            if ((int32_t)i < (int32_t)f.size) {
                ra0_cur<KEEP>(f, cur_rd, cur_rs1, cur_rs2, j);
                const uint32_t rr = f.rr;
                RaStep r;
                r.ord = f.ord; r.rs1 = f.rs1; r.rs2 = f.rs2;
                r.r1r = rr & 0xFFu; r.r2r = (rr >> 8) & 0xFFu; r.rdr = (rr >> 16) & 0xFFu;
                // the three keys read together (registers are below 64), then compared
                const uint32_t kd = wkey[c][r.rdr & 63u], k1 = wkey[c][r.r1r & 63u], k2 = wkey[c][r.r2r & 63u];
                const bool wd = (r.ord != 0u) & (kd == (((j + 1u) << 2) | 1u));
                const bool w1 = (r.r1r != r.rdr) & (k1 == (((j + 1u) << 2) | 2u));
                const bool w2 = (r.r2r != r.rdr) & (k2 == (((j + 1u) << 2) | 3u));
                if (wd) {
                    Sv[n][r.rdr] = (int32_t)r.ord; So[n][r.rdr] = (int32_t)j;
                    if ((int32_t)r.ord >= 64) { won = 1ull << r.rdr; wb[q] = 0x40u | r.rdr; }
                }
                if (w1) { Sv[n][r.r1r] = -1; So[n][r.r1r] = -1; }
                if (w2) { Sv[n][r.r2r] = -1; So[n][r.r2r] = -1; }
                ra0_commit(d, r, (f.pk & 0xFFu) | (((f.pk >> 16) & 1u) << 8), ((f.pk >> 8) & 0xFFu) | (((f.pk >> 17) & 1u) << 8));
                const bool nx = (int32_t)(i + 1) < (int32_t)f.size;
                cur_rd[j] = nx ? f.nw.x : 0u; cur_rs1[j] = nx ? f.nw.y : 0u; cur_rs2[j] = nx ? f.nw.z : 0u;
                if (KEEP) { f.ord = nx ? f.nw.x : 0u; f.rs1 = nx ? f.nw.y : 0u; f.rs2 = nx ? f.nw.z : 0u; }
            }
            if ((int32_t)(i + 1) < (int32_t)f.size) {
                ra0_gather(d, f.nw.y, f.nw.z, g1[q], g2[q]);
                pm[q] = ra0_fpbit(f.nw.w) << 2;  // (not ORed into g1 here: that would wait for the load)
            }
            if (f.own) {
                // the entries it owns now: those it owned that the commit of step
                // i-1 left alone, and the one it won in that commit (a won entry
                // is keyed, and a keyed entry is won by one group alone)
                const uint32_t pw = (f.pk >> 20) & 0x7Fu;
                const uint64_t wprev = (pw & 0x40u) ? 1ull << (pw & 63u) : 0ull;
                const uint64_t o = (f.own & ~kprev) | (f.own & kprev & wprev);
                uint64_t m = o & marked;
                while (m) {
                    const uint32_t reg = (uint32_t)__builtin_ctzll(m);
                    m &= m - 1;
                    const uint32_t v = (uint32_t)Sv[c][reg];
                    if (v - 64u < d.W0) sym_mark(d, v - 64u);
                    else atomicOr(d.flags, EF_OOB);
                    pm[q] |= (v == f.nw.y ? 1u : 0u) | (v == f.nw.z ? 2u : 0u);
                }
                f.own = o;
            }
            f.own |= won;
        }
        if (reg_rows) {
            if (k0 < copies && i + 2 < maxsize) { atomicMin(&Lmin[c][lane], rmn); atomicMax(&Lmax[c][lane], rmx); }
        } else if (i + 2 < maxsize) {
            ra0_stage(d, gmin, gmax, stride, copies, i + 2, pdone, pneed, false, wok, Lmin[c], Lmax[c], wid, nwv, lane);
        }
        // (packed for every slot, finished ones too: no gather stays pending
        // on any path into the next step)
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            fn[q].pk = ra0_pack(g1[q] | ((pm[q] & 1u) << 8) | ((pm[q] >> 2) << 31), g2[q] | ((pm[q] & 2u) << 7)) |
                       (wb[q] << 20);
            asm volatile("" : "+v"(fn[q].pk) : : "memory");  // packed here, before the loads below
        }
        // the walk's row of step i+3, staged at the end of step i+1, and the
        // fields of step i+2, first read at the end of step i+1: issued after
        // every load this step waits for (loads complete in order), so that
        // their latency overlaps step i+1
        if (reg_rows && k0 < copies && i + 3 < maxsize) {
            if (pdone && (i + 3) % WALK_PUB == 0) ra0_wait(d, pdone, pneed, (i + 3) / WALK_PUB, wok);
            const uint64_t o = (uint64_t)k0 * stride + (uint64_t)(i + 3) * 64u + lane;
            rmn = gmin[o]; rmx = gmax[o];
        }
#pragma unroll
        for (int q = 0; q < PER; ++q)
            if ((int32_t)(i + 1) < (int32_t)fn[q].size) fn[q].nw = ra0_word(d, fn[q], i + 2);
    }
}

// Exact fallback for programs in which two functions name the same virtual
// register (see xref_note): the reference's allocation walk as written
// (rv_generator_avx.cpp:1041-1418 at one thread), one wave: steps in order,
// groups of 8 functions in order within a step, the 8 functions of a group on
// lanes 0-7.  Each group body reads its operands' states before any of its
// writes; its writes follow the reference's order (marks, lifetimes,
// register_state, then the symbol entries of rd, rs1, rs2, lane by lane).
// register_state keeps 8 regions of 64 (region = lane); lookups read region 0
// of the step's snapshot.
__global__ __launch_bounds__(64) void k_ra_serial(Dev d, const Ra0Prog *progs, const uint32_t *plist,
                                                  uint64_t *life_all) {
    __shared__ int32_t reg[8][64];
    __shared__ int32_t snap[64];
    const uint32_t lane = threadIdx.x;
    const uint32_t prg = plist[blockIdx.x];
    const Ra0Prog pg = progs[prg];
    const uint32_t maxsize = pg.maxsize, fbase = pg.fbase, nf = pg.nf;
    uint64_t *life = life_all + fbase;
    for (uint32_t k = lane; k < 8 * 64; k += 64) reg[k / 64][k % 64] = -1;
    for (uint32_t f = lane; f < nf; f += 64) life[f] = 0x1Full;
    for (uint32_t e = d.pw[prg] + lane; e < d.pw[prg + 1]; e += 64) d.sym[e] = 0;  // the program's states restart
    __syncthreads();
    const uint32_t G = (nf + 7u) / 8u;
    for (uint32_t i = 0; i < maxsize; ++i) {
        snap[lane] = reg[0][lane];
        __syncthreads();
        for (uint32_t g = 0; g < G; ++g) {
            const uint32_t f = 8u * g + lane;  // within the program
            const bool nc = lane < 8 && f < nf && (int32_t)i < (int32_t)d.fsizes[fbase + f];
            uint4 w = make_uint4(0, 0, 0, 0);
            uint64_t lt = 0;
            if (nc) { w = d.w0_f4[d.fstarts[fbase + f] + i]; lt = life[f]; }
            const uint32_t raw1 = nc && virt_ok(d, w.y) ? d.sym[w.y - 64u] : 0u;
            const uint32_t raw2 = nc && virt_ok(d, w.z) ? d.sym[w.z - 64u] : 0u;
            const RaStep r = ra_eval_raw(w.w, w.x, w.y, w.z, raw1 & 0xFFu, raw2 & 0xFFu, (raw1 >> 8) != 0,
                                         (raw2 >> 8) != 0, lt);
            __syncthreads();  // every lane has read before any lane writes
            if (nc && r.swd) { const int32_t x = snap[r.rdr] - 64; if (x >= 0) sym_mark(d, (uint32_t)x); }
            if (nc && r.sw1) { const int32_t x = snap[r.r1r] - 64; if (x >= 0) sym_mark(d, (uint32_t)x); }
            if (nc && r.sw2) { const int32_t x = snap[r.r2r] - 64; if (x >= 0) sym_mark(d, (uint32_t)x); }
            if (nc) {
                life[f] = r.newlife;
                if (r.ord) reg[lane][r.rdr] = (int32_t)r.ord;
                if (r.r1r != r.rdr) reg[lane][r.r1r] = -1;
                if (r.r2r != r.rdr) reg[lane][r.r2r] = -1;
            }
            const bool oob = nc && (((int32_t)r.ord > 63 && r.ord - 64u >= d.W0) ||
                                    ((int32_t)r.rs1 > 63 && r.rs1 - 64u >= d.W0) ||
                                    ((int32_t)r.rs2 > 63 && r.rs2 - 64u >= d.W0));
            if (oob) atomicOr(d.flags, EF_OOB);
            const bool ok = nc && !oob;
            for (uint32_t k = 0; k < 8; ++k) { if (ok && lane == k && (int32_t)r.ord > 63) d.sym[r.ord - 64u] = (uint16_t)r.rdr; __syncthreads(); }
            for (uint32_t k = 0; k < 8; ++k) { if (ok && lane == k && (int32_t)r.rs1 > 63) d.sym[r.rs1 - 64u] = (uint16_t)r.r1r; __syncthreads(); }
            for (uint32_t k = 0; k < 8; ++k) { if (ok && lane == k && (int32_t)r.rs2 > 63) d.sym[r.rs2 - 64u] = (uint16_t)r.r2r; __syncthreads(); }
        }
    }
}

// Every word naming a register that a function outside the register's own
// function also names must belong to that same function (see xref_note).
__global__ __launch_bounds__(256) void k_xref_check(Dev d) {
    uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i >= d.W0) return;
    const uint4 w = d.w0_f4[i];
    const uint32_t vs[3] = {w.x, w.y, w.z};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const uint32_t v = vs[k];
        if (!((int32_t)v > 63) || v - 64u >= d.W0) continue;
        const uint32_t t = d.xref[v - 64u];
        if (!t) continue;
        uint32_t lo = 0, hi = d.F;  // function of word i: last start <= i
        while (hi - lo > 1) { uint32_t m = (lo + hi) / 2; if (d.fstarts[m] <= i) lo = m; else hi = m; }
        if (lo + 1u != t) { atomicOr(d.flags, EF_XCONFLICT); d.cflag[prog_of(d.pf, d.P, lo)] = 1; }
    }
}

// Instruction counts after allocation (:1473-1495): used words plus one load
// per swapped operand and one store for a swapped destination.
struct RaCountOp {
    Dev d;
    DEV uint32_t sw(uint32_t r) const { return ((int32_t)r > 63 && r - 64u < d.W0 && sym_swapped(d, r - 64u)) ? 1u : 0u; }
    DEV uint32_t value(uint64_t i) const {
        if (!d.usedi[i]) return 0;
        const uint4 w = d.w0_f4[i];  // rd, rs1, rs2, instruction
        return 1u + sw(w.y) + sw(w.z) + sw(w.x);
    }
    DEV void emit(uint64_t i, uint32_t excl, uint32_t) const { d.ra_off[i] = excl; }
};

// spill_offsets (:1450-1469): each function's words copy the flag at its start.
DEV uint32_t spill_of(const Dev &d, uint32_t v) {
    uint32_t lo = 0, hi = d.F;  // last start <= v
    while (hi - lo > 1) { uint32_t m = (lo + hi) / 2; if (d.fstarts[m] <= v) lo = m; else hi = m; }
    return sym_swapped(d, d.fstarts[lo]) ? 1u : 0u;
}

// Emit (:1552-1697).
__global__ __launch_bounds__(256) void k_ra_emit(Dev d) {
    uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    // the function of the wave's first word, looked up once for the wave: a
    // word's operands are mostly registers of its own function, and the
    // wave's words mostly of one function (spill_of of any other register)
    const uint64_t i0 = blockIdx.x * (uint64_t)blockDim.x + (threadIdx.x & ~63u);
    uint32_t fi = 0, fs = 0, fe = 0;
    if (i0 < d.W0 && d.F) {
        uint32_t lo = 0, hi = d.F;  // last start <= i0
        while (hi - lo > 1) { const uint32_t m = (lo + hi) / 2; if (d.fstarts[m] <= (uint32_t)i0) lo = m; else hi = m; }
        fi = lo;
        fs = d.fstarts[lo];
        fe = lo + 1 < d.F ? d.fstarts[lo + 1] : 0xFFFFFFFFu;
    }
    const uint32_t fsp = fe > fs && sym_swapped(d, fs) ? 1u : 0u;
    // the function of this lane's own word, stepping on from the wave's (a
    // wave's words span few functions), for a lane past the wave's function
    uint32_t ls = fs, le = fe, lsp = fsp;
    if (i < d.W0 && d.F && (uint32_t)i >= fe) {
        uint32_t li = fi + 1;
        while (li + 1 < d.F && d.fstarts[li + 1] <= (uint32_t)i) ++li;
        ls = d.fstarts[li];
        le = li + 1 < d.F ? d.fstarts[li + 1] : 0xFFFFFFFFu;
        lsp = sym_swapped(d, ls) ? 1u : 0u;
    }
    auto spill = [&](uint32_t v) -> uint32_t {
        return v >= ls && v < le ? lsp : v >= fs && v < fe ? fsp : spill_of(d, v);
    };
    // slots no word's range covers: the pad at W1 and, when the last word is
    // unused, the reference's last slot W1 - 1 (W1 = last offset + 1)
    auto zero = [&](uint32_t o) {
        d.w1_instr[o] = 0; d.w1_rd[o] = 0; d.w1_rs1[o] = 0; d.w1_rs2[o] = 0; d.w1_jt[o] = 0;
    };
    if (i == 0) zero(d.W1);
    if (i + 1 == d.W0 && !d.usedi[i] && d.ra_off[i] < d.W1) zero(d.ra_off[i]);
    if (i >= d.W0 || !d.usedi[i]) return;
    const int32_t base = (int32_t)d.ra_off[i];
    const uint4 wf = d.w0_f4[i];  // rd, rs1, rs2, instruction
    const uint32_t ins = wf.w;
    const uint32_t r1 = wf.y, r2 = wf.z, rdv = wf.x;
    int32_t l1 = -1, l2 = -1, rdo = -1;
    uint32_t a1 = r1, a2 = r2, ad = rdv, st1 = 0, st2 = 0, std_ = 0;
    if (((int32_t)r1 > 63 && r1 - 64u >= d.W0) || ((int32_t)r2 > 63 && r2 - 64u >= d.W0) ||
        ((int32_t)rdv > 63 && rdv - 64u >= d.W0)) { atomicOr(d.flags, EF_OOB); return; }
    if ((int32_t)r1 > 63) {
        const uint32_t e = d.sym[r1 - 64u];
        const bool sw = (e >> 8) != 0;
        l1 = sw ? base : -1;
        st1 = spill(r1 - 64u);
        a1 = sw ? (((ins == 0xC0007053u) || (ins & 0x7Fu) == 0x53u) ? 37u : 5u) : (e & 0xFFu);
    }
    if ((int32_t)r2 > 63) {
        const uint32_t e = d.sym[r2 - 64u];
        const bool sw = (e >> 8) != 0;
        l2 = sw ? base : -1;  // the reference tests its own, still -1, offset (D6)
        st2 = spill(r2 - 64u);
        a2 = sw ? (((ins & 0x7Fu) == 0x53u) ? 38u : 6u) : (e & 0xFFu);
    }
    const int32_t mo = base + (l1 > 0 ? 1 : 0) + (l2 > 0 ? 1 : 0);
    if ((int32_t)rdv > 63) {
        const uint32_t e = d.sym[rdv - 64u];
        const bool sw = (e >> 8) != 0;
        rdo = sw ? mo + 1 : -1;
        std_ = spill(rdv - 64u);
        ad = e & 0xFFu;
    }
    // The word's slots are [base, base + n) (its count in RaCountOp); the
    // reference's stores land in them as below, later ones over earlier ones,
    // and a slot none of them reaches stays zero.  Each slot of the range is
    // written once, all five fields, so the arrays need no clearing.
    const uint32_t n = 1u + (l1 >= 0 ? 1u : 0u) + (l2 >= 0 ? 1u : 0u) + (rdo >= 0 ? 1u : 0u);
    const uint32_t end = (uint32_t)base + n;
    if ((rdo >= 0 && (uint32_t)rdo >= end) || (uint32_t)mo >= end || end > d.W1) { atomicOr(d.flags, EF_OOB); return; }
    uint32_t wi[4] = {0, 0, 0, 0}, wd[4] = {0, 0, 0, 0}, w1[4] = {0, 0, 0, 0}, w2[4] = {0, 0, 0, 0}, wj[4] = {0, 0, 0, 0};
    auto at = [&](int32_t slot) -> uint32_t { return (uint32_t)(slot - base); };  // < n for every slot stored
    if (l1 >= 0) { wi[at(l1)] = (((st1 - 1u) * 4u) << 20) | 0x2003u; wd[at(l1)] = 5; }
    if (l2 >= 0) { wi[at(l2)] = (((st2 - 1u) * 4u) << 20) | 0x2003u; wd[at(l2)] = 6; }
    {
        const uint32_t k = at(mo);
        wi[k] = ins; wd[k] = ad; w1[k] = a1; w2[k] = a2;
        wj[k] = d.ra_off[d.w0_jt[i] < d.W0 ? d.w0_jt[i] : 0];
        if (d.w0_jt[i] >= d.W0) atomicOr(d.flags, EF_OOB);
    }
    if (rdo >= 0) {
        const uint32_t k = at(rdo);
        uint32_t imm = (std_ - 1u) * (uint32_t)-4;
        wi[k] = 0x2023u | ((imm & 0x1Fu) << 7) | ((imm & 0xFEu) << 19);
        wd[k] = 0; w1[k] = ad; w2[k] = 0; wj[k] = 0;
    }
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k)
        if (k < n) {
            const uint32_t o = (uint32_t)base + k;
            d.w1_instr[o] = wi[k]; d.w1_rd[o] = wd[k]; d.w1_rs1[o] = w1[k]; d.w1_rs2[o] = w2[k]; d.w1_jt[o] = wj[k];
        }
}

// fix_func_tab (:1816-1837) against an offset array; pend[p] is the end of
// program p's range of that array (a function ending there takes the last
// offset + 1, as the reference does at the end of its array).
__global__ void k_fixtab(Dev d, const uint32_t *off, const uint32_t *pend) {
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= d.F) return;
    const uint32_t len = pend[prog_of(d.pf, d.P, k)];
    uint32_t s = d.fstarts[k], e = s + d.fsizes[k];
    uint32_t so = off[s];
    uint32_t eo = (int32_t)e >= (int32_t)len ? off[len - 1] + 1u : off[e];
    d.fstarts[k] = so; d.fends[k] = eo; d.fsizes[k] = eo - so;
}

// Per-program bounds after regalloc (the reference's W1 = last offset + 1,
// :1542) and after fix_jumps (last offset + size of the last word, :1891):
// b1[p] = start of program p's allocated words, e1[p] = its W1 end; b2/e2 the
// same for the final words.
__global__ void k_prog_w1(Dev d, uint32_t *b1, uint32_t *e1) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= d.P) return;
    b1[p] = d.ra_off[d.pw[p]];
    e1[p] = d.ra_off[d.pw[p + 1] - 1] + 1u;
}

// Frame words (:1714-1811): lui s0, 0 / addi s0, s0, 8 after the prologue
// stores and before the epilogue; no callee-saved preservation (never written).
__global__ void k_ra_frame(Dev d) {
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= d.F) return;
    const uint32_t upper = 0x437u, lower = (8u << 20) | 0x40413u;
    uint32_t s = d.fstarts[k], e = s + d.fsizes[k];
    uint32_t pos[4] = {s + 2, s + 3, e - 6, e - 5};
    for (int q = 0; q < 4; ++q) {
        uint32_t x = pos[q];
        if (x >= d.W1) { atomicOr(d.flags, EF_OOB); continue; }
        d.w1_instr[x] = (q & 1) ? lower : upper;
        d.w1_rd[x] = 0; d.w1_rs1[x] = 0; d.w1_rs2[x] = 0; d.w1_jt[x] = 0;
    }
}

// ================================================================ fix_jumps
DEV uint32_t fj_size(uint32_t w) {
    bool jmp = ((w & 0x7Fu) == 0x67u) && w != 0x8067u;
    return jmp ? 2u : (w != 0 ? 1u : 0u);
}

__global__ void k_prog_w2(Dev d, const uint32_t *b1, const uint32_t *e1, uint32_t *b2, uint32_t *e2) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= d.P) return;
    b2[p] = d.fj_off[b1[p]];
    e2[p] = d.fj_off[e1[p] - 1] + fj_size(d.w1_instr[e1[p] - 1]);
}

struct FjSizeOp {
    Dev d;
    DEV uint32_t value(uint64_t i) const { return fj_size(d.w1_instr[i]); }
    DEV void emit(uint64_t i, uint32_t excl, uint32_t) const { d.fj_off[i] = excl; }
};

// Move + jump/branch encoding + postprocess in one pass (:1895-1993, :1995-2011).
__global__ __launch_bounds__(256) void k_fj_move(Dev d) {
    uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i >= d.W1) return;
    const uint32_t w = d.w1_instr[i];
    if (w == 0) return;
    const uint32_t k = d.fj_off[i];
    const bool jmp = ((w & 0x7Fu) == 0x67u) && w != 0x8067u;
    const bool br = (w & 0x7Fu) == 0x63u;
    const uint32_t pp = ((d.w1_rd[i] & 31u) << 7) | ((d.w1_rs1[i] & 31u) << 15) | ((d.w1_rs2[i] & 31u) << 20);
    if (jmp || br) {
        const uint32_t jt1 = d.w1_jt[i];
        if (jt1 >= d.W1) atomicOr(d.flags, EF_OOB);
        const uint32_t tgt = 4u * d.fj_off[jt1 < d.W1 ? jt1 : 0];
        const uint32_t delta = tgt - k * 4u;
        if (jmp) {
            const uint32_t upper = (delta & 0xFFFFF000u) >> 12, lower = delta & 0xFFFu, sign = (delta >> 11) & 1u;
            d.words[k] = ((upper + sign) << 12) | 0x97u;
            d.words[k + 1] = w | (lower << 20) | (1u << 15);
            return;
        }
        const uint32_t sign = (delta >> 12) & 1u, b11 = (delta >> 11) & 1u;
        const uint32_t b105 = (delta >> 5) & 0x1Fu, b14 = (delta >> 1) & 0xFu;
        d.words[k] = w | (sign << 31) | (b11 << 7) | (b105 << 25) | (b14 << 8) | pp;
        return;
    }
    d.words[k] = w | pp;
}

// ============================================================== text output
// rv_generator::to_asm (src/codegen/rv_generator.cpp:114-120): one line per
// word, rvdisasm::instruction(word) (src/disassembler.cpp:453-462) and '\n':
// the mnemonic (instr_name, :83-222), a space, the operands (format_args,
// :224-407, no colour, address 0).  The reference's switches continue with the
// next case in source order wherever an inner switch finds no match; asm_name
// walks its cases as such a chain, so a branch with funct3 2 is named "lw" and
// a SYSTEM word other than ecall/ebreak "flw", exactly as the reference prints
// them.  Lines are at most 48 bytes.  Each word is formatted by one thread: a
// scan of the line lengths gives every line's offset, and the rescan writes
// each thread's 16 consecutive lines (4-byte stores inside its own byte range,
// byte stores at its two ends).
enum : uint8_t {
    M_LUI, M_AUIPC, M_JAL, M_JALR, M_BEQ, M_BNE, M_BLT, M_BGE, M_BLTU, M_BGEU, M_LB, M_LH, M_LW, M_LBU, M_LHU,
    M_SB, M_SH, M_SW, M_ADDI, M_SLLI, M_SLTI, M_SLTIU, M_XORI, M_SRLI, M_SRAI, M_ORI, M_ANDI, M_ADD, M_MUL,
    M_SUB, M_SLL, M_MULH, M_SLT, M_MULHSU, M_SLTU, M_MULHU, M_XOR, M_DIV, M_SRL, M_DIVU, M_SRA, M_OR, M_REM,
    M_AND, M_REMU, M_FENCE, M_ECALL, M_EBREAK, M_FLW, M_FSW, M_FMADD, M_FMSUB, M_FNMSUB, M_FNMADD, M_FADD,
    M_FSUB, M_FMUL, M_FDIV, M_FSQRT, M_FSGNJ, M_FSGNJN, M_FSGNJX, M_FMIN, M_FMAX, M_FCVT_W_S, M_FCVT_WU_S,
    M_FMV_X_W, M_FCLASS, M_FLE, M_FLT, M_FEQ, M_FCVT_S_W, M_FCVT_S_WU, M_FMV_W_X, M_UNKNOWN, M_ZERO, M_COUNT
};
// names and registers in 16- and 8-byte slots, read as 8-byte words
__constant__ __align__(16) char c_asm_name[M_COUNT][16] = {
    "lui", "auipc", "jal", "jalr", "beq", "bne", "blt", "bge", "bltu", "bgeu", "lb", "lh", "lw", "lbu", "lhu",
    "sb", "sh", "sw", "addi", "slli", "slti", "sltiu", "xori", "srli", "srai", "ori", "andi", "add", "mul",
    "sub", "sll", "mulh", "slt", "mulhsu", "sltu", "mulhu", "xor", "div", "srl", "divu", "sra", "or", "rem",
    "and", "remu", "fence", "ecall", "ebreak", "flw", "fsw", "fmadd.s", "fmsub.s", "fnmsub.s", "fnmadd.s",
    "fadd.s", "fsub.s", "fmul.s", "fdiv.s", "fsqrt.s", "fsgnj.s", "fsgnjn.s", "fsgnjx.s", "fmin.s", "fmax.s",
    "fcvt.w.s", "fcvt.wu.s", "fmv.x.w", "fclass.s", "fle.s", "flt.s", "feq.s", "fcvt.s.w", "fcvt.s.wu",
    "fmv.w.x", "unknown", "(zero)"};
// ABI register names (:15-50)
__constant__ __align__(8) char c_asm_x[32][8] = {"zero", "ra", "sp", "gp", "tp", "t0", "t1", "t2", "s0", "s1", "a0",
                                    "a1", "a2", "a3", "a4", "a5", "a6", "a7", "s2", "s3", "s4", "s5",
                                    "s6", "s7", "s8", "s9", "s10", "s11", "t3", "t4", "t5", "t6"};
__constant__ __align__(8) char c_asm_f[32][8] = {"ft0", "ft1", "ft2", "ft3", "ft4", "ft5", "ft6", "ft7", "fs0", "fs1", "fa0",
                                    "fa1", "fa2", "fa3", "fa4", "fa5", "fa6", "fa7", "fs2", "fs3", "fs4", "fs5",
                                    "fs6", "fs7", "fs8", "fs9", "fs10", "fs11", "ft8", "ft9", "ft10", "ft11"};

// instr_name: the case chain entered at the word's opcode.
DEV uint32_t asm_name(uint32_t w) {
    const uint32_t op = w & 0x7Fu, f3 = (w >> 12) & 7u, f7 = (w >> 25) & 0x7Fu, r2 = (w >> 20) & 0x1Fu;
    // cases in source order: 0 lui, 1 auipc, 2 jal, 3 jalr, 4 BRANCH, 5 LOAD,
    // 6 STORE, 7 OP-IMM, 8 OP, 9 fence, 10 SYSTEM, 11 flw, 12 fsw, 13-16 the
    // fused multiply-adds, 17 OP-FP
    int k;
    switch (op) {
        case 0x37: k = 0; break;  case 0x17: k = 1; break;  case 0x6F: k = 2; break;  case 0x67: k = 3; break;
        case 0x63: k = 4; break;  case 0x03: k = 5; break;  case 0x23: k = 6; break;  case 0x13: k = 7; break;
        case 0x33: k = 8; break;  case 0x0F: k = 9; break;  case 0x73: k = 10; break; case 0x07: k = 11; break;
        case 0x27: k = 12; break; case 0x43: k = 13; break; case 0x47: k = 14; break; case 0x4B: k = 15; break;
        case 0x4F: k = 16; break; case 0x53: k = 17; break;
        default: return w ? M_UNKNOWN : M_ZERO;
    }
    for (; k < 18; ++k) {
        switch (k) {
            case 0: return M_LUI;
            case 1: return M_AUIPC;
            case 2: return M_JAL;
            case 3: return M_JALR;
            case 4: { constexpr uint8_t t[8] = {M_BEQ, M_BNE, 0xFF, 0xFF, M_BLT, M_BGE, M_BLTU, M_BGEU};
                      if (t[f3] != 0xFF) return t[f3]; break; }
            case 5: { constexpr uint8_t t[8] = {M_LB, M_LH, M_LW, 0xFF, M_LBU, M_LHU, 0xFF, 0xFF};
                      if (t[f3] != 0xFF) return t[f3]; break; }
            case 6: if (f3 < 3) return M_SB + f3; break;
            case 7: { constexpr uint8_t t[8] = {M_ADDI, M_SLLI, M_SLTI, M_SLTIU, M_XORI, 0xFF, M_ORI, M_ANDI};
                      if (f3 == 5) return f7 == 0 ? M_SRLI : f7 == 0x20 ? M_SRAI : M_ORI;  // else into funct3 6
                      return t[f3]; }
            case 8: {
                // funct3 cases from the word's own on, each matching funct7 0, 1
                // (and 0x20 for add/sub, srl/sra), else the next funct3 case
                constexpr uint8_t base[8] = {M_ADD, M_SLL, M_SLT, M_SLTU, M_XOR, M_SRL, M_OR, M_AND};
                constexpr uint8_t mext[8] = {M_MUL, M_MULH, M_MULHSU, M_MULHU, M_DIV, M_DIVU, M_REM, M_REMU};
                for (uint32_t g = f3; g < 8; ++g) {
                    if (f7 == 0) return base[g];
                    if (f7 == 1) return mext[g];
                    if (f7 == 0x20 && g == 0) return M_SUB;
                    if (f7 == 0x20 && g == 5) return M_SRA;
                }
                break;
            }
            case 9: return M_FENCE;
            case 10: if (f7 == 0) return M_ECALL; if (f7 == 1) return M_EBREAK; break;
            case 11: return M_FLW;
            case 12: return M_FSW;
            case 13: return M_FMADD;
            case 14: return M_FMSUB;
            case 15: return M_FNMSUB;
            case 16: return M_FNMADD;
            case 17: {
                // funct7 cases in source order, entered at the word's own
                constexpr uint8_t sev[12] = {0x00, 0x04, 0x08, 0x0C, 0x2C, 0x10, 0x14, 0x60, 0x70, 0x50, 0x68, 0x78};
                uint32_t t = 0;
                while (t < 12 && sev[t] != f7) ++t;
                for (; t < 12; ++t) {
                    switch (t) {
                        case 0: return M_FADD;
                        case 1: return M_FSUB;
                        case 2: return M_FMUL;
                        case 3: return M_FDIV;
                        case 4: return M_FSQRT;
                        case 5: if (f3 < 3) return M_FSGNJ + f3; break;
                        case 6: if (f3 < 2) return M_FMIN + f3; break;
                        case 7: if (r2 < 2) return M_FCVT_W_S + r2; break;
                        case 8: if (f3 < 2) return M_FMV_X_W + f3; break;
                        case 9: if (f3 < 3) return M_FLE + f3; break;
                        case 10: if (r2 < 2) return M_FCVT_S_W + r2; break;
                        default: return M_FMV_W_X;
                    }
                }
                break;
            }
        }
    }
    return w ? M_UNKNOWN : M_ZERO;
}

// the bytes of v up to the first zero; true when all eight were emitted
template <class S> DEV bool asm_8(S &s, uint64_t v) {
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const char c = (char)(v >> (8 * i));
        if (!c) return false;
        s.put(c);
    }
    return true;
}
template <class S> DEV void asm_str16(S &s, const char *p) {  // a 16-byte slot
    const uint64_t *q = reinterpret_cast<const uint64_t *>(p);
    const uint64_t lo = q[0], hi = q[1];
    if (asm_8(s, lo)) asm_8(s, hi);
}
template <class S> DEV void asm_str8(S &s, const char *p) { asm_8(s, *reinterpret_cast<const uint64_t *>(p)); }
template <class S> DEV void asm_dec(S &s, int32_t x) {
    uint32_t u = (uint32_t)x;
    if (x < 0) { s.put('-'); u = 0u - u; }
    uint64_t dig = 0;  // decimal digits, 4 bits each, least significant first
    int nd = 0;
    do { dig |= (uint64_t)(u % 10u) << (4 * nd); u /= 10u; ++nd; } while (u);
    for (int k = nd - 1; k >= 0; --k) s.put((char)('0' + ((dig >> (4 * k)) & 15u)));
}
// std::hex << std::showbase: "0x" and the digits, a bare "0" for zero
template <class S> DEV void asm_hex(S &s, uint64_t v) {
    if (!v) { s.put('0'); return; }
    s.put('0'); s.put('x');
    for (int k = (67 - __clzll((long long)v)) / 4 - 1; k >= 0; --k) {
        const uint32_t d = (uint32_t)(v >> (4 * k)) & 15u;
        s.put((char)(d < 10 ? '0' + d : 'a' + d - 10));
    }
}

struct AsmCount {
    uint32_t n = 0;
    DEV void put(char) { ++n; }
};

// rvdisasm::disassemble (src/disassembler.cpp:411-451), one line per word for
// rv_generator::print(os, true) (src/codegen/rv_generator.cpp:100-103): before
// a word at a function start a label line (cd + hd + 18 spaces, "func<k>:",
// k in hex after word 0);
// then the word index (decimal, right-aligned in cd columns), three spaces, the
// byte address (hex, zero-padded to hd digits), ": ", the four bytes (hex, high
// byte first), three spaces, the mnemonic left-aligned in 9 columns, a space
// and format_args with colour and the word's address (branch targets are
// absolute).  Colours are the reference's ANSI escapes.
struct AsmList {
    uint64_t first;          // word index of the first word formatted
    const uint32_t *fstarts; // function starts (word indices), sorted, distinct
    uint32_t nf, cd, hd;     // functions; decimal digits of the word count; hex digits of the addresses
};
#define ASM_WHITE "\033[37m"
#define ASM_INSTR "\033[38;5;186m"
#define ASM_REG "\033[38;5;117m"
#define ASM_IMM "\033[38;5;114m"
#define ASM_EXTRA "\033[38;5;114m"
template <class S> DEV void asm_cstr(S &s, const char *p) { while (*p) s.put(*p++); }
// width-padded numbers of the listing's prefix
template <class S> DEV void asm_udec_w(S &s, uint64_t u, uint32_t width, char fill) {
    char b[20];
    int nd = 0;
    do { b[nd++] = (char)('0' + u % 10u); u /= 10u; } while (u);
    for (int k = nd; k < (int)width; ++k) s.put(fill);
    while (nd) s.put(b[--nd]);
}
template <class S> DEV void asm_uhex_w(S &s, uint64_t u, uint32_t width) {
    char b[16];
    int nd = 0;
    do { const uint32_t d = (uint32_t)(u & 15u); b[nd++] = (char)(d < 10 ? '0' + d : 'a' + d - 10); u >>= 4; } while (u);
    for (int k = nd; k < (int)width; ++k) s.put('0');
    while (nd) s.put(b[--nd]);
}

template <bool LIST, class S> DEV void asm_line(S &s, uint32_t w, const AsmList &L, uint64_t idx) {
    const uint32_t op = w & 0x7Fu, f3 = (w >> 12) & 7u, f7 = (w >> 25) & 0x7Fu;
    const uint32_t rd = (w >> 7) & 31u, rs1 = (w >> 15) & 31u, rs2 = (w >> 20) & 31u, rs3 = (w >> 27) & 31u;
    const uint32_t nm = asm_name(w);
    if (LIST) {
        // label: the index of this start among the distinct starts
        uint32_t lo = 0, hi = L.nf;
        while (lo < hi) { const uint32_t m = (lo + hi) >> 1; if (L.fstarts[m] < idx) lo = m + 1; else hi = m; }
        if (lo < L.nf && L.fstarts[lo] == idx) {
            for (uint32_t k = 0; k < L.cd + 3u + L.hd + 2u + 12u + 1u; ++k) s.put(' ');
            // the stream is still in hex from the previous word's line (format_args
            // restores the flags it found), so only a label before word 0 is decimal
            asm_cstr(s, "func");
            if (idx == 0) asm_udec_w(s, lo, 0, ' '); else asm_uhex_w(s, lo, 0);
            s.put(':'); s.put('\n');
        }
        asm_cstr(s, ASM_EXTRA); asm_udec_w(s, idx, L.cd, ' '); asm_cstr(s, ASM_WHITE);
        s.put(' '); s.put(' '); s.put(' ');
        asm_uhex_w(s, 4 * idx, L.hd); s.put(':'); s.put(' ');
        for (int b = 3; b >= 0; --b) {
            s.put(' '); asm_cstr(s, ASM_EXTRA); asm_uhex_w(s, (w >> (8 * b)) & 0xFFu, 2); asm_cstr(s, ASM_WHITE);
        }
        s.put(' '); s.put(' '); s.put(' ');
        asm_cstr(s, ASM_INSTR);
        AsmCount nc;
        asm_str16(nc, c_asm_name[nm]);
        asm_str16(s, c_asm_name[nm]);
        for (uint32_t k = nc.n; k < 9; ++k) s.put(' ');
        asm_cstr(s, ASM_WHITE);
    } else {
        asm_str16(s, c_asm_name[nm]);
    }
    s.put(' ');
    auto X = [&](uint32_t r) { if (LIST) asm_cstr(s, ASM_REG); asm_str8(s, c_asm_x[r]); if (LIST) asm_cstr(s, ASM_WHITE); };
    auto F = [&](uint32_t r) { if (LIST) asm_cstr(s, ASM_REG); asm_str8(s, c_asm_f[r]); if (LIST) asm_cstr(s, ASM_WHITE); };
    auto I0 = [&]() { if (LIST) asm_cstr(s, ASM_IMM); };
    auto I1 = [&]() { if (LIST) asm_cstr(s, ASM_WHITE); };
    auto sep = [&]() { s.put(','); s.put(' '); };
    switch (op) {
        case 0x33:  // R
            X(rd); sep(); X(rs1); sep(); X(rs2);
            break;
        case 0x53:  // R, OP-FP: register files by funct7 (:250-281)
            if (f7 == 0x78 || f7 == 0x68) { F(rd); sep(); X(rs1); }
            else if (f7 == 0x60 || (f7 == 0x70 && f3 == 0)) { X(rd); sep(); F(rs1); }
            else { F(rd); sep(); F(rs1); sep(); F(rs2); }
            break;
        case 0x67: case 0x03: case 0x13: case 0x0F: case 0x73: case 0x07: {  // I
            const int32_t imm = (int32_t)w >> 20;
            if (op == 0x07 || op == 0x03 || op == 0x67) {
                if (op == 0x07) F(rd); else X(rd);
                sep(); I0(); asm_dec(s, imm); I1(); s.put('('); X(rs1); s.put(')');
            } else {
                X(rd); sep(); X(rs1); sep(); I0(); asm_dec(s, imm); I1();
            }
            break;
        }
        case 0x23: case 0x27: {  // S; an fsw names its register by the rs1 field (:357-361)
            const int32_t imm = ((int32_t)(w & 0xFE000000u) >> 20) | (int32_t)((w >> 7) & 0x1Fu);
            if (op == 0x27) F(rs1); else X(rs2);
            sep(); I0(); asm_dec(s, imm); I1(); s.put('('); X(rs1); s.put(')');
            break;
        }
        case 0x63: {  // B: bits 4:1, 9:5 (not 10) and 11, the sign over bits 31:11 (:366-372)
            uint32_t imm = ((w >> 7) & 0x1Eu) | ((w >> 20) & 0x3E0u) | (((w >> 7) & 1u) << 11);
            if (w >> 31) imm |= 0xFFFFF800u;
            const uint64_t addr = LIST ? 4 * idx : 0;
            X(rs1); sep(); X(rs2); sep(); I0(); asm_hex(s, addr + (uint64_t)(int64_t)(int32_t)imm); I1();
            break;
        }
        case 0x37: case 0x17:  // U
            X(rd); sep(); I0(); asm_dec(s, (int32_t)(w & 0xFFFFF000u)); I1();
            break;
        case 0x6F: {  // J: bits 10:1, 11, 19:12, the sign over bits 31:19 (:384-392)
            uint32_t imm = ((w >> 20) & 0x7FEu) | (((w >> 20) & 1u) << 11) | (w & 0xFF000u);
            if (w >> 31) imm |= 0xFFF80000u;
            X(rd); sep(); I0(); asm_dec(s, (int32_t)imm); I1();
            break;
        }
        case 0x43: case 0x47: case 0x4B: case 0x4F:  // R4 (integer names, :395-403)
            X(rd); sep(); X(rs1); sep(); X(rs2); sep(); X(rs3);
            break;
        default:
            break;
    }
    s.put('\n');
}

// Writes bytes from pos on; a 4-byte word lying inside [lo, hi) (the thread's
// own lines) goes out as one store, the bytes of a word shared with a
// neighbour's lines one by one.
struct AsmWrite {
    uint8_t *out;
    uint64_t pos, lo, hi;
    uint32_t acc = 0;
    DEV void put(char ch) {
        const uint64_t wb = pos & ~3ull;
        if (wb >= lo && wb + 4 <= hi) {
            acc |= (uint32_t)(uint8_t)ch << (8u * (uint32_t)(pos & 3u));
            if ((pos & 3u) == 3u) { *reinterpret_cast<uint32_t *>(out + wb) = acc; acc = 0; }
        } else {
            out[pos] = (uint8_t)ch;
        }
        ++pos;
    }
};

// line lengths: formatted once with a counting sink (the tile sums), kept per
// word for the rescan (one byte for to_asm lines, two for listing lines)
template <bool LIST>
struct AsmLenOp {
    using Len = typename std::conditional<LIST, uint16_t, uint8_t>::type;
    const uint32_t *w;
    Len *len;
    AsmList L;
    DEV uint32_t value(uint64_t i) const {
        AsmCount c;
        asm_line<LIST>(c, w[i], L, L.first + i);
        len[i] = (Len)c.n;
        return c.n;
    }
    DEV void emit(uint64_t, uint32_t, uint32_t) const {}
};
template <class Len>
struct AsmLenRead {
    const Len *len;
    DEV uint32_t value(uint64_t i) const { return len[i]; }
};

template <bool LIST>
__global__ __launch_bounds__(SCAN_BLOCK) void k_asm_write(const uint32_t *w, const typename AsmLenOp<LIST>::Len *lens,
                                                          uint64_t n, const uint32_t *tile_pre, uint8_t *out, AsmList L) {
    __shared__ uint32_t s_w[SCAN_BLOCK / 64];
    __shared__ uint32_t s_v[SCAN_TILE + SCAN_TILE / 16];
    const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE + (uint64_t)threadIdx.x * SCAN_ITEMS;
    uint32_t v[SCAN_ITEMS];
    const uint32_t run = scan_tile(AsmLenRead<typename AsmLenOp<LIST>::Len>{lens}, n, tile_pre, s_w, s_v, v);
    uint32_t len = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) len += v[k];
    AsmWrite s{out, run, run, (uint64_t)run + len};
    for (int k = 0; k < SCAN_ITEMS; ++k)
        if (base + k < n) asm_line<LIST>(s, w[base + k], L, L.first + base + k);
}

// ================================================================ host side
#include "ptl_host.inc"

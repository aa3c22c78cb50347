/* CPU oracle: a scalar C restatement of the reference's canonical code
 * generator, rv_generator_avx run with one thread (concurrency 1, sync 0,
 * interleaved partitioning, mininstr 0), built with GCC.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links, loads or
 * calls this file: it is the checker for the parity tests, the smoke check and
 * the bench's CPU baseline leg.
 *
 * Every stage follows src/codegen/rv_generator_avx.cpp (cited per function)
 * at the granularity of its 8-lane AVX2 vectors, because several observable
 * results depend on it:
 *   - avx_buffer (include/codegen/avx_buffer.hpp:42-58): storage is rounded up
 *     to 64 bytes and surrounded by 64 zero bytes, so index -1 reads 0 and
 *     whole-vector stores write the padding lanes past the end;
 *   - scatters inside a vector are not visible to that vector's own loads
 *     (preprocess pass 3, the isn_cnt fix, the regalloc snapshot);
 *   - operators of include/simd.hpp: signed compares, logical >>,
 *     masked gathers that yield 0.
 * The isn_gen depth sort is libstdc++'s std::sort (introsort) restated exactly
 * (sort_introsort below), since nodes of equal depth that write the same word
 * slot resolve by their order in that sort.  ptilopsis::ffz is GCC's build of
 * include/utils.hpp:25-45, whose __builtin_ffs takes an int and so only sees
 * the low 32 bits.
 *
 * Anything that would read or write outside a buffer's allocation in the
 * reference is recorded in ctx->flags instead of being performed.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <stdio.h>
#include "oracle_tables.h"

#define PADE 16 /* 64 bytes of zero padding = 16 u32 before and after */

enum {
    OF_OOB = 1u << 0,          /* access outside an allocation */
    OF_CALL_PATH = 1u << 1,    /* regalloc's JAL branch would run (never for valid input) */
    OF_EMPTY = 1u << 2,        /* no functions: std::ranges::max of an empty range */
    OF_HEAPSORT = 1u << 3,     /* introsort fell back to heapsort somewhere */
    OF_TIE = 1u << 4,          /* two same-depth nodes wrote one word slot */
};

typedef struct { uint32_t *p; int64_t n; int64_t cap; } vb; /* avx_buffer<u32> */
typedef struct { uint64_t *p; int64_t n; int64_t cap; } vb64;

/* per thread: independent oracle runs may share the process (the host baseline
 * compiles many programs on a thread pool) */
static _Thread_local uint32_t g_flags;

static vb vb_new(int64_t n) {
    vb b; b.n = n; b.cap = (n + 15) & ~(int64_t)15;
    uint32_t *raw = (uint32_t *)calloc((size_t)(b.cap + 2 * PADE), 4);
    b.p = raw + PADE; return b;
}
static void vb_free(vb *b) { if (b->p) free(b->p - PADE); b->p = NULL; b->n = b->cap = 0; }
static vb vb_copy(const vb *o) { vb b = vb_new(o->n); memcpy(b.p, o->p, (size_t)o->n * 4); return b; }
static int vb_ok(const vb *b, int64_t i) { return i >= -PADE && i < b->cap + PADE; }
static uint32_t R(const vb *b, int64_t i) { if (!vb_ok(b, i)) { g_flags |= OF_OOB; return 0; } return b->p[i]; }
static void W(vb *b, int64_t i, uint32_t v) { if (!vb_ok(b, i)) { g_flags |= OF_OOB; return; } b->p[i] = v; }
static int64_t vecs(const vb *b) { return (b->n + 7) / 8; } /* size_m256i for 4-byte elements */

static vb64 vb64_new(int64_t n) {
    vb64 b; b.n = n; b.cap = (n + 7) & ~(int64_t)7;
    uint64_t *raw = (uint64_t *)calloc((size_t)(b.cap + 16), 8);
    b.p = raw + 8; return b;
}
static void vb64_free(vb64 *b) { if (b->p) free(b->p - 8); b->p = NULL; }

typedef struct {
    oracle_tables t;
    int64_t n; uint32_t max_depth;
    vb types, rtypes, parents, depth, cidx, nd;
    vb sizes, locs;
    vb fids, fstarts, fends, fsizes;
    uint32_t F;
    vb instr, rd, rs1, rs2, jt;     /* W0, then W1, then W2 */
    vb used;                        /* used_instrs after optimize */
    vb sym_reg, sym_sw;             /* regalloc walk results */
    vb ra_offsets;                  /* regalloc instr_offsets */
    vb fj_offsets;                  /* fix_jumps instr_offsets */
    vb idx_sorted;                  /* isn_gen depth-sorted node order */
    uint32_t opt_iters;
    uint32_t stage_done;
    uint32_t flags;
    uint64_t ties;
} ptlo_ctx;

/* ---------------------------------------------------------------- helpers */

static inline int32_t S(uint32_t v) { return (int32_t)v; }

/* include/utils.hpp:25-45 as compiled by GCC: __builtin_ffs(int) of ~val. */
static inline uint32_t ffz_gcc(uint64_t v) {
    uint32_t low = (uint32_t)(~v);
    if (low == 0) return 64;
    return (uint32_t)__builtin_ctz(low);
}

/* Exclusive scan exactly as the reference's in-register scans do it
 * (rv_generator_avx.cpp:374-392): element -1 is read from the zero padding,
 * every full vector is stored, padding lanes included. */
static void excl_scan_vec(const vb *in, vb *out) {
    uint32_t offset = 0;
    for (int64_t v = 0; v < vecs(in); ++v) {
        uint32_t lane[8];
        for (int l = 0; l < 8; ++l) lane[l] = R(in, v * 8 + l - 1);
        for (int l = 1; l < 8; ++l) lane[l] += lane[l - 1];
        for (int l = 0; l < 8; ++l) { lane[l] += offset; W(out, v * 8 + l, lane[l]); }
        offset = lane[7];
    }
}

/* ------------------------------------------------ libstdc++ std::sort ------
 * GCC's introsort (bits/stl_algo.h, bits/stl_heap.h): depth limit 2*floor(log2 n),
 * median of (first+1, mid, last-1) moved to first, unguarded Hoare partition,
 * segments <= 16 left for a final insertion sort, heapsort when the limit runs
 * out.  The comparator is depth[a] < depth[b] (rv_generator_avx.cpp:539-541). */
static _Thread_local const int32_t *g_key;
#define LESS(a, b) (g_key[(a)] < g_key[(b)])

static void swap_u32(uint32_t *a, uint32_t *b) { uint32_t t = *a; *a = *b; *b = t; }

static void move_median_to_first(uint32_t *result, uint32_t *a, uint32_t *b, uint32_t *c) {
    if (LESS(*a, *b)) {
        if (LESS(*b, *c)) swap_u32(result, b);
        else if (LESS(*a, *c)) swap_u32(result, c);
        else swap_u32(result, a);
    } else if (LESS(*a, *c)) swap_u32(result, a);
    else if (LESS(*b, *c)) swap_u32(result, c);
    else swap_u32(result, b);
}

static uint32_t *unguarded_partition(uint32_t *first, uint32_t *last, uint32_t *pivot) {
    for (;;) {
        while (LESS(*first, *pivot)) ++first;
        --last;
        while (LESS(*pivot, *last)) --last;
        if (!(first < last)) return first;
        swap_u32(first, last);
        ++first;
    }
}

static void push_heap_(uint32_t *first, int64_t hole, int64_t top, uint32_t value) {
    int64_t parent = (hole - 1) / 2;
    while (hole > top && LESS(first[parent], value)) {
        first[hole] = first[parent];
        hole = parent;
        parent = (hole - 1) / 2;
    }
    first[hole] = value;
}

static void adjust_heap(uint32_t *first, int64_t hole, int64_t len, uint32_t value) {
    const int64_t top = hole;
    int64_t child = hole;
    while (child < (len - 1) / 2) {
        child = 2 * (child + 1);
        if (LESS(first[child], first[child - 1])) child--;
        first[hole] = first[child];
        hole = child;
    }
    if ((len & 1) == 0 && child == (len - 2) / 2) {
        child = 2 * (child + 1);
        first[hole] = first[child - 1];
        hole = child - 1;
    }
    push_heap_(first, hole, top, value);
}

static void heap_sort_range(uint32_t *first, uint32_t *last) {
    int64_t len = last - first;
    if (len >= 2) {
        int64_t parent = (len - 2) / 2;
        for (;;) {
            uint32_t value = first[parent];
            adjust_heap(first, parent, len, value);
            if (parent == 0) break;
            parent--;
        }
    }
    while (last - first > 1) {
        --last;
        uint32_t value = *last;
        *last = *first;
        adjust_heap(first, 0, last - first, value);
    }
}

static void introsort_loop(uint32_t *first, uint32_t *last, int64_t limit) {
    while (last - first > 16) {
        if (limit == 0) { g_flags |= OF_HEAPSORT; heap_sort_range(first, last); return; }
        --limit;
        uint32_t *mid = first + (last - first) / 2;
        move_median_to_first(first, first + 1, mid, last - 1);
        uint32_t *cut = unguarded_partition(first + 1, last, first);
        introsort_loop(cut, last, limit);
        last = cut;
    }
}

static void unguarded_linear_insert(uint32_t *last) {
    uint32_t val = *last;
    uint32_t *next = last - 1;
    while (LESS(val, *next)) { *last = *next; last = next; --next; }
    *last = val;
}

static void insertion_sort(uint32_t *first, uint32_t *last) {
    if (first == last) return;
    for (uint32_t *i = first + 1; i != last; ++i) {
        if (LESS(*i, *first)) {
            uint32_t val = *i;
            memmove(first + 1, first, (size_t)(i - first) * 4);
            *first = val;
        } else unguarded_linear_insert(i);
    }
}

static void sort_introsort(uint32_t *first, int64_t n, const int32_t *key) {
    g_key = key;
    if (n <= 0) return;
    int64_t lg = 63 - __builtin_clzll((unsigned long long)n);
    introsort_loop(first, first + n, lg * 2);
    if (n > 16) {
        insertion_sort(first, first + 16);
        for (uint32_t *i = first + 16; i != first + n; ++i) unguarded_linear_insert(i);
    } else insertion_sort(first, first + n);
}

/* ------------------------------------------------------------- construction
 * rv_generator::rv_generator (src/codegen/rv_generator.cpp:41-67): widen to
 * u32, remap pareas node types, zero node_sizes. */
ptlo_ctx *ptlo_create(const uint8_t *types, const uint8_t *rtypes, const int32_t *parents,
                      const int32_t *depth, const int32_t *cidx, const uint32_t *nd,
                      int64_t n, uint32_t max_depth) {
    ptlo_ctx *c = (ptlo_ctx *)calloc(1, sizeof(ptlo_ctx));
    oracle_tables_init(&c->t);
    c->n = n; c->max_depth = max_depth;
    c->types = vb_new(n); c->rtypes = vb_new(n); c->parents = vb_new(n);
    c->depth = vb_new(n); c->cidx = vb_new(n); c->nd = vb_new(n);
    c->sizes = vb_new(n); c->locs = vb_new(n);
    for (int64_t i = 0; i < n; ++i) {
        c->types.p[i] = types[i] < 44 ? c->t.pareas_map[types[i]] : 0;
        c->rtypes.p[i] = rtypes[i];
        c->parents.p[i] = (uint32_t)parents[i];
        c->depth.p[i] = (uint32_t)depth[i];
        c->cidx.p[i] = (uint32_t)cidx[i];
        c->nd.p[i] = nd[i];
    }
    return c;
}

void ptlo_destroy(ptlo_ctx *c) {
    if (!c) return;
    vb *all[] = { &c->types, &c->rtypes, &c->parents, &c->depth, &c->cidx, &c->nd, &c->sizes,
                  &c->locs, &c->fids, &c->fstarts, &c->fends, &c->fsizes, &c->instr, &c->rd,
                  &c->rs1, &c->rs2, &c->jt, &c->used, &c->sym_reg, &c->sym_sw, &c->ra_offsets,
                  &c->fj_offsets, &c->idx_sorted };
    for (unsigned i = 0; i < sizeof(all) / sizeof(all[0]); ++i) vb_free(all[i]);
    free(c);
}

/* ------------------------------------------------------------- preprocess
 * rv_generator_avx.cpp:122-308 */
static void st_preprocess(ptlo_ctx *c) {
    const int64_t nv = vecs(&c->types);
    /* pass 1 (:126-214): argument classification, elementwise */
    for (int64_t v = 0; v < nv; ++v) {
        for (int l = 0; l < 8; ++l) {
            int64_t i = v * 8 + l;
            uint32_t ty = R(&c->types, i);
            if ((ty & O_FARG) != O_FARG) continue;
            uint32_t rt = R(&c->rtypes, i);
            int isf = (rt == D_FLOAT) || (rt == D_FLOAT_REF);
            uint32_t ndv = R(&c->nd, i), ci = R(&c->cidx, i);
            if (isf) {
                uint32_t fd = ndv;
                int comp = S(fd) > 7;
                uint32_t reg = (ci - fd) + fd - 8;
                int on_stack = (S(reg) > 7) && comp;
                int as_int = !on_stack && comp;
                if (on_stack) ty |= 2;
                if (as_int) ty |= 1;
                if (on_stack) reg -= 8;
                if (comp) ndv = reg;
            } else {
                uint32_t idata = ndv;
                uint32_t fargs = ci - idata;
                uint32_t reg = idata + (fargs - (S(fargs) > 7 ? 8u : 0u));
                /* D9: the on-stack "mask" is (reg > 7) ANDed with the value itself */
                uint32_t osm = (S(reg) > 7 ? 0xFFFFFFFFu : 0u) & idata;
                ty |= (osm & 2u);
                reg = reg - (8u & osm);
                ndv = reg;
            }
            W(&c->types, i, ty);
            W(&c->nd, i, ndv);
        }
    }
    /* pass 2 (:217-266): stack-argument count on func_call_arg_list */
    for (int64_t v = 0; v < nv; ++v) {
        int any_list = 0;
        for (int l = 0; l < 8; ++l) any_list |= R(&c->types, v * 8 + l) == O_CALL_ARG_LIST;
        if (!any_list) continue;
        uint32_t pty[8], pnd[8], prt[8], pci[8];
        for (int l = 0; l < 8; ++l) {
            int64_t i = v * 8 + l;
            pty[l] = R(&c->types, i - 1); pnd[l] = R(&c->nd, i - 1);
            prt[l] = R(&c->rtypes, i - 1); pci[l] = R(&c->cidx, i - 1);
        }
        /* the store masks below are not ANDed with the list mask (:236-244) */
        int any_stack = 0;
        for (int l = 0; l < 8; ++l) any_stack |= pty[l] == O_CARG_STACK;
        if (any_stack)
            for (int l = 0; l < 8; ++l)
                if (pty[l] == O_CARG_STACK) W(&c->nd, v * 8 + l, pnd[l] + 1);
        int any_f = 0;
        for (int l = 0; l < 8; ++l) any_f |= (pty[l] == O_CARG) && (prt[l] == D_FLOAT);
        if (any_f)
            for (int l = 0; l < 8; ++l)
                if (pty[l] == O_CARG && prt[l] == D_FLOAT) {
                    int32_t ia = S(pci[l] - pnd[l]) - 8;
                    W(&c->nd, v * 8 + l, (uint32_t)(ia > 0 ? ia : 0));
                }
    }
    /* pass 3 (:270-305): float propagation to comparison-like parents, per vector */
    for (int64_t v = 0; v < nv; ++v) {
        uint32_t rt[8], par[8];
        int st[8];
        for (int l = 0; l < 8; ++l) {
            int64_t i = v * 8 + l;
            par[l] = R(&c->parents, i);
            int valid = S(par[l]) > -1;
            uint32_t pty = R(&c->types, valid ? S(par[l]) : 0);
            int pe = (pty & O_EQ) == O_EQ;
            rt[l] = R(&c->rtypes, i) & (pe ? 0xFFFFFFFFu : 0u);
            st[l] = rt[l] == D_FLOAT;
        }
        for (int l = 0; l < 8; ++l)
            if (st[l]) W(&c->rtypes, S(par[l]), D_FLOAT);
    }
}

/* ---------------------------------------------------------------- isn_cnt
 * rv_generator_avx.cpp:310-518 */
static void st_isn_cnt(ptlo_ctx *c) {
    const int64_t nv = vecs(&c->types);
    const oracle_tables *t = &c->t;
    /* count (:315-369) */
    for (int64_t v = 0; v < nv; ++v) {
        int any_list = 0;
        for (int l = 0; l < 8; ++l) any_list |= R(&c->types, v * 8 + l) == O_CALL_ARG_LIST;
        for (int l = 0; l < 8; ++l) {
            int64_t i = v * 8 + l;
            uint32_t ty = R(&c->types, i), rt = R(&c->rtypes, i);
            uint32_t base = t->size[ty & 255][rt < 6 ? rt : 0];
            if ((ty * 6 + rt) >= 256 * 6) g_flags |= OF_OOB;
            uint32_t delta = 0;
            if (any_list && ty == O_CALL_ARG_LIST) {
                delta += 1;
                uint32_t pty = R(&c->types, i - 1);
                if ((pty & O_CARG) == O_CARG) delta += R(&c->cidx, i - 1) + 1;
            }
            uint32_t par = R(&c->parents, i);
            int valid = !(par == 0xFFFFFFFFu);
            uint32_t pty = R(&c->types, valid ? S(par) : 0);
            uint32_t ci = R(&c->cidx, i);
            int ifelse = pty == O_IF_ELSE;
            if (ci == 0 && (ifelse || pty == O_IF)) delta += 1;
            if (ci == 1 && (ifelse || pty == O_WHILE)) delta += 1;
            W(&c->sizes, i, base + delta);
        }
    }
    /* scan (:374-392) */
    excl_scan_vec(&c->sizes, &c->locs);
    /* fix (:398-447) */
    for (int64_t v = 0; v < nv; ++v) {
        uint32_t par[8], ci[8], val[8]; int cm[8], any = 0;
        for (int l = 0; l < 8; ++l) {
            int64_t i = v * 8 + l;
            par[l] = R(&c->parents, i);
            int valid = S(par[l]) > -1;
            ci[l] = R(&c->cidx, i);
            uint32_t pty = R(&c->types, valid ? S(par[l]) : 0);
            cm[l] = ((pty & O_IF) == O_IF) && (ci[l] == (pty & 1));
            any |= cm[l];
        }
        if (any) {
            for (int l = 0; l < 8; ++l) {
                int64_t i = v * 8 + l;
                val[l] = R(&c->locs, i);
                if (cm[l]) {
                    uint32_t ty = R(&c->types, i), rt = R(&c->rtypes, i);
                    val[l] += t->size[ty & 255][rt < 6 ? rt : 0];
                }
            }
            for (int l = 0; l < 8; ++l) if (cm[l]) W(&c->locs, S(par[l]), val[l]);
        }
        int fm[8], anyf = 0;
        for (int l = 0; l < 8; ++l) {
            fm[l] = (R(&c->types, v * 8 + l) & O_CARG) == O_CARG;
            anyf |= fm[l];
        }
        if (anyf) {
            uint32_t nl[8];
            for (int l = 0; l < 8; ++l)
                nl[l] = fm[l] ? R(&c->locs, S(par[l])) + R(&c->cidx, v * 8 + l) + 1 : 0;
            for (int l = 0; l < 8; ++l) if (fm[l]) W(&c->locs, v * 8 + l, nl[l]);
        }
    }
    /* function table (:451-515) */
    uint32_t F = 0;
    for (int64_t i = 0; i < nv * 8; ++i) F += R(&c->types, i) == O_FDECL;
    c->F = F;
    vb decls = vb_new(F);
    { uint32_t k = 0; for (int64_t i = 0; i < nv * 8; ++i) if (R(&c->types, i) == O_FDECL) W(&decls, k++, (uint32_t)i); }
    c->fids = vb_new(F); c->fstarts = vb_new(F);
    vb offs = vb_new(F);
    for (int64_t v = 0; v < vecs(&decls); ++v)
        for (int l = 0; l < 8; ++l) {
            int64_t k = v * 8 + l;
            uint32_t d = R(&decls, k);
            W(&c->fids, k, R(&c->nd, d));
            uint32_t off = 6 + R(&c->locs, d);
            W(&c->fstarts, k + 1, off);
            W(&offs, k, off);
        }
    c->fsizes = vb_new(F);
    for (int64_t v = 0; v < vecs(&decls); ++v)
        for (int l = 0; l < 8; ++l) {
            int64_t k = v * 8 + l;
            W(&c->fsizes, k, R(&offs, k) - R(&offs, k - 1));
        }
    c->fends = vb_new(F);
    for (int64_t v = 0; v < vecs(&decls); ++v)
        for (int l = 0; l < 8; ++l) {
            int64_t k = v * 8 + l;
            W(&c->fends, k, R(&c->fstarts, k) + R(&c->fsizes, k));
        }
    vb_free(&decls); vb_free(&offs);
}

/* ---------------------------------------------------------------- isn_gen
 * rv_generator_avx.cpp:520-889 */
typedef struct { int hi; uint32_t loc, word, rd, rs1, rs2, jt; int32_t pai; uint32_t dpv; } gen_lane;

static void st_isn_gen(ptlo_ctx *c) {
    const oracle_tables *t = &c->t;
    const int64_t n = c->n;
    vb idx = vb_new(n);
    for (int64_t i = 0; i < n; ++i) idx.p[i] = (uint32_t)i;
    sort_introsort(idx.p, n, (const int32_t *)c->depth.p);
    /* depth_starts (:546-566): stable compaction of level starts; the tail past
     * the kept entries keeps its previous values (shrink_to only moves the end). */
    vb ds = vb_new(n);
    int64_t kept = 0;
    for (int64_t i = 0; i < n; ++i) ds.p[i] = (uint32_t)i;
    {
        /* std::remove_if: keep entries i where i == 0 or the depth changes */
        int64_t w = 0;
        for (int64_t i = 0; i < n; ++i) {
            uint32_t e = ds.p[i];
            int remove = !(e == 0 || c->depth.p[idx.p[e]] != c->depth.p[idx.p[e - 1]]);
            if (!remove) ds.p[w++] = e;
        }
        kept = w;
    }
    (void)kept;
    const uint32_t W0 = R(&c->locs, n - 1);
    vb regs = vb_new(n * 3);
    c->instr = vb_new(W0); c->rd = vb_new(W0); c->rs1 = vb_new(W0); c->rs2 = vb_new(W0); c->jt = vb_new(W0);
    /* winner depth/rank per slot, only to count same-depth collisions */
    uint32_t *owner = (uint32_t *)malloc(sizeof(uint32_t) * (W0 + 1));
    uint32_t *owner_depth = (uint32_t *)malloc(sizeof(uint32_t) * (W0 + 1));
    for (uint32_t i = 0; i <= W0; ++i) { owner[i] = 0xFFFFFFFFu; owner_depth[i] = 0xFFFFFFFFu; }

    for (uint32_t it = 0; it < c->max_depth + 1; ++it) {
        uint32_t d = c->max_depth - it;
        uint32_t start = R(&ds, d);
        uint32_t end = (d == c->max_depth) ? (uint32_t)n : R(&ds, d + 1);
        uint32_t lsize = end - start, steps = (lsize + 1) / 2;
        int uneven = lsize & 1;
        for (uint32_t j = 0; j < steps; ++j) {
            int skip_last = uneven && (steps == j + 1);
            uint32_t i0 = R(&idx, start + 2 * j);
            uint32_t i1 = skip_last ? i0 : R(&idx, start + 2 * j + 1);
            gen_lane L[8];
            for (int lane = 0; lane < 8; ++lane) {
                uint32_t nn = lane < 4 ? i0 : i1;
                uint32_t o = (uint32_t)(lane & 3);
                uint32_t ty = R(&c->types, nn), rt = R(&c->rtypes, nn);
                uint32_t hidx = ty * 24 + o * 6 + rt;
                if (hidx >= 256 * 24) { g_flags |= OF_OOB; hidx = 0; }
                const uint32_t ty8 = hidx / 24, o8 = (hidx / 6) % 4, rt8 = hidx % 6;
                int load = !(lane >= 4 && skip_last);
                uint32_t has = load ? t->has_instr[ty8][o8][rt8] : 0;
                int prop = has || lane == 0 || (lane == 4 && !skip_last);
                uint32_t par = prop ? R(&c->parents, nn) : 0;
                int nopar = par == 0xFFFFFFFFu;
                int lfp = !nopar && prop;
                uint32_t pty = lfp ? R(&c->types, S(par)) : 0;
                uint32_t ci = lfp ? R(&c->cidx, nn) : 0;
                int noncond = ((pty & O_IF) == O_IF) && !(ci == (pty & 1));
                int call = noncond;
                int lookload = (nopar || !noncond) && prop;
                if (lookload) {
                    uint32_t ct = t->parent_arg[ty8][o8][rt8];
                    if (ct == 1) call = 1;
                    if (ct == 2 && S(rt) > 1) call = 1;
                }
                int32_t pai = call ? (int32_t)(par * 3 + ci) : -1;
                uint32_t iib = o + (prop ? R(&c->locs, nn) : 0);
                uint32_t rd = 0;
                if (has) {
                    uint32_t ct = t->get_output[ty8][o8][rt8];
                    if (ct == 3) rd |= 32;
                    if (ct == 4) rd |= 10;
                    if (ct == 1) rd |= R(&c->nd, nn) + 10;
                    if (ct == 2) rd |= R(&c->nd, nn) + 42;
                    int hov = call || (ct == 0 && t->has_output[ty8][o8][rt8]);
                    if (hov) rd |= iib + 64;
                }
                int use_rd = nopar || (!noncond && lfp);
                uint32_t instr_no = iib + (noncond ? t->size[ty8][rt8] : 0);
                uint32_t dpv = (use_rd ? rd : 0) | (noncond ? instr_no : 0);
                uint32_t iloc = has ? iib : 0;
                if (o == 1 && ty == O_IF_ELSE) iloc = R(&regs, (int64_t)nn * 3 + 1);
                if (o == 1 && ty == O_WHILE) iloc = R(&regs, (int64_t)nn * 3 + 2);
                int add2 = (ty == O_FDECL) || (o > 1 && ty == O_FDECL_DUMMY);
                int cal1 = (o == 1) && (ty == O_CALL_ARG_LIST);
                if (add2 || cal1) iloc += 2;
                if (cal1) {
                    uint32_t ppt = R(&c->types, (int64_t)nn - 1);
                    if ((ppt & O_CARG) == O_CARG) iloc += R(&c->cidx, (int64_t)nn - 1);
                }
                uint32_t word = has ? t->instr[ty8][o8][rt8] : 0;
                uint32_t ndv = has ? R(&c->nd, nn) : 0;
                uint32_t cc = has ? t->iconst[ty8][o8] : 0;
                if (cc == 1) word |= ndv & 0xFFFFF000u;
                if (cc == 2) word |= (ndv & 0xFFFu) << 20;
                uint32_t nd4 = ndv * 4;
                if (cc == 3) word |= (uint32_t)(-(nd4 + 8)) << 20;
                if (cc == 4) word |= nd4 << 20;
                if (cc == 5) word |= (uint32_t)(-nd4) << 20;
                uint32_t jc = has ? t->jt[ty8][o8] : 0;
                uint32_t jt = 0;
                if (S(jc) > 0 && !(S(jc) > 5)) {
                    uint32_t rv = R(&regs, (int64_t)nn * 3 + 1 + (S(jc) > 2 ? 1 : 0));
                    rv += ((jc & 1) == 0) ? 1u : 0u;
                    jt = rv;
                }
                if (jc == 6) jt |= R(&c->fends, ndv) - 6;
                if (jc == 7) jt |= R(&c->fstarts, ndv);
                uint32_t rs[2];
                for (int s = 0; s < 2; ++s) {
                    uint32_t oc = has ? t->operand[ty8][o8][rt8][s] : 0;
                    uint32_t arg = 0;
                    if (oc == 1 || oc == 4 || oc == 5 || oc == 6) {
                        int64_t ri = (int64_t)nn * 3;
                        if (oc == 1) ri += s;
                        if (oc == 4 || oc == 5) ri += 1;
                        if (oc == 5) ri -= s;
                        arg = R(&regs, ri);
                    }
                    if (oc == 2) arg |= ndv + 10;
                    if (oc == 3) arg |= ndv + 42;
                    if (oc == 7) arg |= iib + 63;
                    rs[s] = arg;
                }
                L[lane].hi = has != 0; L[lane].loc = iloc; L[lane].word = word; L[lane].rd = rd;
                L[lane].rs1 = rs[0]; L[lane].rs2 = rs[1]; L[lane].jt = jt; L[lane].pai = pai; L[lane].dpv = dpv;
            }
            for (int lane = 0; lane < 8; ++lane) {
                if (!L[lane].hi) continue;
                uint32_t k = L[lane].loc;
                uint32_t nn = lane < 4 ? i0 : i1;
                if (k < W0) {
                    if (owner_depth[k] == d && owner[k] != nn) { g_flags |= OF_TIE; c->ties++; }
                    owner_depth[k] = d; owner[k] = nn;
                }
                W(&c->instr, k, L[lane].word); W(&c->rd, k, L[lane].rd); W(&c->rs1, k, L[lane].rs1);
                W(&c->rs2, k, L[lane].rs2); W(&c->jt, k, L[lane].jt);
            }
            for (int lane = 0; lane < 8; ++lane)
                if (L[lane].pai >= 0) W(&regs, L[lane].pai, L[lane].dpv);
        }
    }
    free(owner); free(owner_depth);
    c->idx_sorted = idx;
    vb_free(&ds); vb_free(&regs);
}

/* --------------------------------------------------------------- optimize
 * rv_generator_avx.cpp:891-1016 */
static void st_optimize(ptlo_ctx *c) {
    const int64_t nv = vecs(&c->instr);
    vb used = vb_new(c->instr.n);
    for (int64_t v = 0; v < nv; ++v) {
        for (int l = 0; l < 8; ++l) { int32_t x = S(R(&c->rs1, v * 8 + l) - 64); if (x >= 0) W(&used, x, 0xFFFFFFFFu); }
        for (int l = 0; l < 8; ++l) { int32_t x = S(R(&c->rs2, v * 8 + l) - 64); if (x >= 0) W(&used, x, 0xFFFFFFFFu); }
    }
    c->opt_iters = 0;
    for (;;) {
        vb nu = vb_copy(&used);
        c->opt_iters++;
        for (int64_t v = 0; v < nv; ++v) {
            int32_t a[8], b[8];
            for (int l = 0; l < 8; ++l) {
                int64_t i = v * 8 + l;
                uint32_t rd = R(&c->rd, i);
                int cr = S(rd) > 63;
                if (cr && R(&used, S(rd - 64))) cr = 0;
                a[l] = cr ? S(R(&c->rs1, i) - 64) : -1;
                b[l] = cr ? S(R(&c->rs2, i) - 64) : -1;
            }
            for (int l = 0; l < 8; ++l) if (a[l] >= 0) W(&nu, a[l], 0);
            for (int l = 0; l < 8; ++l) if (b[l] >= 0) W(&nu, b[l], 0);
        }
        for (int64_t v = 0; v < nv; ++v) {
            int32_t a[8], b[8];
            for (int l = 0; l < 8; ++l) {
                int64_t i = v * 8 + l;
                uint32_t w = R(&c->instr, i) & 0x707Fu;
                int st = (w == 0x2023u) || (w == 0x2027u);
                a[l] = st ? S(R(&c->rs1, i) - 64) : -1;
                b[l] = st ? S(R(&c->rs2, i) - 64) : -1;
            }
            for (int l = 0; l < 8; ++l) if (a[l] >= 0) W(&nu, a[l], 0xFFFFFFFFu);
            for (int l = 0; l < 8; ++l) if (b[l] >= 0) W(&nu, b[l], 0xFFFFFFFFu);
        }
        int same = memcmp(used.p, nu.p, (size_t)used.n * 4) == 0;
        vb_free(&used);
        used = nu;
        if (same) break;
    }
    c->used = vb_new(c->instr.n);
    for (int64_t v = 0; v < nv; ++v)
        for (int l = 0; l < 8; ++l) {
            int64_t i = v * 8 + l;
            uint32_t notv = S(R(&c->rd, i)) > 63 ? 0u : 0xFFFFFFFFu;
            W(&c->used, i, notv | R(&used, i));
        }
    vb_free(&used);
}

/* ---------------------------------------------------------- fix_func_tab
 * rv_generator_avx.cpp:1816-1837 */
static void fix_func_tab(ptlo_ctx *c, const vb *off) {
    for (int64_t v = 0; v < vecs(&c->fstarts); ++v) {
        uint32_t s[8], e[8], z[8];
        for (int l = 0; l < 8; ++l) {
            int64_t k = v * 8 + l;
            uint32_t si = R(&c->fstarts, k);
            s[l] = R(off, S(si));
            uint32_t ei = si + R(&c->fsizes, k);
            int past = S(ei) >= (int32_t)off->n;
            e[l] = past ? R(off, off->n - 1) + 1 : R(off, S(ei));
            z[l] = e[l] - s[l];
        }
        for (int l = 0; l < 8; ++l) {
            int64_t k = v * 8 + l;
            W(&c->fstarts, k, s[l]); W(&c->fends, k, e[l]); W(&c->fsizes, k, z[l]);
        }
    }
}

/* --------------------------------------------------------------- regalloc
 * rv_generator_avx.cpp:1018-1814 */
static void st_regalloc(ptlo_ctx *c) {
    const uint32_t F = c->F;
    const int64_t W0 = c->instr.n;
    uint32_t max_size = 0;
    for (uint32_t f = 0; f < F; ++f) if (c->fsizes.p[f] > max_size) max_size = c->fsizes.p[f];
    const int64_t groups = (F + 7) / 8;
    vb64 life = vb64_new(F);
    for (uint32_t f = 0; f < F; ++f) life.p[f] = 0x1Fu;
    c->sym_reg = vb_new(W0); c->sym_sw = vb_new(W0);
    int32_t regstate[8 * 64];
    for (int i = 0; i < 8 * 64; ++i) regstate[i] = -1;
    int32_t snap[64];
    /* allocation walk (:1041-1418), one thread: steps outer, groups inner */
    for (uint32_t i = 0; i < max_size; ++i) {
        memcpy(snap, regstate, sizeof(snap)); /* only region 0 of the copy is ever read */
        for (int64_t j = 0; j < groups; ++j) {
            int64_t off[8]; int nc[8];
            uint32_t ins[8], rs1[8], rs2[8], r1r[8], r2r[8], rd[8], ord[8];
            uint64_t lt[8], cl[8];
            int any_rdv = 0;
            for (int k = 0; k < 8; ++k) {
                int64_t f = j * 8 + k;
                uint32_t fs = R(&c->fstarts, f), fz = R(&c->fsizes, f);
                int oob = (int32_t)i >= S(fz);
                off[k] = oob ? -1 : (int64_t)(uint32_t)(fs + i);
                if (!oob && off[k] == 0xFFFFFFFFll) off[k] = -1;
                int valid = off[k] != -1;
                ins[k] = valid ? R(&c->instr, off[k]) : 0;
                if (ins[k] == 0xEFu) g_flags |= OF_CALL_PATH;
                nc[k] = valid;
                rs1[k] = nc[k] ? R(&c->rs1, off[k]) : 0;
                rs2[k] = nc[k] ? R(&c->rs2, off[k]) : 0;
                int v1 = S(rs1[k]) > 63, v2 = S(rs2[k]) > 63;
                r1r[k] = v1 ? R(&c->sym_reg, S(rs1[k] - 64)) : 0;
                uint32_t s1 = v1 ? R(&c->sym_sw, S(rs1[k] - 64)) : 0;
                r2r[k] = v2 ? R(&c->sym_reg, S(rs2[k] - 64)) : 0;
                uint32_t s2 = v2 ? R(&c->sym_sw, S(rs2[k] - 64)) : 0;
                if (s1) {
                    int nf = ins[k] == 0xC0007053u || (ins[k] & 0x7F) == 0x53;
                    r1r[k] = nf ? 37 : 5;
                }
                if (s2) {
                    int nf = (ins[k] & 0x7F) == 0x53;
                    r2r[k] = nf ? 38 : 6;
                }
                uint64_t clr = ((1ull << (r1r[k] & 63)) | (1ull << (r2r[k] & 63))) & ~1ull;
                lt[k] = (f < (int64_t)life.cap + 8) ? life.p[f] : 0;
                cl[k] = lt[k] & ~clr;
                rd[k] = nc[k] ? R(&c->rd, off[k]) : 0;
                ord[k] = rd[k];
                any_rdv |= S(rd[k]) > 63;
            }
            if (any_rdv) {
                for (int k = 0; k < 8; ++k) {
                    int rdf = ins[k] == 0xD0007053u || (ins[k] & 0x7F) == 0x53;
                    if (S(ord[k]) > 63) {
                        uint64_t adj = cl[k] | (rdf ? 0xFFFFFFFFull : 0xFFFFFFFF00000000ull);
                        rd[k] = ffz_gcc(adj);
                    }
                    if (rd[k] == 64) rd[k] = rdf ? 37 : 5;
                    if (!nc[k]) rd[k] = 0;
                }
            }
            int sw1[8], sw2[8], swd[8];
            uint64_t rdm[8];
            for (int k = 0; k < 8; ++k) {
                uint64_t m1 = nc[k] ? (1ull << (r1r[k] & 63)) : 0;
                uint64_t m2 = nc[k] ? (1ull << (r2r[k] & 63)) : 0;
                rdm[k] = nc[k] ? (rd[k] < 64 ? (1ull << rd[k]) : 0) : 0;
                sw1[k] = r1r[k] != 0 && (lt[k] & m1) != 0;
                sw2[k] = r2r[k] != 0 && (lt[k] & m2) != 0;
                swd[k] = rd[k] != 0 && (cl[k] & rdm[k]) != 0;
            }
            for (int k = 0; k < 8; ++k) if (swd[k]) { int32_t x = snap[rd[k] & 63] - 64; if (x >= 0) W(&c->sym_sw, x, 0xFFFFFFFFu); }
            for (int k = 0; k < 8; ++k) if (sw1[k]) { int32_t x = snap[r1r[k] & 63] - 64; if (x >= 0) W(&c->sym_sw, x, 0xFFFFFFFFu); }
            for (int k = 0; k < 8; ++k) if (sw2[k]) { int32_t x = snap[r2r[k] & 63] - 64; if (x >= 0) W(&c->sym_sw, x, 0xFFFFFFFFu); }
            for (int k = 0; k < 8; ++k) if (nc[k]) life.p[j * 8 + k] = cl[k] | rdm[k];
            for (int k = 0; k < 8; ++k) if (ord[k]) regstate[k * 64 + (rd[k] & 63)] = (int32_t)ord[k];
            for (int k = 0; k < 8; ++k) if (nc[k] && r1r[k] != rd[k]) regstate[k * 64 + (r1r[k] & 63)] = -1;
            for (int k = 0; k < 8; ++k) if (nc[k] && r2r[k] != rd[k]) regstate[k * 64 + (r2r[k] & 63)] = -1;
            for (int k = 0; k < 8; ++k) if (S(ord[k]) > 63) { W(&c->sym_reg, S(ord[k] - 64), rd[k]); W(&c->sym_sw, S(ord[k] - 64), 0); }
            for (int k = 0; k < 8; ++k) if (S(rs1[k]) > 63) { W(&c->sym_reg, S(rs1[k] - 64), r1r[k]); W(&c->sym_sw, S(rs1[k] - 64), 0); }
            for (int k = 0; k < 8; ++k) if (S(rs2[k]) > 63) { W(&c->sym_reg, S(rs2[k] - 64), r2r[k]); W(&c->sym_sw, S(rs2[k] - 64), 0); }
        }
    }
    vb64_free(&life);

    /* function-start flags and the copy "segmented scan" (:1422-1469) */
    vb starts_b = vb_new(W0);
    for (uint32_t f = 0; f < F; ++f) W(&starts_b, c->fstarts.p[f], 1);
    vb spill = vb_new(W0);
    for (int64_t v = 0; v < vecs(&spill); ++v)
        for (int l = 0; l < 8; ++l) W(&spill, v * 8 + l, 1u & R(&c->sym_sw, v * 8 + l));
    for (int64_t i = 0; i < W0; ++i)
        if (!R(&starts_b, i)) W(&spill, i, i > 0 ? R(&spill, i - 1) : 0);
    /* instruction counts (:1473-1495); preserve counts add 0 (:1498-1519) */
    vb counts = vb_new(W0);
    for (int64_t v = 0; v < vecs(&counts); ++v)
        for (int l = 0; l < 8; ++l) {
            int64_t i = v * 8 + l;
            uint32_t en = R(&c->used, i);
            if (!en) continue;
            uint32_t r1 = R(&c->rs1, i), r2 = R(&c->rs2, i), rdv = R(&c->rd, i);
            uint32_t s1 = S(r1) > 63 ? R(&c->sym_sw, S(r1 - 64)) : 0;
            uint32_t s2 = S(r2) > 63 ? R(&c->sym_sw, S(r2 - 64)) : 0;
            uint32_t sd = S(rdv) > 63 ? R(&c->sym_sw, S(rdv - 64)) : 0;
            uint32_t base = 1;
            W(&counts, i, base + (s1 ? 1 : 0) + (s2 ? 1 : 0) + (sd ? 1 : 0));
        }
    c->ra_offsets = vb_new(W0);
    excl_scan_vec(&counts, &c->ra_offsets);
    const uint32_t W1 = (W0 == 0) ? 0 : R(&c->ra_offsets, W0 - 1) + 1;
    vb ni = vb_new(W1), nrd = vb_new(W1), nr1 = vb_new(W1), nr2 = vb_new(W1), njt = vb_new(W1);
    /* emit (:1552-1697) */
    for (int64_t v = 0; v < vecs(&c->instr); ++v) {
        int any = 0;
        for (int l = 0; l < 8; ++l) any |= R(&c->used, v * 8 + l) != 0;
        if (!any) continue;
        int32_t base[8], l1[8], l2[8], mo[8], rdo[8];
        uint32_t a1[8], a2[8], ad[8], i1w[8], i2w[8], sdw[8], inw[8];
        for (int l = 0; l < 8; ++l) {
            int64_t i = v * 8 + l;
            uint32_t um = R(&c->used, i);
            base[l] = um ? S(R(&c->ra_offsets, i)) : -1;
            uint32_t ins = um ? R(&c->instr, i) : 0;
            inw[l] = ins;
            uint32_t r1 = um ? R(&c->rs1, i) : 0;
            int v1 = S(r1) > 63;
            l1[l] = -1; uint32_t st1 = 0; a1[l] = v1 ? 0 : r1;
            if (v1) {
                uint32_t sw = R(&c->sym_sw, S(r1 - 64));
                l1[l] = sw ? base[l] : -1;
                st1 = R(&spill, S(r1 - 64));
                int nf = ins == 0xC0007053u || (ins & 0x7F) == 0x53;
                a1[l] = sw ? (nf ? 37u : 5u) : R(&c->sym_reg, S(r1 - 64));
            }
            uint32_t r2 = um ? R(&c->rs2, i) : 0;
            int v2 = S(r2) > 63;
            l2[l] = -1; uint32_t st2 = 0; a2[l] = v2 ? 0 : r2;
            if (v2) {
                uint32_t sw = R(&c->sym_sw, S(r2 - 64));
                /* D6: compares its own still -1 load offset */
                l2[l] = sw ? base[l] + (l2[l] > 0 ? 1 : 0) : -1;
                st2 = R(&spill, S(r2 - 64));
                int nf = (ins & 0x7F) == 0x53;
                a2[l] = sw ? (nf ? 38u : 6u) : R(&c->sym_reg, S(r2 - 64));
            }
            mo[l] = base[l] + (l1[l] > 0 ? 1 : 0) + (l2[l] > 0 ? 1 : 0);
            uint32_t rdv = um ? R(&c->rd, i) : 0;
            int vd = S(rdv) > 63;
            rdo[l] = -1; uint32_t std_ = 0; ad[l] = vd ? 0 : rdv;
            if (vd) {
                uint32_t sw = R(&c->sym_sw, S(rdv - 64));
                rdo[l] = sw ? mo[l] + 1 : -1;
                std_ = R(&spill, S(rdv - 64));
                ad[l] = R(&c->sym_reg, S(rdv - 64));
            }
            i1w[l] = (((st1 - 1) * 4) << 20) | 0x2003u;
            i2w[l] = (((st2 - 1) * 4) << 20) | 0x2003u;
            uint32_t imm = (std_ - 1) * (uint32_t)-4;
            sdw[l] = 0x2023u | ((imm & 0x1F) << 7) | ((imm & 0xFE) << 19);
        }
        for (int l = 0; l < 8; ++l) if (l1[l] >= 0) { W(&ni, l1[l], i1w[l]); W(&nrd, l1[l], 5); }
        for (int l = 0; l < 8; ++l) if (l2[l] >= 0) { W(&ni, l2[l], i2w[l]); W(&nrd, l2[l], 6); }
        for (int l = 0; l < 8; ++l) if (mo[l] >= 0) {
            int64_t i = v * 8 + l;
            W(&ni, mo[l], R(&c->instr, i)); W(&nrd, mo[l], ad[l]); W(&nr1, mo[l], a1[l]);
            W(&nr2, mo[l], a2[l]); W(&njt, mo[l], R(&c->ra_offsets, R(&c->jt, i)));
        }
        for (int l = 0; l < 8; ++l) if (rdo[l] >= 0) { W(&ni, rdo[l], sdw[l]); W(&nr1, rdo[l], ad[l]); }
        (void)inw;
    }
    vb_free(&c->instr); vb_free(&c->rd); vb_free(&c->rs1); vb_free(&c->rs2); vb_free(&c->jt);
    c->instr = vb_copy(&ni); c->rd = vb_copy(&nrd); c->rs1 = vb_copy(&nr1); c->rs2 = vb_copy(&nr2); c->jt = vb_copy(&njt);
    vb_free(&ni); vb_free(&nrd); vb_free(&nr1); vb_free(&nr2); vb_free(&njt);
    fix_func_tab(c, &c->ra_offsets);
    /* frame words (:1714-1811): stack = (2 + 0 + 0 + 0) * 4 = 8, no preserve stores */
    for (int64_t v = 0; v < vecs(&c->fstarts); ++v) {
        int valid[8]; uint32_t st[8], en[8];
        for (int l = 0; l < 8; ++l) {
            int64_t k = v * 8 + l;
            valid[l] = k < (int64_t)F;
            st[l] = valid[l] ? R(&c->fstarts, k) : 0;
            en[l] = st[l] + (valid[l] ? R(&c->fsizes, k) : 0);
        }
        const uint32_t stack = 8;
        const uint32_t upper = (stack & 0xFFFFF000u) | 0x437u;
        const uint32_t lower = ((stack & 0xFFFu) << 20) | 0x40413u;
        uint32_t pos[4][8];
        for (int l = 0; l < 8; ++l) { pos[0][l] = st[l] + 2; pos[1][l] = st[l] + 3; pos[2][l] = en[l] - 6; pos[3][l] = en[l] - 5; }
        for (int q = 0; q < 4; ++q)
            for (int l = 0; l < 8; ++l) if (valid[l]) {
                uint32_t x = pos[q][l];
                W(&c->instr, x, (q & 1) ? lower : upper);
                W(&c->rd, x, 0); W(&c->rs1, x, 0); W(&c->rs2, x, 0); W(&c->jt, x, 0);
            }
    }
    vb_free(&starts_b); vb_free(&spill); vb_free(&counts);
}

/* -------------------------------------------------------------- fix_jumps
 * rv_generator_avx.cpp:1853-1993 */
static void st_fix_jumps(ptlo_ctx *c) {
    const int64_t W1 = c->instr.n;
    vb sz = vb_new(W1);
    for (int64_t v = 0; v < vecs(&c->instr); ++v)
        for (int l = 0; l < 8; ++l) {
            uint32_t w = R(&c->instr, v * 8 + l);
            int jmp = ((w & 0x7F) == 0x67) && (w != 0x8067u);
            W(&sz, v * 8 + l, jmp ? 2 : (w != 0 ? 1 : 0));
        }
    c->fj_offsets = vb_new(W1);
    excl_scan_vec(&sz, &c->fj_offsets);
    const uint32_t W2 = (W1 == 0) ? 0 : R(&sz, W1 - 1) + R(&c->fj_offsets, W1 - 1);
    vb ni = vb_new(W2), nrd = vb_new(W2), nr1 = vb_new(W2), nr2 = vb_new(W2), njt = vb_new(W2);
    for (int64_t i = 0; i < W1; ++i) {
        uint32_t k = R(&c->fj_offsets, i);
        W(&ni, k, R(&c->instr, i)); W(&nrd, k, R(&c->rd, i)); W(&nr1, k, R(&c->rs1, i));
        W(&nr2, k, R(&c->rs2, i)); W(&njt, k, R(&c->jt, i));
    }
    for (int64_t v = 0; v < vecs(&njt); ++v)
        for (int l = 0; l < 8; ++l) {
            int64_t k = v * 8 + l;
            W(&njt, k, R(&c->fj_offsets, S(R(&njt, k))));
        }
    for (int64_t v = 0; v < vecs(&c->instr); ++v) {
        uint32_t idx[8], ins[8], tgt[8]; int jm[8], br[8], anyj = 0, anyb = 0;
        for (int l = 0; l < 8; ++l) {
            idx[l] = R(&c->fj_offsets, v * 8 + l);
            ins[l] = R(&ni, S(idx[l]));
            jm[l] = ((ins[l] & 0x7F) == 0x67) && (ins[l] != 0x8067u);
            br[l] = (ins[l] & 0x7F) == 0x63;
            anyj |= jm[l]; anyb |= br[l];
        }
        if (!(anyj || anyb)) continue;
        for (int l = 0; l < 8; ++l) tgt[l] = (jm[l] || br[l]) ? 4 * R(&njt, S(idx[l])) : 0;
        if (anyj) {
            for (int l = 0; l < 8; ++l) if (jm[l]) {
                uint32_t delta = tgt[l] - idx[l] * 4;
                uint32_t upper = (delta & 0xFFFFF000u) >> 12, lower = delta & 0xFFFu, sign = (delta >> 11) & 1;
                uint32_t auipc = ((upper + sign) << 12) | 0x97u;
                uint32_t jw = ins[l] | (lower << 20) | (1u << 15);
                uint32_t k = idx[l];
                W(&ni, k, auipc); W(&njt, k, 0); W(&nrd, k, 0); W(&nr1, k, 0); W(&nr2, k, 0);
                W(&ni, k + 1, jw); W(&njt, k + 1, 0); W(&nrd, k + 1, 0); W(&nr1, k + 1, 0); W(&nr2, k + 1, 0);
            }
        }
        if (anyb) {
            for (int l = 0; l < 8; ++l) if (br[l]) {
                uint32_t delta = tgt[l] - idx[l] * 4;
                uint32_t sign = (delta >> 12) & 1, b11 = (delta >> 11) & 1;
                uint32_t b105 = (delta >> 5) & 0x1F, b14 = (delta >> 1) & 0xF;
                W(&ni, idx[l], ins[l] | (sign << 31) | (b11 << 7) | (b105 << 25) | (b14 << 8));
            }
        }
    }
    vb_free(&c->instr); vb_free(&c->rd); vb_free(&c->rs1); vb_free(&c->rs2); vb_free(&c->jt);
    c->instr = vb_copy(&ni); c->rd = vb_copy(&nrd); c->rs1 = vb_copy(&nr1); c->rs2 = vb_copy(&nr2); c->jt = vb_copy(&njt);
    vb_free(&ni); vb_free(&nrd); vb_free(&nr1); vb_free(&nr2); vb_free(&njt); vb_free(&sz);
    fix_func_tab(c, &c->fj_offsets);
}

/* ------------------------------------------------------------ postprocess
 * rv_generator_avx.cpp:1995-2011 */
static void st_postprocess(ptlo_ctx *c) {
    for (int64_t v = 0; v < vecs(&c->instr); ++v)
        for (int l = 0; l < 8; ++l) {
            int64_t i = v * 8 + l;
            uint32_t w = R(&c->instr, i);
            w |= (R(&c->rd, i) & 31) << 7;
            w |= (R(&c->rs1, i) & 31) << 15;
            w |= (R(&c->rs2, i) & 31) << 20;
            W(&c->instr, i, w);
        }
}

/* Runs stages up to and including `upto` (1 = preprocess ... 7 = postprocess).
 * The stage order is the driver's (src/codegen/rv_generator.cpp:142-148).
 * Returns the accumulated flags. */
uint32_t ptlo_run(ptlo_ctx *c, int upto) {
    g_flags = c->flags;
    if (c->stage_done < 1 && upto >= 1) { st_preprocess(c); c->stage_done = 1; }
    if (c->stage_done < 2 && upto >= 2) { st_isn_cnt(c); c->stage_done = 2; }
    if (c->F == 0 && upto >= 3) { g_flags |= OF_EMPTY; c->flags = g_flags; return g_flags; }
    if (c->stage_done < 3 && upto >= 3) { st_isn_gen(c); c->stage_done = 3; }
    if (c->stage_done < 4 && upto >= 4) { st_optimize(c); c->stage_done = 4; }
    if (c->stage_done < 5 && upto >= 5) { st_regalloc(c); c->stage_done = 5; }
    if (c->stage_done < 6 && upto >= 6) { st_fix_jumps(c); c->stage_done = 6; }
    if (c->stage_done < 7 && upto >= 7) { st_postprocess(c); c->stage_done = 7; }
    c->flags = g_flags;
    return g_flags;
}

/* Named stage-boundary arrays for the tests: returns the element pointer and
 * count (u32 elements), or NULL. */
const uint32_t *ptlo_get(ptlo_ctx *c, const char *name, int64_t *count) {
    struct { const char *n; vb *b; } tab[] = {
        {"types", &c->types}, {"rtypes", &c->rtypes}, {"node_data", &c->nd}, {"sizes", &c->sizes},
        {"locs", &c->locs}, {"func_ids", &c->fids}, {"func_starts", &c->fstarts},
        {"func_ends", &c->fends}, {"func_sizes", &c->fsizes}, {"instr", &c->instr}, {"rd", &c->rd},
        {"rs1", &c->rs1}, {"rs2", &c->rs2}, {"jt", &c->jt}, {"used", &c->used},
        {"sym_reg", &c->sym_reg}, {"sym_sw", &c->sym_sw}, {"ra_offsets", &c->ra_offsets},
        {"fj_offsets", &c->fj_offsets}, {"sorted", &c->idx_sorted},
    };
    for (unsigned i = 0; i < sizeof(tab) / sizeof(tab[0]); ++i)
        if (strcmp(tab[i].n, name) == 0) { *count = tab[i].b->n; return tab[i].b->p; }
    *count = 0; return NULL;
}

uint32_t ptlo_num_functions(ptlo_ctx *c) { return c->F; }
uint32_t ptlo_opt_iters(ptlo_ctx *c) { return c->opt_iters; }
uint64_t ptlo_ties(ptlo_ctx *c) { return c->ties; }

/* The scan primitive on its own, for the known-answer test mirroring
 * benchmarks/prefix_sum.cpp: exclusive scan with element -1 = 0. */
void ptlo_exclusive_scan(const uint32_t *in, uint32_t *out, int64_t n) {
    vb a = vb_new(n), b = vb_new(n);
    memcpy(a.p, in, (size_t)n * 4);
    g_flags = 0;
    excl_scan_vec(&a, &b);
    memcpy(out, b.p, (size_t)n * 4);
    vb_free(&a); vb_free(&b);
}

/* std::sort restated, exposed so a test can compare it with the host's own
 * libstdc++ std::sort on the same keys. */
void ptlo_introsort(uint32_t *idx, int64_t n, const int32_t *key) { sort_introsort(idx, n, key); }

/* ------------------------------------------------------------ text output --
 * rv_generator::to_asm (src/codegen/rv_generator.cpp:114-120): every word as
 * rvdisasm::instruction(word) (src/disassembler.cpp:453-462), then '\n'.  That
 * is instr_name (:63-222), one space, and format_args (:224-407) with no colour
 * and address 0.  The reference's switches fall through wherever an inner
 * switch finds no match; C's own fall-through reproduces that here, case for
 * case in the reference's source order. */
static const char *const ASM_X[32] = {
    "zero", "ra", "sp", "gp", "tp", "t0", "t1", "t2", "s0", "s1", "a0", "a1", "a2", "a3", "a4", "a5",
    "a6", "a7", "s2", "s3", "s4", "s5", "s6", "s7", "s8", "s9", "s10", "s11", "t3", "t4", "t5", "t6"};
static const char *const ASM_F[32] = {
    "ft0", "ft1", "ft2", "ft3", "ft4", "ft5", "ft6", "ft7", "fs0", "fs1", "fa0", "fa1", "fa2", "fa3", "fa4", "fa5",
    "fa6", "fa7", "fs2", "fs3", "fs4", "fs5", "fs6", "fs7", "fs8", "fs9", "fs10", "fs11", "ft8", "ft9", "ft10", "ft11"};

/* instr_name, src/disassembler.cpp:83-222 */
static const char *asm_name(uint32_t w) {
    const uint32_t op = w & 0x7F, f3 = (w >> 12) & 7, f7 = (w >> 25) & 0x7F, r2 = (w >> 20) & 0x1F;
    switch (op) {
        case 0x37: return "lui";
        case 0x17: return "auipc";
        case 0x6F: return "jal";
        case 0x67: return "jalr";
        case 0x63:
            switch (f3) { case 0: return "beq"; case 1: return "bne"; case 4: return "blt"; case 5: return "bge";
                          case 6: return "bltu"; case 7: return "bgeu"; }
            /* fall through */
        case 0x03:
            switch (f3) { case 0: return "lb"; case 1: return "lh"; case 2: return "lw"; case 4: return "lbu";
                          case 5: return "lhu"; }
            /* fall through */
        case 0x23:
            switch (f3) { case 0: return "sb"; case 1: return "sh"; case 2: return "sw"; }
            /* fall through */
        case 0x13:
            switch (f3) {
                case 0: return "addi"; case 1: return "slli"; case 2: return "slti"; case 3: return "sltiu";
                case 4: return "xori";
                case 5: if (f7 == 0x00) return "srli"; if (f7 == 0x20) return "srai"; /* fall through */
                case 6: return "ori";
                case 7: return "andi";
            }
            /* fall through */
        case 0x33:
            switch (f3) {
                case 0: if (f7 == 0) return "add"; if (f7 == 1) return "mul"; if (f7 == 0x20) return "sub"; /* fall through */
                case 1: if (f7 == 0) return "sll"; if (f7 == 1) return "mulh"; /* fall through */
                case 2: if (f7 == 0) return "slt"; if (f7 == 1) return "mulhsu"; /* fall through */
                case 3: if (f7 == 0) return "sltu"; if (f7 == 1) return "mulhu"; /* fall through */
                case 4: if (f7 == 0) return "xor"; if (f7 == 1) return "div"; /* fall through */
                case 5: if (f7 == 0) return "srl"; if (f7 == 1) return "divu"; if (f7 == 0x20) return "sra"; /* fall through */
                case 6: if (f7 == 0) return "or"; if (f7 == 1) return "rem"; /* fall through */
                case 7: if (f7 == 0) return "and"; if (f7 == 1) return "remu";
            }
            /* fall through */
        case 0x0F: return "fence";
        case 0x73:
            if (f7 == 0) return "ecall";
            if (f7 == 1) return "ebreak";
            /* fall through */
        case 0x07: return "flw";
        case 0x27: return "fsw";
        case 0x43: return "fmadd.s";
        case 0x47: return "fmsub.s";
        case 0x4B: return "fnmsub.s";
        case 0x4F: return "fnmadd.s";
        case 0x53:
            switch (f7) {
                case 0x00: return "fadd.s";
                case 0x04: return "fsub.s";
                case 0x08: return "fmul.s";
                case 0x0C: return "fdiv.s";
                case 0x2C: return "fsqrt.s";
                case 0x10: if (f3 == 0) return "fsgnj.s"; if (f3 == 1) return "fsgnjn.s"; if (f3 == 2) return "fsgnjx.s"; /* fall through */
                case 0x14: if (f3 == 0) return "fmin.s"; if (f3 == 1) return "fmax.s"; /* fall through */
                case 0x60: if (r2 == 0) return "fcvt.w.s"; if (r2 == 1) return "fcvt.wu.s"; /* fall through */
                case 0x70: if (f3 == 0) return "fmv.x.w"; if (f3 == 1) return "fclass.s"; /* fall through */
                case 0x50: if (f3 == 0) return "fle.s"; if (f3 == 1) return "flt.s"; if (f3 == 2) return "feq.s"; /* fall through */
                case 0x68: if (r2 == 0) return "fcvt.s.w"; if (r2 == 1) return "fcvt.s.wu"; /* fall through */
                case 0x78: return "fmv.w.x";
            }
    }
    return w ? "unknown" : "(zero)";
}

/* instr_type, :56-81 */
enum { AT_UNKNOWN, AT_R, AT_I, AT_S, AT_B, AT_U, AT_J, AT_R4 };
static int asm_type(uint32_t w) {
    switch (w & 0x7F) {
        case 0x37: case 0x17: return AT_U;
        case 0x6F: return AT_J;
        case 0x67: case 0x03: case 0x13: case 0x0F: case 0x73: case 0x07: return AT_I;
        case 0x63: return AT_B;
        case 0x23: case 0x27: return AT_S;
        case 0x33: case 0x53: return AT_R;
        case 0x43: case 0x47: case 0x4B: case 0x4F: return AT_R4;
    }
    return AT_UNKNOWN;
}

/* format_args (:224-407): the operands, with the reference's colour escapes
 * around registers and immediates when `color`, branch targets relative to
 * `addr`.  Returns the bytes written (at most 200). */
#define C_WHITE "\033[37m"
#define C_INSTR "\033[38;5;186m"
#define C_REG "\033[38;5;117m"
#define C_IMM "\033[38;5;114m"
#define C_EXTRA "\033[38;5;114m"
static int asm_args(uint32_t w, char *out, int color, uint64_t addr) {
    const uint32_t op = w & 0x7F, f3 = (w >> 12) & 7, f7 = (w >> 25) & 0x7F;
    const uint32_t rd = (w >> 7) & 31, rs1 = (w >> 15) & 31, rs2 = (w >> 20) & 31, rs3 = (w >> 27) & 31;
    char x[4][40];
    const char *r0 = color ? C_REG : "", *r1 = color ? C_WHITE : "";
    const char *i0 = color ? C_IMM : "", *i1 = color ? C_WHITE : "";
#define XR(k, name) (sprintf(x[k], "%s%s%s", r0, (name), r1), x[k])
    char *p = out;
    switch (asm_type(w)) {
        case AT_R:
            if (op == 0x53) {
                if (f7 == 0x78 || f7 == 0x68) p += sprintf(p, "%s, %s", XR(0, ASM_F[rd]), XR(1, ASM_X[rs1]));
                else if (f7 == 0x60 || (f7 == 0x70 && f3 == 0)) p += sprintf(p, "%s, %s", XR(0, ASM_X[rd]), XR(1, ASM_F[rs1]));
                else p += sprintf(p, "%s, %s, %s", XR(0, ASM_F[rd]), XR(1, ASM_F[rs1]), XR(2, ASM_F[rs2]));
            } else {
                p += sprintf(p, "%s, %s, %s", XR(0, ASM_X[rd]), XR(1, ASM_X[rs1]), XR(2, ASM_X[rs2]));
            }
            break;
        case AT_I: {
            uint32_t imm = w >> 20;
            if (imm >> 11) imm |= 0xFFFFF000u;
            if (op == 0x07 || op == 0x03 || op == 0x67)
                p += sprintf(p, "%s, %s%d%s(%s)", op == 0x07 ? XR(0, ASM_F[rd]) : XR(0, ASM_X[rd]), i0, (int32_t)imm, i1,
                             XR(1, ASM_X[rs1]));
            else
                p += sprintf(p, "%s, %s, %s%d%s", XR(0, ASM_X[rd]), XR(1, ASM_X[rs1]), i0, (int32_t)imm, i1);
            break;
        }
        case AT_S: {
            uint32_t imm = ((w >> 7) & 0x1F) | ((w >> 20) & 0xFE0);
            if (imm >> 11) imm |= 0xFFFFF000u;
            /* an fsw names its register by the rs1 field (:357-361) */
            p += sprintf(p, "%s, %s%d%s(%s)", op == 0x27 ? XR(0, ASM_F[rs1]) : XR(0, ASM_X[rs2]), i0, (int32_t)imm, i1,
                         XR(1, ASM_X[rs1]));
            break;
        }
        case AT_B: {
            /* bits 4:1, 9:5 (instr[29:25]: bit 10 is not taken) and 11; the
             * sign fills bits 31:11 (:366-372) */
            uint32_t imm = ((w >> 7) & 0x1E) | ((w >> 20) & 0x3E0) | (((w >> 7) & 1u) << 11);
            if (w >> 31) imm |= 0xFFFFF800u;
            const uint64_t tgt = addr + (uint64_t)(int64_t)(int32_t)imm;  /* uint64 address + offset */
            if (tgt) p += sprintf(p, "%s, %s, %s0x%llx%s", XR(0, ASM_X[rs1]), XR(1, ASM_X[rs2]), i0, (unsigned long long)tgt, i1);
            else p += sprintf(p, "%s, %s, %s0%s", XR(0, ASM_X[rs1]), XR(1, ASM_X[rs2]), i0, i1);  /* showbase prints 0 bare */
            break;
        }
        case AT_U:
            p += sprintf(p, "%s, %s%d%s", XR(0, ASM_X[rd]), i0, (int32_t)(w & 0xFFFFF000u), i1);
            break;
        case AT_J: {
            /* bits 10:1, 11, 19:12; the sign fills bits 31:19 (:384-392) */
            uint32_t imm = ((w >> 20) & 0x7FE) | (((w >> 20) & 1u) << 11) | (w & 0xFF000u);
            if (w >> 31) imm |= 0xFFF80000u;
            p += sprintf(p, "%s, %s%d%s", XR(0, ASM_X[rd]), i0, (int32_t)imm, i1);
            break;
        }
        case AT_R4:
            p += sprintf(p, "%s, %s, %s, %s", XR(0, ASM_X[rd]), XR(1, ASM_X[rs1]), XR(2, ASM_X[rs2]), XR(3, ASM_X[rs3]));
            break;
        default:
            break;
    }
#undef XR
    return (int)(p - out);
}

/* One line: instr_name, ' ', format_args (no colour, address 0), '\n'.
 * Returns its length; out needs 64 bytes. */
static int asm_line(uint32_t w, char *out) {
    int n = sprintf(out, "%s ", asm_name(w));
    n += asm_args(w, out + n, 0, 0);
    out[n++] = '\n';
    out[n] = 0;
    return n;
}

/* to_asm of n words into out (cap bytes); returns the text length, or -1 when
 * it does not fit. */
int64_t ptlo_asm(const uint32_t *words, int64_t n, char *out, int64_t cap) {
    int64_t len = 0;
    char line[64];
    for (int64_t i = 0; i < n; ++i) {
        const int k = asm_line(words[i], line);
        if (len + k > cap) return -1;
        memcpy(out + len, line, (size_t)k);
        len += k;
    }
    return len;
}

/* rvdisasm::disassemble(os, words, 0, func_starts) (src/disassembler.cpp:411-451),
 * as rv_generator::print(os, true) calls it (rv_generator.cpp:100-103): a
 * header line, then per word an optional "func<k>:" line (k counts the labels
 * printed so far, in hex after word 0; a word gets one when any entry of
 * func_starts equals its index) and its line: index (decimal, right-aligned, count_digits wide), three
 * spaces, byte address (hex, zero-filled to digits), ": ", the four bytes high
 * first (" " and two hex digits each), three spaces, the name left-aligned in
 * nine columns, a space and the coloured operands.  digits and count_digits as
 * :416-417 compute them.  Returns the length, or -1 when cap is too small. */
int64_t ptlo_listing(const uint32_t *words, int64_t n, const uint32_t *func_starts, int64_t nf, char *out,
                     int64_t cap) {
    if (n <= 0) return 0;
    const size_t digits = (size_t)ceil(log((double)n) / log(16.0)) + 1;
    const size_t count_digits = (size_t)log10((double)n) + 1;
    int64_t len = 0;
    char line[512];
    int k = snprintf(line, sizeof line, C_EXTRA "%lld" C_WHITE " instructions, starting at " C_IMM "0x%0*llx" C_WHITE "\n",
                     (long long)n, (int)digits, 0ull);
    if (k > cap) return -1;
    memcpy(out, line, (size_t)k);
    len = k;
    long long func_index = 0;
    for (int64_t i = 0; i < n; ++i) {
        const uint32_t w = words[i];
        int m = 0;
        int is_start = 0;
        for (int64_t f = 0; f < nf && !is_start; ++f) is_start = func_starts[f] == (uint64_t)i;
        if (is_start) {
            m += sprintf(line + m, "%*s", (int)(count_digits + 3 + digits + 2 + 12 + 1), "");
            /* the stream is still in hex from the previous word's line
             * (format_args restores the flags it found): only a label before
             * word 0 is decimal */
            m += sprintf(line + m, i == 0 ? "func%lld:\n" : "func%llx:\n", func_index++);
        }
        m += sprintf(line + m, C_EXTRA "%*lld" C_WHITE "   %0*llx: ", (int)count_digits, (long long)i, (int)digits,
                     (unsigned long long)(4 * i));
        for (int b = 3; b >= 0; --b) m += sprintf(line + m, " " C_EXTRA "%02x" C_WHITE, (w >> (8 * b)) & 0xFF);
        m += sprintf(line + m, "   " C_INSTR "%-9s" C_WHITE " ", asm_name(w));
        m += asm_args(w, line + m, 1, 4 * (uint64_t)i);
        line[m++] = '\n';
        if (len + m > cap) return -1;
        memcpy(out + len, line, (size_t)m);
        len += m;
    }
    return len;
}

#include "ptl_scalar.inc"

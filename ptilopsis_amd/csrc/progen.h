// Parameters of the seeded synthetic-program generator (progen.cpp).
#ifndef PTL_PROGEN_H
#define PTL_PROGEN_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif
typedef struct {
    uint64_t seed;
    uint32_t n_functions;
    uint32_t min_args, max_args;        /* parameters per function */
    uint32_t min_stmts, max_stmts;      /* top-level statements per body (no size target) */
    uint32_t nested_min, nested_max;    /* statements per if/else/while body */
    uint32_t max_nest;                  /* maximum if/while nesting */
    double p_assign, p_if, p_while, p_return;  /* statement mix; calls take the rest */
    double p_var_leaf;                  /* leaf is a variable (else a literal) */
    double expr_p;                      /* geometric depth parameter: mean depth 1/p */
    double p_leaf_operand;              /* the non-deepest operand is a leaf */
    double p_arith, p_compare;          /* int operator mix; && and || take the rest */
    double float_frac;                  /* share of float-typed variables and operands */
    uint32_t balanced;                  /* 1: wide balanced expressions */
    uint32_t balanced_min, balanced_max;/* their depth range */
    uint32_t chain_min, chain_max;      /* 0, or the depth range of nested if/while chains */
    double p_chain;                     /* a top-level statement is such a chain */
    uint64_t total_target;              /* 0, or stop adding functions once the program has this many nodes */
} ptl_gen_params;

/* Generates one program.  `targets` (may be NULL) gives per-function node-count
 * targets; without it each body has min_stmts..max_stmts statements.
 * Returns 0 on success; *out_src must be released with ptl_free_source. */
int ptl_generate_source(const ptl_gen_params *p, const uint32_t *targets, char **out_src,
                        uint64_t *out_len, uint64_t *out_nodes);
void ptl_free_source(char *s);
#ifdef __cplusplus
}
#endif
#endif

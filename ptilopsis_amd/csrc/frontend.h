// Flattened inverted-tree arrays as produced by the host front end; the same
// six arrays as the reference's DepthTree (pareas-ast include/codegen/depthtree.hpp:12-58).
#ifndef PTL_FRONTEND_H
#define PTL_FRONTEND_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif
typedef struct {
    uint8_t *node_types;     /* pareas NodeType */
    uint8_t *result_types;   /* pareas DataType */
    int32_t *parents;        /* root: -1 */
    int32_t *depth;
    int32_t *child_idx;      /* root: -1 */
    uint32_t *node_data;
    uint64_t n_nodes;
    uint32_t max_depth;
    uint32_t n_functions;
} ptl_host_tree;

/* Parses Pareas-language source; 0 on success, 1 on a parse/type error (message in err). */
int ptl_parse_source(const char *src, uint64_t len, ptl_host_tree *out, char *err, uint64_t err_len);
void ptl_free_host_tree(ptl_host_tree *t);

/* The same parse stopped before flattening: the typed node records in the
 * order the parser made them (children before their parent, implicit DEREF
 * nodes last).  Node v's children, in child order, are
 * kids[first_child[v] .. first_child[v] + n_children[v]); `root` is the
 * program's STATEMENT_LIST.  Every record belongs to the tree. */
typedef struct {
    uint8_t *node_types;
    uint8_t *result_types;
    uint32_t *node_data;
    uint32_t *first_child;
    uint32_t *n_children;
    uint32_t *kids;
    uint64_t n_nodes, n_kids;
    uint32_t root;
    uint32_t n_functions;
} ptl_host_arena;
int ptl_parse_arena(const char *src, uint64_t len, ptl_host_arena *out, char *err, uint64_t err_len);
void ptl_free_host_arena(ptl_host_arena *a);
/* The host flattening of records (what ptl_parse_source does after parsing),
 * one thread; 0 on success.  The records are trusted (no checks). */
int ptl_flatten_records(const ptl_host_arena *a, ptl_host_tree *out);
#ifdef __cplusplus
}
#endif
#endif

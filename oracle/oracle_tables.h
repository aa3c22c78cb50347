/* Oracle copy of the instruction-selection tables of the reference generator.
 *
 * TEST INFRASTRUCTURE ONLY: this header belongs to the CPU oracle under
 * oracle/.  The shipped library carries its own, separately written packed
 * form of the same facts (ptilopsis_amd/csrc/rv_spec.h); a test expands both
 * and requires them to agree entry by entry, and a second test decodes every
 * instruction template with an independent RISC-V decoder.
 *
 * Source of every row: include/codegen/rv_nodetype.hpp
 *   node/data type enums           :13-95
 *   node_size_mapping              :197-259
 *   has_instr_mapping              :266-326
 *   has_output                     :333-393
 *   parent_arg_idx_lookup          :401-461
 *   get_output_table               :468-558
 *   instr_table                    :563-858
 *   instr_constant_table           :863-925
 *   operand_table                  :930-1008
 *   instr_jt_table                 :1013-1075
 *   pareas_to_rv_nodetype          :1087-1134
 * Rows are written per node type, mirroring the reference's layout
 * ([type][offset][data type]); types the reference never lists stay zero, as
 * its value-initialised mapping_helper does.
 */
#ifndef PTL_ORACLE_TABLES_H
#define PTL_ORACLE_TABLES_H
#include <stdint.h>
#include <string.h>

enum {
    O_INVALID = 0, O_STMT_LIST = 1, O_EMPTY = 2, O_EXPR = 3,
    O_ADD = 4, O_SUB = 5, O_MUL = 6, O_DIV = 7, O_MOD = 8, O_BITAND = 9, O_BITOR = 10,
    O_BITXOR = 11, O_LSHIFT = 12, O_RSHIFT = 13, O_URSHIFT = 14, O_LAND = 15, O_LOR = 16,
    O_BITNOT = 17, O_LNOT = 18, O_NEG = 19, O_LIT = 20, O_CAST = 21, O_DEREF = 22,
    O_ASSIGN = 23, O_DECL = 24, O_ID = 25, O_WHILE_DUMMY = 26, O_FDECL_DUMMY = 27,
    O_RETURN = 28, O_FDECL = 29, O_FARG_LIST = 30, O_CALL = 31, O_CALL_ARG_LIST = 32,
    O_IF = 0xD0, O_WHILE = 0xD1, O_IF_ELSE = 0xD2,
    O_EQ = 0xE0, O_NEQ = 0xE1, O_LT = 0xE2, O_GT = 0xE3, O_LTE = 0xE4, O_GTE = 0xE5,
    O_FARG = 0xF0, O_FARG_FI = 0xF1, O_FARG_STACK = 0xF2,
    O_CARG = 0xF4, O_CARG_FI = 0xF5, O_CARG_STACK = 0xF6
};
enum { D_INVALID = 0, D_VOID = 1, D_INT = 2, D_FLOAT = 3, D_INT_REF = 4, D_FLOAT_REF = 5 };

/* RISC-V field builders (R-type layout; I/S immediates are passed in the
 * funct7/rs2 or rd positions exactly where the reference's bit strings put them). */
#define OW_R(f7, rs2, rs1, f3, rd, op) \
    ((((uint32_t)(f7)) << 25) | (((uint32_t)(rs2)) << 20) | (((uint32_t)(rs1)) << 15) | \
     (((uint32_t)(f3)) << 12) | (((uint32_t)(rd)) << 7) | ((uint32_t)(op)))
#define OW_I(imm12, rs1, f3, rd, op) \
    ((((uint32_t)(imm12)) << 20) | (((uint32_t)(rs1)) << 15) | (((uint32_t)(f3)) << 12) | \
     (((uint32_t)(rd)) << 7) | ((uint32_t)(op)))
#define O_ERRW 0x73u

typedef struct {
    uint32_t size[256][6];
    uint32_t has_instr[256][4][6];
    uint32_t has_output[256][4][6];
    uint32_t parent_arg[256][4][6];
    uint32_t get_output[256][4][6];
    uint32_t instr[256][4][6];
    uint32_t iconst[256][4];
    uint32_t operand[256][4][6][2];
    uint32_t jt[256][4];
    uint8_t pareas_map[44];
} oracle_tables;

static void ot_all6(uint32_t *dst, uint32_t v) { for (int i = 0; i < 6; ++i) dst[i] = v; }
static void ot_set6(uint32_t *dst, uint32_t a, uint32_t b, uint32_t c, uint32_t d, uint32_t e, uint32_t f) {
    dst[0] = a; dst[1] = b; dst[2] = c; dst[3] = d; dst[4] = e; dst[5] = f;
}
static void ot_ops6(uint32_t dst[6][2], const uint32_t p[6][2]) { memcpy(dst, p, sizeof(uint32_t) * 12); }
static void ot_opsall(uint32_t dst[6][2], uint32_t a, uint32_t b) {
    for (int i = 0; i < 6; ++i) { dst[i][0] = a; dst[i][1] = b; }
}

static void oracle_tables_init(oracle_tables *t) {
    memset(t, 0, sizeof(*t));
    const uint32_t Y = 0xFFFFFFFFu;

    /* ---- node_size_mapping (rv_nodetype.hpp:197-259) ---- */
    static const uint8_t size1[] = { O_ADD, O_SUB, O_MUL, O_DIV, O_MOD, O_BITAND, O_BITOR, O_BITXOR,
        O_LSHIFT, O_RSHIFT, O_URSHIFT, O_BITNOT, O_LNOT, O_NEG, O_CAST, O_DEREF, O_DECL, O_ID,
        O_WHILE, O_LT, O_GT, O_FARG, O_FARG_FI };
    for (unsigned i = 0; i < sizeof(size1); ++i) ot_all6(t->size[size1[i]], 1);
    ot_set6(t->size[O_LIT], 0, 0, 2, 3, 0, 0);
    ot_all6(t->size[O_ASSIGN], 2);
    ot_all6(t->size[O_FDECL_DUMMY], 6);
    ot_all6(t->size[O_RETURN], 2);
    ot_all6(t->size[O_FDECL], 6);
    ot_set6(t->size[O_CALL], 3, 2, 3, 3, 3, 3);
    ot_set6(t->size[O_EQ], 0, 0, 2, 1, 0, 0);
    ot_all6(t->size[O_NEQ], 2);
    ot_set6(t->size[O_LTE], 0, 0, 2, 1, 0, 0);
    ot_set6(t->size[O_GTE], 0, 0, 2, 1, 0, 0);
    ot_all6(t->size[O_FARG_STACK], 2);

    /* ---- has_instr_mapping (:266-326) ---- */
    static const uint8_t hi0[] = { O_ADD, O_SUB, O_MUL, O_DIV, O_MOD, O_BITAND, O_BITOR, O_BITXOR,
        O_LSHIFT, O_RSHIFT, O_URSHIFT, O_LAND, O_LOR, O_BITNOT, O_LNOT, O_NEG, O_CAST, O_DEREF,
        O_DECL, O_ID, O_WHILE_DUMMY, O_IF, O_LT, O_GT, O_FARG, O_FARG_FI, O_CARG, O_CARG_FI };
    for (unsigned i = 0; i < sizeof(hi0); ++i) ot_all6(t->has_instr[hi0[i]][0], Y);
    ot_all6(t->has_instr[O_LIT][0], Y); ot_all6(t->has_instr[O_LIT][1], Y);
    ot_set6(t->has_instr[O_LIT][2], 0, 0, 0, Y, 0, 0);
    ot_all6(t->has_instr[O_ASSIGN][0], Y); ot_all6(t->has_instr[O_ASSIGN][1], Y);
    for (int o = 0; o < 4; ++o) ot_all6(t->has_instr[O_FDECL_DUMMY][o], Y);
    ot_all6(t->has_instr[O_RETURN][0], Y); ot_all6(t->has_instr[O_RETURN][1], Y);
    for (int o = 0; o < 4; ++o) ot_all6(t->has_instr[O_FDECL][o], Y);
    ot_all6(t->has_instr[O_CALL][0], Y);
    ot_set6(t->has_instr[O_CALL][2], 0, 0, Y, Y, 0, 0);
    ot_all6(t->has_instr[O_CALL_ARG_LIST][0], Y); ot_all6(t->has_instr[O_CALL_ARG_LIST][1], Y);
    ot_all6(t->has_instr[O_WHILE][0], Y); ot_all6(t->has_instr[O_WHILE][1], Y);
    ot_all6(t->has_instr[O_IF_ELSE][0], Y); ot_all6(t->has_instr[O_IF_ELSE][1], Y);
    ot_all6(t->has_instr[O_EQ][0], Y); ot_set6(t->has_instr[O_EQ][1], 0, 0, Y, Y, 0, 0);
    ot_all6(t->has_instr[O_NEQ][0], Y); ot_all6(t->has_instr[O_NEQ][1], Y);
    ot_all6(t->has_instr[O_LTE][0], Y); ot_set6(t->has_instr[O_LTE][1], 0, 0, Y, 0, 0, 0);
    ot_all6(t->has_instr[O_GTE][0], Y); ot_set6(t->has_instr[O_GTE][1], 0, 0, Y, 0, 0, 0);
    ot_all6(t->has_instr[O_FARG_STACK][0], Y); ot_all6(t->has_instr[O_FARG_STACK][1], Y);
    ot_all6(t->has_instr[O_CARG_STACK][0], Y); ot_all6(t->has_instr[O_CARG_STACK][1], Y);

    /* ---- has_output (:333-393) ---- */
    ot_all6(t->has_output[O_LIT][0], Y); ot_set6(t->has_output[O_LIT][1], 0, 0, 0, Y, 0, 0);
    ot_all6(t->has_output[O_RETURN][0], Y);
    ot_set6(t->has_output[O_EQ][0], 0, 0, Y, 0, 0, 0);
    ot_all6(t->has_output[O_NEQ][0], Y);
    ot_set6(t->has_output[O_LTE][0], 0, 0, Y, 0, 0, 0);
    ot_set6(t->has_output[O_GTE][0], 0, 0, Y, 0, 0, 0);
    ot_all6(t->has_output[O_FARG_STACK][0], Y);

    /* ---- parent_arg_idx_lookup (:401-461) ---- */
    static const uint8_t pa2[] = { O_INVALID, O_STMT_LIST, O_EXPR, O_ADD, O_SUB, O_MUL, O_DIV, O_MOD,
        O_BITAND, O_BITOR, O_BITXOR, O_LSHIFT, O_RSHIFT, O_URSHIFT, O_LAND, O_LOR, O_BITNOT,
        O_LNOT, O_NEG, O_CAST, O_DEREF, O_DECL, O_ID, O_WHILE_DUMMY, O_FDECL_DUMMY, O_FARG_LIST,
        O_CALL_ARG_LIST, O_IF, O_WHILE, O_IF_ELSE, O_EQ, O_NEQ, O_LT, O_GT, O_LTE, O_GTE };
    for (unsigned i = 0; i < sizeof(pa2); ++i) ot_all6(t->parent_arg[pa2[i]][0], 2);
    ot_set6(t->parent_arg[O_LIT][1], 1, 1, 1, 0, 1, 1);
    ot_set6(t->parent_arg[O_LIT][2], 0, 0, 0, 1, 0, 0);
    ot_all6(t->parent_arg[O_ASSIGN][1], 1);
    ot_set6(t->parent_arg[O_CALL][2], 0, 0, 1, 1, 0, 0);
    ot_set6(t->parent_arg[O_EQ][1], 0, 0, 1, 0, 0, 0);

    /* ---- get_output_table (:468-558) ---- */
    ot_set6(t->get_output[O_RETURN][0], 4, 4, 4, 3, 4, 4);
    ot_set6(t->get_output[O_CARG][0], 0, 0, 1, 2, 0, 0);
    ot_all6(t->get_output[O_CARG_FI][0], 1);

    /* ---- instr_table (:563-858) ---- */
    static const uint8_t listed[] = { O_INVALID, O_STMT_LIST, O_EMPTY, O_EXPR, O_ADD, O_SUB, O_MUL,
        O_DIV, O_MOD, O_BITAND, O_BITOR, O_BITXOR, O_LSHIFT, O_RSHIFT, O_URSHIFT, O_LAND, O_LOR,
        O_BITNOT, O_LNOT, O_NEG, O_LIT, O_CAST, O_DEREF, O_ASSIGN, O_DECL, O_ID, O_WHILE_DUMMY,
        O_FDECL_DUMMY, O_RETURN, O_FDECL, O_FARG_LIST, O_CALL, O_CALL_ARG_LIST, O_IF, O_WHILE,
        O_IF_ELSE, O_EQ, O_NEQ, O_LT, O_GT, O_LTE, O_GTE, O_FARG, O_FARG_FI, O_FARG_STACK,
        O_CARG, O_CARG_FI, O_CARG_STACK };
    for (unsigned i = 0; i < sizeof(listed); ++i)
        for (int o = 0; o < 4; ++o) ot_all6(t->instr[listed[i]][o], O_ERRW);
    const uint32_t E = O_ERRW;
    ot_set6(t->instr[O_ADD][0], E, E, OW_R(0x00, 0, 0, 0, 0, 0x33), OW_R(0x00, 0, 0, 7, 0, 0x53), E, E);
    ot_set6(t->instr[O_SUB][0], E, E, OW_R(0x20, 0, 0, 0, 0, 0x33), OW_R(0x04, 0, 0, 7, 0, 0x53), E, E);
    ot_set6(t->instr[O_MUL][0], E, E, OW_R(0x01, 0, 0, 0, 0, 0x33), OW_R(0x08, 0, 0, 7, 0, 0x53), E, E);
    ot_set6(t->instr[O_DIV][0], E, E, OW_R(0x01, 0, 0, 4, 0, 0x33), OW_R(0x0C, 0, 0, 7, 0, 0x53), E, E);
    ot_set6(t->instr[O_MOD][0], E, E, OW_R(0x01, 0, 0, 6, 0, 0x33), E, E, E);
    ot_set6(t->instr[O_BITAND][0], E, E, OW_R(0x00, 0, 0, 7, 0, 0x33), E, E, E);
    ot_set6(t->instr[O_BITOR][0], E, E, OW_R(0x00, 0, 0, 6, 0, 0x33), E, E, E);
    ot_set6(t->instr[O_BITXOR][0], E, E, OW_R(0x00, 0, 0, 4, 0, 0x33), E, E, E);
    ot_set6(t->instr[O_LSHIFT][0], E, E, OW_R(0x00, 0, 0, 1, 0, 0x33), E, E, E);
    ot_set6(t->instr[O_RSHIFT][0], E, E, OW_R(0x20, 0, 0, 5, 0, 0x33), E, E, E);
    ot_set6(t->instr[O_URSHIFT][0], E, E, OW_R(0x00, 0, 0, 5, 0, 0x33), E, E, E);
    ot_set6(t->instr[O_BITNOT][0], E, E, OW_I(0xFFF, 0, 4, 0, 0x13), E, E, E);
    ot_set6(t->instr[O_LNOT][0], E, E, OW_I(1, 0, 3, 0, 0x13), E, E, E);
    ot_set6(t->instr[O_NEG][0], E, E, OW_R(0x20, 0, 0, 0, 0, 0x33), OW_R(0x10, 0, 0, 1, 0, 0x53), E, E);
    ot_set6(t->instr[O_LIT][0], E, E, 0x37, 0x37, E, E);
    ot_set6(t->instr[O_LIT][1], E, E, 0x13, 0x13, E, E);
    ot_set6(t->instr[O_LIT][2], E, E, E, OW_R(0x78, 0, 0, 0, 0, 0x53), E, E);
    ot_set6(t->instr[O_CAST][0], E, E, OW_R(0x60, 0, 0, 7, 0, 0x53), OW_R(0x68, 0, 0, 7, 0, 0x53), E, E);
    ot_set6(t->instr[O_DEREF][0], E, E, OW_I(0, 0, 2, 0, 0x03), OW_I(0, 0, 2, 0, 0x07), E, E);
    ot_set6(t->instr[O_ASSIGN][0], E, E, OW_R(0, 0, 0, 2, 0, 0x23), OW_R(0, 0, 0, 2, 0, 0x27), E, E);
    ot_set6(t->instr[O_ASSIGN][1], E, E, OW_R(0x00, 0, 0, 0, 0, 0x33), OW_R(0x10, 0, 0, 0, 0, 0x53), E, E);
    ot_all6(t->instr[O_DECL][0], OW_I(0, 8, 0, 0, 0x13));
    ot_all6(t->instr[O_ID][0], OW_I(0, 8, 0, 0, 0x13));
    ot_all6(t->instr[O_FDECL_DUMMY][0], OW_R(0x7F, 1, 2, 2, 0x1C, 0x23));
    ot_all6(t->instr[O_FDECL_DUMMY][1], OW_R(0x7F, 8, 2, 2, 0x18, 0x23));
    ot_all6(t->instr[O_FDECL_DUMMY][2], OW_R(0x20, 8, 2, 0, 2, 0x33));
    ot_all6(t->instr[O_FDECL_DUMMY][3], OW_R(0x00, 2, 8, 0, 8, 0x33));
    /* return o0: the reference lists only four of six data types; the last two are 0. */
    ot_set6(t->instr[O_RETURN][0], E, OW_R(0, 0, 0, 0, 0, 0x33), OW_R(0, 0, 0, 0, 0, 0x33),
            OW_R(0x10, 0, 0, 0, 0, 0x53), 0, 0);
    ot_all6(t->instr[O_RETURN][1], 0x67);
    ot_all6(t->instr[O_FDECL][0], OW_R(0x00, 8, 2, 0, 2, 0x33));
    ot_all6(t->instr[O_FDECL][1], OW_I(0xFFC, 2, 2, 1, 0x03));
    ot_all6(t->instr[O_FDECL][2], OW_I(0xFF8, 8, 2, 8, 0x03));
    ot_all6(t->instr[O_FDECL][3], OW_I(0, 1, 0, 0, 0x67));
    ot_all6(t->instr[O_CALL][0], OW_I(0, 0, 0, 1, 0x67));
    ot_set6(t->instr[O_CALL][2], E, E, OW_R(0x00, 0, 10, 0, 0, 0x33), OW_R(0x10, 10, 10, 0, 0, 0x53), E, E);
    ot_all6(t->instr[O_CALL_ARG_LIST][0], OW_I(0, 2, 7, 2, 0x13));
    ot_all6(t->instr[O_CALL_ARG_LIST][1], OW_I(0, 2, 6, 2, 0x13));
    ot_all6(t->instr[O_IF][0], 0x63);
    ot_all6(t->instr[O_WHILE][0], 0x63); ot_all6(t->instr[O_WHILE][1], 0x67);
    ot_all6(t->instr[O_IF_ELSE][0], 0x63); ot_all6(t->instr[O_IF_ELSE][1], 0x67);
    ot_set6(t->instr[O_EQ][0], E, E, OW_R(0x20, 0, 0, 0, 0, 0x33), OW_R(0x50, 0, 0, 2, 0, 0x53), E, E);
    ot_set6(t->instr[O_EQ][1], E, E, OW_I(1, 0, 3, 0, 0x13), E, E, E);
    ot_set6(t->instr[O_NEQ][0], E, E, OW_R(0x20, 0, 0, 0, 0, 0x33), OW_R(0x50, 0, 0, 2, 0, 0x53), E, E);
    ot_set6(t->instr[O_NEQ][1], E, E, OW_R(0x00, 0, 0, 3, 0, 0x33), OW_I(1, 0, 3, 0, 0x13), E, E);
    ot_set6(t->instr[O_LT][0], E, E, OW_R(0x00, 0, 0, 2, 0, 0x33), OW_R(0x50, 0, 0, 1, 0, 0x53), E, E);
    ot_set6(t->instr[O_GT][0], E, E, OW_R(0x00, 0, 0, 2, 0, 0x33), OW_R(0x50, 0, 0, 1, 0, 0x53), E, E);
    ot_set6(t->instr[O_LTE][0], E, E, OW_R(0x00, 0, 0, 2, 0, 0x33), OW_R(0x50, 0, 0, 0, 0, 0x53), E, E);
    ot_set6(t->instr[O_LTE][1], E, E, OW_I(1, 0, 3, 0, 0x13), E, E, E);
    ot_set6(t->instr[O_GTE][0], E, E, OW_R(0x00, 0, 0, 2, 0, 0x33), OW_R(0x50, 0, 0, 0, 0, 0x53), E, E);
    ot_set6(t->instr[O_GTE][1], E, E, OW_I(1, 0, 3, 0, 0x13), E, E, E);
    ot_set6(t->instr[O_FARG][0], E, E, E, E, OW_R(0, 0, 0, 2, 0, 0x23), OW_R(0, 0, 0, 2, 0, 0x27));
    ot_all6(t->instr[O_FARG_FI][0], OW_R(0x70, 0, 0, 0, 0, 0x53));
    ot_set6(t->instr[O_FARG_STACK][0], E, E, OW_R(0, 0, 0, 2, 2, 0x23), OW_R(0, 0, 0, 2, 2, 0x27), E, E);
    ot_set6(t->instr[O_CARG][0], E, E, OW_R(0x00, 0, 0, 0, 0, 0x33), OW_R(0x10, 0, 0, 0, 0, 0x53), E, E);
    ot_all6(t->instr[O_CARG_FI][0], OW_R(0x70, 0, 0, 0, 0, 0x53));
    ot_set6(t->instr[O_CARG_STACK][0], E, E, OW_R(0, 0, 0, 2, 2, 0x23), OW_R(0, 0, 0, 2, 2, 0x27), E, E);

    /* ---- instr_constant_table (:863-925) ---- */
    t->iconst[O_LIT][0] = 1; t->iconst[O_LIT][1] = 2;
    t->iconst[O_DECL][0] = 3; t->iconst[O_ID][0] = 3;
    t->iconst[O_CALL_ARG_LIST][0] = 5; t->iconst[O_CALL_ARG_LIST][1] = 4;
    t->iconst[O_FARG_STACK][0] = 4; t->iconst[O_CARG_STACK][0] = 4;

    /* ---- operand_table (:930-1008) ---- */
    static const uint8_t bin11[] = { O_ADD, O_SUB, O_MUL, O_DIV, O_MOD, O_BITAND, O_BITOR, O_BITXOR,
        O_LSHIFT, O_RSHIFT, O_URSHIFT, O_ASSIGN, O_EQ, O_NEQ, O_LT };
    for (unsigned i = 0; i < sizeof(bin11); ++i) ot_opsall(t->operand[bin11[i]][0], 1, 1);
    ot_opsall(t->operand[O_LIT][1], 7, 0);
    { const uint32_t p[6][2] = {{0,0},{0,0},{0,0},{7,0},{0,0},{0,0}}; ot_ops6(t->operand[O_LIT][2], p); }
    ot_opsall(t->operand[O_CAST][0], 1, 0);
    ot_opsall(t->operand[O_DEREF][0], 1, 0);
    { const uint32_t p[6][2] = {{4,0},{4,0},{4,0},{4,4},{4,0},{4,0}}; ot_ops6(t->operand[O_ASSIGN][1], p); }
    { const uint32_t p[6][2] = {{0,0},{0,0},{1,0},{1,0},{0,0},{0,0}}; ot_ops6(t->operand[O_RETURN][0], p); }
    ot_opsall(t->operand[O_IF][0], 1, 0);
    ot_opsall(t->operand[O_WHILE][0], 4, 0);
    ot_opsall(t->operand[O_IF_ELSE][0], 1, 0);
    { const uint32_t p[6][2] = {{0,0},{0,0},{7,0},{0,0},{0,0},{0,0}}; ot_ops6(t->operand[O_EQ][1], p); }
    { const uint32_t p[6][2] = {{0,0},{0,0},{0,7},{7,0},{0,0},{0,0}}; ot_ops6(t->operand[O_NEQ][1], p); }
    ot_opsall(t->operand[O_GT][0], 5, 5);
    { const uint32_t p[6][2] = {{0,0},{0,0},{5,5},{1,1},{0,0},{0,0}}; ot_ops6(t->operand[O_LTE][0], p); }
    { const uint32_t p[6][2] = {{0,0},{0,0},{7,0},{0,0},{0,0},{0,0}}; ot_ops6(t->operand[O_LTE][1], p); }
    { const uint32_t p[6][2] = {{0,0},{0,0},{1,1},{5,5},{0,0},{0,0}}; ot_ops6(t->operand[O_GTE][0], p); }
    { const uint32_t p[6][2] = {{0,0},{0,0},{7,0},{0,0},{0,0},{0,0}}; ot_ops6(t->operand[O_GTE][1], p); }
    { const uint32_t p[6][2] = {{1,0},{1,0},{1,0},{1,0},{1,2},{1,3}}; ot_ops6(t->operand[O_FARG][0], p); }
    ot_opsall(t->operand[O_FARG_FI][0], 1, 2);
    ot_opsall(t->operand[O_FARG_STACK][1], 1, 7);
    { const uint32_t p[6][2] = {{1,0},{1,0},{1,0},{1,6},{1,0},{1,0}}; ot_ops6(t->operand[O_CARG][0], p); }
    ot_opsall(t->operand[O_CARG_FI][0], 1, 0);
    ot_opsall(t->operand[O_CARG_STACK][0], 6, 0);

    /* ---- instr_jt_table (:1013-1075) ---- */
    t->jt[O_RETURN][1] = 6;
    t->jt[O_CALL][0] = 7;
    t->jt[O_IF][0] = 1;
    t->jt[O_WHILE][0] = 4; t->jt[O_WHILE][1] = 5;
    t->jt[O_IF_ELSE][0] = 2; t->jt[O_IF_ELSE][1] = 3;

    /* ---- pareas_to_rv_nodetype (:1087-1134), indexed by pareas NodeType ---- */
    static const uint8_t pm[44] = { O_INVALID, O_STMT_LIST, O_EMPTY, O_FDECL, O_FARG, O_FARG_LIST,
        O_EXPR, O_IF, O_IF_ELSE, O_WHILE, O_CALL, O_CARG, O_CALL_ARG_LIST, O_ADD, O_SUB, O_MUL,
        O_DIV, O_MOD, O_BITAND, O_BITOR, O_BITXOR, O_LSHIFT, O_RSHIFT, O_URSHIFT, O_LAND, O_LOR,
        O_EQ, O_NEQ, O_LT, O_GT, O_LTE, O_GTE, O_BITNOT, O_LNOT, O_NEG, O_LIT, O_CAST, O_DEREF,
        O_ASSIGN, O_DECL, O_ID, O_WHILE_DUMMY, O_FDECL_DUMMY, O_RETURN };
    memcpy(t->pareas_map, pm, 44);
}

#endif

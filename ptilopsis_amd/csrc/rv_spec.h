// Instruction-selection spec of the code generator, packed for the device.
//
// The facts are the reference's nine lookup tables (include/codegen/rv_nodetype.hpp:197-1084),
// restated as one row per emitted instruction slot: which node type and slot offset,
// for which data types, the instruction template, how its destination and
// operands are found, and how the value is handed to the parent.  Only the
// parts the generator can observe are kept: a slot's template, output rule,
// operand rules and has_output bit are read only when the slot emits an
// instruction, so rows exist only for emitting slots; the parent hand-off rule
// of offset 0 is also read for nodes that emit nothing there (pai0 below).
// rv_spec_expand() turns the rows into a dense [48 types][4 offsets][6 data types]
// table that is copied to constant memory.
#ifndef PTL_RV_SPEC_H
#define PTL_RV_SPEC_H
#include <stdint.h>
#include <string.h>

// rv_node_type codes (rv_nodetype.hpp:13-86).  Control nodes carry 0b1101 in the
// top nibble and the condition's child index in bit 0; comparisons 0b1110;
// arguments 0b1111 with bit0 = float passed in an int register, bit1 = on the
// stack, bit2 = call-site argument.
enum : uint8_t {
    RV_INVALID = 0, RV_STMT_LIST, RV_EMPTY, RV_EXPR, RV_ADD, RV_SUB, RV_MUL, RV_DIV, RV_MOD,
    RV_AND, RV_OR, RV_XOR, RV_SHL, RV_SAR, RV_SHR, RV_LAND, RV_LOR, RV_NOT, RV_LNOT, RV_NEG,
    RV_LIT, RV_CAST, RV_DEREF, RV_ASSIGN, RV_DECL, RV_ID, RV_WHILE_DUMMY, RV_FDECL_DUMMY,
    RV_RETURN, RV_FDECL, RV_FARG_LIST, RV_CALL, RV_CALL_ARGS,
    RV_IF = 0xD0, RV_WHILE = 0xD1, RV_IF_ELSE = 0xD2,
    RV_EQ = 0xE0, RV_NE, RV_LT, RV_GT, RV_LE, RV_GE,
    RV_ARG = 0xF0, RV_ARG_FI = 0xF1, RV_ARG_STK = 0xF2,
    RV_CARG = 0xF4, RV_CARG_FI = 0xF5, RV_CARG_STK = 0xF6,
};
enum : uint8_t { DT_INVALID = 0, DT_VOID, DT_INT, DT_FLOAT, DT_INT_REF, DT_FLOAT_REF };

// Pareas NodeType (pareas-ast include/codegen/astnode.hpp:11-56) -> rv_node_type.
static const uint8_t RV_FROM_PAREAS[44] = {
    RV_INVALID, RV_STMT_LIST, RV_EMPTY, RV_FDECL, RV_ARG, RV_FARG_LIST, RV_EXPR, RV_IF, RV_IF_ELSE,
    RV_WHILE, RV_CALL, RV_CARG, RV_CALL_ARGS, RV_ADD, RV_SUB, RV_MUL, RV_DIV, RV_MOD, RV_AND, RV_OR,
    RV_XOR, RV_SHL, RV_SAR, RV_SHR, RV_LAND, RV_LOR, RV_EQ, RV_NE, RV_LT, RV_GT, RV_LE, RV_GE,
    RV_NOT, RV_LNOT, RV_NEG, RV_LIT, RV_CAST, RV_DEREF, RV_ASSIGN, RV_DECL, RV_ID, RV_WHILE_DUMMY,
    RV_FDECL_DUMMY, RV_RETURN,
};

// Data-type masks for rows.
#define M_ALL 0x3Fu
#define M_I (1u << DT_INT)
#define M_F (1u << DT_FLOAT)
#define M_IF (M_I | M_F)

// Instruction encodings by field (RISC-V base + M + F).
#define ENC_R(f7, rs2, rs1, f3, rd, op) \
    ((uint32_t)(f7) << 25 | (uint32_t)(rs2) << 20 | (uint32_t)(rs1) << 15 | (uint32_t)(f3) << 12 | (uint32_t)(rd) << 7 | (uint32_t)(op))
#define ENC_I(imm, rs1, f3, rd, op) ENC_R(((imm) >> 5) & 0x7F, (imm) & 0x1F, rs1, f3, rd, op)
#define ENC_S(imm, rs2, rs1, f3, op) ENC_R(((imm) >> 5) & 0x7F, rs2, rs1, f3, (imm) & 0x1F, op)
enum : uint32_t { OPC_OP = 0x33, OPC_IMM = 0x13, OPC_LUI = 0x37, OPC_FP = 0x53, OPC_LOAD = 0x03,
                  OPC_FLOAD = 0x07, OPC_STORE = 0x23, OPC_FSTORE = 0x27, OPC_JALR = 0x67, OPC_BRANCH = 0x63,
                  OPC_SYSTEM = 0x73 };
#define W_ERR OPC_SYSTEM  // placeholder word the reference emits for slots it never fills

// Destination rules (rv_generator_avx.cpp:675-700): 0 = the slot's own virtual
// register if it has an output, 1 = nd+10 (x10+i), 2 = nd+42 (f10+i), 3 = f0 (32), 4 = x10.
// Operand rules (:817-857): 0 none, 1 child i's value, 2 nd+10, 3 nd+42,
// 4 child 1, 5 child 1-i (swapped), 6 child 0, 7 the previous word's register.
// Parent hand-off (:640-662): 0 none, 1 always, 2 when the data type is above VOID.
// Constants (:757-775): 1 lui upper 20 bits of nd, 2 addi low 12 bits, 3 -(4nd+8)<<20,
// 4 4nd<<20, 5 -4nd<<20.  Jump targets (:778-814): 1..5 from child slots,
// 6 function end - 6, 7 function start.
struct rv_row {
    uint8_t type, off, dmask;
    uint32_t word;
    uint8_t out, op1, op2, has_out, pai;
};

static const rv_row RV_ROWS[] = {
    // binary arithmetic / bitwise: one instruction on child values 0 and 1
    {RV_ADD, 0, M_I, ENC_R(0x00, 0, 0, 0, 0, OPC_OP), 0, 1, 1, 0, 2},
    {RV_ADD, 0, M_F, ENC_R(0x00, 0, 0, 7, 0, OPC_FP), 0, 1, 1, 0, 2},
    {RV_ADD, 0, M_ALL & ~M_IF, W_ERR, 0, 1, 1, 0, 2},
    {RV_SUB, 0, M_I, ENC_R(0x20, 0, 0, 0, 0, OPC_OP), 0, 1, 1, 0, 2},
    {RV_SUB, 0, M_F, ENC_R(0x04, 0, 0, 7, 0, OPC_FP), 0, 1, 1, 0, 2},
    {RV_SUB, 0, M_ALL & ~M_IF, W_ERR, 0, 1, 1, 0, 2},
    {RV_MUL, 0, M_I, ENC_R(0x01, 0, 0, 0, 0, OPC_OP), 0, 1, 1, 0, 2},
    {RV_MUL, 0, M_F, ENC_R(0x08, 0, 0, 7, 0, OPC_FP), 0, 1, 1, 0, 2},
    {RV_MUL, 0, M_ALL & ~M_IF, W_ERR, 0, 1, 1, 0, 2},
    {RV_DIV, 0, M_I, ENC_R(0x01, 0, 0, 4, 0, OPC_OP), 0, 1, 1, 0, 2},
    {RV_DIV, 0, M_F, ENC_R(0x0C, 0, 0, 7, 0, OPC_FP), 0, 1, 1, 0, 2},
    {RV_DIV, 0, M_ALL & ~M_IF, W_ERR, 0, 1, 1, 0, 2},
    {RV_MOD, 0, M_I, ENC_R(0x01, 0, 0, 6, 0, OPC_OP), 0, 1, 1, 0, 2},
    {RV_MOD, 0, M_ALL & ~M_I, W_ERR, 0, 1, 1, 0, 2},
    {RV_AND, 0, M_I, ENC_R(0x00, 0, 0, 7, 0, OPC_OP), 0, 1, 1, 0, 2},
    {RV_AND, 0, M_ALL & ~M_I, W_ERR, 0, 1, 1, 0, 2},
    {RV_OR, 0, M_I, ENC_R(0x00, 0, 0, 6, 0, OPC_OP), 0, 1, 1, 0, 2},
    {RV_OR, 0, M_ALL & ~M_I, W_ERR, 0, 1, 1, 0, 2},
    {RV_XOR, 0, M_I, ENC_R(0x00, 0, 0, 4, 0, OPC_OP), 0, 1, 1, 0, 2},
    {RV_XOR, 0, M_ALL & ~M_I, W_ERR, 0, 1, 1, 0, 2},
    {RV_SHL, 0, M_I, ENC_R(0x00, 0, 0, 1, 0, OPC_OP), 0, 1, 1, 0, 2},
    {RV_SHL, 0, M_ALL & ~M_I, W_ERR, 0, 1, 1, 0, 2},
    {RV_SAR, 0, M_I, ENC_R(0x20, 0, 0, 5, 0, OPC_OP), 0, 1, 1, 0, 2},
    {RV_SAR, 0, M_ALL & ~M_I, W_ERR, 0, 1, 1, 0, 2},
    {RV_SHR, 0, M_I, ENC_R(0x00, 0, 0, 5, 0, OPC_OP), 0, 1, 1, 0, 2},
    {RV_SHR, 0, M_ALL & ~M_I, W_ERR, 0, 1, 1, 0, 2},
    // && and || emit a placeholder word with no operands
    {RV_LAND, 0, M_ALL, W_ERR, 0, 0, 0, 0, 2},
    {RV_LOR, 0, M_ALL, W_ERR, 0, 0, 0, 0, 2},
    // unary operators: no operand rules in the reference
    {RV_NOT, 0, M_I, ENC_I(0xFFF, 0, 4, 0, OPC_IMM), 0, 0, 0, 0, 2},
    {RV_NOT, 0, M_ALL & ~M_I, W_ERR, 0, 0, 0, 0, 2},
    {RV_LNOT, 0, M_I, ENC_I(1, 0, 3, 0, OPC_IMM), 0, 0, 0, 0, 2},
    {RV_LNOT, 0, M_ALL & ~M_I, W_ERR, 0, 0, 0, 0, 2},
    {RV_NEG, 0, M_I, ENC_R(0x20, 0, 0, 0, 0, OPC_OP), 0, 0, 0, 0, 2},
    {RV_NEG, 0, M_F, ENC_R(0x10, 0, 0, 1, 0, OPC_FP), 0, 0, 0, 0, 2},
    {RV_NEG, 0, M_ALL & ~M_IF, W_ERR, 0, 0, 0, 0, 2},
    // literal: lui / addi (+ fmv.w.x for floats)
    {RV_LIT, 0, M_IF, OPC_LUI, 0, 0, 0, 1, 0},
    {RV_LIT, 0, M_ALL & ~M_IF, W_ERR, 0, 0, 0, 1, 0},
    {RV_LIT, 1, M_I, OPC_IMM, 0, 7, 0, 0, 1},
    {RV_LIT, 1, M_F, OPC_IMM, 0, 7, 0, 1, 0},
    {RV_LIT, 1, M_ALL & ~M_IF, W_ERR, 0, 7, 0, 0, 1},
    {RV_LIT, 2, M_F, ENC_R(0x78, 0, 0, 0, 0, OPC_FP), 0, 7, 0, 0, 1},
    {RV_CAST, 0, M_I, ENC_R(0x60, 0, 0, 7, 0, OPC_FP), 0, 1, 0, 0, 2},
    {RV_CAST, 0, M_F, ENC_R(0x68, 0, 0, 7, 0, OPC_FP), 0, 1, 0, 0, 2},
    {RV_CAST, 0, M_ALL & ~M_IF, W_ERR, 0, 1, 0, 0, 2},
    {RV_DEREF, 0, M_I, ENC_I(0, 0, 2, 0, OPC_LOAD), 0, 1, 0, 0, 2},
    {RV_DEREF, 0, M_F, ENC_I(0, 0, 2, 0, OPC_FLOAD), 0, 1, 0, 0, 2},
    {RV_DEREF, 0, M_ALL & ~M_IF, W_ERR, 0, 1, 0, 0, 2},
    // assignment: store then register copy (the copy is the value handed up)
    {RV_ASSIGN, 0, M_I, ENC_S(0, 0, 0, 2, OPC_STORE), 0, 1, 1, 0, 0},
    {RV_ASSIGN, 0, M_F, ENC_S(0, 0, 0, 2, OPC_FSTORE), 0, 1, 1, 0, 0},
    {RV_ASSIGN, 0, M_ALL & ~M_IF, W_ERR, 0, 1, 1, 0, 0},
    {RV_ASSIGN, 1, M_I, ENC_R(0x00, 0, 0, 0, 0, OPC_OP), 0, 4, 0, 0, 1},
    {RV_ASSIGN, 1, M_F, ENC_R(0x10, 0, 0, 0, 0, OPC_FP), 0, 4, 4, 0, 1},
    {RV_ASSIGN, 1, M_ALL & ~M_IF, W_ERR, 0, 4, 0, 0, 1},
    // variable address: addi rd, s0, -(4*(id+2))
    {RV_DECL, 0, M_ALL, ENC_I(0, 8, 0, 0, OPC_IMM), 0, 0, 0, 0, 2},
    {RV_ID, 0, M_ALL, ENC_I(0, 8, 0, 0, OPC_IMM), 0, 0, 0, 0, 2},
    {RV_WHILE_DUMMY, 0, M_ALL, W_ERR, 0, 0, 0, 0, 2},
    // prologue: sw ra,-4(sp); sw s0,-8(sp); sub sp,sp,s0; add s0,s0,sp
    {RV_FDECL_DUMMY, 0, M_ALL, ENC_S(0xFFC, 1, 2, 2, OPC_STORE), 0, 0, 0, 0, 2},
    {RV_FDECL_DUMMY, 1, M_ALL, ENC_S(0xFF8, 8, 2, 2, OPC_STORE), 0, 0, 0, 0, 0},
    {RV_FDECL_DUMMY, 2, M_ALL, ENC_R(0x20, 8, 2, 0, 2, OPC_OP), 0, 0, 0, 0, 0},
    {RV_FDECL_DUMMY, 3, M_ALL, ENC_R(0x00, 2, 8, 0, 8, OPC_OP), 0, 0, 0, 0, 0},
    // return: move the value to a0/fa0, then jump to the epilogue
    {RV_RETURN, 0, 1u << DT_INVALID, W_ERR, 4, 0, 0, 1, 0},
    {RV_RETURN, 0, 1u << DT_VOID, ENC_R(0, 0, 0, 0, 0, OPC_OP), 4, 0, 0, 1, 0},
    {RV_RETURN, 0, M_I, ENC_R(0, 0, 0, 0, 0, OPC_OP), 4, 1, 0, 1, 0},
    {RV_RETURN, 0, M_F, ENC_R(0x10, 0, 0, 0, 0, OPC_FP), 3, 1, 0, 1, 0},
    {RV_RETURN, 0, (1u << DT_INT_REF) | (1u << DT_FLOAT_REF), 0, 4, 0, 0, 1, 0},
    {RV_RETURN, 1, M_ALL, OPC_JALR, 0, 0, 0, 0, 0},
    // epilogue: add sp,sp,s0; lw ra,-4(sp); lw s0,-8(s0); ret
    {RV_FDECL, 0, M_ALL, ENC_R(0x00, 8, 2, 0, 2, OPC_OP), 0, 0, 0, 0, 0},
    {RV_FDECL, 1, M_ALL, ENC_I(0xFFC, 2, 2, 1, OPC_LOAD), 0, 0, 0, 0, 0},
    {RV_FDECL, 2, M_ALL, ENC_I(0xFF8, 8, 2, 8, OPC_LOAD), 0, 0, 0, 0, 0},
    {RV_FDECL, 3, M_ALL, ENC_I(0, 1, 0, 0, OPC_JALR), 0, 0, 0, 0, 0},
    // call: jalr ra,0(zero); the result copy from a0/fa0 in slot 2
    {RV_CALL, 0, M_ALL, ENC_I(0, 0, 0, 1, OPC_JALR), 0, 0, 0, 0, 0},
    {RV_CALL, 2, M_I, ENC_R(0x00, 0, 10, 0, 0, OPC_OP), 0, 0, 0, 0, 1},
    {RV_CALL, 2, M_F, ENC_R(0x10, 10, 10, 0, 0, OPC_FP), 0, 0, 0, 0, 1},
    // argument list: andi sp,sp,imm / ori sp,sp,imm around the call
    {RV_CALL_ARGS, 0, M_ALL, ENC_I(0, 2, 7, 2, OPC_IMM), 0, 0, 0, 0, 2},
    {RV_CALL_ARGS, 1, M_ALL, ENC_I(0, 2, 6, 2, OPC_IMM), 0, 0, 0, 0, 0},
    // control flow
    {RV_IF, 0, M_ALL, OPC_BRANCH, 0, 1, 0, 0, 2},
    {RV_WHILE, 0, M_ALL, OPC_BRANCH, 0, 4, 0, 0, 2},
    {RV_WHILE, 1, M_ALL, OPC_JALR, 0, 0, 0, 0, 0},
    {RV_IF_ELSE, 0, M_ALL, OPC_BRANCH, 0, 1, 0, 0, 2},
    {RV_IF_ELSE, 1, M_ALL, OPC_JALR, 0, 0, 0, 0, 0},
    // comparisons
    {RV_EQ, 0, M_I, ENC_R(0x20, 0, 0, 0, 0, OPC_OP), 0, 1, 1, 1, 2},
    {RV_EQ, 0, M_F, ENC_R(0x50, 0, 0, 2, 0, OPC_FP), 0, 1, 1, 0, 2},
    {RV_EQ, 0, M_ALL & ~M_IF, W_ERR, 0, 1, 1, 0, 2},
    {RV_EQ, 1, M_I, ENC_I(1, 0, 3, 0, OPC_IMM), 0, 7, 0, 0, 1},
    {RV_EQ, 1, M_F, W_ERR, 0, 0, 0, 0, 0},
    {RV_NE, 0, M_I, ENC_R(0x20, 0, 0, 0, 0, OPC_OP), 0, 1, 1, 1, 2},
    {RV_NE, 0, M_F, ENC_R(0x50, 0, 0, 2, 0, OPC_FP), 0, 1, 1, 1, 2},
    {RV_NE, 0, M_ALL & ~M_IF, W_ERR, 0, 1, 1, 1, 2},
    {RV_NE, 1, M_I, ENC_R(0x00, 0, 0, 3, 0, OPC_OP), 0, 0, 7, 0, 0},
    {RV_NE, 1, M_F, ENC_I(1, 0, 3, 0, OPC_IMM), 0, 7, 0, 0, 0},
    {RV_NE, 1, M_ALL & ~M_IF, W_ERR, 0, 0, 0, 0, 0},
    {RV_LT, 0, M_I, ENC_R(0x00, 0, 0, 2, 0, OPC_OP), 0, 1, 1, 0, 2},
    {RV_LT, 0, M_F, ENC_R(0x50, 0, 0, 1, 0, OPC_FP), 0, 1, 1, 0, 2},
    {RV_LT, 0, M_ALL & ~M_IF, W_ERR, 0, 1, 1, 0, 2},
    {RV_GT, 0, M_I, ENC_R(0x00, 0, 0, 2, 0, OPC_OP), 0, 5, 5, 0, 2},
    {RV_GT, 0, M_F, ENC_R(0x50, 0, 0, 1, 0, OPC_FP), 0, 5, 5, 0, 2},
    {RV_GT, 0, M_ALL & ~M_IF, W_ERR, 0, 5, 5, 0, 2},
    {RV_LE, 0, M_I, ENC_R(0x00, 0, 0, 2, 0, OPC_OP), 0, 5, 5, 1, 2},
    {RV_LE, 0, M_F, ENC_R(0x50, 0, 0, 0, 0, OPC_FP), 0, 1, 1, 0, 2},
    {RV_LE, 0, M_ALL & ~M_IF, W_ERR, 0, 0, 0, 0, 2},
    {RV_LE, 1, M_I, ENC_I(1, 0, 3, 0, OPC_IMM), 0, 7, 0, 0, 0},
    {RV_GE, 0, M_I, ENC_R(0x00, 0, 0, 2, 0, OPC_OP), 0, 1, 1, 1, 2},
    {RV_GE, 0, M_F, ENC_R(0x50, 0, 0, 0, 0, OPC_FP), 0, 5, 5, 0, 2},
    {RV_GE, 0, M_ALL & ~M_IF, W_ERR, 0, 0, 0, 0, 2},
    {RV_GE, 1, M_I, ENC_I(1, 0, 3, 0, OPC_IMM), 0, 7, 0, 0, 0},
    // declared parameters: store the incoming register into the variable's slot
    {RV_ARG, 0, M_ALL & ~((1u << DT_INT_REF) | (1u << DT_FLOAT_REF)), W_ERR, 0, 1, 0, 0, 0},
    {RV_ARG, 0, 1u << DT_INT_REF, ENC_S(0, 0, 0, 2, OPC_STORE), 0, 1, 2, 0, 0},
    {RV_ARG, 0, 1u << DT_FLOAT_REF, ENC_S(0, 0, 0, 2, OPC_FSTORE), 0, 1, 3, 0, 0},
    {RV_ARG_FI, 0, M_ALL, ENC_R(0x70, 0, 0, 0, 0, OPC_FP), 0, 1, 2, 0, 0},
    {RV_ARG_STK, 0, M_I, ENC_S(2, 0, 0, 2, OPC_STORE), 0, 0, 0, 1, 0},
    {RV_ARG_STK, 0, M_F, ENC_S(2, 0, 0, 2, OPC_FSTORE), 0, 0, 0, 1, 0},
    {RV_ARG_STK, 0, M_ALL & ~M_IF, W_ERR, 0, 0, 0, 1, 0},
    {RV_ARG_STK, 1, M_ALL, W_ERR, 0, 1, 7, 0, 0},
    // call-site arguments: move into x10+i / f10+i, or fmv.x.w, or store to the stack
    {RV_CARG, 0, (1u << DT_INVALID) | (1u << DT_VOID), W_ERR, 0, 1, 0, 0, 0},
    {RV_CARG, 0, M_I, ENC_R(0x00, 0, 0, 0, 0, OPC_OP), 1, 1, 0, 0, 0},
    {RV_CARG, 0, M_F, ENC_R(0x10, 0, 0, 0, 0, OPC_FP), 2, 1, 6, 0, 0},
    {RV_CARG, 0, (1u << DT_INT_REF) | (1u << DT_FLOAT_REF), W_ERR, 0, 1, 0, 0, 0},
    {RV_CARG_FI, 0, M_ALL, ENC_R(0x70, 0, 0, 0, 0, OPC_FP), 1, 1, 0, 0, 0},
    {RV_CARG_STK, 0, M_I, ENC_S(2, 0, 0, 2, OPC_STORE), 0, 6, 0, 0, 0},
    {RV_CARG_STK, 0, M_F, ENC_S(2, 0, 0, 2, OPC_FSTORE), 0, 6, 0, 0, 0},
    {RV_CARG_STK, 0, M_ALL & ~M_IF, W_ERR, 0, 6, 0, 0, 0},
    {RV_CARG_STK, 1, M_ALL, W_ERR, 0, 0, 0, 0, 0},
};

// Parent hand-off rule of offset 0 for types whose offset 0 emits nothing.
static const uint8_t RV_PAI0[][2] = {
    {RV_INVALID, 2}, {RV_STMT_LIST, 2}, {RV_EXPR, 2}, {RV_FARG_LIST, 2},
};

// Words a node occupies, per data type (no mask = 0 words).
struct rv_size_row { uint8_t type, dmask, words; };
static const rv_size_row RV_SIZES[] = {
    {RV_ADD, M_ALL, 1}, {RV_SUB, M_ALL, 1}, {RV_MUL, M_ALL, 1}, {RV_DIV, M_ALL, 1}, {RV_MOD, M_ALL, 1},
    {RV_AND, M_ALL, 1}, {RV_OR, M_ALL, 1}, {RV_XOR, M_ALL, 1}, {RV_SHL, M_ALL, 1}, {RV_SAR, M_ALL, 1},
    {RV_SHR, M_ALL, 1}, {RV_NOT, M_ALL, 1}, {RV_LNOT, M_ALL, 1}, {RV_NEG, M_ALL, 1},
    {RV_LIT, M_I, 2}, {RV_LIT, M_F, 3}, {RV_CAST, M_ALL, 1}, {RV_DEREF, M_ALL, 1},
    {RV_ASSIGN, M_ALL, 2}, {RV_DECL, M_ALL, 1}, {RV_ID, M_ALL, 1}, {RV_FDECL_DUMMY, M_ALL, 6},
    {RV_RETURN, M_ALL, 2}, {RV_FDECL, M_ALL, 6}, {RV_CALL, M_ALL & ~(1u << DT_VOID), 3},
    {RV_CALL, 1u << DT_VOID, 2}, {RV_WHILE, M_ALL, 1}, {RV_EQ, M_I, 2}, {RV_EQ, M_F, 1},
    {RV_NE, M_ALL, 2}, {RV_LT, M_ALL, 1}, {RV_GT, M_ALL, 1}, {RV_LE, M_I, 2}, {RV_LE, M_F, 1},
    {RV_GE, M_I, 2}, {RV_GE, M_F, 1}, {RV_ARG, M_ALL, 1}, {RV_ARG_FI, M_ALL, 1}, {RV_ARG_STK, M_ALL, 2},
};

// Per (type, offset) constant and jump-target rules.
struct rv_slot_row { uint8_t type, off, iconst, jt; };
static const rv_slot_row RV_SLOTS[] = {
    {RV_LIT, 0, 1, 0}, {RV_LIT, 1, 2, 0}, {RV_DECL, 0, 3, 0}, {RV_ID, 0, 3, 0},
    {RV_CALL_ARGS, 0, 5, 0}, {RV_CALL_ARGS, 1, 4, 0}, {RV_ARG_STK, 0, 4, 0}, {RV_CARG_STK, 0, 4, 0},
    {RV_RETURN, 1, 0, 6}, {RV_CALL, 0, 0, 7}, {RV_IF, 0, 0, 1}, {RV_WHILE, 0, 0, 4}, {RV_WHILE, 1, 0, 5},
    {RV_IF_ELSE, 0, 0, 2}, {RV_IF_ELSE, 1, 0, 3},
};

// ---- dense form -------------------------------------------------------------
// Dense type ids: the 48 node types above, in code order.
#define RV_NTYPES 48
static const uint8_t RV_TYPE_CODES[RV_NTYPES] = {
    0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 21, 22, 23, 24, 25,
    26, 27, 28, 29, 30, 31, 32, 0xD0, 0xD1, 0xD2, 0xE0, 0xE1, 0xE2, 0xE3, 0xE4, 0xE5, 0xF0, 0xF1,
    0xF2, 0xF4, 0xF5, 0xF6,
};

// One entry per (dense type, offset, data type).
//   bits 0      emits an instruction here
//   bit  1      has_output
//   bits 2-3    parent hand-off rule
//   bits 4-6    destination rule
//   bits 7-9    operand rule 1
//   bits 10-12  operand rule 2
//   bits 13-15  constant rule (per type/offset)
//   bits 16-18  jump-target rule (per type/offset)
//   bits 19-21  words the node occupies (per type/data type; same in all offsets)
struct rv_dense {
    uint32_t word[RV_NTYPES][4][6];
    uint32_t bits[RV_NTYPES][4][6];
    uint8_t dense_of[256];  // code -> dense id (0xFF: not a node type)
};

static inline int rv_dense_id(uint8_t code) {
    for (int i = 0; i < RV_NTYPES; ++i) if (RV_TYPE_CODES[i] == code) return i;
    return -1;
}

static inline void rv_spec_expand(rv_dense *d) {
    memset(d, 0, sizeof(*d));
    memset(d->dense_of, 0xFF, sizeof(d->dense_of));
    for (int i = 0; i < RV_NTYPES; ++i) d->dense_of[RV_TYPE_CODES[i]] = (uint8_t)i;
    for (const rv_row &r : RV_ROWS) {
        int t = rv_dense_id(r.type);
        for (int dt = 0; dt < 6; ++dt) {
            if (!(r.dmask >> dt & 1)) continue;
            bool emits = r.word != 0 || true;
            (void)emits;
            d->word[t][r.off][dt] = r.word;
            d->bits[t][r.off][dt] |= 1u | (uint32_t)(r.has_out & 1) << 1 | (uint32_t)(r.pai & 3) << 2 |
                                     (uint32_t)(r.out & 7) << 4 | (uint32_t)(r.op1 & 7) << 7 |
                                     (uint32_t)(r.op2 & 7) << 10;
        }
    }
    for (auto &p : RV_PAI0) {
        int t = rv_dense_id(p[0]);
        for (int dt = 0; dt < 6; ++dt) d->bits[t][0][dt] |= (uint32_t)(p[1] & 3) << 2;
    }
    for (const rv_slot_row &s : RV_SLOTS) {
        int t = rv_dense_id(s.type);
        for (int dt = 0; dt < 6; ++dt)
            d->bits[t][s.off][dt] |= (uint32_t)(s.iconst & 7) << 13 | (uint32_t)(s.jt & 7) << 16;
    }
    for (const rv_size_row &s : RV_SIZES) {
        int t = rv_dense_id(s.type);
        for (int dt = 0; dt < 6; ++dt)
            if (s.dmask >> dt & 1)
                for (int o = 0; o < 4; ++o) d->bits[t][o][dt] |= (uint32_t)(s.words & 7) << 19;
    }
}

#endif

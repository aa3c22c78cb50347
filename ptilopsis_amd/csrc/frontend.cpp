// Host front end: Pareas-language source -> flattened inverted-tree arrays.
//
// This is the input side of the code generator.  Its output must be identical,
// array by array, to what the reference's pareas-ast front end produces
// (subprojects/pareas-ast/src/codegen/{lexer,parser,astnode,symtab,depthtree}.cpp):
//   - tokens and literal conversion (std::stoul truncated to 32 bits, std::stof
//     bit pattern) follow lexer.cpp:48-248, including '=>' and '=<';
//   - grammar and precedence follow parser.cpp:79-540;
//   - type resolution and implicit DEREF insertion follow astnode.cpp:108-244;
//   - symbol ids are global declaration counters and function ids count
//     function declarations (symtab.cpp:10-35);
//   - the flat layout is post-order with depth, child index (root: -1) and
//     parent index (root: -1) as in depthtree.cpp:11-100.
// The node graph lives in one arena and is flattened iteratively, so programs
// with tens of millions of nodes parse in a few seconds.
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>
#include <unordered_map>
#include <stdexcept>

#include "frontend.h"

namespace ptl_fe {

enum NT : uint8_t {
    INVALID, STATEMENT_LIST, EMPTY_STAT, FUNC_DECL, FUNC_ARG, FUNC_ARG_LIST, EXPR_STAT, IF_STAT,
    IF_ELSE_STAT, WHILE_STAT, FUNC_CALL_EXPR, FUNC_CALL_ARG, FUNC_CALL_ARG_LIST, ADD_EXPR, SUB_EXPR,
    MUL_EXPR, DIV_EXPR, MOD_EXPR, BITAND_EXPR, BITOR_EXPR, BITXOR_EXPR, LSHIFT_EXPR, RSHIFT_EXPR,
    URSHIFT_EXPR, LAND_EXPR, LOR_EXPR, EQ_EXPR, NEQ_EXPR, LESS_EXPR, GREAT_EXPR, LESSEQ_EXPR,
    GREATEQ_EXPR, BITNOT_EXPR, LNOT_EXPR, NEG_EXPR, LIT_EXPR, CAST_EXPR, DEREF_EXPR, ASSIGN_EXPR,
    DECL_EXPR, ID_EXPR, WHILE_DUMMY, FUNC_DECL_DUMMY, RETURN_STAT
};
enum DT : uint8_t { DINVALID, DVOID, DINT, DFLOAT, DINT_REF, DFLOAT_REF };

static DT reference_of(DT t) { return t == DINT ? DINT_REF : t == DFLOAT ? DFLOAT_REF : DINVALID; }

struct ParseError : std::runtime_error { using std::runtime_error::runtime_error; };

// ---------------------------------------------------------------- lexer
enum TK {
    T_EOF, T_PLUS, T_MIN, T_MUL, T_DIV, T_MOD, T_BITAND, T_BITOR, T_BITXOR, T_LSHIFT, T_RSHIFT,
    T_URSHIFT, T_ASSIGN, T_NOT, T_BITNOT, T_EQ, T_NEQ, T_LESS, T_GREATER, T_LESSEQ, T_GREATEQ,
    T_LAND, T_LOR, T_OPEN_PAR, T_CLOSE_PAR, T_OPEN_CB, T_CLOSE_CB, T_COMMA, T_SEMICOLON, T_CAST,
    T_DECL, T_IF, T_ELSE, T_WHILE, T_INT, T_FLOAT, T_VOID, T_FUNCTION, T_RETURN, T_INTEGER,
    T_ID
};
// As in token.hpp/lexer.cpp:73-77, a float literal carries the same token type
// as the `float` keyword (which therefore also parses as the literal 0.0, and
// `@2.5` as a cast to float).

struct Token { TK type; uint32_t integer; uint32_t sbeg, slen; };

class Lexer {
    const char *s; size_t n, pos = 0; bool have = false; Token la{};
    static bool ws(char c) { return c == ' ' || c == '\t' || c == '\r' || c == '\n'; }
    static bool idc(char c) { return (c >= 'a' && c <= 'z') || (c >= 'A' && c <= 'Z') || c == '_'; }
    char rd() { return pos < n ? s[pos++] : '\0'; }
    void unrd(char c) { if (c != '\0' || pos < n) --pos; }
    Token number() {
        size_t b = pos; bool isf = false;
        while (pos < n && s[pos] >= '0' && s[pos] <= '9') ++pos;
        if (pos < n && s[pos] == '.') { isf = true; ++pos; while (pos < n && s[pos] >= '0' && s[pos] <= '9') ++pos; }
        std::string txt(s + b, pos - b);
        try {
            if (isf) {
                float f = std::stof(txt);
                uint32_t bits; std::memcpy(&bits, &f, 4);
                return Token{T_FLOAT, bits, 0, 0};
            }
            unsigned long v = std::stoul(txt);
            return Token{T_INTEGER, (uint32_t)v, 0, 0};
        } catch (const std::invalid_argument &) {
            throw ParseError("Invalid number " + txt);
        } catch (const std::out_of_range &) {
            throw ParseError("Number out of range " + txt);
        }
    }
    Token ident() {
        size_t b = pos;
        while (pos < n && (idc(s[pos]) || (s[pos] >= '0' && s[pos] <= '9'))) ++pos;
        size_t len = pos - b; const char *p = s + b;
        auto is = [&](const char *kw) { return strlen(kw) == len && memcmp(kw, p, len) == 0; };
        if (is("if")) return Token{T_IF, 0, 0, 0};
        if (is("else")) return Token{T_ELSE, 0, 0, 0};
        if (is("while")) return Token{T_WHILE, 0, 0, 0};
        if (is("int")) return Token{T_INT, 0, 0, 0};
        if (is("float")) return Token{T_FLOAT, 0, 0, 0};
        if (is("void")) return Token{T_VOID, 0, 0, 0};
        if (is("function")) return Token{T_FUNCTION, 0, 0, 0};
        if (is("return")) return Token{T_RETURN, 0, 0, 0};
        return Token{T_ID, 0, (uint32_t)b, (uint32_t)len};
    }
    Token next() {
        char c = rd();
        while (ws(c)) c = rd();
        switch (c) {
            case '+': return Token{T_PLUS}; case '-': return Token{T_MIN}; case '*': return Token{T_MUL};
            case '/': return Token{T_DIV}; case '%': return Token{T_MOD}; case '^': return Token{T_BITXOR};
            case '(': return Token{T_OPEN_PAR}; case ')': return Token{T_CLOSE_PAR};
            case '{': return Token{T_OPEN_CB}; case '}': return Token{T_CLOSE_CB};
            case ',': return Token{T_COMMA}; case ';': return Token{T_SEMICOLON};
            case '@': return Token{T_CAST}; case '~': return Token{T_BITNOT}; case ':': return Token{T_DECL};
            case '=': { c = rd(); if (c == '=') return Token{T_EQ}; if (c == '>') return Token{T_GREATEQ};
                        if (c == '<') return Token{T_LESSEQ}; unrd(c); return Token{T_ASSIGN}; }
            case '>': { c = rd(); if (c == '=') return Token{T_GREATEQ};
                        if (c == '>') { c = rd(); if (c == '>') return Token{T_URSHIFT}; unrd(c); return Token{T_RSHIFT}; }
                        unrd(c); return Token{T_GREATER}; }
            case '<': { c = rd(); if (c == '=') return Token{T_LESSEQ}; if (c == '<') return Token{T_LSHIFT};
                        unrd(c); return Token{T_LESS}; }
            case '!': { c = rd(); if (c == '=') return Token{T_NEQ}; unrd(c); return Token{T_NOT}; }
            case '&': { c = rd(); if (c == '&') return Token{T_LAND}; unrd(c); return Token{T_BITAND}; }
            case '|': { c = rd(); if (c == '|') return Token{T_LOR}; unrd(c); return Token{T_BITOR}; }
            default: break;
        }
        if (c >= '0' && c <= '9') { unrd(c); return number(); }
        if (idc(c)) { unrd(c); return ident(); }
        if (c == '\0' && pos >= n) return Token{T_EOF};
        throw ParseError(std::string("Unknown token ") + c);
    }
public:
    Lexer(const char *src, size_t len) : s(src), n(len) {}
    Token lex() { if (have) { have = false; return la; } return next(); }
    Token peek() { if (!have) { la = next(); have = true; } return la; }
    std::string text(const Token &t) const { return std::string(s + t.sbeg, t.slen); }
};

// ---------------------------------------------------------------- AST arena
struct Node { NT type; DT rtype; uint32_t integer; uint32_t first_child; uint32_t nchild; };

struct Arena {
    std::vector<Node> nodes;
    std::vector<uint32_t> kids;  // child index lists, contiguous per node
    uint32_t make(NT t, DT d, uint32_t integer, const uint32_t *ch, uint32_t nc) {
        Node x{t, d, integer, (uint32_t)kids.size(), nc};
        kids.insert(kids.end(), ch, ch + nc);
        nodes.push_back(x);
        return (uint32_t)nodes.size() - 1;
    }
    uint32_t make(NT t, DT d = DINVALID, uint32_t integer = 0) { return make(t, d, integer, nullptr, 0); }
    uint32_t make(NT t, DT d, uint32_t integer, std::initializer_list<uint32_t> ch) {
        return make(t, d, integer, ch.begin(), (uint32_t)ch.size());
    }
    uint32_t &child(uint32_t n, uint32_t i) { return kids[nodes[n].first_child + i]; }
};

// ---------------------------------------------------------------- symbols
struct SymTab {
    std::vector<std::unordered_map<std::string, uint32_t>> scopes;
    std::unordered_map<std::string, uint32_t> funcs;
    std::vector<DT> sym_types, func_ret;
    uint32_t declare(const std::string &nm, DT t) {
        auto &last = scopes.back();
        if (last.count(nm)) throw ParseError("Redeclaration of symbol " + nm);
        uint32_t id = (uint32_t)sym_types.size();
        last[nm] = id; sym_types.push_back(t);
        return id;
    }
    uint32_t declare_func(const std::string &nm, DT t) {
        if (funcs.count(nm)) throw ParseError("Redeclaration of function " + nm);
        uint32_t id = (uint32_t)funcs.size();
        funcs[nm] = id; func_ret.push_back(t);
        return id;
    }
    uint32_t resolve(const std::string &nm, DT *t) const {
        for (size_t i = scopes.size(); i > 0; --i) {
            auto it = scopes[i - 1].find(nm);
            if (it != scopes[i - 1].end()) { *t = sym_types[it->second]; return it->second; }
        }
        throw ParseError("Use of undeclared symbol " + nm);
    }
    uint32_t resolve_func(const std::string &nm) const {
        auto it = funcs.find(nm);
        if (it == funcs.end()) throw ParseError("Tried to call undeclared function " + nm);
        return it->second;
    }
};

// ---------------------------------------------------------------- parser
class Parser {
    Lexer &lx; Arena &A; SymTab &st;
    void expect(TK t) { Token k = lx.lex(); if (k.type != t) throw ParseError("unexpected token"); }
    static NT bin(TK t) {
        switch (t) {
            case T_PLUS: return ADD_EXPR; case T_MIN: return SUB_EXPR; case T_MUL: return MUL_EXPR;
            case T_DIV: return DIV_EXPR; case T_MOD: return MOD_EXPR; case T_EQ: return EQ_EXPR;
            case T_NEQ: return NEQ_EXPR; case T_LESS: return LESS_EXPR; case T_GREATER: return GREAT_EXPR;
            case T_LESSEQ: return LESSEQ_EXPR; case T_GREATEQ: return GREATEQ_EXPR;
            case T_BITAND: return BITAND_EXPR; case T_BITOR: return BITOR_EXPR; case T_BITXOR: return BITXOR_EXPR;
            case T_LAND: return LAND_EXPR; case T_LOR: return LOR_EXPR; case T_LSHIFT: return LSHIFT_EXPR;
            case T_RSHIFT: return RSHIFT_EXPR; case T_URSHIFT: return URSHIFT_EXPR; default: return INVALID;
        }
    }
    template <typename Next, typename Pred>
    uint32_t left_assoc(Next next, Pred pred) {
        uint32_t l = (this->*next)();
        for (Token k = lx.peek(); pred(k.type); k = lx.peek()) {
            lx.lex();
            uint32_t r = (this->*next)();
            l = A.make(bin(k.type), DINVALID, 0, {l, r});
        }
        return l;
    }
    uint32_t assign() {
        uint32_t l = logical();
        if (lx.peek().type == T_ASSIGN) { lx.lex(); uint32_t r = assign(); return A.make(ASSIGN_EXPR, DINVALID, 0, {l, r}); }
        return l;
    }
    uint32_t logical() { return left_assoc(&Parser::bitwise, [](TK t) { return t == T_LAND || t == T_LOR; }); }
    uint32_t bitwise() { return left_assoc(&Parser::compare, [](TK t) { return t == T_BITAND || t == T_BITOR || t == T_BITXOR; }); }
    uint32_t compare() {
        return left_assoc(&Parser::shift, [](TK t) {
            return t == T_EQ || t == T_NEQ || t == T_GREATER || t == T_LESS || t == T_GREATEQ || t == T_LESSEQ; });
    }
    uint32_t shift() { return left_assoc(&Parser::add, [](TK t) { return t == T_LSHIFT || t == T_RSHIFT || t == T_URSHIFT; }); }
    uint32_t add() { return left_assoc(&Parser::mul, [](TK t) { return t == T_PLUS || t == T_MIN; }); }
    uint32_t mul() { return left_assoc(&Parser::unary, [](TK t) { return t == T_MUL || t == T_DIV || t == T_MOD; }); }
    uint32_t unary() {
        Token k = lx.peek();
        if (k.type == T_MIN || k.type == T_NOT || k.type == T_BITNOT) {
            lx.lex();
            uint32_t op = unary();
            NT t = k.type == T_MIN ? NEG_EXPR : k.type == T_NOT ? LNOT_EXPR : BITNOT_EXPR;
            return A.make(t, DINVALID, 0, {op});
        }
        return cast();
    }
    uint32_t cast() {
        uint32_t l = atom();
        while (lx.peek().type == T_CAST) {
            lx.lex();
            Token k = lx.lex();
            DT d;
            if (k.type == T_INT) d = DINT; else if (k.type == T_FLOAT) d = DFLOAT;
            else throw ParseError("expecting typename");
            l = A.make(CAST_EXPR, d, 0, {l});
        }
        return l;
    }
    uint32_t atom() {
        Token k = lx.lex();
        switch (k.type) {
            case T_OPEN_PAR: { uint32_t e = assign(); expect(T_CLOSE_PAR); return e; }
            case T_ID: {
                std::string id = lx.text(k);
                Token n = lx.peek();
                if (n.type == T_DECL) {
                    lx.lex();
                    Token ty = lx.lex();
                    DT t;
                    if (ty.type == T_INT) t = DINT_REF; else if (ty.type == T_FLOAT) t = DFLOAT_REF;
                    else throw ParseError("expecting typename");
                    return A.make(DECL_EXPR, t, st.declare(id, t));
                }
                if (n.type == T_OPEN_PAR) {
                    lx.lex();
                    std::vector<uint32_t> params;
                    if (lx.peek().type != T_CLOSE_PAR) {
                        params.push_back(A.make(FUNC_CALL_ARG, DINVALID, 0, {assign()}));
                        while (lx.peek().type == T_COMMA) {
                            lx.lex();
                            params.push_back(A.make(FUNC_CALL_ARG, DINVALID, 0, {assign()}));
                        }
                    }
                    expect(T_CLOSE_PAR);
                    uint32_t al = A.make(FUNC_CALL_ARG_LIST, DINVALID, 0, params.data(), (uint32_t)params.size());
                    uint32_t fid = st.resolve_func(id);
                    return A.make(FUNC_CALL_EXPR, st.func_ret[fid], fid, {al});
                }
                DT t; uint32_t sid = st.resolve(id, &t);
                return A.make(ID_EXPR, t, sid);
            }
            case T_INTEGER: return A.make(LIT_EXPR, DINT, k.integer);
            case T_FLOAT: return A.make(LIT_EXPR, DFLOAT, k.integer);
            default: throw ParseError("expecting atom");
        }
    }
    uint32_t statement() {
        Token k = lx.peek();
        switch (k.type) {
            case T_SEMICOLON: expect(T_SEMICOLON); return A.make(EMPTY_STAT);
            case T_IF: {
                expect(T_IF);
                uint32_t c = assign();
                uint32_t s = statement();
                if (lx.peek().type == T_ELSE) { expect(T_ELSE); uint32_t e = statement(); return A.make(IF_ELSE_STAT, DINVALID, 0, {c, s, e}); }
                return A.make(IF_STAT, DINVALID, 0, {c, s});
            }
            case T_MIN: case T_BITNOT: case T_NOT: case T_OPEN_PAR: case T_INTEGER: case T_FLOAT: case T_ID: {
                uint32_t e = assign(); expect(T_SEMICOLON);
                return A.make(EXPR_STAT, DINVALID, 0, {e});
            }
            case T_OPEN_CB: {
                st.scopes.emplace_back();
                expect(T_OPEN_CB);
                uint32_t l = statement_list();
                expect(T_CLOSE_CB);
                st.scopes.pop_back();
                return l;
            }
            case T_WHILE: {
                expect(T_WHILE);
                uint32_t c = assign();
                uint32_t s = statement();
                uint32_t d = A.make(WHILE_DUMMY);
                return A.make(WHILE_STAT, DINVALID, 0, {d, c, s});
            }
            case T_RETURN: {
                expect(T_RETURN);
                uint32_t fid = (uint32_t)st.funcs.size() - 1;
                if (lx.peek().type == T_SEMICOLON) { expect(T_SEMICOLON); return A.make(RETURN_STAT, DVOID, fid); }
                uint32_t e = assign(); expect(T_SEMICOLON);
                return A.make(RETURN_STAT, DVOID, fid, {e});
            }
            default: throw ParseError("expecting start of statement");
        }
    }
    uint32_t statement_list() {
        std::vector<uint32_t> v;
        for (;;) {
            TK t = lx.peek().type;
            if (t == T_SEMICOLON || t == T_IF || t == T_MIN || t == T_BITNOT || t == T_NOT || t == T_OPEN_PAR ||
                t == T_INTEGER || t == T_FLOAT || t == T_ID || t == T_OPEN_CB || t == T_WHILE || t == T_RETURN)
                v.push_back(statement());
            else break;
        }
        return A.make(STATEMENT_LIST, DINVALID, 0, v.data(), (uint32_t)v.size());
    }
    uint32_t arg_list() {
        expect(T_OPEN_PAR);
        std::vector<uint32_t> args;
        uint32_t nint = 0;
        auto decl = [&]() {
            Token k = lx.lex();
            if (k.type != T_ID) throw ParseError("expecting identifier");
            std::string nm = lx.text(k);
            expect(T_DECL);
            Token ty = lx.lex();
            DT t;
            if (ty.type == T_INT) t = DINT_REF; else if (ty.type == T_FLOAT) t = DFLOAT_REF;
            else throw ParseError("expecting typename");
            uint32_t d = A.make(DECL_EXPR, t, st.declare(nm, t));
            uint32_t idx = t == DINT_REF ? nint++ : (uint32_t)args.size() - nint;
            args.push_back(A.make(FUNC_ARG, DINVALID, idx, {d}));
        };
        if (lx.peek().type != T_CLOSE_PAR) {
            decl();
            while (lx.peek().type == T_COMMA) { lx.lex(); decl(); }
        }
        expect(T_CLOSE_PAR);
        return A.make(FUNC_ARG_LIST, DINVALID, 0, args.data(), (uint32_t)args.size());
    }
    uint32_t function() {
        expect(T_FUNCTION);
        Token id = lx.lex();
        if (id.type != T_ID) throw ParseError("expecting identifier");
        st.scopes.emplace_back();
        uint32_t al = arg_list();
        expect(T_DECL);
        Token ty = lx.lex();
        DT rt;
        if (ty.type == T_INT) rt = DINT; else if (ty.type == T_FLOAT) rt = DFLOAT; else if (ty.type == T_VOID) rt = DVOID;
        else throw ParseError("expecting typename");
        uint32_t fid = st.declare_func(lx.text(id), rt);
        expect(T_OPEN_CB);
        uint32_t body = statement_list();
        expect(T_CLOSE_CB);
        st.scopes.pop_back();
        uint32_t dummy = A.make(FUNC_DECL_DUMMY);
        return A.make(FUNC_DECL, rt, fid, {dummy, al, body});
    }
public:
    Parser(Lexer &l, Arena &a, SymTab &s) : lx(l), A(a), st(s) {}
    uint32_t parse() {
        std::vector<uint32_t> fs;
        while (lx.peek().type == T_FUNCTION) fs.push_back(function());
        uint32_t root = A.make(STATEMENT_LIST, DINVALID, 0, fs.data(), (uint32_t)fs.size());
        expect(T_EOF);
        return root;
    }
};

// ---------------------------------------------------------------- typing
// astnode.cpp:108-244, applied bottom-up (children before parents).
static void resolve_node(Arena &A, uint32_t n) {
    auto rt = [&](uint32_t i) -> DT { return A.nodes[A.child(n, i)].rtype; };
    auto assert_type = [&](uint32_t i, std::initializer_list<DT> ok) {
        DT have = rt(i);
        for (DT t : ok) if (have == t) return;
        for (DT t : ok)
            if (have == reference_of(t)) {
                uint32_t old = A.child(n, i);
                uint32_t d = A.make(DEREF_EXPR, t, 0, {old});
                A.child(n, i) = d;
                return;
            }
        throw ParseError("Mismatched type");
    };
    switch (A.nodes[n].type) {
        case FUNC_DECL: break;
        case IF_STAT: case IF_ELSE_STAT: assert_type(0, {DINT}); A.nodes[n].rtype = DVOID; break;
        case WHILE_STAT: assert_type(1, {DINT}); A.nodes[n].rtype = DVOID; break;
        case ADD_EXPR: case SUB_EXPR: case MUL_EXPR: case DIV_EXPR:
            assert_type(0, {DINT, DFLOAT}); assert_type(1, {DINT, DFLOAT});
            if (rt(0) != rt(1)) throw ParseError("Mismatched types for children");
            A.nodes[n].rtype = rt(0); break;
        case MOD_EXPR: case BITAND_EXPR: case BITOR_EXPR: case BITXOR_EXPR: case LSHIFT_EXPR:
        case RSHIFT_EXPR: case URSHIFT_EXPR: case LAND_EXPR: case LOR_EXPR:
            assert_type(0, {DINT}); assert_type(1, {DINT}); A.nodes[n].rtype = DINT; break;
        case EQ_EXPR: case NEQ_EXPR: case LESS_EXPR: case GREAT_EXPR: case LESSEQ_EXPR: case GREATEQ_EXPR:
            assert_type(0, {DINT, DFLOAT}); assert_type(1, {DINT, DFLOAT});
            if (rt(0) != rt(1)) throw ParseError("Mismatched types for children");
            A.nodes[n].rtype = DINT; break;
        case BITNOT_EXPR: case LNOT_EXPR: assert_type(0, {DINT}); A.nodes[n].rtype = DINT; break;
        case NEG_EXPR: case FUNC_CALL_ARG: assert_type(0, {DINT, DFLOAT}); A.nodes[n].rtype = rt(0); break;
        case CAST_EXPR:
            assert_type(0, {DINT, DFLOAT});
            if (rt(0) == A.nodes[n].rtype) throw ParseError("Invalid type for child");
            break;
        case ASSIGN_EXPR:
            assert_type(0, {DINT_REF, DFLOAT_REF}); assert_type(1, {DINT, DFLOAT});
            if (rt(0) != reference_of(rt(1))) throw ParseError("Mismatched reference type");
            A.nodes[n].rtype = rt(1); break;
        case FUNC_CALL_EXPR: case LIT_EXPR: case ID_EXPR: case DECL_EXPR: break;
        case FUNC_ARG: assert_type(0, {DINT_REF, DFLOAT_REF}); A.nodes[n].rtype = rt(0); break;
        case FUNC_CALL_ARG_LIST: {
            uint32_t nint = 0, nc = A.nodes[n].nchild;
            for (uint32_t i = 0; i < nc; ++i) {
                assert_type(i, {DINT, DFLOAT});
                DT d = rt(i);
                uint32_t idx = d == DINT ? nint++ : i - nint;
                A.nodes[A.child(n, i)].integer = idx;
            }
            break;
        }
        case RETURN_STAT:
            A.nodes[n].rtype = DVOID;
            if (A.nodes[n].nchild > 0) { assert_type(0, {DINT, DFLOAT}); A.nodes[n].rtype = rt(0); }
            break;
        default: A.nodes[n].rtype = DINVALID; break;
    }
}

// Post-order over the arena without recursion; typing runs on the way up.
static void resolve_all(Arena &A, uint32_t root) {
    std::vector<std::pair<uint32_t, uint32_t>> stk;  // node, next child
    stk.push_back({root, 0});
    while (!stk.empty()) {
        auto &top = stk.back();
        uint32_t n = top.first;
        if (top.second < A.nodes[n].nchild) {
            uint32_t c = A.child(n, top.second++);
            stk.push_back({c, 0});
        } else {
            resolve_node(A, n);
            stk.pop_back();
        }
    }
}

// Post-order with children in order (depthtree.cpp:52-100), iteratively.  G
// gives a node's type, result type, integer, child count and i-th child.
template <class G>
static void flatten_with(const G &g, uint32_t root, ptl_host_tree *out) {
    uint64_t total = 0;
    {
        std::vector<uint32_t> s{root};
        while (!s.empty()) { uint32_t n = s.back(); s.pop_back(); ++total;
            for (uint32_t i = 0; i < g.nchild(n); ++i) s.push_back(g.child(n, i)); }
    }
    out->n_nodes = total;
    out->node_types = (uint8_t *)malloc(total);
    out->result_types = (uint8_t *)malloc(total);
    out->parents = (int32_t *)malloc(4 * total);
    out->depth = (int32_t *)malloc(4 * total);
    out->child_idx = (int32_t *)malloc(4 * total);
    out->node_data = (uint32_t *)malloc(4 * total);
    struct Fr { uint32_t n, next, depth; int32_t cidx; };
    std::vector<Fr> stk;
    std::vector<uint32_t> kid_pos;   // flat index of finished children, per open frame
    std::vector<uint32_t> frame_kids; // start offset into kid_pos per frame
    stk.push_back({root, 0, 0, -1});
    frame_kids.push_back(0);
    uint64_t i = 0; uint32_t maxd = 0;
    while (!stk.empty()) {
        Fr &f = stk.back();
        if (f.next < g.nchild(f.n)) {
            uint32_t c = g.child(f.n, f.next);
            int32_t ci = (int32_t)f.next;
            uint32_t d = f.depth + 1;
            f.next++;
            stk.push_back({c, 0, d, ci});
            frame_kids.push_back((uint32_t)kid_pos.size());
            continue;
        }
        out->node_types[i] = g.type(f.n);
        out->result_types[i] = g.rtype(f.n);
        out->depth[i] = (int32_t)f.depth;
        out->child_idx[i] = f.cidx;
        out->node_data[i] = g.integer(f.n);
        out->parents[i] = -1;
        if (f.depth > maxd) maxd = f.depth;
        uint32_t kb = frame_kids.back();
        for (size_t k = kb; k < kid_pos.size(); ++k) out->parents[kid_pos[k]] = (int32_t)i;
        kid_pos.resize(kb);
        frame_kids.pop_back();
        stk.pop_back();
        kid_pos.push_back((uint32_t)i);
        ++i;
    }
    out->max_depth = maxd;
}

struct ArenaView {
    Arena &A;
    uint8_t type(uint32_t n) const { return A.nodes[n].type; }
    uint8_t rtype(uint32_t n) const { return A.nodes[n].rtype; }
    uint32_t integer(uint32_t n) const { return A.nodes[n].integer; }
    uint32_t nchild(uint32_t n) const { return A.nodes[n].nchild; }
    uint32_t child(uint32_t n, uint32_t i) const { return A.kids[A.nodes[n].first_child + i]; }
};

struct RecordView {
    const ptl_host_arena &a;
    uint8_t type(uint32_t n) const { return a.node_types[n]; }
    uint8_t rtype(uint32_t n) const { return a.result_types[n]; }
    uint32_t integer(uint32_t n) const { return a.node_data[n]; }
    uint32_t nchild(uint32_t n) const { return a.n_children[n]; }
    uint32_t child(uint32_t n, uint32_t i) const { return a.kids[a.first_child[n] + i]; }
};

static void flatten(Arena &A, uint32_t root, ptl_host_tree *out) { flatten_with(ArenaView{A}, root, out); }

}  // namespace ptl_fe

extern "C" int ptl_parse_source(const char *src, uint64_t len, ptl_host_tree *out, char *err, uint64_t err_len) {
    using namespace ptl_fe;
    std::memset(out, 0, sizeof(*out));
    try {
        Lexer lx(src, (size_t)len);
        Arena A;
        A.nodes.reserve(len / 3 + 16);
        A.kids.reserve(len / 3 + 16);
        SymTab st;
        Parser p(lx, A, st);
        uint32_t root = p.parse();
        resolve_all(A, root);
        flatten(A, root, out);
        out->n_functions = (uint32_t)st.funcs.size();
        return 0;
    } catch (const std::exception &e) {
        if (err && err_len) { std::strncpy(err, e.what(), err_len - 1); err[err_len - 1] = 0; }
        return 1;
    }
}

extern "C" int ptl_parse_arena(const char *src, uint64_t len, ptl_host_arena *out, char *err, uint64_t err_len) {
    using namespace ptl_fe;
    std::memset(out, 0, sizeof(*out));
    try {
        Lexer lx(src, (size_t)len);
        Arena A;
        A.nodes.reserve(len / 3 + 16);
        A.kids.reserve(len / 3 + 16);
        SymTab st;
        Parser p(lx, A, st);
        uint32_t root = p.parse();
        resolve_all(A, root);
        const uint64_t n = A.nodes.size(), k = A.kids.size();
        out->node_types = (uint8_t *)malloc(n ? n : 1);
        out->result_types = (uint8_t *)malloc(n ? n : 1);
        out->node_data = (uint32_t *)malloc(4 * (n ? n : 1));
        out->first_child = (uint32_t *)malloc(4 * (n ? n : 1));
        out->n_children = (uint32_t *)malloc(4 * (n ? n : 1));
        out->kids = (uint32_t *)malloc(4 * (k ? k : 1));
        if (!out->node_types || !out->result_types || !out->node_data || !out->first_child || !out->n_children ||
            !out->kids)
            throw std::bad_alloc();
        for (uint64_t v = 0; v < n; ++v) {
            const Node &x = A.nodes[v];
            out->node_types[v] = x.type; out->result_types[v] = x.rtype; out->node_data[v] = x.integer;
            out->first_child[v] = x.first_child; out->n_children[v] = x.nchild;
        }
        if (k) std::memcpy(out->kids, A.kids.data(), 4 * k);
        out->n_nodes = n; out->n_kids = k; out->root = root;
        out->n_functions = (uint32_t)st.funcs.size();
        return 0;
    } catch (const std::exception &e) {
        ptl_free_host_arena(out);
        if (err && err_len) { std::strncpy(err, e.what(), err_len - 1); err[err_len - 1] = 0; }
        return 1;
    }
}

extern "C" int ptl_flatten_records(const ptl_host_arena *a, ptl_host_tree *out) {
    using namespace ptl_fe;
    std::memset(out, 0, sizeof(*out));
    if (!a || a->root >= a->n_nodes) return 1;
    flatten_with(RecordView{*a}, a->root, out);
    out->n_functions = a->n_functions;
    return 0;
}

extern "C" void ptl_free_host_arena(ptl_host_arena *a) {
    free(a->node_types); free(a->result_types); free(a->node_data); free(a->first_child); free(a->n_children);
    free(a->kids);
    std::memset(a, 0, sizeof(*a));
}

extern "C" void ptl_free_host_tree(ptl_host_tree *t) {
    free(t->node_types); free(t->result_types); free(t->parents); free(t->depth);
    free(t->child_idx); free(t->node_data);
    std::memset(t, 0, sizeof(*t));
}

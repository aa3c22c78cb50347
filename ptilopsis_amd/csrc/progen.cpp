// Seeded generator of synthetic Pareas-language programs.
//
// It emits source text that the reference's front end accepts
// (grammar: pareas-ast src/codegen/parser.cpp:79-540; typing rules:
// astnode.cpp:148-243; declaration-before-use of functions: symtab.cpp:63-67)
// and counts the AST nodes the front end will build for it, including the
// implicit DEREF nodes inserted on every variable read, so size targets are met
// exactly enough to steer workload shape.
//
// Workload knobs (statement mix, expression depth, argument counts, nesting,
// float share, per-function size targets) are passed in ptl_gen_params; the
// named workloads live in ptilopsis_amd/workloads.py.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <cmath>
#include <string>
#include <vector>

#include "progen.h"

namespace {

struct Rng {
    uint64_t s[4];
    static uint64_t splitmix(uint64_t &x) {
        uint64_t z = (x += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    explicit Rng(uint64_t seed) { uint64_t x = seed; for (auto &v : s) v = splitmix(x); }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {  // xoshiro256**
        uint64_t r = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
        s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3]; s[2] ^= t; s[3] = rotl(s[3], 45);
        return r;
    }
    double uni() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
    uint32_t range(uint32_t lo, uint32_t hi) { return lo + (uint32_t)(next() % (uint64_t)(hi - lo + 1)); }
    bool coin(double p) { return uni() < p; }
};

enum Ty { TI = 0, TF = 1 };

struct Var { uint32_t id; Ty ty; };
struct Func { std::vector<Ty> args; int ret; /* 0 int 1 float 2 void */ };

struct Gen {
    const ptl_gen_params &P;
    Rng rng;
    std::string out;
    uint64_t nodes = 0;
    std::vector<Func> funcs;
    std::vector<std::vector<Var>> scopes;
    uint32_t var_counter = 0;
    int cur_ret = 0;
    uint32_t cur_func = 0;

    Gen(const ptl_gen_params &p) : P(p), rng(p.seed) {}

    Ty pick_ty() { return rng.coin(P.float_frac) ? TF : TI; }

    const Var *pick_var(Ty t) {
        // uniform over visible variables of type t (a few random probes, then a scan)
        size_t total = 0;
        for (auto &s : scopes) total += s.size();
        if (total == 0) return nullptr;
        for (int tries = 0; tries < 6; ++tries) {
            size_t k = rng.next() % total;
            for (auto &s : scopes) {
                if (k < s.size()) { if (s[k].ty == t) return &s[k]; break; }
                k -= s.size();
            }
        }
        for (size_t si = scopes.size(); si > 0; --si)
            for (size_t i = scopes[si - 1].size(); i > 0; --i)
                if (scopes[si - 1][i - 1].ty == t) return &scopes[si - 1][i - 1];
        return nullptr;
    }

    void lit(Ty t) {
        char buf[48];
        if (t == TI) {
            int32_t v = (int32_t)rng.range(0, 65535) - 32768;
            if (v < 0) { snprintf(buf, sizeof buf, "-%d", -v); nodes += 2; }
            else { snprintf(buf, sizeof buf, "%d", v); nodes += 1; }
        } else {
            double v = rng.uni() * 2000.0 - 1000.0;
            if (v < 0) { snprintf(buf, sizeof buf, "-%.6f", -v); nodes += 2; }
            else { snprintf(buf, sizeof buf, "%.6f", v); nodes += 1; }
        }
        out += buf;
    }

    void leaf(Ty t) {
        if (rng.coin(P.p_var_leaf)) {
            const Var *v = pick_var(t);
            if (v) { out += "v"; out += std::to_string(v->id); nodes += 2; return; }  // ID + implicit DEREF
        }
        lit(t);
    }

    uint32_t geo_depth() {
        uint32_t d = 1;
        while (!rng.coin(P.expr_p) && d < 64) ++d;
        return d;
    }

    void binop(const char *op, Ty lt, uint32_t dl, Ty rt, uint32_t dr) {
        out += "(";
        expr(lt, dl);
        out += op;
        expr(rt, dr);
        out += ")";
        nodes += 1;
    }

    void expr(Ty t, uint32_t d) {
        if (d <= 1) { leaf(t); return; }
        uint32_t da, db;
        if (P.balanced && rng.coin(0.6)) { da = db = d - 1; }
        else {
            // the expression depth is carried by one operand; the other is a
            // leaf with probability p_leaf_operand, else of uniform smaller depth
            da = d - 1;
            db = rng.coin(P.p_leaf_operand) ? 1 : 1 + (uint32_t)(rng.next() % (d - 1));
            if (rng.coin(0.5)) { uint32_t x = da; da = db; db = x; }
        }
        if (t == TF) {
            static const char *ops[] = { " + ", " - ", " * ", " / " };
            binop(ops[rng.next() % 4], TF, da, TF, db);
            return;
        }
        double r = rng.uni();
        if (r < P.p_arith) {
            static const char *ops[] = { " + ", " - ", " * ", " / ", " % " };
            binop(ops[rng.next() % 5], TI, da, TI, db);
        } else if (r < P.p_arith + P.p_compare) {
            static const char *ops[] = { " == ", " != ", " < ", " > ", " <= ", " >= " };
            Ty ct = pick_ty();
            binop(ops[rng.next() % 6], ct, da, ct, db);
        } else {
            static const char *ops[] = { " && ", " || " };
            binop(ops[rng.next() % 2], TI, da, TI, db);
        }
    }

    void cond() {
        // conditions are comparisons (or && / || of comparisons) over typed operands
        uint32_t d = P.balanced ? rng.range(P.balanced_min, P.balanced_max) : geo_depth();
        if (d < 2) d = 2;
        static const char *ops[] = { " == ", " != ", " < ", " > ", " <= ", " >= " };
        Ty ct = pick_ty();
        uint32_t db = rng.coin(P.p_leaf_operand) ? 1 : 1 + (uint32_t)(rng.next() % (d - 1));
        binop(ops[rng.next() % 6], ct, d - 1, ct, db);
    }

    uint32_t expr_depth() { return P.balanced ? rng.range(P.balanced_min, P.balanced_max) : geo_depth(); }

    void call_expr(uint32_t callee) {
        const Func &f = funcs[callee];
        out += "f"; out += std::to_string(callee); out += "(";
        nodes += 2;  // FUNC_CALL_EXPR + FUNC_CALL_ARG_LIST
        for (size_t i = 0; i < f.args.size(); ++i) {
            if (i) out += ", ";
            nodes += 1;  // FUNC_CALL_ARG
            uint32_t d = P.balanced ? 1 + (uint32_t)(rng.next() % 3) : geo_depth();
            expr(f.args[i], d);
        }
        out += ")";
    }

    void indent(uint32_t lvl) { out.append(lvl * 2, ' '); }

    void body(uint32_t lvl, uint32_t nstmt, uint32_t nest) {
        scopes.emplace_back();
        for (uint32_t i = 0; i < nstmt; ++i) statement(lvl, nest);
        scopes.pop_back();
    }

    // A comparison of depth 2 (the nesting, not the condition, is the point).
    void small_cond() {
        static const char *ops[] = { " == ", " != ", " < ", " > ", " <= ", " >= " };
        Ty ct = pick_ty();
        binop(ops[rng.next() % 6], ct, 2, ct, 1);
    }

    // k nested if/while levels, each holding one plain statement and the next
    // level: long chains of branches and jumps for fix_jumps.
    void chain(uint32_t lvl, uint32_t k) {
        if (k == 0) { statement(lvl, P.max_nest, false); return; }
        indent(lvl);
        if (rng.coin(0.5)) { out += "while "; nodes += 3; }  // WHILE + WHILE_DUMMY + body STATEMENT_LIST
        else { out += "if "; nodes += 2; }                    // IF + then STATEMENT_LIST
        small_cond();
        out += " {\n";
        scopes.emplace_back();
        statement(lvl + 1, P.max_nest, false);
        chain(lvl + 1, k - 1);
        scopes.pop_back();
        indent(lvl); out += "}\n";
    }

    void statement(uint32_t lvl, uint32_t nest, bool top = false) {
        if (top && P.chain_max && rng.coin(P.p_chain)) { chain(lvl, rng.range(P.chain_min, P.chain_max)); return; }
        indent(lvl);
        double r = rng.uni();
        bool can_nest = nest < P.max_nest;
        double pa = P.p_assign, pi = P.p_if, pw = P.p_while, pr = P.p_return;
        if (!can_nest) { pa += pi + pw; pi = pw = 0; }
        if (r < pa) {
            Ty t = pick_ty();
            const Var *v = rng.coin(0.5) ? pick_var(t) : nullptr;
            nodes += 2;  // EXPR_STAT + ASSIGN
            if (v) { out += "v"; out += std::to_string(v->id); nodes += 1; }
            else {
                uint32_t id = var_counter++;
                out += "v"; out += std::to_string(id); out += t == TI ? " : int" : " : float";
                nodes += 1;  // DECL
                out += " = ";
                expr(t, expr_depth());
                out += ";\n";
                scopes.back().push_back(Var{id, t});
                return;
            }
            out += " = ";
            expr(t, expr_depth());
            out += ";\n";
        } else if (r < pa + pi) {
            bool has_else = rng.coin(0.5);
            out += "if ";
            nodes += 2;  // IF or IF_ELSE + then STATEMENT_LIST
            cond();
            out += " {\n";
            body(lvl + 1, rng.range(P.nested_min, P.nested_max), nest + 1);
            indent(lvl); out += "}";
            if (has_else) {
                out += " else {\n";
                nodes += 1;  // else STATEMENT_LIST
                body(lvl + 1, rng.range(P.nested_min, P.nested_max), nest + 1);
                indent(lvl); out += "}";
            }
            out += "\n";
        } else if (r < pa + pi + pw) {
            out += "while ";
            nodes += 3;  // WHILE + WHILE_DUMMY + body STATEMENT_LIST
            cond();
            out += " {\n";
            body(lvl + 1, rng.range(P.nested_min, P.nested_max), nest + 1);
            indent(lvl); out += "}\n";
        } else if (r < pa + pi + pw + pr) {
            nodes += 1;  // RETURN
            if (cur_ret == 2) { out += "return;\n"; return; }
            out += "return ";
            expr(cur_ret == 0 ? TI : TF, expr_depth());
            out += ";\n";
        } else {
            uint32_t callee = (uint32_t)(rng.next() % (cur_func + 1));
            nodes += 1;  // EXPR_STAT
            call_expr(callee);
            out += ";\n";
        }
    }

    void function(uint32_t fi, uint64_t target) {
        Func f;
        uint32_t na = rng.range(P.min_args, P.max_args);
        for (uint32_t i = 0; i < na; ++i) f.args.push_back(pick_ty());
        double rr = rng.uni();
        f.ret = rr < 0.45 ? 0 : rr < 0.9 ? 1 : 2;
        funcs.push_back(f);
        cur_func = fi; cur_ret = f.ret;
        uint64_t start = nodes;
        out += "function f"; out += std::to_string(fi); out += "(";
        nodes += 4;  // FUNC_DECL + FUNC_DECL_DUMMY + FUNC_ARG_LIST + body STATEMENT_LIST
        scopes.clear();
        scopes.emplace_back();
        for (uint32_t i = 0; i < na; ++i) {
            if (i) out += ", ";
            uint32_t id = var_counter++;
            out += "v"; out += std::to_string(id); out += f.args[i] == TI ? " : int" : " : float";
            nodes += 2;  // FUNC_ARG + DECL
            scopes.back().push_back(Var{id, f.args[i]});
        }
        out += ") : ";
        out += f.ret == 0 ? "int" : f.ret == 1 ? "float" : "void";
        out += " {\n";
        if (target) {
            scopes.emplace_back();
            while (nodes - start < target) statement(1, 0, true);
            scopes.pop_back();
        } else {
            body(1, rng.range(P.min_stmts, P.max_stmts), 0);
        }
        out += "}\n\n";
    }
};

}  // namespace

extern "C" int ptl_generate_source(const ptl_gen_params *p, const uint32_t *targets, char **out_src,
                                   uint64_t *out_len, uint64_t *out_nodes) {
    Gen g(*p);
    g.out.reserve(targets ? 64 * 1024 : 1 << 20);
    g.nodes = 1;  // root STATEMENT_LIST
    for (uint32_t f = 0; f < p->n_functions; ++f) {
        if (p->total_target && g.nodes >= p->total_target) break;
        g.function(f, targets ? targets[f] : 0);
    }
    char *buf = (char *)malloc(g.out.size() + 1);
    if (!buf) return 1;
    memcpy(buf, g.out.data(), g.out.size());
    buf[g.out.size()] = 0;
    *out_src = buf;
    *out_len = g.out.size();
    *out_nodes = g.nodes;
    return 0;
}

extern "C" void ptl_free_source(char *s) { free(s); }

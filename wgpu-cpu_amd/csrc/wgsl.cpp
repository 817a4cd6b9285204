// wgsl.cpp -- WGSL front end (lexer, parser, type checker, C++ emitter).
//
// Semantics follow the reference's shader compiler, which is naga's WGSL
// front end plus naga-cranelift's code generation:
//   - abstract-typed constant expressions are folded at translation time
//     (AbstractInt in 64-bit integers, AbstractFloat in binary64) and
//     converted on use, as naga's constant evaluator does;
//   - operators, conversions and accesses are emitted with the semantics of
//     naga-cranelift/src/expression/{binary,unary,as,access,compose,
//     select,zero_value}.rs (see wgsl_prelude.h for the scalar helpers);
//   - operands are evaluated left to right into temporaries, the order of
//     naga's expression arena that the reference compiles in
//     (statement/emit.rs);
//   - every function returns its abort code (0 = continue) and its value
//     through a pointer, like the reference's compiled functions
//     (function.rs:55-68; tests.rs:400-436).
// Constructs the reference leaves as todo!() are rejected with an error
// that names them.  Where the reference's behaviour is undefined (dynamic
// indices outside a variable, reading stack memory past it), the emitted
// code aborts the invocation with the out-of-bounds code instead.
#include "wgsl.h"

#include <algorithm>
#include <cerrno>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <set>
#include <stdexcept>

namespace wgc {
namespace wgsl {
namespace {

// ------------------------------------------------------------------
// errors
// ------------------------------------------------------------------

struct Error : std::runtime_error {
    explicit Error(const std::string &m) : std::runtime_error(m) {}
};

struct Loc {
    int line = 0, col = 0;
};

[[noreturn]] void fail(Loc l, const std::string &msg)
{
    throw Error(std::to_string(l.line) + ":" + std::to_string(l.col) + ": " + msg);
}

// ------------------------------------------------------------------
// lexer
// ------------------------------------------------------------------

enum TokKind { T_EOF, T_IDENT, T_INT, T_FLOAT, T_PUNCT };

struct Tok {
    TokKind k = T_EOF;
    std::string s;       // identifier / punctuation / literal text
    Loc loc;
    int64_t ival = 0;    // T_INT
    double fval = 0.0;   // T_FLOAT
    char suffix = 0;     // 'i', 'u', 'f', or 0 (abstract)
};

std::vector<Tok> lex(const std::string &src)
{
    static const char *punct3[] = {"<<=", ">>=", nullptr};
    static const char *punct2[] = {"->", "==", "!=", "<=", ">=", "&&", "||", "<<", ">>", "+=", "-=", "*=",
                                   "/=", "%=", "&=", "|=", "^=", "++", "--", nullptr};
    std::vector<Tok> out;
    size_t i = 0, n = src.size();
    int line = 1;
    size_t line_start = 0;
    auto loc_at = [&](size_t p) { return Loc{line, (int)(p - line_start) + 1}; };
    while (i < n) {
        char c = src[i];
        if (c == '\n') {
            ++line;
            line_start = ++i;
            continue;
        }
        if (c == ' ' || c == '\t' || c == '\r' || c == '\f' || c == '\v') {
            ++i;
            continue;
        }
        if (c == '/' && i + 1 < n && src[i + 1] == '/') {
            while (i < n && src[i] != '\n') ++i;
            continue;
        }
        if (c == '/' && i + 1 < n && src[i + 1] == '*') {
            // block comments nest in WGSL
            int depth = 0;
            Loc start = loc_at(i);
            do {
                if (i + 1 < n && src[i] == '/' && src[i + 1] == '*') {
                    ++depth;
                    i += 2;
                } else if (i + 1 < n && src[i] == '*' && src[i + 1] == '/') {
                    --depth;
                    i += 2;
                } else {
                    if (src[i] == '\n') {
                        ++line;
                        line_start = i + 1;
                    }
                    ++i;
                }
            } while (depth > 0 && i < n);
            if (depth > 0) fail(start, "unterminated block comment");
            continue;
        }
        Tok t;
        t.loc = loc_at(i);
        if (isalpha((unsigned char)c) || c == '_') {
            size_t j = i;
            while (j < n && (isalnum((unsigned char)src[j]) || src[j] == '_')) ++j;
            t.k = T_IDENT;
            t.s = src.substr(i, j - i);
            i = j;
            out.push_back(t);
            continue;
        }
        if (isdigit((unsigned char)c) || (c == '.' && i + 1 < n && isdigit((unsigned char)src[i + 1]))) {
            size_t j = i;
            bool is_float = false;
            if (c == '0' && j + 1 < n && (src[j + 1] == 'x' || src[j + 1] == 'X')) {
                j += 2;
                size_t d0 = j;
                while (j < n && isxdigit((unsigned char)src[j])) ++j;
                if (j == d0) fail(t.loc, "malformed hexadecimal literal");
                if (j < n && (src[j] == '.' || src[j] == 'p' || src[j] == 'P'))
                    fail(t.loc, "hexadecimal float literals are not supported");
                t.s = src.substr(i, j - i);
                t.ival = (int64_t)strtoull(t.s.c_str() + 2, nullptr, 16);
                if (t.s.size() > 18 || (uint64_t)t.ival > 0x7FFFFFFFFFFFFFFFull) fail(t.loc, "integer literal too large");
            } else {
                while (j < n && isdigit((unsigned char)src[j])) ++j;
                if (j < n && src[j] == '.') {
                    is_float = true;
                    ++j;
                    while (j < n && isdigit((unsigned char)src[j])) ++j;
                }
                if (j < n && (src[j] == 'e' || src[j] == 'E')) {
                    size_t k = j + 1;
                    if (k < n && (src[k] == '+' || src[k] == '-')) ++k;
                    if (k < n && isdigit((unsigned char)src[k])) {
                        is_float = true;
                        j = k;
                        while (j < n && isdigit((unsigned char)src[j])) ++j;
                    }
                }
                t.s = src.substr(i, j - i);
                if (is_float) {
                    t.fval = strtod(t.s.c_str(), nullptr);
                } else {
                    if (t.s.size() > 1 && t.s[0] == '0' && !(j < n && src[j] == 'f'))
                        fail(t.loc, "leading zeros are not allowed in integer literals");
                    errno = 0;
                    unsigned long long v = strtoull(t.s.c_str(), nullptr, 10);
                    if (errno || v > 0x7FFFFFFFFFFFFFFFull) fail(t.loc, "integer literal too large");
                    t.ival = (int64_t)v;
                }
            }
            if (j < n && (src[j] == 'i' || src[j] == 'u' || src[j] == 'f' || src[j] == 'h')) {
                char sfx = src[j];
                if (sfx == 'h') fail(t.loc, "f16 is not supported (the reference has no f16 path)");
                if ((sfx == 'i' || sfx == 'u') && is_float) fail(t.loc, "integer suffix on a float literal");
                if (sfx == 'f' && !is_float) {
                    is_float = true;
                    t.fval = (double)t.ival;
                }
                t.suffix = sfx;
                ++j;
            }
            if (j < n && (isalnum((unsigned char)src[j]) || src[j] == '_')) fail(t.loc, "malformed number");
            t.k = is_float ? T_FLOAT : T_INT;
            i = j;
            out.push_back(t);
            continue;
        }
        t.k = T_PUNCT;
        bool found = false;
        for (int m = 0; punct3[m] && !found; ++m)
            if (src.compare(i, 3, punct3[m]) == 0) {
                t.s = punct3[m];
                found = true;
            }
        for (int m = 0; punct2[m] && !found; ++m)
            if (src.compare(i, 2, punct2[m]) == 0) {
                t.s = punct2[m];
                found = true;
            }
        if (!found) {
            if (!strchr("(){}[]<>,;:.=+-*/%&|^!~@", c)) fail(t.loc, std::string("unexpected character '") + c + "'");
            t.s = std::string(1, c);
        }
        i += t.s.size();
        out.push_back(t);
    }
    Tok e;
    e.k = T_EOF;
    e.loc = Loc{line, (int)(n - line_start) + 1};
    out.push_back(e);
    return out;
}

// ------------------------------------------------------------------
// syntax tree
// ------------------------------------------------------------------

struct TypeAst;
struct Expr;
struct Stmt;
using TypeP = std::shared_ptr<TypeAst>;
using ExprP = std::shared_ptr<Expr>;
using StmtP = std::shared_ptr<Stmt>;

struct TypeAst {
    Loc loc;
    std::string name;
    std::vector<TypeP> params;   // element types
    ExprP count;                 // array element count (null: runtime-sized)
    std::string access;          // ptr access mode, unused
};

struct Attr {
    Loc loc;
    std::string name;
    std::vector<ExprP> args;
};

struct Expr {
    enum Kind { LIT, IDENT, UNARY, BINARY, CALL, MEMBER, INDEX } k = LIT;
    Loc loc;
    Tok lit;                     // LIT
    std::string name;            // IDENT / CALL callee / MEMBER field / UNARY & BINARY operator
    TypeP tmpl;                  // CALL with a templated type: vec4<f32>(...), array<i32, 4>(...)
    std::vector<ExprP> a;        // operands / arguments
};

struct Stmt {
    enum Kind { BLOCK, VAR, LET, CONST, ASSIGN, COMPOUND, INCR, DECR, PHONY, CALL, IF, SWITCH, LOOP, FOR, WHILE,
                BREAK, BREAK_IF, CONTINUE, RETURN, DISCARD, EMPTY } k = EMPTY;
    Loc loc;
    std::string name;            // VAR / LET / CONST name; COMPOUND operator
    TypeP type;                  // declared type
    ExprP e, e2;                 // init / lhs, rhs / condition / value
    std::vector<StmtP> body;     // BLOCK; IF then; LOOP body
    std::vector<StmtP> alt;      // IF else (one IF or a BLOCK); LOOP continuing
    StmtP init, update;          // FOR
    bool has_continuing = false;
    struct Case {
        std::vector<ExprP> sel;  // empty + dflt: default only
        bool dflt = false;
        std::vector<StmtP> body;
        Loc loc;
    };
    std::vector<Case> cases;     // SWITCH
};

struct Param {
    Loc loc;
    std::string name;
    TypeP type;
    std::vector<Attr> attrs;
};

struct FnDecl {
    Loc loc;
    std::string name;
    std::vector<Attr> attrs;
    std::vector<Param> params;
    TypeP ret;
    std::vector<Attr> ret_attrs;
    std::vector<StmtP> body;
};

struct MemberDecl {
    Loc loc;
    std::string name;
    TypeP type;
    std::vector<Attr> attrs;
};

struct StructDecl {
    Loc loc;
    std::string name;
    std::vector<MemberDecl> members;
};

struct GlobalDecl {
    enum Kind { CONST, VAR, ALIAS, OVERRIDE } k = CONST;
    Loc loc;
    std::string name;
    std::string space, access;   // var<space, access>
    std::vector<Attr> attrs;
    TypeP type;
    ExprP init;
};

struct ModuleAst {
    std::vector<StructDecl> structs;
    std::vector<GlobalDecl> globals;
    std::vector<FnDecl> fns;
    std::vector<std::pair<Loc, ExprP>> asserts;
};

// ------------------------------------------------------------------
// parser
// ------------------------------------------------------------------

bool is_templated_type(const std::string &n)
{
    static const char *names[] = {"vec2", "vec3", "vec4", "mat2x2", "mat2x3", "mat2x4", "mat3x2", "mat3x3",
                                  "mat3x4", "mat4x2", "mat4x3", "mat4x4", "array", "ptr", "atomic",
                                  "texture_2d", "texture_1d", "texture_3d", "texture_cube", "texture_2d_array",
                                  "texture_cube_array", "texture_multisampled_2d", "texture_storage_2d",
                                  "bitcast", nullptr};
    for (int i = 0; names[i]; ++i)
        if (n == names[i]) return true;
    return false;
}

class Parser {
public:
    explicit Parser(std::vector<Tok> toks) : t_(std::move(toks)) {}

    ModuleAst parse_module()
    {
        ModuleAst m;
        while (peek().k != T_EOF) {
            if (is_ident("enable")) fail(peek().loc, "'enable' extensions are not supported");
            if (is_ident("requires") || is_ident("diagnostic")) {
                while (!is_punct(";")) next();
                next();
                continue;
            }
            if (accept(";")) continue;
            std::vector<Attr> attrs = parse_attrs();
            if (is_ident("struct")) {
                m.structs.push_back(parse_struct());
            } else if (is_ident("fn")) {
                FnDecl f = parse_fn();
                f.attrs = attrs;
                m.fns.push_back(std::move(f));
            } else if (is_ident("const") || is_ident("var") || is_ident("alias") || is_ident("override")) {
                GlobalDecl g = parse_global();
                g.attrs = attrs;
                m.globals.push_back(std::move(g));
            } else if (is_ident("const_assert")) {
                Loc l = next().loc;
                m.asserts.push_back({l, parse_expr()});
                expect(";");
            } else {
                fail(peek().loc, "expected a module-scope declaration, found '" + peek().s + "'");
            }
        }
        return m;
    }

private:
    std::vector<Tok> t_;
    size_t p_ = 0;

    const Tok &peek(size_t k = 0) const { return t_[std::min(p_ + k, t_.size() - 1)]; }
    Tok next() { return t_[std::min(p_++, t_.size() - 1)]; }
    bool is_punct(const char *s, size_t k = 0) const { return peek(k).k == T_PUNCT && peek(k).s == s; }
    bool is_ident(const char *s, size_t k = 0) const { return peek(k).k == T_IDENT && peek(k).s == s; }
    bool accept(const char *s)
    {
        if (is_punct(s)) {
            ++p_;
            return true;
        }
        return false;
    }
    void expect(const char *s)
    {
        if (!accept(s)) fail(peek().loc, std::string("expected '") + s + "', found '" + describe(peek()) + "'");
    }
    static std::string describe(const Tok &t) { return t.k == T_EOF ? "end of input" : t.s; }
    std::string ident()
    {
        if (peek().k != T_IDENT) fail(peek().loc, "expected an identifier, found '" + describe(peek()) + "'");
        return next().s;
    }
    // Closing '>' of a template list; splits '>>', '>=' and '>>='.
    void expect_close()
    {
        Tok &t = t_[p_];
        if (t.k == T_PUNCT && t.s == ">") {
            ++p_;
            return;
        }
        if (t.k == T_PUNCT && (t.s == ">>" || t.s == ">=" || t.s == ">>=")) {
            t.s = t.s.substr(1);
            t.loc.col += 1;
            return;
        }
        fail(t.loc, "expected '>' closing a template list");
    }

    std::vector<Attr> parse_attrs()
    {
        std::vector<Attr> out;
        while (is_punct("@")) {
            Attr a;
            a.loc = next().loc;
            a.name = ident();
            if (accept("(")) {
                while (!is_punct(")")) {
                    a.args.push_back(parse_expr());
                    if (!accept(",")) break;
                }
                expect(")");
            }
            out.push_back(std::move(a));
        }
        return out;
    }

    TypeP parse_type()
    {
        auto ty = std::make_shared<TypeAst>();
        ty->loc = peek().loc;
        ty->name = ident();
        if (is_templated_type(ty->name) && accept("<")) {
            if (ty->name == "array") {
                ty->params.push_back(parse_type());
                if (accept(",")) {
                    if (!is_punct(">") && !is_punct(">>")) ty->count = parse_shift_expr();
                    accept(",");
                }
            } else if (ty->name == "ptr") {
                ty->access = ident();   // address space
                expect(",");
                ty->params.push_back(parse_type());
                if (accept(",")) ident();
            } else {
                ty->params.push_back(parse_type());
                while (accept(",")) {
                    if (peek().k == T_IDENT) ty->access = ident();
                }
            }
            expect_close();
        }
        return ty;
    }

    StructDecl parse_struct()
    {
        StructDecl s;
        s.loc = next().loc;
        s.name = ident();
        expect("{");
        while (!accept("}")) {
            MemberDecl m;
            m.attrs = parse_attrs();
            m.loc = peek().loc;
            m.name = ident();
            expect(":");
            m.type = parse_type();
            s.members.push_back(std::move(m));
            if (!accept(",")) {
                expect("}");
                break;
            }
        }
        if (s.members.empty()) fail(s.loc, "a structure must have at least one member");
        return s;
    }

    GlobalDecl parse_global()
    {
        GlobalDecl g;
        g.loc = peek().loc;
        std::string kw = next().s;
        if (kw == "alias") {
            g.k = GlobalDecl::ALIAS;
            g.name = ident();
            expect("=");
            g.type = parse_type();
            expect(";");
            return g;
        }
        if (kw == "override") fail(g.loc, "pipeline-overridable constants are not supported");
        g.k = kw == "const" ? GlobalDecl::CONST : GlobalDecl::VAR;
        if (g.k == GlobalDecl::VAR && accept("<")) {
            g.space = ident();
            if (accept(",")) g.access = ident();
            expect_close();
        }
        g.name = ident();
        if (accept(":")) g.type = parse_type();
        if (accept("=")) g.init = parse_expr();
        expect(";");
        return g;
    }

    FnDecl parse_fn()
    {
        FnDecl f;
        f.loc = next().loc;
        f.name = ident();
        expect("(");
        while (!is_punct(")")) {
            Param p;
            p.attrs = parse_attrs();
            p.loc = peek().loc;
            p.name = ident();
            expect(":");
            p.type = parse_type();
            f.params.push_back(std::move(p));
            if (!accept(",")) break;
        }
        expect(")");
        if (accept("->")) {
            f.ret_attrs = parse_attrs();
            f.ret = parse_type();
        }
        f.body = parse_block();
        return f;
    }

    std::vector<StmtP> parse_block()
    {
        expect("{");
        std::vector<StmtP> out;
        while (!accept("}")) {
            if (peek().k == T_EOF) fail(peek().loc, "unterminated block");
            out.push_back(parse_stmt());
        }
        return out;
    }

    StmtP mk(Stmt::Kind k, Loc l)
    {
        auto s = std::make_shared<Stmt>();
        s->k = k;
        s->loc = l;
        return s;
    }

    // Statements allowed in for-loop headers and as simple statements.
    StmtP parse_simple(bool allow_decl)
    {
        Loc l = peek().loc;
        if (allow_decl && (is_ident("var") || is_ident("let") || is_ident("const"))) {
            std::string kw = next().s;
            auto s = mk(kw == "var" ? Stmt::VAR : kw == "let" ? Stmt::LET : Stmt::CONST, l);
            if (kw == "var" && accept("<")) {
                std::string space = ident();
                if (space != "function") fail(l, "function-scope variables live in the function address space");
                expect_close();
            }
            s->name = ident();
            if (accept(":")) s->type = parse_type();
            if (accept("=")) s->e = parse_expr();
            else if (kw != "var") fail(l, kw + " declarations need an initializer");
            return s;
        }
        if (is_ident("_") && is_punct("=", 1)) {
            next();
            next();
            auto s = mk(Stmt::PHONY, l);
            s->e = parse_expr();
            return s;
        }
        ExprP lhs = parse_expr();
        if (accept("=")) {
            auto s = mk(Stmt::ASSIGN, l);
            s->e = lhs;
            s->e2 = parse_expr();
            return s;
        }
        static const char *cops[] = {"+=", "-=", "*=", "/=", "%=", "&=", "|=", "^=", "<<=", ">>=", nullptr};
        for (int i = 0; cops[i]; ++i)
            if (accept(cops[i])) {
                auto s = mk(Stmt::COMPOUND, l);
                s->name = std::string(cops[i]).substr(0, strlen(cops[i]) - 1);
                s->e = lhs;
                s->e2 = parse_expr();
                return s;
            }
        if (accept("++")) {
            auto s = mk(Stmt::INCR, l);
            s->e = lhs;
            return s;
        }
        if (accept("--")) {
            auto s = mk(Stmt::DECR, l);
            s->e = lhs;
            return s;
        }
        if (lhs->k != Expr::CALL) fail(l, "expression statements must be function calls");
        auto s = mk(Stmt::CALL, l);
        s->e = lhs;
        return s;
    }

    StmtP parse_if()
    {
        auto s = mk(Stmt::IF, next().loc);
        s->e = parse_expr();
        s->body = parse_block();
        if (is_ident("else")) {
            next();
            if (is_ident("if")) {
                s->alt.push_back(parse_if());
            } else {
                auto b = mk(Stmt::BLOCK, peek().loc);
                b->body = parse_block();
                s->alt.push_back(b);
            }
        }
        return s;
    }

    StmtP parse_stmt()
    {
        Loc l = peek().loc;
        std::vector<Attr> attrs = parse_attrs();   // statement attributes are accepted and ignored
        (void)attrs;
        if (accept(";")) return mk(Stmt::EMPTY, l);
        if (is_punct("{")) {
            auto s = mk(Stmt::BLOCK, l);
            s->body = parse_block();
            return s;
        }
        if (is_ident("if")) return parse_if();
        if (is_ident("return")) {
            next();
            auto s = mk(Stmt::RETURN, l);
            if (!is_punct(";")) s->e = parse_expr();
            expect(";");
            return s;
        }
        if (is_ident("discard")) {
            next();
            expect(";");
            return mk(Stmt::DISCARD, l);
        }
        if (is_ident("break")) {
            next();
            if (is_ident("if")) {
                next();
                auto s = mk(Stmt::BREAK_IF, l);
                s->e = parse_expr();
                expect(";");
                return s;
            }
            expect(";");
            return mk(Stmt::BREAK, l);
        }
        if (is_ident("continue")) {
            next();
            expect(";");
            return mk(Stmt::CONTINUE, l);
        }
        if (is_ident("loop")) {
            next();
            auto s = mk(Stmt::LOOP, l);
            expect("{");
            while (!accept("}")) {
                if (is_ident("continuing")) {
                    next();
                    s->has_continuing = true;
                    s->alt = parse_block();
                    expect("}");
                    break;
                }
                if (peek().k == T_EOF) fail(peek().loc, "unterminated loop");
                s->body.push_back(parse_stmt());
            }
            return s;
        }
        if (is_ident("while")) {
            next();
            auto s = mk(Stmt::WHILE, l);
            s->e = parse_expr();
            s->body = parse_block();
            return s;
        }
        if (is_ident("for")) {
            next();
            auto s = mk(Stmt::FOR, l);
            expect("(");
            if (!is_punct(";")) s->init = parse_simple(true);
            expect(";");
            if (!is_punct(";")) s->e = parse_expr();
            expect(";");
            if (!is_punct(")")) s->update = parse_simple(false);
            expect(")");
            s->body = parse_block();
            return s;
        }
        if (is_ident("switch")) {
            next();
            auto s = mk(Stmt::SWITCH, l);
            s->e = parse_expr();
            expect("{");
            while (!accept("}")) {
                Stmt::Case c;
                c.loc = peek().loc;
                if (is_ident("default")) {
                    next();
                    c.dflt = true;
                } else {
                    if (!is_ident("case")) fail(peek().loc, "expected 'case' or 'default'");
                    next();
                    while (true) {
                        if (is_ident("default")) {
                            next();
                            c.dflt = true;
                        } else {
                            c.sel.push_back(parse_expr());
                        }
                        if (!accept(",")) break;
                        if (is_punct(":") || is_punct("{")) break;
                    }
                }
                accept(":");
                c.body = parse_block();
                s->cases.push_back(std::move(c));
            }
            return s;
        }
        StmtP s = parse_simple(true);
        expect(";");
        return s;
    }

    // ---- expressions (WGSL precedence; relational and bitwise operators
    // do not chain with each other, which the grammar enforces) ----

    ExprP mkx(Expr::Kind k, Loc l)
    {
        auto e = std::make_shared<Expr>();
        e->k = k;
        e->loc = l;
        return e;
    }
    ExprP bin(const std::string &op, ExprP a, ExprP b, Loc l)
    {
        auto e = mkx(Expr::BINARY, l);
        e->name = op;
        e->a = {a, b};
        return e;
    }

public:
    ExprP parse_expr()
    {
        Loc l = peek().loc;
        ExprP lhs = parse_relational();
        if (is_punct("&&")) {
            while (accept("&&")) lhs = bin("&&", lhs, parse_relational(), l);
            return lhs;
        }
        if (is_punct("||")) {
            while (accept("||")) lhs = bin("||", lhs, parse_relational(), l);
            return lhs;
        }
        if (is_punct("&") || is_punct("|") || is_punct("^")) {
            std::string op = peek().s;
            while (accept(op.c_str())) lhs = bin(op, lhs, parse_unary(), l);
            return lhs;
        }
        return lhs;
    }

private:
    ExprP parse_relational()
    {
        Loc l = peek().loc;
        ExprP lhs = parse_shift_expr();
        static const char *ops[] = {"==", "!=", "<=", ">=", "<", ">", nullptr};
        for (int i = 0; ops[i]; ++i)
            if (accept(ops[i])) return bin(ops[i], lhs, parse_shift_expr(), l);
        return lhs;
    }

public:
    ExprP parse_shift_expr()
    {
        Loc l = peek().loc;
        ExprP lhs = parse_additive();
        if (accept("<<")) return bin("<<", lhs, parse_unary(), l);
        if (accept(">>")) return bin(">>", lhs, parse_unary(), l);
        return lhs;
    }

private:
    ExprP parse_additive()
    {
        Loc l = peek().loc;
        ExprP lhs = parse_multiplicative();
        while (is_punct("+") || is_punct("-")) {
            std::string op = next().s;
            lhs = bin(op, lhs, parse_multiplicative(), l);
        }
        return lhs;
    }

    ExprP parse_multiplicative()
    {
        Loc l = peek().loc;
        ExprP lhs = parse_unary();
        while (is_punct("*") || is_punct("/") || is_punct("%")) {
            std::string op = next().s;
            lhs = bin(op, lhs, parse_unary(), l);
        }
        return lhs;
    }

    ExprP parse_unary()
    {
        Loc l = peek().loc;
        if (is_punct("-") || is_punct("!") || is_punct("~") || is_punct("&") || is_punct("*")) {
            std::string op = next().s;
            auto e = mkx(Expr::UNARY, l);
            e->name = op;
            e->a.push_back(parse_unary());
            return e;
        }
        return parse_postfix(parse_primary());
    }

    ExprP parse_postfix(ExprP e)
    {
        while (true) {
            Loc l = peek().loc;
            if (accept(".")) {
                auto m = mkx(Expr::MEMBER, l);
                m->name = ident();
                m->a.push_back(e);
                e = m;
            } else if (accept("[")) {
                auto m = mkx(Expr::INDEX, l);
                m->a.push_back(e);
                m->a.push_back(parse_expr());
                expect("]");
                e = m;
            } else {
                return e;
            }
        }
    }

    ExprP parse_primary()
    {
        Loc l = peek().loc;
        const Tok &t = peek();
        if (t.k == T_INT || t.k == T_FLOAT) {
            auto e = mkx(Expr::LIT, l);
            e->lit = next();
            return e;
        }
        if (accept("(")) {
            ExprP e = parse_expr();
            expect(")");
            return e;
        }
        if (t.k != T_IDENT) fail(l, "expected an expression, found '" + describe(t) + "'");
        if (t.s == "true" || t.s == "false") {
            auto e = mkx(Expr::LIT, l);
            e->lit = next();
            return e;
        }
        std::string name = next().s;
        TypeP tmpl;
        if (is_templated_type(name) && is_punct("<")) {
            // a templated type (constructor) or bitcast<T>
            --p_;
            tmpl = parse_type();
        }
        if (accept("(")) {
            auto e = mkx(Expr::CALL, l);
            e->name = name;
            e->tmpl = tmpl;
            while (!is_punct(")")) {
                e->a.push_back(parse_expr());
                if (!accept(",")) break;
            }
            expect(")");
            return e;
        }
        if (tmpl) fail(l, "a type is not an expression");
        auto e = mkx(Expr::IDENT, l);
        e->name = name;
        return e;
    }
};

// ------------------------------------------------------------------
// types
// ------------------------------------------------------------------

enum TK { TK_VOID, TK_BOOL, TK_I32, TK_U32, TK_F32, TK_AINT, TK_AFLOAT, TK_VEC, TK_MAT, TK_ARR, TK_RTARR,
          TK_STRUCT, TK_TEX, TK_SAMP };

struct TypeInfo {
    TK k = TK_VOID;
    int n = 0;      // vector size, matrix columns, array count
    int rows = 0;   // matrix rows
    int elem = -1;  // vector / matrix scalar, array element
    int sid = -1;   // struct index
    bool operator==(const TypeInfo &o) const
    {
        return k == o.k && n == o.n && rows == o.rows && elem == o.elem && sid == o.sid;
    }
};

enum : int { TY_VOID = 0, TY_BOOL, TY_I32, TY_U32, TY_F32, TY_AINT, TY_AFLOAT };

enum Builtin { BI_NONE, BI_POSITION, BI_VERTEX_INDEX, BI_INSTANCE_INDEX, BI_FRONT_FACING, BI_FRAG_DEPTH,
               BI_PRIMITIVE_INDEX, BI_SAMPLE_INDEX, BI_SAMPLE_MASK, BI_OTHER };

// Stage-interface attributes of a parameter, return value or member.
struct IoAttrs {
    int location = -1;
    Builtin builtin = BI_NONE;
    std::string builtin_name;
    uint32_t interp = IO_PERSPECTIVE;
    bool has_interp = false;
    Loc loc;
};

struct MemberInfo {
    std::string name, cname;
    int ty = 0;
    uint32_t offset = 0;    // host-shareable layout
    IoAttrs io;
};

struct StructInfo {
    std::string name, cname;
    std::vector<MemberInfo> members;
    uint32_t align = 4, size = 0;
    Loc loc;
    bool emitted = false;
};

// Constant value: scalars in `i` (bool, i32, u32, AbstractInt) or `f`
// (f32 -- a float held exactly in a double -- or AbstractFloat);
// composites in `el`.
struct Const {
    int ty = 0;
    int64_t i = 0;
    double f = 0.0;
    std::vector<Const> el;
};

// Address spaces of a reference.
enum Space { SP_NONE, SP_FUNCTION, SP_PRIVATE, SP_UNIFORM, SP_STORAGE };

struct Val {
    int ty = 0;
    bool is_const = false;
    Const c;
    std::string code;          // C++ expression of the value (non-constant values)
    // a reference (memory view) instead of a value
    bool is_ref = false;
    Space space = SP_NONE;
    bool writable = false;
    std::string lv;            // function / private: C++ lvalue
    int slot = -1;             // uniform / storage: resource slot
    std::string off;           // ... byte offset expression
    uint32_t rt_base = 0;      // ... for a runtime-sized array: its byte offset in the binding
    int handle = -1;           // texture / sampler variable: resource slot
    bool is_ptr = false;       // &variable (arrayLength's operand)
};

// ------------------------------------------------------------------
// translator
// ------------------------------------------------------------------

struct Sym {
    enum Kind { LOCAL_VAR, LOCAL_LET, CONST, PARAM, PRIV, BUF, HANDLE } k = CONST;
    int ty = 0;
    std::string cname;
    Const c;
    int slot = -1;
    Space space = SP_NONE;
    bool writable = false;
};

struct FnInfo {
    const FnDecl *decl = nullptr;
    std::string cname;
    std::vector<int> param_ty;
    int ret_ty = TY_VOID;
    bool is_entry = false;
    Stage stage = Stage::Vertex;
    std::set<int> calls;           // function indices
    std::set<int> slots;           // resource slots used directly
    bool discards = false;
    uint32_t early_depth = 0;      // fragment entry points: @early_depth_test (EntryPoint::early_depth)
    std::string code;
    int state = 0;                 // call-graph walk: 0 new, 1 visiting, 2 done
};

struct Breakable {
    bool is_loop = false;
    std::string cont_label;        // loops: where 'continue' goes
    bool cont_used = false;
};

struct FnCtx {
    FnInfo *fn = nullptr;
    std::string out;
    int indent = 1;
    int tmp = 0;
    std::vector<std::map<std::string, Sym>> scopes;
    std::vector<Breakable> breakables;
    bool in_continuing = false;
};

class Translator {
public:
    explicit Translator(const ModuleAst &m) : m_(m)
    {
        TypeInfo t;
        for (TK k : {TK_VOID, TK_BOOL, TK_I32, TK_U32, TK_F32, TK_AINT, TK_AFLOAT}) {
            t.k = k;
            types_.push_back(t);
        }
    }

    void run(Module &out);

private:
    const ModuleAst &m_;
    std::vector<TypeInfo> types_;
    std::vector<StructInfo> structs_;
    std::map<std::string, int> struct_by_name_, alias_;
    std::map<std::string, Sym> globals_;
    std::map<std::string, int> fn_by_name_;
    std::vector<FnInfo> fns_;
    std::vector<Resource> resources_;
    std::vector<std::pair<std::string, std::string>> privs_;   // cname, C++ initializer
    std::string priv_types_;
    std::map<int, std::string> loader_code_;
    std::vector<int> loader_order_;
    std::set<std::string> global_names_;

    // ---- type table ----
    int intern(const TypeInfo &t)
    {
        for (size_t i = 0; i < types_.size(); ++i)
            if (types_[i] == t) return (int)i;
        types_.push_back(t);
        return (int)types_.size() - 1;
    }
    const TypeInfo &ti(int ty) const { return types_[ty]; }
    int vec_of(int scalar, int n)
    {
        TypeInfo t;
        t.k = TK_VEC;
        t.n = n;
        t.elem = scalar;
        return intern(t);
    }
    int mat_of(int c, int r, int scalar = TY_F32)
    {
        TypeInfo t;
        t.k = TK_MAT;
        t.n = c;
        t.rows = r;
        t.elem = scalar;
        return intern(t);
    }
    int arr_of(int elem, int n)
    {
        TypeInfo t;
        t.k = n > 0 ? TK_ARR : TK_RTARR;
        t.n = n;
        t.elem = elem;
        return intern(t);
    }
    bool is_scalar(int ty) const { return ty >= TY_BOOL && ty <= TY_AFLOAT; }
    bool is_abstract_scalar(int ty) const { return ty == TY_AINT || ty == TY_AFLOAT; }
    bool is_abstract(int ty) const
    {
        const TypeInfo &t = ti(ty);
        if (t.k == TK_AINT || t.k == TK_AFLOAT) return true;
        if (t.k == TK_VEC || t.k == TK_MAT || t.k == TK_ARR) return is_abstract(t.elem);
        return false;
    }
    bool is_int_scalar(int ty) const { return ty == TY_I32 || ty == TY_U32 || ty == TY_AINT; }
    bool is_float_scalar(int ty) const { return ty == TY_F32 || ty == TY_AFLOAT; }
    bool is_numeric_scalar(int ty) const { return is_int_scalar(ty) || is_float_scalar(ty); }
    // scalar type of a scalar / vector / matrix; -1 otherwise
    int scalar_of(int ty) const
    {
        if (is_scalar(ty)) return ty;
        const TypeInfo &t = ti(ty);
        if (t.k == TK_VEC || t.k == TK_MAT) return t.elem;
        return -1;
    }
    int with_scalar(int ty, int scalar)
    {
        if (is_scalar(ty)) return scalar;
        const TypeInfo &t = ti(ty);
        if (t.k == TK_VEC) return vec_of(scalar, t.n);
        if (t.k == TK_MAT) return mat_of(t.n, t.rows, scalar);
        if (t.k == TK_ARR) return arr_of(with_scalar(t.elem, scalar), t.n);
        return ty;
    }
    int concretize(int ty)
    {
        if (ty == TY_AINT) return TY_I32;
        if (ty == TY_AFLOAT) return TY_F32;
        const TypeInfo &t = ti(ty);
        if (t.k == TK_VEC) return vec_of(concretize(t.elem), t.n);
        if (t.k == TK_MAT) return mat_of(t.n, t.rows, TY_F32);
        if (t.k == TK_ARR) return arr_of(concretize(t.elem), t.n);
        return ty;
    }
    bool constructible(int ty) const
    {
        const TypeInfo &t = ti(ty);
        if (t.k == TK_RTARR || t.k == TK_TEX || t.k == TK_SAMP || t.k == TK_VOID) return false;
        if (t.k == TK_ARR) return constructible(t.elem);
        if (t.k == TK_STRUCT) {
            for (auto &mb : structs_[t.sid].members)
                if (!constructible(mb.ty)) return false;
        }
        return true;
    }
    int lanes(int ty) const
    {
        const TypeInfo &t = ti(ty);
        if (t.k == TK_VEC) return t.n;
        if (t.k == TK_MAT) return t.n * t.rows;
        return 1;
    }

    std::string tname(int ty) const
    {
        const TypeInfo &t = ti(ty);
        switch (t.k) {
        case TK_VOID: return "void";
        case TK_BOOL: return "bool";
        case TK_I32: return "i32";
        case TK_U32: return "u32";
        case TK_F32: return "f32";
        case TK_AINT: return "abstract-int";
        case TK_AFLOAT: return "abstract-float";
        case TK_VEC: return "vec" + std::to_string(t.n) + "<" + tname(t.elem) + ">";
        case TK_MAT: return "mat" + std::to_string(t.n) + "x" + std::to_string(t.rows) + "<" + tname(t.elem) + ">";
        case TK_ARR: return "array<" + tname(t.elem) + ", " + std::to_string(t.n) + ">";
        case TK_RTARR: return "array<" + tname(t.elem) + ">";
        case TK_STRUCT: return structs_[t.sid].name;
        case TK_TEX: return "texture_2d<f32>";
        case TK_SAMP: return "sampler";
        }
        return "?";
    }

    std::string ctype(int ty) const
    {
        const TypeInfo &t = ti(ty);
        switch (t.k) {
        case TK_BOOL: return "bool";
        case TK_I32: return "int32_t";
        case TK_U32: return "uint32_t";
        case TK_F32: return "float";
        case TK_VEC: return "wgsl::vec<" + ctype(t.elem) + ", " + std::to_string(t.n) + ">";
        case TK_MAT: return "wgsl::mat<" + std::to_string(t.n) + ", " + std::to_string(t.rows) + ">";
        case TK_ARR: return "wgsl::arr<" + ctype(t.elem) + ", " + std::to_string(t.n) + ">";
        case TK_STRUCT: return structs_[t.sid].cname;
        default: throw Error("internal: no C++ type for " + tname(ty));
        }
    }

    // WGSL host-shareable layout (alignment and size of each type).
    uint32_t align_of(int ty) const
    {
        const TypeInfo &t = ti(ty);
        switch (t.k) {
        case TK_VEC: return t.n == 2 ? 8u : 16u;
        case TK_MAT: return t.rows == 2 ? 8u : 16u;
        case TK_ARR:
        case TK_RTARR: return align_of(t.elem);
        case TK_STRUCT: return structs_[t.sid].align;
        default: return 4u;
        }
    }
    uint32_t size_of(int ty) const
    {
        const TypeInfo &t = ti(ty);
        switch (t.k) {
        case TK_VEC: return 4u * (uint32_t)t.n;
        case TK_MAT: return (uint32_t)t.n * column_stride(ty);
        case TK_ARR: return (uint32_t)t.n * stride_of(ty);
        case TK_RTARR: return stride_of(ty);
        case TK_STRUCT: return structs_[t.sid].size;
        default: return 4u;
        }
    }
    static uint32_t round_up(uint32_t a, uint32_t v) { return (v + a - 1u) / a * a; }
    uint32_t column_stride(int mat) const { return ti(mat).rows == 2 ? 8u : 16u; }
    uint32_t stride_of(int arr) const
    {
        int e = ti(arr).elem;
        return round_up(align_of(e), size_of(e));
    }

    // ---- declarations ----
    int resolve_type(const TypeP &t);
    void declare_structs();
    void declare_globals();
    void declare_functions();
    IoAttrs io_attrs(const std::vector<Attr> &attrs);
    int64_t attr_int(const Attr &a, size_t k = 0);
    std::string attr_ident(const Attr &a, size_t k = 0);

    // ---- constants ----
    Const eval_const(const ExprP &e);
    Const zero_const(int ty);
    Const convert_const(const Const &c, int ty, Loc l);
    std::string literal(const Const &c);
    static std::string f32_literal(float f);

    // ---- expressions ----
    Val expr(FnCtx *cx, const ExprP &e);
    Val lit(const Tok &t);
    Val ident(FnCtx *cx, const Expr &e);
    Val unary(FnCtx *cx, const Expr &e);
    Val binary(FnCtx *cx, const Expr &e);
    Val logical(FnCtx *cx, const Expr &e);
    Val call(FnCtx *cx, const Expr &e);
    Val construct(FnCtx *cx, int ty, const std::vector<Val> &args, Loc l);
    Val member(FnCtx *cx, const Val &base, const std::string &name, Loc l);
    Val index(FnCtx *cx, const Val &base, const Val &idx, Loc l);
    Val builtin_call(FnCtx *cx, const Expr &e, std::vector<Val> &args);
    Val user_call(FnCtx *cx, int fi, const Expr &e, std::vector<Val> &args);

    Val load(FnCtx *cx, const Val &v, Loc l);
    Val rvalue(FnCtx *cx, const Val &v, Loc l) { return v.is_ref ? load(cx, v, l) : v; }
    Val convert(FnCtx *cx, const Val &v, int ty, Loc l);
    std::string code(FnCtx *cx, const Val &v);
    Val cvalue(const Const &c)
    {
        Val v;
        v.ty = c.ty;
        v.is_const = true;
        v.c = c;
        return v;
    }
    Val tmpval(FnCtx *cx, int ty, const std::string &init);
    std::string new_tmp(FnCtx *cx) { return "_t" + std::to_string(cx->tmp++); }
    void line(FnCtx *cx, const std::string &s) { cx->out += std::string(cx->indent * 4, ' ') + s + "\n"; }
    void check_abort(FnCtx *cx, const std::string &call)
    {
        line(cx, "if (uint32_t _rc = " + call + ") return _rc;");
    }
    std::string loader(int ty);
    Val arith(FnCtx *cx, const std::string &op, Val a, Val b, Loc l);
    std::string scalar_op(FnCtx *cx, const std::string &op, int sty, const std::string &a, const std::string &b);
    Const fold_scalar(const std::string &op, int sty, const Const &a, const Const &b, Loc l);
    Val compare(FnCtx *cx, const std::string &op, Val a, Val b, Loc l);
    Val convert_value(FnCtx *cx, const Val &v, int to, Loc l);
    Const fold_convert(const Const &c, int to, Loc l);
    std::string lane(const std::string &code, int ty, int k) const
    {
        const TypeInfo &t = ti(ty);
        if (t.k == TK_VEC) return code + ".v[" + std::to_string(k) + "]";
        if (t.k == TK_MAT)
            return code + ".c[" + std::to_string(k / t.rows) + "].v[" + std::to_string(k % t.rows) + "]";
        return code;
    }
    Const *const_lane(Const &c, int k) const
    {
        const TypeInfo &t = ti(c.ty);
        if (t.k == TK_VEC) return &c.el[k];
        if (t.k == TK_MAT) return &c.el[k / t.rows].el[k % t.rows];
        return &c;
    }

    // ---- statements ----
    void function(FnInfo &f);
    void block(FnCtx *cx, const std::vector<StmtP> &body, bool new_scope = true);
    void stmt(FnCtx *cx, const Stmt &s);
    void store(FnCtx *cx, const Val &ref, const Val &v, Loc l);
    void declare_local(FnCtx *cx, const Stmt &s);
    const Sym *lookup(FnCtx *cx, const std::string &name);
    void entry_wrapper(FnInfo &f, EntryPoint &ep, std::string &out);
    void finish_reachability(EntryPoint &ep, int fi);
};

// ------------------------------------------------------------------
// declarations
// ------------------------------------------------------------------

int Translator::resolve_type(const TypeP &t)
{
    const std::string &n = t->name;
    auto need_params = [&](size_t k) {
        if (t->params.size() != k) fail(t->loc, "type '" + n + "' needs " + std::to_string(k) + " template argument(s)");
    };
    if (n == "bool") return TY_BOOL;
    if (n == "i32") return TY_I32;
    if (n == "u32") return TY_U32;
    if (n == "f32") return TY_F32;
    if (n == "f16") fail(t->loc, "f16 is not supported (the reference has no f16 path)");
    static const std::map<std::string, std::pair<int, int>> vec_alias = {
        {"vec2f", {TY_F32, 2}}, {"vec3f", {TY_F32, 3}}, {"vec4f", {TY_F32, 4}},
        {"vec2i", {TY_I32, 2}}, {"vec3i", {TY_I32, 3}}, {"vec4i", {TY_I32, 4}},
        {"vec2u", {TY_U32, 2}}, {"vec3u", {TY_U32, 3}}, {"vec4u", {TY_U32, 4}}};
    auto va = vec_alias.find(n);
    if (va != vec_alias.end()) return vec_of(va->second.first, va->second.second);
    if (n == "vec2" || n == "vec3" || n == "vec4") {
        need_params(1);
        int e = resolve_type(t->params[0]);
        if (!is_scalar(e) || is_abstract_scalar(e)) fail(t->loc, "vector elements must be scalars");
        return vec_of(e, n[3] - '0');
    }
    if (n.size() >= 6 && n.compare(0, 3, "mat") == 0 && n[4] == 'x' && n[3] >= '2' && n[3] <= '4' && n[5] >= '2' &&
        n[5] <= '4') {
        int c = n[3] - '0', r = n[5] - '0';
        if (n.size() == 7 && n[6] == 'f') return mat_of(c, r);
        if (n.size() == 6) {
            need_params(1);
            if (resolve_type(t->params[0]) != TY_F32) fail(t->loc, "matrices must have f32 elements");
            return mat_of(c, r);
        }
    }
    if (n == "array") {
        need_params(1);
        int e = resolve_type(t->params[0]);
        if (!constructible(e) && ti(e).k != TK_STRUCT) fail(t->loc, "invalid array element type");
        if (!t->count) return arr_of(e, 0);
        Const c = eval_const(t->count);
        if (!is_int_scalar(c.ty) || c.i <= 0 || c.i > 65536) fail(t->loc, "array element count must be a positive integer constant");
        return arr_of(e, (int)c.i);
    }
    if (n == "texture_2d") {
        need_params(1);
        if (resolve_type(t->params[0]) != TY_F32) fail(t->loc, "only texture_2d<f32> is supported");
        TypeInfo x;
        x.k = TK_TEX;
        return intern(x);
    }
    if (n == "sampler") {
        TypeInfo x;
        x.k = TK_SAMP;
        return intern(x);
    }
    if (n == "sampler_comparison" || n.compare(0, 8, "texture_") == 0)
        fail(t->loc, "'" + n + "' is not supported (only texture_2d<f32> and sampler)");
    if (n == "ptr") fail(t->loc, "pointer types are not supported");
    if (n == "atomic") fail(t->loc, "atomics are todo!() in the reference (statement/atomic.rs)");
    auto s = struct_by_name_.find(n);
    if (s != struct_by_name_.end()) {
        TypeInfo x;
        x.k = TK_STRUCT;
        x.sid = s->second;
        return intern(x);
    }
    auto a = alias_.find(n);
    if (a != alias_.end()) return a->second;
    fail(t->loc, "unknown type '" + n + "'");
}

int64_t Translator::attr_int(const Attr &a, size_t k)
{
    if (a.args.size() <= k) fail(a.loc, "@" + a.name + " needs an argument");
    Const c = eval_const(a.args[k]);
    if (!is_int_scalar(c.ty) || c.i < 0) fail(a.loc, "@" + a.name + " needs a non-negative integer");
    return c.i;
}

std::string Translator::attr_ident(const Attr &a, size_t k)
{
    if (a.args.size() <= k || a.args[k]->k != Expr::IDENT) fail(a.loc, "@" + a.name + " needs a name argument");
    return a.args[k]->name;
}

IoAttrs Translator::io_attrs(const std::vector<Attr> &attrs)
{
    IoAttrs io;
    for (const Attr &a : attrs) {
        io.loc = a.loc;
        if (a.name == "location") {
            io.location = (int)attr_int(a);
        } else if (a.name == "builtin") {
            io.builtin_name = attr_ident(a);
            const std::string &b = io.builtin_name;
            io.builtin = b == "position" ? BI_POSITION
                       : b == "vertex_index" ? BI_VERTEX_INDEX
                       : b == "instance_index" ? BI_INSTANCE_INDEX
                       : b == "front_facing" ? BI_FRONT_FACING
                       : b == "frag_depth" ? BI_FRAG_DEPTH
                       : b == "primitive_index" ? BI_PRIMITIVE_INDEX
                       : b == "sample_index" ? BI_SAMPLE_INDEX
                       : b == "sample_mask" ? BI_SAMPLE_MASK : BI_OTHER;
        } else if (a.name == "interpolate") {
            std::string m = attr_ident(a);
            io.has_interp = true;
            if (m == "flat") io.interp = IO_FLAT;
            else if (m == "linear") io.interp = IO_LINEAR;
            else if (m == "perspective") io.interp = IO_PERSPECTIVE;
            else fail(a.loc, "unknown interpolation '" + m + "'");
            if (a.args.size() > 1) {
                std::string smp = attr_ident(a, 1);
                if (smp != "center" && smp != "centroid" && smp != "sample" && smp != "first" && smp != "either")
                    fail(a.loc, "unknown interpolation sampling '" + smp + "'");
            }
        } else if (a.name == "invariant" || a.name == "align" || a.name == "size") {
            // layout attributes are handled by the struct declaration
        } else if (a.name == "blend_src") {
            fail(a.loc, "dual-source blending is not supported");
        } else {
            fail(a.loc, "unknown attribute @" + a.name);
        }
    }
    return io;
}

void Translator::declare_structs()
{
    for (size_t i = 0; i < m_.structs.size(); ++i) {
        const StructDecl &d = m_.structs[i];
        if (struct_by_name_.count(d.name) || global_names_.count(d.name)) fail(d.loc, "redefinition of '" + d.name + "'");
        global_names_.insert(d.name);
        StructInfo s;
        s.name = d.name;
        s.cname = "S_" + d.name;
        s.loc = d.loc;
        struct_by_name_[d.name] = (int)i;
        structs_.push_back(s);
    }
    for (const GlobalDecl &g : m_.globals)
        if (g.k == GlobalDecl::ALIAS) {
            if (global_names_.count(g.name)) fail(g.loc, "redefinition of '" + g.name + "'");
            global_names_.insert(g.name);
        }
    // members and layouts, in dependency order (a struct's layout needs
    // those of its member types)
    std::vector<int> state(structs_.size(), 0);
    std::function<void(int)> lay = [&](int si) {
        if (state[si] == 2) return;
        if (state[si] == 1) fail(structs_[si].loc, "recursive structure '" + structs_[si].name + "'");
        state[si] = 1;
        const StructDecl &d = m_.structs[si];
        for (const MemberDecl &md : d.members) {
            // visit member struct types first
            std::function<void(const TypeP &)> visit = [&](const TypeP &t) {
                auto it = struct_by_name_.find(t->name);
                if (it != struct_by_name_.end()) lay(it->second);
                for (auto &p : t->params) visit(p);
            };
            visit(md.type);
        }
        StructInfo &s = structs_[si];
        uint32_t off = 0, al = 4;
        std::set<std::string> names;
        for (size_t k = 0; k < d.members.size(); ++k) {
            const MemberDecl &md = d.members[k];
            if (!names.insert(md.name).second) fail(md.loc, "duplicate member '" + md.name + "'");
            MemberInfo mi;
            mi.name = md.name;
            mi.cname = "m_" + md.name;
            mi.ty = resolve_type(md.type);
            if (ti(mi.ty).k == TK_TEX || ti(mi.ty).k == TK_SAMP) fail(md.loc, "handles cannot be structure members");
            if (ti(mi.ty).k == TK_RTARR && k + 1 != d.members.size())
                fail(md.loc, "a runtime-sized array must be the last member");
            mi.io = io_attrs(md.attrs);
            uint32_t a = align_of(mi.ty), sz = size_of(mi.ty);
            for (const Attr &at : md.attrs) {
                if (at.name == "align") a = (uint32_t)attr_int(at);
                if (at.name == "size") sz = (uint32_t)attr_int(at);
            }
            if (a == 0 || (a & (a - 1))) fail(md.loc, "@align must be a power of two");
            off = round_up(a, off);
            mi.offset = off;
            off += sz;
            al = std::max(al, a);
            s.members.push_back(mi);
        }
        s.align = al;
        s.size = round_up(al, off);
        state[si] = 2;
    };
    for (size_t i = 0; i < structs_.size(); ++i) lay((int)i);
    for (const GlobalDecl &g : m_.globals)
        if (g.k == GlobalDecl::ALIAS) alias_[g.name] = resolve_type(g.type);
}

void Translator::declare_globals()
{
    for (const GlobalDecl &g : m_.globals) {
        if (g.k == GlobalDecl::ALIAS) continue;
        if (global_names_.count(g.name)) fail(g.loc, "redefinition of '" + g.name + "'");
        global_names_.insert(g.name);
        Sym s;
        if (g.k == GlobalDecl::CONST) {
            if (!g.init) fail(g.loc, "const declarations need an initializer");
            s.k = Sym::CONST;
            s.c = eval_const(g.init);
            if (g.type) s.c = convert_const(s.c, resolve_type(g.type), g.loc);
            s.ty = s.c.ty;
            globals_[g.name] = s;
            continue;
        }
        // var
        int ty = g.type ? resolve_type(g.type) : -1;
        std::string space = g.space;
        if (space.empty()) {
            if (ty >= 0 && (ti(ty).k == TK_TEX || ti(ty).k == TK_SAMP)) space = "handle";
            else fail(g.loc, "module-scope variables need an address space");
        }
        if (space == "private") {
            Const init;
            if (g.init) {
                init = eval_const(g.init);
                if (ty < 0) ty = concretize(init.ty);
                init = convert_const(init, ty, g.loc);
            } else {
                if (ty < 0) fail(g.loc, "variable needs a type or an initializer");
                init = zero_const(ty);
            }
            if (!constructible(ty)) fail(g.loc, "private variables need a constructible type");
            s.k = Sym::PRIV;
            s.ty = ty;
            s.cname = "pv.g_" + g.name;
            s.space = SP_PRIVATE;
            s.writable = true;
            privs_.push_back({"g_" + g.name, literal(init)});
            priv_types_ += "    " + ctype(ty) + " g_" + g.name + ";\n";
            globals_[g.name] = s;
            continue;
        }
        if (space == "workgroup") fail(g.loc, "workgroup variables exist only in compute shaders");
        if (space == "push_constant" || space == "immediate") fail(g.loc, "immediate data is not supported");
        int group = -1, binding = -1;
        for (const Attr &a : g.attrs) {
            if (a.name == "group") group = (int)attr_int(a);
            else if (a.name == "binding") binding = (int)attr_int(a);
            else fail(a.loc, "unexpected attribute @" + a.name + " on a resource variable");
        }
        if (group < 0 || binding < 0) fail(g.loc, "resource variables need @group and @binding");
        if (group >= 4) fail(g.loc, "@group must be below 4");
        if (g.init) fail(g.loc, "resource variables cannot have initializers");
        if (ty < 0) fail(g.loc, "resource variables need a type");
        Resource r;
        r.group = (uint32_t)group;
        r.binding = (uint32_t)binding;
        r.name = g.name;
        for (const Resource &o : resources_)
            if (o.group == r.group && o.binding == r.binding)
                fail(g.loc, "@group(" + std::to_string(group) + ") @binding(" + std::to_string(binding) + ") is used twice");
        s.slot = (int)resources_.size();
        s.ty = ty;
        if (space == "handle") {
            r.kind = ti(ty).k == TK_TEX ? ResKind::Texture : ResKind::Sampler;
            s.k = Sym::HANDLE;
        } else if (space == "uniform" || space == "storage") {
            if (space == "storage" && !g.access.empty() && g.access != "read")
                fail(g.loc, "read_write storage buffers are not supported: stores from parallel invocations have no defined order");
            if (space == "uniform" && ti(ty).k == TK_RTARR) fail(g.loc, "uniform buffers cannot hold runtime-sized arrays");
            if (!constructible(ty) && ti(ty).k != TK_RTARR && ti(ty).k != TK_STRUCT)
                fail(g.loc, "invalid buffer type " + tname(ty));
            r.kind = space == "uniform" ? ResKind::Uniform : ResKind::Storage;
            r.min_size = size_of(ty);
            s.k = Sym::BUF;
            s.space = space == "uniform" ? SP_UNIFORM : SP_STORAGE;
        } else {
            fail(g.loc, "unknown address space '" + space + "'");
        }
        resources_.push_back(r);
        globals_[g.name] = s;
    }
    for (auto &a : m_.asserts) {
        Const c = eval_const(a.second);
        if (c.ty != TY_BOOL) fail(a.first, "const_assert needs a bool");
        if (!c.i) fail(a.first, "const_assert failed");
    }
}

void Translator::declare_functions()
{
    for (size_t i = 0; i < m_.fns.size(); ++i) {
        const FnDecl &d = m_.fns[i];
        if (global_names_.count(d.name)) fail(d.loc, "redefinition of '" + d.name + "'");
        global_names_.insert(d.name);
        FnInfo f;
        f.decl = &d;
        f.cname = "f_" + d.name;
        for (const Attr &a : d.attrs) {
            if (a.name == "vertex") {
                f.is_entry = true;
                f.stage = Stage::Vertex;
            } else if (a.name == "fragment") {
                f.is_entry = true;
                f.stage = Stage::Fragment;
            } else if (a.name == "compute") {
                fail(a.loc, "compute entry points are not part of a render pipeline");
            } else if (a.name == "early_depth_test") {
                // naga's EarlyDepthTest: `force` tests (and writes) depth
                // before the shader only; a conservative-depth argument
                // allows the early test and keeps the late one
                // (fragment.rs:166-194)
                std::string m = attr_ident(a);
                if (m == "force") f.early_depth = 1;
                else if (m == "greater_equal" || m == "less_equal" || m == "unchanged") f.early_depth = 2;
                else fail(a.loc, "unknown @early_depth_test argument '" + m + "'");
            } else if (a.name == "must_use" || a.name == "diagnostic") {
            } else {
                fail(a.loc, "unknown function attribute @" + a.name);
            }
        }
        for (const Param &p : d.params) {
            int ty = resolve_type(p.type);
            if (!constructible(ty)) fail(p.loc, "parameters must have constructible types (pointers are not supported)");
            f.param_ty.push_back(ty);
        }
        f.ret_ty = d.ret ? resolve_type(d.ret) : TY_VOID;
        if (f.ret_ty != TY_VOID && !constructible(f.ret_ty)) fail(d.loc, "return type must be constructible");
        fn_by_name_[d.name] = (int)i;
        fns_.push_back(std::move(f));
    }
}

// ------------------------------------------------------------------
// constants
// ------------------------------------------------------------------

Const Translator::eval_const(const ExprP &e)
{
    Val v = expr(nullptr, e);
    if (!v.is_const) fail(e->loc, "not a constant expression");
    return v.c;
}

Const Translator::zero_const(int ty)
{
    Const c;
    c.ty = ty;
    const TypeInfo &t = ti(ty);
    switch (t.k) {
    case TK_VEC:
        for (int k = 0; k < t.n; ++k) c.el.push_back(zero_const(t.elem));
        break;
    case TK_MAT:
        for (int k = 0; k < t.n; ++k) c.el.push_back(zero_const(vec_of(t.elem, t.rows)));
        break;
    case TK_ARR:
        for (int k = 0; k < t.n; ++k) c.el.push_back(zero_const(t.elem));
        break;
    case TK_STRUCT:
        for (auto &mb : structs_[t.sid].members) c.el.push_back(zero_const(mb.ty));
        break;
    default:
        break;
    }
    return c;
}

// Implicit conversion of an abstract constant to `ty` (naga's
// concretization; WGSL 'conversion rank').
Const Translator::convert_const(const Const &c, int ty, Loc l)
{
    if (c.ty == ty) return c;
    if (is_scalar(c.ty) && is_scalar(ty)) {
        Const r;
        r.ty = ty;
        if (c.ty == TY_AINT) {
            if (ty == TY_I32) {
                if (c.i < INT32_MIN || c.i > INT32_MAX) fail(l, "value " + std::to_string(c.i) + " does not fit in i32");
                r.i = c.i;
                return r;
            }
            if (ty == TY_U32) {
                if (c.i < 0 || c.i > (int64_t)UINT32_MAX) fail(l, "value " + std::to_string(c.i) + " does not fit in u32");
                r.i = c.i;
                return r;
            }
            if (ty == TY_F32 || ty == TY_AFLOAT) {
                double d = (double)c.i;
                r.f = ty == TY_F32 ? (double)(float)d : d;
                return r;
            }
        }
        if (c.ty == TY_AFLOAT && ty == TY_F32) {
            float f = (float)c.f;
            if (std::isinf(f)) fail(l, "value does not fit in f32");
            r.f = (double)f;
            return r;
        }
        fail(l, "cannot convert " + tname(c.ty) + " to " + tname(ty) + " implicitly");
    }
    const TypeInfo &a = ti(c.ty), &b = ti(ty);
    if (a.k == b.k && (a.k == TK_VEC || a.k == TK_MAT || a.k == TK_ARR) && a.n == b.n && a.rows == b.rows) {
        Const r;
        r.ty = ty;
        int sub = a.k == TK_MAT ? vec_of(b.elem, b.rows) : b.elem;
        for (auto &e : c.el) r.el.push_back(convert_const(e, sub, l));
        return r;
    }
    fail(l, "cannot convert " + tname(c.ty) + " to " + tname(ty) + " implicitly");
}

std::string Translator::f32_literal(float f)
{
    char buf[64];
    if (f == 0.0f) return std::signbit(f) ? "(-0.0f)" : "0.0f";
    snprintf(buf, sizeof buf, "%af", (double)f);
    return f < 0 ? std::string("(") + buf + ")" : std::string(buf);
}

std::string Translator::literal(const Const &c)
{
    const TypeInfo &t = ti(c.ty);
    switch (t.k) {
    case TK_BOOL: return c.i ? "true" : "false";
    case TK_I32:
        if (c.i >= 0) return std::to_string(c.i);
        return "(int32_t)" + std::to_string((uint32_t)(int32_t)c.i) + "u";
    case TK_U32: return std::to_string((uint32_t)c.i) + "u";
    case TK_F32: return f32_literal((float)c.f);
    case TK_AINT: return literal(convert_const(c, TY_I32, Loc{}));
    case TK_AFLOAT: return literal(convert_const(c, TY_F32, Loc{}));
    case TK_VEC:
    case TK_MAT:
    case TK_ARR: {
        if (is_abstract(c.ty)) return literal(convert_const(c, concretize(c.ty), Loc{}));
        std::string s = ctype(c.ty) + "{{";
        for (size_t k = 0; k < c.el.size(); ++k) s += (k ? ", " : "") + literal(c.el[k]);
        return s + "}}";
    }
    case TK_STRUCT: {
        std::string s = ctype(c.ty) + "{";
        for (size_t k = 0; k < c.el.size(); ++k) s += (k ? ", " : "") + literal(c.el[k]);
        return s + "}";
    }
    default: throw Error("internal: literal of " + tname(c.ty));
    }
}

// ------------------------------------------------------------------
// expressions
// ------------------------------------------------------------------

Val Translator::tmpval(FnCtx *cx, int ty, const std::string &init)
{
    Val v;
    v.ty = ty;
    v.code = new_tmp(cx);
    line(cx, "const " + ctype(ty) + " " + v.code + " = " + init + ";");
    return v;
}

std::string Translator::code(FnCtx *cx, const Val &v)
{
    if (v.is_ref) return code(cx, load(cx, v, Loc{}));
    if (v.is_const) return literal(v.c);
    if (v.ty == TY_VOID) throw Error("internal: value of a void call used");
    return v.code;
}

// Loader of a value of type `ty` from host-shareable memory.
std::string Translator::loader(int ty)
{
    const TypeInfo &t = ti(ty);
    switch (t.k) {
    case TK_BOOL: return "wgsl::ld_bool";
    case TK_I32: return "wgsl::ld_i32";
    case TK_U32: return "wgsl::ld_u32";
    case TK_F32: return "wgsl::ld_f32";
    default: break;
    }
    std::string name = "ld_" + std::to_string(ty);
    if (loader_code_.count(ty)) return name;
    loader_code_[ty] = "";   // placeholder against recursion
    std::string body = "WGSL_FN " + ctype(ty) + " " + name + "(const uint8_t *p)\n{\n    " + ctype(ty) + " r;\n";
    if (t.k == TK_VEC) {
        std::string l = loader(t.elem);
        for (int k = 0; k < t.n; ++k) body += "    r.v[" + std::to_string(k) + "] = " + l + "(p + " + std::to_string(4 * k) + ");\n";
    } else if (t.k == TK_MAT) {
        std::string l = loader(vec_of(t.elem, t.rows));
        for (int k = 0; k < t.n; ++k)
            body += "    r.c[" + std::to_string(k) + "] = " + l + "(p + " + std::to_string(k * column_stride(ty)) + ");\n";
    } else if (t.k == TK_ARR) {
        std::string l = loader(t.elem);
        body += "    for (int k = 0; k < " + std::to_string(t.n) + "; ++k) r.v[k] = " + l + "(p + (uint64_t)k * " +
                std::to_string(stride_of(ty)) + "u);\n";
    } else if (t.k == TK_STRUCT) {
        for (auto &mb : structs_[t.sid].members) {
            if (ti(mb.ty).k == TK_RTARR) throw Error("internal: loading a runtime-sized array");
            body += "    r." + mb.cname + " = " + loader(mb.ty) + "(p + " + std::to_string(mb.offset) + ");\n";
        }
    } else {
        throw Error("internal: loader of " + tname(ty));
    }
    body += "    return r;\n}\n";
    loader_code_[ty] = body;
    loader_order_.push_back(ty);
    return name;
}

Val Translator::load(FnCtx *cx, const Val &v, Loc l)
{
    if (!v.is_ref) return v;
    if (!cx) fail(l, "not a constant expression");
    if (v.space == SP_FUNCTION || v.space == SP_PRIVATE) return tmpval(cx, v.ty, v.lv);
    if (ti(v.ty).k == TK_RTARR) fail(l, "a runtime-sized array cannot be loaded as a whole");
    return tmpval(cx, v.ty, loader(v.ty) + "(env.buffer(" + std::to_string(v.slot) + ") + (" + v.off + "))");
}

// Implicit conversion (abstract -> concrete) to `ty`, or a type error.
Val Translator::convert(FnCtx *cx, const Val &v0, int ty, Loc l)
{
    Val v = rvalue(cx, v0, l);
    if (v.ty == ty) return v;
    if (v.is_const && is_abstract(v.ty)) return cvalue(convert_const(v.c, ty, l));
    fail(l, "expected " + tname(ty) + ", found " + tname(v.ty));
}

Val Translator::lit(const Tok &t)
{
    Const c;
    if (t.k == T_IDENT) {
        c.ty = TY_BOOL;
        c.i = t.s == "true";
    } else if (t.k == T_INT) {
        c.i = t.ival;
        if (t.suffix == 'i') {
            if (t.ival > INT32_MAX) fail(t.loc, "literal out of range for i32");
            c.ty = TY_I32;
        } else if (t.suffix == 'u') {
            if (t.ival > (int64_t)UINT32_MAX) fail(t.loc, "literal out of range for u32");
            c.ty = TY_U32;
        } else {
            c.ty = TY_AINT;
        }
    } else {
        if (t.suffix == 'f') {
            float f = (float)t.fval;
            if (std::isinf(f)) fail(t.loc, "literal out of range for f32");
            c.ty = TY_F32;
            c.f = (double)f;
        } else {
            if (std::isinf(t.fval)) fail(t.loc, "float literal out of range");
            c.ty = TY_AFLOAT;
            c.f = t.fval;
        }
    }
    return cvalue(c);
}

const Sym *Translator::lookup(FnCtx *cx, const std::string &name)
{
    if (cx)
        for (auto it = cx->scopes.rbegin(); it != cx->scopes.rend(); ++it) {
            auto f = it->find(name);
            if (f != it->end()) return &f->second;
        }
    auto g = globals_.find(name);
    return g == globals_.end() ? nullptr : &g->second;
}

Val Translator::ident(FnCtx *cx, const Expr &e)
{
    const Sym *s = lookup(cx, e.name);
    if (!s) {
        if (struct_by_name_.count(e.name) || alias_.count(e.name) || fn_by_name_.count(e.name))
            fail(e.loc, "'" + e.name + "' is not a value");
        fail(e.loc, "unknown identifier '" + e.name + "'");
    }
    if (s->k == Sym::CONST) return cvalue(s->c);
    if (!cx) fail(e.loc, "'" + e.name + "' is not a constant");
    Val v;
    v.ty = s->ty;
    switch (s->k) {
    case Sym::LOCAL_VAR:
        v.is_ref = true;
        v.space = SP_FUNCTION;
        v.writable = true;
        v.lv = s->cname;
        break;
    case Sym::PRIV:
        v.is_ref = true;
        v.space = SP_PRIVATE;
        v.writable = true;
        v.lv = s->cname;
        break;
    case Sym::BUF:
        v.is_ref = true;
        v.space = s->space;
        v.slot = s->slot;
        v.off = "0";
        cx->fn->slots.insert(s->slot);
        break;
    case Sym::HANDLE:
        v.handle = s->slot;
        cx->fn->slots.insert(s->slot);
        break;
    default:
        v.code = s->cname;
        break;
    }
    return v;
}

Val Translator::expr(FnCtx *cx, const ExprP &e)
{
    switch (e->k) {
    case Expr::LIT: return lit(e->lit);
    case Expr::IDENT: return ident(cx, *e);
    case Expr::UNARY: return unary(cx, *e);
    case Expr::BINARY:
        if (e->name == "&&" || e->name == "||") return logical(cx, *e);
        return binary(cx, *e);
    case Expr::CALL: return call(cx, *e);
    case Expr::MEMBER: {
        Val b = expr(cx, e->a[0]);
        return member(cx, b, e->name, e->loc);
    }
    case Expr::INDEX: {
        Val b = expr(cx, e->a[0]);
        Val i = rvalue(cx, expr(cx, e->a[1]), e->loc);
        return index(cx, b, i, e->loc);
    }
    }
    fail(e->loc, "internal: unknown expression");
}

// ---- scalar arithmetic ----

static int64_t wrap_i32(int64_t v) { return (int64_t)(int32_t)(uint32_t)(uint64_t)v; }
static int64_t wrap_u32(int64_t v) { return (int64_t)(uint32_t)(uint64_t)v; }

Const Translator::fold_scalar(const std::string &op, int sty, const Const &a, const Const &b, Loc l)
{
    Const r;
    r.ty = sty;
    auto div0 = [&]() { fail(l, "division by zero in a constant expression"); };
    if (sty == TY_BOOL) {
        if (op == "&") r.i = a.i & b.i;
        else if (op == "|") r.i = a.i | b.i;
        else if (op == "^") r.i = a.i ^ b.i;
        else fail(l, "operator '" + op + "' is not defined for bool");
        return r;
    }
    if (sty == TY_F32 || sty == TY_AFLOAT) {
        if (op == "&" || op == "|" || op == "^" || op == "<<" || op == ">>")
            fail(l, "operator '" + op + "' is not defined for floats");
        if (sty == TY_F32) {
            float x = (float)a.f, y = (float)b.f, z;
            if (op == "+") z = x + y;
            else if (op == "-") z = x - y;
            else if (op == "*") z = x * y;
            else if (op == "/") z = x / y;
            else {
                float q = truncf(x / y);
                float m = y * q;
                z = x - m;
            }
            if (std::isinf(z) || std::isnan(z)) fail(l, "constant expression overflows f32");
            r.f = (double)z;
        } else {
            double x = a.f, y = b.f, z;
            if (op == "+") z = x + y;
            else if (op == "-") z = x - y;
            else if (op == "*") z = x * y;
            else if (op == "/") z = x / y;
            else z = x - y * trunc(x / y);
            if (std::isinf(z) || std::isnan(z)) fail(l, "constant expression overflows");
            r.f = z;
        }
        return r;
    }
    int64_t x = a.i, y = b.i, z = 0;
    if (sty == TY_AINT) {
        if (op == "+") {
            if (__builtin_add_overflow(x, y, &z)) fail(l, "constant expression overflows");
        } else if (op == "-") {
            if (__builtin_sub_overflow(x, y, &z)) fail(l, "constant expression overflows");
        } else if (op == "*") {
            if (__builtin_mul_overflow(x, y, &z)) fail(l, "constant expression overflows");
        } else if (op == "/" || op == "%") {
            if (y == 0) div0();
            if (x == INT64_MIN && y == -1) fail(l, "constant expression overflows");
            z = op == "/" ? x / y : x % y;
        } else if (op == "&") z = x & y;
        else if (op == "|") z = x | y;
        else if (op == "^") z = x ^ y;
        else if (op == "<<") {
            if (y < 0 || y >= 63) fail(l, "shift amount out of range");
            z = (int64_t)((uint64_t)x << y);
            if ((z >> y) != x) fail(l, "constant expression overflows");
        } else if (op == ">>") {
            if (y < 0 || y >= 64) fail(l, "shift amount out of range");
            z = x >> y;
        }
        r.i = z;
        return r;
    }
    bool s = sty == TY_I32;
    uint32_t ux = (uint32_t)x, uy = (uint32_t)y;
    if (op == "+") z = (int64_t)(uint32_t)(ux + uy);
    else if (op == "-") z = (int64_t)(uint32_t)(ux - uy);
    else if (op == "*") z = (int64_t)(uint32_t)(ux * uy);
    else if (op == "/" || op == "%") {
        if (y == 0) div0();
        if (s) {
            if (x == INT32_MIN && y == -1) {
                if (op == "/") fail(l, "constant expression overflows");
                z = 0;
            } else {
                z = op == "/" ? x / y : x % y;
            }
        } else {
            z = op == "/" ? (int64_t)(ux / uy) : (int64_t)(ux % uy);
        }
    } else if (op == "&") z = (int64_t)(ux & uy);
    else if (op == "|") z = (int64_t)(ux | uy);
    else if (op == "^") z = (int64_t)(ux ^ uy);
    else if (op == "<<") z = (int64_t)(uint32_t)(ux << (uy & 31u));
    else if (op == ">>") {
        if (s) {
            int32_t sx = (int32_t)ux;
            uint32_t sh = uy & 31u;
            uint32_t u = ux >> sh;
            if (sx < 0 && sh) u |= ~(0xFFFFFFFFu >> sh);
            z = (int64_t)u;
        } else {
            z = (int64_t)(ux >> (uy & 31u));
        }
    }
    r.i = s ? wrap_i32(z) : wrap_u32(z);
    return r;
}

std::string Translator::scalar_op(FnCtx *cx, const std::string &op, int sty, const std::string &a,
                                  const std::string &b)
{
    if (sty == TY_BOOL) return "(bool)(" + a + " " + op + " " + b + ")";
    if (sty == TY_F32) {
        if (op == "%") return "wgsl::frem(" + a + ", " + b + ")";
        return "(" + a + " " + op + " " + b + ")";
    }
    bool s = sty == TY_I32;
    if (op == "/" || op == "%") {
        std::string t = new_tmp(cx);
        line(cx, std::string(s ? "int32_t " : "uint32_t ") + t + ";");
        std::string fn = s ? (op == "/" ? "wgsl::sdiv" : "wgsl::srem") : (op == "/" ? "wgsl::udiv" : "wgsl::urem");
        check_abort(cx, fn + "(" + a + ", " + b + ", " + t + ")");
        return t;
    }
    if (op == "<<") return std::string(s ? "wgsl::ishl(" : "wgsl::ushl(") + a + ", " + b + ")";
    if (op == ">>") return std::string(s ? "wgsl::sshr(" : "wgsl::ushr(") + a + ", " + b + ")";
    if (s) {
        if (op == "+") return "wgsl::iadd(" + a + ", " + b + ")";
        if (op == "-") return "wgsl::isub(" + a + ", " + b + ")";
        if (op == "*") return "wgsl::imul(" + a + ", " + b + ")";
        return "(int32_t)(" + a + " " + op + " " + b + ")";
    }
    return "(uint32_t)(" + a + " " + op + " " + b + ")";
}

// Arithmetic and bitwise operators (binary.rs:44-176): scalars, vectors
// lane by lane, vector-scalar with the scalar repeated, matrices lane by
// lane, matrix x scalar, and matrix x vector as the column sum
// ((c0*v0 + c1*v1) + c2*v2) + c3*v3 (binary.rs:299-321).
Val Translator::arith(FnCtx *cx, const std::string &op, Val a, Val b, Loc l)
{
    a = rvalue(cx, a, l);
    b = rvalue(cx, b, l);
    int sa = scalar_of(a.ty), sb = scalar_of(b.ty);
    if (sa < 0 || sb < 0) fail(l, "operator '" + op + "' is not defined for " + tname(a.ty) + " and " + tname(b.ty));
    // abstract operands take the other operand's scalar type
    if (is_abstract_scalar(sa) && !is_abstract_scalar(sb)) {
        a = convert(cx, a, with_scalar(a.ty, sb), l);
        sa = sb;
    } else if (is_abstract_scalar(sb) && !is_abstract_scalar(sa)) {
        b = convert(cx, b, with_scalar(b.ty, sa), l);
        sb = sa;
    } else if (sa != sb && is_abstract_scalar(sa) && is_abstract_scalar(sb)) {
        if (sa == TY_AINT) a = cvalue(convert_const(a.c, with_scalar(a.ty, TY_AFLOAT), l));
        else b = cvalue(convert_const(b.c, with_scalar(b.ty, TY_AFLOAT), l));
        sa = sb = TY_AFLOAT;
    }
    if (sa != sb) fail(l, "operator '" + op + "': mismatched types " + tname(a.ty) + " and " + tname(b.ty));
    bool shift = op == "<<" || op == ">>";
    bool bitwise = op == "&" || op == "|" || op == "^";
    if (sa == TY_BOOL && !bitwise) fail(l, "operator '" + op + "' is not defined for bool");
    if ((bitwise || shift) && is_float_scalar(sa)) fail(l, "operator '" + op + "' is not defined for floats");
    if (shift && !is_scalar(a.ty)) fail(l, "vector shifts are not implemented by the reference (binary.rs:128-176)");
    const TypeInfo &A = ti(a.ty), &B = ti(b.ty);
    // matrix products
    if (op == "*" && A.k == TK_MAT && B.k == TK_VEC) {
        if (A.n != B.n) fail(l, "matrix columns and vector size differ");
        int rty = vec_of(TY_F32, A.rows);
        if (a.is_const && b.is_const) {
            Const r = zero_const(rty);
            for (int j = 0; j < A.rows; ++j) {
                float s = 0.0f;
                for (int i = 0; i < A.n; ++i) {
                    float x = (float)a.c.el[i].el[j].f * (float)b.c.el[i].f;
                    s = i == 0 ? x : s + x;
                }
                r.el[j].f = (double)s;
            }
            return cvalue(r);
        }
        std::string ca = code(cx, a), cb = code(cx, b);
        std::string t = new_tmp(cx);
        line(cx, ctype(rty) + " " + t + ";");
        for (int j = 0; j < A.rows; ++j) {
            std::string s;
            for (int i = 0; i < A.n; ++i) {
                std::string x = ca + ".c[" + std::to_string(i) + "].v[" + std::to_string(j) + "] * " + cb + ".v[" +
                                std::to_string(i) + "]";
                s = i == 0 ? x : "(" + s + ") + " + x;
            }
            line(cx, t + ".v[" + std::to_string(j) + "] = " + s + ";");
        }
        Val v;
        v.ty = rty;
        v.code = t;
        return v;
    }
    if (op == "*" && A.k == TK_VEC && B.k == TK_MAT) fail(l, "vector * matrix is todo!() in the reference (binary.rs:231-240)");
    if (op == "*" && A.k == TK_MAT && B.k == TK_MAT) fail(l, "matrix * matrix is todo!() in the reference (binary.rs:221-230)");
    // result shape
    int rty;
    if (A.k == TK_MAT || B.k == TK_MAT) {
        bool ok = (op == "+" || op == "-") ? (a.ty == b.ty) : (op == "*" && (is_scalar(a.ty) || is_scalar(b.ty)));
        if (!ok) fail(l, "operator '" + op + "' is not defined for " + tname(a.ty) + " and " + tname(b.ty));
        rty = A.k == TK_MAT ? a.ty : b.ty;
    } else if (A.k == TK_VEC && B.k == TK_VEC) {
        if (a.ty != b.ty) fail(l, "operator '" + op + "': vector sizes differ");
        rty = a.ty;
    } else if (A.k == TK_VEC || B.k == TK_VEC) {
        if (bitwise) fail(l, "operator '" + op + "' needs operands of the same type");
        rty = A.k == TK_VEC ? a.ty : b.ty;
    } else {
        rty = a.ty;
    }
    int n = lanes(rty);
    if (a.is_const && b.is_const) {
        Const r = is_scalar(rty) ? Const{} : zero_const(rty);
        r.ty = rty;
        for (int k = 0; k < n; ++k) {
            const Const &x = is_scalar(a.ty) ? a.c : *const_lane(a.c, k);
            const Const &y = is_scalar(b.ty) ? b.c : *const_lane(b.c, k);
            Const z = fold_scalar(op, sa, x, y, l);
            if (is_scalar(rty)) r = z;
            else *const_lane(r, k) = z;
        }
        return cvalue(r);
    }
    if (!cx) fail(l, "not a constant expression");
    if (is_abstract_scalar(sa)) fail(l, "internal: abstract runtime value");
    std::string ca = code(cx, a), cb = code(cx, b);
    if (is_scalar(rty)) return tmpval(cx, rty, scalar_op(cx, op, sa, ca, cb));
    std::string t = new_tmp(cx);
    line(cx, ctype(rty) + " " + t + ";");
    for (int k = 0; k < n; ++k) {
        std::string x = is_scalar(a.ty) ? ca : lane(ca, a.ty, k);
        std::string y = is_scalar(b.ty) ? cb : lane(cb, b.ty, k);
        std::string z = scalar_op(cx, op, sa, x, y);
        line(cx, lane(t, rty, k) + " = " + z + ";");
    }
    Val v;
    v.ty = rty;
    v.code = t;
    return v;
}

// Comparisons: scalar operands only (binary.rs:545-606); bool operands for
// == and != are allowed as well.
Val Translator::compare(FnCtx *cx, const std::string &op, Val a, Val b, Loc l)
{
    a = rvalue(cx, a, l);
    b = rvalue(cx, b, l);
    if (!is_scalar(a.ty) || !is_scalar(b.ty))
        fail(l, "vector comparisons are not implemented by the reference (binary.rs:172-176)");
    if (is_abstract_scalar(a.ty) && !is_abstract_scalar(b.ty)) a = convert(cx, a, b.ty, l);
    else if (is_abstract_scalar(b.ty) && !is_abstract_scalar(a.ty)) b = convert(cx, b, a.ty, l);
    else if (a.ty != b.ty && is_abstract_scalar(a.ty) && is_abstract_scalar(b.ty)) {
        a = cvalue(convert_const(a.c, TY_AFLOAT, l));
        b = cvalue(convert_const(b.c, TY_AFLOAT, l));
    }
    if (a.ty != b.ty) fail(l, "comparison of " + tname(a.ty) + " with " + tname(b.ty));
    if (a.ty == TY_BOOL && op != "==" && op != "!=") fail(l, "operator '" + op + "' is not defined for bool");
    if (a.is_const && b.is_const) {
        Const r;
        r.ty = TY_BOOL;
        if (is_float_scalar(a.ty)) {
            double x = a.c.f, y = b.c.f;
            r.i = op == "==" ? x == y : op == "!=" ? x != y : op == "<" ? x < y : op == "<=" ? x <= y
                : op == ">" ? x > y : x >= y;
        } else {
            int64_t x = a.c.i, y = b.c.i;
            r.i = op == "==" ? x == y : op == "!=" ? x != y : op == "<" ? x < y : op == "<=" ? x <= y
                : op == ">" ? x > y : x >= y;
        }
        return cvalue(r);
    }
    if (!cx) fail(l, "not a constant expression");
    return tmpval(cx, TY_BOOL, "(" + code(cx, a) + " " + op + " " + code(cx, b) + ")");
}

Val Translator::binary(FnCtx *cx, const Expr &e)
{
    Val a = rvalue(cx, expr(cx, e.a[0]), e.loc);
    Val b = rvalue(cx, expr(cx, e.a[1]), e.loc);
    const std::string &op = e.name;
    if (op == "==" || op == "!=" || op == "<" || op == "<=" || op == ">" || op == ">=")
        return compare(cx, op, a, b, e.loc);
    if (op == "<<" || op == ">>") {
        // the shift amount is u32; the value keeps its type
        if (!is_int_scalar(scalar_of(a.ty)) || !is_int_scalar(scalar_of(b.ty)))
            fail(e.loc, "shifts need integer operands");
        if (a.is_const && b.is_const && is_abstract_scalar(a.ty)) {
            Const bb = b.c;
            bb.ty = TY_AINT;
            return cvalue(fold_scalar(op, TY_AINT, a.c, bb, e.loc));
        }
        if (is_abstract_scalar(b.ty)) b = convert(cx, b, TY_U32, e.loc);
        if (b.ty != TY_U32) fail(e.loc, "the shift amount must be u32");
        if (is_abstract_scalar(a.ty)) a = convert(cx, a, TY_I32, e.loc);
        if (a.is_const && b.is_const) return cvalue(fold_scalar(op, a.ty, a.c, [&] { Const c = b.c; c.ty = a.ty; return c; }(), e.loc));
        if (!cx) fail(e.loc, "not a constant expression");
        return tmpval(cx, a.ty, scalar_op(cx, op, a.ty, code(cx, a), code(cx, b)));
    }
    return arith(cx, op, a, b, e.loc);
}

// && and || evaluate their right operand only when needed.
Val Translator::logical(FnCtx *cx, const Expr &e)
{
    Val a = rvalue(cx, expr(cx, e.a[0]), e.loc);
    if (a.ty != TY_BOOL) fail(e.loc, "operator '" + e.name + "' needs bool operands");
    bool is_and = e.name == "&&";
    if (!cx) {
        Val b = rvalue(cx, expr(cx, e.a[1]), e.loc);
        if (b.ty != TY_BOOL) fail(e.loc, "operator '" + e.name + "' needs bool operands");
        Const r;
        r.ty = TY_BOOL;
        r.i = is_and ? (a.c.i && b.c.i) : (a.c.i || b.c.i);
        return cvalue(r);
    }
    std::string t = new_tmp(cx);
    line(cx, "bool " + t + " = " + code(cx, a) + ";");
    line(cx, std::string("if (") + (is_and ? "" : "!") + t + ") {");
    cx->indent++;
    Val b = rvalue(cx, expr(cx, e.a[1]), e.loc);
    if (b.ty != TY_BOOL) fail(e.loc, "operator '" + e.name + "' needs bool operands");
    line(cx, t + " = " + code(cx, b) + ";");
    cx->indent--;
    line(cx, "}");
    Val v;
    v.ty = TY_BOOL;
    v.code = t;
    return v;
}

// Unary operators (unary.rs:96-160), lane by lane for vectors and matrices.
Val Translator::unary(FnCtx *cx, const Expr &e)
{
    const std::string &op = e.name;
    if (op == "&") {
        Val r = expr(cx, e.a[0]);
        if (!r.is_ref) fail(e.loc, "'&' needs a variable");
        r.is_ptr = true;
        return r;
    }
    if (op == "*") fail(e.loc, "pointers are not supported");
    Val a = rvalue(cx, expr(cx, e.a[0]), e.loc);
    int s = scalar_of(a.ty);
    if (s < 0) fail(e.loc, "operator '" + op + "' is not defined for " + tname(a.ty));
    if (op == "-" && !(s == TY_I32 || s == TY_F32 || s == TY_AINT || s == TY_AFLOAT))
        fail(e.loc, "negation is not valid for " + tname(a.ty));
    if (op == "!" && s != TY_BOOL) fail(e.loc, "logical not is not valid for " + tname(a.ty));
    if (op == "~" && !is_int_scalar(s)) fail(e.loc, "bitwise not is not valid for " + tname(a.ty));
    int n = lanes(a.ty);
    if (a.is_const) {
        Const r = a.c;
        for (int k = 0; k < n; ++k) {
            Const &x = *const_lane(r, k);
            if (op == "-") {
                if (s == TY_AINT) {
                    if (x.i == INT64_MIN) fail(e.loc, "constant expression overflows");
                    x.i = -x.i;
                } else if (s == TY_I32) x.i = wrap_i32(-x.i);
                else x.f = -x.f;
            } else if (op == "!") {
                x.i = !x.i;
            } else {
                x.i = s == TY_AINT ? ~x.i : s == TY_I32 ? wrap_i32(~x.i) : wrap_u32(~x.i);
            }
        }
        return cvalue(r);
    }
    if (!cx) fail(e.loc, "not a constant expression");
    std::string ca = code(cx, a);
    auto one = [&](const std::string &x) -> std::string {
        if (op == "-") return s == TY_I32 ? "wgsl::ineg(" + x + ")" : "(-" + x + ")";
        if (op == "!") return "(!" + x + ")";
        return "(" + ctype(s) + ")(~" + x + ")";
    };
    if (is_scalar(a.ty)) return tmpval(cx, a.ty, one(ca));
    std::string t = new_tmp(cx);
    line(cx, ctype(a.ty) + " " + t + ";");
    for (int k = 0; k < n; ++k) line(cx, lane(t, a.ty, k) + " = " + one(lane(ca, a.ty, k)) + ";");
    Val v;
    v.ty = a.ty;
    v.code = t;
    return v;
}

// ---- conversions T(e) (as.rs:61-172) ----

Const Translator::fold_convert(const Const &c, int to, Loc l)
{
    if (!is_scalar(c.ty)) {
        Const r = zero_const(to);
        int n = lanes(c.ty);
        Const src = c;
        for (int k = 0; k < n; ++k) *const_lane(r, k) = fold_convert(*const_lane(src, k), scalar_of(to), l);
        return r;
    }
    Const r;
    r.ty = to;
    int from = c.ty;
    if (from == to) return c;
    bool from_f = is_float_scalar(from);
    double fv = from_f ? c.f : (double)c.i;
    switch (to) {
    case TY_BOOL:
        r.i = from_f ? (fv != 0.0 || std::isnan(fv)) : c.i != 0;
        break;
    case TY_I32:
        if (from == TY_BOOL) r.i = c.i;
        else if (from_f) {
            float f = (float)fv;
            if (!(f >= -2147483648.0f && f < 2147483648.0f)) fail(l, "value out of range for i32");
            r.i = (int64_t)(int32_t)f;
        } else if (from == TY_AINT) {
            if (c.i < INT32_MIN || c.i > INT32_MAX) fail(l, "value out of range for i32");
            r.i = c.i;
        } else {
            r.i = wrap_i32(c.i);
        }
        break;
    case TY_U32:
        if (from == TY_BOOL) r.i = c.i;
        else if (from_f) {
            float f = (float)fv;
            if (!(f > -1.0f && f < 4294967296.0f)) fail(l, "value out of range for u32");
            r.i = f <= 0.0f ? 0 : (int64_t)(uint32_t)f;
        } else if (from == TY_AINT) {
            if (c.i < 0 || c.i > (int64_t)UINT32_MAX) fail(l, "value out of range for u32");
            r.i = c.i;
        } else {
            r.i = wrap_u32(c.i);
        }
        break;
    case TY_F32: {
        float f;
        if (from == TY_BOOL) f = c.i ? 1.0f : 0.0f;
        else if (from == TY_I32) f = (float)(int32_t)c.i;
        else if (from == TY_U32) f = (float)(uint32_t)c.i;
        else f = (float)fv;
        if (std::isinf(f)) fail(l, "value out of range for f32");
        r.f = (double)f;
        break;
    }
    default:
        fail(l, "cannot convert to " + tname(to));
    }
    return r;
}

Val Translator::convert_value(FnCtx *cx, const Val &v0, int to, Loc l)
{
    Val v = rvalue(cx, v0, l);
    if (v.ty == to) return v;
    int fs = scalar_of(v.ty), ts = scalar_of(to);
    bool same_shape = is_scalar(v.ty) ? is_scalar(to) : (ti(v.ty).k == ti(to).k && ti(v.ty).n == ti(to).n);
    if (fs < 0 || ts < 0 || lanes(v.ty) != lanes(to) || !same_shape)
        fail(l, "cannot convert " + tname(v.ty) + " to " + tname(to));
    if (v.is_const) return cvalue(fold_convert(v.c, to, l));
    if (!cx) fail(l, "not a constant expression");
    std::string cv = code(cx, v);
    auto one = [&](const std::string &x) -> std::string {
        if (fs == ts) return x;
        switch (ts) {
        case TY_BOOL:
            if (fs == TY_F32) return "wgsl::f2b(" + x + ")";
            return "(" + x + " != 0)";
        case TY_I32:
        case TY_U32: {
            const char *ct = ts == TY_I32 ? "int32_t" : "uint32_t";
            if (fs == TY_BOOL) return std::string("(") + ct + ")(" + x + " ? 1 : 0)";
            if (fs == TY_F32) {
                std::string t = new_tmp(cx);
                line(cx, std::string(ct) + " " + t + ";");
                check_abort(cx, std::string(ts == TY_I32 ? "wgsl::f2i(" : "wgsl::f2u(") + x + ", " + t + ")");
                return t;
            }
            return std::string("(") + ct + ")" + x;
        }
        case TY_F32:
            if (fs == TY_BOOL) return "(" + x + " ? 1.0f : 0.0f)";
            return "(float)" + x;
        }
        throw Error("internal: conversion");
    };
    if (is_scalar(to)) return tmpval(cx, to, one(cv));
    std::string t = new_tmp(cx);
    line(cx, ctype(to) + " " + t + ";");
    for (int k = 0; k < lanes(to); ++k) line(cx, lane(t, to, k) + " = " + one(lane(cv, v.ty, k)) + ";");
    Val r;
    r.ty = to;
    r.code = t;
    return r;
}

// ---- construction (compose.rs, zero_value.rs) ----

Val Translator::construct(FnCtx *cx, int ty, const std::vector<Val> &args0, Loc l)
{
    if (!constructible(ty)) fail(l, tname(ty) + " cannot be constructed");
    if (args0.empty()) return cvalue(zero_const(ty));
    std::vector<Val> args;
    for (auto &a : args0) args.push_back(rvalue(cx, a, l));
    const TypeInfo &t = ti(ty);
    std::vector<Val> parts;   // converted components, in order
    int part_ty = -1;
    if (is_scalar(ty)) {
        if (args.size() != 1) fail(l, tname(ty) + "() takes one argument");
        if (!is_scalar(args[0].ty)) fail(l, "cannot convert " + tname(args[0].ty) + " to " + tname(ty));
        return convert_value(cx, args[0], ty, l);
    }
    if (t.k == TK_VEC) {
        if (args.size() == 1 && ti(args[0].ty).k == TK_VEC) {
            if (ti(args[0].ty).n != t.n) fail(l, "vector sizes differ");
            return convert_value(cx, args[0], ty, l);
        }
        part_ty = t.elem;
        if (args.size() == 1 && is_scalar(args[0].ty)) {
            Val s = convert(cx, args[0], t.elem, l);
            for (int k = 0; k < t.n; ++k) parts.push_back(s);
        } else {
            for (auto &a : args) {
                if (is_scalar(a.ty)) {
                    parts.push_back(convert(cx, a, t.elem, l));
                } else if (ti(a.ty).k == TK_VEC) {
                    Val vv = convert(cx, a, vec_of(t.elem, ti(a.ty).n), l);
                    for (int k = 0; k < ti(a.ty).n; ++k) {
                        Val idx = cvalue(Const{TY_U32, k, 0.0, {}});
                        parts.push_back(index(cx, vv, idx, l));
                    }
                } else {
                    fail(l, "invalid vector component " + tname(a.ty));
                }
            }
            if ((int)parts.size() != t.n) fail(l, tname(ty) + " needs " + std::to_string(t.n) + " components");
        }
    } else if (t.k == TK_MAT) {
        int col = vec_of(TY_F32, t.rows);
        part_ty = col;
        if (args.size() == 1 && ti(args[0].ty).k == TK_MAT) return convert(cx, args[0], ty, l);
        if ((int)args.size() == t.n) {
            for (auto &a : args) parts.push_back(convert(cx, a, col, l));
        } else if ((int)args.size() == t.n * t.rows) {
            for (int c = 0; c < t.n; ++c) {
                std::vector<Val> comps(args.begin() + c * t.rows, args.begin() + (c + 1) * t.rows);
                parts.push_back(construct(cx, col, comps, l));
            }
        } else {
            fail(l, tname(ty) + " needs " + std::to_string(t.n) + " columns");
        }
    } else if (t.k == TK_ARR) {
        part_ty = t.elem;
        if ((int)args.size() != t.n) fail(l, tname(ty) + " needs " + std::to_string(t.n) + " elements");
        for (auto &a : args) parts.push_back(convert(cx, a, t.elem, l));
    } else if (t.k == TK_STRUCT) {
        auto &ms = structs_[t.sid].members;
        if (args.size() != ms.size()) fail(l, tname(ty) + " needs " + std::to_string(ms.size()) + " members");
        for (size_t k = 0; k < ms.size(); ++k) parts.push_back(convert(cx, args[k], ms[k].ty, l));
    } else {
        fail(l, tname(ty) + " cannot be constructed");
    }
    (void)part_ty;
    bool all_const = true;
    for (auto &p : parts) all_const = all_const && p.is_const;
    if (all_const) {
        Const r;
        r.ty = ty;
        for (auto &p : parts) r.el.push_back(p.c);
        return cvalue(r);
    }
    if (!cx) fail(l, "not a constant expression");
    std::string init = ctype(ty) + (t.k == TK_STRUCT ? "{" : "{{");
    for (size_t k = 0; k < parts.size(); ++k) init += (k ? ", " : "") + code(cx, parts[k]);
    init += t.k == TK_STRUCT ? "}" : "}}";
    return tmpval(cx, ty, init);
}

// ---- member access (access.rs:82-245) and swizzles ----

static int component_index(char c)
{
    switch (c) {
    case 'x': case 'r': return 0;
    case 'y': case 'g': return 1;
    case 'z': case 'b': return 2;
    case 'w': case 'a': return 3;
    default: return -1;
    }
}

Val Translator::member(FnCtx *cx, const Val &base, const std::string &name, Loc l)
{
    const TypeInfo &t = ti(base.ty);
    if (t.k == TK_STRUCT) {
        const StructInfo &s = structs_[t.sid];
        for (size_t k = 0; k < s.members.size(); ++k) {
            const MemberInfo &mb = s.members[k];
            if (mb.name != name) continue;
            Val r = base;
            r.ty = mb.ty;
            r.is_ptr = false;
            if (base.is_ref) {
                if (base.space == SP_FUNCTION || base.space == SP_PRIVATE) r.lv = base.lv + "." + mb.cname;
                else {
                    r.off = base.off + " + " + std::to_string(mb.offset);
                    if (ti(mb.ty).k == TK_RTARR) r.rt_base = 0;   // offset held in `off`
                }
                return r;
            }
            if (base.is_const) return cvalue(base.c.el[k]);
            r.code = base.code + "." + mb.cname;
            return r;
        }
        fail(l, "'" + tname(base.ty) + "' has no member '" + name + "'");
    }
    if (t.k == TK_VEC) {
        if (name.size() > 4) fail(l, "invalid vector component '" + name + "'");
        std::vector<int> comps;
        bool rgba = false, xyzw = false;
        for (char c : name) {
            int k = component_index(c);
            if (k < 0 || k >= t.n) fail(l, "invalid vector component '" + name + "'");
            (strchr("rgba", c) ? rgba : xyzw) = true;
            comps.push_back(k);
        }
        if (rgba && xyzw) fail(l, "mixed component sets in '" + name + "'");
        if (comps.size() == 1) return index(cx, base, cvalue(Const{TY_U32, comps[0], 0.0, {}}), l);
        // swizzle: a new value
        Val v = rvalue(cx, base, l);
        std::vector<Val> parts;
        for (int k : comps) parts.push_back(index(cx, v, cvalue(Const{TY_U32, k, 0.0, {}}), l));
        return construct(cx, vec_of(t.elem, (int)comps.size()), parts, l);
    }
    fail(l, "'" + tname(base.ty) + "' has no members");
}

// ---- indexing (access.rs:82-380) ----

Val Translator::index(FnCtx *cx, const Val &base, const Val &idx0, Loc l)
{
    const TypeInfo &t = ti(base.ty);
    int n;
    int ety;
    std::string sel;
    if (t.k == TK_VEC) {
        n = t.n;
        ety = t.elem;
        sel = ".v[";
    } else if (t.k == TK_MAT) {
        n = t.n;
        ety = vec_of(TY_F32, t.rows);
        sel = ".c[";
    } else if (t.k == TK_ARR) {
        n = t.n;
        ety = t.elem;
        sel = ".v[";
    } else if (t.k == TK_RTARR) {
        n = -1;
        ety = t.elem;
    } else {
        fail(l, "'" + tname(base.ty) + "' cannot be indexed");
    }
    Val idx = idx0;
    if (!is_int_scalar(idx.ty)) fail(l, "index must be an integer");
    if (idx.is_const) {
        if (idx.c.i < 0 || (n >= 0 && idx.c.i >= n)) fail(l, "index " + std::to_string(idx.c.i) + " out of bounds");
        if (base.is_const) return cvalue(base.c.el[(size_t)idx.c.i]);
    }
    if (!cx) fail(l, "not a constant expression");
    std::string ic = idx.is_const ? std::to_string(idx.c.i) + "u" : "(uint32_t)" + code(cx, idx);
    Val r;
    r.ty = ety;
    if (base.is_ref && (base.space == SP_UNIFORM || base.space == SP_STORAGE)) {
        r = base;
        r.ty = ety;
        r.is_ptr = false;
        uint32_t stride = t.k == TK_VEC ? 4u : t.k == TK_MAT ? column_stride(base.ty) : stride_of(base.ty);
        if (t.k == TK_RTARR) {
            // element count from the bound size (array_length.rs:38-59)
            std::string cnt = "((env.buffer_size(" + std::to_string(base.slot) + ") - (uint64_t)(" + base.off + ")) / " +
                              std::to_string(stride) + "u)";
            line(cx, "if ((uint64_t)" + ic + " >= " + cnt + ") return wgsl::ABORT_OOB;");
        } else if (!idx.is_const) {
            line(cx, "if (" + ic + " >= " + std::to_string(n) + "u) return wgsl::ABORT_OOB;");
        }
        r.off = base.off + " + (uint64_t)" + ic + " * " + std::to_string(stride) + "u";
        return r;
    }
    if (!idx.is_const) line(cx, "if (" + ic + " >= " + std::to_string(n) + "u) return wgsl::ABORT_OOB;");
    if (base.is_ref) {
        r = base;
        r.ty = ety;
        r.is_ptr = false;
        r.lv = base.lv + sel + ic + "]";
        return r;
    }
    r.code = code(cx, base) + sel + ic + "]";
    return r;
}

// ---- calls ----

static bool is_math_builtin(const std::string &n)
{
    static const char *names[] = {
        "abs", "acos", "acosh", "asin", "asinh", "atan", "atanh", "atan2", "ceil", "clamp", "cos", "cosh",
        "countLeadingZeros", "countOneBits", "countTrailingZeros", "cross", "degrees", "determinant", "distance",
        "dot", "dot4U8Packed", "dot4I8Packed", "exp", "exp2", "extractBits", "faceForward", "firstLeadingBit",
        "firstTrailingBit", "floor", "fma", "fract", "frexp", "insertBits", "inverseSqrt", "ldexp", "length",
        "log", "log2", "max", "min", "mix", "modf", "normalize", "pow", "quantizeToF16", "radians", "reflect",
        "refract", "reverseBits", "round", "saturate", "sign", "sin", "sinh", "smoothstep", "sqrt", "step", "tan",
        "tanh", "transpose", "trunc", "pack4x8snorm", "pack4x8unorm", "pack4xI8", "pack4xU8", "pack4xI8Clamp",
        "pack4xU8Clamp", "pack2x16snorm", "pack2x16unorm", "pack2x16float", "unpack4x8snorm", "unpack4x8unorm",
        "unpack4xI8", "unpack4xU8", "unpack2x16snorm", "unpack2x16unorm", "unpack2x16float", nullptr};
    for (int i = 0; names[i]; ++i)
        if (n == names[i]) return true;
    return false;
}

Val Translator::call(FnCtx *cx, const Expr &e)
{
    // type constructors
    int ty = -1;
    bool inferred = false;
    if (e.tmpl) {
        if (e.tmpl->name != "bitcast") ty = resolve_type(e.tmpl);
    } else {
        static const char *scalar_or_alias[] = {"bool", "i32", "u32", "f32", "vec2f", "vec3f", "vec4f", "vec2i",
                                               "vec3i", "vec4i", "vec2u", "vec3u", "vec4u", nullptr};
        for (int i = 0; scalar_or_alias[i]; ++i)
            if (e.name == scalar_or_alias[i]) {
                auto tp = std::make_shared<TypeAst>();
                tp->name = e.name;
                tp->loc = e.loc;
                ty = resolve_type(tp);
            }
        if (ty < 0 && e.name.size() == 7 && e.name.compare(0, 3, "mat") == 0 && e.name[6] == 'f') {
            auto tp = std::make_shared<TypeAst>();
            tp->name = e.name;
            tp->loc = e.loc;
            ty = resolve_type(tp);
        }
        if (ty < 0 && (struct_by_name_.count(e.name) || alias_.count(e.name))) {
            auto tp = std::make_shared<TypeAst>();
            tp->name = e.name;
            tp->loc = e.loc;
            ty = resolve_type(tp);
        }
        if (ty < 0 && (e.name == "vec2" || e.name == "vec3" || e.name == "vec4" || e.name == "array" ||
                       (e.name.size() == 6 && e.name.compare(0, 3, "mat") == 0 && e.name[4] == 'x')))
            inferred = true;
    }
    std::vector<Val> args;
    if (ty >= 0 || inferred) {
        for (auto &a : e.a) args.push_back(rvalue(cx, expr(cx, a), a->loc));
        if (inferred) {
            // element type from the arguments (abstract when all are abstract)
            if (args.empty()) fail(e.loc, e.name + "() without a type needs arguments");
            int s = -1;
            for (auto &a : args) {
                int as = e.name == "array" ? a.ty : scalar_of(a.ty);
                if (as < 0) fail(e.loc, "invalid constructor argument");
                if (s < 0 || (is_abstract(s) && !is_abstract(as))) s = as;
                else if (s == TY_AINT && as == TY_AFLOAT) s = TY_AFLOAT;
            }
            if (e.name == "array") ty = arr_of(s, (int)args.size());
            else if (e.name[0] == 'v') ty = vec_of(s, e.name[3] - '0');
            else ty = mat_of(e.name[3] - '0', e.name[5] - '0');
        }
        return construct(cx, ty, args, e.loc);
    }
    auto f = fn_by_name_.find(e.name);
    if (f != fn_by_name_.end()) {
        if (!cx) fail(e.loc, "function calls are not constant expressions");
        for (auto &a : e.a) args.push_back(rvalue(cx, expr(cx, a), a->loc));
        return user_call(cx, f->second, e, args);
    }
    if (e.name == "arrayLength") {
        if (e.a.size() != 1) fail(e.loc, "arrayLength takes one argument");
        args.push_back(expr(cx, e.a[0]));
    } else {
        for (auto &a : e.a) {
            Val v = expr(cx, a);
            args.push_back(v.handle >= 0 ? v : rvalue(cx, v, a->loc));
        }
    }
    return builtin_call(cx, e, args);
}

Val Translator::builtin_call(FnCtx *cx, const Expr &e, std::vector<Val> &args)
{
    const std::string &n = e.name;
    Loc l = e.loc;
    if (n == "select") {
        // select.rs: scalar condition; the accepted value when true
        if (args.size() != 3) fail(l, "select takes three arguments");
        if (args[2].ty != TY_BOOL) fail(l, "select with a vector condition is not implemented by the reference (select.rs:40-43)");
        Val f = args[0], t = args[1];
        if (is_abstract(f.ty) && !is_abstract(t.ty)) f = convert(cx, f, t.ty, l);
        else if (is_abstract(t.ty) && !is_abstract(f.ty)) t = convert(cx, t, f.ty, l);
        if (f.ty != t.ty) fail(l, "select operands differ in type");
        if (f.is_const && t.is_const && args[2].is_const) return args[2].c.i ? t : f;
        if (!cx) fail(l, "not a constant expression");
        if (is_abstract(f.ty)) {
            f = convert(cx, f, concretize(f.ty), l);
            t = convert(cx, t, concretize(t.ty), l);
        }
        return tmpval(cx, f.ty, code(cx, args[2]) + " ? " + code(cx, t) + " : " + code(cx, f));
    }
    if (n == "arrayLength") {
        const Val &p = args[0];
        if (!p.is_ptr || ti(p.ty).k != TK_RTARR || !(p.space == SP_STORAGE))
            fail(l, "arrayLength needs a pointer to a runtime-sized array in a storage buffer");
        if (!cx) fail(l, "not a constant expression");
        uint32_t stride = stride_of(p.ty);
        return tmpval(cx, TY_U32, "(uint32_t)((env.buffer_size(" + std::to_string(p.slot) + ") - (uint64_t)(" + p.off +
                                      ")) / " + std::to_string(stride) + "u)");
    }
    if (n == "textureSample") {
        if (args.size() != 3) fail(l, "only textureSample(t, s, coords) is supported (no offsets or array layers)");
        if (args[0].handle < 0 || ti(args[0].ty).k != TK_TEX) fail(l, "textureSample needs a texture_2d<f32>");
        if (args[1].handle < 0 || ti(args[1].ty).k != TK_SAMP) fail(l, "textureSample needs a sampler");
        if (!cx) fail(l, "not a constant expression");
        if (cx->fn->stage != Stage::Fragment && cx->fn->is_entry)
            fail(l, "textureSample is only valid in fragment shaders");
        Val c = convert(cx, args[2], vec_of(TY_F32, 2), l);
        std::string cc = code(cx, c);
        std::string t = new_tmp(cx);
        line(cx, "wgsl::vec<float, 4> " + t + ";");
        line(cx, "env.sample(" + std::to_string(args[0].handle) + ", " + std::to_string(args[1].handle) + ", " + cc +
                     ".v[0], " + cc + ".v[1], " + t + ".v);");
        Val v;
        v.ty = vec_of(TY_F32, 4);
        v.code = t;
        return v;
    }
    if (n == "bitcast") {
        if (!e.tmpl || e.tmpl->params.size() != 1 || args.size() != 1) fail(l, "bitcast<T>(e) takes one argument");
        int to = resolve_type(e.tmpl->params[0]);
        Val v = args[0];
        if (is_abstract(v.ty)) v = convert(cx, v, with_scalar(v.ty, TY_I32), l);
        int fs = scalar_of(v.ty), ts = scalar_of(to);
        if (fs < 0 || ts < 0 || lanes(v.ty) != lanes(to)) fail(l, "bitcast needs operands of the same size");
        if (fs == TY_F32 || ts == TY_F32 || fs == TY_BOOL || ts == TY_BOOL) {
            if (fs == ts) return v;
            fail(l, "bitcast between floats and integers panics in the reference (types.rs:217-226)");
        }
        return convert_value(cx, v, to, l);
    }
    if (n == "textureLoad" || n == "textureDimensions" || n == "textureNumLevels" || n == "textureNumLayers" ||
        n == "textureNumSamples")
        fail(l, "'" + n + "' is todo!() in the reference (expression/image.rs:105-120)");
    if (n.compare(0, 7, "texture") == 0) fail(l, "'" + n + "' is not supported");
    if (n == "dpdx" || n == "dpdy" || n == "fwidth" || n.compare(0, 4, "dpdx") == 0 || n.compare(0, 4, "dpdy") == 0 ||
        n.compare(0, 6, "fwidth") == 0)
        fail(l, "derivatives are todo!() in the reference (expression/derivative.rs:15-18)");
    if (n == "all" || n == "any") fail(l, "'" + n + "' is todo!() in the reference (expression/relational.rs:19-28)");
    if (is_math_builtin(n)) fail(l, "math function '" + n + "' is todo!() in the reference (expression/math.rs:21-30)");
    if (n.compare(0, 6, "atomic") == 0) fail(l, "atomics are todo!() in the reference (statement/atomic.rs)");
    if (n.compare(0, 8, "subgroup") == 0 || n.compare(0, 4, "quad") == 0 || n == "workgroupBarrier" ||
        n == "storageBarrier" || n == "textureBarrier" || n == "workgroupUniformLoad")
        fail(l, "'" + n + "' is todo!() in the reference (statement/{barrier,subgroup,work_group}.rs)");
    fail(l, "unknown function '" + n + "'");
}

Val Translator::user_call(FnCtx *cx, int fi, const Expr &e, std::vector<Val> &args)
{
    FnInfo &f = fns_[fi];
    if (f.is_entry) fail(e.loc, "entry points cannot be called");
    if (args.size() != f.param_ty.size())
        fail(e.loc, "'" + e.name + "' takes " + std::to_string(f.param_ty.size()) + " argument(s)");
    std::string call = f.cname + "<Env>(env, pv";
    for (size_t k = 0; k < args.size(); ++k) call += ", " + code(cx, convert(cx, args[k], f.param_ty[k], e.a[k]->loc));
    cx->fn->calls.insert(fi);
    Val v;
    v.ty = f.ret_ty;
    if (f.ret_ty != TY_VOID) {
        v.code = new_tmp(cx);
        line(cx, ctype(f.ret_ty) + " " + v.code + "{};");
        call += ", &" + v.code;
    }
    check_abort(cx, call + ")");
    return v;
}

// ------------------------------------------------------------------
// statements (statement/*.rs)
// ------------------------------------------------------------------

void Translator::store(FnCtx *cx, const Val &ref, const Val &v, Loc l)
{
    if (!ref.is_ref || !ref.writable) fail(l, "cannot assign to this expression");
    Val c = convert(cx, v, ref.ty, l);
    line(cx, ref.lv + " = " + code(cx, c) + ";");
}

void Translator::declare_local(FnCtx *cx, const Stmt &s)
{
    if (cx->scopes.back().count(s.name)) fail(s.loc, "redeclaration of '" + s.name + "' in the same scope");
    if (s.name == "_") fail(s.loc, "'_' cannot be declared");
    Sym sym;
    if (s.k == Stmt::CONST) {
        Const c = eval_const(s.e);
        if (s.type) c = convert_const(c, resolve_type(s.type), s.loc);
        sym.k = Sym::CONST;
        sym.c = c;
        sym.ty = c.ty;
        cx->scopes.back()[s.name] = sym;
        return;
    }
    int ty = s.type ? resolve_type(s.type) : -1;
    Val init;
    bool has_init = (bool)s.e;
    if (has_init) {
        init = rvalue(cx, expr(cx, s.e), s.loc);
        if (init.ty == TY_VOID) fail(s.loc, "a void call has no value");
        if (ty < 0) ty = concretize(init.ty);
        init = convert(cx, init, ty, s.loc);
    }
    if (!constructible(ty)) fail(s.loc, "variables need a constructible type");
    sym.ty = ty;
    sym.cname = "v" + std::to_string(cx->tmp++) + "_" + s.name;
    if (s.k == Stmt::VAR) {
        sym.k = Sym::LOCAL_VAR;
        line(cx, ctype(ty) + " " + sym.cname + (has_init ? " = " + code(cx, init) : "{}") + ";");
    } else {
        sym.k = Sym::LOCAL_LET;
        line(cx, "const " + ctype(ty) + " " + sym.cname + " = " + code(cx, init) + ";");
    }
    cx->scopes.back()[s.name] = sym;
}

void Translator::block(FnCtx *cx, const std::vector<StmtP> &body, bool new_scope)
{
    if (new_scope) cx->scopes.emplace_back();
    for (auto &s : body) stmt(cx, *s);
    if (new_scope) cx->scopes.pop_back();
}

void Translator::stmt(FnCtx *cx, const Stmt &s)
{
    auto open = [&](const std::string &head) {
        line(cx, head + "{");
        cx->indent++;
    };
    auto close = [&](const std::string &tail = "}") {
        cx->indent--;
        line(cx, tail);
    };
    switch (s.k) {
    case Stmt::EMPTY:
        return;
    case Stmt::BLOCK:
        open("");
        block(cx, s.body);
        close();
        return;
    case Stmt::VAR:
    case Stmt::LET:
    case Stmt::CONST:
        declare_local(cx, s);
        return;
    case Stmt::ASSIGN: {
        Val target = expr(cx, s.e);
        Val v = rvalue(cx, expr(cx, s.e2), s.loc);
        store(cx, target, v, s.loc);
        return;
    }
    case Stmt::COMPOUND: {
        Val target = expr(cx, s.e);
        if (!target.is_ref || !target.writable) fail(s.loc, "cannot assign to this expression");
        Val cur = load(cx, target, s.loc);
        Val rhs = rvalue(cx, expr(cx, s.e2), s.loc);
        Val r;
        if (s.name == "<<" || s.name == ">>") {
            if (is_abstract_scalar(rhs.ty)) rhs = convert(cx, rhs, TY_U32, s.loc);
            if (!is_int_scalar(cur.ty) || rhs.ty != TY_U32) fail(s.loc, "shifts need an integer value and a u32 amount");
            r = tmpval(cx, cur.ty, scalar_op(cx, s.name, cur.ty, code(cx, cur), code(cx, rhs)));
        } else {
            r = arith(cx, s.name, cur, rhs, s.loc);
        }
        store(cx, target, r, s.loc);
        return;
    }
    case Stmt::INCR:
    case Stmt::DECR: {
        Val target = expr(cx, s.e);
        if (!target.is_ref || !target.writable) fail(s.loc, "cannot assign to this expression");
        if (target.ty != TY_I32 && target.ty != TY_U32) fail(s.loc, "++ and -- need an integer variable");
        Val cur = load(cx, target, s.loc);
        Val one = cvalue(Const{target.ty, 1, 0.0, {}});
        store(cx, target, arith(cx, s.k == Stmt::INCR ? "+" : "-", cur, one, s.loc), s.loc);
        return;
    }
    case Stmt::PHONY: {
        Val v = expr(cx, s.e);
        if (!v.is_ref && !v.is_const && v.handle < 0 && v.ty != TY_VOID) line(cx, "(void)" + v.code + ";");
        return;
    }
    case Stmt::CALL: {
        auto f = fn_by_name_.find(s.e->name);
        if (f == fn_by_name_.end()) fail(s.loc, "only user functions can be called as statements");
        (void)expr(cx, s.e);
        return;
    }
    case Stmt::IF: {
        Val c = rvalue(cx, expr(cx, s.e), s.loc);
        if (c.ty != TY_BOOL) fail(s.loc, "if condition must be bool");
        open("if (" + code(cx, c) + ") ");
        block(cx, s.body);
        if (s.alt.empty()) {
            close();
            return;
        }
        close("} else {");
        cx->indent++;
        if (s.alt[0]->k == Stmt::BLOCK) block(cx, s.alt[0]->body);
        else stmt(cx, *s.alt[0]);
        close();
        return;
    }
    case Stmt::SWITCH: {
        Val sel = rvalue(cx, expr(cx, s.e), s.loc);
        if (is_abstract_scalar(sel.ty)) sel = convert(cx, sel, TY_I32, s.loc);
        if (sel.ty != TY_I32 && sel.ty != TY_U32) fail(s.loc, "switch selector must be i32 or u32");
        std::set<int64_t> seen;
        bool has_default = false;
        open("switch (" + code(cx, sel) + ") ");
        cx->breakables.push_back(Breakable{false, "", false});
        for (auto &c : s.cases) {
            std::string labels;
            for (auto &ce : c.sel) {
                Const v = convert_const(eval_const(ce), sel.ty, ce->loc);
                if (!seen.insert(v.i).second) fail(ce->loc, "duplicate case value");
                labels += "case " + literal(v) + ": ";
            }
            if (c.dflt) {
                if (has_default) fail(c.loc, "more than one default clause");
                has_default = true;
                labels += "default: ";
            }
            line(cx, labels + "{");
            cx->indent++;
            block(cx, c.body);
            cx->indent--;
            line(cx, "} break;");
        }
        cx->breakables.pop_back();
        if (!has_default) fail(s.loc, "a switch needs a default clause");
        close();
        return;
    }
    case Stmt::LOOP:
    case Stmt::FOR:
    case Stmt::WHILE: {
        // loop.rs: body, then the continuing block (also the target of
        // 'continue'), then back to the body
        std::string label = "_cont" + std::to_string(cx->tmp++);
        cx->scopes.emplace_back();
        if (s.k == Stmt::FOR) {
            open("");
            if (s.init) stmt(cx, *s.init);
        }
        open("for (;;) ");
        if (s.k != Stmt::LOOP && s.e) {
            open("");
            Val c = rvalue(cx, expr(cx, s.e), s.loc);
            if (c.ty != TY_BOOL) fail(s.loc, "loop condition must be bool");
            line(cx, "if (!" + code(cx, c) + ") break;");
            close();
        }
        cx->breakables.push_back(Breakable{true, label, false});
        open("");
        block(cx, s.body);
        close();
        cx->breakables.pop_back();
        line(cx, label + ":;");
        if (s.k == Stmt::FOR && s.update) {
            open("");
            stmt(cx, *s.update);
            close();
        }
        if (s.k == Stmt::LOOP && s.has_continuing) {
            open("");
            bool saved = cx->in_continuing;
            cx->in_continuing = true;
            cx->scopes.emplace_back();
            for (auto &cs : s.alt) {
                if (cs->k == Stmt::BREAK_IF) {
                    Val c = rvalue(cx, expr(cx, cs->e), cs->loc);
                    if (c.ty != TY_BOOL) fail(cs->loc, "break-if condition must be bool");
                    line(cx, "if (" + code(cx, c) + ") break;");
                } else {
                    stmt(cx, *cs);
                }
            }
            cx->scopes.pop_back();
            cx->in_continuing = saved;
            close();
        }
        close();
        if (s.k == Stmt::FOR) close();
        cx->scopes.pop_back();
        return;
    }
    case Stmt::BREAK:
        if (cx->breakables.empty()) fail(s.loc, "break outside a loop or switch");
        line(cx, "break;");
        return;
    case Stmt::BREAK_IF:
        fail(s.loc, "'break if' must be the last statement of a continuing block");
    case Stmt::CONTINUE: {
        for (auto it = cx->breakables.rbegin(); it != cx->breakables.rend(); ++it)
            if (it->is_loop) {
                it->cont_used = true;
                line(cx, "goto " + it->cont_label + ";");
                return;
            }
        fail(s.loc, "continue outside a loop");
    }
    case Stmt::RETURN: {
        if (cx->in_continuing) fail(s.loc, "return is not allowed in a continuing block");
        int rt = cx->fn->ret_ty;
        if (s.e) {
            if (rt == TY_VOID) fail(s.loc, "this function returns no value");
            Val v = convert(cx, rvalue(cx, expr(cx, s.e), s.loc), rt, s.loc);
            line(cx, "*ret = " + code(cx, v) + ";");
        } else if (rt != TY_VOID) {
            fail(s.loc, "missing return value");
        }
        line(cx, "return 0u;");
        return;
    }
    case Stmt::DISCARD:
        cx->fn->discards = true;
        line(cx, "return wgsl::ABORT_KILL;");
        return;
    }
}

void Translator::function(FnInfo &f)
{
    FnCtx cx;
    cx.fn = &f;
    cx.scopes.emplace_back();
    const FnDecl &d = *f.decl;
    std::string sig = "template <class Env>\nWGSL_FN uint32_t " + f.cname + "(Env &env, Priv &pv";
    for (size_t k = 0; k < d.params.size(); ++k) {
        const Param &p = d.params[k];
        if (cx.scopes.back().count(p.name)) fail(p.loc, "duplicate parameter '" + p.name + "'");
        Sym s;
        s.k = Sym::PARAM;
        s.ty = f.param_ty[k];
        s.cname = "a" + std::to_string(k) + "_" + p.name;
        cx.scopes.back()[p.name] = s;
        sig += ", const " + ctype(s.ty) + " " + s.cname;
    }
    if (f.ret_ty != TY_VOID) sig += ", " + ctype(f.ret_ty) + " *ret";
    sig += ")";
    cx.scopes.emplace_back();
    block(&cx, d.body, false);
    f.code = sig + "\n{\n    (void)env;\n    (void)pv;\n" + cx.out + "    return 0u;\n}\n";
}

static bool io_scalar(int s) { return s == TY_F32 || s == TY_I32 || s == TY_U32; }

void Translator::entry_wrapper(FnInfo &f, EntryPoint &ep, std::string &out)
{
    const FnDecl &d = *f.decl;
    bool vs = f.stage == Stage::Vertex;
    std::string body;
    std::set<int> locs_in, locs_out;
    int tmp = 0;
    auto io_type = [&](int ty, Loc l, IoVar &v) {
        int s = scalar_of(ty);
        if (s < 0 || !io_scalar(s) || ti(ty).k == TK_MAT)
            fail(l, "stage inputs and outputs must be scalars or vectors of f32, i32 or u32");
        v.ncomp = (uint32_t)lanes(ty);
        v.scalar = s == TY_F32 ? 'f' : s == TY_I32 ? 'i' : 'u';
    };
    // one input value: returns the C++ statements that fill `dst` of type `ty`
    auto input = [&](const IoAttrs &io, int ty, const std::string &dst, Loc l) {
        if (io.builtin != BI_NONE) {
            if (vs && io.builtin == BI_VERTEX_INDEX && ty == TY_U32) {
                ep.vertex_index = true;
                body += "    " + dst + " = env.vertex_index();\n";
            } else if (vs && io.builtin == BI_INSTANCE_INDEX && ty == TY_U32) {
                ep.instance_index = true;
                body += "    " + dst + " = env.instance_index();\n";
            } else if (!vs && io.builtin == BI_FRONT_FACING && ty == TY_BOOL) {
                ep.front_facing = true;
                body += "    " + dst + " = env.front_facing();\n";
            } else if (!vs && io.builtin == BI_POSITION && ty == vec_of(TY_F32, 4)) {
                ep.position_in = true;
                body += "    env.position(" + dst + ".v);\n";
            } else if (!vs && io.builtin == BI_PRIMITIVE_INDEX && ty == TY_U32) {
                // the primitive's index in its instance's assembly order,
                // culled primitives counted (state.rs:536-541)
                ep.primitive_index = true;
                body += "    " + dst + " = env.primitive_index();\n";
            } else if (!vs && io.builtin == BI_SAMPLE_INDEX && ty == TY_U32) {
                body += "    " + dst + " = 0u;\n";            // fragment.rs:138-140: single sample
            } else if (!vs && io.builtin == BI_SAMPLE_MASK && ty == TY_U32) {
                body += "    " + dst + " = 0xFFFFFFFFu;\n";   // fragment.rs:141: !0
            } else {
                fail(l, "unsupported input @builtin(" + io.builtin_name + ") of type " + tname(ty) + " in this stage");
            }
            return;
        }
        if (io.location < 0) fail(l, "entry-point inputs need @location or @builtin");
        if (!locs_in.insert(io.location).second) fail(l, "duplicate input @location(" + std::to_string(io.location) + ")");
        IoVar v;
        v.location = (uint32_t)io.location;
        io_type(ty, l, v);
        v.interp = io.has_interp ? io.interp : IO_PERSPECTIVE;
        if (!vs && v.scalar != 'f' && v.interp != IO_FLAT)
            fail(l, "integer inputs must use @interpolate(flat) (reference: fragment.rs:335-337)");
        ep.inputs.push_back(v);
        std::string w = "_w" + std::to_string(tmp++);
        body += "    { uint32_t " + w + "[" + std::to_string(v.ncomp) + "]; env.in_location(" + std::to_string(v.location) +
                ", " + w + ", " + std::to_string(v.ncomp) + "u);";
        for (uint32_t k = 0; k < v.ncomp; ++k)
            body += " wgsl::from_word(" + w + "[" + std::to_string(k) + "], " + lane(dst, ty, (int)k) + ");";
        body += " }\n";
    };
    std::string args;
    for (size_t k = 0; k < d.params.size(); ++k) {
        const Param &p = d.params[k];
        int ty = f.param_ty[k];
        std::string a = "_a" + std::to_string(k);
        body += "    " + ctype(ty) + " " + a + "{};\n";
        IoAttrs io = io_attrs(p.attrs);
        if (ti(ty).k == TK_STRUCT && io.builtin == BI_NONE && io.location < 0) {
            for (auto &mb : structs_[ti(ty).sid].members) input(mb.io, mb.ty, a + "." + mb.cname, p.loc);
        } else {
            input(io, ty, a, p.loc);
        }
        args += ", " + a;
    }
    // outputs
    std::string outs;
    auto output = [&](const IoAttrs &io, int ty, const std::string &src, Loc l) {
        if (io.builtin != BI_NONE) {
            if (vs && io.builtin == BI_POSITION && ty == vec_of(TY_F32, 4)) {
                ep.position_out = true;
                outs += "    env.out_position(" + src + ".v);\n";
            } else if (!vs && io.builtin == BI_FRAG_DEPTH && ty == TY_F32) {
                ep.frag_depth = true;
                // outputs are read in member order (runtime.rs:1500-1508):
                // a depth written before the first location is the one the
                // late test sees
                if (ep.outputs.empty()) ep.frag_depth_tested = true;
                outs += "    env.out_frag_depth(" + src + ");\n";
            } else if (!vs && io.builtin == BI_SAMPLE_MASK && ty == TY_U32) {
                outs += "    (void)" + src + ";\n";   // stored and never used (fragment.rs:464-466)
            } else {
                fail(l, "unsupported output @builtin(" + io.builtin_name + ") of type " + tname(ty) + " in this stage");
            }
            return;
        }
        if (io.location < 0) fail(l, "entry-point outputs need @location or @builtin");
        if (!locs_out.insert(io.location).second) fail(l, "duplicate output @location(" + std::to_string(io.location) + ")");
        IoVar v;
        v.location = (uint32_t)io.location;
        io_type(ty, l, v);
        v.interp = io.has_interp ? io.interp : IO_PERSPECTIVE;
        ep.outputs.push_back(v);
        std::string w = "_w" + std::to_string(tmp++);
        outs += "    { uint32_t " + w + "[" + std::to_string(v.ncomp) + "];";
        for (uint32_t k = 0; k < v.ncomp; ++k)
            outs += " " + w + "[" + std::to_string(k) + "] = wgsl::to_word(" + lane(src, ty, (int)k) + ");";
        outs += " env.out_location(" + std::to_string(v.location) + ", " + w + ", " + std::to_string(v.ncomp) + "u); }\n";
    };
    if (f.ret_ty != TY_VOID) {
        body += "    " + ctype(f.ret_ty) + " _r{};\n";
        args += ", &_r";
        IoAttrs io = io_attrs(d.ret_attrs);
        if (ti(f.ret_ty).k == TK_STRUCT && io.builtin == BI_NONE && io.location < 0) {
            for (auto &mb : structs_[ti(f.ret_ty).sid].members) output(mb.io, mb.ty, "_r." + mb.cname, d.loc);
        } else {
            output(io, f.ret_ty, "_r", d.loc);
        }
    }
    if (vs && !ep.position_out) fail(d.loc, "a vertex entry point must output @builtin(position)");
    auto by_loc = [](const IoVar &a, const IoVar &b) { return a.location < b.location; };
    std::sort(ep.inputs.begin(), ep.inputs.end(), by_loc);
    std::sort(ep.outputs.begin(), ep.outputs.end(), by_loc);
    ep.symbol = "ep_" + d.name;
    out += "template <class Env>\nWGSL_FN uint32_t " + ep.symbol + "(Env &env)\n{\n    Priv pv{};\n    priv_init(pv);\n" +
           body + "    if (uint32_t _rc = " + f.cname + "<Env>(env, pv" + args + ")) return _rc;\n" + outs +
           "    return 0u;\n}\n";
}

void Translator::finish_reachability(EntryPoint &ep, int fi)
{
    std::set<int> seen;
    std::set<uint32_t> slots;
    bool discard = false;
    std::function<void(int)> walk = [&](int i) {
        if (!seen.insert(i).second) return;
        for (int s : fns_[i].slots) slots.insert((uint32_t)s);
        discard = discard || fns_[i].discards;
        for (int c : fns_[i].calls) walk(c);
    };
    walk(fi);
    ep.resources.assign(slots.begin(), slots.end());
    ep.may_discard = discard;
    ep.early_depth = fns_[fi].early_depth;
    if (discard && ep.stage != Stage::Fragment) fail(fns_[fi].decl->loc, "discard is only valid in fragment shaders");
    if (ep.early_depth && ep.stage != Stage::Fragment)
        fail(fns_[fi].decl->loc, "@early_depth_test is only valid on fragment entry points");
}

void Translator::run(Module &out)
{
    declare_structs();
    declare_globals();
    declare_functions();
    for (auto &f : fns_) function(f);
    // recursion is not allowed
    std::function<void(int)> visit = [&](int i) {
        if (fns_[i].state == 2) return;
        if (fns_[i].state == 1) fail(fns_[i].decl->loc, "recursion is not allowed ('" + fns_[i].decl->name + "')");
        fns_[i].state = 1;
        for (int c : fns_[i].calls) visit(c);
        fns_[i].state = 2;
    };
    for (size_t i = 0; i < fns_.size(); ++i) visit((int)i);
    // structs, in member dependency order
    std::string structs;
    std::function<void(int)> emit_struct = [&](int si) {
        StructInfo &s = structs_[si];
        if (s.emitted) return;
        s.emitted = true;
        for (auto &mb : s.members) {
            int t = mb.ty;
            while (ti(t).k == TK_ARR || ti(t).k == TK_RTARR) t = ti(t).elem;
            if (ti(t).k == TK_STRUCT) emit_struct(ti(t).sid);
        }
        structs += "struct " + s.cname + " {\n";
        for (auto &mb : s.members)
            if (ti(mb.ty).k != TK_RTARR) structs += "    " + ctype(mb.ty) + " " + mb.cname + ";\n";
        structs += "};\n";
    };
    for (size_t i = 0; i < structs_.size(); ++i) emit_struct((int)i);
    std::string code = structs;
    code += "struct Priv {\n" + priv_types_ + "};\n";
    code += "WGSL_FN void priv_init(Priv &pv)\n{\n    (void)pv;\n";
    for (auto &p : privs_) code += "    pv." + p.first + " = " + p.second + ";\n";
    code += "}\n";
    for (int ty : loader_order_) code += loader_code_[ty];
    for (auto &f : fns_) {
        std::string decl = f.code.substr(0, f.code.find("\n{"));
        code += decl + ";\n";
    }
    for (auto &f : fns_) code += f.code;
    int n_entries = 0;
    for (size_t i = 0; i < fns_.size(); ++i) {
        FnInfo &f = fns_[i];
        if (!f.is_entry) continue;
        EntryPoint ep;
        ep.name = f.decl->name;
        ep.stage = f.stage;
        entry_wrapper(f, ep, code);
        finish_reachability(ep, (int)i);
        out.entries.push_back(ep);
        ++n_entries;
    }
    out.code = code;
    out.resources = resources_;
}

}  // namespace

bool translate(const std::string &src, Module &out, std::string &err)
{
    try {
        Parser p(lex(src));
        ModuleAst m = p.parse_module();
        Translator t(m);
        out = Module{};
        t.run(out);
        return true;
    } catch (const Error &e) {
        err = e.what();
        return false;
    }
}

}  // namespace wgsl
}  // namespace wgc

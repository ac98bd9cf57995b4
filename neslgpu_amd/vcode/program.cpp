// program.cpp -- VCODE opcode table, lexer, parser and linker.
//
// The language is the one accepted by the reference interpreter: tokens as
// in nesl/vcode/tokens.flex:27-107 (identifiers over [-+*/%<>=!._A-Za-z0-9],
// integers, reals, "strings", { comments }, parentheses, I / O markers) and
// the productions of nesl/vcode/grammar.yy:140-472.  The parser is a
// hand-written recursive descent over that grammar; IF / ELSE / ENDIF and
// CALL targets are resolved as in nesl/vcode/actions.c:291-398 (an IF
// branches past its ELSE, an ELSE jumps to the ENDIF, ENDIF emits nothing).
#include "vinterp.hpp"

#include <cctype>
#include <cstdlib>
#include <cstring>
#include <unordered_map>

namespace vcode {

const char *type_name(int t) {
    switch (t) {
        case T_Pair: return "PAIR";
        case T_None: return "NONE";
        case T_Bool: return "BOOL";
        case T_Int: return "INT";
        case T_Float: return "FLOAT";
        case T_Segdes: return "SEGDES";
        case T_String: return "CHAR";
        default: return "ILLEGAL";
    }
}

// ------------------------------------------------------------------ opcodes
// Arguments, results and CVL call shape of every opcode, restated from the
// opcode descriptors of nesl/vcode/vcode_table.c:32-684 and the CVL
// function table nesl/vcode/cvl_table.c:9-292 (function family per opcode).
#define G T_Given
#define B T_Bool
#define I T_Int
#define F T_Float
#define S T_Segdes
#define A1 T_Agree1
#define A2 T_Agree2
#define EW(op, nm, syn, fx, na, a0, a1, a2, res, d, cvl)                                          \
    {op, nm, syn, fx, na, 1, {a0, a1, a2, 0, 0, 0}, {res, 0}, {L_AGREE1, 0}, d, cvl, {1, 0}}
#define UN(op, nm, arg, res, cvl) EW(op, nm, TS_NONE, arg, 1, arg, 0, 0, res, D_Elwise0 + 1, cvl)

static const OpInfo kOps[] = {
    EW(OP_PLUS, "+", TS_1, 0, 2, G, G, 0, A1, D_Elwise0 + 2, "add_wu"),
    EW(OP_MINUS, "-", TS_1, 0, 2, G, G, 0, A1, D_Elwise0 + 2, "sub_wu"),
    EW(OP_TIMES, "*", TS_1, 0, 2, G, G, 0, A1, D_Elwise0 + 2, "mul_wu"),
    EW(OP_DIV, "/", TS_1, 0, 2, G, G, 0, A1, D_Elwise0 + 2, "div_wu"),
    EW(OP_MOD, "%", TS_NONE, I, 2, I, I, 0, I, D_Elwise0 + 2, "mod_wu"),
    EW(OP_LT, "<", TS_1, 0, 2, G, G, 0, B, D_Elwise0 + 2, "les_wu"),
    EW(OP_LEQ, "<=", TS_1, 0, 2, G, G, 0, B, D_Elwise0 + 2, "leq_wu"),
    EW(OP_GT, ">", TS_1, 0, 2, G, G, 0, B, D_Elwise0 + 2, "grt_wu"),
    EW(OP_GEQ, ">=", TS_1, 0, 2, G, G, 0, B, D_Elwise0 + 2, "geq_wu"),
    EW(OP_EQ, "=", TS_1, 0, 2, G, G, 0, B, D_Elwise0 + 2, "eql_wu"),
    EW(OP_NEQ, "!=", TS_1, 0, 2, G, G, 0, B, D_Elwise0 + 2, "neq_wu"),
    EW(OP_LSHIFT, "LSHIFT", TS_NONE, I, 2, I, I, 0, I, D_Elwise0 + 2, "lsh_wu"),
    EW(OP_RSHIFT, "RSHIFT", TS_NONE, I, 2, I, I, 0, I, D_Elwise0 + 2, "rsh_wu"),
    EW(OP_NOT, "NOT", TS_2, 0, 1, G, 0, 0, A1, D_Elwise0 + 1, "not_wu"),
    EW(OP_AND, "AND", TS_2, 0, 2, G, G, 0, A1, D_Elwise0 + 2, "and_wu"),
    EW(OP_OR, "OR", TS_2, 0, 2, G, G, 0, A1, D_Elwise0 + 2, "ior_wu"),
    EW(OP_XOR, "XOR", TS_2, 0, 2, G, G, 0, A1, D_Elwise0 + 2, "xor_wu"),
    EW(OP_SELECT, "SELECT", TS_3, 0, 3, B, G, G, A2, D_Elwise0 + 3, "sel_wu"),
    UN(OP_RAND, "RAND", I, I, "rnd_wu"),
    UN(OP_FLOOR, "FLOOR", F, I, "flr_wu"),
    UN(OP_CEIL, "CEIL", F, I, "cei_wu"),
    UN(OP_TRUNC, "TRUNC", F, I, "trn_wu"),
    UN(OP_ROUND, "ROUND", F, I, "rou_wu"),
    UN(OP_LOG, "LOG", F, F, "log_wu"),
    UN(OP_SQRT, "SQRT", F, F, "sqt_wu"),
    UN(OP_EXP, "EXP", F, F, "exp_wu"),
    UN(OP_SIN, "SIN", F, F, "sin_wu"),
    UN(OP_COS, "COS", F, F, "cos_wu"),
    UN(OP_TAN, "TAN", F, F, "tan_wu"),
    UN(OP_ASIN, "ASIN", F, F, "asn_wu"),
    UN(OP_ACOS, "ACOS", F, F, "acs_wu"),
    UN(OP_ATAN, "ATAN", F, F, "atn_wu"),
    UN(OP_SINH, "SINH", F, F, "snh_wu"),
    UN(OP_COSH, "COSH", F, F, "csh_wu"),
    UN(OP_TANH, "TANH", F, F, "tnh_wu"),
    UN(OP_I_TO_F, "I_TO_F", I, F, "dbl_wu"),
    UN(OP_I_TO_B, "I_TO_B", I, B, "boo_wu"),
    UN(OP_B_TO_I, "B_TO_I", B, I, "int_wu"),
#define SC(op, nm, syn, cvl) {op, nm, syn, 0, 2, 1, {G, S}, {A1, 0}, {L_AGREE1, 0}, D_Scan, cvl, {2, 0}}
#define RD(op, nm, syn, cvl) {op, nm, syn, 0, 2, 1, {G, S}, {A1, 0}, {L_AGREE2, 0}, D_Reduce, cvl, {2, 0}}
    SC(OP_PLUS_SCAN, "+_SCAN", TS_1, "add_se"),
    SC(OP_MULT_SCAN, "*_SCAN", TS_1, "mul_se"),
    SC(OP_MAX_SCAN, "MAX_SCAN", TS_1, "max_se"),
    SC(OP_MIN_SCAN, "MIN_SCAN", TS_1, "min_se"),
    SC(OP_AND_SCAN, "AND_SCAN", TS_2, "and_se"),
    SC(OP_OR_SCAN, "OR_SCAN", TS_2, "ior_se"),
    SC(OP_XOR_SCAN, "XOR_SCAN", TS_2, "xor_se"),
    RD(OP_PLUS_REDUCE, "+_REDUCE", TS_1, "add_re"),
    RD(OP_MULT_REDUCE, "*_REDUCE", TS_1, "mul_re"),
    RD(OP_MAX_REDUCE, "MAX_REDUCE", TS_1, "max_re"),
    RD(OP_MIN_REDUCE, "MIN_REDUCE", TS_1, "min_re"),
    RD(OP_AND_REDUCE, "AND_REDUCE", TS_2, "and_re"),
    RD(OP_OR_REDUCE, "OR_REDUCE", TS_2, "ior_re"),
    RD(OP_XOR_REDUCE, "XOR_REDUCE", TS_2, "xor_re"),
    {OP_PERMUTE, "PERMUTE", TS_3, 0, 3, 1, {G, I, S}, {A1, 0}, {L_AGREE1, 0}, D_Permute, "smp_pe", {3, 0}},
    {OP_DPERMUTE, "DPERMUTE", TS_3, 0, 5, 1, {G, I, G, S, S}, {A1, 0}, {L_AGREE3, 0}, D_Dpermute, "dpe_pe", {4, 5}},
    {OP_FPERMUTE, "FPERMUTE", TS_3, 0, 5, 1, {G, I, B, S, S}, {A1, 0}, {L_COMPAT5, 0}, D_Fpermute, "fpm_pe", {4, 5}},
    {OP_BPERMUTE, "BPERMUTE", TS_3, 0, 4, 1, {G, I, S, S}, {A1, 0}, {L_AGREE2, 0}, D_Bpermute, "bck_pe", {3, 4}},
    {OP_BFPERMUTE, "BFPERMUTE", TS_3, 0, 5, 1, {G, I, B, S, S}, {A1, 0}, {L_COMPAT5, 0}, D_Bfpermute, "bfp_pe", {4, 5}},
    {OP_DFPERMUTE, "DFPERMUTE", TS_3, 0, 6, 1, {G, I, B, G, S, S}, {A1, 0}, {L_AGREE4, 0}, D_Dfpermute, "dfp_pe", {5, 6}},
    {OP_EXTRACT, "EXTRACT", TS_3, 0, 3, 1, {G, I, S}, {A1, 0}, {L_AGREE2, 0}, D_Extract, "ext_ve", {3, 0}},
    {OP_REPLACE, "REPLACE", TS_3, 0, 4, 1, {G, I, G, S}, {A1, 0}, {L_AGREE1, 0}, D_Replace, "rep_ve", {4, 0}},
    {OP_DIST, "DIST", TS_3, 0, 2, 1, {G, S}, {A1, 0}, {L_COMPAT2, 0}, D_Dist, "dis_ve", {2, 0}},
    {OP_INDEX, "INDEX", TS_NONE, I, 3, 1, {I, I, S}, {I, 0}, {L_COMPAT3, 0}, D_Index, "ind_le", {3, 0}},
    {OP_RANK_UP, "RANK_UP", TS_1, 0, 2, 1, {G, S}, {I, 0}, {L_AGREE1, 0}, D_Scan, "rku_le", {2, 0}},
    {OP_RANK_DOWN, "RANK_DOWN", TS_1, 0, 2, 1, {G, S}, {I, 0}, {L_AGREE1, 0}, D_Scan, "rkd_le", {2, 0}},
    {OP_PACK, "PACK", TS_3, 0, 3, 2, {G, B, S}, {A1, S}, {L_AGREE1, L_AGREE2}, D_Pack, "pk2_le", {3, 0}},
    {OP_LENGTH, "LENGTH", TS_3, 0, 1, 1, {G}, {I, 0}, {L_SCALAR, 0}, D_SegOp, "", {0, 0}},
    {OP_MAKE_SEGDES, "MAKE_SEGDES", TS_NONE, S, 1, 1, {I}, {S, 0}, {L_AGREE1, 0}, D_SegOp, "", {0, 0}},
    {OP_LENGTHS, "LENGTHS", TS_NONE, I, 1, 1, {S}, {I, 0}, {L_AGREE1, 0}, D_SegOp, "", {0, 0}},
#define CT(op, nm, syn) {op, nm, syn, T_Int, 0, 0, {0}, {0, 0}, {0, 0}, D_Special, "", {0, 0}}
    CT(OP_COPY, "COPY", TS_NONE),
    CT(OP_POP, "POP", TS_NONE),
    CT(OP_CPOP, "CPOP", TS_NONE),
    CT(OP_PAIR, "PAIR", TS_NONE),
    CT(OP_UNPAIR, "UNPAIR", TS_NONE),
    CT(OP_CALL, "CALL", TS_NONE),
    CT(OP_RET, "RET", TS_NONE),
    CT(OP_IF, "IF", TS_NONE),
    CT(OP_ELSE, "ELSE", TS_NONE),
    CT(OP_CONST, "CONST", TS_NONE),
    CT(OP_EXIT, "EXIT", TS_NONE),
    CT(OP_START_TIMER, "START_TIMER", TS_NONE),
    CT(OP_STOP_TIMER, "STOP_TIMER", TS_NONE),
    CT(OP_READ, "READ", TS_5),
    CT(OP_WRITE, "WRITE", TS_5),
    CT(OP_FOPEN, "FOPEN", TS_NONE),
    CT(OP_FCLOSE, "FCLOSE", TS_NONE),
    CT(OP_FREAD, "FREAD", TS_5),
    CT(OP_FREAD_CHAR, "FREAD_CHAR", TS_NONE),
    CT(OP_FWRITE, "FWRITE", TS_5),
    CT(OP_SPAWN, "SPAWN", TS_NONE),
    CT(OP_SRAND, "SRAND", TS_INT),
    CT(OP_FUSED, "FUSED", TS_NONE),
};
#undef G
#undef B
#undef I
#undef F
#undef S
#undef A1
#undef A2
static_assert(sizeof(kOps) / sizeof(kOps[0]) == OP_COUNT, "opcode table incomplete");

const OpInfo &op_info(Op op) { return kOps[op]; }

// ------------------------------------------------------------------ keywords
enum KwKind { K_OP, K_TYPE, K_BOOL, K_FILE, K_FUNC, K_MAIN, K_ENDIF };

bool lookup_keyword(const std::string &s, int *kind, int *value) {
    static std::unordered_map<std::string, std::pair<int, int>> table;
    if (table.empty()) {
        for (int i = 0; i < OP_COUNT; ++i) table[kOps[i].name] = {K_OP, i};
        table["INT"] = {K_TYPE, T_Int};
        table["BOOL"] = {K_TYPE, T_Bool};
        table["FLOAT"] = {K_TYPE, T_Float};
        table["SEGDES"] = {K_TYPE, T_Segdes};
        table["CHAR"] = {K_TYPE, T_String};
        table["T"] = {K_BOOL, 1};
        table["F"] = {K_BOOL, 0};
        table["NULL_STREAM"] = {K_FILE, -1};
        table["STDIN"] = {K_FILE, 0};
        table["STDOUT"] = {K_FILE, 1};
        table["STDERR"] = {K_FILE, 2};
        table["FUNC"] = {K_FUNC, 0};
        table["MAIN"] = {K_MAIN, 0};
        table["ENDIF"] = {K_ENDIF, 0};
    }
    auto it = table.find(s);
    if (it == table.end()) return false;
    *kind = it->second.first;
    *value = it->second.second;
    return true;
}

// ------------------------------------------------------------------ lexer
enum TokKind { TK_EOF, TK_INT, TK_REAL, TK_STRING, TK_LPAREN, TK_RPAREN, TK_IN, TK_OUT, TK_WORD };

struct Token {
    TokKind kind = TK_EOF;
    std::string text;
    long long ival = 0;
    double dval = 0.0;
    int line = 1;
};

static bool id_char(int c) {
    return std::isalnum(c) || std::strchr("-+*/%<>=!._", c) != nullptr;
}

// [-+]?digit+
static bool full_integer(const std::string &s) {
    size_t i = 0;
    if (i < s.size() && (s[i] == '-' || s[i] == '+')) ++i;
    size_t d = i;
    while (i < s.size() && std::isdigit((unsigned char)s[i])) ++i;
    return i > d && i == s.size();
}

// [-+]?digit+ . digit* ([Ee][-+]?digit+)?
static bool full_real(const std::string &s) {
    size_t i = 0;
    if (i < s.size() && (s[i] == '-' || s[i] == '+')) ++i;
    size_t d = i;
    while (i < s.size() && std::isdigit((unsigned char)s[i])) ++i;
    if (i == d || i >= s.size() || s[i] != '.') return false;
    ++i;
    while (i < s.size() && std::isdigit((unsigned char)s[i])) ++i;
    if (i < s.size() && (s[i] == 'e' || s[i] == 'E')) {
        ++i;
        if (i < s.size() && (s[i] == '-' || s[i] == '+')) ++i;
        size_t e = i;
        while (i < s.size() && std::isdigit((unsigned char)s[i])) ++i;
        if (i == e) return false;
    }
    return i == s.size();
}

class Lexer {
  public:
    explicit Lexer(const std::string &t) : text_(t) {}

    Token next() {
        for (;;) {
            if (pos_ >= text_.size()) return mk(TK_EOF, "");
            char c = text_[pos_];
            if (c == '\n') { ++line_; ++pos_; continue; }
            if (c == ' ' || c == '\t' || c == '\r' || c == '\f' || c == '\v') { ++pos_; continue; }
            if (c == '{') {                              // comment, may span lines
                while (pos_ < text_.size() && text_[pos_] != '}') {
                    if (text_[pos_] == '\n') ++line_;
                    ++pos_;
                }
                if (pos_ < text_.size()) ++pos_;
                continue;
            }
            if (c == '(') { ++pos_; return mk(TK_LPAREN, "("); }
            if (c == ')') { ++pos_; return mk(TK_RPAREN, ")"); }
            if (c == '"') {
                size_t e = text_.find('"', pos_ + 1);
                if (e == std::string::npos) {
                    std::fprintf(stderr, "vinterp: lexer: unterminated string at line %d\n", line_);
                    pos_ = text_.size();
                    return mk(TK_EOF, "");
                }
                std::string s = text_.substr(pos_ + 1, e - pos_ - 1);
                for (char ch : s) if (ch == '\n') ++line_;
                pos_ = e + 1;
                return mk(TK_STRING, s);
            }
            if (id_char((unsigned char)c)) {
                size_t b = pos_;
                while (pos_ < text_.size() && id_char((unsigned char)text_[pos_])) ++pos_;
                std::string w = text_.substr(b, pos_ - b);
                if (full_integer(w)) {
                    Token t = mk(TK_INT, w);
                    t.ival = std::strtoll(w.c_str(), nullptr, 10);
                    return t;
                }
                if (full_real(w)) {
                    Token t = mk(TK_REAL, w);
                    t.dval = std::strtod(w.c_str(), nullptr);
                    return t;
                }
                if (w == "I") return mk(TK_IN, w);
                if (w == "O") return mk(TK_OUT, w);
                return mk(TK_WORD, w);
            }
            std::fprintf(stderr, "vinterp: lexer: Unexpected character in input at line %d\n", line_);
            ++pos_;
        }
    }

  private:
    Token mk(TokKind k, const std::string &s) {
        Token t;
        t.kind = k;
        t.text = s;
        t.line = line_;
        return t;
    }
    const std::string &text_;
    size_t pos_ = 0;
    int line_ = 1;
};

// ------------------------------------------------------------------ parser
class Parser {
  public:
    Parser(const std::string &text, Program &p, bool lex_trace, bool link_trace)
        : lex_(text), prog_(p), lex_trace_(lex_trace), link_trace_(link_trace) {
        advance();
    }

    bool parse() {
        if (tok_.kind == TK_EOF) return fail("empty program");
        while (tok_.kind != TK_EOF) {
            if (!function()) return false;
        }
        return link();
    }

  private:
    void advance() {
        tok_ = lex_.next();
        if (lex_trace_) {
            static const char *kind[] = {"EOF", "INT", "REAL", "STRING", "(", ")", "I", "O", "WORD"};
            std::fprintf(stderr, "lex: line %d %s %s\n", tok_.line, kind[tok_.kind], tok_.text.c_str());
        }
    }

    bool fail(const std::string &msg) {
        std::fprintf(stderr, "vinterp: line %d: parse error: %s near \"%s\"\n", tok_.line, msg.c_str(),
                     tok_.text.c_str());
        return false;
    }

    bool word_is(int kind, int *value = nullptr) {
        if (tok_.kind != TK_WORD) return false;
        int k, v;
        if (!lookup_keyword(tok_.text, &k, &v) || k != kind) return false;
        if (value) *value = v;
        return true;
    }

    bool function() {
        if (!word_is(K_FUNC)) return fail("expected FUNC");
        advance();
        std::string name;
        if (word_is(K_MAIN)) {
            name = "MAIN";
            if (main_seen_) return fail("attempt to redefine main");
            main_seen_ = true;
            prog_.entry = "MAIN";
        } else if (tok_.kind == TK_WORD) {
            name = tok_.text;
            if (prog_.entry.empty()) prog_.entry = name;
        } else {
            return fail("expected a function name");
        }
        if (prog_.labels.count(name))
            std::fprintf(stderr, "vinterp: line %d: warning: label %s being redefined.\n", tok_.line, name.c_str());
        prog_.labels[name] = (int)prog_.code.size();
        advance();
        if (tok_.kind == TK_IN) {                    // I types... O types...
            advance();
            while (word_is(K_TYPE)) advance();
            if (tok_.kind != TK_OUT) return fail("expected O after the input types");
            advance();
            while (word_is(K_TYPE)) advance();
        }
        ifs_.clear();
        for (;;) {
            if (tok_.kind == TK_EOF) return fail("function without RET");
            int v;
            if (word_is(K_OP, &v) && v == OP_RET) {
                if (!ifs_.empty()) return fail("RET inside IF...ENDIF");
                emit(OP_RET, T_Int);
                advance();
                return true;
            }
            if (!statement()) return false;
        }
    }

    Instr &emit(Op op, int type) {
        Instr in;
        in.op = op;
        in.type = type;
        in.line = tok_.line;
        prog_.code.push_back(in);
        return prog_.code.back();
    }

    bool expect_int(long long *v) {
        if (tok_.kind != TK_INT) return fail("expected an integer");
        *v = tok_.ival;
        advance();
        return true;
    }

    bool parse_type(TypeSyntax syn, int *type) {
        int t;
        if (!word_is(K_TYPE, &t)) return fail("expected a type");
        bool ok = false;
        switch (syn) {
            case TS_1: ok = (t == T_Int || t == T_Float); break;
            case TS_2: ok = (t == T_Int || t == T_Bool); break;
            case TS_3: ok = (t == T_Int || t == T_Bool || t == T_Float); break;
            case TS_5: ok = true; break;
            case TS_INT: ok = (t == T_Int); break;
            default: break;
        }
        if (!ok) return fail("type not allowed here");
        *type = t;
        advance();
        return true;
    }

    bool statement() {
        int kind, v;
        if (word_is(K_ENDIF)) {
            if (ifs_.empty()) return fail("improper IF nesting");
            prog_.code[ifs_.back()].branch = (int)prog_.code.size();
            ifs_.pop_back();
            advance();
            return true;
        }
        if (tok_.kind != TK_WORD || !lookup_keyword(tok_.text, &kind, &v) || kind != K_OP)
            return fail("expected an instruction");
        const Op op = (Op)v;
        const OpInfo &oi = kOps[op];
        const int line = tok_.line;
        advance();
        switch (op) {
            case OP_COPY: case OP_POP: case OP_CPOP: {
                long long a, b;
                if (!expect_int(&a) || !expect_int(&b)) return false;
                if (a < 0 || b < 0) return fail("negative stack argument");
                Instr &in = emit(op, T_Int);
                in.line = line;
                in.a = (int)a;
                in.b = (int)b;
                return true;
            }
            case OP_CALL: {
                std::string target;
                if (word_is(K_MAIN)) target = "MAIN";
                else if (tok_.kind == TK_WORD) target = tok_.text;
                else return fail("expected a function name after CALL");
                advance();
                Instr &in = emit(op, T_Int);
                in.line = line;
                in.label = target;
                return true;
            }
            case OP_IF: {
                ifs_.push_back((int)prog_.code.size());
                emit(op, T_Int).line = line;
                return true;
            }
            case OP_ELSE: {
                if (ifs_.empty()) return fail("ELSE without IF");
                prog_.code[ifs_.back()].branch = (int)prog_.code.size() + 1;
                ifs_.back() = (int)prog_.code.size();
                emit(op, T_Int).line = line;
                return true;
            }
            case OP_CONST: return constant(line);
            case OP_FUSED: {
                long long n;
                if (!expect_int(&n)) return false;
                if (n < 0) return fail("negative FUSED number");
                Instr &in = emit(op, T_Int);
                in.line = line;
                in.a = (int)n;
                prog_.uses_fused = true;
                return true;
            }
            default: break;
        }
        int type = oi.fixed_type;
        if (oi.syntax != TS_NONE && !parse_type(oi.syntax, &type)) return false;
        if (type == T_Illegal) type = T_Int;
        emit(op, type).line = line;
        return true;
    }

    bool constant(int line) {
        int t;
        if (!word_is(K_TYPE, &t)) return fail("expected a type after CONST");
        advance();
        if (t == T_String) {
            if (tok_.kind != TK_STRING) return fail("expected a string after CONST CHAR");
            ConstVec cv;
            cv.type = T_String;
            for (unsigned char ch : tok_.text) cv.ints.push_back(ch);
            advance();
            return pooled(line, std::move(cv));
        }
        if (tok_.kind == TK_LPAREN) {                // vector constant
            advance();
            ConstVec cv;
            cv.type = t;
            for (;;) {
                int b = 0;
                if (tok_.kind == TK_RPAREN) { advance(); break; }
                if (t == T_Float && (tok_.kind == TK_REAL || tok_.kind == TK_INT)) {
                    cv.dbls.push_back(tok_.kind == TK_REAL ? tok_.dval : (double)tok_.ival);
                } else if ((t == T_Int || t == T_Segdes) && tok_.kind == TK_INT) {
                    cv.ints.push_back(tok_.ival);
                } else if (t == T_Bool && word_is(K_BOOL, &b)) {
                    cv.ints.push_back(b);
                } else {
                    return fail("illegal type mixing in constant");
                }
                advance();
            }
            return pooled(line, std::move(cv));
        }
        Instr in;
        in.op = OP_CONST;
        in.type = t;
        in.line = line;
        if (t == T_Int) {
            int fd;
            if (tok_.kind == TK_INT) in.ival = tok_.ival;
            else if (word_is(K_FILE, &fd)) in.ival = fd;
            else return fail("expected an integer constant");
        } else if (t == T_Float) {
            if (tok_.kind == TK_REAL) in.dval = tok_.dval;
            else if (tok_.kind == TK_INT) in.dval = (double)tok_.ival;
            else return fail("expected a float constant");
        } else if (t == T_Bool) {
            int b;
            if (!word_is(K_BOOL, &b)) return fail("expected T or F");
            in.ival = b;
        } else if (t == T_Segdes) {                  // one segment of the given length
            if (tok_.kind != TK_INT) return fail("expected a segment length");
            ConstVec cv;
            cv.type = T_Segdes;
            cv.ints.push_back(tok_.ival);
            advance();
            return pooled(line, std::move(cv));
        } else {
            return fail("illegal constant type");
        }
        advance();
        prog_.code.push_back(in);
        return true;
    }

    bool pooled(int line, ConstVec &&cv) {
        Instr in;
        in.op = OP_CONST;
        in.type = cv.type == T_String ? T_Int : cv.type;
        in.line = line;
        in.pooled = true;
        in.pool_index = (int)prog_.pool.size();
        prog_.pool.push_back(std::move(cv));
        prog_.code.push_back(in);
        return true;
    }

    bool link() {
        if (!ifs_.empty()) return fail("IF without ENDIF");
        bool ok = true;
        for (auto &in : prog_.code) {
            if (in.op != OP_CALL) continue;
            auto it = prog_.labels.find(in.label);
            if (it == prog_.labels.end()) {
                std::fprintf(stderr, "vinterp: line %d: link error: function %s not defined\n", in.line,
                             in.label.c_str());
                ok = false;
            } else {
                in.branch = it->second;
                if (link_trace_)
                    std::fprintf(stderr, "link: line %d: CALL %s -> instruction %d\n", in.line, in.label.c_str(),
                                 in.branch);
            }
        }
        if (prog_.entry.empty()) {
            std::fprintf(stderr, "vinterp: no start function\n");
            ok = false;
        }
        return ok;
    }

    Lexer lex_;
    Program &prog_;
    bool lex_trace_, link_trace_;
    Token tok_;
    std::vector<int> ifs_;
    bool main_seen_ = false;
};

bool parse_program(const std::string &text, Program &prog, bool lex_trace, bool link_trace) {
    Parser p(text, prog, lex_trace, link_trace);
    return p.parse();
}

void show_program(const Program &prog, FILE *out) {
    std::map<int, std::string> funcs;
    for (const auto &l : prog.labels) funcs[l.second] = l.first;
    for (size_t pc = 0; pc < prog.code.size(); ++pc) {
        auto f = funcs.find((int)pc);
        if (f != funcs.end()) std::fprintf(out, "FUNC %s\n", f->second.c_str());
        const Instr &in = prog.code[pc];
        std::fprintf(out, "%6zu  line %-6d %s", pc, in.line, op_info(in.op).name);
        if (in.type != T_Illegal && in.type != T_None) std::fprintf(out, " %s", type_name(in.type));
        switch (in.op) {
            case OP_COPY: case OP_POP: case OP_CPOP: std::fprintf(out, " %d %d", in.a, in.b); break;
            case OP_FUSED: std::fprintf(out, " %d", in.a); break;
            case OP_CALL: std::fprintf(out, " %s -> %d", in.label.c_str(), in.branch); break;
            case OP_IF: case OP_ELSE: std::fprintf(out, " -> %d", in.branch); break;
            case OP_CONST:
                if (in.pooled) std::fprintf(out, " <constant %d>", in.pool_index);
                else if (in.type == T_Float) std::fprintf(out, " %.17g", in.dval);
                else std::fprintf(out, " %lld", in.ival);
                break;
            default: break;
        }
        std::fprintf(out, "\n");
    }
}

}  // namespace vcode

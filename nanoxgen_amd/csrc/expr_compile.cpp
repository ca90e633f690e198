// The XGen/SeExpr scalar expression compiler: source text to the float SSA
// IR the device interpreter runs.  The reference compiles in
// src/xgen_expression.cpp (source normalisation 50-118, tokenizer 128-218,
// recursive-descent compiler 220-556, compile_xgen_scalar_expression 771-781,
// make_xgen_float_expression_program 783-818); this compiler accepts the same
// language, emits the same SSA values in the same order (so the programs are
// identical instruction for instruction) and reports the same diagnostics.
// The binary operators are parsed by precedence climbing over one table.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "internal.h"

namespace nxg {
namespace {

const uint8_t kPermutation[256] = {
#include "seexpr_permutation.inc"
};

struct CompileError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

[[noreturn]] void syntax(std::size_t at, const std::string &what) {
    throw CompileError("XGen scalar expression error at byte " + std::to_string(at) + ": " + what);
}

// Comments run from '#' to the line end and become ';'; a line break joins
// the two lines (a space) when either side is an operator, a comma, '(' or,
// after it, ')'; otherwise it ends a statement (';').  A literal "\n" in the
// text is a line break (src/xgen_expression.cpp:50-118).
std::string normalise(const char *src, std::size_t size, std::size_t limit) {
    if (size > limit) throw CompileError("XGen scalar expression source byte limit exceeded");
    if (std::memchr(src, '\0', size)) throw CompileError("XGen scalar expression contains a NUL byte");
    std::string text;
    text.reserve(size);
    bool in_comment = false;
    for (std::size_t i = 0; i < size; ++i) {
        char c = src[i];
        if (c == '\\' && i + 1 < size && src[i + 1] == 'n') {
            c = '\n';
            ++i;
        }
        const bool eol = c == '\n' || c == '\r';
        if (in_comment) {
            if (eol) {
                in_comment = false;
                text.push_back(';');
            }
        } else if (c == '#') {
            in_comment = true;
        } else {
            text.push_back(c);
        }
    }
    auto joins = [](char c) { return std::strchr("+-*/<>=!&|?:,(", c) != nullptr && c != '\0'; };
    std::string out;
    out.reserve(text.size());
    for (std::size_t i = 0; i < text.size(); ++i) {
        const char c = text[i];
        if (c != '\n' && c != '\r') {
            out.push_back(c);
            continue;
        }
        std::size_t back = out.size();
        while (back && (out[back - 1] == ' ' || out[back - 1] == '\t')) --back;
        std::size_t ahead = i + 1;
        while (ahead < text.size() && std::strchr(" \t\n\r", text[ahead]) && text[ahead] != '\0') ++ahead;
        const bool prev = back && joins(out[back - 1]);
        const bool next = ahead < text.size() && (joins(text[ahead]) || text[ahead] == ')');
        out.push_back(prev || next ? ' ' : ';');
    }
    return out;
}

enum class Tok : uint8_t {
    End, Number, Variable, Identifier, String, Semi, Comma, LParen, RParen, Question, Colon, Assign, Plus, Minus,
    Star, Slash, Bang, Arrow, Less, LessEq, Greater, GreaterEq, Eq, NotEq, And, Or
};

struct Token {
    Tok kind;
    std::string text;
    double number;
    std::size_t at;
};

bool ident_start(char c) { return (c >= 'a' && c <= 'z') || (c >= 'A' && c <= 'Z') || c == '_'; }
bool ident_char(char c) { return ident_start(c) || (c >= '0' && c <= '9'); }

std::vector<Token> lex(const std::string &s, std::size_t limit) {
    std::vector<Token> out;
    auto push = [&](Tok k, std::size_t at, std::string text = {}, double v = 0.0) {
        if (out.size() >= limit) throw CompileError("XGen scalar expression token limit exceeded");
        out.push_back({k, std::move(text), v, at});
    };
    std::size_t i = 0;
    while (i < s.size()) {
        const char c = s[i];
        if (c == ' ' || c == '\t' || c == '\f' || c == '\v') {
            ++i;
            continue;
        }
        const bool digit = c >= '0' && c <= '9';
        if (digit || (c == '.' && i + 1 < s.size() && s[i + 1] >= '0' && s[i + 1] <= '9')) {
            char *end = nullptr;
            const double v = std::strtod(s.c_str() + i, &end);
            const std::size_t used = static_cast<std::size_t>(end - (s.c_str() + i));
            if (!used || !std::isfinite(v)) syntax(i, "invalid or non-finite numeric literal");
            push(Tok::Number, i, {}, v);
            i += used;
        } else if (c == '$') {
            const std::size_t at = i++;
            if (i >= s.size() || !ident_start(s[i])) syntax(at, "expected a variable name after '$'");
            const std::size_t from = i;
            while (i < s.size() && ident_char(s[i])) ++i;
            push(Tok::Variable, at, s.substr(from, i - from));
        } else if (ident_start(c)) {
            const std::size_t from = i;
            while (i < s.size() && ident_char(s[i])) ++i;
            push(Tok::Identifier, from, s.substr(from, i - from));
        } else if (c == '\'' || c == '"') {
            const std::size_t at = i++;
            std::string text;
            while (i < s.size() && s[i] != c) {
                if (s[i] == '\\' && i + 1 < s.size()) ++i;
                text.push_back(s[i++]);
            }
            if (i == s.size()) syntax(at, "unterminated string literal");
            ++i;
            push(Tok::String, at, std::move(text));
        } else {
            const char n = i + 1 < s.size() ? s[i + 1] : '\0';
            auto pair = [&](char second, Tok both, Tok single) {
                push(n == second ? both : single, i);
                i += n == second ? 2 : 1;
            };
            switch (c) {
                case ';': push(Tok::Semi, i++); break;
                case ',': push(Tok::Comma, i++); break;
                case '(': push(Tok::LParen, i++); break;
                case ')': push(Tok::RParen, i++); break;
                case '?': push(Tok::Question, i++); break;
                case ':': push(Tok::Colon, i++); break;
                case '+': push(Tok::Plus, i++); break;
                case '*': push(Tok::Star, i++); break;
                case '/': push(Tok::Slash, i++); break;
                case '-': pair('>', Tok::Arrow, Tok::Minus); break;
                case '=': pair('=', Tok::Eq, Tok::Assign); break;
                case '!': pair('=', Tok::NotEq, Tok::Bang); break;
                case '<': pair('=', Tok::LessEq, Tok::Less); break;
                case '>': pair('=', Tok::GreaterEq, Tok::Greater); break;
                case '&':
                    if (n != '&') syntax(i, "single '&' is unsupported; use '&&'");
                    push(Tok::And, i);
                    i += 2;
                    break;
                case '|':
                    if (n != '|') syntax(i, "single '|' is unsupported; use '||'");
                    push(Tok::Or, i);
                    i += 2;
                    break;
                default: syntax(i, "unsupported character");
            }
        }
    }
    push(Tok::End, s.size());
    return out;
}

struct Instruction {
    uint8_t op;
    uint32_t offset;
    uint16_t count;
    uint32_t aux;
    double imm;
};

struct RampPoint {
    double position, value;
    uint8_t interpolation;
};

struct Limits {
    std::size_t tokens, instructions, locals, ramp_points;
};

// Binary operators by precedence level (higher binds tighter), left-associative.
struct BinaryOp {
    Tok tok;
    int level;
    uint8_t op;
};
constexpr BinaryOp kBinary[] = {
    {Tok::Or, 1, NXG_OP_LOGICAL_OR},     {Tok::And, 2, NXG_OP_LOGICAL_AND},
    {Tok::Eq, 3, NXG_OP_EQUAL},          {Tok::NotEq, 3, NXG_OP_NOT_EQUAL},
    {Tok::Less, 4, NXG_OP_LESS},         {Tok::LessEq, 4, NXG_OP_LESS_EQUAL},
    {Tok::Greater, 4, NXG_OP_GREATER},   {Tok::GreaterEq, 4, NXG_OP_GREATER_EQUAL},
    {Tok::Plus, 5, NXG_OP_ADD},          {Tok::Minus, 5, NXG_OP_SUBTRACT},
    {Tok::Star, 6, NXG_OP_MULTIPLY},     {Tok::Slash, 6, NXG_OP_DIVIDE},
};

class Parser {
public:
    Parser(std::vector<Token> tokens, const Limits &limits, uint32_t seed, bool extensions)
        : t_(std::move(tokens)), lim_(limits), seed_(seed), ext_(extensions) {}

    std::vector<Instruction> ins;
    std::vector<uint32_t> operands;
    std::vector<std::string> inputs;
    std::vector<RampPoint> points;
    std::vector<std::pair<uint32_t, uint32_t>> ramps;
    uint32_t result = 0, random_calls = 0;

    void run() {
        bool any = false;
        while (cur().kind != Tok::End) {
            while (eat(Tok::Semi)) {
            }
            if (cur().kind == Tok::End) break;
            const bool assignment =
                (cur().kind == Tok::Variable || cur().kind == Tok::Identifier) && ahead(1).kind == Tok::Assign;
            if (assignment) {
                const std::string name = t_[pos_].text;
                pos_ += 2;
                const uint32_t value = ternary();
                if (!locals_.count(name) && locals_.size() >= lim_.locals)
                    throw CompileError("XGen scalar expression local limit exceeded");
                locals_[name] = value;
                result = value;
            } else {
                result = ternary();
            }
            any = true;
            if (cur().kind != Tok::Semi && cur().kind != Tok::End) syntax(cur().at, "expected ';' after expression");
        }
        if (!any) syntax(0, "expression is empty");
    }

private:
    std::vector<Token> t_;
    Limits lim_;
    uint32_t seed_;
    bool ext_;
    std::size_t pos_ = 0;
    std::unordered_map<std::string, uint32_t> locals_;

    const Token &cur() const { return ahead(0); }
    const Token &ahead(std::size_t k) const { return t_[std::min(pos_ + k, t_.size() - 1)]; }
    bool eat(Tok k) {
        if (cur().kind != k) return false;
        ++pos_;
        return true;
    }
    void need(Tok k, const char *what) {
        if (!eat(k)) syntax(cur().at, what);
    }

    uint32_t emit(uint8_t op, const std::vector<uint32_t> &args = {}, double imm = 0.0, uint32_t aux = 0) {
        if (ins.size() >= lim_.instructions) throw CompileError("XGen scalar expression instruction limit exceeded");
        if (args.size() > 0xffffu) throw CompileError("XGen scalar expression operand limit exceeded");
        const uint32_t offset = static_cast<uint32_t>(operands.size());
        operands.insert(operands.end(), args.begin(), args.end());
        ins.push_back({op, offset, static_cast<uint16_t>(args.size()), aux, imm});
        return static_cast<uint32_t>(ins.size() - 1);
    }

    // cond ? expression : ternary  (src/xgen_expression.cpp:296-304)
    uint32_t ternary() {
        const uint32_t cond = binary(1);
        if (!eat(Tok::Question)) return cond;
        const uint32_t yes = ternary();
        need(Tok::Colon, "expected ':' in conditional expression");
        const uint32_t no = ternary();
        return emit(NXG_OP_SELECT, {cond, yes, no});
    }

    uint32_t binary(int min_level) {
        uint32_t lhs = unary();
        for (;;) {
            const BinaryOp *found = nullptr;
            for (const BinaryOp &b : kBinary)
                if (b.tok == cur().kind && b.level >= min_level) found = &b;
            if (!found) return lhs;
            ++pos_;
            const uint32_t rhs = binary(found->level + 1);
            lhs = emit(found->op, {lhs, rhs});
        }
    }

    uint32_t unary() {
        if (eat(Tok::Plus)) return unary();
        if (eat(Tok::Minus)) return emit(NXG_OP_NEGATE, {unary()});
        if (eat(Tok::Bang)) return emit(NXG_OP_LOGICAL_NOT, {unary()});
        uint32_t value = primary();
        while (eat(Tok::Arrow)) {  // value->method(args): method(value, args)
            if (cur().kind != Tok::Identifier) syntax(cur().at, "expected a method name after '->'");
            const Token name = t_[pos_++];
            need(Tok::LParen, "expected '(' after expression method name");
            std::vector<uint32_t> args{value};
            if (!eat(Tok::RParen)) {
                do args.push_back(ternary());
                while (eat(Tok::Comma));
                need(Tok::RParen, "expected ')' after expression method arguments");
            }
            value = call(name, args);
        }
        return value;
    }

    double signed_number(const char *what) {
        double sign = 1.0;
        if (!eat(Tok::Plus) && eat(Tok::Minus)) sign = -1.0;
        if (cur().kind != Tok::Number) syntax(cur().at, std::string("expected a numeric ramp ") + what);
        return sign * t_[pos_++].number;
    }

    // rampUI(position, value, interpolation : ...)  (src/xgen_expression.cpp:391-441)
    uint32_t ramp(const Token &name) {
        const std::size_t first = points.size();
        do {
            const double p = signed_number("position");
            need(Tok::Comma, "expected ',' after ramp position");
            const double v = signed_number("value");
            need(Tok::Comma, "expected ',' after ramp value");
            const double k = signed_number("interpolation");
            if (k != std::floor(k) || k < 0.0 || k > 3.0)
                syntax(name.at, "ramp interpolation must be an integer in [0, 3]");
            if (p < 0.0 || p > 1.0) syntax(name.at, "ramp positions must be in [0, 1]");
            if (points.size() >= lim_.ramp_points) throw CompileError("XGen scalar expression ramp point limit exceeded");
            if (points.size() != first && p <= points.back().position)
                syntax(name.at, "ramp positions must be strictly increasing");
            points.push_back({p, v, static_cast<uint8_t>(k)});
        } while (eat(Tok::Colon));
        need(Tok::RParen, "expected ')' after rampUI points");
        const std::size_t n = points.size() - first;
        if (n == 0 || n > 0xffffu) syntax(name.at, "ramp point storage limit exceeded");
        ramps.push_back({static_cast<uint32_t>(first), static_cast<uint32_t>(n)});
        return emit(NXG_OP_RAMP, {}, 0.0, static_cast<uint32_t>(ramps.size() - 1));
    }

    uint32_t primary() {
        const Token &tk = cur();
        if (tk.kind == Tok::Number) return emit(NXG_OP_CONSTANT, {}, t_[pos_++].number);
        if (tk.kind == Tok::Variable) {
            const std::string name = t_[pos_++].text;
            if (auto it = locals_.find(name); it != locals_.end()) return it->second;
            uint32_t index = 0;
            while (index < inputs.size() && inputs[index] != name) ++index;
            if (index == inputs.size()) inputs.push_back(name);
            return emit(NXG_OP_INPUT, {}, 0.0, index);
        }
        if (eat(Tok::LParen)) {
            const uint32_t v = ternary();
            need(Tok::RParen, "expected ')'");
            return v;
        }
        if (tk.kind == Tok::String) syntax(tk.at, "string values require an explicit PTEX binding such as map()");
        if (tk.kind != Tok::Identifier) syntax(tk.at, "expected a scalar value");
        const Token name = t_[pos_++];
        if (name.text == "true") return emit(NXG_OP_CONSTANT, {}, 1.0);
        if (name.text == "false") return emit(NXG_OP_CONSTANT, {}, 0.0);
        if (!eat(Tok::LParen)) {
            if (auto it = locals_.find(name.text); it != locals_.end()) return it->second;
            syntax(name.at, "bare identifier is unsupported: " + name.text);
        }
        if (name.text == "rampUI") return ramp(name);
        std::vector<uint32_t> args;
        if (!eat(Tok::RParen)) {
            do args.push_back(ternary());
            while (eat(Tok::Comma));
            need(Tok::RParen, "expected ')' after function arguments");
        }
        return call(name, args);
    }

    uint32_t call(const Token &name, const std::vector<uint32_t> &args) {
        struct Fixed {
            const char *name;
            uint8_t op;
            std::size_t arity;
            bool extension;
        };
        static const Fixed kFixed[] = {
            {"min", NXG_OP_MINIMUM, 2, false},     {"max", NXG_OP_MAXIMUM, 2, false},
            {"clamp", NXG_OP_CLAMP, 3, false},     {"fit", NXG_OP_FIT, 5, false},
            {"gamma", NXG_OP_GAMMA, 2, false},     {"contrast", NXG_OP_CONTRAST, 2, false},
            {"smoothstep", NXG_OP_SMOOTHSTEP, 3, false}, {"pow", NXG_OP_POW, 2, true},
            {"mix", NXG_OP_MIX, 3, true},          {"noise", NXG_OP_NOISE3, 3, true},
        };
        const std::string &n = name.text;
        if (n == "hash") {
            if (args.empty()) syntax(name.at, "hash expects at least one argument");
            return emit(NXG_OP_HASH, args);
        }
        if (n == "rand") {  // call-site seed (src/xgen_expression.cpp:528-538)
            if (args.size() > 3) syntax(name.at, "rand expects zero to three arguments");
            const double site[2] = {static_cast<double>(seed_), static_cast<double>(random_calls)};
            const double seed = seexpr_hash(site, 2);
            ++random_calls;
            return emit(NXG_OP_RANDOM, args, seed);
        }
        for (const Fixed &f : kFixed) {
            if (n != f.name || (f.extension && !ext_)) continue;
            if (args.size() != f.arity)
                syntax(name.at, n + " expects " + std::to_string(f.arity) + " argument(s)");
            return emit(f.op, args);
        }
        if (ext_ && n.size() == 6 && n.compare(0, 5, "texel") == 0 && n[5] >= '0' && n[5] <= '3') {
            if (args.size() != 3) syntax(name.at, n + " expects 3 argument(s)");
            return emit(NXG_OP_TEXEL, args, 0.0, static_cast<uint32_t>(n[5] - '0'));
        }
        if (n == "map") syntax(name.at, "map requires an explicit PTEX binding and is not a scalar intrinsic");
        syntax(name.at, "unsupported function: " + n);
    }
};

void put(char *err, uint64_t cap, const char *text) {
    if (err && cap) std::snprintf(err, cap, "%s", text);
}

}  // namespace

// xgen_seexpr_component / _hash_prefix / _hash_finish / xgen_seexpr_hash
// (src/xgen_expression.cpp:619-661), in double as the reference computes them.
uint32_t seexpr_component(double x) {
    int exponent = 0;
    const double fraction = std::frexp(x * 8.539734222673566, &exponent);
    return static_cast<uint32_t>(fraction * 4294967295.0) ^ static_cast<uint32_t>(exponent);
}

uint32_t seexpr_finish(uint32_t s) {
    s ^= s >> 11u;
    s ^= (s << 7u) & 0x9d2c5680u;
    s ^= (s << 15u) & 0xefc60000u;
    s ^= s >> 18u;
    const uint32_t b3 = kPermutation[s & 0xffu];
    const uint32_t b2 = kPermutation[((s >> 8u) + b3) & 0xffu];
    const uint32_t b1 = kPermutation[((s >> 16u) + b2) & 0xffu];
    const uint32_t b0 = kPermutation[((s >> 24u) + b1) & 0xffu];
    return b0 | (b1 << 8u) | (b2 << 16u) | (b3 << 24u);
}

double seexpr_hash(const double *args, std::size_t n) {
    uint32_t state = 0u;
    for (std::size_t i = 0; i < n; ++i) state = state * 1664525u + seexpr_component(args[i]) + 1013904223u;
    return static_cast<double>(seexpr_finish(state)) * (1.0 / 4294967295.0);
}

// xgen_string_seed (src/xgen_expression.cpp:738-749): PJW hash mod 211.
uint32_t string_seed(const char *text, std::size_t size) {
    uint32_t h = 0u;
    for (std::size_t i = 0; i < size; ++i) {
        h = (h << 4u) + static_cast<unsigned char>(text[i]);
        if (const uint32_t high = h & 0xf0000000u) {
            h ^= high >> 24u;
            h &= ~high;
        }
    }
    return h % 211u;
}

}  // namespace nxg

extern "C" {

uint32_t nxg_expr_compile(const char *source, uint64_t source_bytes, const nxg_expr_compile_options *options,
                          nxg_expr_compiled *out, char *err, uint64_t err_capacity) {
    if (!out || (!source && source_bytes)) return NXG_ERROR_NULL_ARGUMENT;
    nxg::put(err, err_capacity, "");
    nxg_expr_compile_options o{};
    if (options) o = *options;
    auto or_default = [](uint64_t v, uint64_t d) { return v ? v : d; };
    const std::string name = o.expression_name ? o.expression_name : "";
    const std::string type = o.object_type ? o.object_type : "";
    const std::string key = name + type;  // xgen_expression_seed (src/xgen_expression.cpp:751-758)
    const uint32_t seed = nxg::string_seed(key.data(), key.size());
    try {
        const std::string text = nxg::normalise(source ? source : "", source_bytes,
                                                or_default(o.max_source_bytes, 1024u * 1024u));
        nxg::Limits lim{or_default(o.max_tokens, 100000u), or_default(o.max_instructions, 100000u),
                        or_default(o.max_locals, 4096u), or_default(o.max_ramp_points, 4096u)};
        nxg::Parser p(nxg::lex(text, lim.tokens), lim, seed, o.extensions != 0u);
        p.run();
        // make_xgen_float_expression_program (src/xgen_expression.cpp:783-818)
        std::vector<nxg_expr_instruction> ins(p.ins.size());
        for (std::size_t i = 0; i < p.ins.size(); ++i) {
            const nxg::Instruction &s = p.ins[i];
            const float imm = static_cast<float>(s.imm);
            if (!std::isfinite(imm)) throw nxg::CompileError("XGen expression immediate cannot be represented as float");
            ins[i] = {s.op, 0, s.count, s.offset, s.op == NXG_OP_RANDOM ? nxg::seexpr_component(s.imm) : s.aux, imm};
        }
        std::vector<nxg_expr_ramp_point> pts(p.points.size());
        for (std::size_t i = 0; i < p.points.size(); ++i) {
            const float pos = static_cast<float>(p.points[i].position), val = static_cast<float>(p.points[i].value);
            if (!std::isfinite(pos) || !std::isfinite(val))
                throw nxg::CompileError("XGen ramp point cannot be represented as float");
            pts[i] = {pos, val, p.points[i].interpolation};
        }
        uint64_t name_bytes = 0u;
        for (const std::string &s : p.inputs) name_bytes += s.size() + 1u;
        out->instruction_count = static_cast<uint32_t>(ins.size());
        out->operand_count = static_cast<uint32_t>(p.operands.size());
        out->ramp_point_count = static_cast<uint32_t>(pts.size());
        out->ramp_count = static_cast<uint32_t>(p.ramps.size());
        out->input_count = static_cast<uint32_t>(p.inputs.size());
        out->input_name_bytes = static_cast<uint32_t>(name_bytes);
        out->result = p.result;
        out->expression_seed = seed;
        out->random_call_count = p.random_calls;
        if (out->instruction_capacity < ins.size() || out->operand_capacity < p.operands.size() ||
            out->ramp_point_capacity < pts.size() || out->ramp_capacity < p.ramps.size() ||
            out->input_name_capacity < name_bytes)
            return NXG_ERROR_BUFFER_TOO_SMALL;
        if (!ins.empty()) std::memcpy(out->instructions, ins.data(), ins.size() * sizeof(ins[0]));
        if (!p.operands.empty()) std::memcpy(out->operands, p.operands.data(), p.operands.size() * 4u);
        if (!pts.empty()) std::memcpy(out->ramp_points, pts.data(), pts.size() * sizeof(pts[0]));
        for (std::size_t i = 0; i < p.ramps.size(); ++i) out->ramps[i] = {p.ramps[i].first, p.ramps[i].second};
        char *names = out->input_names;
        for (const std::string &s : p.inputs) {
            std::memcpy(names, s.c_str(), s.size() + 1u);
            names += s.size() + 1u;
        }
        return NXG_OK;
    } catch (const nxg::CompileError &e) {
        nxg::put(err, err_capacity, e.what());
        return NXG_ERROR_INVALID_PROGRAM;
    }
}

double nxg_seexpr_hash(const double *arguments, uint32_t count) { return nxg::seexpr_hash(arguments, count); }

uint32_t nxg_string_seed(const char *text) { return nxg::string_seed(text ? text : "", text ? std::strlen(text) : 0u); }

double nxg_face_seed(uint32_t description_id, const char *patch_name, uint32_t face_id) {
    // xgen_face_seed (src/xgen_expression.cpp:760-769): the name after the last ':'
    const char *name = patch_name ? patch_name : "";
    if (const char *colon = std::strrchr(name, ':')) name = colon + 1;
    const double args[3] = {static_cast<double>(description_id),
                            static_cast<double>(nxg::string_seed(name, std::strlen(name))),
                            static_cast<double>(face_id)};
    return nxg::seexpr_hash(args, 3);
}

}  // extern "C"

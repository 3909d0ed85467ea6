// Reader of the reference's configuration format (include/ch_swe_config.h).
// The syntax and value semantics follow geosci_config_parser_mod.f90
// (L21-454) and arithmetic_parser_mod.f90 (L13-369): a value is tried as an
// integer, a real, a logical, a string, a nested section, an array, then an
// arithmetic expression; expressions use integer arithmetic when both
// operands are integers, unary operators (minus, functions) bind tighter
// than every binary operator, and binary operators of equal priority
// associate to the left (** and ^ included).  Own implementation: a
// recursive-descent evaluator over the same grammar.
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>

#include "../../include/ch_swe.h"
#include "../../include/ch_swe_config.h"
#include "swe_error.hpp"

namespace {

struct Val {
    enum Kind { INT, REAL, LOGICAL, STR } kind = INT;
    long long i = 0;
    double r = 0;
    bool b = false;
    std::string s;
};

using Table = std::map<std::string, Val>;

constexpr double kPi = 3.14159265358979323846;

// const_mod.f90 L4-11 and the parser's global constants (L33-42)
const Table &constants()
{
    static Table t = [] {
        Table c;
        auto real = [&](const char *k, double v) {
            Val x;
            x.kind = Val::REAL;
            x.r = v;
            c[k] = x;
        };
        auto integer = [&](const char *k, long long v) {
            Val x;
            x.kind = Val::INT;
            x.i = v;
            c[k] = x;
        };
        real("pi", kPi);
        real("Earth_grav", 9.80616);
        real("Earth_radii", 0.6371229e7);
        real("Earth_omega", 2.0 * kPi / double(23 * 60 * 60 + 56 * 60 + 4));
        real("day", 24.0 * 3600.0);
        integer("min", 60);
        integer("sec", 1);
        integer("hour", 3600);
        return c;
    }();
    return t;
}

bool blank(char c) { return c == ' ' || c == '\t' || c == '\r' || c == '\n'; }

std::string trim(const std::string &s)
{
    size_t a = 0, b = s.size();
    while (a < b && blank(s[a])) a++;
    while (b > a && blank(s[b - 1])) b--;
    return s.substr(a, b - a);
}

std::string lower(std::string s)
{
    for (char &c : s) c = (char)std::tolower((unsigned char)c);
    return s;
}

// ---- arithmetic expressions ----------------------------------------------------
struct Num {
    bool is_int;
    long long i;
    double r;
    double real() const { return is_int ? double(i) : r; }
};

struct Expr {
    const std::string &s;
    size_t p = 0;
    const Table &vars;     // the whole configuration read so far
    const std::string &prefix;  // path of the enclosing section ("" at the top)
    std::string err;

    Expr(const std::string &s_, const Table &v, const std::string &pre) : s(s_), vars(v), prefix(pre) {}

    void skip()
    {
        while (p < s.size() && blank(s[p])) p++;
    }

    bool lookup(const std::string &name, Num &out) const
    {
        auto try_table = [&](const Table &t, const std::string &k) {
            auto it = t.find(k);
            if (it == t.end()) return false;
            if (it->second.kind == Val::INT) out = Num{true, it->second.i, 0};
            else if (it->second.kind == Val::REAL) out = Num{false, 0, it->second.r};
            else return false;
            return true;
        };
        // the enclosing section's entries, the whole configuration, the constants
        if (!prefix.empty() && try_table(vars, prefix + "%" + name)) return true;
        if (try_table(vars, name)) return true;
        return try_table(constants(), name);
    }

    static Num binop(char op, const Num &a, const Num &b, std::string &err)
    {
        if (a.is_int && b.is_int) {
            switch (op) {
            case '+': return Num{true, a.i + b.i, 0};
            case '-': return Num{true, a.i - b.i, 0};
            case '*': return Num{true, a.i * b.i, 0};
            case '/':
                if (b.i == 0) {
                    err = "integer division by zero";
                    return Num{true, 0, 0};
                }
                return Num{true, a.i / b.i, 0};  // truncation, as Fortran
            default: {
                // integer power (Fortran: a**b with b < 0 gives 0 unless |a| = 1)
                long long r = 1;
                if (b.i < 0) r = (a.i == 1) ? 1 : (a.i == -1 ? ((-b.i) % 2 ? -1 : 1) : 0);
                else
                    for (long long k = 0; k < b.i; k++) r *= a.i;
                return Num{true, r, 0};
            }
            }
        }
        const double x = a.real(), y = b.real();
        switch (op) {
        case '+': return Num{false, 0, x + y};
        case '-': return Num{false, 0, x - y};
        case '*': return Num{false, 0, x * y};
        case '/': return Num{false, 0, x / y};
        default: return Num{false, 0, std::pow(x, y)};
        }
    }

    // number literal: digits with an optional '.', exponent e/E/d/D and sign
    bool number(Num &out)
    {
        size_t q = p;
        bool dot = false, digit = false;
        while (q < s.size() && (std::isdigit((unsigned char)s[q]) || (s[q] == '.' && !dot))) {
            if (s[q] == '.') dot = true;
            else digit = true;
            q++;
        }
        if (!digit) return false;
        bool expo = false;
        if (q < s.size() && std::strchr("eEdD", s[q])) {
            size_t r = q + 1;
            if (r < s.size() && (s[r] == '+' || s[r] == '-')) r++;
            if (r < s.size() && std::isdigit((unsigned char)s[r])) {
                while (r < s.size() && std::isdigit((unsigned char)s[r])) r++;
                expo = true;
                q = r;
            }
        }
        std::string t = s.substr(p, q - p);
        for (char &c : t)
            if (c == 'd' || c == 'D') c = 'e';
        if (dot || expo) out = Num{false, 0, std::strtod(t.c_str(), nullptr)};
        else out = Num{true, std::strtoll(t.c_str(), nullptr, 10), 0};
        p = q;
        return true;
    }

    bool identifier(std::string &name)
    {
        if (p >= s.size() || !std::isalpha((unsigned char)s[p])) return false;
        size_t q = p + 1;
        while (q < s.size() && (std::isalnum((unsigned char)s[q]) || s[q] == '_')) q++;
        name = s.substr(p, q - p);
        p = q;
        return true;
    }

    // unary: '-' unary | function unary | '(' expr ')' | number | variable
    bool unary(Num &out)
    {
        skip();
        if (p >= s.size()) {
            err = "unexpected end of expression";
            return false;
        }
        if (s[p] == '-') {
            p++;
            if (!unary(out)) return false;
            out = out.is_int ? Num{true, -out.i, 0} : Num{false, 0, -out.r};
            return true;
        }
        if (s[p] == '(') {
            p++;
            if (!expr(out)) return false;
            skip();
            if (p >= s.size() || s[p] != ')') {
                err = "unbalanced parentheses";
                return false;
            }
            p++;
            return true;
        }
        if (number(out)) return true;
        std::string name;
        if (!identifier(name)) {
            err = "unexpected syntax: " + s.substr(p);
            return false;
        }
        const std::string f = lower(name);
        static const char *funcs[] = {"sin", "cos", "tan", "acos", "asin", "atan", "exp", "log", "int", "sqrt"};
        for (const char *fn : funcs)
            if (f == fn) {
                Num a;
                if (!unary(a)) return false;
                const double x = a.real();
                if (f == "int") out = Num{true, (long long)std::trunc(x), 0};
                else if (f == "sin") out = Num{false, 0, std::sin(x)};
                else if (f == "cos") out = Num{false, 0, std::cos(x)};
                else if (f == "tan") out = Num{false, 0, std::tan(x)};
                else if (f == "acos") out = Num{false, 0, std::acos(x)};
                else if (f == "asin") out = Num{false, 0, std::asin(x)};
                else if (f == "atan") out = Num{false, 0, std::atan(x)};
                else if (f == "exp") out = Num{false, 0, std::exp(x)};
                else if (f == "log") out = Num{false, 0, std::log(x)};
                else out = Num{false, 0, std::sqrt(x)};
                return true;
            }
        if (!lookup(name, out)) {
            err = "unknown variable '" + name + "'";
            return false;
        }
        return true;
    }

    // binary operators by priority (pow 3, mul/div 2, add/sub 1), each level
    // associating to the left
    bool level(int prio, Num &out)
    {
        if (prio > 3) return unary(out);
        if (!level(prio + 1, out)) return false;
        for (;;) {
            skip();
            if (p >= s.size()) return true;
            char op = 0;
            size_t w = 1;
            if (prio == 3 && s.compare(p, 2, "**") == 0) {
                op = '^';
                w = 2;
            } else if (prio == 3 && s[p] == '^') {
                op = '^';
            } else if (prio == 2 && (s[p] == '*' || s[p] == '/') && s.compare(p, 2, "**") != 0) {
                op = s[p];
            } else if (prio == 1 && (s[p] == '+' || s[p] == '-')) {
                op = s[p];
            }
            if (!op) return true;
            p += w;
            Num b;
            if (!level(prio + 1, b)) return false;
            out = binop(op, out, b, err);
            if (!err.empty()) return false;
        }
    }

    bool expr(Num &out) { return level(1, out); }
};

// ---- the configuration parser ----------------------------------------------------
struct Parser {
    Table &t;
    std::string err;

    explicit Parser(Table &table) : t(table) {}

    // end of a scalar token: the first ',' new line or '#' after leading blanks
    static size_t token_end(const std::string &s, size_t from)
    {
        size_t j = from;
        while (j < s.size() && blank(s[j]) && s[j] != '\n') j++;
        size_t e = s.size();
        for (char c : {',', '\n', '#'}) {
            size_t k = s.find(c, j);
            if (k != std::string::npos) e = std::min(e, k);
        }
        return e;
    }

    // after a closing bracket or quote: blanks up to a token end
    bool tail(const std::string &s, size_t &q, const char *what)
    {
        while (q < s.size() && s[q] != ',' && s[q] != '\n' && s[q] != '#') {
            if (!blank(s[q])) {
                err = std::string("trailing characters after ") + what + ": '" + s.substr(q) + "'";
                return false;
            }
            q++;
        }
        return true;
    }

    void set(const std::string &key, const Val &v) { t[key] = v; }

    // parse the value at s[from..], store it under key; end = index of the
    // character that ends it (',', new line, '#', or the end of s)
    bool value(const std::string &s, size_t from, const std::string &key, const std::string &section, size_t &end)
    {
        const size_t te = token_end(s, from);
        const std::string tok = trim(s.substr(from, te - from));
        // integer: digits only
        if (!tok.empty() && tok.find_first_not_of("0123456789") == std::string::npos) {
            Val v;
            v.kind = Val::INT;
            v.i = std::strtoll(tok.c_str(), nullptr, 10);
            set(key, v);
            end = te;
            return true;
        }
        // real: digits, one '.', one exponent letter with an optional sign after it
        if (!tok.empty()) {
            bool ok = true, dot = false, ex = false;
            size_t exp_pos = std::string::npos;
            for (size_t i = 0; i < tok.size() && ok; i++) {
                const char c = tok[i];
                if (std::isdigit((unsigned char)c)) continue;
                if (std::strchr("eEdD", c) && !ex) {
                    ex = true;
                    exp_pos = i;
                } else if (i != 0 && (c == '+' || c == '-') && exp_pos == i - 1) {
                } else if (c == '.' && !ex && !dot) {
                    dot = true;
                } else {
                    ok = false;
                }
            }
            if (ok) {
                std::string t2 = tok;
                for (char &c : t2)
                    if (c == 'd' || c == 'D') c = 'e';
                Val v;
                v.kind = Val::REAL;
                v.r = std::strtod(t2.c_str(), nullptr);
                set(key, v);
                end = te;
                return true;
            }
        }
        // logical
        const std::string lt = lower(tok);
        if (lt == ".true." || lt == ".false.") {
            Val v;
            v.kind = Val::LOGICAL;
            v.b = lt == ".true.";
            set(key, v);
            end = te;
            return true;
        }
        size_t i = from;
        while (i < s.size() && blank(s[i])) i++;
        if (i < s.size() && (s[i] == '\'' || s[i] == '"')) {
            const char q0 = s[i];
            const size_t close = s.find(q0, i + 1);
            if (close != std::string::npos) {
                size_t q = close + 1;
                if (!tail(s, q, "a string")) return false;
                Val v;
                v.kind = Val::STR;
                v.s = s.substr(i + 1, close - i - 1);
                set(key, v);
                end = q;
                return true;
            }
        }
        if (i < s.size() && s[i] == '{') {
            int depth = 1;
            size_t q = i + 1;
            for (; q < s.size() && depth; q++) {
                if (s[q] == '{') depth++;
                else if (s[q] == '}') depth--;
            }
            if (depth) {
                err = "unbalanced {}: '" + s.substr(i) + "'";
                return false;
            }
            const size_t close = q - 1;
            if (!tail(s, q, "'}'")) return false;
            if (!section_body(s.substr(i + 1, close - i - 1), key)) return false;
            end = q;
            return true;
        }
        if (i < s.size() && s[i] == '[') {
            const size_t close = s.find(']', i + 1);
            const size_t nested = s.find('[', i + 1);
            if (close == std::string::npos) {
                err = "unbalanced []: '" + s.substr(i) + "'";
                return false;
            }
            if (nested != std::string::npos && nested < close) {
                err = "multi-level lists are not allowed: '" + s.substr(i) + "'";
                return false;
            }
            size_t q = close + 1;
            if (!tail(s, q, "']'")) return false;
            const std::string body = trim(s.substr(i + 1, close - i - 1));
            int n = 0;
            size_t pos = 0;
            while (pos < body.size()) {
                size_t e = 0;
                if (!value(body, pos, key + "%" + std::to_string(n + 1), section, e)) return false;
                n++;
                pos = e + 1;
            }
            Val len;
            len.kind = Val::INT;
            len.i = n;
            set(key + "%length", len);
            end = q;
            return true;
        }
        // arithmetic expression over the token
        Expr ex(tok, t, section);
        Num r{};
        if (tok.empty() || !ex.expr(r) || (ex.skip(), ex.p != tok.size())) {
            err = "wrong config entry for '" + key + "': '" + tok + "'" + (ex.err.empty() ? "" : " (" + ex.err + ")");
            return false;
        }
        Val v;
        if (r.is_int) {
            v.kind = Val::INT;
            v.i = r.i;
        } else {
            v.kind = Val::REAL;
            v.r = r.r;
        }
        set(key, v);
        end = te;
        return true;
    }

    // name = value entries of a section whose path is prefix ("" at the top)
    bool section_body(const std::string &body, const std::string &prefix)
    {
        const std::string s = trim(body);
        size_t pos = 0;
        while (pos < s.size()) {
            while (pos < s.size() && (blank(s[pos]) || s[pos] == ',')) pos++;
            if (pos >= s.size()) break;
            if (s[pos] == '#') {
                const size_t nl = s.find('\n', pos);
                pos = nl == std::string::npos ? s.size() : nl + 1;
                continue;
            }
            const size_t eq = s.find('=', pos);
            if (eq == std::string::npos) {
                err = "entry without '=': '" + s.substr(pos) + "'";
                return false;
            }
            const std::string name = trim(s.substr(pos, eq - pos));
            if (name.empty()) {
                err = "empty variable name: '" + s.substr(pos) + "'";
                return false;
            }
            for (char c : name)
                if (blank(c)) {
                    err = "incorrect variable name: '" + name + "'";
                    return false;
                }
            const std::string key = prefix.empty() ? name : prefix + "%" + name;
            size_t end = 0;
            if (!value(s, eq + 1, key, prefix, end)) return false;
            pos = end;  // a ',' or new line is skipped above, a '#' starts a comment
        }
        return true;
    }
};

}  // namespace

struct swe_cfg {
    Table t;
};

namespace {
const Val *find(const swe_cfg *cfg, const char *key, const char *who)
{
    if (!cfg || !key) {
        swe::set_error(SWE_EINVAL, std::string(who) + ": NULL argument");
        return nullptr;
    }
    auto it = cfg->t.find(key);
    if (it == cfg->t.end()) {
        swe::set_error(SWE_EINVAL, std::string(who) + ": no entry '" + key + "'");
        return nullptr;
    }
    return &it->second;
}
}  // namespace

extern "C" {

int swe_cfg_parse(const char *text, swe_cfg **out)
{
    if (!text || !out) return swe::set_error(SWE_EINVAL, "swe_cfg_parse: NULL argument");
    *out = nullptr;
    auto *c = new swe_cfg();
    Parser p(c->t);
    if (!p.section_body(text, "")) {
        delete c;
        return swe::set_error(SWE_EINVAL, "configuration: " + p.err);
    }
    *out = c;
    return SWE_OK;
}

int swe_cfg_read(const char *path, swe_cfg **out)
{
    if (!path || !out) return swe::set_error(SWE_EINVAL, "swe_cfg_read: NULL argument");
    FILE *f = std::fopen(path, "rb");
    if (!f) return swe::set_error(SWE_EINVAL, std::string("swe_cfg_read: cannot open '") + path + "'");
    std::string text;
    char buf[4096];
    size_t n;
    while ((n = std::fread(buf, 1, sizeof buf, f)) > 0) text.append(buf, n);
    std::fclose(f);
    return swe_cfg_parse(text.c_str(), out);
}

int swe_cfg_destroy(swe_cfg *cfg)
{
    delete cfg;
    return SWE_OK;
}

int swe_cfg_has(const swe_cfg *cfg, const char *key)
{
    return cfg && key && cfg->t.count(key) ? 1 : 0;
}

int swe_cfg_get_real(const swe_cfg *cfg, const char *key, double *value)
{
    const Val *v = find(cfg, key, "swe_cfg_get_real");
    if (!v) return SWE_EINVAL;
    // strict, as the reference's typed get (geosci_config_get_var.fpp L52-74):
    // an integer entry is a type error for a real read
    if (v->kind != Val::REAL)
        return swe::set_error(SWE_EINVAL, std::string("swe_cfg_get_real: '") + key + "' is not a real");
    *value = v->r;
    return SWE_OK;
}

int swe_cfg_get_int(const swe_cfg *cfg, const char *key, int *value)
{
    const Val *v = find(cfg, key, "swe_cfg_get_int");
    if (!v) return SWE_EINVAL;
    if (v->kind != Val::INT) return swe::set_error(SWE_EINVAL, std::string("swe_cfg_get_int: '") + key + "' is not an integer");
    *value = (int)v->i;
    return SWE_OK;
}

int swe_cfg_get_logical(const swe_cfg *cfg, const char *key, int *value)
{
    const Val *v = find(cfg, key, "swe_cfg_get_logical");
    if (!v) return SWE_EINVAL;
    if (v->kind != Val::LOGICAL)
        return swe::set_error(SWE_EINVAL, std::string("swe_cfg_get_logical: '") + key + "' is not a logical");
    *value = v->b ? 1 : 0;
    return SWE_OK;
}

int swe_cfg_get_str(const swe_cfg *cfg, const char *key, char *buf, int buflen, int *len)
{
    const Val *v = find(cfg, key, "swe_cfg_get_str");
    if (!v) return SWE_EINVAL;
    if (v->kind != Val::STR) return swe::set_error(SWE_EINVAL, std::string("swe_cfg_get_str: '") + key + "' is not a string");
    if (len) *len = (int)v->s.size();
    if (buf && buflen > 0) {
        const size_t n = std::min((size_t)(buflen - 1), v->s.size());
        std::memcpy(buf, v->s.data(), n);
        buf[n] = 0;
    }
    return SWE_OK;
}

}  // extern "C"

// Minimal strict JSON reader (see json.hpp).
#include "json.hpp"

#include <cerrno>
#include <cstdlib>
#include <cstring>

namespace pxl {
namespace {

struct Parser {
    const std::string &s;
    size_t i = 0;
    JsonError *err;
    int depth = 0;

    explicit Parser(const std::string &src, JsonError *e) : s(src), err(e) {}

    bool fail(const char *msg) {
        size_t line = 1, col = 1;
        for (size_t k = 0; k < i && k < s.size(); k++) {
            if (s[k] == '\n') { line++; col = 1; } else { col++; }
        }
        err->message = msg;
        err->line = line;
        err->column = col;
        return false;
    }
    void ws() {
        while (i < s.size() && (s[i] == ' ' || s[i] == '\t' || s[i] == '\n' || s[i] == '\r')) i++;
    }
    bool lit(const char *w) {
        size_t n = std::strlen(w);
        if (s.compare(i, n, w) != 0) return false;
        i += n;
        return true;
    }
    static void put_utf8(std::string &o, uint32_t cp) {
        if (cp < 0x80) {
            o += char(cp);
        } else if (cp < 0x800) {
            o += char(0xC0 | (cp >> 6));
            o += char(0x80 | (cp & 0x3F));
        } else if (cp < 0x10000) {
            o += char(0xE0 | (cp >> 12));
            o += char(0x80 | ((cp >> 6) & 0x3F));
            o += char(0x80 | (cp & 0x3F));
        } else {
            o += char(0xF0 | (cp >> 18));
            o += char(0x80 | ((cp >> 12) & 0x3F));
            o += char(0x80 | ((cp >> 6) & 0x3F));
            o += char(0x80 | (cp & 0x3F));
        }
    }
    bool hex4(uint32_t &v) {
        if (i + 4 > s.size()) return fail("EOF while parsing a string");
        v = 0;
        for (int k = 0; k < 4; k++) {
            char c = s[i++];
            v <<= 4;
            if (c >= '0' && c <= '9') v |= uint32_t(c - '0');
            else if (c >= 'a' && c <= 'f') v |= uint32_t(c - 'a' + 10);
            else if (c >= 'A' && c <= 'F') v |= uint32_t(c - 'A' + 10);
            else return fail("invalid escape");
        }
        return true;
    }
    bool string(std::string &o) {
        i++;  // opening quote
        while (true) {
            if (i >= s.size()) return fail("EOF while parsing a string");
            unsigned char c = (unsigned char)s[i];
            if (c == '"') { i++; return true; }
            if (c < 0x20) return fail("control character (\\u0000-\\u001F) found while parsing a string");
            if (c != '\\') { o += char(c); i++; continue; }
            i++;
            if (i >= s.size()) return fail("EOF while parsing a string");
            char e = s[i++];
            switch (e) {
                case '"': o += '"'; break;
                case '\\': o += '\\'; break;
                case '/': o += '/'; break;
                case 'b': o += '\b'; break;
                case 'f': o += '\f'; break;
                case 'n': o += '\n'; break;
                case 'r': o += '\r'; break;
                case 't': o += '\t'; break;
                case 'u': {
                    uint32_t cp;
                    if (!hex4(cp)) return false;
                    if (cp >= 0xD800 && cp < 0xDC00) {
                        uint32_t lo;
                        if (!(lit("\\u") && hex4(lo)) || lo < 0xDC00 || lo > 0xDFFF)
                            return fail("lone leading surrogate in hex escape");
                        cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
                    } else if (cp >= 0xDC00 && cp < 0xE000) {
                        return fail("lone trailing surrogate in hex escape");
                    }
                    put_utf8(o, cp);
                    break;
                }
                default: return fail("invalid escape");
            }
        }
    }
    bool number(JsonValue &v) {
        size_t start = i;
        bool integer = true;
        if (s[i] == '-') i++;
        if (i >= s.size()) return fail("EOF while parsing a value");
        if (s[i] == '0') {
            i++;
        } else if (s[i] >= '1' && s[i] <= '9') {
            while (i < s.size() && s[i] >= '0' && s[i] <= '9') i++;
        } else {
            return fail("invalid number");
        }
        if (i < s.size() && s[i] == '.') {
            integer = false;
            i++;
            if (i >= s.size() || !(s[i] >= '0' && s[i] <= '9')) return fail("invalid number");
            while (i < s.size() && s[i] >= '0' && s[i] <= '9') i++;
        }
        if (i < s.size() && (s[i] == 'e' || s[i] == 'E')) {
            integer = false;
            i++;
            if (i < s.size() && (s[i] == '+' || s[i] == '-')) i++;
            if (i >= s.size() || !(s[i] >= '0' && s[i] <= '9')) return fail("invalid number");
            while (i < s.size() && s[i] >= '0' && s[i] <= '9') i++;
        }
        v.kind = JsonValue::Number;
        v.text = s.substr(start, i - start);
        v.is_integer = integer;
        v.negative = v.text[0] == '-';
        errno = 0;
        v.num = std::strtod(v.text.c_str(), nullptr);  // correctly rounded decimal -> f64
        if (integer && !v.negative) {
            errno = 0;
            unsigned long long u = std::strtoull(v.text.c_str(), nullptr, 10);
            v.uint_fits = errno != ERANGE;
            v.uint_value = u;
        }
        return true;
    }
    bool value(JsonValue &v) {
        if (++depth > 128) return fail("recursion limit exceeded");
        ws();
        if (i >= s.size()) return fail("EOF while parsing a value");
        char c = s[i];
        bool ok;
        if (c == '{') {
            v.kind = JsonValue::Object;
            i++;
            ws();
            if (i < s.size() && s[i] == '}') { i++; depth--; return true; }
            while (true) {
                ws();
                if (i >= s.size()) return fail("EOF while parsing an object");
                if (s[i] != '"') return fail("key must be a string");
                std::string key;
                if (!string(key)) return false;
                ws();
                if (i >= s.size() || s[i] != ':') return fail("expected `:`");
                i++;
                JsonValue item;
                if (!value(item)) return false;
                v.members.emplace_back(std::move(key), std::move(item));
                ws();
                if (i < s.size() && s[i] == ',') { i++; continue; }
                if (i < s.size() && s[i] == '}') { i++; break; }
                return fail("expected `,` or `}`");
            }
            ok = true;
        } else if (c == '[') {
            v.kind = JsonValue::Array;
            i++;
            ws();
            if (i < s.size() && s[i] == ']') { i++; depth--; return true; }
            while (true) {
                JsonValue item;
                if (!value(item)) return false;
                v.items.push_back(std::move(item));
                ws();
                if (i < s.size() && s[i] == ',') { i++; continue; }
                if (i < s.size() && s[i] == ']') { i++; break; }
                return fail("expected `,` or `]`");
            }
            ok = true;
        } else if (c == '"') {
            v.kind = JsonValue::String;
            ok = string(v.text);
        } else if (c == 't') {
            v.kind = JsonValue::Bool;
            v.b = true;
            ok = lit("true") || fail("expected value");
        } else if (c == 'f') {
            v.kind = JsonValue::Bool;
            ok = lit("false") || fail("expected value");
        } else if (c == 'n') {
            v.kind = JsonValue::Null;
            ok = lit("null") || fail("expected value");
        } else if (c == '-' || (c >= '0' && c <= '9')) {
            ok = number(v);
        } else {
            ok = fail("expected value");
        }
        depth--;
        return ok;
    }
};

}  // namespace

bool json_parse(const std::string &text, JsonValue &out, JsonError &err) {
    Parser p(text, &err);
    if (!p.value(out)) return false;
    p.ws();
    if (p.i != text.size()) return p.fail("trailing characters");
    return true;
}

std::string json_describe(const JsonValue &v) {
    switch (v.kind) {
        case JsonValue::Null: return "null";
        case JsonValue::Bool: return std::string("boolean `") + (v.b ? "true" : "false") + "`";
        case JsonValue::Number:
            if (v.is_integer) return std::string("integer `") + v.text + "`";
            return std::string("floating point `") + v.text + "`";
        case JsonValue::String: return "string \"" + v.text + "\"";
        case JsonValue::Array: return "a sequence";
        case JsonValue::Object: return "a map";
    }
    return "value";
}

}  // namespace pxl

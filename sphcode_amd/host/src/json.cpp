#include "json.hpp"

#include <cerrno>
#include <cstdlib>

#include "sph/exception.hpp"

namespace sph {
namespace json {

const Value *Value::find(const std::string &key) const {
    if (kind != OBJECT) return nullptr;
    // property trees answer with the first member of that name
    for (const auto &m : members)
        if (m.first == key) return &m.second;
    return nullptr;
}

namespace {

struct Parser {
    const std::string &s;
    size_t p = 0;

    [[noreturn]] void fail(const char *what) const {
        throw SPHException("JSON parse error at offset " + std::to_string(p) + ": " + what);
    }
    void ws() {
        while (p < s.size() && (s[p] == ' ' || s[p] == '\t' || s[p] == '\n' || s[p] == '\r')) ++p;
    }
    bool eat(char c) {
        ws();
        if (p < s.size() && s[p] == c) {
            ++p;
            return true;
        }
        return false;
    }
    void expect(char c) {
        if (!eat(c)) fail((std::string("expected '") + c + "'").c_str());
    }
    bool literal(const char *lit) {
        size_t k = 0;
        while (lit[k]) ++k;
        if (s.compare(p, k, lit) == 0) {
            p += k;
            return true;
        }
        return false;
    }
    static void utf8(std::string &out, unsigned cp) {
        if (cp < 0x80) {
            out += (char)cp;
        } else if (cp < 0x800) {
            out += (char)(0xC0 | (cp >> 6));
            out += (char)(0x80 | (cp & 0x3F));
        } else if (cp < 0x10000) {
            out += (char)(0xE0 | (cp >> 12));
            out += (char)(0x80 | ((cp >> 6) & 0x3F));
            out += (char)(0x80 | (cp & 0x3F));
        } else {
            out += (char)(0xF0 | (cp >> 18));
            out += (char)(0x80 | ((cp >> 12) & 0x3F));
            out += (char)(0x80 | ((cp >> 6) & 0x3F));
            out += (char)(0x80 | (cp & 0x3F));
        }
    }
    unsigned hex4() {
        if (p + 4 > s.size()) fail("short \\u escape");
        unsigned v = 0;
        for (int k = 0; k < 4; ++k) {
            const char c = s[p++];
            v <<= 4;
            if (c >= '0' && c <= '9') v |= (unsigned)(c - '0');
            else if (c >= 'a' && c <= 'f') v |= (unsigned)(c - 'a' + 10);
            else if (c >= 'A' && c <= 'F') v |= (unsigned)(c - 'A' + 10);
            else fail("bad \\u escape");
        }
        return v;
    }
    std::string str() {
        expect('"');
        std::string out;
        while (true) {
            if (p >= s.size()) fail("unterminated string");
            const char c = s[p++];
            if (c == '"') break;
            if (c != '\\') {
                out += c;
                continue;
            }
            if (p >= s.size()) fail("unterminated escape");
            const char e = s[p++];
            switch (e) {
            case '"': out += '"'; break;
            case '\\': out += '\\'; break;
            case '/': out += '/'; break;
            case 'b': out += '\b'; break;
            case 'f': out += '\f'; break;
            case 'n': out += '\n'; break;
            case 'r': out += '\r'; break;
            case 't': out += '\t'; break;
            case 'u': {
                unsigned cp = hex4();
                if (cp >= 0xD800 && cp < 0xDC00 && s.compare(p, 2, "\\u") == 0) {
                    p += 2;
                    const unsigned lo = hex4();
                    cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
                }
                utf8(out, cp);
                break;
            }
            default: fail("bad escape");
            }
        }
        return out;
    }
    Value value() {
        ws();
        if (p >= s.size()) fail("unexpected end");
        Value v;
        const char c = s[p];
        if (c == '{') {
            ++p;
            v.kind = Value::OBJECT;
            if (eat('}')) return v;
            do {
                ws();
                std::string key = str();
                expect(':');
                v.members.emplace_back(std::move(key), value());
            } while (eat(','));
            expect('}');
        } else if (c == '[') {
            ++p;
            v.kind = Value::ARRAY;
            if (eat(']')) return v;
            do {
                v.items.push_back(value());
            } while (eat(','));
            expect(']');
        } else if (c == '"') {
            v.kind = Value::STRING;
            v.text = str();
        } else if (literal("true")) {
            v.kind = Value::BOOL;
            v.text = "true";
        } else if (literal("false")) {
            v.kind = Value::BOOL;
            v.text = "false";
        } else if (literal("null")) {
            v.kind = Value::NUL;
        } else {
            const size_t b = p;
            if (s[p] == '-' || s[p] == '+') ++p;
            while (p < s.size() && ((s[p] >= '0' && s[p] <= '9') || s[p] == '.' || s[p] == 'e' || s[p] == 'E' ||
                                    s[p] == '-' || s[p] == '+'))
                ++p;
            if (p == b) fail("unexpected character");
            v.kind = Value::NUMBER;
            v.text = s.substr(b, p - b);
        }
        return v;
    }
};

const Value &need(const Value &obj, const std::string &key) {
    const Value *v = obj.find(key);
    if (!v) throw SPHException("No such node (" + key + ")");
    return *v;
}

[[noreturn]] void bad(const std::string &key) {
    throw SPHException("conversion of data to type failed (" + key + ")");
}

bool scalar(const Value &v) { return v.kind == Value::BOOL || v.kind == Value::NUMBER || v.kind == Value::STRING; }

double to_real(const Value &v, const std::string &key) {
    if (!scalar(v) || v.text.empty()) bad(key);
    const char *b = v.text.c_str();
    char *e = nullptr;
    errno = 0;
    const double x = std::strtod(b, &e);
    while (*e == ' ') ++e;
    if (*e != '\0' || e == b) bad(key);
    return x;
}

int to_int(const Value &v, const std::string &key) {
    if (!scalar(v) || v.text.empty()) bad(key);
    const char *b = v.text.c_str();
    char *e = nullptr;
    errno = 0;
    const long x = std::strtol(b, &e, 10);
    while (*e == ' ') ++e;
    if (*e != '\0' || e == b || errno == ERANGE || x < -2147483647L - 1 || x > 2147483647L) bad(key);
    return (int)x;
}

bool to_bool(const Value &v, const std::string &key) {
    if (!scalar(v)) bad(key);
    if (v.text == "true" || v.text == "1") return true;
    if (v.text == "false" || v.text == "0") return false;
    bad(key);
}

}  // namespace

Value parse(const std::string &text) {
    Parser ps{text};
    Value v = ps.value();
    ps.ws();
    if (ps.p != text.size()) ps.fail("trailing characters");
    return v;
}

double get_real(const Value &obj, const std::string &key) { return to_real(need(obj, key), key); }
double get_real(const Value &obj, const std::string &key, double def) {
    const Value *v = obj.find(key);
    return v ? to_real(*v, key) : def;
}
int get_int(const Value &obj, const std::string &key) { return to_int(need(obj, key), key); }
int get_int(const Value &obj, const std::string &key, int def) {
    const Value *v = obj.find(key);
    return v ? to_int(*v, key) : def;
}
bool get_bool(const Value &obj, const std::string &key, bool def) {
    const Value *v = obj.find(key);
    return v ? to_bool(*v, key) : def;
}
std::string get_string(const Value &obj, const std::string &key) {
    const Value &v = need(obj, key);
    if (!scalar(v)) bad(key);
    return v.text;
}
std::string get_string(const Value &obj, const std::string &key, const std::string &def) {
    const Value *v = obj.find(key);
    if (!v) return def;
    if (!scalar(*v)) bad(key);
    return v->text;
}
std::vector<double> get_real_array(const Value &obj, const std::string &key) {
    const Value &v = need(obj, key);
    std::vector<double> out;
    if (v.kind == Value::ARRAY) {
        for (const auto &it : v.items) out.push_back(to_real(it, key));
    } else if (v.kind == Value::OBJECT) {
        for (const auto &m : v.members) out.push_back(to_real(m.second, key));
    } else {
        bad(key);
    }
    return out;
}

}  // namespace json
}  // namespace sph

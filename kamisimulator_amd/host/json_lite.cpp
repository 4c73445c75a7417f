// json_lite.cpp -- recursive-descent reader for json_lite.h.
#include "json_lite.h"

#include <cctype>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <stdexcept>

namespace kami_amd {
namespace {

struct Reader {
    const std::string& s;
    size_t p = 0;

    [[noreturn]] void fail(const char* what) const {
        throw std::runtime_error(std::string("JSON: ") + what + " at byte " + std::to_string(p));
    }
    void ws() {
        while (p < s.size() && std::isspace((unsigned char)s[p])) p++;
    }
    bool eat(char c) {
        ws();
        if (p < s.size() && s[p] == c) {
            p++;
            return true;
        }
        return false;
    }
    void expect(char c) {
        if (!eat(c)) fail((std::string("expected '") + c + "'").c_str());
    }
    std::string string_lit() {
        ws();
        if (p >= s.size() || s[p] != '"') fail("expected a string");
        p++;
        std::string out;
        while (p < s.size() && s[p] != '"') {
            char c = s[p++];
            if (c == '\\') {
                if (p >= s.size()) fail("unterminated escape");
                const char e = s[p++];
                switch (e) {
                    case 'n': out += '\n'; break;
                    case 't': out += '\t'; break;
                    case 'r': out += '\r'; break;
                    case 'b': out += '\b'; break;
                    case 'f': out += '\f'; break;
                    case 'u':
                        if (p + 4 > s.size()) fail("short \\u escape");
                        out += '?';  // settings never need non-ASCII text
                        p += 4;
                        break;
                    default: out += e;
                }
            } else {
                out += c;
            }
        }
        if (p >= s.size()) fail("unterminated string");
        p++;
        return out;
    }
    double number() {
        ws();
        const char* b = s.c_str() + p;
        char* e = nullptr;
        const double v = std::strtod(b, &e);
        if (e == b) fail("expected a value");
        p += (size_t)(e - b);
        return v;
    }
    bool word(const char* w) {
        ws();
        const size_t n = std::char_traits<char>::length(w);
        if (s.compare(p, n, w) == 0) {
            p += n;
            return true;
        }
        return false;
    }
    JsonValue value() {
        ws();
        if (p >= s.size()) fail("unexpected end");
        JsonValue v;
        const char c = s[p];
        if (c == '"') {
            v.kind = JsonValue::STRING;
            v.str = string_lit();
        } else if (c == '{') {
            p++;
            v.kind = JsonValue::OBJECT;
            if (!eat('}')) {
                do {
                    std::string k = string_lit();
                    expect(':');
                    v.members[k] = value();  // a repeated key keeps the last value
                } while (eat(','));
                expect('}');
            }
        } else if (c == '[') {
            p++;
            bool all_numbers = true;
            if (!eat(']')) {
                do {
                    v.items.push_back(value());
                    if (v.items.back().kind == JsonValue::NUMBER)
                        v.nums.push_back(v.items.back().num);
                    else
                        all_numbers = false;
                } while (eat(','));
                expect(']');
            }
            v.kind = all_numbers ? JsonValue::NUMBERS : JsonValue::ARRAY;
        } else if (word("true")) {
            v.kind = JsonValue::BOOL;
            v.b = true;
        } else if (word("false")) {
            v.kind = JsonValue::BOOL;
            v.b = false;
        } else if (word("null")) {
            v.kind = JsonValue::NUL;
        } else {
            v.kind = JsonValue::NUMBER;
            v.num = number();
        }
        return v;
    }
};

}  // namespace

JsonFile JsonFile::parse(const std::string& text) {
    JsonFile f;
    f.text_ = text;
    Reader r{f.text_};
    r.ws();
    if (r.p >= f.text_.size() || f.text_[r.p] != '{') r.fail("expected '{'");
    f.root_ = r.value();
    r.ws();
    if (r.p != f.text_.size()) r.fail("trailing characters");
    return f;
}

JsonFile JsonFile::load(const std::string& path) {
    std::ifstream in(path, std::ios::binary);
    if (!in) throw std::runtime_error("cannot open " + path);
    std::stringstream ss;
    ss << in.rdbuf();
    return parse(ss.str());
}

const JsonValue& JsonFile::at(const std::string& k, JsonValue::Kind want) const {
    auto it = root_.members.find(k);
    if (it == root_.members.end()) throw std::runtime_error("JSON key " + k + " is missing");
    const JsonValue& v = it->second;
    const bool ok = v.kind == want || (want == JsonValue::NUMBER && v.kind == JsonValue::BOOL) ||
                    (want == JsonValue::BOOL && v.kind == JsonValue::NUMBER);
    if (!ok) throw std::runtime_error("JSON key " + k + " has the wrong type");
    return v;
}

bool JsonFile::GetBool(const std::string& k) const {
    const JsonValue& v = at(k, JsonValue::BOOL);
    return v.kind == JsonValue::BOOL ? v.b : v.num != 0;
}
int JsonFile::GetInteger(const std::string& k) const {
    const JsonValue& v = at(k, JsonValue::NUMBER);
    return v.kind == JsonValue::BOOL ? (int)v.b : (int)v.num;
}
float JsonFile::GetFloat(const std::string& k) const {
    const JsonValue& v = at(k, JsonValue::NUMBER);
    return v.kind == JsonValue::BOOL ? (float)v.b : (float)v.num;
}
std::string JsonFile::GetString(const std::string& k) const { return at(k, JsonValue::STRING).str; }
std::vector<double> JsonFile::GetNumbers(const std::string& k) const { return at(k, JsonValue::NUMBERS).nums; }

namespace {
const JsonValue& null_value() {
    static const JsonValue v;
    return v;
}
[[noreturn]] void not_convertible(const char* what, const char* to) {
    throw std::runtime_error(std::string("JSON: ") + what + " is not convertible to " + to);
}
}  // namespace

size_t JsonValue::size() const {
    if (is_array()) return items.size();
    if (is_object()) return members.size();
    return 0;
}

const JsonValue& JsonValue::operator[](size_t i) const {
    if (kind == NUL) return null_value();
    if (!is_array()) throw std::runtime_error("JSON: element access on a value that is not an array");
    return i < items.size() ? items[i] : null_value();
}

const JsonValue& JsonValue::operator[](const std::string& key) const {
    if (kind == NUL) return null_value();
    if (!is_object()) throw std::runtime_error("JSON: member " + key + " read from a value that is not an object");
    auto it = members.find(key);
    return it == members.end() ? null_value() : it->second;
}

float JsonValue::as_float(const char* what) const {
    switch (kind) {
        case NUL: return 0.f;
        case BOOL: return b ? 1.f : 0.f;
        case NUMBER: return (float)num;
        default: not_convertible(what, "a number");
    }
}

int JsonValue::as_int(const char* what) const {
    switch (kind) {
        case NUL: return 0;
        case BOOL: return b ? 1 : 0;
        case NUMBER:
            if (!(num >= -2147483648.0 && num <= 2147483647.0)) not_convertible(what, "an int (out of range)");
            return (int)num;
        default: not_convertible(what, "an int");
    }
}

bool JsonValue::as_bool(const char* what) const {
    switch (kind) {
        case NUL: return false;
        case BOOL: return b;
        case NUMBER: return num != 0;
        default: not_convertible(what, "a bool");
    }
}

std::string JsonValue::as_string(const char* what) const {
    if (kind == NUL) return std::string();
    if (kind != STRING) not_convertible(what, "a string");
    return str;
}

}  // namespace kami_amd

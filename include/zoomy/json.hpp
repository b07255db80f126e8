// Minimal JSON reader for settings files: objects, arrays, strings (with \" \\ \/ \b \f \n \r
// \t \uXXXX escapes), numbers, true / false / null.  Numbers keep whether they were written
// as integers, so typed reads follow the usual JSON-library conventions: an integer or a
// floating value reads as either C++ number type, any other type mismatch is an error.
#pragma once
#include <cstdlib>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace zoomy {

struct JsonValue {
    enum Kind { Null, Bool, Number, String, Array, Object } kind = Null;
    bool b = false;
    double num = 0;
    bool is_int = false;
    std::string str;
    std::vector<JsonValue> arr;
    std::vector<std::pair<std::string, JsonValue>> obj;   // insertion order kept

    bool contains(const std::string& k) const { return find(k) != nullptr; }
    const JsonValue* find(const std::string& k) const {
        if (kind != Object) return nullptr;
        for (const auto& kv : obj)
            if (kv.first == k) return &kv.second;
        return nullptr;
    }
    const JsonValue& operator[](const std::string& k) const {
        const JsonValue* v = find(k);
        if (!v) throw std::runtime_error("json: no key '" + k + "'");
        return *v;
    }
    bool is_number_integer() const { return kind == Number && is_int; }
    bool is_number() const { return kind == Number; }

    template <typename T>
    T get() const;

    // obj.value(key, default): the default when the key is absent, else the typed value
    template <typename T>
    T value(const std::string& k, const T& dflt) const {
        const JsonValue* v = find(k);
        return v ? v->get<T>() : dflt;
    }
    std::string value(const std::string& k, const char* dflt) const { return value<std::string>(k, dflt); }
};

template <>
inline double JsonValue::get<double>() const {
    if (kind != Number) throw std::runtime_error("json: type must be number");
    return num;
}
template <>
inline int JsonValue::get<int>() const {
    if (kind != Number) throw std::runtime_error("json: type must be number");
    return (int)num;
}
template <>
inline bool JsonValue::get<bool>() const {
    if (kind != Bool) throw std::runtime_error("json: type must be boolean");
    return b;
}
template <>
inline std::string JsonValue::get<std::string>() const {
    if (kind != String) throw std::runtime_error("json: type must be string");
    return str;
}

class JsonParser {
public:
    explicit JsonParser(const std::string& s) : s_(s) {}
    JsonValue parse() {
        JsonValue v = value();
        ws();
        if (i_ != s_.size()) err("trailing characters");
        return v;
    }

private:
    const std::string& s_;
    size_t i_ = 0;
    [[noreturn]] void err(const std::string& m) const {
        throw std::runtime_error("json: " + m + " at offset " + std::to_string(i_));
    }
    void ws() {
        while (i_ < s_.size() && (s_[i_] == ' ' || s_[i_] == '\t' || s_[i_] == '\n' || s_[i_] == '\r')) ++i_;
    }
    bool lit(const char* w) {
        size_t n = std::char_traits<char>::length(w);
        if (s_.compare(i_, n, w) == 0) {
            i_ += n;
            return true;
        }
        return false;
    }
    static void utf8(std::string& o, unsigned cp) {
        if (cp < 0x80) {
            o += (char)cp;
        } else if (cp < 0x800) {
            o += (char)(0xC0 | (cp >> 6));
            o += (char)(0x80 | (cp & 0x3F));
        } else if (cp < 0x10000) {
            o += (char)(0xE0 | (cp >> 12));
            o += (char)(0x80 | ((cp >> 6) & 0x3F));
            o += (char)(0x80 | (cp & 0x3F));
        } else {
            o += (char)(0xF0 | (cp >> 18));
            o += (char)(0x80 | ((cp >> 12) & 0x3F));
            o += (char)(0x80 | ((cp >> 6) & 0x3F));
            o += (char)(0x80 | (cp & 0x3F));
        }
    }
    unsigned hex4() {
        if (i_ + 4 > s_.size()) err("bad \\u escape");
        unsigned v = 0;
        for (int k = 0; k < 4; ++k) {
            const char c = s_[i_++];
            v <<= 4;
            if (c >= '0' && c <= '9') v |= (unsigned)(c - '0');
            else if (c >= 'a' && c <= 'f') v |= (unsigned)(c - 'a' + 10);
            else if (c >= 'A' && c <= 'F') v |= (unsigned)(c - 'A' + 10);
            else err("bad \\u escape");
        }
        return v;
    }
    std::string string() {
        if (s_[i_] != '"') err("expected string");
        ++i_;
        std::string o;
        while (true) {
            if (i_ >= s_.size()) err("unterminated string");
            const char c = s_[i_++];
            if (c == '"') break;
            if (c != '\\') {
                o += c;
                continue;
            }
            if (i_ >= s_.size()) err("unterminated escape");
            const char e = s_[i_++];
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
                    unsigned cp = hex4();
                    if (cp >= 0xD800 && cp < 0xDC00 && s_.compare(i_, 2, "\\u") == 0) {
                        i_ += 2;
                        const unsigned lo = hex4();
                        cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
                    }
                    utf8(o, cp);
                    break;
                }
                default: err("bad escape");
            }
        }
        return o;
    }
    JsonValue value() {
        ws();
        if (i_ >= s_.size()) err("unexpected end");
        JsonValue v;
        const char c = s_[i_];
        if (c == '{') {
            v.kind = JsonValue::Object;
            ++i_;
            ws();
            if (i_ < s_.size() && s_[i_] == '}') {
                ++i_;
                return v;
            }
            while (true) {
                ws();
                std::string k = string();
                ws();
                if (i_ >= s_.size() || s_[i_] != ':') err("expected ':'");
                ++i_;
                JsonValue item = value();
                bool replaced = false;   // a repeated key keeps the last value
                for (auto& kv : v.obj)
                    if (kv.first == k) {
                        kv.second = item;
                        replaced = true;
                    }
                if (!replaced) v.obj.emplace_back(std::move(k), std::move(item));
                ws();
                if (i_ < s_.size() && s_[i_] == ',') {
                    ++i_;
                    continue;
                }
                if (i_ < s_.size() && s_[i_] == '}') {
                    ++i_;
                    break;
                }
                err("expected ',' or '}'");
            }
        } else if (c == '[') {
            v.kind = JsonValue::Array;
            ++i_;
            ws();
            if (i_ < s_.size() && s_[i_] == ']') {
                ++i_;
                return v;
            }
            while (true) {
                v.arr.push_back(value());
                ws();
                if (i_ < s_.size() && s_[i_] == ',') {
                    ++i_;
                    continue;
                }
                if (i_ < s_.size() && s_[i_] == ']') {
                    ++i_;
                    break;
                }
                err("expected ',' or ']'");
            }
        } else if (c == '"') {
            v.kind = JsonValue::String;
            v.str = string();
        } else if (lit("true")) {
            v.kind = JsonValue::Bool;
            v.b = true;
        } else if (lit("false")) {
            v.kind = JsonValue::Bool;
            v.b = false;
        } else if (lit("null")) {
            v.kind = JsonValue::Null;
        } else {
            const size_t st = i_;
            if (s_[i_] == '-') ++i_;
            bool frac = false;
            while (i_ < s_.size()) {
                const char d = s_[i_];
                if (d >= '0' && d <= '9') {
                    ++i_;
                } else if (d == '.' || d == 'e' || d == 'E' || d == '+' || d == '-') {
                    frac = true;
                    ++i_;
                } else {
                    break;
                }
            }
            if (i_ == st) err("unexpected character");
            const std::string tok = s_.substr(st, i_ - st);
            char* end = nullptr;
            v.kind = JsonValue::Number;
            v.num = std::strtod(tok.c_str(), &end);
            if (!end || *end) err("bad number '" + tok + "'");
            v.is_int = !frac;
        }
        return v;
    }
};

inline JsonValue parse_json(const std::string& text) { return JsonParser(text).parse(); }

}  // namespace zoomy

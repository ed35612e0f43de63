// Minimal JSON for the request/reply layer (rcwa_json.hpp): a value tree that keeps object keys in
// their written order, a strict RFC 8259 parser that reports line, column and byte offset of the first
// error, and a streaming writer.  Self-contained C++17, no device code.
//
// Writer format: with indent < 0 everything is on one line as {"key": value, "list": [1, 2]}; with
// indent >= 0 each member of a multi-line container starts on its own line, indented by indent spaces per
// level, while containers opened with one_line = true stay on one line.  Doubles are written with the
// fewest significant digits (general notation) that read back to the same double; NaN and +-Inf are
// written as null.
#pragma once

#include <charconv>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace srcwa_amd {
namespace json {

class Value {
public:
    enum class Type { Null, Bool, Number, String, Array, Object };

    Value() = default;
    static Value boolean(bool b) {
        Value v;
        v.type_ = Type::Bool;
        v.b_ = b;
        return v;
    }
    static Value number(double x, bool integral) {
        Value v;
        v.type_ = Type::Number;
        v.d_ = x;
        v.integral_ = integral;
        return v;
    }
    static Value string(std::string s) {
        Value v;
        v.type_ = Type::String;
        v.s_ = std::move(s);
        return v;
    }
    static Value array() {
        Value v;
        v.type_ = Type::Array;
        return v;
    }
    static Value object() {
        Value v;
        v.type_ = Type::Object;
        return v;
    }

    Type type() const { return type_; }
    bool is_null() const { return type_ == Type::Null; }
    bool is_bool() const { return type_ == Type::Bool; }
    bool is_number() const { return type_ == Type::Number; }
    bool is_string() const { return type_ == Type::String; }
    bool is_array() const { return type_ == Type::Array; }
    bool is_object() const { return type_ == Type::Object; }

    bool as_bool() const { return b_; }
    double as_double() const { return d_; }
    const std::string& as_string() const { return s_; }
    // An integer-valued number that fits int (written without fraction or exponent, or with a fraction
    // or exponent that still makes it a whole number).
    bool to_int(int& out) const {
        if (!is_number() || !std::isfinite(d_) || d_ != std::floor(d_) || d_ < -2147483648.0 || d_ > 2147483647.0)
            return false;
        out = static_cast<int>(d_);
        return true;
    }

    // arrays and objects
    std::size_t size() const { return items_.size(); }
    const Value& at(std::size_t i) const { return items_[i].second; }
    const std::string& key(std::size_t i) const { return items_[i].first; }
    const Value* find(const std::string& k) const {
        for (const auto& it : items_)
            if (it.first == k) return &it.second;
        return nullptr;
    }
    bool contains(const std::string& k) const { return find(k) != nullptr; }
    const Value& at(const std::string& k) const {
        const Value* v = find(k);
        if (!v) throw std::out_of_range("json: no member " + k);
        return *v;
    }
    void push(Value v) { items_.emplace_back(std::string(), std::move(v)); }
    void put(std::string k, Value v) { items_.emplace_back(std::move(k), std::move(v)); }

private:
    Type type_ = Type::Null;
    bool b_ = false;
    double d_ = 0.0;
    bool integral_ = false;
    std::string s_;
    std::vector<std::pair<std::string, Value>> items_;
};

class ParseError : public std::runtime_error {
public:
    ParseError(const std::string& reason, std::size_t line, std::size_t column, std::size_t offset)
        : std::runtime_error(reason), reason_(reason), line_(line), column_(column), offset_(offset) {}
    const std::string& reason() const { return reason_; }
    std::size_t line() const { return line_; }       // 1-based
    std::size_t column() const { return column_; }   // 1-based, in bytes
    std::size_t offset() const { return offset_; }   // 0-based byte offset

private:
    std::string reason_;
    std::size_t line_, column_, offset_;
};

namespace detail {

constexpr int kMaxDepth = 512;

class Parser {
public:
    explicit Parser(const std::string& text) : t_(text) {}

    Value document() {
        if (t_.size() >= 3 && static_cast<unsigned char>(t_[0]) == 0xEF && static_cast<unsigned char>(t_[1]) == 0xBB &&
            static_cast<unsigned char>(t_[2]) == 0xBF)
            i_ = 3;
        space();
        Value v = value(0);
        space();
        if (i_ != t_.size()) error("unexpected text after the JSON value");
        return v;
    }

private:
    [[noreturn]] void error(const std::string& why) const {
        std::size_t line = 1, col = 1;
        for (std::size_t k = 0; k < i_ && k < t_.size(); ++k) {
            if (t_[k] == '\n') {
                ++line;
                col = 1;
            } else {
                ++col;
            }
        }
        throw ParseError(why, line, col, i_);
    }
    bool end() const { return i_ >= t_.size(); }
    char peek() const { return end() ? '\0' : t_[i_]; }
    void space() {
        while (!end() && (t_[i_] == ' ' || t_[i_] == '\t' || t_[i_] == '\n' || t_[i_] == '\r')) ++i_;
    }
    void expect_word(const char* w) {
        for (const char* p = w; *p; ++p) {
            if (peek() != *p) error(std::string("invalid literal (expected ") + w + ")");
            ++i_;
        }
    }

    Value value(int depth) {
        if (depth > kMaxDepth) error("nesting too deep");
        if (end()) error("unexpected end of input");
        const char c = peek();
        if (c == '{') return object(depth);
        if (c == '[') return array(depth);
        if (c == '"') return Value::string(string());
        if (c == 't') {
            expect_word("true");
            return Value::boolean(true);
        }
        if (c == 'f') {
            expect_word("false");
            return Value::boolean(false);
        }
        if (c == 'n') {
            expect_word("null");
            return Value();
        }
        if (c == '-' || (c >= '0' && c <= '9')) return number();
        error("unexpected character");
    }

    Value object(int depth) {
        ++i_;
        Value v = Value::object();
        space();
        if (peek() == '}') {
            ++i_;
            return v;
        }
        for (;;) {
            space();
            if (peek() != '"') error("expected a string key");
            const std::size_t key_at = i_;
            std::string k = string();
            if (v.contains(k)) {
                i_ = key_at;
                error("duplicate key \"" + k + "\"");
            }
            space();
            if (peek() != ':') error("expected ':' after the key");
            ++i_;
            space();
            Value item = value(depth + 1);
            v.put(std::move(k), std::move(item));
            space();
            if (peek() == ',') {
                ++i_;
                continue;
            }
            if (peek() == '}') {
                ++i_;
                return v;
            }
            error("expected ',' or '}' in an object");
        }
    }

    Value array(int depth) {
        ++i_;
        Value v = Value::array();
        space();
        if (peek() == ']') {
            ++i_;
            return v;
        }
        for (;;) {
            space();
            v.push(value(depth + 1));
            space();
            if (peek() == ',') {
                ++i_;
                continue;
            }
            if (peek() == ']') {
                ++i_;
                return v;
            }
            error("expected ',' or ']' in an array");
        }
    }

    Value number() {
        const std::size_t start = i_;
        bool integral = true;
        if (peek() == '-') ++i_;
        if (peek() == '0') {
            ++i_;
        } else if (peek() >= '1' && peek() <= '9') {
            while (peek() >= '0' && peek() <= '9') ++i_;
        } else {
            error("invalid number");
        }
        if (peek() == '.') {
            integral = false;
            ++i_;
            if (!(peek() >= '0' && peek() <= '9')) error("invalid number: digits must follow '.'");
            while (peek() >= '0' && peek() <= '9') ++i_;
        }
        if (peek() == 'e' || peek() == 'E') {
            integral = false;
            ++i_;
            if (peek() == '+' || peek() == '-') ++i_;
            if (!(peek() >= '0' && peek() <= '9')) error("invalid number: digits must follow the exponent");
            while (peek() >= '0' && peek() <= '9') ++i_;
        }
        double x = 0.0;
        const char* first = t_.data() + start;
        const char* last = t_.data() + i_;
        const std::from_chars_result r = std::from_chars(first, last, x);
        if (r.ec == std::errc::result_out_of_range) {
            // a magnitude beyond double would read as +-inf, which no JSON text can stand for
            i_ = start;
            error("number out of the range of a double");
        }
        if (r.ec != std::errc() || r.ptr != last) {
            i_ = start;
            error("invalid number");
        }
        return Value::number(x, integral);
    }

    static void put_utf8(std::string& out, std::uint32_t cp) {
        if (cp < 0x80) {
            out.push_back(static_cast<char>(cp));
        } else if (cp < 0x800) {
            out.push_back(static_cast<char>(0xC0 | (cp >> 6)));
            out.push_back(static_cast<char>(0x80 | (cp & 0x3F)));
        } else if (cp < 0x10000) {
            out.push_back(static_cast<char>(0xE0 | (cp >> 12)));
            out.push_back(static_cast<char>(0x80 | ((cp >> 6) & 0x3F)));
            out.push_back(static_cast<char>(0x80 | (cp & 0x3F)));
        } else {
            out.push_back(static_cast<char>(0xF0 | (cp >> 18)));
            out.push_back(static_cast<char>(0x80 | ((cp >> 12) & 0x3F)));
            out.push_back(static_cast<char>(0x80 | ((cp >> 6) & 0x3F)));
            out.push_back(static_cast<char>(0x80 | (cp & 0x3F)));
        }
    }
    std::uint32_t hex4() {
        std::uint32_t v = 0;
        for (int k = 0; k < 4; ++k) {
            const char c = peek();
            v <<= 4;
            if (c >= '0' && c <= '9') v |= static_cast<std::uint32_t>(c - '0');
            else if (c >= 'a' && c <= 'f') v |= static_cast<std::uint32_t>(c - 'a' + 10);
            else if (c >= 'A' && c <= 'F') v |= static_cast<std::uint32_t>(c - 'A' + 10);
            else error("invalid \\u escape");
            ++i_;
        }
        return v;
    }
    // one UTF-8 sequence starting at i_, checked for shortest form and the scalar range
    void copy_utf8(std::string& out) {
        const unsigned char c = static_cast<unsigned char>(t_[i_]);
        int len = 0;
        std::uint32_t cp = 0, min = 0;
        if (c < 0x80) {
            out.push_back(static_cast<char>(c));
            ++i_;
            return;
        } else if ((c & 0xE0) == 0xC0) {
            len = 2;
            cp = c & 0x1F;
            min = 0x80;
        } else if ((c & 0xF0) == 0xE0) {
            len = 3;
            cp = c & 0x0F;
            min = 0x800;
        } else if ((c & 0xF8) == 0xF0) {
            len = 4;
            cp = c & 0x07;
            min = 0x10000;
        } else {
            error("invalid UTF-8 in a string");
        }
        if (i_ + len > t_.size()) error("truncated UTF-8 in a string");
        for (int k = 1; k < len; ++k) {
            const unsigned char d = static_cast<unsigned char>(t_[i_ + k]);
            if ((d & 0xC0) != 0x80) error("invalid UTF-8 in a string");
            cp = (cp << 6) | (d & 0x3F);
        }
        if (cp < min || cp > 0x10FFFF || (cp >= 0xD800 && cp <= 0xDFFF)) error("invalid UTF-8 in a string");
        out.append(t_, i_, static_cast<std::size_t>(len));
        i_ += static_cast<std::size_t>(len);
    }
    std::string string() {
        ++i_;   // opening quote
        std::string out;
        for (;;) {
            if (end()) error("unterminated string");
            const char c = t_[i_];
            if (c == '"') {
                ++i_;
                return out;
            }
            if (static_cast<unsigned char>(c) < 0x20) error("control character in a string");
            if (c != '\\') {
                copy_utf8(out);
                continue;
            }
            ++i_;
            const char e = peek();
            ++i_;
            switch (e) {
            case '"': out.push_back('"'); break;
            case '\\': out.push_back('\\'); break;
            case '/': out.push_back('/'); break;
            case 'b': out.push_back('\b'); break;
            case 'f': out.push_back('\f'); break;
            case 'n': out.push_back('\n'); break;
            case 'r': out.push_back('\r'); break;
            case 't': out.push_back('\t'); break;
            case 'u': {
                std::uint32_t cp = hex4();
                if (cp >= 0xD800 && cp <= 0xDBFF) {
                    if (peek() != '\\') error("unpaired surrogate in a \\u escape");
                    ++i_;
                    if (peek() != 'u') error("unpaired surrogate in a \\u escape");
                    ++i_;
                    const std::uint32_t lo = hex4();
                    if (lo < 0xDC00 || lo > 0xDFFF) error("unpaired surrogate in a \\u escape");
                    cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
                } else if (cp >= 0xDC00 && cp <= 0xDFFF) {
                    error("unpaired surrogate in a \\u escape");
                }
                put_utf8(out, cp);
                break;
            }
            default:
                --i_;
                error("invalid escape in a string");
            }
        }
    }

    const std::string& t_;
    std::size_t i_ = 0;
};

inline void append_string(std::string& out, const std::string& s) {
    static const char* hex = "0123456789abcdef";
    out.push_back('"');
    for (const char c : s) {
        switch (c) {
        case '"': out += "\\\""; break;
        case '\\': out += "\\\\"; break;
        case '\b': out += "\\b"; break;
        case '\f': out += "\\f"; break;
        case '\n': out += "\\n"; break;
        case '\r': out += "\\r"; break;
        case '\t': out += "\\t"; break;
        default:
            if (static_cast<unsigned char>(c) < 0x20) {
                out += "\\u00";
                out.push_back(hex[(c >> 4) & 0xF]);
                out.push_back(hex[c & 0xF]);
            } else {
                out.push_back(c);
            }
        }
    }
    out.push_back('"');
}

// The fewest significant digits (general notation) that read back to x.
inline void append_double(std::string& out, double x) {
    if (!std::isfinite(x)) {
        out += "null";
        return;
    }
    char buf[64];
    for (int digits = 1; digits <= 17; ++digits) {
        const std::to_chars_result r = std::to_chars(buf, buf + sizeof(buf), x, std::chars_format::general, digits);
        double back = 0.0;
        std::from_chars(buf, r.ptr, back);
        if (back == x || digits == 17) {
            out.append(buf, r.ptr);
            return;
        }
    }
}

}  // namespace detail

inline Value parse(const std::string& text) { return detail::Parser(text).document(); }

class Writer {
public:
    explicit Writer(int indent = -1) : indent_(indent) {}

    void begin_object(bool one_line = false) { open('{', true, one_line); }
    void end_object() { close('}', true); }
    void begin_array(bool one_line = false) { open('[', false, one_line); }
    void end_array() { close(']', false); }
    void key(const std::string& k) {
        if (stack_.empty() || !stack_.back().object || pending_key_) throw std::logic_error("json::Writer: misplaced key");
        separate();
        detail::append_string(out_, k);
        out_ += ": ";
        pending_key_ = true;
    }
    void null() {
        before_value();
        out_ += "null";
    }
    void value(bool b) {
        before_value();
        out_ += b ? "true" : "false";
    }
    void value(int x) { value(static_cast<long long>(x)); }
    void value(long long x) {
        before_value();
        out_ += std::to_string(x);
    }
    void value(double x) {
        before_value();
        detail::append_double(out_, x);
    }
    void value(const std::string& s) {
        before_value();
        detail::append_string(out_, s);
    }
    void value(const char* s) { value(std::string(s)); }
    std::string take() {
        std::string s = std::move(out_);
        out_.clear();
        stack_.clear();
        pending_key_ = false;
        return s;
    }

private:
    struct Level {
        bool object, flat, empty;
    };
    void newline(std::size_t depth) {
        out_.push_back('\n');
        out_.append(depth * static_cast<std::size_t>(indent_), ' ');
    }
    void separate() {
        Level& l = stack_.back();
        if (!l.empty) out_.push_back(',');
        if (!l.flat) newline(stack_.size());
        else if (!l.empty) out_.push_back(' ');
        l.empty = false;
    }
    void before_value() {
        if (stack_.empty()) return;
        if (stack_.back().object) {
            if (!pending_key_) throw std::logic_error("json::Writer: a value in an object needs a key");
            pending_key_ = false;
        } else {
            separate();
        }
    }
    void open(char c, bool object, bool one_line) {
        before_value();
        const bool flat = indent_ < 0 || one_line || (!stack_.empty() && stack_.back().flat);
        stack_.push_back(Level{object, flat, true});
        out_.push_back(c);
    }
    void close(char c, bool object) {
        if (stack_.empty() || stack_.back().object != object || pending_key_)
            throw std::logic_error("json::Writer: mismatched close");
        const Level l = stack_.back();
        stack_.pop_back();
        if (!l.flat && !l.empty) newline(stack_.size());
        out_.push_back(c);
    }

    int indent_;
    std::string out_;
    std::vector<Level> stack_;
    bool pending_key_ = false;
};

}  // namespace json
}  // namespace srcwa_amd

// json_lite.h -- minimal JSON reader for Aurora scene files.
//
// Object members are kept in a std::map, so iteration visits keys in sorted
// order.  That is the order in which the reference (jsoncpp 0.5.0, whose
// objects are std::map-backed) walks "objects", "materials" and "shaders"
// (src/translator/jsonParser.cpp:507-531, 580-617, 619-661), and it fixes the
// global triangle numbering.  Arrays keep document order.
#pragma once

#include <cstdlib>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace aur {

struct JsonValue {
    enum Kind { Null, Bool, Number, String, Array, Object } kind = Null;
    bool b = false;
    double num = 0.0;
    std::string str;
    std::vector<JsonValue> arr;
    std::map<std::string, JsonValue> obj;

    bool has(const std::string &k) const { return kind == Object && obj.count(k) != 0; }
    const JsonValue &at(const std::string &k) const {
        auto it = obj.find(k);
        if (kind != Object || it == obj.end()) throw std::runtime_error("missing json attribute: " + k);
        return it->second;
    }
    double as_double() const {
        if (kind == Number) return num;
        if (kind == Bool) return b ? 1.0 : 0.0;
        throw std::runtime_error("json value is not a number");
    }
    int as_int() const { return (int)as_double(); }
    const std::string &as_string() const {
        if (kind != String) throw std::runtime_error("json value is not a string");
        return str;
    }
};

class JsonReader {
public:
    explicit JsonReader(const std::string &text) : s_(text), i_(0) {}

    JsonValue parse() {
        skip();
        JsonValue v = value();
        skip();
        if (i_ != s_.size()) fail("trailing characters");
        return v;
    }

private:
    const std::string &s_;
    size_t i_;

    [[noreturn]] void fail(const char *what) {
        throw std::runtime_error(std::string("json parse error at offset ") + std::to_string(i_) + ": " + what);
    }
    void skip() {
        for (;;) {
            while (i_ < s_.size() && (s_[i_] == ' ' || s_[i_] == '\t' || s_[i_] == '\n' || s_[i_] == '\r')) ++i_;
            // jsoncpp accepts C/C++ comments; PyAurora's exporter writes a "//" header
            if (i_ + 1 < s_.size() && s_[i_] == '/' && s_[i_ + 1] == '/') {
                while (i_ < s_.size() && s_[i_] != '\n') ++i_;
                continue;
            }
            if (i_ + 1 < s_.size() && s_[i_] == '/' && s_[i_ + 1] == '*') {
                size_t e = s_.find("*/", i_ + 2);
                if (e == std::string::npos) fail("unterminated comment");
                i_ = e + 2;
                continue;
            }
            break;
        }
    }
    JsonValue value() {
        if (i_ >= s_.size()) fail("unexpected end");
        char c = s_[i_];
        if (c == '{') return object();
        if (c == '[') return array();
        if (c == '"') { JsonValue v; v.kind = JsonValue::String; v.str = string(); return v; }
        if (c == 't' && s_.compare(i_, 4, "true") == 0) { i_ += 4; JsonValue v; v.kind = JsonValue::Bool; v.b = true; return v; }
        if (c == 'f' && s_.compare(i_, 5, "false") == 0) { i_ += 5; JsonValue v; v.kind = JsonValue::Bool; return v; }
        if (c == 'n' && s_.compare(i_, 4, "null") == 0) { i_ += 4; return JsonValue(); }
        return number();
    }
    JsonValue number() {
        const char *start = s_.c_str() + i_;
        char *end = nullptr;
        double d = std::strtod(start, &end);
        if (end == start) fail("bad value");
        i_ += (size_t)(end - start);
        JsonValue v; v.kind = JsonValue::Number; v.num = d;
        return v;
    }
    std::string string() {
        std::string out;
        ++i_;
        while (i_ < s_.size() && s_[i_] != '"') {
            char c = s_[i_++];
            if (c == '\\') {
                if (i_ >= s_.size()) fail("bad escape");
                char e = s_[i_++];
                switch (e) {
                case 'n': out += '\n'; break;
                case 't': out += '\t'; break;
                case 'r': out += '\r'; break;
                case 'b': out += '\b'; break;
                case 'f': out += '\f'; break;
                case 'u': {
                    if (i_ + 4 > s_.size()) fail("bad unicode escape");
                    unsigned cp = (unsigned)std::strtoul(s_.substr(i_, 4).c_str(), nullptr, 16);
                    i_ += 4;
                    if (cp < 0x80) out += (char)cp;
                    else if (cp < 0x800) { out += (char)(0xC0 | (cp >> 6)); out += (char)(0x80 | (cp & 0x3F)); }
                    else { out += (char)(0xE0 | (cp >> 12)); out += (char)(0x80 | ((cp >> 6) & 0x3F)); out += (char)(0x80 | (cp & 0x3F)); }
                    break;
                }
                default: out += e;
                }
            } else {
                out += c;
            }
        }
        if (i_ >= s_.size()) fail("unterminated string");
        ++i_;
        return out;
    }
    JsonValue array() {
        JsonValue v; v.kind = JsonValue::Array;
        ++i_; skip();
        if (i_ < s_.size() && s_[i_] == ']') { ++i_; return v; }
        for (;;) {
            skip();
            v.arr.push_back(value());
            skip();
            if (i_ >= s_.size()) fail("unterminated array");
            if (s_[i_] == ',') { ++i_; continue; }
            if (s_[i_] == ']') { ++i_; return v; }
            fail("expected , or ]");
        }
    }
    JsonValue object() {
        JsonValue v; v.kind = JsonValue::Object;
        ++i_; skip();
        if (i_ < s_.size() && s_[i_] == '}') { ++i_; return v; }
        for (;;) {
            skip();
            if (i_ >= s_.size() || s_[i_] != '"') fail("expected key");
            std::string k = string();
            skip();
            if (i_ >= s_.size() || s_[i_] != ':') fail("expected :");
            ++i_; skip();
            v.obj[k] = value();
            skip();
            if (i_ >= s_.size()) fail("unterminated object");
            if (s_[i_] == ',') { ++i_; continue; }
            if (s_[i_] == '}') { ++i_; return v; }
            fail("expected , or }");
        }
    }
};

}  // namespace aur

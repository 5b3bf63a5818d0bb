// Minimal JSON value, parser and serializer for the calculator hyper-parameters.
//
// featomic parses hyper-parameters with serde_json (featomic/src/calculator.rs:556-588);
// this is a small self-contained replacement: objects keep their insertion order,
// numbers are doubles, and parse errors carry the byte offset.
#pragma once

#include <cmath>
#include <cstdio>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace rascaline_amd {

struct JsonError : std::runtime_error {
    explicit JsonError(const std::string& what) : std::runtime_error(what) {}
};

class Json {
public:
    enum class Kind { Null, Bool, Number, String, Array, Object };

    Json() = default;
    static Json null() { return Json(); }
    static Json boolean(bool b) { Json j; j.kind_ = Kind::Bool; j.bool_ = b; return j; }
    static Json number(double v) { Json j; j.kind_ = Kind::Number; j.number_ = v; return j; }
    static Json string(std::string s) { Json j; j.kind_ = Kind::String; j.string_ = std::move(s); return j; }
    static Json array() { Json j; j.kind_ = Kind::Array; return j; }
    static Json object() { Json j; j.kind_ = Kind::Object; return j; }

    Kind kind() const { return kind_; }
    bool is_null() const { return kind_ == Kind::Null; }
    bool is_object() const { return kind_ == Kind::Object; }
    bool is_number() const { return kind_ == Kind::Number; }
    bool is_string() const { return kind_ == Kind::String; }
    bool is_array() const { return kind_ == Kind::Array; }

    bool as_bool() const {
        if (kind_ != Kind::Bool) throw JsonError("expected a boolean");
        return bool_;
    }
    double as_number() const {
        if (kind_ != Kind::Number) throw JsonError("expected a number");
        return number_;
    }
    const std::string& as_string() const {
        if (kind_ != Kind::String) throw JsonError("expected a string");
        return string_;
    }
    const std::vector<Json>& as_array() const {
        if (kind_ != Kind::Array) throw JsonError("expected an array");
        return array_;
    }
    const std::vector<std::pair<std::string, Json>>& as_object() const {
        if (kind_ != Kind::Object) throw JsonError("expected an object");
        return object_;
    }

    const Json* find(const std::string& key) const {
        if (kind_ != Kind::Object) return nullptr;
        for (auto& kv : object_) {
            if (kv.first == key) return &kv.second;
        }
        return nullptr;
    }

    void push(Json v) { array_.push_back(std::move(v)); }
    void set(const std::string& key, Json v) {
        for (auto& kv : object_) {
            if (kv.first == key) { kv.second = std::move(v); return; }
        }
        object_.emplace_back(key, std::move(v));
    }

    static Json parse(const std::string& text) {
        size_t pos = 0;
        Json value = parse_value(text, pos);
        skip_ws(text, pos);
        if (pos != text.size()) fail(text, pos, "trailing characters");
        return value;
    }

    std::string dump() const {
        std::string out;
        dump_to(out);
        return out;
    }

private:
    Kind kind_ = Kind::Null;
    bool bool_ = false;
    double number_ = 0.0;
    std::string string_;
    std::vector<Json> array_;
    std::vector<std::pair<std::string, Json>> object_;

    [[noreturn]] static void fail(const std::string& text, size_t pos, const char* msg) {
        (void)text;
        throw JsonError(std::string(msg) + " at byte " + std::to_string(pos));
    }

    static void skip_ws(const std::string& t, size_t& p) {
        while (p < t.size() && (t[p] == ' ' || t[p] == '\n' || t[p] == '\t' || t[p] == '\r')) p++;
    }

    static Json parse_value(const std::string& t, size_t& p) {
        skip_ws(t, p);
        if (p >= t.size()) fail(t, p, "unexpected end of input");
        char c = t[p];
        if (c == '{') {
            Json obj = object();
            p++;
            skip_ws(t, p);
            if (p < t.size() && t[p] == '}') { p++; return obj; }
            for (;;) {
                skip_ws(t, p);
                if (p >= t.size() || t[p] != '"') fail(t, p, "expected a string key");
                std::string key = parse_string(t, p);
                skip_ws(t, p);
                if (p >= t.size() || t[p] != ':') fail(t, p, "expected ':'");
                p++;
                Json v = parse_value(t, p);
                if (obj.find(key)) fail(t, p, "duplicate key");
                obj.object_.emplace_back(key, std::move(v));
                skip_ws(t, p);
                if (p < t.size() && t[p] == ',') { p++; continue; }
                if (p < t.size() && t[p] == '}') { p++; break; }
                fail(t, p, "expected ',' or '}'");
            }
            return obj;
        }
        if (c == '[') {
            Json arr = array();
            p++;
            skip_ws(t, p);
            if (p < t.size() && t[p] == ']') { p++; return arr; }
            for (;;) {
                arr.array_.push_back(parse_value(t, p));
                skip_ws(t, p);
                if (p < t.size() && t[p] == ',') { p++; continue; }
                if (p < t.size() && t[p] == ']') { p++; break; }
                fail(t, p, "expected ',' or ']'");
            }
            return arr;
        }
        if (c == '"') return string(parse_string(t, p));
        if (t.compare(p, 4, "true") == 0) { p += 4; return boolean(true); }
        if (t.compare(p, 5, "false") == 0) { p += 5; return boolean(false); }
        if (t.compare(p, 4, "null") == 0) { p += 4; return null(); }
        if (c == '-' || (c >= '0' && c <= '9')) {
            size_t start = p;
            if (t[p] == '-') p++;
            while (p < t.size() && ((t[p] >= '0' && t[p] <= '9') || t[p] == '.' || t[p] == 'e' ||
                                    t[p] == 'E' || t[p] == '+' || t[p] == '-')) p++;
            std::string s = t.substr(start, p - start);
            char* end = nullptr;
            double v = std::strtod(s.c_str(), &end);
            if (end != s.c_str() + s.size()) fail(t, start, "invalid number");
            return number(v);
        }
        fail(t, p, "unexpected character");
    }

    static std::string parse_string(const std::string& t, size_t& p) {
        std::string out;
        p++;  // opening quote
        while (p < t.size() && t[p] != '"') {
            char c = t[p++];
            if (c == '\\') {
                if (p >= t.size()) fail(t, p, "bad escape");
                char e = t[p++];
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
                    if (p + 4 > t.size()) fail(t, p, "bad unicode escape");
                    unsigned code = std::stoul(t.substr(p, 4), nullptr, 16);
                    p += 4;
                    if (code < 0x80) {
                        out += static_cast<char>(code);
                    } else if (code < 0x800) {
                        out += static_cast<char>(0xC0 | (code >> 6));
                        out += static_cast<char>(0x80 | (code & 0x3F));
                    } else {
                        out += static_cast<char>(0xE0 | (code >> 12));
                        out += static_cast<char>(0x80 | ((code >> 6) & 0x3F));
                        out += static_cast<char>(0x80 | (code & 0x3F));
                    }
                    break;
                }
                default: fail(t, p, "bad escape");
                }
            } else {
                out += c;
            }
        }
        if (p >= t.size()) fail(t, p, "unterminated string");
        p++;
        return out;
    }

    void dump_to(std::string& out) const {
        switch (kind_) {
        case Kind::Null: out += "null"; break;
        case Kind::Bool: out += bool_ ? "true" : "false"; break;
        case Kind::Number: {
            if (std::isfinite(number_) && number_ == std::floor(number_) && std::fabs(number_) < 1e15) {
                char buf[64];
                std::snprintf(buf, sizeof(buf), "%.1f", number_);
                // serde prints integers stored as f64 with a trailing ".0"
                out += buf;
            } else {
                char buf[64];
                std::snprintf(buf, sizeof(buf), "%.17g", number_);
                out += buf;
            }
            break;
        }
        case Kind::String: {
            out += '"';
            for (char c : string_) {
                if (c == '"' || c == '\\') { out += '\\'; out += c; }
                else if (c == '\n') out += "\\n";
                else out += c;
            }
            out += '"';
            break;
        }
        case Kind::Array: {
            out += '[';
            for (size_t i = 0; i < array_.size(); i++) {
                if (i) out += ',';
                array_[i].dump_to(out);
            }
            out += ']';
            break;
        }
        case Kind::Object: {
            out += '{';
            for (size_t i = 0; i < object_.size(); i++) {
                if (i) out += ',';
                out += '"';
                out += object_[i].first;
                out += "\":";
                object_[i].second.dump_to(out);
            }
            out += '}';
            break;
        }
        }
    }
};

}  // namespace rascaline_amd

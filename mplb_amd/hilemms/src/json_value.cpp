// json_value.cpp -- recursive-descent JSON reader for the case files.
#include "json_value.h"

#include <cctype>
#include <cstdlib>

namespace mplb_json {

namespace {
const Value& null_value() {
    static const Value v;
    return v;
}
}  // namespace

const Value& Value::operator[](const std::string& key) const {
    if (kind_ != Object) return null_value();
    auto it = obj_.find(key);
    return it == obj_.end() ? null_value() : it->second;
}
const Value& Value::operator[](size_t i) const {
    if (kind_ != Array || i >= arr_.size()) return null_value();
    return arr_[i];
}
double Value::number() const {
    if (kind_ == Bool) return b_ ? 1.0 : 0.0;
    if (kind_ != Number) throw ParseError("JSON value is not a number");
    return n_;
}
bool Value::boolean() const {
    if (kind_ == Number) return n_ != 0.0;
    if (kind_ != Bool) throw ParseError("JSON value is not a boolean");
    return b_;
}
const std::string& Value::str() const {
    if (kind_ != String) throw ParseError("JSON value is not a string");
    return s_;
}
const std::vector<Value>& Value::array() const {
    if (kind_ != Array) throw ParseError("JSON value is not an array");
    return arr_;
}

class Parser {
   public:
    explicit Parser(const std::string& t) : t_(t) {}
    Value document() {
        Value v = value();
        ws();
        if (p_ != t_.size()) fail("trailing characters");
        return v;
    }

   private:
    const std::string& t_;
    size_t p_{0};
    [[noreturn]] void fail(const char* what) {
        throw ParseError(std::string("JSON: ") + what + " at offset " + std::to_string(p_));
    }
    void ws() {
        while (p_ < t_.size() && std::isspace((unsigned char)t_[p_])) p_++;
    }
    bool eat(char c) {
        ws();
        if (p_ < t_.size() && t_[p_] == c) {
            p_++;
            return true;
        }
        return false;
    }
    void expect(char c) {
        if (!eat(c)) fail("unexpected character");
    }
    std::string string_lit() {
        expect('"');
        std::string out;
        while (p_ < t_.size() && t_[p_] != '"') {
            char c = t_[p_++];
            if (c == '\\') {
                if (p_ >= t_.size()) fail("bad escape");
                char e = t_[p_++];
                switch (e) {
                    case 'n': out += '\n'; break;
                    case 't': out += '\t'; break;
                    case 'r': out += '\r'; break;
                    case 'b': out += '\b'; break;
                    case 'f': out += '\f'; break;
                    case 'u': {
                        if (p_ + 4 > t_.size()) fail("bad \\u escape");
                        unsigned cp = (unsigned)std::strtoul(t_.substr(p_, 4).c_str(), nullptr, 16);
                        p_ += 4;
                        if (cp < 0x80) out += (char)cp;
                        else if (cp < 0x800) {
                            out += (char)(0xC0 | (cp >> 6));
                            out += (char)(0x80 | (cp & 0x3F));
                        } else {
                            out += (char)(0xE0 | (cp >> 12));
                            out += (char)(0x80 | ((cp >> 6) & 0x3F));
                            out += (char)(0x80 | (cp & 0x3F));
                        }
                    } break;
                    default: out += e;
                }
            } else {
                out += c;
            }
        }
        if (p_ >= t_.size()) fail("unterminated string");
        p_++;
        return out;
    }
    Value value() {
        ws();
        if (p_ >= t_.size()) fail("unexpected end");
        Value v;
        const char c = t_[p_];
        if (c == '{') {
            p_++;
            v.kind_ = Value::Object;
            if (eat('}')) return v;
            do {
                ws();
                std::string k = string_lit();
                expect(':');
                v.obj_[k] = value();
            } while (eat(','));
            expect('}');
        } else if (c == '[') {
            p_++;
            v.kind_ = Value::Array;
            if (eat(']')) return v;
            do {
                v.arr_.push_back(value());
            } while (eat(','));
            expect(']');
        } else if (c == '"') {
            v.kind_ = Value::String;
            v.s_ = string_lit();
        } else if (t_.compare(p_, 4, "true") == 0) {
            p_ += 4;
            v.kind_ = Value::Bool;
            v.b_ = true;
        } else if (t_.compare(p_, 5, "false") == 0) {
            p_ += 5;
            v.kind_ = Value::Bool;
        } else if (t_.compare(p_, 4, "null") == 0) {
            p_ += 4;
        } else {
            const char* s = t_.c_str() + p_;
            char* e = nullptr;
            v.n_ = std::strtod(s, &e);
            if (e == s) fail("bad value");
            p_ += (size_t)(e - s);
            v.kind_ = Value::Number;
        }
        return v;
    }
};

Value Value::parse(const std::string& text) { return Parser(text).document(); }

}  // namespace mplb_json

// json_value.h -- a small JSON document model and reader for the HiLeMMS
// case files (objects, arrays, strings, numbers, true/false/null).
#ifndef MPLB_JSON_VALUE_H
#define MPLB_JSON_VALUE_H
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace mplb_json {

class Value {
   public:
    enum Kind { Null, Bool, Number, String, Array, Object };
    Value() = default;
    Kind kind() const { return kind_; }
    bool is_null() const { return kind_ == Null; }
    bool contains(const std::string& key) const {
        return kind_ == Object && obj_.count(key) != 0;
    }
    const Value& operator[](const std::string& key) const;
    const Value& operator[](size_t i) const;
    size_t size() const { return kind_ == Array ? arr_.size() : obj_.size(); }
    double number() const;
    bool boolean() const;
    const std::string& str() const;
    const std::vector<Value>& array() const;
    const std::map<std::string, Value>& object() const { return obj_; }

    static Value parse(const std::string& text);

   private:
    friend class Parser;
    Kind kind_{Null};
    bool b_{false};
    double n_{0};
    std::string s_;
    std::vector<Value> arr_;
    std::map<std::string, Value> obj_;
};

class ParseError : public std::runtime_error {
   public:
    using std::runtime_error::runtime_error;
};

}  // namespace mplb_json
#endif

// json.h -- minimal JSON reader for the condition files.
//
// Values: null, bool, number (double), string, array, object.  Object key
// order is kept.  Parse errors throw std::runtime_error with the offset.
#pragma once
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace gample {

struct Json {
    enum Type { Null, Bool, Number, String, Array, Object } type = Null;
    bool b = false;
    double num = 0.0;
    std::string str;
    std::vector<Json> arr;
    std::vector<std::pair<std::string, Json>> obj;

    bool is_object() const { return type == Object; }
    bool is_array() const { return type == Array; }
    bool is_number() const { return type == Number; }
    bool is_string() const { return type == String; }
    const Json *find(const std::string &k) const
    {
        for (auto &kv : obj)
            if (kv.first == k) return &kv.second;
        return nullptr;
    }
    // overwrite-or-append (NSMutableDictionary addEntriesFromDictionary:)
    void set(const std::string &k, const Json &v)
    {
        for (auto &kv : obj)
            if (kv.first == k) { kv.second = v; return; }
        obj.emplace_back(k, v);
    }
    static Json number(double v) { Json j; j.type = Number; j.num = v; return j; }
    static Json string(const std::string &s) { Json j; j.type = String; j.str = s; return j; }
    static Json array(const std::vector<double> &v)
    {
        Json j; j.type = Array;
        for (double x : v) j.arr.push_back(number(x));
        return j;
    }
    static Json object() { Json j; j.type = Object; return j; }
};

Json json_parse(const std::string &text);
Json json_load(const std::string &path);

}  // namespace gample

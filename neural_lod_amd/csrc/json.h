// Minimal JSON reader for tcnn-style model configurations (objects, arrays,
// numbers, strings, booleans, null).  Keys are looked up case-sensitively,
// string values that name types are compared case-insensitively by callers.
#pragma once

#include <cctype>
#include <cstdlib>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "nlod_common.h"

namespace nlod {

struct Json {
    enum Type { Null, Bool, Number, String, Array, Object } type = Null;
    bool b = false;
    double num = 0.0;
    std::string str;
    std::vector<Json> arr;
    std::vector<std::pair<std::string, Json>> obj;  // insertion order kept

    bool is_object() const { return type == Object; }
    bool is_array() const { return type == Array; }
    bool contains(const std::string& k) const {
        if (type != Object) return false;
        for (auto& kv : obj)
            if (kv.first == k) return true;
        return false;
    }
    const Json& at(const std::string& k) const {
        for (auto& kv : obj)
            if (kv.first == k) return kv.second;
        fail(NLOD_ERR_CONFIG, "missing key '" + k + "'");
    }
    double number() const {
        if (type == Number) return num;
        if (type == Bool) return b ? 1.0 : 0.0;
        fail(NLOD_ERR_CONFIG, "expected a number");
    }
    double value(const std::string& k, double dflt) const { return contains(k) ? at(k).number() : dflt; }
    uint32_t value_u32(const std::string& k, uint32_t dflt) const {
        if (!contains(k)) return dflt;
        double v = at(k).number();
        if (v < 0 || v != (double)(uint32_t)v) fail(NLOD_ERR_CONFIG, "key '" + k + "' must be a non-negative integer");
        return (uint32_t)v;
    }
    bool value_bool(const std::string& k, bool dflt) const { return contains(k) ? at(k).number() != 0.0 : dflt; }
    std::string value_str(const std::string& k, const std::string& dflt) const {
        if (!contains(k)) return dflt;
        const Json& v = at(k);
        if (v.type != String) fail(NLOD_ERR_CONFIG, "key '" + k + "' must be a string");
        return v.str;
    }

    static Json parse(const std::string& text) {
        size_t i = 0;
        Json j = parse_value(text, i);
        skip_ws(text, i);
        if (i != text.size()) fail(NLOD_ERR_CONFIG, "trailing characters in JSON at offset " + std::to_string(i));
        return j;
    }

    std::string dump() const {
        std::ostringstream o;
        o.precision(17);
        dump_to(o);
        return o.str();
    }

private:
    void dump_to(std::ostringstream& o) const {
        switch (type) {
            case Null: o << "null"; break;
            case Bool: o << (b ? "true" : "false"); break;
            case Number: o << num; break;
            case String: o << '"' << str << '"'; break;
            case Array:
                o << '[';
                for (size_t k = 0; k < arr.size(); ++k) {
                    if (k) o << ',';
                    arr[k].dump_to(o);
                }
                o << ']';
                break;
            case Object:
                o << '{';
                for (size_t k = 0; k < obj.size(); ++k) {
                    if (k) o << ',';
                    o << '"' << obj[k].first << "\":";
                    obj[k].second.dump_to(o);
                }
                o << '}';
                break;
        }
    }
    static void skip_ws(const std::string& s, size_t& i) {
        while (i < s.size() && std::isspace((unsigned char)s[i])) ++i;
    }
    [[noreturn]] static void bad(size_t i, const char* what) {
        fail(NLOD_ERR_CONFIG, std::string("malformed JSON: ") + what + " at offset " + std::to_string(i));
    }
    static std::string parse_string(const std::string& s, size_t& i) {
        if (s[i] != '"') bad(i, "expected string");
        ++i;
        std::string out;
        while (i < s.size() && s[i] != '"') {
            char c = s[i++];
            if (c == '\\') {
                if (i >= s.size()) bad(i, "bad escape");
                char e = s[i++];
                switch (e) {
                    case 'n': out += '\n'; break;
                    case 't': out += '\t'; break;
                    case 'r': out += '\r'; break;
                    case 'b': out += '\b'; break;
                    case 'f': out += '\f'; break;
                    case 'u': {
                        if (i + 4 > s.size()) bad(i, "bad \\u escape");
                        unsigned v = (unsigned)std::strtoul(s.substr(i, 4).c_str(), nullptr, 16);
                        i += 4;
                        out += (v < 0x80) ? (char)v : '?';
                        break;
                    }
                    default: out += e;
                }
            } else {
                out += c;
            }
        }
        if (i >= s.size()) bad(i, "unterminated string");
        ++i;
        return out;
    }
    static Json parse_value(const std::string& s, size_t& i) {
        skip_ws(s, i);
        if (i >= s.size()) bad(i, "unexpected end");
        Json j;
        char c = s[i];
        if (c == '{') {
            j.type = Object;
            ++i;
            skip_ws(s, i);
            if (i < s.size() && s[i] == '}') { ++i; return j; }
            while (true) {
                skip_ws(s, i);
                std::string k = parse_string(s, i);
                skip_ws(s, i);
                if (i >= s.size() || s[i] != ':') bad(i, "expected ':'");
                ++i;
                j.obj.emplace_back(k, parse_value(s, i));
                skip_ws(s, i);
                if (i < s.size() && s[i] == ',') { ++i; continue; }
                if (i < s.size() && s[i] == '}') { ++i; break; }
                bad(i, "expected ',' or '}'");
            }
        } else if (c == '[') {
            j.type = Array;
            ++i;
            skip_ws(s, i);
            if (i < s.size() && s[i] == ']') { ++i; return j; }
            while (true) {
                j.arr.push_back(parse_value(s, i));
                skip_ws(s, i);
                if (i < s.size() && s[i] == ',') { ++i; continue; }
                if (i < s.size() && s[i] == ']') { ++i; break; }
                bad(i, "expected ',' or ']'");
            }
        } else if (c == '"') {
            j.type = String;
            j.str = parse_string(s, i);
        } else if (s.compare(i, 4, "true") == 0) {
            j.type = Bool; j.b = true; i += 4;
        } else if (s.compare(i, 5, "false") == 0) {
            j.type = Bool; j.b = false; i += 5;
        } else if (s.compare(i, 4, "null") == 0) {
            j.type = Null; i += 4;
        } else {
            const char* begin = s.c_str() + i;
            char* end = nullptr;
            j.num = std::strtod(begin, &end);
            if (end == begin) bad(i, "unexpected character");
            j.type = Number;
            i += (size_t)(end - begin);
        }
        return j;
    }
};

inline std::string lower(std::string s) {
    for (auto& c : s) c = (char)std::tolower((unsigned char)c);
    return s;
}

}  // namespace nlod

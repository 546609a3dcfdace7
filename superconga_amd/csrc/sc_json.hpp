// sc_json.hpp -- minimal JSON reader for simulation_config.json.
//
// Supports objects, arrays, strings (with the common escapes), numbers,
// true/false/null.  Enough for the configuration schema the reference's
// frontend writes (frontend/simulation/arguments.py, src/simulation.cu).
#pragma once

#include <cctype>
#include <cstdlib>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace scamd {
namespace json {

struct Value {
  enum Kind { Null, Bool, Number, String, Array, Object } kind = Null;
  bool b = false;
  double num = 0.0;
  std::string str;
  std::vector<Value> arr;
  std::vector<std::pair<std::string, Value>> obj; // keeps insertion order

  bool has(const std::string &k) const {
    if (kind != Object) return false;
    for (const auto &p : obj)
      if (p.first == k) return true;
    return false;
  }
  const Value &operator[](const std::string &k) const {
    static const Value null;
    if (kind == Object)
      for (const auto &p : obj)
        if (p.first == k) return p.second;
    return null;
  }
  const Value &operator[](size_t i) const {
    static const Value null;
    return (kind == Array && i < arr.size()) ? arr[i] : null;
  }
  size_t size() const {
    return kind == Array ? arr.size() : (kind == Object ? obj.size() : 0);
  }
  double asDouble(double def = 0.0) const {
    if (kind == Number) return num;
    if (kind == Bool) return b ? 1.0 : 0.0;
    return def;
  }
  int asInt(int def = 0) const {
    return kind == Number ? static_cast<int>(num) : (kind == Bool ? b : def);
  }
  bool asBool(bool def = false) const {
    if (kind == Bool) return b;
    if (kind == Number) return num != 0.0;
    return def;
  }
  std::string asString(const std::string &def = "") const {
    return kind == String ? str : def;
  }
};

class Parser {
public:
  explicit Parser(const std::string &s) : s_(s) {}
  Value parse() {
    Value v = value();
    ws();
    if (p_ != s_.size()) fail("trailing characters");
    return v;
  }

private:
  const std::string &s_;
  size_t p_ = 0;

  [[noreturn]] void fail(const char *msg) {
    throw std::runtime_error(std::string("config JSON: ") + msg + " at offset " +
                             std::to_string(p_));
  }
  void ws() {
    while (p_ < s_.size() && std::isspace(static_cast<unsigned char>(s_[p_])))
      ++p_;
  }
  bool eat(char c) {
    ws();
    if (p_ < s_.size() && s_[p_] == c) {
      ++p_;
      return true;
    }
    return false;
  }
  void expect(char c) {
    if (!eat(c)) fail("unexpected character");
  }
  std::string string() {
    expect('"');
    std::string out;
    while (p_ < s_.size() && s_[p_] != '"') {
      char c = s_[p_++];
      if (c == '\\') {
        if (p_ >= s_.size()) fail("bad escape");
        const char e = s_[p_++];
        switch (e) {
        case 'n': out += '\n'; break;
        case 't': out += '\t'; break;
        case 'r': out += '\r'; break;
        case 'b': out += '\b'; break;
        case 'f': out += '\f'; break;
        case 'u': {
          if (p_ + 4 > s_.size()) fail("bad unicode escape");
          const unsigned code = std::stoul(s_.substr(p_, 4), nullptr, 16);
          p_ += 4;
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
        default: out += e;
        }
      } else {
        out += c;
      }
    }
    expect('"');
    return out;
  }
  Value value() {
    ws();
    if (p_ >= s_.size()) fail("unexpected end");
    Value v;
    const char c = s_[p_];
    if (c == '{') {
      ++p_;
      v.kind = Value::Object;
      if (eat('}')) return v;
      do {
        ws();
        std::string k = string();
        expect(':');
        v.obj.emplace_back(k, value());
      } while (eat(','));
      expect('}');
    } else if (c == '[') {
      ++p_;
      v.kind = Value::Array;
      if (eat(']')) return v;
      do {
        v.arr.push_back(value());
      } while (eat(','));
      expect(']');
    } else if (c == '"') {
      v.kind = Value::String;
      v.str = string();
    } else if (s_.compare(p_, 4, "true") == 0) {
      p_ += 4;
      v.kind = Value::Bool;
      v.b = true;
    } else if (s_.compare(p_, 5, "false") == 0) {
      p_ += 5;
      v.kind = Value::Bool;
    } else if (s_.compare(p_, 4, "null") == 0) {
      p_ += 4;
    } else {
      const char *start = s_.c_str() + p_;
      char *end = nullptr;
      v.num = std::strtod(start, &end);
      if (end == start) fail("bad value");
      p_ += static_cast<size_t>(end - start);
      v.kind = Value::Number;
    }
    return v;
  }
};

inline Value parse(const std::string &text) { return Parser(text).parse(); }

} // namespace json
} // namespace scamd

// json_min.hpp -- a small JSON reader for the microlc input files (objects,
// arrays, numbers, strings, true/false/null; RFC 8259 syntax, UTF-8 passed
// through).  Replaces the jsoncpp dependency of the reference driver.
#pragma once

#include <cstdlib>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace mlcjson {

struct Value {
  enum Kind { Null, Bool, Number, String, Array, Object } kind = Null;
  bool b = false;
  double num = 0.0;
  std::string str;
  std::vector<Value> arr;
  std::map<std::string, Value> obj;

  bool has(const std::string& k) const { return kind == Object && obj.count(k); }
  const Value& operator[](const std::string& k) const {
    if (kind != Object) throw std::runtime_error("JSON: not an object where key \"" + k + "\" is read");
    auto it = obj.find(k);
    if (it == obj.end()) throw std::runtime_error("JSON: missing key \"" + k + "\"");
    return it->second;
  }
  const Value& operator[](size_t i) const {
    if (kind != Array) throw std::runtime_error("JSON: not an array");
    if (i >= arr.size()) throw std::runtime_error("JSON: index out of range");
    return arr[i];
  }
  size_t size() const { return kind == Array ? arr.size() : kind == Object ? obj.size() : 0; }
  double asDouble() const {
    if (kind != Number) throw std::runtime_error("JSON: number expected");
    return num;
  }
  int asInt() const {
    const double v = asDouble();
    if (v != (double)(int)v) throw std::runtime_error("JSON: integer expected");
    return (int)v;
  }
  const std::string& asString() const {
    if (kind != String) throw std::runtime_error("JSON: string expected");
    return str;
  }
};

class Parser {
 public:
  explicit Parser(const std::string& text) : s_(text) {}
  Value parse() {
    Value v = value();
    ws();
    if (p_ != s_.size()) fail("trailing characters");
    return v;
  }

 private:
  const std::string& s_;
  size_t p_ = 0;

  [[noreturn]] void fail(const std::string& what) const {
    size_t line = 1, col = 1;
    for (size_t i = 0; i < p_ && i < s_.size(); ++i) {
      if (s_[i] == '\n') {
        ++line;
        col = 1;
      } else {
        ++col;
      }
    }
    throw std::runtime_error("JSON: " + what + " at line " + std::to_string(line) + ", column " +
                             std::to_string(col));
  }
  void ws() {
    while (p_ < s_.size() && (s_[p_] == ' ' || s_[p_] == '\t' || s_[p_] == '\n' || s_[p_] == '\r')) ++p_;
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
    if (!eat(c)) fail(std::string("'") + c + "' expected");
  }
  bool word(const char* w) {
    size_t n = std::char_traits<char>::length(w);
    if (s_.compare(p_, n, w) == 0) {
      p_ += n;
      return true;
    }
    return false;
  }
  Value value() {
    ws();
    if (p_ >= s_.size()) fail("value expected");
    Value v;
    const char c = s_[p_];
    if (c == '{') {
      ++p_;
      v.kind = Value::Object;
      if (eat('}')) return v;
      do {
        ws();
        if (p_ >= s_.size() || s_[p_] != '"') fail("object key expected");
        std::string k = string();
        expect(':');
        v.obj[k] = value();
      } while (eat(','));
      expect('}');
    } else if (c == '[') {
      ++p_;
      v.kind = Value::Array;
      if (eat(']')) return v;
      do v.arr.push_back(value());
      while (eat(','));
      expect(']');
    } else if (c == '"') {
      v.kind = Value::String;
      v.str = string();
    } else if (word("true")) {
      v.kind = Value::Bool;
      v.b = true;
    } else if (word("false")) {
      v.kind = Value::Bool;
    } else if (word("null")) {
      v.kind = Value::Null;
    } else {
      v.kind = Value::Number;
      const char* b = s_.c_str() + p_;
      char* e = nullptr;
      v.num = std::strtod(b, &e);
      if (e == b) fail("value expected");
      p_ += (size_t)(e - b);
    }
    return v;
  }
  std::string string() {
    ++p_;  // opening quote
    std::string out;
    while (p_ < s_.size() && s_[p_] != '"') {
      char c = s_[p_++];
      if (c == '\\') {
        if (p_ >= s_.size()) fail("bad escape");
        char e = s_[p_++];
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
            if (p_ + 4 > s_.size()) fail("bad \\u escape");
            unsigned cp = (unsigned)std::strtoul(s_.substr(p_, 4).c_str(), nullptr, 16);
            p_ += 4;
            if (cp < 0x80) {
              out += (char)cp;
            } else if (cp < 0x800) {
              out += (char)(0xC0 | (cp >> 6));
              out += (char)(0x80 | (cp & 0x3F));
            } else {
              out += (char)(0xE0 | (cp >> 12));
              out += (char)(0x80 | ((cp >> 6) & 0x3F));
              out += (char)(0x80 | (cp & 0x3F));
            }
            break;
          }
          default: fail("bad escape");
        }
      } else {
        out += c;
      }
    }
    if (p_ >= s_.size()) fail("unterminated string");
    ++p_;
    return out;
  }
};

inline Value parse(const std::string& text) { return Parser(text).parse(); }

}  // namespace mlcjson

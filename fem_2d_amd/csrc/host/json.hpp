// Minimal JSON reader for mesh files.
//
// Mesh files (Mesh::from_file, reference src/fem_domain/domain/mesh.rs:100-327)
// hold an object with an "Elements" array of {"materials": [4 numbers],
// "node_ids": [4 integers]} and a "Nodes" array of [x, y] pairs. This reader
// accepts general JSON (objects, arrays, numbers, strings, literals) but keeps
// only what those files need: numbers are parsed with strtod.
#pragma once

#include <cstdlib>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace fem2d {
namespace json {

struct Value {
  enum class Kind { Null, Bool, Number, String, Array, Object };
  Kind kind = Kind::Null;
  bool boolean = false;
  double number = 0.0;
  bool is_integer = false;  // number was written without fraction/exponent
  std::string text;
  std::vector<Value> items;
  std::map<std::string, Value> fields;

  bool is_array() const { return kind == Kind::Array; }
  bool is_number() const { return kind == Kind::Number; }
  const Value* get(const std::string& key) const {
    if (kind != Kind::Object) return nullptr;
    auto it = fields.find(key);
    return it == fields.end() ? nullptr : &it->second;
  }
};

class Parser {
 public:
  explicit Parser(const std::string& s) : s_(s) {}

  Value parse_document() {
    Value v = parse_value();
    skip_ws();
    if (pos_ != s_.size()) fail("trailing characters after JSON value");
    return v;
  }

 private:
  const std::string& s_;
  size_t pos_ = 0;

  [[noreturn]] void fail(const std::string& msg) {
    throw std::runtime_error("JSON parse error at byte " + std::to_string(pos_) + ": " + msg);
  }
  void skip_ws() {
    while (pos_ < s_.size() && (s_[pos_] == ' ' || s_[pos_] == '\n' || s_[pos_] == '\r' || s_[pos_] == '\t'))
      ++pos_;
  }
  char peek() {
    skip_ws();
    if (pos_ >= s_.size()) fail("unexpected end of input");
    return s_[pos_];
  }
  void expect(char c) {
    if (peek() != c) fail(std::string("expected '") + c + "'");
    ++pos_;
  }

  Value parse_value() {
    char c = peek();
    if (c == '{') return parse_object();
    if (c == '[') return parse_array();
    if (c == '"') {
      Value v;
      v.kind = Value::Kind::String;
      v.text = parse_string();
      return v;
    }
    if (c == 't' || c == 'f' || c == 'n') return parse_literal();
    return parse_number();
  }

  Value parse_object() {
    Value v;
    v.kind = Value::Kind::Object;
    expect('{');
    if (peek() == '}') {
      ++pos_;
      return v;
    }
    while (true) {
      if (peek() != '"') fail("expected object key");
      std::string key = parse_string();
      expect(':');
      v.fields[key] = parse_value();
      char c = peek();
      ++pos_;
      if (c == '}') break;
      if (c != ',') fail("expected ',' or '}'");
    }
    return v;
  }

  Value parse_array() {
    Value v;
    v.kind = Value::Kind::Array;
    expect('[');
    if (peek() == ']') {
      ++pos_;
      return v;
    }
    while (true) {
      v.items.push_back(parse_value());
      char c = peek();
      ++pos_;
      if (c == ']') break;
      if (c != ',') fail("expected ',' or ']'");
    }
    return v;
  }

  std::string parse_string() {
    expect('"');
    std::string out;
    while (pos_ < s_.size() && s_[pos_] != '"') {
      char c = s_[pos_++];
      if (c == '\\') {
        if (pos_ >= s_.size()) fail("bad escape");
        char e = s_[pos_++];
        switch (e) {
          case 'n': out.push_back('\n'); break;
          case 't': out.push_back('\t'); break;
          case 'r': out.push_back('\r'); break;
          case 'b': out.push_back('\b'); break;
          case 'f': out.push_back('\f'); break;
          case 'u': pos_ += 4; out.push_back('?'); break;
          default: out.push_back(e);
        }
      } else {
        out.push_back(c);
      }
    }
    if (pos_ >= s_.size()) fail("unterminated string");
    ++pos_;
    return out;
  }

  Value parse_literal() {
    Value v;
    auto take = [&](const char* w) {
      size_t n = std::char_traits<char>::length(w);
      if (s_.compare(pos_, n, w) != 0) fail("bad literal");
      pos_ += n;
    };
    char c = peek();
    if (c == 't') {
      take("true");
      v.kind = Value::Kind::Bool;
      v.boolean = true;
    } else if (c == 'f') {
      take("false");
      v.kind = Value::Kind::Bool;
    } else {
      take("null");
    }
    return v;
  }

  Value parse_number() {
    skip_ws();
    const char* begin = s_.c_str() + pos_;
    char* end = nullptr;
    double d = std::strtod(begin, &end);
    if (end == begin) fail("expected a value");
    Value v;
    v.kind = Value::Kind::Number;
    v.number = d;
    v.is_integer = true;
    for (const char* p = begin; p < end; ++p)
      if (*p == '.' || *p == 'e' || *p == 'E') v.is_integer = false;
    pos_ += static_cast<size_t>(end - begin);
    return v;
  }
};

inline Value parse(const std::string& text) { return Parser(text).parse_document(); }

}  // namespace json
}  // namespace fem2d

// Minimal JSON value, parser and writer for the C ABI's request/response
// protocol (no external dependency is available in this toolchain).
#pragma once

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace qmbed {
namespace json {

struct Value;
using Object = std::vector<std::pair<std::string, Value>>;
using Array = std::vector<Value>;

struct Value {
  enum Kind { Null, Bool, Number, String, Arr, Obj } kind = Null;
  bool b = false;
  double num = 0.0;
  std::string raw;  // number text (exact integers) or string contents
  std::shared_ptr<Array> arr;
  std::shared_ptr<Object> obj;

  bool is_null() const { return kind == Null; }
  const Value* get(const std::string& key) const {
    if (kind != Obj) return nullptr;
    for (auto& kv : *obj)
      if (kv.first == key) return &kv.second;
    return nullptr;
  }
  const Array& array() const {
    static const Array empty;
    return kind == Arr ? *arr : empty;
  }
};

struct ParseError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

class Parser {
 public:
  explicit Parser(const std::string& s) : s_(s) {}
  Value parse() {
    Value v = value();
    ws();
    if (p_ != s_.size()) fail("trailing characters");
    return v;
  }

 private:
  const std::string& s_;
  size_t p_ = 0;
  [[noreturn]] void fail(const std::string& what) {
    throw ParseError(what + " at column " + std::to_string(p_ + 1));
  }
  void ws() {
    while (p_ < s_.size() && (s_[p_] == ' ' || s_[p_] == '\n' || s_[p_] == '\r' || s_[p_] == '\t')) ++p_;
  }
  // nesting limit of arrays and objects (serde_json's default recursion
  // limit): deeper requests are rejected instead of exhausting the stack
  static constexpr int kMaxDepth = 128;
  int depth_ = 0;
  struct Nest {
    Parser& p;
    explicit Nest(Parser& q) : p(q) {
      if (++p.depth_ > kMaxDepth) p.fail("recursion limit exceeded");
    }
    ~Nest() { --p.depth_; }
  };
  Value value() {
    ws();
    if (p_ >= s_.size()) fail("EOF while parsing a value");
    char c = s_[p_];
    Value v;
    if (c == '{') {
      Nest nest(*this);
      ++p_;
      v.kind = Value::Obj;
      v.obj = std::make_shared<Object>();
      ws();
      if (p_ < s_.size() && s_[p_] == '}') {
        ++p_;
        return v;
      }
      for (;;) {
        ws();
        if (p_ >= s_.size() || s_[p_] != '"') fail("key must be a string");
        std::string k = str();
        ws();
        if (p_ >= s_.size() || s_[p_] != ':') fail("expected `:`");
        ++p_;
        v.obj->emplace_back(k, value());
        ws();
        if (p_ < s_.size() && s_[p_] == ',') {
          ++p_;
          continue;
        }
        if (p_ < s_.size() && s_[p_] == '}') {
          ++p_;
          return v;
        }
        fail("expected `,` or `}`");
      }
    }
    if (c == '[') {
      Nest nest(*this);
      ++p_;
      v.kind = Value::Arr;
      v.arr = std::make_shared<Array>();
      ws();
      if (p_ < s_.size() && s_[p_] == ']') {
        ++p_;
        return v;
      }
      for (;;) {
        v.arr->push_back(value());
        ws();
        if (p_ < s_.size() && s_[p_] == ',') {
          ++p_;
          continue;
        }
        if (p_ < s_.size() && s_[p_] == ']') {
          ++p_;
          return v;
        }
        fail("expected `,` or `]`");
      }
    }
    if (c == '"') {
      v.kind = Value::String;
      v.raw = str();
      return v;
    }
    if (s_.compare(p_, 4, "true") == 0) {
      p_ += 4;
      v.kind = Value::Bool;
      v.b = true;
      return v;
    }
    if (s_.compare(p_, 5, "false") == 0) {
      p_ += 5;
      v.kind = Value::Bool;
      return v;
    }
    if (s_.compare(p_, 4, "null") == 0) {
      p_ += 4;
      return v;
    }
    size_t st = p_;
    if (c == '-') ++p_;
    while (p_ < s_.size() && (isdigit(s_[p_]) || s_[p_] == '.' || s_[p_] == 'e' || s_[p_] == 'E' || s_[p_] == '+' ||
                              s_[p_] == '-'))
      ++p_;
    if (st == p_) fail("expected value");
    v.kind = Value::Number;
    v.raw = s_.substr(st, p_ - st);
    char* end = nullptr;
    v.num = std::strtod(v.raw.c_str(), &end);
    if (!end || *end) fail("invalid number");
    return v;
  }
  unsigned hex4() {
    if (p_ + 4 > s_.size()) fail("EOF while parsing a string");
    unsigned cp = 0;
    for (int k = 0; k < 4; ++k) {
      const char h = s_[p_ + size_t(k)];
      const int d = h >= '0' && h <= '9' ? h - '0' : h >= 'a' && h <= 'f' ? h - 'a' + 10
                                                   : h >= 'A' && h <= 'F' ? h - 'A' + 10 : -1;
      if (d < 0) fail("invalid escape");
      cp = cp * 16 + unsigned(d);
    }
    p_ += 4;
    return cp;
  }
  static void utf8(std::string& out, unsigned cp) {
    if (cp < 0x80) {
      out += char(cp);
    } else if (cp < 0x800) {
      out += char(0xC0 | (cp >> 6));
      out += char(0x80 | (cp & 0x3F));
    } else if (cp < 0x10000) {
      out += char(0xE0 | (cp >> 12));
      out += char(0x80 | ((cp >> 6) & 0x3F));
      out += char(0x80 | (cp & 0x3F));
    } else {
      out += char(0xF0 | (cp >> 18));
      out += char(0x80 | ((cp >> 12) & 0x3F));
      out += char(0x80 | ((cp >> 6) & 0x3F));
      out += char(0x80 | (cp & 0x3F));
    }
  }
  std::string str() {
    ++p_;  // opening quote
    std::string out;
    while (p_ < s_.size() && s_[p_] != '"') {
      char c = s_[p_++];
      if (c == '\\') {
        if (p_ >= s_.size()) fail("EOF in escape");
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
            unsigned cp = hex4();
            if (cp >= 0xDC00 && cp <= 0xDFFF) fail("lone trailing surrogate in hex escape");
            if (cp >= 0xD800 && cp <= 0xDBFF) {  // a leading surrogate needs its trailing half
              if (p_ + 2 > s_.size() || s_[p_] != '\\' || s_[p_ + 1] != 'u')
                fail("lone leading surrogate in hex escape");
              p_ += 2;
              const unsigned lo = hex4();
              if (lo < 0xDC00 || lo > 0xDFFF) fail("lone leading surrogate in hex escape");
              cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
            }
            utf8(out, cp);
            break;
          }
          default: fail("bad escape");
        }
      } else {
        if (static_cast<unsigned char>(c) < 0x20) fail("control character (\\u0000-\\u001F) found while parsing a string");
        out += c;
      }
    }
    if (p_ >= s_.size()) fail("EOF while parsing a string");
    ++p_;
    return out;
  }
};

inline Value parse(const std::string& s) { return Parser(s).parse(); }

// ---------------------------------------------------------------- writer --
inline void write_string(std::string& o, const std::string& s) {
  o += '"';
  for (char c : s) {
    switch (c) {
      case '"': o += "\\\""; break;
      case '\\': o += "\\\\"; break;
      case '\n': o += "\\n"; break;
      case '\r': o += "\\r"; break;
      case '\t': o += "\\t"; break;
      default:
        if (static_cast<unsigned char>(c) < 0x20) {
          char buf[8];
          std::snprintf(buf, sizeof buf, "\\u%04x", c);
          o += buf;
        } else {
          o += c;
        }
    }
  }
  o += '"';
}

// shortest representation that round-trips; non-finite values become null
inline void write_number(std::string& o, double v) {
  if (!std::isfinite(v)) {
    o += "null";
    return;
  }
  char buf[40];
  for (int prec = 1; prec <= 17; ++prec) {
    std::snprintf(buf, sizeof buf, "%.*g", prec, v);
    if (std::strtod(buf, nullptr) == v) break;
  }
  std::string s(buf);
  if (s.find_first_of(".eEn") == std::string::npos) s += ".0";
  o += s;
}

inline void write_uint(std::string& o, uint64_t v) { o += std::to_string(v); }

}  // namespace json
}  // namespace qmbed

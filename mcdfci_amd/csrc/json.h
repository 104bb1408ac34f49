// Minimal JSON value, parser, RFC 7386 merge-patch and an indented writer
// whose layout follows the reference's `option.dump(4)` output (sorted
// keys, 4-space indent, floats in shortest round-trip form with ".0" on
// integral values, exponents with a sign and at least two digits).
#pragma once

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace mcjson {

struct Value {
  enum Kind { Null, Bool, Int, UInt, Float, String, Array, Object } kind = Null;
  bool b = false;
  long long i = 0;
  unsigned long long u = 0;
  double f = 0.0;
  std::string s;
  std::vector<Value> a;
  std::map<std::string, Value> o;

  static Value object() {
    Value v;
    v.kind = Object;
    return v;
  }
  static Value array() {
    Value v;
    v.kind = Array;
    return v;
  }
  static Value of(double x) {
    Value v;
    v.kind = Float;
    v.f = x;
    return v;
  }
  static Value of_int(long long x) {
    Value v;
    v.kind = Int;
    v.i = x;
    return v;
  }
  static Value of_uint(unsigned long long x) {
    Value v;
    v.kind = UInt;
    v.u = x;
    return v;
  }
  static Value of_bool(bool x) {
    Value v;
    v.kind = Bool;
    v.b = x;
    return v;
  }
  static Value of_str(const std::string& x) {
    Value v;
    v.kind = String;
    v.s = x;
    return v;
  }
  bool is_number() const { return kind == Int || kind == UInt || kind == Float; }
  double num() const {
    if (kind == Int) return (double)i;
    if (kind == UInt) return (double)u;
    if (kind == Float) return f;
    if (kind == Bool) return b ? 1.0 : 0.0;
    throw std::invalid_argument("[json.exception.type_error.302] type must be number");
  }
  long long integer() const {
    if (kind == Int) return i;
    if (kind == UInt) return (long long)u;
    if (kind == Float) return (long long)f;
    if (kind == Bool) return b ? 1 : 0;
    throw std::invalid_argument("[json.exception.type_error.302] type must be number");
  }
  bool boolean() const {
    if (kind == Bool) return b;
    throw std::invalid_argument("[json.exception.type_error.302] type must be boolean");
  }
  const std::string& str() const {
    if (kind == String) return s;
    throw std::invalid_argument("[json.exception.type_error.302] type must be string");
  }
  Value& operator[](const std::string& k) {
    if (kind == Null) kind = Object;
    return o[k];
  }
  const Value* find(const std::string& k) const {
    if (kind != Object) return nullptr;
    auto it = o.find(k);
    return it == o.end() ? nullptr : &it->second;
  }
};

class Parser {
 public:
  explicit Parser(const std::string& t) : t_(t) {}
  Value parse() {
    Value v = value();
    ws();
    if (p_ != t_.size()) throw std::invalid_argument("trailing characters in JSON");
    return v;
  }

 private:
  const std::string& t_;
  size_t p_ = 0;
  void ws() {
    while (p_ < t_.size() && (t_[p_] == ' ' || t_[p_] == '\n' || t_[p_] == '\r' || t_[p_] == '\t')) ++p_;
  }
  char peek() {
    ws();
    if (p_ >= t_.size()) throw std::invalid_argument("unexpected end of JSON");
    return t_[p_];
  }
  void expect(char c) {
    if (peek() != c) throw std::invalid_argument(std::string("expected '") + c + "' in JSON");
    ++p_;
  }
  std::string string_lit() {
    expect('"');
    std::string out;
    while (p_ < t_.size() && t_[p_] != '"') {
      char c = t_[p_++];
      if (c == '\\') {
        if (p_ >= t_.size()) break;
        char e = t_[p_++];
        switch (e) {
          case 'n': out += '\n'; break;
          case 't': out += '\t'; break;
          case 'r': out += '\r'; break;
          case 'b': out += '\b'; break;
          case 'f': out += '\f'; break;
          case 'u': {
            unsigned code = (unsigned)std::strtoul(t_.substr(p_, 4).c_str(), nullptr, 16);
            p_ += 4;
            if (code < 0x80) out += (char)code;
            else if (code < 0x800) {
              out += (char)(0xC0 | (code >> 6));
              out += (char)(0x80 | (code & 0x3F));
            } else {
              out += (char)(0xE0 | (code >> 12));
              out += (char)(0x80 | ((code >> 6) & 0x3F));
              out += (char)(0x80 | (code & 0x3F));
            }
            break;
          }
          default: out += e;
        }
      } else {
        out += c;
      }
    }
    if (p_ >= t_.size()) throw std::invalid_argument("unterminated string in JSON");
    ++p_;
    return out;
  }
  Value value() {
    char c = peek();
    if (c == '{') {
      ++p_;
      Value v = Value::object();
      if (peek() == '}') {
        ++p_;
        return v;
      }
      while (true) {
        std::string k = string_lit();
        expect(':');
        v.o[k] = value();
        char d = peek();
        ++p_;
        if (d == '}') break;
        if (d != ',') throw std::invalid_argument("expected ',' or '}' in JSON");
      }
      return v;
    }
    if (c == '[') {
      ++p_;
      Value v = Value::array();
      if (peek() == ']') {
        ++p_;
        return v;
      }
      while (true) {
        v.a.push_back(value());
        char d = peek();
        ++p_;
        if (d == ']') break;
        if (d != ',') throw std::invalid_argument("expected ',' or ']' in JSON");
      }
      return v;
    }
    if (c == '"') return Value::of_str(string_lit());
    if (t_.compare(p_, 4, "true") == 0) {
      p_ += 4;
      return Value::of_bool(true);
    }
    if (t_.compare(p_, 5, "false") == 0) {
      p_ += 5;
      return Value::of_bool(false);
    }
    if (t_.compare(p_, 4, "null") == 0) {
      p_ += 4;
      return Value();
    }
    size_t start = p_;
    bool is_float = false;
    if (t_[p_] == '-' || t_[p_] == '+') ++p_;
    while (p_ < t_.size() && (isdigit((unsigned char)t_[p_]) || t_[p_] == '.' || t_[p_] == 'e' ||
                              t_[p_] == 'E' || t_[p_] == '-' || t_[p_] == '+')) {
      if (t_[p_] == '.' || t_[p_] == 'e' || t_[p_] == 'E') is_float = true;
      ++p_;
    }
    std::string num = t_.substr(start, p_ - start);
    if (num.empty() || num == "-") throw std::invalid_argument("invalid JSON value");
    if (is_float) return Value::of(std::strtod(num.c_str(), nullptr));
    if (num[0] == '-') return Value::of_int(std::strtoll(num.c_str(), nullptr, 10));
    return Value::of_uint(std::strtoull(num.c_str(), nullptr, 10));
  }
};

inline Value parse(const std::string& text) { return Parser(text).parse(); }

// RFC 7386 merge patch applied to target.
inline void merge_patch(Value& target, const Value& patch) {
  if (patch.kind != Value::Object) {
    target = patch;
    return;
  }
  if (target.kind != Value::Object) target = Value::object();
  for (auto& kv : patch.o) {
    if (kv.second.kind == Value::Null) target.o.erase(kv.first);
    else merge_patch(target.o[kv.first], kv.second);
  }
}

inline std::string format_double(double x) {
  if (x == 0.0) return std::signbit(x) ? "-0.0" : "0.0";
  if (std::isnan(x) || std::isinf(x)) return "null";
  char buf[64];
  int prec = 1;
  for (; prec <= 17; ++prec) {
    std::snprintf(buf, sizeof buf, "%.*e", prec - 1, x);
    if (std::strtod(buf, nullptr) == x) break;
  }
  // buf = [-]d.ddde[+-]XX ; extract digits and exponent
  std::string s(buf);
  bool neg = s[0] == '-';
  if (neg) s = s.substr(1);
  size_t epos = s.find('e');
  std::string mant = s.substr(0, epos);
  int e10 = std::atoi(s.c_str() + epos + 1);
  std::string digits;
  for (char c : mant)
    if (c != '.') digits += c;
  while (digits.size() > 1 && digits.back() == '0') digits.pop_back();
  const int k = (int)digits.size();
  const int n = e10 + 1;  // position of the decimal point relative to the digits
  std::string out;
  if (k <= n && n <= 15) {
    out = digits + std::string(n - k, '0') + ".0";
  } else if (0 < n && n <= 15) {
    out = digits.substr(0, n) + "." + digits.substr(n);
  } else if (-4 < n && n <= 0) {
    out = "0." + std::string(-n, '0') + digits;
  } else {
    out = digits.substr(0, 1);
    if (k > 1) out += "." + digits.substr(1);
    int ex = n - 1;
    char eb[16];
    std::snprintf(eb, sizeof eb, "e%c%02d", ex < 0 ? '-' : '+', ex < 0 ? -ex : ex);
    out += eb;
  }
  return neg ? "-" + out : out;
}

inline std::string escape(const std::string& s) {
  std::string o = "\"";
  for (char c : s) {
    switch (c) {
      case '"': o += "\\\""; break;
      case '\\': o += "\\\\"; break;
      case '\n': o += "\\n"; break;
      case '\t': o += "\\t"; break;
      default: o += c;
    }
  }
  return o + "\"";
}

inline void dump_to(const Value& v, std::string& out, int indent, int level) {
  std::string pad((size_t)indent * (level + 1), ' '), pad0((size_t)indent * level, ' ');
  switch (v.kind) {
    case Value::Null: out += "null"; break;
    case Value::Bool: out += v.b ? "true" : "false"; break;
    case Value::Int: out += std::to_string(v.i); break;
    case Value::UInt: out += std::to_string(v.u); break;
    case Value::Float: out += format_double(v.f); break;
    case Value::String: out += escape(v.s); break;
    case Value::Array:
      if (v.a.empty()) {
        out += "[]";
        break;
      }
      out += "[\n";
      for (size_t i = 0; i < v.a.size(); ++i) {
        out += pad;
        dump_to(v.a[i], out, indent, level + 1);
        out += i + 1 < v.a.size() ? ",\n" : "\n";
      }
      out += pad0 + "]";
      break;
    case Value::Object: {
      if (v.o.empty()) {
        out += "{}";
        break;
      }
      out += "{\n";
      size_t i = 0;
      for (auto& kv : v.o) {
        out += pad + escape(kv.first) + ": ";
        dump_to(kv.second, out, indent, level + 1);
        out += ++i < v.o.size() ? ",\n" : "\n";
      }
      out += pad0 + "}";
      break;
    }
  }
}

inline std::string dump(const Value& v, int indent = 4) {
  std::string out;
  dump_to(v, out, indent, 0);
  return out;
}

}  // namespace mcjson

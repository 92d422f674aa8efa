// json.hpp — a small JSON reader for the reference's receptor, ligand and
// scoring files (RFC 8259 values: objects, arrays, numbers, strings, true,
// false, null).  Numbers are kept as double.  Errors throw pxdock::conf_error.
#pragma once

#include <cstdlib>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace pxdock {

// failed_conf_error (engine/include/bytedock/lib/error.h:38-48): malformed input
class conf_error : public std::runtime_error {
 public:
  explicit conf_error(const std::string& m) : std::runtime_error(m) {}
};

class json {
 public:
  enum kind_t { NUL, BOOL, NUM, STR, ARR, OBJ };

  json() = default;
  static json parse(const std::string& text) {
    size_t i = 0;
    json v = value(text, i);
    skip(text, i);
    if (i != text.size()) throw conf_error("json: trailing characters at " + std::to_string(i));
    return v;
  }

  kind_t kind() const { return k_; }
  bool is_array() const { return k_ == ARR; }
  bool is_object() const { return k_ == OBJ; }
  double num() const {
    if (k_ != NUM) throw conf_error("json: number expected");
    return n_;
  }
  long long integer() const { return (long long)num(); }
  const std::string& str() const {
    if (k_ != STR) throw conf_error("json: string expected");
    return s_;
  }
  size_t size() const { return k_ == ARR ? a_.size() : (k_ == OBJ ? o_.size() : 0); }
  const json& operator[](size_t i) const {
    if (k_ != ARR || i >= a_.size()) throw conf_error("json: index out of range");
    return a_[i];
  }
  bool has(const std::string& key) const { return k_ == OBJ && o_.count(key) > 0; }
  const json& operator[](const std::string& key) const {
    if (k_ != OBJ) throw conf_error("json: object expected for key " + key);
    auto it = o_.find(key);
    if (it == o_.end()) throw conf_error("json: missing key " + key);
    return it->second;
  }
  const std::vector<json>& items() const { return a_; }
  const std::map<std::string, json>& fields() const { return o_; }
  // insertion order of an object's keys
  const std::vector<std::string>& keys() const { return order_; }

 private:
  kind_t k_ = NUL;
  double n_ = 0.0;
  std::string s_;
  std::vector<json> a_;
  std::map<std::string, json> o_;
  std::vector<std::string> order_;

  static void skip(const std::string& t, size_t& i) {
    while (i < t.size() && (t[i] == ' ' || t[i] == '\n' || t[i] == '\r' || t[i] == '\t')) ++i;
  }
  static void expect(const std::string& t, size_t& i, char c) {
    skip(t, i);
    if (i >= t.size() || t[i] != c)
      throw conf_error(std::string("json: expected '") + c + "' at " + std::to_string(i));
    ++i;
  }
  static std::string string_lit(const std::string& t, size_t& i) {
    expect(t, i, '"');
    std::string out;
    while (i < t.size() && t[i] != '"') {
      char c = t[i++];
      if (c == '\\') {
        if (i >= t.size()) break;
        char e = t[i++];
        switch (e) {
          case 'n': out += '\n'; break;
          case 't': out += '\t'; break;
          case 'r': out += '\r'; break;
          case 'b': out += '\b'; break;
          case 'f': out += '\f'; break;
          case 'u': {  // keep ASCII, replace others
            if (i + 4 > t.size()) throw conf_error("json: bad escape");
            unsigned v = (unsigned)std::strtoul(t.substr(i, 4).c_str(), nullptr, 16);
            out += v < 128 ? (char)v : '?';
            i += 4;
            break;
          }
          default: out += e;
        }
      } else {
        out += c;
      }
    }
    if (i >= t.size()) throw conf_error("json: unterminated string");
    ++i;
    return out;
  }
  static json value(const std::string& t, size_t& i) {
    skip(t, i);
    if (i >= t.size()) throw conf_error("json: unexpected end");
    json v;
    char c = t[i];
    if (c == '{') {
      v.k_ = OBJ;
      ++i;
      skip(t, i);
      if (i < t.size() && t[i] == '}') {
        ++i;
        return v;
      }
      for (;;) {
        std::string key = string_lit(t, i);
        expect(t, i, ':');
        if (!v.o_.count(key)) v.order_.push_back(key);
        v.o_[key] = value(t, i);
        skip(t, i);
        if (i < t.size() && t[i] == ',') {
          ++i;
          continue;
        }
        expect(t, i, '}');
        return v;
      }
    }
    if (c == '[') {
      v.k_ = ARR;
      ++i;
      skip(t, i);
      if (i < t.size() && t[i] == ']') {
        ++i;
        return v;
      }
      for (;;) {
        v.a_.push_back(value(t, i));
        skip(t, i);
        if (i < t.size() && t[i] == ',') {
          ++i;
          continue;
        }
        expect(t, i, ']');
        return v;
      }
    }
    if (c == '"') {
      v.k_ = STR;
      v.s_ = string_lit(t, i);
      return v;
    }
    if (t.compare(i, 4, "true") == 0) {
      v.k_ = BOOL;
      v.n_ = 1.0;
      i += 4;
      return v;
    }
    if (t.compare(i, 5, "false") == 0) {
      v.k_ = BOOL;
      i += 5;
      return v;
    }
    if (t.compare(i, 4, "null") == 0) {
      i += 4;
      return v;
    }
    const char* b = t.c_str() + i;
    char* e = nullptr;
    double d = std::strtod(b, &e);
    if (e == b) throw conf_error("json: bad value at " + std::to_string(i));
    i += (size_t)(e - b);
    v.k_ = NUM;
    v.n_ = d;
    return v;
  }
};

}  // namespace pxdock

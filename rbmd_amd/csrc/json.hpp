// Minimal JSON reader for RBMD configuration documents: objects, arrays,
// strings, numbers, booleans, null, plus // and /* */ comments (RBMD parses
// its input with comments allowed).  Read-only tree; no serialisation.
#pragma once
#include <cstdlib>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace rbmdx {

struct Json {
  enum Kind { Null, Bool, Number, String, Array, Object } kind = Null;
  bool b = false;
  double num = 0.0;
  std::string str;
  std::vector<Json> arr;
  std::vector<std::pair<std::string, Json>> obj;  // insertion order

  const Json* get(const std::string& k) const {
    if (kind != Object) return nullptr;
    for (auto& kv : obj)
      if (kv.first == k) return &kv.second;
    return nullptr;
  }
  // Path lookup: get({"hyper_parameters", "neighbor", "cut_off"}).
  const Json* at(std::initializer_list<const char*> path) const {
    const Json* j = this;
    for (const char* p : path) {
      if (!j) return nullptr;
      j = j->get(p);
    }
    return j;
  }
};

class JsonParser {
 public:
  explicit JsonParser(const std::string& s) : s_(s) {}
  Json parse() {
    Json j = value();
    ws();
    if (p_ != s_.size()) fail("trailing characters");
    return j;
  }

 private:
  const std::string& s_;
  size_t p_ = 0;

  [[noreturn]] void fail(const char* what) {
    throw std::runtime_error(std::string("JSON parse error at offset ") +
                             std::to_string(p_) + ": " + what);
  }
  void ws() {
    for (;;) {
      while (p_ < s_.size() && (s_[p_] == ' ' || s_[p_] == '\t' ||
                                s_[p_] == '\n' || s_[p_] == '\r'))
        ++p_;
      if (p_ + 1 < s_.size() && s_[p_] == '/' && s_[p_ + 1] == '/') {
        while (p_ < s_.size() && s_[p_] != '\n') ++p_;
        continue;
      }
      if (p_ + 1 < s_.size() && s_[p_] == '/' && s_[p_ + 1] == '*') {
        const size_t e = s_.find("*/", p_ + 2);
        if (e == std::string::npos) fail("unterminated comment");
        p_ = e + 2;
        continue;
      }
      return;
    }
  }
  bool lit(const char* w) {
    size_t n = 0;
    while (w[n]) ++n;
    if (s_.compare(p_, n, w) == 0) {
      p_ += n;
      return true;
    }
    return false;
  }
  Json value() {
    ws();
    if (p_ >= s_.size()) fail("unexpected end");
    Json j;
    const char c = s_[p_];
    if (c == '{') {
      j.kind = Json::Object;
      ++p_;
      ws();
      if (p_ < s_.size() && s_[p_] == '}') {
        ++p_;
        return j;
      }
      for (;;) {
        ws();
        if (p_ >= s_.size() || s_[p_] != '"') fail("expected key");
        std::string k = string();
        ws();
        if (p_ >= s_.size() || s_[p_] != ':') fail("expected ':'");
        ++p_;
        j.obj.emplace_back(k, value());
        ws();
        if (p_ < s_.size() && s_[p_] == ',') {
          ++p_;
          continue;
        }
        if (p_ < s_.size() && s_[p_] == '}') {
          ++p_;
          return j;
        }
        fail("expected ',' or '}'");
      }
    }
    if (c == '[') {
      j.kind = Json::Array;
      ++p_;
      ws();
      if (p_ < s_.size() && s_[p_] == ']') {
        ++p_;
        return j;
      }
      for (;;) {
        j.arr.push_back(value());
        ws();
        if (p_ < s_.size() && s_[p_] == ',') {
          ++p_;
          continue;
        }
        if (p_ < s_.size() && s_[p_] == ']') {
          ++p_;
          return j;
        }
        fail("expected ',' or ']'");
      }
    }
    if (c == '"') {
      j.kind = Json::String;
      j.str = string();
      return j;
    }
    if (lit("true")) {
      j.kind = Json::Bool;
      j.b = true;
      return j;
    }
    if (lit("false")) {
      j.kind = Json::Bool;
      return j;
    }
    if (lit("null")) return j;
    const char* b = s_.c_str() + p_;
    char* e = nullptr;
    j.num = std::strtod(b, &e);
    if (e == b) fail("bad value");
    p_ += (size_t)(e - b);
    j.kind = Json::Number;
    return j;
  }
  std::string string() {
    ++p_;  // opening quote
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
          case 'u':
            if (p_ + 4 > s_.size()) fail("bad \\u escape");
            out += '?';
            p_ += 4;
            break;
          default: out += e;
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

inline Json parse_json(const std::string& s) { return JsonParser(s).parse(); }

}  // namespace rbmdx

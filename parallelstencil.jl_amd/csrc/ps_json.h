// Minimal JSON reader/writer for the option strings of the C-ABI
// (kernel keywords in, kernel metadata out).
#pragma once
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <memory>
#include <string>
#include <vector>

namespace ps {

struct JVal {
  enum T { Null, Bool, Num, Str, Arr, Obj } t = Null;
  bool b = false;
  double n = 0;
  std::string s;
  std::vector<JVal> arr;
  std::vector<std::pair<std::string, JVal>> obj;

  const JVal* get(const std::string& k) const {
    for (auto& kv : obj)
      if (kv.first == k) return &kv.second;
    return nullptr;
  }
  bool has(const std::string& k) const { return get(k) != nullptr && get(k)->t != Null; }
};

struct JParser {
  const std::string& s;
  size_t p = 0;
  explicit JParser(const std::string& s_) : s(s_) {}
  void ws() {
    while (p < s.size() && std::isspace((unsigned char)s[p])) ++p;
  }
  [[noreturn]] void fail(const char* m) { throw std::runtime_error(std::string("invalid JSON options: ") + m); }
  JVal value() {
    ws();
    if (p >= s.size()) fail("unexpected end");
    char c = s[p];
    JVal v;
    if (c == '{') {
      v.t = JVal::Obj;
      ++p;
      ws();
      if (s[p] == '}') {
        ++p;
        return v;
      }
      for (;;) {
        ws();
        JVal k = value();
        if (k.t != JVal::Str) fail("object key must be a string");
        ws();
        if (s[p] != ':') fail("expected ':'");
        ++p;
        JVal x = value();
        v.obj.emplace_back(k.s, x);
        ws();
        if (s[p] == ',') {
          ++p;
          continue;
        }
        if (s[p] == '}') {
          ++p;
          break;
        }
        fail("expected ',' or '}'");
      }
      return v;
    }
    if (c == '[') {
      v.t = JVal::Arr;
      ++p;
      ws();
      if (s[p] == ']') {
        ++p;
        return v;
      }
      for (;;) {
        v.arr.push_back(value());
        ws();
        if (s[p] == ',') {
          ++p;
          continue;
        }
        if (s[p] == ']') {
          ++p;
          break;
        }
        fail("expected ',' or ']'");
      }
      return v;
    }
    if (c == '"') {
      v.t = JVal::Str;
      ++p;
      while (p < s.size() && s[p] != '"') {
        if (s[p] == '\\' && p + 1 < s.size()) {
          ++p;
          char e = s[p];
          v.s += (e == 'n') ? '\n' : (e == 't') ? '\t' : e;
        } else {
          v.s += s[p];
        }
        ++p;
      }
      ++p;
      return v;
    }
    if (s.compare(p, 4, "true") == 0) {
      v.t = JVal::Bool;
      v.b = true;
      p += 4;
      return v;
    }
    if (s.compare(p, 5, "false") == 0) {
      v.t = JVal::Bool;
      p += 5;
      return v;
    }
    if (s.compare(p, 4, "null") == 0) {
      p += 4;
      return v;
    }
    char* e = nullptr;
    v.t = JVal::Num;
    v.n = std::strtod(s.c_str() + p, &e);
    if (e == s.c_str() + p) fail("bad value");
    p = (size_t)(e - s.c_str());
    return v;
  }
};

inline JVal json_parse(const std::string& s) {
  if (s.empty()) {
    JVal v;
    v.t = JVal::Obj;
    return v;
  }
  JParser P(s);
  JVal v = P.value();
  return v;
}

inline std::string json_str(const std::string& s) {
  std::string o = "\"";
  for (char c : s) {
    if (c == '"' || c == '\\') o += '\\';
    if (c == '\n') {
      o += "\\n";
      continue;
    }
    o += c;
  }
  return o + "\"";
}

}  // namespace ps

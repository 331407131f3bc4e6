#include "json.h"

#include <cstdlib>
#include <stdexcept>

namespace rad::json {
namespace {

struct Parser {
  const std::string& s;
  size_t i = 0;

  [[noreturn]] void Fail(const char* what) {
    size_t line = 1, col = 1;
    for (size_t k = 0; k < i && k < s.size(); k++) {
      if (s[k] == '\n') {
        line++;
        col = 1;
      } else {
        col++;
      }
    }
    throw std::runtime_error(std::string("JSON parse error: ") + what + " at line " + std::to_string(line) + ", column " +
                             std::to_string(col));
  }
  void Ws() {
    while (i < s.size()) {
      char c = s[i];
      if (c == ' ' || c == '\t' || c == '\n' || c == '\r') {
        i++;
      } else if (c == '/' && i + 1 < s.size() && s[i + 1] == '/') {  // tolerate line comments
        while (i < s.size() && s[i] != '\n') i++;
      } else {
        break;
      }
    }
  }
  bool Lit(const char* w) {
    size_t n = 0;
    while (w[n]) n++;
    if (s.compare(i, n, w) == 0) {
      i += n;
      return true;
    }
    return false;
  }
  std::string Str() {
    if (s[i] != '"') Fail("expected string");
    i++;
    std::string out;
    while (i < s.size() && s[i] != '"') {
      char c = s[i++];
      if (c == '\\') {
        if (i >= s.size()) Fail("bad escape");
        char e = s[i++];
        switch (e) {
          case 'n': out += '\n'; break;
          case 't': out += '\t'; break;
          case 'r': out += '\r'; break;
          case 'b': out += '\b'; break;
          case 'f': out += '\f'; break;
          case 'u': {
            if (i + 4 > s.size()) Fail("bad unicode escape");
            unsigned cp = std::stoul(s.substr(i, 4), nullptr, 16);
            i += 4;
            if (cp < 0x80) {
              out += char(cp);
            } else if (cp < 0x800) {
              out += char(0xC0 | (cp >> 6));
              out += char(0x80 | (cp & 0x3F));
            } else {
              out += char(0xE0 | (cp >> 12));
              out += char(0x80 | ((cp >> 6) & 0x3F));
              out += char(0x80 | (cp & 0x3F));
            }
            break;
          }
          default: out += e;
        }
      } else {
        out += c;
      }
    }
    if (i >= s.size()) Fail("unterminated string");
    i++;
    return out;
  }
  Value Val(int depth) {
    if (depth > 256) Fail("nesting too deep");
    Ws();
    if (i >= s.size()) Fail("unexpected end");
    Value v;
    char c = s[i];
    if (c == '{') {
      v.kind = Value::Object;
      i++;
      Ws();
      if (s[i] == '}') {
        i++;
        return v;
      }
      for (;;) {
        Ws();
        std::string k = Str();
        Ws();
        if (s[i] != ':') Fail("expected ':'");
        i++;
        v.obj.emplace_back(std::move(k), Val(depth + 1));
        Ws();
        if (s[i] == ',') {
          i++;
          continue;
        }
        if (s[i] == '}') {
          i++;
          break;
        }
        Fail("expected ',' or '}'");
      }
    } else if (c == '[') {
      v.kind = Value::Array;
      i++;
      Ws();
      if (s[i] == ']') {
        i++;
        return v;
      }
      for (;;) {
        v.arr.push_back(Val(depth + 1));
        Ws();
        if (s[i] == ',') {
          i++;
          continue;
        }
        if (s[i] == ']') {
          i++;
          break;
        }
        Fail("expected ',' or ']'");
      }
    } else if (c == '"') {
      v.kind = Value::String;
      v.str = Str();
    } else if (Lit("true")) {
      v.kind = Value::Bool;
      v.b = true;
    } else if (Lit("false")) {
      v.kind = Value::Bool;
      v.b = false;
    } else if (Lit("null")) {
      v.kind = Value::Null;
    } else {
      const char* b = s.c_str() + i;
      char* e = nullptr;
      v.num = std::strtod(b, &e);
      if (e == b) Fail("unexpected character");
      v.kind = Value::Number;
      i += size_t(e - b);
    }
    return v;
  }
};

}  // namespace

Value Parse(const std::string& text) {
  Parser p{text};
  Value v = p.Val(0);
  p.Ws();
  if (p.i != text.size()) p.Fail("trailing characters");
  return v;
}

}  // namespace rad::json

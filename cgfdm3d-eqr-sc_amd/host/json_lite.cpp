#include "json_lite.h"

#include <cctype>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <strings.h>

namespace cgh {

namespace {

struct Reader {
  const std::string& s;
  size_t p = 0;
  std::string err;

  explicit Reader(const std::string& t) : s(t) {}

  void ws() {
    while (p < s.size() && (s[p] == ' ' || s[p] == '\t' || s[p] == '\n' || s[p] == '\r')) ++p;
  }
  bool fail(const char* m) {
    if (err.empty()) err = std::string(m) + " at offset " + std::to_string(p);
    return false;
  }
  bool lit(const char* w) {
    size_t n = std::strlen(w);
    if (s.compare(p, n, w) != 0) return false;
    p += n;
    return true;
  }
  static void put_utf8(std::string& o, unsigned cp) {
    if (cp < 0x80) {
      o += (char)cp;
    } else if (cp < 0x800) {
      o += (char)(0xC0 | (cp >> 6));
      o += (char)(0x80 | (cp & 0x3F));
    } else if (cp < 0x10000) {
      o += (char)(0xE0 | (cp >> 12));
      o += (char)(0x80 | ((cp >> 6) & 0x3F));
      o += (char)(0x80 | (cp & 0x3F));
    } else {
      o += (char)(0xF0 | (cp >> 18));
      o += (char)(0x80 | ((cp >> 12) & 0x3F));
      o += (char)(0x80 | ((cp >> 6) & 0x3F));
      o += (char)(0x80 | (cp & 0x3F));
    }
  }
  bool hex4(unsigned& v) {
    if (p + 4 > s.size()) return fail("short \\u escape");
    v = 0;
    for (int q = 0; q < 4; ++q) {
      char c = s[p++];
      v <<= 4;
      if (c >= '0' && c <= '9') v |= (unsigned)(c - '0');
      else if (c >= 'a' && c <= 'f') v |= (unsigned)(c - 'a' + 10);
      else if (c >= 'A' && c <= 'F') v |= (unsigned)(c - 'A' + 10);
      else return fail("bad \\u escape");
    }
    return true;
  }
  bool string(std::string& o) {
    if (p >= s.size() || s[p] != '"') return fail("expected string");
    ++p;
    o.clear();
    while (p < s.size()) {
      char c = s[p++];
      if (c == '"') return true;
      if (c != '\\') {
        o += c;
        continue;
      }
      if (p >= s.size()) break;
      char e = s[p++];
      switch (e) {
        case '"': o += '"'; break;
        case '\\': o += '\\'; break;
        case '/': o += '/'; break;
        case 'b': o += '\b'; break;
        case 'f': o += '\f'; break;
        case 'n': o += '\n'; break;
        case 'r': o += '\r'; break;
        case 't': o += '\t'; break;
        case 'u': {
          unsigned cp = 0;
          if (!hex4(cp)) return false;
          if (cp >= 0xD800 && cp < 0xDC00 && p + 6 <= s.size() && s[p] == '\\' && s[p + 1] == 'u') {
            p += 2;
            unsigned lo = 0;
            if (!hex4(lo)) return false;
            cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
          }
          put_utf8(o, cp);
          break;
        }
        default: return fail("bad escape");
      }
    }
    return fail("unterminated string");
  }
  bool number(JValue& v) {
    size_t b = p;
    if (p < s.size() && s[p] == '-') ++p;
    if (p >= s.size() || !std::isdigit((unsigned char)s[p])) return fail("bad number");
    while (p < s.size() && std::isdigit((unsigned char)s[p])) ++p;
    if (p < s.size() && s[p] == '.') {
      ++p;
      if (p >= s.size() || !std::isdigit((unsigned char)s[p])) return fail("bad fraction");
      while (p < s.size() && std::isdigit((unsigned char)s[p])) ++p;
    }
    if (p < s.size() && (s[p] == 'e' || s[p] == 'E')) {
      ++p;
      if (p < s.size() && (s[p] == '+' || s[p] == '-')) ++p;
      if (p >= s.size() || !std::isdigit((unsigned char)s[p])) return fail("bad exponent");
      while (p < s.size() && std::isdigit((unsigned char)s[p])) ++p;
    }
    v.kind = JValue::Num;
    v.num = std::strtod(s.substr(b, p - b).c_str(), nullptr);
    return true;
  }
  bool value(JValue& v, int depth) {
    if (depth > 200) return fail("nesting too deep");
    ws();
    if (p >= s.size()) return fail("unexpected end");
    char c = s[p];
    if (c == '{') {
      ++p;
      v.kind = JValue::Obj;
      ws();
      if (p < s.size() && s[p] == '}') {
        ++p;
        return true;
      }
      for (;;) {
        ws();
        std::string k;
        if (!string(k)) return false;
        ws();
        if (p >= s.size() || s[p] != ':') return fail("expected ':'");
        ++p;
        v.keys.push_back(k);
        v.items.emplace_back();
        if (!value(v.items.back(), depth + 1)) return false;
        ws();
        if (p < s.size() && s[p] == ',') {
          ++p;
          continue;
        }
        if (p < s.size() && s[p] == '}') {
          ++p;
          return true;
        }
        return fail("expected ',' or '}'");
      }
    }
    if (c == '[') {
      ++p;
      v.kind = JValue::Arr;
      ws();
      if (p < s.size() && s[p] == ']') {
        ++p;
        return true;
      }
      for (;;) {
        v.items.emplace_back();
        if (!value(v.items.back(), depth + 1)) return false;
        ws();
        if (p < s.size() && s[p] == ',') {
          ++p;
          continue;
        }
        if (p < s.size() && s[p] == ']') {
          ++p;
          return true;
        }
        return fail("expected ',' or ']'");
      }
    }
    if (c == '"') {
      v.kind = JValue::Str;
      return string(v.str);
    }
    if (lit("true")) {
      v.kind = JValue::Bool;
      v.b = true;
      return true;
    }
    if (lit("false")) {
      v.kind = JValue::Bool;
      v.b = false;
      return true;
    }
    if (lit("null")) {
      v.kind = JValue::Null;
      return true;
    }
    return number(v);
  }
};

}  // namespace

const JValue* JValue::get(const char* key) const {
  if (kind != Obj) return nullptr;
  for (size_t q = 0; q < keys.size(); ++q)
    if (strcasecmp(keys[q].c_str(), key) == 0) return &items[q];
  return nullptr;
}

int JValue::as_int() const {
  double v = as_double();
  if (v >= (double)INT_MAX) return INT_MAX;
  if (v <= (double)INT_MIN) return INT_MIN;
  return (int)v;
}

double JValue::as_double() const {
  if (kind == Num) return num;
  if (kind == Bool) return b ? 1.0 : 0.0;
  return 0.0;
}

bool parse_json(const std::string& text, JValue& out, std::string& err) {
  Reader r(text);
  out = JValue();
  if (!r.value(out, 0)) {
    err = r.err;
    return false;
  }
  r.ws();
  // cJSON_Parse accepts trailing bytes after the first value; so does this reader
  return true;
}

}  // namespace cgh

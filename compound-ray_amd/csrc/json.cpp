#include "json.h"

#include <cstdlib>
#include <cstring>

namespace eye {

static const JsonValue& null_value() {
  static const JsonValue v;
  return v;
}

const JsonValue& JsonValue::operator[](const std::string& key) const {
  if (kind != Object) return null_value();
  auto it = obj.find(key);
  return it == obj.end() ? null_value() : it->second;
}

const JsonValue& JsonValue::operator[](size_t i) const {
  if (kind != Array || i >= arr.size()) return null_value();
  return arr[i];
}

namespace {

struct Parser {
  const char* p;
  const char* end;
  std::string err;

  void ws() {
    while (p < end && (*p == ' ' || *p == '\n' || *p == '\r' || *p == '\t')) ++p;
  }
  bool fail(const char* msg) {
    if (err.empty()) err = msg;
    return false;
  }
  static void put_utf8(std::string& s, unsigned cp) {
    if (cp < 0x80) {
      s += (char)cp;
    } else if (cp < 0x800) {
      s += (char)(0xC0 | (cp >> 6));
      s += (char)(0x80 | (cp & 0x3F));
    } else if (cp < 0x10000) {
      s += (char)(0xE0 | (cp >> 12));
      s += (char)(0x80 | ((cp >> 6) & 0x3F));
      s += (char)(0x80 | (cp & 0x3F));
    } else {
      s += (char)(0xF0 | (cp >> 18));
      s += (char)(0x80 | ((cp >> 12) & 0x3F));
      s += (char)(0x80 | ((cp >> 6) & 0x3F));
      s += (char)(0x80 | (cp & 0x3F));
    }
  }
  bool hex4(unsigned* out) {
    if (end - p < 4) return fail("truncated \\u escape");
    unsigned v = 0;
    for (int i = 0; i < 4; ++i) {
      const char c = *p++;
      v <<= 4;
      if (c >= '0' && c <= '9') v |= (unsigned)(c - '0');
      else if (c >= 'a' && c <= 'f') v |= (unsigned)(c - 'a' + 10);
      else if (c >= 'A' && c <= 'F') v |= (unsigned)(c - 'A' + 10);
      else return fail("bad \\u escape");
    }
    *out = v;
    return true;
  }
  bool string(std::string* out) {
    if (p >= end || *p != '"') return fail("expected string");
    ++p;
    out->clear();
    while (p < end && *p != '"') {
      char c = *p++;
      if (c != '\\') {
        *out += c;
        continue;
      }
      if (p >= end) return fail("truncated escape");
      c = *p++;
      switch (c) {
        case '"': *out += '"'; break;
        case '\\': *out += '\\'; break;
        case '/': *out += '/'; break;
        case 'b': *out += '\b'; break;
        case 'f': *out += '\f'; break;
        case 'n': *out += '\n'; break;
        case 'r': *out += '\r'; break;
        case 't': *out += '\t'; break;
        case 'u': {
          unsigned cp;
          if (!hex4(&cp)) return false;
          if (cp >= 0xD800 && cp < 0xDC00 && end - p >= 6 && p[0] == '\\' && p[1] == 'u') {
            p += 2;
            unsigned lo;
            if (!hex4(&lo)) return false;
            cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
          }
          put_utf8(*out, cp);
          break;
        }
        default: return fail("bad escape");
      }
    }
    if (p >= end) return fail("unterminated string");
    ++p;
    return true;
  }
  bool value(JsonValue* v, int depth) {
    if (depth > 256) return fail("nesting too deep");
    ws();
    if (p >= end) return fail("unexpected end of input");
    const char c = *p;
    if (c == '{') {
      ++p;
      v->kind = JsonValue::Object;
      ws();
      if (p < end && *p == '}') { ++p; return true; }
      for (;;) {
        ws();
        std::string key;
        if (!string(&key)) return false;
        ws();
        if (p >= end || *p != ':') return fail("expected ':'");
        ++p;
        JsonValue child;
        if (!value(&child, depth + 1)) return false;
        v->obj[key] = std::move(child);
        ws();
        if (p < end && *p == ',') { ++p; continue; }
        if (p < end && *p == '}') { ++p; return true; }
        return fail("expected ',' or '}'");
      }
    }
    if (c == '[') {
      ++p;
      v->kind = JsonValue::Array;
      ws();
      if (p < end && *p == ']') { ++p; return true; }
      for (;;) {
        JsonValue child;
        if (!value(&child, depth + 1)) return false;
        v->arr.push_back(std::move(child));
        ws();
        if (p < end && *p == ',') { ++p; continue; }
        if (p < end && *p == ']') { ++p; return true; }
        return fail("expected ',' or ']'");
      }
    }
    if (c == '"') {
      v->kind = JsonValue::String;
      return string(&v->str);
    }
    if (end - p >= 4 && !strncmp(p, "true", 4)) { p += 4; v->kind = JsonValue::Bool; v->b = true; return true; }
    if (end - p >= 5 && !strncmp(p, "false", 5)) { p += 5; v->kind = JsonValue::Bool; v->b = false; return true; }
    if (end - p >= 4 && !strncmp(p, "null", 4)) { p += 4; v->kind = JsonValue::Null; return true; }
    if (c == '-' || (c >= '0' && c <= '9')) {
      // strtod needs a terminated buffer: copy the numeric token
      const char* q = p;
      while (q < end && (strchr("+-0123456789.eE", *q) != nullptr)) ++q;
      std::string tok(p, q);
      char* e = nullptr;
      v->kind = JsonValue::Number;
      v->num = strtod(tok.c_str(), &e);
      if (e == tok.c_str()) return fail("bad number");
      p = q;
      return true;
    }
    return fail("unexpected character");
  }
};

}  // namespace

bool json_parse(const std::string& text, JsonValue* out, std::string* err) {
  Parser ps;
  ps.p = text.data();
  ps.end = text.data() + text.size();
  *out = JsonValue();
  if (!ps.value(out, 0)) {
    if (err) *err = "JSON: " + ps.err + " at byte " + std::to_string(ps.p - text.data());
    return false;
  }
  ps.ws();
  if (ps.p != ps.end) {
    if (err) *err = "JSON: trailing characters at byte " + std::to_string(ps.p - text.data());
    return false;
  }
  return true;
}

}  // namespace eye

// Minimal JSON reader for glTF 2.0 scene files (RFC 8259 subset: objects,
// arrays, strings with escapes, numbers, true/false/null).  Host only.
#pragma once

#include <map>
#include <memory>
#include <string>
#include <vector>

namespace eye {

struct JsonValue {
  enum Kind { Null, Bool, Number, String, Array, Object };
  Kind kind = Null;
  bool b = false;
  double num = 0.0;
  std::string str;
  std::vector<JsonValue> arr;
  std::map<std::string, JsonValue> obj;

  bool is_null() const { return kind == Null; }
  bool is_object() const { return kind == Object; }
  bool is_array() const { return kind == Array; }
  bool is_number() const { return kind == Number; }
  bool is_string() const { return kind == String; }
  bool is_bool() const { return kind == Bool; }

  // Member access; returns a shared null value when absent or not an object.
  const JsonValue& operator[](const std::string& key) const;
  const JsonValue& operator[](size_t i) const;
  bool has(const std::string& key) const { return kind == Object && obj.count(key) != 0; }
  size_t size() const { return kind == Array ? arr.size() : (kind == Object ? obj.size() : 0); }

  double as_number(double dflt = 0.0) const { return kind == Number ? num : dflt; }
  // Truncated to int; values outside the int range (and NaN) give dflt.
  int as_int(int dflt = -1) const {
    return kind == Number && num > -2147483649.0 && num < 2147483648.0 ? (int)num : dflt;
  }
  std::string as_string(const std::string& dflt = "") const { return kind == String ? str : dflt; }
};

// Parses `text`; on failure returns false and sets `err`.
bool json_parse(const std::string& text, JsonValue* out, std::string* err);

}  // namespace eye

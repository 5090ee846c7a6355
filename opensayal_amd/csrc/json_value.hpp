// json_value.hpp - a small JSON reader for OpenSayal.conf.json.
//
// The reference reads its config with nlohmann/json (src/config_parser.cpp:18),
// an un-vendored submodule that is absent here.  This reader keeps the parts
// of its behaviour the config path depends on:
//   * numbers keep their kind: integers as int64 (uint64 when above INT64_MAX),
//     anything with a fraction or exponent as double (strtod rounding);
//   * a duplicated object key keeps the last value;
//   * conversions follow nlohmann's from_json: an arithmetic target accepts
//     any number or a boolean (static_cast), a bool target accepts only a
//     boolean, anything else throws.
#pragma once

#include <cstdint>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace osy {

struct JsonError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

class Json {
 public:
  enum class Kind { Null, Bool, Int, UInt, Float, String, Array, Object };

  Json() = default;
  static Json parse(const std::string& text);  // throws JsonError

  Kind kind() const { return kind_; }
  bool is_object() const { return kind_ == Kind::Object; }
  const char* type_name() const;

  // nlohmann::json::contains: false for anything that is not an object.
  bool contains(const std::string& key) const;
  // nlohmann::json::at: throws unless an object holding `key`.
  const Json& at(const std::string& key) const;

  int64_t to_int64() const;  // arithmetic conversion (number or boolean)
  double to_double() const;  // arithmetic conversion (number or boolean)
  bool to_bool() const;      // boolean only
  const std::string& str() const;

 private:
  Kind kind_ = Kind::Null;
  bool b_ = false;
  int64_t i_ = 0;
  uint64_t u_ = 0;
  double f_ = 0.0;
  std::string s_;
  std::vector<Json> arr_;
  std::map<std::string, Json> obj_;

  friend class JsonParser;
};

// ConfigParser::get_or (inc/config_parser.hpp:131-150): a direct hit on the
// whole key wins; otherwise split at the first '.' and look the rest up in
// the child object.  A conversion error on a direct hit propagates (at the
// top level the reference dies with an uncaught exception; the config loader
// turns it into OSY_E_CONFIG); any error below the top level yields the
// default.
template <typename T, typename Conv>
T json_get_or(const Json& j, const std::string& key, T def, Conv conv) {
  if (j.contains(key)) return conv(j.at(key));
  if (key.empty()) return def;
  size_t dot = key.find('.');
  if (dot == std::string::npos) return def;
  std::string parent = key.substr(0, dot), child = key.substr(dot + 1);
  try {
    return json_get_or(j.at(parent), child, def, conv);
  } catch (...) {
    return def;
  }
}

}  // namespace osy

// json.hpp — strict RFC 8259 JSON document model for scene files.
//
// Same contract as the reference's reader (include/raytracer/io/json.hpp):
// objects keep insertion order, duplicate keys and trailing garbage are
// errors, numbers must be finite doubles, nesting is limited, and every
// parse error names its line and column. Values live in a tagged node tree
// owned by the document.
#pragma once
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace rtm {

class JsonError : public std::runtime_error {
 public:
  using std::runtime_error::runtime_error;
};

class Json {
 public:
  enum class Type : uint8_t { Null, Bool, Number, String, Array, Object };

  Json() = default;
  static Json null() { return Json(); }
  static Json boolean(bool b);
  static Json number(double d);
  static Json string(std::string s);
  static Json array();
  static Json object();

  Type type() const { return type_; }
  bool is_null() const { return type_ == Type::Null; }
  bool is_bool() const { return type_ == Type::Bool; }
  bool is_number() const { return type_ == Type::Number; }
  bool is_string() const { return type_ == Type::String; }
  bool is_array() const { return type_ == Type::Array; }
  bool is_object() const { return type_ == Type::Object; }

  bool as_bool() const;
  double as_number() const;
  const std::string& as_string() const;
  const std::vector<Json>& items() const;                           // array
  const std::vector<std::pair<std::string, Json>>& members() const;  // object

  // Object access: nullptr when absent (or not an object).
  const Json* get(const std::string& key) const;
  // Builders.
  Json& push(Json v);                         // array append
  Json& set(const std::string& key, Json v);  // object insert/replace

  // indent < 0: compact. Numbers use the shortest round-trip form.
  std::string dump(int indent = -1) const;
  static Json parse(const std::string& text, int max_depth = 128);

  bool operator==(const Json& o) const;

 private:
  Type type_ = Type::Null;
  bool b_ = false;
  double n_ = 0;
  std::string s_;
  std::vector<Json> a_;
  std::vector<std::pair<std::string, Json>> o_;
};

}  // namespace rtm

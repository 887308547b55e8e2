// Minimal JSON document model and parser for the scene format.
// Objects keep key order sorted (as nlohmann::json's std::map does); arrays
// keep element order.  Numbers are held as double.
#pragma once

#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace ash {

struct JsonError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

class Json {
 public:
  enum class Type { Null, Bool, Number, String, Array, Object };

  Json() = default;
  static Json parse(const std::string& text);  // throws JsonError

  Type type() const { return type_; }
  bool isNull() const { return type_ == Type::Null; }
  bool isNumber() const { return type_ == Type::Number; }
  bool isString() const { return type_ == Type::String; }
  bool isArray() const { return type_ == Type::Array; }
  bool isObject() const { return type_ == Type::Object; }
  bool isBool() const { return type_ == Type::Bool; }

  // nlohmann-like accessors (throw JsonError on type mismatch).
  bool contains(const std::string& key) const;
  const Json& operator[](const std::string& key) const;
  const Json& operator[](size_t i) const;
  size_t size() const;  // array/object element count, 0 otherwise
  double num() const;
  float f() const { return float(num()); }
  int i() const { return int(num()); }
  bool b() const;  // bool, or number != 0
  const std::string& str() const;
  const std::vector<Json>& arr() const;
  const std::map<std::string, Json>& obj() const;
  std::string dump() const;

 private:
  Type type_ = Type::Null;
  bool bool_ = false;
  double num_ = 0;
  std::string str_;
  std::vector<Json> arr_;
  std::map<std::string, Json> obj_;
  friend class JsonParser;
};

}  // namespace ash

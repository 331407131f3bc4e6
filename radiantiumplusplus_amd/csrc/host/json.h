// Minimal JSON document model for scene files (the reference parses them with
// nlohmann::json, module/rad.offline.cli/main.cpp:35-42).
#pragma once

#include <map>
#include <memory>
#include <string>
#include <utility>
#include <vector>

namespace rad::json {

struct Value {
  enum Kind { Null, Bool, Number, String, Array, Object } kind = Null;
  bool b = false;
  double num = 0;
  std::string str;
  std::vector<Value> arr;
  std::vector<std::pair<std::string, Value>> obj;

  bool is_number() const { return kind == Number; }
  bool is_array() const { return kind == Array; }
  bool is_object() const { return kind == Object; }
  bool is_string() const { return kind == String; }
  bool is_bool() const { return kind == Bool; }
  const Value* find(const std::string& key) const {
    if (kind != Object) return nullptr;
    for (const auto& kv : obj)
      if (kv.first == key) return &kv.second;
    return nullptr;
  }
  size_t size() const { return kind == Array ? arr.size() : (kind == Object ? obj.size() : 0); }
};

// Throws std::runtime_error with a line/column on malformed input.
Value Parse(const std::string& text);

}  // namespace rad::json

// Minimal TOML reader for scene files: the subset the reference's example
// scenes and generators use (tables, arrays of tables, dotted headers that
// address the last array element, inline tables, nested arrays, strings,
// integers, floats, booleans, comments).  The reference parses with the
// `toml` 0.8.23 crate into serde structs (src/scene.rs:847); this reader
// produces an ordered value tree that scene_load.cpp maps onto the same
// schema with the same required/optional rules.
#pragma once
#include <cstdint>
#include <memory>
#include <string>
#include <utility>
#include <vector>

namespace lmn {

struct TomlValue {
  enum Kind { STRING, INTEGER, FLOAT, BOOLEAN, ARRAY, TABLE } kind = TABLE;
  std::string str;
  int64_t integer = 0;
  double flt = 0.0;
  bool boolean = false;
  bool is_array_of_tables = false;  // created by [[header]]
  bool defined_by_header = false;   // a [header] already named this table
  std::vector<TomlValue> array;
  std::vector<std::pair<std::string, TomlValue>> table;

  const TomlValue* get(const std::string& key) const;
  TomlValue* get_mut(const std::string& key);
  TomlValue& put(const std::string& key, TomlValue v);
  bool is_number() const { return kind == INTEGER || kind == FLOAT; }
  double as_double() const { return kind == INTEGER ? double(integer) : flt; }
};

// Parses `text`; on failure returns false and writes "line N: message" to err.
bool toml_parse(const std::string& text, TomlValue& root, std::string& err);

}  // namespace lmn

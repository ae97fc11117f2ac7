// JSON values for the search report: objects keep their keys sorted (the
// reference builds its report from serde_json values without key order
// preservation, so every object is printed in key order), the printer
// follows serde_json's pretty format (two-space indent, one element per
// line, "[]" and "{}" when empty) and doubles print as the shortest text
// that reads back to the same value ("0.0", "1024.0", "1e-12",
// "2904.8351186807267"); non-finite doubles print as null.
#pragma once
#include <cstdint>
#include <map>
#include <memory>
#include <string>
#include <vector>

namespace fbls {
namespace json {

struct Value {
  enum Kind { Null, Bool, Int, Uint, Float, String, Array, Object };
  Kind kind = Null;
  bool b = false;
  int64_t i = 0;
  uint64_t u = 0;
  double f = 0.0;
  std::string s;
  std::vector<Value> a;
  std::map<std::string, Value> o;

  Value() = default;
  static Value null() { return Value(); }
  static Value boolean(bool x) {
    Value v;
    v.kind = Bool;
    v.b = x;
    return v;
  }
  static Value integer(int64_t x) {
    Value v;
    v.kind = Int;
    v.i = x;
    return v;
  }
  static Value unsigned_(uint64_t x) {
    Value v;
    v.kind = Uint;
    v.u = x;
    return v;
  }
  static Value number(double x) {
    Value v;
    v.kind = Float;
    v.f = x;
    return v;
  }
  // NaN (an absent optional value) prints as null as well
  static Value optional(double x) { return number(x); }
  static Value string(std::string x) {
    Value v;
    v.kind = String;
    v.s = std::move(x);
    return v;
  }
  static Value array() {
    Value v;
    v.kind = Array;
    return v;
  }
  static Value object() {
    Value v;
    v.kind = Object;
    return v;
  }
  Value& operator[](const std::string& key) {
    kind = Object;
    return o[key];
  }
  void push(Value x) {
    kind = Array;
    a.push_back(std::move(x));
  }
};

Value numbers(const double* x, size_t n);

// Shortest round-trip text of a finite double in serde_json's layout.
std::string format_double(double x);
// The value as serde_json::to_string_pretty prints it.
std::string pretty(const Value& v);
// Parses JSON text (numbers with '.', 'e' or 'E' are doubles, others
// integers); returns false with a message on malformed input.
bool parse(const std::string& text, Value* out, std::string* error);

}  // namespace json
}  // namespace fbls

// Small JSON reader for the parameter and station files.
//
// The reference parses params.json and station.json with cJSON and looks keys
// up with cJSON_GetObjectItem, which compares key names without regard to case
// (so "out" in a file fills params.OUT, reference getParams.cpp:186-187, and
// "srtm90" fills SRTM90, :192-193). Lookups here keep that rule and return the
// first matching member, as cJSON does. Numbers read as integers follow cJSON's
// valueint: the double value truncated toward zero and saturated to the int
// range.
#ifndef CGFDM_HOST_JSON_LITE_H
#define CGFDM_HOST_JSON_LITE_H

#include <string>
#include <utility>
#include <vector>

namespace cgh {

struct JValue {
  enum Kind { Null, Bool, Num, Str, Arr, Obj };
  Kind kind = Null;
  double num = 0.0;
  bool b = false;
  std::string str;
  std::vector<JValue> items;                                  // array elements, or object member values
  std::vector<std::string> keys;                              // object member names (same order as items)

  // member lookup without regard to case (cJSON_GetObjectItem); nullptr if absent
  const JValue* get(const char* key) const;
  // element count of an array or member count of an object (cJSON_GetArraySize)
  int size() const { return (kind == Arr || kind == Obj) ? (int)items.size() : 0; }
  // i-th element or member (cJSON_GetArrayItem)
  const JValue* at(int i) const { return (i >= 0 && i < size()) ? &items[(size_t)i] : nullptr; }
  // cJSON valueint / valuedouble of a number (booleans read as 1/0)
  int as_int() const;
  double as_double() const;
  bool is_number() const { return kind == Num || kind == Bool; }
};

// Parse a whole document. Returns false with a message on malformed input.
bool parse_json(const std::string& text, JValue& out, std::string& err);

}  // namespace cgh

#endif

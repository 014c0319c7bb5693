// Minimal strict JSON reader for the fluid scene block.
//
// Parses RFC 8259 JSON into a small tree.  Numbers keep their source text so the caller can
// apply serde_json's typing rules (an integer field rejects "1.5"; a float field accepts
// integers); duplicate keys are kept in order so the caller can reject them like serde does.
#pragma once

#include <cstdint>
#include <memory>
#include <string>
#include <utility>
#include <vector>

namespace pxl {

struct JsonValue {
    enum Kind { Null, Bool, Number, String, Array, Object } kind = Null;
    bool b = false;
    double num = 0.0;
    bool is_integer = false;  // number written without fraction or exponent
    bool negative = false;
    uint64_t uint_value = 0;  // valid when is_integer && !negative && fits
    bool uint_fits = false;
    std::string text;         // string contents, or the number's source text
    std::vector<JsonValue> items;
    std::vector<std::pair<std::string, JsonValue>> members;
};

struct JsonError {
    std::string message;
    size_t line = 1, column = 1;
};

// Returns false and fills err on malformed input.
bool json_parse(const std::string &text, JsonValue &out, JsonError &err);

// serde-style type name of a value for "invalid type" messages.
std::string json_describe(const JsonValue &v);

}  // namespace pxl

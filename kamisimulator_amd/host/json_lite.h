// json_lite.h -- JSON reader for the solver's setting files and YarnBall
// scene files.  Plays the role of the two JSON readers of the reference: the
// flat accessors of JsonParser (src/utilities/JsonParser.h: CheckKey /
// GetBool / GetInteger / GetFloat / GetString) over the top-level object, and
// the tree access of jsoncpp's Json::Value that YarnBall::buildFromJSON uses
// (YarnBall/io/jsonBuilder.cpp): a missing member or array element reads as
// null, and null converts to 0 / false.
#pragma once

#include <map>
#include <string>
#include <vector>

namespace kami_amd {

struct JsonValue {
    // NUMBERS: an array whose elements are all numbers (nums and items both
    // hold them); ARRAY: any other array.
    enum Kind { NUL, BOOL, NUMBER, STRING, NUMBERS, ARRAY, OBJECT } kind = NUL;
    bool b = false;
    double num = 0;
    std::string str;
    std::vector<double> nums;
    std::vector<JsonValue> items;
    std::map<std::string, JsonValue> members;

    bool is_null() const { return kind == NUL; }
    bool is_array() const { return kind == NUMBERS || kind == ARRAY; }
    bool is_object() const { return kind == OBJECT; }
    // elements of an array, members of an object, 0 otherwise
    size_t size() const;
    // Json::Value::operator[] const: the element or member, null when absent
    // or when this value is null; other kinds throw std::runtime_error.
    const JsonValue& operator[](size_t i) const;
    const JsonValue& operator[](const std::string& key) const;
    // asFloat / asInt / asBool: numbers and bools convert, null reads as 0 or
    // false, other kinds throw std::runtime_error naming `what`.
    float as_float(const char* what = "value") const;
    int as_int(const char* what = "value") const;
    bool as_bool(const char* what = "value") const;
    std::string as_string(const char* what = "value") const;  // null reads as ""
};

class JsonFile {
   public:
    // Parses the text; throws std::runtime_error with the byte offset on a
    // syntax error or when the top level is not an object.
    static JsonFile parse(const std::string& text);
    static JsonFile load(const std::string& path);

    bool CheckKey(const std::string& k) const { return root_.members.count(k) != 0; }
    bool GetBool(const std::string& k) const;
    int GetInteger(const std::string& k) const;
    float GetFloat(const std::string& k) const;
    std::string GetString(const std::string& k) const;
    std::vector<double> GetNumbers(const std::string& k) const;
    // The top-level object for tree access.
    const JsonValue& root() const { return root_; }
    // The text as read (GetInfoString prints it, as the reference does).
    const std::string& dump() const { return text_; }

   private:
    const JsonValue& at(const std::string& k, JsonValue::Kind want) const;
    JsonValue root_;
    std::string text_;
};

}  // namespace kami_amd

// json_lite.h -- reader for the flat JSON setting files of the solver
// (key -> number | bool | string | array of numbers).  Nested objects and
// arrays of other types are parsed and skipped.  Plays the role of the
// JsonParser the reference solver reads its settings with
// (src/utilities/JsonParser.h: CheckKey / GetBool / GetInteger / GetFloat /
// GetString).
#pragma once

#include <map>
#include <string>
#include <vector>

namespace kami_amd {

struct JsonValue {
    enum Kind { NUL, BOOL, NUMBER, STRING, NUMBERS, OTHER } kind = NUL;
    bool b = false;
    double num = 0;
    std::string str;
    std::vector<double> nums;
};

class JsonFile {
   public:
    // Parses the text; throws std::runtime_error with the byte offset on a
    // syntax error or when the top level is not an object.
    static JsonFile parse(const std::string& text);
    static JsonFile load(const std::string& path);

    bool CheckKey(const std::string& k) const { return values_.count(k) != 0; }
    bool GetBool(const std::string& k) const;
    int GetInteger(const std::string& k) const;
    float GetFloat(const std::string& k) const;
    std::string GetString(const std::string& k) const;
    std::vector<double> GetNumbers(const std::string& k) const;
    // The text as read (GetInfoString prints it, as the reference does).
    const std::string& dump() const { return text_; }

   private:
    const JsonValue& at(const std::string& k, JsonValue::Kind want) const;
    std::map<std::string, JsonValue> values_;
    std::string text_;
};

}  // namespace kami_amd

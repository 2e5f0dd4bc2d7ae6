// A small JSON reader for parameter files.  Values keep their source text,
// and typed access converts on demand, which gives the lookup rules of a
// string-valued property tree: get<int> on "1.5" is an error, get<bool>
// accepts true/false/1/0, and a number is readable as a string.
#pragma once

#include <map>
#include <string>
#include <vector>

namespace sph {
namespace json {

struct Value {
    enum Kind { NUL, BOOL, NUMBER, STRING, ARRAY, OBJECT } kind = NUL;
    std::string text;                          // scalars: the literal (strings unescaped)
    std::vector<Value> items;                  // ARRAY
    std::vector<std::pair<std::string, Value>> members;  // OBJECT, in file order

    const Value *find(const std::string &key) const;
};

// Throws SPHException("JSON parse error ...") on malformed input.
Value parse(const std::string &text);

// Typed access to a member of an object; the two-argument form throws
// SPHException("No such node (key)") when the key is missing, the
// three-argument form returns the default.  A value that does not convert
// throws SPHException("conversion of data to type failed (key)").
double get_real(const Value &obj, const std::string &key);
double get_real(const Value &obj, const std::string &key, double def);
int get_int(const Value &obj, const std::string &key);
int get_int(const Value &obj, const std::string &key, int def);
bool get_bool(const Value &obj, const std::string &key, bool def);
std::string get_string(const Value &obj, const std::string &key);
std::string get_string(const Value &obj, const std::string &key, const std::string &def);
// Elements of an array member as reals; throws when missing.
std::vector<double> get_real_array(const Value &obj, const std::string &key);

}  // namespace json
}  // namespace sph

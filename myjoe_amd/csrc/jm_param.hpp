// Joe.in parameter map with the reference's semantics (Param.h:60-317,
// Param.cpp:247-298): each line is split on " =\t", '#' starts a comment,
// the first token is the key, repeated keys append, lookups use the first
// entry, and sub-keys return the token after the name ("SCHEME=LEASTSQUARES").
#pragma once

#include <cstdlib>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace jm {

// an option of the reference that this build does not provide (JM_ERR_UNSUPPORTED)
struct Unsupported : std::invalid_argument {
  using std::invalid_argument::invalid_argument;
};

// a bad argument of a call (a size or pointer), reported as JM_ERR_ARG
struct ArgError : std::invalid_argument {
  using std::invalid_argument::invalid_argument;
};


class ParamMap {
 public:
  using Tokens = std::vector<std::string>;

  ParamMap() = default;
  explicit ParamMap(const std::string &text) { add_text(text); }

  void add_text(const std::string &text) {
    std::istringstream is(text);
    std::string line;
    while (std::getline(is, line)) add_line(line);
  }

  void add_line(std::string line) {
    size_t h = line.find('#');
    if (h != std::string::npos) line = line.substr(0, h);
    Tokens t;
    size_t pos = line.find_first_not_of(" =\t\r");
    while (pos != std::string::npos) {
      size_t end = line.find_first_of(" =\t\r", pos);
      t.push_back(line.substr(pos, end == std::string::npos ? std::string::npos : end - pos));
      pos = end == std::string::npos ? end : line.find_first_not_of(" =\t\r", end);
    }
    if (!t.empty()) map_[t[0]].push_back(t);
  }

  bool has(const std::string &k) const { return map_.count(k) != 0; }
  const Tokens *get(const std::string &k) const {
    auto it = map_.find(k);
    return it == map_.end() ? nullptr : &it->second.front();
  }
  std::string str(const std::string &k, const std::string &def) const {
    const Tokens *t = get(k);
    return (t && t->size() > 1) ? (*t)[1] : def;
  }
  std::string str(const std::string &k) const {
    const Tokens *t = get(k);
    if (!t || t->size() < 2) throw std::invalid_argument("missing parameter " + k);
    return (*t)[1];
  }
  double dbl(const std::string &k, double def) const { return has(k) ? to_d(str(k), k) : def; }
  double dbl(const std::string &k) const { return to_d(str(k), k); }
  int integer(const std::string &k, int def) const { return has(k) ? to_i(str(k), k) : def; }

  // value after a named token of line k, e.g. sub("CFL_RAMP", "MAX_CFL")
  std::string sub(const std::string &k, const std::string &name) const {
    const Tokens *t = get(k);
    if (!t) throw std::invalid_argument("missing parameter " + k);
    for (size_t i = 1; i + 1 < t->size(); i++)
      if ((*t)[i] == name) return (*t)[i + 1];
    throw std::invalid_argument("no " + name + " in parameter " + k);
  }
  double sub_d(const std::string &k, const std::string &name) const { return to_d(sub(k, name), k); }
  int sub_i(const std::string &k, const std::string &name) const { return to_i(sub(k, name), k); }

  static double to_d(const std::string &s, const std::string &k) {
    char *end = nullptr;
    double v = std::strtod(s.c_str(), &end);
    if (end == s.c_str()) throw std::invalid_argument("parameter " + k + ": '" + s + "' is not a number");
    return v;
  }
  static int to_i(const std::string &s, const std::string &k) {
    char *end = nullptr;
    long v = std::strtol(s.c_str(), &end, 10);
    if (end == s.c_str()) throw std::invalid_argument("parameter " + k + ": '" + s + "' is not an integer");
    return (int)v;
  }

 private:
  std::map<std::string, std::vector<Tokens>> map_;
};

}  // namespace jm

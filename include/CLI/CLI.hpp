/*
 * CLI/CLI.hpp -- a small command-line parser exposing the part of CLI11's
 * interface that the example programs use (examples/include/utils.hpp:40-70,
 * examples/bfs/bfs.cpp:102-112): App, add_flag, add_option, Option::check /
 * excludes / required / count, the NonNegativeNumber and PositiveNumber
 * validators, IsMember with ignore_case, and CLI11_PARSE.  CLI11 itself is a
 * third-party dependency of the reference that this image does not hold;
 * this header is an independent implementation of that subset only.
 *
 * Accepted syntax: `--name value`, `--name=value`, `-x value`, `-xvalue`,
 * grouped short flags (`-cu`), positionals in declaration order, `--` ends
 * options, `-h/--help` prints the usage and exits 0.
 */
#pragma once

#include <cstdlib>
#include <functional>
#include <iostream>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

namespace CLI {

class ParseError : public std::runtime_error {
public:
  ParseError(const std::string& what, int code) : std::runtime_error(what), _code(code) {}
  int get_exit_code() const { return _code; }

private:
  int _code;
};

/* A validator returns an empty string when the value is accepted, else the
 * reason; it may rewrite the value (IsMember maps to the canonical member). */
class Validator {
public:
  Validator() = default;
  Validator(std::string desc, std::function<std::string(std::string&)> fn) : _desc(std::move(desc)), _fn(std::move(fn)) {}
  std::string operator()(std::string& v) const { return _fn ? _fn(v) : std::string(); }
  const std::string& description() const { return _desc; }

private:
  std::string _desc;
  std::function<std::string(std::string&)> _fn;
};

namespace detail {
inline bool to_number(const std::string& s, double& out) {
  if (s.empty()) return false;
  char* end = nullptr;
  out = std::strtod(s.c_str(), &end);
  return end && *end == '\0';
}
inline std::string lower(std::string s) {
  for (auto& c : s) c = static_cast<char>(std::tolower(static_cast<unsigned char>(c)));
  return s;
}
} // namespace detail

inline const Validator NonNegativeNumber{"NONNEGATIVE", [](std::string& v) -> std::string {
                                           double d;
                                           if (!detail::to_number(v, d) || d < 0) return "value " + v + " is not a non-negative number";
                                           return {};
                                         }};
inline const Validator PositiveNumber{"POSITIVE", [](std::string& v) -> std::string {
                                        double d;
                                        if (!detail::to_number(v, d) || d <= 0) return "value " + v + " is not a positive number";
                                        return {};
                                      }};

inline std::string ignore_case(std::string s) { return detail::lower(std::move(s)); }

inline Validator IsMember(std::vector<std::string> members, std::function<std::string(std::string)> filter = nullptr) {
  std::string desc = "{";
  for (size_t i = 0; i < members.size(); ++i) desc += (i ? "," : "") + members[i];
  desc += "}";
  return Validator(desc, [members, filter](std::string& v) -> std::string {
    const std::string key = filter ? filter(v) : v;
    for (const auto& m : members)
      if ((filter ? filter(m) : m) == key) {
        v = m;
        return {};
      }
    return "value " + v + " not in the allowed set";
  });
}

class Option {
public:
  Option(std::string names, std::string desc, bool is_flag, std::function<bool(const std::string&)> set)
      : _desc(std::move(desc)), _flag(is_flag), _set(std::move(set)) {
    std::stringstream ss(names);
    std::string n;
    while (std::getline(ss, n, ',')) {
      if (n.rfind("--", 0) == 0) _long.push_back(n.substr(2));
      else if (n.rfind("-", 0) == 0) _short.push_back(n.substr(1));
      else _positional = n;
    }
    _display = names;
  }

  Option* check(const Validator& v) {
    _checks.push_back(v);
    return this;
  }
  Option* excludes(Option* other) {
    if (other) _excludes.push_back(other);
    return this;
  }
  Option* required(bool r = true) {
    _required = r;
    return this;
  }
  size_t count() const { return _count; }

  /* internal interface used by App */
  bool matches_long(const std::string& n) const {
    for (const auto& l : _long)
      if (l == n) return true;
    return false;
  }
  bool matches_short(const std::string& n) const {
    for (const auto& s : _short)
      if (s == n) return true;
    return false;
  }
  bool is_flag() const { return _flag; }
  bool is_positional() const { return !_positional.empty(); }
  bool is_required() const { return _required; }
  const std::string& display() const { return _display; }
  const std::string& desc() const { return _desc; }
  const std::vector<Option*>& excluded() const { return _excludes; }
  void apply(std::string value) {
    for (const auto& c : _checks) {
      const std::string err = c(value);
      if (!err.empty()) throw ParseError(_display + ": " + err, 105);
    }
    if (!_set(value)) throw ParseError(_display + ": could not convert '" + value + "'", 106);
    ++_count;
  }

private:
  std::string _desc, _display, _positional;
  std::vector<std::string> _long, _short;
  bool _flag = false, _required = false;
  size_t _count = 0;
  std::function<bool(const std::string&)> _set;
  std::vector<Validator> _checks;
  std::vector<Option*> _excludes;
};

class App {
public:
  explicit App(std::string description = "", std::string name = "") : _description(std::move(description)), _name(std::move(name)) {}

  Option* add_flag(const std::string& names, bool& target, const std::string& desc = "") {
    _opts.push_back(std::make_unique<Option>(names, desc, true, [&target](const std::string&) {
      target = true;
      return true;
    }));
    return _opts.back().get();
  }

  template<typename T>
  Option* add_option(const std::string& names, T& target, const std::string& desc = "") {
    _opts.push_back(std::make_unique<Option>(names, desc, false, [&target](const std::string& v) { return convert(v, target); }));
    return _opts.back().get();
  }

  void parse(int argc, char** argv) {
    if (_name.empty() && argc > 0) _name = argv[0];
    std::vector<std::string> pos;
    bool only_pos = false;
    for (int i = 1; i < argc; ++i) {
      std::string a = argv[i];
      auto next_value = [&](Option* o) -> std::string {
        if (i + 1 >= argc) throw ParseError(o->display() + " requires a value", 107);
        return argv[++i];
      };
      if (only_pos || a.size() < 2 || a[0] != '-' || is_number(a)) {
        pos.push_back(a);
      } else if (a == "--") {
        only_pos = true;
      } else if (a == "-h" || a == "--help") {
        throw ParseError(help(), 0);
      } else if (a.rfind("--", 0) == 0) {
        std::string name = a.substr(2), val;
        const size_t eq = name.find('=');
        const bool has_eq = eq != std::string::npos;
        if (has_eq) {
          val = name.substr(eq + 1);
          name = name.substr(0, eq);
        }
        Option* o = find_long(name);
        if (!o) throw ParseError("unknown option --" + name, 109);
        if (o->is_flag()) o->apply("true");
        else o->apply(has_eq ? val : next_value(o));
      } else {
        for (size_t k = 1; k < a.size(); ++k) {
          Option* o = find_short(a.substr(k, 1));
          if (!o) throw ParseError("unknown option -" + a.substr(k, 1), 109);
          if (o->is_flag()) {
            o->apply("true");
            continue;
          }
          o->apply(k + 1 < a.size() ? a.substr(k + 1) : next_value(o));
          break;
        }
      }
    }
    size_t p = 0;
    for (auto& o : _opts) {
      if (!o->is_positional()) continue;
      if (p < pos.size()) o->apply(pos[p++]);
    }
    if (p < pos.size()) throw ParseError("unexpected argument " + pos[p], 109);
    for (auto& o : _opts) {
      if (o->is_required() && o->count() == 0) throw ParseError(o->display() + " is required", 106);
      if (o->count())
        for (Option* x : o->excluded())
          if (x->count()) throw ParseError(o->display() + " excludes " + x->display(), 109);
    }
  }

  int exit(const ParseError& e) const {
    if (e.get_exit_code() == 0) std::cout << e.what();
    else std::cerr << e.what() << "\nRun with --help for more information." << std::endl;
    return e.get_exit_code();
  }

  std::string help() const {
    std::ostringstream s;
    if (!_description.empty()) s << _description << "\n";
    s << "Usage: " << _name << " [OPTIONS]";
    for (auto& o : _opts)
      if (o->is_positional()) s << " " << o->display();
    s << "\n\nOptions:\n  -h,--help  Print this help message and exit\n";
    for (auto& o : _opts) s << "  " << o->display() << "  " << o->desc() << "\n";
    return s.str();
  }

private:
  static bool is_number(const std::string& a) {
    double d;
    return detail::to_number(a, d);
  }
  template<typename T>
  static bool convert(const std::string& v, T& out) {
    if constexpr (std::is_same_v<T, std::string>) {
      out = v;
      return true;
    } else if constexpr (std::is_same_v<T, bool>) {
      const std::string l = detail::lower(v);
      out = !(l == "false" || l == "0" || l == "off" || l == "no");
      return true;
    } else {
      std::istringstream ss(v);
      if constexpr (std::is_integral_v<T> && std::is_unsigned_v<T>)
        if (!v.empty() && v[0] == '-') return false;
      T t{};
      ss >> t;
      if (ss.fail() || !ss.eof()) return false;
      out = t;
      return true;
    }
  }
  Option* find_long(const std::string& n) {
    for (auto& o : _opts)
      if (o->matches_long(n)) return o.get();
    return nullptr;
  }
  Option* find_short(const std::string& n) {
    for (auto& o : _opts)
      if (o->matches_short(n)) return o.get();
    return nullptr;
  }

  std::string _description, _name;
  std::vector<std::unique_ptr<Option>> _opts;
};

} // namespace CLI

#define CLI11_PARSE(app, argc, argv)    \
  try {                                 \
    (app).parse((argc), (argv));        \
  } catch (const CLI::ParseError& e_) { \
    return (app).exit(e_);              \
  }

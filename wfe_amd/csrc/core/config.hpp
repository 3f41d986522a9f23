#pragma once
// Derived in part from Xassemblianist/WFE, Copyright (c) 2026 Omer Kaan (Xassemblianist),
// MIT License; the licence text is in THIRD_PARTY_NOTICES at the repository root.

#include <map>
#include <set>
#include <string>
#include <vector>

namespace wfe {

// "key = value" case file reader with the reference's semantics
// (src/core/config.hpp:15-30, src/core/config.cpp:17-60):
//  - '#' and ';' start a comment anywhere on a line;
//  - lines without '=' are ignored; keys and values are trimmed;
//  - a later duplicate key wins; command-line overrides use set();
//  - numbers are parsed with atof/atoi (so "12abc" -> 12, "" -> 0);
//  - every getter marks its key as used; unused() lists keys never read,
//    except proj_* and terrain_source, which only the preprocessing reads.
class Config {
 public:
  bool load(const std::string& path);
  bool load_text(const std::string& text);
  void set(const std::string& key, const std::string& val) { kv_[key] = val; }
  bool has(const std::string& key) const { return kv_.count(key) > 0; }
  double get_real(const std::string& key, double def) const;
  int get_int(const std::string& key, int def) const;
  std::string get_str(const std::string& key, const std::string& def) const;
  const std::map<std::string, std::string>& raw() const { return kv_; }
  std::vector<std::string> unused() const;

 private:
  void parse_line(std::string line);
  std::map<std::string, std::string> kv_;
  mutable std::set<std::string> used_;
};

}  // namespace wfe

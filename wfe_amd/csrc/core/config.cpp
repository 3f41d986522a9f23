// Derived in part from Xassemblianist/WFE, Copyright (c) 2026 Omer Kaan (Xassemblianist),
// MIT License; the licence text is in THIRD_PARTY_NOTICES at the repository root.
#include "core/config.hpp"

#include <cstdlib>
#include <fstream>
#include <sstream>

namespace wfe {

namespace {
std::string strip(const std::string& s) {
  const char* ws = " \t\r\n";
  size_t a = s.find_first_not_of(ws);
  if (a == std::string::npos) return std::string();
  size_t b = s.find_last_not_of(ws);
  return s.substr(a, b - a + 1);
}
}  // namespace

void Config::parse_line(std::string line) {
  size_t cmt = line.find_first_of("#;");
  if (cmt != std::string::npos) line.resize(cmt);
  size_t eq = line.find('=');
  if (eq == std::string::npos) return;
  std::string key = strip(line.substr(0, eq));
  if (key.empty()) return;
  kv_[key] = strip(line.substr(eq + 1));
}

bool Config::load(const std::string& path) {
  std::ifstream in(path);
  if (!in) return false;
  std::string line;
  while (std::getline(in, line)) parse_line(line);
  return true;
}

bool Config::load_text(const std::string& text) {
  std::istringstream in(text);
  std::string line;
  while (std::getline(in, line)) parse_line(line);
  return true;
}

double Config::get_real(const std::string& key, double def) const {
  used_.insert(key);
  auto it = kv_.find(key);
  if (it == kv_.end()) return def;
  return std::atof(it->second.c_str());
}

int Config::get_int(const std::string& key, int def) const {
  used_.insert(key);
  auto it = kv_.find(key);
  if (it == kv_.end()) return def;
  return std::atoi(it->second.c_str());
}

std::string Config::get_str(const std::string& key, const std::string& def) const {
  used_.insert(key);
  auto it = kv_.find(key);
  return it == kv_.end() ? def : it->second;
}

std::vector<std::string> Config::unused() const {
  std::vector<std::string> keys;
  for (const auto& kv : kv_) {
    const std::string& k = kv.first;
    if (used_.count(k)) continue;
    if (k.compare(0, 5, "proj_") == 0) continue;
    if (k == "terrain_source") continue;
    keys.push_back(k);
  }
  return keys;
}

}  // namespace wfe

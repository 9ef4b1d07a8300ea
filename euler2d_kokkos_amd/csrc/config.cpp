// config.cpp -- see config.hpp for the reference semantics reproduced here.
#include "config.hpp"

#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>

namespace e2d {

namespace {

constexpr size_t kLineChunk = 200;  // fgets buffer: at most 199 characters per read
constexpr size_t kNameCap = 50;     // section / continuation name storage

bool is_space(char c) { return std::isspace(static_cast<unsigned char>(c)) != 0; }

std::string strip(const std::string &s) {
  size_t b = 0, e = s.size();
  while (b < e && is_space(s[b])) ++b;
  while (e > b && is_space(s[e - 1])) --e;
  return s.substr(b, e - b);
}

// Position of `stop`, or of a ';' preceded by whitespace, or npos.
size_t find_stop(const std::string &s, size_t from, char stop) {
  bool prev_space = false;
  for (size_t k = from; k < s.size(); ++k) {
    if (s[k] == stop) return k;
    if (prev_space && s[k] == ';') return k;
    prev_space = is_space(s[k]);
  }
  return std::string::npos;
}

std::string clip(const std::string &s) { return s.size() < kNameCap ? s : s.substr(0, kNameCap - 1); }

}  // namespace

std::string ConfigMap::key(const std::string &section, const std::string &name) {
  std::string k = section + "." + name;
  for (auto &c : k) c = static_cast<char>(std::tolower(static_cast<unsigned char>(c)));
  return k;
}

int ConfigMap::load(const std::string &path) {
  FILE *f = std::fopen(path.c_str(), "r");
  if (!f) return -1;
  std::string section, last_name;
  int lineno = 0, first_error = 0;
  char chunk[kLineChunk];
  while (std::fgets(chunk, static_cast<int>(sizeof(chunk)), f)) {
    ++lineno;
    const std::string raw(chunk);
    const std::string line = strip(raw);
    const bool indented = !line.empty() && is_space(raw[0]);
    if (!last_name.empty() && indented) {
      values_[key(section, last_name)] = line;  // continuation replaces the value
      continue;
    }
    if (line.empty() || line[0] == ';' || line[0] == '#') continue;
    if (line[0] == '[') {
      const size_t close = find_stop(line, 1, ']');
      if (close != std::string::npos && line[close] == ']') {
        section = clip(line.substr(1, close - 1));
        last_name.clear();
      } else if (!first_error) {
        first_error = lineno;
      }
      continue;
    }
    const size_t eq = find_stop(line, 0, '=');
    if (eq == std::string::npos || line[eq] != '=') {
      if (!first_error) first_error = lineno;
      continue;
    }
    const std::string name = strip(line.substr(0, eq));
    std::string value = line.substr(eq + 1);
    size_t b = 0;
    while (b < value.size() && is_space(value[b])) ++b;
    value = value.substr(b);
    const size_t cm = find_stop(value, 0, '\0');
    if (cm != std::string::npos && value[cm] == ';') value = value.substr(0, cm);
    value = strip(value);
    last_name = clip(name);
    values_[key(section, name)] = value;
  }
  std::fclose(f);
  return first_error;
}

std::string ConfigMap::getString(const std::string &section, const std::string &name,
                                 const std::string &default_value) const {
  auto it = values_.find(key(section, name));
  return it == values_.end() ? default_value : it->second;
}

void ConfigMap::setString(const std::string &section, const std::string &name,
                          const std::string &value) {
  values_[key(section, name)] = value;
}

long ConfigMap::getInteger(const std::string &section, const std::string &name,
                           long default_value) const {
  const std::string v = getString(section, name, "");
  const char *s = v.c_str();
  char *end = nullptr;
  const long n = std::strtol(s, &end, 0);
  return end > s ? n : default_value;
}

float ConfigMap::getFloat(const std::string &section, const std::string &name,
                          float default_value) const {
  const std::string v = getString(section, name, "");
  const char *s = v.c_str();
  char *end = nullptr;
  const float x = std::strtof(s, &end);
  return end > s ? x : default_value;
}

bool ConfigMap::getBool(const std::string &section, const std::string &name,
                        bool default_value) const {
  const std::string v = getString(section, name, "");
  if (v.empty()) return default_value;
  if (v == "1" || v == "yes" || v == "true" || v == "on") return true;
  if (v == "0" || v == "no" || v == "false" || v == "off") return false;
  return default_value;
}

}  // namespace e2d

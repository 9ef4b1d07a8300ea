// config.hpp -- .ini reader with the semantics of the reference's ConfigMap.
//
// Reference behaviour reproduced (pkestene/euler2d_kokkos):
//   config/inih/ini.cpp:16-152   line reader: 199-char line chunks, ';' comments
//                                only after whitespace, '#' / ';' comments at line
//                                start, indented lines continue the previous name
//                                (INI_ALLOW_MULTILINE=1, ini.h:37-39), section
//                                names clipped to 49 characters
//   config/inih/INIReader.cpp:36-100  lower-cased "section.name" keys, last
//                                assignment wins, getInteger = strtol base 0
//   config/ConfigMap.cpp:31-76   getFloat = strtof (float!), getBool
#pragma once

#include <map>
#include <string>

namespace e2d {

class ConfigMap {
 public:
  ConfigMap() = default;
  explicit ConfigMap(const std::string &path) { parse_error_ = load(path); }

  // 0 on success, first malformed line number, or -1 when the file is missing
  int ParseError() const { return parse_error_; }

  std::string getString(const std::string &section, const std::string &name,
                        const std::string &default_value) const;
  long getInteger(const std::string &section, const std::string &name, long default_value) const;
  float getFloat(const std::string &section, const std::string &name, float default_value) const;
  bool getBool(const std::string &section, const std::string &name, bool default_value) const;
  void setString(const std::string &section, const std::string &name, const std::string &value);

 private:
  int load(const std::string &path);
  static std::string key(const std::string &section, const std::string &name);
  std::map<std::string, std::string> values_;
  int parse_error_ = 0;
};

}  // namespace e2d

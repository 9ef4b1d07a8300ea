// ref_config_binding.cpp -- TEST INFRASTRUCTURE ONLY.
//
// C-ABI binding over the reference's own configuration reader
// (config/ConfigMap.{h,cpp}, config/inih/INIReader.{h,cpp}, config/inih/ini.cpp),
// compiled unmodified from /root/reference by oracle/Makefile into
// oracle/_ref/libref_config.so.  Tests use it to pin the oracle's and the
// product's .ini parsing (ConfigMap::getFloat uses strtof: ConfigMap.cpp:31-40).
#include "ConfigMap.h"

#include <cstring>
#include <string>

extern "C" {

void *refcfg_open(const char *path) { return new ConfigMap(std::string(path)); }

int refcfg_parse_error(void *h) { return static_cast<ConfigMap *>(h)->ParseError(); }

void refcfg_close(void *h) { delete static_cast<ConfigMap *>(h); }

int refcfg_get_string(void *h, const char *section, const char *name, const char *def, char *out,
                      int out_len) {
  std::string v = static_cast<ConfigMap *>(h)->getString(section, name, def);
  if (out && out_len > 0) {
    std::strncpy(out, v.c_str(), static_cast<size_t>(out_len));
    out[out_len - 1] = '\0';
  }
  return static_cast<int>(v.size());
}

long refcfg_get_integer(void *h, const char *section, const char *name, long def) {
  return static_cast<ConfigMap *>(h)->getInteger(section, name, def);
}

float refcfg_get_float(void *h, const char *section, const char *name, float def) {
  return static_cast<ConfigMap *>(h)->getFloat(section, name, def);
}

int refcfg_get_bool(void *h, const char *section, const char *name, int def) {
  return static_cast<ConfigMap *>(h)->getBool(section, name, def != 0) ? 1 : 0;
}
}

// C shim over the reference's own INI reader (src/config/ConfigMap.cpp,
// inih/INIReader.cpp, inih/ini.cpp), compiled from the reference tree by
// oracle/ref.mk into oracle/_ref/libref_ini.so.  TEST INFRASTRUCTURE ONLY:
// tests/test_config.py uses it to check the C++ facade's ConfigMap.
#include <cstring>
#include <string>

#include "ConfigMap.h"

extern "C" {
void *ref_ini_open(const char *path) { return new ConfigMap(std::string(path)); }
void ref_ini_close(void *h) { delete static_cast<ConfigMap *>(h); }
int ref_ini_error(void *h) { return static_cast<ConfigMap *>(h)->ParseError(); }
long ref_ini_get_integer(void *h, const char *s, const char *n, long def) {
    return static_cast<ConfigMap *>(h)->getInteger(s, n, def);
}
double ref_ini_get_float(void *h, const char *s, const char *n, float def) {
    return static_cast<ConfigMap *>(h)->getFloat(s, n, def);
}
int ref_ini_get_bool(void *h, const char *s, const char *n, int def) {
    return static_cast<ConfigMap *>(h)->getBool(s, n, def != 0) ? 1 : 0;
}
int ref_ini_get_string(void *h, const char *s, const char *n, const char *def, char *out, int len) {
    const std::string v = static_cast<ConfigMap *>(h)->getString(s, n, def);
    std::strncpy(out, v.c_str(), (size_t)len - 1);
    out[len - 1] = '\0';
    return (int)v.size();
}
}

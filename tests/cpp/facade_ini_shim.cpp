// The same C API as oracle/ref_ini_shim.cpp, over the C++ facade's ConfigMap
// and ADVParams (parallel-advection_amd/cpp/bkma/config.hpp), so the tests can
// feed both readers the same files.
#include <cstring>
#include <string>

#include "bkma/config.hpp"

using bkma::ConfigMap;

extern "C" {
void *fac_ini_open(const char *path) { return new ConfigMap(std::string(path)); }
void fac_ini_close(void *h) { delete static_cast<ConfigMap *>(h); }
int fac_ini_error(void *h) { return static_cast<ConfigMap *>(h)->ParseError(); }
long fac_ini_get_integer(void *h, const char *s, const char *n, long def) {
    return static_cast<ConfigMap *>(h)->getInteger(s, n, def);
}
double fac_ini_get_float(void *h, const char *s, const char *n, float def) {
    return static_cast<ConfigMap *>(h)->getFloat(s, n, def);
}
int fac_ini_get_bool(void *h, const char *s, const char *n, int def) {
    return static_cast<ConfigMap *>(h)->getBool(s, n, def != 0) ? 1 : 0;
}
int fac_ini_get_string(void *h, const char *s, const char *n, const char *def, char *out, int len) {
    const std::string v = static_cast<ConfigMap *>(h)->getString(s, n, def);
    std::strncpy(out, v.c_str(), (size_t)len - 1);
    out[len - 1] = '\0';
    return (int)v.size();
}
// ADVParamsNonCopyable::setup on a file; fills the POD parameters.
int fac_adv_setup(const char *path, bkma_adv_params *out, char *impl, int len) {
    ConfigMap c{std::string(path)};
    bkma::ADVParamsNonCopyable p;
    p.setup(c);
    *out = p;
    std::strncpy(impl, p.kernelImpl.c_str(), (size_t)len - 1);
    impl[len - 1] = '\0';
    return c.ParseError();
}
}

extern "C" {
// Conv1dParamsNonCopyable::setup on a file: out = {length, channel_in,
// channel_out, k, total_batch_size, batch_size_n2, n0, n1, n2, n_write, gpu,
// pref_wg_size, seq_size0, seq_size2, inplace}.
int fac_conv_setup(const char *path, unsigned long long out[15], char *impl, int len) {
    ConfigMap c{std::string(path)};
    bkma::Conv1dParamsNonCopyable p;
    p.setup(c);
    const unsigned long long v[15] = {p.length, p.channel_in, p.channel_out, p.k, p.total_batch_size,
                                      p.batch_size_n2, p.n0, p.n1, p.n2, p.n_write, (unsigned long long)p.gpu,
                                      p.pref_wg_size, p.seq_size0, p.seq_size2, (unsigned long long)p.inplace};
    for (int i = 0; i < 15; ++i) out[i] = v[i];
    std::strncpy(impl, p.kernelImpl.c_str(), (size_t)len - 1);
    impl[len - 1] = '\0';
    return c.ParseError();
}
}

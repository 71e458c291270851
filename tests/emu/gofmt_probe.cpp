// Test access to the Go-formatting helpers the library and the geofdtd
// command share (test infrastructure).
#include <string.h>

#include "gofmt.h"

static int put(const std::string &s, char *out, int cap) {
  if (cap <= 0) return -1;
  strncpy(out, s.c_str(), (size_t)cap - 1);
  out[cap - 1] = 0;
  return (int)s.size();
}

extern "C" int emu_go_g(double x, char *out, int cap) { return put(gofmt::g(x), out, cap); }
extern "C" int emu_go_g_prec(double x, int prec, char *out, int cap) { return put(gofmt::g_prec(x, prec), out, cap); }
extern "C" int emu_go_quote(const char *s, char *out, int cap) { return put(gofmt::quote(s), out, cap); }

// gofmt.h — Go's number and string formatting, for messages and files that
// must read byte for byte as the reference's Go code writes them.
//
//   gofmt::g(x)        fmt's %g / strconv.FormatFloat(x, 'g', -1, 64)
//   gofmt::g_prec(x,p) strconv.FormatFloat(x, 'g', p, 64)
//   gofmt::quote(s)    fmt's %q of a string
#pragma once

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <string>

namespace gofmt {

inline const char *special(double x) {
  if (isnan(x)) return "NaN";
  if (isinf(x)) return x > 0 ? "+Inf" : "-Inf";
  return nullptr;
}

// Shortest digits that read back exactly, in exponent form only when the
// decimal exponent is below -4 or at least 6 (C's %g switches at the digit
// count instead).  Go writes at least two exponent digits, as C does.
inline std::string g(double x) {
  char buf[64];
  if (const char *s = special(x)) return s;
  if (x == 0) return signbit(x) ? "-0" : "0";
  int p = 1;
  for (; p < 17; p++) {
    snprintf(buf, sizeof buf, "%.*e", p - 1, x);
    if (strtod(buf, nullptr) == x) break;
  }
  snprintf(buf, sizeof buf, "%.*e", p - 1, x);
  const int e = atoi(strchr(buf, 'e') + 1);
  if (e >= -4 && e < 6) snprintf(buf, sizeof buf, "%.*f", std::max(p - 1 - e, 0), x);
  return buf;
}

// FormatFloat(x, 'g', prec, 64), prec >= 1.  Go rounds the exact binary value
// to prec significant digits (ties to even, as glibc's printf does), drops
// trailing zeros, and uses exponent form when the exponent is below -4 or at
// least prec.  Go's extra rule, lowering the threshold to the digit count when
// the digits cover the integer part, never changes that decision for an
// explicit precision (the exponent is then below the digit count), so for
// finite values this is C's %.{prec}g.
inline std::string g_prec(double x, int prec) {
  if (const char *s = special(x)) return s;
  char buf[64];
  snprintf(buf, sizeof buf, "%.*g", prec, x);
  return buf;
}

// %q of a byte string: double quotes, with \" \\ and the C-style escapes Go
// uses for control bytes (\a \b \f \n \r \t \v, else \x..).  Printable ASCII
// passes through; bytes >= 0x80 are copied as they are, which is Go's output
// for valid UTF-8 text.
inline std::string quote(const std::string &s) {
  std::string out = "\"";
  for (unsigned char c : s) {
    switch (c) {
      case '"': out += "\\\""; break;
      case '\\': out += "\\\\"; break;
      case '\a': out += "\\a"; break;
      case '\b': out += "\\b"; break;
      case '\f': out += "\\f"; break;
      case '\n': out += "\\n"; break;
      case '\r': out += "\\r"; break;
      case '\t': out += "\\t"; break;
      case '\v': out += "\\v"; break;
      default:
        if (c < 0x20 || c == 0x7f) {
          char b[8];
          snprintf(b, sizeof b, "\\x%02x", c);
          out += b;
        } else {
          out += (char)c;
        }
    }
  }
  return out + "\"";
}

}  // namespace gofmt

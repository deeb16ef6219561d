// fmt_general.h -- a double as fmt's "{:.P}" prints it with the fmt 7.1.3 the
// reference bundles (the energy.dat line of src/main.cpp:121, the headers of
// src/utils/Printer.cpp): general notation with P significant digits that
// keeps its trailing zeros, and at least one digit after the point
// ("0.0000", "0.10000", "12345.0", "1.0000e+05" at P = 5).
//
// Built from %.*e (P significant digits, correctly rounded, which also gives
// the decimal exponent X after rounding) and %.*f with P - 1 - X decimals:
// glibc's own "%#.*g" drops the zeros when rounding carries into a new digit
// (99999.5 at P = 5 prints "1.e+05").
#pragma once

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

namespace ch {

inline std::string fmt_general(double v, int prec) {
    char buf[512];
    if(!std::isfinite(v)) {
        std::snprintf(buf, sizeof(buf), "%g", v);
        return buf;
    }
    if(prec < 1) prec = 1;
    std::snprintf(buf, sizeof(buf), "%.*e", prec - 1, v);
    const char *e = std::strchr(buf, 'e');
    const int X = e ? std::atoi(e + 1) : 0;
    if(X < -4 || X >= prec) return buf;  // scientific: mantissa with its zeros, exponent of 2+ digits
    std::snprintf(buf, sizeof(buf), "%.*f", prec - 1 - X, v);
    std::string s(buf);
    if(s.find('.') == std::string::npos) s += ".0";
    return s;
}

}  // namespace ch

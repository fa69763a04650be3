// Exact fixed-point formatting for the per-height output files.
//
// The reference writes every record with printf "%6.3f" / "%8.4f" / "%d"
// (simulation.cpp:183-194).  snprintf costs about a microsecond per line, so
// at 10^7 records formatting alone takes seconds.  fmt_fixed produces the
// same characters as glibc's "%W.Df": the exact binary value of the double
// is rounded to D decimals to nearest, ties to even (glibc's behaviour in
// the default rounding mode), a negative value (also -0.0 and values that
// round to zero) keeps its '-', and the field is right-aligned in W
// columns.  Non-finite and very large values fall back to snprintf.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstdio>

namespace mafm {

inline char* fmt_uint(char* p, uint64_t v) {
    char tmp[24];
    int n = 0;
    do {
        tmp[n++] = (char)('0' + v % 10);
        v /= 10;
    } while (v);
    while (n) *p++ = tmp[--n];
    return p;
}

inline char* fmt_int(char* p, int64_t v) {
    if (v < 0) {
        *p++ = '-';
        return fmt_uint(p, (uint64_t)(-(v + 1)) + 1);
    }
    return fmt_uint(p, (uint64_t)v);
}

// "%W.Df" of v into p (no terminator); returns the end.  D <= 6.
inline char* fmt_fixed(char* p, double v, int decimals, int width) {
    if (!std::isfinite(v) || std::fabs(v) >= 1e12) {
        char buf[400];
        int n = snprintf(buf, sizeof buf, "%*.*f", width, decimals, v);
        for (int i = 0; i < n; ++i) *p++ = buf[i];
        return p;
    }
    static const uint64_t P10[] = {1, 10, 100, 1000, 10000, 100000, 1000000};
    const bool neg = std::signbit(v);
    const double a = std::fabs(v);
    // a = m * 2^e exactly, m < 2^53
    int e = 0;
    const double fr = std::frexp(a, &e);
    const uint64_t m = (uint64_t)std::ldexp(fr, 53);
    e -= 53;
    const unsigned __int128 x = (unsigned __int128)m * P10[decimals];   // a * 10^D = x * 2^e, x < 2^73
    uint64_t q;
    if (e >= 0) {
        q = (uint64_t)(x << e);   // not reached for a < 1e12 (e <= -13)
    } else {
        const int s = -e;
        if (s >= 100) {
            q = 0;                // x * 2^-s < 2^-26: rounds to zero
        } else {
            const unsigned __int128 qq = x >> s;
            const unsigned __int128 rem = x - (qq << s);
            const unsigned __int128 half = (unsigned __int128)1 << (s - 1);
            q = (uint64_t)qq;
            if (rem > half || (rem == half && (q & 1))) ++q;
        }
    }
    const uint64_t ip = q / P10[decimals];
    uint64_t fp = q % P10[decimals];
    char digits[48];
    char* d = digits;
    if (neg) *d++ = '-';
    d = fmt_uint(d, ip);
    if (decimals > 0) {
        *d++ = '.';
        for (int k = decimals - 1; k >= 0; --k) {
            d[k] = (char)('0' + fp % 10);
            fp /= 10;
        }
        d += decimals;
    }
    const int len = (int)(d - digits);
    for (int k = len; k < width; ++k) *p++ = ' ';
    for (int k = 0; k < len; ++k) *p++ = digits[k];
    return p;
}

}  // namespace mafm

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
//
// The exact product of the double's 53-bit significand (read from its bits)
// and 10^D is formed and rounded in 128-bit integers; the decimal count is a
// template parameter so every division by a power of ten is a
// multiplication, and digits are written in place, two at a time from a
// table.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>

namespace mafm {

namespace fmt_detail {
// "00" "01" ... "99"
struct Pairs {
    char d[200];
    constexpr Pairs() : d() {
        for (int i = 0; i < 100; ++i) {
            d[2 * i] = (char)('0' + i / 10);
            d[2 * i + 1] = (char)('0' + i % 10);
        }
    }
};
inline constexpr Pairs kPairs{};

constexpr uint64_t pow10(int d) { return d == 0 ? 1 : 10 * pow10(d - 1); }
}  // namespace fmt_detail

// Decimal digits of v (at least 1).
inline int n_digits(uint64_t v) {
    int n = 1;
    while (v >= 100) {
        v /= 100;
        n += 2;
    }
    return n + (v >= 10);
}

// The digits of v written backwards so that they end at `end`.
inline void put_digits(char* end, uint64_t v) {
    while (v >= 100) {
        end -= 2;
        memcpy(end, fmt_detail::kPairs.d + 2 * (v % 100), 2);
        v /= 100;
    }
    if (v >= 10) memcpy(end - 2, fmt_detail::kPairs.d + 2 * v, 2);
    else end[-1] = (char)('0' + v);
}

inline char* fmt_uint(char* p, uint64_t v) {
    const int n = n_digits(v);
    put_digits(p + n, v);
    return p + n;
}

inline char* fmt_int(char* p, int64_t v) {
    if (v < 0) {
        *p++ = '-';
        return fmt_uint(p, (uint64_t)(-(v + 1)) + 1);
    }
    return fmt_uint(p, (uint64_t)v);
}

template <int D>
inline char* emit(char* p, bool neg, uint64_t q, int width);

// "%W.Df" of v into p (no terminator); returns the end.  0 <= D <= 6.
template <int D>
inline char* fmt_fixed(char* p, double v, int width) {
    static_assert(D >= 0 && D <= 6, "decimals");
    if (!std::isfinite(v) || std::fabs(v) >= 1e12) {
        char buf[400];
        int n = snprintf(buf, sizeof buf, "%*.*f", width, D, v);
        for (int i = 0; i < n; ++i) *p++ = buf[i];
        return p;
    }
    constexpr uint64_t P = fmt_detail::pow10(D);
    uint64_t bits;
    memcpy(&bits, &v, sizeof bits);
    const bool neg = (bits >> 63) != 0;
    const int be = (int)((bits >> 52) & 0x7ff);
    uint64_t m = bits & ((1ull << 52) - 1);
    int e;                                     // |v| = m * 2^e exactly, m < 2^53
    if (be == 0) {
        e = -1074;                             // zero and subnormals
    } else {
        m |= 1ull << 52;
        e = be - 1075;
    }
    const unsigned __int128 x = (unsigned __int128)m * P;   // |v| * 10^D = x * 2^e, x < 2^73
    uint64_t q;
    if (e >= 0) {
        q = (uint64_t)(x << e);                // not reached for |v| < 1e12 (e <= -12)
    } else {
        const int s = -e;
        if (s >= 100) {
            q = 0;                             // x * 2^-s < 2^-26: rounds to zero
        } else {
            const unsigned __int128 qq = x >> s;
            const unsigned __int128 rem = x - (qq << s);
            const unsigned __int128 half = (unsigned __int128)1 << (s - 1);
            q = (uint64_t)qq;
            if (rem > half || (rem == half && (q & 1))) ++q;
        }
    }
    return emit<D>(p, neg, q, width);
}

// The digits of q / 10^D with D decimals, sign and left padding to width,
// written right to left in place.
template <int D>
inline char* emit(char* p, bool neg, uint64_t q, int width) {
    constexpr uint64_t P = fmt_detail::pow10(D);
    const uint64_t ip = q / P;
    const int nd = n_digits(ip);
    const int len = (int)neg + nd + (D > 0 ? D + 1 : 0);
    for (int k = len; k < width; ++k) *p++ = ' ';
    char* e = p + len;
    if constexpr (D > 0) {
        uint64_t fp = q % P;
        char* f = e;
        int k = D;
        while (k >= 2) {
            k -= 2;
            f -= 2;
            memcpy(f, fmt_detail::kPairs.d + 2 * (fp % 100), 2);
            fp /= 100;
        }
        if (k == 1) *--f = (char)('0' + fp);
        f[-1] = '.';
    }
    put_digits(p + (int)neg + nd, ip);
    if (neg) *p = '-';
    return e;
}

// The same with the decimal count given at run time.
inline char* fmt_fixed(char* p, double v, int decimals, int width) {
    switch (decimals) {
        case 0: return fmt_fixed<0>(p, v, width);
        case 1: return fmt_fixed<1>(p, v, width);
        case 2: return fmt_fixed<2>(p, v, width);
        case 3: return fmt_fixed<3>(p, v, width);
        case 4: return fmt_fixed<4>(p, v, width);
        case 5: return fmt_fixed<5>(p, v, width);
        default: return fmt_fixed<6>(p, v, width);
    }
}

}  // namespace mafm

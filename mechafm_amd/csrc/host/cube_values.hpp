// Volumetric section of a Gaussian cube file: the first `total`
// whitespace-separated numbers, parsed by up to 16 threads on chunks cut at
// whitespace.  std::from_chars rounds decimal to binary correctly, as glibc's
// strtod and the reference's `>>` loop do (cube_io.cpp:40-54); text after
// the `total` numbers is ignored, as that loop ignores it.
#pragma once

#include <algorithm>
#include <charconv>
#include <functional>
#include <system_error>
#include <thread>
#include <vector>

// false when fewer than `total` numbers precede the first text that is not a
// number (or the end).  Two passes over each chunk: count its tokens, then,
// at the offsets the counts give, parse them straight into `out`.
inline bool parse_cube_values(const char* p, const char* end, size_t total, std::vector<double>& out,
                              int max_threads = 16) {
    auto space = [](char ch) { return ch == ' ' || ch == '\n' || ch == '\t' || ch == '\r' || ch == '\v' || ch == '\f'; };
    const size_t bytes = (size_t)(end - p);
    int T = (int)std::min<unsigned>((unsigned)std::max(1, max_threads), std::max(1u, std::thread::hardware_concurrency()));
    if (total < (1u << 16)) T = 1;
    std::vector<const char*> cut(T + 1);
    cut[0] = p;
    for (int t = 1; t < T; ++t) {
        const char* e = std::max(cut[t - 1], p + bytes * t / T);
        while (e < end && !space(*e)) ++e;               // between two tokens
        cut[t] = e;
    }
    cut[T] = end;
    std::vector<size_t> count(T + 1, 0), bad(T, (size_t)-1);
    auto run = [&](auto&& f) {
        if (T == 1) {
            f(0);
            return;
        }
        std::vector<std::thread> pool;
        for (int t = 0; t < T; ++t) pool.emplace_back(f, t);
        for (auto& th : pool) th.join();
    };
    run([&](int t) {                                     // tokens per chunk
        size_t n = 0;
        bool in = false;
        for (const char* q = cut[t]; q < cut[t + 1]; ++q) {
            const bool s = space(*q);
            n += !s && !in;
            in = !s;
        }
        count[t + 1] = n;
    });
    for (int t = 0; t < T; ++t) count[t + 1] += count[t];
    if (count[T] < total) return false;
    out.resize(total);
    run([&](int t) {                                     // parse this chunk's tokens into place
        size_t i = count[t];
        const char* q = cut[t];
        while (i < total) {
            while (q < cut[t + 1] && space(*q)) ++q;
            if (q >= cut[t + 1]) return;
            const char* s = *q == '+' ? q + 1 : q;       // strtod accepts a leading '+'
            double v;
            const std::from_chars_result r = std::from_chars(s, cut[t + 1], v);
            if ((r.ec != std::errc() && r.ec != std::errc::result_out_of_range) ||
                (r.ptr < cut[t + 1] && !space(*r.ptr))) {
                bad[t] = i;                              // first token that is not a number
                return;
            }
            out[i++] = v;
            q = r.ptr;
        }
    });
    for (int t = 0; t < T; ++t)
        if (bad[t] < total) return false;
    return true;
}

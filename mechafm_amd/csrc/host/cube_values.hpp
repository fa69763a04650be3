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
// number (or the end)
inline bool parse_cube_values(const char* p, const char* end, size_t total, std::vector<double>& out,
                              int max_threads = 16) {
    auto space = [](char ch) { return ch == ' ' || ch == '\n' || ch == '\t' || ch == '\r' || ch == '\v' || ch == '\f'; };
    struct Chunk {
        const char* b;
        const char* e;
        std::vector<double> v;
        bool stopped = false;     // met text that is not a number before the chunk's end
    };
    const size_t bytes = (size_t)(end - p);
    int T = (int)std::min<unsigned>((unsigned)std::max(1, max_threads), std::max(1u, std::thread::hardware_concurrency()));
    if (total < (1u << 16)) T = 1;
    std::vector<Chunk> ch(T);
    const char* b = p;
    for (int t = 0; t < T; ++t) {
        const char* e = t + 1 == T ? end : p + bytes * (t + 1) / T;
        if (e < b) e = b;
        while (e < end && !space(*e)) ++e;               // cut between numbers
        ch[t].b = b;
        ch[t].e = e;
        b = e;
    }
    auto work = [&](Chunk& k) {
        const char* q = k.b;
        k.v.reserve((size_t)(k.e - k.b) / 8 + 1);
        for (;;) {
            while (q < k.e && space(*q)) ++q;
            if (q >= k.e) return;
            if (*q == '+') ++q;                          // strtod accepts a leading '+'
            double v;
            const std::from_chars_result r = std::from_chars(q, k.e, v);
            if (r.ec != std::errc() && r.ec != std::errc::result_out_of_range) {
                k.stopped = true;
                return;
            }
            k.v.push_back(v);
            q = r.ptr;
        }
    };
    if (T == 1) {
        work(ch[0]);
    } else {
        std::vector<std::thread> pool;
        for (int t = 0; t < T; ++t) pool.emplace_back(work, std::ref(ch[t]));
        for (auto& th : pool) th.join();
    }
    out.resize(total);
    size_t n = 0;
    for (const Chunk& k : ch) {
        const size_t take = std::min(k.v.size(), total - n);
        std::copy(k.v.begin(), k.v.begin() + take, out.begin() + n);
        n += take;
        if (n == total) return true;
        if (k.stopped) break;
    }
    return total == 0;
}

// Checks parse_cube_values (mechafm_amd/csrc/host/cube_values.hpp) against a
// sequential strtod parse: bit-identical values for every number format a
// cube writer emits, with 1 and 16 threads; trailing text ignored; short or
// broken data rejected.  Prints "ok" and exits 0 on success.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../mechafm_amd/csrc/host/cube_values.hpp"

static int fails = 0;
#define CHECK(c, msg)                                     \
    do {                                                  \
        if (!(c)) {                                       \
            std::fprintf(stderr, "FAIL: %s\n", msg);      \
            ++fails;                                      \
        }                                                 \
    } while (0)

static std::vector<double> strtod_all(const std::string& s, size_t n) {
    std::vector<double> v;
    const char* p = s.c_str();
    for (size_t i = 0; i < n; ++i) {
        char* q;
        v.push_back(std::strtod(p, &q));
        p = q;
    }
    return v;
}

static bool same(const std::vector<double>& a, const std::vector<double>& b) {
    return a.size() == b.size() && std::memcmp(a.data(), b.data(), a.size() * sizeof(double)) == 0;
}

int main() {
    std::mt19937_64 rng(7);
    std::uniform_real_distribution<double> mant(-1.0, 1.0);
    std::uniform_int_distribution<int> ex(-30, 30), fmt(0, 5), sep(0, 5);
    const size_t N = 200000;
    std::string s;
    char tmp[80];
    for (size_t i = 0; i < N; ++i) {
        const double x = mant(rng) * std::pow(10.0, ex(rng));
        switch (fmt(rng)) {
            case 0: std::snprintf(tmp, sizeof tmp, "%13.5E", x); break;
            case 1: std::snprintf(tmp, sizeof tmp, "%.17g", x); break;
            case 2: std::snprintf(tmp, sizeof tmp, "%+.6e", x); break;
            case 3: std::snprintf(tmp, sizeof tmp, "%.3f", x); break;
            case 4: std::snprintf(tmp, sizeof tmp, "%d", (int)(x * 1000) % 1000); break;
            default: std::snprintf(tmp, sizeof tmp, "%.9e", x); break;
        }
        s += tmp;
        static const char* seps[] = {" ", "  ", "\n", "\t", "\r\n", " \n "};
        s += seps[sep(rng)];
    }
    const std::vector<double> want = strtod_all(s, N);
    for (int threads : {1, 16}) {
        std::vector<double> got;
        CHECK(parse_cube_values(s.data(), s.data() + s.size(), N, got, threads), "parse");
        CHECK(same(got, want), threads == 1 ? "values, 1 thread" : "values, 16 threads");
        // a shorter read takes the first numbers only
        CHECK(parse_cube_values(s.data(), s.data() + s.size(), N / 3, got, threads) && got.size() == N / 3 &&
                  std::memcmp(got.data(), want.data(), got.size() * sizeof(double)) == 0,
              "prefix");
        // too few numbers
        CHECK(!parse_cube_values(s.data(), s.data() + s.size(), N + 1, got, threads), "short data rejected");
        // text after the numbers is ignored
        const std::string tail = s + "\n end of data 1.0 x\n";
        CHECK(parse_cube_values(tail.data(), tail.data() + tail.size(), N, got, threads) && same(got, want),
              "trailing text");
        // text before the last number is an error
        std::string broken = s;
        broken.replace(broken.size() / 2, 0, " abc ");
        CHECK(!parse_cube_values(broken.data(), broken.data() + broken.size(), N, got, threads), "broken data");
    }
    std::vector<double> none;
    CHECK(parse_cube_values(s.data(), s.data(), 0, none, 16) && none.empty(), "empty");
    if (fails) return 1;
    std::printf("ok\n");
    return 0;
}

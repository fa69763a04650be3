// Compares mafm::fmt_fixed / fmt_int (the output writer's formatter) with
// glibc's snprintf "%W.Df" / "%d" on special values, exact decimal ties and
// random doubles; prints the mismatch count and exits non-zero on any.
#include <cstring>
#include <random>

#include "../mechafm_amd/csrc/host/format.hpp"

static long n_checked = 0, n_bad = 0;

static void check(double v, int D, int W) {
    char a[512], b[512];
    char* e = mafm::fmt_fixed(a, v, D, W);
    *e = 0;
    snprintf(b, sizeof b, "%*.*f", W, D, v);
    ++n_checked;
    if (strcmp(a, b) != 0) {
        if (n_bad < 10) printf("mismatch %.17g %%%d.%df: '%s' vs '%s'\n", v, W, D, a, b);
        ++n_bad;
    }
}

static void check_int(long long v) {
    char a[64], b[64];
    char* e = mafm::fmt_int(a, v);
    *e = 0;
    snprintf(b, sizeof b, "%lld", v);
    ++n_checked;
    if (strcmp(a, b) != 0) ++n_bad;
}

int main(int argc, char** argv) {
    const long n = argc > 1 ? atol(argv[1]) : 1000000;
    std::mt19937_64 g(12345);
    const double special[] = {0.0, -0.0, 0.5, 1.5, 2.5, -0.5, -2.5, 0.125, 0.0625, 0.0005, -0.0005, 0.00015,
                              1e-300, -1e-300, 5e-324, 999.9995, 999.99949999, 123456.78905, 1e11, -1e11, 1e12,
                              1e15, -1e300, NAN, -NAN, INFINITY, -INFINITY, 0.1, 0.2, 0.3, 9.9995, -9.99949};
    for (double v : special)
        for (int D = 0; D <= 6; ++D) check(v, D, 8);
    std::uniform_real_distribution<double> u(-1, 1);
    for (long i = 0; i < n; ++i) {
        const double v = u(g) * std::pow(10.0, (int)(g() % 14) - 7);
        check(v, 3, 6);
        check(v, 4, 8);
    }
    for (long i = 0; i < n; ++i) {       // exact binary values on and near decimal ties
        const double v = (double)((int64_t)(g() % 4000000) - 2000000) / 4096.0;
        check(v, 3, 6);
        check(v, 4, 8);
        check(v, 1, 4);
    }
    for (long i = 0; i < n; ++i) {       // random bit patterns
        uint64_t bits = g();
        double v;
        memcpy(&v, &bits, 8);
        check(v, 3, 6);
        check(v, 4, 8);
    }
    const long long ints[] = {0, 1, -1, 9, 10, -10, 2147483647LL, -2147483648LL, 9223372036854775807LL,
                              -9223372036854775807LL - 1};
    for (long long v : ints) check_int(v);
    for (long i = 0; i < n / 10; ++i) check_int((long long)g());
    printf("checked %ld, mismatches %ld\n", n_checked, n_bad);
    return n_bad != 0;
}

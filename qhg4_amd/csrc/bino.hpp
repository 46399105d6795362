// Binomial mutation-count table (host): BinomialDist::buildTable with the
// Numerical-Recipes regularised incomplete beta function
// (QHG4/utils/BinomialDist.cpp:60-104, QHG4/utils/bino_tools.cpp).
// Entry k-1 of the table is P(X <= k-1) = 1 - I_p(k, n-k+1) for X ~ Bin(n, p);
// the table stops at the first k whose tail I_p(k, n-k+1) <= eps.
#pragma once
#include <cmath>
#include <cstdint>

namespace qhg {

inline double nr_gammaln(double xx) {
    static const double cof[6] = {76.18009172947146, -86.50532032941677, 24.01409824083091,
                                  -1.231739572450155, 0.1208650973866179e-2, -0.5395239384953e-5};
    double x = xx, y = xx + 1, tmp = x + 5.5;
    tmp -= (x + 0.5) * std::log(tmp);
    double ser = 1.000000000190015;
    for (int j = 0; j <= 5; j++) ser += cof[j] / y++;
    return -tmp + std::log(2.5066282746310005 * ser / x);
}

// continued fraction of the incomplete beta (modified Lentz)
inline double nr_betacf(double a, double b, double x) {
    const int kMaxIt = 100;
    const double kEps = 3.0e-7;
    const float kTiny = 1.0e-30f;
    const double qab = a + b, qap = a + 1.0, qam = a - 1.0;
    double c = 1.0, d = 1.0 - qab * x / qap;
    if (std::fabs(d) < kTiny) d = kTiny;
    d = 1.0 / d;
    double h = d;
    int m = 1;
    for (; m <= kMaxIt; m++) {
        const int m2 = 2 * m;
        double aa = m * (b - m) * x / ((qam + m2) * (a + m2));
        d = 1.0 + aa * d;
        if (std::fabs(d) < kTiny) d = kTiny;
        c = 1.0 + aa / c;
        if (std::fabs(c) < kTiny) c = kTiny;
        d = 1.0 / d;
        h *= d * c;
        aa = -(a + m) * (qab + m) * x / ((a + m2) * (qap + m2));
        d = 1.0 + aa * d;
        if (std::fabs(d) < kTiny) d = kTiny;
        c = 1.0 + aa / c;
        if (std::fabs(c) < kTiny) c = kTiny;
        d = 1.0 / d;
        const double del = d * c;
        h *= del;
        if (std::fabs(del - 1.0) < kEps) break;
    }
    return (m > kMaxIt) ? -1.0 : h;
}

inline double nr_betai(double a, double b, double x) {
    if (x < 0.0 || x > 1.0) return -1.0;
    const double bt = (x == 0.0 || x == 1.0)
                          ? 0.0
                          : std::exp(nr_gammaln(a + b) - nr_gammaln(a) - nr_gammaln(b) + a * std::log(x) +
                                     b * std::log(1.0 - x));
    if (x < (a + 1.0) / (a + b + 2.0)) return bt * nr_betacf(a, b, x) / a;
    return 1.0 - bt * nr_betacf(b, a, 1.0 - x) / b;
}

// returns the number of entries, or -1
inline int32_t bino_table(double p, int32_t n, double eps, double *cdf, int32_t max) {
    if (!(p >= 0 && p < 1 && n > 0)) return -1;
    int32_t cnt = 0;
    int k = 1;
    double tail = nr_betai(k, n - k + 1, p);
    while (tail > eps && k < n) {
        if (cnt >= max) return -1;
        cdf[cnt++] = 1 - tail;
        k++;
        tail = nr_betai(k, n - k + 1, p);
    }
    if (cnt >= max) return -1;
    cdf[cnt++] = 1 - tail;
    return cnt;
}

}  // namespace qhg

// Host build of libtp's FP64 special functions (csrc/tp_host_math.hpp) for the
// CPU test suite.  Reads commands on stdin, prints results with 17 digits.
//   w l1 m1 l2 m2 l3 m3   -> wigner3j
//   p l m theta           -> sphPlm
//   s l m LMAX            -> seanindex
//   g n                   -> n Gauss-Legendre nodes and weights
//   j n x                 -> J_{l+1/2}(x), l = 0..n-1 (bessel_half_all)
//   l n                   -> worst error (ulp of the exact value, long double
//                            reference) of tp_fastlog over n arguments
//   t n                   -> the same for tp_log_series
//   e n                   -> worst error of the energy epilogues' tp_fastlog_e
//                            in ulp of max(|log x|, 1) over n arguments in
//                            1e-30..1e30 and around 1, and its count of
//                            results that differ from tp_fastlog away from 1
#include <cmath>
#include <cstdio>
#include <random>
#include <vector>
#include "tp_fastlog.hpp"
#include "tp_host_math.hpp"

static double ulp_err(double x, const double *tab)
{
    const long double ref = logl((long double)x);
    const double got = tab ? tp_fastlog(x, tab) : tp_log_series(x);
    const double rd = (double)ref;
    const double u = std::nextafter(std::fabs(rd), INFINITY) - std::fabs(rd);
    return (double)(fabsl((long double)got - ref) / (long double)u);
}

int main()
{
    char c;
    while (std::scanf(" %c", &c) == 1) {
        if (c == 'w') {
            int a[6];
            if (std::scanf("%d %d %d %d %d %d", a, a + 1, a + 2, a + 3, a + 4, a + 5) != 6) return 1;
            std::printf("%.17g\n", tph::wigner3j(a[0], a[1], a[2], a[3], a[4], a[5]));
        } else if (c == 'p') {
            int l, m;
            double th;
            if (std::scanf("%d %d %lf", &l, &m, &th) != 3) return 1;
            std::printf("%.17g\n", tph::sphPlm(l, m, th));
        } else if (c == 's') {
            int l, m, L;
            if (std::scanf("%d %d %d", &l, &m, &L) != 3) return 1;
            std::printf("%lld\n", (long long)tph::seanindex(l, m, L));
        } else if (c == 'j') {
            int n;
            double x;
            if (std::scanf("%d %lf", &n, &x) != 2 || n < 1) return 1;
            std::vector<double> J((size_t)n);
            tph::bessel_half_all(x, n, J.data());
            for (int l = 0; l < n; ++l) std::printf("%.17g%c", J[(size_t)l], l + 1 < n ? ' ' : '\n');
        } else if (c == 'l' || c == 't') {
            int n;
            if (std::scanf("%d", &n) != 1) return 1;
            double tabv[kLogTabSize];
            tp_fastlog_table(tabv);
            const double *tab = c == 'l' ? tabv : nullptr;
            std::mt19937_64 rng(12345);
            std::uniform_real_distribution<double> U(0.0, 1.0);
            double worst = 0.0;
            for (int i = 0; i < n; ++i) {
                double x;
                switch (i % 4) {
                case 0: x = std::exp(-69.0 + 138.0 * U(rng)); break;             // 1e-30 .. 1e30
                case 1: x = 1.0 + (U(rng) - 0.5) * 0.03125; break;               // around 1
                case 2: x = std::ldexp(1.0 + U(rng), (int)(U(rng) * 80) - 40); break;
                default:
                    if (tab) x = std::ldexp(1.0 + (std::floor(U(rng) * 128) + (U(rng) < 0.5 ? 0.0 : 1.0)) / 128.0, 3) *
                                 (1.0 + (U(rng) - 0.5) * 1e-15);                // table-interval edges
                    else x = std::ldexp(1.4142135623730951 * (1.0 + (U(rng) - 0.5) * 1e-14),
                                        (int)(U(rng) * 20) - 10);               // the sqrt2 fold
                    break;
                }
                worst = std::fmax(worst, ulp_err(x, tab));
            }
            std::printf("%.6g\n", worst);
        } else if (c == 'e') {
            int n;
            if (std::scanf("%d", &n) != 1) return 1;
            double tab[kLogTabSize];
            tp_fastlog_table(tab);
            const LogConsts k = tp_fastlog_consts(tab);
            std::mt19937_64 rng(777);
            std::uniform_real_distribution<double> U(0.0, 1.0);
            double worst = 0.0;
            long differ = 0;
            for (int i = 0; i < n; ++i) {
                const double x = i % 2 ? std::exp(-69.0 + 138.0 * U(rng)) : 1.0 + (U(rng) - 0.5) * 0.03;
                const double got = tp_fastlog_e(x, tab, k);
                const long double ref = logl((long double)x);
                const double big = std::fmax(std::fabs((double)ref), 1.0);
                const double u = std::nextafter(big, INFINITY) - big;   // ulp of max(|log x|, 1)
                worst = std::fmax(worst, (double)(fabsl((long double)got - ref) / (long double)u));
                if (std::fabs(x - 1.0) >= 0.0078125 && got != tp_fastlog(x, tab)) ++differ;
            }
            std::printf("%.6g %ld\n", worst, differ);
        } else if (c == 'g') {
            int n;
            if (std::scanf("%d", &n) != 1) return 1;
            std::vector<double> x, w;
            tph::gauss_legendre(n, x, w);
            for (int i = 0; i < n; ++i) std::printf("%.17g %.17g\n", x[i], w[i]);
        }
    }
    return 0;
}

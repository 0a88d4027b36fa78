// tp_host_math.hpp -- FP64 host-side special functions for the basis
// precompute of libtp: Wigner 3j, normalised associated Legendre, the
// seanindex coefficient layout and Gauss-Legendre nodes.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <vector>

namespace tph {

constexpr double kPi = 3.14159265358979323846;

// Position (1-based) of (l,m) in a length-LMAX^2 coefficient vector: all
// m >= 0 entries ordered by (m, l), then the m < 0 entries ordered by
// (m from -(LMAX-1) upwards, l).  Same layout as spherical_harmonics.jl:163-170.
inline int64_t seanindex(int l, int m, int LMAX)
{
    const int64_t B = LMAX;   // bandwidth
    if (m >= 0) return 1 + (int64_t)m * B - (int64_t)m * (m - 1) / 2 + (l - m);
    const int64_t pos_block = B * (B + 1) / 2;              // number of m >= 0 entries
    const int64_t a = B - 1 + m;                            // m = -(B-1) -> 0
    return 1 + pos_block + a * (a + 1) / 2 + (l + m);
}

// Exact binomial coefficient for n <= 62.
inline unsigned __int128 binom_u(int n, int k)
{
    if (k < 0 || k > n) return 0;
    if (2 * k > n) k = n - k;
    unsigned __int128 r = 1;
    for (int i = 0; i < k; ++i) r = r * (unsigned)(n - i) / (unsigned)(i + 1);
    return r;
}

// Wigner 3j symbol (j1 j2 j3; m1 m2 m3) for integer arguments.  The Racah
// sum is rewritten with binomials and evaluated exactly in 128-bit integers:
//   3j = (-1)^(j1-j2-m3) * sqrt(P) * sum_k (-1)^k C(a,k) C(b,j1-m1-k) C(c,j2+m2-k)
// with a=j1+j2-j3, b=j1-j2+j3, c=-j1+j2+j3 and
//   P = prod_i (j_i+m_i)!(j_i-m_i)! / ((j1+j2+j3+1)! a! b! c!).
// P is formed as a product of ratios in long double.  Valid for j <= 22.
inline double wigner3j(int j1, int m1, int j2, int m2, int j3, int m3)
{
    if (m1 + m2 + m3 != 0) return 0.0;
    if (std::abs(m1) > j1 || std::abs(m2) > j2 || std::abs(m3) > j3) return 0.0;
    if (j3 > j1 + j2 || j3 < std::abs(j1 - j2)) return 0.0;
    const int a = j1 + j2 - j3, b = j1 - j2 + j3, c = -j1 + j2 + j3;
    __int128 pos = 0, neg = 0;
    const int kmin = std::max(0, std::max(j1 - m1 - b, j2 + m2 - c));
    const int kmax = std::min(a, std::min(j1 - m1, j2 + m2));
    for (int k = kmin; k <= kmax; ++k) {
        __int128 t = (__int128)(binom_u(a, k) * binom_u(b, j1 - m1 - k) * binom_u(c, j2 + m2 - k));
        if (k & 1) neg += t; else pos += t;
    }
    const __int128 S = pos - neg;
    if (S == 0) return 0.0;
    // P = [ (2j1)!(2j2)!(2j3)! / ((J+1)! a! b! c!) ] / [ C(2j1,j1+m1) C(2j2,j2+m2) C(2j3,j3+m3) ]
    // and (2j1)!(2j2)!(2j3)!/((J+1)! a! b! c!) is evaluated as a running product.
    long double lp = 0.0L;
    {
        // log of factorial ratio computed by summing logs of integers (exactly
        // representable), long double accumulation
        auto lfact = [](int n) { long double s = 0.0L; for (int i = 2; i <= n; ++i) s += logl((long double)i); return s; };
        lp = lfact(2 * j1) + lfact(2 * j2) + lfact(2 * j3) - lfact(j1 + j2 + j3 + 1) - lfact(a) - lfact(b) - lfact(c);
        lp -= logl((long double)binom_u(2 * j1, j1 + m1)) + logl((long double)binom_u(2 * j2, j2 + m2)) +
              logl((long double)binom_u(2 * j3, j3 + m3));
    }
    long double v = (long double)S * expl(0.5L * lp);
    if ((j1 - j2 - m3) & 1) v = -v;
    return (double)v;
}

// Normalised associated Legendre function with Condon-Shortley phase,
// sqrt((2l+1)/(4pi) (l-m)!/(l+m)!) P_l^m(cos theta); for m < 0 the value of
// |m| times (-1)^|m| (spherical_harmonics.jl:69-76).  cos(theta) is taken in
// double as the reference does before calling GSL.
inline double sphPlm(int l, int m, double theta)
{
    const int am = m < 0 ? -m : m;
    if (am > l) return 0.0;
    const long double x = (long double)std::cos(theta);
    const long double sn = sqrtl((1.0L - x) * (1.0L + x));
    // diagonal start: Y_m^m = (-1)^m sqrt((2m+1)!! / (4 pi (2m)!!)) sin^m
    long double ymm = 1.0L / sqrtl(4.0L * (long double)kPi);
    for (int k = 1; k <= am; ++k) ymm *= -sn * sqrtl((long double)(2 * k + 1) / (long double)(2 * k));
    long double prev = 0.0L, cur = ymm;
    for (int ll = am + 1; ll <= l; ++ll) {
        // Y_ll = x a_ll Y_{ll-1} - (a_ll / a_{ll-1}) Y_{ll-2}, a_ll = sqrt((4ll^2-1)/(ll^2-m^2))
        const long double all = sqrtl((4.0L * ll * ll - 1.0L) / ((long double)ll * ll - (long double)am * am));
        long double nxt;
        if (ll == am + 1) {
            nxt = x * all * cur;
        } else {
            const int lp = ll - 1;
            const long double alp = sqrtl((4.0L * lp * lp - 1.0L) / ((long double)lp * lp - (long double)am * am));
            nxt = all * (x * cur - prev / alp);
        }
        prev = cur;
        cur = nxt;
    }
    double r = (double)cur;
    if (m < 0 && (am & 1)) r = -r;
    return r;
}

// n-point Gauss-Legendre nodes x (descending) and weights on [-1, 1].
inline void gauss_legendre(int n, std::vector<double> &x, std::vector<double> &w)
{
    x.assign(n, 0.0);
    w.assign(n, 0.0);
    for (int i = 0; i < n; ++i) {
        long double z = cosl((long double)kPi * (i + 0.75L) / (n + 0.5L)), dp = 1.0L;
        for (int it = 0; it < 100; ++it) {
            long double p0 = 1.0L, p1 = z;
            for (int k = 2; k <= n; ++k) {
                long double p2 = ((2.0L * k - 1.0L) * z * p1 - (k - 1.0L) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            if (n == 1) { p1 = z; p0 = 1.0L; }
            dp = n * (z * p1 - p0) / (z * z - 1.0L);
            long double dz = p1 / dp;
            z -= dz;
            if (fabsl(dz) < 1e-20L) break;
        }
        x[i] = (double)z;
        w[i] = (double)(2.0L / ((1.0L - z * z) * dp * dp));
    }
}

// Bessel functions of half-integer order, J[l] = J_{l+1/2}(x) for l = 0..n-1,
// x > 0: the spherical j_l by Miller's downward recurrence
// j_{l-1} = (2l+1)/x j_l - j_{l+1} (stable downwards for every order), started
// far above max(n, x) and normalised by the closed form of j_0 = sin x / x or
// j_1 = sin x / x^2 - cos x / x, whichever is larger; J_{l+1/2} = sqrt(2x/pi) j_l.
// Long double throughout: the unnormalised values stay far inside its range.
inline void bessel_half_all(double x, int n, double *J)
{
    const double mx = n > x ? (double)n : x;
    const int top = n + 40 + (int)(mx + std::sqrt(40.0 * mx));
    const int ns = n > 2 ? n : 2;
    std::vector<long double> j((size_t)ns, 0.0L);
    const long double xl = x;
    long double jp = 0.0L, jc = 1.0L;   // j_{l+1}, j_l (unnormalised), l = top
    for (int l = top; l >= 1; --l) {
        const long double jm = (2.0L * l + 1.0L) / xl * jc - jp;
        jp = jc;
        jc = jm;   // j_{l-1}
        if (l - 1 < ns) j[(size_t)(l - 1)] = jc;
    }
    const long double sx = sinl(xl), cx = cosl(xl);
    const long double j0 = sx / xl, j1 = sx / (xl * xl) - cx / xl;
    const long double sc = fabsl(j0) >= fabsl(j1) ? j0 / j[0] : j1 / j[1];
    const long double f = sqrtl(2.0L * xl / (long double)kPi);
    for (int l = 0; l < n; ++l) J[l] = (double)(f * sc * j[(size_t)l]);
}

}  // namespace tph

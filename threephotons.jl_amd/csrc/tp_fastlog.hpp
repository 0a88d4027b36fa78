// tp_fastlog.hpp -- table-driven FP64 natural logarithm for the energy
// epilogues (the Bayes sum needs one log per histogram bin per evaluation).
//
// Method: P. T. P. Tang, "Table-driven implementation of the logarithm
// function in IEEE floating-point arithmetic", ACM TOMS 16(4), 1990.
// For normal positive x = 2^e m, m in [1, 2): j = the top 7 bits of m's
// fraction, c_j = 1 + (j + 1/2)/128, inv_j = RN(1/c_j), t_j = -log(inv_j),
// and for c_j >= sqrt2 (j >= 53) the reduction is to m/2 in [1/sqrt2, 1):
// t_j -= ln2, a_j = 1 (so e ln2 and t_j never cancel by more than a factor
// 2); tables built on the host in long double by tp_fastlog_table, (inv_j,
// t_j) adjacent (one 16-byte LDS read per logarithm), with ln 2 split as
// hi + lo, hi having 21 significant bits so that (e + a_j) hi is exact;
// r = m inv_j - 1 by one FMA (|r| < 2^-8, rounding error < 2^-60), and
//   log x = (e + a_j) ln2 + t_j + log1p(r),
// log1p by its Taylor series (to r^8 here, truncation < 2^-66).  Near 1
// (|x - 1| < 2^-7, where e ln2 + t_j would cancel) log1p(x - 1) is summed
// directly to r^10.  Error: a few ulp (checked against libm in
// tests/test_host_math.py).  Non-positive, infinite, NaN and subnormal
// arguments go to the library log.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>

#ifdef __HIPCC__
#define TP_HD __host__ __device__
#else
#define TP_HD
#endif

constexpr int kLogTab = 128;
constexpr int kLogUpJ = 53;                    // c_j >= sqrt2 exactly from j = 53: 128 (sqrt2 - 1) - 1/2 = 52.52
constexpr int kLogPoly = 7;                    // log1p coefficients -1/2, 1/3, .., -1/8
constexpr int kLogTabSize = 2 * kLogTab + 2 + kLogPoly;   // (inv_j, t_j) pairs, ln2 hi, ln2 lo, coefficients

inline void tp_fastlog_table(double *tab)
{
    const long double ln2 = logl(2.0L);
    for (int j = 0; j < kLogTab; ++j) {
        const long double c = 1.0L + (j + 0.5L) / kLogTab;
        const double inv = (double)(1.0L / c);
        const bool up = j >= kLogUpJ;
        tab[2 * j] = inv;
        tab[2 * j + 1] = (double)(-logl((long double)inv) - (up ? ln2 : 0.0L));
    }
    double hi = (double)ln2;
    uint64_t b;
    std::memcpy(&b, &hi, 8);
    b &= ~0xFFFFFFFFull;   // keep 21 significant bits
    std::memcpy(&hi, &b, 8);
    tab[2 * kLogTab] = hi;
    tab[2 * kLogTab + 1] = (double)(ln2 - (long double)hi);
    for (int i = 0; i < kLogPoly; ++i) tab[2 * kLogTab + 2 + i] = (i % 2 ? 1.0 : -1.0) / (i + 2);
}

// ln2hi, ln2lo: tab[2 kLogTab], tab[2 kLogTab + 1], read once by the caller.
TP_HD inline double tp_fastlog(double x, const double *tab, double ln2hi, double ln2lo)
{
    uint64_t bits;
    std::memcpy(&bits, &x, 8);
    const int ex = (int)((bits >> 52) & 0x7FF);
    if (ex == 0 || ex == 0x7FF || (bits >> 63)) return std::log(x);   // 0, subnormal, inf/NaN, negative
    const double d = x - 1.0;                                        // exact for x in [1/2, 2]
    if (std::fabs(d) < 0.0078125) {
        double p = -0.1;
        p = std::fma(p, d, 1.0 / 9.0);
        p = std::fma(p, d, -0.125);
        p = std::fma(p, d, 1.0 / 7.0);
        p = std::fma(p, d, -1.0 / 6.0);
        p = std::fma(p, d, 0.2);
        p = std::fma(p, d, -0.25);
        p = std::fma(p, d, 1.0 / 3.0);
        p = std::fma(p, d, -0.5);
        return std::fma(p * d, d, d);
    }
    const int e = ex - 1023;
    const uint64_t mb = (bits & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull;
    double m;
    std::memcpy(&m, &mb, 8);
    const int j = (int)((bits >> 45) & 0x7F);
    const double inv = tab[2 * j], tj = tab[2 * j + 1];
    const double r = std::fma(m, inv, -1.0);
    double p = -0.125;
    p = std::fma(p, r, 1.0 / 7.0);
    p = std::fma(p, r, -1.0 / 6.0);
    p = std::fma(p, r, 0.2);
    p = std::fma(p, r, -0.25);
    p = std::fma(p, r, 1.0 / 3.0);
    p = std::fma(p, r, -0.5);
    p = std::fma(p * r, r, r);
    const double ee = (double)e + (j >= kLogUpJ ? 1.0 : 0.0);
    return std::fma(ee, ln2hi, tj) + std::fma(ee, ln2lo, p);
}

TP_HD inline double tp_fastlog(double x, const double *tab)
{
    return tp_fastlog(x, tab, tab[2 * kLogTab], tab[2 * kLogTab + 1]);
}

// Constants of tp_fastlog_e, read once by a kernel from the table in global
// memory (wave-uniform loads: they stay in scalar registers instead of being
// rebuilt in vector registers at every call).
struct LogConsts {
    double ln2hi, ln2lo, c[kLogPoly];
};

TP_HD inline LogConsts tp_fastlog_consts(const double *tab)
{
    LogConsts k;
    k.ln2hi = tab[2 * kLogTab];
    k.ln2lo = tab[2 * kLogTab + 1];
    for (int i = 0; i < kLogPoly; ++i) k.c[i] = tab[2 * kLogTab + 2 + i];
    return k;
}

// The energy epilogues' logarithm: tp_fastlog's main path without branches,
// for finite x >= 1e-30 (the energy clamps its argument there; +inf returns
// +inf).  It equals tp_fastlog bit for bit except for |x - 1| < 2^-7, where it
// keeps the table reduction instead of summing log1p(x - 1): there the
// absolute error stays within a few 2^-53 (what the energy sum sees), the
// relative error is not bounded.
TP_HD inline double tp_fastlog_e(double x, const double *tab, const LogConsts &k)
{
    uint64_t bits;
    std::memcpy(&bits, &x, 8);
    const int ex = (int)((bits >> 52) & 0x7FF);
    const int e = ex - 1023;
    const uint64_t mb = (bits & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull;
    double m;
    std::memcpy(&m, &mb, 8);
    const int j = (int)((bits >> 45) & 0x7F);
    const double inv = tab[2 * j], tj = tab[2 * j + 1];
    const double r = std::fma(m, inv, -1.0);
    double p = k.c[6];
    p = std::fma(p, r, k.c[5]);
    p = std::fma(p, r, k.c[4]);
    p = std::fma(p, r, k.c[3]);
    p = std::fma(p, r, k.c[2]);
    p = std::fma(p, r, k.c[1]);
    p = std::fma(p, r, k.c[0]);
    p = std::fma(p * r, r, r);
    const double ee = (double)e + (j >= kLogUpJ ? 1.0 : 0.0);
    const double y = std::fma(ee, k.ln2hi, tj) + std::fma(ee, k.ln2lo, p);
    return ex == 0x7FF ? x : y;
}

// Table-free FP64 natural logarithm for epilogues where a table in LDS would
// cost occupancy.  For normal positive x = 2^e m, m in [1, 2), m > sqrt2 is
// halved (e + 1) so that m is in [1/sqrt2, sqrt2); then with f = m - 1
// (exact) and s = f / (2 + f), |s| <= 0.1716,
//   log m = 2 atanh s = 2s + s z (2/3 + 2z/5 + 2z^2/7 + ... + 2z^10/23),  z = s^2,
// the atanh series to z^10 (truncation < z^11 / 23 < 1e-18 of the result),
// and log x = e ln2_hi + (log m + e ln2_lo) with ln2_hi of 32 significant
// bits (e ln2_hi exact for |e| < 2^21).  About 35 instructions, a few ulp
// (checked against long double in tests/test_host_math.py).  Non-positive,
// infinite, NaN and subnormal arguments go to the library log.
TP_HD inline double tp_log_series(double x)
{
    uint64_t bits;
    std::memcpy(&bits, &x, 8);
    const int ex = (int)((bits >> 52) & 0x7FF);
    if (ex == 0 || ex == 0x7FF || (bits >> 63)) return std::log(x);
    int e = ex - 1023;
    uint64_t mb = (bits & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull;   // m in [1, 2)
    if (mb > 0x3FF6A09E667F3BCCull) {   // m > sqrt2 (the double just below sqrt2 is 0x3FF6A09E667F3BCC)
        mb -= 0x0010000000000000ull;    // m / 2
        ++e;
    }
    double m;
    std::memcpy(&m, &mb, 8);
    const double f = m - 1.0;
    const double sv = f / (2.0 + f);
    const double z = sv * sv;
    double q = 2.0 / 23.0;
    q = std::fma(q, z, 2.0 / 21.0);
    q = std::fma(q, z, 2.0 / 19.0);
    q = std::fma(q, z, 2.0 / 17.0);
    q = std::fma(q, z, 2.0 / 15.0);
    q = std::fma(q, z, 2.0 / 13.0);
    q = std::fma(q, z, 2.0 / 11.0);
    q = std::fma(q, z, 2.0 / 9.0);
    q = std::fma(q, z, 2.0 / 7.0);
    q = std::fma(q, z, 2.0 / 5.0);
    q = std::fma(q, z, 2.0 / 3.0);
    const double lm = std::fma(sv * z, q, 2.0 * sv);
    constexpr double ln2_hi = 0.6931471803691238;       // ln 2 truncated to 32 significant bits
    constexpr double ln2_lo = 1.9082149292705877e-10;   // ln 2 - ln2_hi, rounded
    const double de = (double)e;
    return std::fma(de, ln2_hi, std::fma(de, ln2_lo, lm));
}

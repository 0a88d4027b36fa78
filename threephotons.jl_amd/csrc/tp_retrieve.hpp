// tp_retrieve.hpp -- retrieveSolution (correlations.jl:262-319) on the host:
// a start structure from a two-photon histogram.  Small dense least squares,
// a symmetric eigensolver (K2 x K2) and the negativity measure of the result;
// everything here is O(K2^3 + K2 LMAX^3) and runs once per reconstruction.
#pragma once

#include <algorithm>
#include <cmath>
#include <complex>
#include <string>
#include <vector>

#include "tp_c2.hpp"
#include "tp_host_math.hpp"

namespace tpc2 {

// Least squares min |A f - s| by Householder QR (A: rows x cols, column-major).
inline void lstsq(std::vector<double> A, int rows, int cols, std::vector<double> s, std::vector<double> &f)
{
    for (int j = 0; j < cols; ++j) {
        double nrm = 0.0;
        for (int i = j; i < rows; ++i) nrm += A[i + rows * j] * A[i + rows * j];
        nrm = std::sqrt(nrm);
        if (nrm == 0.0) continue;
        const double alpha = A[j + rows * j] > 0 ? -nrm : nrm;
        std::vector<double> v(rows, 0.0);
        for (int i = j; i < rows; ++i) v[i] = A[i + rows * j];
        v[j] -= alpha;
        double vn = 0.0;
        for (int i = j; i < rows; ++i) vn += v[i] * v[i];
        if (vn == 0.0) continue;
        for (int c = j; c < cols; ++c) {
            double d = 0.0;
            for (int i = j; i < rows; ++i) d += v[i] * A[i + rows * c];
            d = 2.0 * d / vn;
            for (int i = j; i < rows; ++i) A[i + rows * c] -= d * v[i];
        }
        double d = 0.0;
        for (int i = j; i < rows; ++i) d += v[i] * s[i];
        d = 2.0 * d / vn;
        for (int i = j; i < rows; ++i) s[i] -= d * v[i];
    }
    f.assign(cols, 0.0);
    for (int j = cols - 1; j >= 0; --j) {
        double t = s[j];
        for (int c = j + 1; c < cols; ++c) t -= A[j + rows * c] * f[c];
        f[j] = A[j + rows * j] != 0.0 ? t / A[j + rows * j] : 0.0;
    }
}

// Eigenvalues (ascending) and eigenvectors (columns of V, n x n column-major)
// of the symmetric matrix S by cyclic Jacobi rotations.  Each eigenvector's
// sign is fixed so that its entry of largest magnitude (first such) is
// positive (LAPACK, used by the reference, leaves the sign unspecified).
inline void sym_eig(std::vector<double> S, int n, std::vector<double> &w, std::vector<double> &V)
{
    V.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) V[i + (size_t)n * i] = 1.0;
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0, tot = 0.0;
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) {
                tot += S[i + (size_t)n * j] * S[i + (size_t)n * j];
                if (i != j) off += S[i + (size_t)n * j] * S[i + (size_t)n * j];
            }
        if (off <= 1e-30 * tot || off == 0.0) break;
        for (int p = 0; p < n - 1; ++p)
            for (int q = p + 1; q < n; ++q) {
                const double apq = S[p + (size_t)n * q];
                if (apq == 0.0) continue;
                const double app = S[p + (size_t)n * p], aqq = S[q + (size_t)n * q];
                const double th = (aqq - app) / (2.0 * apq);
                const double t = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), sn = t * c;
                for (int k = 0; k < n; ++k) {
                    const double skp = S[k + (size_t)n * p], skq = S[k + (size_t)n * q];
                    S[k + (size_t)n * p] = c * skp - sn * skq;
                    S[k + (size_t)n * q] = sn * skp + c * skq;
                }
                for (int k = 0; k < n; ++k) {
                    const double spk = S[p + (size_t)n * k], sqk = S[q + (size_t)n * k];
                    S[p + (size_t)n * k] = c * spk - sn * sqk;
                    S[q + (size_t)n * k] = sn * spk + c * sqk;
                }
                for (int k = 0; k < n; ++k) {
                    const double vkp = V[k + (size_t)n * p], vkq = V[k + (size_t)n * q];
                    V[k + (size_t)n * p] = c * vkp - sn * vkq;
                    V[k + (size_t)n * q] = sn * vkp + c * vkq;
                }
            }
    }
    std::vector<int> ord(n);
    for (int i = 0; i < n; ++i) ord[i] = i;
    std::sort(ord.begin(), ord.end(), [&](int a, int b) { return S[a + (size_t)n * a] < S[b + (size_t)n * b]; });
    std::vector<double> V2((size_t)n * n);
    w.resize(n);
    for (int j = 0; j < n; ++j) {
        const int src = ord[j];
        w[j] = S[src + (size_t)n * src];
        int imax = 0;
        for (int i = 1; i < n; ++i)
            if (std::fabs(V[i + (size_t)n * src]) > std::fabs(V[imax + (size_t)n * src])) imax = i;
        const double sg = V[imax + (size_t)n * src] < 0 ? -1.0 : 1.0;
        for (int i = 0; i < n; ++i) V2[i + (size_t)n * j] = sg * V[i + (size_t)n * src];
    }
    V.swap(V2);
}

}  // namespace tpc2

namespace tpc2 {

// Fraction-of-negativity measure of one shell, negativityCheck
// (spherical_harmonics.jl:625-628): the shell's SH series is evaluated on the
// (2 LMAX)^2 grid of its surface (theta_j = pi (2j+1) / (4 LMAX), phi_i =
// 2 pi i / (2 LMAX)) with Ylm = sphPlm e^{i m phi} (spherical_harmonics.jl:89-91),
// and sum |min(Re f, 0)| / sum |f| is returned (0/0 gives NaN, which then
// never triggers the flip, as in the reference).  P holds sphPlm(l, m>=0,
// theta_j) at P[(j * LMAX + l) * LMAX + m].
inline double shell_negativity(const std::complex<double> *c, int LMAX, const std::vector<double> &P)
{
    const int S = 2 * LMAX;
    std::vector<std::complex<double>> g((size_t)S * (2 * LMAX - 1));
    // g[j][m + LMAX - 1] = sum_l c_lm sphPlm(l, m, theta_j)
    for (int j = 0; j < S; ++j)
        for (int m = -(LMAX - 1); m <= LMAX - 1; ++m) {
            const int am = m < 0 ? -m : m;
            const double sg = (m < 0 && (am & 1)) ? -1.0 : 1.0;
            std::complex<double> s = 0.0;
            for (int l = am; l < LMAX; ++l)
                s += c[tph::seanindex(l, m, LMAX) - 1] * (sg * P[((size_t)j * LMAX + l) * LMAX + am]);
            g[(size_t)j * (2 * LMAX - 1) + m + LMAX - 1] = s;
        }
    double neg = 0.0, tot = 0.0;
    for (int j = 0; j < S; ++j)
        for (int i = 0; i < S; ++i) {
            const double phi = 2.0 * tph::kPi * i / S;
            std::complex<double> f = 0.0;
            for (int m = -(LMAX - 1); m <= LMAX - 1; ++m)
                f += g[(size_t)j * (2 * LMAX - 1) + m + LMAX - 1] * std::polar(1.0, m * phi);
            neg += std::fabs(std::min(f.real(), 0.0));
            tot += std::abs(f);
        }
    return neg / tot;
}

// retrieveSolution (correlations.jl:262-319).  c2: Julia
// layout [N][K2dim][K2dim] (c2[a + N (k2-1 + K2dim (k1-1))]); shells lo..hi
// (1-based) are fitted; out: hi shells of ld complex coefficients (every other
// entry 0).  Per even l <= L: for k1 in lo..hi, k2 in lo..k1 the symmetrised
// slice 0.5 (s + reverse(s)) is fitted by least squares with the columns
// P_l'(cos alpha_a) / (4 pi), l' = 0:2:upperL, upperL = min(L, 2 min(k1,k2) - 2);
// G[k1,k2] = G[k2,k1] = the coefficient of l (when l <= upperL).  The
// 2l+1 largest eigenpairs of G (ascending) give x_k[j] = sqrt(max(0, w_j))
// V[k, j] in the real basis, m = j - l; real_to_comp (spherical_harmonics.jl:
// 472-481) maps x to c = Umat(l)^H x.  If the summed shell negativity of the
// result exceeds 0.33 every coefficient changes sign.  The fit does not
// depend on l, so each (k1, k2) is fitted once.  With lambda != 0 the slices
// are not symmetrised and the columns are P_l'(cos alpha_star) / (4 pi) of the
// pair (datagen.jl:49-53) with the shell spacing mdq.  Returns an error
// message, or an empty string; *flipped (if given) reports the sign change.
inline std::string retrieve_solution(const double *c2, int N, int K2dim, int lo, int hi, int L, int LMAX,
                                     std::complex<double> *out, int64_t ld, int *flipped, double *negativity,
                                     double lambda = 0.0, double mdq = 0.0)
{
    const int K2 = hi - lo + 1;
    if (lo < 1 || hi > K2dim || K2 < 1) return "need 1 <= k2lo <= k2hi <= K2";
    if (!(2 * L < K2)) return "need L < K2/2 (K2 = k2hi - k2lo + 1)";
    if (L < 0 || L > LMAX - 1) return "need 0 <= L <= LMAX - 1";
    if (ld < (int64_t)LMAX * LMAX) return "coefficient stride shorter than LMAX^2";
    const int nlf = L / 2 + 1;
    if (N < nlf) return "need N >= L/2 + 1 angles for the Legendre fit";
    // Legendre columns / (4 pi) on the alpha grid
    std::vector<double> Pl((size_t)N * nlf);
    for (int a = 0; a < N; ++a) {
        const double x = std::cos((a + 0.5) * tph::kPi / N);
        tp_legendre(2 * (nlf - 1), x, [&](int l, double p) {
            if ((l & 1) == 0) Pl[(size_t)a + (size_t)N * (l / 2)] = p / (4.0 * tph::kPi);
        });
    }
    // fitted coefficients per (k1, k2): fit[(k1-lo) * K2 + (k2-lo)][l/2]
    std::vector<double> fit((size_t)K2 * K2 * nlf, 0.0);
    std::vector<int> up((size_t)K2 * K2, -1);
    std::vector<double> s(N), f;
    for (int k1 = lo; k1 <= hi; ++k1)
        for (int k2 = lo; k2 <= k1; ++k2) {
            const double *sl = c2 + (size_t)N * ((k2 - 1) + (size_t)K2dim * (k1 - 1));
            if (lambda == 0.0)   // symmetrised on the flat detector only (correlations.jl:282-285)
                for (int a = 0; a < N; ++a) s[a] = 0.5 * (sl[a] + sl[N - 1 - a]);
            else
                for (int a = 0; a < N; ++a) s[a] = sl[a];
            const int upperL = std::min(L, 2 * std::min(k1, k2) - 2);
            const int nc = upperL / 2 + 1;
            std::vector<double> A(Pl.begin(), Pl.begin() + (size_t)N * nc);
            if (lambda != 0.0) {
                const double t1 = std::acos(k1 * mdq * lambda / (4 * tph::kPi));
                const double t2 = std::acos(k2 * mdq * lambda / (4 * tph::kPi));
                for (int a = 0; a < N; ++a) {
                    const double al = (a + 0.5) * tph::kPi / N;
                    const double x = std::cos(std::acos(std::sin(t1) * std::sin(t2) * std::cos(al) +
                                                        std::cos(t1) * std::cos(t2)));
                    tp_legendre(2 * (nc - 1), x, [&](int l, double p) {
                        if ((l & 1) == 0) A[(size_t)a + (size_t)N * (l / 2)] = p / (4.0 * tph::kPi);
                    });
                }
            }
            lstsq(A, N, nc, s, f);
            const size_t o = (size_t)(k1 - lo) * K2 + (k2 - lo);
            up[o] = upperL;
            for (int j = 0; j < nc; ++j) fit[o * nlf + j] = f[j];
        }
    for (int k = 0; k < hi; ++k)
        for (int64_t e = 0; e < ld; ++e) out[(size_t)k * ld + e] = 0.0;
    const double r2 = std::sqrt(0.5);
    std::vector<double> G, w, V;
    for (int l = 0; l <= L; l += 2) {
        G.assign((size_t)K2 * K2, 0.0);
        for (int i1 = 0; i1 < K2; ++i1)
            for (int i2 = 0; i2 <= i1; ++i2) {
                const size_t o = (size_t)i1 * K2 + i2;
                if (l <= up[o]) G[i1 + (size_t)K2 * i2] = G[i2 + (size_t)K2 * i1] = fit[o * nlf + l / 2];
            }
        sym_eig(G, K2, w, V);
        const int n = 2 * l + 1, j0 = K2 - n;
        for (int kk = 0; kk < K2; ++kk) {
            std::vector<double> x(n);
            for (int j = 0; j < n; ++j) x[j] = std::sqrt(std::max(0.0, w[j0 + j])) * V[kk + (size_t)K2 * (j0 + j)];
            std::complex<double> *c = out + (size_t)(kk + lo - 1) * ld;
            // c = Umat(l)^H x: c_0 = x_0; c_p = (x_p + i x_-p)/sqrt2; c_-p = (-1)^p (x_p - i x_-p)/sqrt2
            c[tph::seanindex(l, 0, LMAX) - 1] = x[l];
            for (int p = 1; p <= l; ++p) {
                const double xp = x[l + p], xm = x[l - p], sg = (p & 1) ? -1.0 : 1.0;
                c[tph::seanindex(l, p, LMAX) - 1] = std::complex<double>(r2 * xp, r2 * xm);
                c[tph::seanindex(l, -p, LMAX) - 1] = std::complex<double>(sg * r2 * xp, -sg * r2 * xm);
            }
        }
    }
    // negativityCheck over the fitted shells
    const int S = 2 * LMAX;
    std::vector<double> P((size_t)S * LMAX * LMAX, 0.0);
    for (int j = 0; j < S; ++j) {
        const double th = tph::kPi * (2 * j + 1) / (4.0 * LMAX);
        for (int l = 0; l < LMAX; ++l)
            for (int m = 0; m <= l; ++m) P[((size_t)j * LMAX + l) * LMAX + m] = tph::sphPlm(l, m, th);
    }
    double neg = 0.0;
    for (int k = lo; k <= hi; ++k) {
        const double r = shell_negativity(out + (size_t)(k - 1) * ld, LMAX, P);
        neg += std::isnan(r) ? r : std::max(r, 0.0);
    }
    const bool flip = neg > 0.33;
    if (flip)
        for (int k = 0; k < hi; ++k)
            for (int64_t e = 0; e < ld; ++e) out[(size_t)k * ld + e] = -out[(size_t)k * ld + e];
    if (flipped) *flipped = flip ? 1 : 0;
    if (negativity) *negativity = neg;
    return std::string();
}

}  // namespace tpc2

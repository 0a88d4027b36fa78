// tp_c2.hpp -- two-photon correlation of a structure (twoPhotons,
// correlations.jl:229-243) on the device.  On the flat detector (lambda = 0)
// alpha_star(alpha) = alpha and the Legendre table is shared by every shell
// pair; otherwise the series is taken at alpha_star (datagen.jl:49-53), the
// angle between the two shells' Ewald-sphere points.
#pragma once

// Legendre P_l(x), l = 0..lmax, by Bonnet's recurrence
// (n + 1) P_{n+1} = (2n + 1) x P_n - n P_{n-1}; emit(l, P_l(x)).
template <typename F>
__host__ __device__ inline void tp_legendre(int lmax, double x, F &&emit)
{
    double p0 = 1.0, p1 = x;
    emit(0, p0);
    if (lmax >= 1) emit(1, p1);
    for (int n = 1; n < lmax; ++n) {
        const double p2 = ((2.0 * n + 1.0) * x * p1 - n * p0) / (n + 1.0);
        emit(n + 1, p2);
        p0 = p1;
        p1 = p2;
    }
}

constexpr int kC2Chunk = 8;   // shells k2 per workgroup

// LDS bytes of k_two_photons: shell k1 (n_c double2), the Legendre table
// (N x nl) and the dot products of one chunk (kC2Chunk x nl).
inline size_t two_photons_lds_bytes(int n_c, int N, int nl)
{
    return sizeof(double) * (2 * (size_t)n_c + (size_t)N * nl + (size_t)kC2Chunk * nl);
}

// c2[a, k2, k1] = scale * sum_{l in 0:lstep:L} Re(sum_m conj(c_k1[l,m]) c_k2[l,m]) P_l(cos alpha_a),
// alpha_a = (a + 1/2) pi / N (alpharange, correlations.jl:247-250).  Workgroup
// (k1, chunk of kC2Chunk shells k2): shell k1 and the table P_l(cos alpha_a)
// go to LDS, one thread per (k2, l) forms the dot product, then one thread
// per (k2, a) sums the Legendre series from LDS.  With minimal only k2 <= k1
// is computed (the rest is written 0, as the reference leaves it).  coef:
// compact [K2][n_c], order l at coff[l], m = -l..l.  out: [K2][K2][N] (a
// fastest) == Julia c2[:, k2, k1].  ewc = dq lambda / (4 pi): 0 on the flat
// detector; otherwise theta_k = acos(k ewc) and the series is taken at
// P_l(cos alpha_star), cos alpha_star = sin th1 sin th2 cos alpha + cos th1 cos th2.
__global__ __launch_bounds__(256) void k_two_photons(const double2 *__restrict__ coef, int n_c,
                                                     const int *__restrict__ coff, int L, int lstep, int N, int K2,
                                                     int minimal, double scale, double ewc, double *__restrict__ out)
{
    extern __shared__ double lds[];
    const int nl = L / lstep + 1;
    double2 *sa = reinterpret_cast<double2 *>(lds);
    double *pt = lds + 2 * n_c;        // [N][nl]
    double *dl = pt + N * nl;          // [kC2Chunk][nl]
    const int k1 = blockIdx.x, c0 = blockIdx.y * kC2Chunk;
    const int nk = min(kC2Chunk, K2 - c0);
    const int k2n = minimal ? min(nk, k1 + 1 - c0) : nk;   // shells of the chunk to compute
    double *o = out + (size_t)N * (c0 + (size_t)K2 * k1);
    if (k2n <= 0) {
        for (int t = threadIdx.x; t < nk * N; t += blockDim.x) o[t] = 0.0;
        return;
    }
    for (int t = threadIdx.x; t < n_c; t += blockDim.x) sa[t] = coef[(size_t)k1 * n_c + t];
    for (int a = threadIdx.x; a < N; a += blockDim.x) {
        const double x = cos((a + 0.5) * M_PI / N);
        tp_legendre(L, x, [&](int l, double p) {
            if (l % lstep == 0) pt[a * nl + l / lstep] = p;
        });
    }
    __syncthreads();
    for (int t = threadIdx.x; t < k2n * nl; t += blockDim.x) {
        const int j = t / nl, l = (t - j * nl) * lstep, off = coff[l];
        const double2 *b = coef + (size_t)(c0 + j) * n_c + off;
        double s = 0.0;
        for (int m = 0; m <= 2 * l; ++m) s += sa[off + m].x * b[m].x + sa[off + m].y * b[m].y;
        dl[t] = s;
    }
    __syncthreads();
    for (int t = threadIdx.x; t < nk * N; t += blockDim.x) {
        const int j = t / N, a = t - j * N;
        double v = 0.0;
        if (j < k2n) {
            const double *d = dl + j * nl;
            if (ewc == 0.0) {
                const double *p = pt + a * nl;
                for (int li = 0; li < nl; ++li) v += d[li] * p[li];
            } else {
                const double t1 = acos((k1 + 1) * ewc), t2 = acos((c0 + j + 1) * ewc);
                const double al = (a + 0.5) * M_PI / N;
                const double x = cos(acos(sin(t1) * sin(t2) * cos(al) + cos(t1) * cos(t2)));
                tp_legendre(L, x, [&](int l, double p) {
                    if (l % lstep == 0) v += d[l / lstep] * p;
                });
            }
            v *= scale;
        }
        o[t] = v;
    }
}

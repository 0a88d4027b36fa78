// tp_hist.hpp -- photon-pair and photon-triplet histograms of sparse images
// (histogramCorrelationsInPicture_alltoall, datagen.jl:228-282; angle_between,
// utilities.jl:61-72), one workgroup per picture.
//
// Per picture: keep the photons with k = round(|p|/dq) in 1..K2 (rint: ties
// to even, as Julia's round), order them by |p| ascending keeping the input
// order among equal norms (a stable sort, as Julia's sort), then for every
// i > j > m (so k1 >= k2 >= k3):
//   c1[k1]                  one count (weight 1/k1 applied when read);
//   alpha = angle_between(p_i, p_j), alpha_s = 2 pi - alpha if alpha > pi,
//   ai = floor(alpha_s / da) mod N, da = pi / N;  c2[ai, k2, k1] one count;
//   if k1, k2, k3 <= K3: beta = angle_between(p_i, p_m) (2 pi - beta if
//   alpha > pi), bi = floor(beta / da) mod 2N;  c3[ai, bi, k3, k2, k1] one count.
// The reference adds the weight of a bin (1/(doubletFactor k1 k2),
// 1/(tripletFactor k1 k2 k3)) once per contribution; every contribution of a
// bin has the same weight, so the histograms are kept as integer counts
// (64-bit atomics: exact and independent of the order in which workgroups
// finish) and weighted when read.  The angle of p_i with every lower photon is
// computed once per i (an LDS row): both alpha and beta of the triplets of
// photon i are entries of that row.  Arithmetic is written in the order of the
// reference (no fused multiply-adds) so that bins agree with the oracle.
#pragma once

constexpr int kHistMaxPhotons = 512;   // per picture (scattered photons of one image)
constexpr int kHistThreads = 256;

__device__ __forceinline__ double hist_angle(double x1, double y1, double z1, double n1, double x2, double y2,
                                             double z2, double n2)
{
#pragma clang fp contract(off)
    double val = (x1 * x2 + y1 * y2 + z1 * z2) / (n1 * n2);
    val = val < -1.0 ? -1.0 : (val > 1.0 ? 1.0 : val);
    if (fabs(val) > 0.0) {
        double a = acos(val);
        if (x1 * y2 - y1 * x2 > 0.0) a = 2.0 * M_PI - a;
        return a;
    }
    return 0.0;   // orthogonal photons: the reference returns 0 here
}

__device__ __forceinline__ int hist_mod_floor(double x, int m)
{
    const long long f = (long long)floor(x);
    long long r = f % m;
    return (int)(r < 0 ? r + m : r);
}

// Pictures: photons [off[b], off[b+1]) of ph, or with cnt given the cnt[b]
// photons of slot b (ph[b cap .. b cap + cap), generated pictures); a slot
// whose picture accepted more than cap photons lost some of them and is
// skipped whole, so that no picture is histogrammed with a subset.
__global__ __launch_bounds__(kHistThreads) void k_hist_pictures(const double *__restrict__ ph,
                                                                const int64_t *__restrict__ off,
                                                                const int *__restrict__ cnt, int cap, double dq,
                                                                int N, int K2, int K3,
                                                                unsigned long long *__restrict__ n1,
                                                                unsigned long long *__restrict__ n2,
                                                                unsigned long long *__restrict__ n3)
{
#pragma clang fp contract(off)
    __shared__ double sx[kHistMaxPhotons], sy[kHistMaxPhotons], sz[kHistMaxPhotons], sn[kHistMaxPhotons];
    __shared__ double ox[kHistMaxPhotons], oy[kHistMaxPhotons], oz[kHistMaxPhotons], on_[kHistMaxPhotons];
    __shared__ int sk[kHistMaxPhotons], ok_[kHistMaxPhotons];
    __shared__ double ang[kHistMaxPhotons];
    __shared__ int s_l;
    const int tid = threadIdx.x;
    const int64_t p0 = cnt ? (int64_t)blockIdx.x * cap : off[blockIdx.x];
    const int n = cnt ? (cnt[blockIdx.x] > cap ? 0 : cnt[blockIdx.x]) : (int)(off[blockIdx.x + 1] - p0);
    const double da = M_PI / N;
    if (tid == 0) s_l = 0;
    for (int i = tid; i < n; i += kHistThreads) {
        const double x = ph[3 * (p0 + i)], y = ph[3 * (p0 + i) + 1], z = ph[3 * (p0 + i) + 2];
        const double nm = sqrt(x * x + y * y + z * z);
        const double kr = rint(nm / dq);
        ox[i] = x;
        oy[i] = y;
        oz[i] = z;
        on_[i] = nm;
        ok_[i] = (kr >= 1.0 && kr <= (double)K2) ? (int)kr : 0;
    }
    __syncthreads();
    // stable rank by norm among the kept photons
    for (int i = tid; i < n; i += kHistThreads) {
        if (ok_[i] == 0) continue;
        const double ni = on_[i];
        int r = 0;
        for (int j = 0; j < n; ++j)
            if (ok_[j] != 0 && (on_[j] < ni || (on_[j] == ni && j < i))) ++r;
        sx[r] = ox[i];
        sy[r] = oy[i];
        sz[r] = oz[i];
        sn[r] = ni;
        sk[r] = ok_[i];
        atomicAdd(&s_l, 1);
    }
    __syncthreads();
    const int l = s_l;
    if (l < 3) return;   // the reference records nothing for such a picture
    for (int i = tid; i < l; i += kHistThreads) atomicAdd(&n1[sk[i] - 1], 1ull);
    const int R = 2 * N * N;
    for (int i = 1; i < l; ++i) {
        const double xi = sx[i], yi = sy[i], zi = sz[i], ni = sn[i];
        const int k1 = sk[i];
        for (int m = tid; m < i; m += kHistThreads) ang[m] = hist_angle(xi, yi, zi, ni, sx[m], sy[m], sz[m], sn[m]);
        __syncthreads();
        // pairs (i, j)
        for (int j = tid; j < i; j += kHistThreads) {
            const double al = ang[j];
            const double als = al > M_PI ? 2.0 * M_PI - al : al;
            const int ai = hist_mod_floor(als / da, N);
            atomicAdd(&n2[ai + (size_t)N * ((sk[j] - 1) + (size_t)K2 * (k1 - 1))], 1ull);
        }
        // triplets (i, j, m), m < j < i, flattened over t = j (j - 1) / 2 + m
        if (k1 <= K3) {
            const int tp = i * (i - 1) / 2;
            for (int t = tid; t < tp; t += kHistThreads) {
                int j = (int)((1.0 + sqrt(1.0 + 8.0 * t)) * 0.5);
                while (j * (j - 1) / 2 > t) --j;
                while ((j + 1) * j / 2 <= t) ++j;
                const int m = t - j * (j - 1) / 2;
                const int k2 = sk[j], k3 = sk[m];
                if (k2 > K3 || k3 > K3) continue;
                const double al = ang[j];
                const double als = al > M_PI ? 2.0 * M_PI - al : al;
                const int ai = hist_mod_floor(als / da, N);
                double be = ang[m];
                if (al > M_PI) be = 2.0 * M_PI - be;
                const int bi = hist_mod_floor(be / da, 2 * N);
                const long long kt = (long long)(k1 - 1) * k1 * (k1 + 1) / 6 + (long long)(k2 - 1) * k2 / 2 + (k3 - 1);
                atomicAdd(&n3[kt * R + (long long)bi * N + ai], 1ull);
            }
        }
        __syncthreads();
    }
}

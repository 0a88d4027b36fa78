// tp_fit.hpp -- rotational fit of a structure to a reference (fitStructures_full
// / fitStructures_random, data_processing.jl:26-79): the similarity of the
// reference with many rotated copies of a structure, scored on the device.
//
// similarity(reference, volume, K_range) (spherical_harmonics.jl:426-435) is
// the mean over the shells k in K_range of Pearson's correlation
// (cor_ISC, utilities.jl:45-47) between the real parts of the two shells'
// surfaces on the (2 LMAX)^2 grid of the inverse SH transform (backtransform,
// spherical_harmonics.jl:139-148): theta_j = pi (2j+1) / (4 LMAX),
// phi_i = 2 pi i / (2 LMAX), f = sum_{l < LMAX} sum_m c_lm Y_lm with
// Y_lm = sphPlm e^{i m phi} (spherical_harmonics.jl:89-91).
//
// The phi sums are done in closed form: with g_j(m) = sum_l c_lm sphPlm(l, m, theta_j)
// and B = LMAX, over the 2B uniform phi_i (|m|, |m'| < B, so m + m' = 0 is the
// only alias):
//   sum_i Re f            = 2B Re g_j(0)
//   sum_i Re f Re h       = B Re[sum_m g_j(m) h_j(-m) + sum_m g_j(m) conj(h_j(m))]
//   sum_i (Re f)^2        = B Re[sum_m g_j(m) g_j(-m) + sum_m |g_j(m)|^2]
// so a shell costs 2B x LMAX^2 / 2 complex MACs and no transform.
#pragma once

constexpr int kFitThreads = 256;
constexpr int kFitMaxLMAX = 32;

__device__ __forceinline__ double fit_block_sum(double v, double *red)
{
    v = wave_sum(v);
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[w] = v;
    __syncthreads();
    double s = 0.0;
    for (int i = 0; i < kFitThreads / 64; ++i) s += red[i];
    return s;
}

// One workgroup per (shell kk of the range, rotation r).  The shell's
// coefficients x_{l,m} (l < LMAX) are gathered into LDS: rotated entries
// (even l <= L, compact layout) from rot[r], all others from vol.  Work item
// (m >= 0, j) with j fastest, so the lanes of a wave share m (LDS broadcast of
// x) and read consecutive entries of P: sphPlm(l, m, theta_j) at
// P[(m * LMAX + l) * 2 LMAX + j].  sean:
// seanindex - 1 of (l, m) at sean[l * l + l + m].
// Reference mode (rot == nullptr, ref_out != nullptr): g_j(m) of the shell is
// written to gref[kk][m + LMAX - 1][j] and (sum Re h, sum (Re h)^2) to
// ref_out[kk]; cor is not written.
// Otherwise cor[r * nk + kk] = the clamped Pearson correlation with the
// reference shell kk (gref, ref_out as written by the reference mode).
__global__ __launch_bounds__(kFitThreads) void k_fit_similarity(
    const double2 *__restrict__ vol, int64_t ldv, const double2 *__restrict__ rot, int n_c,
    const int *__restrict__ coff, int L, int LMAX, int K, int klo, int nk, const int *__restrict__ sean,
    const double *__restrict__ P, double2 *__restrict__ gref, double *__restrict__ ref_out,
    double *__restrict__ cor)
{
    __shared__ double2 xs[kFitMaxLMAX * kFitMaxLMAX];
    __shared__ double red[kFitThreads / 64];
    const int kk = blockIdx.x, r = blockIdx.y, k = klo + kk;
    const int B = LMAX, nq = LMAX * LMAX, nmm = 2 * LMAX - 1;
    const bool refmode = rot == nullptr;
    for (int q = threadIdx.x; q < nq; q += kFitThreads) {
        const int l = (int)sqrt((double)q + 0.5), m = q - l * l - l;
        double2 x;
        if (!refmode && (l & 1) == 0 && l <= L)
            x = rot[((size_t)r * K + k) * n_c + coff[l] + m + l];
        else
            x = vol[(size_t)k * ldv + sean[q]];
        xs[q] = x;
    }
    __syncthreads();
    const double2 *h = gref + (size_t)kk * (2 * B) * nmm;
    double sf = 0.0, sff = 0.0, sfh = 0.0;
    for (int it = threadIdx.x; it < 2 * B * LMAX; it += kFitThreads) {
        const int mm = it / (2 * B), j = it - mm * (2 * B);
        const double *pj = P + (size_t)mm * LMAX * (2 * B) + j;
        double gpr = 0.0, gpi = 0.0, gmr = 0.0, gmi = 0.0;
        for (int l = mm; l < LMAX; ++l) {
            const double p = pj[(size_t)l * (2 * B)];
            const double2 xp = xs[l * l + l + mm], xm = xs[l * l + l - mm];
            gpr += xp.x * p;
            gpi += xp.y * p;
            gmr += xm.x * p;
            gmi += xm.y * p;
        }
        if (mm & 1) { gmr = -gmr; gmi = -gmi; }   // sphPlm(l, -m) = (-1)^m sphPlm(l, m)
        double2 *gw = refmode ? gref + (size_t)kk * (2 * B) * nmm + (size_t)(LMAX - 1) * (2 * B) + j : nullptr;
        if (mm == 0) {
            sf += 2.0 * B * gpr;
            sff += B * (gpr * gpr - gpi * gpi + gpr * gpr + gpi * gpi);
            if (refmode) {
                gw[0] = make_double2(gpr, gpi);
            } else {
                const double2 h0 = h[(size_t)(LMAX - 1) * (2 * B) + j];
                // g0 h0 + g0 conj(h0)
                sfh += B * ((gpr * h0.x - gpi * h0.y) + (gpr * h0.x + gpi * h0.y));
            }
        } else {
            // g_m g_-m (twice) + |g_m|^2 + |g_-m|^2
            sff += B * (2.0 * (gpr * gmr - gpi * gmi) + gpr * gpr + gpi * gpi + gmr * gmr + gmi * gmi);
            if (refmode) {
                gw[(ptrdiff_t)mm * (2 * B)] = make_double2(gpr, gpi);
                gw[-(ptrdiff_t)mm * (2 * B)] = make_double2(gmr, gmi);
            } else {
                const double2 hp = h[(size_t)(LMAX - 1 + mm) * (2 * B) + j], hm = h[(size_t)(LMAX - 1 - mm) * (2 * B) + j];
                // g_m h_-m + g_-m h_m + g_m conj(h_m) + g_-m conj(h_-m)
                sfh += B * ((gpr * hm.x - gpi * hm.y) + (gmr * hp.x - gmi * hp.y) + (gpr * hp.x + gpi * hp.y) +
                            (gmr * hm.x + gmi * hm.y));
            }
        }
    }
    sf = fit_block_sum(sf, red);
    sff = fit_block_sum(sff, red);
    if (refmode) {
        if (threadIdx.x == 0) {
            ref_out[2 * kk] = sf;
            ref_out[2 * kk + 1] = sff;
        }
        return;
    }
    sfh = fit_block_sum(sfh, red);
    if (threadIdx.x == 0) {
        const double npts = 4.0 * B * B;
        const double sh = ref_out[2 * kk], shh = ref_out[2 * kk + 1];
        const double num = sfh - sf * sh / npts;
        const double den = sqrt((sff - sf * sf / npts) * (shh - sh * sh / npts));
        double c = num / den;
        c = c > 1.0 ? 1.0 : (c < -1.0 ? -1.0 : c);
        cor[(size_t)r * nk + kk] = c;
    }
}

// tp_pictures.hpp -- synthetic sparse pictures: pointsPerOrientation
// (datagen.jl:88-144) for many orientations at once, one workgroup per picture.
//
// Per picture: a Haar-random rotation R (a uniform unit quaternion from four
// normals: the same distribution as random_rotation(3), utilities.jl:165-176);
// n incident photons p = (x, y, 0), x, y ~ N(0, sigma^2); those outside the
// beamstop (|x| > w and |y| > w) with |p| <= qcut are mapped to the Ewald
// sphere (detector_to_Ewald_sphere, datagen.jl:33-39) and rotated, r = R e(p);
// the intensity there is interpolated from the shells' surfaces
// (getVolumeInterpolated for a SurfaceVolume, spherical_harmonics.jl:484-530,
// radial interpolation on) and p is accepted when u < I(r) / (maxI / g(0) g(r)),
// g the isotropic 3-D normal density of width sigma.  Accepted photons keep
// the incident order and their detector position p.  Random numbers: Philox
// streams of (seed, picture index): rotation normals, photon normals
// (x = normal 3i, y = normal 3i + 1 of the picture's stream) and one uniform
// per photon.  Arithmetic is written in the reference's order, FMA contraction off.
#pragma once

#define TP_TAG_PROT 0x50524f54u   // picture rotation normals
#define TP_TAG_PPHO 0x5050484fu   // incident photon normals
#define TP_TAG_PACC 0x50414343u   // acceptance uniforms
#define TP_TAG_PCNT 0x50434e54u   // photon-count normals (incident_photon_variance)
#define TP_TAG_PBST 0x50425354u   // beamstop-shot uniforms

constexpr int kPicThreads = 256;

// What one launch of k_pictures draws (passed by value).
struct PicArgs {
    const double2 *surf;    // (K, (2 LMAX)^2) intensity surfaces
    const double2 *nsurf;   // noise surfaces (n_noise > 0)
    int LMAX, K;
    double dr;              // shell spacing rmax / K
    double maxI, nmaxI;     // envelope bounds 1.3 max(surface)
    double qcut, sigma, nsigma, lambda, bw;
    int n_inc, n_noise, variance, particles, beamstop_only, cap, first;
    uint64_t seed;
};

// acceptance uniform of photon i of segment seg of picture pic
__device__ inline double pic_uniform(uint64_t seed, uint32_t pic, uint32_t i, uint32_t seg)
{
    uint32_t c[4] = {i, 0xFFFFFFFEu - seg, pic, TP_TAG_PACC};
    philox10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    return u01(c[0], c[1]);
}

// incident photons of segment seg: round(Normal(n, variance)) (ties to even,
// as Julia's round), at least 0
__device__ inline int pic_count(uint64_t seed, uint32_t pic, uint32_t seg, int n, int variance)
{
    if (variance <= 0) return n;
    const double v = rint((double)n + (double)variance * dev_normal(seed, pic, seg, 0, TP_TAG_PCNT));
    return v > 0.0 ? (int)v : 0;
}

// surface index (0-based) of ang_to_i (spherical_harmonics.jl:213-221)
__device__ inline int pic_ang_to_i(double phi, double theta, int S)
{
#pragma clang fp contract(off)
    phi = fmod(phi, 2.0 * kPiD);   // mod(phi, 2 pi) for phi >= 0
    theta = fmod(theta, kPiD);     // mod(theta, pi) for theta >= 0
    const int ip = (int)rint(phi / (2.0 * kPiD) * S);
    const int it = (int)rint(((theta * 2 * S) / kPiD - 1) / 2);
    int i = it * S + ip + 1;
    i = i < 1 ? 1 : (i > S * S ? S * S : i);
    return i - 1;
}

// real part of the trilinear interpolation of the surfaces at r (radial_interp = true)
__device__ inline double pic_interp(const double2 *__restrict__ surf, int LMAX, int K, double dr, double x, double y,
                                    double z)
{
#pragma clang fp contract(off)
    const int S = 2 * LMAX;
    const double q = sqrt(x * x + y * y + z * z);
    double p = atan2(y, x);
    if (p < 0.0) p = p + 2.0 * kPiD;
    double cz = q > 0.0 ? z / q : 0.0;
    cz = cz < -1.0 ? -1.0 : (cz > 1.0 ? 1.0 : cz);
    const double t = q > 0.0 ? acos(cz) : 0.0;
    const double dp = 2.0 * kPiD / S, dt = kPiD / S;
    const double k = q / dr;
    double fk0 = floor(k), fk1 = ceil(k);
    fk0 = fk0 < 1.0 ? 1.0 : (fk0 > K ? (double)K : fk0);
    fk1 = fk1 < 1.0 ? 1.0 : (fk1 > K ? (double)K : fk1);
    const int k0 = (int)fk0, k1 = (int)fk1;
    const double p0 = floor(p / dp) * dp, p1 = ceil(p / dp) * dp;
    double t0 = floor(((t * 2 * S) / kPiD - 1) / 2) * dt + dt / 2;
    double t1 = ceil(((t * 2 * S) / kPiD - 1) / 2) * dt + dt / 2;
    t0 = t0 > 0.0 ? t0 : 0.0;
    t1 = t1 < kPiD ? t1 : kPiD;
    const double a = k - fk0, b = (p - p0) / dp, c = (t - t0) / dt;
    const double V000 = (1 - a) * (1 - b) * (1 - c), V100 = a * (1 - b) * (1 - c), V010 = (1 - a) * b * (1 - c);
    const double V001 = (1 - a) * (1 - b) * c, V101 = a * (1 - b) * c, V011 = (1 - a) * b * c;
    const double V110 = a * b * (1 - c), V111 = a * b * c;
    const double2 *s0 = surf + (size_t)(k0 - 1) * S * S, *s1 = surf + (size_t)(k1 - 1) * S * S;
    const int i00 = pic_ang_to_i(p0, t0, S), i10 = pic_ang_to_i(p1, t0, S);
    const int i01 = pic_ang_to_i(p0, t1, S), i11 = pic_ang_to_i(p1, t1, S);
    return V000 * s0[i00].x + V100 * s1[i00].x + V010 * s0[i10].x + V001 * s0[i01].x + V101 * s1[i01].x +
           V011 * s0[i11].x + V110 * s1[i10].x + V111 * s1[i11].x;
}

// One workgroup per picture a.first + blockIdx.x, made of segments in the
// reference's order (generateHistogram, datagen.jl:399-407): for each of the
// `particles` particles, its photons from the intensity (segment 2p) and, if
// n_noise > 0, noise photons from the noise surfaces with the same rotation
// (segment 2p + 1, envelope nsigma).  Particle p's rotation uses the normals
// of index 0xFFFFFFFF - p; segment s the photon normals of index s, the
// uniforms with word 0xFFFFFFFE - s and, with variance > 0, its photon count
// from the count normal of index s (noise segments keep n_noise).  With
// beamstop_only, one segment of beamstop_shot (datagen.jl:147-161): photons
// uniform on [-qcut, qcut]^2 (two uniforms of the Philox block {i,
// 0xFFFFFFFE, picture, PBST}) kept outside the beamstop and inside |p| <= qcut,
// identity rotation.  Accepted photons go to ph[(blockIdx.x cap + j) 3 +
// (0,1,2)] for j < cap, their number to cnt (it may exceed cap: the host
// reports the overflow), particle 0's rotation to rot (row-major) if given.
__global__ __launch_bounds__(kPicThreads) void k_pictures(const PicArgs a, double *__restrict__ ph,
                                                          int *__restrict__ cnt, double *__restrict__ rot)
{
#pragma clang fp contract(off)
    __shared__ double R[9];
    __shared__ int s_base, s_wave[kPicThreads / 64];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t pic = (uint32_t)(a.first + (int)blockIdx.x);
    const uint64_t seed = a.seed;
    if (tid == 0) s_base = 0;
    const int nseg = a.beamstop_only ? 1 : a.particles * (a.n_noise > 0 ? 2 : 1);
    for (int seg = 0; seg < nseg; ++seg) {
        const int part = a.n_noise > 0 ? seg >> 1 : seg;
        const bool noise = !a.beamstop_only && a.n_noise > 0 && (seg & 1);
        __syncthreads();   // R of the previous segment is no longer read
        if (tid == 0 && !noise) {
            if (a.beamstop_only) {
                for (int i = 0; i < 9; ++i) R[i] = (i % 4 == 0) ? 1.0 : 0.0;
            } else {
                const uint32_t rl = 0xFFFFFFFFu - (uint32_t)part;
                double w = dev_normal(seed, pic, rl, 0, TP_TAG_PROT), x = dev_normal(seed, pic, rl, 1, TP_TAG_PROT);
                double y = dev_normal(seed, pic, rl, 2, TP_TAG_PROT), z = dev_normal(seed, pic, rl, 3, TP_TAG_PROT);
                const double nq = sqrt(w * w + x * x + y * y + z * z);
                w /= nq; x /= nq; y /= nq; z /= nq;
                R[0] = 1 - 2 * (y * y + z * z); R[1] = 2 * (x * y - w * z);     R[2] = 2 * (x * z + w * y);
                R[3] = 2 * (x * y + w * z);     R[4] = 1 - 2 * (x * x + z * z); R[5] = 2 * (y * z - w * x);
                R[6] = 2 * (x * z - w * y);     R[7] = 2 * (y * z + w * x);     R[8] = 1 - 2 * (x * x + y * y);
            }
            if (rot && part == 0)
                for (int i = 0; i < 9; ++i) rot[(size_t)blockIdx.x * 9 + i] = R[i];
        }
        __syncthreads();
        const double2 *sv = noise ? a.nsurf : a.surf;
        const double sg = noise ? a.nsigma : a.sigma, mI = noise ? a.nmaxI : a.maxI;
        const int n = noise ? a.n_noise
                            : (a.beamstop_only ? a.n_inc : pic_count(seed, pic, (uint32_t)seg, a.n_inc, a.variance));
        // log of the envelope density at 0 (MvNormal with standard deviations sg)
        const double lg0 = -1.5 * log(2.0 * kPiD) - 3.0 * log(sg);
        const double max_gauss = exp(lg0);
        for (int i0 = 0; i0 < n; i0 += kPicThreads) {
            const int i = i0 + tid;
            bool acc = false;
            double x = 0.0, y = 0.0;
            if (i < n && a.beamstop_only) {
                uint32_t c[4] = {(uint32_t)i, 0xFFFFFFFEu, pic, TP_TAG_PBST};
                philox10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
                x = 2 * a.qcut * (u01(c[0], c[1]) - 0.5);
                y = 2 * a.qcut * (u01(c[2], c[3]) - 0.5);
                acc = fabs(x) > a.bw && fabs(y) > a.bw && sqrt(x * x + y * y) <= a.qcut;
            } else if (i < n) {
                x = sg * dev_normal(seed, pic, (uint32_t)seg, (uint32_t)(3 * i), TP_TAG_PPHO);
                y = sg * dev_normal(seed, pic, (uint32_t)seg, (uint32_t)(3 * i + 1), TP_TAG_PPHO);
                if (fabs(x) > a.bw && fabs(y) > a.bw) {
                    const double qn = sqrt(x * x + y * y);
                    if (qn <= a.qcut) {
                        const double th = acos(qn * a.lambda / (4 * kPiD));
                        const double c = cos(th), s = sin(th);
                        const double e0 = s * x, e1 = s * y, e2 = c * qn;
                        const double r0 = R[0] * e0 + R[1] * e1 + R[2] * e2;
                        const double r1 = R[3] * e0 + R[4] * e1 + R[5] * e2;
                        const double r2 = R[6] * e0 + R[7] * e1 + R[8] * e2;
                        const double I = pic_interp(sv, a.LMAX, a.K, a.dr, r0, r1, r2);
                        const double gx = exp(lg0 - 0.5 * ((r0 / sg) * (r0 / sg) + (r1 / sg) * (r1 / sg) +
                                                           (r2 / sg) * (r2 / sg)));
                        const double ratio = I / (mI / max_gauss * gx);
                        acc = pic_uniform(seed, pic, (uint32_t)i, (uint32_t)seg) < ratio;
                    }
                }
            }
            // in-order compaction of the accepted photons of this chunk
            const uint64_t m = __ballot(acc);
            const int before = __popcll(m & ((1ull << lane) - 1ull));
            if (lane == 0) s_wave[wave] = __popcll(m);
            __syncthreads();
            int off = s_base;
            for (int w = 0; w < wave; ++w) off += s_wave[w];
            const int pos = off + before;
            if (acc && pos < a.cap) {
                double *o = ph + ((size_t)blockIdx.x * a.cap + pos) * 3;
                o[0] = x;
                o[1] = y;
                o[2] = 0.0;
            }
            __syncthreads();
            if (tid == 0)
                for (int w = 0; w < kPicThreads / 64; ++w) s_base += s_wave[w];
            __syncthreads();
        }
    }
    if (tid == 0) cnt[blockIdx.x] = s_base;
}

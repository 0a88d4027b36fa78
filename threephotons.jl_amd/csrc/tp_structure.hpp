// tp_structure.hpp -- the structure step before the hot path:
// createSphericalHarmonicsStructure (structure.jl:124-134) after loadPDB.
//   density   getDensity (spherical_harmonics.jl:234-244): the electron density
//             of Gaussian atoms (getDensity_cartesian, structure.jl:74-87)
//             sampled on every shell r_k = k rmax / K at the (2 LMAX)^2 grid
//             points (i_to_ang, spherical_harmonics.jl:203-210;
//             sphericalToCartesian, utilities.jl:105-110), then analysed;
//   fourier   forward(density), the spherical Bessel transform;
//   intensity absoluteSquare(fourier) (spherical_harmonics.jl:393-397): the
//             squared modulus on the surfaces, analysed again.
// The analysis, synthesis and Bessel transforms are the kernels of
// tp_phase.hpp; this file adds the density sampling and the squared modulus.
#pragma once

constexpr int kStrThreads = 256;
constexpr int kStrAtomChunk = 1024;   // atoms staged in LDS per pass: 40 KB

// Atoms as the density kernel reads them: position, the amplitude
// B / (R sqrt(2 pi))^3 and the width 2 R^2.
struct StrAtom {
    double x, y, z, amp, w;
};

// surf[k][j S + i] = sum over atoms with |p - a|^2 < 5 of
// amp exp(-|p - a|^2 / w), p the grid point (phi_i, theta_j) of shell k + 1;
// one thread per grid point, atoms in LDS chunks.  The atoms come sorted by
// |a|, and range[k] = [first, end) holds every atom with ||a| - r_k| < sqrt 5
// (the others are farther than sqrt 5 from the whole shell, since
// |p - a| >= |r_k - |a||); a workgroup loops over the union of its shells'
// ranges and sums in that order.
__global__ __launch_bounds__(kStrThreads) void k_density_surface(const StrAtom *__restrict__ atoms,
                                                                 const int2 *__restrict__ range, int LMAX, int K,
                                                                 double dr, double2 *__restrict__ surf)
{
    __shared__ StrAtom sa[kStrAtomChunk];
    const int S = 2 * LMAX, NP = S * S;
    const int64_t t = (int64_t)blockIdx.x * kStrThreads + threadIdx.x;
    const bool live = t < (int64_t)K * NP;
    double x = 0.0, y = 0.0, z = 0.0;
    if (live) {
        const int k = (int)(t / NP), p = (int)(t - (int64_t)k * NP);
        const int j = p / S, i = p - j * S;
        const double r = (k + 1) * dr;
        const double phi = (double)i / S * 2.0 * kPiD;
        const double theta = kPiD * (2 * j + 1) / (2 * S);
        const double st = sin(theta);
        x = r * cos(phi) * st;
        y = r * sin(phi) * st;
        z = r * cos(theta);
    }
    const int64_t t0 = (int64_t)blockIdx.x * kStrThreads;
    const int kf = (int)(t0 / NP), kl = (int)min((int64_t)K - 1, (t0 + kStrThreads - 1) / NP);
    const int abeg = range[kf].x, aend = range[kl].y;
    double rho = 0.0;
    for (int a0 = abeg; a0 < aend; a0 += kStrAtomChunk) {
        const int na = min(kStrAtomChunk, aend - a0);
        __syncthreads();
        for (int a = threadIdx.x; a < na; a += kStrThreads) sa[a] = atoms[a0 + a];
        __syncthreads();
        for (int a = 0; a < na; ++a) {
            const StrAtom at = sa[a];
            const double dx = x - at.x, dy = y - at.y, dz = z - at.z;
            const double rsq = dx * dx + dy * dy + dz * dz;
            if (rsq < 5.0) rho += at.amp * exp(-rsq / at.w);
        }
    }
    if (live) surf[t] = make_double2(rho, 0.0);
}

// f <- |f|^2 (complex(abs.(f).^2))
__global__ __launch_bounds__(kStrThreads) void k_abs_square(double2 *__restrict__ f, int64_t n)
{
    const int64_t t = (int64_t)blockIdx.x * kStrThreads + threadIdx.x;
    if (t >= n) return;
    const double2 v = f[t];
    f[t] = make_double2(v.x * v.x + v.y * v.y, 0.0);
}

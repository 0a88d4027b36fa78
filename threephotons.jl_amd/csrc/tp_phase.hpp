// tp_phase.hpp -- phase retrieval in the spherical-harmonics representation
// (sRAAR, phases.jl:31-95) and the transforms it is made of:
//   surfaces <-> coefficients on the (2 LMAX)^2 grid of s2kit's transform
//     (getSurfaceVolume / getSphericalHarmonicsVolume, spherical_harmonics.jl:139-160):
//     theta_j = pi (2j+1) / (4 LMAX), phi_i = 2 pi i / (2 LMAX), samples
//     theta-major (p = j * 2 LMAX + i); f = sum_{l < LMAX, m} c_lm Y_lm with
//     Y_lm = sphPlm(l, m, theta) e^{i m phi} (spherical_harmonics.jl:89-91);
//     the analysis c_lm = (2 pi / 2 LMAX) sum_j w_j sum_i f conj(Y_lm) with the
//     Driscoll-Healy weights w_j, exact for band-limited surfaces;
//   the spherical Bessel (Hankel) transform of every (l, m) radial profile
//     (sphBesselTransform, spherical_harmonics.jl:296-377), a K x K real
//     matrix per l and the factor (+-i)^l.
// Coefficients on the device: [K][LMAX^2] complex, (l, m) at l^2 + l + m;
// surfaces: [K][(2 LMAX)^2] complex.  The transforms run one workgroup per
// (shell, 5 rings theta_j): the phi sums by a twiddle table in LDS, the theta
// sums against the sphPlm table P[m][l][j]; the analysis's theta sums run
// one thread per (shell, l, m) over the rings' phi sums in HBM.
#pragma once

constexpr int kPhThreads = 256;
constexpr int kPhMaxLMAX = 48;   // k_sh_synth then holds <= 46 KB of LDS
constexpr int kPhRings = 5;   // rings theta_j per workgroup of the transforms

// LDS bytes of k_sh_synth: the shell's coefficients, g of its rings, twiddles.
inline size_t sh_synth_lds_bytes(int LMAX)
{
    const int S = 2 * LMAX;
    return sizeof(double2) * ((size_t)LMAX * LMAX + (size_t)kPhRings * (2 * LMAX - 1) + S);
}
// LDS bytes of k_sh_rings: the workgroup's rings of the surface, twiddles.
inline size_t sh_rings_lds_bytes(int LMAX)
{
    const int S = 2 * LMAX;
    return sizeof(double2) * ((size_t)kPhRings * S + S);
}

// e^{2 pi i (m i mod S) / S} from the table tw[0..S)
__device__ __forceinline__ double2 ph_tw(const double2 *tw, int m, int i, int S)
{
    int r = (m * i) % S;
    if (r < 0) r += S;
    return tw[r];
}

// coefficients -> surface: workgroup (shell blockIdx.x, rings kPhRings
// blockIdx.y .. + kPhRings - 1)
__global__ __launch_bounds__(kPhThreads) void k_sh_synth(const double2 *__restrict__ c, int LMAX,
                                                         const double *__restrict__ P,
                                                         const double2 *__restrict__ twg, double2 *__restrict__ surf)
{
    extern __shared__ double2 ph_lds[];
    const int S = 2 * LMAX, nq = LMAX * LMAX, nm = 2 * LMAX - 1;
    double2 *cs = ph_lds, *g = cs + nq, *tw = g + (size_t)kPhRings * nm;
    const int k = blockIdx.x, j0 = blockIdx.y * kPhRings, nr = min(kPhRings, S - j0), tid = threadIdx.x;
    for (int q = tid; q < nq; q += kPhThreads) cs[q] = c[(size_t)k * nq + q];
    for (int i = tid; i < S; i += kPhThreads) tw[i] = twg[i];
    __syncthreads();
    // g[j][m] = sum_l c_lm sphPlm(l, m, theta_j)  (items (|m|, ring), ring fastest)
    for (int it = tid; it < nr * LMAX; it += kPhThreads) {
        const int mm = it / nr, jl = it - mm * nr;
        const double *pj = P + (size_t)mm * LMAX * S + j0 + jl;
        double pr = 0.0, pi = 0.0, mr = 0.0, mi = 0.0;
        for (int l = mm; l < LMAX; ++l) {
            const double p = pj[(size_t)l * S];
            const double2 xp = cs[l * l + l + mm], xm = cs[l * l + l - mm];
            pr += xp.x * p;
            pi += xp.y * p;
            mr += xm.x * p;
            mi += xm.y * p;
        }
        const double sg = (mm & 1) ? -1.0 : 1.0;   // sphPlm(l, -m) = (-1)^m sphPlm(l, m)
        g[(size_t)jl * nm + LMAX - 1 + mm] = make_double2(pr, pi);
        if (mm > 0) g[(size_t)jl * nm + LMAX - 1 - mm] = make_double2(sg * mr, sg * mi);
    }
    __syncthreads();
    // f[j][i] = sum_m g[j][m] e^{i m phi_i}
    for (int p = tid; p < nr * S; p += kPhThreads) {
        const int jl = p / S, i = p - jl * S;
        const double2 *gj = g + (size_t)jl * nm;
        double fr = 0.0, fi = 0.0;
        for (int m = -(LMAX - 1); m <= LMAX - 1; ++m) {
            const double2 a = gj[m + LMAX - 1], e = ph_tw(tw, m, i, S);
            fr += a.x * e.x - a.y * e.y;
            fi += a.x * e.y + a.y * e.x;
        }
        surf[(size_t)k * S * S + (size_t)(j0 + jl) * S + i] = make_double2(fr, fi);
    }
}

// surface -> coefficients, step 1: G[k][j][m] = sum_i f[j][i] e^{-i m phi_i}
// for the workgroup's rings (shell blockIdx.x, rings from kPhRings blockIdx.y)
__global__ __launch_bounds__(kPhThreads) void k_sh_rings(const double2 *__restrict__ surf, int LMAX,
                                                         const double2 *__restrict__ twg, double2 *__restrict__ G)
{
    extern __shared__ double2 ph_lds[];
    const int S = 2 * LMAX, nm = 2 * LMAX - 1;
    double2 *fs = ph_lds, *tw = fs + (size_t)kPhRings * S;
    const int k = blockIdx.x, j0 = blockIdx.y * kPhRings, nr = min(kPhRings, S - j0), tid = threadIdx.x;
    for (int p = tid; p < nr * S; p += kPhThreads) fs[p] = surf[(size_t)k * S * S + (size_t)j0 * S + p];
    for (int i = tid; i < S; i += kPhThreads) tw[i] = twg[i];
    __syncthreads();
    for (int it = tid; it < nr * nm; it += kPhThreads) {
        const int jl = it / nm, m = it - jl * nm - (LMAX - 1);
        const double2 *fj = fs + (size_t)jl * S;
        double gr = 0.0, gi = 0.0;
        for (int i = 0; i < S; ++i) {
            const double2 a = fj[i], e = ph_tw(tw, m, i, S);   // times conj(e)
            gr += a.x * e.x + a.y * e.y;
            gi += a.y * e.x - a.x * e.y;
        }
        G[((size_t)k * S + j0 + jl) * nm + m + LMAX - 1] = make_double2(gr, gi);
    }
}

// surface -> coefficients, step 2: c_lm = sum_j W_j sphPlm(l, m, theta_j) G[j][m];
// one thread per (shell, l, m); W[j] = (2 pi / S) w_j
__global__ __launch_bounds__(kPhThreads) void k_sh_analysis(const double2 *__restrict__ G, int LMAX, int K,
                                                            const double *__restrict__ P,
                                                            const double *__restrict__ W, double2 *__restrict__ c)
{
    const int S = 2 * LMAX, nq = LMAX * LMAX, nm = 2 * LMAX - 1;
    const int t = blockIdx.x * kPhThreads + threadIdx.x;
    if (t >= K * nq) return;
    const int k = t / nq, q = t - k * nq;
    const int l = (int)sqrt((double)q + 0.5), m = q - l * l - l, am = m < 0 ? -m : m;
    const double *pl = P + ((size_t)am * LMAX + l) * S;
    const double2 *Gk = G + (size_t)k * S * nm + m + LMAX - 1;
    double cr = 0.0, ci = 0.0;
    for (int j = 0; j < S; ++j) {
        const double wp = W[j] * pl[j];
        const double2 a = Gk[(size_t)j * nm];
        cr += wp * a.x;
        ci += wp * a.y;
    }
    const double sg = (m < 0 && (am & 1)) ? -1.0 : 1.0;
    c[(size_t)k * nq + q] = make_double2(sg * cr, sg * ci);
}

// out[ko][q] = (+-i)^l sum_ki H[l][ko][ki] in[ki][q], l of q; one thread per (ko, q)
__global__ __launch_bounds__(kPhThreads) void k_sph_bessel(const double2 *__restrict__ in, int LMAX, int K,
                                                           const double *__restrict__ H, int forward,
                                                           double2 *__restrict__ out)
{
    const int nq = LMAX * LMAX;
    const int t = blockIdx.x * kPhThreads + threadIdx.x;
    if (t >= K * nq) return;
    const int ko = t / nq, q = t - ko * nq, l = (int)sqrt((double)q + 0.5);
    const double *h = H + ((size_t)l * K + ko) * K;
    double sr = 0.0, si = 0.0;
    for (int ki = 0; ki < K; ++ki) {
        const double2 a = in[(size_t)ki * nq + q];
        sr += h[ki] * a.x;
        si += h[ki] * a.y;
    }
    // (i)^l forward, (-i)^l backward
    const int r = forward ? (l & 3) : ((4 - (l & 3)) & 3);
    double2 o;
    if (r == 0) o = make_double2(sr, si);
    else if (r == 1) o = make_double2(-si, sr);
    else if (r == 2) o = make_double2(-sr, -si);
    else o = make_double2(si, -sr);
    out[(size_t)ko * nq + q] = o;
}

// ---- sRAAR element-wise steps (n = K (2 LMAX)^2 points, shell = point / NP) ----

// amp = sqrt(max(Re I, 0)); F = amp e^{i phase}
__global__ __launch_bounds__(kPhThreads) void k_raar_init(const double2 *__restrict__ Isurf,
                                                          const double *__restrict__ phase, int64_t n,
                                                          double *__restrict__ amp, double2 *__restrict__ F)
{
    const int64_t t = (int64_t)blockIdx.x * kPhThreads + threadIdx.x;
    if (t >= n) return;
    const double a = sqrt(fmax(Isurf[t].x, 0.0));
    amp[t] = a;
    double s, co;
    sincos(phase[t], &s, &co);
    F[t] = make_double2(a * co, a * s);
}

// F = amp e^{i angle(F)} (B, phases.jl:17-28)
__global__ __launch_bounds__(kPhThreads) void k_raar_modulus(double2 *__restrict__ F, const double *__restrict__ amp,
                                                             int64_t n)
{
    const int64_t t = (int64_t)blockIdx.x * kPhThreads + threadIdx.x;
    if (t >= n) return;
    const double2 z = F[t];
    double s, co;
    sincos(atan2(z.y, z.x), &s, &co);
    F[t] = make_double2(amp[t] * co, amp[t] * s);
}

// x = 2 a - b
__global__ __launch_bounds__(kPhThreads) void k_raar_reflect(const double2 *__restrict__ a,
                                                             const double2 *__restrict__ b, int64_t n,
                                                             double2 *__restrict__ x)
{
    const int64_t t = (int64_t)blockIdx.x * kPhThreads + threadIdx.x;
    if (t >= n) return;
    x[t] = make_double2(2.0 * a[t].x - b[t].x, 2.0 * a[t].y - b[t].y);
}

// tmp3 = A(tmp1) (support k <= Koff and positivity, phases.jl:6-14);
// tmp_u = 0.5 (beta (2 tmp3 - tmp1) + (1 - beta) tmp1 + density);
// per-block partial sums of |density - tmp_u| and |density| into part[2 block + {0,1}]
__global__ __launch_bounds__(kPhThreads) void k_raar_update(const double2 *__restrict__ tmp1,
                                                            const double2 *__restrict__ dens, int64_t n, int NP,
                                                            int Koff, double beta, double2 *__restrict__ tmpu,
                                                            double *__restrict__ part)
{
    __shared__ double red[2][kPhThreads / 64];
    const int64_t t = (int64_t)blockIdx.x * kPhThreads + threadIdx.x;
    double dd = 0.0, dn = 0.0;
    if (t < n) {
        const double2 t1 = tmp1[t], d = dens[t];
        const int k = (int)(t / NP) + 1;
        const double t3 = k > Koff ? 0.0 : fmax(t1.x, 0.0);
        const double ur = 0.5 * (beta * (2.0 * t3 - t1.x) + (1.0 - beta) * t1.x + d.x);
        const double ui = 0.5 * (beta * (2.0 * 0.0 - t1.y) + (1.0 - beta) * t1.y + d.y);
        tmpu[t] = make_double2(ur, ui);
        dd = hypot(d.x - ur, d.y - ui);
        dn = hypot(d.x, d.y);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        dd += __shfl_xor(dd, o);
        dn += __shfl_xor(dn, o);
    }
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = dd;
        red[1][threadIdx.x >> 6] = dn;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0.0, b = 0.0;
        for (int w = 0; w < kPhThreads / 64; ++w) {
            a += red[0][w];
            b += red[1][w];
        }
        part[2 * blockIdx.x] = a;
        part[2 * blockIdx.x + 1] = b;
    }
}

// diff[n_it] = sum of the nb block partials of |density - tmp_u| / those of |density| (fixed order)
__global__ __launch_bounds__(64) void k_raar_diff(const double *__restrict__ part, int nb, double *__restrict__ diff,
                                                  int n_it)
{
    double a = 0.0, b = 0.0;
    for (int i = threadIdx.x; i < nb; i += 64) {
        a += part[2 * i];
        b += part[2 * i + 1];
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        a += __shfl_xor(a, o);
        b += __shfl_xor(b, o);
    }
    if (threadIdx.x == 0) diff[n_it] = a / b;
}

// density = A(x): Re clamped at 0, shells past Koff zeroed
__global__ __launch_bounds__(kPhThreads) void k_raar_support(const double2 *__restrict__ x, int64_t n, int NP,
                                                             int Koff, double2 *__restrict__ out)
{
    const int64_t t = (int64_t)blockIdx.x * kPhThreads + threadIdx.x;
    if (t >= n) return;
    const int k = (int)(t / NP) + 1;
    out[t] = make_double2(k > Koff ? 0.0 : fmax(x[t].x, 0.0), 0.0);
}

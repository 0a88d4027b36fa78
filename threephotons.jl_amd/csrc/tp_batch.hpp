// tp_batch.hpp -- gfx950 device code of the batched evaluation path (included
// by tp_hip.hip after tp_kernels.hpp).
//
// A batch is a set of evaluations (coefficient sets, each with its own shell
// count K and its own histogram) scored together: every stage is one launch
// over all of them, so a sweep of many annealing runs with uneven K is a
// fixed sequence of launches that is captured once as a HIP graph.
//
// Column layout: the As / c3 columns of an evaluation are cut into 64-column
// k-triple tiles; tile j of proposal p of a group (one annealing run, whose
// proposals share the run's histogram) is 64-column block
// cbase + j * nprop + p, so the blocks that read the same histogram columns
// are neighbours in the contraction's tile order.
#pragma once

#include "tp_kernels.hpp"
#include "tp_triple16.hpp"

#define TP_TAG_QAXIS 0x51415849u
#define TP_TAG_QANGLE 0x51414E47u

struct EvalDesc {
    const double2 *src;    // rotation input [K][n_c] (the group's current state), sweep mode only
    double2 *coef;         // coefficients evaluated [K][n_c]
    const double *qsum;    // [ncols] histogram column sums (after normalisation)
    const double *tf;      // [ncols] triplet factors
    int K, ncols;          // shells; energy columns (K3_range walk length)
    int cbase, nprop, p;   // column blocks (as_col)
    int group;             // annealing run
    int pad0, pad1;
};

// One annealing run of a sweep batch.  Its proposals are evaluations
// e0 .. e0 + nprop - 1.
struct GroupState {
    double E;       // current energy
    double T;       // temperature
    double amax;    // proposal angle bound, as a fraction of pi/4
    double pad;
    uint64_t seed;
    int nacc, nsweep, K, nprop, e0, nlearn;
};

// ---------------------------------------------------------------------------
// Random rotation proposals: for every even l >= 2 an independent rotation
// of SO(3) with uniformly random axis and angle theta <= theta_max drawn from
// the Haar measure restricted to that ball (density ~ sin^2(theta/2), i.e.
// CDF ~ theta - sin theta, inverted by safeguarded Newton).  Written as unit
// quaternions (w, x, y, z).  Slots p >= nprop - nlearn belong to the learned
// proposer and are left alone; probe mode makes slot 0 the identity (its
// energy is the run's current energy).
// ---------------------------------------------------------------------------
__device__ inline void philox_u01x2(uint64_t seed, uint32_t a, uint32_t b, uint32_t c, uint32_t tag, double &u0,
                                    double &u1)
{
    uint32_t x[4] = {a, b, c, tag};
    philox10(x, (uint32_t)seed, (uint32_t)(seed >> 32));
    u0 = u01(x[0], x[1]);
    u1 = u01(x[2], x[3]);
}

__device__ inline double haar_angle(double u, double tmax)
{
    const double target = u * (tmax - sin(tmax));
    double lo = 0.0, hi = tmax, t = tmax * cbrt(u);
    for (int it = 0; it < 40; ++it) {
        const double f = t - sin(t) - target;
        if (f > 0.0) hi = t; else lo = t;
        const double d = 1.0 - cos(t);
        double tn = d > 0.0 ? t - f / d : 0.5 * (lo + hi);
        if (!(tn > lo && tn < hi)) tn = 0.5 * (lo + hi);
        if (fabs(tn - t) <= 1e-15 * tmax) { t = tn; break; }
        t = tn;
    }
    return t;
}

__global__ __launch_bounds__(256) void k_quat_random(const EvalDesc *__restrict__ ev, int ne, int nl,
                                                     const GroupState *__restrict__ gs, int probe,
                                                     double *__restrict__ quat)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ne * nl) return;
    const int e = i / nl, li = i % nl;
    const EvalDesc d = ev[e];
    const GroupState g = gs[d.group];
    double4 q = make_double4(1.0, 0.0, 0.0, 0.0);
    if (!probe && d.p >= d.nprop - g.nlearn) return;
    if (!(probe && d.p == 0)) {
        double u0, u1, u2, u3, u4, u5;
        const uint32_t sw = probe ? 0xFFFFFFFFu : (uint32_t)g.nsweep;   // the probe has its own stream
        philox_u01x2(g.seed, (uint32_t)li, (uint32_t)d.p, sw, TP_TAG_QAXIS, u0, u1);
        philox_u01x2(g.seed, (uint32_t)li, (uint32_t)d.p, sw, TP_TAG_QAXIS + 1u, u2, u3);
        philox_u01x2(g.seed, (uint32_t)li, (uint32_t)d.p, sw, TP_TAG_QANGLE, u4, u5);
        // Box-Muller normals -> uniform axis
        const double r0 = sqrt(-2.0 * log(1.0 - u0)), r1 = sqrt(-2.0 * log(1.0 - u2));
        double ax = r0 * cos(2.0 * kPiD * u1), ay = r0 * sin(2.0 * kPiD * u1), az = r1 * cos(2.0 * kPiD * u3);
        const double n = sqrt(ax * ax + ay * ay + az * az);
        if (n > 0.0) { ax /= n; ay /= n; az /= n; } else { ax = 0.0; ay = 0.0; az = 1.0; }
        const double th = haar_angle(u4, g.amax * 0.25 * kPiD);
        const double sh = sin(0.5 * th);
        q = make_double4(cos(0.5 * th), sh * ax, sh * ay, sh * az);
    }
    double *o = quat + (size_t)i * 4;
    o[0] = q.x; o[1] = q.y; o[2] = q.z; o[3] = q.w;
}

// ---------------------------------------------------------------------------
// Rotation of the l-blocks of all shells of an evaluation by quaternions, in
// FP64.  Workgroup (li, e): l = 2(li+1).  The real-SH rotation matrices D^1..
// D^l of R(q) are built by the recurrence of Ivanic and Ruedenberg (J. Phys.
// Chem. 100, 6342 (1996); 102, 9099 (1998)) in LDS; the coefficients are
// taken to the real basis
//   x_m = (s c_m + c_-m)/sqrt2,  x_-m = i (s c_m - c_-m)/sqrt2,  s = (-1)^m (m > 0),
// in which y(R r) = D^l(R) y(r) for the real harmonics y, rotated x' = D^l x
// (so f'(r) = f(R^-1 r)), and taken back: c_m = s (x_m - i x_-m)/sqrt2,
// c_-m = (x_m + i x_-m)/sqrt2.  Workgroup li = 0 also copies the l = 0 entry.
// ---------------------------------------------------------------------------
constexpr int kRotMaxL = 22;
constexpr int kRotShellChunk = 16;

// Move families of the random slots.
#define TP_MOVE_QUAT 0    // SO(3) rotation per l (D^l of a quaternion)
#define TP_MOVE_PLANE 1   // rotation in the plane of a Haar-random 2-frame of R^(2l+1)

// With gs != nullptr and family == TP_MOVE_PLANE, every slot (in probe mode
// all but slot 0) instead gets the reference's move (get_MC_step /
// random_rotation_step, determination.jl:232-241, utilities.jl:180-190,
// written as in k_propose):
//   D = I + (cos b - 1)(u u^T + v v^T) + sin b (v u^T - u v^T),
// (u, v) Gram-Schmidt of two Gaussian vectors; random slots take
// b = amax (pi/4)(l-1) z (the reference's step sizes (pi/4)(l-1),
// determination.jl:295), learned slots (p >= nprop - nlearn) the angle of
// their quaternion, 2 atan2(|(x,y,z)|, |w|), with a random sign; b is clipped
// to [-pi/4, pi/4].  The angle actually used is written back as the
// quaternion (cos |b|/2, sin |b|/2, 0, 0), the label the move log records.
// A Haar-random frame is Haar-random in any fixed orthonormal basis, so the
// real basis used here serves as well as the reference's Umat basis.
// x <- D x on the l-block of every shell of evaluation d (real basis x =
// Umat(l) c, spherical_harmonics.jl:445-457): d.src -> d.coef; the l = 2
// workgroup (li = 0) also copies the l = 0 entries.  D: n x n row-major in LDS.
__device__ inline void block_rotate(const double *D, int l, int li, const EvalDesc &d, int n_c,
                                    double2 (*xs)[2 * kRotMaxL + 1], double2 (*xo)[2 * kRotMaxL + 1])
{
    const int tid = threadIdx.x;
    const int n = 2 * l + 1, off = l * (l - 1) / 2;
    const double h = 0.70710678118654752440;
    if (li == 0)
        for (int k = tid; k < d.K; k += 256) d.coef[(size_t)k * n_c] = d.src[(size_t)k * n_c];
    for (int k0 = 0; k0 < d.K; k0 += kRotShellChunk) {
        const int nk = min(kRotShellChunk, d.K - k0);
        for (int i = tid; i < nk * n; i += 256) {
            const int kk = i / n, m = i % n - l;
            const double2 *c = d.src + (size_t)(k0 + kk) * n_c + off;
            const int am = m < 0 ? -m : m;
            const double s = (am & 1) ? -1.0 : 1.0;
            const double2 cp = c[l + am], cm = c[l - am];
            double2 x;
            if (m == 0) x = cp;
            else if (m > 0) x = make_double2(h * (s * cp.x + cm.x), h * (s * cp.y + cm.y));
            else x = make_double2(-h * (s * cp.y - cm.y), h * (s * cp.x - cm.x));   // i (s c_m - c_-m)/sqrt2
            xs[kk][m + l] = x;
        }
        __syncthreads();
        for (int i = tid; i < nk * n; i += 256) {
            const int kk = i / n, r = i % n;
            double ar = 0.0, ai = 0.0;
            for (int c = 0; c < n; ++c) {
                const double dv = D[r * n + c];
                ar += dv * xs[kk][c].x;
                ai += dv * xs[kk][c].y;
            }
            xo[kk][r] = make_double2(ar, ai);
        }
        __syncthreads();
        for (int i = tid; i < nk * n; i += 256) {
            const int kk = i / n, m = i % n - l;
            const int am = m < 0 ? -m : m;
            const double s = (am & 1) ? -1.0 : 1.0;
            const double2 xp = xo[kk][l + am], xm = xo[kk][l - am];
            double2 c;
            if (m == 0) c = xp;
            else if (m > 0) c = make_double2(h * s * (xp.x + xm.y), h * s * (xp.y - xm.x));   // s (x_m - i x_-m)/sqrt2
            else c = make_double2(h * (xp.x - xm.y), h * (xp.y + xm.x));                     // (x_m + i x_-m)/sqrt2
            d.coef[(size_t)(k0 + kk) * n_c + off + m + l] = c;
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void k_rotate_batch(const EvalDesc *__restrict__ ev, double *__restrict__ quat,
                                                      int nl, int n_c, const GroupState *__restrict__ gs, int family,
                                                      int probe)
{
    __shared__ double Dm[2][(2 * kRotMaxL + 1) * (2 * kRotMaxL + 1)];
    __shared__ double R1[9];
    __shared__ double2 xs[kRotShellChunk][2 * kRotMaxL + 1], xo[kRotShellChunk][2 * kRotMaxL + 1];
    __shared__ double uv[2][64];
    __shared__ double s_beta;
    const int li = blockIdx.x, e = blockIdx.y, l = 2 * (li + 1);
    const int tid = threadIdx.x;
    const EvalDesc d = ev[e];
    bool plane = false, learned = false;
    GroupState g;
    if (gs && family == TP_MOVE_PLANE) {
        g = gs[d.group];
        plane = !(probe && d.p == 0);
        learned = !probe && d.p >= d.nprop - g.nlearn;
    }
    if (plane) {
        const int n = 2 * l + 1;
        const uint32_t step = probe ? 0xFFFFFFFFu : (uint32_t)g.nsweep;
        const uint32_t lid = (uint32_t)li | ((uint32_t)d.p << 8);
        if (tid < 64) {
            const int lane = tid;
            const double g1 = lane < n ? dev_normal(g.seed, step, lid, (uint32_t)lane, TP_TAG_MOVE) : 0.0;
            const double g2 = lane < n ? dev_normal(g.seed, step, lid, (uint32_t)(n + lane), TP_TAG_MOVE) : 0.0;
            const double u = g1 / sqrt(wave_sum(g1 * g1));
            const double dd = wave_sum(u * g2);
            const double vt = g2 - dd * u;
            const double v = vt / sqrt(wave_sum(vt * vt));
            uv[0][lane] = lane < n ? u : 0.0;
            uv[1][lane] = lane < n ? v : 0.0;
            if (lane == 0) {
                const double q4 = 0.25 * kPiD;
                const double z = dev_normal(g.seed, step, lid, (uint32_t)(2 * n), TP_TAG_MOVE);
                double b;
                double *q = quat + ((size_t)e * nl + li) * 4;
                if (learned) {
                    const double vn = sqrt(q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
                    b = 2.0 * atan2(vn, fabs(q[0])) * (z < 0.0 ? -1.0 : 1.0);
                } else {
                    b = g.amax * q4 * (l - 1) * z;
                }
                b = b > q4 ? q4 : (b < -q4 ? -q4 : b);
                s_beta = b;
                q[0] = cos(0.5 * fabs(b));
                q[1] = sin(0.5 * fabs(b));
                q[2] = 0.0;
                q[3] = 0.0;
            }
        }
        __syncthreads();
        const double cm = cos(s_beta) - 1.0, sb = sin(s_beta);
        for (int idx = tid; idx < n * n; idx += 256) {
            const int r = idx / n, c = idx % n;
            const double ur = uv[0][r], vr = uv[1][r], uc = uv[0][c], vc = uv[1][c];
            Dm[0][idx] = (r == c ? 1.0 : 0.0) + cm * (ur * uc + vr * vc) + sb * (vr * uc - ur * vc);
        }
        __syncthreads();
    }
    if (!plane && tid == 0) {
        const double *q4 = quat + ((size_t)e * nl + li) * 4;
        double w = q4[0], x = q4[1], y = q4[2], z = q4[3];
        const double nq = sqrt(w * w + x * x + y * y + z * z);
        if (nq > 0.0) { w /= nq; x /= nq; y /= nq; z /= nq; } else { w = 1.0; x = y = z = 0.0; }
        double Rc[3][3] = {{1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)},
                           {2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)},
                           {2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)}};
        const int cix[3] = {1, 2, 0};   // real l = 1 harmonics m = -1, 0, 1 ~ y, z, x
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) R1[a * 3 + b] = Rc[cix[a]][cix[b]];
    }
    __syncthreads();
    if (!plane)
        for (int i = tid; i < 9; i += 256) Dm[1][i] = R1[i];
    __syncthreads();
    for (int lp = 2; lp <= (plane ? 1 : l); ++lp) {
        const double *pv = Dm[(lp - 1) & 1];
        double *cur = Dm[lp & 1];
        const int n = 2 * lp + 1, np = 2 * lp - 1;
        auto P = [&](int i, int a, int b) -> double {
            const double ri1 = R1[(i + 1) * 3 + 2], rim1 = R1[(i + 1) * 3 + 0], ri0 = R1[(i + 1) * 3 + 1];
            const int ra = (a + lp - 1) * np;
            if (b == lp) return ri1 * pv[ra + 2 * lp - 2] - rim1 * pv[ra];
            if (b == -lp) return ri1 * pv[ra] + rim1 * pv[ra + 2 * lp - 2];
            return ri0 * pv[ra + b + lp - 1];
        };
        for (int idx = tid; idx < n * n; idx += 256) {
            const int m = idx / n - lp, nn = idx % n - lp;
            const int am = m < 0 ? -m : m;
            const double dlt = m == 0 ? 1.0 : 0.0;
            const double den = (nn == lp || nn == -lp) ? (double)(2 * lp) * (2 * lp - 1) : (double)(lp + nn) * (lp - nn);
            const double u = sqrt((double)(lp + m) * (lp - m) / den);
            const double v = 0.5 * sqrt((1.0 + dlt) * (lp + am - 1) * (lp + am) / den) * (1.0 - 2.0 * dlt);
            const double w = -0.5 * sqrt((double)(lp - am - 1) * (lp - am) / den) * (1.0 - dlt);
            double s = 0.0;
            if (u != 0.0) s += u * P(0, m, nn);
            if (v != 0.0) {
                double V;
                if (m == 0) V = P(1, 1, nn) + P(-1, -1, nn);
                else if (m > 0) V = m == 1 ? P(1, 0, nn) * 1.4142135623730951 : P(1, m - 1, nn) - P(-1, -m + 1, nn);
                else V = m == -1 ? P(-1, 0, nn) * 1.4142135623730951 : P(1, m + 1, nn) + P(-1, -m - 1, nn);
                s += v * V;
            }
            if (w != 0.0) {
                const double W = m > 0 ? P(1, m + 1, nn) + P(-1, -m - 1, nn) : P(1, m - 1, nn) - P(-1, -m + 1, nn);
                s += w * W;
            }
            cur[idx] = s;
        }
        __syncthreads();
    }
    block_rotate(Dm[0], l, li, d, n_c, xs, xo);   // l is even: D^l is in Dm[0]
}

// Given rotations of a batch's proposals (tp_batch_eval_rotations), one
// workgroup per (block l, evaluation e).  Rotations are given in the basis
// x = Umat(l) c of the reference (spherical_harmonics.jl:445-457); the
// real basis of block_rotate differs from it by the signs s_u = (-1)^u
// (u > 0), -(-1)^|u| (u < 0), 1 (u = 0), so D is applied as S D S.
//   slot p < ngeo of run g: D = exp(-theta G) with G run g's real
//     skew-symmetric generator of block l (row-major, |G|_F <= 1) and
//     theta = theta[e][li]: A = -theta G / 2^s with |A|_F <= 1/4, the
//     degree-12 Taylor polynomial of exp(A) by Horner's rule (truncation
//     below 1/4^13 / 13! < 1e-18 of |exp(A)|), squared s times;
//   other slots: D = I + (cos b - 1)(u u^T + v v^T) + sin b (v u^T - u v^T)
//     from uvb[e][li] = (u[n], v[n], b): a rotation by b in the plane (u, v).
constexpr int kExpDegree = 12;

__device__ inline void lds_matmul(const double *A, const double *B, double *C, int n)
{
    for (int idx = threadIdx.x; idx < n * n; idx += blockDim.x) {
        const int r = idx / n, c = idx % n;
        double acc = 0.0;
        for (int j = 0; j < n; ++j) acc = fma(A[r * n + j], B[j * n + c], acc);
        C[idx] = acc;
    }
}

__global__ __launch_bounds__(256) void k_rotation_batch(const EvalDesc *__restrict__ ev, int nl, int n_c, int ngeo,
                                                        const double *__restrict__ gen, int64_t gstride,
                                                        const double *__restrict__ theta,
                                                        const double *__restrict__ uvb, int uvstride)
{
    constexpr int NM = (2 * kRotMaxL + 1) * (2 * kRotMaxL + 1);
    __shared__ double Dm[NM], Am[NM], Tm[NM];
    __shared__ double sg[2 * kRotMaxL + 1];
    __shared__ double2 xs[kRotShellChunk][2 * kRotMaxL + 1], xo[kRotShellChunk][2 * kRotMaxL + 1];
    const int li = blockIdx.x, e = blockIdx.y, l = 2 * (li + 1), n = 2 * l + 1;
    const int tid = threadIdx.x;
    const EvalDesc d = ev[e];
    for (int r = tid; r < n; r += 256) {
        const int u = r - l, a = u < 0 ? -u : u;
        const double par = (a & 1) ? -1.0 : 1.0;
        sg[r] = u > 0 ? par : (u < 0 ? -par : 1.0);
    }
    if (d.p < ngeo) {
        int64_t goff = 0;
        for (int j = 0; j < li; ++j) goff += (int64_t)(4 * j + 5) * (4 * j + 5);
        const double *G = gen + (int64_t)d.group * gstride + goff;
        const double th = theta[(int64_t)e * nl + li];
        int s = 0;
        for (double a = fabs(th); a > 0.25 && s < 64; a *= 0.5) ++s;
        const double sc = -ldexp(th, -s);
        for (int i = tid; i < n * n; i += 256) {
            Am[i] = sc * G[i];
            Dm[i] = (i / n == i % n ? 1.0 : 0.0) + Am[i] * (1.0 / kExpDegree);
        }
        __syncthreads();
        for (int k = kExpDegree - 1; k >= 1; --k) {   // Dm <- I + A Dm / k
            lds_matmul(Am, Dm, Tm, n);
            __syncthreads();
            const double ik = 1.0 / k;
            for (int i = tid; i < n * n; i += 256) Dm[i] = (i / n == i % n ? 1.0 : 0.0) + Tm[i] * ik;
            __syncthreads();
        }
        for (int q = 0; q < s; ++q) {
            lds_matmul(Dm, Dm, Tm, n);
            __syncthreads();
            for (int i = tid; i < n * n; i += 256) Dm[i] = Tm[i];
            __syncthreads();
        }
        for (int i = tid; i < n * n; i += 256) Dm[i] *= sg[i / n] * sg[i % n];
    } else {
        const double *q = uvb + ((int64_t)e * nl + li) * uvstride;
        const double b = q[2 * n], cm = cos(b) - 1.0, sb = sin(b);
        __syncthreads();
        for (int idx = tid; idx < n * n; idx += 256) {
            const int r = idx / n, c = idx % n;
            const double ur = q[r], vr = q[n + r], uc = q[c], vc = q[n + c];
            Dm[idx] = sg[r] * sg[c] * ((r == c ? 1.0 : 0.0) + cm * (ur * uc + vr * vc) + sb * (vr * uc - ur * vc));
        }
    }
    __syncthreads();
    block_rotate(Dm, l, li, d, n_c, xs, xo);
}

// ---------------------------------------------------------------------------
// Triple products of a batch: work item = (evaluation, l-triple | half << 16),
// most expensive first.
// ---------------------------------------------------------------------------
// Mixed-precision triple products of a batch (triple16_body); shell scales
// of evaluation e at shsc + e * kst.
__global__ __launch_bounds__(512) void k_triple16_batch(const EvalDesc *__restrict__ ev, const int2 *__restrict__ work,
                                                        const double *__restrict__ shsc, int kst, int n_c,
                                                        const float *__restrict__ ew,
                                                        const uint32_t *__restrict__ eidx,
                                                        const int *__restrict__ gstart,
                                                        const int4 *__restrict__ ltinfo, int Lw, int Emax, int Prows,
                                                        double *__restrict__ As, int64_t ldA)
{
    extern __shared__ __align__(16) char smem16b[];
    const int2 w = work[blockIdx.x];
    const EvalDesc &d = ev[w.x];
    triple16_body(smem16b, d.coef, shsc + (size_t)w.x * kst, n_c, d.K, w.y & 0xffff, w.y >> 16, ew, eidx, gstart,
                  ltinfo, Lw, Emax, Prows, As, ldA, d.cbase, d.nprop, d.p);
}

// Per-column power-of-two scales of an evaluation's As columns (the bound of
// k_split_as) and its shell scales (triple16_body), one workgroup per
// evaluation; the padding columns of its last tile get scale 0.
__global__ __launch_bounds__(256) void k_colscale_batch(const EvalDesc *__restrict__ ev, int n_c, int L,
                                                        double *__restrict__ sc, double *__restrict__ isc,
                                                        double *__restrict__ shsc, int kst)
{
    __shared__ double s_n2[64 * 12];
    __shared__ double s_nk[64];
    const EvalDesc d = ev[blockIdx.x];
    const int tid = threadIdx.x, K = d.K, nl = L / 2 + 1;
    for (int i = tid; i < K * nl; i += 256) {
        const int k = i / nl, l = 2 * (i % nl);
        const double2 *v = d.coef + (size_t)k * n_c + l * (l - 1) / 2;
        double s = 0.0;
        for (int m = 0; m <= 2 * l; ++m) s += v[m].x * v[m].x + v[m].y * v[m].y;
        s_n2[i] = s;
    }
    __syncthreads();
    if (tid < K) {
        double m = 0.0;
        for (int l = 0; l < nl; ++l) m = fmax(m, s_n2[tid * nl + l]);
        const double nk = sqrt(m);
        s_nk[tid] = nk;
        shsc[(size_t)blockIdx.x * kst + tid] = nk > 0.0 ? ldexp(1.0, 12 - ilogb(nk)) : 1.0;
    }
    __syncthreads();
    const int klength = K * (K + 1) * (K + 2) / 6, ncol = (klength + 63) / 64 * 64;
    for (int kt = tid; kt < ncol; kt += 256) {
        double s = 0.0;
        if (kt < klength) {
            int k1 = 1;
            while ((k1 + 1) * (k1 + 2) * k1 / 6 <= kt) ++k1;
            const int r1 = kt - (k1 - 1) * k1 * (k1 + 1) / 6;
            int k2 = 1;
            while ((k2 + 1) * k2 / 2 <= r1) ++k2;
            const int k3 = r1 - (k2 - 1) * k2 / 2 + 1;
            const double bnd = s_nk[k1 - 1] * s_nk[k2 - 1] * s_nk[k3 - 1];
            s = bnd > 0.0 ? ldexp(1.0, 13 - ilogb(bnd)) : 1.0;
        }
        const int64_t col = as_col(kt, d.cbase, d.nprop, d.p);
        sc[col] = s;
        isc[col] = s > 0.0 ? 1.0 / s : 0.0;
    }
}

// As [lt][col] (FP64, or FP32 from k_triple_sym_batch) -> scaled 16-bit
// copies [col][lt], one 64 x 64 tile per block.
template <int MODE, typename AT = double>
__global__ __launch_bounds__(256) void k_split_batch(const AT *__restrict__ As, int64_t ldA, int nlt, int ldk,
                                                     const double *__restrict__ sc, uint16_t *__restrict__ Ahi,
                                                     uint16_t *__restrict__ Alo)
{
    __shared__ double tile[64][65];
    const int64_t c0 = (int64_t)blockIdx.x * 64;
    const int lt0 = blockIdx.y * 64;
    const int tid = threadIdx.x, c = tid & 63, g = tid >> 6;
    for (int r = g; r < 64; r += 4) {
        const int lt = lt0 + r;
        tile[r][c] = lt < nlt ? (double)As[(int64_t)lt * ldA + c0 + c] : 0.0;
    }
    __syncthreads();
    const int kr = tid >> 2, q = tid & 3;
    const double s = sc[c0 + kr];
    u16x8 h0, h1, l0, l1;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const double x = tile[q * 16 + i][kr] * s;
        uint16_t hb, lb = 0;
        if (MODE == 1) {
            f16_split(x, hb, lb);
        } else {
            hb = bf16_bits(x);
        }
        if (i < 8) { h0[i] = hb; l0[i] = lb; } else { h1[i - 8] = hb; l1[i - 8] = lb; }
    }
    u16x8 *ph = reinterpret_cast<u16x8 *>(Ahi + (c0 + kr) * ldk + lt0 + q * 16);
    ph[0] = h0;
    ph[1] = h1;
    if (MODE == 1) {
        u16x8 *pl = reinterpret_cast<u16x8 *>(Alo + (c0 + kr) * ldk + lt0 + q * 16);
        pl[0] = l0;
        pl[1] = l1;
    }
}

// Energy of every evaluation from the contraction's per-row-tile partial
// sums: E = sum_kt tf (qsum log S - Q) over the K3_range walk
// (correlations.jl:203-220), one workgroup per evaluation, fixed-order sums.
__global__ __launch_bounds__(256) void k_final_batch(const EvalDesc *__restrict__ ev, const double *__restrict__ part,
                                                     int n_rt, int64_t ldp, double *__restrict__ Eout)
{
    __shared__ double red[256];
    const EvalDesc d = ev[blockIdx.x];
    const int tid = threadIdx.x;
    double acc = 0.0;
    for (int kt = tid; kt < d.ncols; kt += 256) {
        const int64_t col = as_col(kt, d.cbase, d.nprop, d.p);
        double S = 0.0, Q = 0.0;
        for (int q = 0; q < n_rt; ++q) {
            S += part[((int64_t)q * 2 + 0) * ldp + col];
            Q += part[((int64_t)q * 2 + 1) * ldp + col];
        }
        acc += d.tf[kt] * (d.qsum[kt] * log(S) - Q);
    }
    red[tid] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (tid < s) red[tid] += red[tid + s];
        __syncthreads();
    }
    if (tid == 0) Eout[blockIdx.x] = red[0];
}

// ---------------------------------------------------------------------------
// Sweep decision per annealing run (one workgroup each).
// probe: E <- energy of slot 0 (identity), T <- tfac * std(energies of slots
//        1..nprop-1) (calculate_temperature, determination.jl:244-261, with the
//        sweep's proposals as the samples).
// step : T *= decay; the lowest-energy proposal is accepted with the Metropolis
//        rule on the exact energy difference (evaluate_step,
//        determination.jl:491-515); on acceptance it becomes the run's state
//        and the angle bound grows by `factor` (capped at pi/4), otherwise it
//        shrinks.  The run's energy after the sweep goes to Ehist.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_select_batch(GroupState *__restrict__ gs, const EvalDesc *__restrict__ ev,
                                                      const double *__restrict__ Eprop, double2 *__restrict__ cur,
                                                      int64_t cstride, int n_c, int probe, double decay, double factor,
                                                      double tfac, double *__restrict__ Ehist, int hist_cap,
                                                      int ngroups)
{
    __shared__ double sE[256];
    __shared__ int sI[256];
    __shared__ int s_acc, s_best;
    const int g = blockIdx.x, tid = threadIdx.x;
    GroupState st = gs[g];
    const int np = st.nprop;
    if (probe) {
        // two-pass mean / variance over slots 1..np-1 in a fixed order
        if (tid == 0) {
            double m = 0.0;
            for (int p = 1; p < np; ++p) m += Eprop[st.e0 + p];
            m /= (np - 1);
            double v = 0.0;
            for (int p = 1; p < np; ++p) {
                const double dv = Eprop[st.e0 + p] - m;
                v += dv * dv;
            }
            v /= (np - 2 > 0 ? np - 2 : 1);
            st.E = Eprop[st.e0];
            st.T = tfac * sqrt(v);
            gs[g] = st;
        }
        return;
    }
    double best = 1e308;
    int bi = 0x7fffffff;
    for (int p = tid; p < np; p += 256) {
        const double e = Eprop[st.e0 + p];
        if (e < best) { best = e; bi = p; }
    }
    sE[tid] = best;
    sI[tid] = bi;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (tid < s) {
            const double e2 = sE[tid + s];
            const int i2 = sI[tid + s];
            if (e2 < sE[tid] || (e2 == sE[tid] && i2 < sI[tid])) { sE[tid] = e2; sI[tid] = i2; }
        }
        __syncthreads();
    }
    if (tid == 0) {
        st.T *= decay;
        const double dE = sE[0] - st.E;
        const double u = dev_uniform(st.seed, (uint32_t)st.nsweep, TP_TAG_ACC);
        const int acc = (dE < 0.0 || (dE > 0.0 && u < exp(-dE / st.T))) ? 1 : 0;
        if (acc) {
            st.E = sE[0];
            st.amax = fmin(1.0, st.amax * factor);
            st.nacc += 1;
        } else {
            st.amax /= factor;
        }
        if (st.nsweep < hist_cap) Ehist[(int64_t)st.nsweep * ngroups + g] = st.E;
        st.nsweep += 1;
        s_acc = acc;
        s_best = sI[0];
        gs[g] = st;
    }
    __syncthreads();
    if (s_acc) {
        const double2 *src = ev[st.e0 + s_best].coef;
        double2 *dst = cur + (int64_t)g * cstride;
        for (int i = tid; i < st.K * n_c; i += 256) dst[i] = src[i];
    }
}
